"""Learned anneal guidance: a message-passing network trained on GPU TEBD runs.

The network (tessera_gnn_* in libtessera_qa.so, kernels in csrc/gnn.hip) reads
an Ising graph -- node features (h_i, deg_i / 3), edge feature J_ij -- and
predicts

* z_i: the final <Z_i> of the plain anneal (trained on <Z_i> that the GPU TEBD
  engine produces, tessera_tebd_expect_z), used as a per-spin longitudinal
  warm-start bias b_i = strength * z_i: the anneal starts in the product ground
  state of -sum_i (X_i + b_i Z_i) and runs with the single-site generator
  -A X - (B h_i + A b_i) Z, a field that vanishes at s = 1;
* s_g: the schedule fraction at which the qubit-0 entanglement entropy of the
  plain anneal peaks (trained on the GPU TEBD entropy trace), turned into a
  schedule stretch: an instance whose predicted peak comes later than the
  training mean gets proportionally more steps.

Plain TEBD stays the fallback and the reference point: bias None and stretch 1 give the
plain anneal exactly, and the evaluation compares guided energies against the
best plain energy. The structural template is python/gnn_guide.py:81-103
(IsingGNN); targets, heads and the guidance rule are this library's.
"""
from __future__ import annotations

import atexit
import ctypes
import time
import weakref
from dataclasses import dataclass

import numpy as np

from .instances import Instance, Ising, Rng, Schedule, linear, sparse_banded
from .qa import TebdSession, _check, _p, batch_run, lib


def _bind():
    L = lib()
    if getattr(L, "_gnn_bound", False):
        return L
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    L.tessera_gnn_create.restype = ctypes.c_int
    L.tessera_gnn_create.argtypes = [ctypes.c_uint64, vp]
    L.tessera_gnn_param_count.restype = sz
    L.tessera_gnn_param_count.argtypes = []
    L.tessera_gnn_get_params.restype = ctypes.c_int
    L.tessera_gnn_get_params.argtypes = [vp, vp]
    L.tessera_gnn_set_params.restype = ctypes.c_int
    L.tessera_gnn_set_params.argtypes = [vp, vp]
    L.tessera_gnn_set_data.restype = ctypes.c_int
    L.tessera_gnn_set_data.argtypes = [vp, sz, vp, vp, vp, vp, vp, vp, vp]
    L.tessera_gnn_predict.restype = ctypes.c_int
    L.tessera_gnn_predict.argtypes = [vp, sz, sz, vp, vp]
    L.tessera_gnn_grad.restype = ctypes.c_int
    L.tessera_gnn_grad.argtypes = [vp, sz, sz, vp, vp]
    L.tessera_gnn_train_step.restype = ctypes.c_int
    L.tessera_gnn_train_step.argtypes = [vp, sz, sz, ctypes.c_float, vp]
    L.tessera_gnn_destroy.restype = None
    L.tessera_gnn_destroy.argtypes = [vp]
    L._gnn_bound = True
    return L


# ---------------------------------------------------------------- graphs

def pack(models: list[Ising], tz=None, ts=None) -> dict:
    """Concatenate graphs: node_off (B+1), feat (N x 2: h, deg/3), symmetric
    adjacency adj_off/adj_nbr/adj_j with global node ids, optional targets
    tz (per node, list of arrays or one array) and ts (per graph)."""
    node_off = [0]
    feat, adj_off, nbr, jj = [], [0], [], []
    for model in models:
        n = model.n
        base = node_off[-1]
        lists = [[] for _ in range(n)]
        for i, j, w in zip(model.ei.tolist(), model.ej.tolist(), model.ew.tolist()):
            if i == j:
                continue
            lists[i].append((j, w))
            lists[j].append((i, w))
        for i in range(n):
            feat.append((model.h[i], len(lists[i]) / 3.0))
            for j, w in lists[i]:
                nbr.append(base + j)
                jj.append(w)
            adj_off.append(len(nbr))
        node_off.append(base + n)
    out = dict(node_off=np.asarray(node_off, np.int32), feat=np.asarray(feat, np.float32).reshape(-1),
               adj_off=np.asarray(adj_off, np.int32), adj_nbr=np.asarray(nbr, np.int32),
               adj_j=np.asarray(jj, np.float32))
    if tz is not None:
        out["tz"] = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).ravel() for t in tz])
                                         if isinstance(tz, (list, tuple)) else np.asarray(tz, np.float32))
        out["ts"] = np.ascontiguousarray(ts, np.float32)
        if out["tz"].shape[0] != node_off[-1] or out["ts"].shape[0] != len(models):
            raise ValueError("targets do not match the graphs")
    return out


def sub_batch(batch: dict, g0: int, g1: int) -> dict:
    """Graphs [g0, g1) of a packed batch with node ids renumbered from 0."""
    off = batch["node_off"]
    n0, n1 = int(off[g0]), int(off[g1])
    a0, a1 = int(batch["adj_off"][n0]), int(batch["adj_off"][n1])
    out = dict(node_off=(off[g0:g1 + 1] - n0).astype(np.int32), feat=batch["feat"][2 * n0:2 * n1].copy(),
               adj_off=(batch["adj_off"][n0:n1 + 1] - a0).astype(np.int32),
               adj_nbr=(batch["adj_nbr"][a0:a1] - n0).astype(np.int32), adj_j=batch["adj_j"][a0:a1].copy())
    if "tz" in batch:
        out["tz"] = batch["tz"][n0:n1].copy()
        out["ts"] = batch["ts"][g0:g1].copy()
    return out


# ---------------------------------------------------------------- network

# Networks still open at interpreter exit are destroyed by an atexit hook, before
# the HIP runtime's own exit handlers run.
_live_nets: "weakref.WeakSet[GuidanceNet]" = weakref.WeakSet()


def _close_live_nets():
    for g in list(_live_nets):
        g.close()


atexit.register(_close_live_nets)


class GuidanceNet:
    """The device network: FP32 master weights, BF16 MFMA forward/backward, Adam."""

    def __init__(self, seed: int = 0):
        L = _bind()
        h = ctypes.c_void_p()
        self._h = None
        _check(L.tessera_gnn_create(int(seed), ctypes.byref(h)))
        self._h = h
        self.count = int(L.tessera_gnn_param_count())
        self.graphs = 0
        self._batch = None
        _live_nets.add(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def close(self):
        if self._h:
            _bind().tessera_gnn_destroy(self._h)
            self._h = None
        _live_nets.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.count, np.float32)
        _check(_bind().tessera_gnn_get_params(self._h, _p(out)))
        return out

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.shape != (self.count,):
            raise ValueError(f"expected {self.count} parameters")
        _check(_bind().tessera_gnn_set_params(self._h, _p(w)))

    def set_data(self, batch: dict):
        tz = batch.get("tz")
        ts = batch.get("ts")
        B = batch["node_off"].shape[0] - 1
        _check(_bind().tessera_gnn_set_data(self._h, B, _p(batch["node_off"]), _p(batch["feat"]),
                                            _p(batch["adj_off"]), _p(batch["adj_nbr"]), _p(batch["adj_j"]),
                                            _p(tz), _p(ts)))
        self.graphs = B
        self._batch = batch  # the arrays stay referenced while the device copy is used

    def predict(self, g0: int = 0, g1: int | None = None):
        g1 = self.graphs if g1 is None else g1
        off = self._batch["node_off"]
        z = np.zeros(int(off[g1] - off[g0]), np.float32)
        s = np.zeros(g1 - g0, np.float32)
        _check(_bind().tessera_gnn_predict(self._h, g0, g1, _p(z), _p(s)))
        return z, s

    def grad(self, g0: int = 0, g1: int | None = None):
        g1 = self.graphs if g1 is None else g1
        g = np.zeros(self.count, np.float32)
        loss = ctypes.c_double()
        _check(_bind().tessera_gnn_grad(self._h, g0, g1, _p(g), ctypes.addressof(loss)))
        return loss.value, g

    def train_step(self, g0: int, g1: int, lr: float) -> float:
        loss = ctypes.c_double()
        _check(_bind().tessera_gnn_train_step(self._h, g0, g1, float(lr), ctypes.addressof(loss)))
        return loss.value

    def fit(self, epochs: int, batch_graphs: int = 64, lr: float = 2e-3, seed: int = 0) -> list[float]:
        """Adam over mini-batches of consecutive graphs in a shuffled block
        order; returns the mean loss of every epoch."""
        rng = np.random.default_rng(seed)
        starts = np.arange(0, self.graphs, batch_graphs)
        hist = []
        for _ in range(epochs):
            tot = 0.0
            for st in rng.permutation(starts):
                tot += self.train_step(int(st), int(min(st + batch_graphs, self.graphs)), lr)
            hist.append(tot / len(starts))
        return hist


# ---------------------------------------------------------------- data

def guidance_instances(count: int, base_seed: int, n_lo: int = 16, n_hi: int = 40, chi: int = 32,
                       steps: int = 200, dt: float = 0.05) -> list[Instance]:
    """Instances generated like the batched workload (banded sparse graphs of
    mean degree 3 and range <= 3, J ~ N(0,1), h ~ N(0,0.25)) with n ~ U{n_lo..n_hi}
    and one anneal length."""
    out = []
    for i in range(count):
        rng = Rng(base_seed + i)
        n = rng.randint(n_lo, n_hi)
        model = sparse_banded(n, 3, 3.0, 0, rng=rng)
        out.append(Instance(model, linear(steps), dt, chi=chi, name=f"guide_{base_seed + i}"))
    return out


@dataclass
class Targets:
    z: list            # per instance: final <Z_i> of the plain anneal
    s_peak: np.ndarray  # per instance: schedule fraction of the entropy peak
    energy: np.ndarray  # readout energy of the plain anneal
    entropy: list      # per instance: entropy at the end of every chunk
    seconds: float
    updates: int = 0   # two-site updates the anneals performed


def _chunks(steps: int, chunks: int):
    edges = np.linspace(0, steps, chunks + 1).round().astype(int)
    return [(int(a), int(b)) for a, b in zip(edges[:-1], edges[1:]) if b > a]


def tebd_targets(instances: list[Instance], chunks: int = 10, group: int = 128, groups: int = -1,
                 log=None) -> Targets:
    """Plain anneals of every instance on the GPU, batched; the entropy is
    sampled at the end of each of `chunks` schedule segments. `log` (optional)
    receives a line per finished chunk."""
    z, sp, en, ents = [], [], [], []
    updates = 0
    t0 = time.perf_counter()
    for k in range(0, len(instances), group):
        part = instances[k:k + group]
        sess = [TebdSession(x.model, x.chi, x.dt) for x in part]
        updates += sum(s.updates_per_step * x.sched.steps for s, x in zip(sess, part))
        spans = [_chunks(x.sched.steps, chunks) for x in part]
        trace = [[] for _ in part]
        for c in range(chunks):
            act = [i for i in range(len(part)) if c < len(spans[i])]
            scheds = [Schedule(part[i].sched.a[spans[i][c][0]:spans[i][c][1]],
                               part[i].sched.b[spans[i][c][0]:spans[i][c][1]]) for i in act]
            batch_run([sess[i] for i in act], scheds, groups)
            for i in act:
                trace[i].append(sess[i].entropy_now())
            if log:
                log(f"targets: instances {k}..{k + len(part) - 1}, chunk {c + 1}/{chunks}, "
                    f"{time.perf_counter() - t0:.1f}s")
        for i, s in enumerate(sess):
            z.append(s.expect_z())
            _, sc = s.readout()
            en.append(sc[0])
            tr = np.asarray(trace[i])
            ents.append(tr)
            peak_end = spans[i][int(np.argmax(tr))][1]
            sp.append((peak_end - 1) / max(part[i].sched.steps - 1, 1))
            s.close()
    return Targets(z=z, s_peak=np.asarray(sp), energy=np.asarray(en), entropy=ents,
                   seconds=time.perf_counter() - t0, updates=updates)


# ---------------------------------------------------------------- guidance

@dataclass
class Guidance:
    bias: list          # per instance (None for plain)
    steps: np.ndarray   # per instance
    infer_seconds: float


def guidance(net: GuidanceNet, models: list[Ising], base_steps, fraction: float, strength: float,
             s_mean: float, stretch_range=(0.5, 2.0)) -> Guidance:
    """Bias b = strength * z and steps = fraction * stretch * base_steps with
    stretch = clip(s_pred / s_mean, *stretch_range)."""
    t0 = time.perf_counter()
    batch = pack(models)
    net.set_data(batch)
    z, s = net.predict()
    secs = time.perf_counter() - t0
    off = batch["node_off"]
    bias = [strength * z[off[g]:off[g + 1]].astype(np.float64) for g in range(len(models))]
    stretch = np.clip(s / s_mean, *stretch_range)
    steps = np.maximum(2, np.round(fraction * stretch * np.asarray(base_steps)).astype(int))
    return Guidance(bias=bias, steps=steps, infer_seconds=secs)


def anneal_batch(instances: list[Instance], steps, bias=None, group: int = 128, groups: int = -1, log=None):
    """Batched GPU anneals on linear schedules of the given lengths; with bias,
    warm-started and biased. Returns (readout energies, spins, wall seconds)."""
    energies, spins = [], []
    t0 = time.perf_counter()
    for k in range(0, len(instances), group):
        part = instances[k:k + group]
        sess = [TebdSession(x.model, x.chi, x.dt) for x in part]
        for i, s in enumerate(sess):
            if bias is not None and bias[k + i] is not None:
                s.reset_warm(bias[k + i])
                s.set_bias(bias[k + i])
        batch_run(sess, [linear(int(steps[k + i])) for i in range(len(part))], groups)
        for s in sess:
            sp, sc = s.readout()
            energies.append(sc[0])
            spins.append(sp)
            s.close()
        if log:
            log(f"anneals: {k + len(part)}/{len(instances)} done, {time.perf_counter() - t0:.1f}s")
    return np.asarray(energies), spins, time.perf_counter() - t0


def within(energy, best, tol: float = 0.01):
    """E within tol of the best energy (energies are negative or zero)."""
    energy = np.asarray(energy)
    best = np.asarray(best)
    return energy <= best + tol * np.abs(best) + 1e-12
