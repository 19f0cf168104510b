"""Seeded Ising instances and annealing schedules.

The model is the transverse-field Ising annealer of TESSERA:

    H(s) = -A(s) * sum_i X_i  -  B(s) * ( sum_<ij> J_ij Z_i Z_j + sum_i h_i Z_i )

with the classical energy E(s) = -sum_e J_e s_i s_j - sum_i h_i s_i summed over
the edge list in order, then the fields in order (rust/src/ising.rs:23-32).

Schedules sample s_k = k/(steps-1) including both ends (rust/src/schedule.rs:21-46).

Randomness comes from a SplitMix64-seeded xoshiro256** stream (uniforms from the
top 53 bits, normals by Box-Muller), so the CPU oracle and the GPU library consume
byte-identical instances. Edge lists are sorted lexicographically with i < j,
which fixes the gate order of the swap network.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_MASK = (1 << 64) - 1


def _rotl(x: int, k: int) -> int:
    return ((x << k) | (x >> (64 - k))) & _MASK


class Rng:
    """xoshiro256** seeded through SplitMix64."""

    def __init__(self, seed: int):
        z = seed & _MASK
        s = []
        for _ in range(4):
            z = (z + 0x9E3779B97F4A7C15) & _MASK
            t = z
            t = ((t ^ (t >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
            t = ((t ^ (t >> 27)) * 0x94D049BB133111EB) & _MASK
            s.append(t ^ (t >> 31))
        self.s = s
        self._spare = None

    def next_u64(self) -> int:
        s = self.s
        result = (_rotl((s[1] * 5) & _MASK, 7) * 9) & _MASK
        t = (s[1] << 17) & _MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (inclusive)."""
        span = hi - lo + 1
        return lo + int(self.uniform() * span) % span

    def normal(self, mean: float = 0.0, std: float = 1.0) -> float:
        if self._spare is not None:
            z, self._spare = self._spare, None
            return mean + std * z
        while True:
            u1 = self.uniform()
            if u1 > 0.0:
                break
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        self._spare = r * math.sin(2.0 * math.pi * u2)
        return mean + std * r * math.cos(2.0 * math.pi * u2)


@dataclass
class Ising:
    """Fields h (length n) and couplings (ei[e], ej[e], ew[e]) with ei < ej."""

    h: np.ndarray
    ei: np.ndarray
    ej: np.ndarray
    ew: np.ndarray

    @property
    def n(self) -> int:
        return int(self.h.shape[0])

    @property
    def m(self) -> int:
        return int(self.ew.shape[0])

    def energy(self, spins) -> float:
        """E(s) in the reference summation order (edges in list order, then fields)."""
        s = [int(x) for x in spins]
        e = 0.0
        for i, j, w in zip(self.ei.tolist(), self.ej.tolist(), self.ew.tolist()):
            e -= w * float(s[i]) * float(s[j])
        for hi, si in zip(self.h.tolist(), s):
            e -= hi * float(si)
        return e

    def max_range(self) -> int:
        return int(np.max(self.ej - self.ei)) if self.m else 0


def make_ising(h, edges) -> Ising:
    """Build an Ising model from fields and (i, j, w) triples (kept in the given order)."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    if edges:
        ei = np.array([int(e[0]) for e in edges], dtype=np.uint64)
        ej = np.array([int(e[1]) for e in edges], dtype=np.uint64)
        ew = np.array([float(e[2]) for e in edges], dtype=np.float64)
    else:
        ei = np.zeros(0, np.uint64)
        ej = np.zeros(0, np.uint64)
        ew = np.zeros(0, np.float64)
    return Ising(h=h, ei=ei, ej=ej, ew=ew)


@dataclass
class Schedule:
    a: np.ndarray
    b: np.ndarray

    @property
    def steps(self) -> int:
        return int(self.a.shape[0])


def linear(steps: int) -> Schedule:
    """A(s) = 1 - s, B(s) = s (rust/src/schedule.rs:21-31)."""
    n = max(steps, 2)
    a = np.empty(n)
    b = np.empty(n)
    for k in range(n):
        s = k / (n - 1)
        a[k] = 1.0 - s
        b[k] = s
    return Schedule(a, b)


def smooth(steps: int) -> Schedule:
    """A(s) = cos^2(pi s/2), B(s) = sin^2(pi s/2) (rust/src/schedule.rs:35-46)."""
    n = max(steps, 2)
    a = np.empty(n)
    b = np.empty(n)
    for k in range(n):
        s = k / (n - 1)
        th = (math.pi / 2) * s
        c, sn = math.cos(th), math.sin(th)
        a[k] = c * c
        b[k] = sn * sn
    return Schedule(a, b)


@dataclass
class Instance:
    """One annealing run: model, schedule, Trotter step and bond dimension."""

    model: Ising
    sched: Schedule
    dt: float
    chi: int = 64
    zz_substeps: int = 1
    name: str = ""
    meta: dict = field(default_factory=dict)


def chain_nnn(n: int, seed: int) -> Ising:
    """Chain with nearest and next-nearest couplings J ~ U[-1,1], h ~ U[-0.5,0.5]."""
    rng = Rng(seed)
    h = [rng.uniform(-0.5, 0.5) for _ in range(n)]
    edges = []
    for i in range(n):
        for r in (1, 2):
            if i + r < n:
                edges.append((i, i + r, rng.uniform(-1.0, 1.0)))
    edges.sort(key=lambda t: (t[0], t[1]))
    return make_ising(h, edges)


def random_regular3(n: int, seed: int) -> Ising:
    """Random 3-regular graph (pairing model with rejection), J ~ N(0,1), h ~ N(0,0.25)."""
    if (3 * n) % 2:
        raise ValueError("3-regular graph needs even n")
    rng = Rng(seed)
    while True:
        stubs = [v for v in range(n) for _ in range(3)]
        for k in range(len(stubs) - 1, 0, -1):  # Fisher-Yates
            j = rng.randint(0, k)
            stubs[k], stubs[j] = stubs[j], stubs[k]
        pairs = set()
        ok = True
        for k in range(0, len(stubs), 2):
            i, j = stubs[k], stubs[k + 1]
            if i == j:
                ok = False
                break
            key = (min(i, j), max(i, j))
            if key in pairs:
                ok = False
                break
            pairs.add(key)
        if ok:
            break
    h = [rng.normal(0.0, 0.5) for _ in range(n)]
    edges = [(i, j, rng.normal(0.0, 1.0)) for (i, j) in sorted(pairs)]
    return make_ising(h, edges)


def chain_range(n: int, rmax: int, seed: int) -> Ising:
    """Edges (i, i+r) for r = 1..rmax, J ~ U[-1,1]/r, h ~ U[-0.5,0.5]."""
    rng = Rng(seed)
    h = [rng.uniform(-0.5, 0.5) for _ in range(n)]
    edges = []
    for i in range(n):
        for r in range(1, rmax + 1):
            if i + r < n:
                edges.append((i, i + r, rng.uniform(-1.0, 1.0) / r))
    edges.sort(key=lambda t: (t[0], t[1]))
    return make_ising(h, edges)


def sparse_banded(n: int, rmax: int, mean_degree: float, seed: int, rng: Rng | None = None) -> Ising:
    """Random sparse graph restricted to |i-j| <= rmax with the given mean degree,
    J ~ N(0,1), h ~ N(0,0.25)."""
    rng = rng or Rng(seed)
    cand = [(i, i + r) for i in range(n) for r in range(1, rmax + 1) if i + r < n]
    p = min(1.0, mean_degree * n / 2.0 / max(1, len(cand)))
    h = [rng.normal(0.0, 0.5) for _ in range(n)]
    edges = []
    for (i, j) in cand:
        if rng.uniform() < p:
            edges.append((i, j, rng.normal(0.0, 1.0)))
    if not edges:
        edges.append((0, 1, rng.normal(0.0, 1.0)))
    edges.sort(key=lambda t: (t[0], t[1]))
    return make_ising(h, edges)


def chain12_anneal(seed: int = 1) -> Instance:
    """n=12 chain (nn + nnn), linear schedule, T=10, dt=0.05 (200 steps), chi=64 (exact)."""
    return Instance(chain_nnn(12, seed), linear(200), 0.05, chi=64, name="tfim_n12_chain_nnn")


def regular3_n30_anneal(seed: int = 2, steps: int = 1000) -> Instance:
    """n=30 random 3-regular, linear schedule, T=20, dt=0.02 (1000 steps)."""
    return Instance(random_regular3(30, seed), linear(steps), 0.02, chi=0, name="sv_n30_3regular")


def chain128_range4_anneal(seed: int = 3, steps: int = 1000, n: int = 128, chi: int = 256) -> Instance:
    """n=128 chain with couplings of range <= 4, linear schedule, T=40, dt=0.04, chi=256."""
    return Instance(chain_range(n, 4, seed), linear(steps), 0.04, chi=chi, name=f"tebd_n{n}_range4_chi{chi}")


def batch_instances(count: int = 512, base_seed: int = 4000) -> list[Instance]:
    """Batched instances: n ~ U{16..96}, chi in {32,64,128}, banded sparse graphs of mean
    degree 3 and range <= 3, own T ~ U[10,40] at dt = 0.05."""
    out = []
    for i in range(count):
        rng = Rng(base_seed + i)
        n = rng.randint(16, 96)
        chi = (32, 64, 128)[rng.randint(0, 2)]
        T = rng.uniform(10.0, 40.0)
        steps = max(2, int(round(T / 0.05)))
        model = sparse_banded(n, 3, 3.0, 0, rng=rng)
        out.append(Instance(model, linear(steps), 0.05, chi=chi, name=f"batch_{i}"))
    return out
