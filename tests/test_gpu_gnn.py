"""GPU guidance network (csrc/gnn.hip) vs its CPU oracle (oracle/gnn_ref.py).

The oracle with BF16 operand rounding (bf=True) computes what the MFMA path
computes, in FP64: typical outputs agree to FP32 accumulation level (~1e-8).
The GPU rounds FP32 activations to BF16 and the oracle rounds FP64 ones, so a
few of the ~1e5 rounded elements land on the other side of a rounding
boundary (one BF16 ulp, 2^-8 relative); the maxima allow for those. Against the
plain FP64 oracle the gap is BF16-sized.
"""
import numpy as np
import pytest

import gnn_ref as R
from tessera_amd import instances as I
from tessera_amd import learned as Lrn

pytestmark = pytest.mark.gpu


def _data(count=40, seed=11, n_lo=6, n_hi=30):
    rng = I.Rng(seed)
    models = [I.sparse_banded(rng.randint(n_lo, n_hi), 3, 3.0, 0, rng=rng) for _ in range(count)]
    batch = Lrn.pack(models)
    r = np.random.default_rng(seed)
    batch["tz"] = r.uniform(-1, 1, batch["node_off"][-1]).astype(np.float32)
    batch["ts"] = r.uniform(0, 1, count).astype(np.float32)
    return batch


def _blocks():
    out = {"W_in": (R.W_IN, R.B_IN), "b_in": (R.B_IN, R.B_IN + R.F), "w_z": (R.W_Z, R.C_Z + 1),
           "w_s": (R.W_S, R.C_S + 1)}
    for l in range(R.L):
        b = R.layer_base(l)
        out.update({f"W_sd{l}": (b + R.W_SD, b + R.W_J), f"w_J{l}": (b + R.W_J, b + R.B_M),
                    f"b_m{l}": (b + R.B_M, b + R.W_U), f"W_u{l}": (b + R.W_U, b + R.B_U),
                    f"b_u{l}": (b + R.B_U, b + R.LAYER_SIZE)})
    return out


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_forward_matches_oracle():
    batch = _data()
    net = Lrn.GuidanceNet(seed=5)
    net.set_data(batch)
    w = net.params.astype(np.float64)
    z, s = net.predict()
    zo, so = R.forward(w, batch, bf=True)
    zf, sf = R.forward(w, batch, bf=False)
    print(f"z err bf16-oracle {np.abs(z - zo).max():.2e} fp64-oracle {np.abs(z - zf).max():.2e}; "
          f"s err {np.abs(s - so).max():.2e} / {np.abs(s - sf).max():.2e}", flush=True)
    assert np.median(np.abs(z - zo)) <= 1e-6
    assert np.abs(z - zo).max() <= 3e-3
    assert np.abs(s - so).max() <= 1e-3
    assert np.abs(z - zf).max() <= 3e-2


def test_gradient_matches_oracle():
    batch = _data()
    net = Lrn.GuidanceNet(seed=6)
    net.set_data(batch)
    w = net.params.astype(np.float64)
    loss, g = net.grad()
    lo, go = R.loss_and_grad(w, batch, bf=True)
    lf, gf = R.loss_and_grad(w, batch, bf=False)
    assert loss == pytest.approx(lo, rel=1e-4)
    worst = 0.0
    for name, (a, b) in _blocks().items():
        e = _rel(g[a:b].astype(np.float64), go[a:b])
        ef = _rel(g[a:b].astype(np.float64), gf[a:b])
        print(f"{name:6s} rel err vs bf16 oracle {e:.2e}, vs fp64 oracle {ef:.2e}", flush=True)
        worst = max(worst, e)
    for name, (a, b) in _blocks().items():
        assert _rel(g[a:b].astype(np.float64), gf[a:b]) <= 5e-2, name
    assert worst <= 2e-2
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos >= 0.99999


def test_gradient_is_deterministic_and_range_local():
    """Same batch, same bits; graphs [g0, g1) of a dataset give the bits of the
    same graphs uploaded on their own."""
    batch = _data(count=30, seed=12)
    net = Lrn.GuidanceNet(seed=7)
    net.set_data(batch)
    l1, g1 = net.grad(10, 25)
    l2, g2 = net.grad(10, 25)
    assert l1 == l2 and np.array_equal(g1, g2)
    other = Lrn.GuidanceNet(seed=99)
    other.params = net.params
    other.set_data(Lrn.sub_batch(batch, 10, 25))
    l3, g3 = other.grad()
    assert l3 == l1 and np.array_equal(g3, g1)
    z, s = net.predict(10, 25)
    z3, s3 = other.predict()
    assert np.array_equal(z, z3) and np.array_equal(s, s3)


def test_adam_step_matches_oracle():
    batch = _data(count=20, seed=13)
    net = Lrn.GuidanceNet(seed=8)
    net.set_data(batch)
    w0 = net.params.astype(np.float64)
    m = np.zeros(R.COUNT)
    v = np.zeros(R.COUNT)
    w = w0
    for t in (1, 2):
        _, g = net.grad()
        net.train_step(0, net.graphs, 1e-3)
        w, m, v = R.adam(w, g.astype(np.float64), m, v, t, 1e-3)
        got = net.params.astype(np.float64)
        assert np.abs(got - w).max() <= 1e-6, t


def test_training_reduces_loss():
    """A learnable target (z = tanh(2h + sum_j J_ij h_j), s = mean |h| + 0.2):
    thirty epochs of Adam lower the loss at least fourfold."""
    rng = I.Rng(21)
    models = [I.sparse_banded(rng.randint(8, 24), 3, 3.0, 0, rng=rng) for _ in range(128)]
    tz, ts = [], []
    for m in models:
        f = 2.0 * m.h.copy()
        for i, j, w in zip(m.ei.tolist(), m.ej.tolist(), m.ew.tolist()):
            f[i] += w * m.h[j]
            f[j] += w * m.h[i]
        tz.append(np.tanh(f))
        ts.append(np.abs(m.h).mean() + 0.2)
    batch = Lrn.pack(models, tz, ts)
    net = Lrn.GuidanceNet(seed=9)
    net.set_data(batch)
    l0, _ = net.grad()
    hist = net.fit(epochs=30, batch_graphs=16, lr=2e-3, seed=1)
    l1, _ = net.grad()
    print(f"loss {l0:.4f} -> {l1:.4f} (epoch means {hist[0]:.4f} .. {hist[-1]:.4f})", flush=True)
    assert l1 <= 0.25 * l0


def test_bad_arguments_raise():
    net = Lrn.GuidanceNet(seed=1)
    batch = _data(count=3)
    bad = dict(batch)
    bad["adj_nbr"] = batch["adj_nbr"].copy()
    bad["adj_nbr"][0] = batch["node_off"][-1] - 1  # neighbour in another graph
    with pytest.raises(Exception):
        net.set_data(bad)
    net.set_data(batch)
    with pytest.raises(Exception):
        net.predict(2, 5)
    nt = {k: v for k, v in batch.items() if k not in ("tz", "ts")}
    net.set_data(nt)
    net.predict()
    with pytest.raises(Exception):
        net.grad()
