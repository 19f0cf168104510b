"""CPU checks of the guidance network's oracle and of the graph packing.

The oracle's analytic gradient is pinned by central finite differences of its
own forward pass (FP64), its BF16 rounding by hand-worked values, and Adam by
its closed form. The GPU network is then compared against this oracle in
tests/test_gpu_gnn.py.
"""
import numpy as np
import pytest

import gnn_ref as R
from tessera_amd import instances as I
from tessera_amd import learned as Lrn


def _batch(seed=0, count=3):
    models = [I.sparse_banded(n, 3, 3.0, seed + k) for k, n in enumerate([5, 7, 6][:count])]
    rng = np.random.default_rng(seed)
    batch = Lrn.pack(models)
    n = batch["node_off"][-1]
    batch["tz"] = rng.uniform(-1, 1, n).astype(np.float32)
    batch["ts"] = rng.uniform(0, 1, count).astype(np.float32)
    return batch


def _params(seed=1):
    """Weights ~ N(0, 1/fan_in) and non-zero biases, so that no output saturates
    and the relu kinks are exercised both ways."""
    rng = np.random.default_rng(seed)
    w = np.zeros(R.COUNT)
    w[R.W_IN:R.B_IN] = rng.normal(0, 0.7, 2 * R.F)
    w[R.B_IN:R.B_IN + R.F] = rng.normal(0, 0.1, R.F)
    for l in range(R.L):
        b = R.layer_base(l)
        w[b + R.W_SD:b + R.W_J] = rng.normal(0, R.F ** -0.5, 2 * R.F * R.F)
        w[b + R.W_J:b + R.B_M] = rng.normal(0, 0.5, R.F)
        w[b + R.B_M:b + R.W_U] = rng.normal(0, 0.2, R.F)
        w[b + R.W_U:b + R.B_U] = rng.normal(0, (2 * R.F) ** -0.5, 2 * R.F * R.F)
        w[b + R.B_U:b + R.LAYER_SIZE] = rng.normal(0, 0.1, R.F)
    w[R.W_Z:R.C_Z] = rng.normal(0, 0.1, R.F)
    w[R.W_S:R.C_S] = rng.normal(0, 0.1, R.F)
    return w


def test_layout_matches_library():
    from tessera_amd.learned import _bind
    assert int(_bind().tessera_gnn_param_count()) == R.COUNT == 50050


def _fd_sample(w):
    rng = np.random.default_rng(7)
    blocks = [(R.W_IN, R.B_IN + R.F), (R.W_Z, R.COUNT)]
    for l in range(R.L):
        b = R.layer_base(l)
        blocks += [(b + R.W_SD, b + R.W_J), (b + R.W_J, b + R.B_M), (b + R.B_M, b + R.W_U),
                   (b + R.W_U, b + R.B_U), (b + R.B_U, b + R.LAYER_SIZE)]
    return np.concatenate([rng.choice(np.arange(a, e), size=min(6, e - a), replace=False) for a, e in blocks])


def _fd_errors(w, batch, g, idx, eps=1e-6):
    errs = []
    for k in idx:
        wp = w.copy()
        wm = w.copy()
        wp[k] += eps
        wm[k] -= eps
        fd = (R.loss_of(wp, batch) - R.loss_of(wm, batch)) / (2 * eps)
        errs.append(abs(fd - g[k]) / (1e-6 * max(1.0, abs(fd)) + 1e-8))
    return np.asarray(errs)  # <= 1 passes


def test_oracle_gradient_matches_finite_differences():
    batch = _batch()
    w = _params()
    loss, g = R.loss_and_grad(w, batch)
    assert loss == pytest.approx(R.loss_of(w, batch), rel=1e-14)
    errs = _fd_errors(w, batch, g, _fd_sample(w))
    assert errs.max() <= 1.0, errs.max()


@pytest.mark.parametrize("plant", ["drop_residual", "drop_source"])
def test_gradient_check_catches_planted_mistake(plant):
    """The same finite-difference check fails for a gradient with a dropped term."""
    batch = _batch()
    w = _params()
    _, g = R.loss_and_grad(w, batch, plant=plant)
    errs = _fd_errors(w, batch, g, _fd_sample(w))
    assert errs.max() > 100.0


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 1.0 + 2.0 ** -7 + 2.0 ** -8, 3.0e-3],
                 np.float32)
    got = R.bf16(x)
    # ties go to even: 1 + 2^-8 -> 1; 1 + 3*2^-8 -> 1 + 2^-6; 1 + 2^-7 + 2^-8 -> 1 + 2^-6;
    # 3e-3 lies in [2^-9, 2^-8) where the BF16 spacing is 2^-16: 3e-3 / 2^-16 = 196.6 -> 197
    exp = np.array([1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 1.0 + 2.0 ** -6, 197 * 2.0 ** -16])
    assert np.array_equal(got, exp)


def test_adam_closed_form():
    w = np.array([1.0])
    g = np.array([0.25])
    m = np.zeros(1)
    v = np.zeros(1)
    w1, m1, v1 = R.adam(w, g, m, v, 1, 0.01)
    # first step: m_hat = g, v_hat = g^2 -> w - lr * g / (|g| + eps)
    assert w1[0] == pytest.approx(1.0 - 0.01 * 0.25 / (0.25 + 1e-8), rel=1e-15)
    w2, _, _ = R.adam(w1, g, m1, v1, 2, 0.01)
    assert w2[0] == pytest.approx(1.0 - 2 * 0.01 * 0.25 / (0.25 + 1e-8), rel=1e-12)


def test_pack_adjacency_and_sub_batch():
    models = [I.sparse_banded(n, 3, 3.0, 40 + n) for n in (6, 9, 4)]
    b = Lrn.pack(models)
    assert b["node_off"].tolist() == [0, 6, 15, 19]
    for g, model in enumerate(models):
        base = b["node_off"][g]
        deg = np.zeros(model.n)
        for i, j in zip(model.ei.tolist(), model.ej.tolist()):
            deg[i] += 1
            deg[j] += 1
        assert np.allclose(b["feat"].reshape(-1, 2)[base:base + model.n, 1], deg / 3.0)
        assert np.allclose(b["feat"].reshape(-1, 2)[base:base + model.n, 0], model.h.astype(np.float32))
    # symmetric: (u in adj(v)) <=> (v in adj(u)) with the same coupling
    pairs = {}
    for v in range(19):
        for e in range(b["adj_off"][v], b["adj_off"][v + 1]):
            pairs[(int(b["adj_nbr"][e]), v)] = float(b["adj_j"][e])
    for (u, v), j in pairs.items():
        assert pairs[(v, u)] == j
    b["tz"] = np.arange(19, dtype=np.float32)
    b["ts"] = np.arange(3, dtype=np.float32)
    s = Lrn.sub_batch(b, 1, 3)
    ref = Lrn.pack(models[1:], tz=b["tz"][6:], ts=b["ts"][1:])
    for k in ("node_off", "feat", "adj_off", "adj_nbr", "adj_j", "tz", "ts"):
        assert np.array_equal(s[k], ref[k]), k


def test_within_rule():
    best = np.array([-10.0, -2.0])
    assert Lrn.within([-9.91, -1.97], best).tolist() == [True, False]
