"""Warm-start bias, per-spin <Z_i>, current entropy and the GPU target
generation that the guidance network trains on.

The biased anneal replaces the single-site generator of mps.cpp:313-322 by
-A X - (B h_i + A b_i) Z and starts in the product ground state of
-sum_i (X_i + b_i Z_i); with b = 0 it is the plain anneal bit for bit.
"""
import numpy as np
import pytest

from tessera_amd import instances as I
from tessera_amd import learned as Lrn
from tessera_amd import qa

from test_gpu_tebd import _fid, dense_mps_splitting

pytestmark = pytest.mark.gpu


def _warm_dense(bias):
    """Product of the lowest eigenvectors of -(X + b Z) (numpy eigh), bit i = spin i."""
    psi = np.ones(1, complex)
    for b in bias:
        w, V = np.linalg.eigh(-np.array([[b, 1.0], [1.0, -b]]))
        v = V[:, 0] * np.sign(V[0, 0])
        psi = np.kron(v, psi)  # spin i is bit i: later spins are more significant
    return psi


def _dense_z(psi, n):
    p = np.abs(psi) ** 2
    x = np.arange(p.shape[0])
    return np.array([np.sum(p * (1 - 2 * ((x >> i) & 1))) for i in range(n)]) / p.sum()


@pytest.mark.parametrize("executor", [0, -1])
def test_zero_bias_reproduces_plain_bitwise(executor):
    model = I.chain_nnn(8, 31)
    sched = I.linear(12)
    out = []
    for warm in (False, True):
        s = qa.TebdSession(model, 16, 0.1)
        s.set_executor(executor)
        if warm:
            s.reset_warm(np.zeros(8))
            s.set_bias(np.zeros(8))
        s.run(sched.a, sched.b)
        out.append((s.dense(), s.sync()[0]))
        s.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("executor", [0, -1])
def test_biased_warm_anneal_matches_dense(executor):
    model = I.chain_nnn(8, 32)
    sched = I.linear(16)
    bias = np.random.default_rng(3).uniform(-1.5, 1.5, 8)
    s = qa.TebdSession(model, 16, 0.1)
    s.set_executor(executor)
    s.reset_warm(bias)
    s.set_bias(bias)
    s.run(sched.a, sched.b)
    psi = s.dense()
    ent = s.sync()[0]
    z = s.expect_z()
    e_now = s.entropy_now()
    s.close()
    ref, ents = dense_mps_splitting(model, sched, 0.1, bias=bias, psi0=_warm_dense(bias))
    inf = 1 - _fid(ref, psi)
    print(f"executor={executor} infidelity={inf:.2e}", flush=True)
    assert inf <= 1e-10
    assert np.abs(ent - ents).max() <= 1e-9
    assert np.abs(z - _dense_z(ref, 8)).max() <= 1e-9
    assert abs(e_now - ents[-1]) <= 1e-9


def test_batched_bias_matches_single_runs():
    cases = [(I.chain_nnn(7, 40), 0.1), (I.chain_range(9, 3, 41), 0.08), (I.sparse_banded(10, 3, 3.0, 42), 0.05)]
    rng = np.random.default_rng(4)
    biases = [rng.uniform(-1, 1, m.n) for m, _ in cases]
    steps = [10, 14, 8]
    solo = []
    for (m, dt), b, k in zip(cases, biases, steps):
        s = qa.TebdSession(m, 32, dt)
        s.reset_warm(b)
        s.set_bias(b)
        sch = I.linear(k)
        s.run(sch.a, sch.b, record=False)
        solo.append(s.dense())
        s.close()
    sess = [qa.TebdSession(m, 32, dt) for m, dt in cases]
    for s, b in zip(sess, biases):
        s.reset_warm(b)
        s.set_bias(b)
    qa.batch_run(sess, [I.linear(k) for k in steps])
    for s, ref in zip(sess, solo):
        assert 1 - _fid(ref, s.dense()) <= 1e-12
        s.close()


def test_targets_match_single_runs():
    insts = Lrn.guidance_instances(5, 900, n_lo=6, n_hi=10, chi=32, steps=20, dt=0.1)
    t = Lrn.tebd_targets(insts, chunks=4)
    for k, x in enumerate(insts):
        s = qa.TebdSession(x.model, x.chi, x.dt)
        s.run(x.sched.a, x.sched.b)
        ent = s.sync()[0]
        z = s.expect_z()
        _, sc = s.readout()
        s.close()
        assert np.abs(t.z[k] - z).max() <= 1e-9
        assert t.energy[k] == pytest.approx(sc[0], abs=1e-12)
        ends = [b - 1 for _, b in Lrn._chunks(20, 4)]
        assert np.abs(t.entropy[k] - ent[ends]).max() <= 1e-9
        assert t.s_peak[k] == pytest.approx(ends[int(np.argmax(ent[ends]))] / 19)
