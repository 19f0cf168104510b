"""Lockstep TEBD executor vs the per-update launch path and the exact evolution.

The executor runs a whole `run()` call as one cooperative launch: groups of
workgroups run the two-site updates of several Trotter steps in super-steps of
mutually independent updates (consecutive steps overlap), with the single-site
half layers folded into the two-site gates and the per-step entropy taken from
site snapshots. It computes the same Trotter splitting (mps.cpp:309-356), so at
exact chi it must reproduce the dense evolution to the bars of the per-update
path, and under truncation keep the same ranks.
"""
import numpy as np
import pytest

import pyoracle
from tessera_amd import instances as I
from tessera_amd import qa

from test_gpu_tebd import _fid, dense_mps_splitting

pytestmark = pytest.mark.gpu


def _run(model, sched, dt, chi, groups, ns=1, chunks=1, record=True):
    s = qa.TebdSession(model, chi, dt, ns)
    s.set_executor(groups)
    edges = np.linspace(0, sched.steps, chunks + 1).astype(int)
    for lo, hi in zip(edges[:-1], edges[1:]):
        s.run(sched.a[lo:hi], sched.b[lo:hi], record=record)
    ent, tr, bonds = s.sync()
    spins, sc = s.readout()
    state = s.dense() if model.n <= 16 else None
    s.close()
    return dict(entropy=ent, max_trunc=tr, bonds=bonds, spins=spins, expected_energy=sc[1], state=state)


@pytest.mark.parametrize("groups", [1, 8, -1])
def test_executor_exact_chi_matches_dense(groups):
    model = I.chain_nnn(8, 48)
    sched = I.linear(20)
    g = _run(model, sched, 0.1, 16, groups)
    psi, ents = dense_mps_splitting(model, sched, 0.1)
    inf = 1 - _fid(psi, g["state"])
    err = np.abs(g["entropy"] - ents).max()
    print(f"groups={groups} infidelity={inf:.2e} ent_err={err:.2e}", flush=True)
    assert inf <= 1e-10
    assert err <= 1e-9
    assert g["max_trunc"] <= 1e-20


def test_executor_chunks_substeps_and_no_record():
    """Two run() calls with two ZZ slices per step equal one stream-path run;
    a run without recording leaves the same state."""
    model = I.chain_range(9, 3, 5)
    sched = I.linear(16)
    base = _run(model, sched, 0.08, 32, 0, ns=2)
    ex = _run(model, sched, 0.08, 32, -1, ns=2, chunks=2)
    quiet = _run(model, sched, 0.08, 32, 4, ns=2, record=False)
    print(f"chunks infidelity={1 - _fid(base['state'], ex['state']):.2e}", flush=True)
    assert ex["entropy"].shape == base["entropy"].shape == (16,)
    assert 1 - _fid(base["state"], ex["state"]) <= 1e-12
    assert np.abs(base["entropy"] - ex["entropy"]).max() <= 1e-10
    assert np.array_equal(base["bonds"], ex["bonds"])
    assert 1 - _fid(base["state"], quiet["state"]) <= 1e-12
    assert quiet["entropy"].shape == (0,)


def test_executor_truncated_matches_accurate_oracle():
    """chi = 8 on a range-3 chain of 12 spins: ranks and spins identical to the
    accurate oracle (svd_mode 1), entropy and <H> to the truncated-run bars."""
    model = I.chain_range(12, 3, 17)
    sched = I.linear(20)
    g = _run(model, sched, 0.1, 8, -1)
    o = pyoracle.tebd(model, sched, 0.1, 8, svd_mode=1)
    print(f"trunc gpu={g['max_trunc']:.6e} oracle={o['max_trunc']:.6e} "
          f"ent_err={np.abs(g['entropy'] - o['entropy']).max():.2e}", flush=True)
    assert int(g["bonds"].max()) == int(o["max_bond"][-1])
    assert np.abs(g["entropy"] - o["entropy"]).max() <= 1e-4
    assert abs(g["expected_energy"] - o["expected_energy"]) <= 1e-4
    assert np.array_equal(g["spins"], o["spins"])


def test_executor_config1_prefix_matches_accurate_oracle():
    inst = I.config1()
    sched = I.linear(200)
    sched = I.Schedule(sched.a[:12], sched.b[:12])
    g = _run(inst.model, sched, inst.dt, inst.chi, -1)
    o = pyoracle.tebd(inst.model, sched, inst.dt, inst.chi, svd_mode=1, want_state=True)
    inf = 1 - _fid(o["state"], g["state"])
    print(f"executor config1[:12] infidelity={inf:.2e}", flush=True)
    assert inf <= 1e-10
    assert np.abs(g["entropy"] - o["entropy"]).max() <= 1e-9
    assert abs(g["expected_energy"] - o["expected_energy"]) <= 1e-9
    assert np.array_equal(g["spins"], o["spins"])


def test_executor_saturated_n128_step_matches_stream():
    """One step of the n=128, chi=64 range-4 chain from a saturated random
    state: executor and per-update path agree (bond dims, entropy, trunc)."""
    inst = I.chain_range(128, 4, 3)
    sched = I.linear(1000)
    a, b = sched.a[500:501], sched.b[500:501]
    out = []
    for groups in (0, -1):
        s = qa.TebdSession(inst, 64, 0.04)
        s.set_executor(groups)
        s.randomize(7, 40.0)
        s.run(a, b)
        out.append(s.sync())
        s.close()
    (e0, t0, b0), (e1, t1, b1) = out
    print(f"n128 chi64 ent {e0[0]:.12f} {e1[0]:.12f} trunc {t0:.3e} {t1:.3e}", flush=True)
    assert np.array_equal(b0, b1)
    assert abs(e0[0] - e1[0]) <= 1e-8
    assert abs(t0 - t1) <= 1e-8 * max(1.0, t0)
