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


def _run(model, sched, dt, chi, groups, ns=1, chunks=1, record=True, workers=0):
    s = qa.TebdSession(model, chi, dt, ns)
    s.set_executor(groups, workers)
    edges = np.linspace(0, sched.steps, chunks + 1).astype(int)
    for lo, hi in zip(edges[:-1], edges[1:]):
        s.run(sched.a[lo:hi], sched.b[lo:hi], record=record)
    ent, tr, bonds = s.sync()
    spins, sc = s.readout()
    state = s.dense() if model.n <= 16 else None
    s.close()
    return dict(entropy=ent, max_trunc=tr, bonds=bonds, spins=spins, expected_energy=sc[1], state=state)


@pytest.mark.parametrize("groups,workers", [(1, 0), (8, 0), (-1, 0), (-1, 1), (4, 2)])
def test_executor_exact_chi_matches_dense(groups, workers):
    """workers > 0: fewer workgroups per group than block pairs per round, so
    each takes several pairs (workers = 1: one workgroup runs a whole update)."""
    model = I.chain_nnn(8, 48)
    sched = I.linear(20)
    g = _run(model, sched, 0.1, 16, groups, workers=workers)
    psi, ents = dense_mps_splitting(model, sched, 0.1)
    inf = 1 - _fid(psi, g["state"])
    err = np.abs(g["entropy"] - ents).max()
    print(f"groups={groups} workers={workers} infidelity={inf:.2e} ent_err={err:.2e}", flush=True)
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
    """chi = 8 on a range-3 chain of 12 spins (max discarded weight ~1e-2). The
    truncated dynamics amplifies rounding: two accurate CPU implementations (the
    oracle's Jacobi SVD and LAPACK, tests/tebd_numpy.py) drift apart from 1e-10
    at step 3 to ~5e-2 in entropy by step 15. Bars: kept ranks identical, and the
    executor's distance from the oracle no larger than 10x that spread (running
    maximum over the steps) for the entropy trace and the max discarded weight."""
    from tebd_numpy import tebd_numpy
    model = I.chain_range(12, 3, 17)
    sched = I.linear(20)
    g = _run(model, sched, 0.1, 8, -1)
    o = pyoracle.tebd(model, sched, 0.1, 8, svd_mode=1)
    tr_np, bonds_np, ent_np = tebd_numpy(model, sched, 0.1, 8, want_entropy=True)
    err = np.maximum.accumulate(np.abs(g["entropy"] - o["entropy"]))
    spread = np.maximum.accumulate(np.abs(ent_np - o["entropy"]))
    print(f"trunc gpu={g['max_trunc']:.6e} oracle={o['max_trunc']:.6e} lapack={tr_np:.6e}", flush=True)
    print("entropy err / spread:", " ".join(f"{e:.1e}/{s:.1e}" for e, s in zip(err, spread)), flush=True)
    assert int(g["bonds"].max()) == int(o["max_bond"][-1]) == bonds_np[-1]
    assert np.all(err <= 10 * spread + 1e-12)
    assert abs(g["max_trunc"] - o["max_trunc"]) <= 10 * abs(tr_np - o["max_trunc"]) + 1e-12


def test_executor_config1_prefix_matches_accurate_oracle():
    inst = I.chain12_anneal()
    sched = I.linear(200)
    sched = I.Schedule(sched.a[:12], sched.b[:12])
    g = _run(inst.model, sched, inst.dt, inst.chi, -1)
    o = pyoracle.tebd(inst.model, sched, inst.dt, inst.chi, svd_mode=1, want_state=True)
    inf = 1 - _fid(o["state"], g["state"])
    print(f"executor n=12 chain[:12] infidelity={inf:.2e}", flush=True)
    assert inf <= 1e-10
    assert np.abs(g["entropy"] - o["entropy"]).max() <= 1e-9
    assert abs(g["expected_energy"] - o["expected_energy"]) <= 1e-9
    assert np.array_equal(g["spins"], o["spins"])


def test_executor_n24_truncated_matches_oracle():
    """Two mid-anneal steps of a 24-spin range-4 chain from the product state at
    chi = 16 (668 updates, max discarded weight 5.6e-4, a well-conditioned cut):
    kept ranks identical to the accurate oracle, entropy within 1e-9, discarded
    weight no further from the oracle than 10x an independent LAPACK-SVD TEBD."""
    from tebd_numpy import tebd_numpy
    model = I.chain_range(24, 4, 3)
    full = I.linear(1000)
    sched = I.Schedule(full.a[500:502], full.b[500:502])
    g = _run(model, sched, 0.04, 16, -1)
    o = pyoracle.tebd(model, sched, 0.04, 16, svd_mode=1)
    tr_np, bonds_np = tebd_numpy(model, sched, 0.04, 16)
    print(f"n24 chi16 trunc gpu={g['max_trunc']:.9e} oracle={o['max_trunc']:.9e} lapack={tr_np:.9e}", flush=True)
    assert int(g["bonds"].max()) == int(o["max_bond"][-1]) == bonds_np[-1]
    assert np.abs(g["entropy"] - o["entropy"]).max() <= 1e-9
    assert abs(g["max_trunc"] - o["max_trunc"]) <= 10 * abs(tr_np - o["max_trunc"]) + 1e-12


def test_executor_n128_matches_stream():
    """Two mid-anneal steps of the n=128 range-4 chain from the product state at
    chi = 256, where nothing is truncated (bonds reach ~216; the weight dropped
    by the S < 1e-14 rule is ~1e-30): 3996 updates, up to 8 per super-step across
    the two steps. Executor and per-update path give the same entropy trace."""
    inst = I.chain_range(128, 4, 3)
    sched = I.linear(1000)
    a, b = sched.a[500:502], sched.b[500:502]
    out = []
    for groups, workers in ((0, 0), (-1, 0), (-1, 8)):
        s = qa.TebdSession(inst, 256, 0.04)
        s.set_executor(groups, workers)
        s.run(a, b)
        out.append(s.sync())
        s.close()
    (e0, t0, b0), (e1, t1, b1), (e2, t2, b2) = out
    print(f"n128 chi256 ent {e0} {e1} {e2} trunc {t0:.3e} {t1:.3e} {t2:.3e} max bond {b0.max()} {b1.max()}",
          flush=True)
    assert t0 <= 1e-20 and t1 <= 1e-20 and t2 <= 1e-20
    assert np.abs(e0 - e1).max() <= 1e-9
    assert np.abs(e0 - e2).max() <= 1e-9


@pytest.mark.parametrize("chi,workers", [(32, 0), (8, 2)])
def test_group_barriers_match_lockstep_grid_barrier(monkeypatch, chi, workers):
    """The grouped layout runs each group's update on the group's own barrier
    (own tournament length, LQ panel count and convergence), with point-to-point
    tournament rounds when every worker holds one block pair; the knobs switch
    the rounds back to full per-group barriers (TESSERA_EXEC_P2P=0) or run all
    groups' SVDs in lockstep on the grid barrier (TESSERA_EXEC_LOCKSTEP=1, the
    extra rounds and panels of a smaller matrix doing no work). Each matrix sees
    the same rotations, so all three give the same run: a range-3 chain of 12
    sites with up to 8 updates per super-step, exact at chi = 32 (4 workers, one
    pair each: point-to-point rounds) and truncated at chi = 8 (two workers per
    group, several block pairs each: full barriers)."""
    model = I.chain_range(12, 3, 23)
    sched = I.linear(5)
    grp = _run(model, sched, 0.1, chi, 8, workers=workers)
    monkeypatch.setenv("TESSERA_EXEC_P2P", "0")
    full = _run(model, sched, 0.1, chi, 8, workers=workers)
    monkeypatch.delenv("TESSERA_EXEC_P2P")
    monkeypatch.setenv("TESSERA_EXEC_LOCKSTEP", "1")
    lock = _run(model, sched, 0.1, chi, 8, workers=workers)
    for name, other in (("full barriers", full), ("lockstep", lock)):
        inf = 1 - _fid(grp["state"], other["state"])
        same = np.array_equal(grp["state"], other["state"])
        print(f"chi={chi} workers={workers} vs {name}: infidelity={inf:.2e} bit-identical={same} "
              f"trunc {grp['max_trunc']:.6e} / {other['max_trunc']:.6e}", flush=True)
        assert np.array_equal(grp["bonds"], other["bonds"])
        assert inf <= 1e-12
        assert np.abs(grp["entropy"] - other["entropy"]).max() <= 1e-10
        assert abs(grp["max_trunc"] - other["max_trunc"]) <= 1e-12 + 1e-9 * other["max_trunc"]


@pytest.mark.skip(reason="XCD-local group synchronisation is switched off: its first run at n=128, chi=256 "
                         "left the GPU faulted with no cause established")
@pytest.mark.parametrize("chi,workers,steps", [(32, 0, 5), (8, 2, 5), (64, 0, 3)])
def test_xcd_local_groups_match_device_scope_barriers(monkeypatch, chi, workers, steps):
    """With 8 (or a multiple of 8) groups each group sits on one XCD and the
    grouped layout synchronises it XCD-locally: barriers and round counters
    through the shared L2, the group's data read with L1-bypassing loads, no L2
    write-back or invalidation (TESSERA_EXEC_XCD_LOCAL=1; the default keeps the
    device-scope barriers). The arithmetic is the same, so the runs are bit-identical: a
    range-3 chain of 12 sites (exact at chi = 32 and 64 with point-to-point
    rounds; truncated at chi = 8 with two workers per group, full barriers)."""
    model = I.chain_range(12, 3, 23)
    sched = I.linear(steps)
    dev = _run(model, sched, 0.1, chi, 8, workers=workers)
    monkeypatch.setenv("TESSERA_EXEC_XCD_LOCAL", "1")
    xl = _run(model, sched, 0.1, chi, 8, workers=workers)
    same = np.array_equal(xl["state"], dev["state"])
    print(f"chi={chi} workers={workers}: bit-identical={same} infidelity={1 - _fid(xl['state'], dev['state']):.2e} "
          f"trunc {xl['max_trunc']:.6e} / {dev['max_trunc']:.6e}", flush=True)
    assert np.array_equal(xl["bonds"], dev["bonds"])
    assert same
    assert np.array_equal(xl["entropy"], dev["entropy"])
    assert xl["max_trunc"] == dev["max_trunc"]


@pytest.mark.parametrize("groups", [8, 0])
def test_svd_without_lq_preconditioning_matches(monkeypatch, groups):
    """The tuner's LQ switch: the multi-workgroup Jacobi SVD straight on theta
    (tessera_tebd_set_svd_lq(0)) factors the same matrices with another sweep
    count, so at exact chi (chi = 64 on this 12-site chain: discarded weight
    ~1e-30) the run equals the preconditioned one to rounding; in the executor
    (groups = 8) and on the per-update path forced onto the cooperative SVD
    (groups = 0). (At chi = 32 the run truncates 8e-3 of the weight and the two
    factorisations' rounding is amplified, as between any two accurate SVDs.)"""
    if groups == 0:
        monkeypatch.setenv("TESSERA_FORCE_DIST_SVD", "1")
    model = I.chain_range(12, 3, 23)
    sched = I.linear(5)
    out = []
    for lq in (True, False):
        s = qa.TebdSession(model, 64, 0.1)
        s.set_executor(groups)
        s.set_svd_lq(lq)
        s.run(sched.a, sched.b)
        ent, tr, bonds = s.sync()
        st = s.svd_stats()
        out.append((s.dense(), ent, tr, st["sweeps"] / max(st["calls"], 1)))
        s.close()
    (p1, e1, t1, w1), (p0, e0, t0, w0) = out
    inf = 1 - _fid(p1, p0)
    print(f"groups={groups}: LQ on vs off infidelity {inf:.2e}, entropy {np.abs(e1 - e0).max():.2e}, "
          f"sweeps per SVD {w1:.2f} / {w0:.2f}", flush=True)
    assert inf <= 1e-12
    assert np.abs(e1 - e0).max() <= 1e-10
    assert t1 <= 1e-20 and t0 <= 1e-20


def test_exec_profile_phases_are_consistent():
    """tessera_tebd_exec_profile: workgroup 0's phase timers partition each launch
    (snapshots + theta, SVD, wait after it, reconstruction, super-step-end wait),
    the LQ, Jacobi and round-barrier times lie inside the SVD time, and the totals
    accumulate over launches."""
    model = I.chain_range(14, 3, 29)
    sched = I.linear(4)
    s = qa.TebdSession(model, 32, 0.1)
    s.set_executor(-1)
    s.run(sched.a[:2], sched.b[:2])
    s.sync()
    p1 = s.exec_profile()
    s.run(sched.a[2:], sched.b[2:])
    s.sync()
    p2 = s.exec_profile()
    s.close()
    print({k: round(v, 3) for k, v in p2.items()}, flush=True)
    assert p1["kernel_ms"] > 0 and p1["rounds"] > 0 and p1["sweeps"] > 0 and p1["group_updates"] > 0
    parts = p1["theta_ms"] + p1["svd_ms"] + p1["post_svd_wait_ms"] + p1["recon_ms"] + p1["wait_ms"]
    assert parts <= p1["kernel_ms"] * (1 + 1e-9) and parts >= 0.95 * p1["kernel_ms"]
    assert p1["jacobi_ms"] + p1["lq_ms"] <= p1["svd_ms"] * (1 + 1e-9)
    assert p1["round_barrier_ms"] <= p1["jacobi_ms"]
    for k in p1:
        assert p2[k] >= p1[k]
    assert p2["kernel_ms"] > p1["kernel_ms"] and p2["rounds"] > p1["rounds"]


def test_group_barrier_wait_that_gives_up_is_reported(monkeypatch):
    """Every wait of the per-group barrier is bounded: a wait that gives up raises
    the executor's error word, and the session reports it instead of returning
    results. Forced from the host with a limit of zero polls (each wait not
    released at its first poll gives up at once, later waits stop on the error
    word, and the launch still ends at its grid barriers)."""
    monkeypatch.setenv("TESSERA_EXEC_SPIN_LIMIT", "0")
    model = I.chain_range(12, 3, 23)
    sched = I.linear(2)
    s = qa.TebdSession(model, 32, 0.1)
    s.set_executor(8)
    with pytest.raises(qa.TesseraError, match="timed out"):
        s.run(sched.a, sched.b, record=False)
        s.sync()
    s.close()
    # a fresh session with the default limit runs normally afterwards
    monkeypatch.delenv("TESSERA_EXEC_SPIN_LIMIT")
    g = _run(model, sched, 0.1, 32, 8)
    assert np.isfinite(g["max_trunc"]) and g["bonds"].max() <= 32


@pytest.mark.parametrize("chi,record", [(32, False), (8, False), (32, True), (8, True)])
def test_dataflow_order_matches_superstep_barriers(monkeypatch, chi, record):
    """In the dataflow order (the default; TESSERA_EXEC_DATAFLOW=0
    keeps the super-step barriers) the per-group layout has no super-step
    barriers: each update waits only for the done flags of the latest earlier
    updates on its two sites (every dependency, bond dimensions included, goes
    through a shared site), and each entropy snapshot is taken by the latest
    update on its site before it releases its done flag. The same rotations run
    on the same data, so the run equals the one with super-step barriers and the
    lockstep one, entropy trace included: a range-3 chain of 14 sites, 5 steps in
    one launch, exact at chi = 32 and truncated at chi = 8. Groups take their
    updates from a shared queue in the flat super-step order (TESSERA_EXEC_QUEUE=0
    gives group g slot g of every super-step instead); the two runs are
    bit-identical."""
    model = I.chain_range(14, 3, 31)
    sched = I.linear(5)
    monkeypatch.setenv("TESSERA_EXEC_DATAFLOW", "0")
    steps = _run(model, sched, 0.1, chi, 8, record=record)
    monkeypatch.delenv("TESSERA_EXEC_DATAFLOW")
    flow = _run(model, sched, 0.1, chi, 8, record=record)
    monkeypatch.setenv("TESSERA_EXEC_LOCKSTEP", "1")
    lock = _run(model, sched, 0.1, chi, 8, record=record)
    monkeypatch.delenv("TESSERA_EXEC_LOCKSTEP")
    monkeypatch.setenv("TESSERA_EXEC_QUEUE", "0")
    slots = _run(model, sched, 0.1, chi, 8, record=record)
    monkeypatch.delenv("TESSERA_EXEC_QUEUE")
    # which group runs an update changes nothing in its arithmetic
    for key in ("state", "bonds", "spins", "entropy"):
        assert np.array_equal(flow[key], slots[key]), key
    assert flow["max_trunc"] == slots["max_trunc"]
    for name, other in (("super-step barriers", steps), ("lockstep", lock)):
        inf = 1 - _fid(flow["state"], other["state"])
        same = np.array_equal(flow["state"], other["state"])
        print(f"chi={chi} vs {name}: infidelity={inf:.2e} bit-identical={same} "
              f"trunc {flow['max_trunc']:.6e} / {other['max_trunc']:.6e}", flush=True)
        assert np.array_equal(flow["bonds"], other["bonds"])
        assert inf <= 1e-12
        assert abs(flow["max_trunc"] - other["max_trunc"]) <= 1e-12 + 1e-9 * other["max_trunc"]
        assert np.array_equal(flow["spins"], other["spins"])
        assert flow["entropy"].shape == other["entropy"].shape == ((5,) if record else (0,))
        if record:
            assert np.abs(flow["entropy"] - other["entropy"]).max() <= 1e-10
