"""Batched anneal (tessera_tebd_batch_run) vs the same sessions run one at a time.

The batched scheduler places the two-site updates of every instance, each in
its own upstream gate order (mps.cpp:309-356), into shared super-steps of one
cooperative launch. At exact chi every instance must end in the state the
per-instance path reaches (to rounding), with the same readout.
"""
import numpy as np
import pytest

from tessera_amd import instances as I
from tessera_amd import qa

from test_gpu_tebd import _fid

pytestmark = pytest.mark.gpu


def _mixed():
    """Instances of different size, range, bond dimension and step count."""
    return [
        (I.chain_nnn(8, 3), I.linear(14), 0.1, 16),
        (I.chain_range(10, 3, 4), I.linear(9), 0.08, 32),
        (I.sparse_banded(12, 3, 3.0, 5), I.linear(11), 0.05, 64),
        (I.chain_nnn(6, 7), I.linear(20), 0.1, 8),
        (I.chain_range(9, 2, 8), I.linear(5), 0.1, 16),
    ]


def _rank_at_rounding_level(model, sched, dt, chi):
    """Whether the oracle meets kept-rank decisions set by singular values at the
    rounding level of theta (between 1e-15 and 1e-13, next to the 1e-14 cut of
    mps.cpp:82) on this run: there two accurate SVDs may keep different numbers
    of directions of weight ~1e-28, and only the states are comparable."""
    import pyoracle
    return int(np.sum(pyoracle.tebd(model, sched, dt, chi, svd_mode=1)["ambiguous"])) > 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_batch_matches_one_at_a_time(mode):
    """mode 0: automatic layout (grouped here: bonds up to 64), 1: grouped, 2: solo
    (thetas up to 64 x 64 in LDS, larger ones through the block SVD). Bond
    dimensions are compared exactly unless the instance has rank decisions at the
    rounding level (the states are compared in every case)."""
    cases = _mixed()
    amb = [_rank_at_rounding_level(*c) for c in cases]
    solo = []
    for model, sched, dt, chi in cases:
        s = qa.TebdSession(model, chi, dt)
        s.run(sched.a, sched.b, record=False)
        _, tr, bonds = s.sync()
        spins, sc = s.readout()
        solo.append((s.dense(), tr, bonds, spins, sc))
        s.close()
    sessions = [qa.TebdSession(model, chi, dt) for model, _, dt, chi in cases]
    qa.batch_run(sessions, [c[1] for c in cases], mode=mode)
    for k, s in enumerate(sessions):
        _, tr, bonds = s.sync()
        spins, sc = s.readout()
        psi = s.dense()
        inf = 1 - _fid(solo[k][0], psi)
        print(f"instance {k}: infidelity={inf:.2e} trunc={tr:.2e} bonds={bonds.max()} "
              f"(one at a time {solo[k][2].max()}; rounding-level ranks: {amb[k]})", flush=True)
        assert inf <= 1e-12
        if not amb[k]:
            assert np.array_equal(bonds, solo[k][2])
        assert tr <= 1e-20 and solo[k][1] <= 1e-20
        assert np.array_equal(spins, solo[k][3])
        assert abs(sc[1] - solo[k][4][1]) <= 1e-10
        s.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_continues_and_groups_agree(mode):
    """Two consecutive batch calls equal one, and the group count does not
    change the result (1 group: strictly sequential super-steps)."""
    cases = _mixed()[:3]
    out = []
    for groups, split in ((1, False), (-1, True)):
        sessions = [qa.TebdSession(model, chi, dt) for model, _, dt, chi in cases]
        if split:
            first = [I.Schedule(c[1].a[:4], c[1].b[:4]) for c in cases]
            rest = [I.Schedule(c[1].a[4:], c[1].b[4:]) for c in cases]
            qa.batch_run(sessions, first, groups, mode)
            qa.batch_run(sessions, rest, groups, mode)
        else:
            qa.batch_run(sessions, [c[1] for c in cases], groups, mode)
        out.append([s.dense() for s in sessions])
        for s in sessions:
            s.close()
    for a, b in zip(*out):
        assert 1 - _fid(a, b) <= 1e-12


def test_batch_config4_sample_runs():
    """Eight batched-workload instances (n up to 96, chi in {32, 64, 128}) for three
    steps: the batched run completes and every session reads out finite values
    with bonds within chi; the solo and grouped layouts agree on the expected
    energy and discard nothing. Their ranks may differ: three steps from the
    product state leave many singular values at the rounding level of theta
    (~1e-14 of its norm), and the two SVD variants (in-LDS pair Jacobi for
    small thetas, LQ-preconditioned block Jacobi) round them to different sides
    of the S < 1e-14 keep cut (mps.cpp:77-85); what either drops weighs below
    1e-20. (On the mixed exact cases above the ranks are identical wherever the
    oracle sees no rank decision at the rounding level.)"""
    insts = I.batch_instances(8)
    scheds = [I.Schedule(x.sched.a[:3], x.sched.b[:3]) for x in insts]
    res = []
    for mode in (2, 1):
        sessions = [qa.TebdSession(x.model, x.chi, x.dt) for x in insts]
        qa.batch_run(sessions, scheds, mode=mode)
        out = []
        for x, s in zip(insts, sessions):
            _, tr, bonds = s.sync()
            spins, sc = s.readout()
            assert np.all(np.isfinite(sc))
            assert bonds.max() <= x.chi
            assert set(np.unique(spins)) <= {-1, 1}
            out.append((bonds, sc[1], tr))
            s.close()
        res.append(out)
    for (b2, e2, t2), (b1, e1, t1) in zip(*res):
        print(f"rank difference solo - grouped: max {np.abs(b2.astype(int) - b1).max()}", flush=True)
        assert abs(e2 - e1) <= 1e-9
        assert t2 <= 1e-20 and t1 <= 1e-20


def test_layout_tuner_picks_a_measured_layout_and_runs_it(tmp_path):
    """The tuner measures every candidate configuration (solo layout; grouped with
    one or half a workgroup per block pair; with and without LQ preconditioning)
    on saturated probe chains, caches the fastest (also in a JSON file) and the
    tuned batch equals a direct batch in that configuration."""
    from tessera_amd import tuning

    cache = tmp_path / "layouts.json"
    rec = tuning.tune(16, 6, cache_path=cache)
    names = {c[0] for c in tuning._configs(16)}
    print(f"tuner record: {rec}", flush=True)
    assert rec["mode"] in (1, 2) and set(rec["updates_per_s"]) == names and rec["config"] in names
    assert {"solo", "grouped_w2_lq", "grouped_w2_nolq"} <= names
    assert all(v > 0 for v in rec["updates_per_s"].values())
    assert cache.exists() and tuning.tune(16, 5, cache_path=cache) is rec  # same class, cached
    cases = _mixed()[:3]
    out = []
    for tuned in (True, False):
        sessions = [qa.TebdSession(model, chi, dt) for model, _, dt, chi in cases]
        if tuned:
            tuning.tuned_batch_run(sessions, [c[1] for c in cases], [c[3] for c in cases])
        else:
            r = tuning.tune(64, 3)
            sessions[0].set_executor(0, r["workers"])
            sessions[0].set_svd_lq(r["lq"])
            qa.batch_run(sessions, [c[1] for c in cases], -1, r["mode"])
        out.append([s.dense() for s in sessions])
        for s in sessions:
            s.close()
    for a, b in zip(*out):
        assert 1 - _fid(a, b) <= 1e-12
