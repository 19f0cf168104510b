"""The batched-annealing workload at its stated shape: all 512 generator
instances (n ~ U{16..96}, chi in {32, 64, 128}, banded sparse graphs of mean
degree 3 and range <= 3, J ~ N(0,1), h ~ N(0,0.25), each with its own
T ~ U[10, 40] at dt = 0.05) advanced together by the first three Trotter steps
of their own schedules in one batched executor launch (tessera_tebd_batch_run).

Checks: every session reads out finite values, bonds within its chi, spins in
{-1, +1}, <H> within +-(sum |J| + sum |h|); a sample of instances of every chi
and of the smallest and largest n equals the same sessions run one at a time on
the per-update path (<H> within 1e-9, kept ranks identical unless the oracle
meets rank decisions at the rounding level of theta); and every instance has a
per-instance completion time inside the launch (the batched p50 / p99).
"""
import numpy as np
import pytest

import pyoracle
from tessera_amd import instances as I
from tessera_amd import qa

pytestmark = pytest.mark.gpu

STEPS = 3


def test_all_512_instances_advance_in_one_batched_launch():
    insts = I.batch_instances(512)
    ns = np.array([x.model.n for x in insts])
    chis = np.array([x.chi for x in insts])
    assert ns.min() >= 16 and ns.max() <= 96 and set(chis) == {32, 64, 128}
    assert all(200 <= x.sched.steps <= 800 for x in insts)  # T in [10, 40] at dt = 0.05
    scheds = [I.Schedule(x.sched.a[:STEPS], x.sched.b[:STEPS]) for x in insts]
    sessions = [qa.TebdSession(x.model, x.chi, x.dt) for x in insts]
    try:
        qa.batch_run(sessions, scheds)
        out = []
        for x, s in zip(insts, sessions):
            _, tr, bonds = s.sync()
            spins, sc = s.readout()
            bound = float(np.abs(x.model.ew).sum() + np.abs(x.model.h).sum())
            assert np.all(np.isfinite(sc)) and np.isfinite(tr)
            assert bonds.max() <= x.chi
            assert set(np.unique(spins)) <= {-1, 1}
            assert abs(sc[1]) <= bound + 1e-9
            out.append((bonds.copy(), sc[1], tr, s.last_run_ms()))
    finally:
        for s in sessions:
            s.close()
    ms = np.array([o[3] for o in out])
    print(f"512 instances x {STEPS} steps, one launch: per-instance completion p50 {np.percentile(ms, 50):.1f} ms, "
          f"p99 {np.percentile(ms, 99):.1f} ms, last {ms.max():.1f} ms; max bond {max(o[0].max() for o in out)}",
          flush=True)
    assert np.all(ms > 0)
    # a sample against one-at-a-time runs: each chi, smallest and largest n
    picks = sorted({int(np.argmin(ns)), int(np.argmax(ns))} |
                   {int(np.flatnonzero(chis == c)[0]) for c in (32, 64, 128)})
    for k in picks:
        x = insts[k]
        s = qa.TebdSession(x.model, x.chi, x.dt)
        try:
            s.run(scheds[k].a, scheds[k].b, record=False)
            _, tr, bonds = s.sync()
            _, sc = s.readout()
        finally:
            s.close()
        amb = int(np.sum(pyoracle.tebd(x.model, scheds[k], x.dt, x.chi, svd_mode=1)["ambiguous"])) > 0
        print(f"instance {k} (n={x.model.n}, chi={x.chi}): <H> batched {out[k][1]:.12f} one at a time {sc[1]:.12f}; "
              f"bonds {'equal' if np.array_equal(bonds, out[k][0]) else 'differ'} (rounding-level ranks: {amb})",
              flush=True)
        assert abs(out[k][1] - sc[1]) <= 1e-9
        assert tr <= 1e-20 and out[k][2] <= 1e-20
        if not amb:
            assert np.array_equal(bonds, out[k][0])
