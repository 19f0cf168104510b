"""Measured choice of the batched executor's layout and Jacobi schedule.

`tessera_tebd_batch_run_ex` runs the updates of many MPS in super-steps of one
cooperative launch, in one of two layouts: grouped (each update's SVD spread
over a group of workgroups that synchronise on the group's own barrier; few concurrent updates,
large thetas) or solo (one workgroup per update, thetas up to 64 x 64 or 128 x 32 in LDS;
up to one update per resident workgroup). Which is faster depends on the bond
dimension and on how many independent updates the batch offers, so the tuner
measures it: for a shape class -- the largest chi of the batch and the instance
count rounded up to a power of two -- it builds that many saturated benchmark
states (random isometries with decaying bond weights, every bond at capacity)
on a short range-3 chain, advances them one Trotter step in every candidate
configuration and keeps the fastest. The candidates: the solo layout; the
grouped layout with one workgroup per block pair of a tournament round or with
half as many (each taking two pairs per round: fewer workgroups in every
barrier); each with and without the Householder LQ preconditioning of the
multi-workgroup Jacobi SVD (a different sweep count for the same
factorisation). Choices are cached for the process and, when a path is given,
in a JSON file.
"""
from __future__ import annotations

import json
import time
from pathlib import Path

from . import instances as I
from . import qa

_cache: dict[str, dict] = {}
MODES = {1: "grouped", 2: "solo"}


def _key(chi: int, count: int) -> str:
    bucket = 1
    while bucket < count:
        bucket *= 2
    return f"chi{chi}_count{bucket}"


def _configs(chi: int):
    """Candidate (name, mode, workers per group, LQ preconditioning)."""
    nb = (2 * chi + 7) // 8
    pairs = max(2, nb + (nb & 1)) // 2  # block pairs of a tournament round at full bonds
    out = [("solo", 2, 0, True)]
    for lq in (True, False):
        out.append((f"grouped_w{pairs}_{'lq' if lq else 'nolq'}", 1, 0, lq))
        if pairs >= 4:
            out.append((f"grouped_w{pairs // 2}_{'lq' if lq else 'nolq'}", 1, pairs // 2, lq))
    return out


def _measure(chi: int, count: int, n: int, mode: int, workers: int = 0, lq: bool = True) -> float:
    """Two-site updates per second of one saturated Trotter step of `count` chains."""
    sessions = []
    for k in range(count):
        s = qa.TebdSession(I.chain_range(n, 3, 9000 + k), chi, 0.05)
        s.randomize(9000 + k, 40.0)
        sessions.append(s)
    # batched runs take the first session's executor shape and SVD options
    sessions[0].set_executor(0, workers)
    sessions[0].set_svd_lq(lq)
    sched = I.linear(200)
    one = I.Schedule(sched.a[100:101], sched.b[100:101])
    updates = sum(s.updates_per_step for s in sessions)
    t0 = time.perf_counter()
    qa.batch_run(sessions, [one] * count, -1, mode)
    for s in sessions:
        s.sync()
    dt = time.perf_counter() - t0
    for s in sessions:
        s.close()
    return updates / dt


def tune(chi: int, count: int, cache_path: str | Path | None = None, max_updates: int = 4000) -> dict:
    """Best configuration for a batch of `count` instances with bond dimension up
    to `chi`: {"mode": 1 or 2, "layout": name, "workers": workgroups per group (0:
    one per block pair), "lq": LQ preconditioning, "updates_per_s": {candidate:
    rate}}. The probe chains have n = 2 log2(chi) + 2 sites (the middle bond
    reaches chi) and at most `max_updates` updates per probe (fewer chains for
    large chi)."""
    key = _key(chi, count)
    if cache_path is not None and not _cache:
        p = Path(cache_path)
        if p.exists():
            _cache.update(json.loads(p.read_text()))
    if key in _cache:
        return _cache[key]
    bucket = int(key.split("count")[1])
    n = 2 * max(1, int(chi).bit_length() - 1) + 2
    per_chain = sum((n - r) * (2 * r - 1) for r in (1, 2, 3) if r < n)
    probe = max(2, min(bucket, max_updates // max(per_chain, 1)))
    cfgs = _configs(chi)
    rates = {name: _measure(chi, probe, n, mode, workers, lq) for name, mode, workers, lq in cfgs}
    best = max(rates, key=rates.get)
    _, mode, workers, lq = next(c for c in cfgs if c[0] == best)
    _cache[key] = {"mode": mode, "layout": MODES[mode], "config": best, "workers": workers, "lq": lq,
                   "updates_per_s": rates}
    if cache_path is not None:
        Path(cache_path).write_text(json.dumps(_cache, indent=1, sort_keys=True))
    return _cache[key]


def tuned_batch_run(sessions, schedules, chis, groups: int = -1, cache_path: str | Path | None = None) -> dict:
    """batch_run with the configuration the tuner measured best for this batch's
    shape class (largest chi, instance count); returns the tuner's record. The
    first session's executor shape and SVD option are reset to the defaults
    afterwards."""
    rec = tune(max(chis), len(sessions), cache_path=cache_path)
    first = sessions[0]
    first.set_executor(0, rec.get("workers", 0))
    first.set_svd_lq(rec.get("lq", True))
    try:
        qa.batch_run(sessions, schedules, groups, rec["mode"])
    finally:  # the first session's own settings back to the defaults
        first.set_executor(0, 0)
        first.set_svd_lq(True)
    return rec
