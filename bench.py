#!/usr/bin/env python3
"""TEBD two-site updates/s on n=128 spins at bond dimension chi=256.

Workload (one "step" = one symmetric Trotter step of the MPS-TEBD anneal):
  chain of 128 spins with couplings (i, i+r), r = 1..4 (502 edges, J ~ U[-1,1]/r,
  h ~ U[-0.5,0.5], seed 3), routed by swap networks: 1998 two-site updates per
  step, plus two single-site layers and the qubit-0 entropy. dt = 0.04 on the
  linear schedule of 1000 steps; the timed steps are schedule steps 500.. (mid
  anneal). The MPS starts saturated: random site tensors with every bond at its
  capacity min(256, 2^b, 2^(128-b)), brought by one sweep of identity updates to
  left-isometries with bond weights ~exp(-k/40), so every interior update
  factorises a 512 x 512 complex128 theta and keeps 256 of its singular values.

By default the timed steps (12) run in the lockstep executor: one cooperative
launch in which consecutive Trotter steps overlap (1874 of a step's 1998 updates
form one dependency chain, so a launch of K steps has ~K updates per super-step
of the host's schedule while the pipeline fills; 6 steps give 5.46 on average, 12
give 6.0, and the library runs a 1000-step anneal in launches of 64 steps, 7.5).
Each group takes the next update from a shared queue, ordered by dependency depth,
and waits only for the done flags of its predecessors on its two sites (the
dataflow order; TESSERA_EXEC_QUEUE=0 gives group g slot g of every super-step,
TESSERA_EXEC_DATAFLOW=0 adds a grid barrier after every super-step). --steps 6
times the shorter launch, whose rate the pipeline's fill and drain set.
--executor 0 runs one launch sequence per update.

Output: one JSON line. `value` is two-site updates per second over the timed
steps (inputs resident on the device). `roofline` covers the dominant kernel
(the executor launch, or the SVD launches on the per-update path): algorithmic
FP64 flops of its one-sided Jacobi = 16 per visited column-pair row + 18 per
rotated one (counted on the device), divided by the kernel's HIP-event time.
`cpu_baseline` times the bit-exact CPU port of the upstream engine on one host
core for a bounded sample (one Jacobi sweep of a saturated update), scaled by
the 25 sweeps the upstream eigensolver needs on it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

N_SPINS, RANGE, CHI, DT, SCHED_STEPS, SEED = 128, 4, 256, 0.04, 1000, 3
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# One saturated chi=256 update of the upstream engine runs its 1024 x 1024 real
# Jacobi eigensolver for this many sweeps (measured once with the bit-exact CPU
# port on this matrix; the per-sweep time is re-measured on the benchmark host).
REF_SWEEPS_CHI256 = 25


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed Trotter steps (tebd: one launch, default 12; consecutive steps overlap in its "
                         "super-steps; the other workloads: default 6)")
    ap.add_argument("--warmup", type=int, default=0,
                    help="untimed steps (the saturated-state preparation already runs every kernel)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--n", type=int, default=N_SPINS)
    ap.add_argument("--chi", type=int, default=CHI)
    ap.add_argument("--executor", type=int, default=-1,
                    help="lockstep executor groups (-1: as many as fit, 0: one launch sequence per update)")
    ap.add_argument("--workload", choices=("tebd", "sv", "batch", "learned", "gnn", "chain12"), default="tebd",
                    help="tebd: the headline metric; sv: the n=30 state-vector sweep (HBM roofline); "
                         "batch: batched anneals of the batched-workload instances; learned: guidance network "
                         "(GPU TEBD targets, GPU training, guided vs plain anneals); gnn: training "
                         "throughput of the guidance network alone; chain12: the n=12 chain case, "
                         "MPS-TEBD at chi=64 vs the exact state vector")
    ap.add_argument("--train", type=int, default=512, help="learned workload: training instances")
    ap.add_argument("--eval", type=int, default=32, help="learned workload: held-out instances")
    ap.add_argument("--epochs", type=int, default=60, help="learned workload: training epochs")
    ap.add_argument("--fractions", default="0.25,0.5", help="learned workload: shortened anneal lengths")
    ap.add_argument("--strengths", default="0.5,1.0", help="learned workload: warm-start bias strengths")
    ap.add_argument("--no-lq", action="store_true",
                    help="tebd: Jacobi directly on theta, without the LQ preconditioning of the block SVD")
    ap.add_argument("--step-profile", action="store_true",
                    help="tebd: add the executor's cycle profile of the Jacobi rotation steps (slows workgroup 0; "
                         "zeros unless the library was built with TESSERA_STEP_PROFILE=1)")
    ap.add_argument("--exec-workers", type=int, default=0,
                    help="lockstep executor: workgroups per group (0: one per block pair of a round)")
    ap.add_argument("--instances", type=int, default=32, help="batch workload: number of batched-workload instances")
    ap.add_argument("--exec-mode", type=int, default=0, choices=(-1, 0, 1, 2),
                    help="batch workloads: executor layout 0 auto, 1 grouped, 2 solo (one workgroup per update), "
                         "-1 measured by the layout tuner (tessera_amd.tuning)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (PMC byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 12 if args.workload == "tebd" else 6
    return args


def progress(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_info() -> dict:
    """CPU model and logical core count of the host the baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu": model, "logical_cores": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def cpu_baseline(chi: int):
    """Bounded sample of the upstream update on one host core (bit-exact port)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the CPU oracle: only the baseline leg uses it

    t_setup, sweeps, t_sweep = pyoracle.update_probe(chi, 1)
    info = {"cores": 1, "kind": "port", "host": host_info()}
    if REF_SWEEPS_CHI256 and chi == 256:
        per_update = t_setup + REF_SWEEPS_CHI256 * t_sweep
        info.update(value=1.0 / per_update, unit="two-site updates/s",
                    sample=(f"theta + Gram + 1 of {REF_SWEEPS_CHI256} Jacobi sweeps of one saturated chi=256 update "
                            f"(1024x1024 real embedding) on 1 core: setup {t_setup:.2f}s, sweep {t_sweep:.2f}s"))
    else:
        info.update(value=None, unit="two-site updates/s",
                    sample=f"theta + Gram {t_setup:.2f}s, 1 Jacobi sweep {t_sweep:.2f}s (sweep count unknown)")
    return info


def copy_calibration(mib: int):
    """A plain device copy of known size: the PMC byte counters of this run,
    divided by 2 * mib MiB, give the counter scale used for `traffic`."""
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n, "launches": 3}), flush=True)


SV_N, SV_SEED, SV_DT, SV_STEPS = 30, 2, 0.02, 1000


def sv_bench(args):
    """n=30 random 3-regular TFIM (J ~ N(0,1), h ~ N(0,0.25)), linear schedule
    T=20, dt=0.02: per-step time of the phase + X-rotation sweep from the
    difference of two anneal lengths (allocation, initial state and readout
    cancel)."""
    from tessera_amd import instances as I
    from tessera_amd import qa

    inst = I.regular3_n30_anneal(SV_SEED, SV_STEPS)
    n = inst.model.n
    w, k = max(args.warmup, 1), args.steps
    full = inst.sched

    def timed(steps):
        sched = I.Schedule(full.a[:steps], full.b[:steps])
        t0 = time.perf_counter()
        qa.anneal_statevector(inst.model, sched, SV_DT)
        return time.perf_counter() - t0

    timed(1)  # code objects, first allocation
    t_w = timed(w)
    progress(f"sv warm run of {w} steps: {t_w:.3f}s")
    t_wk = timed(w + k)
    progress(f"sv run of {w + k} steps: {t_wk:.3f}s")
    per_step = (t_wk - t_w) / k
    algo = 32.0 * (1 << n)  # read + write every complex128 amplitude once per step
    achieved = algo / per_step / 1e9
    out = {
        "metric": "state-vector Trotter steps/s @n=30 (HBM GB/s of the phase + X sweep)",
        "value": 1.0 / per_step, "unit": "Trotter steps/s", "n_gpus": 1, "steps": k, "warmup": w,
        "ms_per_step": per_step * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)", "data": "synthetic: seeded random 3-regular graph",
        "config": {"workload": f"exact state vector n={n}, 2^{n} amplitudes, dt={SV_DT}"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sv_pass (2 passes per step at n=30, algorithmic bytes = 32 * 2^n per step)"},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = sv_cpu_baseline()
    print(json.dumps(out), flush=True)


def sv_cpu_baseline(n: int = 22):
    """The oracle's restatement of quantum.rs (one core) on a random 3-regular
    graph at n = 22: per-step time from the difference of a 6- and a 2-step run,
    extrapolated to n = 30 by 2^(30-n) amplitudes and (4 + 30) / (4 + n) passes
    per step (the upstream step makes 4 + n sweeps over the state)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle  # the CPU oracle: only the baseline leg uses it

    from tessera_amd import instances as I

    model = I.random_regular3(n, SV_SEED)
    full = I.linear(SV_STEPS)

    def timed(steps):
        t0 = time.perf_counter()
        pyoracle.sv(model, I.Schedule(full.a[:steps], full.b[:steps]), SV_DT)
        return time.perf_counter() - t0

    t2 = timed(2)
    t6 = timed(6)
    per_step = max(t6 - t2, 1e-9) / 4
    ns_amp = per_step / (1 << n) * 1e9
    per_step_30 = per_step * (1 << (30 - n)) * (4 + 30) / (4 + n)
    progress(f"cpu sv n={n}: {per_step:.3f}s per step, {ns_amp:.2f} ns per amplitude-step")
    return {"value": 1.0 / per_step_30, "unit": "Trotter steps/s (extrapolated to n=30)", "cores": 1, "kind": "port",
            "sample": f"oracle restatement of quantum.rs, random 3-regular n={n}, 4 steps: {per_step:.3f} s per step "
                      f"= {ns_amp:.2f} ns per amplitude-step; x 2^{30 - n} amplitudes x (34/{4 + n}) passes "
                      f"-> {per_step_30:.1f} s per n=30 step"}


def config1_bench(args):
    """The CPU-runnable case: n = 12 chain with nearest and next-nearest couplings
    (seed 1), linear schedule of 200 steps at dt = 0.05, as an exact state vector
    and as MPS-TEBD at chi = 64 (exact for n = 12). Reports the fidelity of the
    two final states (different Trotter splittings, so O(dt^2) apart), both
    energies, the entropy traces and the discarded weight; `value` is MPS-TEBD
    anneals per second on the GPU."""
    import numpy as np

    from tessera_amd import instances as I
    from tessera_amd import qa

    inst = I.chain12_anneal()
    psi_sv, ex_sv, ent_sv = qa.statevector_final(inst.model, inst.sched, inst.dt)
    g_sv = qa.anneal_statevector(inst.model, inst.sched, inst.dt)
    walls = []
    for rep_ in range(max(args.steps, 1)):
        s = qa.TebdSession(inst.model, inst.chi, inst.dt)
        t0 = time.perf_counter()
        s.run(inst.sched.a, inst.sched.b)
        ent, trunc, bonds = s.sync()
        walls.append(time.perf_counter() - t0)
        spins, sc = s.readout()
        psi_mps = s.dense()
        s.close()
    wall = float(np.median(walls))
    # MPS physical index 0 <-> spin +1, state-vector bit set <-> spin +1
    psi_m = psi_mps[::-1]
    fid = abs(np.vdot(psi_m, psi_sv)) ** 2 / (np.vdot(psi_m, psi_m).real * np.vdot(psi_sv, psi_sv).real)
    progress(f"n=12 chain: fidelity {fid:.9f}, TEBD anneal {wall:.3f}s")
    out = {
        "metric": "n=12 chain MPS-TEBD anneals/s (n=12 chain, chi=64, 200 steps) with fidelity to the exact state vector",
        "value": 1.0 / wall, "unit": "anneals/s", "n_gpus": 1, "steps": max(args.steps, 1), "warmup": 0,
        "ms_per_step": wall / inst.sched.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128 (fp64)", "data": "synthetic: seeded n=12 chain (nn + nnn couplings)",
        "config": {"workload": "n=12 chain, linear(200), dt=0.05, chi=64 (exact) vs 2^12 state vector",
                   "fidelity_mps_sv": fid, "mps_energy": float(sc[0]), "mps_expected_energy": float(sc[1]),
                   "sv_energy": g_sv.energy, "sv_expected_energy": ex_sv, "sv_success_prob": g_sv.success_prob,
                   "spins_agree": bool(np.array_equal(spins, g_sv.spins)), "max_trunc": trunc,
                   "max_bond": int(bonds.max()), "entropy_max_abs_diff": float(np.abs(ent - ent_sv).max()),
                   "entropy_mps_every_20": [round(float(x), 6) for x in ent[::20]],
                   "entropy_sv_every_20": [round(float(x), 6) for x in ent_sv[::20]],
                   "two_site_updates": 41 * inst.sched.steps},
        "roofline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle  # the CPU oracle: only the baseline leg uses it

        # bounded sample: the first 30 of the 200 steps on both sides; the full
        # anneal (~200 s on one core) is estimated with the GPU's own ratio
        pre = I.Schedule(inst.sched.a[:30], inst.sched.b[:30])
        t0 = time.perf_counter()
        pyoracle.tebd(inst.model, pre, inst.dt, inst.chi, svd_mode=0)
        c30 = time.perf_counter() - t0
        s = qa.TebdSession(inst.model, inst.chi, inst.dt)
        t0 = time.perf_counter()
        s.run(pre.a, pre.b)
        s.sync()
        g30 = time.perf_counter() - t0
        s.close()
        out["cpu_baseline"] = {"value": (1.0 / wall) * g30 / c30, "unit": "anneals/s (estimated)", "cores": 1,
                               "kind": "port",
                               "sample": f"bit-exact port of the upstream engine (Gram SVD) on the first 30 of 200 "
                                         f"steps: {c30:.2f} s vs {g30:.3f} s on the GPU; full-anneal rate scaled by "
                                         f"that ratio"}
    print(json.dumps(out), flush=True)


def batch_bench(args):
    """Batched-workload instances (n ~ U{16..96}, chi in {32, 64, 128}, banded sparse
    graphs), each from a saturated state (bonds at capacity, bond weights
    exp(-k/40)), advanced by `steps` mid-anneal Trotter steps: all together in
    one batched launch vs one at a time on the per-update path."""
    import numpy as np

    from tessera_amd import instances as I
    from tessera_amd import qa

    insts = I.batch_instances(args.instances)

    def sessions():
        out = []
        for k, x in enumerate(insts):
            s = qa.TebdSession(x.model, x.chi, x.dt)
            s.randomize(4000 + k, 40.0)
            out.append(s)
        return out

    def sched(x):
        k0 = x.sched.steps // 2
        return I.Schedule(x.sched.a[k0:k0 + args.steps], x.sched.b[k0:k0 + args.steps])

    solo_times = []
    ss = sessions()
    for s, x in zip(ss, insts):
        sc = sched(x)
        t0 = time.perf_counter()
        s.run(sc.a, sc.b, record=False)
        s.sync()
        solo_times.append(time.perf_counter() - t0)
    total_updates = sum(s.updates_per_step * args.steps for s in ss)
    for s in ss:
        s.close()
    t_solo = sum(solo_times)
    progress(f"one at a time: {len(insts)} instances x {args.steps} steps in {t_solo:.2f}s")
    mode = args.exec_mode
    rec = None
    if mode == -1:
        from tessera_amd import tuning

        rec = tuning.tune(max(x.chi for x in insts), len(insts))
        mode = rec["mode"]
        progress(f"tuner: {rec}")
    ss = sessions()
    if rec is not None:  # batched runs take the first session's executor shape and SVD option
        ss[0].set_executor(0, rec["workers"])
        ss[0].set_svd_lq(rec["lq"])
    t0 = time.perf_counter()
    qa.batch_run(ss, [sched(x) for x in insts], args.executor or -1, mode)
    for s in ss:
        s.sync()
    t_batch = time.perf_counter() - t0
    ex = ss[0].exec_stats()
    # per-instance completion inside the batched launch (device clock: launch start to
    # the end of the super-step holding the instance's last update)
    batched_ms = np.array([s.last_run_ms() for s in ss])
    for s in ss:
        s.close()
    progress(f"batched: {t_batch:.2f}s")
    st = np.array(solo_times)
    out = {
        "metric": "batched TEBD instances/s (batched-workload sample, saturated, mid-anneal steps)",
        "value": len(insts) / t_batch, "unit": "instances/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": t_batch / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128 (fp64)", "data": "synthetic: batched-workload generator (seeds 4000+i), saturated states",
        "config": {"workload": f"{len(insts)} instances x {args.steps} Trotter steps, chi in {{32,64,128}}",
                   "two_site_updates": total_updates,
                   "batched_updates_per_s": total_updates / t_batch,
                   "one_at_a_time_updates_per_s": total_updates / t_solo,
                   "one_at_a_time_instances_per_s": len(insts) / t_solo,
                   "speedup": t_solo / t_batch,
                   "one_at_a_time_instance_s_p50": float(np.percentile(st, 50)),
                   "one_at_a_time_instance_s_p99": float(np.percentile(st, 99)),
                   "batched_instance_s_p50": float(np.percentile(batched_ms, 50)) * 1e-3,
                   "batched_instance_s_p99": float(np.percentile(batched_ms, 99)) * 1e-3,
                   "batched_launch_s": t_batch,
                   "exec_mode": mode, "tuner": rec, "supersteps": ex["supersteps"],
                   "updates_per_superstep": ex["updates"] / max(ex["supersteps"], 1)},
    }
    print(json.dumps(out), flush=True)


def learned_bench(args):
    """Learned guidance, a reduced sample of the learned-guidance workload: training instances generated
    like the batched workload (n ~ U{16..40}, chi = 32, T = 10 at dt = 0.05),
    targets from batched GPU TEBD anneals (final <Z_i>, entanglement-peak
    fraction), the network trained on the GPU, then held-out instances (n ~
    U{32..64}) annealed plain at full length and at shorter lengths, plain and
    guided (warm-start bias b = strength * z, stretch from the predicted peak).
    A run succeeds when its readout energy is within 1% of the best plain
    energy of its instance; guided wall times include the network inference."""
    import numpy as np

    from tessera_amd import learned as Lrn

    steps, dt, chi = 200, 0.05, 32
    fractions = [float(x) for x in args.fractions.split(",")]
    strengths = [float(x) for x in args.strengths.split(",")]
    train = Lrn.guidance_instances(args.train, 5_000_000, 16, 40, chi, steps, dt)
    evals = Lrn.guidance_instances(args.eval, 6_000_000, 32, 64, chi, steps, dt)
    progress(f"targets for {len(train)} training instances")
    tg = Lrn.tebd_targets(train, chunks=10, log=progress)
    progress(f"targets: {tg.seconds:.1f}s, {tg.updates / tg.seconds:.0f} two-site updates/s")
    batch = Lrn.pack([x.model for x in train], tg.z, tg.s_peak)
    n_val = max(1, len(train) // 8)  # last eighth held out of training: validation loss
    net = Lrn.GuidanceNet(seed=5)
    net.set_data(batch)
    net.fit(1, batch_graphs=64, lr=2e-3, seed=0)  # warm-up epoch (code objects, buffers)
    net.params = Lrn.GuidanceNet(seed=5).params
    l0, _ = net.grad(len(train) - n_val, len(train))
    fit_graphs = len(train) - n_val
    t0 = time.perf_counter()
    rng = np.random.default_rng(1)
    starts = np.arange(0, fit_graphs, 64)
    hist = []
    for _ in range(args.epochs):
        tot = 0.0
        for st in rng.permutation(starts):
            tot += net.train_step(int(st), int(min(st + 64, fit_graphs)), 2e-3)
        hist.append(tot / len(starts))
    t_train = time.perf_counter() - t0
    l1, _ = net.grad(len(train) - n_val, len(train))
    progress(f"trained {args.epochs} epochs in {t_train:.2f}s: train loss {hist[0]:.4f} -> {hist[-1]:.4f}, "
             f"validation {l0:.4f} -> {l1:.4f}")
    base = np.array([x.sched.steps for x in evals])
    e_plain, _, w_plain = Lrn.anneal_batch(evals, base, log=progress)
    progress(f"plain full-length anneals: {w_plain:.1f}s")
    runs, plain_e = [], [e_plain]
    models = [x.model for x in evals]
    for frac in fractions:
        short = np.maximum(2, np.round(frac * base).astype(int))
        e_p, _, w_p = Lrn.anneal_batch(evals, short, log=progress)
        plain_e.append(e_p)
        runs.append(dict(kind="plain", fraction=frac, strength=0.0, steps=int(short.sum()), wall_s=w_p, e=e_p))
        for strength in strengths:
            g = Lrn.guidance(net, models, base, frac, strength, float(tg.s_peak.mean()))
            e_g, _, w_g = Lrn.anneal_batch(evals, g.steps, g.bias, log=progress)
            runs.append(dict(kind="guided", fraction=frac, strength=strength, steps=int(g.steps.sum()),
                             wall_s=w_g + g.infer_seconds, infer_s=g.infer_seconds, e=e_g))
            progress(f"fraction {frac} strength {strength}: guided {w_g:.1f}s")
    best = np.min(np.stack(plain_e), axis=0)
    ok_full = float(Lrn.within(e_plain, best).mean())
    table, pick = [], None
    for r in runs:
        ok = Lrn.within(r.pop("e"), best)
        r.update(within_1pct=float(ok.mean()), wall_ratio=r["wall_s"] / w_plain)
        table.append(r)
        if r["kind"] == "guided" and r["within_1pct"] >= ok_full and (pick is None or r["wall_s"] < pick["wall_s"]):
            pick = r
    # CPU baseline: the FP64 oracle's forward + backward of one 64-graph batch on one core
    sys.path.insert(0, str(ROOT / "oracle"))
    import gnn_ref  # the CPU oracle: only the baseline leg uses it

    sub = Lrn.sub_batch(batch, 0, 64)
    w = net.params.astype(np.float64)
    tc = time.perf_counter()
    gnn_ref.loss_and_grad(w, sub)
    t_cpu = time.perf_counter() - tc
    out = {
        "metric": "time-to-1% energy, learned vs plain (guided wall / plain wall; reduced learned-guidance sample)",
        "value": pick["wall_ratio"] if pick else None, "unit": "ratio", "n_gpus": 1, "steps": steps,
        "warmup": 0, "ms_per_step": None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 MFMA / fp32 accumulate (network); c128 (TEBD)",
        "data": "synthetic: banded sparse Ising graphs (seeds 5000000+i train, 6000000+i eval)",
        "config": {"workload": f"{len(train)} training instances n 16..40 ({n_val} for validation), "
                               f"{len(evals)} held-out n 32..64, chi={chi}, T={steps * dt:g} at dt={dt}",
                   "target_seconds": tg.seconds, "target_updates_per_s": tg.updates / tg.seconds,
                   "train_seconds": t_train,
                   "train_graphs_per_s": args.epochs * fit_graphs / t_train,
                   "train_loss_first_epoch": hist[0], "train_loss_last_epoch": hist[-1],
                   "validation_loss_initial": l0, "validation_loss_final": l1,
                   "plain_full_wall_s": w_plain, "plain_full_within_1pct": ok_full,
                   "picked": None if pick is None else {k: pick[k] for k in ("fraction", "strength")},
                   "runs": table},
        "roofline": None,
        "cpu_baseline": {"value": 64.0 / t_cpu, "unit": "training graphs/s (forward + backward)", "cores": 1,
                         "kind": "port", "sample": "FP64 numpy oracle, one 64-graph batch"},
    }
    print(json.dumps(out), flush=True)


def gnn_bench(args):
    """Training throughput of the guidance network alone: 2048 banded sparse
    graphs (n ~ U{16..40}, the training distribution of the learned workload)
    with synthetic targets, Adam over mini-batches of 64 graphs for `steps`
    epochs after `warmup` epochs. The dense products are tiny (N x 64 x 128),
    so the network is launch- and latency-bound; the MFMA figure is reported
    for completeness, not as a roofline claim."""
    import numpy as np

    from tessera_amd import instances as I
    from tessera_amd import learned as Lrn

    rng = I.Rng(77)
    models = [I.sparse_banded(rng.randint(16, 40), 3, 3.0, 0, rng=rng) for _ in range(2048)]
    tz = [np.tanh(2.0 * m.h) for m in models]
    ts = [float(np.abs(m.h).mean()) for m in models]
    batch = Lrn.pack(models, tz, ts)
    net = Lrn.GuidanceNet(seed=3)
    net.set_data(batch)
    nodes = int(batch["node_off"][-1])
    w = max(args.warmup, 1)
    net.fit(w, batch_graphs=64, lr=1e-3, seed=0)
    progress(f"gnn warm-up {w} epoch(s) done")
    t0 = time.perf_counter()
    hist = net.fit(args.steps, batch_graphs=64, lr=1e-3, seed=1)
    wall = time.perf_counter() - t0
    progress(f"gnn {args.steps} epochs: {wall:.2f}s, loss {hist[0]:.4f} -> {hist[-1]:.4f}")
    # MFMA work per epoch: forward [P Q] = H W_sd and Z = [H M] W_u per layer (2 * N * 64 * 128
    # flops each), backward two products per forward product
    flops = 3 * 2 * (2.0 * nodes * 64 * 128) * 3 * args.steps
    sys.path.insert(0, str(ROOT / "oracle"))
    import gnn_ref  # the CPU oracle: only the baseline leg uses it

    sub = Lrn.sub_batch(batch, 0, 64)
    tc = time.perf_counter()
    gnn_ref.loss_and_grad(net.params.astype(np.float64), sub)
    t_cpu = time.perf_counter() - tc
    out = {
        "metric": "guidance-network training graphs/s (forward + backward + Adam)",
        "value": args.steps * len(models) / wall, "unit": "graphs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": w, "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16 MFMA / fp32 accumulate, fp32 master weights",
        "data": "synthetic: 2048 banded sparse graphs, synthetic targets",
        "config": {"workload": f"3-layer message passing width 64, {len(models)} graphs, {nodes} nodes, "
                               "batches of 64 graphs", "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1]},
        "roofline": {"bound": "mfma", "achieved": flops / wall / 1e12, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": flops / wall / 1e12 / 2500.0, "traffic": None,
                     "kernel": "k_gemm_bf16 (latency-bound: N x 64 x 128 products)"},
        "cpu_baseline": {"value": 64.0 / t_cpu, "unit": "graphs/s (forward + backward)", "cores": 1, "kind": "port",
                         "sample": "FP64 numpy oracle, one 64-graph batch"},
    }
    print(json.dumps(out), flush=True)


def gemm_flops_per_step(model, chi: int):
    """Algorithmic flops of one Trotter step on a saturated MPS: per two-site
    update on sites (p, p+1), F = 32 Dl Dr (mid + keep) real flops -- the theta
    contraction (4 Dl mid Dr complex MACs, mps.cpp:49-70) plus the
    reconstruction (4 Dl Dr keep complex MACs, linalg.cpp:133-139, mps.cpp:87-99)
    at 8 flops per complex MAC -- with every bond at its capacity
    cap(b) = min(chi, 2^b, 2^(n-b)) and keep = cap(p+1). The updates are those of
    the swap-network routing of mps.cpp:126-145 (2r - 1 per edge of range r).
    Returns (flops per step, updates per step)."""
    n = model.n
    cap = [min(chi, 2 ** min(b, 60), 2 ** min(n - b, 60)) for b in range(n + 1)]
    flops, ups = 0.0, 0
    for i, j in zip(model.ei, model.ej):
        i, j = sorted((int(i), int(j)))
        if i == j:
            continue
        sites = [k - 1 for k in range(j, i + 1, -1)] + [i] + list(range(i + 1, j))
        for p in sites:
            dl, mid, dr = cap[p], cap[p + 1], cap[p + 2]
            flops += 32.0 * dl * dr * (mid + mid)
            ups += 1
    return flops, ups


def pmc_traffic(kernel: str, updates: int):
    """HBM bytes per launch of `kernel` from the committed PMC summary
    (profiles/pmc_bytes.json: FETCH_SIZE and WRITE_SIZE per two-site update,
    each scaled by the counter's reading for a plain copy of known size), or
    None when no summary covers the kernel."""
    f = ROOT / "profiles" / "pmc_bytes.json"
    if not f.exists():
        return None, None
    rec = json.loads(f.read_text()).get(kernel)
    if not rec:
        return None, None
    per = rec["fetch_bytes_per_update"] + rec["write_bytes_per_update"]
    return per * updates, rec


def sweep_summary(h0, h1, rounds):
    """Per Jacobi sweep index: SVDs of workgroup 0 that ran it and its rotations per
    SVD there (the executor's convergence profile over the timed steps), and its
    thread 0's shader cycles per tournament round in the parts of the row-layout
    rotation steps."""
    import numpy as np

    svds = h1["svds"] - h0["svds"]
    rot = h1["rotations"] - h0["rotations"]
    last = int(np.flatnonzero(svds)[-1]) + 1 if svds.any() else 0
    cyc = {k: round((h1["step_cycles"][k] - h0["step_cycles"][k]) / max(rounds, 1), 1) for k in h1["step_cycles"]}
    out = {"svds": [int(x) for x in svds[:last]],
           "rotations_per_svd": [round(float(r / v), 1) if v else 0.0 for r, v in zip(rot[:last], svds[:last])],
           "rounds_ms": round(h1["rounds_ms"] - h0["rounds_ms"], 3),
           "sweep_end_ms": round(h1["sweep_end_ms"] - h0["sweep_end_ms"], 3),
           "round_pre_ms": round(h1["round_pre_ms"] - h0["round_pre_ms"], 3),
           "round_post_ms": round(h1["round_post_ms"] - h0["round_post_ms"], 3)}
    if any(cyc.values()):
        out["step_cycles_per_round"] = cyc
    return out


def tebd_timed(args, model, sched, k0, sess):
    """The headline measurement on an open session (closed by the caller)."""
    sess.set_executor(args.executor, args.exec_workers)
    if args.no_lq:
        sess.set_svd_lq(False)
    sess.randomize(SEED, 40.0)
    ups = sess.updates_per_step
    a = sched.a[k0:k0 + args.warmup + args.steps]
    b = sched.b[k0:k0 + args.warmup + args.steps]
    progress(f"n={args.n} chi={args.chi} updates/step={ups} executor={args.executor}")
    if args.warmup:
        tw = time.perf_counter()
        sess.run(a[:args.warmup], b[:args.warmup], record=False)
        sess.sync()
        progress(f"warmup {args.warmup} step(s): {time.perf_counter() - tw:.2f}s")
    sess.set_timing(True)
    st0 = sess.svd_stats()
    ex0 = sess.exec_stats()
    pr0 = sess.exec_profile()
    sw0 = sess.sweep_profile()
    t0 = time.perf_counter()
    # the timed steps are the anneal's own: single-site layers, routed ZZ block and
    # the qubit-0 entropy of every step (recorded, as mps.cpp:355 does)
    if args.executor:
        sess.run(a[args.warmup:], b[args.warmup:], record=True)  # one launch: the steps overlap
    else:
        for k in range(args.steps):  # a progress line per step
            sess.run(a[args.warmup + k:args.warmup + k + 1], b[args.warmup + k:args.warmup + k + 1], record=True)
            sess.sync()
            progress(f"step {k + 1}/{args.steps}: {time.perf_counter() - t0:.2f}s")
    ent, trunc, bonds = sess.sync()
    t1 = time.perf_counter()
    progress(f"timed {args.steps} step(s): {t1 - t0:.2f}s")
    st1 = sess.svd_stats()
    wall = t1 - t0
    updates = ups * args.steps
    calls = st1["calls"] - st0["calls"]
    vis = st1["visited_rows"] - st0["visited_rows"]
    rot = st1["rotated_rows"] - st0["rotated_rows"]
    kernel_ms = st1["svd_ms"]  # executor: the persistent kernel; else the SVD launches
    jacobi_flops = 16.0 * vis + 18.0 * rot
    f_step, ups_host = gemm_flops_per_step(model, args.chi)
    caps = [min(args.chi, 2 ** min(x, 60), 2 ** min(args.n - x, 60)) for x in range(args.n + 1)]
    saturated = [int(x) for x in bonds] == caps
    gemm_flops = f_step * args.steps
    achieved = gemm_flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0
    if args.executor:
        ex1 = sess.exec_stats()
        ss = ex1["supersteps"] - ex0["supersteps"]
        kname = "k_tebd_lockstep"
        kernel = ("k_tebd_lockstep (updates in dataflow order from a shared queue: theta and reconstruction on FP64 "
                  "MFMA, block-Jacobi SVDs (FP64 VALU) on per-group barriers with point-to-point tournament rounds)")
        pr1 = sess.exec_profile()
        # the theta and reconstruction phases alone (workgroup 0's timers cover its
        # group's updates): the mean F per update x the updates the group ran (it takes
        # them from the shared queue, a mix of the whole schedule) / phase time, against the FP64 MFMA
        # peak of the group's CUs (one workgroup per CU)
        nb = (2 * args.chi + 7) // 8
        per_group = args.exec_workers if args.exec_workers > 0 else max(2, nb + (nb & 1)) // 2
        g_updates = pr1["group_updates"] - pr0["group_updates"]
        g_ms = (pr1["theta_ms"] - pr0["theta_ms"]) + (pr1["recon_ms"] - pr0["recon_ms"])
        g_tflops = g_updates * (f_step / max(ups_host, 1)) / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
        g_peak = FP64_PEAK_TFLOPS * per_group / 256.0
        timing = {"gemm_phases": {"tflops_per_group": g_tflops, "group_peak_tflops": g_peak,
                                  "frac_of_group_peak": g_tflops / g_peak,
                                  "note": "theta + reconstruction phases of workgroup 0's group on v_mfma_f64_16x16x4"},
                  "kernel_ms": kernel_ms, "kernel_share_of_wall": kernel_ms * 1e-3 / wall, "supersteps": ss,
                  "updates_per_superstep": (ex1["updates"] - ex0["updates"]) / max(ss, 1),
                  "group_barriers": os.environ.get("TESSERA_EXEC_LOCKSTEP", "0") != "1",
                  "point_to_point_rounds": (os.environ.get("TESSERA_EXEC_LOCKSTEP", "0") != "1"
                                            and os.environ.get("TESSERA_EXEC_P2P", "1") != "0"),
                  "workgroup0_phases": {k: round(pr1[k] - pr0[k], 3) for k in pr1},
                  "workgroup0_sweeps": sweep_summary(sw0, sess.sweep_profile(), pr1["rounds"] - pr0["rounds"]),
                  "svd_stop_ratio": float(os.environ.get("TESSERA_SVD_STOP", "0") or 0),
                  "svd_lq": not args.no_lq}
    else:
        kname = "k_svd_dist"
        kernel = "per-update launches; timed: the SVD launches (k_svd_dist / k_svd)"
        timing = {"svd_ms_per_update": kernel_ms / max(calls, 1), "svd_share_of_wall": kernel_ms * 1e-3 / wall}
    sweeps = (st1["sweeps"] - st0["sweeps"]) / max(calls, 1)
    traffic, pmc = pmc_traffic(f"{kname} n={args.n} chi={args.chi}", updates)
    return {
        "metric": "TEBD two-site updates/s @n=128 chi=256",
        "value": updates / wall,
        "unit": "two-site updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (fp64)",
        "data": "synthetic: seeded range-4 chain, saturated MPS (random isometries, bond weights exp(-k/40))",
        "config": {"workload": f"MPS-TEBD Trotter step n={args.n} chi={args.chi} range<=4 dt={DT}",
                   "updates_per_step": ups, "max_bond": int(bonds.max()), "bonds_at_capacity": saturated,
                   "svd_calls": calls, "jacobi_sweeps_per_svd": round(sweeps, 2),
                   "max_trunc": trunc, "entropy_last": float(ent[-1]) if len(ent) else None,
                   "executor_groups": args.executor, "executor_workers": args.exec_workers},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "basis": ("F = 32 Dl Dr (mid + keep) flops per update (theta + reconstruction GEMMs) summed "
                               f"on the host over {args.steps} x {ups_host} updates of the saturated chain, "
                               "/ the kernel's HIP-event time, / the FP64 peak (the arithmetic used)"),
                     "flops_per_update_saturated": f_step / max(ups_host, 1),
                     "jacobi_flops": jacobi_flops,
                     "jacobi_tflops": jacobi_flops / (kernel_ms * 1e-3) / 1e12 if kernel_ms > 0 else 0.0,
                     "traffic_source": None if pmc is None else pmc.get("source"),
                     "kernel": kernel, **timing},
    }


def main():
    args = parse()
    if getattr(args, "step_profile", False):
        os.environ["TESSERA_EXEC_STEP_PROFILE"] = "1"  # read when the session is created
    if args.gpus != 1:
        print(json.dumps({"error": "this implementation runs on one GPU; multi-GPU sharding is designed, not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.workload == "sv":
        sv_bench(args)
        return
    if args.workload == "batch":
        batch_bench(args)
        return
    if args.workload == "learned":
        learned_bench(args)
        return
    if args.workload == "gnn":
        gnn_bench(args)
        return
    if args.workload == "chain12":
        config1_bench(args)
        return
    import numpy as np

    from tessera_amd import instances as I
    from tessera_amd import qa

    model = I.chain_range(args.n, RANGE, SEED)
    sched = I.linear(SCHED_STEPS)
    k0 = SCHED_STEPS // 2
    with qa.TebdSession(model, args.chi, DT) as sess:
        out = tebd_timed(args, model, sched, k0, sess)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.chi)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
