"""GPU MPS-TEBD vs the oracle.

The oracle restates the upstream engine bit-for-bit (svd_mode 0, pinned against
the upstream build in test_oracle_vs_ref.py) and also runs the same algorithm
with an accurate SVD (svd_mode 1). The upstream Gram SVD keeps noise singular
values (~1e-8) whose vectors it normalises by that noise, which blows the MPS
norm up once bonds fill (see test_oracle_vs_ref.py); the GPU computes the
upstream update in exact arithmetic, so it is pinned against svd_mode 1 and the
exact dense evolution of the same Trotter splitting, and against the upstream
itself wherever the upstream stays numerically sound (its readouts).

Bars: state infidelity <= 1e-10 at exact chi, entropy and <H> within 1e-9,
kept ranks identical, discarded weight within 1e-10, spins identical.
"""
import math

import numpy as np
import pytest

import pyoracle
from tessera_amd import instances as I
from tessera_amd import qa

pytestmark = pytest.mark.gpu


def dense_mps_splitting(model, sched, dt, bias=None, psi0=None):
    """Exact evolution of the MPS engine's splitting U1(dt/2) ZZ(dt) U1(dt/2)
    (mps.cpp:309-356), physical index 0 <-> +1 stored in bit i. With bias the
    single-site generator is -A X - (B h_i + A bias_i) Z; psi0 replaces |+>^n."""
    n = model.n
    dim = 1 << n
    x = np.arange(dim)
    spins = np.array([1 - 2 * ((x >> i) & 1) for i in range(n)])
    zz = np.zeros(dim)
    for i, j, w in zip(model.ei, model.ej, model.ew):
        zz += w * spins[int(i)] * spins[int(j)]
    psi = np.ones(dim, complex) / math.sqrt(dim) if psi0 is None else np.asarray(psi0, complex).copy()
    bias = np.zeros(n) if bias is None else np.asarray(bias, float)
    ents = []

    def layer(psi, A, B):
        for i in range(n):
            hz = B * model.h[i] + A * bias[i]
            g = np.array([[-hz, -A], [-A, hz]])
            w, V = np.linalg.eigh(g)
            U = V @ np.diag(np.exp(-1j * dt / 2 * w)) @ V.conj().T
            psi = np.einsum("ab,xbz->xaz", U, psi.reshape(-1, 2, 1 << i)).reshape(dim)
        return psi

    for A, B in zip(sched.a, sched.b):
        psi = layer(psi, A, B)
        psi = psi * np.exp(1j * B * dt * zz)
        psi = layer(psi, A, B)
        p = psi.reshape(-1, 2)
        rho = p.T @ p.conj()
        ev = np.clip(np.linalg.eigvalsh(rho / np.trace(rho).real), 0, 1)
        ents.append(-sum(l * math.log2(l) for l in ev if l > 1e-12))
    return psi, np.array(ents)


def _fid(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


def _gpu_run(model, sched, dt, chi, ns=1):
    s = qa.TebdSession(model, chi, dt, ns)
    s.run(sched.a, sched.b, record=True)
    ent, tr, bonds = s.sync()
    spins, sc = s.readout()
    state = s.dense() if model.n <= 16 else None
    s.close()
    return dict(entropy=ent, max_trunc=tr, bonds=bonds, spins=spins, energy=sc[0], expected_energy=sc[1],
                prob=sc[2], state=state)


@pytest.mark.parametrize("n,chi,steps", [(6, 8, 30), (8, 16, 20)])
def test_exact_chi_matches_dense(n, chi, steps):
    model = I.chain_nnn(n, 40 + n)
    sched = I.linear(steps)
    g = _gpu_run(model, sched, 0.1, chi)
    psi, ents = dense_mps_splitting(model, sched, 0.1)
    inf = 1 - _fid(psi, g["state"])
    # amplitude_prob (mps.cpp:197-216) of the readout: physical index 0 <-> +1 in bit i
    x = sum(1 << i for i, s_ in enumerate(g["spins"]) if s_ < 0)
    prob = abs(psi[x]) ** 2 / np.vdot(psi, psi).real
    print(f"n={n} chi={chi} infidelity={inf:.2e} ent_err={np.abs(g['entropy'] - ents).max():.2e} "
          f"prob={g['prob']:.12f} dense={prob:.12f}", flush=True)
    assert inf <= 1e-10
    assert np.abs(g["entropy"] - ents).max() <= 1e-9
    assert g["max_trunc"] <= 1e-20
    assert abs(g["prob"] - prob) <= 1e-10


def test_exact_chi_long_range_matches_dense():
    """Random 3-regular graph on 10 spins (edges of range up to 9, routed by long
    swap chains; mps.cpp:126-145) at chi = 32, the largest bond of 10 sites:
    nothing is truncated, so the routed evolution equals the dense one."""
    model = I.random_regular3(10, 31)
    assert model.max_range() >= 5
    sched = I.linear(12)
    g = _gpu_run(model, sched, 0.1, 32)
    psi, ents = dense_mps_splitting(model, sched, 0.1)
    inf = 1 - _fid(psi, g["state"])
    print(f"range {model.max_range()}: infidelity={inf:.2e} ent_err={np.abs(g['entropy'] - ents).max():.2e}",
          flush=True)
    assert inf <= 1e-10
    assert np.abs(g["entropy"] - ents).max() <= 1e-9
    assert g["max_trunc"] <= 1e-20


def test_config1_prefix_matches_accurate_oracle():
    inst = I.chain12_anneal()
    sched = I.linear(200)
    sched = I.Schedule(sched.a[:12], sched.b[:12])  # first 12 steps of the T=10 run
    g = _gpu_run(inst.model, sched, inst.dt, inst.chi)
    o = pyoracle.tebd(inst.model, sched, inst.dt, inst.chi, svd_mode=1, want_state=True)
    inf = 1 - _fid(o["state"], g["state"])
    print(f"n=12 chain[:12] infidelity={inf:.2e} ent_err={np.abs(g['entropy'] - o['entropy']).max():.2e}", flush=True)
    assert inf <= 1e-10
    assert np.abs(g["entropy"] - o["entropy"]).max() <= 1e-9
    assert abs(g["expected_energy"] - o["expected_energy"]) <= 1e-9
    assert abs(g["prob"] - o["prob"]) <= 1e-10
    assert np.array_equal(g["spins"], o["spins"])


def test_truncated_ranks_and_weight_match_accurate_oracle():
    """chi = 4 on a range-3 chain: heavy truncation with near-degenerate
    singular values at the cut, so the max discarded weight is only defined up
    to the conditioning of the cut. Bar: kept ranks identical step by step, and
    the GPU no further from the accurate oracle than an independent LAPACK-SVD
    TEBD is (tests/tebd_numpy.py)."""
    from tebd_numpy import tebd_numpy
    model = I.chain_range(10, 3, 11)
    sched = I.linear(25)
    chi = 4
    s = qa.TebdSession(model, chi, 0.1)
    bonds_gpu = []
    for k in range(sched.steps):
        s.run(sched.a[k:k + 1], sched.b[k:k + 1])
        bonds_gpu.append(int(s.sync()[2].max()))
    ent, tr, bonds = s.sync()
    spins, sc = s.readout()
    s.close()
    o = pyoracle.tebd(model, sched, 0.1, chi, svd_mode=1)
    tr_np, bonds_np = tebd_numpy(model, sched, 0.1, chi)
    spread = abs(tr_np - o["max_trunc"])
    print(f"trunc gpu={tr:.9e} oracle={o['max_trunc']:.9e} lapack={tr_np:.9e} spread={spread:.2e}", flush=True)
    assert bonds_gpu == [int(x) for x in o["max_bond"]] == bonds_np
    assert abs(tr - o["max_trunc"]) <= 10 * spread + 1e-12
    assert np.abs(ent - o["entropy"]).max() <= 1e-4
    assert np.array_equal(spins, o["spins"])
    assert abs(sc[1] - o["expected_energy"]) <= 1e-4
    assert abs(sc[2] - o["prob"]) <= 1e-4 * max(o["prob"], 1e-3)


def test_distributed_svd_path_matches(monkeypatch):
    """The cooperative multi-workgroup SVD (used when theta exceeds one
    workgroup's LDS) forced at small sizes gives the same run."""
    model = I.chain_nnn(8, 9)
    sched = I.linear(15)
    base = _gpu_run(model, sched, 0.1, 16)
    monkeypatch.setenv("TESSERA_FORCE_DIST_SVD", "1")
    dist = _gpu_run(model, sched, 0.1, 16)
    print(f"dist infidelity={1 - _fid(base['state'], dist['state']):.2e}", flush=True)
    assert 1 - _fid(base["state"], dist["state"]) <= 1e-12
    assert np.abs(base["entropy"] - dist["entropy"]).max() <= 1e-10
    assert np.array_equal(base["bonds"], dist["bonds"])


@pytest.mark.parametrize("n,chi,steps", [(12, 64, 12), (10, 8, 25)])
def test_small_svd_kernel_matches_workgroup_svd(monkeypatch, n, chi, steps):
    """The per-update SVD of thetas up to 64 columns x 128 rows runs on 16-lane
    column-pair groups (k_svd_small); the one-pair-per-wave workgroup SVD
    (k_svd) gives the same run. At exact chi (n=12 chain, thetas up to 128 x 32
    and 64 x 64, both row counts of the small kernel) the states agree to
    rounding. A truncated run of this scheme (no canonical form) is
    ill-conditioned: two LAPACK drivers (gesdd, gesvd) on the chi=8 case end
    7.5e-3 apart in fidelity and 1.6e-2 in <H>, with discarded weights equal to
    1.3e-4 relative. There the bar is identical bond dimensions and a discarded
    weight within ten times the LAPACK spread."""
    import scipy.linalg as sl
    from tebd_numpy import tebd_numpy
    model = I.chain_nnn(n, 70 + n)
    sched = I.linear(steps)
    small = _gpu_run(model, sched, 0.1, chi)
    monkeypatch.setenv("TESSERA_SMALL_SVD", "0")
    wg = _gpu_run(model, sched, 0.1, chi)
    inf = 1 - _fid(small["state"], wg["state"])
    print(f"n={n} chi={chi}: infidelity={inf:.2e} max bond={small['bonds'].max()} "
          f"trunc {small['max_trunc']:.6e} / {wg['max_trunc']:.6e}", flush=True)
    assert np.array_equal(small["bonds"], wg["bonds"])
    if wg["max_trunc"] <= 1e-20:  # exact chi
        assert inf <= 1e-12
        assert np.abs(small["entropy"] - wg["entropy"]).max() <= 1e-10
        assert np.array_equal(small["spins"], wg["spins"])
        return
    t_dd, _ = tebd_numpy(model, sched, 0.1, chi)
    t_vd, _ = tebd_numpy(model, sched, 0.1, chi, svd=lambda a: sl.svd(a, full_matrices=False, lapack_driver="gesvd"))
    spread = abs(t_dd - t_vd)
    print(f"LAPACK trunc gesdd={t_dd:.6e} gesvd={t_vd:.6e} spread={spread:.2e}", flush=True)
    assert abs(small["max_trunc"] - wg["max_trunc"]) <= 10 * spread + 1e-12


def test_upstream_mps_suite():
    """cuda/tests/test_mps.cpp:78-152 through tessera_qa_anneal_ex."""
    sched = I.smooth(300)
    r = qa.mps_anneal(I.make_ising([1.0], []), sched, 0.1, 16)
    assert r.spins[0] == 1 and abs(r.energy + 1.0) < 1e-9
    r = qa.mps_anneal(I.make_ising([0.0, 0.0], [(0, 1, 1.0)]), I.smooth(400), 0.1, 16)
    assert abs(r.energy + 1.0) < 1e-9 and r.spins[0] == r.spins[1]
    # frustrated 3-ring: <H> -> exact ground energy at long T
    ring = I.make_ising([0.5, -0.3, 0.2], [(0, 1, -1.0), (1, 2, -1.0), (2, 0, -1.0)])
    e0 = min(ring.energy([1 if (x >> i) & 1 else -1 for i in range(3)]) for x in range(8))
    r = qa.mps_anneal(ring, I.smooth(3200), 0.1, 4)
    assert abs(r.expected_energy - e0) < 0.05
    r = qa.mps_anneal(I.make_ising([0.3, -0.3], [(0, 1, -1.0)]), I.smooth(400), 0.1, 16)
    assert 0.05 < r.entropy.max() <= 1.0 + 1e-9


def test_upstream_random_instances_and_readout_parity():
    """cuda/tests/test_mps.cpp:96-123 on the same std::mt19937_64(123) instances:
    >= 5/6 exact hits, worst trunc < 1e-9, and readout spins identical to the
    upstream engine's (oracle svd_mode 0 is bit-exact with it)."""
    hits, worst = 0, 0.0
    sched = I.smooth(1200)
    for t in range(6):
        model = pyoracle.upstream_instance(123, t, 6)
        e_exact = min(model.energy([1 if (x >> i) & 1 else -1 for i in range(6)]) for x in range(64))
        r = qa.mps_anneal(model, sched, 0.08, 8)
        up = pyoracle.tebd(model, sched, 0.08, 8, svd_mode=0)
        worst = max(worst, r.max_trunc)
        hits += abs(r.energy - e_exact) < 1e-6
        assert np.array_equal(r.spins, up["spins"]), f"instance {t}"
        assert r.energy == up["energy"]
    assert hits >= 5
    assert worst < 1e-9


def test_drop_in_entry_runs_large_bonds_in_the_executor():
    """tessera_qa_anneal_ex runs chains whose bond capacity reaches 128 in the
    lockstep executor (the per-step max bonds then come from the entropy
    snapshots). n = 14 nn + nnn chain at chi = 128, exact (capacity 2^7 = 128):
    the entropy trace equals the dense evolution of the same splitting and the
    readout's prob its amplitude, and the max bond after step k equals the largest
    bond that a session running the same executor over the first k steps in one
    launch reads back at its end (the same arithmetic up to step k; a launch
    boundary or the per-update path's SVD kernels round differently, so their
    rank decisions at the 1e-14 cut can differ by a few)."""
    model = I.chain_nnn(14, 61)
    sched = I.linear(8)
    r = qa.mps_anneal(model, sched, 0.1, 128)
    psi, ents = dense_mps_splitting(model, sched, 0.1)
    x = sum(1 << i for i, s_ in enumerate(r.spins) if s_ < 0)
    prob = abs(psi[x]) ** 2 / np.vdot(psi, psi).real
    bonds = []
    for k in range(1, sched.steps + 1):
        s = qa.TebdSession(model, 128, 0.1)
        s.set_executor(-1)
        s.run(sched.a[:k], sched.b[:k], record=True)
        bonds.append(int(s.sync()[2][1:model.n].max()))
        s.close()
    print(f"max bonds {r.max_bond.tolist()}, at the end of launches of 1..{sched.steps} steps {bonds}; "
          f"ent_err={np.abs(r.entropy - ents).max():.2e} prob={r.prob:.12f} dense={prob:.12f}", flush=True)
    assert r.max_bond.tolist() == bonds
    assert np.abs(r.entropy - ents).max() <= 1e-9
    assert abs(r.prob - prob) <= 1e-10
    assert r.max_trunc <= 1e-20


def test_plain_abi_matches_ex_defaults():
    model = I.chain_nnn(5, 3)
    sched = I.smooth(40)
    r = qa.quantum_anneal(model, sched, 8)
    ex = qa.mps_anneal(model, sched, 0.1, 8)
    assert np.array_equal(r.spins, ex.spins) and r.energy == ex.energy
    assert np.array_equal(r.entropy, ex.entropy)


@pytest.mark.parametrize("ns,biased", [(1, False), (2, True)])
def test_step_graph_replay_matches_direct_launches(monkeypatch, ns, biased):
    """The per-update path captures one Trotter step (single-site layers, every
    routed two-site update, the entropy sweep, a step-index node) into a hipGraph
    and replays it; the step's schedule values, ZZ gate phasors and entropy slot
    come from device memory, computed on the host by the same calls as the
    direct launches. Same kernels, same values: bit-identical to the direct path
    (TESSERA_STEP_GRAPH=0), over run() calls of different lengths with and
    without recording, with two ZZ slices and with a warm-start bias."""
    import time
    model = I.chain_range(10, 3, 77)
    sched = I.linear(24)
    bias = 0.3 * np.tanh(model.h) if biased else None
    parts = [(0, 7, True), (7, 9, False), (9, 24, True)]

    def go():
        s = qa.TebdSession(model, 16, 0.1, ns)
        try:
            if bias is not None:
                s.reset_warm(bias)
                s.set_bias(bias)
            t0 = time.perf_counter()
            for lo, hi, rec in parts:
                s.run(sched.a[lo:hi], sched.b[lo:hi], record=rec)
            ent, tr, bonds = s.sync()
            wall = time.perf_counter() - t0
            spins, sc = s.readout()
            return dict(ent=ent, tr=tr, bonds=bonds, state=s.dense(), sc=sc, wall=wall)
        finally:
            s.close()

    graph = go()
    monkeypatch.setenv("TESSERA_STEP_GRAPH", "0")
    direct = go()
    print(f"ns={ns} biased={biased}: graph {graph['wall']:.3f}s, direct {direct['wall']:.3f}s, "
          f"bit-identical state {np.array_equal(graph['state'], direct['state'])}", flush=True)
    assert np.array_equal(graph["state"], direct["state"])
    assert np.array_equal(graph["ent"], direct["ent"])
    assert np.array_equal(graph["bonds"], direct["bonds"])
    assert graph["tr"] == direct["tr"]
    assert np.array_equal(graph["sc"], direct["sc"])
