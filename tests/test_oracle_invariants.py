"""Oracle checks that do not depend on the upstream build: closed forms, dense
matrix products of the same Trotter splittings, invariants, and the upstream
test expectations (cuda/tests/test_linalg.cpp, test_mps.cpp, rust quantum.rs,
schedule.rs). One test plants a plausible mistake in the oracle's inputs to show
the dense check detects it."""
import math

import numpy as np
import pytest

import pyoracle
from tessera_amd import instances as I


def dense_sv_splitting(model, sched, dt):
    """quantum.rs:80-86 as dense matrices: e^{-iD b dt/2} (x)_i e^{+i a dt X} e^{-iD b dt/2}."""
    n = model.n
    dim = 1 << n
    x = np.arange(dim)
    diag = np.array([model.energy([1 if (v >> i) & 1 else -1 for i in range(n)]) for v in range(dim)])
    psi = np.full(dim, 1 / math.sqrt(dim), complex)
    for a, b in zip(sched.a, sched.b):
        ph = np.exp(-1j * b * dt * 0.5 * diag)
        rx = np.array([[math.cos(a * dt), 1j * math.sin(a * dt)], [1j * math.sin(a * dt), math.cos(a * dt)]])
        full = np.array([[1.0]])
        for _ in range(n):
            full = np.kron(rx, full)
        psi = ph * (full @ (ph * psi))
    return psi, diag


def dense_mps_splitting(model, sched, dt):
    """mps.cpp:309-356 as dense matrices; physical index 0 <-> +1 in bit i."""
    n = model.n
    dim = 1 << n
    x = np.arange(dim)
    s = np.array([1 - 2 * ((x >> i) & 1) for i in range(n)])
    zz = np.zeros(dim)
    for i, j, w in zip(model.ei, model.ej, model.ew):
        zz += w * s[int(i)] * s[int(j)]
    psi = np.ones(dim, complex) / math.sqrt(dim)
    for A, B in zip(sched.a, sched.b):
        for half in range(2):
            for i in range(n):
                g = np.array([[-B * model.h[i], -A], [-A, B * model.h[i]]])
                w, v = np.linalg.eigh(g)
                u = v @ np.diag(np.exp(-1j * dt / 2 * w)) @ v.conj().T
                psi = np.einsum("ab,xbz->xaz", u, psi.reshape(-1, 2, 1 << i)).reshape(dim)
            if half == 0:
                psi = psi * np.exp(1j * B * dt * zz)
    return psi


def fid(a, b):
    return abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


# ---- linalg (cuda/tests/test_linalg.cpp:30-98) through the oracle
def test_pauli_spectra_and_expm_identities():
    w, _ = pyoracle.herm_eig(np.diag([1.0, -1.0]).astype(complex))
    assert sorted(w) == pytest.approx([-1, 1], abs=1e-9)
    w, _ = pyoracle.herm_eig(np.array([[0, -1j], [1j, 0]]))
    assert abs(abs(w[0]) - 1) < 1e-9 and abs(w[0] + w[1]) < 1e-9
    X = np.array([[0, 1], [1, 0]], complex)
    assert np.abs(pyoracle.expm(X, -1j * math.pi / 2) - (-1j * X)).max() < 1e-9
    H = np.array([[2, 1 + 2j, -0.5 + 1j], [1 - 2j, -1, -1j], [-0.5 - 1j, 1j, 0.5]])
    w, v = pyoracle.herm_eig(H)
    assert np.abs(v @ np.diag(w) @ v.conj().T - H).max() < 1e-7
    U = pyoracle.expm(H, -0.83j)
    assert np.abs(U @ U.conj().T - np.eye(3)).max() < 1e-9
    assert np.abs(U - __import__("scipy.linalg", fromlist=["expm"]).expm(-0.83j * H)).max() < 1e-9


def test_gram_svd_reconstructs_both_branches():
    rng = np.random.default_rng(1)
    for r, c in ((6, 4), (4, 7)):
        a = rng.normal(size=(r, c)) + 1j * rng.normal(size=(r, c))
        u, s, vh = pyoracle.gram_svd(a)
        assert np.abs(u @ np.diag(s) @ vh - a).max() < 1e-9
        assert np.all(np.diff(s) <= 0)
        assert np.abs(s - np.linalg.svd(a, compute_uv=False)).max() < 1e-9


# ---- state vector (rust/src/quantum.rs)
@pytest.mark.parametrize("n", [3, 6, 8])
def test_sv_oracle_matches_dense_product(n):
    model = I.chain_nnn(n, 7 + n)
    sched = I.linear(15)
    o = pyoracle.sv(model, sched, 0.1, want_state=True)
    psi, diag = dense_sv_splitting(model, sched, 0.1)
    assert np.abs(o["state"] - psi).max() < 1e-12
    p = np.abs(psi) ** 2
    assert abs(o["expected_energy"] - (p @ diag) / p.sum()) < 1e-12
    best = int(np.argmax(p))
    assert o["success_prob"] == pytest.approx(p[best] / p.sum(), abs=1e-12)
    assert [int(v) for v in o["spins"]] == [1 if (best >> i) & 1 else -1 for i in range(n)]


def test_planted_sign_error_is_detected():
    """A plausible mistake (the X rotation taken as e^{-i a dt X}) must fail the
    dense check: planted by flipping the sign of the schedule's A."""
    model = I.chain_nnn(5, 2)
    sched = I.linear(15)
    wrong = I.Schedule(-sched.a, sched.b)
    psi, _ = dense_sv_splitting(model, sched, 0.1)
    o = pyoracle.sv(model, wrong, 0.1, want_state=True)
    assert np.abs(o["state"] - psi).max() > 1e-3


def test_sv_single_spin_closed_form():
    model = I.make_ising([1.0], [])
    sched = I.smooth(400)
    o = pyoracle.sv(model, sched, 0.1, want_state=True)
    psi = np.array([1, 1], complex) / math.sqrt(2)
    for a, b in zip(sched.a, sched.b):
        ph = np.exp(-1j * b * 0.05 * np.array([1.0, -1.0]))
        rx = np.array([[math.cos(a * 0.1), 1j * math.sin(a * 0.1)], [1j * math.sin(a * 0.1), math.cos(a * 0.1)]])
        psi = ph * (rx @ (ph * psi))
    assert np.abs(o["state"] - psi).max() < 1e-13
    assert o["spins"][0] == 1 and abs(o["energy"] + 1) < 1e-9 and o["success_prob"] > 0.9


def test_quantum_rs_expectations_on_oracle():
    o = pyoracle.sv(I.make_ising([0.0, 0.0], [(0, 1, 1.0)]), I.smooth(500), 0.1)
    assert abs(o["energy"] + 1) < 1e-9
    o = pyoracle.sv(I.make_ising([0.3, -0.3], [(0, 1, -1.0)]), I.smooth(400), 0.1)
    tr = o["entropy"]
    assert len(tr) == 400 and tr.max() > 0.05 and tr.max() > tr[0] and tr.max() >= tr[-1] and tr.max() <= 1 + 1e-9
    o = pyoracle.sv(I.make_ising([0.5, -0.5, 0.2], [(0, 1, 0.4)]), I.smooth(50), 0.1)
    assert o["entropy"][0] < 0.2


# ---- TEBD in exact arithmetic (svd_mode 1) vs the dense product of its splitting
@pytest.mark.parametrize("n,chi", [(4, 4), (6, 8)])
def test_tebd_accurate_mode_matches_dense(n, chi):
    model = I.chain_range(n, 3, 20 + n)
    sched = I.linear(12)
    o = pyoracle.tebd(model, sched, 0.1, chi, svd_mode=1, want_state=True)
    psi = dense_mps_splitting(model, sched, 0.1)
    assert 1 - fid(psi, o["state"]) < 1e-12
    assert o["max_trunc"] < 1e-20


def test_gram_svd_noise_floor_on_rank_deficient_input():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(12, 9)) + 1j * rng.normal(size=(12, 9))
    a[:, 8] = a[:, 0] * 2  # rank deficient
    u, s, vh = pyoracle.gram_svd(a)  # upstream Gram SVD: noise-level last value
    assert abs(s[-1]) < 1e-6
    s_ref = np.linalg.svd(a, compute_uv=False)
    assert np.abs(s[:-1] - s_ref[:-1]).max() < 1e-9


# ---- schedules (rust/src/schedule.rs:48-75)
@pytest.mark.parametrize("make", [I.linear, I.smooth])
def test_schedule_endpoints_monotone(make):
    s = make(64)
    assert abs(s.a[0] - 1) < 1e-12 and abs(s.b[0]) < 1e-12 and abs(s.a[-1]) < 1e-12 and abs(s.b[-1] - 1) < 1e-12
    assert np.all(np.diff(s.a) <= 1e-12) and np.all(np.diff(s.b) >= -1e-12)
    if make is I.smooth:
        assert np.abs(s.a + s.b - 1).max() < 1e-12
    assert make(1).steps == 2
