"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
reference, never as the product path.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libtessera_ref.so"
REF_HARNESS = HERE / "_ref" / "ref_harness"
_lib = None
vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        sz = ctypes.c_size_t
        L.oracle_qa_anneal.restype = ctypes.c_double
        L.oracle_qa_anneal.argtypes = [vp, sz, vp, vp, vp, sz, vp, vp, sz, sz, vp, vp]
        L.oracle_tebd.restype = ctypes.c_int
        L.oracle_tebd.argtypes = [vp, sz, vp, vp, vp, sz, vp, vp, sz, ctypes.c_double, sz, ctypes.c_int, vp, vp, vp,
                                  vp, vp, vp, ctypes.c_int, vp]
        L.oracle_sv.restype = ctypes.c_int
        L.oracle_sv.argtypes = [vp, sz, vp, vp, vp, sz, vp, vp, sz, ctypes.c_double, vp, vp, vp, vp, vp]
        L.oracle_herm_eig.restype = ctypes.c_int
        L.oracle_herm_eig.argtypes = [vp, sz, vp, vp]
        L.oracle_expm.restype = ctypes.c_int
        L.oracle_expm.argtypes = [vp, sz, ctypes.c_double, ctypes.c_double, vp]
        L.oracle_svd.restype = ctypes.c_int
        L.oracle_svd.argtypes = [vp, sz, sz, vp, vp, vp]
        L.oracle_update_probe.restype = ctypes.c_double
        L.oracle_update_probe.argtypes = [ctypes.c_int, ctypes.c_int, vp, vp]
        L.oracle_upstream_instance.restype = sz
        L.oracle_upstream_instance.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, sz]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


def _edges(model):
    return (np.ascontiguousarray(model.ei, np.uintp), np.ascontiguousarray(model.ej, np.uintp),
            np.ascontiguousarray(model.ew, np.float64))


def tebd(model, sched, dt, chi, zz_substeps=1, svd_mode=0, want_norms=False, want_state=False):
    """Full TEBD run. svd_mode 0: upstream Gram SVD (bit-exact with the upstream
    engine); 1: accurate one-sided Jacobi SVD (the upstream update in exact
    arithmetic). `ambiguous[k]` counts the updates of step k whose kept rank is
    set by singular values at the rounding level (cut 1e-13 and 1e-15 keep
    different ranks)."""
    n, st = model.n, sched.steps
    ei, ej, ew = _edges(model)
    spins = np.zeros(n, np.int8)
    sc = np.zeros(4)
    ent = np.zeros(max(st, 1))
    mb = np.zeros(max(st, 1), np.int32)
    amb = np.zeros(max(st, 1), np.int32)
    nr = np.zeros(max(st, 1)) if want_norms else None
    state = np.zeros(2 << n) if want_state else None
    a = np.ascontiguousarray(sched.a, np.float64)
    b = np.ascontiguousarray(sched.b, np.float64)
    rc = lib().oracle_tebd(_p(np.ascontiguousarray(model.h, np.float64)), n, _p(ei), _p(ej), _p(ew), model.m, _p(a),
                           _p(b), st, dt, chi, zz_substeps, _p(spins), _p(sc), _p(ent), _p(mb), _p(nr), _p(state),
                           svd_mode, _p(amb))
    if rc != 0:
        raise RuntimeError(f"oracle_tebd failed ({rc})")
    out = dict(spins=spins, energy=sc[0], expected_energy=sc[1], prob=sc[2], max_trunc=sc[3], entropy=ent[:st],
               max_bond=mb[:st], ambiguous=amb[:st])
    if want_norms:
        out["norm2"] = nr[:st]
    if want_state:
        out["state"] = state[0::2] + 1j * state[1::2]
    return out


def qa_anneal(model, sched, chi):
    """tessera_qa_anneal semantics (dt = 0.1)."""
    ei, ej, ew = _edges(model)
    spins = np.zeros(model.n, np.int8)
    ent = np.zeros(sched.steps)
    e = lib().oracle_qa_anneal(_p(np.ascontiguousarray(model.h, np.float64)), model.n, _p(ei), _p(ej), _p(ew),
                               model.m, _p(np.ascontiguousarray(sched.a)), _p(np.ascontiguousarray(sched.b)),
                               sched.steps, chi, _p(spins), _p(ent))
    return e, spins, ent


def sv(model, sched, dt, want_state=False):
    """rust/src/quantum.rs anneal_statevector restated."""
    n, st = model.n, sched.steps
    ei, ej, ew = _edges(model)
    spins = np.zeros(n, np.int8)
    sc = np.zeros(3)
    ent = np.zeros(st)
    en = np.zeros(st)
    state = np.zeros(2 << n) if want_state else None
    rc = lib().oracle_sv(_p(np.ascontiguousarray(model.h, np.float64)), n, _p(ei), _p(ej), _p(ew), model.m,
                         _p(np.ascontiguousarray(sched.a)), _p(np.ascontiguousarray(sched.b)), st, dt, _p(spins),
                         _p(sc), _p(ent), _p(en), _p(state))
    if rc != 0:
        raise RuntimeError(f"oracle_sv failed ({rc})")
    out = dict(spins=spins, energy=sc[0], expected_energy=sc[1], success_prob=sc[2], entropy=ent, energy_trace=en)
    if want_state:
        out["state"] = state[0::2] + 1j * state[1::2]
    return out


def upstream_instance(seed: int, skip: int, n: int):
    """Instance number `skip` of the std::mt19937_64(seed) stream used by
    cuda/tests/test_mps.cpp:100-111 and 159-167."""
    from tessera_amd.instances import Ising
    h = np.zeros(n)
    cap = n * n
    ei = np.zeros(cap, np.uintp)
    ej = np.zeros(cap, np.uintp)
    ew = np.zeros(cap)
    m = lib().oracle_upstream_instance(seed, skip, n, _p(h), _p(ei), _p(ej), _p(ew), cap)
    return Ising(h, ei[:m].astype(np.uint64), ej[:m].astype(np.uint64), ew[:m].copy())


def herm_eig(hm):
    hm = np.ascontiguousarray(hm, np.complex128)
    n = hm.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n), np.complex128)
    lib().oracle_herm_eig(_p(hm.view(np.float64)), n, _p(w), _p(v.view(np.float64)))
    return w, v


def expm(hm, coeff):
    hm = np.ascontiguousarray(hm, np.complex128)
    n = hm.shape[0]
    out = np.zeros((n, n), np.complex128)
    lib().oracle_expm(_p(hm.view(np.float64)), n, coeff.real, coeff.imag, _p(out.view(np.float64)))
    return out


def gram_svd(a):
    a = np.ascontiguousarray(a, np.complex128)
    r, c = a.shape
    k = min(r, c)
    u = np.zeros((r, k), np.complex128)
    s = np.zeros(k)
    vh = np.zeros((k, c), np.complex128)
    lib().oracle_svd(_p(a.view(np.float64)), r, c, _p(u.view(np.float64)), _p(s), _p(vh.view(np.float64)))
    return u, s, vh


def update_probe(chi: int, max_sweeps: int):
    sw = ctypes.c_int()
    sweep_s = ctypes.c_double()
    t = lib().oracle_update_probe(chi, max_sweeps, ctypes.byref(sw), ctypes.byref(sweep_s))
    return t, sw.value, sweep_s.value
