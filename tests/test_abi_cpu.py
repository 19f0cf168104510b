"""The C-ABI library without a GPU: it loads, exports every symbol
include/tessera_qa.h declares, and rejects bad arguments before any device work
(upstream semantics for tessera_solve_qa: 0.0 on null pointers, NaN for n > 24,
rust/src/capi.rs:121-126)."""
import ctypes
import math
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tessera_amd import instances as I
from tessera_amd import qa

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "tessera_qa.h"


def declared_symbols():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(tessera_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", str(qa.library_path())], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (tessera_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert "tessera_qa_anneal" in exported  # cuda/src/ffi.cpp:10 drop-in


def test_library_has_gfx950_code_object():
    data = qa.library_path().read_bytes()
    assert b"gfx950" in data


def _args(h, ei, ej, ew, a, b):
    P = qa._p
    return P(h), h.shape[0], P(ei), P(ej), P(ew), ew.shape[0], P(a), P(b), a.shape[0]


def test_extended_entry_rejects_bad_arguments():
    L = qa.lib()
    h = np.zeros(3)
    ew = np.ones(2)
    a = np.ones(4)
    b = np.ones(4)
    sp = np.zeros(3, np.int8)
    sc = np.zeros(4)
    bad_idx = np.array([0, 5], np.uintp)
    ok = np.array([1, 2], np.uintp)
    rc = L.tessera_qa_anneal_ex(*_args(h, bad_idx, ok, ew, a, b), 0.1, 8, 1, None, qa._p(sp), qa._p(sc), None, None)
    assert rc == -1 and b"edge index" in L.tessera_last_error()
    hn = np.array([0.0, math.nan, 0.0])
    rc = L.tessera_qa_anneal_ex(*_args(hn, ok - 1, ok, ew, a, b), 0.1, 8, 1, None, qa._p(sp), qa._p(sc), None, None)
    assert rc == -4
    rc = L.tessera_qa_anneal_ex(*_args(h, ok - 1, ok, ew, a, b), 0.1, 0, 1, None, qa._p(sp), qa._p(sc), None, None)
    assert rc == -1  # chi == 0
    rc = L.tessera_qa_anneal_ex(None, 3, None, None, None, 0, None, None, 0, 0.1, 8, 1, None, qa._p(sp), qa._p(sc),
                                None, None)
    assert rc == -1


def test_plain_entry_never_throws_and_reports_nan():
    h = np.array([0.0, math.inf])
    sp = np.zeros(2, np.int8)
    sp[:] = 7
    a = np.ones(3)
    e = qa.lib().tessera_qa_anneal(qa._p(h), 2, None, None, None, 0, qa._p(a), qa._p(a), 3, 4, qa._p(sp), None)
    assert math.isnan(e) and list(sp) == [1, 1]


def test_solve_qa_upstream_guards():
    L = qa.lib()
    sp = np.zeros(30, np.int8)
    assert L.tessera_solve_qa(None, 2, None, None, None, 0, 10, 0.1, qa._p(sp), None, None, None) == 0.0
    e, *_ = qa.solve_qa(np.zeros(30), None, None, None, 100, 0.1)
    assert math.isnan(e)


def test_sv_entry_rejects_too_large_and_bad_steps():
    L = qa.lib()
    h = np.zeros(40)
    sp = np.zeros(40, np.int8)
    a = np.ones(2)
    assert L.tessera_sv_anneal(qa._p(h), 40, None, None, None, 0, qa._p(a), qa._p(a), 2, 0.1, qa._p(sp), None, None,
                               None, None, None) == -2
    assert L.tessera_sv_anneal(qa._p(h), 3, None, None, None, 0, qa._p(a), qa._p(a), 0, 0.1, qa._p(sp), None, None,
                               None, None, None) == -1


def test_quantum_anneal_rejects_degenerate_schedule():
    # rust/src/ffi.rs:102-104
    with pytest.raises(qa.QaFailure) as ei:
        qa.quantum_anneal(I.make_ising([1.0], []), I.Schedule(np.array([1.0]), np.array([0.0])), 32)
    assert ei.value.kind is qa.QaError.BadSchedule


def test_session_entries_reject_null_sessions():
    """Session, executor, batch and diagnostic entries validate their arguments
    before touching a device."""
    L = qa.lib()
    assert L.tessera_tebd_set_executor(None, 0) == -1
    assert L.tessera_tebd_exec_stats(None, None, None) == -1
    assert L.tessera_tebd_exec_profile(None, None, 0) == -1
    r = np.zeros(1, np.int32)
    assert L.tessera_tebd_theta(None, 0, None, 0, qa._p(r), qa._p(r)) == -1
    assert L.tessera_tebd_batch_run(None, 1, None, None, None, -1) == -1
    handles = (ctypes.c_void_p * 1)(None)
    steps = np.zeros(1, np.uintp)
    ptrs = (ctypes.c_void_p * 1)(None)
    rc = L.tessera_tebd_batch_run(ctypes.addressof(handles), 1, ctypes.addressof(ptrs), ctypes.addressof(ptrs),
                                  qa._p(steps), -1)
    assert rc == -1 and b"null session" in L.tessera_last_error()
    assert L.tessera_tebd_batch_run(ctypes.addressof(handles), 0, None, None, None, -2) == -1
    assert L.tessera_tebd_batch_run(ctypes.addressof(handles), 0, None, None, None, -1) == 0  # nothing to do
    assert qa.lib().tessera_tebd_updates_per_step(None) == 0


def test_guidance_and_hook_entries_reject_bad_arguments():
    """The learned-guidance, warm-start and executor-shape entries validate their
    arguments before touching a device."""
    from tessera_amd import learned

    L = learned._bind()
    assert L.tessera_gnn_create(1, None) == -1
    assert L.tessera_gnn_param_count() == 50050
    out = np.zeros(4, np.float32)
    assert L.tessera_gnn_get_params(None, qa._p(out)) == -1
    assert L.tessera_gnn_set_params(None, qa._p(out)) == -1
    assert L.tessera_gnn_predict(None, 0, 1, None, None) == -1
    assert L.tessera_gnn_grad(None, 0, 1, None, None) == -1
    assert L.tessera_gnn_train_step(None, 0, 1, ctypes.c_float(1e-3), None) == -1
    off = np.zeros(2, np.int32)
    assert L.tessera_gnn_set_data(None, 1, qa._p(off), qa._p(out), qa._p(off), None, None, None, None) == -1
    L.tessera_gnn_destroy(None)
    Lq = qa.lib()
    z = np.zeros(4)
    assert Lq.tessera_tebd_expect_z(None, qa._p(z)) == -1
    assert Lq.tessera_tebd_entropy_now(None, qa._p(z)) == -1
    assert Lq.tessera_tebd_set_bias(None, None) == -1
    assert Lq.tessera_tebd_reset_warm(None, None) == -1
    assert Lq.tessera_tebd_set_executor_shape(None, -1, 0) == -1
    handles = (ctypes.c_void_p * 1)(None)
    assert Lq.tessera_tebd_batch_run_ex(ctypes.addressof(handles), 0, None, None, None, -1, 3) == -1
    h = np.zeros(2)
    spins = np.zeros(2, np.int8)
    st = np.zeros(8)
    assert Lq.tessera_sv_anneal_state(qa._p(h), 2, None, None, None, 0, None, None, 0, 0.1, qa._p(spins), None, None,
                                      qa._p(st)) == -1  # steps < 1


def test_static_scratch_entry_rejects_sizes_beyond_its_limits():
    """tessera_sv_anneal_static checks n, steps and m against its static scratch
    before any device work."""
    L = qa.lib()
    sp = np.zeros(17, np.int8)
    for n, steps, m in ((17, 4, 0), (4, 16385, 0), (4, 4, 4097)):
        h = np.zeros(n)
        ei = np.zeros(m, np.uintp)
        ej = np.ones(m, np.uintp)
        ew = np.ones(m)
        a = np.ones(steps)
        b = np.ones(steps)
        rc = L.tessera_sv_anneal_static(*_args(h, ei, ej, ew, a, b), 0.1, qa._p(sp), None, None, None, None, None)
        assert rc == -2, (n, steps, m)
        assert b"static-scratch" in L.tessera_last_error()

