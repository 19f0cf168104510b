"""The plain entry point called from compiled code linked like the Rust crate.

tests/c_abi/qa_mps_bench (built by __graft_entry__.build with gcc, -ltessera_qa)
declares tessera_qa_anneal as rust/src/ffi.rs:20-35 does and walks the QA(mps)
column of rust/src/bin/bench.rs:41-56 (smooth schedule, chi = 32, n = 6, 8, 10
here with 150 steps instead of 800 and 12: the swap chains of dense random graphs
make the full column a few minutes of per-update launches). Each
result must equal the ctypes path on the same instance bit for bit, the returned
energy must be the Ising energy of the returned spins (ffi.cpp:36), and no
energy may undercut the brute-force ground energy.
"""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tessera_amd import instances as I
from tessera_amd import qa

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
HARNESS = ROOT / "tests" / "c_abi" / "qa_mps_bench"


def test_compiled_caller_matches_ctypes_path():
    if not HARNESS.exists():
        pytest.fail("tests/c_abi/qa_mps_bench is not built (run __graft_entry__.build())")
    steps = 150
    proc = subprocess.Popen([str(HARNESS), str(steps), "10"], stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                            text=True)
    rows = []
    for line in proc.stdout:  # one line per finished instance
        print(line[:160], flush=True)
        if line.startswith("{"):
            rows.append(json.loads(line))
    assert proc.wait(timeout=240) == 0, proc.stderr.read()
    assert [r["n"] for r in rows] == [6, 8, 10]
    hits = 0
    for r in rows:
        model = I.make_ising(r["h"], [tuple(e) for e in r["edges"]])
        py = qa.quantum_anneal(model, I.smooth(steps), 32)
        print(f"n={r['n']}: energy {r['energy']:.6f} exact {r['exact']:.6f} (ctypes {py.energy:.6f})", flush=True)
        assert py.energy == r["energy"]
        assert np.array_equal(py.spins, np.array(r["spins"], np.int8))
        assert r["energy"] >= r["exact"] - 1e-9
        hits += abs(r["energy"] - r["exact"]) < 1e-9
    assert hits >= 2
