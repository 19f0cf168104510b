#!/usr/bin/env python3
"""Writes tests/golden/*.json from the upstream TESSERA C++ engine.

The engine is compiled from its own sources under /root/reference by
`make -C oracle ref` into oracle/_ref/ and driven through its public API by
oracle/ref_harness.cpp. Only inputs and outputs are stored here (vectors, not
source). Run in the build container:

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from tessera_amd import instances as I  # noqa: E402

HARNESS = ROOT / "oracle" / "_ref" / "ref_harness"


def inst_text(model, sched, dt, chi, ns=1):
    lines = [f"{model.n} {model.m} {sched.steps} {chi} {ns} {dt!r}",
             " ".join(repr(float(x)) for x in model.h)]
    for i, j, w in zip(model.ei, model.ej, model.ew):
        lines.append(f"{int(i)} {int(j)} {float(w)!r}")
    lines.append(" ".join(repr(float(x)) for x in sched.a))
    lines.append(" ".join(repr(float(x)) for x in sched.b))
    return "\n".join(lines) + "\n"


def run(mode, text=""):
    out = subprocess.run([str(HARNESS), *mode.split()], input=text, capture_output=True, text=True, check=True).stdout
    res = {}
    for line in out.splitlines():
        key, *vals = line.split()
        res[key] = [float(v) for v in vals]
    return res


def model_json(model):
    return {"h": [float(x) for x in model.h], "ei": [int(x) for x in model.ei], "ej": [int(x) for x in model.ej],
            "ew": [float(x) for x in model.ew]}


def case(name, model, sched_spec, dt, chi, ns=1, trace=True):
    sched = {"linear": I.linear, "smooth": I.smooth}[sched_spec[0]](sched_spec[1])
    text = inst_text(model, sched, dt, chi, ns)
    r = run("anneal", text)
    d = {"name": name, "model": model_json(model), "sched": list(sched_spec), "dt": dt, "chi": chi,
         "zz_substeps": ns,
         "spins": [int(x) for x in r["spins"]], "energy": r["energy"][0], "expected_energy": r["expected_energy"][0],
         "prob": r["prob"][0], "max_trunc": r["max_trunc"][0], "entropy": r["entropy"]}
    if trace:
        t = run("trace", text)
        d["max_bond"] = [int(x) for x in t["max_bond"]]
        d["norm2"] = t["norm2"]
    return d


def main():
    if not HARNESS.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"], check=True)
    import pyoracle

    cases = []
    cases.append(case("single_spin_smooth300", I.make_ising([1.0], []), ("smooth", 300), 0.1, 16))
    cases.append(case("ferro_pair_smooth400", I.make_ising([0.0, 0.0], [(0, 1, 1.0)]), ("smooth", 400), 0.1, 16))
    cases.append(case("afm_pair_entropy_smooth400", I.make_ising([0.3, -0.3], [(0, 1, -1.0)]), ("smooth", 400), 0.1,
                      16))
    cases.append(case("frustrated_ring_T20", I.make_ising([0.5, -0.3, 0.2], [(0, 1, -1.0), (1, 2, -1.0),
                                                                             (2, 0, -1.0)]), ("smooth", 200), 0.1, 4))
    cases.append(case("chain_nnn_n6_chi8", I.chain_nnn(6, 1), ("linear", 30), 0.1, 8))
    cases.append(case("chain_nnn_n8_chi4_trunc", I.chain_nnn(8, 2), ("linear", 20), 0.1, 4))
    cases.append(case("range3_n7_chi4_zz2", I.chain_range(7, 3, 5), ("linear", 12), 0.08, 4, ns=2))
    cases.append(case("chain12_prefix8", I.chain12_anneal().model, ("linear", 8), 0.05, 64))
    for t in range(6):  # cuda/tests/test_mps.cpp:96-123 instances
        model = pyoracle.upstream_instance(123, t, 6)
        cases.append(case(f"upstream_seed123_{t}", model, ("smooth", 1200), 0.08, 8, trace=False))
    (HERE / "tebd_ref.json").write_text(json.dumps(cases))

    lin = run("linalg")
    (HERE / "linalg_ref.json").write_text(json.dumps(lin))
    print(f"wrote {len(cases)} TEBD cases and {len(lin)} linalg arrays")


if __name__ == "__main__":
    main()
