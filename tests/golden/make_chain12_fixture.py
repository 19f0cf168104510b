#!/usr/bin/env python3
"""Writes tests/golden/chain12_tebd.json: the full n = 12 chain anneal (linear
schedule, 200 steps, dt = 0.05, chi = 64, exact) run by the oracle's accurate
SVD mode (svd_mode 1, the upstream update of cuda/src/mps.cpp:41-100 with an
accurate SVD), plus the exact state vector of the same instance (oracle
restatement of rust/src/quantum.rs:63-112).

Stored: per-step qubit-0 entropy and largest bond, the count of updates per
step whose kept rank is set by a singular value at the rounding level of theta
(between 1e-15 and 1e-13), the readout (spins, energy,
<H>, prob, max discarded weight), the per-step norm of the MPS, the final-state
fidelity between the MPS and the state vector (different Trotter splittings,
so O(dt^2) apart), and the MPS-vs-dense infidelity of the oracle itself. The
state vector's own traces are stored too, for the MPS-vs-SV error statement.
Run in the build container (about a minute):

    python tests/golden/make_chain12_fixture.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from tessera_amd import instances as I  # noqa: E402


def main():
    import pyoracle

    inst = I.chain12_anneal()
    sched = I.linear(200)
    o = pyoracle.tebd(inst.model, sched, inst.dt, inst.chi, svd_mode=1, want_norms=True, want_state=True)
    s = pyoracle.sv(inst.model, sched, inst.dt, want_state=True)
    # MPS index: bit i of x = physical index of site i (0 <-> +1); SV index: bit i set <-> +1
    n = inst.model.n
    flip = (1 << n) - 1
    psi_mps = o["state"]
    psi_sv = s["state"][np.arange(1 << n) ^ flip]
    fid = abs(np.vdot(psi_mps, psi_sv)) ** 2 / (np.vdot(psi_mps, psi_mps).real * np.vdot(psi_sv, psi_sv).real)
    d = {
        "name": "chain12_linear200_chi64",
        "model": {"h": [float(x) for x in inst.model.h], "ei": [int(x) for x in inst.model.ei],
                  "ej": [int(x) for x in inst.model.ej], "ew": [float(x) for x in inst.model.ew]},
        "sched": ["linear", 200], "dt": inst.dt, "chi": inst.chi, "zz_substeps": 1,
        "spins": [int(x) for x in o["spins"]], "energy": float(o["energy"]),
        "expected_energy": float(o["expected_energy"]), "prob": float(o["prob"]),
        "max_trunc": float(o["max_trunc"]), "entropy": [float(x) for x in o["entropy"]],
        "max_bond": [int(x) for x in o["max_bond"]], "norm2": [float(x) for x in o["norm2"]],
        "ambiguous": [int(x) for x in o["ambiguous"]],
        "sv": {"spins": [int(x) for x in s["spins"]], "energy": float(s["energy"]),
               "expected_energy": float(s["expected_energy"]), "success_prob": float(s["success_prob"]),
               "entropy": [float(x) for x in s["entropy"]]},
        "mps_sv_fidelity": float(fid),
    }
    (HERE / "chain12_tebd.json").write_text(json.dumps(d))
    print(f"wrote chain12_tebd.json: <H>={d['expected_energy']:.12f} fidelity to SV={fid:.12f} "
          f"max bond={max(d['max_bond'])} max|norm2-1|={max(abs(x - 1) for x in d['norm2']):.2e}")


if __name__ == "__main__":
    main()
