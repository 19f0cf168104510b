"""tessera_amd: MI355X (gfx950) implementation of TESSERA's quantum-annealing hot
path (MPS-TEBD two-site update and exact state-vector Trotter sweep) behind the
upstream tessera_qa C ABI."""
from .instances import Ising, Schedule, Instance, Rng, linear, smooth, make_ising  # noqa: F401
from .qa import (  # noqa: F401
    AnnealResult,
    QaError,
    QaFailure,
    QaOutcome,
    QaResult,
    TebdSession,
    TesseraError,
    anneal_statevector,
    lib,
    library_path,
    mps_anneal,
    quantum_anneal,
    solve_qa,
)

__all__ = [
    "Ising", "Schedule", "Instance", "Rng", "linear", "smooth", "make_ising",
    "AnnealResult", "QaError", "QaFailure", "QaOutcome", "QaResult", "TebdSession", "TesseraError",
    "anneal_statevector", "lib", "library_path", "mps_anneal", "quantum_anneal", "solve_qa",
]
