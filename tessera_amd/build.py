"""Build libtessera_qa.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m tessera_amd.build          # or tessera_amd.build.build()

Output: tessera_amd/_lib/libtessera_qa.so. The oracle (test infrastructure) is
built separately by `make -C oracle`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIBDIR = HERE / "_lib"
LIB = LIBDIR / "libtessera_qa.so"

SOURCES = ["sv.hip", "tebd.hip", "gnn.hip", "abi.cpp"]
ARCH = os.environ.get("TESSERA_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions stay off (only scalar/plain vector FP32 ops are emitted)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Per-source flags. tebd.hip contracts a*b+c into an FMA only inside one source
# expression (-ffp-contract=on), never across statements after inlining: its SVD
# code is instantiated once per executor layout, and with cross-statement
# contraction each instantiation could round differently, while the layouts are
# meant to apply bit-identical rotations.
SOURCE_FLAGS = {"tebd.hip": ["-ffp-contract=on"]}
# TESSERA_STEP_PROFILE=1 in the environment builds the executor's cycle profile of
# the Jacobi rotation steps (it costs the steps registers, so it is off by default)
if os.environ.get("TESSERA_STEP_PROFILE") == "1":
    SOURCE_FLAGS["tebd.hip"] = SOURCE_FLAGS["tebd.hip"] + ["-DTESSERA_STEP_PROFILE=1"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    LIBDIR.mkdir(exist_ok=True)
    objdir = LIBDIR / "obj"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    objs = []
    procs = []
    for src in SOURCES:
        obj = objdir / (src + ".o")
        cmd = [cc, *FLAGS, *SOURCE_FLAGS.get(src, []), "-I", str(CSRC), "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        # hipcc's host pass warns that it does not know the packed-fp32 feature; harmless
        lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append((src, "\n".join(lines)))
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(f"[{s}]\n{o}" for s, o in failed))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, *FLAGS, "-shared", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
