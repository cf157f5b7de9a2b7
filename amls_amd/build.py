"""Build libamls_hip.so in-tree with hipcc for gfx950.

Usage: python -m amls_amd.build   (or amls_amd.build.build())
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libamls_hip.so"
ARCH = "gfx950"

SOURCES = [
    CSRC / "kernels" / "gemm.hip",
    CSRC / "kernels" / "factor.hip",
    CSRC / "kernels" / "eig.hip",
    CSRC / "kernels" / "sparse.hip",
    CSRC / "kernels" / "adjoint.hip",
    CSRC / "kernels" / "lanczos.hip",
    CSRC / "kernels" / "chfsi.hip",
    CSRC / "kernels" / "store.hip",
    CSRC / "host" / "tree.cpp",
    CSRC / "sil.cpp",
    CSRC / "noalloc.cpp",
    CSRC / "operators.cpp",
    CSRC / "chfsi.cpp",
    CSRC / "trace.cpp",
    CSRC / "store.cpp",
    CSRC / "engine.cpp",
]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # hardware FP64 atomics on coarse-grained (hipMalloc) memory
    "-munsafe-fp-atomics",
    # no packed FP32 instructions from the compiler
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in ordinary VGPRs (gfx950 reads and writes
    # them there): otherwise the GEMM loop copies every accumulator between
    # the two register files on each k-chunk
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, jobs: int = 6) -> Path:
    """Compile every source to an object (in parallel) and link the library."""
    objdir = HERE / "_build"
    objdir.mkdir(exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = objdir / (src.stem + ("_hip" if src.suffix == ".hip" else "_cpp") + ".o")
        objs.append(obj)
        cmd = [cc, *HIP_FLAGS, "-x", "hip", "-c", str(src), "-o", str(obj),
               "-I", str(CSRC), "-I", str(HERE.parent / "include")]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"---- {s}\n{o}" for s, o in failed)
        raise RuntimeError("HIP build failed:\n" + msg)
    tmp = LIB.with_suffix(".so.tmp")
    link = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
