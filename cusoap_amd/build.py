"""Build libcusoap_amd.so for gfx950 with hipcc (in-tree, next to this file).

python -m cusoap_amd.build [--jobs N]
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libcusoap_amd.so")
SOURCES = ["capi.hip", "neighbor.hip", "coeffs.hip", "spectrum.hip", "jacobian.hip", "adjoint.hip", "surrogate.hip",
           "chebyshev.hip", "order.hip"]
# compiled once per l_max value (-DCUSOAP_ONLY_L=l, that degree's kernels) plus once for the rest
PER_DEGREE = ("coeffs.hip", "adjoint.hip", "chebyshev.hip")
MAX_L = 9
ARCH = os.environ.get("CUSOAP_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: never emit packed FP32 VALU ops (v_pk_fma_f32 & co).
HIPFLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build cusoap_amd)")


# per-source flags: the Chebyshev coefficient kernel keeps its matrix-core accumulators in
# ordinary VGPRs (no accumulator-register copies around the hi/lo split of the epilogue;
# same register total and occupancy)
SOURCE_FLAGS = {"chebyshev.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def units():
    """(source, extra flags, object name) of every compilation."""
    out = []
    for src in SOURCES:
        extra = SOURCE_FLAGS.get(src, [])
        out.append((src, list(extra), src.replace(".hip", ".o")))
        if src in PER_DEGREE:
            out += [(src, [*extra, f"-DCUSOAP_ONLY_L={l}"], src.replace(".hip", f"_l{l}.o")) for l in range(MAX_L + 1)]
    return out


def _compile(unit) -> str:
    src, flags, name = unit
    obj = os.path.join(OBJ, name)
    cmd = [hipcc(), *HIPFLAGS, *flags, "-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {name}:\n{res.stderr[-6000:]}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    todo = units()
    jobs = jobs or min(len(todo), os.cpu_count() or 4, 16)
    # the heaviest compilations first, so the pool ends together
    def degree(u):
        return max([int(f.split("=")[1]) for f in u[1] if f.startswith("-DCUSOAP_ONLY_L=")] or [0])
    todo.sort(key=lambda u: (u[0] not in PER_DEGREE, -degree(u)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, todo))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    a = ap.parse_args()
    build(a.jobs, verbose=True)
    sys.exit(0)
