"""Build step for the HIP library (gfx950) and the Fortran host shim.

``python -m flash5_amd.build`` compiles ``csrc/fx_hydro.hip`` into
``flash5_amd/libfx_hydro.so`` in-tree with hipcc.  All HIP code is built
with packed-FP32 instructions disabled.  FP32 division and square root use
the hardware reciprocal and root (about 2.5 ulp) instead of correctly rounded
sequences, and FP32 subnormals are flushed to zero, which drops the
frexp/ldexp scaling around each FP32 reciprocal and root: only the
mixed-precision reconstruction and the seed network compute in FP32, and
their error bars are set against the FP64 path.  FP64 code is not affected
by either flag (its machine code is the same with or without them).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfx_hydro.so")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_lib(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "fx_hydro.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, src, "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "(ignore" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    print(build_lib(verbose=True))


if __name__ == "__main__":
    main()
