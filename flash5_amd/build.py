"""Build step for the HIP library (gfx950) and the Fortran host shim.

``python -m flash5_amd.build`` compiles ``csrc/fx_hydro.hip`` and the
column-march units (``csrc/fx_march_inst.hip``, once per solver family and
real type) in parallel and links them into ``flash5_amd/libfx_hydro.so``
in-tree with hipcc.  All HIP code is built
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
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-hip-fp32-correctly-rounded-divide-sqrt",
    "-fgpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]

# The column-march kernels of each (solver family, real type) pair form a
# translation unit of their own (fx_march_inst.hip), so the units build in
# parallel; the main unit holds every other kernel and the C ABI.
MARCH_UNITS = [(rs, r) for rs in ("FX_HLLC", "FX_HYBR", "RS_STATES")
               for r in ("double", "float")]
OBJ_DIR = os.path.join(HERE, "_build")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _filter(stderr: str) -> str:
    return "\n".join(l for l in stderr.splitlines()
                     if "packed-fp32-ops' is not a recognized feature" not in l
                     and "(ignore" not in l)


def _units():
    main = os.path.join(CSRC, "fx_hydro.hip")
    inst = os.path.join(CSRC, "fx_march_inst.hip")
    out = [(main, [], os.path.join(OBJ_DIR, "fx_hydro.o"))]
    for rs, r in MARCH_UNITS:
        out.append((inst, [f"-DFX_MARCH_RS={rs}", f"-DFX_MARCH_R={r}"],
                    os.path.join(OBJ_DIR, f"fx_march_{rs.lower()}_{r}.o")))
    return out


def build_lib(verbose: bool = False, jobs: int = 0) -> str:
    """Compile every unit (in parallel, ``jobs`` at a time; 0 = the CPU
    count, at most 16) and link libfx_hydro.so in-tree."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    jobs = jobs or min(16, os.cpu_count() or 1)
    pending = list(_units())
    running = []
    errors = []
    while pending or running:
        while pending and len(running) < jobs:
            src, defs, obj = pending.pop(0)
            cmd = [hipcc(), *HIP_FLAGS, *defs, "-c", src, "-o", obj]
            running.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                             text=True), obj))
        proc, obj = running.pop(0)
        _, err = proc.communicate()
        err = _filter(err)
        if proc.returncode != 0:
            errors.append(f"{os.path.basename(obj)}:\n{err}")
        elif verbose and err.strip():
            print(err, file=sys.stderr)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    tmp = LIB + ".tmp"
    objs = [obj for _, _, obj in _units()]
    res = subprocess.run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{_filter(res.stderr)}")
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    print(build_lib(verbose=True))


if __name__ == "__main__":
    main()
