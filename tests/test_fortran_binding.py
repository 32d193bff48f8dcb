"""The Fortran side of the boundary: include/fx_hydro_c.F90 compiles with
amdflang, its bind(C) types have the C structs' sizes and field offsets (the
C side from a gcc probe of include/fx_hydro.h), and a Fortran program linked
against libfx_hydro.so calls through the module (argument checks of
fx_hydro_init that return before any GPU call, and fx_hydro_version).
No GPU is used."""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "flash5_amd")
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

STRUCTS = {
    "fx_params": ["ndim", "nxb", "nyb", "nzb", "ng", "nunk", "var", "gamma", "limiter",
                  "limitedSlopeBeta", "charLimiting", "flattening", "steepening",
                  "use3dFullCTU", "useAuxEint", "riemann", "fP", "smlrho", "smallp", "smallE",
                  "eintSwitch", "tiny", "gasConstant", "singleSpeciesA", "exactMaxIter",
                  "exactTol", "precision", "learnedSeed", "shockDetect", "shockLowerCFL",
                  "shokVar", "fallbackLowerCFL", "cflFallbackFactor"],
    "fx_tile": ["U", "flux", "pavg", "del", "level", "block_class"],
    "fx_amr_block": ["level", "pos"],
    "fx_status": ["code", "first_tile", "first_i", "first_j", "first_k", "first_dir",
                  "fallbacks", "newton_iters", "newton_fallbacks", "min_cell_cfl",
                  "fallback_cfl"],
    "fx_fluxreg_face": ["coarse_tile", "fine_tile", "dir", "ci", "cj", "ck", "fi", "fj", "fk"],
}

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="amdflang not installed")


def _c_layout(d):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "fx_hydro.h"',
             "int main(void){"]
    for s, fields in STRUCTS.items():
        lines.append(f'printf("{s} size %zu\\n", sizeof({s}));')
        for f in fields:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    src = os.path.join(d, "probe.c")
    exe = os.path.join(d, "probe_c")
    open(src, "w").write("\n".join(lines))
    subprocess.run(["gcc", "-I", INC, src, "-o", exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return {tuple(l.split()[:2]): int(l.split()[2]) for l in out.strip().splitlines()}


def _fortran_program():
    decl, body = [], []
    for s, fields in STRUCTS.items():
        v = "x_" + s
        decl.append(f"  type({s}), target :: {v}")
        body.append(f'  write(*, "(A,1X,A,1X,I0)") "{s}", "size", c_sizeof({v})')
        for f in fields:
            first = "(1)" if f in ("var", "flux", "del", "pos") else ""
            body.append(f'  write(*, "(A,1X,A,1X,I0)") "{s}", "{f}", '
                        f"transfer(c_loc({v}%{f}{first}), 0_c_intptr_t) - "
                        f"transfer(c_loc({v}), 0_c_intptr_t)")
    return "\n".join([
        "program probe",
        "  use iso_c_binding",
        "  use fx_hydro_c",
        "  implicit none",
        *decl,
        "  type(fx_params) :: p",
        "  character(kind=c_char), pointer :: ver(:)",
        "  integer :: i",
        *body,
        # argument checks that return before any GPU call
        "  p = fx_params(3, 16, 16, 16, 4, 10, [(i, i = 0, 9)], 5d0/3d0, FX_MC, 1d0, 1, 1, 0, 1,"
        " 1, FX_HLLC, 0d0, 1d-10, 1d-10, 1d-10, 1d-4, 1d-10, 8.3144598d7, 1d0, 40, 1d-12,"
        " FX_PREC_FP64, 0, 0, 0, -1, 0, 0.9d0)",
        "  p%ng = 3",
        '  write(*, "(A,1X,I0)") "init_ng3", fx_hydro_init(p)',
        "  p%ng = 4",
        "  p%gamma = 1d0",
        '  write(*, "(A,1X,I0)") "init_gamma1", fx_hydro_init(p)',
        "  p%gamma = 5d0/3d0",
        "  p%precision = 1",
        '  write(*, "(A,1X,I0)") "init_prec1", fx_hydro_init(p)',
        "  p%precision = FX_PREC_FP64",
        "  p%fallbackLowerCFL = 1",
        "  p%cflFallbackFactor = 0d0",
        '  write(*, "(A,1X,I0)") "init_fbfactor0", fx_hydro_init(p)',
        "  call c_f_pointer(fx_hydro_version(), ver, [64])",
        "  i = 1",
        "  do while (ver(i) /= c_null_char .and. i < 64)",
        "    i = i + 1",
        "  end do",
        '  write(*, "(A,1X,64A)") "version", ver(1:i-1)',
        "end program probe",
    ])


def test_fortran_module_matches_the_c_abi():
    lib = os.path.join(LIBDIR, "libfx_hydro.so")
    if not os.path.exists(lib):
        pytest.skip("libfx_hydro.so not built")
    with tempfile.TemporaryDirectory() as d:
        c = _c_layout(d)
        prog = os.path.join(d, "probe.f90")
        open(prog, "w").write(_fortran_program())
        exe = os.path.join(d, "probe_f")
        r = subprocess.run([FLANG, "-J", d, os.path.join(INC, "fx_hydro_c.F90"), prog, "-o", exe,
                            "-L", LIBDIR, "-lfx_hydro", f"-Wl,-rpath,{LIBDIR}",
                            "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True, cwd=d)
        assert r.returncode == 0, r.stderr
        out = subprocess.run([exe], capture_output=True, text=True, check=True, cwd=d).stdout
    f = {}
    for line in out.strip().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] in STRUCTS:
            f[(parts[0], parts[1])] = int(parts[2])
        else:
            f[parts[0]] = " ".join(parts[1:])
    for key, off in c.items():
        assert f[key] == off, f"{key}: Fortran {f[key]} vs C {off}"
    assert f["init_ng3"] == "3" and f["init_gamma1"] == "3" and f["init_prec1"] == "3"
    assert f["init_fbfactor0"] == "3"
    assert "gfx950" in f["version"]
