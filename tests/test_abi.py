"""C ABI checks that need no GPU: the HIP library loads, exports every entry
point declared in include/fx_hydro.h, and the ctypes mirrors of the ABI
structs match the C layout (sizes and field offsets from a gcc probe)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from flash5_amd import hydro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fx_hydro.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_]+\**\s+\**(fx_[a-z0-9_]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "fx_hydro_compute_fluxes" in names and "fx_hydro_update_solution" in names
    assert len(names) >= 10


def test_library_loads_and_exports_every_declared_symbol():
    L = hydro.load_library()
    for name in declared_functions():
        assert hasattr(L, name), f"{name} declared in fx_hydro.h but not exported"
    assert set(declared_functions()) <= set(hydro.EXPORTS)
    assert b"gfx950" in L.fx_hydro_version()


def _c_layout():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "fx_hydro.h"
#define O(T,f) printf(#T "." #f " %zu\n", offsetof(T, f))
int main(void){
  printf("fx_params %zu\nfx_tile %zu\nfx_status %zu\n", sizeof(fx_params), sizeof(fx_tile), sizeof(fx_status));
  O(fx_params,gamma); O(fx_params,limiter); O(fx_params,fP); O(fx_params,exactTol);
  O(fx_params,precision); O(fx_params,learnedSeed); O(fx_params,shokVar);
  O(fx_params,fallbackLowerCFL); O(fx_params,cflFallbackFactor);
  O(fx_tile,pavg); O(fx_tile,del); O(fx_tile,level);
  O(fx_status,fallbacks); O(fx_status,min_cell_cfl); O(fx_status,fallback_cfl);
  return 0;}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return dict(line.rsplit(" ", 1) for line in out.strip().splitlines())


def test_ctypes_structs_match_c_layout():
    lay = {k: int(v) for k, v in _c_layout().items()}
    assert C.sizeof(hydro.FxParams) == lay["fx_params"]
    assert C.sizeof(hydro.FxTile) == lay["fx_tile"]
    assert C.sizeof(hydro.FxStatus) == lay["fx_status"]
    for f in ("gamma", "limiter", "fP", "exactTol", "precision", "learnedSeed", "shokVar",
              "fallbackLowerCFL", "cflFallbackFactor"):
        assert getattr(hydro.FxParams, f).offset == lay[f"fx_params.{f}"], f
    assert hydro.FxTile.pavg.offset == lay["fx_tile.pavg"]
    assert hydro.FxTile.del_.offset == lay["fx_tile.del"]
    assert hydro.FxTile.level.offset == lay["fx_tile.level"]
    assert hydro.FxStatus.fallbacks.offset == lay["fx_status.fallbacks"]
    assert hydro.FxStatus.min_cell_cfl.offset == lay["fx_status.min_cell_cfl"]
    assert hydro.FxStatus.fallback_cfl.offset == lay["fx_status.fallback_cfl"]


def test_init_rejects_bad_parameters():
    L = hydro.load_library()
    p = hydro.HydroParams(ndim=3, ng=2).to_c()
    assert L.fx_hydro_init(C.byref(p)) != 0          # NGUARD must be 4
    p = hydro.HydroParams(ndim=3, gamma=1.0).to_c()
    assert L.fx_hydro_init(C.byref(p)) != 0          # gamma > 1
    p = hydro.HydroParams(ndim=3, hy_fallbackLowerCFL=True, hy_cflFallbackFactor=0.0).to_c()
    assert L.fx_hydro_init(C.byref(p)) != 0          # a positive fallback factor


def test_hydro_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(RuntimeError):
        hydro.Hydro(hydro.HydroParams())
