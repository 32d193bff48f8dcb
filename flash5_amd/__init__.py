"""flash5_amd: MI355X-native unsplit hydro block update (FLASH5 drop-in).

The product is the HIP library ``libfx_hydro.so`` (C ABI in
``include/fx_hydro.h``).  This package holds its Python host mirror
(:mod:`flash5_amd.hydro`), the build step (:mod:`flash5_amd.build`) and the
synthetic workload generators (:mod:`flash5_amd.workloads`).
"""
from .hydro import HydroParams, Hydro, Tile, FxError  # noqa: F401

__all__ = ["HydroParams", "Hydro", "Tile", "FxError"]
