"""Python host mirror of FLASH5's unsplit Hydro entry points over the C ABI.

Mirrors (paths relative to source/physics/Hydro):
  * ``Hydro_init``       HydroMain/unsplit/Hydro_init.F90:24  -> :class:`Hydro`
  * ``hy_computeFluxes`` HydroMain/unsplit/hy_computeFluxes.F90:76
                                                       -> :meth:`Hydro.compute_fluxes`
  * ``hy_updateSolution`` HydroMain/unsplit/hy_updateSolution.F90:104
                                                       -> :meth:`Hydro.update_solution`
  * ``Hydro_computeDt``  HydroMain/unsplit/Hydro_computeDt.F90:186
                                                       -> :meth:`Hydro.compute_dt`
Every call goes through ``libfx_hydro.so`` (``include/fx_hydro.h``); PyTorch
only provides device memory.  There is no CPU fallback: a missing library
or a missing GPU raises.

Block data follow FLASH's index-reordered layout U(i,j,k,var): a torch
tensor of shape (nunk, nz, ny, nx) per block (guard cells included), or
(nblocks, nunk, nz, ny, nx) for a batch.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import List, Optional, Sequence

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfx_hydro.so")

NVAR_USED = 10
NFLUX = 8
DENS, VELX, VELY, VELZ, PRES, ENER, GAMC, GAME, TEMP, EINT = range(10)
MINMOD, MC, HYBRID, VANLEER, LIMITED = 1, 2, 3, 4, 5
ROE, HLL, HLLC, HYBR, EXACT = 1, 3, 4, 8, 20
PREC_FP64, PREC_MIXED = 0, 2

_LIMITERS = {"minmod": MINMOD, "mc": MC, "hybrid": HYBRID, "vanleer": VANLEER,
             "limited": LIMITED}
_SOLVERS = {"roe": ROE, "hll": HLL, "hllc": HLLC, "hybrid": HYBR, "exact": EXACT}


class FxParams(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("nxb", C.c_int), ("nyb", C.c_int), ("nzb", C.c_int),
        ("ng", C.c_int), ("nunk", C.c_int), ("var", C.c_int * NVAR_USED),
        ("gamma", C.c_double), ("limiter", C.c_int), ("limitedSlopeBeta", C.c_double),
        ("charLimiting", C.c_int), ("flattening", C.c_int), ("steepening", C.c_int),
        ("use3dFullCTU", C.c_int), ("useAuxEint", C.c_int), ("riemann", C.c_int),
        ("fP", C.c_double), ("smlrho", C.c_double), ("smallp", C.c_double),
        ("smallE", C.c_double), ("eintSwitch", C.c_double), ("tiny", C.c_double),
        ("gasConstant", C.c_double), ("singleSpeciesA", C.c_double),
        ("exactMaxIter", C.c_int), ("exactTol", C.c_double),
        ("precision", C.c_int), ("learnedSeed", C.c_int),
        ("shockDetect", C.c_int), ("shockLowerCFL", C.c_int), ("shokVar", C.c_int),
        ("fallbackLowerCFL", C.c_int), ("cflFallbackFactor", C.c_double),
    ]


_DP = C.POINTER(C.c_double)


class FxTile(C.Structure):
    _fields_ = [("U", _DP), ("flux", _DP * 3), ("pavg", _DP), ("del_", C.c_double * 3),
                ("level", C.c_int), ("block_class", C.c_int)]


class FxAmrBlock(C.Structure):
    _fields_ = [("level", C.c_int), ("pos", C.c_int * 3)]


class FxTuneResult(C.Structure):
    _fields_ = [("tile_x", C.c_int), ("persist", C.c_int), ("order", C.c_int),
                ("occupancy", C.c_int), ("chosen", C.c_int), ("ms", C.c_float)]


class FxStatus(C.Structure):
    _fields_ = [("code", C.c_int), ("first_tile", C.c_int), ("first_i", C.c_int),
                ("first_j", C.c_int), ("first_k", C.c_int), ("first_dir", C.c_int),
                ("fallbacks", C.c_longlong), ("newton_iters", C.c_longlong),
                ("newton_fallbacks", C.c_longlong), ("min_cell_cfl", C.c_double),
                ("fallback_cfl", C.c_double)]


class FxFluxRegFace(C.Structure):
    """fx_fluxreg_face: one copy of a coarse face on a fine-coarse boundary."""
    _fields_ = [("coarse_tile", C.c_int), ("fine_tile", C.c_int), ("dir", C.c_int),
                ("ci", C.c_int), ("cj", C.c_int), ("ck", C.c_int),
                ("fi", C.c_int), ("fj", C.c_int), ("fk", C.c_int)]


# Exported symbols and their signatures (include/fx_hydro.h).
EXPORTS = {
    "fx_hydro_init": ([C.POINTER(FxParams)], C.c_int),
    "fx_hydro_finalize": ([], None),
    "fx_hydro_compute_fluxes": ([C.POINTER(FxTile), C.c_int, C.c_double, C.c_double,
                                 C.POINTER(C.c_ubyte), C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_update_solution": ([C.POINTER(FxTile), C.c_int, C.c_double, C.c_double,
                                  C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_advance_batch": ([_DP, C.c_long, C.c_long, C.c_double * 3, C.c_double,
                                C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_advance_batch_noalloc": ([_DP, C.c_long, C.c_long, C.c_double * 3, C.c_double,
                                        C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_shock_detect": ([C.POINTER(FxTile), C.c_int, C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_shock_detect_batch": ([_DP, C.c_long, C.c_long, C.c_double * 3,
                                     C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_compute_dt": ([_DP, C.c_long, C.c_long, C.c_double * 3, C.c_double, _DP], C.c_int),
    "fx_fluxreg_put": ([C.POINTER(FxTile), C.c_int, C.POINTER(FxFluxRegFace), C.c_int, _DP,
                        C.POINTER(FxStatus)], C.c_int),
    "fx_fluxreg_conserve": ([C.POINTER(FxTile), C.c_int, C.POINTER(FxFluxRegFace), C.c_int, _DP,
                             C.POINTER(FxStatus)], C.c_int),
    "fx_mlp_features": ([_DP, C.c_long, C.c_double, C.POINTER(C.c_float)], C.c_int),
    "fx_mlp_forward": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_long,
                        C.POINTER(C.c_float)], C.c_int),
    "fx_mlp_grad": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_long,
                     C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)], C.c_int),
    "fx_mlp_adam": ([C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                     C.POINTER(C.c_float), C.c_int, C.c_float], C.c_int),
    "fx_riemann_batch": ([_DP, C.c_long, C.c_double, C.POINTER(C.c_float), C.c_int, C.c_double,
                          _DP, C.POINTER(C.c_int), C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_set_seed_params": ([C.POINTER(C.c_float)], C.c_int),
    "fx_mlp_set_gate": ([C.c_double], C.c_int),
    "fx_hydro_face_states": ([_DP, C.c_long, C.POINTER(C.c_long)], C.c_int),
    "fx_grid_fill_guards": ([_DP, C.c_long, C.c_long, C.POINTER(C.c_int), C.POINTER(FxStatus)],
                            C.c_int),
    "fx_grid_run_graph": ([_DP, C.c_long, C.c_long, C.POINTER(C.c_int), C.c_double * 3, C.c_double,
                           C.c_int, C.c_int, _DP, C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_set_stream": ([C.c_void_p], None),
    "fx_hydro_set_async": ([C.c_int], None),
    "fx_hydro_set_timing": ([C.c_int], None),
    "fx_hydro_set_kernel_path": ([C.c_int], None),
    "fx_hydro_set_block_classes": ([C.POINTER(C.c_int), C.c_int], C.c_int),
    "fx_hydro_autotune": ([C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_int,
                           C.POINTER(C.c_int)], C.c_int),
    "fx_amr_autotune": ([C.c_double, C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "fx_hydro_autotune_reset": ([], None),
    "fx_amr_setup": ([C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_int),
                      C.POINTER(C.c_int), C.POINTER(FxStatus)], C.c_int),
    "fx_amr_fill_guards": ([C.POINTER(FxStatus)], C.c_int),
    "fx_amr_restrict": ([C.POINTER(FxStatus)], C.c_int),
    "fx_amr_compute_dt": ([C.c_double, C.POINTER(C.c_double)], C.c_int),
    "fx_amr_step": ([C.c_double, C.POINTER(FxStatus)], C.c_int),
    "fx_amr_run_graph": ([C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double),
                          C.POINTER(FxStatus)], C.c_int),
    "fx_hydro_stage_times": ([C.POINTER(C.c_float), C.c_int], C.c_int),
    "fx_hydro_version": ([], C.c_char_p),
}

_lib = None


class FxError(RuntimeError):
    """A nonzero return of the C ABI (the reference would Driver_abortFlash)."""

    def __init__(self, what: str, status: FxStatus | None = None, rc: int = 0):
        self.status = status
        self.rc = rc
        msg = f"{what} failed (rc={rc})"
        if status is not None and status.code:
            msg += (f": code {status.code} at tile {status.first_tile} cell "
                    f"({status.first_i},{status.first_j},{status.first_k}) dir {status.first_dir}")
        super().__init__(msg)


def load_library(path: str = LIB_PATH):
    """Load libfx_hydro.so and bind every exported symbol (raises if absent)."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} is missing: build it with `python -m flash5_amd.build` (hipcc, gfx950)")
    L = C.CDLL(path)
    for name, (args, res) in EXPORTS.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if path == LIB_PATH:
        _lib = L
    return L


@dataclasses.dataclass
class HydroParams:
    """Runtime knobs of unsplit Hydro (unsplit/Config defaults, order=3 PPM)."""
    ndim: int = 3
    nxb: int = 16
    nyb: int = 16
    nzb: int = 16
    ng: int = 4
    nunk: int = NVAR_USED
    var: Sequence[int] = tuple(range(NVAR_USED))
    gamma: float = 1.4
    slopeLimiter: str = "vanLeer"
    LimitedSlopeBeta: float = 1.0
    charLimiting: bool = True
    use_flattening: bool = True
    use_steepening: bool = False
    use_3dFullCTU: bool = True
    use_auxEintEqn: bool = True
    RiemannSolver: str = "HLLC"
    hy_fPresInMomFlux: float = 1.0
    smlrho: float = 1.e-10
    smallp: float = 1.e-10
    smallE: float = 1.e-10
    eintSwitch: float = 0.0
    tiny: float = 1.e-32
    gasConstant: float = 8.3144598e7
    eos_singleSpeciesA: float = 1.0
    exactMaxIter: int = 20
    exactTol: float = 1.e-12
    precision: int = PREC_FP64
    learnedSeed: bool = False
    shockDetect: bool = False
    shockLowerCFL: bool = False
    shokVar: int = -1            # slot of SHOK_VAR in U's var axis (-1: none)
    hy_fallbackLowerCFL: bool = False
    hy_cflFallbackFactor: float = 0.9

    def to_c(self) -> FxParams:
        p = FxParams()
        p.ndim, p.nxb = self.ndim, self.nxb
        p.nyb = self.nyb if self.ndim > 1 else 1
        p.nzb = self.nzb if self.ndim > 2 else 1
        p.ng, p.nunk = self.ng, self.nunk
        for i, s in enumerate(self.var):
            p.var[i] = s
        p.gamma = self.gamma
        p.limiter = _LIMITERS[self.slopeLimiter.lower()]
        p.limitedSlopeBeta = self.LimitedSlopeBeta
        p.charLimiting = int(self.charLimiting)
        p.flattening = int(self.use_flattening)
        p.steepening = int(self.use_steepening)
        p.use3dFullCTU = int(self.use_3dFullCTU)
        p.useAuxEint = int(self.use_auxEintEqn)
        p.riemann = _SOLVERS[self.RiemannSolver.lower()]
        p.fP = self.hy_fPresInMomFlux
        p.smlrho, p.smallp, p.smallE = self.smlrho, self.smallp, self.smallE
        p.eintSwitch, p.tiny = self.eintSwitch, self.tiny
        p.gasConstant, p.singleSpeciesA = self.gasConstant, self.eos_singleSpeciesA
        p.exactMaxIter, p.exactTol = self.exactMaxIter, self.exactTol
        p.precision, p.learnedSeed = self.precision, int(self.learnedSeed)
        p.shockDetect, p.shockLowerCFL = int(self.shockDetect), int(self.shockLowerCFL)
        p.shokVar = self.shokVar
        p.fallbackLowerCFL = int(self.hy_fallbackLowerCFL)
        p.cflFallbackFactor = self.hy_cflFallbackFactor
        return p

    def block_shape(self):
        nx = self.nxb + 2 * self.ng
        ny = self.nyb + 2 * self.ng if self.ndim > 1 else 1
        nz = self.nzb + 2 * self.ng if self.ndim > 2 else 1
        return (self.nunk, nz, ny, nx)

    def flux_shape(self, d: int):
        fx = self.nxb + (d == 0)
        fy = (self.nyb if self.ndim > 1 else 1) + (d == 1 and self.ndim > 1)
        fz = (self.nzb if self.ndim > 2 else 1) + (d == 2 and self.ndim > 2)
        return (NFLUX, fz, fy, fx)

    def interior_shape(self):
        return (self.nzb if self.ndim > 2 else 1, self.nyb if self.ndim > 1 else 1, self.nxb)


@dataclasses.dataclass
class Tile:
    """One tile of the reference's tile iterator: its U and optional fluxes."""
    U: "object"                       # torch tensor (nunk, nz, ny, nx), float64, on device
    deltas: Sequence[float]
    flux: Optional[List[object]] = None   # per-direction flux tensors or None
    pavg: Optional[object] = None         # interior-cell pressure average or None
    level: int = 1
    block_class: int = 0                  # index into Hydro.set_block_classes's shapes


def _ptr(t) -> _DP:
    if t is None:
        return _DP()
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if t.dtype != torch.float64 or not t.is_contiguous() or not t.is_cuda:
        raise ValueError("expected a contiguous float64 tensor on the GPU")
    return C.cast(C.c_void_p(t.data_ptr()), _DP)


class Hydro:
    """Hydro unit: Hydro_init on construction, then per-step entry points."""

    def __init__(self, params: HydroParams, lib_path: str = LIB_PATH):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("flash5_amd.Hydro needs a GPU (no CPU fallback)")
        self.params = params
        self.lib = load_library(lib_path)
        self._cp = params.to_c()
        torch.cuda.current_device()
        rc = self.lib.fx_hydro_init(C.byref(self._cp))
        if rc:
            raise FxError("fx_hydro_init", rc=rc)
        self.last_status = FxStatus()
        self._classes = None

    def _reinit(self):
        rc = self.lib.fx_hydro_init(C.byref(self._cp))
        if rc:
            raise FxError("fx_hydro_init", rc=rc)
        if self._classes is not None:
            self._apply_classes()

    def _apply_classes(self, classes=None):
        classes = self._classes if classes is None else classes
        flat = [int(v) for shape in classes for v in shape]
        arr = (C.c_int * max(1, len(flat)))(*flat)
        rc = self.lib.fx_hydro_set_block_classes(arr, len(classes))
        if rc:
            raise FxError("fx_hydro_set_block_classes", rc=rc)

    def set_block_classes(self, shapes) -> None:
        """Interior shapes (nxb, nyb, nzb) of the block classes; a Tile picks
        its shape by block_class.  Tiles of several classes go through one
        persistent launch per stage.  An empty list returns to the params
        shape.  A rejected list leaves the classes as they were."""
        classes = [tuple(int(v) for v in sh) + (1,) * (3 - len(sh)) for sh in shapes]
        self._apply_classes(classes)
        self._classes = classes or None

    def class_params(self, c: int) -> HydroParams:
        """HydroParams of block class c (the params with that class's shape)."""
        nx, ny, nz = self._classes[c]
        return dataclasses.replace(self.params, nxb=nx, nyb=ny, nzb=nz)

    # ---- autotuner ----
    @staticmethod
    def _tune_list(buf, n):
        return [{"tile_x": buf[k].tile_x, "persist": buf[k].persist, "order": buf[k].order,
                 "occupancy": buf[k].occupancy, "chosen": bool(buf[k].chosen),
                 "ms": float(buf[k].ms)} for k in range(n)]

    def autotune(self, tiles: Sequence[Tile], dt: float, reps: int = 3):
        """Times the flux stages' candidate launch shapes on these tiles and
        keeps the fastest for later calls; returns every candidate's time."""
        self._reinit()
        arr = self._tiles(tiles)
        buf = (FxTuneResult * 32)()
        n = C.c_int(0)
        rc = self.lib.fx_hydro_autotune(C.cast(arr, C.c_void_p), len(tiles), dt, reps,
                                        C.cast(buf, C.c_void_p), 32, C.byref(n))
        if rc:
            raise FxError("fx_hydro_autotune", rc=rc)
        return self._tune_list(buf, n.value)

    def autotune_batch(self, U, dt: float, deltas: Sequence[float], reps: int = 3):
        """autotune on a batch tensor (nBlocks, nunk, nz, ny, nx) of blocks."""
        tiles = [Tile(U=U[b], deltas=deltas) for b in range(U.shape[0])]
        return self.autotune(tiles, dt, reps)

    def amr_autotune(self, dt: float, reps: int = 3):
        buf = (FxTuneResult * 32)()
        n = C.c_int(0)
        rc = self.lib.fx_amr_autotune(dt, reps, C.cast(buf, C.c_void_p), 32, C.byref(n))
        if rc:
            raise FxError("fx_amr_autotune", rc=rc)
        return self._tune_list(buf, n.value)

    def autotune_reset(self) -> None:
        self.lib.fx_hydro_autotune_reset()

    # ---- multi-level grids held on the device (fx_amr_*) ----
    def amr_setup(self, tiles: Sequence[Tile], levels: Sequence[int], positions,
                  level_extent, periodic=(0, 0, 0)) -> None:
        """Hands the library a multi-level grid: tiles (with block_class and
        their level's deltas), each block's 1-based level and position on its
        level, the blocks per axis covering the domain on each level."""
        self._amr_tiles = self._tiles(tiles)          # keeps the ctypes array alive
        blk = (FxAmrBlock * len(tiles))()
        for n, (L, pos) in enumerate(zip(levels, positions)):
            blk[n].level = int(L)
            for a in range(3):
                blk[n].pos[a] = int(pos[a]) if a < len(pos) else 0
        ext = (C.c_int * (3 * len(level_extent)))(
            *[int(e[a]) if a < len(e) else 1 for e in level_extent for a in range(3)])
        per = (C.c_int * 3)(*[int(bool(p)) for p in periodic])
        st = FxStatus()
        rc = self.lib.fx_amr_setup(C.cast(self._amr_tiles, C.c_void_p), C.cast(blk, C.c_void_p),
                                   len(tiles), len(level_extent), ext, per, C.byref(st))
        if rc:
            raise FxError("fx_amr_setup", st, rc)
        self._amr_blocks = blk

    def _amr_call(self, name, *args) -> FxStatus:
        st = FxStatus()
        rc = getattr(self.lib, name)(*args, C.byref(st))
        self.last_status = st
        if rc:
            raise FxError(name, st, rc)
        return st

    def amr_fill_guards(self) -> FxStatus:
        return self._amr_call("fx_amr_fill_guards")

    def amr_restrict(self) -> FxStatus:
        return self._amr_call("fx_amr_restrict")

    def amr_compute_dt(self, cfl: float) -> float:
        out = C.c_double()
        rc = self.lib.fx_amr_compute_dt(cfl, C.byref(out))
        if rc:
            raise FxError("fx_amr_compute_dt", rc=rc)
        return out.value

    def amr_step(self, dt: float) -> FxStatus:
        return self._amr_call("fx_amr_step", C.c_double(dt))

    def amr_run_graph(self, cfl: float, steps_per_graph: int, replays: int) -> float:
        """Runs replays x steps_per_graph graph-captured steps; returns t."""
        t = C.c_double()
        self._amr_call("fx_amr_run_graph", C.c_double(cfl), int(steps_per_graph), int(replays),
                       C.byref(t))
        return t.value

    def set_kernel_path(self, path: int) -> None:
        """0 = fused column-march kernel (default; exact-solver batch steps of
        3-D blocks up to 16 x 16 across solve and update in one kernel),
        1 = three-kernel path, 2 = as 0 with the update always separate."""
        self.lib.fx_hydro_set_kernel_path(int(path))

    def set_timing(self, on: bool) -> None:
        self.lib.fx_hydro_set_timing(int(on))

    def stage_times(self):
        buf = (C.c_float * 3)()
        n = self.lib.fx_hydro_stage_times(buf, 3)
        return [float(buf[k]) for k in range(n)]

    def set_stream(self, stream) -> None:
        self.lib.fx_hydro_set_stream(C.c_void_p(stream.cuda_stream if stream is not None else 0))

    def _tiles(self, tiles: Sequence[Tile]):
        arr = (FxTile * len(tiles))()
        for n, t in enumerate(tiles):
            arr[n].U = _ptr(t.U)
            for d in range(3):
                f = t.flux[d] if (t.flux is not None and d < len(t.flux)) else None
                arr[n].flux[d] = _ptr(f)
            arr[n].pavg = _ptr(t.pavg)
            for d in range(3):
                arr[n].del_[d] = float(t.deltas[d]) if d < len(t.deltas) else 1.0
            arr[n].level = t.level
            arr[n].block_class = t.block_class
        return arr

    def compute_fluxes(self, tiles: Sequence[Tile], dt: float, dtOld: float = 0.0,
                       update_in_place: Optional[Sequence[bool]] = None) -> FxStatus:
        self._reinit()
        arr = self._tiles(tiles)
        upd = None
        if update_in_place is not None:
            upd = (C.c_ubyte * len(tiles))(*[1 if u else 0 for u in update_in_place])
        st = FxStatus()
        rc = self.lib.fx_hydro_compute_fluxes(arr, len(tiles), dt, dtOld, upd, C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_compute_fluxes", st, rc)
        return st

    def update_solution(self, tiles: Sequence[Tile], dt: float, dtOld: float = 0.0) -> FxStatus:
        self._reinit()
        arr = self._tiles(tiles)
        st = FxStatus()
        rc = self.lib.fx_hydro_update_solution(arr, len(tiles), dt, dtOld, C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_update_solution", st, rc)
        return st

    @staticmethod
    def _faces(faces):
        arr = (FxFluxRegFace * max(1, len(faces)))()
        for e, f in enumerate(faces):
            arr[e] = FxFluxRegFace(*[int(x) for x in f])
        return arr

    def fluxreg_put(self, fine_tiles: Sequence[Tile], faces, reg) -> FxStatus:
        """Grid_putFluxData for a fine level: reg (device, FX_NFLUX x nFaces)
        := coef * sum of the covering fine face fluxes.  faces: sequences
        (coarse_tile, fine_tile, dir, ci, cj, ck, fi, fj, fk), e.g. from
        flash5_amd.amr.BlockMesh.register_faces."""
        self._reinit()
        st = FxStatus()
        rc = self.lib.fx_fluxreg_put(self._tiles(fine_tiles), len(fine_tiles), self._faces(faces),
                                     len(faces), _ptr(reg), C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_fluxreg_put", st, rc)
        return st

    def fluxreg_conserve(self, coarse_tiles: Sequence[Tile], faces, reg) -> FxStatus:
        """Grid_conserveFluxes for a coarse level: listed coarse face fluxes := reg."""
        self._reinit()
        st = FxStatus()
        rc = self.lib.fx_fluxreg_conserve(self._tiles(coarse_tiles), len(coarse_tiles),
                                          self._faces(faces), len(faces), _ptr(reg), C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_fluxreg_conserve", st, rc)
        return st

    def set_seed_params(self, params) -> None:
        """Seed network for learnedSeed=True (float32 device tensor, FX_MLP_NPARAM)."""
        ptr = None
        if params is not None:
            import torch
            assert params.is_cuda and params.dtype == torch.float32 and params.is_contiguous()
            ptr = C.cast(C.c_void_p(params.data_ptr()), C.POINTER(C.c_float))
        rc = self.lib.fx_hydro_set_seed_params(ptr)
        if rc:
            raise FxError("fx_hydro_set_seed_params", rc=rc)

    def face_states(self):
        """Riemann problems of the last exact-solver step, (6, nfaces) float64
        (FP32 values after a mixed-precision step, which stores them in FP32)."""
        import torch
        n = C.c_long(0)
        self.lib.fx_hydro_face_states(None, 0, C.byref(n))
        if n.value <= 0:
            raise FxError("fx_hydro_face_states", rc=3)
        S = torch.empty((6, n.value), dtype=torch.float64, device="cuda")
        rc = self.lib.fx_hydro_face_states(_ptr(S), n.value, C.byref(n))
        if rc:
            raise FxError("fx_hydro_face_states", rc=rc)
        return S

    def shock_detect(self, tiles: Sequence[Tile]) -> FxStatus:
        """hy_shockDetect over the tiles: writes SHOK_VAR in place; the
        status's min_cell_cfl is the CFL cap under shockLowerCFL."""
        self._reinit()
        arr = self._tiles(tiles)
        st = FxStatus()
        rc = self.lib.fx_hydro_shock_detect(arr, len(tiles), C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_shock_detect", st, rc)
        return st

    def shock_detect_batch(self, U, deltas: Sequence[float]) -> FxStatus:
        """shock_detect over a (nblocks, nunk, nz, ny, nx) batch."""
        self._reinit()
        nb = U.shape[0]
        st = FxStatus()
        rc = self.lib.fx_hydro_shock_detect_batch(_ptr(U), nb, U[0].numel() if nb else 0,
                                                  (C.c_double * 3)(*[float(x) for x in deltas]),
                                                  C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_shock_detect_batch", st, rc)
        return st

    def advance_batch(self, U, dt: float, deltas: Sequence[float]) -> FxStatus:
        """In-place step of a (nblocks, nunk, nz, ny, nx) batch of independent blocks."""
        self._reinit()
        nb = U.shape[0]
        stride = U[0].numel() if nb else 0
        st = FxStatus()
        rc = self.lib.fx_hydro_advance_batch(_ptr(U), nb, stride,
                                             (C.c_double * 3)(*[float(x) for x in deltas]),
                                             dt, C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_advance_batch", st, rc)
        return st

    def advance_batch_noalloc(self, U, dt: float, deltas: Sequence[float]) -> FxStatus:
        """advance_batch through the entry point that allocates nothing (its
        scratch is a fixed device array: at most 77 blocks of 16^3 a call)."""
        self._reinit()
        nb = U.shape[0]
        stride = U[0].numel() if nb else 0
        st = FxStatus()
        rc = self.lib.fx_hydro_advance_batch_noalloc(_ptr(U), nb, stride,
                                                     (C.c_double * 3)(*[float(x) for x in deltas]),
                                                     dt, C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_hydro_advance_batch_noalloc", st, rc)
        return st

    def compute_dt(self, U, deltas: Sequence[float], cfl: float) -> float:
        self._reinit()
        out = C.c_double(0.0)
        nb = U.shape[0]
        rc = self.lib.fx_hydro_compute_dt(_ptr(U), nb, U[0].numel(),
                                          (C.c_double * 3)(*[float(x) for x in deltas]),
                                          cfl, C.byref(out))
        if rc:
            raise FxError("fx_hydro_compute_dt", rc=rc)
        return out.value
