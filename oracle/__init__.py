"""TEST INFRASTRUCTURE: ctypes bindings for the CPU FP64 oracle.

The oracle (``hydro_oracle.c``) restates FLASH5's unsplit hydro block update
(source/physics/Hydro/HydroMain/unsplit) on the CPU.  It is the checker for
the HIP product in ``flash5_amd``; only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import it.

Parity status of each oracle function and the check that pins it is listed
in the module docstrings of tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

DENS, VELX, VELY, VELZ, PRES, ENER, GAMC, GAME, TEMP, EINT = range(10)
NVAR_USED = 10
F_DENS, F_XMOM, F_YMOM, F_ZMOM, F_ENER, F_PRES, F_EINT, F_VOLU = range(8)
NFLUX = 8
MINMOD, MC, HYBRID, VANLEER, LIMITED = 1, 2, 3, 4, 5
ROE, HLL, HLLC, HYBR, EXACT = 1, 3, 4, 8, 20
SHOK = 10        # slot of SHOK_VAR when a layout carries it after the ten used variables


class Params(C.Structure):
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
        ("shockDetect", C.c_int), ("shokVar", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [("fallbacks", C.c_longlong), ("riemannErrors", C.c_longlong),
                ("eigenErrors", C.c_longlong), ("exactIters", C.c_longlong),
                ("exactFallbacks", C.c_longlong)]


_lib = None
_DP = C.POINTER(C.c_double)


def build() -> None:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.ox_default_params.argtypes = [C.POINTER(Params), C.c_int, C.c_int, C.c_double]
        L.ox_block_cells.argtypes = [C.POINTER(Params)]
        L.ox_block_cells.restype = C.c_long
        L.ox_flux_faces.argtypes = [C.POINTER(Params), C.c_int]
        L.ox_flux_faces.restype = C.c_long
        L.ox_eos_box.argtypes = [C.POINTER(Params), _DP, C.c_int, C.c_int * 3, C.c_int * 3]
        L.ox_compute_fluxes.argtypes = [C.POINTER(Params), _DP, _DP * 3, _DP, C.c_double,
                                        C.c_double * 3, C.c_int, C.POINTER(Stats)]
        L.ox_update_solution.argtypes = [C.POINTER(Params), _DP, _DP * 3, _DP, C.c_double,
                                         C.c_double * 3, C.POINTER(Stats)]
        L.ox_compute_fluxes_batch.argtypes = [C.POINTER(Params), _DP, C.c_long, C.c_double,
                                              C.c_double * 3, C.c_int, C.POINTER(Stats)]
        L.ox_block_dt.argtypes = [C.POINTER(Params), _DP, C.c_double * 3, C.c_double]
        L.ox_block_dt.restype = C.c_double
        L.ox_eigen.argtypes = [_DP, C.c_int, _DP, _DP, _DP]
        L.ox_limiter.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.ox_limiter.restype = C.c_double
        L.ox_tvd_slope.argtypes = [C.POINTER(Params), _DP, _DP, _DP, _DP, _DP, _DP]
        L.ox_upwind_trans_flux.argtypes = [C.POINTER(Params), _DP, _DP, _DP, _DP, _DP, _DP, _DP]
        L.ox_ppm.argtypes = [C.POINTER(Params), C.c_int, C.c_double, C.c_double, _DP, C.c_double,
                             C.c_int, _DP, _DP, _DP, _DP, _DP, _DP]
        L.ox_hllc.argtypes = [C.POINTER(Params), C.c_int, _DP, _DP, _DP, _DP]
        L.ox_hll.argtypes = [C.POINTER(Params), C.c_int, _DP, _DP, _DP, _DP]
        L.ox_shock_detect.argtypes = [C.POINTER(Params), _DP]
        L.ox_shock_detect.restype = C.c_long
        L.ox_exact_riemann.argtypes = [C.POINTER(Params), C.c_int, _DP, _DP, _DP, _DP,
                                       C.POINTER(Stats)]
        L.ox_exact_pstar.argtypes = [C.c_double] * 8 + [C.c_double, C.c_int, _DP, _DP]
        L.ox_exact_pstar.restype = C.c_int
        L.ox_exact_sample.argtypes = [C.c_double] * 8 + [_DP]
        L.ox_exact_sample.restype = C.c_int
        L.ox_fluxreg_put.argtypes = [C.POINTER(Params), C.POINTER(_DP), C.c_void_p, C.c_int, _DP]
        L.ox_fluxreg_conserve.argtypes = [C.POINTER(Params), C.POINTER(_DP), C.c_void_p, C.c_int,
                                          _DP]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def params(ndim: int, nb: int, gamma: float, **kw) -> Params:
    p = Params()
    lib().ox_default_params(C.byref(p), ndim, nb, gamma)
    for k, v in kw.items():
        if k == "var":
            for i, s in enumerate(v):
                p.var[i] = s
        else:
            setattr(p, k, v)
    return p


def block_shape(p: Params):
    """(nunk, nz, ny, nx) of one guard-inclusive block."""
    nx = p.nxb + 2 * p.ng
    ny = p.nyb + 2 * p.ng if p.ndim > 1 else 1
    nz = p.nzb + 2 * p.ng if p.ndim > 2 else 1
    return (p.nunk, nz, ny, nx)


def flux_shape(p: Params, d: int):
    """(NFLUX, fz, fy, fx) of one direction's interior-face flux array."""
    fx = p.nxb + (d == 0)
    fy = (p.nyb if p.ndim > 1 else 1) + (d == 1 and p.ndim > 1)
    fz = (p.nzb if p.ndim > 2 else 1) + (d == 2 and p.ndim > 2)
    return (NFLUX, fz, fy, fx)


def interior_shape(p: Params):
    return (p.nzb if p.ndim > 2 else 1, p.nyb if p.ndim > 1 else 1, p.nxb)


def eos(p: Params, U: np.ndarray, mode: int, lo=None, hi=None) -> None:
    """Gamma-law EOS on a box (mode 1 = DENS_PRES, 2 = DENS_EI); whole block by default."""
    _, nz, ny, nx = block_shape(p)
    lo = lo if lo is not None else (0, 0, 0)
    hi = hi if hi is not None else (nx - 1, ny - 1, nz - 1)
    lib().ox_eos_box(C.byref(p), _p(U), mode, (C.c_int * 3)(*lo), (C.c_int * 3)(*hi))


def compute_fluxes(p: Params, U: np.ndarray, dt: float, dx, update: bool = True,
                   want_flux: bool = True, want_pavg: bool = False):
    """hy_computeFluxes on one block; returns (rc, fluxes, pavg, stats)."""
    fl = [np.zeros(flux_shape(p, d)) if (want_flux and d < p.ndim) else None for d in range(3)]
    arr = (_DP * 3)(*[(_p(f) if f is not None else _DP()) for f in fl])
    pavg = np.zeros(interior_shape(p)) if want_pavg else None
    st = Stats()
    rc = lib().ox_compute_fluxes(C.byref(p), _p(U), arr, _p(pavg) if pavg is not None else _DP(),
                                 dt, (C.c_double * 3)(*dx), int(update), C.byref(st))
    return rc, fl, pavg, st


def update_solution(p: Params, U: np.ndarray, fluxes, pavg: np.ndarray, dt: float, dx):
    arr = (_DP * 3)(*[(_p(f) if f is not None else _DP()) for f in fluxes])
    st = Stats()
    return lib().ox_update_solution(C.byref(p), _p(U), arr, _p(pavg), dt, (C.c_double * 3)(*dx),
                                    C.byref(st))


def compute_fluxes_batch(p: Params, U: np.ndarray, dt: float, dx, nthreads: int = 0):
    st = Stats()
    nb = U.shape[0]
    rc = lib().ox_compute_fluxes_batch(C.byref(p), _p(U), nb, dt, (C.c_double * 3)(*dx),
                                       nthreads, C.byref(st))
    return rc, st


def block_dt(p: Params, U: np.ndarray, dx, cfl: float) -> float:
    return lib().ox_block_dt(C.byref(p), _p(U), (C.c_double * 3)(*dx), cfl)


# ---- per-stage entry points (primitive vectors: dens, vx, vy, vz, pres,
# gamc, game, eint, grav) ----
def eigen(V, d: int):
    V = np.ascontiguousarray(V, dtype=np.float64)
    lam, le, re = np.zeros(5), np.zeros((5, 5)), np.zeros((5, 5))
    rc = lib().ox_eigen(_p(V), d, _p(lam), _p(le), _p(re))
    return rc, lam, le, re          # le[wave][var], re[wave][var]


def limiter(kind: int, a: float, b: float, beta: float = 1.0) -> float:
    return lib().ox_limiter(kind, a, b, beta)


def tvd_slope(p: Params, VL, V0, VR, d: int):
    rc, lam, le, re = eigen(V0, d)
    out = np.zeros(9)
    lib().ox_tvd_slope(C.byref(p), _p(np.ascontiguousarray(VL, dtype=np.float64)),
                       _p(np.ascontiguousarray(V0, dtype=np.float64)),
                       _p(np.ascontiguousarray(VR, dtype=np.float64)), _p(lam), _p(le), _p(out))
    return out


def trans_flux(p: Params, vm1, vc0, vp1, V0, d: int):
    """sigma of a cell with eigen system of V0 (hy_upwindTransverseFlux)."""
    rc, lam, le, re = eigen(V0, d)
    sig = np.zeros(9)
    f = lambda a: _p(np.ascontiguousarray(a, dtype=np.float64))
    lib().ox_upwind_trans_flux(C.byref(p), f(vm1), f(vc0), f(vp1), _p(lam), _p(le), _p(re), _p(sig))
    return sig


def ppm(p: Params, d: int, dt: float, delta: float, data, flat: float = 0.0, trans: bool = False):
    data = np.ascontiguousarray(data, dtype=np.float64).reshape(5, 9)
    lam, le, re = np.zeros(5), np.zeros((5, 5)), np.zeros((5, 5))
    Wp, Wm, sig = np.zeros(9), np.zeros(9), np.zeros(9)
    rc = lib().ox_ppm(C.byref(p), d, dt, delta, _p(data), flat, int(trans), _p(lam), _p(le), _p(re),
                      _p(Wp), _p(Wm), _p(sig))
    return rc, Wp, Wm, sig


def hllc(p: Params, d: int, VL, VR):
    F, sp = np.zeros(6), C.c_double(0.0)
    f = lambda a: _p(np.ascontiguousarray(a, dtype=np.float64))
    rc = lib().ox_hllc(C.byref(p), d, f(VL), f(VR), _p(F), C.byref(sp))
    return rc, F, sp.value


def hll(p: Params, d: int, VL, VR):
    F, sp = np.zeros(6), C.c_double(0.0)
    f = lambda a: _p(np.ascontiguousarray(a, dtype=np.float64))
    rc = lib().ox_hll(C.byref(p), d, f(VL), f(VR), _p(F), C.byref(sp))
    return rc, F, sp.value


def shock_detect(p: Params, U: np.ndarray) -> int:
    """hy_shockDetectBlk on one block in place; returns the flagged cells
    (-1 without a SHOK slot)."""
    return lib().ox_shock_detect(C.byref(p), _p(U))


def exact_flux(p: Params, d: int, VL, VR):
    F, sp = np.zeros(6), C.c_double(0.0)
    f = lambda a: _p(np.ascontiguousarray(a, dtype=np.float64))
    st = Stats()
    rc = lib().ox_exact_riemann(C.byref(p), d, f(VL), f(VR), _p(F), C.byref(sp), C.byref(st))
    return rc, F, st


def exact_pstar(gamma, rl, ul, pl, rr, ur, pr, guess=-1.0, tol=1e-12, maxit=50):
    ps, us = C.c_double(0.0), C.c_double(0.0)
    it = lib().ox_exact_pstar(gamma, rl, ul, pl, rr, ur, pr, guess, tol, maxit,
                              C.byref(ps), C.byref(us))
    return it, ps.value, us.value


def prim(rho, u, v, w, p, gamma):
    """Primitive reconstruction vector with gamc = game = gamma, eint from the EOS."""
    return np.array([rho, u, v, w, p, gamma, gamma, p / ((gamma - 1.0) * rho), 0.0])


# ---- flux register (faces: sequence of 9-tuples (coarse_tile, fine_tile, dir,
# ci, cj, ck, fi, fj, fk)); fluxes: per tile a list of 3 arrays or None ----
def _faces_c(faces):
    arr = (C.c_int * (9 * max(1, len(faces))))()
    for e, f in enumerate(faces):
        for k in range(9):
            arr[9 * e + k] = int(f[k])
    return arr


def _flux_ptrs(fluxes):
    ptrs = (_DP * (3 * max(1, len(fluxes))))()
    for t, fl in enumerate(fluxes):
        for d in range(3):
            ptrs[3 * t + d] = _p(fl[d]) if fl[d] is not None else _DP()
    return ptrs


def fluxreg_put(p: Params, fine_fluxes, faces):
    """Grid_putFluxData on a level: register values, shape (NFLUX, nFaces)."""
    reg = np.zeros((NFLUX, max(1, len(faces))))
    lib().ox_fluxreg_put(C.byref(p), _flux_ptrs(fine_fluxes), C.cast(_faces_c(faces), C.c_void_p),
                         len(faces), _p(reg))
    return reg


def fluxreg_conserve(p: Params, coarse_fluxes, faces, reg) -> None:
    """Grid_conserveFluxes on a level: overwrite coarse fluxes in place."""
    reg = np.ascontiguousarray(reg, dtype=np.float64)
    lib().ox_fluxreg_conserve(C.byref(p), _flux_ptrs(coarse_fluxes),
                              C.cast(_faces_c(faces), C.c_void_p), len(faces), _p(reg))


def exact_sample(gamma, rl, ul, pl, rr, ur, pr, S):
    """Exact (rho, u, p) at x/t = S."""
    out = np.zeros(3)
    lib().ox_exact_sample(gamma, rl, ul, pl, rr, ur, pr, S, _p(out))
    return out
