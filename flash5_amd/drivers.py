"""Multi-step drivers on uniform block grids: the Sod shock tube along the
diagonal (FLASH's Simulation/SimulationMain/Sod) and a Kelvin-Helmholtz
shear layer.

Each step does what FLASH's driver and Hydro do on one level:
guard-cell fill (Grid_fillGuardCells, Hydro.F90:121), the unsplit update of
every block in place, and the CFL time step from the updated interior cells
(Hydro_computeDt.F90:186-282, min over blocks).  Everything runs on the GPU
through libfx_hydro.so; the host only loops.  FLASH's driver also limits how
fast dt may grow between steps (Driver_computeDt); that belongs to the
driver, not the hydro path, and is left out: every step uses the CFL dt, and
the last one is shortened to land on tmax.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from .hydro import FxError, FxStatus, Hydro, HydroParams, load_library

BC_OUTFLOW, BC_REFLECT = -1, -2
DENS, VELX, VELY, VELZ, PRES, ENER, GAMC, GAME, TEMP, EINT = range(10)


def uniform_neighbors(nblk, periodic, ndim: int) -> np.ndarray:
    """Neighbour table [nBlocks][6] (-x, +x, -y, +y, -z, +z) of a uniform
    block grid, blocks numbered b = bi + nbx*(bj + nby*bk)."""
    nblk = tuple(nblk) + (1,) * (3 - len(nblk))
    periodic = tuple(periodic) + (False,) * (3 - len(periodic))
    nb = np.full((nblk[0] * nblk[1] * nblk[2], 6), BC_OUTFLOW, dtype=np.int32)
    for bk in range(nblk[2]):
        for bj in range(nblk[1]):
            for bi in range(nblk[0]):
                b = bi + nblk[0] * (bj + nblk[1] * bk)
                idx = [bi, bj, bk]
                for a in range(ndim):
                    for side in (0, 1):
                        n = list(idx)
                        n[a] += 1 if side else -1
                        if n[a] < 0 or n[a] >= nblk[a]:
                            if not periodic[a]:
                                continue
                            n[a] %= nblk[a]
                        nb[b, 2 * a + side] = n[0] + nblk[0] * (n[1] + nblk[1] * n[2])
    return nb


def eos_fill(U, gamma: float, gas_constant: float = 8.3144598e7, abar: float = 1.0):
    """Ideal-gas EOS (MODE_DENS_PRES) on numpy blocks: eint, ener, temp,
    gamc, game from dens, velocities and pres (eos_idealGamma.F90)."""
    ekin = 0.5 * (U[:, VELX] ** 2 + U[:, VELY] ** 2 + U[:, VELZ] ** 2)
    U[:, EINT] = U[:, PRES] * (1.0 / (gamma - 1.0)) / U[:, DENS]
    U[:, ENER] = U[:, EINT] + ekin
    U[:, TEMP] = U[:, EINT] / ((1.0 / (gamma - 1.0)) * gas_constant) * abar
    U[:, GAMC] = gamma
    U[:, GAME] = gamma
    return U


def sod_blocks(nblk: int = 16, nb: int = 8, ng: int = 4, gamma: float = 1.4,
               angle_deg: float = 45.0, posn: float = 1.0):
    """FLASH's Sod setup on [0,1]^2 (Sod/Simulation_initBlock.F90:140-223,
    Simulation_init.F90:106-121): left state (1, 0, 1), right (0.125, 0, 0.1),
    the discontinuity x = posn - y*yCos/xCos, cells it cuts take the average.
    Returns numpy blocks [nBlocks][10][1][ny][nx] (guards left for the fill)."""
    xcos = math.cos(angle_deg * 0.0174532925)
    ycos = math.sqrt(1.0 - xcos * xcos)
    nx = nb + 2 * ng
    U = np.zeros((nblk * nblk, 10, 1, nx, nx))
    h = 1.0 / (nblk * nb)
    left = (1.0, 1.0)           # rho, p
    right = (0.125, 0.1)
    for bj in range(nblk):
        for bi in range(nblk):
            b = bi + nblk * bj
            for j in range(nx):
                y = (bj * nb + j - ng + 0.5) * h
                lposn = posn - y * ycos / xcos
                for i in range(nx):
                    xl = (bi * nb + i - ng) * h
                    xr = xl + h
                    if xr <= lposn:
                        rho, p = left
                    elif xl < lposn < xr:
                        rho, p = 0.5 * (left[0] + right[0]), 0.5 * (left[1] + right[1])
                    else:
                        rho, p = right
                    U[b, DENS, 0, j, i] = rho
                    U[b, PRES, 0, j, i] = p
    return eos_fill(U, gamma)


def kh_blocks(ncells: int = 2048, nb: int = 16, ng: int = 4, gamma: float = 1.4,
              seed: int = 0x5EED0003, width: float = 0.01):
    """Kelvin-Helmholtz shear layer on the periodic unit square: rho = 2 for
    |y - 0.5| < 0.25 else 1, u = +0.5 inside and -0.5 outside joined by tanh
    profiles of the given width, v = 0.01 sin(4 pi x) plus 1e-4 seeded white
    noise, p = 2.5."""
    nblk = ncells // nb
    nx = nb + 2 * ng
    rng = np.random.default_rng(seed)
    h = 1.0 / ncells
    U = np.zeros((nblk * nblk, 10, 1, nx, nx))
    xs = (np.arange(nx) - ng + 0.5) * h
    for bj in range(nblk):
        y = bj * nb * h + xs
        inside = np.abs(y - 0.5) < 0.25
        u = 0.5 * (np.tanh((y - 0.25) / width) - np.tanh((y - 0.75) / width)) - 0.5
        for bi in range(nblk):
            b = bi + nblk * bj
            x = bi * nb * h + xs
            U[b, DENS, 0] = np.where(inside, 2.0, 1.0)[:, None]
            U[b, VELX, 0] = u[:, None]
            U[b, VELY, 0] = 0.01 * np.sin(4 * np.pi * x)[None, :] + 1e-4 * rng.standard_normal((nx, nx))
            U[b, PRES, 0] = 2.5
    return eos_fill(U, gamma)


@dataclass
class RunStats:
    steps: int
    time: float
    wall_s: float
    zone_updates: int
    fallbacks: int

    @property
    def zone_updates_per_s(self) -> float:
        return self.zone_updates / self.wall_s if self.wall_s > 0 else 0.0


def shock_cfl(cfl: float, cap: float) -> float:
    """CFL of the next dt under shockLowerCFL (Hydro.F90:104-108 and 128-141,
    hy_shockDetectBlk.F90:228-240).  Each step resets the CFL to cfl before
    the detector, a flagged cell lowers it to min(cfl, cap), and the
    recovery before the fluxes raises it by 1.2 up to cfl: min(cfl, 1.2 cap)
    after a step with a flagged cell.  cap is fx_status.min_cell_cfl, 0 when
    nothing was flagged or shockLowerCFL is off."""
    return min(cfl, 1.2 * cap) if cap > 0 else cfl


def step_cfl(cfl: float, last: float, detect: bool, shock_cap: float,
             fallback_cap: float) -> float:
    """hy_cfl at the end of a step, which the next dt takes.  cfl is the
    original CFL, last the value the previous step left (cfl before the
    first step), shock_cap and fallback_cap the step's fx_status
    min_cell_cfl and fallback_cfl (0: none).  The reference resets hy_cfl to
    the original before the detector when detection is on
    (Hydro.F90:104-108), lowers it to the shock cap when a cell is flagged
    (hy_shockDetectBlk.F90:228-240), raises it by 1.2 up to the original
    before the fluxes (Hydro.F90:128-141) and, under hy_fallbackLowerCFL,
    lowers it to the fallback cap when a state fell back to first order
    (hy_getRiemannState.F90:934-941, 1303-1319).  Without fallbacks and with
    detection on this is shock_cfl(cfl, shock_cap)."""
    c = cfl if detect else last
    if shock_cap > 0:
        c = min(c, shock_cap)
    c = 1.2 * c if 1.2 * c < cfl else cfl
    if fallback_cap > 0:
        c = min(c, fallback_cap)
    return c


class UniformGrid:
    """A single-level grid of equal blocks advanced by the hot path."""

    def __init__(self, params: HydroParams, U, nbr: np.ndarray, deltas, cfl: float = 0.8):
        import torch
        self.params = params
        self.hydro = Hydro(params)
        self.lib = load_library()
        self.U = U if isinstance(U, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(U))
        self.U = self.U.to("cuda").contiguous()
        self.nbr = np.ascontiguousarray(nbr, dtype=np.int32)
        self.deltas = tuple(float(d) for d in deltas)
        self.cfl = cfl
        self.t = 0.0
        self.nsteps = 0
        self.detect = bool(params.shockDetect or params.shockLowerCFL or
                           params.RiemannSolver.lower() == "hybrid")
        self._cfl_next = cfl          # the CFL the last step's flags and fallbacks left

    def fill_guards(self) -> None:
        import ctypes as C
        st = FxStatus()
        nb = self.U.shape[0]
        rc = self.lib.fx_grid_fill_guards(
            C.cast(C.c_void_p(self.U.data_ptr()), C.POINTER(C.c_double)), nb, self.U[0].numel(),
            self.nbr.ctypes.data_as(C.POINTER(C.c_int)), C.byref(st))
        if rc:
            raise FxError("fx_grid_fill_guards", st, rc)

    def step(self, dt_max: float = math.inf) -> float:
        """Guard fill, CFL dt (capped by dt_max), unsplit update in place.
        With shock detection on: fill, flag, fill again (SHOK_VAR's guard
        cells too), as Hydro.F90:94-123 orders them.  Under shockLowerCFL or
        hy_fallbackLowerCFL the dt takes the CFL that the previous step's
        flags and fallbacks left (step_cfl)."""
        cfl = self._cfl_next
        self.fill_guards()
        shock_cap = 0.0
        if self.detect:
            shock_cap = self.hydro.shock_detect_batch(self.U, self.deltas).min_cell_cfl
            self.fill_guards()
        dt = min(self.hydro.compute_dt(self.U, self.deltas, cfl), dt_max)
        st = self.hydro.advance_batch(self.U, dt, self.deltas)
        self._cfl_next = step_cfl(self.cfl, cfl, self.detect, shock_cap, st.fallback_cfl)
        self.t += dt
        self.nsteps += 1
        return dt

    def run_graph(self, steps_per_graph: int, replays: int) -> float:
        """steps_per_graph whole steps (guard fill, device CFL dt, update)
        captured once in a HIP graph and replayed; returns the simulated time
        the replays covered.  The run starts at the original CFL; under
        shockLowerCFL the device carries the lowered CFL from step to step."""
        import ctypes as C
        st = FxStatus()
        t = C.c_double(0.0)
        self.hydro._reinit()
        rc = self.lib.fx_grid_run_graph(
            C.cast(C.c_void_p(self.U.data_ptr()), C.POINTER(C.c_double)), self.U.shape[0],
            self.U[0].numel(), self.nbr.ctypes.data_as(C.POINTER(C.c_int)),
            (C.c_double * 3)(*self.deltas), self.cfl, steps_per_graph, replays, C.byref(t),
            C.byref(st))
        self.last_status = st
        if rc:
            raise FxError("fx_grid_run_graph", st, rc)
        self.t += t.value
        self.nsteps += steps_per_graph * replays
        return t.value

    def run(self, tmax: float = math.inf, nsteps: int | None = None) -> RunStats:
        import torch
        p = self.params
        zones = self.U.shape[0] * p.nxb * (p.nyb if p.ndim > 1 else 1) * (p.nzb if p.ndim > 2 else 1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        n = 0
        fb = 0
        while (nsteps is None or n < nsteps) and self.t < tmax * (1 - 1e-14):
            self.step(tmax - self.t)
            fb += int(self.hydro.last_status.fallbacks)
            n += 1
        torch.cuda.synchronize()
        return RunStats(n, self.t, time.perf_counter() - t0, n * zones, fb)


def sod(nblk: int = 16, nb: int = 8, tmax: float = 0.2, **kw):
    """Config-1 Sod problem: 128x128 cells as 16x16 blocks of 8x8, PPM with
    MC limiting (Sod/flash.par) and HLLC, flattening on, eintSwitch 1e-4,
    outflow on every side, CFL 0.8, to t = 0.2."""
    hp = HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, slopeLimiter="mc",
                     use_flattening=True, eintSwitch=1e-4, **kw)
    U = sod_blocks(nblk, nb)
    grid = UniformGrid(hp, U, uniform_neighbors((nblk, nblk), (False, False), 2),
                       (1.0 / (nblk * nb),) * 2 + (1.0,))
    return grid, grid.run(tmax=tmax)


def kh(ncells: int = 2048, nb: int = 16, nsteps: int = 200, **kw):
    """Config-3 Kelvin-Helmholtz run: 2048^2 cells as 16x16 blocks, periodic."""
    hp = HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, **kw)
    nblk = ncells // nb
    grid = UniformGrid(hp, kh_blocks(ncells, nb), uniform_neighbors((nblk, nblk), (True, True), 2),
                       (1.0 / ncells,) * 2 + (1.0,))
    return grid, grid.run(nsteps=nsteps)


def kh_precision_study(ncells: int = 2048, nb: int = 16, nsteps: int = 200):
    """Config-3 precision study: the same Kelvin-Helmholtz run in FP64 and in
    the mixed mode (FP32 reconstruction and fluxes, FP64 update), stepped with
    the FP64 run's dt sequence so both integrate the same times, then the
    per-variable L1 / Linf (sfocu norms) of mixed against FP64 over all
    interior cells.  The FP64 GPU path agrees with the FP64 oracle bit for
    bit, so it stands in for the oracle at this size."""
    import torch
    from . import sfocu
    from .hydro import PREC_FP64, PREC_MIXED
    nblk = ncells // nb
    U0 = kh_blocks(ncells, nb)
    nbr = uniform_neighbors((nblk, nblk), (True, True), 2)
    deltas = (1.0 / ncells,) * 2 + (1.0,)
    runs = {}
    dts = []
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        hp = HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, precision=prec)
        grid = UniformGrid(hp, U0, nbr, deltas)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(nsteps):
            if name == "fp64":
                dts.append(grid.step())
            else:
                grid.fill_guards()
                grid.hydro.advance_batch(grid.U, dts[s], deltas)
                grid.t += dts[s]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        runs[name] = (grid.U.cpu().numpy(), wall, grid.t)
        del grid
    names = {"dens": DENS, "velx": VELX, "vely": VELY, "pres": PRES, "ener": ENER, "eint": EINT,
             "temp": TEMP}
    Ua, Ub = runs["mixed"][0], runs["fp64"][0]
    norms = sfocu.compare(Ua, Ub, 1, names, 2, 4)
    zones = ncells * ncells * nsteps
    return {"cells": ncells * ncells, "steps": nsteps, "t_end": runs["fp64"][2],
            "l1": {k: v[0] for k, v in norms.items()}, "linf": {k: v[1] for k, v in norms.items()},
            "fp64_zone_updates_per_s": zones / runs["fp64"][1],
            "mixed_zone_updates_per_s": zones / runs["mixed"][1]}


# ---- three levels of uneven block sizes: the four-quadrant Riemann problem ----
def quadrant_states(seed: int | None = None):
    """Seeded quadrant states (dens, velx, vely, pres) for quadrants 0: x<1/2,
    y<1/2; 1: x>1/2, y<1/2; 2: x<1/2, y>1/2; 3: x>1/2, y>1/2, drawn with dens
    in [0.1, 2], pres in [0.03, 1.5], velocities in [-1.2, 1.2]; velz = 0."""
    from .workloads import SEED_BASE
    rng = np.random.default_rng(SEED_BASE + 5 if seed is None else seed)
    dens = rng.uniform(0.1, 2.0, 4)
    pres = rng.uniform(0.03, 1.5, 4)
    vel = rng.uniform(-1.2, 1.2, (4, 2))
    return dens, vel, pres


class LevelStorage:
    """Device storage of a LevelGrid: one contiguous float64 tensor per level
    (blocks, nunk, nz, ny, nx); the tiles are views of its blocks."""

    def __init__(self, grid, nunk: int = 10, device="cuda"):
        import torch
        self.grid = grid
        self.index = []                  # per block: (level, slot)
        self.levels = []
        count = [0] * len(grid.nb)
        for L, _ in grid.blocks:
            self.index.append((L, count[L]))
            count[L] += 1
        for L in range(len(grid.nb)):
            self.levels.append(torch.zeros((count[L],) + grid.block_shape(L, nunk),
                                           dtype=torch.float64, device=device))

    def block(self, b):
        L, s = self.index[b]
        return self.levels[L][s]

    def clone(self):
        other = LevelStorage.__new__(LevelStorage)
        other.grid, other.index = self.grid, self.index
        other.levels = [t.clone() for t in self.levels]
        return other

    def leaf_masks(self):
        """Per level: which blocks are leaves (refined blocks are covered whole)."""
        import torch
        g = self.grid
        refined = [set() for _ in g.nb]
        for L, pos in g.blocks:
            if L == 0:
                continue
            r = [g.ext[L][a] // g.ext[L - 1][a] for a in range(3)]
            refined[L - 1].add(tuple(pos[a] // r[a] for a in range(3)))
        masks = [torch.ones(t.shape[0], dtype=torch.bool, device=t.device) for t in self.levels]
        for b, (L, pos) in enumerate(g.blocks):
            if tuple(pos) in refined[L]:
                masks[L][self.index[b][1]] = False
        return masks


def quadrant_init(grid, gamma: float = 1.4, states=None, nunk: int = 10) -> LevelStorage:
    """The four-quadrant Riemann problem lifted to 3-D (uniform in z) on
    every level, guard cells included, with the ideal-gas EOS."""
    import torch
    dens, vel, pres = quadrant_states() if states is None else states
    S = LevelStorage(grid, nunk)
    for L, U in enumerate(S.levels):
        blocks = [b for b, (Lb, _) in enumerate(grid.blocks) if Lb == L]
        pos = torch.tensor([grid.blocks[b][1] for b in blocks], dtype=torch.float64,
                           device=U.device)
        d = grid.dels(L)
        nb = grid.nb[L]
        g = grid.ng
        ix = torch.arange(-g, nb[0] + g, dtype=torch.float64, device=U.device)
        iy = torch.arange(-g, nb[1] + g, dtype=torch.float64, device=U.device)
        x = (pos[:, 0, None] * nb[0] + ix[None, :] + 0.5) * d[0]      # (blocks, nx)
        y = (pos[:, 1, None] * nb[1] + iy[None, :] + 0.5) * d[1]      # (blocks, ny)
        q = (x[:, None, :] > 0.5).long() + 2 * (y[:, :, None] > 0.5).long()   # (blocks, ny, nx)
        q = q[:, None, :, :].expand(-1, U.shape[2], -1, -1)
        tab = lambda v: torch.tensor(v, dtype=torch.float64, device=U.device)[q]
        U[:, DENS] = tab(dens)
        U[:, VELX] = tab(vel[:, 0])
        U[:, VELY] = tab(vel[:, 1])
        U[:, VELZ] = 0.0
        U[:, PRES] = tab(pres)
        eos_fill(U, gamma)
    return S


class AmrRun:
    """A LevelGrid stepped on the device (fx_amr_*): one block class per
    level shape, the grid handed to the library once."""

    def __init__(self, grid, params: HydroParams, storage: LevelStorage, seed_params=None):
        from .hydro import Tile
        self.grid, self.S = grid, storage
        shapes = []
        for n in grid.nb:
            if n not in shapes:
                shapes.append(n)
        self.hydro = Hydro(params)
        if seed_params is not None:
            self.hydro.set_seed_params(seed_params)
        self.hydro.set_block_classes(shapes)
        tiles = [Tile(U=storage.block(b), deltas=grid.dels(L), level=L + 1,
                      block_class=shapes.index(grid.nb[L]))
                 for b, (L, _) in enumerate(grid.blocks)]
        self.hydro.amr_setup(tiles, [L + 1 for L, _ in grid.blocks], [p for _, p in grid.blocks],
                             grid.ext, grid.periodic)
        self.t = 0.0
        self.detect = bool(params.shockDetect or params.shockLowerCFL or
                           params.RiemannSolver.lower() == "hybrid")
        self.cfl = None               # the original CFL, from the first dt() call
        self._cfl_next = None         # the CFL the last step's flags and fallbacks left
        self._flag_cap = 0.0          # shockLowerCFL: the cap the last step's flags set

    def step(self, dt: float) -> FxStatus:
        st = self.hydro.amr_step(dt)
        self.t += dt
        self._flag_cap = st.min_cell_cfl
        if self.cfl is not None:
            last = self._cfl_next if self._cfl_next is not None else self.cfl
            self._cfl_next = step_cfl(self.cfl, last, self.detect, st.min_cell_cfl,
                                      st.fallback_cfl)
        return st

    def dt(self, cfl: float = 0.8) -> float:
        """CFL dt over every level; under shockLowerCFL or hy_fallbackLowerCFL
        with the CFL that the last step's flags and fallbacks left
        (step_cfl).  cfl is the original CFL; a run keeps the first one."""
        if self.cfl is None:
            self.cfl = cfl
        c = self._cfl_next if self._cfl_next is not None else self.cfl
        return self.hydro.amr_compute_dt(c)

    def run_graph(self, cfl: float, steps_per_graph: int, replays: int) -> float:
        self.t += self.hydro.amr_run_graph(cfl, steps_per_graph, replays)
        return self.t


def leaf_l1(a: LevelStorage, b: LevelStorage, names) -> dict:
    """sfocu L1 (sum |a - b| / sum |b|) per variable over the leaf interior cells."""
    out = {}
    masks = b.leaf_masks()
    g = a.grid
    for name, v in names.items():
        num = den = 0.0
        for L, (Ua, Ub) in enumerate(zip(a.levels, b.levels)):
            nb = g.nb[L]
            sl = (masks[L], v, slice(4, 4 + nb[2]), slice(4, 4 + nb[1]), slice(4, 4 + nb[0]))
            num += float((Ua[sl] - Ub[sl]).abs().sum())
            den += float(Ub[sl].abs().sum())
        out[name] = num / max(den, 1e-99)
    return out


def amr_study(steps: int = 100, steps_per_graph: int = 10, l1_steps: int = 20, grid=None,
              only: str | None = None, checkpoint_dir: str | None = None):
    """Config-5 run: three levels (512 x 32^3, 2048 x 16^3, 8192 x 8^3 by
    default) with the four-quadrant data, 100 graph-captured steps in each of
    plain FP64 HLLC, FP64 exact, FP64 with the hybrid switch (fill, flag,
    fill, then HLL next to flagged cells and HLLC elsewhere; the blocks carry
    a shock variable), and mixed precision with the exact solver
    seeded by the learned network (trained on this grid's own face states);
    then l1_steps eager steps of one fixed dt in FP64 exact and in the mixed
    learned form, compared per variable over the leaf cells."""
    import dataclasses
    import torch
    from . import learned as Lr
    from .amr import quadrant_grid
    from .hydro import PREC_MIXED
    grid = quadrant_grid() if grid is None else grid
    gamma = 1.4
    S0 = quadrant_init(grid, gamma)
    zones = grid.zones()
    base = HydroParams(ndim=3, nxb=grid.nb[0][0], nyb=grid.nb[0][1], nzb=grid.nb[0][2],
                       gamma=gamma)
    out = {"grid": {"levels": [sum(1 for L, _ in grid.blocks if L == lev) for lev in range(3)],
                    "block_cells": [list(n) for n in grid.nb], "zones": zones},
           "steps": steps, "steps_per_graph": steps_per_graph}
    # the seed network: one FP64 exact step of the grid, its face states solved
    # and learned (the exact solver's own states)
    t0 = time.perf_counter()
    exact64 = dataclasses.replace(base, RiemannSolver="exact")
    run = AmrRun(grid, exact64, S0.clone())
    run.step(run.dt(0.8))
    faces = run.hydro.face_states()
    del run
    gen = torch.Generator(device=faces.device).manual_seed(7)
    pick = torch.randperm(faces.shape[1], generator=gen, device=faces.device)[:1 << 22]
    Sf = faces[:, pick].contiguous()
    del faces
    ps, _, _ = Lr.solve(Sf, gamma)
    net = Lr.SeedNet(seed=3)
    net.train(net.features(Sf, gamma), Lr.targets(Sf, ps, gamma), steps=1500, batch=8192,
              lr=3e-3, log_every=500)
    strength = Lr.strength(Sf, gamma)
    out["seed"] = {"train_states": int(Sf.shape[1]), "train_s": time.perf_counter() - t0,
                   "faces_at_or_above_gate": float((strength >= 0.2).double().mean())}
    del Sf, ps
    modes = {"fp64_hllc": base, "fp64_exact": exact64,
             "fp64_hybrid": dataclasses.replace(base, RiemannSolver="hybrid", nunk=11, shokVar=10),
             "mixed_exact_learned": dataclasses.replace(base, RiemannSolver="exact",
                                                        precision=PREC_MIXED, learnedSeed=True)}
    for name, hp in modes.items():
        if only is not None and name != only:
            continue
        S = S0.clone()
        if hp.nunk == 11:                  # the same data with a SHOK_VAR slot
            S11 = LevelStorage(grid, 11)
            for a, b in zip(S11.levels, S.levels):
                a[:, :10] = b
            S = S11
        run = AmrRun(grid, hp, S, seed_params=net.params if hp.learnedSeed else None)
        tune = run.hydro.amr_autotune(run.dt(0.8), reps=2)   # launch shapes for this grid
        run.run_graph(0.8, 1, 1)                       # warm-up step (captures once)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run.run_graph(0.8, steps_per_graph, steps // steps_per_graph)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = run.hydro.last_status
        flagged = None
        if hp.nunk == 11:                  # flagged interior cells after the last step's fill
            n_in = sum(t.shape[0] * grid.nb[L][0] * grid.nb[L][1] * grid.nb[L][2]
                       for L, t in enumerate(run.S.levels))
            flagged = sum(float((t[:, 10, 4:-4, 4:-4, 4:-4] > 0).sum())
                          for t in run.S.levels) / n_in
        out[name] = {"zone_updates_per_s": zones * steps / wall, "ms_per_step": 1e3 * wall / steps,
                     "t_end": run.t, "code": st.code, "flagged_cell_fraction": flagged,
                     "last_step_newton_iters": int(st.newton_iters),
                     "last_step_newton_fallbacks": int(st.newton_fallbacks),
                     "autotune": tune}
        run.hydro.autotune_reset()
        del run
        torch.cuda.empty_cache()
    if l1_steps <= 0:
        return out
    # precision: the same fixed dt in FP64 exact and in the mixed learned form
    run = AmrRun(grid, exact64, S0.clone())
    dt = run.dt(0.8)
    for _ in range(l1_steps):
        run.step(dt)
    ref = run.S
    del run
    run = AmrRun(grid, modes["mixed_exact_learned"], S0.clone(), seed_params=net.params)
    for _ in range(l1_steps):
        run.step(dt)
    names = {"dens": DENS, "velx": VELX, "vely": VELY, "velz": VELZ, "pres": PRES, "ener": ENER,
             "eint": EINT}
    out["l1_mixed_learned_vs_fp64_exact"] = {"steps": l1_steps, "dt": dt,
                                            "l1": leaf_l1(run.S, ref, names)}
    # the same two states as FLASH-layout checkpoints through the sfocu
    # comparison (benchmark: FP64 exact); its L1 is the leaf L1 above
    from . import checkpoint as CK
    from .sfocu import compare_checkpoints
    slots = list(names.values())
    ck_mixed = CK.from_level_storage(grid, run.S, slots, list(names), time=l1_steps * dt, dt=dt,
                                     nstep=l1_steps)
    ck_ref = CK.from_level_storage(grid, ref, slots, list(names), time=l1_steps * dt, dt=dt,
                                   nstep=l1_steps)
    rep = compare_checkpoints(ck_mixed, ck_ref, err_tol=1e-6)
    out["l1_mixed_learned_vs_fp64_exact"]["sfocu"] = {
        "blocks_compared": rep["blocks_compared"], "result": "FAILURE" if rep["fail"] else "SUCCESS",
        "err_tol": 1e-6,
        "vars": {n: {"l1": v["ln_norm"], "mag_error": v["mag_error"], "bad_blocks": v["bad_blocks"]}
                 for n, v in rep["vars"].items()}}
    if checkpoint_dir:
        import os
        os.makedirs(checkpoint_dir, exist_ok=True)
        CK.write(os.path.join(checkpoint_dir, "quadrant_mixed_learned_chk.npz"), ck_mixed)
        CK.write(os.path.join(checkpoint_dir, "quadrant_fp64_exact_chk.npz"), ck_ref)
    return out


def main() -> None:
    import argparse
    import json
    ap = argparse.ArgumentParser(description="multi-step runs: Sod (config 1), KH precision "
                                             "study, three-level quadrant problem")
    ap.add_argument("case", choices=["sod", "kh", "amr"])
    ap.add_argument("--mode", choices=["fp64_hllc", "fp64_exact", "fp64_hybrid",
                                       "mixed_exact_learned"],
                    default=None, help="amr: run only this mode")
    ap.add_argument("--l1-steps", type=int, default=20, help="amr: steps of the precision study")
    ap.add_argument("--checkpoint-dir", default=None,
                    help="amr: write the precision study's two end states as FLASH-layout "
                         "checkpoints (python -m flash5_amd.sfocu compares them)")
    ap.add_argument("--ncells", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=200)
    a = ap.parse_args()
    if a.case == "sod":
        import torch
        grid, st = sod()
        # the same grid advanced 60 steps eagerly and 60 steps as replays of
        # 20 graph-captured steps, from the initial state
        rates = {}
        for mode in ("eager", "graph"):
            g2, _ = sod(tmax=0.0)
            zones = g2.U.shape[0] * 64
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode == "eager":
                for _ in range(60):
                    g2.step()
            else:
                g2.run_graph(steps_per_graph=20, replays=3)
            torch.cuda.synchronize()
            rates[mode] = 60 * zones / (time.perf_counter() - t0)
        print(json.dumps({"case": "sod 128^2 16x16 blocks of 8x8", "steps": st.steps,
                          "t": grid.t, "zone_updates_per_s": st.zone_updates_per_s,
                          "fallbacks": st.fallbacks,
                          "first_60_steps_zone_updates_per_s": rates}))
    elif a.case == "amr":
        print(json.dumps(amr_study(a.steps if a.steps != 200 else 100, l1_steps=a.l1_steps,
                                   only=a.mode, checkpoint_dir=a.checkpoint_dir)))
    else:
        print(json.dumps(kh_precision_study(a.ncells, 16, a.steps)))


if __name__ == "__main__":
    main()
