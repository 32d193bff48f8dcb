"""GPU parity: libfx_hydro.so (HIP, gfx950) against the CPU FP64 oracle.

Each test runs the same seeded blocks through the C ABI on the GPU and
through the oracle on the CPU and compares every output variable over the
interior with sfocu norms.  The FP64 path (FX_PREC_FP64, built without
contraction) must agree to 1e-12 relative L1 per variable; the stricter
bitwise check is reported and asserted where the operation order is shared.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd import sfocu, workloads as W
from flash5_amd.hydro import Hydro, HydroParams, Tile

pytestmark = pytest.mark.gpu

NAMES = {"dens": O.DENS, "velx": O.VELX, "vely": O.VELY, "velz": O.VELZ, "pres": O.PRES,
         "ener": O.ENER, "eint": O.EINT, "temp": O.TEMP, "game": O.GAME, "gamc": O.GAMC}
TOL = 1e-12


def _oracle_params(hp: HydroParams):
    kw = dict(limiter={"vanleer": O.VANLEER, "mc": O.MC, "minmod": O.MINMOD,
                       "hybrid": O.HYBRID, "limited": O.LIMITED}[hp.slopeLimiter.lower()],
              charLimiting=int(hp.charLimiting), flattening=int(hp.use_flattening),
              steepening=int(hp.use_steepening), use3dFullCTU=int(hp.use_3dFullCTU),
              useAuxEint=int(hp.use_auxEintEqn), eintSwitch=hp.eintSwitch,
              riemann={"hllc": O.HLLC, "exact": O.EXACT}[hp.RiemannSolver.lower()],
              exactMaxIter=hp.exactMaxIter, exactTol=hp.exactTol)
    return O.params(hp.ndim, hp.nxb, hp.gamma, **kw)


def _blocks(hp: HydroParams, nblocks: int, seed: int):
    U3 = W.fourier_blocks(nblocks, n=hp.nxb, gamma=hp.gamma, seed=seed).numpy()
    if hp.ndim == 3:
        return U3
    k0 = hp.ng + hp.nxb // 3
    U2 = np.ascontiguousarray(U3[:, :, k0:k0 + 1, :, :])
    if hp.ndim == 1:
        j0 = hp.ng + hp.nxb // 2
        U2 = np.ascontiguousarray(U2[:, :, :, j0:j0 + 1, :])
    return U2


def _run_both(hp: HydroParams, U: np.ndarray, dt: float, dx: float, path: int = 0,
              stats: list | None = None):
    deltas = (dx, dx, dx)
    po = _oracle_params(hp)
    Uo = U.copy()
    for b in range(U.shape[0]):
        rc, _, _, st = O.compute_fluxes(po, Uo[b], dt, deltas, update=True, want_flux=False)
        assert rc == 0
        if stats is not None:
            stats.append(st)
    hy = Hydro(hp)
    hy.set_kernel_path(path)
    try:
        Ug = torch.from_numpy(U.copy()).cuda()
        st = hy.advance_batch(Ug, dt, deltas)
    finally:
        hy.set_kernel_path(0)
    assert st.code == 0
    if stats is not None:
        stats.append(st)
    return Uo, Ug.cpu().numpy()


def _check(hp, Uo, Ug, bitwise):
    norms = sfocu.compare(Ug, Uo, 1, NAMES, hp.ndim, hp.ng)
    for n, (l1, linf) in norms.items():
        assert l1 <= TOL, f"{n}: L1 {l1:.3e} Linf {linf:.3e}"
    if bitwise:
        diff = np.argwhere(Ug != Uo)
        assert diff.size == 0, f"{len(diff)} values differ, first {diff[:4]}"
    return norms


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "three-kernel"])
@pytest.mark.parametrize("ndim,n", [(3, 16), (3, 8), (3, 32), (2, 16), (2, 8), (1, 16)])
def test_fp64_block_update_matches_oracle(ndim, n, path):
    gamma = 5.0 / 3.0
    hp = HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=gamma)
    nb = {3: 6, 2: 12, 1: 12}[ndim] if n < 32 else 2
    U = _blocks(hp, nb, seed=W.SEED_BASE + 2)
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), ndim, 4, (dx,) * 3, 0.8)
    Uo, Ug = _run_both(hp, U, dt, dx, path)
    _check(hp, Uo, Ug, bitwise=True)


@pytest.mark.parametrize("lim", ["mc", "minmod", "hybrid", "limited"])
def test_fp64_limiter_variants(lim):
    hp = HydroParams(ndim=3, nxb=8, nyb=8, nzb=8, gamma=1.4, slopeLimiter=lim,
                     use_steepening=True, eintSwitch=1e-4)
    U = _blocks(hp, 4, seed=W.SEED_BASE + 7)
    dx = 1.0 / 8
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    Uo, Ug = _run_both(hp, U, dt, dx)
    _check(hp, Uo, Ug, bitwise=True)


def test_fp64_primitive_limiting_no_ctu():
    hp = HydroParams(ndim=3, nxb=8, nyb=8, nzb=8, gamma=1.4, charLimiting=False,
                     use_3dFullCTU=False, use_flattening=False)
    U = _blocks(hp, 4, seed=W.SEED_BASE + 9)
    dx = 1.0 / 8
    dt = 0.4 * W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    Uo, Ug = _run_both(hp, U, dt, dx)
    _check(hp, Uo, Ug, bitwise=True)


def test_fluxes_then_update_solution_matches_oracle():
    """hy_computeFluxes(Uout=null) + hy_updateSolution path, with a flux edit
    in between standing in for a flux-register overwrite."""
    hp = HydroParams(ndim=3, nxb=8, nyb=8, nzb=8, gamma=5.0 / 3.0)
    U = _blocks(hp, 3, seed=W.SEED_BASE + 11)
    dx = 1.0 / 8
    deltas = (dx,) * 3
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, deltas, 0.8)
    po = _oracle_params(hp)
    hy = Hydro(hp)
    tiles, Uo = [], U.copy()
    ofl = []
    for b in range(U.shape[0]):
        rc, fl, pavg, _ = O.compute_fluxes(po, Uo[b], dt, deltas, update=False, want_pavg=True)
        assert rc == 0
        ofl.append((fl, pavg))
        t = Tile(U=torch.from_numpy(U[b].copy()).cuda(), deltas=deltas,
                 flux=[torch.zeros(hp.flux_shape(d), dtype=torch.float64, device="cuda")
                       for d in range(3)],
                 pavg=torch.zeros(hp.interior_shape(), dtype=torch.float64, device="cuda"))
        tiles.append(t)
    hy.compute_fluxes(tiles, dt, 0.0, update_in_place=[False] * len(tiles))
    for b, t in enumerate(tiles):
        fl, pavg = ofl[b]
        for d in range(3):
            np.testing.assert_array_equal(t.flux[d].cpu().numpy(), fl[d])
        np.testing.assert_array_equal(t.pavg.cpu().numpy(), pavg)
        # edit one face like a register overwrite, identically on both sides
        fl[0][:, :, :, 0] *= 0.75
        t.flux[0][:, :, :, 0] *= 0.75
        assert O.update_solution(po, Uo[b], fl, pavg, dt, deltas) == 0
    hy.update_solution(tiles, dt)
    Ug = np.stack([t.U.cpu().numpy() for t in tiles])
    _check(hp, Uo, Ug, bitwise=True)


@pytest.mark.parametrize("path", [0, 1, 2], ids=["fused", "three-kernel", "separate-update"])
@pytest.mark.parametrize("ndim,n", [(3, 8), (3, 16), (2, 16)])
def test_fp64_exact_riemann_solver_matches_oracle(ndim, n, path):
    """Exact solver in place of HLLC.  The device takes powers as
    exp2(e log2 x) and the oracle calls libm pow; the two round differently
    in the last places, so the bar is the 1e-12 L1 one, not bitwise;
    iteration counts must agree closely."""
    hp = HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=1.4, RiemannSolver="exact")
    U = _blocks(hp, 4 if ndim == 3 else 8, seed=W.SEED_BASE + 13)
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), ndim, 4, (dx,) * 3, 0.8)
    stats = []
    Uo, Ug = _run_both(hp, U, dt, dx, path, stats)
    _check(hp, Uo, Ug, bitwise=False)
    o_iters = sum(st.exactIters for st in stats[:-1])
    g = stats[-1]
    assert g.newton_iters > 0 and g.newton_fallbacks == 0
    assert abs(g.newton_iters - o_iters) <= 1e-3 * o_iters, (g.newton_iters, o_iters)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 16, 8), (16, 8, 32)])
def test_exact_solve_with_update_matches_separate_update(prec, shape):
    """Exact-solver batch steps of 3-D blocks up to 16 x 16 across run the
    face solve and the update as one kernel (path 0); path 2 keeps the face
    solve and the update apart.  Same fluxes, same update order: bitwise."""
    from flash5_amd.hydro import PREC_MIXED
    nx, ny, nz = shape
    hp = HydroParams(ndim=3, nxb=nx, nyb=ny, nzb=nz, gamma=1.4, RiemannSolver="exact",
                     precision=PREC_MIXED if prec == "mixed" else 0)
    # non-cubic blocks: a corner of cubic ones, ghost cells included
    g = hp.ng
    U = W.fourier_blocks(5, n=max(shape), gamma=hp.gamma, seed=W.SEED_BASE + 17).numpy()
    U = np.ascontiguousarray(U[:, :, :nz + 2 * g, :ny + 2 * g, :nx + 2 * g])
    dx = 1.0 / 16
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    out, iters = [], []
    hy = Hydro(hp)
    try:
        for path in (0, 2):
            hy.set_kernel_path(path)
            Ug = torch.from_numpy(U.copy()).cuda()
            st = hy.advance_batch(Ug, dt, (dx,) * 3)
            assert st.code == 0
            out.append(Ug.cpu().numpy())
            iters.append(st.newton_iters)
    finally:
        hy.set_kernel_path(0)
    assert iters[0] == iters[1] and iters[0] > 0
    diff = np.argwhere(out[0] != out[1])
    assert diff.size == 0, f"{len(diff)} values differ, first {diff[:4]}"
    assert not np.array_equal(out[0], U)


@pytest.mark.parametrize("riemann", ["hllc", "exact"])
@pytest.mark.parametrize("ndim,n", [(3, 16), (2, 16)])
def test_mixed_precision_l1_against_fp64_oracle(ndim, n, riemann):
    """FP32 reconstruction and fluxes (HLLC, or the exact solver with its
    FP64 Newton iterate and convergence test) with the FP64 update: the L1
    relative error of every variable against the FP64 oracle after one step
    of the batch workload, density at most 1e-6."""
    from flash5_amd.hydro import PREC_MIXED
    hp = HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=5.0 / 3.0, precision=PREC_MIXED,
                     RiemannSolver=riemann)
    U = _blocks(hp, 6 if ndim == 3 else 12, seed=W.SEED_BASE + 2)
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), ndim, 4, (dx,) * 3, 0.8)
    stats = []
    Uo, Ug = _run_both(hp, U, dt, dx, stats=stats)
    if riemann == "exact":
        g = stats[-1]
        assert g.newton_iters > 0 and g.newton_fallbacks == 0
    norms = sfocu.compare(Ug, Uo, 1, NAMES, ndim, 4)
    print({k: f"{v[0]:.2e}" for k, v in norms.items()})
    assert norms["dens"][0] <= 1e-6, norms["dens"]
    for name in ("velx", "vely", "velz", "pres", "ener", "eint"):
        assert norms[name][0] <= 1e-5, (name, norms[name])
    assert not np.array_equal(Ug, Uo)          # the FP32 path really ran
