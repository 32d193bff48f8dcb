"""hy_fallbackLowerCFL on the GPU.

A step in which a reconstructed state falls back to first order reports the
cap hy_cflFallbackFactor / NDIM in fx_status.fallback_cfl (2-D and 3-D; the
reference never copies the lowered cell CFL into hy_cfl in 1-D,
hy_getRiemannState.F90:934-941).  The drivers carry hy_cfl from step to step
as the reference does: lowered after a step with a fallback, raised by 1.2 per
step up to the original (Hydro.F90:128-141).  A 2-D blast whose shock front
drives states negative runs eagerly bit for bit against the oracle stepped
with the reference's CFL sequence, and graph-captured runs (uniform and
three-level) equal their eager runs.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd import drivers as D
from flash5_amd.hydro import Hydro, HydroParams

from test_oracle_fallback import random_block

pytestmark = pytest.mark.gpu

G = 1.4


def _bitwise(a, b, what):
    diff = np.argwhere(a != b)
    assert diff.size == 0, f"{what}: {len(diff)} values differ, first {diff[:4]}"


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_status_reports_the_fallback_cap(ndim):
    n = 8
    p = O.params(ndim, n, G)
    U = np.stack([random_block(p, 700 + ndim + b) for b in range(2)])
    dt = min(O.block_dt(p, U[b], (1.0 / n,) * 3, 0.8) for b in range(2))
    for lower, factor in ((False, 0.9), (True, 0.9), (True, 0.6)):
        hy = Hydro(HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=G,
                               hy_fallbackLowerCFL=lower, hy_cflFallbackFactor=factor))
        st = hy.advance_batch(torch.from_numpy(U.copy()).cuda(), dt, (1.0 / n,) * 3)
        assert st.code == 0 and st.fallbacks > 0
        want = factor / ndim if (lower and ndim > 1) else 0.0
        assert st.fallback_cfl == want, (lower, factor, st.fallback_cfl)
    # smooth data: no fallback, no cap
    smooth = np.stack([random_block(p, 800 + b, lo=-0.05) for b in range(2)])
    for b in range(2):
        for a in range(ndim):
            smooth[b, O.VELX + a] *= 0.05
        O.eos(p, smooth[b], 1)
    st = hy.advance_batch(torch.from_numpy(smooth).cuda(), dt * 1e-3, (1.0 / n,) * 3)
    assert st.code == 0 and st.fallbacks == 0 and st.fallback_cfl == 0.0


NBLK, NB = 4, 8
DX = 1.0 / (NBLK * NB)


def _blast_blocks():
    """Pressure 1e4 inside r = 0.1 of the centre, 1e-6 outside, density 1,
    at rest; blocks in the order of uniform_neighbors."""
    p = O.params(2, NB, G)
    shp = O.block_shape(p)
    U = np.zeros((NBLK * NBLK,) + shp)
    nz, ny, nx = shp[1:]
    _, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    for b in range(NBLK * NBLK):
        bi, bj = b % NBLK, b // NBLK
        X = (bi * NB + x - 4 + 0.5) * DX
        Y = (bj * NB + y - 4 + 0.5) * DX
        r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
        U[b, O.DENS] = 1.0
        U[b, O.PRES] = np.where(r2 < 0.01, 1e4, 1e-6)
        O.eos(p, U[b], 1)
    return p, U


def _blast_grid(lower=True):
    _, U = _blast_blocks()
    hp = HydroParams(ndim=2, nxb=NB, nyb=NB, nzb=1, gamma=G, hy_fallbackLowerCFL=lower)
    return D.UniformGrid(hp, U, D.uniform_neighbors((NBLK, NBLK), (False, False), 2),
                         (DX, DX, 1.0))


def test_fallback_lower_cfl_driver_matches_oracle_bitwise():
    """Ten steps of fill, dt, update.  The oracle's hy_cfl follows the
    reference: recovery by 1.2 up to 0.8 before the fluxes, then
    min(hy_cfl, 0.9 / 2) after a step with a fallback."""
    from test_drivers import pad_fill
    nsteps = 10
    grid = _blast_grid()
    dts = [grid.step() for _ in range(nsteps)]
    p, U = _blast_blocks()
    t, hy_cfl, cfls, fbs = 0.0, 0.8, [], []
    for _ in range(nsteps):
        U = pad_fill(U, (NBLK, NBLK), 2, 4, ("outflow", "outflow"))
        dt = min(O.block_dt(p, U[b], (DX, DX, 1.0), hy_cfl) for b in range(U.shape[0]))
        cfls.append(hy_cfl)
        fb = 0
        for b in range(U.shape[0]):
            rc, _, _, st = O.compute_fluxes(p, U[b], dt, (DX, DX, 1.0), update=True,
                                            want_flux=False)
            assert rc == 0
            fb += st.fallbacks
        fbs.append(fb)
        t += dt
        hy_cfl = 1.2 * hy_cfl if 1.2 * hy_cfl < 0.8 else 0.8
        if fb:
            hy_cfl = min(hy_cfl, 0.9 / 2)
    print("oracle fallbacks per step", fbs, "CFL per step", cfls)
    assert sum(fbs) > 0 and min(cfls) == 0.45 and cfls[-1] > 0.45
    assert grid.t == t
    torch.cuda.synchronize()
    g = 4
    _bitwise(grid.U.cpu().numpy()[:, :, :, g:g + NB, g:g + NB], U[:, :, :, g:g + NB, g:g + NB],
             "blast")
    plain = _blast_grid(lower=False)
    assert [plain.step() for _ in range(nsteps)] != dts


def test_fallback_lower_cfl_graph_matches_eager():
    """Eight graph-captured steps (four per graph, two replays) against eight
    eager steps: the lowered CFL and the fallback count of the previous step
    are carried on the device."""
    eager, graph = _blast_grid(), _blast_grid()
    t = graph.run_graph(steps_per_graph=4, replays=2)
    for _ in range(8):
        eager.step()
    plain = _blast_grid(lower=False)
    for _ in range(8):
        plain.step()
    torch.cuda.synchronize()
    assert t > 0 and graph.t == pytest.approx(eager.t, rel=0, abs=1e-15)
    assert plain.t != eager.t
    _bitwise(graph.U.cpu().numpy()[:, :, :, 4:4 + NB, 4:4 + NB],
             eager.U.cpu().numpy()[:, :, :, 4:4 + NB, 4:4 + NB], "graph")


def test_fallback_lower_cfl_amr_run_driver_matches_graph():
    """The blast on the three-level 2-D grid under hy_fallbackLowerCFL:
    AmrRun's eager dt()/step() loop against its graph-captured steps."""
    from flash5_amd.drivers import AmrRun, LevelStorage
    from test_oracle_amr import GAMMA, _grid2d, _state

    grid = _grid2d()
    blast = lambda X, Y, Z: (1.0 + 0 * X, 0 * X, 0 * X, 0 * X,
                             np.where((X - 0.5) ** 2 + (Y - 0.5) ** 2 < 0.01, 1e4, 1e-6))
    U = _state(grid, 43, fn=blast)
    hp = HydroParams(ndim=2, nxb=16, nyb=16, nzb=1, gamma=GAMMA, hy_fallbackLowerCFL=True)

    def run():
        S = LevelStorage(grid)
        for b in range(len(grid.blocks)):
            S.block(b).copy_(torch.from_numpy(U[b]))
        return S, AmrRun(grid, hp, S)

    S1, eager = run()
    fbs, cfls = [], []
    for _ in range(6):
        st = eager.step(eager.dt(0.8))
        assert st.code == 0
        fbs.append(int(st.fallbacks))
        cfls.append(eager._cfl_next)
    print("fallbacks per step", fbs, "next CFL", cfls)
    assert sum(fbs) > 0 and min(cfls) == 0.45
    e = [S1.block(b).cpu().numpy() for b in range(len(grid.blocks))]
    S2, graph = run()
    graph.run_graph(0.8, 3, 2)
    assert graph.t == eager.t
    for b in range(len(grid.blocks)):
        _bitwise(S2.block(b).cpu().numpy(), e[b], f"block {b}")
