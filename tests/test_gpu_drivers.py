"""GPU guard fill and multi-step drivers: guard cells bitwise against the
np.pad reference, 20 Sod steps bitwise against the oracle with the same guard
fill, and the full Sod run against the exact solution."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd import drivers as D
from flash5_amd.hydro import HydroParams
from test_drivers import blocks_to_global, neighbors_with_bc, pad_fill

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ndim,nblk,bc", [
    (2, (3, 2), ("periodic", "outflow")),
    (2, (2, 3), ("reflect", "periodic")),
    (3, (2, 2, 2), ("outflow", "reflect", "periodic"))])
def test_guard_fill_matches_reference(ndim, nblk, bc):
    nb, ng = 6, 4
    rng = np.random.default_rng(2)
    nbk = nblk + (1,) * (3 - len(nblk))
    shape = (int(np.prod(nbk)), 10, nb + 2 * ng if ndim > 2 else 1, nb + 2 * ng, nb + 2 * ng)
    U = rng.normal(size=shape)
    hp = HydroParams(ndim=ndim, nxb=nb, nyb=nb, nzb=nb if ndim > 2 else 1)
    grid = D.UniformGrid(hp, U.copy(), neighbors_with_bc(nblk, ndim, bc), (0.1, 0.1, 0.1))
    grid.fill_guards()
    np.testing.assert_array_equal(grid.U.cpu().numpy(), pad_fill(U, nblk, ndim, ng, bc))


def _oracle_steps(U, nblk, nb, nsteps, p, dx):
    U = U.copy()
    t = 0.0
    for _ in range(nsteps):
        U = pad_fill(U, (nblk, nblk), 2, 4, ("outflow", "outflow"))
        dt = min(O.block_dt(p, U[b], (dx, dx, 1.0), 0.8) for b in range(U.shape[0]))
        for b in range(U.shape[0]):
            rc, _, _, _ = O.compute_fluxes(p, U[b], dt, (dx, dx, 1.0), update=True, want_flux=False)
            assert rc == 0
        t += dt
    return U, t


def test_sod_driver_matches_oracle_bitwise():
    nblk, nb, nsteps = 4, 8, 20
    dx = 1.0 / (nblk * nb)
    U0 = D.sod_blocks(nblk, nb)
    hp = HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, slopeLimiter="mc",
                     use_flattening=True, eintSwitch=1e-4)
    grid = D.UniformGrid(hp, U0.copy(), D.uniform_neighbors((nblk, nblk), (False, False), 2),
                         (dx, dx, 1.0))
    st = grid.run(nsteps=nsteps)
    p = O.params(2, nb, 1.4, limiter=O.MC, eintSwitch=1e-4)
    Uo, to = _oracle_steps(U0, nblk, nb, nsteps, p, dx)
    assert grid.t == to and st.steps == nsteps
    Ug = grid.U.cpu().numpy()
    g = 4
    np.testing.assert_array_equal(Ug[:, :, :, g:g + nb, g:g + nb], Uo[:, :, :, g:g + nb, g:g + nb])


def test_sod_config_run_against_exact_solution():
    """Config 1 (128^2 cells, 16x16 blocks of 8x8) to t = 0.2: density L1
    error against the exact solution along the diagonal."""
    grid, st = D.sod()
    assert abs(grid.t - 0.2) < 1e-12 and st.fallbacks == 0
    nblk, nb, g = 16, 8, 4
    G = blocks_to_global(grid.U.cpu().numpy(), (nblk, nblk), g, 2)
    n = nblk * nb
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc)
    s = (X + Y - 1.0) / math.sqrt(2.0)                  # signed distance, right state > 0
    exact = np.vectorize(lambda v: O.exact_sample(1.4, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1, v)[0])
    rho_ex = exact(s / grid.t)
    err = np.abs(G[D.DENS, 0] - rho_ex).mean()
    print("Sod 128^2 t=0.2: steps", st.steps, "L1(dens)", err, "zone-updates/s",
          st.zone_updates_per_s)
    assert err < 2e-2, err


def _sod_grid(nblk=4, nb=8, **kw):
    dx = 1.0 / (nblk * nb)
    hp = HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, slopeLimiter="mc",
                     use_flattening=True, eintSwitch=1e-4, **kw)
    return D.UniformGrid(hp, D.sod_blocks(nblk, nb), D.uniform_neighbors((nblk, nblk),
                                                                       (False, False), 2),
                         (dx, dx, 1.0))


@pytest.mark.parametrize("kw", [{}, {"RiemannSolver": "exact"}], ids=["hllc", "exact"])
def test_graph_captured_steps_match_eager_steps(kw):
    """Every replay of the captured steps against the same number of eager
    steps, with eager work and synchronisation between the replays."""
    eager = _sod_grid(**kw)
    graph = _sod_grid(**kw)
    for rep in range(3):
        t = graph.run_graph(steps_per_graph=3, replays=2)
        for _ in range(6):
            eager.step()
        torch.cuda.synchronize()
        junk = torch.randn(1000, device="cuda").sum()          # unrelated eager work
        torch.cuda.synchronize()
        assert t > 0 and graph.t == pytest.approx(eager.t, rel=0, abs=1e-15)
        np.testing.assert_array_equal(graph.U.cpu().numpy()[:, :, :, 4:12, 4:12],
                                      eager.U.cpu().numpy()[:, :, :, 4:12, 4:12],
                                      err_msg=f"replay round {rep}")
        del junk


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_compute_dt_keeps_the_first_cell_of_a_rounding_tie(ndim):
    """Blocks with two cells whose (|v| + c) / dx times the CFL round alike:
    the device takes the earlier cell of each block, as the reference's loop
    does (Hydro_computeDt.F90:264-272), and the smallest block dt, bit for
    bit with the oracle; alone and in a batch."""
    from flash5_amd.hydro import Hydro
    from test_oracle_block import tie_block
    p = O.params(ndim, 8, 1.4)
    U1, va, vb = tie_block(p, 0.8, first_smaller=True)
    U2, _, _ = tie_block(p, 0.8, first_smaller=False)
    dels = (1.0 / 8,) * 3
    hy = Hydro(HydroParams(ndim=ndim, nxb=8, nyb=8, nzb=8, gamma=1.4))
    want = [O.block_dt(p, U, dels, 0.8) for U in (U1, U2)]
    assert want[0] == 0.8 / va and want[0] != 0.8 / vb
    for U, w in zip((U1, U2), want):
        assert hy.compute_dt(torch.from_numpy(U[None].copy()).cuda(), dels, 0.8) == w
    both = torch.from_numpy(np.stack([U1, U2])).cuda()
    assert hy.compute_dt(both, dels, 0.8) == min(want)
