"""GPU parity of the first-order positivity fallback.

Blocks of log-uniform densities and pressures with Mach-3 velocities drive
reconstructed states negative after the 2-D CTU corrections
(hy_dataReconstOneStep.F90:1117-1154) and, in 3-D, after the corner terms
(hy_getRiemannState.F90:1037-1051).  The GPU step must equal the CPU oracle
bit for bit there, on both kernel paths, and report the same number of
resets: one per (state cell, direction), a state cell on a column-tile seam
counted once.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd.hydro import Hydro, HydroParams, Tile

from test_oracle_fallback import random_block

pytestmark = pytest.mark.gpu

G = 1.4


def _batch(ndim, n, nblocks, seed):
    p = O.params(ndim, n, G)
    return p, np.stack([random_block(p, seed + b) for b in range(nblocks)])


def _dt(p, U, dx):
    return min(O.block_dt(p, U[b], (dx,) * 3, 0.8) for b in range(U.shape[0]))


@pytest.mark.parametrize("path", [0, 1], ids=["fused", "three-kernel"])
@pytest.mark.parametrize("ndim,n", [(1, 16), (2, 16), (2, 8), (3, 16), (3, 8)])
def test_fallback_step_matches_oracle_bitwise(ndim, n, path):
    if ndim == 1 and path == 0:
        pytest.skip("1-D blocks take the three-kernel path")
    p, U = _batch(ndim, n, 3, 500 + 10 * ndim + n)
    dx = 1.0 / n
    dt = _dt(p, U, dx)
    Uo = U.copy()
    fb = 0
    for b in range(U.shape[0]):
        rc, _, _, st = O.compute_fluxes(p, Uo[b], dt, (dx,) * 3, update=True, want_flux=False)
        assert rc == 0
        fb += st.fallbacks
    assert fb > 0
    hy = Hydro(HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=G))
    hy.set_kernel_path(path)
    try:
        Ug = torch.from_numpy(U.copy()).cuda()
        st = hy.advance_batch(Ug, dt, (dx,) * 3)
    finally:
        hy.set_kernel_path(0)
    assert st.code == 0
    assert st.fallbacks == fb, (st.fallbacks, fb)
    Ug = Ug.cpu().numpy()
    diff = np.argwhere(Ug != Uo)
    assert diff.size == 0, f"{len(diff)} values differ, first {diff[:4]}"


def test_fallback_fluxes_then_update_solution_match_oracle():
    """The fluxes-only call (hy_computeFluxes with Uout null) and
    hy_updateSolution on fallback data: fluxes, pavg and U bit for bit."""
    ndim, n = 3, 8
    p, U = _batch(ndim, n, 2, 900)
    dx = 1.0 / n
    dt = _dt(p, U, dx)
    hp = HydroParams(ndim=ndim, nxb=n, nyb=n, nzb=n, gamma=G)
    hy = Hydro(hp)
    tiles = [Tile(U=torch.from_numpy(U[b].copy()).cuda(), deltas=(dx,) * 3,
                  flux=[torch.zeros(hp.flux_shape(d), dtype=torch.float64, device="cuda")
                        for d in range(3)],
                  pavg=torch.zeros(hp.interior_shape(), dtype=torch.float64, device="cuda"))
             for b in range(U.shape[0])]
    st = hy.compute_fluxes(tiles, dt, 0.0, update_in_place=[False] * len(tiles))
    assert st.code == 0 and st.fallbacks > 0
    assert hy.update_solution(tiles, dt).code == 0
    fb = 0
    for b, t in enumerate(tiles):
        Uo = U[b].copy()
        rc, fl, pavg, so = O.compute_fluxes(p, Uo, dt, (dx,) * 3, update=False, want_pavg=True)
        assert rc == 0
        fb += so.fallbacks
        for d in range(ndim):
            np.testing.assert_array_equal(t.flux[d].cpu().numpy(), fl[d])
        np.testing.assert_array_equal(t.pavg.cpu().numpy(), pavg)
        assert O.update_solution(p, Uo, fl, pavg, dt, (dx,) * 3) == 0
        np.testing.assert_array_equal(t.U.cpu().numpy(), Uo)
    assert st.fallbacks == fb
