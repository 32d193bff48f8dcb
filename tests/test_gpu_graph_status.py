"""Status of graph-captured runs: one status for the whole run.

The captured steps keep the first error of any step and sum the fallback
counters over every step of every replay (fx_grid_run_graph), where an eager
caller would see each step's status in turn.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd import drivers as D
from flash5_amd.hydro import FxError, HydroParams

from test_oracle_fallback import random_block

pytestmark = pytest.mark.gpu

NB, NBLK = 16, 2
DX = 1.0 / (NB * NBLK)
DELTAS = (DX, DX, 1.0)


def _grid(U):
    hp = HydroParams(ndim=2, nxb=NB, nyb=NB, nzb=1, gamma=1.4)
    return D.UniformGrid(hp, U.copy(), D.uniform_neighbors((NBLK, NBLK), (True, True), 2), DELTAS)


def _fallback_blocks():
    p = O.params(2, NB, 1.4)
    return np.stack([random_block(p, 700 + b, lo=-2.0) for b in range(NBLK * NBLK)])


def test_graph_run_sums_fallbacks_over_its_steps():
    U = _fallback_blocks()
    eager = _grid(U)
    per_step = []
    for _ in range(3):
        eager.step()
        per_step.append(int(eager.hydro.last_status.fallbacks))
    assert min(per_step) > 0
    graph = _grid(U)
    graph.run_graph(steps_per_graph=3, replays=1)
    st = graph.last_status
    assert st.code == 0
    assert st.fallbacks == sum(per_step), (st.fallbacks, per_step)
    np.testing.assert_array_equal(graph.U.cpu().numpy()[:, :, :, 4:4 + NB, 4:4 + NB],
                                  eager.U.cpu().numpy()[:, :, :, 4:4 + NB, 4:4 + NB])


def test_graph_run_reports_an_error_of_its_first_step():
    """A zero pressure in the middle of block 3 gives an imaginary-sound-speed
    (or Riemann) error in the first captured step; the run returns it with
    that block as its location."""
    U = D.kh_blocks(NB * NBLK, NB)
    c = 4 + NB // 2
    U[3, D.PRES, 0, c, c] = 0.0
    U[3, D.EINT, 0, c, c] = 0.0
    U[3, D.ENER, 0, c, c] = 0.5 * (U[3, D.VELX, 0, c, c] ** 2 + U[3, D.VELY, 0, c, c] ** 2)
    grid = _grid(U)
    with pytest.raises(FxError) as e:
        grid.run_graph(steps_per_graph=3, replays=1)
    torch.cuda.synchronize()
    st = grid.last_status
    assert st.code in (1, 2)
    assert st.first_tile == 3
