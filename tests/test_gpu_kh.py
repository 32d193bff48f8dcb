"""Config 3 (Kelvin-Helmholtz shear layer) on the GPU against the CPU oracle.

A 128^2 periodic grid of 8 x 8 blocks of 16^2 cells with the config's
initial data (drivers.kh_blocks) is stepped 20 times:

* FP64, eager and graph-captured: bit for bit the oracle loop (periodic
  guard fill by np.pad, CFL dt over all blocks, block update), the same dt
  sequence and end time;
* the mixed mode (FP32 reconstruction and fluxes, FP64 update), stepped with
  the oracle's dt sequence: per-variable sfocu L1 against the FP64 oracle,
  density within the 1e-6 bar.

The full 2048^2 x 200-step study against the oracle is kh_precision_oracle.py
in this directory.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as O
from flash5_amd import drivers as D, sfocu
from flash5_amd.hydro import PREC_MIXED, HydroParams
from test_drivers import pad_fill

pytestmark = pytest.mark.gpu

NC, NB, STEPS = 128, 16, 20
NBLK = NC // NB
DX = 1.0 / NC
DELTAS = (DX, DX, 1.0)
BC = ("periodic", "periodic")
NAMES = {"dens": D.DENS, "velx": D.VELX, "vely": D.VELY, "pres": D.PRES, "ener": D.ENER,
         "eint": D.EINT, "temp": D.TEMP}
# mixed mode against FP64 after 20 steps: the density bar is the 1e-6 one.
# vely is a 0.01-amplitude perturbation beside O(1) velocities, so the FP32
# rounding of the face states (about 1e-7 absolute) is about 1e-5 of it.
MIXED_L1 = {"dens": 1e-6, "velx": 1e-6, "vely": 5e-5, "pres": 1e-6, "ener": 1e-6, "eint": 1e-6,
            "temp": 1e-6}

_cache = {}


def _oracle():
    """20 oracle steps from the KH data: (interior-compared U, dts, t)."""
    if "run" not in _cache:
        p = O.params(2, NB, 1.4)
        U = D.kh_blocks(NC, NB)
        dts = []
        for _ in range(STEPS):
            U = pad_fill(U, (NBLK, NBLK), 2, 4, BC)
            dt = min(O.block_dt(p, U[b], DELTAS, 0.8) for b in range(U.shape[0]))
            rc, _ = O.compute_fluxes_batch(p, U, dt, DELTAS)
            assert rc == 0
            dts.append(dt)
        t = 0.0
        for dt in dts:
            t += dt
        _cache["run"] = (U, dts, t)
    return _cache["run"]


def _grid(**kw):
    hp = HydroParams(ndim=2, nxb=NB, nyb=NB, nzb=1, gamma=1.4, **kw)
    return D.UniformGrid(hp, D.kh_blocks(NC, NB), D.uniform_neighbors((NBLK, NBLK), (True, True), 2),
                         DELTAS)


def _interior(U):
    return U[:, :, :, 4:4 + NB, 4:4 + NB]


def test_kh_fp64_steps_match_oracle_bitwise():
    Uo, dts, to = _oracle()
    grid = _grid()
    got = [grid.step() for _ in range(STEPS)]
    assert got == dts and grid.t == to
    np.testing.assert_array_equal(_interior(grid.U.cpu().numpy()), _interior(Uo))


def test_kh_graph_captured_fp64_steps_match_oracle_bitwise():
    Uo, dts, to = _oracle()
    grid = _grid()
    t = 0.0
    for _ in range(2):
        t += grid.run_graph(steps_per_graph=5, replays=2)
        torch.cuda.synchronize()
    assert grid.nsteps == STEPS
    assert t == pytest.approx(to, rel=1e-15)
    np.testing.assert_array_equal(_interior(grid.U.cpu().numpy()), _interior(Uo))


def test_kh_mixed_precision_l1_against_oracle():
    Uo, dts, _ = _oracle()
    grid = _grid(precision=PREC_MIXED)
    for dt in dts:
        grid.fill_guards()
        st = grid.hydro.advance_batch(grid.U, dt, DELTAS)
        assert st.code == 0
    norms = sfocu.compare(grid.U.cpu().numpy(), Uo, 1, NAMES, 2, 4)
    print("KH 128^2 x 20 steps, mixed vs FP64 oracle L1:",
          {k: f"{v[0]:.2e}" for k, v in norms.items()})
    for k, (l1, _) in norms.items():
        assert l1 <= MIXED_L1[k], f"{k}: L1 {l1:.3e}"
