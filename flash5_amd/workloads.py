"""Synthetic block inputs for the hydro path (device-agnostic torch code).

* :func:`fourier_blocks` -- independent 3-D blocks with seeded smooth fields
  (rho = 1 + 0.3 S, p = 1 + 0.2 S, v_d = 0.5 sqrt(gamma) S, each S the mean of
  8 Fourier modes with wavevectors in {0..4}^3 minus 0 and uniform phases),
  evaluated analytically on every cell centre including guard cells, so
  blocks need no guard-cell fill.  Default: the 4096 x 16^3 batch, gamma 5/3.
* :func:`eos_dens_pres` -- gamma-law EOS in MODE_DENS_PRES on whole blocks,
  following physics/Eos/EosMain/Gamma/eos_idealGamma.F90:343-356 and
  Eos_putData.F90:197-210 (fills eint, ener, temp, gamc, game).
* :func:`cfl_dt` -- unsplit/Hydro_computeDt.F90:186-282 over interior cells.

Seeds default to 0x5EED0000 + workload number.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .hydro import DENS, VELX, VELY, VELZ, PRES, ENER, GAMC, GAME, TEMP, EINT, NVAR_USED

SEED_BASE = 0x5EED0000
GAS_CONSTANT = 8.3144598e7


def _modes(rng: np.random.Generator, nblocks: int, nfield: int = 5, nmode: int = 8):
    k = rng.integers(0, 5, size=(nblocks, nfield, nmode, 3))
    zero = (k == 0).all(axis=-1)
    while zero.any():
        k[zero] = rng.integers(0, 5, size=(int(zero.sum()), 3))
        zero = (k == 0).all(axis=-1)
    phi = rng.uniform(0.0, 2.0 * math.pi, size=(nblocks, nfield, nmode))
    return k, phi


def fourier_blocks(nblocks: int = 4096, n: int = 16, ng: int = 4, gamma: float = 5.0 / 3.0,
                   seed: int = SEED_BASE + 2, device="cpu", nunk: int = NVAR_USED,
                   var=tuple(range(NVAR_USED)), chunk: int = 256,
                   block_offset: int = 0) -> torch.Tensor:
    """Batch (nblocks, nunk, n+2ng, n+2ng, n+2ng) float64 of independent 3-D blocks.

    ``block_offset`` skips that many blocks of the seeded sequence, so a
    sub-range of a large batch can be regenerated on its own.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    kk, phi = _modes(rng, block_offset + nblocks)
    kk, phi = kk[block_offset:], phi[block_offset:]
    nt = n + 2 * ng
    x = (torch.arange(nt, dtype=torch.float64, device=device) - ng + 0.5) / n
    U = torch.zeros((nblocks, nunk, nt, nt, nt), dtype=torch.float64, device=device)
    kt = torch.as_tensor(kk, device=device)
    pt = torch.as_tensor(phi, dtype=torch.float64, device=device)
    two_pi = 2.0 * math.pi
    c0 = math.sqrt(gamma)
    for b0 in range(0, nblocks, chunk):
        b1 = min(nblocks, b0 + chunk)
        kb, pb = kt[b0:b1], pt[b0:b1]
        S = torch.zeros((b1 - b0, 5, nt, nt, nt), dtype=torch.float64, device=device)
        for m in range(kb.shape[2]):
            ax = two_pi * kb[:, :, m, 0, None].double() * x            # (b,5,nt)
            ay = two_pi * kb[:, :, m, 1, None].double() * x
            az = two_pi * kb[:, :, m, 2, None].double() * x
            arg = (pb[:, :, m, None, None, None] + az[:, :, :, None, None]
                   + ay[:, :, None, :, None] + ax[:, :, None, None, :])
            S += torch.sin(arg)
        S *= 1.0 / 8.0
        U[b0:b1, var[DENS]] = 1.0 + 0.3 * S[:, 0]
        U[b0:b1, var[PRES]] = 1.0 + 0.2 * S[:, 1]
        U[b0:b1, var[VELX]] = 0.5 * c0 * S[:, 2]
        U[b0:b1, var[VELY]] = 0.5 * c0 * S[:, 3]
        U[b0:b1, var[VELZ]] = 0.5 * c0 * S[:, 4]
    eos_dens_pres(U, gamma, var)
    return U


def eos_dens_pres(U: torch.Tensor, gamma: float, var=tuple(range(NVAR_USED)),
                  gas_constant: float = GAS_CONSTANT, abar: float = 1.0) -> None:
    """MODE_DENS_PRES on every cell of a (..., nunk, nz, ny, nx) tensor, in place."""
    gammam1 = 1.0 / (gamma - 1.0)
    ggprodinv = 1.0 / (gammam1 * gas_constant)
    d, p = U[..., var[DENS], :, :, :], U[..., var[PRES], :, :, :]
    vx, vy, vz = (U[..., var[VELX], :, :, :], U[..., var[VELY], :, :, :],
                  U[..., var[VELZ], :, :, :])
    ekin = 0.5 * (vx * vx + vy * vy + vz * vz)
    eint = p * gammam1 / d
    U[..., var[EINT], :, :, :] = eint
    U[..., var[TEMP], :, :, :] = eint * ggprodinv * abar
    U[..., var[GAMC], :, :, :] = gamma
    U[..., var[ENER], :, :, :] = eint + ekin
    U[..., var[GAME], :, :, :] = p / (eint * d) + 1


def cfl_dt(U: torch.Tensor, ndim: int, ng: int, deltas, cfl: float,
           var=tuple(range(NVAR_USED))) -> float:
    """cfl / max over interior cells of max_d (|v_d| + c)/del_d."""
    sl = [slice(None)] * U.dim()
    for a in range(ndim):
        sl[U.dim() - 1 - a] = slice(ng, U.shape[U.dim() - 1 - a] - ng)
    V = U[tuple(sl)]
    vidx = U.dim() - 4
    def field(v):
        return V.select(vidx, var[v])
    cs = torch.sqrt(field(GAMC) * field(PRES) / field(DENS))
    m = (field(VELX).abs() + cs) / deltas[0]
    if ndim > 1:
        m = torch.maximum(m, (field(VELY).abs() + cs) / deltas[1])
    if ndim > 2:
        m = torch.maximum(m, (field(VELZ).abs() + cs) / deltas[2])
    return float(cfl / m.max())
