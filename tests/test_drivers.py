"""Host logic of the multi-step drivers (CPU): neighbour tables, FLASH's Sod
initial condition, and an independent guard-fill reference (whole-grid
np.pad per axis, then cut into blocks) used by the GPU driver tests."""
import math

import numpy as np
import pytest

from flash5_amd import drivers as D


def blocks_to_global(U, nblk, ng, ndim):
    """Interior cells of all blocks as one array [var][Z][Y][X]."""
    nbk = nblk + (1,) * (3 - len(nblk))
    _, nunk, nz, ny, nx = U.shape
    bz, by, bx = (nz - 2 * ng if ndim > 2 else 1), (ny - 2 * ng if ndim > 1 else 1), nx - 2 * ng
    G = np.zeros((nunk, bz * nbk[2], by * nbk[1], bx * nbk[0]))
    gz, gy = (ng if ndim > 2 else 0), (ng if ndim > 1 else 0)
    for bk in range(nbk[2]):
        for bj in range(nbk[1]):
            for bi in range(nbk[0]):
                b = bi + nbk[0] * (bj + nbk[1] * bk)
                G[:, bk * bz:(bk + 1) * bz, bj * by:(bj + 1) * by, bi * bx:(bi + 1) * bx] = \
                    U[b, :, gz:gz + bz, gy:gy + by, ng:ng + bx]
    return G


def pad_fill(U, nblk, ndim, ng, bc):
    """Reference guard fill: pad the global interior along x, then y, then z
    ('wrap' periodic, 'edge' outflow, 'symmetric' reflect with the normal
    velocity negated in the new layers), then cut every block's window."""
    G = blocks_to_global(U, nblk, ng, ndim)
    for a in range(ndim):
        ax = 3 - a
        pad = [(0, 0)] * 4
        pad[ax] = (ng, ng)
        mode = {"periodic": "wrap", "outflow": "edge", "reflect": "symmetric"}[bc[a]]
        G = np.pad(G, pad, mode=mode)
        if bc[a] == "reflect":
            sl_lo = [slice(None)] * 4
            sl_hi = [slice(None)] * 4
            sl_lo[ax] = slice(0, ng)
            sl_hi[ax] = slice(G.shape[ax] - ng, G.shape[ax])
            for sl in (sl_lo, sl_hi):
                G[(D.VELX + a,) + tuple(sl[1:])] *= -1
    out = U.copy()
    nbk = nblk + (1,) * (3 - len(nblk))
    _, nunk, nz, ny, nx = U.shape
    bz, by, bx = (nz - 2 * ng if ndim > 2 else 1), (ny - 2 * ng if ndim > 1 else 1), nx - 2 * ng
    for bk in range(nbk[2]):
        for bj in range(nbk[1]):
            for bi in range(nbk[0]):
                b = bi + nbk[0] * (bj + nbk[1] * bk)
                z0, y0, x0 = bk * bz, bj * by, bi * bx
                out[b] = G[:, z0:z0 + nz, y0:y0 + ny, x0:x0 + nx]
    return out


def neighbors_with_bc(nblk, ndim, bc):
    nb = D.uniform_neighbors(nblk, [b == "periodic" for b in bc], ndim)
    if "reflect" in bc:
        nbk = tuple(nblk) + (1,) * (3 - len(nblk))
        for b in range(nb.shape[0]):
            idx = (b % nbk[0], (b // nbk[0]) % nbk[1], b // (nbk[0] * nbk[1]))
            for a in range(ndim):
                if bc[a] == "reflect":
                    if idx[a] == 0:
                        nb[b, 2 * a] = D.BC_REFLECT
                    if idx[a] == nbk[a] - 1:
                        nb[b, 2 * a + 1] = D.BC_REFLECT
    return nb


def test_uniform_neighbors():
    nb = D.uniform_neighbors((3, 2), (True, False), 2)
    assert nb.shape == (6, 6)
    assert list(nb[0]) == [2, 1, -1, 3, -1, -1]        # x wraps, y outflow below
    assert list(nb[5]) == [4, 3, 2, -1, -1, -1]
    nb3 = D.uniform_neighbors((2, 2, 2), (True, True, True), 3)
    assert (nb3 >= 0).all()


def test_sod_initial_condition_follows_the_diagonal():
    nblk, nb, ng = 4, 8, 4
    U = D.sod_blocks(nblk, nb)
    h = 1.0 / (nblk * nb)
    for b in range(nblk * nblk):
        bi, bj = b % nblk, b // nblk
        for j in range(ng, ng + nb):
            for i in range(ng, ng + nb):
                xl, yc = (bi * nb + i - ng) * h, (bj * nb + j - ng + 0.5) * h
                cut = 1.0 - yc
                rho = U[b, D.DENS, 0, j, i]
                if xl + h <= cut - 1e-12:
                    assert rho == 1.0
                elif xl >= cut + 1e-12:
                    assert rho == 0.125
                else:
                    assert rho == pytest.approx(0.5625)
    assert U[:, D.GAMC].min() == 1.4
    np.testing.assert_allclose(U[:, D.EINT], U[:, D.PRES] / (0.4 * U[:, D.DENS]))


def test_pad_fill_reference_is_consistent():
    """Periodic padding of a periodic field equals the field itself shifted."""
    rng = np.random.default_rng(1)
    nblk, ng, nb = (3, 2), 4, 6
    U = rng.normal(size=(6, 10, 1, nb + 2 * ng, nb + 2 * ng))
    F = pad_fill(U, nblk, 2, ng, ("periodic", "periodic"))
    G = blocks_to_global(U, nblk, ng, 2)
    # block 0's low x guards are the last ng columns of the global grid
    np.testing.assert_array_equal(F[0, :, 0, ng:ng + nb, :ng], G[:, 0, :nb, -ng:])
