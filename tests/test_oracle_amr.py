"""Pins of the CPU restatement of the multi-level grid operations.

The interpolator (AMReX's cell-conservative linear, not in the reference
tree) is pinned by what it must satisfy: the children of a coarse cell
average back to it in conserved form, linear data is reproduced exactly,
constants stay constant, and restriction after prolongation is the identity.
The flux-register face list is pinned against the octree face list of
BlockMesh (itself pinned by TestFluxCorrection's cases) on equal-size levels.
The EOS restatement is pinned against the oracle's C EOS.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import oracle as O
import oracle.amr as OA
from flash5_amd.amr import BlockMesh, LevelGrid, quadrant_grid

GAMMA = 1.4


def _grid2d(periodic=(False, False, False)):
    """Level 0: 2 x 2 blocks of 16^2; level 1 (8^2 blocks) under the lower
    row; level 2 (8^2) under two level-1 blocks in its middle."""
    return LevelGrid.nested(2, [(16, 16), (8, 8), (8, 8)], (2, 2),
                            [lambda p: p[1] == 0, lambda p: p[0] in (3, 4) and p[1] == 1],
                            periodic=periodic)


def _state(grid, seed, fn=None):
    """Per-block states: smooth random fields (or fn(x, y, z) -> dens, vx,
    vy, vz, pres) with consistent EOS."""
    rng = np.random.default_rng(seed)
    kx, ky, kz = rng.integers(1, 3, size=3)
    ph = rng.uniform(0, 2 * np.pi, size=5)
    c = OA.eos_consts(GAMMA)
    U = []
    for b, (L, _) in enumerate(grid.blocks):
        x, y, z = grid.cell_centers(b)
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        X, Y, Z = X.transpose(2, 1, 0), Y.transpose(2, 1, 0), Z.transpose(2, 1, 0)
        if fn is None:
            s = [np.sin(2 * np.pi * (kx * X + ky * Y + kz * Z) + ph[q]) for q in range(5)]
            dens, pres = 1.0 + 0.3 * s[0], 1.0 + 0.2 * s[4]
            vx, vy, vz = 0.3 * s[1], 0.3 * s[2], 0.3 * s[3]
        else:
            dens, vx, vy, vz, pres = fn(X, Y, Z)
        A = np.zeros(grid.block_shape(L, O.NVAR_USED))
        eint = pres / (dens * (GAMMA - 1.0))
        A[O.DENS], A[O.VELX], A[O.VELY], A[O.VELZ] = dens, vx, vy, vz
        A[O.EINT] = eint
        A[O.ENER] = eint + 0.5 * (vx * vx + vy * vy + vz * vz)
        pr, tp, en, ei, ge = OA.eos_dens_ei(c, dens, vx, vy, vz, A[O.EINT], A[O.ENER])
        A[O.PRES], A[O.TEMP], A[O.ENER], A[O.EINT], A[O.GAME] = pr, tp, en, ei, ge
        A[O.GAMC] = GAMMA
        U.append(A)
    return U


def _children(grid, L):
    """Global fine cells (3, n) of level L under every cell of level L-1,
    ordered so that consecutive 2^d cells share a parent."""
    d = grid.as_dict()
    idx = OA._index(d)
    cells = []
    for b, (Lb, pos) in enumerate(grid.blocks):
        if Lb != L - 1:
            continue
        nb = grid.nb[L - 1]
        for k, j, i in itertools.product(*[range(nb[a]) for a in (2, 1, 0)]):
            gc = [pos[0] * nb[0] + i, pos[1] * nb[1] + j, pos[2] * nb[2] + k]
            for oz, oy, ox in itertools.product(range(2 if grid.ndim > 2 else 1),
                                                range(2 if grid.ndim > 1 else 1), range(2)):
                cells.append((2 * gc[0] + ox, 2 * gc[1] + oy, 2 * gc[2] + oz if grid.ndim > 2 else 0))
    return np.array(cells, dtype=np.int64).T, idx, d


def test_level_grid_counts():
    g = quadrant_grid()
    counts = [sum(1 for L, _ in g.blocks if L == lev) for lev in range(3)]
    assert counts == [512, 2048, 8192]
    assert g.zones() == 512 * 32 ** 3 + 2048 * 16 ** 3 + 8192 * 8 ** 3


def test_prolongation_conserves_and_keeps_constants():
    grid = _grid2d()
    U = _state(grid, 3)
    g, idx, d = _children(grid, 1)
    vals = OA.prolong(d, U, idx, 1, g, GAMMA)
    nchild = 4
    # conserved form of the children averages back to the parent
    for v in range(O.NVAR_USED):
        if v in (O.PRES, O.TEMP, O.GAME, O.GAMC, O.EINT):
            continue       # set by the EOS after interpolation (eint = ener - ekin)
        q = vals[v] * vals[O.DENS] if v in OA.PER_MASS else vals[v]
        q = q.reshape(-1, nchild).mean(axis=1)
        gc = g[:, ::nchild] // 2
        cid, cloc = OA._lookup(d, idx, 0, gc)
        p = OA._gather(U, cid, cloc, v, True)
        if v == O.ENER:    # the EOS re-forms ener from eint + ekin
            np.testing.assert_allclose(q, p, rtol=1e-12)
        else:
            np.testing.assert_allclose(q, p, rtol=1e-14, atol=1e-15)
    # constants stay constant
    Uc = _state(grid, 3, fn=lambda X, Y, Z: (1.5 + 0 * X, 0.25 + 0 * X, -0.5 + 0 * X,
                                             0 * X, 2.0 + 0 * X))
    vc = OA.prolong(d, Uc, idx, 1, g, GAMMA)
    assert np.all(vc[O.DENS] == 1.5) and np.all(vc[O.VELX] == 0.25) and np.all(vc[O.VELY] == -0.5)


def test_prolongation_reproduces_linear_data():
    """Dyadic linear density (exact in binary): fine values are exact."""
    grid = _grid2d()
    U = _state(grid, 5, fn=lambda X, Y, Z: (2.0 + 0.5 * X + 0.25 * Y, 0 * X, 0 * X, 0 * X,
                                            1.0 + 0 * X))
    g, idx, d = _children(grid, 1)
    # interior parents only (the limiter flattens slopes at the clamped domain edge)
    dx = grid.dels(1)
    x = (g[0] + 0.5) * dx[0]
    y = (g[1] + 0.5) * dx[1]
    vals = OA.prolong(d, U, idx, 1, g, GAMMA)
    n0, n1 = grid.ext[0][0] * grid.nb[0][0], grid.ext[0][1] * grid.nb[0][1]
    inner = (g[0] // 2 > 0) & (g[0] // 2 < n0 - 1) & (g[1] // 2 > 0) & (g[1] // 2 < n1 - 1)
    np.testing.assert_array_equal(vals[O.DENS][inner], (2.0 + 0.5 * x + 0.25 * y)[inner])


def test_restriction_after_prolongation_is_identity():
    grid = _grid2d()
    U = _state(grid, 7)
    d = grid.as_dict()
    idx = OA._index(d)
    coarse = [u.copy() for u in U]
    # fill the level-1 interiors from level 0, then restrict back
    for b, (L, pos) in enumerate(grid.blocks):
        if L != 1:
            continue
        nb = grid.nb[1]
        j, i = np.indices((nb[1], nb[0])).reshape(2, -1)
        g = np.array([pos[0] * nb[0] + i, pos[1] * nb[1] + j, 0 * i])
        vals = OA.prolong(d, U, idx, 1, g, GAMMA)
        U[b][:, 0, j + 4, i + 4] = vals
    OA.restrict({**d, "nb": d["nb"][:2], "ext": d["ext"][:2],
                 "blocks": [b for b in d["blocks"] if b[0] < 2]}, U[:len([b for b in d["blocks"] if b[0] < 2])])
    for b, (L, pos) in enumerate(grid.blocks):
        if L != 0 or pos[1] != 0:
            continue
        for v in (O.DENS, O.VELX, O.VELY, O.ENER):   # eint is re-formed by the EOS
            np.testing.assert_allclose(U[b][v, 0, 4:-4, 4:-4], coarse[b][v, 0, 4:-4, 4:-4],
                                       rtol=1e-12 if v == O.ENER else 1e-14, atol=1e-15)


@pytest.mark.parametrize("periodic", [False, True])
def test_register_faces_match_octree_faces(periodic):
    """Equal block sizes (8 children per refined block): the face list equals
    BlockMesh's, whose cases follow TestFluxCorrection."""
    per = (periodic, periodic, False)
    mesh = BlockMesh(ndim=2, nb=8, base=(2, 2), periodic=per)
    mesh.refine(1, (0, 0, 0))
    mesh.refine(1, (1, 0, 0))
    want = mesh.register_faces(1)
    grid = LevelGrid.nested(2, [(8, 8), (8, 8)], (2, 2), [lambda p: p[1] == 0], periodic=per)
    got = OA.register_faces(grid.as_dict(), 0)
    # renumber: BlockMesh tiles are per level, sorted by (bk, bj, bi)
    tile = {b: mesh.blocks(L + 1).index(tuple(pos)) for b, (L, pos) in enumerate(grid.blocks)}
    renum = sorted((tile[f[0]], tile[f[1]]) + tuple(f[2:]) for f in got)
    assert len(want) > 0
    assert renum == sorted(want)


def test_eos_restatement_matches_oracle_eos():
    p = O.params(3, 8, GAMMA)
    rng = np.random.default_rng(11)
    U = np.zeros(O.block_shape(p))
    U[O.DENS] = rng.uniform(0.2, 2.0, U.shape[1:])
    for v in (O.VELX, O.VELY, O.VELZ):
        U[v] = rng.uniform(-1, 1, U.shape[1:])
    U[O.EINT] = rng.uniform(0.5, 3.0, U.shape[1:])
    U[O.ENER] = U[O.EINT] + 0.5 * (U[O.VELX] ** 2 + U[O.VELY] ** 2 + U[O.VELZ] ** 2)
    U[O.GAMC] = U[O.GAME] = GAMMA
    V = U.copy()
    O.eos(p, U, 2)
    c = OA.eos_consts(GAMMA)
    pr, tp, en, ei, ge = OA.eos_dens_ei(c, V[O.DENS], V[O.VELX], V[O.VELY], V[O.VELZ], V[O.EINT],
                                        V[O.ENER])
    for v, x in ((O.PRES, pr), (O.TEMP, tp), (O.ENER, en), (O.EINT, ei), (O.GAME, ge)):
        np.testing.assert_array_equal(U[v], x)


def _grid3d_small():
    """Level 0: 2^3 blocks of 8^3 (cells 1/16); level 1 (8^3) under block (0,0,0)."""
    return LevelGrid.nested(3, [(8, 8, 8), (8, 8, 8)], (2, 2, 2), [lambda p: p == (0, 0, 0)])


def _coarse_hood(d, idx, U, gc, v):
    """min and max of variable v (conserved form) over the 3^ndim
    neighbourhood of coarse cells gc, brute force."""
    ndim = d["ndim"]
    lo = hi = None
    for off in itertools.product(*[(-1, 0, 1)] * ndim):
        gn = gc.copy()
        for a in range(ndim):
            gn[a] += off[a]
        ids, loc = OA._lookup(d, idx, 0, OA._domain(d, 0, gn))
        x = OA._gather(U, ids, loc, v, True)
        lo = x if lo is None else np.minimum(lo, x)
        hi = x if hi is None else np.maximum(hi, x)
    return lo, hi


def test_prolongation_bound_lowers_slopes_in_3d():
    """Hand-computed AMReX case: coarse cell (5,5,5) holds 1, its -x/-y/-z
    neighbours 0.5, its +x/+y/+z neighbours 1.25, every other cell 1.  Each
    MC slope is min(2*0.25, 2*0.5, 0.375) = 0.375; the corner child would
    reach 1 + 3*0.375/4 = 1.28125 > 1.25, so alpha = 0.25/0.28125 = 8/9 and
    the corner child lands on the neighbourhood maximum 1.25."""
    grid = _grid3d_small()
    n0 = grid.ext[0][0] * grid.nb[0][0]

    def fn(X, Y, Z):
        ic = [np.floor(c * n0).astype(int) for c in (X, Y, Z)]
        dens = np.ones(X.shape)
        for a in range(3):
            other = [ic[b] == 5 for b in range(3) if b != a]
            dens = np.where(other[0] & other[1] & (ic[a] == 4), 0.5, dens)
            dens = np.where(other[0] & other[1] & (ic[a] == 6), 1.25, dens)
        return dens, 0 * X, 0 * X, 0 * X, 1.0 + 0 * X

    U = _state(grid, 1, fn=fn)
    d = grid.as_dict()
    idx = OA._index(d)
    kids = np.array([[10 + ox, 10 + oy, 10 + oz] for oz in (0, 1) for oy in (0, 1)
                     for ox in (0, 1)], dtype=np.int64).T
    vals = OA.prolong(d, U, idx, 1, kids, GAMMA)[O.DENS]
    alpha = 0.25 / 0.28125
    s = 0.375 * alpha
    assert vals[-1] == pytest.approx(1.25, rel=0, abs=5e-16)      # (+,+,+) child at the bound
    assert vals[-1] <= 1.25
    assert vals[0] == pytest.approx(1.0 - 0.75 * s, abs=5e-16)    # (-,-,-) child
    assert vals.mean() == pytest.approx(1.0, abs=1e-15)
    plain = 1.0 + 3 * 0.25 * 0.375                                 # unbounded MC corner child
    assert plain > 1.25


@pytest.mark.parametrize("ndim", [2, 3])
def test_prolongation_stays_inside_the_neighbourhood(ndim):
    """Rough random data: every child (conserved form) lies inside the min and
    max of its parent's 3^ndim neighbourhood.  In 3-D the bound binds for
    some parents (alpha < 1), and there the extreme child sits on the bound;
    in 2-D it never binds (MC slopes alone keep the children inside)."""
    grid = _grid2d() if ndim == 2 else _grid3d_small()
    rng = np.random.default_rng(11 + ndim)

    def fn(X, Y, Z):
        return (1.0 + rng.uniform(0, 1, X.shape), rng.uniform(-1, 1, X.shape),
                rng.uniform(-1, 1, X.shape), rng.uniform(-1, 1, X.shape),
                1.0 + rng.uniform(0, 1, X.shape))

    U = _state(grid, 0, fn=fn)
    g, idx, d = _children(grid, 1)
    vals = OA.prolong(d, U, idx, 1, g, GAMMA)
    gc = g // 2
    nchild = 2 ** ndim
    bound = 0
    for v in (O.DENS, O.VELX, O.VELY, O.VELZ):
        q = vals[v] * vals[O.DENS] if v in OA.PER_MASS else vals[v]
        lo, hi = _coarse_hood(d, idx, U, gc, v)
        tol = 4e-16 * np.maximum(np.abs(lo), np.abs(hi))
        assert np.all(q >= lo - tol) and np.all(q <= hi + tol), f"var {v}"
        # the factor per parent, from the slopes of its axis neighbours
        par = gc[:, ::nchild]
        cid, cloc = OA._lookup(d, idx, 0, par)
        q0 = OA._gather(U, cid, cloc, v, True)
        slopes = []
        for a in range(ndim):
            m, p_ = par.copy(), par.copy()
            m[a] -= 1
            p_[a] += 1
            um = OA._gather(U, *OA._lookup(d, idx, 0, OA._domain(d, 0, m)), v, True)
            up = OA._gather(U, *OA._lookup(d, idx, 0, OA._domain(d, 0, p_)), v, True)
            slopes.append(OA._mcslope(q0, um, up))
        plo, phi = lo[::nchild], hi[::nchild]
        alpha = OA._bound_factor(q0, slopes, plo, phi)
        lim = alpha < 1.0
        bound += int(lim.sum())
        qc = q.reshape(-1, nchild)
        ext = np.where(np.abs(qc.max(axis=1) - phi) < np.abs(qc.min(axis=1) - plo),
                       qc.max(axis=1) - phi, qc.min(axis=1) - plo)
        assert np.all(np.abs(ext[lim]) <= 8e-16 * np.abs(phi[lim]) + 1e-15), f"var {v}"
        # children average back to the parent
        np.testing.assert_allclose(qc.mean(axis=1), q0, rtol=1e-14, atol=1e-15)
    if ndim == 3:
        assert bound > 0
    else:
        assert bound == 0


def test_mcslope_matches_the_mc_limiter_on_a_table():
    """AMReX's slope equals the textbook MC limiter of the one-sided
    differences, (sgn a + sgn b)/2 min(2|a|, 2|b|, |a+b|/2), on a table
    (up to the rounding of the central difference)."""
    vals = np.array([-2.0, -1.0, -0.25, 0.0, 0.3, 1.0, 3.0])
    for um, u0, up in itertools.product(vals, vals, vals):
        a, b = u0 - um, up - u0
        want = 0.5 * (np.sign(a) + np.sign(b)) * min(2 * abs(a), 2 * abs(b), abs(a + b) / 2)
        got = OA._mcslope(np.array([u0]), np.array([um]), np.array([up]))[0]
        assert got == pytest.approx(want, abs=1e-15), (um, u0, up)
