"""CPU restatement of the multi-level grid operations (test infrastructure).

Checker only: tests and the bench's CPU baseline call it; the product path
never does.  FP64 numpy, in the operation order of the device kernels, so
that a step of a multi-level grid can be compared bit for bit.

Reference semantics (paths relative to source/Grid/GridMain/AMR/Amrex):
  guard fill     Grid_fillGuardCells.F90:287-369 -- copies from the same
                 level, amrex_fillpatch from the next coarser level with the
                 CELL_CONSERVATIVE_LINEAR interpolator (Config:39-40) on data
                 in conserved form (gr_primitiveToConserve.F90:107-116), then
                 EOS on the guard cells (Hydro.F90:121; Eos_wrapped MODE_DENS_EI,
                 Eos_getData.F90:189-215, eos_idealGamma.F90:323-333).
  restriction    gr_restrictAllLevels.F90:86-130 (PER_MASS variables times
                 density, amrex_average_down finest to coarsest, divide back).
  flux register  Grid_putFluxData.F90:51-80 / Grid_conserveFluxes.F90:45-108
                 for each level pair, finest pair first; the face list covers
                 every coarse copy of a face on the boundary of the finer level
                 (TestFluxCorrection's Driver_evolveFlash.F90 states the cases).
AMReX is not part of the reference tree.  FLASH maps CELL_CONSERVATIVE_LINEAR
to amrex_interp_cell_cons (RuntimeParameters_mapStrToInt.F90:482-485), the
cell-conservative linear interpolator without AMReX's cross-component
("lincc") limiting.  Its published algorithm (AMReX_Interp_{2,3}D_C.H,
mf_cell_cons_lin_interp_mcslope and mf_cell_cons_lin_interp) is restated in
prolong: per component and axis the MC slope
    s = sign(dc) min(|dc|, min(|df|, |db|))  (0 unless df db >= 0),
    dc = (u+ - u-)/2, df = 2 (u+ - u0), db = 2 (u0 - u-),
then one factor alpha per component and coarse cell that keeps every fine
value inside the min / max of the coarse cell's 3^ndim neighbourhood:
    dumax = sum_a |s_a| (r-1)/(2r)  (1/4 at refinement ratio r = 2),
    alpha = 1, lowered to (umax - u0)/dumax, then to (u0 - umin)/dumax,
    where dumax alpha exceeds the room above, then below,
and the fine value u0 + sum_a off_a alpha s_a at the offsets off = +-1/4.
Pinned by (tests/test_oracle_amr.py): children averaging back to their
parent, linear data reproduced exactly, constants kept, restriction after
prolongation being the identity, every fine value inside its neighbourhood
bounds with equality at the limited extreme, and a hand-computed 3-D case
where the bound lowers the slopes (plain MC would overshoot).  In 1-D and
2-D the factor never binds: an MC slope is at most twice either one-sided
difference, so the neighbourhood always has room for sum_a |s_a|/4.

A grid is a dict: ndim, ng, nb[L] (cells per block along x, y, z on level
L, 0-based), ext[L] (blocks per axis covering the domain on level L),
blocks [(L, (px, py, pz)), ...], periodic (3 flags), del[L] (cell widths);
U is a list of per-block arrays (nunk, nz, ny, nx), guard cells included.
"""
from __future__ import annotations

import numpy as np

import oracle as O

PER_MASS = (O.VELX, O.VELY, O.VELZ, O.ENER, O.EINT)   # hydro unsplit Config:44-53


def _index(grid):
    """Dense block index per level: idx[L][bz, by, bx] = block or -1."""
    out = []
    for L, ext in enumerate(grid["ext"]):
        a = -np.ones((ext[2], ext[1], ext[0]), dtype=np.int64)
        out.append(a)
    for b, (L, pos) in enumerate(grid["blocks"]):
        out[L][pos[2], pos[1], pos[0]] = b
    return out


def _domain(grid, L, g):
    """Global cells g (3, n) on level L wrapped (periodic) or clamped."""
    g = g.copy()
    for a in range(grid["ndim"]):
        n = grid["ext"][L][a] * grid["nb"][L][a]
        if grid["periodic"][a]:
            g[a] = np.mod(g[a], n)
        else:
            g[a] = np.clip(g[a], 0, n - 1)
    return g


def _lookup(grid, idx, L, g):
    """Block ids and local guard-inclusive (i, j, k) of global cells g."""
    nb, ng, ndim = grid["nb"][L], grid["ng"], grid["ndim"]
    bp = [g[a] // nb[a] for a in range(3)]
    ids = idx[L][bp[2], bp[1], bp[0]]
    loc = [g[a] - bp[a] * nb[a] + (ng if a < ndim else 0) for a in range(3)]
    return ids, loc


def _gather(U, ids, loc, v, cons):
    """Variable v at the cells (conserved form when cons)."""
    out = np.empty(ids.shape[0])
    for b in np.unique(ids):
        m = ids == b
        x = U[b][v, loc[2][m], loc[1][m], loc[0][m]]
        if cons and v in PER_MASS:
            x = U[b][O.DENS, loc[2][m], loc[1][m], loc[0][m]] * x
        out[m] = x
    return out


def _mcslope(u0, um, up):
    """AMReX's MC slope of one axis (mf_cell_cons_lin_interp_mcslope)."""
    dc = 0.5 * (up - um)
    df = 2.0 * (up - u0)
    db = 2.0 * (u0 - um)
    s = np.where(df * db >= 0.0, np.minimum(np.abs(df), np.abs(db)), 0.0)
    return np.copysign(1.0, dc) * np.minimum(s, np.abs(dc))


def _bound_factor(u0, slopes, umin, umax):
    """AMReX's alpha: the largest factor <= 1 keeping u0 +- sum |s_a|/4 inside
    [umin, umax], in the published operation order."""
    dumax = np.zeros_like(u0)
    for s in slopes:
        dumax = dumax + np.abs(s) * 0.25
    alpha = np.ones_like(u0)
    live = dumax > 0.0
    safe = np.where(live, dumax, 1.0)
    hi = live & (dumax * alpha > (umax - u0))
    alpha = np.where(hi, (umax - u0) / safe, alpha)
    lo = live & (dumax * alpha > (u0 - umin))
    alpha = np.where(lo, (u0 - umin) / safe, alpha)
    return alpha


def eos_consts(gamma, gas_constant=8.3144598e7):
    gammam1 = 1.0 / (gamma - 1.0)
    ggprod = gammam1 * gas_constant
    return dict(gamma=gamma, gam1inv=1.0 / gammam1, ggprodinv=1.0 / ggprod)


def eos_dens_ei(c, dens, vx, vy, vz, eint_in, ener_in, small_e=1e-10, eint_switch=0.0, abar=1.0):
    """MODE_DENS_EI on arrays (Eos_getData.F90:189-215, eos_idealGamma.F90:323-333,
    Eos_putData.F90:197-210): returns pres, temp, ener, eint, game."""
    ekin = 0.5 * (vx * vx + vy * vy + vz * vz)
    ie = np.where(ener_in - ekin > np.maximum(small_e, eint_switch * ekin), ener_in - ekin, eint_in)
    ie = np.maximum(ie, small_e)
    pres = dens * ie * c["gam1inv"]
    temp = ie * c["ggprodinv"] * abar
    ener = ie + ekin
    game = pres / (ie * dens) + 1
    return pres, temp, ener, ie, game


def prolong(grid, U, idx, L, g, gamma, nunk=O.NVAR_USED, **eos_kw):
    """Values (nunk, n) at global cells g (3, n) of level L from level L - 1:
    AMReX's cell-conservative linear interpolation (MC slopes bounded by the
    3^ndim neighbourhood) in conserved form, per-mass form back, then
    MODE_DENS_EI EOS."""
    ndim = grid["ndim"]
    c = eos_consts(gamma)
    gc = g // 2
    off = g - 2 * gc
    cid, cloc = _lookup(grid, idx, L - 1, gc)
    assert (cid >= 0).all(), "levels not properly nested"
    nbr = []
    for a in range(ndim):
        lo, hi = gc.copy(), gc.copy()
        lo[a] -= 1
        hi[a] += 1
        lo, hi = _domain(grid, L - 1, lo), _domain(grid, L - 1, hi)
        nbr.append((_lookup(grid, idx, L - 1, lo), _lookup(grid, idx, L - 1, hi)))
    hood = []                         # the 3^ndim neighbourhood, centre included
    for d in np.ndindex(*([3] * ndim)):
        gn = gc.copy()
        for a in range(ndim):
            gn[a] += d[a] - 1
        hood.append(_lookup(grid, idx, L - 1, _domain(grid, L - 1, gn)))

    def interp(v):
        q0 = _gather(U, cid, cloc, v, True)
        slopes = []
        for a in range(ndim):
            (mi, ml), (pi, pl) = nbr[a]
            slopes.append(_mcslope(q0, _gather(U, mi, ml, v, True), _gather(U, pi, pl, v, True)))
        umin, umax = q0.copy(), q0.copy()
        for hi_, hl in hood:
            x = _gather(U, hi_, hl, v, True)
            umin, umax = np.minimum(umin, x), np.maximum(umax, x)
        alpha = _bound_factor(q0, slopes, umin, umax)
        q = q0
        for a in range(ndim):
            q = q + np.where(off[a] == 1, 0.25, -0.25) * (slopes[a] * alpha)
        return q
    vals = np.empty((nunk, g.shape[1]))
    dens = interp(O.DENS)
    for v in range(nunk):
        q = interp(v)
        vals[v] = q / dens if v in PER_MASS else q
    pres, temp, ener, eint, game = eos_dens_ei(c, dens, vals[O.VELX], vals[O.VELY], vals[O.VELZ],
                                               vals[O.EINT], vals[O.ENER], **eos_kw)
    vals[O.PRES], vals[O.TEMP], vals[O.ENER], vals[O.EINT] = pres, temp, ener, eint
    vals[O.GAMC], vals[O.GAME] = gamma, game
    return vals


def fill_guards(grid, U, gamma, nunk=O.NVAR_USED, **eos_kw):
    """Guard cells of every block, in place."""
    idx = _index(grid)
    ndim, ng = grid["ndim"], grid["ng"]
    new = []
    for b, (L, pos) in enumerate(grid["blocks"]):
        sh = U[b].shape[1:]
        k, j, i = np.indices(sh).reshape(3, -1)
        loc = [i, j, k]
        nb = grid["nb"][L]
        guard = np.zeros(i.shape, dtype=bool)
        for a in range(ndim):
            guard |= (loc[a] < ng) | (loc[a] >= ng + nb[a])
        li, lj, lk = i[guard], j[guard], k[guard]
        gl = [li, lj, lk]
        g = np.zeros((3, li.size), dtype=np.int64)
        for a in range(ndim):
            g[a] = pos[a] * nb[a] + gl[a] - ng
        g = _domain(grid, L, g)
        ids, sloc = _lookup(grid, idx, L, g)
        vals = np.empty((nunk, li.size))
        have = ids >= 0
        for v in range(nunk):
            if have.any():
                vals[v, have] = _gather(U, ids[have], [s[have] for s in sloc], v, False)
        miss = ~have
        if miss.any():
            assert L > 0, "a coarsest-level block has no neighbour"
            vals[:, miss] = prolong(grid, U, idx, L, g[:, miss], gamma, nunk, **eos_kw)
        new.append((b, lk, lj, li, vals))
    for b, lk, lj, li, vals in new:      # every source cell is interior: no order dependence
        U[b][:nunk, lk, lj, li] = vals


def restrict(grid, U, nunk=O.NVAR_USED):
    """Covered cells of each level from their children, finest level first."""
    idx = _index(grid)
    ndim, ng = grid["ndim"], grid["ng"]
    nlev = len(grid["nb"])
    for L in range(nlev - 1, 0, -1):
        for b, (Lb, pos) in enumerate(grid["blocks"]):
            if Lb != L:
                continue
            nb = grid["nb"][L]
            h = [nb[a] // 2 if a < ndim else 1 for a in range(3)]
            cz, cy, cx = np.indices((h[2], h[1], h[0])).reshape(3, -1)
            c = [cx, cy, cz]
            G = np.zeros((3, cx.size), dtype=np.int64)
            for a in range(ndim):
                G[a] = pos[a] * h[a] + c[a]
            pid, ploc = _lookup(grid, idx, L - 1, G)
            assert (pid >= 0).all()
            nz, ny = (2 if ndim > 2 else 1), (2 if ndim > 1 else 1)
            frac = 1.0 / float(2 * ny * nz)
            F = U[b]

            def child_sum(v):
                s = np.zeros(cx.size)
                for oz in range(nz):
                    for oy in range(ny):
                        for ox in range(2):
                            fi = 2 * cx + ox + ng
                            fj = 2 * cy + oy + (ng if ndim > 1 else 0)
                            fk = 2 * cz + oz + (ng if ndim > 2 else 0)
                            x = F[v, fk, fj, fi]
                            if v in PER_MASS:
                                x = F[O.DENS, fk, fj, fi] * x
                            s = s + x
                return s
            dens = child_sum(O.DENS) * frac
            for v in range(nunk):
                avg = child_sum(v) * frac
                val = avg / dens if v in PER_MASS else avg
                for p in np.unique(pid):
                    m = pid == p
                    U[p][v, ploc[2][m], ploc[1][m], ploc[0][m]] = val[m]


def register_faces(grid, L):
    """Flux-register faces of level pair (L, L+1) as 9-tuples (coarse block,
    fine block, dir, ci, cj, ck, fi, fj, fk), block numbers into grid
    ["blocks"]: every coarse copy of every face on the finer level's
    boundary, periodic wrap included, none on a non-periodic domain face."""
    idx = _index(grid)
    ndim = grid["ndim"]
    C, Fn = grid["nb"][L], grid["nb"][L + 1]
    Ce, Fe = grid["ext"][L], grid["ext"][L + 1]
    out = []
    for fb, (Lb, pos) in enumerate(grid["blocks"]):
        if Lb != L + 1:
            continue
        for d in range(ndim):
            for side in (0, 1):
                npos = list(pos)
                npos[d] += 1 if side else -1
                if not 0 <= npos[d] < Fe[d]:
                    if not grid["periodic"][d]:
                        continue
                    npos[d] %= Fe[d]
                if idx[L + 1][npos[2], npos[1], npos[0]] >= 0:
                    continue
                ncell = Ce[d] * C[d]
                gcf = (pos[d] * Fn[d] + (Fn[d] if side else 0)) // 2
                t1, t2 = (1 if d == 0 else 0), (1 if d == 2 else 2)
                n1 = Fn[t1] // 2 if t1 < ndim else 1
                n2 = Fn[t2] // 2 if t2 < ndim else 1
                for c2 in range(n2):
                    for c1 in range(n1):
                        gct = [0, 0, 0]
                        if t1 < ndim:
                            gct[t1] = pos[t1] * (Fn[t1] // 2) + c1
                        if t2 < ndim:
                            gct[t2] = pos[t2] * (Fn[t2] // 2) + c2
                        fi = [0, 0, 0]
                        fi[d] = Fn[d] if side else 0
                        if t1 < ndim:
                            fi[t1] = 2 * c1
                        if t2 < ndim:
                            fi[t2] = 2 * c2
                        last = -1
                        for k in (0, 1):
                            gc = list(gct)
                            gc[d] = gcf if k else gcf - 1
                            if not 0 <= gc[d] < ncell:
                                if not grid["periodic"][d]:
                                    continue
                                gc[d] %= ncell
                            bp = [gc[a] // C[a] for a in range(3)]
                            cb = int(idx[L][bp[2], bp[1], bp[0]])
                            assert cb >= 0, "the coarse level does not cover the face"
                            if cb == last:
                                continue
                            last = cb
                            ci = [gc[a] - bp[a] * C[a] for a in range(3)]
                            if not k:
                                ci[d] += 1
                            out.append((cb, fb, d, ci[0], ci[1], ci[2], fi[0], fi[1], fi[2]))
    return out


def level_params(grid, L, gamma, **kw):
    nb = grid["nb"][L]
    assert nb[0] == nb[1] or grid["ndim"] < 2
    return O.params(grid["ndim"], nb[0], gamma, **kw)


def step(grid, U, dt, gamma, **kw):
    """One step of every level: guard fill, per-block fluxes, flux register
    (finest pair first), per-block update, restriction.  With shock
    detection (kw shockDetect, or riemann = HYBR): fill, flag every block,
    fill again so that SHOK_VAR's guard cells hold their neighbours' flags
    (Hydro.F90:94-123)."""
    nunk = kw.get("nunk", O.NVAR_USED)
    fill_guards(grid, U, gamma, nunk)
    nlev = len(grid["nb"])
    params = [level_params(grid, L, gamma, **kw) for L in range(nlev)]
    if kw.get("shockDetect") or kw.get("riemann") == O.HYBR:
        for b, (L, _) in enumerate(grid["blocks"]):
            assert O.shock_detect(params[L], U[b]) >= 0
        fill_guards(grid, U, gamma, nunk)
    fluxes, pavgs = [], []
    for b, (L, _) in enumerate(grid["blocks"]):
        rc, fl, pavg, _ = O.compute_fluxes(params[L], U[b], dt, grid["del"][L], update=False,
                                           want_pavg=True)
        assert rc == 0
        fluxes.append(fl)
        pavgs.append(pavg)
    for L in range(nlev - 2, -1, -1):
        faces = register_faces(grid, L)
        if faces:
            reg = O.fluxreg_put(params[L + 1], fluxes, faces)
            O.fluxreg_conserve(params[L], fluxes, faces, reg)
    for b, (L, _) in enumerate(grid["blocks"]):
        assert O.update_solution(params[L], U[b], fluxes[b], pavgs[b], dt, grid["del"][L]) == 0
    restrict(grid, U, nunk)
