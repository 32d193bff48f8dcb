"""Oracle pins of the first-order positivity fallback.

FLASH resets a cell's face states to the cell-centre primitives (first-order
Godunov) in two places:

* after the 2-D CTU transverse corrections, per direction, when a
  reconstructed density or pressure is negative
  (hy_dataReconstOneStep.F90:1117-1154);
* in 3-D, after the full-CTU corner terms, when one is zero or negative
  (hy_getRiemannState.F90:1037-1051, 1139-1153, 1244-1258, through
  fallbackToFirstOrder at :1456-1475).

The tests build blocks whose states go negative in each stage (log-uniform
densities and pressures with Mach-3 velocities) and check that the oracle's
block fluxes equal the composition of its separately pinned stage routines
(ox_ppm, ox_upwind_trans_flux, ox_hllc; see test_oracle_stages.py) with the
rule above, face by face and bit for bit, and that the fallback count is the
number of (state cell, direction) pairs the rule reset.  A dropped check, a
check before the transverse terms, a reset to another cell or to only one of
the two states, or a missing corner-term check changes some face flux.
Every cell's density still changes by its face-flux differences (the
conservative form), except where it is floored at smlrho.
"""
import numpy as np
import pytest

import oracle as O

G = 1.4


def random_block(p, seed, lo=-4.0):
    """Log-uniform density and pressure on [10^lo, 1], velocities up to
    three sound speeds per axis."""
    r = np.random.default_rng(seed)
    shp = O.block_shape(p)
    U = np.zeros(shp)
    rho = 10.0 ** r.uniform(lo, 0.0, shp[1:])
    pres = 10.0 ** r.uniform(lo, 0.0, shp[1:])
    c = np.sqrt(G * pres / rho)
    U[O.DENS], U[O.PRES] = rho, pres
    for a in range(p.ndim):
        U[O.VELX + a] = r.uniform(-3.0, 3.0, shp[1:]) * c
    O.eos(p, U, 1)
    return U


def cfl_dt(p, U, dx, cfl=0.8):
    return O.block_dt(p, U, (dx,) * 3, cfl)


def prim(U, k, j, i):
    return np.array([U[O.DENS, k, j, i], U[O.VELX, k, j, i], U[O.VELY, k, j, i],
                     U[O.VELZ, k, j, i], U[O.PRES, k, j, i], U[O.GAMC, k, j, i],
                     U[O.GAME, k, j, i], U[O.EINT, k, j, i], 0.0])


def step(d):
    o = [0, 0, 0]
    o[d] = 1
    return o


def composed_fluxes(p, U, dt, dx):
    """Block face fluxes built from the pinned stages and FLASH's fallback
    rule (flattening off: it is pinned in 1-D by test_oracle_block.py).
    Returns (fluxes per direction, resets in the 2-D stage, resets in the
    3-D corner stage)."""
    nd, g, n = p.ndim, p.ng, p.nxb
    ext = [1, int(nd > 1), int(nd > 2)]
    lo = [g, g if nd > 1 else 0, g if nd > 2 else 0]
    hi = [g + n - 1, g + n - 1 if nd > 1 else 0, g + n - 1 if nd > 2 else 0]
    hdt = [0.5 * dt / dx] * 3
    sig = {}      # (k, j, i, d) -> transverse sigma of the cell along d
    W = {}        # (k, j, i, d) -> [Wp, Wn] of a state cell
    resets2 = 0
    # every cell the corner terms reach: interior +- 2 (3-D), +- 1 otherwise
    reach = 2 if nd == 3 else 1
    rng = [range(lo[a] - reach * ext[a], hi[a] + reach * ext[a] + 1) for a in range(3)]
    for k in rng[2]:
        for j in rng[1]:
            for i in rng[0]:
                c = (k, j, i)
                Vc = prim(U, *c)
                states = {}
                for d in range(nd):
                    o = step(d)
                    data = np.stack([prim(U, k + s * o[2], j + s * o[1], i + s * o[0])
                                     for s in range(-2, 3)])
                    rc, wp, wm, sg = O.ppm(p, d, dt, dx, data)
                    assert rc == 0
                    sig[c + (d,)] = sg
                    states[d] = [wp.copy(), wm.copy()]
                # the cell is a state cell of d when its d-coordinate lies in
                # [lo-1, hi+1] and the others are interior
                ci = [i, j, k]
                def is_state(d):
                    if not (lo[d] - 1 <= ci[d] <= hi[d] + 1):
                        return False
                    return all(lo[a] <= ci[a] <= hi[a] for a in range(nd) if a != d)
                if not any(is_state(d) for d in range(nd)):
                    continue
                # 2-D CTU corrections in the reference's order
                # (hy_dataReconstOneStep.F90:906-914, 940-951)
                if nd > 1:
                    for s in (1, 0):   # Wn then Wp, both with the other direction's sigma
                        states[0][s] = states[0][s] - hdt[1] * sig[c + (1,)]
                        states[1][s] = states[1][s] - hdt[0] * sig[c + (0,)]
                    if nd > 2:
                        for s in (1, 0):
                            states[0][s] = states[0][s] - hdt[2] * sig[c + (2,)]
                            states[1][s] = states[1][s] - hdt[2] * sig[c + (2,)]
                            states[2][s] = states[2][s] - hdt[0] * sig[c + (0,)] - \
                                hdt[1] * sig[c + (1,)]
                for d in range(nd):
                    if not is_state(d):
                        continue
                    wp, wn = states[d]
                    if min(wn[0], wp[0], wn[4], wp[4]) < 0.0:        # :1120-1154
                        wp, wn = Vc.copy(), Vc.copy()
                        resets2 += 1
                    W[c + (d,)] = [wp, wn]
    resets3 = 0
    if nd == 3 and p.use3dFullCTU:
        fac = dt * dt / (6.0 * dx * dx)
        # (sigma direction, eigen direction, neighbour step) of the two terms
        tab = {0: ((2, 1, 1), (1, 2, 2)), 1: ((0, 2, 2), (2, 0, 0)), 2: ((1, 0, 0), (0, 1, 1))}
        for key in sorted(W):
            k, j, i, d = key
            c = (k, j, i)
            Vc = prim(U, *c)
            tf = []
            for sd, ed, nb in tab[d]:
                o = step(nb)
                cm = (k - o[2], j - o[1], i - o[0])
                cp = (k + o[2], j + o[1], i + o[0])
                tf.append(O.trans_flux(p, sig[cm + (sd,)], sig[c + (sd,)], sig[cp + (sd,)], Vc, ed))
            wp, wn = W[key]
            wp = wp + (tf[0] + tf[1]) * fac
            wn = wn + (tf[0] + tf[1]) * fac
            if min(wp[0], wn[0], wp[4], wn[4]) <= 0.0:              # :1037-1051
                wp, wn = Vc.copy(), Vc.copy()
                resets3 += 1
            W[key] = [wp, wn]
    flux = []
    for d in range(nd):
        shp = O.flux_shape(p, d)
        F = np.zeros(shp)
        for fk in range(shp[1]):
            for fj in range(shp[2]):
                for fi in range(shp[3]):
                    i, j, k = fi + lo[0], fj + lo[1], fk + lo[2]
                    o = step(d)
                    L = W[(k - o[2], j - o[1], i - o[0], d)][0]
                    R = W[(k, j, i, d)][1]
                    rc, Fs, _ = O.hllc(p, d, L, R)
                    assert rc == 0
                    up = L if Fs[0] > 0 else R                       # hy_getFaceFlux.F90:433-442
                    F[:, fk, fj, fi] = np.concatenate([Fs, [Fs[0] * up[7], Fs[0] / up[0]]])
        flux.append(F)
    return flux, resets2, resets3


CASES = [(1, 16, 11), (2, 8, 12), (3, 8, 13)]


@pytest.mark.parametrize("ndim,n,seed", CASES)
def test_block_fluxes_are_the_composition_with_first_order_resets(ndim, n, seed):
    p = O.params(ndim, n, G, flattening=0)
    U = random_block(p, seed)
    dx = 1.0 / n
    dt = cfl_dt(p, U, dx)
    rc, fl, _, st = O.compute_fluxes(p, U.copy(), dt, (dx,) * 3, update=False)
    assert rc == 0
    want, r2, r3 = composed_fluxes(p, U, dt, dx)
    assert r2 > 0, "the data must trigger the reset after the transverse corrections"
    if ndim == 3:
        assert r3 > 0, "the data must trigger the reset after the corner terms"
    for d in range(ndim):
        np.testing.assert_array_equal(fl[d], want[d])
    assert st.fallbacks == r2 + r3


@pytest.mark.parametrize("ndim,n,seed", CASES[:2])
def test_fallback_faces_carry_cell_centre_states(ndim, n, seed):
    """Where every state feeding a face was reset, the face flux is the
    first-order Godunov (HLLC) flux of the two cell-centre states."""
    p = O.params(ndim, n, G, flattening=0)
    U = random_block(p, seed)
    dx = 1.0 / n
    dt = cfl_dt(p, U, dx)
    rc, fl, _, st = O.compute_fluxes(p, U.copy(), dt, (dx,) * 3, update=False)
    assert rc == 0 and st.fallbacks > 0
    # recompute with the stages to find which states were reset
    g = p.ng
    lo = [g, g if ndim > 1 else 0, g if ndim > 2 else 0]
    want, _, _ = composed_fluxes(p, U, dt, dx)
    hits = 0
    for d in range(ndim):
        shp = O.flux_shape(p, d)
        o = step(d)
        for fk in range(shp[1]):
            for fj in range(shp[2]):
                for fi in range(shp[3]):
                    i, j, k = fi + lo[0], fj + lo[1], fk + lo[2]
                    L = prim(U, k - o[2], j - o[1], i - o[0])
                    R = prim(U, k, j, i)
                    rc, Fs, _ = O.hllc(p, d, L, R)
                    up = L if Fs[0] > 0 else R
                    god = np.concatenate([Fs, [Fs[0] * up[7], Fs[0] / up[0]]])
                    if np.array_equal(want[d][:, fk, fj, fi], god):
                        hits += 1
                        np.testing.assert_array_equal(fl[d][:, fk, fj, fi], god)
    assert hits > 0


@pytest.mark.parametrize("ndim,n,seed", CASES)
def test_conservative_form_holds_with_fallbacks(ndim, n, seed):
    """Every interior cell's density changes by its face-flux differences
    (hy_unsplitUpdate.F90:1187-1242), or sits at the smlrho floor (:811)."""
    p = O.params(ndim, n, G)
    U = random_block(p, seed)
    dx = 1.0 / n
    dt = cfl_dt(p, U, dx)
    U1 = U.copy()
    rc, fl, _, st = O.compute_fluxes(p, U1, dt, (dx,) * 3, update=True)
    assert rc == 0 and st.fallbacks > 0
    g = p.ng
    sl = tuple([slice(g, g + n) if ndim > a else slice(None) for a in (2, 1, 0)])
    want = U[O.DENS][sl].copy()
    for d in range(ndim):
        F = fl[d][O.F_DENS]
        ax = 2 - d
        hi = np.take(F, np.arange(1, F.shape[ax]), axis=ax)
        lo = np.take(F, np.arange(0, F.shape[ax] - 1), axis=ax)
        want = want - dt / dx * (hi - lo)
    got = U1[O.DENS][sl]
    floored = got == p.smlrho
    assert np.all(want[floored] <= p.smlrho * (1 + 1e-12) + 1e-12)
    ok = ~floored
    assert ok.sum() > 0.9 * ok.size
    np.testing.assert_allclose(got[ok], want[ok], rtol=1e-11, atol=1e-14)


def test_1d_flattened_fallback_matches_the_stages():
    """1-D with flattening on: the reset acts on the flattened PPM states."""
    n = 16
    p = O.params(1, n, G)
    U = random_block(p, 11)
    g = p.ng
    dx = 1.0 / n
    dt = cfl_dt(p, U, dx)
    rc, fl, _, st = O.compute_fluxes(p, U.copy(), dt, (dx,) * 3, update=False)
    assert rc == 0 and st.fallbacks > 0
    P, Vx = U[O.PRES, 0, 0], U[O.VELX, 0, 0]

    def tilde(i):
        dp1, dp2 = P[i + 1] - P[i - 1], P[i + 2] - P[i - 2]
        dv1 = Vx[i + 1] - Vx[i - 1]
        sp = dp1 / dp2 - 0.75 if abs(dp2) > 1e-15 else 0.0
        ft = max(0.0, min(1.0, 10.0 * sp))
        if abs(dp1) / min(P[i - 1], P[i + 1]) < 1.0 / 3.0 or dv1 > 0.0:
            ft = 0.0
        return ft

    W, resets = {}, 0
    for i in range(g - 1, g + n + 1):
        dp1 = P[i + 1] - P[i - 1]
        ft = tilde(i)
        fc = ft if dp1 == 0 else max(ft, tilde(i + 1) if dp1 < 0 else tilde(i - 1))
        data = np.stack([prim(U, 0, 0, i + s) for s in range(-2, 3)])
        rc, wp, wm, _ = O.ppm(p, 0, dt, dx, data, flat=fc)
        if min(wp[0], wm[0], wp[4], wm[4]) < 0.0:
            wp, wm = prim(U, 0, 0, i), prim(U, 0, 0, i)
            resets += 1
        W[i] = (wp, wm)
    assert resets == st.fallbacks > 0
    for f in range(n + 1):
        L, R = W[g + f - 1][0], W[g + f][1]
        rc, F, _ = O.hllc(p, 0, L, R)
        up = L if F[0] > 0 else R
        np.testing.assert_array_equal(fl[0][:, 0, 0, f],
                                      np.concatenate([F, [F[0] * up[7], F[0] / up[0]]]))
