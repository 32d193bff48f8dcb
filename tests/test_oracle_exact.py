"""Pinning tests for the oracle's exact Riemann solver (CPU only).

The exact solver is not part of the reference (it is the new RiemannSolver
option with hy_HLLC.F90's contract).  It is pinned by the literature:

* Toro (2009) Table 4.2 star pressures, velocities and densities for tests
  1-5 (golden fixture tests/golden/toro_exact_riemann.json), each compared to
  the printed precision; the densities through the Godunov flux at x/t = 0,
  which lies in a different wave region for each test;
* a transonic rarefaction, where the state at x/t = 0 is fixed by three
  invariants (sonic point u = c, isentrope, Riemann invariant);
* Galilean invariance of p*, mirror symmetry of the flux;
* p* against bisection of the pressure function on random states;
* iteration-limit and vacuum reporting.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TORO = json.load(open(os.path.join(HERE, "golden", "toro_exact_riemann.json")))
G = TORO["gamma"]


def printed(s):
    """Value and half a unit in the last printed digit."""
    dec = len(s.split(".")[1]) if "." in s else 0
    return float(s), 0.5 * 10.0 ** (-dec)


def euler_flux(rho, u, p, g=G, v=0.0, w=0.0):
    E = 0.5 * rho * (u * u + v * v + w * w) + p / (g - 1)
    return np.array([rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u, p])


@pytest.mark.parametrize("case", TORO["tests"], ids=[t["name"] for t in TORO["tests"]])
def test_star_state_matches_toro_table(case):
    (rl, ul, pl), (rr, ur, pr) = case["left"], case["right"]
    it, ps, us = O.exact_pstar(G, rl, ul, pl, rr, ur, pr, tol=1e-12, maxit=50)
    assert 0 < it <= 20
    p_t, dp = printed(case["pstar"])
    u_t, du = printed(case["ustar"])
    # test 5's data are themselves rounded test-3/4 star values: add how far
    # half a unit in each input's last printed digit moves the solution
    sp = su = 0.0
    if case["name"].startswith("test5"):
        x = [rl, ul, pl, rr, ur, pr]
        for i in range(6):
            h = 0.5 * 10.0 ** -len(repr(x[i]).split(".")[1])
            y = list(x)
            y[i] += h
            _, ps2, us2 = O.exact_pstar(G, *y, tol=1e-12, maxit=50)
            sp, su = sp + abs(ps2 - ps), su + abs(us2 - us)
    assert abs(ps - p_t) <= dp * 1.0001 + sp, (ps, p_t)
    assert abs(us - u_t) <= du * 1.0001 + su, (us, u_t)


def _godunov(case):
    (rl, ul, pl), (rr, ur, pr) = case["left"], case["right"]
    p = O.params(1, 8, G, riemann=O.EXACT, exactMaxIter=50)
    rc, F, st = O.exact_flux(p, 0, O.prim(rl, ul, 0, 0, pl, G), O.prim(rr, ur, 0, 0, pr, G))
    assert rc == 0 and st.exactFallbacks == 0
    return F


def test_godunov_flux_samples_the_right_wave_region():
    cases = {t["name"]: t for t in TORO["tests"]}
    # test 1 and 3: x/t = 0 inside the left star region (left rarefaction's tail
    # moves left, contact right)
    for name in ("test1-sod", "test3-blast-left"):
        c = cases[name]
        F = _godunov(c)
        rho, drho = printed(c["rhoL"])
        u, du = printed(c["ustar"])
        p, dp = printed(c["pstar"])
        want = euler_flux(rho, u, p)
        tol = np.abs(euler_flux(rho + drho, u + du, p + dp) - want) * 3 + 1e-12
        assert (np.abs(F - want) <= tol).all(), (name, F, want)
    # test 4: right star region (contact moves left, right rarefaction's tail right)
    c = cases["test4-blast-right"]
    F = _godunov(c)
    rho, drho = printed(c["rhoR"])
    u, du = printed(c["ustar"])
    p, dp = printed(c["pstar"])
    want = euler_flux(rho, u, p)
    tol = np.abs(euler_flux(rho + drho, u - du, p + dp) - want) * 3 + 1e-12
    assert (np.abs(F - want) <= tol).all(), (F, want)
    # test 2: x/t = 0 on the stationary contact: no mass flux, momentum flux p*
    c = cases["test2-123"]
    F = _godunov(c)
    p, dp = printed(c["pstar"])
    assert abs(F[0]) < 1e-12 and abs(F[4]) < 1e-12
    assert abs(F[1] - p) <= dp
    # test 5: the left shock moves right: x/t = 0 sees the unshocked left state
    c = cases["test5-shock-collision"]
    F = _godunov(c)
    np.testing.assert_allclose(F, euler_flux(*c["left"]), rtol=1e-14)


def test_transonic_rarefaction_state_is_sonic_isentropic_invariant():
    """Left state moving so that x/t = 0 falls inside the left fan."""
    rl, ul, pl, rr, ur, pr = 1.0, 0.4, 1.0, 0.125, 0.4, 0.1
    cl = np.sqrt(G * pl / rl)
    p = O.params(1, 8, G, riemann=O.EXACT)
    rc, F, _ = O.exact_flux(p, 0, O.prim(rl, ul, 0, 0, pl, G), O.prim(rr, ur, 0, 0, pr, G))
    assert rc == 0
    # state from the invariants: u = c, p / rho^g = K, u + 2c/(g-1) = J
    K = pl / rl ** G
    J = ul + 2 * cl / (G - 1)
    u = c = J * (G - 1) / (G + 1)
    rho = (c * c / (G * K)) ** (1 / (G - 1))
    np.testing.assert_allclose(F, euler_flux(rho, u, K * rho ** G), rtol=1e-13)


def test_galilean_invariance_and_mirror_symmetry():
    rng = np.random.default_rng(7)
    p = O.params(3, 8, G, riemann=O.EXACT, exactMaxIter=50)
    for _ in range(200):
        rl, rr, pl, pr = 10 ** rng.uniform(-2, 2, 4)
        cl, cr = np.sqrt(G * pl / rl), np.sqrt(G * pr / rr)
        ul, ur = rng.uniform(-1, 1) * cl, rng.uniform(-1, 1) * cr
        it, ps, us = O.exact_pstar(G, rl, ul, pl, rr, ur, pr, maxit=50)
        shift = rng.normal() * (cl + cr)
        it2, ps2, us2 = O.exact_pstar(G, rl, ul + shift, pl, rr, ur + shift, pr, maxit=50)
        assert it > 0 and it2 > 0
        assert ps2 == pytest.approx(ps, rel=1e-11)
        assert us2 == pytest.approx(us + shift, rel=1e-11, abs=1e-11 * (cl + cr))
        d = int(rng.integers(3))
        vt = rng.normal(size=2)
        VL = O.prim(rl, 0, 0, 0, pl, G)
        VR = O.prim(rr, 0, 0, 0, pr, G)
        VL[1 + d], VR[1 + d] = ul, ur
        VL[1 + (d + 1) % 3], VR[1 + (d + 2) % 3] = vt
        _, Fa, _ = O.exact_flux(p, d, VL, VR)
        flip = np.ones(9)
        flip[1 + d] = -1
        _, Fb, _ = O.exact_flux(p, d, VR * flip, VL * flip)
        sgn = -np.ones(6)
        sgn[1 + d] = 1
        sgn[5] = 1
        np.testing.assert_allclose(Fb, sgn * Fa, atol=1e-12 * np.abs(Fa).max())


def _f_textbook(ps, rho, p, g=G):
    """Toro eqs. 4.6 (shock) and 4.7 (rarefaction)."""
    c = np.sqrt(g * p / rho)
    if ps > p:
        A, B = 2 / ((g + 1) * rho), (g - 1) / (g + 1) * p
        return (ps - p) * np.sqrt(A / (ps + B))
    return 2 * c / (g - 1) * ((ps / p) ** ((g - 1) / (2 * g)) - 1)


def test_pstar_matches_bisection_of_pressure_function():
    rng = np.random.default_rng(11)
    for _ in range(300):
        rl, rr, pl, pr = 10 ** rng.uniform(-3, 3, 4)
        cl, cr = np.sqrt(G * pl / rl), np.sqrt(G * pr / rr)
        ul, ur = rng.uniform(-3, 3) * cl, rng.uniform(-3, 3) * cr
        if 2 / (G - 1) * (cl + cr) <= ur - ul:
            continue
        it, ps, us = O.exact_pstar(G, rl, ul, pl, rr, ur, pr, tol=1e-13, maxit=100)
        assert it > 0
        fun = lambda q: _f_textbook(q, rl, pl) + _f_textbook(q, rr, pr) + ur - ul
        lo, hi = 1e-300, max(pl, pr)
        while fun(hi) < 0:
            hi *= 2
        for _ in range(2000):
            mid = 0.5 * (lo + hi) if hi / lo < 4 else np.sqrt(lo * hi)
            if mid in (lo, hi):
                break
            lo, hi = (mid, hi) if fun(mid) < 0 else (lo, mid)
        assert ps == pytest.approx(0.5 * (lo + hi), rel=1e-10)


def test_iteration_limit_and_vacuum_are_reported():
    it, ps, _ = O.exact_pstar(G, 1.0, 0.0, 1000.0, 1.0, 0.0, 0.01, maxit=1)
    assert it == -1
    p = O.params(1, 8, G, riemann=O.EXACT)
    # generated vacuum: 2(cl+cr)/(g-1) <= ur - ul
    rc, F, _ = O.exact_flux(p, 0, O.prim(1, -20, 0, 0, 1, G), O.prim(1, 20, 0, 0, 1, G))
    assert rc == 2
