"""GPU kernels of the learned Newton seed: BF16-MFMA forward and backward
against the float64 oracle (oracle/mlp.py, BF16 rounding emulated), the
batched exact solver against the C oracle, and the seed's effect on the
Newton iteration count."""
import numpy as np
import pytest
import torch

import oracle as O
from oracle import mlp as M
from flash5_amd import learned as L

pytestmark = pytest.mark.gpu
G = 1.4


def _params(seed):
    p = L.init_params(seed).cpu()
    gen = torch.Generator().manual_seed(seed + 100)
    p[L.OFF["b1"]:L.OFF["W2"]] = torch.randn(64, generator=gen) * 0.1
    p[L.OFF["b2"]:L.OFF["w3"]] = torch.randn(64, generator=gen) * 0.1
    p[L.OFF["b3"]] = 0.3
    return p.cuda()


def test_mfma_forward_matches_oracle():
    net = L.SeedNet(params=_params(1))
    S = L.sample_interfaces(1000, seed=11)            # not a multiple of 64
    x = net.features(S, G)
    y = net.forward(x).cpu().numpy()
    yo, _ = M.forward(net.params.cpu().numpy(), x.cpu().numpy())
    assert np.isfinite(y).all()
    np.testing.assert_allclose(y, yo, rtol=2e-3, atol=2e-3)


def test_mfma_gradient_matches_oracle():
    net = L.SeedNet(params=_params(2))
    S = L.sample_interfaces(700, seed=12)
    x = net.features(S, G)
    gen = torch.Generator(device="cuda").manual_seed(3)
    t = torch.randn(700, generator=gen, device="cuda", dtype=torch.float32)
    loss, g = net.grad(x, t)
    go, lo = M.grad(net.params.cpu().numpy(), x.cpu().numpy(), t.cpu().numpy(), 1.0 / 700)
    g = g.cpu().numpy()
    assert loss == pytest.approx(lo, rel=1e-3)
    for name, (a, b) in {"W1": (0, 384), "b1": (384, 448), "W2": (448, 4544),
                         "b2": (4544, 4608), "w3": (4608, 4672), "b3": (4672, 4673)}.items():
        err = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        assert err < 1e-2, (name, err)


def test_riemann_batch_matches_oracle():
    S = L.sample_interfaces(4096, seed=13)
    ps, it, st = L.solve(S, G)
    assert st.code == 0 and st.newton_fallbacks == 0
    Sh = S.cpu().numpy()
    ps = ps.cpu().numpy()
    it = it.cpu().numpy()
    for i in range(0, 4096, 7):
        k, po, _ = O.exact_pstar(G, *Sh[:, i], tol=1e-12, maxit=50)
        assert ps[i] == pytest.approx(po, rel=4e-12), i
        assert abs(it[i] - k) <= 1, (i, it[i], k)
    assert int(st.newton_iters) == int(it.sum())


def test_learned_seed_cuts_newton_iterations():
    """Train briefly on solver-generated states; on held-out interfaces the
    seeded Newton must converge to the same p* in fewer iterations."""
    Str = L.sample_interfaces(1 << 16, seed=21)
    ps, _, _ = L.solve(Str, G)
    net = L.SeedNet(seed=4)
    x = net.features(Str, G)
    t = L.targets(Str, ps, G)
    hist = net.train(x, t, steps=400, batch=4096, lr=3e-3, log_every=100)
    assert hist[-1][1] < 0.5 * hist[0][1], hist
    Ste = L.sample_interfaces(1 << 14, seed=22)
    p0, i0, s0 = L.solve(Ste, G)
    p1, i1, s1 = L.solve(Ste, G, params=net.params)
    assert s0.code == 0 and s1.code == 0
    rel = ((p1 - p0).abs() / p0).max().item()
    assert rel < 4e-12, rel
    m0 = i0.abs().float().mean().item()
    m1 = i1.abs().float().mean().item()
    assert m1 < m0, (m0, m1)
    assert s1.newton_fallbacks <= 0.01 * Ste.shape[1]


def test_learned_seed_inside_hydro_step():
    """Train on the face states an exact-solver step of the batch workload
    produced, then run the same step seeded by the network: the same converged
    answer, and no more Newton iterations (most faces of smooth data are weak
    and keep the two-shock start)."""
    from flash5_amd import sfocu, workloads as W
    from flash5_amd.hydro import Hydro, HydroParams
    n = 8
    U = W.fourier_blocks(16, n=n, gamma=G, seed=W.SEED_BASE + 4).numpy()
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    hy = Hydro(HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=G, RiemannSolver="exact"))
    U0 = torch.from_numpy(U.copy()).cuda()
    st0 = hy.advance_batch(U0, dt, (dx,) * 3)
    S = hy.face_states()
    assert S.shape[1] == 16 * 3 * 9 * 8 * 8
    ps, _, _ = L.solve(S, G)
    net = L.SeedNet(seed=5)
    net.train(net.features(S, G), L.targets(S, ps, G), steps=400, batch=4096, lr=3e-3)
    hy2 = Hydro(HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=G, RiemannSolver="exact",
                            learnedSeed=True))
    hy2.set_seed_params(net.params)
    U1 = torch.from_numpy(U.copy()).cuda()
    st1 = hy2.advance_batch(U1, dt, (dx,) * 3)
    assert st0.code == 0 and st1.code == 0
    names = {"dens": 0, "velx": 1, "vely": 2, "velz": 3, "pres": 4, "ener": 5, "eint": 9}
    norms = sfocu.compare(U1.cpu().numpy(), U0.cpu().numpy(), 1, names, 3, 4)
    for k, (l1, _) in norms.items():
        assert l1 <= 1e-12, (k, l1)
    print("newton iterations per face: two-shock", st0.newton_iters / S.shape[1],
          "learned", st1.newton_iters / S.shape[1], "fallbacks", st1.newton_fallbacks)
    # weak faces keep the two-shock start; the seed may only help the rest
    assert st1.newton_iters <= st0.newton_iters
    hy2.set_seed_params(None)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("gate", [0.0, 0.05, 0.2])
def test_learned_batch_solve_update_matches_face_stage(prec, gate):
    """Learned-seed batch steps of 3-D blocks solve the listed strong faces
    seeded first, then every other face unseeded inside the solve+update
    kernel (kernel path 0); path 2 runs the face stage that reads a seed
    per face, then the update.  Same seeds, same solves: bitwise, with the
    same Newton iteration and fallback counts, at gate 0 (every face
    seeded), at a gate some faces pass, and at the default gate."""
    from flash5_amd import workloads as W
    from flash5_amd.hydro import PREC_MIXED, Hydro, HydroParams
    n = 16
    U = W.fourier_blocks(6, n=n, gamma=G, seed=W.SEED_BASE + 4).numpy()
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    kw = {"precision": PREC_MIXED} if prec == "mixed" else {}
    hy = Hydro(HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=G, RiemannSolver="exact",
                           learnedSeed=True, **kw))
    hy.set_seed_params(_params(7))
    out, stats = [], []
    L.set_gate(gate)
    try:
        for path in (0, 2, 0):
            hy.set_kernel_path(path)
            Ug = torch.from_numpy(U.copy()).cuda()
            st = hy.advance_batch(Ug, dt, (dx,) * 3)
            assert st.code == 0
            out.append(Ug.cpu().numpy())
            stats.append((int(st.newton_iters), int(st.newton_fallbacks)))
    finally:
        L.set_gate(0.2)
        hy.set_kernel_path(0)
        hy.set_seed_params(None)
    for i in (1, 2):
        diff = np.argwhere(out[i] != out[0])
        assert diff.size == 0, f"run {i}: {len(diff)} values differ, first {diff[:4]}"
    assert stats[0] == stats[1] == stats[2], stats
    assert not np.array_equal(out[0], U)


def test_mixed_mode_face_states_are_fp32():
    """The mixed-precision exact step stores its face states in FP32: the
    collected states are FP32 numbers, within FP32 reconstruction error of the
    FP64 step's states on the same blocks; the seeded mixed step converges
    with no fallbacks."""
    from flash5_amd import workloads as W
    from flash5_amd.hydro import PREC_MIXED, Hydro, HydroParams
    n = 8
    U = W.fourier_blocks(8, n=n, gamma=G, seed=W.SEED_BASE + 4).numpy()
    dx = 1.0 / n
    dt = W.cfl_dt(torch.from_numpy(U), 3, 4, (dx,) * 3, 0.8)
    states = {}
    for prec in ("fp64", PREC_MIXED):
        kw = {} if prec == "fp64" else {"precision": PREC_MIXED}
        hy = Hydro(HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=G, RiemannSolver="exact", **kw))
        st = hy.advance_batch(torch.from_numpy(U.copy()).cuda(), dt, (dx,) * 3)
        assert st.code == 0 and st.newton_fallbacks == 0
        states[prec] = hy.face_states().cpu().numpy()
    s64, s32 = states["fp64"], states[PREC_MIXED]
    assert s32.shape == s64.shape == (6, 8 * 3 * 9 * 8 * 8)
    np.testing.assert_array_equal(s32, s32.astype(np.float32).astype(np.float64))
    assert not np.array_equal(s32, s64)
    scale = np.abs(s64).max(axis=1, keepdims=True)
    assert (np.abs(s32 - s64) <= 1e-5 * scale).all(), np.abs(s32 - s64).max()
