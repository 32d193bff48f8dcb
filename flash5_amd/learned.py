"""Learned Newton seed for the exact Riemann solver (no counterpart in the
reference): sampling of Riemann problems, training of the 6->64->64->1 seed
network with the library's own BF16-MFMA forward/backward kernels, and the
batched exact solver with or without the seed.

All compute runs in libfx_hydro.so on the GPU; torch only holds device
buffers and draws random numbers.
"""
from __future__ import annotations

import ctypes as C
import math

from .hydro import FxError, FxStatus, load_library

NPARAM = 4673
OFF = {"W1": 0, "b1": 384, "W2": 448, "b2": 4544, "w3": 4608, "b3": 4672}


def _fp(t):
    import torch
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.float32
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_float))


def _dp(t):
    import torch
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.float64
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double))


def _ip(t):
    import torch
    assert t.is_cuda and t.is_contiguous() and t.dtype == torch.int32
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_int))


def sample_interfaces(n: int, seed: int, gamma: float = 1.4, device="cuda"):
    """Riemann problems of the learned-seed study: rho and p log-uniform on
    [1e-3, 1e3], normal velocities uniform in [-3, 3] times the local sound
    speed.  Returns a (6, n) float64 tensor rhoL, uL, pL, rhoR, uR, pR."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    u = torch.rand((6, n), generator=gen, device=device, dtype=torch.float64)
    S = torch.empty((6, n), device=device, dtype=torch.float64)
    for side in (0, 3):
        S[side] = 10.0 ** (6.0 * u[side] - 3.0)                 # rho
        S[side + 2] = 10.0 ** (6.0 * u[side + 2] - 3.0)         # p
        c = torch.sqrt(gamma * S[side + 2] / S[side])
        S[side + 1] = (6.0 * u[side + 1] - 3.0) * c             # u
    return S.contiguous()


def init_params(seed: int = 0, device="cuda"):
    """Glorot-uniform weights, zero biases, FP32 master copy."""
    import torch
    gen = torch.Generator(device="cpu")
    gen.manual_seed(seed)
    p = torch.zeros(NPARAM, dtype=torch.float32)
    def glorot(off, fan_out, fan_in):
        a = math.sqrt(6.0 / (fan_in + fan_out))
        p[off:off + fan_in * fan_out] = (torch.rand(fan_in * fan_out, generator=gen) * 2 - 1) * a
    glorot(OFF["W1"], 64, 6)
    glorot(OFF["W2"], 64, 64)
    glorot(OFF["w3"], 1, 64)
    return p.to(device)


class SeedNet:
    """Seed network and its training on the GPU."""

    def __init__(self, params=None, seed: int = 0):
        import torch
        self.lib = load_library()
        self.params = params if params is not None else init_params(seed)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0

    def features(self, S, gamma: float):
        import torch
        n = S.shape[1]
        x = torch.empty((n, 6), dtype=torch.float32, device=S.device)
        rc = self.lib.fx_mlp_features(_dp(S), n, gamma, _fp(x))
        if rc:
            raise FxError("fx_mlp_features", rc=rc)
        return x

    def forward(self, x):
        import torch
        n = x.shape[0]
        y = torch.empty(n, dtype=torch.float32, device=x.device)
        rc = self.lib.fx_mlp_forward(_fp(self.params), _fp(x), n, _fp(y))
        if rc:
            raise FxError("fx_mlp_forward", rc=rc)
        return y

    def grad(self, x, t):
        """Mean squared error over the batch and its gradient."""
        import torch
        n = x.shape[0]
        g = torch.zeros(NPARAM, dtype=torch.float32, device=x.device)
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        rc = self.lib.fx_mlp_grad(_fp(self.params), _fp(x), _fp(t), n, 1.0 / n, _fp(g), _fp(loss))
        if rc:
            raise FxError("fx_mlp_grad", rc=rc)
        return float(loss.item()), g

    def adam(self, g, lr: float):
        self.step_count += 1
        rc = self.lib.fx_mlp_adam(_fp(self.params), _fp(g), _fp(self.m), _fp(self.v),
                                  self.step_count, lr)
        if rc:
            raise FxError("fx_mlp_adam", rc=rc)

    def train(self, x, t, steps: int, batch: int = 4096, lr: float = 2e-3, seed: int = 1,
              log_every: int = 0):
        """Adam on random minibatches of (x, t); cosine-decayed learning rate."""
        import torch
        gen = torch.Generator(device=x.device)
        gen.manual_seed(seed)
        n = x.shape[0]
        hist = []
        for k in range(steps):
            idx = torch.randint(0, n, (batch,), generator=gen, device=x.device)
            xb, tb = x[idx].contiguous(), t[idx].contiguous()
            loss, g = self.grad(xb, tb)
            self.adam(g, lr * 0.5 * (1 + math.cos(math.pi * k / steps)))
            if log_every and (k % log_every == 0 or k == steps - 1):
                hist.append((k, loss))
        return hist


def set_gate(gate: float) -> None:
    """Interfaces whose largest |feature| is below gate keep the two-shock start."""
    rc = load_library().fx_mlp_set_gate(gate)
    if rc:
        raise FxError("fx_mlp_set_gate", rc=rc)


def solve(S, gamma: float, params=None, maxit: int = 50, tol: float = 1e-12):
    """Batched exact Riemann problems: (pstar float64, iters int32, status)."""
    import torch
    lib = load_library()
    n = S.shape[1]
    ps = torch.empty(n, dtype=torch.float64, device=S.device)
    it = torch.empty(n, dtype=torch.int32, device=S.device)
    st = FxStatus()
    rc = lib.fx_riemann_batch(_dp(S), n, gamma, _fp(params) if params is not None else None,
                              maxit, tol, _dp(ps), _ip(it), C.byref(st))
    if rc:
        raise FxError("fx_riemann_batch", st, rc)
    return ps, it, st


def targets(S, pstar, gamma: float = 1.4):
    """Training target ln(p*/P0), P0 = sqrt(pL pR)."""
    import torch
    del gamma
    return (torch.log(pstar) - 0.5 * (torch.log(S[2]) + torch.log(S[5]))).float().contiguous()


def study(n_test_log2: int = 24, n_train_log2: int = 22, steps: int = 3000, gamma: float = 1.4,
          seed: int = 0x5EED0004, keep_params: bool = False):
    """Learned-seed study: train on 2^n_train solver-emitted states, then solve
    2^n_test held-out interfaces from the two-shock start and from the seed.
    Returns a dict of iteration counts, fallbacks, p* agreement and rates."""
    import time
    import torch
    t0 = time.perf_counter()
    Str = sample_interfaces(1 << n_train_log2, seed=seed, gamma=gamma)
    ps, _, _ = solve(Str, gamma)
    net = SeedNet(seed=1)
    x = net.features(Str, gamma)
    t = targets(Str, ps, gamma)
    hist = net.train(x, t, steps=steps, batch=8192, lr=3e-3, log_every=max(1, steps // 5))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    del Str, ps, x, t
    Ste = sample_interfaces(1 << n_test_log2, seed=seed + 1, gamma=gamma)
    solve(Ste[:, :4096].contiguous(), gamma)                      # warm-up
    solve(Ste[:, :4096].contiguous(), gamma, params=net.params)
    out = {}
    res = {}
    for name, prm in (("two_shock", None), ("learned", net.params)):
        torch.cuda.synchronize()
        a = time.perf_counter()
        p, it, st = solve(Ste, gamma, params=prm)
        torch.cuda.synchronize()
        dt = time.perf_counter() - a
        res[name] = p
        itf = it.abs().float()
        out[name] = {"iters_mean": itf.mean().item(), "iters_max": int(itf.max().item()),
                     "newton_fallbacks": int(st.newton_fallbacks),
                     "fallback_rate": int(st.newton_fallbacks) / Ste.shape[1],
                     "not_converged": int((it < 0).sum().item()),
                     "interfaces_per_s": Ste.shape[1] / dt, "ms": dt * 1e3}
    rel = ((res["learned"] - res["two_shock"]).abs() / res["two_shock"])
    r = {"interfaces": Ste.shape[1], "train_states": 1 << n_train_log2, "train_steps": steps,
         "train_s": train_s, "loss": [h[1] for h in hist], "gamma": gamma,
         "pstar_rel_diff_max": rel.max().item(), "pstar_rel_diff_mean": rel.mean().item(), **out}
    if keep_params:
        r["params"] = net.params
    return r


def strength(S, gamma: float):
    """Largest |feature| of each interface (the gate's measure)."""
    import torch
    rl, ul, pl, rr, ur, pr = S
    P0, R0 = torch.sqrt(pl * pr), torch.sqrt(rl * rr)
    C0 = torch.sqrt(gamma * P0 / R0)
    return torch.maximum(torch.maximum((0.5 * torch.log(rl / rr)).abs(),
                                       (0.5 * torch.log(pl / pr)).abs()),
                         (0.5 * (ur - ul) / C0).abs())


def strength_bins(params, n_log2: int = 22, gamma: float = 1.4, seed: int = 0x5EED0005,
                  edges=(0, 0.02, 0.05, 0.1, 0.2, 0.5, 1, 2, 100)):
    """Mean Newton iterations per strength bin from the two-shock start and
    from the seed with the gate off (held-out interfaces)."""
    S = sample_interfaces(1 << n_log2, seed=seed, gamma=gamma)
    m = strength(S, gamma)
    _, i0, _ = solve(S, gamma)
    set_gate(0.0)
    try:
        _, i1, _ = solve(S, gamma, params=params)
    finally:
        set_gate(0.2)
    rows = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = (m >= lo) & (m < hi)
        if int(sel.sum()):
            rows.append({"bin": [lo, hi], "n": int(sel.sum()),
                         "two_shock": i0[sel].abs().float().mean().item(),
                         "seeded": i1[sel].abs().float().mean().item()})
    return rows


def batch_face_strengths(blocks: int = 256, n: int = 16, gamma: float = 5.0 / 3.0):
    """Strength quantiles of the face states of one exact-solver step of the
    3-D batch workload."""
    from . import workloads as W
    from .hydro import Hydro, HydroParams
    U = W.fourier_blocks(blocks, n=n, gamma=gamma, device="cuda")
    dt = W.cfl_dt(U, 3, 4, (1.0 / n,) * 3, 0.8)
    hy = Hydro(HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=gamma, RiemannSolver="exact"))
    hy.advance_batch(U, dt, (1.0 / n,) * 3)
    m = strength(hy.face_states(), gamma)
    return {"faces": int(m.numel()),
            "quantiles": {str(q): m.quantile(q).item() for q in (0.5, 0.9, 0.99, 0.999)},
            "frac_at_or_above": {str(t): (m >= t).float().mean().item() for t in (0.05, 0.1, 0.2, 0.5)}}


def main() -> None:
    import argparse
    import json
    ap = argparse.ArgumentParser(description="learned Newton seed study (2^24 interfaces)")
    ap.add_argument("--test-log2", type=int, default=24)
    ap.add_argument("--train-log2", type=int, default=22)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--bins", action="store_true",
                    help="also: iterations per strength bin, and batch-workload face strengths")
    a = ap.parse_args()
    res = study(a.test_log2, a.train_log2, a.steps, keep_params=a.bins)
    if a.bins:
        params = res.pop("params")
        res["strength_bins"] = strength_bins(params)
        res["batch_face_strengths"] = batch_face_strengths()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
