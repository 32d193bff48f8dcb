"""Benchmark: unsplit PPM + HLLC hydro step on a batch of 4096 independent
3-D blocks (16^3 interior + 4 guard cells, gamma = 5/3, FP64), one MI355X.

A step is one pass of the hot path over the whole batch: face states
(PPM, characteristic tracing, CTU), HLLC fluxes, conservative update,
energy fix and EOS, in place.  Inputs are resident in HBM before timing.
Prints one JSON line (metric, value, roofline, cpu_baseline, ...).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--blocks N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, AMD published
FP64_PEAK_TFLOPS = 78.6        # MI355X FP64 vector, AMD published
N_IN, N_OUT = 7, 8             # FP64 variables read (with halo) / written per zone
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def bytes_per_zone(d: int, n: int, g: int) -> float:
    """Algorithmic bytes per zone-update of the whole step: 8*[N_in*((n+2g)/n)^d + N_out]."""
    return 8.0 * (N_IN * ((n + 2 * g) / n) ** d + N_OUT)


def march_bytes_per_zone(d: int, n: int, g: int) -> float:
    """Algorithmic bytes per zone of the face-state/flux kernel alone: the
    compulsory read of the 7 input fields with their guard cells.  Its face
    fluxes are scratch for the update kernel, so they are not counted."""
    return 8.0 * N_IN * ((n + 2 * g) / n) ** d


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each divided by the ratio the
    same counter reads for a plain copy of known size), or None when the
    summary was taken on another workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("workload") != workload:
        return None
    return (rec["fetch_kib"] * 1024.0 / rec["copy_fetch_factor"] +
            rec["write_kib"] * 1024.0 / rec["copy_write_factor"])


def measured_flops_per_zone(kernel: str, workload: str):
    """FP64 flops per zone of a kernel from the committed VALU counter summary
    (64 lanes x (ADD + MUL + 2 FMA) per dispatch / zones), or None when the
    summary was taken on another workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if rec.get("workload") != workload or "fp64_flops_per_zone" not in rec:
        return None
    return float(rec["fp64_flops_per_zone"])


def host_cpu():
    """Host CPU model, logical CPUs and the CPUs this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def copy_probe(nbytes: int = 1 << 30, reps: int = 3) -> None:
    """Plain device copies of known size, for scaling the PMC byte counters."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copy probe: {reps} copies of {nbytes} bytes", file=sys.stderr)


def train_seed(hp, U, dt, dx, gamma, n_train: int = 1 << 22, steps: int = 1500):
    """Untimed: one exact-solver step on a copy of the batch, 2^22 of its face
    states solved for ln(p*/P0), the seed network trained on them with the
    library's MFMA kernels."""
    import dataclasses
    import torch
    from flash5_amd import learned as Lr
    from flash5_amd.hydro import Hydro
    t0 = time.perf_counter()
    hy0 = Hydro(dataclasses.replace(hp, learnedSeed=False))
    Uc = U.clone()
    hy0.advance_batch(Uc, dt, (dx,) * 3)
    S = hy0.face_states()
    del Uc
    gen = torch.Generator(device=S.device).manual_seed(7)
    idx = torch.randperm(S.shape[1], generator=gen, device=S.device)[:n_train]
    S = S[:, idx].contiguous()
    ps, _, _ = Lr.solve(S, gamma)
    net = Lr.SeedNet(seed=3)
    hist = net.train(net.features(S, gamma), Lr.targets(S, ps, gamma), steps=steps, batch=8192,
                     lr=3e-3, log_every=steps // 4)
    torch.cuda.synchronize()
    return {"params": net.params, "train_states": int(S.shape[1]), "steps": steps,
            "loss_first_last": [hist[0][1], hist[-1][1]],
            "train_s": time.perf_counter() - t0}


def mixed_learned_leg(U0, dt, dx, gamma, steps: int, warmup: int, plain_value: float):
    """The headline comparison on the same batch: FP32 reconstruction and
    fluxes with the FP64 update, the exact solver seeded by the network
    (trained untimed on this batch's own face states), launch shapes
    autotuned; K timed steps from the initial state.  Its density L1 after
    one step is taken against the FP64 exact-solver step of the same dt."""
    import dataclasses
    import torch
    from flash5_amd.hydro import PREC_MIXED, Hydro, HydroParams
    n = 16
    base = HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=gamma, RiemannSolver="exact")
    info = train_seed(base, U0, dt, dx, gamma)
    hp = dataclasses.replace(base, precision=PREC_MIXED, learnedSeed=True)
    hy = Hydro(hp)
    hy.set_seed_params(info.pop("params"))
    stream = torch.cuda.current_stream()
    hy.set_stream(stream)
    # precision: one step of each from the initial state, same dt
    U = U0.clone()
    hy.advance_batch(U, dt, (dx,) * 3)
    ref = U0.clone()
    Hydro(base).advance_batch(ref, dt, (dx,) * 3)
    g = 4
    sl = (slice(None), slice(None), slice(g, g + n), slice(g, g + n), slice(g, g + n))
    l1 = {}
    for name, v in (("dens", 0), ("velx", 1), ("pres", 4), ("ener", 5), ("eint", 9)):
        a, b = U[sl][:, v], ref[sl][:, v]
        l1[name] = float((a - b).abs().sum() / b.abs().sum())
    del ref
    hy = Hydro(hp)
    hy.set_stream(stream)
    U.copy_(U0)
    tune = hy.autotune_batch(U, dt, (dx,) * 3, reps=3)
    for _ in range(warmup):
        hy.advance_batch(U, dt, (dx,) * 3)
    U.copy_(U0)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(steps):
        hy.advance_batch(U, dt, (dx,) * 3)
    end.record(stream)
    torch.cuda.synchronize()
    ms = start.elapsed_time(end) / steps
    value = U.shape[0] * n ** 3 / (ms * 1e-3)
    st = hy.last_status
    hy.autotune_reset()
    return {"value": value, "unit": "zone-updates/s", "ms_per_step": ms,
            "vs_plain_fp64": value / plain_value,
            "l1_rel_vs_fp64_exact_one_step": l1,
            "dtype": "f32+f64", "riemann": "exact", "learned_seed": info,
            "newton_iters_per_face": int(st.newton_iters) / (3 * U.shape[0] * 17 * 16 * 16),
            "newton_fallbacks": int(st.newton_fallbacks),
            "autotune": [t for t in tune if t["chosen"]]}


def cpu_baseline(U_gpu, dt, dx, riemann: str = "hllc", target_s: float = 12.0):
    """Time the FP64 oracle (OpenMP over blocks) on a bounded sample of the
    batch: as many whole rounds of the sample as fit in about target_s, the
    median per-round rate reported.  The hybrid solver's rounds include the
    oracle's shock detection."""
    import numpy as np
    import oracle as O
    model, nproc, usable = host_cpu()
    # every CPU the process may use, unless OMP_NUM_THREADS sets the host's share
    nthreads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or usable
    if riemann in ("hll", "hybrid"):
        p = O.params(3, 16, 5.0 / 3.0, nunk=11, shokVar=O.SHOK,
                     riemann={"hll": O.HLL, "hybrid": O.HYBR}[riemann])
    else:
        p = O.params(3, 16, 5.0 / 3.0)
    nb = nthreads                       # calibration: one block per thread
    U = U_gpu[:nb].cpu().numpy().copy()
    t0 = time.perf_counter()
    O.compute_fluxes_batch(p, U, dt, (dx,) * 3, nthreads)
    per_block_round = max(time.perf_counter() - t0, 1e-3)
    nb = min(U_gpu.shape[0], nthreads * max(1, int(0.3 * target_s / per_block_round)))
    U0 = U_gpu[:nb].cpu().numpy()
    rates, spent = [], 0.0
    while spent < target_s or len(rates) < 3:
        U = U0.copy()
        t0 = time.perf_counter()
        rc, _ = O.compute_fluxes_batch(p, U, dt, (dx,) * 3, nthreads)
        t = time.perf_counter() - t0
        spent += t
        rates.append(nb * 16 ** 3 / t)
        if len(rates) >= 20:
            break
    return {"value": float(np.median(rates)), "unit": "zone-updates/s", "cores": nthreads,
            "kind": "port", "cpu_model": model, "nproc": nproc, "affinity_cpus": usable,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"{nb} of the 4096 blocks (same seeded fields), one step each, "
                      f"median of {len(rates)} rounds, FP64 C oracle, OpenMP over blocks on "
                      f"{nthreads} threads ({usable} CPUs usable of {nproc}), "
                      f"{spent:.1f} s CPU wall"}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--blocks", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--riemann", choices=["hllc", "hll", "hybrid", "exact"], default="hllc",
                    help="face flux solver: HLLC, HLL, the hybrid switch (HLL next to "
                         "cells the shock detector flags, HLLC elsewhere; the blocks carry a "
                         "shock variable and each step flags them first), or the exact "
                         "iterative solver")
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64",
                    help="fp64, or FP32 reconstruction/fluxes with the FP64 update")
    ap.add_argument("--learned", action="store_true",
                    help="seed the exact solver's Newton iteration with the network, trained "
                         "(untimed) on face states of this workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run plain 1 GiB copies (PMC byte-counter calibration)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in launch shapes instead of timing the candidates")
    ap.add_argument("--no-mixed-leg", action="store_true",
                    help="skip the mixed-precision + learned-seed leg of the default run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}); "
              "the path shards by block across GPUs with no exchange but the dt min-reduction",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from flash5_amd import workloads as W
    from flash5_amd.hydro import PREC_FP64, PREC_MIXED, Hydro, HydroParams

    dev = torch.device("cuda:0")
    n, g, gamma = 16, 4, 5.0 / 3.0
    dx = 1.0 / n
    prec = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[args.precision]
    hp = HydroParams(ndim=3, nxb=n, nyb=n, nzb=n, gamma=gamma, RiemannSolver=args.riemann,
                     precision=prec)
    U = W.fourier_blocks(args.blocks, n=n, ng=g, gamma=gamma, device=dev)
    cfl = 0.5 if args.riemann == "hll" else 0.8
    dt = W.cfl_dt(U, 3, g, (dx,) * 3, cfl)
    hybrid = args.riemann == "hybrid"
    if args.riemann in ("hll", "hybrid"):      # a SHOK_VAR slot after the ten hydro variables
        hp.nunk, hp.shokVar = 11, 10
        U = torch.cat([U, torch.zeros_like(U[:, :1])], dim=1).contiguous()
    default_run = (args.precision == "fp64" and args.riemann == "hllc" and not args.learned)
    U0 = U.clone() if default_run and not args.no_mixed_leg else None
    train_info = None
    if args.learned:
        if args.riemann != "exact":
            print("bench.py: --learned needs --riemann exact", file=sys.stderr)
            sys.exit(2)
        train_info = train_seed(hp, U, dt, dx, gamma)
        hp.learnedSeed = True
    hy = Hydro(hp)
    if args.learned:
        hy.set_seed_params(train_info.pop("params"))
    stream = torch.cuda.current_stream(dev)
    hy.set_stream(stream)
    tune = None
    if not args.no_autotune:   # untimed: candidate launch shapes timed on this batch
        tune = hy.autotune_batch(U, dt, (dx,) * 3, reps=3)
    for _ in range(args.warmup):
        if hybrid:
            hy.shock_detect_batch(U, (dx,) * 3)
        hy.advance_batch(U, dt, (dx,) * 3)
    torch.cuda.synchronize()

    # timed region: exactly K steps (with the hybrid solver, each step flags
    # shocks first, as Hydro does)
    hy.set_timing(True)
    stage_sum = [0.0, 0.0, 0.0]
    detect_ms = 0.0
    d0 = torch.cuda.Event(enable_timing=True)
    d1 = torch.cuda.Event(enable_timing=True)
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        if hybrid:
            d0.record(stream)
            hy.shock_detect_batch(U, (dx,) * 3)
            d1.record(stream)
        hy.advance_batch(U, dt, (dx,) * 3)
        if hybrid:
            d1.synchronize()
            detect_ms += d0.elapsed_time(d1)
        st = hy.stage_times()
        for k in range(3):
            stage_sum[k] += st[k]
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    hy.set_timing(False)
    ms_total = start.elapsed_time(end)
    ms_step = ms_total / args.steps
    zones = args.blocks * n ** 3
    value = zones / (ms_step * 1e-3)
    stage_ms = [s / args.steps for s in stage_sum]
    march_ms = stage_ms[0]
    bz = bytes_per_zone(3, n, g)
    bmz = march_bytes_per_zone(3, n, g)
    achieved = bmz * zones / (march_ms * 1e-3) / 1e9
    fallbacks = int(hy.last_status.fallbacks)
    workload = (f"{args.blocks}x16^3+4g {args.precision} {args.riemann}"
                + (" learned" if args.learned else ""))
    flops_zone = measured_flops_per_zone("k_march", workload)
    flagged = float((U[:, 10, g:g + n, g:g + n, g:g + n] > 0).double().mean()) if hybrid else None
    if args.copy_probe:
        copy_probe()
    mixed = None
    if U0 is not None:
        hy.autotune_reset()
        del U
        torch.cuda.empty_cache()
        mixed = mixed_learned_leg(U0, dt, dx, gamma, args.steps, args.warmup, value)
        U = U0

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(U, dt, dx, args.riemann)
        except Exception as e:  # report, do not hide
            cpu = {"value": None, "unit": "zone-updates/s", "cores": 0, "kind": "port",
                   "sample": f"failed: {e}"}

    line = {
        "metric": "zone-updates/s, mixed+learned vs plain FP64, at L1 rel err<=1e-6; % HBM peak",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.blocks} independent 3-D blocks, 16^3 interior + 4 guard "
                               "cells, gamma=5/3, seeded Fourier-mode fields, one unsplit "
                               f"PPM+{args.riemann.upper()} step ({args.precision}"
                               + (", learned Newton seed)" if args.learned else ")"),
                   "blocks": args.blocks, "block": [16, 16, 16], "nguard": 4,
                   "precision": args.precision, "riemann": args.riemann, "dt": dt, "cfl": cfl,
                   "newton_iters_per_face": (int(hy.last_status.newton_iters) /
                                             (3 * args.blocks * 17 * 16 * 16)),
                   "newton_fallbacks": int(hy.last_status.newton_fallbacks),
                   "flagged_cell_fraction": flagged,
                   "learned_seed": train_info, "autotune": tune},
        "roofline": {"bound": "fp64-valu" if flops_zone else "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "compute": None if not flops_zone else {
                         "achieved": flops_zone * zones / (march_ms * 1e-3) / 1e12,
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": flops_zone * zones / (march_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                         "fp64_flops_per_zone": flops_zone,
                         "source": "SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 per dispatch, "
                                   "profiles/traffic.json"},
                     "traffic": measured_traffic("k_march", workload),
                     "kernel": "k_march (face states + fluxes, fused column march)",
                     "kernel_ms": march_ms, "bytes_per_zone": bmz, "zones_per_launch": zones,
                     "stage_ms": {"k_march": stage_ms[0], "riemann_stage": stage_ms[1],
                                  "k_update": stage_ms[2],
                                  "shock_detect": detect_ms / args.steps if hybrid else None},
                     "step_bytes_per_zone": bz,
                     "step_frac": bz * zones / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "mixed_learned": mixed,
        "fallbacks": fallbacks,
        "wall_s": wall,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
