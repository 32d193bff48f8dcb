"""Config 3 precision study against the FP64 oracle (needs a GPU).

The Kelvin-Helmholtz shear layer at 2048^2 cells as 16^2-cell blocks,
periodic, 200 steps, run three ways in lockstep:

* the CPU oracle (FP64, OpenMP over blocks): the reference run and its dt
  sequence;
* the GPU FP64 path with its own device dt: checked bit for bit against the
  oracle at every 20th step and at the end;
* the GPU mixed mode (FP32 reconstruction and fluxes, FP64 update) stepped
  with the oracle's dt sequence: per-variable sfocu L1 and Linf against the
  oracle at the end.

Test infrastructure, not collected by pytest (it takes minutes of host CPU).
Usage: python tests/kh_precision_oracle.py [--cells 2048] [--steps 200]
                                          [--threads 16] [--out FILE.md]
"""
from __future__ import annotations

import argparse
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from flash5_amd import drivers as D, sfocu  # noqa: E402
from flash5_amd.hydro import PREC_FP64, PREC_MIXED, HydroParams  # noqa: E402

NG = 4


def periodic_fill(U, nblk, nb):
    """Guard cells of a periodic 2-D grid of nblk x nblk blocks (block
    b = bi + nblk * bj): pad the global interior by wrapping, then cut each
    block's guard-inclusive window."""
    nbk, nunk, _, ny, nx = U.shape
    G = U[:, :, 0, NG:NG + nb, NG:NG + nb].reshape(nblk, nblk, nunk, nb, nb)
    G = G.transpose(2, 0, 3, 1, 4).reshape(nunk, nblk * nb, nblk * nb)
    Gp = np.pad(G, ((0, 0), (NG, NG), (NG, NG)), mode="wrap")
    win = np.lib.stride_tricks.sliding_window_view(Gp, (ny, nx), axis=(1, 2))[:, ::nb, ::nb]
    out = np.ascontiguousarray(win.transpose(1, 2, 0, 3, 4)).reshape(nbk, nunk, 1, ny, nx)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--cells", type=int, default=2048)
    ap.add_argument("--nb", type=int, default=16)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--out", default="")
    a = ap.parse_args()
    import torch

    nc, nb = a.cells, a.nb
    nblk = nc // nb
    deltas = (1.0 / nc, 1.0 / nc, 1.0)
    p = O.params(2, nb, 1.4)
    U0 = D.kh_blocks(nc, nb)
    nbr = D.uniform_neighbors((nblk, nblk), (True, True), 2)
    g64 = D.UniformGrid(HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4, precision=PREC_FP64),
                        U0, nbr, deltas)
    gmx = D.UniformGrid(HydroParams(ndim=2, nxb=nb, nyb=nb, nzb=1, gamma=1.4,
                                    precision=PREC_MIXED), U0, nbr, deltas)
    Uo = U0.copy()
    t_oracle = t_gpu64 = t_mixed = 0.0
    bitwise_steps = []
    t = 0.0
    for s in range(a.steps):
        t0 = time.perf_counter()
        Uo = periodic_fill(Uo, nblk, nb)
        dt = min(O.block_dt(p, Uo[b], deltas, 0.8) for b in range(Uo.shape[0]))
        rc, _ = O.compute_fluxes_batch(p, Uo, dt, deltas, nthreads=a.threads)
        if rc:
            raise SystemExit(f"oracle step {s} failed")
        t_oracle += time.perf_counter() - t0
        t += dt
        t0 = time.perf_counter()
        dt64 = g64.step()
        torch.cuda.synchronize()
        t_gpu64 += time.perf_counter() - t0
        if dt64 != dt:
            raise SystemExit(f"step {s}: GPU FP64 dt {dt64!r} != oracle dt {dt!r}")
        t0 = time.perf_counter()
        gmx.fill_guards()
        gmx.hydro.advance_batch(gmx.U, dt, deltas)
        gmx.t += dt
        torch.cuda.synchronize()
        t_mixed += time.perf_counter() - t0
        if (s + 1) % 20 == 0 or s + 1 == a.steps:
            same = bool(np.array_equal(g64.U.cpu().numpy()[:, :, :, NG:NG + nb, NG:NG + nb],
                                       Uo[:, :, :, NG:NG + nb, NG:NG + nb]))
            bitwise_steps.append((s + 1, same))
            print(f"step {s + 1}: t = {t:.6f}, GPU FP64 bitwise vs oracle: {same}", flush=True)
    names = {"dens": D.DENS, "velx": D.VELX, "vely": D.VELY, "pres": D.PRES, "ener": D.ENER,
             "eint": D.EINT, "temp": D.TEMP}
    norms = sfocu.compare(gmx.U.cpu().numpy(), Uo, 1, names, 2, NG)
    zones = nc * nc * a.steps
    lines = [
        f"# Config 3: Kelvin-Helmholtz {nc}^2 x {a.steps} steps, mixed mode against the FP64 oracle",
        "",
        f"Command: `python tests/kh_precision_oracle.py --cells {nc} --steps {a.steps} "
        f"--threads {a.threads}{' --out ' + a.out if a.out else ''}`",
        "",
        f"- Grid: {nc}^2 cells as {nblk}^2 blocks of {nb}^2, 4 guard cells, periodic, gamma 1.4, "
        "PPM + HLLC, CFL 0.8 (dt from the oracle).",
        f"- End time t = {t:.9f}.",
        f"- GPU FP64 path (own device dt) against the oracle, interior cells: bitwise at steps "
        f"{', '.join(str(s) for s, ok in bitwise_steps if ok) or 'none'}"
        f"{'' if all(ok for _, ok in bitwise_steps) else '; DIFFERS at ' + ', '.join(str(s) for s, ok in bitwise_steps if not ok)}.",
        "- Mixed mode (FP32 reconstruction and fluxes, FP64 update), stepped with the oracle's "
        "dt sequence, against the FP64 oracle (sfocu norms over all interior cells):",
        "",
        "| variable | L1 rel | Linf (mag err) |",
        "|---|---|---|",
    ]
    for k, (l1, linf) in norms.items():
        lines.append(f"| {k} | {l1:.3e} | {linf:.3e} |")
    lines += [
        "",
        f"- Wall time of the {a.steps} steps (guard fill, dt, update; host-timed with "
        "synchronisation): "
        f"oracle {t_oracle:.1f} s on {a.threads} threads ({zones / t_oracle:.3e} zone-updates/s; "
        f"host CPU {cpu_model()}, nproc {os.cpu_count()}), "
        f"GPU FP64 {t_gpu64:.2f} s ({zones / t_gpu64:.3e}), "
        f"GPU mixed {t_mixed:.2f} s ({zones / t_mixed:.3e}).",
        "",
    ]
    text = "\n".join(lines)
    print(text)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text)
    if not all(ok for _, ok in bitwise_steps):
        raise SystemExit(1)
    if norms["dens"][0] > 1e-6:
        raise SystemExit(f"density L1 {norms['dens'][0]:.3e} above 1e-6")


if __name__ == "__main__":
    main()
