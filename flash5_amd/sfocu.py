"""sfocu-style comparison norms (tools/sfocu/sfocu.c:31-33, 812-884).

For each variable over interior leaf cells, unweighted:
  * L1 (sfocu's default ``-L1``, benchmark B): sum|B - A| / max(sum|B|, 1e-99)
  * Linf ("magErr"): sup|A - B| / max(max(|A|, |B|) over cells, 1e-99)
  * per-cell norm 2|a - b| / |a + b| (sfocu.c:31-33), reported as its max
"""
from __future__ import annotations

import numpy as np


def interior(U: np.ndarray, ndim: int, ng: int) -> np.ndarray:
    """View of the interior cells of (..., nz, ny, nx) block data."""
    sl = [slice(None)] * U.ndim
    for a in range(ndim):
        sl[U.ndim - 1 - a] = slice(ng, U.shape[U.ndim - 1 - a] - ng)
    return U[tuple(sl)]


def l1_rel(a: np.ndarray, b: np.ndarray) -> float:
    return float(np.abs(b - a).sum() / max(np.abs(b).sum(), 1e-99))


def linf_rel(a: np.ndarray, b: np.ndarray) -> float:
    scale = max(float(np.abs(a).max(initial=0.0)), float(np.abs(b).max(initial=0.0)), 1e-99)
    return float(np.abs(a - b).max(initial=0.0) / scale)


def cell_norm_max(a: np.ndarray, b: np.ndarray) -> float:
    den = np.abs(a + b)
    num = 2.0 * np.abs(a - b)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    return float(r.max(initial=0.0))


def compare(A: np.ndarray, B: np.ndarray, var_axis: int, names, ndim: int, ng: int):
    """Per-variable (L1, Linf) of block data A against benchmark B."""
    out = {}
    for n, v in names.items():
        a = interior(np.take(A, v, axis=var_axis), ndim, ng)
        b = interior(np.take(B, v, axis=var_axis), ndim, ng)
        out[n] = (l1_rel(a, b), linf_rel(a, b))
    return out


# ---------------------------------------------------------------------------
# File-level comparison of two checkpoints (tools/sfocu/sfocu.c:360-990,
# sameblock.c, namecmp.c), on flash5_amd.checkpoint.Checkpoint objects.
# ---------------------------------------------------------------------------
def _sameblock(dim, A, B, na, nb_, mesh_tol, ignore_lref):
    """sameblock.c: same refine level (unless ignored), then bit-equal
    bounding box, centre and size, or a centre distance below mesh_tol in
    units of the summed block sizes."""
    if not ignore_lref and A.lref[na] != B.lref[nb_]:
        return False
    if mesh_tol == 0.0:
        return (np.array_equal(A.bbox[na, :dim], B.bbox[nb_, :dim]) and
                np.array_equal(A.coord[na, :dim], B.coord[nb_, :dim]) and
                np.array_equal(A.size[na, :dim], B.size[nb_, :dim]))
    d = ((A.coord[na, :dim] - B.coord[nb_, :dim]) ** 2 /
         (A.size[na, :dim] + B.size[nb_, :dim]) ** 2).sum()
    return bool(np.sqrt(d) < mesh_tol)


def _strip(name):
    return "".join(str(name).split())


def compare_checkpoints(A, B, err_tol: float = 0.0, mesh_tol: float = 0.0,
                        ignore_lref: bool = False, ignore=(), extra=(), norm_order: int = 1,
                        benchmark: str = "B") -> dict:
    """sfocu's comparison of checkpoint A against B.  Leaf blocks of A are
    matched to blocks of B (search order b, b-1, b+1, b-2, ...); for every
    variable in both: bad blocks, min/max of 2|a-b|/max(|a+b|, 1e-99), the
    largest |a-b|, the Ln norm ||B-A|| / max(||bench||, 1e-99) and the mag
    error sup|a-b| / max(sup|a|, sup|b|, 1e-99).  The comparison fails on
    unmatched leaf blocks, on a variable in one file only, or on a mag error
    above err_tol."""
    rep = {"fail": False, "messages": [], "vars": {}, "blocks_compared": 0,
           "extra_blocks_a": 0, "extra_blocks_b": 0}
    if tuple(A.nb) != tuple(B.nb):
        rep["fail"] = True
        rep["messages"].append(f"Block shapes do not match: {list(A.nb)} != {list(B.nb)}")
        return rep
    dim = A.ndim
    nA, nB = A.nblocks, B.nblocks
    a2b = np.full(nA, -1, dtype=np.int64)
    for ba in range(nA):
        if A.node[ba] != 1:
            continue
        for i in range(nB):
            bb = (ba + nB + ((i + 1) >> 1) * (-1 if i & 1 else 1)) % nB
            if _sameblock(dim, A, B, ba, bb, mesh_tol, ignore_lref):
                a2b[ba] = bb
                rep["blocks_compared"] += 1
                break
        else:
            rep["extra_blocks_a"] += 1
    nodeB = np.array(B.node).copy()
    nodeB[a2b[a2b >= 0]] = -1
    rep["extra_blocks_b"] = int((nodeB == 1).sum())
    if rep["extra_blocks_a"] or rep["extra_blocks_b"]:
        rep["fail"] = True
    ign = {_strip(v) for v in ignore}
    names_a = [] if "*" in ign else [n for n in A.names if _strip(n) not in ign]
    names_b = [] if "*" in ign else [n for n in B.names if _strip(n) not in ign]
    for e in extra:
        if _strip(e) not in map(_strip, names_a):
            names_a.append(e)
        if _strip(e) not in map(_strip, names_b):
            names_b.append(e)
    common = [n for n in names_a if _strip(n) in map(_strip, names_b)]
    for n in names_a:
        if _strip(n) not in map(_strip, names_b):
            rep["fail"] = True
            rep["messages"].append(f"{n} present in A but not in B")
    for n in names_b:
        if _strip(n) not in map(_strip, names_a):
            rep["fail"] = True
            rep["messages"].append(f"{n} present in B but not in A")
    p = float(norm_order)
    pairs = [(ba, int(a2b[ba])) for ba in range(nA) if a2b[ba] >= 0]
    for n in common:
        key_b = next(m for m in B.names if _strip(m) == _strip(n))
        v = {"bad_blocks": 0, "min_error": 0.0, "max_error": 0.0, "max_error_vals": (0.0, 0.0),
             "min_error_vals": (0.0, 0.0), "abs_error": 0.0, "abs_error_vals": (0.0, 0.0),
             "sum_a": 0.0, "max_a": 0.0, "min_a": 0.0, "sum_b": 0.0, "max_b": 0.0, "min_b": 0.0}
        diff_sum = bench_sum = 0.0
        first = True
        for ba, bb in pairs:
            a = np.asarray(A.data[n][ba], dtype=np.float64).ravel()
            b = np.asarray(B.data[key_b][bb], dtype=np.float64).ravel()
            if first:
                first = False
                v["max_a"] = v["min_a"] = float(a[0])
                v["max_b"] = v["min_b"] = float(b[0])
                v["min_error"] = float(cell_norm(a[0], b[0])) if a[0] != b[0] else 0.0
            ne = a != b
            bench = np.abs(a if benchmark == "A" else b) ** p
            bench_sum += float(bench.sum())
            if ne.any():
                v["bad_blocks"] += 1
                da, db = a[ne], b[ne]
                err = cell_norm(da, db)
                diff_sum += float((np.abs(da - db) ** p).sum())
                k = int(np.argmax(err))
                if err[k] > v["max_error"]:
                    v["max_error"], v["max_error_vals"] = float(err[k]), (float(da[k]), float(db[k]))
                k = int(np.argmin(err))
                if err[k] < v["min_error"]:
                    v["min_error"], v["min_error_vals"] = float(err[k]), (float(da[k]), float(db[k]))
                ad = np.abs(da - db)
                k = int(np.argmax(ad))
                if ad[k] > v["abs_error"]:
                    v["abs_error"], v["abs_error_vals"] = float(ad[k]), (float(da[k]), float(db[k]))
            v["sum_a"] += float(a.sum())
            v["sum_b"] += float(b.sum())
            v["max_a"], v["min_a"] = max(v["max_a"], float(a.max())), min(v["min_a"], float(a.min()))
            v["max_b"], v["min_b"] = max(v["max_b"], float(b.max())), min(v["min_b"], float(b.min()))
        scale = max(abs(v["max_a"]), abs(v["min_a"]), abs(v["max_b"]), abs(v["min_b"]), 1e-99)
        v["mag_error"] = v["abs_error"] / scale
        r = bench_sum ** (1.0 / p)
        v["ln_norm"] = diff_sum ** (1.0 / p) / (r if r > 1e-99 else 1e-99)
        if v["mag_error"] > err_tol:
            rep["fail"] = True
        rep["vars"][_strip(n)] = v
    return rep


def cell_norm(a, b):
    """sfocu.c:31-33: 2|a - b| / max(|a + b|, 1e-99)."""
    return np.abs(2 * (a - b)) / np.maximum(np.abs(a + b), 1e-99)


def report_text(rep: dict, norm_order: int = 1, benchmark: str = "B") -> str:
    """The report sfocu prints (sfocu.c:891-990), ending SUCCESS or FAILURE."""
    out = list(rep["messages"])
    if rep["extra_blocks_a"]:
        out.append(f"A has {rep['extra_blocks_a']} leaf blocks that don't exist in B")
    if rep["extra_blocks_b"]:
        out.append(f"B has {rep['extra_blocks_b']} leaf blocks that don't exist in A")
    out.append(f"Total leaf blocks compared: {rep['blocks_compared']} (all other blocks are ignored)")
    out.append("Var  | Bad Blocks | Min Error || Max Error |     A     |     B     "
               "|| Abs Error |     A     |     B     |")
    for n, v in rep["vars"].items():
        out.append(f"{n:<4} | {v['bad_blocks']:<10d} | {v['min_error']:<9.4g} || "
                   f"{v['max_error']:<9.4g} | {v['max_error_vals'][0]: 9.3g} | "
                   f"{v['max_error_vals'][1]: 9.3g} || {v['abs_error']:<9.4g} | "
                   f"{v['abs_error_vals'][0]: 9.3g} | {v['abs_error_vals'][1]: 9.3g} |")
    out.append(f"Var  | L{norm_order}-ErrNorm | Mag Error ||    Sum A  |    Max A  |    Min A  "
               f"||    Sum B  |    Max B  |    Min B  |   (Rel / {benchmark})")
    for n, v in rep["vars"].items():
        out.append(f"{n:<4} | {v['ln_norm']:<10.4g} | {v['mag_error']:<9.4g} || "
                   f"{v['sum_a']: 9.3g} | {v['max_a']: 9.3g} | {v['min_a']: 9.3g} || "
                   f"{v['sum_b']: 9.3g} | {v['max_b']: 9.3g} | {v['min_b']: 9.3g} |")
    out.append("FAILURE" if rep["fail"] else "SUCCESS")
    return "\n".join(out)


def main(argv=None) -> int:
    """python -m flash5_amd.sfocu [-t mesh_tol] [-e err_tol] [-i] [-I var]...
    [-s var]... [-L n] [-b A|B] file1 file2 (the sfocu options)."""
    import argparse
    from . import checkpoint as ck
    ap = argparse.ArgumentParser(prog="sfocu")
    ap.add_argument("-t", dest="mesh_tol", type=float, default=0.0)
    ap.add_argument("-e", dest="err_tol", type=float, default=0.0)
    ap.add_argument("-i", dest="ignore_lref", action="store_true")
    ap.add_argument("-I", dest="ignore", action="append", default=[])
    ap.add_argument("-s", dest="extra", action="append", default=[])
    ap.add_argument("-L", dest="norm_order", type=int, default=1)
    ap.add_argument("-b", dest="benchmark", default="B")
    ap.add_argument("file1")
    ap.add_argument("file2")
    a = ap.parse_args(argv)
    bench = a.benchmark.upper()[:1]
    rep = compare_checkpoints(ck.read(a.file1), ck.read(a.file2), a.err_tol, a.mesh_tol,
                              a.ignore_lref, a.ignore, a.extra, a.norm_order, bench)
    print(report_text(rep, a.norm_order, bench))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
