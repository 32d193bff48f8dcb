"""Checkpoint files in FLASH's checkpoint layout, for sfocu-style comparison.

The datasets are those sfocu reads from a FLASH HDF5 checkpoint
(tools/sfocu/flash_reader_hdf5.c:403-496):

  "integer scalars"  (name, value) pairs: globalnumblocks, nxb, nyb, nzb,
                     dimensionality, nstep
  "real scalars"     (name, value) pairs: time, dt
  "unknown names"    (nvar, 1), four-character names
  "node type"        (nblocks,), 1 = leaf, 2 = parent
  "refine level"     (nblocks,), 1-based
  "coordinates"      (nblocks, 3), block centres
  "block size"       (nblocks, 3)
  "bounding box"     (nblocks, 3, 2)
  one dataset per variable, (nblocks, nzb, nyb, nxb), interior cells only.

This image has no HDF5 library, so a checkpoint is written as a .npz archive
holding those datasets under the same names (the scalar lists as parallel
name/value arrays).  With h5py importable, a path ending in .h5 or .hdf5 is
written as HDF5 with FLASH's dataset names instead.

FLASH's layout has one block shape.  Grids whose levels use different block
shapes are cut into sub-blocks of one common shape (per axis, the greatest
common divisor of the shapes), each with its own bounding box, refine level
and node type; cutting is lossless, so two grids cut the same way compare
block by block.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# the unsplit hydro's cell variables (unsplit/Config) in the layout's order
VARNAMES = ("dens", "velx", "vely", "velz", "pres", "ener", "gamc", "game", "temp", "eint",
            "shok")
LEAF, PARENT = 1, 2


@dataclass
class Checkpoint:
    ndim: int
    nb: tuple                       # interior cells per block (nxb, nyb, nzb)
    names: list                     # variable names, at most four characters
    lref: np.ndarray                # (nblocks,) refine level, 1-based
    node: np.ndarray                # (nblocks,) LEAF or PARENT
    coord: np.ndarray               # (nblocks, 3) block centres
    size: np.ndarray                # (nblocks, 3) block extents
    bbox: np.ndarray                # (nblocks, 3, 2) lower and upper corners
    data: dict = field(default_factory=dict)   # name -> (nblocks, nzb, nyb, nxb)
    time: float = 0.0
    dt: float = 0.0
    nstep: int = 0

    @property
    def nblocks(self) -> int:
        return int(self.lref.shape[0])


def _common_shape(shapes, ndim):
    out = []
    for a in range(3):
        g = 0
        for s in shapes:
            g = math.gcd(g, int(s[a]))
        out.append(g if a < ndim else 1)
    return tuple(out)


def from_blocks(ndim, arrays, levels, lower, dels, leaf, slots, names=None, ng=4,
                time=0.0, dt=0.0, nstep=0) -> Checkpoint:
    """Checkpoint of guard-inclusive blocks: arrays[b] is (nunk, nz, ny, nx)
    (numpy or torch), levels[b] 1-based, lower[b] the block's lower corner,
    dels[b] its cell widths, leaf[b] whether it is a leaf; slots[k] is the
    var-axis slot of names[k]."""
    names = list(names if names is not None else VARNAMES[:len(slots)])
    shapes = []
    blocks = []
    for b, A in enumerate(arrays):
        A = A.detach().cpu().numpy() if hasattr(A, "detach") else np.asarray(A)
        nz = A.shape[1] - 2 * ng if ndim > 2 else 1
        ny = A.shape[2] - 2 * ng if ndim > 1 else 1
        nx = A.shape[3] - 2 * ng
        shapes.append((nx, ny, nz))
        blocks.append(A)
    c = _common_shape(shapes, ndim)
    lref, node, coord, size, bbox = [], [], [], [], []
    data = {n: [] for n in names}
    for b, A in enumerate(blocks):
        nx, ny, nz = shapes[b]
        d = [float(x) for x in dels[b]] + [0.0] * (3 - len(dels[b]))
        lo = [float(x) for x in lower[b]] + [0.0] * (3 - len(lower[b]))
        gz, gy, gx = (ng if ndim > 2 else 0), (ng if ndim > 1 else 0), ng
        for kz in range(nz // c[2]):
            for ky in range(ny // c[1]):
                for kx in range(nx // c[0]):
                    off = (kx * c[0], ky * c[1], kz * c[2])
                    box = np.zeros((3, 2))
                    for a in range(3):
                        if a < ndim:
                            box[a, 0] = lo[a] + off[a] * d[a]
                            box[a, 1] = lo[a] + (off[a] + c[a]) * d[a]
                    lref.append(int(levels[b]))
                    node.append(LEAF if leaf[b] else PARENT)
                    bbox.append(box)
                    coord.append(0.5 * (box[:, 0] + box[:, 1]))
                    size.append(box[:, 1] - box[:, 0])
                    sl = (slice(gz + off[2], gz + off[2] + c[2]),
                          slice(gy + off[1], gy + off[1] + c[1]),
                          slice(gx + off[0], gx + off[0] + c[0]))
                    for n, v in zip(names, slots):
                        data[n].append(A[(v,) + sl])
    return Checkpoint(ndim=ndim, nb=c, names=names, lref=np.array(lref, dtype=np.int32),
                      node=np.array(node, dtype=np.int32), coord=np.array(coord),
                      size=np.array(size), bbox=np.array(bbox),
                      data={n: np.ascontiguousarray(np.stack(v)) for n, v in data.items()},
                      time=float(time), dt=float(dt), nstep=int(nstep))


def from_level_storage(grid, storage, slots=None, names=None, **kw) -> Checkpoint:
    """A LevelGrid held in a LevelStorage (flash5_amd.drivers): levels
    1-based, blocks covered by a finer level are parents."""
    masks = [m.cpu().numpy() for m in storage.leaf_masks()]
    host = None
    if hasattr(storage, "levels"):      # one device-to-host copy per level
        host = [t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
                for t in storage.levels]
    arrays, levels, lower, dels, leaf = [], [], [], [], []
    for b, (L, pos) in enumerate(grid.blocks):
        d = grid.dels(L)
        nb = grid.nb[L]
        arrays.append(host[L][storage.index[b][1]] if host is not None else storage.block(b))
        levels.append(L + 1)
        lower.append([pos[a] * nb[a] * d[a] for a in range(3)])
        dels.append(d)
        leaf.append(bool(masks[L][storage.index[b][1]]))
    nunk = arrays[0].shape[0]
    slots = list(slots if slots is not None else range(min(nunk, len(VARNAMES))))
    return from_blocks(grid.ndim, arrays, levels, lower, dels, leaf, slots, names, grid.ng, **kw)


def from_uniform(U, nblk, ndim, deltas, slots=None, names=None, ng=4, **kw) -> Checkpoint:
    """A uniform grid of blocks (nblocks, nunk, nz, ny, nx) numbered with x
    fastest over nblk = (bx, by, bz) blocks per axis; all leaves, level 1."""
    nblk = tuple(nblk) + (1,) * (3 - len(nblk))
    U = U.detach().cpu().numpy() if hasattr(U, "detach") else np.asarray(U)
    nx = U.shape[4] - 2 * ng
    ny = U.shape[3] - 2 * ng if ndim > 1 else 1
    nz = U.shape[2] - 2 * ng if ndim > 2 else 1
    nb = (nx, ny, nz)
    lower = []
    for b in range(U.shape[0]):
        p = (b % nblk[0], (b // nblk[0]) % nblk[1], b // (nblk[0] * nblk[1]))
        lower.append([p[a] * nb[a] * deltas[a] for a in range(3)])
    slots = list(slots if slots is not None else range(min(U.shape[1], len(VARNAMES))))
    return from_blocks(ndim, list(U), [1] * U.shape[0], lower, [deltas] * U.shape[0],
                       [True] * U.shape[0], slots, names, ng, **kw)


def _scalars(ck: Checkpoint):
    ints = [("globalnumblocks", ck.nblocks), ("nxb", ck.nb[0]), ("nyb", ck.nb[1]),
            ("nzb", ck.nb[2]), ("dimensionality", ck.ndim), ("nstep", ck.nstep)]
    reals = [("time", ck.time), ("dt", ck.dt)]
    return ints, reals


def write(path: str, ck: Checkpoint) -> None:
    ints, reals = _scalars(ck)
    if path.endswith((".h5", ".hdf5")):
        import h5py                                  # only when the library is present
        with h5py.File(path, "w") as f:
            f["integer scalars"] = np.array([(n.ljust(80).encode(), v) for n, v in ints],
                                            dtype=[("name", "S80"), ("value", "<i4")])
            f["real scalars"] = np.array([(n.ljust(80).encode(), v) for n, v in reals],
                                         dtype=[("name", "S80"), ("value", "<f8")])
            f["unknown names"] = np.array([[n.ljust(4).encode()] for n in ck.names], dtype="S4")
            f["node type"] = ck.node
            f["refine level"] = ck.lref
            f["coordinates"] = ck.coord
            f["block size"] = ck.size
            f["bounding box"] = ck.bbox
            for n in ck.names:
                f[n.ljust(4)] = ck.data[n]
        return
    arrays = {
        "integer scalars/name": np.array([n for n, _ in ints]),
        "integer scalars/value": np.array([v for _, v in ints], dtype=np.int64),
        "real scalars/name": np.array([n for n, _ in reals]),
        "real scalars/value": np.array([v for _, v in reals], dtype=np.float64),
        "unknown names": np.array([[n.ljust(4)] for n in ck.names]),
        "node type": ck.node, "refine level": ck.lref, "coordinates": ck.coord,
        "block size": ck.size, "bounding box": ck.bbox,
    }
    for n in ck.names:
        arrays[n.ljust(4)] = ck.data[n]
    with open(path, "wb") as fh:
        np.savez(fh, **arrays)


def read(path: str) -> Checkpoint:
    if path.endswith((".h5", ".hdf5")):
        import h5py
        with h5py.File(path, "r") as f:
            ints = {r["name"].decode().strip(): int(r["value"]) for r in f["integer scalars"][()]}
            reals = {r["name"].decode().strip(): float(r["value"]) for r in f["real scalars"][()]}
            names = [x[0].decode().strip() for x in f["unknown names"][()]]
            ds = {k: f[k][()] for k in ("node type", "refine level", "coordinates", "block size",
                                        "bounding box")}
            data = {n: f[n.ljust(4)][()] for n in names}
    else:
        with np.load(path, allow_pickle=False) as z:
            ints = dict(zip([str(x) for x in z["integer scalars/name"]],
                            [int(x) for x in z["integer scalars/value"]]))
            reals = dict(zip([str(x) for x in z["real scalars/name"]],
                             [float(x) for x in z["real scalars/value"]]))
            names = [str(x[0]).strip() for x in z["unknown names"]]
            ds = {k: z[k] for k in ("node type", "refine level", "coordinates", "block size",
                                    "bounding box")}
            data = {n: z[n.ljust(4)] for n in names}
    return Checkpoint(ndim=ints["dimensionality"], nb=(ints["nxb"], ints["nyb"], ints["nzb"]),
                      names=names, lref=ds["refine level"], node=ds["node type"],
                      coord=ds["coordinates"], size=ds["block size"], bbox=ds["bounding box"],
                      data=data, time=reals.get("time", 0.0), dt=reals.get("dt", 0.0),
                      nstep=ints.get("nstep", 0))

