"""Block bookkeeping for octree-refined meshes of equal-size blocks.

FLASH keeps this in its Grid (AMReX BoxArrays per level, or Paramesh's
tree).  The standalone drivers and tests need the same facts: which blocks
exist on each level, who neighbours whom (with periodic wrap), and which
coarse faces lie on a fine-coarse boundary.  The latter is the face list
``fx_fluxreg_put`` / ``fx_fluxreg_conserve`` take (as 9-tuples
(coarse_tile, fine_tile, dir, ci, cj, ck, fi, fj, fk), tile numbers in the
order of ``blocks(level)``).  It reproduces AMReX's
flux-register coverage for refinement ratio 2: every copy of a coarse face on
the boundary of the fine level's region is listed, including faces across a
periodic boundary, and none on a non-periodic domain boundary
(TestFluxCorrection's Driver_evolveFlash.F90 states these cases).

Blocks are named (level, (bi, bj, bk)) with 1-based levels; a refined block
at (L, b) has the 2^d children (L+1, 2b + o), o in {0,1}^d.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field


@dataclass
class BlockMesh:
    ndim: int
    nb: int                                   # interior cells per block and axis
    base: tuple                               # blocks per axis on level 1
    periodic: tuple = (False, False, False)
    levels: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.nb % 2:
            raise ValueError("block size must be even for refinement ratio 2")
        base = tuple(self.base) + (1,) * (3 - len(self.base))
        self.base = tuple(base[d] if d < self.ndim else 1 for d in range(3))
        self.periodic = tuple(self.periodic) + (False,) * (3 - len(self.periodic))
        self.levels = {1: set(itertools.product(*[range(n) for n in self.base]))}

    # ---- structure ----
    def extent(self, level: int):
        return tuple(self.base[d] * 2 ** (level - 1) if d < self.ndim else 1 for d in range(3))

    def exists(self, level: int, b) -> bool:
        return tuple(b) in self.levels.get(level, ())

    def children(self, level: int, b):
        offs = itertools.product(*[(0, 1) if d < self.ndim else (0,) for d in range(3)])
        return [tuple(2 * b[d] + o[d] for d in range(3)) for o in offs]

    def refine(self, level: int, b) -> None:
        b = tuple(b)
        if not self.exists(level, b):
            raise ValueError(f"no block {b} on level {level}")
        self.levels.setdefault(level + 1, set()).update(self.children(level, b))

    def is_refined(self, level: int, b) -> bool:
        return any(self.exists(level + 1, c) for c in self.children(level, b))

    def finest(self) -> int:
        return max(self.levels)

    def blocks(self, level: int):
        """Tile order of a level: sorted by (bk, bj, bi)."""
        return sorted(self.levels.get(level, ()), key=lambda b: (b[2], b[1], b[0]))

    def leaves(self, level: int):
        return [b for b in self.blocks(level) if not self.is_refined(level, b)]

    def neighbor(self, level: int, b, d: int, side: int):
        """Block index across face side (0 low, 1 high) along axis d, with
        periodic wrap; None across a non-periodic domain boundary."""
        n = list(b)
        n[d] += 1 if side else -1
        ext = self.extent(level)[d]
        if n[d] < 0 or n[d] >= ext:
            if not self.periodic[d]:
                return None
            n[d] %= ext
        return tuple(n)

    # ---- flux register faces between level L (coarse) and L+1 ----
    def register_faces(self, coarse_level: int):
        L = coarse_level
        ctiles = {b: t for t, b in enumerate(self.blocks(L))}
        ftiles = {b: t for t, b in enumerate(self.blocks(L + 1))}
        nb, h = self.nb, self.nb // 2
        out = []
        for P in self.blocks(L):
            if not self.is_refined(L, P):
                continue
            for d in range(self.ndim):
                for side in (0, 1):
                    N = self.neighbor(L, P, d, side)
                    if N is None or not self.exists(L, N) or self.is_refined(L, N):
                        continue
                    tr = [a for a in range(3) if a != d]
                    rng = [range(nb) if a < self.ndim else range(1) for a in tr]
                    for t0 in rng[0]:
                        for t1 in rng[1]:
                            tv = {tr[0]: t0, tr[1]: t1}
                            child = [0, 0, 0]
                            fstart = [0, 0, 0]
                            child[d] = 2 * P[d] + side
                            fstart[d] = nb if side else 0
                            for a in tr:
                                if a < self.ndim:
                                    child[a] = 2 * P[a] + tv[a] // h
                                    fstart[a] = 2 * (tv[a] % h)
                                else:
                                    child[a] = P[a]
                            ft = ftiles[tuple(child)]
                            for cb, face_d in ((P, nb if side else 0), (N, 0 if side else nb)):
                                cidx = [0, 0, 0]
                                cidx[d] = face_d
                                for a in tr:
                                    cidx[a] = tv[a]
                                out.append((ctiles[cb], ft, d, *cidx, *fstart))
        return out


@dataclass
class LevelGrid:
    """Blocks on refinement levels whose block shapes differ (refinement
    ratio 2 in resolution; a refined parent is covered by r^d children with
    r = 2 nb_parent / nb_child).  Level 0 covers the domain.  Blocks are
    (L, (px, py, pz)) with 0-based levels; the fx_amr_* entry points take
    1-based levels."""
    ndim: int
    nb: list                          # per level: interior cells (nx, ny, nz)
    ext: list                         # per level: blocks per axis covering the domain
    blocks: list                      # [(L, pos)], coarsest level first
    periodic: tuple = (False, False, False)
    domain: tuple = (1.0, 1.0, 1.0)
    ng: int = 4

    @classmethod
    def nested(cls, ndim, nb, base, refine, periodic=(False, False, False)):
        """nb[L]: block shape of level L; base: level-0 blocks per axis;
        refine[L](pos) -> bool: refine that block of level L."""
        nb = [tuple(n) + (1,) * (3 - len(n)) for n in nb]
        nb = [tuple(n[a] if a < ndim else 1 for a in range(3)) for n in nb]
        base = tuple(base) + (1,) * (3 - len(base))
        ext = [tuple(base[a] if a < ndim else 1 for a in range(3))]
        for L in range(1, len(nb)):
            r = []
            for a in range(3):
                if a >= ndim:
                    r.append(1)
                    continue
                q, m = divmod(2 * nb[L - 1][a], nb[L][a])
                if m or q < 1:
                    raise ValueError("block sizes must nest with refinement ratio 2")
                r.append(q)
            ext.append(tuple(ext[L - 1][a] * r[a] for a in range(3)))
        blocks = [(0, p) for p in itertools.product(*[range(n) for n in ext[0]])]
        blocks = [(0, (p[0], p[1], p[2])) for _, p in blocks]
        for L in range(1, len(nb)):
            r = [ext[L][a] // ext[L - 1][a] for a in range(3)]
            for (Lp, pos) in [b for b in blocks if b[0] == L - 1]:
                if not refine[L - 1](pos):
                    continue
                for o in itertools.product(*[range(r[a]) for a in range(3)]):
                    blocks.append((L, tuple(pos[a] * r[a] + o[a] for a in range(3))))
        order = sorted(range(len(blocks)),
                       key=lambda b: (blocks[b][0], blocks[b][1][2], blocks[b][1][1], blocks[b][1][0]))
        return cls(ndim, nb, ext, [blocks[b] for b in order], tuple(periodic))

    def dels(self, L):
        return tuple(self.domain[a] / (self.ext[L][a] * self.nb[L][a]) if a < self.ndim else 1.0
                     for a in range(3))

    def level_blocks(self, L):
        return [b for b, (Lb, _) in enumerate(self.blocks) if Lb == L]

    def block_shape(self, L, nunk):
        nx, ny, nz = self.nb[L]
        g = self.ng
        return (nunk, nz + 2 * g if self.ndim > 2 else 1, ny + 2 * g if self.ndim > 1 else 1,
                nx + 2 * g)

    def cell_centers(self, b):
        """Guard-inclusive cell-centre coordinates (x, y, z) of block b."""
        import numpy as np
        L, pos = self.blocks[b]
        d = self.dels(L)
        out = []
        for a in range(3):
            n = self.nb[L][a]
            g = self.ng if a < self.ndim else 0
            idx = np.arange(-g, n + g) if a < self.ndim else np.zeros(1)
            out.append((pos[a] * n + idx + 0.5) * d[a] if a < self.ndim else np.full(1, 0.5 * d[a]))
        return out

    def as_dict(self):
        """Plain description (the CPU restatement's grid format)."""
        return dict(ndim=self.ndim, ng=self.ng, nb=list(self.nb), ext=list(self.ext),
                    blocks=list(self.blocks), periodic=tuple(self.periodic),
                    **{"del": [self.dels(L) for L in range(len(self.nb))]})

    def zones(self):
        n = 0
        for L, _ in self.blocks:
            nx, ny, nz = self.nb[L]
            n += nx * ny * nz
        return n


def quadrant_grid(nb=((32, 32, 32), (16, 16, 16), (8, 8, 8)), base=(8, 8, 8)):
    """Three levels around the line x = y = 1/2 where the quadrants meet:
    level 0 is base blocks of nb[0]; level 1 refines the level-0 blocks
    nearest the line in x and y (2 x 2 columns through z), level 2 the
    level-1 blocks nearest it.  With the defaults: 512 blocks of 32^3, 2048
    of 16^3 and 8192 of 8^3."""
    ndim = 3

    def near(ext_xy):
        lo, hi = ext_xy // 2 - 1, ext_xy // 2
        return lambda p: lo <= p[0] <= hi and lo <= p[1] <= hi
    # level-L extents follow from the shapes; computed once to build the predicates
    probe = LevelGrid.nested(ndim, list(nb[:1]), base, [])
    ext1 = [probe.ext[0][a] * (2 * nb[0][a] // nb[1][a]) for a in range(3)]
    return LevelGrid.nested(ndim, list(nb), base, [near(base[0]), near(ext1[0])])
