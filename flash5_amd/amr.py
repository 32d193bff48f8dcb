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
