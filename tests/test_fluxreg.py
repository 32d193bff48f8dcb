"""Flux-register correction (Grid_putFluxData / Grid_conserveFluxes,
Cartesian, ratio 2) against FLASH's TestFluxCorrection known-answer test
(fixture tests/golden/flux_correction_kat.json), on the CPU oracle; the GPU
entry points are checked against the same KAT and the oracle in
test_gpu_fluxreg.py."""
import json
import os

import numpy as np
import pytest

import oracle as O
from flash5_amd.amr import BlockMesh

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "flux_correction_kat.json")))


def kat_mesh():
    m = BlockMesh(ndim=KAT["ndim"], nb=KAT["nb"], base=tuple(KAT["base"]),
                  periodic=tuple(KAT["periodic"]))
    for lev, blocks in KAT["refine"].items():
        for b in blocks:
            m.refine(int(lev), tuple(b) + (0,))
    return m


def kat_fluxes(m, p):
    """X fluxes = level, Y fluxes = -level (every flux variable)."""
    out = {}
    for lev in range(1, m.finest() + 1):
        out[lev] = []
        for _ in m.blocks(lev):
            out[lev].append([np.full(O.flux_shape(p, 0), float(lev)),
                             np.full(O.flux_shape(p, 1), -float(lev)), None])
    return out


def kat_expected(m, p, lev):
    """Expected flux arrays after the cascade, per tile of level lev."""
    exp = []
    xs = set()
    for i, (j0, j1) in zip(KAT["level2_overwritten_x_faces"]["i"],
                           KAT["level2_overwritten_x_faces"]["j_ranges"]):
        xs |= {(i, j) for j in range(j0, j1 + 1)}
    ys = set()
    for j, (i0, i1) in zip(KAT["level2_overwritten_y_faces"]["j"],
                           KAT["level2_overwritten_y_faces"]["i_ranges"]):
        ys |= {(i, j) for i in range(i0, i1 + 1)}
    nb = KAT["nb"]
    for b in m.blocks(lev):
        fx = np.full(O.flux_shape(p, 0), float(lev))
        fy = np.full(O.flux_shape(p, 1), -float(lev))
        if lev == 2:
            for a in range(nb + 1):
                for c in range(nb):
                    if (b[0] * nb + 1 + a, b[1] * nb + 1 + c) in xs:
                        fx[:, 0, c, a] = lev + 1
            for a in range(nb):
                for c in range(nb + 1):
                    if (b[0] * nb + 1 + a, b[1] * nb + 1 + c) in ys:
                        fy[:, 0, c, a] = -(lev + 1)
        exp.append((fx, fy))
    return exp


def run_cascade(m, p, fl, put, conserve):
    """Driver_evolveFlash.F90's loop: finest to coarsest, conserve then put."""
    reg = {}
    for lev in range(m.finest(), 0, -1):
        if lev < m.finest():
            conserve(lev, fl[lev], m.register_faces(lev), reg[lev + 1])
        if lev > 1:
            reg[lev] = put(lev, fl[lev], m.register_faces(lev - 1))


def test_kat_mesh_matches_the_reference_layout():
    m = kat_mesh()
    assert m.finest() == 3
    assert len(m.blocks(1)) == 4 and len(m.blocks(2)) == 16 and len(m.blocks(3)) == 16
    assert m.register_faces(1) == []               # level 2 covers the domain; x is outflow
    faces = m.register_faces(2)
    # every entry is one copy of a coarse face; each copy appears once
    keys = {(f[0], f[2], f[3], f[4]) for f in faces}
    assert len(keys) == len(faces)


def test_oracle_flux_correction_kat():
    m = kat_mesh()
    p = O.params(2, 8, 1.4)
    fl = kat_fluxes(m, p)
    run_cascade(m, p, fl,
                lambda lev, f, faces, : O.fluxreg_put(p, f, faces),
                lambda lev, f, faces, reg: O.fluxreg_conserve(p, f, faces, reg))
    for lev in (1, 2, 3):
        for t, (ex, ey) in enumerate(kat_expected(m, p, lev)):
            np.testing.assert_array_equal(fl[lev][t][0], ex, err_msg=f"level {lev} tile {t} X")
            np.testing.assert_array_equal(fl[lev][t][1], ey, err_msg=f"level {lev} tile {t} Y")


@pytest.mark.parametrize("ndim", [2, 3])
def test_register_averages_fine_flux_densities(ndim):
    """The register holds the mean of the covering fine flux densities, so a
    coarse face of area 2^(d-1) fine faces carries their total flux."""
    nb = 4
    m = BlockMesh(ndim=ndim, nb=nb, base=(2,) * ndim, periodic=(True,) * 3)
    m.refine(1, (0, 0, 0))
    p = O.params(ndim, nb, 1.4)
    rng = np.random.default_rng(3)
    fine = [[rng.normal(size=O.flux_shape(p, d)) if d < ndim else None for d in range(3)]
            for _ in m.blocks(2)]
    faces = m.register_faces(1)
    reg = O.fluxreg_put(p, fine, faces)
    nfine = 2 ** (ndim - 1)
    for e, f in enumerate(faces):
        ct, ft, d = f[:3]
        idx = list(f[6:9])
        tr = [a for a in range(3) if a != d and a < ndim]
        vals = []
        for offs in np.ndindex(*([2] * len(tr))):
            k = list(idx)
            for a, o in zip(tr, offs):
                k[a] += o
            vals.append(fine[ft][d][:, k[2], k[1], k[0]])
        np.testing.assert_allclose(reg[:, e], np.mean(vals, axis=0), rtol=1e-15, atol=1e-15)
    # each refined-parent side facing an unrefined neighbour contributes nb^(d-1)
    # faces, listed once for the parent's copy and once for the neighbour's
    assert len(faces) == 2 * (2 * ndim) * nb ** (ndim - 1)
    assert nfine == 2 ** (ndim - 1)
