"""Checkpoint files in FLASH's layout and the sfocu comparison (CPU).

* Round trips through the .npz form keep every dataset bit for bit; levels of
  different block shapes are cut into sub-blocks of the common shape whose
  bounding boxes tile each parent block and whose data are its interior.
* The comparison (tools/sfocu/sfocu.c:360-990) is checked on hand-computed
  cases: one changed cell gives the known min/max/abs/mag errors and L1 and
  L2 norms against benchmark B (and A); permuted blocks still match; a
  missing leaf block, a variable in one file only, or a mag error above the
  tolerance fails; -I drops a variable; mesh_tol matches shifted blocks that
  the bit-level test does not; parent blocks are not compared; the leaf L1
  equals a direct sum over leaf cells.
"""
import numpy as np
import pytest

from flash5_amd import checkpoint as CK
from flash5_amd import sfocu as S
from flash5_amd.amr import LevelGrid
from test_oracle_amr import _grid2d, _state


def _tiny(values, names=("dens", "pres")):
    """Two 2x2 leaf blocks side by side on [0,2]x[0,1]; values[name] is
    (2, 2, 2) = (block, j, i)."""
    return CK.Checkpoint(
        ndim=2, nb=(2, 2, 1), names=list(names), lref=np.array([1, 1], np.int32),
        node=np.array([1, 1], np.int32), coord=np.array([[0.5, 0.5, 0.], [1.5, 0.5, 0.]]),
        size=np.array([[1., 1., 0.], [1., 1., 0.]]),
        bbox=np.array([[[0., 1.], [0., 1.], [0., 0.]], [[1., 2.], [0., 1.], [0., 0.]]]),
        data={n: np.asarray(values[n], dtype=np.float64).reshape(2, 1, 2, 2) for n in names})


def _base():
    return {"dens": np.arange(1.0, 9.0).reshape(2, 2, 2), "pres": np.full((2, 2, 2), 2.0)}


def test_one_changed_cell_gives_the_hand_computed_norms():
    a = _base()
    b = {k: v.copy() for k, v in a.items()}
    b["dens"][1, 0, 1] = 8.0            # A holds 6 there
    rep = S.compare_checkpoints(_tiny(a), _tiny(b))
    d = rep["vars"]["dens"]
    assert rep["blocks_compared"] == 2 and d["bad_blocks"] == 1
    assert d["max_error"] == pytest.approx(2 * 2 / 14) and d["max_error_vals"] == (6.0, 8.0)
    assert d["abs_error"] == 2.0 and d["abs_error_vals"] == (6.0, 8.0)
    assert d["min_error"] == 0.0         # the first cell compared is equal
    assert d["mag_error"] == pytest.approx(2.0 / 8.0)
    assert d["ln_norm"] == pytest.approx(2.0 / (36.0 - 6.0 + 8.0))
    assert d["sum_a"] == 36.0 and d["sum_b"] == 38.0 and (d["min_a"], d["max_b"]) == (1.0, 8.0)
    assert rep["vars"]["pres"]["bad_blocks"] == 0 and rep["vars"]["pres"]["ln_norm"] == 0.0
    assert rep["fail"]                   # mag error 0.25 > err_tol 0
    assert not S.compare_checkpoints(_tiny(a), _tiny(b), err_tol=0.25)["fail"]
    l2 = S.compare_checkpoints(_tiny(a), _tiny(b), norm_order=2)["vars"]["dens"]["ln_norm"]
    assert l2 == pytest.approx(2.0 / np.sqrt((np.arange(1.0, 9.0) ** 2).sum() - 36 + 64))
    la = S.compare_checkpoints(_tiny(a), _tiny(b), benchmark="A")["vars"]["dens"]["ln_norm"]
    assert la == pytest.approx(2.0 / 36.0)
    text = S.report_text(rep)
    assert text.endswith("FAILURE") and "Total leaf blocks compared: 2" in text


def test_first_cell_sets_the_min_error():
    a = _base()
    b = {k: v.copy() for k, v in a.items()}
    b["dens"][0, 0, 0] = 1.5
    b["dens"][1, 1, 1] = 8.5
    d = S.compare_checkpoints(_tiny(a), _tiny(b))["vars"]["dens"]
    assert d["min_error"] == pytest.approx(1.0 / 16.5)      # below the first cell's 1/2.5
    assert d["max_error"] == pytest.approx(1.0 / 2.5)


def test_permuted_missing_and_parent_blocks():
    a, b = _tiny(_base()), _tiny(_base())
    perm = [1, 0]
    for f in ("lref", "node", "coord", "size", "bbox"):
        setattr(b, f, getattr(b, f)[perm])
    b.data = {k: v[perm] for k, v in b.data.items()}
    rep = S.compare_checkpoints(a, b)
    assert rep["blocks_compared"] == 2 and not rep["fail"]
    # the second block missing from B
    c = _tiny(_base())
    for f in ("lref", "node", "coord", "size", "bbox"):
        setattr(c, f, getattr(c, f)[:1])
    c.data = {k: v[:1] for k, v in c.data.items()}
    rep = S.compare_checkpoints(a, c)
    assert rep["extra_blocks_a"] == 1 and rep["fail"]
    rep = S.compare_checkpoints(c, a)
    assert rep["extra_blocks_b"] == 1 and rep["fail"]
    # a parent block in A is not compared, even where its data differ
    p = _tiny(_base())
    p.node = np.array([1, 2], np.int32)
    q = _tiny(_base())
    q.node = np.array([1, 2], np.int32)
    q.data["dens"][1] += 5.0
    rep = S.compare_checkpoints(p, q)
    assert rep["blocks_compared"] == 1 and not rep["fail"]


def test_variables_in_one_file_ignore_and_mesh_tolerance():
    a = _tiny(_base())
    b = _tiny(_base(), names=("dens",))
    b.data = {"dens": a.data["dens"].copy()}
    rep = S.compare_checkpoints(a, b)
    assert rep["fail"] and "pres present in A but not in B" in rep["messages"]
    assert not S.compare_checkpoints(a, b, ignore=["pres"])["fail"]
    c = _tiny(_base())
    c.coord = c.coord + 1e-9
    c.bbox = c.bbox + 1e-9
    assert S.compare_checkpoints(a, c)["extra_blocks_a"] == 2
    rep = S.compare_checkpoints(a, c, mesh_tol=1e-6)
    assert rep["blocks_compared"] == 2 and not rep["fail"]
    d = _tiny(_base())
    d.lref = np.array([2, 2], np.int32)
    assert S.compare_checkpoints(a, d)["blocks_compared"] == 0
    assert S.compare_checkpoints(a, d, ignore_lref=True)["blocks_compared"] == 2


class _Storage:
    """LevelStorage stand-in over numpy blocks."""

    def __init__(self, grid, U):
        self.grid, self.U = grid, U
        self.index = []
        count = [0] * len(grid.nb)
        for L, _ in grid.blocks:
            self.index.append((L, count[L]))
            count[L] += 1

    def block(self, b):
        return self.U[b]

    def leaf_masks(self):
        import torch
        g = self.grid
        refined = [set() for _ in g.nb]
        for L, pos in g.blocks:
            if L:
                r = [g.ext[L][a] // g.ext[L - 1][a] for a in range(3)]
                refined[L - 1].add(tuple(pos[a] // r[a] for a in range(3)))
        out = [torch.ones(sum(1 for L, _ in g.blocks if L == lev), dtype=torch.bool)
               for lev in range(len(g.nb))]
        for b, (L, pos) in enumerate(g.blocks):
            if tuple(pos) in refined[L]:
                out[L][self.index[b][1]] = False
        return out


def test_level_grid_round_trip_and_reblocking(tmp_path):
    grid = _grid2d()
    U = _state(grid, 3)
    ck = CK.from_level_storage(grid, _Storage(grid, U), time=0.25, dt=1e-3, nstep=7)
    assert ck.nb == (8, 8, 1)
    n16 = sum(1 for L, _ in grid.blocks if L == 0)
    assert ck.nblocks == 4 * n16 + (len(grid.blocks) - n16)
    path = str(tmp_path / "chk_0000.npz")
    CK.write(path, ck)
    r = CK.read(path)
    assert (r.ndim, r.nb, r.names, r.time, r.dt, r.nstep) == (2, (8, 8, 1), ck.names, 0.25,
                                                             1e-3, 7)
    for f in ("lref", "node", "coord", "size", "bbox"):
        np.testing.assert_array_equal(getattr(r, f), getattr(ck, f))
    for n in ck.names:
        np.testing.assert_array_equal(r.data[n], ck.data[n])
    # the four sub-blocks of the first 16^2 block tile it and carry its interior
    d = grid.dels(0)
    np.testing.assert_array_equal(ck.bbox[:4, 0, 0], [0.0, 8 * d[0], 0.0, 8 * d[0]])
    np.testing.assert_array_equal(ck.bbox[:4, 1, 0], [0.0, 0.0, 8 * d[1], 8 * d[1]])
    np.testing.assert_array_equal(ck.data["dens"][1, 0], U[0][0, 0, 4:12, 12:20])
    # leaves: level-0 blocks under level 1 are parents
    assert set(ck.node.tolist()) == {1, 2}
    assert not S.compare_checkpoints(ck, r)["fail"]


def test_leaf_l1_equals_the_direct_sum(tmp_path):
    grid = _grid2d()
    U = _state(grid, 5)
    V = [u.copy() for u in U]
    rng = np.random.default_rng(1)
    for v in V:
        v[0] *= 1 + 1e-3 * rng.standard_normal(v[0].shape)
    st_u, st_v = _Storage(grid, U), _Storage(grid, V)
    rep = S.compare_checkpoints(CK.from_level_storage(grid, st_v), CK.from_level_storage(grid, st_u))
    masks = st_u.leaf_masks()
    num = den = 0.0
    for b, (L, _) in enumerate(grid.blocks):
        if not masks[L][st_u.index[b][1]]:
            continue
        nb = grid.nb[L]
        a = V[b][0, 0, 4:4 + nb[1], 4:4 + nb[0]]
        c = U[b][0, 0, 4:4 + nb[1], 4:4 + nb[0]]
        num += np.abs(a - c).sum()
        den += np.abs(c).sum()
    assert rep["vars"]["dens"]["ln_norm"] == pytest.approx(num / den, rel=1e-12)


def test_uniform_grid_and_command_line(tmp_path, capsys):
    rng = np.random.default_rng(2)
    U = rng.uniform(1, 2, size=(6, 10, 1, 16, 16))
    ck = CK.from_uniform(U, (3, 2), 2, (0.125, 0.125, 1.0))
    assert ck.nblocks == 6 and ck.nb == (8, 8, 1)
    np.testing.assert_array_equal(ck.bbox[4, :2], [[0.125 * 8, 0.125 * 16], [0.125 * 8, 2.0]])
    p1, p2 = str(tmp_path / "a.npz"), str(tmp_path / "b.npz")
    CK.write(p1, ck)
    CK.write(p2, ck)
    assert S.main([p1, p2]) == 0
    out = capsys.readouterr().out
    assert out.strip().endswith("SUCCESS") and "Total leaf blocks compared: 6" in out
