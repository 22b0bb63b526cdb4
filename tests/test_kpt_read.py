"""nkPT read: an explicit list of crystal k points in the input (ReadInput.h:421-445)
instead of the uniform grid.  Shells come from the listed points nearest Gamma
and neighbours from the reference's search rule (CalculateWeigths.h:145-226), so
the uniform grid written out as a list reproduces the grid's setup exactly, a
shuffled list reproduces it up to the permutation, and a list that misses a k + b
is an error."""
import numpy as np
import pytest

from helpers import N, O
from test_gpu_engine import tb_input


def _listed(path_in, order):
    """The input at path_in with its nkPT grid replaced by an explicit list of the
    same grid's points in the given order (oracle k-grid, BZ_parameters order)."""
    text = open(path_in).read()
    a = text.index("nkPT{")
    b = text.index("}", a)
    dims = [int(x) for x in text[a + 5:b].split()]
    kpt = O.kgrid(*dims)[order]
    body = "\n".join(" ".join(repr(float(x)) for x in k) for k in kpt)
    return text[:a] + f"nkPT read {{\n{dims[0]} {dims[1]} {dims[2]}\n{body}\n}}" + text[b + 1:], kpt


def _setup(tmp_path, name, text):
    d = tmp_path / name
    d.mkdir()
    p = d / "input.txt"
    p.write_text(text)
    return N.Setup(str(p))


@pytest.mark.parametrize("model", ["gen2d_hexagonal a= 2.50 gap= 5.62 t1= 2.92 t2= 0.1 phi0= 0.3 core= -20 chv_dipole= 0.05 "
                                   "chc_dipole= 0.07", "GeS"])
def test_listed_grid_reproduces_the_grid(tmp_path, model):
    g = N.Setup(tb_input(str(tmp_path), nk=12, t2=0.01, model=model))
    nk = g.info["nk"]
    rng = np.random.default_rng(5)
    for label, order in (("inorder", np.arange(nk)), ("shuffled", rng.permutation(nk))):
        text, kpt = _listed(str(tmp_path / "input.txt"), order)
        s = _setup(tmp_path, label, text)
        assert s.info["nk"] == nk
        # the engine's stencil width: every b vector of every shell, as on the grid
        assert s.info["n_nbr"] == g.info["n_nbr"] == s.problem().n_nbr == s.nbr().shape[1]
        np.testing.assert_array_equal(s.array("kpt").reshape(-1, 3), kpt)
        for name in ("H", "U", "D", "rho0"):
            np.testing.assert_array_equal(s.array(name, True), g.array(name, True).reshape(nk, -1)[order].reshape(-1))
        # the b vectors of a shell are taken in list order (CalculateWeigths.h:40-60 scans
        # the k points), so the columns of a shuffled list are the grid's, permuted
        wl, wg = s.array("nbr_wb").reshape(-1, 3), g.array("nbr_wb").reshape(-1, 3)
        # (the shell sums then run in another order: weights equal to rounding)
        col = [int(np.flatnonzero(np.all(np.abs(wg - row) <= 1e-12 * np.abs(wg).max(), axis=1))[0]) for row in wl]
        assert sorted(col) == list(range(len(wg)))
        if label == "inorder":
            assert col == list(range(len(wg)))
        inv = np.argsort(order)
        np.testing.assert_array_equal(s.nbr(), inv[g.nbr()[order][:, col]])
        # gradient of H from the permuted neighbours: the same values, permuted
        np.testing.assert_allclose(s.array("gradH", True), g.array("gradH", True).reshape(nk, -1)[order].reshape(-1),
                                   rtol=0, atol=1e-15)


def test_list_missing_a_neighbour_is_an_error(tmp_path):
    tb_input(str(tmp_path), nk=8)
    nk = 64
    keep = np.delete(np.arange(nk), 9)  # one point of the grid gone (not Gamma)
    text, _ = _listed(str(tmp_path / "input.txt"), keep)
    text = text.replace("nkPT read {\n1 8 8\n", "nkPT read {\n1 9 7\n")  # 63 points listed
    with pytest.raises(N.EdusError, match="is not in the k list"):
        _setup(tmp_path, "missing", text)


def test_oracle_search_on_a_shuffled_list(tmp_path):
    """The oracle's own shell and neighbour search (brute force, pinned on grids by
    the closed form) on the shuffled list: the same shells as on the grid (b vectors
    taken in list order) and the grid's neighbour table, permuted like the points."""
    import helpers as Hh
    g = N.Setup(tb_input(str(tmp_path), nk=12, t2=0.01))
    nk = g.info["nk"]
    order = np.random.default_rng(11).permutation(nk)
    text, kpt = _listed(str(tmp_path / "input.txt"), order)
    s = _setup(tmp_path, "shuffled", text)
    og = Hh.oracle_system(g, use_oracle_tables=True, bruteforce_neighbours=True)
    ol = Hh.oracle_system(s, use_oracle_tables=False, bruteforce_neighbours=True)
    np.testing.assert_array_equal(ol["sizes"], og["sizes"])
    np.testing.assert_allclose(ol["w"], og["w"], rtol=1e-12)
    wl, wg = ol["bvec"], og["bvec"]
    col = [int(np.flatnonzero(np.all(np.abs(wg - row) <= 1e-12 * np.abs(wg).max(), axis=1))[0]) for row in wl]
    assert sorted(col) == list(range(len(wg)))
    inv = np.argsort(order)
    np.testing.assert_array_equal(ol["nbr"], inv[og["nbr"][order][:, col]])
    # and the host's hashed search gives the oracle's table
    np.testing.assert_array_equal(s.nbr(), ol["nbr"])
