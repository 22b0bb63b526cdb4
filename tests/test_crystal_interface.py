"""Crystal interface (Crystal_interface.h, Interpolation.h): the host reads the four
files and interpolates H and the Berry connections on the run's grid; the oracle
restates FittedData's coefficient form.  No reference data ship for this mode, so
the oracle is pinned by what bicubic convolution fixes: the table values at the
nodes and exact reproduction of quadratic data away from the periodic wrap."""
import numpy as np
import pytest

import helpers as Hh
from helpers import N, O


def test_oracle_interpolation_exact_at_nodes_and_for_quadratics():
    n = 21
    rng = np.random.default_rng(2)
    tab = rng.normal(size=(1, n, n)) + 1j * rng.normal(size=(1, n, n))
    k = -0.5 + np.arange(n) / (n - 1)
    pts = np.array([[0.0, k[i], k[j]] for i in range(n - 1) for j in range(n - 1)])
    got = O.cy_interp(tab, pts)[:, 0]
    ref = np.array([tab[0, i, j] for i in range(n - 1) for j in range(n - 1)])
    # a node can land at t = 1 - eps of the previous cell ((k + 0.5)/spacing rounds down)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13 * np.max(np.abs(tab)))
    # quadratic in the index coordinates (y, z): reproduced where the 4x4 stencil does not wrap
    Y, Z = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    c = rng.normal(size=6) + 1j * rng.normal(size=6)
    quad = lambda y, z: c[0] + c[1] * y + c[2] * z + c[3] * y * y + c[4] * z * z + c[5] * y * z
    tab = quad(Y, Z)[None].astype(np.complex128)
    yz = rng.uniform(1.0, n - 3.0, size=(50, 2))
    pts = np.column_stack([np.zeros(50), -0.5 + yz[:, 0] / (n - 1), -0.5 + yz[:, 1] / (n - 1)])
    got = O.cy_interp(tab, pts)[:, 0]
    np.testing.assert_allclose(got, quad(yz[:, 0], yz[:, 1]), rtol=1e-12, atol=1e-12)


def test_setup_tables_match_oracle(tmp_path):
    p, H, B, (a1, a2, b1, b2) = Hh.write_crystal_input(str(tmp_path), nkf=25, nk=18)
    s = N.Setup(p)
    info = s.info
    nb = info["nb"]
    assert nb == 3 and info["model"] == 2
    # the lattice comes from kmesh_data.dat (a0 = 15, b0 = 1/15 out of plane)
    a = s.array("a").reshape(3, 3)
    b = s.array("b").reshape(3, 3)
    np.testing.assert_array_equal(a, [[15, 0, 0], [0, a1[0], a1[1]], [0, a2[0], a2[1]]])
    np.testing.assert_array_equal(b, [[1 / 15, 0, 0], [0, b1[0], b1[1]], [0, b2[0], b2[1]]])
    kpt = s.array("kpt").reshape(-1, 3)
    tabs = [H[i, j] for i in range(nb) for j in range(i + 1)]
    tabs += [B[i, j, x] for i in range(nb) for j in range(i + 1) for x in range(2)]
    vals = O.cy_interp(np.array(tabs), kpt)
    Hs = s.array("H", True).reshape(-1, nb, nb)
    Ds = s.array("D", True).reshape(-1, nb, nb, 3)
    q = 0
    for i in range(nb):
        for j in range(i + 1):
            np.testing.assert_allclose(Hs[:, i, j], vals[:, q], rtol=0, atol=1e-15)
            np.testing.assert_allclose(Hs[:, j, i], np.conj(vals[:, q]), rtol=0, atol=1e-15)
            q += 1
    for i in range(nb):
        for j in range(i + 1):
            for x in range(2):
                v = vals[:, q].copy()
                if i == j:
                    v = v.real + 0j
                np.testing.assert_allclose(Ds[:, i, j, x + 1], v, rtol=0, atol=1e-15)
                np.testing.assert_allclose(Ds[:, j, i, x + 1], np.conj(v), rtol=0, atol=1e-15)
                q += 1
    assert np.max(np.abs(Ds[..., 0])) == 0.0
    # U diagonalises H
    U = s.array("U", True).reshape(-1, nb, nb)
    E = np.einsum("kai,kij,kbj->kab", U, Hs, np.conj(U))
    assert np.max(np.abs(E - np.einsum("kaa->ka", E)[:, :, None] * np.eye(nb))) <= 1e-13
