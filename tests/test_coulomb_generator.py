"""Coulomb coefficient generator (Fourier_Coulomb.cpp:95-500) in the oracle: the
lattice sum is pinned by the reference's own V_Hartree file (the example's G shell
of 502657 vectors), the Simpson/Fourier integration by trigonometric modes whose
coefficients are known in closed form."""
import os

import numpy as np
import pytest

import helpers as Hh
from helpers import N, O


@pytest.fixture(scope="module")
def example_setup(tmp_path_factory):
    d = tmp_path_factory.mktemp("cgen")
    return N.Setup(Hh.coulomb_input(str(d), nk=12))


def test_g_shell_of_the_example(example_setup):
    p = Hh.coulomb_gen_params(example_setup)
    assert len(p["g"]) == 502657  # the count in the example's file names
    np.testing.assert_allclose(p["disp"][2], [0.0, 2.72758, 0.0], atol=1e-5)


def test_oracle_hartree_matches_reference_file(example_setup):
    p = Hh.coulomb_gen_params(example_setup, n_sum=3, n_peak=3, ncut=1)
    out = O.coulomb_coefficients(p)
    ref = np.loadtxt(os.path.join(Hh.COULOMB, "V_Hartree_" + Hh.COULOMB_LABEL + ".txt"))
    assert np.max(np.abs(out["vh"].imag)) == 0.0
    assert np.max(np.abs(out["vh"].real - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_oracle_integration_recovers_trigonometric_modes():
    """W00 = cos(2 pi a qx) cos(2 pi b qy) + sin(2 pi c qx) sin(2 pi d qy) and
    W01 = cos(...)cos(...) + i sin(2 pi c qx) cos(2 pi d qy) have unit coefficients
    A[a,b], D[c,d] and A[a,b], C[c,d] = i (4 x the m/n-dependent 1/4, 1/2, 1 factor
    times the integral of the squared modes)."""
    ncut, n = 4, 201
    p = {"nb": 2, "ncut": ncut, "n_sum": n, "n_peak": 3, "rytova": 1, "g": np.zeros((1, 2), np.int32), "b1": [0, 1, 0],
         "b2": [0, 0, 1], "eps_area": 1.0, "qtf": 0.0, "r0": 0.0, "q_peak": 0.0, "disp": np.zeros((2, 3))}
    q = -0.5 + np.arange(n) / (n - 1)
    QX, QY = np.meshgrid(q, q, indexing="ij")
    for (a, b, c, d) in ((0, 0, 1, 1), (2, 0, 3, 1), (1, 3, 2, 0), (3, 2, 1, 3)):
        cc = np.cos(2 * np.pi * a * QX) * np.cos(2 * np.pi * b * QY)
        W = np.zeros((n * n, 4), dtype=np.complex128)
        W[:, 0] = (cc + np.sin(2 * np.pi * c * QX) * np.sin(2 * np.pi * d * QY)).ravel()
        W[:, 1] = (cc + 1j * np.sin(2 * np.pi * c * QX) * np.cos(2 * np.pi * d * QY)).ravel()
        A, B, C, D = O.cgen_integrate(p, W)
        expA = np.zeros((ncut * ncut, 4), complex)
        expD = np.zeros_like(expA)
        expC = np.zeros_like(expA)
        expA[a * ncut + b, 0] = 1
        expA[a * ncut + b, 1] = 1
        if c > 0 and d > 0:
            expD[c * ncut + d, 0] = 1
        if c > 0:
            expC[c * ncut + d, 1] = 1j
        for got, exp in ((A, expA), (B, np.zeros_like(expA)), (C, expC), (D, expD)):
            np.testing.assert_allclose(got, exp, atol=2e-6)
