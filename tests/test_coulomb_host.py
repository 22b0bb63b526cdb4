"""Coulomb mean-field input path on the CPU: the host reads the reference's
coefficient files (Fourier_Coulomb.cpp:505-622) exactly as numpy parses them,
and rejects the Coulomb variants this library does not carry.  (Without
Read_Coulomb_from_files the host computes the coefficients on the GPU; that path
is tested in test_gpu_coulomb_gen.py.)"""
import os

import numpy as np
import pytest

import helpers as Hh
from helpers import N


def test_coefficient_files_read_like_numpy(tmp_path):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=12))
    info = s.info
    assert info["ncut"] == 20 and info["nb"] == 3
    ref = Hh.read_coulomb_files(Hh.COULOMB, Hh.COULOMB_LABEL, 3, 20)
    for q in "ABCD":
        got = s.array("coul_" + q, True).reshape(400, 9)
        np.testing.assert_array_equal(got, ref[q])
    np.testing.assert_array_equal(s.array("coul_screen", True).reshape(9), ref["screen"])
    np.testing.assert_array_equal(s.array("coul_vh", True).reshape(3, 3), ref["vh"])
    # the files' own structure: V_Hartree symmetric, A diagonal entries real
    assert np.allclose(ref["vh"], ref["vh"].T)
    assert np.max(np.abs(ref["A"][:, [0, 4, 8]].imag)) == 0.0


def test_coulomb_diag_basis_accepted_with_the_same_tables(tmp_path):
    """Coulomb_diag_basis changes only where the mean-field sums run, not the
    setup's tables."""
    (tmp_path / "w").mkdir()
    (tmp_path / "d").mkdir()
    pw = Hh.coulomb_input(str(tmp_path / "w"), nk=12)
    pd = Hh.coulomb_input(str(tmp_path / "d"), nk=12)
    text = open(pd).read().replace("Read_Coulomb_from_files\n", "Read_Coulomb_from_files\nCoulomb_diag_basis\n")
    open(pd, "w").write(text)
    sw, sd = N.Setup(pw), N.Setup(pd)
    for name in ("H", "U", "D", "rho0"):
        np.testing.assert_array_equal(sd.array(name), sw.array(name))
    assert sd.info["ncut"] == sw.info["ncut"] == 20


def test_missing_coefficient_file(tmp_path):
    p = Hh.coulomb_input(str(tmp_path), nk=12)
    text = open(p).read().replace(Hh.COULOMB_LABEL, "no_such_label_")
    open(p, "w").write(text)
    with pytest.raises(RuntimeError, match="cannot open Coulomb coefficient file"):
        N.Setup(p)


def test_oracle_coulomb_rhs_reduces_to_bare_rhs_at_equilibrium(tmp_path):
    """At P = P0 the mean field minus its equilibrium vanishes (X - X0 = 0,
    EH - EH0 = 0, P - P_W0 = 0), so the Coulomb RHS equals the bare RHS."""
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=12))
    osys = Hh.oracle_system(s, coulomb=True)
    P0 = s.array("rho0", True).reshape(-1, 3, 3)
    f6 = np.array([1e-3, 2e-3, -3e-3, 1e-4, 0, 0])
    a = osys["sys"].rhs_coulomb(P0, P0, f6, True, 0.3)
    b = osys["sys"].rhs(P0, f6, True)
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))
    # away from equilibrium the term is present
    rng = np.random.default_rng(3)
    X = rng.normal(size=P0.shape) + 1j * rng.normal(size=P0.shape)
    P = P0 + 0.02 * (X + np.conj(np.swapaxes(X, 1, 2)))
    a = osys["sys"].rhs_coulomb(P0, P, f6, True, 0.3)
    b = osys["sys"].rhs(P, f6, True)
    assert np.max(np.abs(a - b)) > 1e-6 * np.max(np.abs(b))
