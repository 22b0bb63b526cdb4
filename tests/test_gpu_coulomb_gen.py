"""GPU Coulomb coefficient generator (edus_coulomb_coefficients) against the oracle
on a reduced problem, and the drop-in's full-size generation (4999^2 Simpson
points x 502657 G vectors, 39999^2 correction grid) against the coefficient files
the reference ships for its hBN example."""
import os
import time

import numpy as np
import pytest

import helpers as Hh
from helpers import N, O

pytestmark = pytest.mark.gpu


def _close(got, ref, rel):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return float(np.max(np.abs(got - ref)) / scale) <= rel


def test_generator_matches_oracle_on_reduced_grids(tmp_path):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=12))
    p = Hh.coulomb_gen_params(s, g_distance=6.0, n_sum=31, n_peak=801, ncut=5, g_scan=60)
    p["q_peak"] = 0.1
    ora = O.coulomb_coefficients(p)
    gpu = N.coulomb_coefficients(p)
    for k in ("A", "B", "C", "D", "vh"):
        assert _close(gpu[k], ora[k], 1e-11), k
    assert _close(gpu["screen"], ora["screen"], 1e-10), "screen"
    assert np.max(np.abs(ora["screen"])) > 0  # the correction grid has points


def test_dropin_generates_the_reference_coefficient_files(tmp_path):
    """Coulomb{} without Read_Coulomb_from_files and with the example's G_distance
    400: the drop-in computes the coefficients at setup and writes them under
    Output/Coulomb/ with the reference's label and formats."""
    p = Hh.coulomb_input(str(tmp_path), nk=6)
    text = open(p).read().replace("Read_Coulomb_from_files\n", "G_distance 400\n").replace(
        "labelInput " + Hh.COULOMB_LABEL, "labelInput _")
    os.unlink(os.path.join(str(tmp_path), "Output", "Coulomb"))
    open(p, "w").write(text)
    t0 = time.time()
    s = N.Setup(p)
    print(f"generation + setup: {time.time() - t0:.1f} s", flush=True)
    d = os.path.join(str(tmp_path), "Output", "Coulomb")
    label = "G_num502657_Ncut20_"
    for name in ("A_coeff", "B_coeff", "C_coeff", "D_coeff"):
        mine = np.loadtxt(os.path.join(d, name + label + ".txt"))
        ref = np.loadtxt(os.path.join(Hh.COULOMB, name + label + ".txt"))
        assert mine.shape == ref.shape == (400, 18)
        err = np.max(np.abs(mine - ref)) / np.max(np.abs(ref))
        print(name, "max rel err", err, flush=True)
        assert err <= 1e-9, name
    for name in ("V_Hartree_", "Screen_const_"):
        mine = np.loadtxt(os.path.join(d, name + label + ".txt"))
        ref = np.loadtxt(os.path.join(Hh.COULOMB, name + label + ".txt"))
        err = np.max(np.abs(mine - ref)) / np.max(np.abs(ref))
        print(name, "max rel err", err, flush=True)
        assert mine.shape == ref.shape and err <= 1e-9, name
    # the file layout: 20-digit scientific pairs for A..D, plain for V_Hartree
    first = open(os.path.join(d, "A_coeff" + label + ".txt")).readline()
    ref_first = open(os.path.join(Hh.COULOMB, "A_coeff" + label + ".txt")).readline()
    assert len(first.split()) == len(ref_first.split()) and "e" in first.split()[0]
