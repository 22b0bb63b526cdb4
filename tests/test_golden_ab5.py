"""The reference's only committed run (example/Adams-Bashforth.../AB5_*, 70x70
Haldane_CoreBN_Nedge, AB5, 16591 steps) pins the oracle, the host's input
parser, k-grid, shells, tight-binding tables and laser schedule.

Tolerances:
  J1, Losses (valence/conduction): 1e-10 relative to the series' max |value|.
  J2: the printed sum cancels ~1e7-fold (|terms| ~ 1, |J2| ~ 1e-7), so both the
      reference and any restatement carry ~1e-16 absolute rounding; the bar is
      5e-16 absolute (<= 4e-9 relative).
  EF, TransientAbs: printed with 8 significant digits -> 5e-8 relative.
  Core-band Losses: the hole population is ~1e-13, i.e. rounding of 1-P; 1e-13 absolute.
"""
import os
import re

import numpy as np
import pytest

import helpers as Hh
from helpers import N, O


@pytest.fixture(scope="module")
def setup():
    return N.Setup(os.path.join(Hh.AB5, "input.txt"))


@pytest.fixture(scope="module")
def oracle_rows(setup):
    osys = Hh.oracle_system(setup)
    rows, _ = Hh.oracle_run(setup, osys, setup.info["itfi"] + 1)
    return rows


def test_parse_matches_banner(setup):
    """Values the reference printed in output.log for this input."""
    inf = setup.info
    assert (inf["nk0"], inf["nk1"], inf["nk2"], inf["nb"]) == (1, 70, 70, 3)
    assert inf["stepper"] == N.STEPPER_AB and inf["order"] == 5
    assert inf["itfi"] + 1 == 16591
    assert abs(inf["det_a"] - 19.329) < 5e-4          # "unit cell area = 19.329"
    assert abs(inf["jacobian"] - 2.0272) < 5e-5       # "Jacobian determinant: 2.0272"
    assert inf["n_diss"] == 5                          # "n_diss_terms 5"
    pump, probe = setup.pulse(0), setup.pulse(1)
    assert abs(pump["wl"] - 0.25713) < 5e-6 and abs(pump["period"] - 24.43618) < 5e-6
    assert abs(pump["E0"] / 3.77426e-04 - 1) < 2e-6 and abs(pump["A0"] / 1.46786e-03 - 1) < 4e-6
    assert abs(pump["tf"] / 1.71053e+02 - 1) < 4e-6
    assert abs(probe["t0"] / 8.48367e+01 - 1) < 2e-6 and abs(probe["E0"] / 1.68790e-05 - 1) < 3e-6
    assert pump["pol1"] == [0.0, 1.0, 0.0] and pump["pol2"] == [0.0, 0.0, 1.0]


def test_kgrid_matches_k_grid_txt(setup):
    kpt = setup.array("kpt").reshape(-1, 3)
    ref = open(os.path.join(Hh.AB5, "k_grid.txt")).read().split("\n")
    mine = ["%g %g %g" % tuple(k) for k in kpt]
    assert mine == [r for r in ref if r]
    np.testing.assert_array_equal(O.kgrid(1, 70, 70), kpt)


def test_shells_match_output_log(setup):
    rows = []
    for line in open(os.path.join(Hh.AB5, "shells_from_output_log.txt")):
        m = re.match(r"\*\s+\|\s+(\d+)\s+\|\s+(\S+)\s+(\S+)\s+(\S+)\s+\|\s+(\S+)\s+\|\s+(\S+)", line)
        if m:
            rows.append([float(m.group(i)) for i in range(1, 7)])
    rows = np.array(rows)
    bvec = setup.array("bvec").reshape(-1, 3)
    assert len(bvec) == len(rows) == 12
    np.testing.assert_allclose(bvec, rows[:, 1:4], atol=6e-6)
    kpt = setup.array("kpt").reshape(-1, 3)
    ob, sizes, w, res = O.shells(kpt, [1, 70, 70], Hh.reciprocal_rows(setup))
    np.testing.assert_array_equal(sizes, [6, 6])
    np.testing.assert_allclose(ob, bvec, rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, setup.array("weights"), rtol=1e-11)
    assert res < 1e-7 and setup.info["shell_residual"] < 1e-7
    # B1 condition sum_b w b_a b_b = delta_ab in the sample plane (y, z)
    wb = setup.array("nbr_wb").reshape(-1, 3)
    B1 = wb[:, 1:].T @ bvec[:, 1:]
    np.testing.assert_allclose(B1, np.eye(2), atol=1e-12)


def _matrix_block(text, title):
    lines = text.split("\n")
    i = [j for j, l in enumerate(lines) if l.startswith(title)][0]
    out = []
    for l in lines[i + 1:i + 4]:
        out.append([complex(float(a), float(b)) for a, b in re.findall(r"\(([^,]+),([^)]+)\)", l)])
    return np.array(out)


def test_tight_binding_matches_test_matrices(setup):
    text = open(os.path.join(Hh.AB5, "Test_matrices.txt")).read()
    k = np.array([[0.0, 0.3, 0.1]])
    H, U, D = setup.eval_tb(k)
    name, par = setup.model()
    oH, oU, oD = O.tb_tables(name, par, k, 3)
    for mine in ((H, oH), (U, oU), (D, oD)):
        np.testing.assert_array_equal(mine[0], mine[1])
    np.testing.assert_allclose(H[0], _matrix_block(text, "Energy"), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(U[0], _matrix_block(text, "Unitary"), rtol=1e-5, atol=1e-6)
    for x, t in enumerate(("dx", "dy", "dz")):
        np.testing.assert_allclose(D[0, :, :, x], _matrix_block(text, t), rtol=1e-5, atol=1e-6)
    # U H U^+ diagonal and U U^+ = 1 (test_matrix_k.cpp checks)
    E2 = U[0] @ H[0] @ U[0].conj().T
    assert np.max(np.abs(E2 - np.diag(np.diag(E2)))) < 1e-15
    np.testing.assert_allclose(U[0] @ U[0].conj().T, np.eye(3), atol=1e-15)


def test_setup_tables_equal_oracle_tables(setup):
    osys = Hh.oracle_system(setup)
    np.testing.assert_array_equal(osys["nbr"], setup.nbr())
    for name in ("H", "U", "D"):
        np.testing.assert_array_equal(setup.array(name, True).reshape(osys[name].shape), osys[name])
    np.testing.assert_array_equal(setup.array("rho0", True).reshape(osys["U"].shape), O.rho0(osys["U"], 2))


def test_field_schedule_matches_ef_txt(setup):
    """EF.txt prints the pump field left by the previous step (AB: E(t_{n-1}))."""
    fx = np.loadtxt(os.path.join(Hh.AB5, "EF.txt"))
    pump = setup.pulse(0)
    res = setup.info["it_resolution"]
    for row in range(1, len(fx), 97):
        it = row * res
        f, _ = setup.fields(it - 1, 1)
        np.testing.assert_allclose(f[0, 0, :3], fx[row, 1:4], rtol=5e-8, atol=1e-14)
        np.testing.assert_allclose(O.field(pump, setup.time(it - 1)), f[0, 0, :3], rtol=0, atol=0)


def test_oracle_reproduces_ab5_outputs(oracle_rows):
    fx = Hh.ab5_fixture()
    r = Hh.split_rows(oracle_rows, 3)
    n = len(oracle_rows)
    assert n == len(fx["J1"]) == 3319
    np.testing.assert_allclose(r["t"], fx["Losses"][:, 0], rtol=0, atol=1e-13)
    for c in range(6):
        assert Hh.rel_err(r["J1"][:, c], fx["J1"][:, 1 + c]) <= 1e-10
    assert np.max(np.abs(r["J2"] - fx["J2"][:, 1:4])) <= 5e-16
    assert np.max(np.abs(r["pop"][:, 0] - fx["Losses"][:, 1])) <= 1e-13
    for c in (1, 2):
        assert Hh.rel_err(r["pop"][:, c], fx["Losses"][:, 1 + c]) <= 1e-10
    assert Hh.rel_err(r["EF"], fx["EF"][:, 1:4]) <= 5e-8
    assert Hh.rel_err(r["EFprobe"], fx["TransientAbs"][:, 1:4]) <= 5e-8
    for c in range(6):
        assert Hh.rel_err(r["TAbs"][:, c], fx["TransientAbs"][:, 4 + c]) <= 5e-8
