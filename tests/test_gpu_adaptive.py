"""GPU parity of the adaptive-dt drop-in (Source_Main/dyn_dt_evolution.cpp): the host
loop over edus_step_dyn writes the reference's Output files; they must match the
oracle's adaptive loop (pinned in test_adaptive_dt.py) on the same input."""
import os

import numpy as np
import pytest

import helpers as Hh
from helpers import N
from test_adaptive_dt import dyn_input

pytestmark = pytest.mark.gpu


def _compare_files(out, rows, nb, what, rel=1e-10):
    """rel: bar relative to the series max.  The step-size measure divides the change
    of P_ii over one step (~1e-8 of P_ii under a weak field) by P_ii-sized terms, so
    its rounding is ~1e-8 relative and the dt sequences of two correct
    implementations agree only to that level; runs whose dt control engages often
    are compared at the resulting shift in time.  On the mean-field input below a
    1e-12 relative change of epsStepAbs alone moves J1 by 1.8e-10 of its max in the
    oracle itself."""
    o = Hh.split_rows(rows, nb)
    L = np.loadtxt(os.path.join(out, "Losses.txt"), ndmin=2)
    J1 = np.loadtxt(os.path.join(out, "J1.txt"), ndmin=2)
    TA = np.loadtxt(os.path.join(out, "TransientAbs.txt"), ndmin=2)
    EF = np.loadtxt(os.path.join(out, "EF.txt"), ndmin=2)
    assert L.shape[0] == rows.shape[0], (what, L.shape, rows.shape)
    # times: the dt sequence follows max_dP, whose rounding differs by ~1e-11 relative
    np.testing.assert_allclose(L[:, 0], o["t"], rtol=1e-9, atol=1e-12)
    for c in range(nb):
        ref = o["pop"][:, c]
        assert np.max(np.abs(L[:, 1 + c] - ref)) <= max(rel * np.max(np.abs(ref)), 1e-14), (what, "pop", c)
    j1s = max(np.max(np.abs(o["J1"])), 1e-300)
    assert np.max(np.abs(J1[:, 1:] - o["J1"])) <= rel * j1s, (what, "J1")
    tas = max(np.max(np.abs(o["TAbs"])), 1e-300)
    assert np.max(np.abs(TA[:, 4:] - o["TAbs"])) <= 5e-8 * tas, (what, "TAbs")
    efs = max(np.max(np.abs(o["EF"])), 1e-300)
    assert np.max(np.abs(EF[:, 1:] - o["EF"])) <= 5e-8 * efs, (what, "EF")


@pytest.mark.parametrize("stepper", ["", "Taylor\nTaylorOrder 6\n"])
def test_adaptive_tb_run_matches_oracle(tmp_path, stepper):
    s = N.Setup(dyn_input(str(tmp_path), stepper=stepper, eps=2e-3, nk=20, dt=0.2, res=2))
    osys = Hh.oracle_system(s)
    n = 120
    rows, _, dts = Hh.oracle_run_dyn(s, osys, n)
    assert dts.min() < 0.5 * s.info["dt"]
    out = str(tmp_path / "run")
    steps = s.run(out, max_steps=n)
    assert steps == len(dts)
    _compare_files(os.path.join(out, "Output"), rows, 3, "tb " + stepper.split("\n")[0])


def test_adaptive_taylor_with_mean_field_matches_oracle(tmp_path):
    """The reference's hBN Coulomb example mode: Taylor-6, adaptive dt, mean field."""
    p = Hh.coulomb_input(str(tmp_path), nk=24, stepper="Taylor\nTaylorOrder 6\ndynamical_dt_evolution\nepsStepAbs 0.02\n",
                         dt=0.1, nsteps=400, pump_int=5e13)
    s = N.Setup(p)
    assert s.info["dyn_dt"] == 1
    osys = Hh.oracle_system(s, coulomb=True)
    n = 150
    rows, _, dts = Hh.oracle_run_dyn(s, osys, n)
    out = str(tmp_path / "run")
    steps = s.run(out, max_steps=n)
    assert steps == len(dts)
    _compare_files(os.path.join(out, "Output"), rows, 3, "coulomb taylor", rel=1e-8)


def test_adaptive_wannier_run_matches_oracle(tmp_path):
    """Adaptive dt on the wave kernel (8 Wannier bands, RK4)."""
    from edus_amd import workloads
    p = workloads.wannier_synthetic(str(tmp_path), nw=8, rmax=2, nk=10, nsteps=400, dt=0.2, stepper="RK4", nvalence=4,
                                    seed=11, e_v_ang=2.0, it_resolution=2)
    text = open(p).read().replace("num_pump_pulses 1\n", "num_pump_pulses 1\ndynamical_dt_evolution\nepsStepAbs 2e-3\n")
    open(p, "w").write(text)
    s = N.Setup(p)
    assert s.info["dyn_dt"] == 1
    osys = Hh.oracle_system(s, use_oracle_tables=False, bruteforce_neighbours=False)
    n = 80
    rows, _, dts = Hh.oracle_run_dyn(s, osys, n)
    out = str(tmp_path / "run")
    steps = s.run(out, max_steps=n)
    assert steps == len(dts)
    _compare_files(os.path.join(out, "Output"), rows, 8, "wannier8 adaptive")
