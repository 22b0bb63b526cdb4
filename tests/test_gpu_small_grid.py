"""GPU parity of the small-grid workload family: the 3-band hBN model
(gen2d_hexagonal, T2 = 10 fs, pump along Gamma-K) through the thread-per-k
kernel (k_stage_small<3>), against the oracle, with the three steppers the
reference offers (RungeKutta.cpp:11-137, Linear_Multistep_ODE_solvers.cpp:1-93,
Taylor_DE_solver_deriv_Polynom.cpp:61-177), and the workload itself at its own
size (60 x 60).  The uniform grids run with neighbours formed from the grid
offsets; the shuffled k list of test_gpu_band_classes.py runs the table.

Bars as in test_gpu_engine.py: observable series 1e-10 of the series max,
final states 1e-11.
"""
import numpy as np
import pytest

import helpers as Hh
from helpers import N
from test_gpu_engine import check_series, run_engine_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stepper", ["RK4", "AB", "Taylor"])
def test_hbn_trajectory_matches_oracle(tmp_path, stepper):
    """24 x 24 grid, 600 steps around the pump peak, RK4, Adams-Bashforth 4 and Taylor 4."""
    from edus_amd import workloads
    s = N.Setup(workloads.hbn_tb(str(tmp_path), nk=24, nsteps=600, dt=0.05, stepper=stepper, order=4, it_resolution=10))
    osys = Hh.oracle_system(s)
    n = s.info["itfi"] + 1
    print(f"hbn 24x24 {stepper} steps={n}", flush=True)
    ora, Pora = Hh.oracle_run(s, osys, n)
    gpu, eng = run_engine_rows(s, n)
    check_series(gpu, ora, 3, f"hbn {stepper}")
    assert np.max(np.abs(eng.get_state() - Pora)) <= 1e-11 * np.max(np.abs(Pora))


def test_small_grid_workload_matches_oracle(tmp_path):
    """The workload itself: hBN 60 x 60, RK4, dt 0.05, T2 = 10 fs, 400 steps
    around the pump peak."""
    from edus_amd import workloads
    s = N.Setup(workloads.hbn_tb(str(tmp_path), nk=60, nsteps=400, dt=0.05, stepper="RK4"))
    assert s.info["nk"] == 3600
    osys = Hh.oracle_system(s)
    n = s.info["itfi"] + 1
    print(f"hbn 60x60 RK4 steps={n}", flush=True)
    ora, Pora = Hh.oracle_run(s, osys, n)
    gpu, eng = run_engine_rows(s, n)
    check_series(gpu, ora, 3, "hbn60")
    assert np.max(np.abs(eng.get_state() - Pora)) <= 1e-11 * np.max(np.abs(Pora))
