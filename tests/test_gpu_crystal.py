"""GPU parity of a run on the crystal interface (Hermitian dipole tables from
interpolated Berry connections, Crystal_interface.h) against the oracle on the
setup's tables, which test_crystal_interface.py checks against the oracle's own
interpolation."""
import numpy as np
import pytest

import helpers as Hh
from helpers import N
from test_gpu_engine import check_series, run_engine_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("stepper", ["", "Taylor\nTaylorOrder 4\n"])
def test_crystal_trajectory_matches_oracle(tmp_path, stepper):
    # 1700 steps of 0.1 au reach the peak of the 3-cycle pump, so the current is
    # well above the rounding of its sum over k (J1 up to 3e-4)
    p, _, _, _ = Hh.write_crystal_input(str(tmp_path), nkf=25, nk=20, stepper=stepper, nsteps=1800, dt=0.1)
    s = N.Setup(p)
    osys = Hh.oracle_system(s, use_oracle_tables=False, bruteforce_neighbours=False)
    n = 1700
    ora, Pora = Hh.oracle_run(s, osys, n)
    gpu, eng = run_engine_rows(s, n)
    check_series(gpu, ora, 3, "crystal " + stepper.split("\n")[0])
    assert np.max(np.abs(eng.get_state() - Pora)) <= 1e-11 * np.max(np.abs(Pora))
