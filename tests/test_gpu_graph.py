"""The captured-step path (edus_set_field_table + edus_step_rk4_graph): one RK4
step (RungeKutta.cpp:11-137) captured into a HIP graph that reads its fields
from a device-resident table through a device step counter, replayed once per
step.  It must advance the state bit for bit like edus_step_rk4 with the same
fields, over several calls, with and without the mean field."""
import numpy as np
import pytest

import helpers as Hh
from helpers import N
from test_gpu_engine import _wannier_setup, tb_input

pytestmark = pytest.mark.gpu


def _compare(s, nsteps, chunks):
    a = N.Engine(s.problem())
    b = N.Engine(s.problem())
    f, g = s.fields(0, nsteps)
    b.set_field_table(f, g)
    done = 0
    for n in chunks:
        a.step(f[done:done + n], g[done:done + n])
        b.step_graph(n)
        done += n
        np.testing.assert_array_equal(b.get_state(), a.get_state())
        oa, ob = a.observables(), b.observables()
        for key in oa:
            np.testing.assert_array_equal(ob[key], oa[key])
    ms, launches = b.last_step_time()
    assert launches == 4 * chunks[-1] and ms > 0
    return a, b


@pytest.mark.parametrize("case", ["tb3", "wannier16", "wannier32"])
def test_graph_steps_equal_stream_steps(tmp_path, case):
    if case == "tb3":
        s = N.Setup(tb_input(str(tmp_path), nk=16, t1=0.004, t2=0.01, tch=0.02, nsteps=120, pump_int=8e12))
    else:
        s = _wannier_setup(str(tmp_path), 16 if case == "wannier16" else 32, nk=12, t2=10.0)
    n = min(60, s.info["itfi"] + 1)
    a, b = _compare(s, n, [7, 1, n - 8])
    # the table is used up: one more step is an error, and a new table restarts it
    with pytest.raises(N.EdusError, match="beyond the end of the field table"):
        b.step_graph(1)


def test_graph_steps_with_mean_field(tmp_path):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=24, dt=0.1, nsteps=200, pump_int=5e13))
    _compare(s, 40, [10, 30])


def test_graph_path_rejects_other_steppers(tmp_path):
    s = N.Setup(tb_input(str(tmp_path), nk=8, stepper="Adams_Bashforth 4\n"))
    e = N.Engine(s.problem())
    f, g = s.fields(0, 4)
    with pytest.raises(N.EdusError, match="runs RK4 contexts"):
        e.set_field_table(np.zeros((4, 3, 6)), g)
