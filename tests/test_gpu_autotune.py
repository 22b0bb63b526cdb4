"""edus_autotune: the launch schedule picked by timing (strip width of the wave
kernels' workgroup walk, workgroup size of the thread-per-k kernel) changes no
result.  Every candidate forms the same sums for every k, so a tuned context and
an untuned one step to bit-identical states; the kept choice is one of the
candidates, and a captured step graph made before the tuning is rebuilt."""
import numpy as np
import pytest

from helpers import N
from test_gpu_engine import _wannier_setup

pytestmark = pytest.mark.gpu


def _field(s, it):
    f, _ = s.fields(it, 1)
    return f[0, 0]


def _twins_after_tuning(s, nsteps, it0):
    tuned, plain = N.Engine(s.problem()), N.Engine(s.problem())
    chosen = tuned.autotune(_field(s, it0), True, reps=3)
    f, g = s.fields(it0, nsteps)
    g[0] = 1
    for e in (tuned, plain):
        e.step(f, g)
    return chosen, tuned, plain


def test_autotune_wave_kernel_keeps_results(tmp_path):
    s = _wannier_setup(str(tmp_path), 16, nk=40)  # rows of 40 k: strip widths 0, 16 and 32
    it0 = s.info["itfi"] // 2
    chosen, tuned, plain = _twins_after_tuning(s, 12, it0)
    print(f"16 bands, 40x40: strip width {chosen}", flush=True)
    assert chosen in (0, 16, 32)
    np.testing.assert_array_equal(tuned.get_state(), plain.get_state())


def test_autotune_small_kernel_keeps_results(tmp_path):
    from edus_amd import workloads
    s = N.Setup(workloads.hbn_tb(str(tmp_path), nk=60, nsteps=400, dt=0.05, stepper="RK4"))
    it0 = s.info["itfi"] // 2
    chosen, tuned, plain = _twins_after_tuning(s, 40, it0)
    print(f"3 bands, 60x60: {chosen} threads per workgroup", flush=True)
    assert chosen in (64, 128, 256)
    np.testing.assert_array_equal(tuned.get_state(), plain.get_state())


def test_autotune_after_a_captured_step(tmp_path):
    """Tuning drops a graph captured with the old schedule; the graph steps that
    follow still equal the stream steps."""
    from edus_amd import workloads
    s = N.Setup(workloads.hbn_tb(str(tmp_path), nk=24, nsteps=400, dt=0.05, stepper="RK4"))
    it0 = s.info["itfi"] // 2
    f, g = s.fields(it0, 20)
    a, b = N.Engine(s.problem()), N.Engine(s.problem())
    a.set_field_table(f, g)
    a.step_graph(10)
    b.step(f[:10], g[:10])
    a.autotune(_field(s, it0), True, reps=2)
    a.step_graph(10)
    b.step(f[10:], g[10:])
    np.testing.assert_array_equal(a.get_state(), b.get_state())
