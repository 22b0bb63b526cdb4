"""Batched stepping (edus_batch_*): several independent RK4 problems of
different band counts, grids, models and pulses advanced by one launch per stage
and kernel class.  Each member must step bit-identically to the same problem in
its own context (the single-context path is the one pinned to the oracle by
test_gpu_engine.py)."""
import numpy as np
import pytest

from helpers import N
from test_gpu_engine import _wannier_setup, tb_input

pytestmark = pytest.mark.gpu


def _setups(tmp_path):
    def d(name):
        (tmp_path / name).mkdir(exist_ok=True)
        return str(tmp_path / name)

    return [
        N.Setup(tb_input(d("tb_a"), nk=12, t2=0.01, pump_int=3e12)),
        N.Setup(tb_input(d("tb_b"), nk=16, t1=0.004, t2=0.02, pump_int=8e12)),
        _wannier_setup(d("w5"), 5, nk=10),
        _wannier_setup(d("w8"), 8, nk=12, t2=10.0),
        _wannier_setup(d("w16"), 16, nk=8),
    ]


@pytest.mark.parametrize("precision", [None, "mixed"])
def test_batch_members_match_their_own_contexts(tmp_path, precision):
    setups = _setups(tmp_path)
    if precision == "mixed":  # mixed precision is for nb >= 5
        setups = setups[2:]
    prec = N.PRECISION_MIXED if precision == "mixed" else None
    nsteps, chunks = 12, 3
    batch = N.Batch([s.problem() for s in setups], precision=prec)
    singles = [N.Engine(s.problem(), precision=prec) for s in setups]
    it = 0
    for _ in range(chunks):
        fg = [s.fields(it, nsteps) for s in setups]
        batch.step([f for f, _ in fg], [g for _, g in fg])
        for e, (f, g) in zip(singles, fg):
            e.step(f, g)
        it += nsteps
    ms, launches = batch.last_step_time()
    # kernel classes: band count (small kernels) or tile (wave kernels), split by dephasing
    classes = len({(s.info["nb"] if s.info["nb"] <= 4 else (8 if s.info["nb"] <= 8 else 16 if s.info["nb"] <= 16 else 32),
                    s.info["n_diss"] > 0) for s in setups})
    assert launches == 4 * nsteps * classes and ms > 0
    for m, e, s in zip(batch.members, singles, setups):
        np.testing.assert_array_equal(m.get_state(), e.get_state())
        om, oe = m.observables(), e.observables()
        for k in om:
            np.testing.assert_array_equal(om[k], oe[k])
        # the state moved
        assert np.max(np.abs(e.get_state() - s.array("rho0", True).reshape(e.get_state().shape))) > 0
    batch.close()
