"""GPU parity of the Coulomb mean-field term (RungeKutta.h:69-138, Coulomb.h:273-541)
against the oracle, with the reference's own coefficient files (hBN example,
Ncut 20, 3 bands) on a 24 x 24 grid.  The oracle's Coulomb term is pinned by the
brute-force convolution test of the CPU suite."""
import numpy as np
import pytest

import helpers as Hh
from helpers import N
from test_gpu_engine import check_series, random_hermitian_state, run_engine_rows

pytestmark = pytest.mark.gpu


def test_rhs_with_mean_field_matches_oracle(tmp_path):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=24))
    osys = Hh.oracle_system(s, coulomb=True)
    eng = N.Engine(s.problem())
    P0 = s.array("rho0", True).reshape(-1, 3, 3)
    P = random_hermitian_state(P0, 0.02, 5)
    # plus 1% population in the conduction band, so that the peak-correction gate is open
    Uc = s.array("U", True).reshape(-1, 3, 3)[:, 2, :]
    P = P + 0.01 * np.einsum("ki,kj->kij", Uc, np.conj(Uc))
    eng.set_state(P)
    # the engine's peak-correction gate is the largest printed population of P
    ob = eng.observables()
    pop = ob["pop"].copy()
    pop[:s.info["ncc"]] = 1.0 - pop[:s.info["ncc"]]
    n_cond = max(0.0, float(np.max(pop)))
    assert n_cond > 1e-20
    for f6, on in (([1e-3, 2e-3, -3e-3, 1e-4, 0, 0], True), ([0, 0, 0, 0, 0, 0], False)):
        pv = eng.rhs(np.array(f6), on)
        ref = osys["sys"].rhs_coulomb(P0, P, np.array(f6), on, n_cond)
        assert np.max(np.abs(pv - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("stepper", ["", "Taylor\nTaylorOrder 6\n", "Adams_Bashforth 4\n"])
def test_trajectory_with_mean_field_matches_oracle(tmp_path, stepper):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=24, stepper=stepper, dt=0.1, nsteps=400, pump_int=5e13))
    osys = Hh.oracle_system(s, coulomb=True)
    n = 300
    ora, Pora = Hh.oracle_run(s, osys, n)
    gpu, eng = run_engine_rows(s, n)
    check_series(gpu, ora, 3, "coulomb " + stepper.split("\n")[0])
    assert np.max(np.abs(eng.get_state() - Pora)) <= 1e-11 * np.max(np.abs(Pora))


def _diag_input(d, **kw):
    p = Hh.coulomb_input(d, **kw)
    text = open(p).read().replace("Read_Coulomb_from_files\n", "Read_Coulomb_from_files\nCoulomb_diag_basis\n")
    open(p, "w").write(text)
    return p


def test_rhs_with_mean_field_in_the_diagonal_basis_matches_oracle(tmp_path):
    """Coulomb_diag_basis (RungeKutta.h:41-55, 97-124): sums over P in the
    eigenbasis against the filled-band equilibrium, corrections there, then
    U^+ X_d U; no n_cond gate."""
    s = N.Setup(_diag_input(str(tmp_path), nk=24))
    osys = Hh.oracle_system(s, coulomb=True, diag_basis=True)
    eng = N.Engine(s.problem())
    P0 = s.array("rho0", True).reshape(-1, 3, 3)
    P = random_hermitian_state(P0, 0.02, 6)
    Uc = s.array("U", True).reshape(-1, 3, 3)[:, 2, :]
    P = P + 0.01 * np.einsum("ki,kj->kij", Uc, np.conj(Uc))
    eng.set_state(P)
    for f6, on in (([1e-3, 2e-3, -3e-3, 1e-4, 0, 0], True), ([0, 0, 0, 0, 0, 0], False)):
        pv = eng.rhs(np.array(f6), on)
        ref = osys["sys"].rhs_coulomb(P0, P, np.array(f6), on, 0.0)
        assert np.max(np.abs(pv - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("stepper", ["", "Adams_Bashforth 4\n"])
def test_trajectory_with_mean_field_in_the_diagonal_basis_matches_oracle(tmp_path, stepper):
    s = N.Setup(_diag_input(str(tmp_path), nk=24, stepper=stepper, dt=0.1, nsteps=400, pump_int=5e13))
    osys = Hh.oracle_system(s, coulomb=True, diag_basis=True)
    n = 240
    ora, Pora = Hh.oracle_run(s, osys, n)
    gpu, eng = run_engine_rows(s, n)
    check_series(gpu, ora, 3, "coulomb diag " + stepper.strip())
    assert np.max(np.abs(eng.get_state() - Pora)) <= 1e-11 * np.max(np.abs(Pora))
