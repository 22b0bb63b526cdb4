"""k-resolved snapshots of the drop-in (if_resolution.cpp:19-35, 232-273;
Observables_MPI.h:3-130, 440-477): the trigger times, the file names and the
values, against numpy on the oracle's states at the same steps."""
import os

import numpy as np
import pytest

import helpers as Hh
from helpers import N, O
from test_gpu_engine import random_hermitian_state, tb_input

pytestmark = pytest.mark.gpu


def eigen_element(P, U, ic, jc):
    """(conj(U) P_sym U^T)[ic][jc], P_sym from the stored upper triangle."""
    Ps = np.triu(P) + np.conj(np.swapaxes(np.triu(P, 1), 1, 2))
    v = np.einsum("ki,kij,kj->k", np.conj(U[:, ic, :]), Ps, U[:, jc, :])
    if ic == jc:
        v = v.real + 0j
    return v


def test_eigen_element_matches_numpy(tmp_path):
    s = N.Setup(tb_input(str(tmp_path), nk=12))
    eng = N.Engine(s.problem())
    P = random_hermitian_state(s.array("rho0", True).reshape(-1, 3, 3), 0.05, 3)
    eng.set_state(P)
    U = s.array("U", True).reshape(-1, 3, 3)
    for ic, jc in ((1, 2), (1, 1), (2, 2), (2, 0)):
        got = eng.eigen_element(ic, jc)
        ref = eigen_element(P, U, ic, jc)
        assert np.max(np.abs(got - ref)) <= 1e-14


def test_print_population_files(tmp_path):
    p = tb_input(str(tmp_path), nk=12, stepper="", nsteps=200, dt=0.1, res=5)
    text = open(p).read().replace("it_resolution 5\n", "it_resolution 5\nPrintPopulation 2.0 12.0 3.0\n")
    open(p, "w").write(text)
    s = N.Setup(p)
    info = s.info
    assert info["print_flags"] == 1
    n = 160
    out = str(tmp_path / "run")
    assert s.run(out, max_steps=n) == n
    od = os.path.join(out, "Output")
    # trigger rule: at an observation, t > start + icont*step and t < end -> icont += 1
    icont, shots = 0, []
    for it in range(0, n, info["it_resolution"]):
        t = s.time(it)
        if t > 2.0 + icont * 3.0 and t < 12.0:
            icont += 1
            shots.append((icont, it, t))
    assert len(shots) >= 2
    tek = np.loadtxt(os.path.join(od, "t_Ek.txt"), ndmin=1)
    np.testing.assert_allclose(tek, [t * O.TIME_AU_FS for _, _, t in shots], rtol=1e-5)
    osys = Hh.oracle_system(s)
    U = osys["U"]
    ncc = info["ncc"]
    for icont, it, _ in shots:
        _, P = Hh.oracle_run(s, osys, it)
        for label, ic, jc in (("P_val_cond_in_Eigenstate", ncc - 1, ncc), ("P_val_in_Eigenstate", ncc - 1, ncc - 1),
                              ("P_cond_in_Eigenstate", ncc, ncc)):
            f = os.path.join(od, f"{label}_{ic}{jc}_{icont}.txt")
            got = np.loadtxt(f)
            ref = eigen_element(P, U, ic, jc)
            assert got.shape == (info["nk"], 2)
            scale = np.max(np.abs(ref))
            # printed with 6 significant digits
            assert np.all(np.abs(got[:, 0] - ref.real) <= 6e-6 * np.abs(ref.real) + 1e-12 * scale), (label, icont)
            assert np.all(np.abs(got[:, 1] - ref.imag) <= 6e-6 * np.abs(ref.imag) + 1e-12 * scale), (label, icont)
    assert not os.path.exists(os.path.join(od, f"P_cond_in_Eigenstate_22_{len(shots) + 1}.txt"))


def test_band_dispersion_files_with_mean_field(tmp_path):
    p = Hh.coulomb_input(str(tmp_path), nk=12, dt=0.1, nsteps=400, res=10)
    text = open(p).read().replace("Rytova_Keldysh\n", "Rytova_Keldysh\nPrint_new_band_dispersion\n")
    open(p, "w").write(text)
    s = N.Setup(p)
    assert s.info["print_flags"] == 2
    out = str(tmp_path / "run")
    s.run(out, max_steps=200)
    od = os.path.join(out, "Output")
    tek = np.loadtxt(os.path.join(od, "t_Ek.txt"), ndmin=1)
    assert len(tek) >= 1
    E1 = np.loadtxt(os.path.join(od, "Edia_1.txt"))
    assert E1.shape == (s.info["nk"], 3)
    assert np.all(np.diff(E1, axis=1) >= 0)
    # the mean field is a small shift of the bare bands early in the pulse
    H = s.array("H", True).reshape(-1, 3, 3)
    bare = np.linalg.eigvalsh(H)
    assert np.max(np.abs(E1 - bare)) < 0.05 * np.max(np.abs(bare))
    assert not os.path.exists(os.path.join(od, "P_cond_in_Eigenstate_22_1.txt"))
