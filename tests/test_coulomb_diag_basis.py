"""The Coulomb_diag_basis form of the mean field (ReadInput.h:1112-1114,
RungeKutta.h:41-55 and 97-124, main_MPI.cpp:291-331, InitParameters.h:264-370):
the Fourier sums run on P in the eigenbasis, Pd = conj(U) P U^T, against the
equilibrium of filled bands; the screening corrections act in the eigenbasis;
the mean-field matrix returns to the orbital basis as U^+ X_d U.

The oracle's restatement is pinned by composing it from parts pinned elsewhere:
the sums' reconstruction X(k) and Hartree shift (the coulomb_x hook, pinned by
brute-force convolution in test_oracle_invariants.py), which are linear in the
state, and the plain right-hand side with H replaced by the renormalised H; the
eigenbasis transforms and corrections are done here in numpy."""
import numpy as np

import helpers as Hh
from helpers import N, O


def _systems(s):
    base = Hh.oracle_system(s, coulomb=True)
    info = s.info
    cf = Hh.read_coulomb_files(Hh.COULOMB, Hh.COULOMB_LABEL, info["nb"], info["ncut"])
    sysd = O.System(base["H"], base["D"], base["U"], base["T"], base["nbr"], np.repeat(base["w"], base["sizes"]),
                    base["bvec"], info["nch"], info["ncc"], info["wx2"])
    sysd.set_coulomb(info["ncut"], base["kpt"], cf["A"], cf["B"], cf["C"], cf["D"], cf["screen"], cf["vh"],
                     diag_basis=True)
    return base, sysd, cf


def test_diag_basis_rhs_is_the_composition_of_its_parts(tmp_path):
    s = N.Setup(Hh.coulomb_input(str(tmp_path), nk=12))
    info = s.info
    nb, ncc = info["nb"], info["ncc"]
    base, sysd, cf = _systems(s)
    H, D, U = base["H"], base["D"], base["U"]
    P0 = O.rho0(U, ncc)
    rng = np.random.default_rng(4)
    X = rng.normal(size=P0.shape) + 1j * rng.normal(size=P0.shape)
    P = P0 + 0.02 * 0.5 * (X + np.conj(np.swapaxes(X, 1, 2)))
    Uc = U[:, 2, :]
    P = P + 0.01 * np.einsum("ki,kj->kij", Uc, np.conj(Uc))
    for f6, on in (([1e-3, 2e-3, -3e-3, 1e-4, 0, 0], True), ([0, 0, 0, 0, 0, 0], False)):
        ref = sysd.rhs_coulomb(P0, P, np.array(f6), on, 0.0)
        # the parts: P in the eigenbasis, X and the Hartree shift of Pd minus the filled bands
        Pd = np.einsum("kai,kij,kbj->kab", np.conj(U), P, U)
        idx = np.arange(nb)
        Pd[:, idx, idx] = np.real(Pd[:, idx, idx])
        Peq = np.zeros_like(Pd)
        Peq[:, idx[:ncc], idx[:ncc]] = 1.0
        Xd, EH = base["sys"].coulomb_x(Pd - Peq)
        scr = cf["screen"]
        for i in range(nb):
            Xd[:, i, i] -= scr[0] * (Pd[:, i, i] - (1.0 if i < ncc else 0.0))
            for j in range(i + 1, nb):
                Xd[:, i, j] -= np.conj(scr[1]) * Pd[:, j, i]
                Xd[:, j, i] = np.conj(Xd[:, i, j])
        XW = np.einsum("kai,kab,kbj->kij", np.conj(U), Xd, U)
        XW[:, idx, idx] = np.real(XW[:, idx, idx])
        Hr = H + XW
        Hr[:, idx, idx] += EH[None, :]
        plain = O.System(Hr, D, U, base["T"], base["nbr"], np.repeat(base["w"], base["sizes"]), base["bvec"], info["nch"],
                         ncc, info["wx2"])
        ref2 = plain.rhs(P, np.array(f6), on)
        assert np.max(np.abs(ref - ref2)) <= 1e-13 * np.max(np.abs(ref2))
        # and it is not the orbital-basis form
        wan = base["sys"].rhs_coulomb(P0, P, np.array(f6), on, 1.0)
        assert np.max(np.abs(ref - wan)) > 1e-6 * np.max(np.abs(wan))
