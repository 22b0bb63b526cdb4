"""Shared helpers of the test-suite: golden fixtures and oracle problems."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
AB5 = os.path.join(GOLDEN, "ab5")

from edus_amd import native as N  # noqa: E402
from oracle import oracle_py as O  # noqa: E402


def load_cols(path):
    return np.loadtxt(path)


def ab5_fixture():
    """The reference's committed AB5 run (Haldane_CoreBN_Nedge, 70x70, AB5)."""
    return {name: load_cols(os.path.join(AB5, name + ".txt")) for name in ("J1", "J2", "Losses", "TransientAbs", "EF")}


def oracle_system(setup, use_oracle_tables=True, bruteforce_neighbours=True):
    """Oracle tables for a setup: k-grid, shells, neighbours and model from the oracle's
    own restatement (tight binding), the setup's parameters and pulses."""
    info = setup.info
    N3 = [info["nk0"], info["nk1"], info["nk2"]]
    nb = info["nb"]
    if use_oracle_tables:
        kpt = O.kgrid(*N3)
    else:
        kpt = setup.array("kpt").reshape(-1, 3)
    b = reciprocal_rows(setup)
    bvec, sizes, w, res = O.shells(kpt, N3, b)
    wshell = np.repeat(w, sizes)
    if bruteforce_neighbours:
        nbr = O.neighbours(kpt, b, bvec)
    else:
        nbr = setup.nbr()
    name, par = setup.model()
    if par is not None and use_oracle_tables:
        H, U, D = O.tb_tables(name, par, kpt, nb)
    else:
        H = setup.array("H", True).reshape(-1, nb, nb)
        U = setup.array("U", True).reshape(-1, nb, nb)
        D = setup.array("D", True).reshape(-1, nb, nb, 3)
    T = setup.array("T").reshape(nb, nb)
    sys_ = O.System(H, D, U, T, nbr, wshell, bvec, info["nch"], info["ncc"], info["wx2"])
    return {"sys": sys_, "kpt": kpt, "bvec": bvec, "w": w, "sizes": sizes, "res": res, "nbr": nbr, "H": H, "U": U, "D": D,
            "T": T}


def reciprocal_rows(setup):
    """b vectors (rows) from the direct lattice: b_i = 2 pi (a_j x a_k) / (a_i . (a_j x a_k))
    (ReadInput.h:388-391), computed here independently of the host."""
    a = setup.array("a").reshape(3, 3)
    b = np.zeros((3, 3))
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        c = np.cross(a[j], a[k])
        b[i] = c / (np.dot(a[i], c) / 2.0 / np.pi)
    return b


def gradH_from(H, nbr, wshell, bvec):
    """GradientEnergy of save_HUD.cpp:245-253: sum w b_x (H(k+b) - H(k))."""
    nk, nb, _ = H.shape
    G = np.zeros((nk, nb, nb, 3), dtype=np.complex128)
    for n in range(nbr.shape[1]):
        dH = H[nbr[:, n]] - H
        for x in range(3):
            G[..., x] += wshell[n] * bvec[n, x] * dH
    return G


def oracle_run(setup, osys, nsteps, it_res=None, gradH=True):
    info = setup.info
    pumps = [setup.pulse(i) for i in range(info["n_pumps"])]
    probe = setup.pulse(info["n_pumps"])
    sysd = osys["sys"]
    G = gradH_from(osys["H"], osys["nbr"], np.repeat(osys["w"], osys["sizes"]), osys["bvec"]) if gradH else None
    P0 = O.rho0(osys["U"], info["ncc"])
    dk3 = 1.0 / (info["nk0"] * info["nk1"] * info["nk2"])
    rows, P = sysd.run(P0, info["stepper"], info["order"], info["dt"], info["t_start"], nsteps,
                       it_res or info["it_resolution"], pumps, probe, gradH=G, j_norm=info["j_norm"],
                       jacobian=info["jacobian"], dk3=dk3, nktot=info["nktot"])
    return rows, P


def split_rows(rows, nb):
    """Oracle rows -> dict of the printed columns."""
    return {"t": rows[:, 0], "EF": rows[:, 1:4], "pop": rows[:, 4:4 + nb], "J1": rows[:, 4 + nb:10 + nb],
            "J2": rows[:, 10 + nb:13 + nb], "EFprobe": rows[:, 13 + nb:16 + nb], "TAbs": rows[:, 16 + nb:22 + nb]}


def rel_err(a, b, floor=0.0):
    """max|a-b| / max(max|b|, floor) over the whole series."""
    a, b = np.asarray(a), np.asarray(b)
    den = max(np.max(np.abs(b)), floor)
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)
