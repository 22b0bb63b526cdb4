"""Shared helpers of the test-suite: golden fixtures and oracle problems."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
AB5 = os.path.join(GOLDEN, "ab5")
COULOMB = os.path.join(GOLDEN, "coulomb")
COULOMB_LABEL = "G_num502657_Ncut20_"

from edus_amd import native as N  # noqa: E402
from oracle import oracle_py as O  # noqa: E402


def load_cols(path):
    return np.loadtxt(path)


def ab5_fixture():
    """The reference's committed AB5 run (Haldane_CoreBN_Nedge, 70x70, AB5)."""
    return {name: load_cols(os.path.join(AB5, name + ".txt")) for name in ("J1", "J2", "Losses", "TransientAbs", "EF")}


def read_coulomb_files(directory, label, nb, ncut):
    """The coefficient files of Fourier_Coulomb.cpp:505-622, parsed with numpy
    (independently of the host's reader)."""
    def cx(name, rows):
        a = np.loadtxt(os.path.join(directory, name + label + ".txt"), ndmin=2)[:rows]
        return a[:, 0::2] + 1j * a[:, 1::2]
    out = {q: cx(q + "_coeff", ncut * ncut) for q in "ABCD"}
    out["screen"] = cx("Screen_const_", nb * nb)[:, 0]
    out["vh"] = np.loadtxt(os.path.join(directory, "V_Hartree_" + label + ".txt"), ndmin=2)[:nb, :nb].astype(np.complex128)
    return out


def coulomb_input(d, nk=24, stepper="", dt=0.1, nsteps=400, res=2, pump_int=5e12, probe_int=1e9, cycles=1):
    """Input of the reference's hBN Coulomb example (Haldane_CoreBN_Nedge, Ncut 20,
    coefficient files read from Output/Coulomb) on a small grid with constant dt."""
    t0 = -0.21 / O.TIME_AU_FS
    t_fin = t0 + (nsteps - 0.5) * dt
    text = ("NEWINPUT\n\ntdse{\ntightbinding Haldane_CoreBN_Nedge gap=  7.25 t1=  -2.3 t2= 0 phi0= 0.0\n"
            f"dt {dt!r} au\nt_fin {t_fin!r} au\n{stepper}}}\n\n"
            "Coulomb{\nqTF 0.01\nNcut 20\nepsilon_static 1.0\nRead_Coulomb_from_files\n"
            f"labelInput {COULOMB_LABEL}\n\nRytova_Keldysh\n}}\n\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n\ndecoherence{{\ncorehole 0.00396892 au\n}}\n\n"
            f"laserpump sin2  {{\ncycles {cycles}\nwavelength 206.66666666666666 nm\n"
            f"intensity {pump_int!r} wcm2\npolarization 0 1 1\n}}\n\n"
            "laserprobe gaussian {\nsigma 0.08 fs\ndelay  0.0 fs\n"
            f"intensity  {probe_int!r}\nfrequency 15.06\npolarization 0 1 0\n}}\n\n"
            f"observables{{\nit_resolution {res}\nTAbs\nCurrent\n}}\n")
    os.makedirs(os.path.join(d, "Output"), exist_ok=True)
    link = os.path.join(d, "Output", "Coulomb")
    if not os.path.exists(link):
        os.symlink(COULOMB, link)
    p = os.path.join(d, "input.txt")
    with open(p, "w") as f:
        f.write(text)
    return p


def coulomb_gen_params(setup, g_distance=400.0, n_sum=4999, n_peak=39999, ncut=20, qtf=0.01, eps=1.0, r0=18.89726,
                       rytova=1, g_scan=4999):
    """Generator parameters of Fourier_Coulomb.cpp:1-120 for a setup ("011"): the G
    shell |G| < g_distance sqrt(det_b) scanned first offset outer, orbital positions
    from the real dipole diagonal at k = (0, 0.3, 0.1), eps * in-plane cell area."""
    info = setup.info
    b = setup.array("b").reshape(3, 3)
    gd = g_distance * np.sqrt(info["det_b"])
    gl = []
    iy = np.arange(-g_scan, g_scan + 1)
    for ix in range(-g_scan, g_scan + 1):
        cart = ix * b[1][None, :] + iy[:, None] * b[2][None, :]
        m = np.sqrt(np.sum(cart * cart, axis=1)) < gd
        if m.any():
            gl.append(np.stack([np.full(m.sum(), ix), iy[m]], axis=1))
    g = np.concatenate(gl).astype(np.int32)
    _, _, D = setup.eval_tb(np.array([[0.0, 0.3, 0.1]]))
    disp = np.real(np.einsum("iix->ix", D[0]))
    return {"nb": info["nb"], "ncut": ncut, "n_sum": n_sum, "n_peak": n_peak, "rytova": rytova, "g": g, "b1": b[1].copy(),
            "b2": b[2].copy(), "eps_area": eps * info["det_a"], "qtf": qtf, "r0": r0, "q_peak": max(qtf, 0.02), "disp": disp}


def crystal_tables(nkf, norb=3, seed=5):
    """Smooth periodic H(k) and Berry connections (x, y) on the crystal interface's
    grid k = -0.5 + i/(nkf-1): [norb][norb][nkf][nkf] (H, lower triangle used) and
    [norb][norb][2][nkf][nkf]."""
    rng = np.random.default_rng(seed)
    k = -0.5 + np.arange(nkf) / (nkf - 1)
    K1, K2 = np.meshgrid(k, k, indexing="ij")
    e1, e2 = np.exp(2j * np.pi * K1), np.exp(2j * np.pi * K2)
    H = np.zeros((norb, norb, nkf, nkf), dtype=np.complex128)
    B = np.zeros((norb, norb, 2, nkf, nkf), dtype=np.complex128)
    for i in range(norb):
        H[i, i] = -0.3 + 0.25 * i + 0.02 * (np.cos(2 * np.pi * K1) + np.cos(2 * np.pi * K2))
        for x in range(2):
            B[i, i, x] = 0.1 * (i + 1) * np.sin(2 * np.pi * (K1 if x == 0 else K2))
        for j in range(i):
            c = rng.normal(size=3) + 1j * rng.normal(size=3)
            H[i, j] = 0.03 * (c[0] + c[1] * e1 + c[2] * e2)
            for x in range(2):
                d = rng.normal(size=2) + 1j * rng.normal(size=2)
                B[i, j, x] = 0.2 * (d[0] * e1 + d[1] * np.conj(e2))
    return H, B


def write_crystal_input(d, nkf=25, nk=24, norb=3, valence=1, stepper="", nsteps=400, dt=0.1, a=4.7243):
    """The crystal interface's four files (Crystal_interface.h:58-200 layout) for
    crystal_tables(), and an input.txt that selects them."""
    os.makedirs(d, exist_ok=True)
    H, B = crystal_tables(nkf, norb)
    c30, s30 = np.cos(np.pi / 6), np.sin(np.pi / 6)
    a1, a2 = [float(a * c30), float(a * s30)], [float(a * c30), float(-a * s30)]
    det = a1[0] * a2[1] - a1[1] * a2[0]
    b1 = [2 * np.pi * a2[1] / det, -2 * np.pi * a2[0] / det]
    b2 = [-2 * np.pi * a1[1] / det, 2 * np.pi * a1[0] / det]
    with open(os.path.join(d, "kmesh_data.dat"), "w") as f:
        f.write(f"a0\n{a1[0]!r} {a1[1]!r}\na1\n{a2[0]!r} {a2[1]!r}\nb0\n{b1[0]!r} {b1[1]!r}\nb1\n{b2[0]!r} {b2[1]!r}\n")
        f.write("atomic positions\n0.0 0.0\nnumber of orbitals in the basis\n")
        f.write(f"{norb}\nnumber of k points in each direction\n{nkf}\nk points (crystal)\n")
        for i in range(nkf * nkf):
            f.write(f"{i} {-0.5 + (i // nkf) / (nkf - 1)!r} {-0.5 + (i % nkf) / (nkf - 1)!r}\n")
    with open(os.path.join(d, "hk_orthonormal.dat"), "w") as f:
        for i1 in range(nkf):
            for i2 in range(nkf):
                for part in (np.real, np.imag):
                    for ic in range(norb):
                        f.write(" ".join(repr(float(part(H[ic, jc, i1, i2]))) for jc in range(ic + 1)) + "\n")
                    f.write("\n")
                f.write("\n")
    with open(os.path.join(d, "bc_betabetap_real.dat"), "w") as fr, open(os.path.join(d, "bc_betabetap_imag.dat"), "w") as fi:
        for i1 in range(nkf):
            for i2 in range(nkf):
                for x in range(2):
                    for ic in range(norb):
                        fr.write(" ".join(repr(float(B[ic, jc, x, i1, i2].real)) for jc in range(ic + 1)) + "\n")
                        fi.write(" ".join(repr(float(B[ic, jc, x, i1, i2].imag)) for jc in range(ic + 1)) + "\n")
                    fr.write("\n")
                    fi.write("\n")
                fr.write("\n")
                fi.write("\n")
    t0 = -0.21 / O.TIME_AU_FS
    t_fin = t0 + (nsteps - 0.5) * dt
    text = ("NEWINPUT\ntdse{\nCrystal\n"
            f"dt {dt} au\nt_fin {t_fin!r} au\n{stepper}num_pump_pulses 1\n}}\n"
            f"bands{{\nvalence {valence}\nconduction {norb - valence}\n}}\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n"
            "laserpump sin2 {\ncycles 3\nwavelength 800 nm\nintensity 5e12 wcm2\npolarization 0 1 1\n}\n"
            "laserprobe gaussian {\nsigma 0.3 fs\nintensity 0\nfrequency 15.06\npolarization 0 1 0\n}\n"
            "observables{\nit_resolution 4\nCurrent\n}\n")
    p = os.path.join(d, "input.txt")
    with open(p, "w") as f:
        f.write(text)
    return p, H, B, (a1, a2, b1, b2)


def oracle_system(setup, use_oracle_tables=True, bruteforce_neighbours=True, coulomb=False, coul_dir=COULOMB,
                  coul_label=COULOMB_LABEL, diag_basis=False):
    """Oracle tables for a setup: k-grid, shells, neighbours and model from the oracle's
    own restatement (tight binding), the setup's parameters and pulses.  With
    coulomb, the mean-field coefficients are parsed from the files in coul_dir
    (the reference's example files by default)."""
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
        H, U, D = O.tb_tables(name, par, kpt, nb, kcart=b.T)
    else:
        H = setup.array("H", True).reshape(-1, nb, nb)
        U = setup.array("U", True).reshape(-1, nb, nb)
        D = setup.array("D", True).reshape(-1, nb, nb, 3)
    T = setup.array("T").reshape(nb, nb)
    sys_ = O.System(H, D, U, T, nbr, wshell, bvec, info["nch"], info["ncc"], info["wx2"])
    if coulomb:
        ncut = info["ncut"]
        cf = read_coulomb_files(coul_dir, coul_label, nb, ncut)
        sys_.set_coulomb(ncut, kpt, cf["A"], cf["B"], cf["C"], cf["D"], cf["screen"], cf["vh"], diag_basis=diag_basis)
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


def oracle_run_dyn(setup, osys, max_steps, it_res=None, gradH=True):
    """The oracle's adaptive-dt loop with the setup's parameters."""
    info = setup.info
    pumps = [setup.pulse(i) for i in range(info["n_pumps"])]
    probe = setup.pulse(info["n_pumps"])
    G = gradH_from(osys["H"], osys["nbr"], np.repeat(osys["w"], osys["sizes"]), osys["bvec"]) if gradH else None
    P0 = O.rho0(osys["U"], info["ncc"])
    dk3 = 1.0 / (info["nk0"] * info["nk1"] * info["nk2"])
    return osys["sys"].run_dyn(P0, info["stepper"], info["order"], info["dt"], info["t_start"], info["t_fin"],
                               info["eps_step"], max_steps, it_res or info["it_resolution"], pumps, probe, gradH=G,
                               j_norm=info["j_norm"], jacobian=info["jacobian"], dk3=dk3, nktot=info["nktot"])


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
