"""Input files (reference grammar) for the benchmark workloads.

Each function writes an ``input.txt`` (and, for Wannier models, the seeded
``<seed>_tb.dat``) into a directory and returns the input path.  Field strengths
are converted with E0[a.u.] = E[V/A] * 1e10 / 5.14221e11 and
I = E0^2 * 3.51e16 W/cm^2 (the reference's Laser::set_IntensityWcm2 inverse).
"""
import math
import os

TIME_AU_FS = 0.02418884326505
EV_AU = 1.0 / 27.211396
EFIELD_VM_AU = 5.14221e11


def intensity_for(e_v_per_ang):
    e0 = e_v_per_ang * 1e10 / EFIELD_VM_AU
    return e0 * e0 * 3.51e16


def sigma_fs_for_intensity_fwhm(fwhm_fs):
    # |E|^2 ~ exp(-t^2/sigma^2): intensity FWHM = 2 sqrt(ln 2) sigma
    return fwhm_fs / (2.0 * math.sqrt(math.log(2.0)))


def t_fin_for(nsteps, dt):
    """t_fin (a.u.) whose loop runs exactly nsteps steps from t = -0.21 fs.  The
    reference ignores t_fin <= 0 (runge_kutta_init.cpp:366-370), so short runs
    must be cut by the caller (max_steps)."""
    t0 = -0.21 / TIME_AU_FS
    t = t0 + (nsteps - 0.5) * dt
    if t <= 0:
        raise ValueError("t_fin would be <= 0; use a longer window and cut the run")
    return t


def _pump(freq_au, e_v_ang, fwhm_fs, pol, t_peak_au, phase=0.0):
    sig = sigma_fs_for_intensity_fwhm(fwhm_fs)
    # The gaussian peak sits at t0_pulse + 4 sigma (ReadInput.h:322).
    t0_pulse_fs = (t_peak_au - 4.0 * sig / TIME_AU_FS) * TIME_AU_FS
    return ("laserpump gaussian {\n"
            f"frequency {freq_au!r}\n"
            f"intensity {intensity_for(e_v_ang)!r} wcm2\n"
            f"sigma {sig!r} fs\n"
            f"polarization {pol[0]} {pol[1]} {pol[2]}\n"
            f"phase {phase!r}\n"
            f"t0_pulse {t0_pulse_fs!r} fs\n"
            "}\n")


PROBE_OFF = ("laserprobe gaussian {\nsigma 0.08 fs\nintensity 0\nfrequency 15.06\npolarization 1 0 0\n}\n")


def write(path, text):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)
    return path


def hbn_tb(directory, nk=60, nsteps=8000, dt=0.05, stepper="RK4", order=4, e_v_ang=0.3, t2_fs=10.0,
           it_resolution=10, photon_ev=0.6, fwhm_fs=15.0):
    """Config 0 family: gen2d_hexagonal hBN (a=2.50 A, t1=2.92 eV, gap 5.62 eV,
    inert core at -20 eV), linearly polarised along Gamma-K (z), gaussian pump."""
    st = {"RK4": "", "AB": f"Adams_Bashforth {order}\n", "Taylor": f"Taylor\nTaylorOrder {order}\n"}[stepper]
    dec = f"decoherence{{\noffdiagonal {1.0 / (t2_fs / TIME_AU_FS)!r} au\n}}\n" if t2_fs else ""
    t_peak = t_fin_for(nsteps, dt) / 2.0 - 0.105 / TIME_AU_FS
    text = ("NEWINPUT\n\n"
            "tdse{\n"
            "tightbinding gen2d_hexagonal a= 2.50 gap= 5.62 t1= 2.92 t2= 0 phi0= 0 core= -20 chv_dipole= 0 chc_dipole= 0\n"
            f"dt {dt!r} au\n"
            f"t_fin {t_fin_for(nsteps, dt)!r} au\n"
            f"{st}"
            "num_pump_pulses 1\n"
            "}\n\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n\n"
            f"{dec}\n"
            + _pump(photon_ev * EV_AU, e_v_ang, fwhm_fs, (0, 0, 1), t_peak)
            + "\n" + PROBE_OFF + "\n"
            f"observables{{\nit_resolution {it_resolution}\nCurrent\nTAbs\n}}\n")
    return write(os.path.join(directory, "input.txt"), text)


def wannier_synthetic(directory, nw=16, rmax=10, nk=200, nsteps=4000, dt=0.05, stepper="RK4", order=4,
                      e_v_ang=0.5, nvalence=4, seed=1001, decay_ang=2.0, t2_fs=0.0, it_resolution=10,
                      photon_ev=0.6, fwhm_fs=15.0, pol=(0, 0, 1), phase=0.0, extra_pumps=()):
    """Config 1 / 4 family: seeded random Hermitian H(R), r(R) on a hexagonal
    lattice (a = 2.5 A), `nvalence` filled bands, pump along z.  `extra_pumps`
    lists further (photon eV, V/A) gaussian pumps of the same width, polarisation
    and peak time (num_pump_pulses > 1, ReadInput.h:485-490, 965-967)."""
    from . import native

    seedname = f"syn_nw{nw}_r{rmax}_s{seed}"
    tb = os.path.join(directory, seedname + "_tb.dat")
    if not os.path.exists(tb):
        os.makedirs(directory, exist_ok=True)
        native.write_synthetic_tb(tb, nw, rmax, 2.5, 1.5, 0.3, decay_ang, seed)
    st = {"RK4": "", "AB": f"Adams_Bashforth {order}\n", "Taylor": f"Taylor\nTaylorOrder {order}\n"}[stepper]
    dec = f"decoherence{{\noffdiagonal {1.0 / (t2_fs / TIME_AU_FS)!r} au\n}}\n" if t2_fs else ""
    t_peak = t_fin_for(nsteps, dt) / 2.0 - 0.105 / TIME_AU_FS
    a = 2.5
    c30, s30 = math.cos(math.pi / 6), math.sin(math.pi / 6)
    text = (f"{seedname}\n\n"
            "unitcell angstrom {\n"
            f"3.35 0 0\n0 {a * c30!r} {a * s30!r}\n0 {a * c30!r} {-a * s30!r}\n"
            "}\n\n"
            "tdse{\nWannier\n"
            f"dt {dt!r} au\n"
            f"t_fin {t_fin_for(nsteps, dt)!r} au\n"
            f"{st}"
            f"num_pump_pulses {1 + len(extra_pumps)}\n"
            "}\n\n"
            f"bands{{\nvalence {nvalence}\nconduction {nw - nvalence}\n}}\n\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n\n"
            f"{dec}\n"
            + _pump(photon_ev * EV_AU, e_v_ang, fwhm_fs, pol, t_peak, phase)
            + "".join(_pump(ev * EV_AU, e0, fwhm_fs, pol, t_peak, phase) for ev, e0 in extra_pumps)
            + "\n" + PROBE_OFF + "\n"
            f"observables{{\nit_resolution {it_resolution}\nCurrent\nTAbs\n}}\n")
    return write(os.path.join(directory, "input.txt"), text)


def boron_nitride(directory, nk=48, nsteps=3000, dt=0.05, e_v_ang=0.3, photon_ev=0.6, fwhm_fs=15.0, pol=(0, 0, 1),
                  phase=0.0, it_resolution=10):
    """The reference's 2-band BoronNitride tight-binding model (a = 2.5 A), RK4,
    gaussian pump."""
    t_peak = t_fin_for(nsteps, dt) / 2.0 - 0.105 / TIME_AU_FS
    text = ("NEWINPUT\n\n"
            "tdse{\ntightbinding BoronNitride\n"
            f"dt {dt!r} au\nt_fin {t_fin_for(nsteps, dt)!r} au\nnum_pump_pulses 1\n}}\n\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n\n"
            + _pump(photon_ev * EV_AU, e_v_ang, fwhm_fs, pol, t_peak, phase)
            + "\n" + PROBE_OFF + "\n"
            f"observables{{\nit_resolution {it_resolution}\nCurrent\nTAbs\n}}\n")
    return write(os.path.join(directory, "input.txt"), text)


def batch_scan(directory, n=256, seed=3003, nsteps=3000, dt=0.05, grids=(48, 96, 144, 192), rmax=10):
    """Config 3: n independent instances.  Grids from {48, 96, 144, 192}^2, band
    counts from {2 (BoronNitride), 4, 8 (seeded Wannier, seed + i)}, and per
    instance a gaussian pump with peak field U[0.05, 0.6] V/A, CEP U[0, 2 pi),
    polarisation angle U[0, pi/3) in the yz plane, photon energy U[0.4, 1.0] eV
    and intensity FWHM U[10, 20] fs; RK4.  Returns the input paths.  `grids` and
    `rmax` shrink the same draw for parity tests (the draws do not depend on
    them)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    paths = []
    for i in range(n):
        nk = int(grids[rng.choice(len(grids))])
        nb = int(rng.choice([2, 4, 8]))
        e0 = float(rng.uniform(0.05, 0.6))
        cep = float(rng.uniform(0.0, 2.0 * math.pi))
        ang = float(rng.uniform(0.0, math.pi / 3.0))
        w = float(rng.uniform(0.4, 1.0))
        fwhm = float(rng.uniform(10.0, 20.0))
        pol = (0.0, math.cos(ang), math.sin(ang))
        d = os.path.join(directory, f"inst{i:03d}")
        if nb == 2:
            paths.append(boron_nitride(d, nk=nk, nsteps=nsteps, dt=dt, e_v_ang=e0, photon_ev=w, fwhm_fs=fwhm, pol=pol,
                                       phase=cep))
        else:
            paths.append(wannier_synthetic(d, nw=nb, rmax=rmax, nk=nk, nsteps=nsteps, dt=dt, e_v_ang=e0, nvalence=nb // 2,
                                           seed=seed + i, photon_ev=w, fwhm_fs=fwhm, pol=pol, phase=cep))
    return paths


def hbn_coulomb(directory, coef_dir, label, nk=300, nsteps=6000, dt=0.03, order=4, e_v_ang=0.3, it_resolution=10,
                photon_ev=0.6, fwhm_fs=15.0):
    """Config 2 family: the 3-band Haldane_CoreBN_Nedge hBN model of the
    reference's Coulomb example with the Hartree-Fock mean field (Ncut 20,
    Rytova-Keldysh), coefficients read from `coef_dir` (files named with
    `label`), Adams-Bashforth `order`, gaussian pump along z."""
    os.makedirs(os.path.join(directory, "Output"), exist_ok=True)
    link = os.path.join(directory, "Output", "Coulomb")
    if not os.path.exists(link):
        os.symlink(os.path.abspath(coef_dir), link)
    t_peak = t_fin_for(nsteps, dt) / 2.0 - 0.105 / TIME_AU_FS
    text = ("NEWINPUT\n\n"
            "tdse{\ntightbinding Haldane_CoreBN_Nedge gap=  7.25 t1=  -2.3 t2= 0 phi0= 0.0\n"
            f"dt {dt!r} au\nt_fin {t_fin_for(nsteps, dt)!r} au\nAdams_Bashforth {order}\nnum_pump_pulses 1\n}}\n\n"
            "Coulomb{\nqTF 0.01\nNcut 20\nepsilon_static 1.0\nRead_Coulomb_from_files\n"
            f"labelInput {label}\n\nRytova_Keldysh\n}}\n\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\n\n"
            + _pump(photon_ev * EV_AU, e_v_ang, fwhm_fs, (0, 0, 1), t_peak)
            + "\n" + PROBE_OFF + "\n"
            f"observables{{\nit_resolution {it_resolution}\nCurrent\nTAbs\n}}\n")
    return write(os.path.join(directory, "input.txt"), text)
