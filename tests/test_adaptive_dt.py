"""Adaptive-dt loop (Source_Main/dyn_dt_evolution.cpp) in the oracle and the host.

The oracle's adaptive loop is pinned three ways:
- with a step bound it never reaches, it reproduces the constant-dt loop (pinned
  by the reference's AB5 run and the stepper invariants) bit for bit;
- replaying its recorded dt sequence one RK4 step at a time, an independent
  numpy computation of the step-size measure reproduces every dt update and the
  time axis (time_loop advances by the *updated* dt);
- the Taylor update with a non-uniform dt history equals 2 f0 dt minus the
  integral of the interpolating polynomial through the nodes the reference uses
  (t_0 = dt_prev[0], t_i = t_(i-1) + dt_prev[i-1]).
"""
import ctypes as C
import os

import numpy as np
import numpy.polynomial.polynomial as P
import pytest

import helpers as Hh
from helpers import N, O

TIME_AU_FS = O.TIME_AU_FS


def dyn_input(d, stepper="", eps=1e-9, nk=12, dt=0.2, t_fin=30.0, res=2, pump_int=3e13, dyn=True):
    """gen2d hBN with a gaussian pump centred at t = 0 (window -8.68 .. t_fin au)."""
    text = ("NEWINPUT\ntdse{\ntightbinding gen2d_hexagonal a= 2.50 gap= 5.62 t1= 2.92 t2= 0.1 phi0= 0.3 core= -20 "
            "chv_dipole= 0.05 chc_dipole= 0.07\n"
            f"dt {dt!r} au\nt_fin {t_fin!r} au\n{stepper}" + ("dynamical_dt_evolution\n" if dyn else "") +
            f"epsStepAbs {eps!r}\nnum_pump_pulses 1\n}}\n"
            f"nkPT{{\n1 {nk} {nk}\n}}\ndecoherence{{\noffdiagonal 0.01 au\ncorehole 0.004 au\n}}\n"
            "laserpump gaussian {\nfrequency 0.05\n"
            f"intensity {pump_int!r} wcm2\nsigma 0.3 fs\npolarization 0 0 1\nt0_pulse -1.2 fs\n}}\n"
            "laserprobe gaussian {\nsigma 0.08 fs\nintensity 1e9\nfrequency 15.06\npolarization 0 1 0\n}\n"
            f"observables{{\nit_resolution {res}\nCurrent\nTAbs\n}}\n")
    p = os.path.join(d, "input.txt")
    with open(p, "w") as f:
        f.write(text)
    return p


def test_setup_reads_adaptive_flags(tmp_path):
    s = N.Setup(dyn_input(str(tmp_path), eps=0.02))
    assert s.info["dyn_dt"] == 1 and s.info["eps_step"] == 0.02 and s.info["stepper"] == N.STEPPER_RK4
    # the adaptive loop runs RK4 unless Taylor is asked for (dyn_dt_evolution.cpp:51-58)
    s = N.Setup(dyn_input(str(tmp_path), stepper="Adams_Bashforth 4\n"))
    assert s.info["stepper"] == N.STEPPER_RK4
    s = N.Setup(dyn_input(str(tmp_path), stepper="Taylor\nTaylorOrder 6\n"))
    assert s.info["stepper"] == N.STEPPER_TAYLOR and s.info["order"] == 6


@pytest.mark.parametrize("stepper", ["", "Taylor\nTaylorOrder 5\n"])
def test_unreachable_bound_reproduces_constant_loop(tmp_path, stepper):
    s = N.Setup(dyn_input(str(tmp_path), stepper=stepper, eps=1e9))
    osys = Hh.oracle_system(s)
    n = 60
    rows_d, P_d, dts = Hh.oracle_run_dyn(s, osys, n)
    rows_c, P_c = Hh.oracle_run(s, osys, n)
    assert len(dts) == n and np.all(dts == s.info["dt"])
    np.testing.assert_array_equal(rows_d, rows_c)
    np.testing.assert_array_equal(P_d, P_c)


def _excited_max(P, U, ncc):
    """P_cond_max (Observables_MPI.h:330-347): largest excited population over k in
    the eigenbasis, holes for bands < ncc."""
    Ps = np.triu(P) + np.conj(np.swapaxes(np.triu(P, 1), 1, 2))
    pop = np.real(np.einsum("kai,kij,kaj->ka", np.conj(U), Ps, U))
    ex = np.abs(pop)
    ex[:, :ncc] = np.abs(1.0 - pop[:, :ncc])
    return float(ex.max())


def test_rk4_dt_control_replayed_in_numpy(tmp_path):
    eps = 2e-3
    s = N.Setup(dyn_input(str(tmp_path), eps=eps, dt=0.2, res=1))
    osys = Hh.oracle_system(s)
    info = s.info
    nsteps = 50
    rows, Pf, dts = Hh.oracle_run_dyn(s, osys, nsteps, it_res=1)
    assert len(dts) == nsteps
    assert dts.min() < 0.5 * info["dt"], "the control never engaged; make the pump stronger"
    pumps = [s.pulse(i) for i in range(info["n_pumps"])]
    probe = s.pulse(info["n_pumps"])
    G = Hh.gradH_from(osys["H"], osys["nbr"], np.repeat(osys["w"], osys["sizes"]), osys["bvec"])
    dk3 = 1.0 / (info["nk0"] * info["nk1"] * info["nk2"])
    Pst = O.rho0(osys["U"], info["ncc"])
    t = info["t_start"]
    dt0 = info["dt"]
    for i in range(nsteps - 1):
        np.testing.assert_allclose(rows[i, 0], t * TIME_AU_FS, rtol=1e-13)
        pc = _excited_max(Pst, osys["U"], info["ncc"])
        n_cond = max(0.0, float(np.max(rows[i, 4:4 + info["nb"]])))
        _, Pn = osys["sys"].run(Pst, 0, 4, dts[i], t, 1, 1, pumps, probe, gradH=G, j_norm=info["j_norm"],
                                jacobian=info["jacobian"], dk3=dk3, nktot=info["nktot"])
        d_old = np.real(np.einsum("kii->ki", Pst))
        d_new = np.einsum("kii->ki", Pn)
        mdp = float(np.max(2 * np.abs(d_old - d_new) / (pc + n_cond + np.abs(d_new) + 1e-7)))
        dt = dts[i]
        dt = dt * eps / mdp if (mdp > eps and dt > 0.01) else (dt * 1.02 if dt < dt0 else dt)
        np.testing.assert_allclose(dts[i + 1], dt, rtol=1e-9)
        t += dts[i + 1]
        Pst = Pn


@pytest.mark.parametrize("order", [3, 4, 6])
def test_taylor_update_with_dt_history(order):
    rng = np.random.default_rng(20 + order)
    n = 4
    dt_prev = np.array([0.05, 0.04, 0.047, 0.03, 0.05, 0.033, 0.05, 0.05][:order])
    dt = dt_prev[0]
    nodes = np.zeros(order)
    nodes[1] = dt_prev[0]
    for i in range(2, order):
        nodes[i] = nodes[i - 1] + dt_prev[i - 2]
    f = rng.normal(size=(order, n)) + 1j * rng.normal(size=(order, n))
    d = np.zeros(n, dtype=np.complex128)
    dp = np.ascontiguousarray(dt_prev)
    O.lib.orc_taylor_update_hist(order, dt, dp.ctypes.data_as(C.POINTER(C.c_double)),
                                 f.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), n,
                                 d.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    for e in range(n):
        V = np.vander(nodes, order, increasing=True)
        c = np.linalg.solve(V, f[:, e])          # q(s) through (t_i, f_i), s = backward offset
        exact = 2 * f[0, e] * dt - P.polyval(dt, P.polyint(c))
        assert abs(d[e] - exact) <= 1e-9 * max(1.0, abs(exact))
