"""Pins of the oracle by what the mathematics fixes, for the parts no reference
output covers:

  * RK4 under a time-dependent field (RungeKutta.cpp:11-137): a driven two-level
    mode whose exact solution is a phase with the field's time integral; the
    error must fall ~16x per halving of dt.  A stage field taken at the wrong
    time (t instead of t + dt/2, say) leaves a first- or second-order method and
    fails the ratio.
  * The E_non0 gate of the gradient term (RungeKutta.cpp:151-162): on the
    step during which a ramped field switches on, the gate still holds the
    previous step's last field (zero), so the drift is off for all four stages.
    A scalar RK4 of the same mode with that gate schedule must agree to
    rounding, and the schedule without the lag must not.
  * Initialize_T (InitParameters.h:203-224) restated here from its definition
    and compared with the host's table; the dephasing and hole source of the
    right-hand side (RungeKutta.h:219-254) against the exact RK4 amplification
    factor of each element at zero field.
  * The 2-band BoronNitride and 3-band gen2d_hexagonal tables
    (TightBinding.h:271-290, 387-440, 761-800, 924-1004): U U^+ = 1, U H U^+ =
    diag of the closed-form bands, gaps at Gamma and K, periodicity in k, and the
    host's tables equal to the oracle's.

Each test builds its tables here, independently of the product host, except
where it compares the two.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from helpers import N, O
from test_gpu_engine import tb_input


def _pulse(kind="gaussian", E0=0.02, t0=0.0, sigma=6.0, wl=0.35, phase=0.3, tf=0.0, w1=0.0, w2=0.0, envelope=0):
    """A pump along z in the oracle's pulse layout (Laser.h fields)."""
    return {"gaussian": int(kind == "gaussian"), "sin2": int(kind == "sin2"), "pw": int(kind == "pw"),
            "envelope": envelope, "wl": wl, "E0": E0, "t0": t0, "tf": tf, "w1": w1, "w2": w2, "sigma": sigma,
            "phase": phase, "pol1": [0.0, 0.0, 1.0], "pol2": [0.0, 0.0, 0.0]}


OFF = _pulse(E0=0.0)


def _ring_mode_system(nk=8, mode=1, eps=(0.0, 0.3), dz=(0.4, -0.5), beta=1.3, w=0.25):
    """nk k points on a ring with the two neighbours k +- 1 at b = +-beta z, shell
    weight w; H = diag(eps), D_z = diag(dz) at every k, U = 1, no dephasing.
    For the state P[k][0][1] = a exp(2 pi i m k / nk) (rho_10 in the transposed
    storage) every term of the right-hand side is a scalar times P[k][0][1]:
    dP01/dt = -i (M11 - M00) P01 + 2 i w beta E_z sin(theta) P01,
    theta = 2 pi m / nk, M_ii = eps_i + E_z dz_i; the diagonal is uniform in k,
    so neither the commutator nor the stencil moves it."""
    nb = 2
    H = np.zeros((nk, nb, nb), dtype=np.complex128)
    D = np.zeros((nk, nb, nb, 3), dtype=np.complex128)
    for i in range(nb):
        H[:, i, i] = eps[i]
        D[:, i, i, 2] = dz[i]
    U = np.broadcast_to(np.eye(nb, dtype=np.complex128), (nk, nb, nb)).copy()
    nbr = np.stack([(np.arange(nk) + 1) % nk, (np.arange(nk) - 1) % nk], axis=1).astype(np.int32)
    bvec = np.array([[0, 0, beta], [0, 0, -beta]], dtype=np.float64)
    sysd = O.System(H, D, U, np.zeros((nb, nb)), nbr, np.array([w, w]), bvec, 0, 1, 0.0)
    theta = 2 * np.pi * mode / nk
    P = np.zeros((nk, nb, nb), dtype=np.complex128)
    P[:, 0, 0], P[:, 1, 1] = 0.8, 0.2
    a = 0.3 + 0.1j
    P[:, 0, 1] = a * np.exp(1j * theta * np.arange(nk))
    P[:, 1, 0] = np.conj(P[:, 0, 1])
    rate_e = eps[1] - eps[0]
    kappa = (dz[1] - dz[0]) - 2 * w * beta * math.sin(theta)
    return sysd, P, rate_e, kappa


def _efield_z(p, t):
    return O.field(p, t)[2]


def test_rk4_driven_mode_converges_at_fourth_order():
    """Exact: P01(t) = P01(ts) exp(-i de (t - ts) - i kappa int_ts^t E_z).  RK4
    errors at dt, dt/2, dt/4 fall by 16 each (within 13-19)."""
    sysd, P, de, kappa = _ring_mode_system()
    pulse = _pulse(E0=0.6, t0=12.0, sigma=5.0, wl=0.45, phase=0.7)
    ts, T = 0.0, 24.0  # inside 6.3 sigma of the gaussian's centre (no cutoff)
    A, _ = integrate.quad(lambda t: _efield_z(pulse, t), ts, T, epsabs=1e-15, epsrel=1e-14, limit=400)
    assert abs(A) > 0.1  # the field's integral shapes the phase
    exact = P[:, 0, 1] * np.exp(-1j * de * (T - ts) - 1j * kappa * A)
    errs = []
    for dt in (0.4, 0.2, 0.1):
        n = int(round((T - ts) / dt))
        _, Pn = sysd.run(P, 0, 4, dt, ts, n, 10 ** 9, [pulse], OFF)
        errs.append(np.max(np.abs(Pn[:, 0, 1] - exact)))
        # populations do not move; Hermiticity is kept
        np.testing.assert_allclose(Pn[:, 0, 0].real, 0.8, rtol=0, atol=1e-14)
        np.testing.assert_array_equal(Pn[:, 1, 0], np.conj(Pn[:, 0, 1]))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    print("RK4 errors", errs, "ratios", r1, r2)
    assert errs[2] > 1e-12  # above rounding, so the ratios mean something
    assert 13.0 < r1 < 19.0 and 13.0 < r2 < 19.0


def _scalar_rk4(lam, p0, ts, dt, n, gate_rule, per_stage=None):
    """RK4 of dp/dt = lam(t, g) p with the gradient gate g chosen per step by
    gate_rule(t_n, t_prev_last) where t_prev_last is the time of the previous
    step's last stage field (None on the first step); per_stage(t) instead gates
    each stage on its own field."""
    p = complex(p0)
    prev = None
    for i in range(n):
        t = ts + i * dt
        g = gate_rule(t, prev)
        if per_stage is None:
            f = lambda tt, q: lam(tt, g) * q  # noqa: E731
        else:
            f = lambda tt, q: lam(tt, per_stage(tt)) * q  # noqa: E731
        k1 = f(t, p)
        k2 = f(t + dt / 2, p + dt / 2 * k1)
        k3 = f(t + dt / 2, p + dt / 2 * k2)
        k4 = f(t + dt, p + dt * k3)
        p = p + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        prev = t + dt
    return p


def test_rk4_gradient_gate_lags_one_step():
    """A field ramped linearly from zero at t_on (the 'pw' envelope, Laser.h),
    with t_on inside a step: that step's drift is gated off by the previous
    step's last field E(t_n) = 0; from the next step on it is on."""
    sysd, P, de, kappa_full = _ring_mode_system(dz=(0.0, 0.0), beta=1.5, w=0.3)
    theta = 2 * np.pi / 8
    kd = 2 * 0.3 * 1.5 * math.sin(theta)  # the stencil's part of kappa
    dt, ts, n = 0.1, 0.0, 60
    t_on = 1.23  # inside step [1.2, 1.3]
    pulse = _pulse(kind="pw", E0=0.5, t0=t_on, w1=t_on + 2.0, w2=t_on + 50.0, tf=t_on + 60.0, envelope=1)
    ez = lambda t: _efield_z(pulse, t)  # noqa: E731
    assert ez(1.2) == 0.0 and ez(1.25) > 0.0

    def lam(t, g):
        return -1j * de + (1j * kd * ez(t) if g else 0.0)

    def gate_lag(t, prev):  # RungeKutta.cpp:151-162: |E| of the previous step's last field
        return True if prev is None else abs(ez(prev)) > 1e-20

    def stage_gate(t):  # the alternative: each stage gated on its own field
        return abs(ez(t)) > 1e-20

    _, Pn = sysd.run(P, 0, 4, dt, ts, n, 10 ** 9, [pulse], OFF)
    got = Pn[0, 0, 1]
    want = _scalar_rk4(lam, P[0, 0, 1], ts, dt, n, gate_lag)
    alt = _scalar_rk4(lam, P[0, 0, 1], ts, dt, n, lambda t, prev: True)
    other = _scalar_rk4(lam, P[0, 0, 1], ts, dt, n, gate_lag, per_stage=stage_gate)
    print("oracle", got, "lagged gate", want, "ungated", alt, "per-stage gate", other)
    assert abs(got - want) <= 1e-13 * abs(want)
    assert abs(other - want) > 1e-7  # the schedule matters at this size
    assert abs(alt - want) > 1e-7


def initialize_T(nb, nch, T1, T2, Tch):
    """InitParameters.h:203-224 from its definition: non-core block T1 on the
    diagonal and T2 off it; core rows Tch against core columns and Tch/2 against
    non-core columns; then the upper triangle mirrored into the lower."""
    T = np.zeros((nb, nb))
    for i in range(nch, nb):
        for j in range(nch, nb):
            T[i, j] = T1 if i == j else T2
    for i in range(nch):
        T[i, :nch] = Tch
        T[i, nch:] = Tch / 2.0
    for i in range(nb):
        for j in range(i + 1, nb):
            T[j, i] = T[i, j]
    return T


def test_host_T_is_initialize_T(tmp_path):
    T1, T2, Tch = 0.0041, 0.013, 0.0203
    s = N.Setup(tb_input(str(tmp_path), nk=6, t1=T1, t2=T2, tch=Tch))
    nb, nch = s.info["nb"], s.info["nch"]
    assert (nb, nch) == (3, 1)
    np.testing.assert_array_equal(s.array("T").reshape(nb, nb), initialize_T(nb, nch, T1, T2, Tch))
    (tmp_path / "bn").mkdir()
    s2 = N.Setup(tb_input(str(tmp_path / "bn"), nk=6, model="BoronNitride", t1=T1, t2=T2))
    np.testing.assert_array_equal(s2.array("T").reshape(2, 2), initialize_T(2, 0, T1, T2, 0.0))


def test_zero_field_dephasing_is_the_rk4_amplification():
    """U = 1, H = diag(eps), no field, no stencil: every element decays on its own,
    dP_ij/dt = -i (eps_j - eps_i) P_ij - T_ij P_ij + T_ii [i = j < ncc], so after n
    steps (P - Pinf)_ij = R(z_ij)^n (P - Pinf)_ij(0), R(z) = 1 + z + z^2/2 + z^3/6
    + z^4/24, z = (-i (eps_j - eps_i) - T_ij) dt and Pinf = diag(filled)."""
    nb, nch, ncc, nk = 3, 1, 2, 4
    T = initialize_T(nb, nch, 0.011, 0.037, 0.052)
    eps = np.array([-0.7, -0.2, 0.25])
    H = np.broadcast_to(np.diag(eps).astype(np.complex128), (nk, nb, nb)).copy()
    D = np.zeros((nk, nb, nb, 3), dtype=np.complex128)
    U = np.broadcast_to(np.eye(nb, dtype=np.complex128), (nk, nb, nb)).copy()
    nbr = np.zeros((nk, 1), dtype=np.int32)
    sysd = O.System(H, D, U, T, nbr, np.zeros(1), np.zeros((1, 3)), nch, ncc, 0.0)
    rng = np.random.default_rng(4)
    X = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    P = 0.1 * (X + np.conj(np.swapaxes(X, 1, 2)))
    P[:, np.arange(nb), np.arange(nb)] = [0.95, 0.6, 0.3]
    dt, n = 0.2, 50
    _, Pn = sysd.run(P, 0, 4, dt, 0.0, n, 10 ** 9, [OFF], OFF)
    Pinf = np.diag([1.0 if i < ncc else 0.0 for i in range(nb)])
    z = (-1j * (eps[None, :] - eps[:, None]) - T) * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    want = Pinf + (P - Pinf) * R[None] ** n
    err = np.max(np.abs(Pn - want))
    print("max err", err)
    assert err <= 1e-13
    # each element decayed at its own rate: the pins are not vacuous
    assert np.max(np.abs(Pn - P)) > 0.1


def _f_bn(k):
    return 1 + np.exp(2j * np.pi * k[:, 1]) + np.exp(2j * np.pi * k[:, 2])


@pytest.mark.parametrize("model", ["BoronNitride",
                                   "gen2d_hexagonal a= 2.50 gap= 5.62 t1= 2.92 t2= 0 phi0= 0 core= -20 chv_dipole= 0.05 "
                                   "chc_dipole= 0.07",
                                   "gen2d_hexagonal a= 2.50 gap= 5.62 t1= 2.92 t2= 0.15 phi0= 0.4 core= -20 chv_dipole= 0 "
                                   "chc_dipole= 0"])
def test_tight_binding_tables_fixed_by_their_closed_forms(tmp_path, model):
    s = N.Setup(tb_input(str(tmp_path), nk=6, model=model))
    name, par = s.model()
    nb = s.info["nb"]
    rng = np.random.default_rng(8)
    K = np.array([0.0, 1.0 / 3.0, -1.0 / 3.0])
    k = np.concatenate([np.zeros((1, 3)), K[None], np.c_[np.zeros(40), rng.uniform(-0.5, 0.5, size=(40, 2))]])
    H, U, D = s.eval_tb(k)
    oH, oU, oD = O.tb_tables(name, par, k, nb)
    # the host's and the oracle's restatements agree (BoronNitride: model 0)
    np.testing.assert_allclose(H, oH, rtol=0, atol=1e-15)
    np.testing.assert_allclose(U, oU, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(D, oD)
    for (h, u) in ((H, U), (oH, oU)):
        np.testing.assert_allclose(np.einsum("kai,kbi->kab", u, np.conj(u)), np.broadcast_to(np.eye(nb), u.shape), atol=1e-14)
        E = np.einsum("kai,kij,kbj->kab", u, h, np.conj(u))
        off = E - np.einsum("kaa->ka", E)[:, :, None] * np.eye(nb)
        assert np.max(np.abs(off)) <= 1e-14
        ev = np.real(np.einsum("kaa->ka", E))
        np.testing.assert_allclose(ev, np.linalg.eigvalsh(h), atol=1e-14)
    g, b = par["g"], par["b"]
    f = _f_bn(k)
    if name == "gen2d_hexagonal":
        f = np.conj(f)  # the displaced lattice of gen2d (TightBinding.h:400-415) conjugates f; |f| is the same
    # Haldane term off: the two upper bands are -+sqrt(g^2/4 + b^2 |f|^2)
    ev = np.linalg.eigvalsh(H)
    if name == "BoronNitride" or par["t2"] == 0.0:
        band = np.sqrt(g * g / 4 + b * b * np.abs(f) ** 2)
        np.testing.assert_allclose(ev[:, -2], -band, atol=1e-14)
        np.testing.assert_allclose(ev[:, -1], band, atol=1e-14)
        # gaps: 2 sqrt(g^2/4 + 9 b^2) at Gamma, g at K
        assert abs((ev[0, -1] - ev[0, -2]) - 2 * math.sqrt(g * g / 4 + 9 * b * b)) <= 1e-14
        assert abs((ev[1, -1] - ev[1, -2]) - g) <= 1e-14
    if nb == 3:
        np.testing.assert_allclose(ev[:, 0], par["core"], atol=1e-14)  # the inert core
    # periodicity in the reciprocal lattice: k and k + G give the same tables
    G = np.c_[np.zeros(len(k)), rng.integers(-2, 3, size=(len(k), 2))]
    H2, U2, D2 = s.eval_tb(k + G)
    np.testing.assert_allclose(H2, H, atol=1e-13)
    np.testing.assert_allclose(np.abs(U2), np.abs(U), atol=1e-12)
    np.testing.assert_array_equal(D2, D)
    # dipoles are the orbital positions and couplings, constant in k and symmetric
    np.testing.assert_array_equal(D, np.broadcast_to(D[:1], D.shape))
    np.testing.assert_array_equal(D, np.swapaxes(D, 1, 2))
