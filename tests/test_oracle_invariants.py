"""Checks of the oracle against what the mathematics fixes, for the parts the AB5
fixture leaves open (RK4, Adams-Bashforth orders 1-4, the Taylor multistep
scheme, the Wannier phase sum, the initial projector, the neighbour search)."""
import ctypes as C

import numpy as np
import numpy.polynomial.polynomial as P
import pytest

import helpers as Hh
from helpers import N, O

O.lib.orc_ab_update.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.c_long, C.POINTER(C.c_double)]
O.lib.orc_taylor_update.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.c_long, C.POINTER(C.c_double)]


def _hist_from_poly(coefs, t, dt, order):
    """f_s = p(t - s dt) for s = 0..order-1, complex polynomial per element."""
    n = coefs.shape[0]
    f = np.zeros((order, n), dtype=np.complex128)
    for s in range(order):
        f[s] = [P.polyval(t - s * dt, c) for c in coefs]
    return f


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5])
def test_adams_bashforth_integrates_interpolant_exactly(order):
    """AB-s: delta = int_t^{t+dt} p(u) du for every polynomial of degree <= s-1
    (the defining property of the Adams-Bashforth weights)."""
    rng = np.random.default_rng(order)
    n, dt, t = 7, 0.13, 1.7
    coefs = rng.normal(size=(n, order)) + 1j * rng.normal(size=(n, order))
    f = _hist_from_poly(coefs, t, dt, order)
    d = np.zeros(n, dtype=np.complex128)
    O.lib.orc_ab_update(order, dt, f.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), n,
                        d.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    exact = np.array([P.polyval(t + dt, P.polyint(c)) - P.polyval(t, P.polyint(c)) for c in coefs])
    np.testing.assert_allclose(d, exact, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("order", [2, 3, 4, 5, 6])
def test_taylor_scheme_as_written(order):
    """The reference's Taylor step fits f(t) - f(t - i dt) = sum_p alpha_p (i dt)^(p+1)
    and adds f dt + sum dt^(i+1) alpha_(i-1)/(i+1).  For any polynomial f of degree
    <= s-1 that equals 2 f(t) dt - int_(t-dt)^t f(u) du (signs as written; exact
    for degree <= 1, second order in general)."""
    rng = np.random.default_rng(10 + order)
    n, dt, t = 5, 0.05, 0.3
    coefs = rng.normal(size=(n, order)) + 1j * rng.normal(size=(n, order))
    f = _hist_from_poly(coefs, t, dt, order)
    d = np.zeros(n, dtype=np.complex128)
    O.lib.orc_taylor_update(order, dt, f.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), n,
                            d.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    back = np.array([P.polyval(t, P.polyint(c)) - P.polyval(t - dt, P.polyint(c)) for c in coefs])
    exact = 2 * f[0] * dt - back
    # tau is a scaled Vandermonde matrix: its inverse carries cond(tau) * eps
    np.testing.assert_allclose(d, exact, rtol=1e-9, atol=1e-12)


def _tiny_system(eps, nk=2):
    nb = len(eps)
    H = np.zeros((nk, nb, nb), dtype=np.complex128)
    H[:, np.arange(nb), np.arange(nb)] = eps
    D = np.zeros((nk, nb, nb, 3), dtype=np.complex128)
    U = np.tile(np.eye(nb, dtype=np.complex128), (nk, 1, 1))
    T = np.zeros((nb, nb))
    nbr = np.arange(nk, dtype=np.int32).reshape(nk, 1)
    return O.System(H, D, U, T, nbr, np.zeros(1), np.zeros((1, 3)), 0, 1, 0.0)


ZERO_PULSE = {"gaussian": 0, "sin2": 0, "pw": 0, "envelope": 0, "wl": 0.0, "E0": 0.0, "t0": 0.0, "tf": 0.0, "w1": 0.0,
              "w2": 0.0, "sigma": 0.0, "phase": 0.0, "pol1": [0, 0, 0], "pol2": [0, 0, 0]}


def test_rk4_free_precession_closed_form():
    """Zero field, diagonal H: P_ij' = -i(e_j - e_i) P_ij, so n RK4 steps multiply
    P_ij by R(z)^n, R(z) = 1 + z + z^2/2 + z^3/6 + z^4/24, z = -i(e_j - e_i) dt."""
    eps = np.array([-0.3, 0.05, 0.42])
    sysd = _tiny_system(eps)
    rng = np.random.default_rng(3)
    X = rng.normal(size=(2, 3, 3)) + 1j * rng.normal(size=(2, 3, 3))
    P0 = X + np.conj(np.swapaxes(X, 1, 2))
    dt, n = 0.7, 50
    _, Pn = sysd.run(P0, 0, 4, dt, 0.0, n, 10 ** 9, [ZERO_PULSE], ZERO_PULSE)
    z = -1j * (eps[None, :] - eps[:, None]) * dt
    R = 1 + z + z ** 2 / 2 + z ** 3 / 6 + z ** 4 / 24
    np.testing.assert_allclose(Pn, P0 * R[None] ** n, rtol=1e-12, atol=1e-13)


def test_wannier_sum_equals_fft():
    """H(k) = sum_R e^{2 pi i k.R} H(R) on the grid k = -1/2 + i/N is (-1)^R-weighted
    inverse DFT of H(R): numpy's FFT is the independent reference."""
    rng = np.random.default_rng(5)
    nw, N1, N2, rmax = 3, 6, 8, 2
    R = np.array([(0, a, b) for a in range(-rmax, rmax + 1) for b in range(-rmax, rmax + 1)], dtype=np.int32)
    HR = rng.normal(size=(len(R), nw, nw)) + 1j * rng.normal(size=(len(R), nw, nw))
    rR = rng.normal(size=(len(R), nw, nw, 3)) + 1j * rng.normal(size=(len(R), nw, nw, 3))
    kpt = O.kgrid(1, N1, N2)
    H, D = O.wannier_tables(kpt, R, HR, rR, nw)
    grid = np.zeros((N1, N2, nw, nw), dtype=np.complex128)
    gridr = np.zeros((N1, N2, nw, nw, 3), dtype=np.complex128)
    for r, (_, a, b) in enumerate(R):
        sgn = (-1.0) ** (a + b)
        grid[a % N1, b % N2] += sgn * HR[r]
        gridr[a % N1, b % N2] += sgn * rR[r]
    Hf = np.fft.ifft2(grid, axes=(0, 1)) * N1 * N2
    Df = np.fft.ifft2(gridr, axes=(0, 1)) * N1 * N2
    Hf = Hf.reshape(-1, nw, nw).copy()
    Df = Df.reshape(-1, nw, nw, 3)
    idx = np.arange(nw)
    Hf[:, idx, idx] = Hf[:, idx, idx].real
    Df = Df.copy()
    Df[:, idx, idx, :] = Df[:, idx, idx, :].real
    Df = 0.5 * (Df + np.swapaxes(Df, 1, 2))
    np.testing.assert_allclose(H, Hf, rtol=0, atol=1e-12)
    np.testing.assert_allclose(D, Df, rtol=0, atol=1e-12)


def test_rho0_is_the_filled_band_projector(tmp_path):
    s = N.Setup(Hh.AB5 + "/input.txt")
    U = s.array("U", True).reshape(-1, 3, 3)
    P0 = O.rho0(U, 2)
    Pt = np.swapaxes(P0, 1, 2)  # rho = P^T
    np.testing.assert_allclose(Pt @ Pt, Pt, atol=1e-14)
    np.testing.assert_allclose(np.trace(Pt, axis1=1, axis2=2), 2.0, atol=1e-14)
    # in the eigenbasis (rows of U) rho is diag(1, 1, 0)
    rho_e = U @ Pt @ np.conj(np.swapaxes(U, 1, 2))
    np.testing.assert_allclose(rho_e, np.tile(np.diag([1.0, 1.0, 0.0]), (len(U), 1, 1)), atol=1e-14)


SQUARE_INPUT = """NEWINPUT
unitcell angstrom {
10 0 0
0 %s 0
0 0 %s
}
tdse{
tightbinding BoronNitride
dt 0.1 au
t_fin 10 au
}
nkPT{
1 %d %d
}
laserpump sin2 {
cycles 3
wavelength 400 nm
intensity 1e11 wcm2
polarization 0 0 1
}
laserprobe gaussian {
sigma 0.08 fs
intensity 0
frequency 15.06
polarization 1 0 0
}
"""


@pytest.mark.parametrize("n1,n2", [(8, 8), (12, 12), (10, 16), (30, 30)])
def test_neighbours_closed_form_equals_bruteforce(tmp_path, n1, n2):
    """Integer tables must be bit-exact: the closed form (host) against the
    reference's O(Nk^2) search (oracle) on hexagonal grids."""
    p = tmp_path / "input.txt"
    text = open(Hh.AB5 + "/input.txt").read().replace("1 70 70", f"1 {n1} {n2}")
    p.write_text(text)
    s = N.Setup(str(p), verify_neighbours=True)
    kpt = s.array("kpt").reshape(-1, 3)
    ob, sizes, w, res = O.shells(kpt, [1, n1, n2], Hh.reciprocal_rows(s))
    np.testing.assert_array_equal(O.neighbours(kpt, Hh.reciprocal_rows(s), ob), s.nbr())
    # every neighbour list is a permutation of the grid per direction
    for n in range(s.nbr().shape[1]):
        assert len(np.unique(s.nbr()[:, n])) == len(kpt)


def test_coulomb_term_equals_bruteforce_convolution():
    """The separable Fourier form of the exchange term (Coulomb.h:273-541) is the
    convolution X_ij(k) = -(1/N) sum_k' sum_mn (1+corr_mn) V_mn(k-k') P_ij(k') with
    V_mn(d) = A cos(m dx)cos(n dy) + D sin(m dx)sin(n dy) - B cos(m dx)sin(n dy)
    - C sin(m dx)cos(n dy) (B, C on off-diagonal elements; the diagonal uses the
    element-0 coefficients), stored as X[j][i]; the Hartree shift is
    Re V_H . <P_jj>.  Checked by brute force on an 8x6 grid."""
    rng = np.random.default_rng(21)
    N1, N2, nb, ncut = 8, 6, 2, 3
    kpt = O.kgrid(1, N1, N2)
    nk = len(kpt)
    X = rng.normal(size=(nk, nb, nb)) + 1j * rng.normal(size=(nk, nb, nb))
    Pm = X + np.conj(np.swapaxes(X, 1, 2))
    A, B, Cc, D = (rng.normal(size=(ncut * ncut, nb * nb)) + 1j * rng.normal(size=(ncut * ncut, nb * nb)) for _ in range(4))
    screen = rng.normal(size=nb * nb) + 0j
    vh = rng.normal(size=(nb, nb)) + 0j
    H = np.zeros((nk, nb, nb), dtype=np.complex128)
    sysd = O.System(H, np.zeros((nk, nb, nb, 3)), np.tile(np.eye(nb), (nk, 1, 1)), np.zeros((nb, nb)),
                    np.zeros((nk, 1), dtype=np.int32), np.zeros(1), np.zeros((1, 3)), 0, 1, 0.0)
    sysd.set_coulomb(ncut, kpt, A, B, Cc, D, screen, vh)
    Xk, EH = sysd.coulomb_x(Pm)
    ref = np.zeros((nk, nb, nb), dtype=np.complex128)
    dx = 2 * np.pi * (kpt[:, None, 1] - kpt[None, :, 1])
    dy = 2 * np.pi * (kpt[:, None, 2] - kpt[None, :, 2])
    for i in range(nb):
        for j in range(i, nb):
            e = i * nb + j
            acc = np.zeros(nk, dtype=np.complex128)
            for m in range(ncut):
                for n in range(ncut):
                    mm = m * ncut + n
                    corr = -4 * np.pi ** 2 * (m * m + n * n) / (nk * 27)
                    if i == j:
                        V = A[mm, 0] * np.cos(m * dx) * np.cos(n * dy) + D[mm, 0] * np.sin(m * dx) * np.sin(n * dy)
                    else:
                        V = (A[mm, e] * np.cos(m * dx) * np.cos(n * dy) + D[mm, e] * np.sin(m * dx) * np.sin(n * dy)
                             - B[mm, e] * np.cos(m * dx) * np.sin(n * dy) - Cc[mm, e] * np.sin(m * dx) * np.cos(n * dy))
                    acc += (1 + corr) * (V @ Pm[:, i, j]) / nk
            ref[:, j, i] = -acc
            if i != j:
                ref[:, i, j] = np.conj(-acc)
    np.testing.assert_allclose(Xk, ref, rtol=0, atol=1e-12 * np.max(np.abs(ref)))
    np.testing.assert_allclose(EH, np.real(vh) @ np.real(np.mean(np.einsum("kii->ki", Pm), axis=0)), rtol=1e-13)
