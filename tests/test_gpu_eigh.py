"""Batched Hermitian eigensolve of the setup (edus_eigh_batched, replacing the
per-k eig_sym of Wannier.h:496-542) against numpy's eigh with the same phase
convention: ascending eigenvalues, U rows = conj(eigenvectors), each
eigenvector's largest-modulus component real positive."""
import numpy as np
import pytest

from helpers import N

pytestmark = pytest.mark.gpu


def _random_hermitian(n, nb, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, nb, nb)) + 1j * rng.normal(size=(n, nb, nb))
    return 0.5 * (X + np.conj(np.swapaxes(X, 1, 2)))


def _reference(H):
    w, V = np.linalg.eigh(H)                  # columns v_a, ascending
    U = np.conj(np.swapaxes(V, 1, 2))         # rows conj(v_a)
    best = np.argmax(np.abs(U), axis=2)
    ub = np.take_along_axis(U, best[:, :, None], axis=2)
    return w, U * (np.conj(ub) / np.abs(ub))


@pytest.mark.parametrize("nb,n", [(5, 300), (16, 2000), (32, 40000)])
def test_eigh_matches_numpy(nb, n):
    H = _random_hermitian(n, nb, nb)
    w, U = N.eigh_batched(H)
    wr, Ur = _reference(H)
    scale = np.max(np.abs(wr))
    assert np.max(np.abs(w - wr)) <= 1e-12 * scale
    # U H U^+ is diagonal with the eigenvalues, U is unitary
    E = U @ H @ np.conj(np.swapaxes(U, 1, 2))
    assert np.max(np.abs(E - w[:, :, None] * np.eye(nb))) <= 1e-12 * scale
    assert np.max(np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - np.eye(nb))) <= 1e-12
    # random spectra are non-degenerate: the phased eigenvectors are unique
    assert np.max(np.abs(U - Ur)) <= 1e-9
    # the convention itself: the largest component of every row is real positive
    best = np.argmax(np.abs(U), axis=2)
    ub = np.take_along_axis(U, best[:, :, None], axis=2)
    assert np.max(np.abs(ub.imag)) <= 1e-14 and np.min(ub.real) > 0
