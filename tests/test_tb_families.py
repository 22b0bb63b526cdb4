"""The tight-binding families beyond boron nitride (TightBinding.h:44-147 init,
175-459 energy, 892-1213 unitary, 1256-1421 dipoles; lattices and band counts of
ReadInput.h:762-835): Graphene, CoreGraphene, Graphene_costD,
Graphene_symmetricD, GrapheneZurron, DLGraphene, BN1b, Vampa2015, GeS and
GeS_HSE06.

Each model's tables are pinned by what the mathematics fixes, independently of
both restatements: U U^+ = 1 and U H U^+ diagonal; the closed-form bands
(graphene +-b|f| with the Dirac point at K, the bilayer's +-(S +- d)/2, BN1b's
single band, the GeS two-band form); GrapheneZurron's cartesian structure factor
against the crystal one; Vampa2015's gap at Gamma; periodicity in the reciprocal
lattice; dipoles constant and symmetric.  Then the product host's tables equal
the oracle's restatement."""
import math

import numpy as np
import pytest

from helpers import N, O, reciprocal_rows
from test_gpu_engine import tb_input

EV = 1.0 / 27.211396
MODELS = ["Graphene", "CoreGraphene", "Graphene_costD", "Graphene_symmetricD", "GrapheneZurron", "DLGraphene", "BN1b",
          "Vampa2015", "GeS", "GeS_HSE06"]
NB = {"Graphene": 2, "CoreGraphene": 4, "Graphene_costD": 4, "Graphene_symmetricD": 4, "GrapheneZurron": 2,
      "DLGraphene": 4, "BN1b": 1, "Vampa2015": 2, "GeS": 2, "GeS_HSE06": 2}


def _fhex(k):
    return 1 + np.exp(2j * np.pi * k[:, 1]) + np.exp(2j * np.pi * k[:, 2])


def _kpoints(n=40, seed=8):
    """Gamma plus random in-plane crystal k (never exactly K, where the graphene
    eigenvectors are undefined)."""
    rng = np.random.default_rng(seed)
    return np.concatenate([np.zeros((1, 3)), np.c_[np.zeros(n), rng.uniform(-0.5, 0.5, size=(n, 2))]])


@pytest.mark.parametrize("model", MODELS)
def test_tables_unitary_diagonalising_and_equal_to_oracle(tmp_path, model):
    s = N.Setup(tb_input(str(tmp_path), nk=8, model=model))
    name, _ = s.model()
    nb = s.info["nb"]
    assert name == model and nb == NB[model]
    k = _kpoints()
    H, U, D = s.eval_tb(k)
    oH, oU, oD = O.tb_tables(name, None, k, nb, kcart=reciprocal_rows(s).T)
    np.testing.assert_allclose(H, oH, rtol=0, atol=1e-15)
    np.testing.assert_allclose(U, oU, rtol=0, atol=1e-14)
    np.testing.assert_array_equal(D, oD)
    for h, u in ((H, U), (oH, oU)):
        np.testing.assert_allclose(np.einsum("kai,kbi->kab", u, np.conj(u)), np.broadcast_to(np.eye(nb), u.shape), atol=1e-14)
        E = np.einsum("kai,kij,kbj->kab", u, h, np.conj(u))
        off = E - np.einsum("kaa->ka", E)[:, :, None] * np.eye(nb)
        assert np.max(np.abs(off)) <= 1e-14
        np.testing.assert_allclose(np.sort(np.real(np.einsum("kaa->ka", E)), axis=1), np.linalg.eigvalsh(h), atol=1e-14)
    np.testing.assert_array_equal(H, np.conj(np.swapaxes(H, 1, 2)))
    # dipoles: orbital positions and couplings, constant in k and symmetric
    np.testing.assert_array_equal(D, np.broadcast_to(D[:1], D.shape))
    np.testing.assert_array_equal(D, np.swapaxes(D, 1, 2))


@pytest.mark.parametrize("model", MODELS)
def test_bands_fixed_by_closed_forms(tmp_path, model):
    s = N.Setup(tb_input(str(tmp_path), nk=8, model=model))
    k = _kpoints(seed=9)
    H, _, _ = s.eval_tb(k)
    ev = np.linalg.eigvalsh(H)
    f = np.abs(_fhex(k))
    b_gr, b_bn, g_bn = 2.97 * EV, 3.16 * EV, 6.2 * EV
    core = -10.892178
    if model in ("Graphene", "GrapheneZurron"):
        # +-b|f|; for GrapheneZurron f is written in cartesian k, and its modulus is the crystal one
        np.testing.assert_allclose(ev, np.c_[-b_gr * f, b_gr * f], atol=1e-14)
    elif model in ("CoreGraphene", "Graphene_costD", "Graphene_symmetricD"):
        np.testing.assert_allclose(ev, np.c_[np.full(len(k), core), np.full(len(k), core), -b_gr * f, b_gr * f], atol=1e-14)
    elif model == "DLGraphene":
        d = 0.39 * EV
        S = np.sqrt(d * d + 4 * b_gr * b_gr * f * f)
        np.testing.assert_allclose(ev, np.sort(np.c_[-(S + d) / 2, -(S - d) / 2, (S - d) / 2, (S + d) / 2], axis=1), atol=1e-14)
    elif model == "BN1b":
        np.testing.assert_allclose(ev[:, 0], 0.5 * np.sqrt(g_bn ** 2 + 4 * b_bn ** 2 * f * f), atol=1e-14)
    elif model == "Vampa2015":
        # the cosine series of each band sums to zero at Gamma: the band edges there are 0 and the gap g
        assert abs(ev[0, 0]) <= 1e-15 and abs(ev[0, 1] - 0.1213) <= 1e-15
    else:  # GeS: mean diagonal +- sqrt(g^2 + |H_01|^2), g the gap parameter
        g = (0.41 if model == "GeS" else 0.62) * EV
        mid = 0.5 * np.real(H[:, 0, 0] + H[:, 1, 1])
        r = np.sqrt(g * g + np.abs(H[:, 0, 1]) ** 2)
        np.testing.assert_allclose(ev, np.c_[mid - r, mid + r], atol=1e-14)
        np.testing.assert_allclose(np.real(H[:, 0, 0] - H[:, 1, 1]), 2 * g, atol=1e-15)
    if model in ("Graphene", "CoreGraphene", "Graphene_costD", "Graphene_symmetricD", "GrapheneZurron"):
        # the Dirac point: the two graphene bands meet at K = (1/3, -1/3)
        HK, _, _ = s.eval_tb(np.array([[0.0, 1.0 / 3.0, -1.0 / 3.0]]))
        eK = np.linalg.eigvalsh(HK)[0]
        assert abs(eK[-1] - eK[-2]) <= 1e-13
    if model != "Vampa2015":  # Vampa2015's series is in k without 2 pi (not periodic in crystal units)
        rng = np.random.default_rng(3)
        G = np.c_[np.zeros(len(k)), rng.integers(-2, 3, size=(len(k), 2))]
        H2, _, _ = s.eval_tb(k + G)
        np.testing.assert_allclose(np.linalg.eigvalsh(H2), ev, atol=1e-13)


def test_lattices_of_the_families(tmp_path):
    """Band counts and lattices: the in-plane vectors the reference sets for each
    family (ReadInput.h:762-835), through the reciprocal vectors' norms."""
    ang = 1.0 / 0.52917721092
    for model, (nch, ncc, nb), a1 in (("Graphene", (0, 1, 2), 2.46 * ang), ("CoreGraphene", (2, 3, 4), 2.46 * ang),
                                      ("DLGraphene", (0, 2, 4), math.hypot(4.02574074, 2.324263)),
                                      ("BN1b", (0, 1, 1), 2.5 * ang * math.sqrt(1.5)), ("Vampa2015", (0, 1, 2), 6.14),
                                      ("GeS", (0, 1, 2), math.hypot(4.53, 3.63) / 2 * ang)):
        d = tmp_path / model
        d.mkdir()
        s = N.Setup(tb_input(str(d), nk=8, model=model))
        info = s.info
        assert (info["nch"], info["ncc"], info["nb"]) == (nch, ncc, nb), model
        a = s.array("a").reshape(3, 3)
        assert abs(np.linalg.norm(a[1]) - a1) <= 1e-9 * a1, model


@pytest.mark.parametrize("model", ["DLGraphene", "BN1b", "Vampa2015", "GeS"])
def test_shells_and_neighbours_of_the_other_lattices(tmp_path, model):
    """The bilayer, BN1b, orthorhombic (Vampa2015) and rhombic (GeS) lattices need
    other shell sets than the hexagonal one: the host's closed-form neighbour
    table equals the oracle's brute-force search (CalculateWeigths.h:145-226) and
    the B1 system is solved (CalculateWeigths.h:18-133)."""
    import helpers as Hh
    s = N.Setup(tb_input(str(tmp_path), nk=16, model=model))
    o = Hh.oracle_system(s)
    np.testing.assert_array_equal(o["nbr"], s.nbr())
    assert o["res"] < 1e-7
    w = np.repeat(o["w"], o["sizes"])
    # the weights are pinv(A) q, the minimum-norm solution: these lattices' shell
    # systems are rank-deficient, where another least-squares solution differs
    np.testing.assert_allclose(s.array("nbr_wb").reshape(-1, 3), w[:, None] * o["bvec"], rtol=0, atol=1e-13)
    # sum_b w b_a b_b = delta_ab in the plane of the grid
    M = np.einsum("n,na,nb->ab", w, o["bvec"], o["bvec"])
    np.testing.assert_allclose(M[1:, 1:], np.eye(2), atol=1e-8)
