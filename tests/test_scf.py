"""Full-SCF energies of the SCF driver (serenity_amd.scf) against the reference's own tests,
with the oracle's XC build standing in for the GPU build (the GPU form is in test_gpu_scf.py).

Pins (total energies of Serenity's ScfTask: SMALL-grid iterations, final DEFAULT-grid
energy from the converged density, ScfTask.cpp:127-153, 358-359):

* H2/STO-6G LDA without density fitting, src/scf/Scf_test.cpp:169-190: -1.1298876147476988
  (tolerance 1e-6 as there).
* H2/STO-6G LDA with RI-J, Scf_test.cpp:144-167: -1.1299388663978864.  The value matches
  a Cartesian RI_J_WEIGEND basis (H carries a d fitting function): the test geometry's
  atoms are shared objects that keep the first auxiliary basis loaded on them, and the
  stored TestSystem_H2_MINBAS settings (data/testresources) have MAKESPHERICALBASIS false.
  A spherical auxiliary basis lands 6.9e-7 lower.
* water/STO-3G PBE with RI-J, qcserenity/tests/test_tasks.py:50-76: -75.2218044178 with
  LibXC's PBE and -75.22180876683022 with XCFun's (the unrestricted run of a closed shell;
  places=7 there, i.e. 5e-8).
* triplet O2 / 6-31G / PBE, unrestricted, exact J, accuracy-7 grids
  (data/testresources/TestSystem_O2_TRIP_6_31G_PBE_TDA): the total of the stored energy
  components (ids 3 + 5 + 6 + 202) -150.1286472735, to 1e-8.
* ethane/def2-SVP Hartree-Fock with exact four-centre J and K, Cartesian and spherical d
  shells, src/tasks/ScfTask_test.cpp:43-88: -79.1786812939 and -79.1728833551 at 1e-8 --
  no XC involved, so this pins the integrals (s, p, d; both AO conventions) and the SCF.
* the ATOM_SCF initial guess, src/scf/initialGuess/AtomicDensityGuessCalculator_test.cpp:42-56:
  O2/STO-6G density elements after one HF Fock diagonalisation of the guess, at 1e-6.
"""
import json

import numpy as np
import pytest

from serenity_amd import scf, xc
from oracle import oracle as O

from test_integrals import water_geometry


def oracle_xc(basis, grid, functional):
    def build(P):
        F, E, _, _ = O.fock_build(basis, grid.points, grid.weights, P, functional)
        return E, F
    return build


def scf_energy(Z, X, basis, functional, aux=None, aux_spherical=True, spherical=True):
    b = xc.BasisController.from_file(basis, Z, X, spherical=spherical)
    a = None if aux is None else xc.BasisController.from_file(aux, Z, X, spherical=aux_spherical)
    ints = scf.ScfIntegrals(b, Z, X, a)
    small = xc.molecular_grid(Z, X, accuracy=2)
    default = xc.molecular_grid(Z, X, accuracy=4)
    r = scf.perform(ints, int(np.sum(Z)) // 2, oracle_xc(b, small, functional))
    assert r["converged"]
    _, e_xc, _, _ = O.fock_build(b, default.points, default.weights, r["P"], functional)
    return scf.total_energy(ints, r["P"], e_xc), r


def h2():
    """TestSystem_H2_MINBAS geometry: H at z = -0.7 and +0.7 bohr (SystemController__TEST_SUPPLY.cpp:1744-1745)."""
    return np.array([1, 1]), np.array([[0.0, 0.0, -0.7], [0.0, 0.0, 0.7]])


def test_h2_sto6g_lda_no_ri():
    Z, X = h2()
    e, r = scf_energy(Z, X, "STO-6G", "LDA")
    assert abs(e - (-1.1298876147476988)) < 1e-7
    assert r["cycles"] < 20


def test_h2_sto6g_lda_ri():
    Z, X = h2()
    e, _ = scf_energy(Z, X, "STO-6G", "LDA", aux="RI_J_WEIGEND", aux_spherical=False)
    assert abs(e - (-1.1299388663978864)) < 1e-7
    e_sph, _ = scf_energy(Z, X, "STO-6G", "LDA", aux="RI_J_WEIGEND", aux_spherical=True)
    assert 5e-7 < (-1.1299388663978864) - e_sph < 9e-7


@pytest.mark.parametrize("functional,ref", [("PBE", -75.22180876683022), ("PBE_LIBXC", -75.2218044178)])
def test_water_sto3g_pbe(functional, ref):
    Z, X = water_geometry()
    e, r = scf_energy(Z, X, "STO-3G", functional, aux="RI_J_WEIGEND")
    assert abs(e - ref) < 5e-8
    # the SMALL-grid energy differs from the final DEFAULT-grid one: the protocol matters
    assert abs(r["energy"] - e) > 1e-6


def test_scf_converges_from_any_guess_to_the_same_energy():
    Z, X = water_geometry()
    b = xc.BasisController.from_file("STO-3G", Z, X)
    ints = scf.ScfIntegrals(b, Z, X, xc.BasisController.from_file("RI_J_WEIGEND", Z, X))
    g = xc.molecular_grid(Z, X, accuracy=2)
    build = oracle_xc(b, g, "LDA")
    r0 = scf.perform(ints, 5, build)
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((b.nao, 5)))
    Xo = scf.canonical_orthogonalization(ints.S)
    Co = Xo @ Q
    r1 = scf.perform(ints, 5, build, P0=2.0 * Co @ Co.T)
    assert r0["converged"] and r1["converged"]
    assert abs(r0["energy"] - r1["energy"]) < 1e-9
    # idempotency and electron count of the converged density
    P, S = r0["P"], ints.S
    assert abs(np.trace(P @ S) - 10.0) < 1e-10
    assert np.abs(P @ S @ P - 2.0 * P).max() < 1e-10


def ethane():
    """EthaneA of SystemController__TEST_SUPPLY.cpp:2613-2634 (Angstrom)."""
    A = xc.ANGSTROM_TO_BOHR
    Z = np.array([6, 6, 1, 1, 1, 1, 1, 1])
    X = np.array([[-3.86799, 1.86633, 0.0], [-2.662850, 2.791330, 0.0], [-4.69801, 2.32927, 0.57419],
                  [-3.59828, 0.89649, 0.46861], [-4.20382, 1.68541, -1.04280], [-2.32703, 2.97225, 1.04280],
                  [-2.93257, 3.76118, -0.46861], [-1.83284, 2.32840, -0.57419]]) * A
    return Z, X


@pytest.mark.parametrize("spherical,ref", [(False, -79.1786812939), (True, -79.1728833551)])
def test_ethane_svp_hartree_fock(spherical, ref):
    Z, X = ethane()
    b = xc.BasisController.from_file("DEF2-SVP", Z, X, spherical=spherical)
    ints = scf.ScfIntegrals(b, Z, X)
    r = scf.perform(ints, 9, None, exchange=1.0)
    assert r["converged"]
    assert abs(r["energy"] - ref) < 1e-8


def test_o2_triplet_unrestricted_pbe():
    from test_unrestricted import o2_system
    Z, X, b, _, _, _, _, E = o2_system()
    ints = scf.ScfIntegrals(b, Z, X)
    g = xc.molecular_grid(Z, X, accuracy=7)

    def build(Pa, Pb):
        Fa, Fb, e = O.fock_build_pol(b, g.points, g.weights, Pa, Pb, "PBE")
        return e, Fa, Fb
    r = scf.perform_unrestricted(ints, 9, 7, build)
    assert r["converged"]
    ref = E[3] + E[5] + E[6] + E[202]
    assert abs(r["energy"] - ref) < 1e-8
    S = ints.S
    assert abs(np.trace(r["Pa"] @ S) - 9) < 1e-10 and abs(np.trace(r["Pb"] @ S) - 7) < 1e-10


def test_nuclear_repulsion():
    Z, X = water_geometry()
    e = scf.nuclear_repulsion(Z, X)
    d = lambda a, b: np.linalg.norm(X[a] - X[b])
    assert abs(e - (1 / d(0, 1) + 8 / d(0, 2) + 8 / d(1, 2))) < 1e-13


def test_atom_scf_guess_matches_reference():
    """ATOM_SCF initial guess (AtomicDensityGuessCalculator_test.cpp:42-56): O2 / STO-6G at 1.21 A; the
    stored MINAO free-atom densities projected into STO-6G, one HF Fock matrix of that density
    (Settings' default method) diagonalised; the reference's density elements to 1e-6.  The projection
    of MINAO onto itself is the identity, and the guess of water holds ~10 electrons."""
    Z = np.array([8, 8])
    X = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.21 * xc.ANGSTROM_TO_BOHR]])
    b = xc.BasisController.from_file("STO-6G", Z, X)
    ints = scf.ScfIntegrals(b, Z, X)
    P0 = scf.atomic_density_guess("STO-6G", Z)
    assert np.abs(P0[:5, 5:]).max() == 0.0  # atom blocks only
    F = ints.h + ints.coulomb(P0) - 0.5 * ints.exchange(P0)
    Xo = scf.canonical_orthogonalization(ints.S)
    _, Cp = np.linalg.eigh(Xo.T @ F @ Xo)
    C = Xo @ Cp[:, :8]
    P = 2.0 * C @ C.T
    ref = {(0, 0): 2.1071094855194836, (1, 0): -0.46128984211922747, (3, 3): 0.88024781372049266,
           (5, 0): -7.3034439855780155e-05, (0, 2): 0.0}
    for k, v in ref.items():
        assert abs(P[k] - v) < 1e-6, (k, P[k], v)
    Pm = scf.atomic_density_guess("MINAO", [8])
    stored = json.load(open(scf.GUESS_FILE))["spherical"]["O"]
    assert np.abs(Pm - np.reshape(stored["P"], (5, 5))).max() < 1e-14
    Zw, Xw = water_geometry()
    bw = xc.BasisController.from_file("DEF2-SVP", Zw, Xw)
    assert abs(np.vdot(scf.atomic_density_guess("DEF2-SVP", Zw), bw.overlap()) - 10.0) < 1e-2


@pytest.mark.parametrize("molecule", ["water", "hcl"])
def test_atom_scf_guess_converges_to_the_same_energy(molecule):
    """Water and HCl (1.2746 A) / def2-SVP / LDA with RI-J from the ATOM_SCF and the core-Hamiltonian
    guesses: the same SCF energy, in fewer cycles from the atomic densities."""
    if molecule == "water":
        Z, X = water_geometry()
    else:
        Z, X = np.array([17, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.2746 * xc.ANGSTROM_TO_BOHR]])
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    a = xc.BasisController.from_file("RI_J_WEIGEND", Z, X)
    ints = scf.ScfIntegrals(b, Z, X, a)
    grid = xc.molecular_grid(Z, X, accuracy=2)
    build = oracle_xc(b, grid, "LDA")
    nocc = int(np.sum(Z)) // 2
    r_core = scf.perform(ints, nocc, build)
    r_atom = scf.perform(ints, nocc, build, P0=scf.atomic_density_guess("DEF2-SVP", Z))
    assert abs(r_core["energy"] - r_atom["energy"]) < 1e-9
    assert r_atom["cycles"] < r_core["cycles"]


@pytest.mark.parametrize("z", range(1, 19))
def test_free_atom_guess_holds_z_electrons(z):
    """The stored free-atom density of every element H..Ar holds Z electrons in the MINAO basis it is
    stored in (tr P S = Z to rounding); projected into def2-TZVP it keeps Z to 2e-3 (def2-TZVP spans the
    MINAO orbitals of H..Ar to that accuracy).  A wrong element, AO order or normalisation in the
    basis or guess tables changes the count."""
    Z, X = np.array([z]), np.zeros((1, 3))
    for basis, tol in (("MINAO", 1e-10), ("DEF2-TZVP", 2e-3)):
        b = xc.BasisController.from_file(basis, Z, X)
        n = np.vdot(scf.atomic_density_guess(basis, Z), b.overlap())
        assert abs(n - z) < tol, (basis, z, n)

