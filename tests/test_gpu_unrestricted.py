"""GPU unrestricted XC build (xc_fock_build_unrestricted) against the oracle (FP64: 1e-9),
the reference's stored triplet O2 Fock matrices, and the restricted build for a closed shell."""
import numpy as np
import pytest

from serenity_amd import scf, workloads as W, xc
from serenity_amd._lib import XCError
from oracle import oracle as O

from test_unrestricted import o2_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


def test_o2_triplet_matches_oracle_and_stored_fock(ctx):
    Z, X, b, Pa, Pb, Fa_ref, Fb_ref, _ = o2_system()
    g = xc.molecular_grid(Z, X, accuracy=7)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1)
    pot = xc.FuncPotential(bf, "PBE", "FP64")
    Fa, Fb = pot.getMatrixUnrestricted(Pa, Pb)
    Oa, Ob, Eo = O.fock_build_pol(b, g.points, g.weights, Pa, Pb, "PBE")
    print("O2 triplet: max|dF| %.2e %.2e |dE| %.2e" % (np.abs(Fa - Oa).max(), np.abs(Fb - Ob).max(),
                                                      abs(pot.getEnergy() - Eo)))
    assert np.abs(Fa - Oa).max() < 1e-9 and np.abs(Fb - Ob).max() < 1e-9
    assert abs(pot.getEnergy() - Eo) < 1e-9
    ints = scf.ScfIntegrals(b, Z, X)
    J = ints.coulomb(Pa + Pb)
    assert np.abs(ints.h + J + Fa - Fa_ref).max() < 1e-8
    assert np.abs(ints.h + J + Fb - Fb_ref).max() < 1e-8
    bf._release()


@pytest.mark.parametrize("functional", ["PBE", "LDA", "BP86"])
def test_open_shell_alkane_matches_oracle(ctx, functional):
    """C20H42/def2-SVP with 81 alpha and 80 beta synthetic orbitals: FP64 vs oracle, FP32 vs FP64."""
    d = W.build_system(2)
    b, g, S = d["basis"], d["grid"], d["S"]
    Pa = W.synthetic_density(S, 81, seed=11) / 2.0
    Pb = W.synthetic_density(S, 80, seed=12) / 2.0
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    p64 = xc.FuncPotential(bf, functional, "FP64")
    Fa, Fb = p64.getMatrixUnrestricted(Pa, Pb)
    Oa, Ob, Eo = O.fock_build_pol(b, g.points, g.weights, Pa, Pb, functional)
    dF = max(np.abs(Fa - Oa).max(), np.abs(Fb - Ob).max())
    p32 = xc.FuncPotential(bf, functional, "FP32")
    Ga, Gb = p32.getMatrixUnrestricted(Pa, Pb)
    dF32 = max(np.abs(Ga - Fa).max(), np.abs(Gb - Fb).max())
    print("C20H42 open shell %s: FP64 max|dF| %.2e |dE| %.2e; FP32 vs FP64 max|dF| %.2e |dE| %.2e" %
          (functional, dF, abs(p64.getEnergy() - Eo), dF32, abs(p32.getEnergy() - p64.getEnergy())))
    assert dF < 1e-9 and abs(p64.getEnergy() - Eo) < 1e-9
    assert dF32 < 1e-5 and abs(p32.getEnergy() - p64.getEnergy()) < 1e-4
    bf._release()


def test_open_shell_second_row_matches_oracle(ctx):
    """PH2 and ClO radicals 4 A apart / def2-TZVP / PBE (f shells on P, Cl and O), 22 alpha and 20 beta
    synthetic orbitals: FP64 vs oracle."""
    Z = np.array([15, 1, 1, 17, 8])
    X = np.array([[0, 0, 0], [1.42, 0, 0], [-0.3, 1.38, 0], [4.0, 0, 0], [4.0, 1.57, 0]]) * xc.ANGSTROM_TO_BOHR
    b = xc.BasisController.from_file("DEF2-TZVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=3)
    S = b.overlap()
    Pa = W.synthetic_density(S, 22, seed=21) / 2.0
    Pb = W.synthetic_density(S, 20, seed=22) / 2.0
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1)
    pot = xc.FuncPotential(bf, "PBE", "FP64")
    Fa, Fb = pot.getMatrixUnrestricted(Pa, Pb)
    Oa, Ob, Eo = O.fock_build_pol(b, g.points, g.weights, Pa, Pb, "PBE")
    dF = max(np.abs(Fa - Oa).max(), np.abs(Fb - Ob).max())
    print("PH2 + ClO open shell PBE: max|dF| %.2e |dE| %.2e" % (dF, abs(pot.getEnergy() - Eo)))
    assert dF < 1e-9 and abs(pot.getEnergy() - Eo) < 1e-9
    bf._release()


def test_closed_shell_equals_restricted(ctx):
    d = W.build_system(1)
    bf = xc.BasisFunctionOnGridController(ctx, d["basis"], d["grid"], 128, 1e-9, 1)
    pot = xc.FuncPotential(bf, "PBE", "FP64")
    F = pot.getMatrix(d["P"])
    E = pot.getEnergy()
    Fa, Fb = pot.getMatrixUnrestricted(d["P"] / 2, d["P"] / 2)
    assert np.abs(Fa - F).max() < 1e-12 and np.abs(Fb - F).max() < 1e-12
    assert abs(pot.getEnergy() - E) < 1e-11
    with pytest.raises(XCError, match="keep_fp32"):
        xc.FuncPotential(bf, "PBE", "FP32").getMatrixUnrestricted(d["P"] / 2, d["P"] / 2)
    bf._release()
