"""GPU nuclear XC gradient (xc_geom_gradient) and AO second-derivative tiles against the
oracle, and the reference's gradient values through the GPU SCF.  Bar: 1e-9 against the
oracle (FP64 throughout); the reference's own tolerances for its values."""
import numpy as np
import pytest

from serenity_amd import scf, workloads as W, xc
from serenity_amd._lib import XCError
from oracle import oracle as O

from test_gradient import H6_31GS, h2, shifted_water

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("spherical", [True, False])
def test_hessian_tiles_match_oracle(ctx, spherical):
    Z, X = shifted_water()
    b = xc.BasisController.from_file("DEF2-SVP", Z, X, spherical=spherical)
    g = xc.molecular_grid(Z, X, accuracy=2)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 2)
    worst = 0.0
    for blk in range(0, bf.getNBlocks(), max(1, bf.getNBlocks() // 12)):
        d = bf.getBlockOnGridData(blk)
        B = d["functionValues"].shape[0]
        o = O.ao_block(b, g.points[128 * blk:128 * blk + B], thr=1e-9, deriv=2)
        assert np.array_equal(d["negligible"], o["negligible"])
        for c in range(6):
            worst = max(worst, np.abs(d["secondDerivativeValues"][c] - o["hess"][c]).max())
        assert np.abs(d["functionValues"] - o["values"]).max() < 1e-13
    bf._release()
    assert worst < 1e-11


@pytest.mark.parametrize("functional", ["LDA", "PBE"])
def test_gradient_matches_oracle_water(ctx, functional):
    Z, X = shifted_water()
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=3)
    P = W.synthetic_density(b.overlap(), 5, seed=3)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 2)
    pot = xc.FuncPotential(bf, functional, "FP64")
    G = pot.getGeomGradients(P)
    Go, Eo = O.xc_gradient(b, g.points, g.weights, P, functional, b.getAtomIndicesOfBasis(), 3)
    print("%s water: max|dG| %.2e |dE| %.2e" % (functional, np.abs(G - Go).max(), abs(pot.getEnergy() - Eo)))
    assert np.abs(G - Go).max() < 1e-9
    assert abs(pot.getEnergy() - Eo) < 1e-9
    bf._release()


def test_gradient_matches_oracle_water8_pbe(ctx):
    """(H2O)8/def2-SVP (192 AOs, blocks up to 160 significant AOs) on the accuracy-2 grid."""
    Z, X = W.geometry(W.CONFIGS[1]["system"])
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=2)
    P = W.synthetic_density(b.overlap(), 40, seed=7)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 2)
    pot = xc.FuncPotential(bf, "PBE", "FP64")
    G = pot.getGeomGradients(P)
    Go, _ = O.xc_gradient(b, g.points, g.weights, P, "PBE", b.getAtomIndicesOfBasis(), len(Z))
    print("(H2O)8 PBE: max|dG| %.2e max|G| %.3f" % (np.abs(G - Go).max(), np.abs(G).max()))
    assert np.abs(G - Go).max() < 1e-9
    bf._release()


def second_row_molecules():
    """NaF, H2S, HCl and SiH4 about 3.5 A apart (Angstrom geometry, converted to bohr)."""
    Z = np.array([11, 9, 16, 1, 1, 17, 1, 14, 1, 1, 1, 1])
    X = np.array([[0, 0, 0], [0, 0, 1.93], [3.5, 0, 0], [3.5, 1.34, 0], [4.8, 0, 0.4], [0, 3.6, 0], [0, 3.6, 1.27],
                  [-3.5, 0, 0], [-3.5, 1.48, 0], [-4.0, -0.7, 1.2], [-4.0, -0.7, -1.2], [-2.0, -0.5, 0]])
    return Z, X * xc.ANGSTROM_TO_BOHR


def test_gradient_matches_oracle_second_row_tzvp(ctx):
    """Second-row atoms in def2-TZVP (f shells on every heavy atom): AO second-derivative tiles and
    the PBE gradient against the oracle."""
    Z, X = second_row_molecules()
    b = xc.BasisController.from_file("DEF2-TZVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=2)
    P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2, seed=5)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 2)
    worst = 0.0
    for blk in range(0, bf.getNBlocks(), max(1, bf.getNBlocks() // 12)):
        d = bf.getBlockOnGridData(blk)
        B = d["functionValues"].shape[0]
        o = O.ao_block(b, g.points[128 * blk:128 * blk + B], thr=1e-9, deriv=2)
        assert np.array_equal(d["negligible"], o["negligible"])
        for c in range(6):
            worst = max(worst, np.abs(d["secondDerivativeValues"][c] - o["hess"][c]).max())
    assert worst < 1e-10
    pot = xc.FuncPotential(bf, "PBE", "FP64")
    G = pot.getGeomGradients(P)
    Go, Eo = O.xc_gradient(b, g.points, g.weights, P, "PBE", b.getAtomIndicesOfBasis(), len(Z))
    print("NaF/H2S/HCl/SiH4 TZVP PBE: tiles %.2e max|dG| %.2e |dE| %.2e" % (worst, np.abs(G - Go).max(),
                                                                            abs(pot.getEnergy() - Eo)))
    assert np.abs(G - Go).max() < 1e-9
    assert abs(pot.getEnergy() - Eo) < 1e-9
    bf._release()


def test_h2_tzvp_pbe_gradient_pin(ctx):
    Z, X = h2()
    t = scf.ScfTask(ctx, Z, X, basis="DEF2-TZVP", functional="PBE", ri=False, accuracy=7, small_accuracy=7,
                    radial_threshold=0.0, block_ave_threshold=0.0, highest_derivative=2)
    out = t.run("FP64")
    G = t.geom_gradient(out["P_final"])
    t.close()
    ref = np.array([[0.0, 0.0, -0.25023729], [0.0, 0.0, 0.25023729]])
    print("H2/TZVP PBE gradient z: %.9f" % G[0, 2])
    assert np.abs(G - ref).max() < 5e-8


def test_h2_631gs_lda_gradient_pin(ctx):
    Z, X = h2()
    t = scf.ScfTask(ctx, Z, X, basis=H6_31GS, functional="LDA")
    out = t.run("FP64")
    G = t.geom_gradient(out["P_final"])
    t.close()
    ref = np.array([[0.0, 0.0, -2.379533e-01], [0.0, 0.0, 2.379533e-01]])
    print("H2/6-31G* LDA gradient z: %.9f" % G[0, 2])
    assert np.abs(G - ref).max() < 1e-6


@pytest.mark.parametrize("functional", ["LDA", "PBE"])
def test_unrestricted_gradient(ctx, functional):
    """Open-shell spin densities: the GPU gradient against finite differences of the oracle's
    unrestricted E_xc with the basis moved at a fixed grid and densities; a closed shell equals
    the restricted gradient."""
    Z, X = shifted_water()
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=2)
    S = b.overlap()
    Pa = W.synthetic_density(S, 5, seed=3) / 2
    Pb = W.synthetic_density(S, 4, seed=4) / 2
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 0.0, 2)
    pot = xc.FuncPotential(bf, functional, "FP64", 0.0)
    G = pot.getGeomGradients((Pa, Pb))
    h = 1e-4
    fd = np.zeros((3, 3))
    for a in range(3):
        for c in range(3):
            E = []
            for sgn in (1, -1):
                Xs = X.copy()
                Xs[a, c] += sgn * h
                bs = xc.BasisController.from_file("DEF2-SVP", Z, Xs)
                E.append(O.fock_build_pol(bs, g.points, g.weights, Pa, Pb, functional, thr=0.0, block_ave_thr=0.0)[2])
            fd[a, c] = (E[0] - E[1]) / (2 * h)
    print("%s unrestricted gradient: max|G - FD| %.2e" % (functional, np.abs(G - fd).max()))
    assert np.abs(G - fd).max() < 5e-8
    P = W.synthetic_density(S, 5, seed=3)
    Gr = pot.getGeomGradients(P)
    Gu = pot.getGeomGradients((P / 2, P / 2))
    assert np.abs(Gr - Gu).max() < 1e-12
    bf._release()


def test_gradient_errors(ctx):
    Z, X = shifted_water()
    b = xc.BasisController.from_file("DEF2-SVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=1)
    P = W.synthetic_density(b.overlap(), 5, seed=3)
    bf1 = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1)
    with pytest.raises(XCError, match="second derivatives"):
        xc.FuncPotential(bf1, "PBE", "FP64").getGeomGradients(P)
    G = xc.FuncPotential(bf1, "LDA", "FP64").getGeomGradients(P)  # LDA needs first derivatives only
    assert np.isfinite(G).all()
    bad = b.getAtomIndicesOfBasis().copy()
    bad[0] = 7
    with pytest.raises(XCError, match="atom index"):
        xc.FuncPotential(bf1, "LDA", "FP64").getGeomGradients(P, ao_atom=bad, natoms=3)
    bf1._release()
