"""GPU density-fitted Coulomb matrix (J engine, xc_coulomb_*) against the host's analytic
3-centre integrals: b = B p and J = B^T M^-1 B p to 1e-10 relative, for spherical and
Cartesian basis sets with d and f shells and auxiliary shells up to g, and the SCF energy
with either J."""
import numpy as np
import pytest

from serenity_amd import scf, workloads as W, xc

from test_integrals import water_geometry

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("basis,spherical,nwater", [("DEF2-TZVP", True, 1), ("DEF2-SVP", False, 1),
                                                    ("DEF2-TZVP", True, 2)])
def test_j_engine_matches_host_integrals(ctx, basis, spherical, nwater):
    if nwater == 1:
        Z, X = water_geometry()
    else:
        Z, X = W.water_cluster(2, 1, 1)
    b = xc.BasisController.from_file(basis, Z, X, spherical=spherical)
    a = xc.BasisController.from_file("RI_J_WEIGEND", Z, X, spherical=spherical)
    P = W.synthetic_density(b.overlap(), int(Z.sum()) // 2, seed=5)
    B3, M = scf.ri_j_integrals(b, a, 0.0)
    b_ref = B3 @ scf.pack_density(P)
    J_ref = scf.unpack_symmetric(B3.T @ np.linalg.solve(M, b_ref), b.nao)
    jc = xc.DensityFittedCoulomb(ctx, b, a, threshold=0.0)
    assert np.abs(jc.metric - M).max() == 0.0
    bg = jc.rhs(P)
    Jg = jc.coulomb(P)
    eb = np.abs(bg - b_ref).max() / np.abs(b_ref).max()
    eJ = np.abs(Jg - J_ref).max() / np.abs(J_ref).max()
    print("%s %s (H2O)%d: %d pairs, rel |db| %.2e |dJ| %.2e" % (basis, "sph" if spherical else "cart", nwater,
                                                                 jc.npairs, eb, eJ))
    assert eb < 1e-11 and eJ < 1e-11
    jc._release()


def test_scf_with_gpu_coulomb_equals_host(ctx):
    Z, X = water_geometry()
    t_host = scf.ScfTask(ctx, Z, X, basis="DEF2-SVP", functional="PBE")
    t_gpu = scf.ScfTask(ctx, Z, X, basis="DEF2-SVP", functional="PBE", gpu_coulomb=True)
    e_host = t_host.run("FP64")["energy"]
    e_gpu = t_gpu.run("FP64")["energy"]
    t_host.close()
    t_gpu.close()
    print("water/SVP/PBE SCF: host J %.12f GPU J %.12f" % (e_host, e_gpu))
    assert abs(e_host - e_gpu) < 1e-10
