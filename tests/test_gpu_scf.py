"""Full SCF with the GPU XC build (serenity_amd.scf.ScfTask) against the reference's energies
and against the same SCF driven by the oracle's CPU XC build.

Bars: FP64 GPU SCF total energy within 1e-8 Eh of the oracle-driven SCF; the mixed form
(FP32-MFMA iterations, then an FP64 finish from the FP32-converged density) within 1e-6
Eh before the finish and 1e-8 Eh after it.
"""
import numpy as np
import pytest

from serenity_amd import scf, workloads as W, xc
from oracle import oracle as O

from test_integrals import water_geometry
from test_scf import h2, oracle_xc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


def oracle_scf(task):
    """The same SCF (integrals, grids, protocol) with the oracle's XC build."""
    r = scf.perform(task.ints, task.nocc, oracle_xc(task.basis, task.grid_small, task.functional), P0=task.P_guess)
    _, e_xc, _, _ = O.fock_build(task.basis, task.grid_default.points, task.grid_default.weights, r["P"],
                                 task.functional)
    return scf.total_energy(task.ints, r["P"], e_xc), r


def test_h2_sto6g_lda_no_ri(ctx):
    Z, X = h2()
    t = scf.ScfTask(ctx, Z, X, basis="STO-6G", functional="LDA", ri=False)
    out = t.run("FP64")
    t.close()
    assert out["converged"]
    assert abs(out["energy"] - (-1.1298876147476988)) < 1e-7


@pytest.mark.parametrize("functional,ref", [("PBE", -75.22180876683022), ("PBE_LIBXC", -75.2218044178)])
def test_water_sto3g_pbe_fp64_and_mixed(ctx, functional, ref):
    Z, X = water_geometry()
    t = scf.ScfTask(ctx, Z, X, basis="STO-3G", functional=functional)
    e_oracle, _ = oracle_scf(t)
    fp64 = t.run("FP64")
    mixed = t.run("MIXED")
    t.close()
    print("water %s: oracle %.10f fp64 %.10f (%d cycles) mixed before %.10f after %.10f (%d+%d cycles)" %
          (functional, e_oracle, fp64["energy"], fp64["cycles"], mixed["energy_before_fp64_finish"],
           mixed["energy"], mixed["cycles_fp32"], mixed["cycles_fp64_finish"]), flush=True)
    assert fp64["converged"] and mixed["converged"]
    assert abs(fp64["energy"] - ref) < 5e-8
    assert abs(fp64["energy"] - e_oracle) < 1e-8
    assert abs(mixed["energy_before_fp64_finish"] - fp64["energy"]) < 1e-6
    assert abs(mixed["energy"] - fp64["energy"]) < 1e-8


def test_o2_triplet_unrestricted_pbe(ctx):
    """Unrestricted SCF with the GPU unrestricted build: the reference's stored energy total."""
    import common
    from test_unrestricted import o2_system
    Z, X, b, _, _, _, _, E = o2_system()
    t = scf.ScfTask(ctx, Z, X, basis=common.GOLDEN + "/O-6-31G", functional="PBE", ri=False, accuracy=7,
                    small_accuracy=7, spin=2)
    fp64 = t.run("FP64")
    fp32 = t.run("FP32")
    t.close()
    ref = E[3] + E[5] + E[6] + E[202]
    print("O2 triplet UKS: FP64 %.10f (%d cycles) FP32 %.10f ref %.10f" % (fp64["energy"], fp64["cycles"],
                                                                          fp32["energy"], ref))
    assert fp64["converged"] and abs(fp64["energy"] - ref) < 1e-8
    assert abs(fp32["energy"] - fp64["energy"]) < 1e-5


def test_water8_svp_lda_full_scf(ctx):
    """(H2O)8/def2-SVP/LDA with RI-J: the full SCF of the first benchmark configuration."""
    Z, X = W.geometry(W.CONFIGS[1]["system"])
    t = scf.ScfTask(ctx, Z, X, basis="DEF2-SVP", functional="LDA")
    assert t.basis.nao == 192
    e_oracle, r_oracle = oracle_scf(t)
    fp64 = t.run("FP64")
    mixed = t.run("MIXED")
    # one more XC build on the converged density, on the default grid, against the oracle
    pot = t._potential("default", "FP64")
    F = pot.getMatrix(fp64["P"])
    Fo, Eo, _, _ = O.fock_build(t.basis, t.grid_default.points, t.grid_default.weights, fp64["P"], "LDA")
    t.close()
    print("(H2O)8: oracle %.10f (%d cycles) fp64 %.10f (%d) mixed before %.10f after %.10f (%d+%d)" %
          (e_oracle, r_oracle["cycles"], fp64["energy"], fp64["cycles"], mixed["energy_before_fp64_finish"],
           mixed["energy"], mixed["cycles_fp32"], mixed["cycles_fp64_finish"]), flush=True)
    assert fp64["converged"] and mixed["converged"]
    assert abs(fp64["energy"] - e_oracle) < 1e-8
    assert abs(mixed["energy_before_fp64_finish"] - fp64["energy"]) < 1e-6
    assert abs(mixed["energy"] - fp64["energy"]) < 1e-8
    assert abs(pot.getEnergy() - Eo) < 1e-9
    assert np.abs(F - Fo).max() < 1e-9
