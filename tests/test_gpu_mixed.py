"""Mixed precision (XC_PREC_MIXED): FP32-MFMA contractions with the density of the high-density
blocks contracted again in FP64.  The FP32 rounding of the AO values and of phi P near the heavy
nuclei biases E_xc coherently (the same core values on every atom of an element), so the pure FP32
error grows with the number of heavy atoms; the refinement removes that part.  The bars are the
ones stated for the mixed form: SCF energy within 1e-6 Eh of FP64 before the FP64 finish and
within 1e-8 Eh after it, on (H2O)64/def2-TZVP/PBE."""
import numpy as np
import pytest

from serenity_amd import scf as SCF, xc
from serenity_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


def test_mixed_build_removes_the_fp32_bias(ctx):
    """(H2O)8/def2-TZVP/PBE on the ATOM_SCF guess density (real atomic cores): the mixed build's
    E_xc is far closer to FP64 than the FP32 build's, its F within the FP32 level."""
    Z, X = W.water_cluster(2, 2, 2)
    b = xc.BasisController.from_file("DEF2-TZVP", Z, X)
    g = xc.molecular_grid(Z, X, accuracy=4)
    P = SCF.atomic_density_guess("DEF2-TZVP", Z)
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    out = {}
    for prec in ("FP64", "FP32", "MIXED"):
        pot = xc.FuncPotential(bf, "PBE", prec)
        F = pot.getMatrix(P)
        out[prec] = (pot.getEnergy(), F)
    e64, f64 = out["FP64"]
    d32, dmx = abs(out["FP32"][0] - e64), abs(out["MIXED"][0] - e64)
    f32, fmx = np.abs(out["FP32"][1] - f64).max(), np.abs(out["MIXED"][1] - f64).max()
    say("(H2O)8/TZVP guess density: |dE| FP32 %.2e mixed %.2e; max|dF| FP32 %.2e mixed %.2e" % (d32, dmx, f32, fmx))
    assert dmx < 1e-7 and dmx < d32
    assert fmx < 1e-5
    # the same mixed build again: the blocks found dense by the first build are now skipped by the
    # FP32 pass and contracted in FP64 directly; the result is the first build's
    pot = xc.FuncPotential(bf, "PBE", "MIXED")
    F2 = pot.getMatrix(P)
    assert abs(pot.getEnergy() - out["MIXED"][0]) < 1e-12
    assert np.abs(F2 - out["MIXED"][1]).max() < 1e-12
    # a different density next (half the charge: fewer dense blocks): the carried-over blocks are
    # contracted in FP64 whatever their density, so the build stays within the mixed bar of FP64
    # for that density; back to P, the dense blocks are found again and the first result returns
    e64h = xc.FuncPotential(bf, "PBE", "FP64")
    e64h.getMatrix(0.5 * P)
    pot = xc.FuncPotential(bf, "PBE", "MIXED")
    pot.getMatrix(0.5 * P)
    dh = abs(pot.getEnergy() - e64h.getEnergy())
    pot = xc.FuncPotential(bf, "PBE", "MIXED")
    F3 = pot.getMatrix(P)
    say("mixed builds carrying dense blocks over: P/2 |dE| vs FP64 %.2e; back to P: |dE| vs the first %.2e" %
        (dh, abs(pot.getEnergy() - out["MIXED"][0])))
    assert dh < 1e-7
    assert abs(pot.getEnergy() - out["MIXED"][0]) < 1e-12
    assert np.abs(F3 - out["MIXED"][1]).max() < 1e-12
    only32 = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32="only")
    with pytest.raises(xc.XCError, match="both tile sets"):
        xc.FuncPotential(only32, "PBE", "MIXED").getMatrix(P)
    with pytest.raises(xc.XCError, match="restricted single-system"):
        xc.FuncPotential(bf, "PBE", "MIXED").getMatrixUnrestricted(P / 2, P / 2)
    only32._release()
    bf._release()


def test_config3_scf_mixed_meets_the_bars(ctx):
    """(H2O)64/def2-TZVP/PBE, RI-J on the GPU J engine, ATOM_SCF guess: the mixed SCF's energy
    before the FP64 finish is within 1e-6 Eh of the FP64 SCF, and within 1e-8 Eh after it."""
    cfg = W.CONFIGS[3]
    Z, X = W.geometry(cfg["system"])
    task = SCF.ScfTask(ctx, Z, X, basis=cfg["basis"], functional=cfg["functional"], gpu_coulomb=True)
    r64 = task.run("FP64")
    say("config 3 FP64 SCF: E %.10f, %d cycles, %.1f s" % (r64["energy"], r64["cycles"], r64["wall_s"]))
    rmx = task.run("MIXED")
    before = abs(rmx["energy_before_fp64_finish"] - r64["energy"])
    after = abs(rmx["energy"] - r64["energy"])
    say("config 3 mixed SCF: %d + %d cycles, %.1f s; |dE| before the FP64 finish %.2e, after %.2e" %
        (rmx["cycles_fp32"], rmx["cycles_fp64_finish"], rmx["wall_s"], before, after))
    task.close()
    assert r64["converged"] and rmx["converged"]
    assert before < 1e-6
    assert after < 1e-8
