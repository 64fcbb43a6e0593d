"""Bit-for-bit repeatable Fock matrices (xc_ctx_set_deterministic): the row reduce adds each
element's terms as exact integer parts, so repeated builds, and packed against alone builds, give
identical F; the values agree with the ordinary FP64 reduce to rounding."""
import numpy as np
import pytest

from serenity_amd import xc
from serenity_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _system(Z, X, basis, seed=0):
    b = xc.BasisController.from_file(basis, Z, X)
    g = xc.molecular_grid(Z, X)
    P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2, seed=seed)
    return b, g, P


@pytest.fixture(scope="module")
def dctx():
    c = xc.Context(0)
    c.set_deterministic(True)
    yield c
    c.close()


def test_repeated_builds_are_bitwise_equal(ctx, dctx):
    Z, X = W.alkane(20)  # C20H42/def2-SVP: 490 AOs, ~1900 blocks
    b, g, P = _system(Z, X, "DEF2-SVP")
    bf_d = xc.BasisFunctionOnGridController(dctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    bf_n = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    for prec in ("FP64", "FP32", "MIXED"):
        pot = xc.FuncPotential(bf_d, "PBE", prec)
        F0 = pot.getMatrix(P).copy()
        E0 = pot.getEnergy()
        for _ in range(3):
            F = pot.getMatrix(P)
            assert np.array_equal(F, F0), prec
            assert pot.getEnergy() == E0
        assert np.array_equal(F0, F0.T)
        ref = xc.FuncPotential(bf_n, "PBE", prec)
        Fr = ref.getMatrix(P)
        dF = np.abs(Fr - F0).max()
        say("%s: deterministic F vs the FP64-addition reduce max|dF| %.2e, |dE| %.2e" % (prec, dF,
                                                                                         abs(ref.getEnergy() - E0)))
        assert dF < 1e-12
        assert abs(ref.getEnergy() - E0) < 1e-12
    with pytest.raises(xc.XCError, match="before any basis"):
        dctx.set_deterministic(False)
    bf_d._release()
    bf_n._release()


def test_packed_equals_alone_bitwise(dctx):
    systems = []
    for k, ((Z, X), basis) in enumerate(((W.water_chain(3), "DEF2-TZVP"), (W.alkane(6), "DEF2-SVP"))):
        b, g, P = _system(Z, X, basis, seed=k + 1)
        bf = xc.BasisFunctionOnGridController(dctx, b, g, 128, 1e-9, 1, keep_fp32=True)
        n = b.nao
        dP, dF = dctx.alloc(8 * n * n), dctx.alloc(8 * n * n)
        dctx.upload(dP, np.asfortranarray(P))
        systems.append(dict(bf=bf, n=n, dP=dP, dF=dF))
    for prec in ("FP64", "FP32"):
        alone = []
        for s in systems:
            xc.FuncPotential(s["bf"], "PBE", prec).build_device(s["dP"], s["dF"])
            F = np.zeros((s["n"], s["n"]), order="F")
            dctx.download(F, s["dF"])
            alone.append(F)
        for s in systems:
            dctx.upload(s["dF"], np.full((s["n"], s["n"]), 7.0))
        xc.fock_build_packed([xc.FuncPotential(s["bf"], "PBE", prec) for s in systems],
                             [s["dP"] for s in systems], [s["dF"] for s in systems])
        for s, Fa in zip(systems, alone):
            F = np.zeros((s["n"], s["n"]), order="F")
            dctx.download(F, s["dF"])
            assert np.array_equal(F, Fa), prec
        say("%s: packed F equals alone F bit for bit (2 systems)" % prec)
    for s in systems:
        dctx.free(s["dP"])
        dctx.free(s["dF"])
        s["bf"]._release()
