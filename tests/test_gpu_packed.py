"""Packed builds (xc_fock_build_packed): the XC matrices and energies of several
controllers computed with one launch per stage and work class equal the builds of
each controller alone (the same kernels on the same items; only the order of the
FP64 atomic additions differs) and, in FP64, the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from serenity_amd import xc
from serenity_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


@pytest.fixture(scope="module")
def ctx():
    c = xc.Context(0)
    yield c
    c.close()


def _systems(ctx, which=None):
    out = []
    specs = ((W.water_chain(1), "DEF2-SVP"), (W.alkane(1), "DEF2-TZVP"), (W.water_chain(3), "DEF2-TZVP"),
             (W.alkane(3), "DEF2-SVP"), (W.water_cluster(1, 1, 2), "DEF2-SVP"), (W.alkane(6), "DEF2-SVP"))
    for k, ((Z, X), basis) in enumerate(specs):
        if which is not None and k not in which:
            continue
        b = xc.BasisController.from_file(basis, Z, X)
        g = xc.molecular_grid(Z, X, accuracy=3)
        P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2, seed=len(Z))
        bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
        n = b.nao
        dP, dF = ctx.alloc(8 * n * n), ctx.alloc(8 * n * n)
        ctx.upload(dP, np.asfortranarray(P))
        out.append(dict(basis=b, grid=g, P=P, bf=bf, n=n, dP=dP, dF=dF))
    return out


def _release(ctx, systems):
    for s in systems:
        ctx.free(s["dP"])
        ctx.free(s["dF"])
        s["bf"]._release()


@pytest.mark.parametrize("functional", ["LDA", "PBE"])
def test_packed_equals_alone_and_oracle(ctx, functional):
    systems = _systems(ctx)
    worst_oracle = 0.0
    for prec in ("FP64", "FP32"):
        alone = []
        for s in systems:
            pot = xc.FuncPotential(s["bf"], functional, prec)
            pot.build_device(s["dP"], s["dF"])
            F = np.zeros((s["n"], s["n"]), order="F")
            ctx.download(F, s["dF"])
            alone.append((F, pot.getEnergy()))
        pots = [xc.FuncPotential(s["bf"], functional, prec) for s in systems]
        # F buffers are overwritten by the packed build
        for s in systems:
            ctx.upload(s["dF"], np.full((s["n"], s["n"]), 7.0))
        E = xc.fock_build_packed(pots, [s["dP"] for s in systems], [s["dF"] for s in systems])
        dE = dF = 0.0
        for s, (Fa, Ea), e, p in zip(systems, alone, E, pots):
            F = np.zeros((s["n"], s["n"]), order="F")
            ctx.download(F, s["dF"])
            dE = max(dE, abs(e - Ea))
            dF = max(dF, np.abs(F - Fa).max())
            assert p.getEnergy() == e
            if prec == "FP64":
                Fo, Eo, _, _ = O.fock_build(s["basis"], s["grid"].points, s["grid"].weights, s["P"], functional)
                worst_oracle = max(worst_oracle, abs(e - Eo), np.abs(F - Fo).max())
        say("%s %s packed (%d systems) vs alone: |dE| %.2e max|dF| %.2e" % (functional, prec, len(systems), dE, dF))
        assert dE < 1e-12 and dF < 1e-12
    say("%s packed FP64 vs oracle: %.2e" % (functional, worst_oracle))
    assert worst_oracle < 1e-9
    _release(ctx, systems)


def test_packed_single_system_and_errors(ctx):
    systems = _systems(ctx, (0, 1))
    s = systems[0]
    pot = xc.FuncPotential(s["bf"], "PBE", "FP64")
    e_alone = pot.build_device(s["dP"], s["dF"])
    e = xc.fock_build_packed([xc.FuncPotential(s["bf"], "PBE", "FP64")], [s["dP"]], [s["dF"]])
    assert abs(e[0] - e_alone) < 1e-13
    with pytest.raises(xc.XCError, match="appears twice"):
        xc.fock_build_packed([pot, pot], [s["dP"], s["dP"]], [s["dF"], s["dF"]])
    with pytest.raises(xc.XCError, match="differ"):
        xc.fock_build_packed([pot, xc.FuncPotential(systems[1]["bf"], "LDA", "FP64")],
                             [s["dP"], systems[1]["dP"]], [s["dF"], systems[1]["dF"]])
    _release(ctx, systems)


def test_packed_repeated_composition_replays_graph(ctx):
    """A composition built again is captured once into a HIP graph and then replayed: every replay,
    with other builds, new density matrices and (from the third build on) new device buffers for
    P and F in between, equals the builds alone."""
    systems = _systems(ctx, (0, 1, 2, 3))
    pots = [xc.FuncPotential(s["bf"], "PBE", "FP32") for s in systems]
    other = _systems(ctx, (4,))[0]
    for rep in range(4):
        if rep >= 2:
            # the replayed graph reads the matrices through the per-call system table: move them
            # (allocated before the old ones are freed, so that the addresses differ)
            for s in systems:
                old = (s["dP"], s["dF"])
                s["dP"], s["dF"] = ctx.alloc(8 * s["n"] ** 2), ctx.alloc(8 * s["n"] ** 2)
                assert s["dP"].value not in [o.value for o in old] and s["dF"].value not in [o.value for o in old]
                ctx.free(old[0])
                ctx.free(old[1])
        alone = []
        for i, s in enumerate(systems):
            Pr = W.synthetic_density(s["basis"].overlap(), max(1, int(round(np.vdot(s["P"], s["basis"].overlap()))) // 2),
                                     seed=100 * rep + i)
            ctx.upload(s["dP"], np.asfortranarray(Pr))
            p = xc.FuncPotential(s["bf"], "PBE", "FP32")
            p.build_device(s["dP"], s["dF"])
            F = np.zeros((s["n"], s["n"]), order="F")
            ctx.download(F, s["dF"])
            alone.append((F, p.getEnergy()))
        xc.FuncPotential(other["bf"], "LDA", "FP64").build_device(other["dP"], other["dF"])  # eager work between
        E = xc.fock_build_packed(pots, [s["dP"] for s in systems], [s["dF"] for s in systems])
        worst = 0.0
        for s, (Fa, Ea), e in zip(systems, alone, E):
            F = np.zeros((s["n"], s["n"]), order="F")
            ctx.download(F, s["dF"])
            worst = max(worst, np.abs(F - Fa).max(), abs(e - Ea))
        say("packed build %d of one composition (%s): max diff %.2e" % (rep, "eager" if rep == 0 else
                                                                         "captured" if rep == 1 else "replay", worst))
        assert worst < 1e-12
    _release(ctx, systems)
    ctx.free(other["dP"])
    ctx.free(other["dF"])
    other["bf"]._release()
