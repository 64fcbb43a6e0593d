"""Serving: asynchronous builds on resident controllers overlap on the device
(several slots, each its own stream, HIP-graph replay) and give the same F and
E_xc as the same builds run alone, and as the oracle."""
import time

import numpy as np
import pytest

from oracle import oracle as O
from serenity_amd import xc
from serenity_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _systems():
    out = []
    for name, (Z, X), basis, func in (("h2o2/svp/LDA", W.water_cluster(1, 1, 2), "DEF2-SVP", "LDA"),
                                      ("butane/svp/PBE", W.alkane(4), "DEF2-SVP", "PBE"),
                                      ("h2o4/tzvp/PBE", W.water_cluster(2, 2, 1), "DEF2-TZVP", "PBE"),
                                      ("hexane/svp/PBE", W.alkane(6), "DEF2-SVP", "PBE")):
        b = xc.BasisController.from_file(basis, Z, X)
        g = xc.molecular_grid(Z, X, accuracy=3)
        P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2)
        out.append(dict(name=name, basis=b, grid=g, P=P, func=func))
    return out


def test_server_same_alone_and_in_mix(ctx):
    systems = _systems()
    for s in systems:
        s["bf"] = xc.BasisFunctionOnGridController(ctx, s["basis"], s["grid"], 128, 1e-9, 1, keep_fp32=True)
    alone = {}
    for i, s in enumerate(systems):
        for prec in ("FP64", "FP32", "MIXED"):
            pot = xc.FuncPotential(s["bf"], s["func"], prec)
            alone[(i, prec)] = (pot.getMatrix(s["P"]), pot.getEnergy())
        Fo, Eo, _, _ = O.fock_build(s["basis"], s["grid"].points, s["grid"].weights, s["P"], s["func"])
        s["oracle"] = (Fo, Eo)
    srv = xc.Server(ctx, nslots=3, use_graphs=True)
    # three rounds over every (system, precision), interleaved so that builds of different
    # systems overlap; a controller with a build in flight waits for its turn (the mixed builds
    # run on the slots' single stream, with the dense blocks carried from the build alone)
    queue = [(r, i, prec) for r in range(3) for prec in ("FP64", "FP32", "MIXED") for i in range(len(systems))]
    tickets, results = {}, {}
    t0 = time.time()
    while queue or tickets:
        for item in list(queue):
            r, i, prec = item
            t = srv.submit(systems[i]["bf"], systems[i]["func"], prec, systems[i]["P"])
            if t is not None:
                tickets[item] = t
                queue.remove(item)
        for key, t in list(tickets.items()):
            out = srv.poll(t)
            if out is not None:
                results[key] = out
                del tickets[key]
        assert time.time() - t0 < 120, "server builds did not finish"
    st = srv.stats()
    say("server stats", st)
    assert st["submitted"] == len(results) == 3 * 3 * len(systems)
    assert st["graph_launches"] == st["submitted"]
    for (r, i, prec), (F, E) in results.items():
        Fa, Ea = alone[(i, prec)]
        tol = 1e-12 if prec == "FP64" else 1e-10
        dF, dE = np.abs(F - Fa).max(), abs(E - Ea)
        assert dF <= tol and dE <= tol, (systems[i]["name"], prec, r, dF, dE)
        if prec == "FP64":
            Fo, Eo = systems[i]["oracle"]
            assert np.abs(F - Fo).max() <= 1e-9 and abs(E - Eo) <= 1e-9, systems[i]["name"]
    srv._release()


def test_server_busy_and_errors(ctx):
    s = _systems()[1]
    bf = xc.BasisFunctionOnGridController(ctx, s["basis"], s["grid"], 128, 1e-9, 1)
    bf2 = xc.BasisFunctionOnGridController(ctx, s["basis"], s["grid"], 128, 1e-9, 1)
    srv = xc.Server(ctx, nslots=1, use_graphs=False)
    t = srv.submit(bf, "PBE", "FP64", s["P"])
    assert t is not None
    assert srv.submit(bf, "PBE", "FP64", s["P"]) is None  # controller in flight
    assert srv.submit(bf2, "PBE", "FP64", s["P"]) is None  # no free slot
    with pytest.raises(xc.XCError, match="in flight"):
        xc.FuncPotential(bf, "PBE").getMatrix(s["P"])
    # every path that writes the controller's work buffers refuses it, before any copy or launch,
    # whether it is the active subsystem or an environment one
    n = s["basis"].nao
    dP = ctx.alloc(8 * n * n)
    ctx.upload(dP, np.asfortranarray(s["P"]))
    for act, env in ((bf, bf2), (bf2, bf)):
        nadd = xc.NAddFuncPotential(act, [env], "PBE")
        with pytest.raises(xc.XCError, match="in flight"):
            nadd.getMatrix(s["P"], [s["P"]])
        with pytest.raises(xc.XCError, match="in flight"):
            nadd.getMatrixUnrestricted(0.5 * s["P"], 0.5 * s["P"], [0.5 * s["P"]], [0.5 * s["P"]])
    with pytest.raises(xc.XCError, match="in flight"):
        bf.update_density(dP)
    dF = ctx.alloc(8 * n * n)
    with pytest.raises(xc.XCError, match="in flight"):
        xc.NAddFuncPotential(bf, [bf2], "PBE").build_cached(dF)
    with pytest.raises(xc.XCError, match="in flight"):
        xc.NAddFuncPotential(bf2, [bf], "PBE").build_device(dP, [dP], dF)
    F, E = srv.wait(t)
    ctx.free(dP)
    ctx.free(dF)
    Fs = xc.FuncPotential(bf2, "PBE").getMatrix(s["P"])
    assert np.abs(F - Fs).max() < 1e-12
    with pytest.raises(xc.XCError, match="ticket"):
        srv.poll(t)
    with pytest.raises(xc.XCError, match="FP32"):
        srv.submit(bf, "PBE", "FP32", s["P"])
    srv._release()


def test_fp32_only_controller(ctx):
    """keep_fp32="only": FP32 builds equal those of a controller holding both tile sets; FP64 paths refuse."""
    s = _systems()[1]
    both = xc.BasisFunctionOnGridController(ctx, s["basis"], s["grid"], 128, 1e-9, 1, keep_fp32=True)
    only = xc.BasisFunctionOnGridController(ctx, s["basis"], s["grid"], 128, 1e-9, 1, keep_fp32="only")
    assert only.stats()[2] * 3 == both.stats()[2]
    Fb = xc.FuncPotential(both, "PBE", "FP32").getMatrix(s["P"])
    Fo = xc.FuncPotential(only, "PBE", "FP32").getMatrix(s["P"])
    assert np.abs(Fb - Fo).max() < 1e-13
    with pytest.raises(xc.XCError, match="FP64"):
        xc.FuncPotential(only, "PBE", "FP64").getMatrix(s["P"])
    with pytest.raises(xc.XCError, match="FP32 tiles only"):
        only.getBlockOnGridData(0)


@pytest.mark.parametrize("grid_type", ["SSF", "BECKE", "VORONOI"])
def test_device_grid_partition_matches_host(ctx, grid_type):
    """The fuzzy-cell partition on the GPU gives the host grid: same points in the same (Hilbert)
    order, weights equal to rounding."""
    for Z, X in (W.water_cluster(2, 2, 2), W.alkane(6)):
        gh, nh = xc.molecular_grid(Z, X, grid_type=grid_type, return_before_cut=True)
        t0 = time.time()
        gd, nd = xc.molecular_grid(Z, X, grid_type=grid_type, return_before_cut=True, ctx=ctx)
        say("device grid", grid_type, len(Z), "atoms", gd.getNGridPoints(), "points %.3f s" % (time.time() - t0))
        assert nh == nd and gh.getNGridPoints() == gd.getNGridPoints()
        assert np.array_equal(gh.points, gd.points)
        assert np.abs(gh.weights - gd.weights).max() <= 1e-14 * np.abs(gh.weights).max()


def test_serving_with_packing_equals_unpacked(ctx):
    """A short burst of the serving mix served with and without packing of small requests:
    every request's E_xc is the same (packed builds run the same kernels on the same items)."""
    from serenity_amd import serving as SV
    stream = SV.make_stream(48, 1.0e6)  # all arrive at once: small requests wait together
    host = SV.HostSystems(ctx=ctx)
    plain = SV.run_stream(ctx, stream, nslots=2, pack_nao=0, host=host)
    packed = SV.run_stream(ctx, stream, nslots=2, pack_nao=SV.PACK_NAO, host=host)
    e0, e1 = plain.pop("_energies"), packed.pop("_energies")
    say("serving burst: %d requests packed in %d launches; p50 %.2f ms packed vs %.2f ms" %
        (packed["packed_requests"], packed["packed_launches"], packed["p50_ms"], plain["p50_ms"]))
    assert packed["packed_requests"] >= 2 and packed["packed_launches"] >= 1
    assert set(e0) == set(e1)
    assert max(abs(e0[k] - e1[k]) for k in e0) < 1e-10


def test_serving_failed_request_is_isolated(ctx):
    """A request that cannot be built (an unknown functional, rejected on the host before any
    launch) fails alone: in a packed group it sends its group back to one-by-one builds, and every
    other request of the burst gets the energy of the clean run."""
    from serenity_amd import serving as SV
    stream = SV.make_stream(24, 1.0e6)
    host = SV.HostSystems(ctx=ctx)
    small = [j for j, (_, key, _) in enumerate(stream) if host.get(key)["basis"].nao <= SV.PACK_NAO]
    bad = small[0]
    broken = list(stream)
    broken[bad] = (stream[bad][0], stream[bad][1], "NOPE")
    clean = SV.run_stream(ctx, stream, nslots=2, pack_nao=SV.PACK_NAO, host=host)
    out = SV.run_stream(ctx, broken, nslots=2, pack_nao=SV.PACK_NAO, host=host)
    say("failed requests: %d (%s)" % (out["failed"], list(out["_failed"].values())[:1]))
    assert out["failed"] == 1 and bad in out["_failed"] and "unknown functional" in out["_failed"][bad]
    e0, e1 = clean["_energies"], out["_energies"]
    for k, e in e1.items():
        assert abs(e - e0[k]) < 1e-10, k
    assert len(e1) >= len(e0) - 1


def test_serving_under_memory_pressure(ctx):
    """A device budget far below the working set: systems are evicted and admitted again inside the
    stream (their grids built on the GPU when first admitted, inside the latency of that request),
    and every request's E_xc equals the same build done alone."""
    from serenity_amd import serving as SV
    stream = SV.make_stream(60, 1.0e6)
    host = SV.HostSystems(ctx=ctx)
    out = SV.run_stream(ctx, stream, nslots=2, budget_gb=0.3, host=host, pack_nao=SV.PACK_NAO)
    energies = out.pop("_energies")
    say("pressure: %d requests, %d controllers built, %d evicted, %d cold (p50 %s ms) / warm p50 %s ms" %
        (out["requests"], out["created"], out["evicted"], out["cold_requests"], out["cold_p50_ms"],
         out["warm_p50_ms"]))
    assert out["failed"] == 0 and out["evicted"] > 0 and out["cold_requests"] > 0
    worst = 0.0
    for (key, func), e in energies.items():
        d = host.get(key)
        bf = xc.BasisFunctionOnGridController(ctx, d["basis"], host.grid(key), 128, 1e-9, 1, keep_fp32="only")
        pot = xc.FuncPotential(bf, func, "FP32")
        pot.getMatrix(d["P"])
        worst = max(worst, abs(pot.getEnergy() - e))
        bf._release()
    say("pressure: max |dE| against the builds alone %.2e over %d (system, functional)" % (worst, len(energies)))
    assert worst < 1e-10
