"""Serving stream of XC Fock builds on one MI355X.

Request mix (seeded): half water clusters (H2O)_n with n uniform in 1..48, half
linear alkanes CkH2k+2 with k uniform in 1..30; def2-SVP or def2-TZVP (50/50);
LDA or PBE (50/50); Poisson arrivals at a fixed rate.  A system is a (molecule,
basis) pair: its grid, basis, density matrix (P = 2 C C^T, C = S^-1/2 Q, seeded
per system) and FP32 AO tiles are built once and stay resident in HBM, admitted
against a device-memory budget by their measured footprint (least recently used
systems without a build in flight are evicted first).  A request is one XC Fock
build of its system with FP32-MFMA contractions.

Continuous batching: arrived requests are dispatched in arrival order whenever a
server slot is free and their system has no build in flight; builds of
different systems overlap on the device (one stream per slot, HIP-graph replay
of each system's build).  With packing on, waiting requests of small systems
(idle, one functional) are first gathered into one packed build: one launch
per stage and work class for all of them.  Latency = completion - arrival.

    python -m serenity_amd.serving [--requests 2000] [--qps 400] [--slots 4]
"""
import argparse
import collections
import json
import sys
import time

import numpy as np

from . import workloads as W
from . import xc

BASES = ("DEF2-SVP", "DEF2-TZVP")
PACK_NAO = 250  # largest system (basis functions) whose waiting requests are packed
FUNCTIONALS = ("LDA", "PBE")


def make_stream(n_requests, qps, seed=W.SEED):
    """[(arrival_s, (kind, size, basis), functional)] with exponential inter-arrival times."""
    rng = np.random.default_rng(seed + 5)
    t, out = 0.0, []
    for _ in range(n_requests):
        t += rng.exponential(1.0 / qps)
        if rng.random() < 0.5:
            key = ("water", int(rng.integers(1, 49)))
        else:
            key = ("alkane", int(rng.integers(1, 31)))
        basis = BASES[int(rng.integers(0, 2))]
        func = FUNCTIONALS[int(rng.integers(0, 2))]
        out.append((t, key + (basis,), func))
    return out


class HostSystems:
    """Host-side data of every system: geometry, basis and density matrix (what a request brings),
    built once.  The molecular grid is built on the GPU (xc_molecular_grid_device) the first time a
    system is admitted to the device, inside that request's latency, and kept for later admissions."""

    def __init__(self, seed=W.SEED, ctx=None):
        self.seed = seed
        self.ctx = ctx
        self.grids = {}
        self.data = {}

    def get(self, key):
        if key not in self.data:
            kind, size, basis = key
            Z, X = W.water_chain(size, seed=self.seed) if kind == "water" else W.alkane(size, seed=self.seed)
            b = xc.BasisController.from_file(basis, Z, X)
            sys_seed = self.seed + 100000 * (kind == "alkane") + 100 * size + BASES.index(basis)
            P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2, seed=sys_seed)
            self.data[key] = dict(basis=b, P=P, Z=Z, X=X)
        return self.data[key]

    def grid(self, key):
        """The molecular grid of a system (one per molecule, shared by its bases)."""
        kind, size, _ = key
        if (kind, size) not in self.grids:
            d = self.get(key)
            self.grids[(kind, size)] = xc.molecular_grid(d["Z"], d["X"], ctx=self.ctx)
        return self.grids[(kind, size)]


class ResidentSystems:
    """Device-resident controllers (FP32 tiles) admitted against a byte budget, LRU eviction."""

    def __init__(self, ctx, host, budget_bytes, tiles="only"):
        self.ctx, self.host, self.budget, self.tiles = ctx, host, budget_bytes, tiles
        self.lru = collections.OrderedDict()  # key -> (bfogc, footprint bytes)
        self.bytes = 0
        self.created = self.evicted = 0
        self.last_created = False  # whether the last get() had to build the controller

    def fits(self, key):
        """Whether a controller for key could be admitted without evicting another."""
        return key in self.lru or self.bytes < self.budget

    def get(self, key, busy):
        self.last_created = False
        if key in self.lru:
            self.lru.move_to_end(key)
            return self.lru[key][0]
        self.last_created = True
        d = self.host.get(key)
        g = self.host.grid(key)
        before = self.ctx.memory_in_use()
        bf = xc.BasisFunctionOnGridController(self.ctx, d["basis"], g, 128, 1e-9, 1, keep_fp32=self.tiles)
        foot = self.ctx.memory_in_use() - before
        self.lru[key] = (bf, foot)
        self.bytes += foot
        self.created += 1
        while self.bytes > self.budget:
            victim = next((k for k in self.lru if k != key and k not in busy), None)
            if victim is None:
                break
            vbf, vfoot = self.lru.pop(victim)
            vbf._release()
            self.bytes -= vfoot
            self.evicted += 1
        return bf


def run_stream(ctx, stream, nslots=4, budget_gb=250.0, prewarm=True, precision="FP32", log=None, pack_nao=0,
               host=None, use_graphs=True):
    """Serve the stream in real time; returns latency percentiles, throughput and counts.

    pack_nao > 0: pending requests of idle systems with at most that many basis functions are
    gathered (per functional) and built together with one packed launch per stage
    (xc.fock_build_packed) when two or more are waiting; the rest go to the server slots."""
    say = log or (lambda *a: None)
    host = host or HostSystems(ctx=ctx)
    res = ResidentSystems(ctx, host, budget_gb * 1e9, tiles="only" if precision == "FP32" else False)
    keys = sorted({k for _, k, _ in stream})
    t0 = time.perf_counter()
    for k in keys:
        host.get(k)
    t_host = time.perf_counter() - t0
    say("host setup of %d systems: %.1f s" % (len(keys), t_host))
    t0 = time.perf_counter()
    if prewarm:
        # admit systems in key order while they fit the budget (nothing is evicted before the stream)
        for k in keys:
            if not res.fits(k):
                break
            res.get(k, ())
        ctx.synchronize()
    t_dev = time.perf_counter() - t0
    say("resident: %d systems, %.1f GB of device memory, setup %.1f s" % (len(res.lru), res.bytes / 1e9, t_dev))
    srv = xc.Server(ctx, nslots=nslots, use_graphs=use_graphs)
    # capture every resident system's build once (first-call graph capture is setup, not serving)
    if prewarm:
        for k in list(res.lru):
            d = host.get(k)
            for f in {f for _, kk, f in stream if kk == k}:
                try:
                    t = srv.submit(res.get(k, ()), f, precision, d["P"])
                except xc.XCError:
                    continue  # the request fails when it is served
                srv.wait(t)
    n = len(stream)
    arrive = np.array([a for a, _, _ in stream])
    done_at = np.zeros(n)
    pending = collections.deque()
    inflight = {}  # ticket -> request index
    busy = set()
    energies = {}
    failed = {}     # request -> error message: a failing request ends alone, the others go on
    cold = set()    # requests whose system had to be (re)built on the device when they were served
    nopack = set()  # requests of a pack that failed: served one by one
    packbuf = {}  # key -> (dP, dF) device matrices of systems built in packs
    npacked = npacks = 0

    def run_pack(group):
        """One packed build of the pending requests in group (distinct idle systems, one functional)."""
        pots, dPs, dFs = [], [], []
        keep = busy | {stream[j][1] for j in group}  # no member of the pack may be evicted for another
        for j in group:
            _, key, func = stream[j]
            d = host.get(key)
            nao = d["basis"].nao
            if key not in packbuf:
                dP, dF = ctx.alloc(8 * nao * nao), ctx.alloc(8 * nao * nao)
                ctx.upload(dP, np.asfortranarray(d["P"]))
                packbuf[key] = (dP, dF)
            pots.append(xc.FuncPotential(res.get(key, keep), func, precision))
            if res.last_created:
                cold.add(j)
            dPs.append(packbuf[key][0])
            dFs.append(packbuf[key][1])
        E = xc.fock_build_packed(pots, dPs, dFs)
        for j, e, dF in zip(group, E, dFs):
            _, key, func = stream[j]
            nao = host.get(key)["basis"].nao
            F = np.empty((nao, nao), order="F")
            ctx.download(F, dF)
            energies[(key, func)] = e
        return time.perf_counter() - t0

    i = nxt = 0
    t0 = time.perf_counter()
    while nxt < n:
        now = time.perf_counter() - t0
        while i < n and arrive[i] <= now:
            pending.append(i)
            i += 1
        if pack_nao:
            groups = {}
            for j in pending:
                _, key, func = stream[j]
                if j in nopack or key in busy or host.get(key)["basis"].nao > pack_nao:
                    continue
                g = groups.setdefault(func, {})
                g.setdefault(key, j)  # one request per system in a pack
            for func, g in groups.items():
                if len(g) < 2:
                    continue
                group = list(g.values())
                try:
                    t_done = run_pack(group)
                except xc.XCError:
                    nopack.update(group)  # each request then succeeds or fails on its own
                    continue
                for j in group:
                    done_at[j] = t_done
                    pending.remove(j)
                    nxt += 1
                npacked += len(group)
                npacks += 1
        # dispatch in arrival order; a request whose system is busy waits, later ones may pass it
        for j in list(pending):
            _, key, func = stream[j]
            if key in busy:
                continue
            bf = res.get(key, busy)
            if res.last_created:
                cold.add(j)
            try:
                t = srv.submit(bf, func, precision, host.get(key)["P"])
            except xc.XCError as e:
                failed[j] = str(e)
                done_at[j] = time.perf_counter() - t0
                pending.remove(j)
                nxt += 1
                continue
            if t is None:
                break
            inflight[t] = j
            busy.add(key)
            pending.remove(j)
        for t in list(inflight):
            out = srv.poll(t)
            if out is not None:
                j = inflight.pop(t)
                done_at[j] = time.perf_counter() - t0
                _, key, func = stream[j]
                busy.discard(key)
                energies[(key, func)] = out[1]
                nxt += 1
    wall = time.perf_counter() - t0
    ok = np.array([j not in failed for j in range(n)])
    lat_all = (done_at - arrive) * 1e3
    lat = lat_all[ok]
    is_cold = np.array([j in cold for j in range(n)])
    lat_cold, lat_warm = lat_all[ok & is_cold], lat_all[ok & ~is_cold]

    def pct(x, q):
        return round(float(np.percentile(x, q)), 3) if len(x) else None
    st = srv.stats()
    srv._release()
    for dP, dF in packbuf.values():
        ctx.free(dP)
        ctx.free(dF)
    for bf, _ in res.lru.values():
        bf._release()
    return dict(requests=n, qps_offered=round(n / arrive[-1], 2), slots=nslots, precision=precision,
                p50_ms=round(float(np.percentile(lat, 50)), 3), p99_ms=round(float(np.percentile(lat, 99)), 3),
                mean_ms=round(float(lat.mean()), 3), max_ms=round(float(lat.max()), 3),
                builds_per_s=round(n / wall, 2), wall_s=round(wall, 3), distinct_systems=len(keys),
                budget_gb=budget_gb, resident_gb=round(res.bytes / 1e9, 2), created=res.created, evicted=res.evicted,
                cold_requests=int(is_cold[ok].sum()), cold_p50_ms=pct(lat_cold, 50), cold_p99_ms=pct(lat_cold, 99),
                warm_p50_ms=pct(lat_warm, 50), warm_p99_ms=pct(lat_warm, 99),
                graph_launches=st["graph_launches"], host_setup_s=round(t_host, 1), device_setup_s=round(t_dev, 1),
                pack_max_nao=pack_nao, packed_requests=npacked, packed_launches=npacks, failed=len(failed),
                _failed=failed,
                _energies=energies, _host=host)


def fp64_check(ctx, host, energies, sample=None, seed=W.SEED):
    """Max |dE_xc| of the served FP32 builds against FP64 builds on the GPU: over every distinct
    (system, functional) served (sample None), or over a seeded sample of them."""
    keys = sorted(energies)
    if sample is None or sample >= len(keys):
        pick = keys
    else:
        rng = np.random.default_rng(seed + 11)
        pick = [keys[j] for j in rng.choice(len(keys), size=sample, replace=False)]
    worst = 0.0
    by_system = collections.defaultdict(list)
    for key, func in pick:
        by_system[key].append(func)
    for key, funcs in by_system.items():
        d = host.get(key)
        bf = xc.BasisFunctionOnGridController(ctx, d["basis"], host.grid(key), 128, 1e-9, 1)
        for func in funcs:
            pot = xc.FuncPotential(bf, func, "FP64")
            pot.getMatrix(d["P"])
            worst = max(worst, abs(pot.getEnergy() - energies[(key, func)]))
        bf._release()
    return worst, len(pick)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--qps", type=float, default=400.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=250.0)
    ap.add_argument("--fp64-sample", type=int, default=16)
    ap.add_argument("--pack-nao", type=int, default=0,
                    help="pack waiting requests of systems with at most this many basis functions (0: off)")
    args = ap.parse_args()
    log = lambda *a: print("[serving]", *a, file=sys.stderr, flush=True)
    ctx = xc.Context(0)
    stream = make_stream(args.requests, args.qps)
    out = run_stream(ctx, stream, nslots=args.slots, budget_gb=args.budget_gb, log=log, pack_nao=args.pack_nao)
    energies, host = out.pop("_energies"), out.pop("_host")
    out.pop("_failed")
    out["max_dE_xc_fp32_vs_fp64"], out["fp64_sample"] = fp64_check(ctx, host, energies, args.fp64_sample)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
