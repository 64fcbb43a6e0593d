"""XC Fock-build benchmark on one MI355X.

Workload (default): (H2O)64 on a 4 x 4 x 4 lattice (seeded jitter and
orientations) / def2-TZVP (2752 AOs) / PBE on the default SSF accuracy-4 grid,
with the synthetic density P = 2 C_occ C_occ^T (320 occupied orbitals,
C_occ = S^-1/2 Q).  A step is one XC Fock build P -> (F, E_xc) with the grid,
basis and AO tiles resident in HBM and P already on the device.  Three
contraction precisions are timed on the same P: the mixed form (FP32-MFMA with
the dense blocks' density contracted again in FP64; the headline value), plain
FP32-MFMA and the all-FP64 path; dE_xc and max|dF| of each against FP64 are
reported.  The CPU baseline is the oracle timed over a bounded sample of the
workload (its first grid points), and the GPU builds of the same sample are
checked against it.  The same build is also timed through the serving path
(host P in, host F out, replayed graph).

Beside the headline line the run reports (in the same JSON object) the same
timings for C20H42/def2-SVP/PBE, the embedding sweep of C6H6 + 24 H2O
(non-additive PBE + TF potentials), packed builds of small systems, a serving
stream of 2000 Poisson-arriving builds at a fixed QPS (p50/p99 latency,
builds/s), and the full restricted SCFs with RI-J (SMALL-grid iterations,
DEFAULT-grid final energy) of (H2O)8/def2-SVP/LDA and (H2O)64/def2-TZVP/PBE in
FP64 and in the mixed form (FP32 iterations + FP64 finish), beside the
(H2O)8 SCF driven by the CPU oracle's XC build (the (H2O)64 one with
--scf3-oracle).

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--config 3] [--no-fde] [--no-serving] [--no-scf]
"""
import argparse
import json
import os
import sys
import time

# OpenMP placement of the CPU baseline (read by the OpenMP runtimes when they load)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_TFLOPS = {"FP64": 78.6, "FP32": 157.3, "MIXED": 157.3}  # MI355X dense matrix peaks (MIXED: its FP32 part)
PEAK_HBM_GBS = 8000.0
TRAFFIC_PROFILE = {2: os.path.join(HERE, "profiles", "xc_fock_build_c20h42_svp_pbe_hbm_traffic.md"),
                   3: os.path.join(HERE, "profiles", "xc_fock_build_h2o64_tzvp_pbe_hbm_traffic.md")}


def measured_traffic(config):
    """HBM bytes per build of each contraction stage, from the rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes summarised in the profile (lines "traffic <stage> <precision> <bytes>"), or {}."""
    out = {}
    path = TRAFFIC_PROFILE.get(config)
    if path and os.path.exists(path):
        for line in open(path):
            p = line.split()
            if len(p) == 4 and p[0] == "traffic":
                out[(p[1], p[2])] = float(p[3])
    return out


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def time_path(ctx, pot, dP, dF, steps, warmup):
    for _ in range(warmup):
        pot.build_device(dP, dF, sync=False)
    ctx.synchronize()
    ctx.profile_collect()
    ctx.profile(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pot.build_device(dP, dF, sync=False)
    ctx.synchronize()
    t1 = time.perf_counter()
    ctx.profile(False)
    stages, nb = ctx.profile_collect()
    assert nb == steps, (nb, steps)
    return (t1 - t0) * 1e3 / steps, {k: v / steps for k, v in stages.items()}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sysd, target_seconds):
    """The oracle's XC Fock build over a bounded sample of the workload: every k-th 128-point block
    of the Hilbert-ordered grid (the same mix of dense and sparse regions as the whole), k sized
    from a probe over every 50th block so that the timed build takes about target_seconds."""
    from oracle import oracle as O
    from serenity_amd import xc
    b, g, P, func = sysd["basis"], sysd["grid"], sysd["P"], sysd["cfg"]["functional"]
    pts, w = np.asarray(g.points).reshape(-1, 3), np.asarray(g.weights)
    N = len(w)
    nblk = (N + 127) // 128

    def sample(k):
        sel = np.concatenate([np.arange(128 * i, min(N, 128 * i + 128)) for i in range(0, nblk, k)])
        return np.ascontiguousarray(pts[sel]), np.ascontiguousarray(w[sel])
    sp, sw = sample(50)
    t0 = time.perf_counter()
    O.fock_build(b, sp, sw, P, func)
    t_probe = time.perf_counter() - t0
    k = max(1, int(round(50 * t_probe / target_seconds)))
    sp, sw = sample(k)
    t0 = time.perf_counter()
    F, E, _, _ = O.fock_build(b, sp, sw, P, func)
    el = time.perf_counter() - t0
    return dict(seconds=el, points=len(sw), stride=k, F=F, E=E, threads=O.max_threads(),
                grid=xc.GridController(sp, sw))


def served_build(ctx, bf, functional, P, reps=10):
    """The same build through the serving path: host P in, host F and E_xc out, one slot, the
    controller's build replayed from its captured graph (first submission captures it)."""
    from serenity_amd import xc
    out = {}
    srv = xc.Server(ctx, nslots=1, use_graphs=True)
    try:
        for prec in ("FP64", "FP32", "MIXED"):
            srv.wait(srv.submit(bf, functional, prec, P))
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                F, E = srv.wait(srv.submit(bf, functional, prec, P))
                ts.append(time.perf_counter() - t0)
            ms = 1e3 * float(np.median(ts))
            out[prec.lower()] = dict(ms_per_build_median=round(ms, 4), builds_per_s=round(1e3 / ms, 2), E=E)
    finally:
        srv._release()
    return out


def time_config(ctx, config, steps, warmup):
    """FP64 and FP32 builds of another benchmark configuration on its synthetic P."""
    from serenity_amd import xc, workloads as W
    sysd = W.build_system(config)
    b, g, P, cfg = sysd["basis"], sysd["grid"], sysd["P"], sysd["cfg"]
    N, n = g.getNGridPoints(), b.nao
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    dP, dF = ctx.alloc(8 * n * n), ctx.alloc(8 * n * n)
    ctx.upload(dP, np.asfortranarray(P).ravel(order="F"))
    res = {}
    for prec in ("FP64", "FP32"):
        pot = xc.FuncPotential(bf, cfg["functional"], prec)
        ms, stages = time_path(ctx, pot, dP, dF, steps, warmup)
        E = pot.build_device(dP, dF, sync=True)
        F = np.zeros(n * n)
        ctx.download(F, dF)
        res[prec] = (ms, stages, E, F)
    ctx.free(dP)
    ctx.free(dF)
    bf._release()
    (m64, s64, e64, f64), (m32, s32, e32, f32) = res["FP64"], res["FP32"]
    return dict(workload=cfg["name"], npoints=N, nao=n,
                fp64=dict(ms_per_build=round(m64, 5), grid_pts_per_s=N / (m64 * 1e-3), stages_ms=s64),
                fp32_mfma=dict(ms_per_build=round(m32, 5), grid_pts_per_s=N / (m32 * 1e-3), stages_ms=s32),
                speedup_fp32_over_fp64=m64 / m32, dE_xc_fp32_vs_fp64=abs(e32 - e64),
                max_dF_fp32_vs_fp64=float(np.abs(f32 - f64).max()))


def fde_sweep(ctx, steps, warmup, check_oracle):
    """C6H6 + 24 H2O / def2-SVP on the supersystem grid (FP64).  One freeze-and-thaw sweep visits every
    subsystem as the active one: its density is updated (as after its SCF step) and the PBE
    non-additive XC and TF non-additive kinetic potentials are built on the cached subsystem
    densities (25 density updates + 50 NAdd builds), every density matrix resident on the device.
    A direct NAdd build (all 25 densities recomputed inside the call) is timed beside it."""
    from serenity_amd import xc, workloads as W
    emb = W.build_embedding(24, accuracy=4)
    g, subs = emb["grid"], emb["subsystems"]
    bfs = [xc.BasisFunctionOnGridController(ctx, s["basis"], g, 128, 1e-9, 1, keep_fp32=True) for s in subs]
    dP, dF = [], []
    for s in subs:
        n = s["basis"].nao
        d = ctx.alloc(8 * n * n)
        ctx.upload(d, np.asfortranarray(s["P"]).ravel(order="F"))
        dP.append(d)
        dF.append(ctx.alloc(8 * n * n))
    steps_a = []
    for a in range(len(subs)):
        env = [i for i in range(len(subs)) if i != a]
        steps_a.append((a, env, [xc.NAddFuncPotential(bfs[a], [bfs[i] for i in env], f, "FP64")
                                 for f in ("PBE", "TF")]))
    for i in range(len(subs)):
        bfs[i].update_density(dP[i])

    def sweep():
        for a, env, pots in steps_a:
            bfs[a].update_density(dP[a])
            for pot in pots:
                pot.build_cached(dF[a], sync=False)

    for _ in range(warmup):
        sweep()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sweep()
    ctx.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    a, env, pots = steps_a[0]
    direct = pots[0]
    direct.build_device(dP[a], [dP[i] for i in env], dF[a], sync=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        direct.build_device(dP[a], [dP[i] for i in env], dF[a], sync=False)
    ctx.synchronize()
    ms_direct = (time.perf_counter() - t0) * 1e3 / steps
    # each subsystem's own PBE XC matrix, all 25 in one packed build per cycle (FP32 contractions; an
    # FP64 packed build finishes), against the same builds issued one by one
    dG = [ctx.alloc(8 * s["basis"].nao ** 2) for s in subs]
    own = {}
    reps = max(10, steps)
    for prec in ("FP32", "FP64"):
        pots_own = [xc.FuncPotential(bf, "PBE", prec) for bf in bfs]
        alone = [p.build_device(dP[i], dG[i]) for i, p in enumerate(pots_own)]
        for _ in range(3):  # first calls grow the merged-list buffers (small packs also capture a graph)
            xc.fock_build_packed(pots_own, dP, dG)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for i, p in enumerate(pots_own):
                p.build_device(dP[i], dG[i], sync=False)
        ctx.synchronize()
        t_one = (time.perf_counter() - t0) * 1e3 / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            E = xc.fock_build_packed(pots_own, dP, dG)
        t_pack = (time.perf_counter() - t0) * 1e3 / reps
        own[prec] = dict(E=E, ms_one_by_one=round(t_one, 4), ms_packed=round(t_pack, 4),
                         max_dE_packed_vs_alone=float(max(abs(a - b) for a, b in zip(E, alone))))
    own_xc = dict(fp32=dict((k, v) for k, v in own["FP32"].items() if k != "E"),
                  fp64=dict((k, v) for k, v in own["FP64"].items() if k != "E"),
                  max_dE_fp32_vs_fp64=float(max(abs(a - b) for a, b in zip(own["FP32"]["E"], own["FP64"]["E"]))))
    out = dict(workload="C6H6+24H2O/def2-SVP, PBE + TF non-additive potentials, supersystem SSF accuracy-4 grid, "
                        "25 subsystems, synthetic subsystem densities",
               npoints=g.getNGridPoints(), nao_total=int(sum(s["basis"].nao for s in subs)),
               density_updates_per_sweep=len(subs), nadd_builds_per_sweep=2 * len(subs), ms_per_sweep=round(ms, 4),
               ms_direct_nadd_build=round(ms_direct, 4), subsystem_xc_builds=own_xc)
    if check_oracle:
        from oracle import oracle as O
        E = pots[0].build_cached(dF[a], sync=True)
        n = subs[a]["basis"].nao
        F = np.zeros(n * n)
        ctx.download(F, dF[a])
        Fo, Eo = O.nadd_fock_build(subs[a]["basis"], [subs[i]["basis"] for i in env], g.points, g.weights,
                                   subs[a]["P"], [subs[i]["P"] for i in env], "PBE")
        out["benzene_pbe_nadd_dE_vs_oracle"] = abs(E - Eo)
        out["benzene_pbe_nadd_max_dF_vs_oracle"] = float(np.abs(F.reshape(n, n, order="F") - Fo).max())
    for d in dP + dF + dG:
        ctx.free(d)
    return out


def packed_leg(ctx, reps=20):
    """32 small systems of the serving mix ((H2O)n n = 1..8 and CkH2k+2 k = 1..8, def2-SVP and
    def2-TZVP, PBE): one build each on the context stream (launches of one build after another)
    vs one packed build of all 32 (one launch per stage and work class), per precision."""
    from serenity_amd import workloads as W, xc
    systems = []
    for basis in ("DEF2-SVP", "DEF2-TZVP"):
        for kind in ("water", "alkane"):
            for size in range(1, 9):
                Z, X = W.water_chain(size) if kind == "water" else W.alkane(size)
                b = xc.BasisController.from_file(basis, Z, X)
                g = xc.molecular_grid(Z, X, ctx=ctx)
                P = W.synthetic_density(b.overlap(), W.n_electrons(Z) // 2, seed=size)
                bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
                n = b.nao
                dP, dF = ctx.alloc(8 * n * n), ctx.alloc(8 * n * n)
                ctx.upload(dP, np.asfortranarray(P))
                systems.append(dict(bf=bf, n=n, dP=dP, dF=dF, npts=g.getNGridPoints()))
    out = dict(workload="32 small systems: (H2O)n n=1..8 and CkH2k+2 k=1..8, def2-SVP and def2-TZVP, PBE, "
                        "accuracy-4 grids", systems=len(systems), nao_range=[min(s["n"] for s in systems),
                                                                            max(s["n"] for s in systems)],
               grid_points=int(sum(s["npts"] for s in systems)))
    for prec in ("FP32", "FP64"):
        pots = [xc.FuncPotential(s["bf"], "PBE", prec) for s in systems]
        dPs, dFs = [s["dP"] for s in systems], [s["dF"] for s in systems]
        alone = [p.build_device(s["dP"], s["dF"]) for p, s in zip(pots, systems)]
        for _ in range(3):
            xc.fock_build_packed(pots, dPs, dFs)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            for p, s in zip(pots, systems):
                p.build_device(s["dP"], s["dF"], sync=False)
        ctx.synchronize()
        t_seq = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        for _ in range(reps):
            E = xc.fock_build_packed(pots, dPs, dFs)
        t_pack = (time.perf_counter() - t0) / reps
        out[prec.lower()] = dict(ms_one_by_one=round(t_seq * 1e3, 4), ms_packed=round(t_pack * 1e3, 4),
                                 builds_per_s_one_by_one=round(len(systems) / t_seq, 1),
                                 builds_per_s_packed=round(len(systems) / t_pack, 1),
                                 speedup=round(t_seq / t_pack, 3),
                                 max_dE_packed_vs_alone=float(max(abs(a - b) for a, b in zip(E, alone))))
    for s in systems:
        ctx.free(s["dP"])
        ctx.free(s["dF"])
        s["bf"]._release()
    return out


def scf_leg(ctx, check_oracle, config=1, cycle_log=None):
    """Full SCF of a benchmark configuration (RI-J): FP64, mixed with FP64 finish, and (optionally) the
    oracle-driven SCF.  Config 1 computes J on the host, config 3 with the GPU J engine."""
    from serenity_amd import scf as SCF, workloads as W
    cfg = W.CONFIGS[config]
    Z, X = W.geometry(cfg["system"])
    t0 = time.perf_counter()
    gpu_j = config != 1
    task = SCF.ScfTask(ctx, Z, X, basis=cfg["basis"], functional=cfg["functional"], gpu_coulomb=gpu_j,
                       log=cycle_log)
    out = dict(workload="%s, RI-J (RI_J_WEIGEND, %s), SCF on the accuracy-2 grid, final energy on the accuracy-4 "
                        "grid" % (cfg["name"], "GPU J engine" if gpu_j else "host 3-centre integrals"),
               nao=task.basis.nao, naux=task.auxbasis.nao,
               npts_small=task.grid_small.getNGridPoints(), npts_default=task.grid_default.getNGridPoints(),
               setup_s=round(time.perf_counter() - t0, 3))
    if config == 1:
        task.run("FP64")  # first run builds the resident tiles of both grids
    r64 = task.run("FP64")
    rmx = task.run("MIXED")
    out["fp64"] = dict(energy=r64["energy"], cycles=r64["cycles"], wall_s=r64["wall_s"], xc_s=r64["t_xc_s"],
                       j_s=r64["t_j_s"], host_algebra_s=r64["t_alg_s"])
    out["mixed"] = dict(energy_before_fp64_finish=rmx["energy_before_fp64_finish"], energy=rmx["energy"],
                        cycles_fp32=rmx["cycles_fp32"], cycles_fp64_finish=rmx["cycles_fp64_finish"],
                        wall_s=rmx["wall_s"])
    out["mixed"]["dE_before_finish_vs_fp64"] = abs(rmx["energy_before_fp64_finish"] - r64["energy"])
    # the FP32-converged density's energy in FP64: its convergence error alone
    out["mixed"]["dE_fp32_density_in_fp64_vs_fp64"] = abs(task.final_energy(rmx["P_fp32"], "FP64") - r64["energy"])
    out["mixed"]["dE_after_finish_vs_fp64"] = abs(rmx["energy"] - r64["energy"])
    if check_oracle:
        from oracle import oracle as O
        t0 = time.perf_counter()
        b, gs, gd = task.basis, task.grid_small, task.grid_default

        f = cfg["functional"]

        def build(P):
            F, E, _, _ = O.fock_build(b, gs.points, gs.weights, P, f)
            return E, F
        r = SCF.perform(task.ints, task.nocc, build, P0=task.P_guess, log=cycle_log)
        _, e_xc, _, _ = O.fock_build(b, gd.points, gd.weights, r["P"], f)
        e = SCF.total_energy(task.ints, r["P"], e_xc)
        out["cpu_oracle"] = dict(energy=e, cycles=r["cycles"], wall_s=round(time.perf_counter() - t0, 3),
                                 threads=O.max_threads())
        out["fp64"]["dE_vs_cpu_oracle"] = abs(r64["energy"] - e)
        out["mixed"]["dE_after_finish_vs_cpu_oracle"] = abs(rmx["energy"] - e)
    task.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=(1, 2, 3))
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fde", action="store_true", help="skip the embedding (non-additive potential) sweep")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving stream")
    ap.add_argument("--no-scf", action="store_true", help="skip the full SCFs of configs 1 and 3")
    ap.add_argument("--no-packed", action="store_true", help="skip the packed-build comparison of small systems")
    ap.add_argument("--no-extra-config", action="store_true",
                    help="skip the timing of the second configuration (C20H42/def2-SVP/PBE beside config 3)")
    ap.add_argument("--no-served", action="store_true", help="skip the build timed through the serving path")
    ap.add_argument("--no-scf3", action="store_true",
                    help="skip the full SCF of (H2O)64/def2-TZVP/PBE (FP64 and mixed; about a minute)")
    ap.add_argument("--scf3-oracle", action="store_true",
                    help="also run the (H2O)64 SCF driven by the CPU oracle's XC build (about 10 min on 16 threads)")
    ap.add_argument("--serving-requests", type=int, default=2000)
    ap.add_argument("--serving-qps", type=float, default=400.0)
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="issue one device copy of this size first (HBM byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the XC Fock build runs on one MI355X here "
              "(multi-GPU block sharding is described in the design notes, not implemented).", file=sys.stderr)
        sys.exit(2)

    from serenity_amd import xc, workloads as W

    t0 = time.perf_counter()
    sysd = W.build_system(args.config)
    b, g, P, cfg = sysd["basis"], sysd["grid"], sysd["P"], sysd["cfg"]
    N, n = g.getNGridPoints(), b.nao
    log("workload %s: %d grid points, %d AOs, host setup %.2fs" % (cfg["name"], N, n, time.perf_counter() - t0))

    ctx = xc.Context(0)
    t0 = time.perf_counter()
    bf = xc.BasisFunctionOnGridController(ctx, b, g, 128, 1e-9, 1, keep_fp32=True)
    ctx.synchronize()
    t_ao = time.perf_counter() - t0
    sum_nb, sum_Bnb2, tile_bytes = bf.stats()
    nblocks = bf.getNBlocks()
    log("AO tiles resident: %d blocks, mean n_b %.1f, %.2f GB, setup %.2fs" %
        (nblocks, sum_nb / nblocks, tile_bytes / 1e9, t_ao))

    if args.copy_calibration_mb > 0:
        nb = args.copy_calibration_mb << 20
        a, c = ctx.alloc(nb), ctx.alloc(nb)
        ctx.copy(c, a, nb)
        ctx.free(a)
        ctx.free(c)
        log("calibration copy: %d MiB" % args.copy_calibration_mb)

    dP, dF = ctx.alloc(8 * n * n), ctx.alloc(8 * n * n)
    ctx.upload(dP, np.asfortranarray(P).ravel(order="F"))
    results = {}
    for prec in ("FP64", "FP32", "MIXED"):
        pot = xc.FuncPotential(bf, cfg["functional"], prec)
        ms, stages = time_path(ctx, pot, dP, dF, args.steps, args.warmup)
        E = pot.build_device(dP, dF, sync=True)
        F = np.zeros(n * n)
        ctx.download(F, dF)
        results[prec] = dict(ms=ms, stages=stages, E=E, F=F.reshape(n, n, order="F"))
        log("%s: %.3f ms/build  (%.3e pts/s)  stages(ms) %s" %
            (prec, ms, N / (ms * 1e-3), {k: round(v, 4) for k, v in stages.items()}))
    ctx.free(dP)
    ctx.free(dF)

    # algorithmic work: 2 sum_b B_b n_b^2 flops per contraction (phi.P and phi^T.G)
    flops_per_kernel = 2.0 * sum_Bnb2

    traffic = measured_traffic(args.config)

    def roofline(prec):
        st = results[prec]["stages"]
        kern = max(("density", "fock"), key=lambda k: st[k])
        achieved = flops_per_kernel / (st[kern] * 1e-3) / 1e12
        tb = traffic.get((kern, prec))
        return dict(bound="mfma", kernel=kern, achieved=round(achieved, 3), peak=PEAK_TFLOPS[prec], unit="TFLOP/s",
                    frac=round(achieved / PEAK_TFLOPS[prec], 5), traffic=tb,
                    hbm_gbs=round(tb / (st[kern] * 1e-3) / 1e9, 1) if tb else None,
                    flops_per_launch=flops_per_kernel, avg_launch_ms=round(st[kern], 5))

    r64, r32, rmx = results["FP64"], results["FP32"], results["MIXED"]
    out_extra = dict(
        dE_xc_mixed_vs_fp64=abs(rmx["E"] - r64["E"]),
        max_dF_mixed_vs_fp64=float(np.abs(rmx["F"] - r64["F"]).max()),
        dE_xc_fp32_vs_fp64=abs(r32["E"] - r64["E"]),
        max_dF_fp32_vs_fp64=float(np.abs(r32["F"] - r64["F"]).max()),
        E_xc_fp64=r64["E"],
        fp64=dict(ms_per_build=round(r64["ms"], 5), grid_pts_per_s=N / (r64["ms"] * 1e-3),
                  builds_per_s=1e3 / r64["ms"], stages_ms=r64["stages"], roofline=roofline("FP64")),
        fp32_mfma=dict(ms_per_build=round(r32["ms"], 5), grid_pts_per_s=N / (r32["ms"] * 1e-3),
                       builds_per_s=1e3 / r32["ms"], stages_ms=r32["stages"], roofline=roofline("FP32")),
        mixed=dict(ms_per_build=round(rmx["ms"], 5), grid_pts_per_s=N / (rmx["ms"] * 1e-3),
                   builds_per_s=1e3 / rmx["ms"], stages_ms=rmx["stages"],
                   form="FP32-MFMA contractions; the density of the blocks whose largest density exceeds 2 bohr^-3 "
                        "contracted again in FP64; FP64 functional, reductions and F accumulation"),
        speedup_fp32_over_fp64=r64["ms"] / r32["ms"],
        speedup_mixed_over_fp64=r64["ms"] / rmx["ms"],
        ao_setup_s=round(t_ao, 3),
    )

    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline: oracle (restatement of Serenity's per-block path, 3 AO passes per build)")
        cb = cpu_baseline(sysd, args.cpu_seconds)
        pts_s = cb["points"] / cb["seconds"]
        cpu = dict(value=pts_s, unit="grid pts/s", cores=cb["threads"], kind="port",
                   sample="one XC Fock build of the workload restricted to every %d-th 128-point block of its "
                          "Hilbert-ordered grid, %d of %d points (%.1f s; CPU restatement of Serenity's per-block "
                          "path, OpenMP schedule(dynamic), FP64, register-blocked built-in loops, no BLAS)"
                          % (cb["stride"], cb["points"], N, cb["seconds"]),
                   builds_per_s=pts_s / N, cpu_model=cpu_model(),
                   omp=dict(OMP_NUM_THREADS=os.environ.get("OMP_NUM_THREADS"),
                            OMP_PROC_BIND=os.environ.get("OMP_PROC_BIND"),
                            OMP_PLACES=os.environ.get("OMP_PLACES")))
        # the same sample built on the GPU, checked against the oracle
        sbf = xc.BasisFunctionOnGridController(ctx, sysd["basis"], cb["grid"], 128, 1e-9, 1, keep_fp32=True)
        for prec in ("FP64", "FP32", "MIXED"):
            pot = xc.FuncPotential(sbf, cfg["functional"], prec)
            Fs = pot.getMatrix(P)
            out_extra["sample_dE_xc_%s_vs_oracle" % prec.lower()] = abs(pot.getEnergy() - cb["E"])
            out_extra["sample_max_dF_%s_vs_oracle" % prec.lower()] = float(np.abs(Fs - cb["F"]).max())
        sbf._release()
        log("oracle: %.0f grid pts/s on %d threads over %d points; FP64 on the same points |dE| %.2e max|dF| %.2e" %
            (pts_s, cb["threads"], cb["points"], out_extra["sample_dE_xc_fp64_vs_oracle"],
             out_extra["sample_max_dF_fp64_vs_oracle"]))

    if not args.no_served:
        sv1 = served_build(ctx, bf, cfg["functional"], P)
        for k in ("fp64", "fp32", "mixed"):
            sv1[k]["dE_vs_device_build"] = abs(sv1[k].pop("E") - results[k.upper()]["E"])
        sv1["note"] = ("one slot; each build copies P host->device and F device->host (%.1f MB each way) "
                       "around the replayed graph" % (8.0 * n * n / 1e6))
        log("served build: FP64 %.3f ms, FP32 %.3f ms, mixed %.3f ms (median, host P/F)" %
            (sv1["fp64"]["ms_per_build_median"], sv1["fp32"]["ms_per_build_median"],
             sv1["mixed"]["ms_per_build_median"]))
        out_extra["served_build"] = sv1

    if not args.no_extra_config:
        other = 2 if args.config == 3 else 3
        ex = time_config(ctx, other, args.steps, args.warmup)
        log("%s: FP64 %.3f ms, FP32 %.3f ms (%.2fx), dE %.2e" %
            (ex["workload"], ex["fp64"]["ms_per_build"], ex["fp32_mfma"]["ms_per_build"],
             ex["speedup_fp32_over_fp64"], ex["dE_xc_fp32_vs_fp64"]))
        out_extra["config%d" % other] = ex

    if not args.no_fde:
        fde = fde_sweep(ctx, max(1, args.steps // 4), 1, not args.no_cpu_baseline)
        log("FDE sweep: %.3f ms (%d density updates + %d cached NAdd builds); direct NAdd build %.3f ms" %
            (fde["ms_per_sweep"], fde["density_updates_per_sweep"], fde["nadd_builds_per_sweep"],
             fde["ms_direct_nadd_build"]))
        ox = fde["subsystem_xc_builds"]
        log("FDE subsystem XC builds (25, packed): FP32 %.3f ms packed vs %.3f ms one by one; FP64 %.3f / %.3f ms; "
            "max dE FP32 vs FP64 %.2e" % (ox["fp32"]["ms_packed"], ox["fp32"]["ms_one_by_one"],
                                         ox["fp64"]["ms_packed"], ox["fp64"]["ms_one_by_one"],
                                         ox["max_dE_fp32_vs_fp64"]))
        out_extra["fde"] = fde

    if not args.no_packed:
        pk = packed_leg(ctx)
        log("packed builds of %d small systems (%d-%d AOs, %d points): FP32 %.3f ms one by one, %.3f ms packed "
            "(%.2fx); FP64 %.3f / %.3f ms (%.2fx)" %
            (pk["systems"], pk["nao_range"][0], pk["nao_range"][1], pk["grid_points"], pk["fp32"]["ms_one_by_one"],
             pk["fp32"]["ms_packed"], pk["fp32"]["speedup"], pk["fp64"]["ms_one_by_one"], pk["fp64"]["ms_packed"],
             pk["fp64"]["speedup"]))
        out_extra["packed"] = pk

    if not args.no_scf:
        sc = scf_leg(ctx, not args.no_cpu_baseline)
        log("SCF (H2O)8/SVP/LDA: FP64 E %.10f (%d cycles, %.2f s); mixed before finish dE %.2e, after %.2e "
            "(%d+%d cycles, %.2f s)%s" %
            (sc["fp64"]["energy"], sc["fp64"]["cycles"], sc["fp64"]["wall_s"],
             sc["mixed"]["dE_before_finish_vs_fp64"], sc["mixed"]["dE_after_finish_vs_fp64"],
             sc["mixed"]["cycles_fp32"], sc["mixed"]["cycles_fp64_finish"], sc["mixed"]["wall_s"],
             "" if "cpu_oracle" not in sc else "; CPU oracle SCF %.2f s, |dE| %.2e" %
             (sc["cpu_oracle"]["wall_s"], sc["fp64"]["dE_vs_cpu_oracle"])))
        out_extra["scf"] = sc
    if not args.no_scf3 and not args.no_scf:
        sc3 = scf_leg(ctx, args.scf3_oracle and not args.no_cpu_baseline, config=3,
                      cycle_log=lambda *a: log("  scf3", *a))
        log("SCF (H2O)64/TZVP/PBE: FP64 E %.10f (%d cycles, %.1f s; XC %.2f s, J %.2f s, host algebra %.2f s); "
            "mixed before finish dE %.2e, after %.2e (%d+%d cycles, %.1f s)" %
            (sc3["fp64"]["energy"], sc3["fp64"]["cycles"], sc3["fp64"]["wall_s"], sc3["fp64"]["xc_s"],
             sc3["fp64"]["j_s"], sc3["fp64"]["host_algebra_s"],
             sc3["mixed"]["dE_before_finish_vs_fp64"], sc3["mixed"]["dE_after_finish_vs_fp64"],
             sc3["mixed"]["cycles_fp32"], sc3["mixed"]["cycles_fp64_finish"], sc3["mixed"]["wall_s"]))
        out_extra["scf_config3"] = sc3

    if not args.no_serving:
        from serenity_amd import serving as SV
        bf._release()  # the headline workload's controller: the serving legs get the device to themselves
        stream = SV.make_stream(args.serving_requests, args.serving_qps)
        sv = SV.run_stream(ctx, stream, nslots=4, log=log)
        energies, host = sv.pop("_energies"), sv.pop("_host")
        sv.pop("_failed")
        # every distinct (system, functional) served, against FP64 builds of the same density
        sv["max_dE_xc_fp32_vs_fp64"], sv["fp64_checked"] = SV.fp64_check(ctx, host, energies)
        sv["workload"] = ("Poisson stream: (H2O)_n n~U{1..48} / CkH2k+2 k~U{1..30}, def2-SVP/TZVP, LDA/PBE, "
                          "FP32-MFMA builds on resident FP32 tiles")
        log("serving: %d requests at %.0f QPS: p50 %.3f ms, p99 %.3f ms, %.1f builds/s, max dE vs FP64 %.2e" %
            (sv["requests"], sv["qps_offered"], sv["p50_ms"], sv["p99_ms"], sv["builds_per_s"],
             sv["max_dE_xc_fp32_vs_fp64"]))
        # the same stream with waiting small requests packed into one launch per stage
        svp = SV.run_stream(ctx, stream, nslots=4, log=log, pack_nao=SV.PACK_NAO, host=host)
        ep = svp.pop("_energies")
        svp.pop("_host")
        svp.pop("_failed")
        svp["max_dE_packed_vs_slots"] = float(max(abs(ep[k] - energies[k]) for k in ep))
        log("serving with packing (<= %d AOs): p50 %.3f ms, p99 %.3f ms, %.1f builds/s, %d requests in %d packs, "
            "max dE vs the unpacked run %.2e" %
            (svp["pack_max_nao"], svp["p50_ms"], svp["p99_ms"], svp["builds_per_s"], svp["packed_requests"],
             svp["packed_launches"], svp["max_dE_packed_vs_slots"]))
        sv["packed"] = {k: v for k, v in svp.items() if k in ("p50_ms", "p99_ms", "mean_ms", "max_ms", "builds_per_s",
                                                               "pack_max_nao", "packed_requests", "packed_launches",
                                                               "max_dE_packed_vs_slots")}
        # offered load sweep (packing on): where p99 turns up
        keep = ("qps_offered", "p50_ms", "p99_ms", "builds_per_s", "packed_requests")
        sweep = []
        for qps in (200.0, 800.0, 1600.0):
            r = SV.run_stream(ctx, SV.make_stream(args.serving_requests, qps), nslots=4, log=log,
                              pack_nao=SV.PACK_NAO, host=host)
            for k in ("_energies", "_host", "_failed"):
                r.pop(k)
            sweep.append({k: r[k] for k in keep})
            log("serving at %.0f QPS: p50 %.3f ms, p99 %.3f ms, %.1f builds/s" %
                (r["qps_offered"], r["p50_ms"], r["p99_ms"], r["builds_per_s"]))
        sweep.append({k: svp[k] for k in keep})
        sv["qps_sweep"] = sorted(sweep, key=lambda x: x["qps_offered"])
        # under memory pressure: a 64 GB budget against the ~210 GB working set, so systems are evicted
        # and built again (grid on the GPU, AO tiles) inside the timed stream; offered at 30 QPS, below
        # what cold builds sustain, so that the latencies are build times rather than queueing
        pr = SV.run_stream(ctx, SV.make_stream(300, 30.0), nslots=4, budget_gb=64.0, log=log,
                           pack_nao=SV.PACK_NAO, host=host)
        for k in ("_energies", "_host", "_failed"):
            pr.pop(k)
        sv["pressure_64gb"] = {k: pr[k] for k in ("qps_offered", "requests", "p50_ms", "p99_ms", "builds_per_s", "created", "evicted",
                                                   "cold_requests", "cold_p50_ms", "cold_p99_ms", "warm_p50_ms",
                                                   "warm_p99_ms", "resident_gb")}
        log("serving under a 64 GB budget: p50 %.3f ms, p99 %.3f ms, %d evictions, %d cold requests "
            "(p50 %s ms, p99 %s ms)" % (pr["p50_ms"], pr["p99_ms"], pr["evicted"], pr["cold_requests"],
                                        pr["cold_p50_ms"], pr["cold_p99_ms"]))
        out_extra["serving"] = sv

    line = dict(
        metric="XC Fock-build grid pts/s, FP32-MFMA vs FP64 at stated dE; p50/p99 latency @QPS",
        value=N / (rmx["ms"] * 1e-3),
        unit="grid pts/s",
        n_gpus=1,
        steps=args.steps,
        warmup=args.warmup,
        ms_per_step=round(rmx["ms"], 5),
        higher_is_better=True,
        scaling="weak",
        vs_baseline=None,
        dtype="f32",
        data="synthetic",
        config=dict(workload=cfg["name"] + ", SSF/Ahlrichs/Lebedev accuracy-4 grid, seeded geometry, "
                    "synthetic P (P = 2 C C^T, C = S^-1/2 Q)", npoints=N, nao=n, nblocks=nblocks,
                    mean_nb=round(sum_nb / nblocks, 2), seed=W.SEED, parallelism="single GPU"),
        roofline=roofline("MIXED"),
        cpu_baseline=cpu,
    )
    line.update(out_extra)
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
