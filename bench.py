"""Benchmark: one tick (trace + sort + deposit + merge + FP64 chemistry
update) of the 256^3 H + He workload (16,777,216 particles, Zel'dovich
lattice, temperature evolution, 1024 sources with a 16-bin blackbody
spectrum, 65,536 rays per tick) on one MI355X, with exact chemistry and
with the MLP surrogate from the same state.

Prints one JSON line.  value = ray-particle intersections per second of the
whole tick (sum of full hit-list lengths, the reference's
stat.total_deposit_count, src/loop.c:423, over the wall time of the timed
ticks); the trace+deposit-only rate, ticks/s and per-kernel times are extra
fields.  Inputs (field, tree, particle state, ray lists of every tick) are
resident in HBM before the timed region starts.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per intersection: trace 16 (pos + sml); trace + deposit
# 52 (H-only) / 84 (H + He).  The deposit kernel's share: mass f32 + x_HI f64
# + read-modify-write of yGdepHI, heat, hits (3 x 8) = 36; H + He adds x_HeI,
# x_HeII (16) and yGdepHeI / yGdepHeII (16)
BYTES_TRACE = 16
BYTES_DEPOSIT_H = 36
BYTES_DEPOSIT_HE = 68
# HBM bytes per deposit launch from rocprofv3 FETCH_SIZE + WRITE_SIZE passes,
# scaled by the same counters on a copy of known size (FETCH_SIZE x 2,
# WRITE_SIZE x 1), the last dispatch of each several-ray kernel (its LDS
# size names it) in the traffic passes of each workload
# (profiles/k_trace_and_deposit_traffic.md); other kernels: null
TRAFFIC_DEPOSIT = {(2, "k_deposit_multi<false, 2>"): (56229 * 2 + 48219) * 1024,
                   (2, "k_deposit_multi<false, 4>"): (55851 * 2 + 47922) * 1024,
                   (2, "k_deposit_multi<false, 8>"): (55127 * 2 + 54072) * 1024,
                   (3, "k_deposit_multi<true, 2>"): (3397351 * 2 + 4979083) * 1024,
                   (3, "k_deposit_multi<true, 4>"): (3213322 * 2 + 5058251) * 1024,
                   (3, "k_deposit_multi<true, 8>"): (3128020 * 2 + 6280064) * 1024,
                   (5, "k_deposit_multi<false, 2>"): (95600 * 2 + 55917) * 1024,
                   (5, "k_deposit_multi<false, 4>"): (88597 * 2 + 88264) * 1024,
                   (5, "k_deposit_multi<false, 8>"): (90718 * 2 + 87264) * 1024}
WORKLOADS = {
    1: "32^3 jittered lattice, H-only, isothermal 1e4 K, one source",
    2: "128^3 lognormal, H-only, isothermal 1e4 K, 64 sources",
    3: "256^3 Zel'dovich, H+He, temperature evolution, 16-bin blackbody, 1024 sources",
    5: "96 batched lognormal fields n^3, n in {8..64}, H-only, isothermal",
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def physical_cores():
    """physical cores from lscpu (unique core, socket pairs), None without lscpu"""
    import subprocess
    try:
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                             timeout=20).stdout
    except Exception:
        return None
    pairs = {ln for ln in out.splitlines() if ln and not ln.startswith("#")}
    return len(pairs) or None


def cpu_threads():
    """threads for the CPU baseline: the physical cores, capped by this
    process's CPU share (affinity) and OMP_NUM_THREADS when it is set"""
    share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    n = physical_cores() or share
    n = min(n, share or n)
    if os.environ.get("OMP_NUM_THREADS"):
        n = min(n, int(os.environ["OMP_NUM_THREADS"]))
    return max(1, n)


def cpu_baseline(wl, ray_ticks, seconds_target=15.0):
    """the CPU oracle (compiled -O3 -fopenmp -ffp-contract=off) on the host
    cores: full ticks (trace+sort, deposit, merge, update) over a bounded
    sample of the bench's own ray lists"""
    from oracle import pyoracle as po
    nth = cpu_threads()
    u = wl.physics.units
    f = wl.physics.flags
    # the workload's own run flags (H + He with temperature for the 256^3 field)
    of = po.OrFlags(f.isothermal, f.adiabatic, f.on_the_spot, f.h_only, f.secondary_ionization,
                    f.trace_only, f.no_photon, f.comoving, f.redshift, f.fake_temperature)
    orc = po.Oracle(po.units(hmf=u.C_HMF, hemf=u.C_HEMF), of,
                    rates=wl.physics.rates_path, kernel=po.reference_kernel_table())
    t0 = time.perf_counter()
    tree = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    t_tree = time.perf_counter() - t0
    ps = wl.ps.copy()
    # untimed: the oracle's one-time set-up on its first tick
    po.tick(orc, tree, ps.copy(), ray_ticks[0][:8], 1, wl.tick_time, nthreads=nth)
    nsamp = 64
    total_hits, total_t, tdep_t = 0, 0.0, 0.0
    used, ticks_used = 0, 0
    # batches of the bench's tick-1 ray list, then of its tick-2 list and so
    # on (each batch a first tick from the initial state), until the target
    for t, rays in enumerate(ray_ticks, start=1):
        pos = 0
        while pos < len(rays) and total_t < seconds_target:
            sub = rays[pos:pos + nsamp]
            ps2 = ps.copy()
            t0 = time.perf_counter()
            rc, st = po.tick(orc, tree, ps2, sub, 1, wl.tick_time, nthreads=nth)
            dt = time.perf_counter() - t0
            total_hits += st.total_deposit_count
            total_t += dt
            tdep_t += st.trace_time + st.deposit_time
            used += len(sub)
            pos += len(sub)
            nsamp = min(nsamp * 2, 1024)
        ticks_used = t
        if total_t >= seconds_target:
            break
    # one thread on the first 64 rays, for scale
    ps1 = ps.copy()
    t0 = time.perf_counter()
    rc, st1 = po.tick(orc, tree, ps1, ray_ticks[0][:64], 1, wl.tick_time, nthreads=1)
    one = st1.total_deposit_count / (time.perf_counter() - t0)
    span = "the tick-1 list" if ticks_used <= 1 else f"the lists of ticks 1..{ticks_used}"
    return {
        "value": total_hits / total_t,
        "unit": "intersections/s",
        "cores": nth,
        "kind": "port",
        "sample": f"{used} rays taken in order from {span} of the bench ({len(ray_ticks[0])} rays per "
                  f"tick), in batches of 64..1024, each batch a first tick from the initial state of "
                  f"the full {wl.ps.npar}-particle field; full ticks (trace+sort, deposit, merge, "
                  f"update; oracle tree build {t_tree:.1f}s excluded), {total_t:.1f}s of CPU work",
        "trace_deposit_value": total_hits / max(tdep_t, 1e-12),
        "one_thread_value": one,
        "lscpu_physical_cores": physical_cores(),
        "host_cpus": os.cpu_count(),
    }


MLP_FLOP_PER_ROW = 2 * (22 * 64 + 64 * 64 + 64 * 64 + 64 * 4)   # 19,712
BF16_PEAK_TFLOPS = 2500.0


def mlp_forward_only(n):
    """the network's forward kernel alone (k_mlp_forward) on n random rows;
    prints the rate from the wall time of the calls (host copies included)
    -- the kernel's own time comes from rocprofv3"""
    import psphray_amd as psr
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1, (n, 22)).astype(np.float32)
    x[:, 1:4] = rng.uniform(0, 1, (n, 3))
    with psr.Context(0) as c:
        c.mlp_set_params(rng.normal(0, 0.1, psr._abi.lib().psr_mlp_param_count()).astype(np.float32))
        c.mlp_set_norm(np.zeros(22, np.float32), np.ones(22, np.float32))
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            c.mlp_forward(x)
            best = min(best, time.perf_counter() - t0)
    print(json.dumps({"mlp_forward_rows": n, "flop_per_row": MLP_FLOP_PER_ROW,
                      "best_call_ms_incl_copies": best * 1e3,
                      "tflops_incl_copies": n * MLP_FLOP_PER_ROW / best / 1e12}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10,
                    help="untimed ticks; more follow until the autotuner has decided (four trials "
                         "for the deposit accumulation, then four for graph vs kernel-by-kernel)")
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--n", type=int, default=None, help="override lattice size")
    ap.add_argument("--nray", type=int, default=None)
    ap.add_argument("--chem", choices=["both", "exact", "surrogate"], default="both",
                    help="both (default): exact ticks are timed (the line's value), then the same "
                         "ticks from the same state with the MLP surrogate; the network is trained "
                         "first on tuples the exact solver writes")
    ap.add_argument("--train-ticks", type=int, default=8,
                    help="exact ticks whose steps train the surrogate")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--mlp-residual", type=float, default=None,
                    help="surrogate residual threshold (library default 1e-4)")
    ap.add_argument("--accum", choices=["auto", "atomic", "binned"], default="auto",
                    help="deposit accumulation (auto: timed on the first ticks)")
    ap.add_argument("--walk", choices=["auto", "single", "multi", "multi2", "multi8"], default="auto",
                    help="deposit: one ray per 3-wave workgroup, or four (two, eight) rays per 4-wave "
                         "workgroup (auto: timed)")
    ap.add_argument("--surrogate-scope", choices=["all", "stiff"], default="stiff",
                    help="surrogate chemistry for every active particle, or only for the particles the "
                         "exact solver's first attempt leaves unfinished")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mlp-forward-rows", type=int, default=0,
                    help="only time the surrogate network's forward kernel on this many random "
                         "feature rows (random weights) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run plain device copies of a known size (scale for the PMC "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "design (ray sharding + reduce-scatter of deposits) is described, not built",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        nbytes = 512 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": nbytes, "bytes_written": nbytes, "dispatches": 5}))
        return
    if args.mlp_forward_rows:
        mlp_forward_only(args.mlp_forward_rows)
        return

    import psphray_amd as psr
    from psphray_amd.fields import make_workload

    rates = os.path.join(ROOT, "tests", "golden", "atomic_rates_Hui.txt")
    t0 = time.perf_counter()
    wl = make_workload(args.config, rates, n=args.n, nray=args.nray)
    log(f"workload {wl.name}: {wl.ps.npar} particles, {wl.nray} rays/tick "
        f"({time.perf_counter() - t0:.1f}s)")
    exact_timed = args.chem in ("both", "exact")
    surrogate = args.chem in ("both", "surrogate")
    K = args.steps
    ntrain = args.train_ticks if surrogate else 0
    tune_extra = 48      # ticks allowed beyond --warmup for the autotuner's trials (restarted trials included)
    # ray lists of every tick the run may need, resident before any timing
    nt = args.warmup + tune_extra + ntrain + 3 * K
    ray_ticks = [wl.rays(t) for t in range(1, nt + 1)]
    firsts = np.cumsum([0] + [len(r) for r in ray_ticks])
    pool = np.concatenate(ray_ticks)

    ctx = psr.Context(0)
    accum_ids = {"auto": psr.PSR_ACCUM_AUTO, "atomic": psr.PSR_ACCUM_ATOMIC, "binned": psr.PSR_ACCUM_BINNED}
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_WALK, {"auto": psr.PSR_WALK_AUTO, "single": psr.PSR_WALK_SINGLE,
                                               "multi": psr.PSR_WALK_MULTI, "multi2": psr.PSR_WALK_MULTI2,
                                               "multi8": psr.PSR_WALK_MULTI8}[args.walk])
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, accum_ids[args.accum])
    ctx.set_physics(wl.physics)
    t0 = time.perf_counter()
    ctx.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)
    ctx.sync()
    nnode, depth, _ = ctx.tree_info()
    log(f"device tree: {nnode} nodes, depth {depth} ({time.perf_counter() - t0:.2f}s incl. upload)")
    ctx.rays_upload(pool)
    next_tick = [1]

    def run(n, mode):
        for _ in range(n):
            t = next_tick[0]
            ctx.tick(t, wl.tick_time, ray_first=int(firsts[t - 1]), nray=len(ray_ticks[t - 1]), chem=mode)
            next_tick[0] += 1

    def decided():
        return (ctx.tuning(psr.PSR_TUNE_DEPOSIT_ACCUM) != psr.PSR_ACCUM_AUTO
                and ctx.tuning(psr.PSR_TUNE_DEPOSIT_WALK) != psr.PSR_WALK_AUTO
                and ctx.tuning(psr.PSR_TUNE_TICK_GRAPH) != -1
                and ctx.tuning(psr.PSR_TUNE_TRACE_THREADS) in (128, 256))

    run(args.warmup, psr.PSR_CHEM_EXACT)
    ctx.sync()
    extra = 0
    # further ticks back to back, as the timed ones run (the tick-graph trials
    # time each tick from the end of the one before), until the autotuner
    # has decided
    while not decided() and extra < tune_extra:
        run(1, psr.PSR_CHEM_EXACT)
        extra += 1
    ctx.sync()
    if not decided():
        raise SystemExit("autotuner undecided after the warm-up ticks; no line is printed")
    warmup_ticks = args.warmup + extra
    accum = {v: k for k, v in accum_ids.items()}[ctx.tuning()]
    walk = {psr.PSR_WALK_SINGLE: "one ray per workgroup", psr.PSR_WALK_MULTI2: "two rays per workgroup",
            psr.PSR_WALK_MULTI: "four rays per workgroup",
            psr.PSR_WALK_MULTI8: "eight rays per workgroup"}[ctx.tuning(psr.PSR_TUNE_DEPOSIT_WALK)]
    graph = {0: "kernel by kernel", 1: "graph"}[ctx.tuning(psr.PSR_TUNE_TICK_GRAPH)]
    ctx_trace_threads = int(ctx.tuning(psr.PSR_TUNE_TRACE_THREADS))
    log(f"{warmup_ticks} warm-up ticks; trace: {ctx_trace_threads} threads per ray; deposit: {walk}, "
        f"{accum} accumulation; ticks: {graph}")

    sur = None
    if surrogate:
        # training tuples from the exact solver, then the network trained on them
        ctx.mlp_capture(min(ntrain * int(wl.ps.npar), 40_000_000))
        run(ntrain, psr.PSR_CHEM_EXACT)
        ctx.sync()
        ntup, _, _ = ctx.mlp_captured()
        t0 = time.perf_counter()
        loss = ctx.mlp_train(n=0, epochs=args.epochs, batch=8192, lr=2e-3, seed=1)
        t_train = time.perf_counter() - t0
        ctx.mlp_capture(0)
        if args.mlp_residual is not None:
            ctx.mlp_set_residual(args.mlp_residual)
        log(f"surrogate: {ntup} tuples, {args.epochs} epochs in {t_train:.2f}s, loss {loss:.3e}")
        sur = {"tuples": int(ntup), "epochs": args.epochs, "train_s": t_train, "train_loss": loss,
               "residual": args.mlp_residual if args.mlp_residual is not None else 1e-4}
    snap = wl.ps.copy()
    ctx.state_download(snap)
    ctx.sync()
    t_first = next_tick[0]

    def timed(mode):
        """K ticks from the snapshot state; (wall s, stats, graph launches)"""
        next_tick[0] = t_first
        ctx.state_upload(snap)
        ctx.sync()
        ctx.reset_stats()
        ctx.sync()
        g0 = ctx.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES)
        t0 = time.perf_counter()
        run(K, mode)
        ctx.sync()
        wall = time.perf_counter() - t0
        return wall, ctx.stats(), ctx.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES) - g0

    smode = psr.PSR_CHEM_SURROGATE_STIFF if args.surrogate_scope == "stiff" else psr.PSR_CHEM_SURROGATE
    res = {}
    if exact_timed:
        wall, st, gl = timed(psr.PSR_CHEM_EXACT)
        ex = wl.ps.copy()
        ctx.state_download(ex)
        res["exact"] = (wall, st, gl)
        log(f"exact: {K} ticks in {wall * 1e3:.2f} ms")
    if surrogate:
        wall_s, st_s, gl_s = timed(smode)
        sg = wl.ps.copy()
        ctx.state_download(sg)
        res["surrogate"] = (wall_s, st_s, gl_s)
        log(f"surrogate: {K} ticks in {wall_s * 1e3:.2f} ms")
        sur.update({"scope": args.surrogate_scope, "ticks_per_s": K / wall_s, "ms_per_step": wall_s / K * 1e3,
                    "intersections_per_s": st_s["total_deposit_count"] / wall_s,
                    "accepted": int(st_s["surrogate_count"]),
                    "fallbacks": int(st_s["surrogate_fallbacks"]),
                    "active": int(st_s["active_count"])})
        if exact_timed:
            dx = np.abs(sg.lambdaH - ex.lambdaH)
            sur.update({"x_HII_err_max": float(dx.max()), "x_HII_err_p99": float(np.percentile(dx, 99)),
                        "tick_speedup_vs_exact": res["exact"][0] / wall_s})
            if not wl.physics.flags.h_only:
                for f in ("lambdaHeI", "lambdaHeII"):
                    sur[f"{f}_err_max"] = float(np.abs(getattr(sg, f) - getattr(ex, f)).max())
                sur["T_rel_err_p99"] = float(np.percentile(np.abs(sg.ie / ex.ie - 1), 99))
        log(f"surrogate: {sur}")
    main_mode = "exact" if exact_timed else "surrogate"
    wall, st, graph_launches = res[main_mode]
    # per-phase pass: HIP events on the context stream around each kernel
    # (they synchronise every tick and keep its phases in order, so they are
    # kept out of the timed run)
    next_tick[0] = t_first
    ctx.state_upload(snap)
    ctx.reset_stats()
    ctx.set_timing(True)
    run(K, psr.PSR_CHEM_EXACT if exact_timed else smode)
    ctx.sync()
    tm = ctx.timers()
    st_t = ctx.stats()
    ctx.set_timing(False)
    inter = st["total_deposit_count"]
    value = inter / wall
    td_ms = tm["trace"] + tm["sort"] + tm["deposit"]
    # dominant kernel: the deposit; its share of the per-intersection bytes
    he = not wl.physics.flags.h_only
    bytes_dep = BYTES_DEPOSIT_HE if he else BYTES_DEPOSIT_H
    dep_launch_s = tm["deposit"] / 1e3 / K
    roof_achieved = st_t["total_deposit_count"] / K * bytes_dep / dep_launch_s / 1e9
    # the deposit kernel the autotuner chose; traffic is the counters' figure
    # for that kernel on this workload, where measured
    he_t = "true" if he else "false"
    if walk == "one ray per workgroup":
        dep_kernel = f"k_deposit<{he_t}, {'true' if accum == 'binned' else 'false'}>"
    else:
        mr = {"two rays per workgroup": 2, "four rays per workgroup": 4, "eight rays per workgroup": 8}[walk]
        dep_kernel = f"k_deposit_multi<{he_t}, {mr}>"
    out = {
        "metric": "ray-particle intersections/s + ticks/s, exact FP64 chemistry (value) vs "
                  "MLP-surrogate chemistry (surrogate) at <=1e-3 x_HII error",
        "value": value,
        "unit": "intersections/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 geometry / f64 deposit+chemistry",
        "data": "synthetic (seeded generator: jittered or Zel'dovich lattice, lognormal density, "
                "Poisson sources; counter-based ray lists)",
        "config": {
            "workload": f"{wl.name}: {wl.ps.npar} particles ({WORKLOADS.get(args.config, '')}), "
                        f"{wl.nray} rays/tick, {main_mode} chemistry timed for the value",
            "particles": int(wl.ps.npar),
            "rays_per_tick": int(wl.nray),
            "chemistry": main_mode,
        },
        "ticks_per_s": K / wall,
        "warmup_ticks_run": warmup_ticks,
        "intersections_per_tick": inter / K,
        "trace_deposit_intersections_per_s": st_t["total_deposit_count"] / (td_ms / 1e3) if td_ms > 0 else None,
        "kernel_ms_per_tick": {k: v / K for k, v in tm.items()},
        "active_per_tick": st["active_count"] / K,
        "deposit_walked_per_tick": st["deposit_processed"] / K,
        "solver_steps_per_update": st["solver_steps"] / max(st["active_count"], 1),
        "deposit_accumulation": accum,
        "deposit_walk": walk,
        "trace_threads_per_ray": ctx_trace_threads,
        "tick_launch": graph,
        # kernel-by-kernel ticks run each update beside the next tick's trace
        # and sort (the per-phase pass above runs them one after another)
        "tick_overlap": graph == "kernel by kernel" and ctx.tuning(psr.PSR_TUNE_TICK_OVERLAP) == 1,
        "graph_launches_timed": graph_launches,
        "surrogate": sur,
        "roofline": {
            "kernel": dep_kernel,
            "bound": "hbm",
            "achieved": roof_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": roof_achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC_DEPOSIT.get((args.config, dep_kernel)),
            "bytes_per_intersection": bytes_dep,
            "launch_us": dep_launch_s * 1e6,
        },
    }
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)...")
        out["cpu_baseline"] = cpu_baseline(wl, ray_ticks)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
