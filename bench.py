"""Benchmark: one tick (trace + sort + deposit + merge + exact FP64 chemistry
update) of workload 2 (128^3 particles, lognormal density, 64 sources,
4096 rays per tick) on one MI355X.

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
# HBM bytes per k_deposit launch from rocprofv3 FETCH_SIZE + WRITE_SIZE
# passes, scaled by the same counters on a copy of known size (FETCH_SIZE x 2,
# WRITE_SIZE x 1), the timed-tick dispatch of each workload (the 128^3,
# 256^3 and batched-field traffic summaries under profiles/)
TRAFFIC_DEPOSIT = {2: (84968 * 2 + 50806) * 1024, 3: (5984385 * 2 + 4237966) * 1024,
                   5: (151006 * 2 + 92589) * 1024}
WORKLOADS = {
    1: "32^3 jittered lattice, H-only, isothermal 1e4 K, one source",
    2: "128^3 lognormal, H-only, isothermal 1e4 K, 64 sources",
    3: "256^3 Zel'dovich, H+He, temperature evolution, 16-bin blackbody, 1024 sources",
    5: "96 batched lognormal fields n^3, n in {8..64}, H-only, isothermal",
}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(wl, rays, seconds_target=15.0):
    """the CPU oracle on the host cores: full ticks (trace+sort, deposit,
    merge, update) over a bounded sample of the tick's rays"""
    from oracle import pyoracle as po
    nth = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    u = wl.physics.units
    f = wl.physics.flags
    # the workload's own run flags (H + He with temperature for the 256^3 field)
    of = po.OrFlags(f.isothermal, f.adiabatic, f.on_the_spot, f.h_only, f.secondary_ionization,
                    f.trace_only, f.no_photon, f.comoving, f.redshift, f.fake_temperature)
    orc = po.Oracle(po.units(hmf=u.C_HMF, hemf=u.C_HEMF), of,
                    rates=wl.physics.rates_path, kernel=wl.physics.kernel_table())
    t0 = time.perf_counter()
    tree = po.Tree(wl.ps.pos, wl.ps.sml, wl.boxsize)
    t_tree = time.perf_counter() - t0
    ps = wl.ps.copy()
    # untimed: the oracle's one-time set-up on its first tick
    po.tick(orc, tree, ps.copy(), rays[:8], 1, wl.tick_time, nthreads=nth)
    nsamp = 64
    total_hits, total_t, tdep_t = 0, 0.0, 0.0
    used = 0
    # the first tick's rays, then (for a short tick) the next ticks' rays,
    # each batch a first tick from the initial state, until the time target
    pool, tick_rays, ticks_used = rays, 1, 1
    while total_t < seconds_target:
        if used >= len(pool):
            if tick_rays >= 200:
                break
            tick_rays += 1
            pool = np.concatenate([pool, wl.rays(tick_rays)])
            ticks_used = tick_rays
        sub = pool[used:used + nsamp]
        ps2 = ps.copy()
        t0 = time.perf_counter()
        rc, st = po.tick(orc, tree, ps2, sub, 1, wl.tick_time, nthreads=nth)
        dt = time.perf_counter() - t0
        total_hits += st.total_deposit_count
        total_t += dt
        tdep_t += st.trace_time + st.deposit_time
        used += len(sub)
        nsamp = min(nsamp * 2, 1024)
    # one thread on the first 64 rays, for scale
    ps1 = ps.copy()
    t0 = time.perf_counter()
    rc, st1 = po.tick(orc, tree, ps1, rays[:64], 1, wl.tick_time, nthreads=1)
    one = st1.total_deposit_count / (time.perf_counter() - t0)
    return {
        "value": total_hits / total_t,
        "unit": "intersections/s",
        "cores": nth,
        "kind": "port",
        "sample": f"{used} rays ({len(rays)} per tick, ticks 1..{ticks_used} of the ray lists, each batch "
                  f"from the initial state) on the full {wl.ps.npar}-particle field, full ticks "
                  f"(trace+sort, deposit, merge, update; tree build {t_tree:.1f}s excluded), "
                  f"{total_t:.1f}s of CPU work",
        "trace_deposit_value": total_hits / max(tdep_t, 1e-12),
        "one_thread_value": one,
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
                    help="untimed ticks (they also run the autotuner's trials: four for the deposit "
                         "accumulation, then four for graph vs kernel-by-kernel ticks)")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="override lattice size")
    ap.add_argument("--nray", type=int, default=None)
    ap.add_argument("--chem", choices=["exact", "surrogate"], default="exact",
                    help="surrogate: exact ticks write training tuples, the network is trained, then "
                         "the same ticks are timed with the exact solver and with the surrogate")
    ap.add_argument("--train-ticks", type=int, default=8,
                    help="exact ticks whose steps train the surrogate (--chem surrogate)")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--mlp-residual", type=float, default=None,
                    help="surrogate residual threshold (--chem surrogate; library default 1e-4)")
    ap.add_argument("--accum", choices=["auto", "atomic", "binned"], default="auto",
                    help="deposit accumulation (auto: timed on the first ticks)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mlp-forward-rows", type=int, default=0,
                    help="only time the surrogate network's forward kernel on this many random "
                         "feature rows (random weights) and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run plain device copies of a known size (scale for the PMC "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
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
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "design (ray sharding + reduce-scatter of deposits) is described, not built",
              file=sys.stderr)
        sys.exit(2)

    import psphray_amd as psr
    from psphray_amd.fields import make_workload

    rates = os.path.join(ROOT, "tests", "golden", "atomic_rates_Hui.txt")
    t0 = time.perf_counter()
    wl = make_workload(args.config, rates, n=args.n, nray=args.nray)
    log(f"workload {wl.name}: {wl.ps.npar} particles, {wl.nray} rays/tick "
        f"({time.perf_counter() - t0:.1f}s)")
    # warmup, the timed ticks, then as many again with per-phase event timers
    # (those synchronise every tick, so they are kept out of the timed run);
    # the surrogate run adds the training ticks after the warmup
    surrogate = args.chem == "surrogate"
    ntrain = args.train_ticks if surrogate else 0
    nt = args.warmup + ntrain + 2 * args.steps
    ray_ticks = [wl.rays(t) for t in range(1, nt + 1)]
    firsts = np.cumsum([0] + [len(r) for r in ray_ticks])
    pool = np.concatenate(ray_ticks)

    ctx = psr.Context(0)
    accum_ids = {"auto": psr.PSR_ACCUM_AUTO, "atomic": psr.PSR_ACCUM_ATOMIC, "binned": psr.PSR_ACCUM_BINNED}
    ctx.set_tuning(psr.PSR_TUNE_DEPOSIT_ACCUM, accum_ids[args.accum])
    ctx.set_physics(wl.physics)
    t0 = time.perf_counter()
    ctx.load(wl.ps, wl.field_offsets, wl.boxsizes, boxsize=wl.boxsize)
    ctx.sync()
    nnode, depth, _ = ctx.tree_info()
    log(f"device tree: {nnode} nodes, depth {depth} ({time.perf_counter() - t0:.2f}s incl. upload)")
    ctx.rays_upload(pool)
    chem = psr.PSR_CHEM_EXACT if args.chem == "exact" else psr.PSR_CHEM_SURROGATE

    def run(t0_, t1_, mode):
        for t in range(t0_, t1_ + 1):
            ctx.tick(t, wl.tick_time, ray_first=int(firsts[t - 1]), nray=len(ray_ticks[t - 1]), chem=mode)

    run(1, args.warmup, psr.PSR_CHEM_EXACT)
    ctx.sync()
    sur = None
    t_first = args.warmup + ntrain + 1
    t_end = args.warmup + ntrain + args.steps
    if surrogate:
        # training tuples from the exact solver, then the network trained on them
        ctx.mlp_capture(min(ntrain * int(wl.ps.npar), 40_000_000))
        run(args.warmup + 1, args.warmup + ntrain, psr.PSR_CHEM_EXACT)
        ctx.sync()
        ntup, _, _ = ctx.mlp_captured()
        t0 = time.perf_counter()
        loss = ctx.mlp_train(n=0, epochs=args.epochs, batch=8192, lr=2e-3, seed=1)
        t_train = time.perf_counter() - t0
        ctx.mlp_capture(0)
        if args.mlp_residual is not None:
            ctx.mlp_set_residual(args.mlp_residual)
        log(f"surrogate: {ntup} tuples, {args.epochs} epochs in {t_train:.2f}s, loss {loss:.3e}")
        # the plain method over the timed ticks, from the same state
        snap = wl.ps.copy()
        ctx.state_download(snap)
        ctx.sync()
        t0 = time.perf_counter()
        run(t_first, t_end, psr.PSR_CHEM_EXACT)
        ctx.sync()
        wall_exact = time.perf_counter() - t0
        ex = wl.ps.copy()
        ctx.state_download(ex)
        ctx.state_upload(snap)
        sur = {"tuples": int(ntup), "epochs": args.epochs, "train_s": t_train, "train_loss": loss,
               "residual": args.mlp_residual if args.mlp_residual is not None else 1e-4,
               "exact_ms_per_step": wall_exact / args.steps * 1e3}
    ctx.sync()
    ctx.reset_stats()
    ctx.sync()
    accum = {v: k for k, v in accum_ids.items()}[ctx.tuning()]
    graph = {-1: "undecided", 0: "kernel by kernel", 1: "graph"}[ctx.tuning(psr.PSR_TUNE_TICK_GRAPH)]
    g0 = ctx.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES)
    log(f"deposit accumulation: {accum}; ticks: {graph}")
    t0 = time.perf_counter()
    run(t_first, t_end, chem)
    ctx.sync()
    wall = time.perf_counter() - t0
    log(f"{args.steps} ticks in {wall * 1e3:.2f} ms")
    st = ctx.stats()
    ctx_graph_launches = ctx.tuning(psr.PSR_TUNE_GRAPH_LAUNCHES) - g0
    if surrogate:
        sg = wl.ps.copy()
        ctx.state_download(sg)
        dx = np.abs(sg.lambdaH - ex.lambdaH)
        sur.update({"x_HII_err_max": float(dx.max()), "x_HII_err_p99": float(np.percentile(dx, 99)),
                    "accepted": int(st["surrogate_count"]), "fallbacks": int(st["surrogate_fallbacks"]),
                    "tick_speedup_vs_exact": sur["exact_ms_per_step"] / (wall / args.steps * 1e3)})
        if not wl.physics.flags.h_only:
            for f in ("lambdaHeI", "lambdaHeII"):
                sur[f"{f}_err_max"] = float(np.abs(getattr(sg, f) - getattr(ex, f)).max())
            sur["T_rel_err_p99"] = float(np.percentile(np.abs(sg.ie / ex.ie - 1), 99))
        log(f"surrogate: {sur}")
    # per-phase pass: HIP events on the context stream around each kernel
    ctx.reset_stats()
    ctx.set_timing(True)
    run(t_end + 1, nt, chem)
    ctx.sync()
    tm = ctx.timers()
    st_t = ctx.stats()
    ctx.set_timing(False)
    K = args.steps
    inter = st["total_deposit_count"]
    value = inter / wall
    td_ms = tm["trace"] + tm["sort"] + tm["deposit"]
    hits_per_tick = inter / K
    # dominant kernel: the deposit; its share of the per-intersection bytes
    he = not wl.physics.flags.h_only
    bytes_dep = BYTES_DEPOSIT_HE if he else BYTES_DEPOSIT_H
    dep_launch_s = tm["deposit"] / 1e3 / K
    roof_achieved = st_t["total_deposit_count"] / K * bytes_dep / dep_launch_s / 1e9
    out = {
        "metric": "ray-particle intersections/s (whole tick, exact FP64 chemistry)" if not surrogate else
                  "ray-particle intersections/s (whole tick, MLP-surrogate chemistry, exact FP64 fallback)",
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
        "data": "synthetic (seeded generator: jittered lattice, lognormal density, Poisson sources; "
                "counter-based ray lists)",
        "config": {
            "workload": f"{wl.name}: {wl.ps.npar} particles ({WORKLOADS.get(args.config, '')}), "
                        f"{wl.nray} rays/tick, {args.chem} chemistry",
            "particles": int(wl.ps.npar),
            "rays_per_tick": int(wl.nray),
            "chemistry": args.chem,
        },
        "ticks_per_s": K / wall,
        "intersections_per_tick": hits_per_tick,
        "trace_deposit_intersections_per_s": inter / (td_ms / 1e3) if td_ms > 0 else None,
        "kernel_ms_per_tick": {k: v / K for k, v in tm.items()},
        "active_per_tick": st["active_count"] / K,
        "deposit_walked_per_tick": st["deposit_processed"] / K,
        "solver_steps_per_update": st["solver_steps"] / max(st["active_count"], 1),
        "deposit_accumulation": accum,
        "tick_launch": graph,
        "graph_launches_timed": ctx_graph_launches,
        "surrogate": sur,
        "roofline": {
            "kernel": "k_deposit",
            "bound": "hbm",
            "achieved": roof_achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": roof_achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC_DEPOSIT.get(args.config),
            "bytes_per_intersection": bytes_dep,
            "launch_us": dep_launch_s * 1e6,
        },
    }
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on host cores)...")
        out["cpu_baseline"] = cpu_baseline(wl, ray_ticks[0])
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
