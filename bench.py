#!/usr/bin/env python3
"""Benchmark of the global-TGV hot path on one MI355X.

Workload: the depth-10 scene -- torus (R=1, r=0.3) union a box rotated 30
degrees, 96 seeded views of 1280x960 on the radius-3 sphere, depth noise
sigma 0.003, 3% outliers, octree clamped at depth 10 (about 6 M leaves), 10
histogram bins, alpha1 = 1, alpha0 = 2, lambda = 1, FP32 state.

Step: one coarse-to-fine sweep over every level of that octree with 50
primal-dual iterations per level (the hot path's unit of work: reset, parent
initialisation, 50 iterations and one FP64 gap evaluation per level), inputs
resident in HBM. value = leaf-iterations / s over the K timed steps.

Also reported: the roofline of the iteration at the finest level (measured
live with HIP events on the library's stream); the time to relative duality
gap <= 1e-4 of the plain cold loop with the reference's steps, of the cold
loop with steps sigma / tau = 1024, with the dual recovered from the initial
field, and warm-started by the trained network; the same sweep with the
reference's own update rules (the drop-in tgv_minimize's iteration); and the
CPU restatement of the reference's run_level timed on this host (one core,
and the process's cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ELEM = {"fp32": 4, "bf16": 2, "fp64": 8}
# algorithmic bytes per leaf-iteration of the iteration, whatever kernels run
# it: the 17 state values (u, v[3], p[3], q[6], u_bar, v_bar[3]) read and
# written once, (W, m) = 8 B and 6 neighbour indices; neighbour re-reads are
# not counted. The fused kernels keep u_bar, v_bar on chip and move fewer
# bytes than this (26 state values, a 4 B info word), so for them the
# figure is the work done, not the traffic.
ALG_BYTES = {k: 34 * e + 8 + 24 for k, e in ELEM.items()}
# p, q in BF16, u, v and the extrapolations in FP32: 16 values of 4 B and 18 of 2 B
ALG_BYTES["bf16-dual"] = 16 * 4 + 18 * 2 + 8 + 24
KERNEL_NAMES = {0: "k_dual + k_primal_adj", 1: "k_fused_tile", 2: "k_fused_adj",
                3: "k_fused_tile (p, q staged in LDS)", 4: "k_fused_halo", 5: "k_fused2", 6: "k_level1",
                7: "k_persist", 8: "k_fused4"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp32", "bf16", "bf16-dual", "fp64"], default="fp32")
    ap.add_argument("--views", type=int, default=96)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=960)
    ap.add_argument("--max-depth", type=int, default=10)
    ap.add_argument("--iters", type=int, default=50, help="iterations per level in one step")
    ap.add_argument("--block", type=int, default=0,
                    help="workgroup / tile size (0: 1024 for the two-iteration kernel, 512 for the other tiled "
                         "fused variants, else 256); with the variant, the autotuner's base configuration")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--fused", type=int, choices=[0, 1, 2, 3, 4, 5, 6, 7, 8], default=5,
                    help="0: dual + primal kernels per iteration; 1: LDS-tiled fused kernel; 2: untiled fused "
                         "kernel; 3: LDS-tiled with the tile's p, q staged in LDS; 4: LDS-tiled with a "
                         "tabulated halo; 5: two iterations per launch over tabulated tile sets; 6: levels of at "
                         "most 1024 cubes in one workgroup for all their iterations, larger ones as 5; 7: a level "
                         "in one cooperative launch with grid barriers (levels too large for it as 6); 8: four "
                         "iterations per launch over tabulated tile sets where they fit, else as 5")
    ap.add_argument("--no-fused", action="store_true", help="same as --fused 0")
    ap.add_argument("--no-warm", action="store_true", help="skip the warm-start time-to-gap runs")
    ap.add_argument("--no-oc", action="store_true",
                    help="skip timing the drop-in tgv_minimize_oc (frozen-border treetop leaves)")
    ap.add_argument("--no-octree", action="store_true",
                    help="skip timing the octree build (depth maps -> balanced cubes) on the device and the host")
    ap.add_argument("--step-ratio", type=float, default=1024.0,
                    help="sigma / tau of the tuned to-gap runs (sigma * tau = 1/16 as the reference's 0.25 * 0.25)")
    ap.add_argument("--gap-tol", type=float, default=1e-4)
    ap.add_argument("--no-gap-run", action="store_true")
    ap.add_argument("--no-ref", action="store_true", help="skip the sweep with the reference's update rules")
    ap.add_argument("--cpu-seconds", type=float, default=30.0,
                    help="CPU baseline budget over its three timings (0 = skip)")
    ap.add_argument("--workload", choices=["depth10", "batched", "streaming", "variants"], default="depth10",
                    help="batched: 256 independent blocks of log-uniform size in one problem vs one problem per "
                         "block; streaming: the depth-10 scene's streaming solve resident and out of core; "
                         "variants: per-level iteration time of each launch variant of the depth-10 scene, "
                         "grid-barrier cost, whole sweeps with one variant")
    ap.add_argument("--leaf-cubes", type=int, default=1 << 20, help="streaming: cubes per treetop leaf at most")
    ap.add_argument("--budgets-mb", default="1024,256",
                    help="streaming: device budgets of the out-of-core solve, MiB, comma-separated")
    ap.add_argument("--blocks", type=int, default=256)
    ap.add_argument("--tune", type=int, default=1, help="1: autotune launch configurations before timing")
    ap.add_argument("--batched-gap", type=float, default=0.0,
                    help="batched workload: also time solves to this relative gap (0 = skip)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.no_fused:
        args.fused = 0
    if not args.block:
        args.block = {5: 1024, 1: 512, 3: 512, 4: 512}.get(args.fused, 256)

    if args.calibrate_copy:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = args.calibrate_copy << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
        assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
        for _ in range(3):
            assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
        hip.hipDeviceSynchronize()
        hip.hipFree(a)
        hip.hipFree(b)
        print(json.dumps({"calibrate_copy_bytes": nbytes}))
        return

    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): the multi-GPU path, sharding by treetop leaves, "
              "is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import tgvreconclaude_amd as T
    from tgvreconclaude_amd import scene

    prec = {"fp32": T.FP32, "bf16": T.BF16, "bf16-dual": T.BF16_DUAL, "fp64": T.FP64}[args.precision]
    if args.workload == "batched":
        return batched(args, T, scene, prec)
    if args.workload == "streaming":
        return streaming(args, T, scene)
    if args.workload == "variants":
        return variants(args, T, scene)
    t0 = time.time()
    log(f"[bench] generating scene: {args.views} views {args.width}x{args.height}, depth <= {args.max_depth}")
    sc = scene.torus_box_scene(views=args.views, W=args.width, H=args.height, max_depth=args.max_depth)
    cubes = sc.cubes
    n = len(cubes)
    t_scene = time.time() - t0
    log(f"[bench] scene: {n} leaves in {t_scene:.1f} s")

    ctx = T.Context(0)
    arch, cus = ctx.device_info()

    # the input's stages 1-3 (per-pixel cubes, merge, 2:1 balance) on the
    # device, against the host generator (OpenMP per-view cubes, serial merge
    # and FIFO balance) on the same depth maps; outside the timed TGV steps
    octree = None
    if not args.no_octree:
        t1 = time.time()
        host_cubes = scene.build_octree(sc.depths, sc.cams, args.max_depth)
        t_host = time.time() - t1
        ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, args.max_depth)  # first call allocates
        t1 = time.time()
        dev_cubes, orep = ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, args.max_depth)
        t_dev = time.time() - t1
        same = len(dev_cubes) == len(host_cubes) and bool(np.array_equal(
            np.ascontiguousarray(dev_cubes).view(np.uint8), np.ascontiguousarray(host_cubes).view(np.uint8)))
        octree = {"gpu_ms": round(orep["total_ms"], 2), "gpu_wall_ms": round(t_dev * 1e3, 2),
                  "host_ms": round(t_host * 1e3, 1),
                  "host_threads": int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0)),
                  "speedup": round(t_host / t_dev, 2), "identical_to_host": same, "pixels": orep["pixels"],
                  "raw_cubes": orep["raw"], "merged": orep["merged"], "bridges": orep["bridges"],
                  "balance_rounds": orep["rounds"],
                  "stages_ms": {k: round(orep[k], 2) for k in ("upload_ms", "cubes_ms", "merge_ms", "balance_ms")}}
        log(f"[bench] octree build: device {orep['total_ms']:.1f} ms (upload {orep['upload_ms']:.1f}, cubes "
            f"{orep['cubes_ms']:.1f}, merge {orep['merge_ms']:.1f}, balance {orep['balance_ms']:.1f} in "
            f"{orep['rounds']} rounds), host {t_host * 1e3:.0f} ms, identical {same}")
        del host_cubes, dev_cubes

    t1 = time.time()
    hist = ctx.histograms(cubes, sc.views(), sc.origin, sc.r_root, 10)
    t_hist = time.time() - t1
    log(f"[bench] histograms on the GPU: {t_hist * 1e3:.1f} ms (incl. upload of {sc.depths.nbytes / 1e6:.0f} MB)")

    t1 = time.time()
    ctx.load(cubes, hist, precision=prec, op_mode=T.OP_ADJ)
    t_load = time.time() - t1
    levels = ctx.levels()
    log(f"[bench] problem resident in {t_load:.2f} s; levels {levels}")

    variant = 0 if args.no_fused else args.fused
    # the step is the reference's tgv_minimize: each level's iterations, no
    # energy evaluation (the per-level P, D report is timed separately below)
    prm = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_ADJ, iters=args.iters,
                      use_graph=0 if args.no_graph else 1, block=args.block, fused=variant, no_energy=1)
    tuning = None
    if args.tune:
        tuning = ctx.autotune(prm, iters=args.iters)  # the step's own chunk of iterations per level
        prm.tune = 1
        log(f"[bench] autotuned (depth, variant, block, xcd, us): "
            f"{[(l['depth'], l['variant'], l['block'], l['xcd'], round(l['best_us'], 1)) for l in tuning]}")
    for i in range(args.warmup):
        ctx.run(prm)
    log(f"[bench] {args.warmup} warmup steps done")
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    ts = time.perf_counter()
    leaf_iters = 0.0
    for i in range(args.steps):
        rep = ctx.run(prm)
        leaf_iters += rep.leaf_iters
    hip.hipDeviceSynchronize()
    dt = time.perf_counter() - ts
    value = leaf_iters / dt
    log(f"[bench] {args.steps} steps in {dt:.3f} s: {value:.4g} leaf-iter/s")
    # the same steps with each level's FP64 energy and dual value evaluated
    # after its iterations (this library's report; the reference has none)
    prm.no_energy = 0
    ctx.run(prm)
    hip.hipDeviceSynchronize()
    t1 = time.perf_counter()
    li_e = 0.0
    for i in range(args.steps):
        li_e += ctx.run(prm).leaf_iters
    hip.hipDeviceSynchronize()
    dt_e = time.perf_counter() - t1
    with_energy = {"value": li_e / dt_e, "unit": "leaf-iter/s", "ms_per_step": round(dt_e / args.steps * 1e3, 4),
                   "definition": "the same steps with each level's FP64 energy P and dual value D evaluated "
                                 "after its iterations"}
    log(f"[bench] with the per-level energy report: {li_e / dt_e:.4g} leaf-iter/s")
    prm.no_energy = 1
    # where a step's time goes: each level's iterations (events around its
    # launches, per-level graphs so that the events sit between them)
    ug = prm.use_graph
    if ug:
        prm.use_graph = 2
    rep = ctx.run(prm)
    prm.use_graph = ug
    level_ms = [(l["depth"], round(l["ms"], 4)) for l in rep.levels()]
    # the same leaf-iterations with the one-time setup counted in: histogram
    # voting on the GPU (with the depth maps' upload) and the problem load
    setup = t_hist + t_load
    with_setup = {"value": leaf_iters / (dt + setup), "unit": "leaf-iter/s", "setup_s": round(setup, 3),
                  "definition": f"leaf-iterations of the {args.steps} timed steps / (their time + histogram "
                                f"voting + problem load)"}

    # roofline of the iteration at the finest level: HIP events on the library's stream
    finest, n_f = levels[-1]
    total_ms, dual_ms, primal_ms, fused_ms = ctx.level_bench(finest, prm, 100)
    if tuning:
        variant = tuning[-1]["variant"]
    fused = variant != 0
    tile = tuning[-1]["block"] if tuning else args.block
    xcd = tuning[-1]["xcd"] if tuning else 1
    if variant in (6, 7, 8):
        # the variants for small levels time the finest level's fused pass as
        # the two-iteration kernel they fall back to there
        variant, tile, xcd = 5, (1024 if n_f >= 1 << 18 else 512), 1
    # fused_ms is per iteration; the two-iteration kernel's launch is two
    two = variant == 5 and ctx.tile_sets(finest, tile)[0]
    if variant == 5 and not two:
        variant = 1  # the level's tile sets do not fit: one-iteration tiled passes
    it_ms = fused_ms if fused else dual_ms + primal_ms
    kernel = f"{KERNEL_NAMES[variant]}, depth {finest}, {n_f} leaves" + (
        " (one launch = two iterations)" if two else (" (one launch = one iteration)" if fused else ""))
    if fused:
        kernel += f", workgroup {tile}, {'XCD-contiguous' if xcd else 'plain'} tile order"
    ab = ALG_BYTES[args.precision]
    achieved = ab * n_f / (it_ms * 1e-3) / 1e9
    log(f"[bench] depth {finest} ({n_f} leaves): fused {fused_ms * 1e3:.1f} us; two-kernel dual {dual_ms * 1e3:.1f} us "
        f"+ primal {primal_ms * 1e3:.1f} us; 100 its {total_ms:.2f} ms -> {achieved:.0f} GB/s algorithmic")

    # time to relative gap <= tol (exact-adjoint mode): the plain cold loop
    # with the reference's steps (sigma = tau = 0.25, p = q = 0) is the
    # baseline; then the same cold loop with steps sigma / tau = step_ratio,
    # with the dual recovered from the initial field, and with the trained
    # network (committed weights) and the per-level fallback
    gap, warm = None, None
    wpath = os.path.join(ROOT, "tgvreconclaude_amd", "weights", "warmstart_mlp.npy")
    if not args.no_gap_run:
        sg, ta = 0.25 * args.step_ratio ** 0.5, 0.25 / args.step_ratio ** 0.5
        # per-depth steps: sigma / tau = 256 up to four levels above the finest,
        # 1024 for the next three, 4096 at the finest (sigma * tau = 1/16)
        top = max(d for d, _ in levels)
        ratios = [256 if d <= top - 4 else (1024 if d < top else 4096) for d in range(top + 1)]
        sd, td = T.step_schedule(ratios)
        sched = dict(sigma=sg, tau=ta, sigma_depth=sd, tau_depth=td)
        runs = [("cold_reference_steps", {}), ("cold", dict(sigma=sg, tau=ta)),
                ("cold_dual", dict(sigma=sg, tau=ta, dual_eps=1e-4)),
                ("cold_sched", dict(sched)), ("cold_dual_sched", dict(sched, dual_eps=1e-4))]
        if not args.no_warm and os.path.exists(wpath):
            ctx.mlp_set(np.load(wpath))
            runs.append(("warm", dict(sigma=sg, tau=ta, dual_eps=1e-4, warm_start=1, warm_fallback=1)))
            runs.append(("warm_sched", dict(sched, dual_eps=1e-4, warm_start=1, warm_fallback=1)))
        gap = {"tol": args.gap_tol, "step_ratio": args.step_ratio, "step_ratio_per_depth": ratios,
               "setup_ms": round(setup * 1e3, 2)}
        for name, kw in runs:
            pw = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_ADJ, gap_tol=args.gap_tol,
                             chunk=50, max_iters=20000, use_graph=prm.use_graph, block=args.block, fused=prm.fused,
                             tune=prm.tune, **kw)
            rw = ctx.run(pw)
            lv = rw.levels()
            gap[name] = {"wall_ms": round(rw.solve_ms, 2), "iterate_ms": round(rw.iterate_ms, 2),
                         "gap_eval_ms": round(rw.gap_ms, 2), "leaf_iters": rw.leaf_iters,
                         "iters_per_level": [l["iters"] for l in lv], "warm_levels": [l["warm"] for l in lv],
                         "final_rel_gap": [l["rel_gap"] for l in lv],
                         "reached": all(l["rel_gap"] <= args.gap_tol for l in lv)}
            base = gap["cold_reference_steps"]["wall_ms"]
            gap[name]["speedup_vs_cold_reference_steps"] = round(base / rw.solve_ms, 3)
            if "cold" in gap and name != "cold_reference_steps":
                gap[name]["speedup_vs_cold"] = round(gap["cold"]["wall_ms"] / rw.solve_ms, 3)
            log(f"[bench] {name} to gap {args.gap_tol}: {rw.solve_ms:.1f} ms "
                f"({gap[name]['speedup_vs_cold_reference_steps']}x the reference-step cold loop), "
                f"iters {gap[name]['iters_per_level']}")
        if "warm" in gap:
            # the network's own gain leads: the same steps, with and without the
            # dual recovered from the initial field; against the reference's
            # steps the factor mostly measures the steps
            warm = {"vs_cold_same_steps": gap["warm"]["speedup_vs_cold"],
                    "vs_cold_dual_same_steps": round(gap["cold_dual"]["wall_ms"] / gap["warm"]["wall_ms"], 3),
                    "leaf_iters_vs_cold_same_steps": round(gap["cold"]["leaf_iters"] / gap["warm"]["leaf_iters"], 3),
                    "vs_cold_reference_steps": gap["warm"]["speedup_vs_cold_reference_steps"],
                    "target": 3.0,
                    "definition": "wall time to relative gap <= tol of the cold-started loop / of the loop "
                                  "warm-started by the network (its (u, v) per level, dual recovered from it, "
                                  "cold fallback per level)"}
        if "warm_sched" in gap:
            warm["per_depth_steps"] = {
                "vs_cold_reference_steps": gap["warm_sched"]["speedup_vs_cold_reference_steps"],
                "vs_cold_same_steps": round(gap["cold_sched"]["wall_ms"] / gap["warm_sched"]["wall_ms"], 3),
                "vs_cold_dual_same_steps": round(gap["cold_dual_sched"]["wall_ms"] / gap["warm_sched"]["wall_ms"], 3)}

    # the drop-in path: the reference's own update rules (tgv_minimize's
    # iteration, tgv.cpp:137-232, bit for bit) over the same octree and sweep,
    # two iterations per launch over tile sets of the reference's table
    ref_sweep = None
    if not args.no_ref:
        ctx.load(cubes, hist, precision=prec, op_mode=T.OP_REF)
        pr = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_REF, iters=args.iters,
                         use_graph=prm.use_graph, fused=6)
        rtune = None
        if args.tune:
            rtune = ctx.autotune(pr, iters=args.iters)
            pr.tune = 1
        ctx.run(pr)
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        li_ref = 0.0
        for _ in range(args.steps):
            li_ref += ctx.run(pr).leaf_iters
        hip.hipDeviceSynchronize()
        dt_ref = time.perf_counter() - t1
        _, rd_ms, rp_ms, rf_ms = ctx.level_bench(finest, pr, 100)
        # the drop-in tgv_minimize end to end (host arrays in, state out: load,
        # one sweep of the reference's rules, read-back), the better of two
        pd = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_REF, iters=args.iters, fused=6,
                         no_energy=1)
        call_s = []
        for _ in range(2):
            t2 = time.perf_counter()
            ctx.solve(cubes, hist, pd)
            call_s.append(time.perf_counter() - t2)
        r_var = rtune[-1]["variant"] if rtune else 5
        r_it = rf_ms if (r_var != 0 and rf_ms > 0) else rd_ms + rp_ms
        ref_sweep = {"value": li_ref / dt_ref, "unit": "leaf-iter/s", "ms_per_step": dt_ref / args.steps * 1e3,
                     "op_mode": "reference", "finest_kernel": "k_fused2 (reference rules), two iterations per launch"
                     if r_var == 5 else "k_dual + k_primal_ref",
                     "finest_us_per_iteration": r_it * 1e3, "finest_two_kernel_us": (rd_ms + rp_ms) * 1e3,
                     "finest_achieved_gbs": ab * n_f / (r_it * 1e-3) / 1e9,
                     "finest_frac": ab * n_f / (r_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "autotune": [(l["depth"], l["variant"], l["block"], l["xcd"]) for l in rtune] if rtune else None,
                     "dropin_call_s": min(call_s),
                     "dropin_call_definition": "one tgv_minimize-equivalent call (tgvx_solve: the cubes and histograms "
                                               "from host memory, the problem load, one sweep of the reference's "
                                               "rules, the state read back into host memory), the better of two"}
        log(f"[bench] reference rules: {ref_sweep['value']:.4g} leaf-iter/s; depth {finest} "
            f"{r_it * 1e3:.1f} us per iteration (two kernels {(rd_ms + rp_ms) * 1e3:.1f} us); drop-in call "
            f"{min(call_s) * 1e3:.1f} ms")

    # the drop-in tgv_minimize_oc (tgv.cpp:482-507): treetop leaves of at
    # most 2^20 cubes solved one after another per level, out-of-leaf
    # u_bar, v_bar frozen at the level start; one call = problem load +
    # sweep (the iteration time is the report's, events around each level's
    # leaf loop). Replaces the resident problem.
    if not args.no_ref and not args.no_oc:
        leaves = scene.treetop(cubes, 1 << 20)
        po = T.TGVParams(alpha0=2.0, alpha1=1.0, iters=args.iters, precision=prec, op_mode=T.OP_REF, fused=6)
        ctx.solve_oc(cubes, hist, leaves, po)
        t1 = time.perf_counter()
        _, orep = ctx.solve_oc(cubes, hist, leaves, po)
        dt_oc = time.perf_counter() - t1
        ref_sweep["oc"] = {"leaves": int(len(leaves)), "wall_s": round(dt_oc, 3), "iterate_ms": round(orep.iterate_ms, 2),
                           "leaf_iter_per_s": orep.leaf_iters / (orep.iterate_ms * 1e-3),
                           "definition": "tgv_minimize_oc, treetop leaves of at most 2^20 cubes, one call (load "
                                         "included in wall_s); leaf_iter_per_s over the iterations only"}
        log(f"[bench] reference rules, frozen-border leaves ({len(leaves)}): iterations {orep.iterate_ms:.1f} ms "
            f"({ref_sweep['oc']['leaf_iter_per_s']:.4g} leaf-iter/s), call {dt_oc:.3f} s")

    cpu = None
    if args.cpu_seconds > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        model, _ = O.host_cpu()
        _, march = O.timing_lib()
        # the process's CPU share: the cores it may run on, as the job's
        # thread limit caps them (OMP_NUM_THREADS)
        share = len(os.sched_getaffinity(0))
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        ncpu = min(share, omp) if omp > 0 else share

        def timed(op_mode, threads, budget):
            op = O.params(alpha0=2.0, alpha1=1.0, op_mode=op_mode, threads=threads)
            t1, na = O.time_level(cubes, hist, op, finest, 1)
            k = max(1, int(budget / max(t1, 1e-3)))
            tk, na = O.time_level(cubes, hist, op, finest, k)
            return na * k / tk, k, tk, na

        budget = args.cpu_seconds / 3
        v1, k1, t1, na = timed(0, 1, budget)
        cpu = {"value": v1, "unit": "leaf-iter/s", "cores": 1, "kind": "port",
               "sample": f"{k1} iterations of the depth-{finest} level ({na} leaves) of this workload: the CPU "
                         f"restatement of run_level (tgv.cpp:137-232) with the reference's update rules, AoS "
                         f"TGVCube, active[] indirection, the median recomputed in prox_data, FP32, "
                         f"-O3 -march={march}, 1 thread, {t1:.1f} s",
               "host_cpu": model, "nproc": os.cpu_count(), "cpu_share": ncpu, "march": march}
        log(f"[bench] CPU baseline (reference rules, 1 thread): {v1:.4g} leaf-iter/s ({k1} its, {t1:.1f} s)")
        # the same loops over the process's cores (OpenMP over cubes, Jacobi passes)
        vm, km, tm, _ = timed(0, ncpu, budget)
        cpu["all_cores"] = {"value": vm, "cores": ncpu,
                            "sample": f"{km} iterations of the same level, reference rules, {ncpu} OpenMP threads "
                                      f"(the process's CPU share of {share} cores), {tm:.1f} s"}
        log(f"[bench] CPU, reference rules, {ncpu} threads: {vm:.4g} leaf-iter/s")
        # the exact-adjoint iteration the GPU sweep runs, same loops, one thread
        va, ka, ta, _ = timed(1, 1, budget)
        cpu["adjoint_1_thread"] = {"value": va, "cores": 1,
                                   "sample": f"{ka} iterations of the same level, exact-adjoint rules (the "
                                             f"GPU sweep's iteration), 1 thread, {ta:.1f} s"}
        log(f"[bench] CPU, exact-adjoint rules, 1 thread: {va:.4g} leaf-iter/s")

    # HBM traffic per iteration from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # summary of this workload (counters cannot be read from inside the run)
    traffic = traffic_it = None
    its_launch = 2 if two else 1
    pj = os.path.join(ROOT, "profiles", "tgv_iteration_depth10_torus_box.json")
    if args.precision == "fp32" and os.path.exists(pj):
        name = KERNEL_NAMES[variant]
        if variant in (1, 3, 4, 5):
            name += f", tile {tile}"
        rec = json.load(open(pj)).get(name)
        if rec and rec["leaves"] == n_f:
            traffic_it = rec["traffic_bytes_per_iteration"]
            traffic = traffic_it * rec.get("iterations_per_launch", 1)

    out = {
        "metric": "TGV leaf-iter/s + time-to-gap<=1e-4 (learned warm start vs cold), 1 GPU",
        "value": value,
        "unit": "leaf-iter/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16-storage/f32", "bf16-dual": "bf16-dual-storage/f32",
                  "fp64": "f64"}[args.precision],
        "data": "synthetic",
        "config": {"workload": "torus_box_depth10", "leaves": n, "views": args.views,
                   "image": [args.width, args.height], "max_depth": args.max_depth, "bins": 10,
                   "iters_per_level_per_step": args.iters, "op_mode": "adjoint", "levels": levels,
                   "parallelism": "single-gpu", "device": arch, "cus": cus},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_unit": "HBM bytes per launch of the kernel at the finest level (rocprofv3 "
                                     "FETCH_SIZE x 2 + WRITE_SIZE, committed profile)",
                     "traffic_per_iteration": traffic_it, "iterations_per_launch": its_launch,
                     "launch_us": (fused_ms if fused else dual_ms + primal_ms) * 1e3 * its_launch,
                     "alg_bytes_per_launch": ab * n_f * its_launch,
                     "kernel": kernel, "fused_us": fused_ms * 1e3,
                     "alg_bytes_per_leaf_iter": ab, "dual_us": dual_ms * 1e3, "primal_us": primal_ms * 1e3},
        "time_to_gap": gap,
        "warm_start_speedup": warm,
        "reference_rules_sweep": ref_sweep,
        "with_energy_report": with_energy,
        "level_ms_per_step": level_ms,
        "cpu_baseline": cpu,
        "autotune": tuning,
        "setup_s": {"scene": round(t_scene, 2), "histograms": round(t_hist, 3), "load": round(t_load, 3)},
        "value_with_setup": with_setup,
        "octree_build": octree,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def variants(args, T, scene):
    """Where launches cost: on the depth-10 scene (exact-adjoint rules, FP32),
    each level of at most 2^18 cubes timed for 50 iterations with each launch
    variant (the better of three runs, graphs for the variants that launch
    per iteration), the cost of one grid barrier of the persistent level
    kernel by grid size, and whole sweeps (50 iterations per level) with one
    variant for every level."""
    t0 = time.time()
    sc = scene.torus_box_scene(views=args.views, W=args.width, H=args.height, max_depth=args.max_depth)
    log(f"[bench] scene: {len(sc.cubes)} cubes in {time.time() - t0:.1f} s")
    ctx = T.Context(0)
    arch, cus = ctx.device_info()
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    ctx.load(sc.cubes, hist, precision=T.FP32, op_mode=T.OP_ADJ)
    cfgs = [(0, 256), (4, 128), (5, 512), (5, 1024), (6, 1024), (7, 1024), (7, 256)]
    per_level = []
    for d, n in ctx.levels():
        if n > 1 << 18:
            continue
        row = {"depth": d, "cubes": n}
        for v, b in cfgs:
            p = T.TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=T.FP32, op_mode=T.OP_ADJ, fused=v, block=b)
            ctx.level_bench(d, p, 50)  # graphs and tables built
            row[f"{KERNEL_NAMES[v]}/{b}"] = round(min(ctx.level_bench(d, p, 50)[0] for _ in range(3)) * 1e3 / 50, 2)
        per_level.append(row)
        log(f"[bench] depth {d} ({n} cubes), us per iteration: "
            f"{ {k: v for k, v in row.items() if k not in ('depth', 'cubes')} }")
    barrier = []
    for wgs, b in [(1, 1024), (4, 1024), (16, 1024), (64, 1024), (256, 1024), (64, 256), (256, 256), (1024, 256)]:
        us = min(ctx.grid_barrier_us(wgs, b) for _ in range(2))
        barrier.append({"workgroups": wgs, "block": b, "us_per_barrier": round(us, 3)})
        log(f"[bench] grid barrier, {wgs} workgroups of {b}: {us:.2f} us")
    hip = __import__("ctypes").CDLL("libamdhip64.so")
    sweeps = []
    for v, b in [(6, 1024), (7, 1024), (7, 256)]:
        p = T.TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=T.FP32, op_mode=T.OP_ADJ, fused=v, block=b)
        ctx.run(p)
        hip.hipDeviceSynchronize()
        t1 = time.perf_counter()
        li = 0.0
        for _ in range(args.steps):
            rep = ctx.run(p)
            li += rep.leaf_iters
        hip.hipDeviceSynchronize()
        dt = time.perf_counter() - t1
        sweeps.append({"variant": KERNEL_NAMES[v], "block": b, "ms_per_sweep": round(dt / args.steps * 1e3, 3),
                       "leaf_iter_per_s": li / dt,
                       "level_ms": [(l["depth"], round(l["ms"], 3)) for l in rep.levels()]})
        log(f"[bench] sweep, {KERNEL_NAMES[v]}/{b} where it fits: {dt / args.steps * 1e3:.2f} ms")
    out = {"metric": "per-level TGV iteration time by launch variant, 1 GPU", "unit": "us per iteration",
           "higher_is_better": False, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "torus_box_depth10", "leaves": int(len(sc.cubes)), "op_mode": "adjoint",
                      "iters": 50, "device": arch, "cus": cus},
           "per_level": per_level, "grid_barrier": barrier, "sweeps": sweeps}
    print(json.dumps(out), flush=True)
    ctx.close()


def streaming(args, T, scene):
    """tgv_minimize_streaming (the reference's stage 6, tgv.cpp:802-1070) on the
    depth-10 scene split into treetop leaves: the resident solve (every cube's
    state in HBM) against the out-of-core solve with the state in pinned host
    memory and at most a given budget of per-cube data on the device. A step
    is one coarse-to-fine sweep with `iters` iterations per level, uploads,
    host work and downloads included (the reference reads and writes its
    files in the same place)."""
    import numpy as np

    t0 = time.time()
    sc = scene.torus_box_scene(views=args.views, W=args.width, H=args.height, max_depth=args.max_depth)
    cubes = sc.cubes
    log(f"[bench] scene: {len(cubes)} cubes in {time.time() - t0:.1f} s")
    ctx = T.Context(0)
    arch, cus = ctx.device_info()
    hist = ctx.histograms(cubes, sc.views(), sc.origin, sc.r_root, 10)
    leaves = scene.treetop(cubes, args.leaf_cubes)
    p = T.TGVParams(alpha0=2.0, alpha1=1.0, iters=args.iters)
    ctx.solve_streaming(cubes, hist, leaves, p)  # warm-up
    t1 = time.perf_counter()
    ref, rrep = ctx.solve_streaming(cubes, hist, leaves, p)
    t_res = time.perf_counter() - t1
    li = rrep.leaf_iters
    log(f"[bench] resident streaming solve: {t_res:.3f} s ({li / t_res:.4g} leaf-iter/s incl. load)")
    runs = []
    for mb in [int(x) for x in args.budgets_mb.split(",") if x]:
        st, rep, orep = ctx.solve_streaming_ooc(cubes, hist, leaves, p, device_bytes=mb << 20)
        t1 = time.perf_counter()
        st, rep, orep = ctx.solve_streaming_ooc(cubes, hist, leaves, p, device_bytes=mb << 20)
        dt = time.perf_counter() - t1
        same = bool(np.array_equal(np.ascontiguousarray(st).view(np.uint8), np.ascontiguousarray(ref).view(np.uint8)))
        runs.append({"budget_mb": mb, "s": round(dt, 3), "leaf_iter_per_s": li / dt, "identical_to_resident": same,
                     "iterate_and_transfer_ms": round(rep.iterate_ms, 1), **orep,
                     "pcie_gbs": orep["moved_bytes"] / max(rep.iterate_ms, 1e-9) / 1e6})
        log(f"[bench] out of core, {mb} MiB: {dt:.3f} s, {orep['batches']} batches, "
            f"{orep['moved_bytes'] / 1e9:.2f} GB moved, identical {same}")
    resident_bytes = len(cubes) * (17 * 4 + 4 + 4 * 4 + 6 * 4 + 1 + 4 + 12)
    out = {"metric": "streaming TGV leaf-iter/s, resident and out of core, 1 GPU",
           "value": runs[-1]["leaf_iter_per_s"] if runs else li / t_res, "unit": "leaf-iter/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "torus_box_depth10_streaming", "cubes": int(len(cubes)), "leaves": int(len(leaves)),
                      "leaf_cubes_max": args.leaf_cubes, "iters_per_level": args.iters, "op_mode": "reference",
                      "device": arch, "cus": cus},
           "resident": {"s": round(t_res, 3), "leaf_iter_per_s": li / t_res,
                        "device_state_bytes_estimate": resident_bytes},
           "out_of_core": runs}
    print(json.dumps(out), flush=True)
    ctx.close()


def batched(args, T, scene, prec):
    """256 independent octree blocks (seeded unions of 1-6 spheres and boxes,
    about 4k-512k cubes each, log-uniform, 10 bins), solved as one resident
    problem -- every iteration of a depth is one launch over that depth's
    cubes of all blocks, 50-iteration chunks as graphs -- against one problem
    per block. A step is one coarse-to-fine sweep of 50 iterations per level."""
    import numpy as np

    ctx = T.Context(0)
    arch, cus = ctx.device_info()
    targets = scene.batch_targets(args.blocks)
    blocks, t0 = [], time.time()
    for b, tgt in enumerate(targets):
        sc = scene.batch_block(4000 + b, tgt)
        blocks.append((sc.cubes, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)))
        if b % 16 == 15:
            log(f"[bench] {b + 1}/{len(targets)} blocks generated ({time.time() - t0:.0f} s)")
    counts = np.array([len(c) for c, _ in blocks])
    n = int(counts.sum())
    log(f"[bench] {n} cubes in {len(blocks)} blocks (min {counts.min()}, median {int(np.median(counts))}, "
        f"max {counts.max()}) in {time.time() - t0:.0f} s")

    prm = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_ADJ, iters=args.iters,
                      use_graph=0 if args.no_graph else 1, fused=args.fused, block=args.block)
    ctx.load_blocks(blocks, precision=prec, op_mode=T.OP_ADJ)
    tuning = None
    if args.tune:
        tuning = ctx.autotune(prm, iters=args.iters)
        prm.tune = 1
        log(f"[bench] autotuned: {[(l['depth'], l['variant'], l['block'], l['xcd']) for l in tuning]}")
    for _ in range(args.warmup):
        ctx.run(prm)
    ts = time.perf_counter()
    li = 0.0
    for _ in range(args.steps):
        li += ctx.run(prm).leaf_iters
    dt = time.perf_counter() - ts
    value = li / dt
    log(f"[bench] batched: {args.steps} sweeps in {dt:.3f} s: {value:.4g} leaf-iter/s")

    # one problem per block: the same sweep, each block loaded and run alone
    per_t, per_li = 0.0, 0.0
    for cubes, hist in blocks:
        ctx.load(cubes, hist, precision=prec, op_mode=T.OP_ADJ)
        ctx.run(prm)  # builds this block's graphs
        t1 = time.perf_counter()
        per_li += ctx.run(prm).leaf_iters
        per_t += time.perf_counter() - t1
    per_value = per_li / per_t
    log(f"[bench] one problem per block: {per_t * 1e3:.1f} ms per sweep of all blocks: {per_value:.4g} leaf-iter/s")

    gap = None
    if args.batched_gap > 0:
        pg = T.TGVParams(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=T.OP_ADJ, gap_tol=args.batched_gap,
                         chunk=50, max_iters=20000, use_graph=prm.use_graph, fused=prm.fused, block=prm.block,
                         tune=prm.tune)
        ctx.load_blocks(blocks, precision=prec, op_mode=T.OP_ADJ)
        ctx.run(pg)
        t1 = time.perf_counter()
        rb = ctx.run(pg)
        tb = time.perf_counter() - t1
        it, _ = ctx.block_report()
        tp = 0.0
        for cubes, hist in blocks:
            ctx.load(cubes, hist, precision=prec, op_mode=T.OP_ADJ)
            ctx.run(pg)
            t1 = time.perf_counter()
            ctx.run(pg)
            tp += time.perf_counter() - t1
        gap = {"tol": args.batched_gap, "batched_s": round(tb, 3), "per_block_s": round(tp, 3),
               "speedup": round(tp / tb, 3), "batched_leaf_iters": rb.leaf_iters,
               "max_block_iters_per_level": it.max(axis=1).tolist()}
        log(f"[bench] to gap {args.batched_gap}: batched {tb:.2f} s, one problem per block {tp:.2f} s")

    out = {
        "metric": "TGV leaf-iter/s, 256 independent blocks batched, 1 GPU",
        "value": value, "unit": "leaf-iter/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp32": "f32", "bf16": "bf16-storage/f32", "bf16-dual": "bf16-dual-storage/f32",
                  "fp64": "f64"}[args.precision], "data": "synthetic",
        "config": {"workload": f"batched_{len(blocks)}_blocks", "cubes": n, "blocks": len(blocks),
                   "cubes_min_median_max": [int(counts.min()), int(np.median(counts)), int(counts.max())],
                   "iters_per_level_per_step": args.iters, "op_mode": "adjoint", "bins": 10,
                   "parallelism": "single-gpu", "device": arch, "cus": cus},
        "one_problem_per_block": {"value": per_value, "ms_per_sweep": per_t * 1e3,
                                  "speedup_batched": round(value / per_value, 3)},
        "to_gap": gap,
        "autotune": tuning,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
