"""Storage-precision accuracy of the TGV solve against FP64 (test support).

For a storage precision X (FP32, or BF16 storage with FP32 arithmetic) the
scene is solved cold to relative duality gap `tol` in the exact-adjoint mode,
or run for given per-level iteration counts. The same problem is then run in
FP64 for exactly the per-level iteration counts the X run took (the FP64 path
is bit-identical to the FP64 CPU oracle, pinned by the parity tests), and the
two are compared: L_inf of u over all cubes and over the cubes with votes,
and the Hausdorff distance between the surfaces extracted from both
(marching_cubes.cpp:63-181 over all cubes, faces only between cubes with
votes, by the library's device extraction, which the surface tests pin
bit for bit to the oracle's), in units of the finest leaf edge. The distance
is taken between the mesh vertex sets, both ways. The FP64 solve to the same
gap is also run, for its iteration counts.

    python tests/accuracy.py --out gpurun_out/accuracy_depth10.json   # the depth-10 torus+box scene

With --oracle the FP64 reference is the CPU oracle itself (the restatement of
tgv.cpp's iteration in FP64, OpenMP over the host cores), solved to the same gap
with steps sigma / tau = 1024: the GPU FP64 solve must equal it bit for bit,
with the same per-level iteration counts, and the GPU FP32 solve is compared
with it (u and surfaces).

    python tests/accuracy.py --oracle --out gpurun_out/accuracy_oracle_depth10.json

With --config3 the sphere workload (24 ring views of 320x240, depth <= 7) is
solved to the gap in FP32, BF16-dual (p, q in BF16) and BF16 storage and each
is compared with the FP64 oracle's solution to the same gap.

    python tests/accuracy.py --config3 --out gpurun_out/accuracy_sphere_precisions.json
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

from tgvreconclaude_amd import BF16, BF16_DUAL, FP32, FP64, OP_ADJ, TGVParams  # noqa: E402

NAMES = {FP32: "fp32", BF16: "bf16", FP64: "fp64", BF16_DUAL: "bf16-dual"}


def _params(prec, **kw):
    base = dict(alpha0=2.0, alpha1=1.0, precision=prec, op_mode=OP_ADJ, chunk=50, max_iters=20000)
    base.update(kw)
    return TGVParams(**base)


def hausdorff(a, b):
    """Symmetric Hausdorff distance between two point sets (inf if exactly one is empty)."""
    from scipy.spatial import cKDTree
    if len(a) == 0 and len(b) == 0:
        return 0.0
    if len(a) == 0 or len(b) == 0:
        return float("inf")
    da, _ = cKDTree(b).query(a, k=1, workers=-1)
    db, _ = cKDTree(a).query(b, k=1, workers=-1)
    return float(max(da.max(), db.max()))


def surface(ctx, sc, u, has_data):
    """The library's extraction over the loaded problem (bit-identical to the
    oracle's, pinned by the surface tests)."""
    return ctx.surface(sc.origin, sc.r_root, u=np.ascontiguousarray(u, dtype=np.float32), has_data=has_data)


def log(msg):
    print(f"[accuracy] {msg}", flush=True)


def u_of(st):
    return st["u"].astype(np.float64) if st.dtype.names else st[:, 0].copy()


def compare(ctx, sc, hist, prec, tol=1e-4, level_iters=None, max_iters=20000, **steps):
    """Report for one storage precision against FP64 run for the same
    iterations: solved to relative gap tol, or for the given per-level
    iteration counts. steps: sigma, tau, sigma_depth, tau_depth for both runs."""
    has = (hist.sum(axis=1) > 0).astype(np.uint8)
    leaf = 2.0 * sc.r_root / 2 ** int(sc.cubes["depth"].max())
    ctx.load(sc.cubes, hist, precision=prec, op_mode=OP_ADJ)
    t0 = time.time()
    if level_iters is None:
        rep = ctx.run(_params(prec, gap_tol=tol, max_iters=max_iters, **steps))
    else:
        rep = ctx.run_levels(_params(prec, **steps), level_iters)
    its = [l["iters"] for l in rep.levels()]
    gaps = [l["rel_gap"] for l in rep.levels()]
    reached = all(g <= tol for g in gaps)
    u = u_of(ctx.state())
    log(f"{NAMES[prec]}: {rep.solve_ms:.0f} ms, iterations {its}")
    ctx.load(sc.cubes, hist, precision=FP64, op_mode=OP_ADJ)
    r64 = ctx.run_levels(_params(FP64, **steps), its)
    u64 = u_of(ctx.state())
    log(f"fp64 for the same iterations: {r64.solve_ms:.0f} ms")
    d = np.abs(u - u64)
    va, _ = surface(ctx, sc, u, has)
    vb, _ = surface(ctx, sc, u64, has)
    h = hausdorff(va, vb)
    log(f"surfaces of {len(va)} and {len(vb)} vertices, Hausdorff {h:.3g}")
    return {"precision": NAMES[prec], "gap_tol": tol if level_iters is None else None, "reached": reached,
            "iters_per_level": its, "final_rel_gap": gaps,
            "solve_ms": rep.solve_ms, "linf_u": float(d.max()), "linf_u_voted": float(d[has > 0].max()),
            "mean_abs_u": float(d.mean()), "rms_u": float(np.sqrt(np.mean(d * d))),
            "l2_u": float(np.linalg.norm(d)), "hausdorff": h, "leaf_edge": leaf, "hausdorff_leaf": h / leaf,
            "surface_vertices": [int(len(va)), int(len(vb))], "wall_s": round(time.time() - t0, 2)}


def fp64_to_gap(ctx, sc, hist, tol=1e-4):
    ctx.load(sc.cubes, hist, precision=FP64, op_mode=OP_ADJ)
    rep = ctx.run(_params(FP64, gap_tol=tol))
    return {"iters_per_level": [l["iters"] for l in rep.levels()],
            "reached": all(l["rel_gap"] <= tol for l in rep.levels()), "solve_ms": rep.solve_ms}


def against_oracle(ctx, sc, hist, tol, sigma, tau, threads, fused=None, block=None):
    """GPU FP32 and FP64 solves to the gap against the FP64 CPU oracle's;
    fused / block select the GPU's launch variant (default: the library's)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    has = (hist.sum(axis=1) > 0).astype(np.uint8)
    leaf = 2.0 * sc.r_root / 2 ** int(sc.cubes["depth"].max())
    kw = dict(gap_tol=tol, sigma=sigma, tau=tau)
    launch = {}
    if fused is not None:
        launch["fused"] = fused
    if block is not None:
        launch["block"] = block
    gpu = {}
    for prec in (FP32, FP64):
        ctx.load(sc.cubes, hist, precision=prec, op_mode=OP_ADJ)
        rep = ctx.run(_params(prec, **kw, **launch))
        st = ctx.state()
        gpu[prec] = (st, [l["iters"] for l in rep.levels()], rep.solve_ms)
        log(f"gpu {NAMES[prec]} to gap ({launch or 'default launch'}): {rep.solve_ms:.0f} ms, "
            f"iterations {gpu[prec][1]}")
    same32 = None
    if launch:  # the FP32 state of the chosen launch against the default launch's
        ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
        ctx.run(_params(FP32, **kw))
        a, b = np.asarray(ctx.state()), np.asarray(gpu[FP32][0])
        same32 = bool(a.tobytes() == b.tobytes())
        log(f"fp32 state equal to the default launch's bit for bit: {same32}")
    t0 = time.time()
    log(f"oracle fp64 to gap, {threads} host threads ...")
    ref, _, its, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=tol, chunk=50,
                                                             max_iters=20000, threads=threads, sigma=sigma, tau=tau),
                                    f64=True)
    t_or = time.time() - t0
    log(f"oracle fp64: {t_or:.1f} s, iterations {its.tolist()}")
    st64 = gpu[FP64][0]
    g64 = (np.concatenate([st64["u"][:, None], st64["v"], st64["p"], st64["q"]], axis=1) if st64.dtype.names
           else np.ascontiguousarray(st64[:, :13], dtype=np.float64))
    bit_exact = bool(np.array_equal(g64.view(np.uint64), np.ascontiguousarray(ref).view(np.uint64)))
    u32 = u_of(gpu[FP32][0])
    d = np.abs(u32 - ref[:, 0])
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    va, _ = surface(ctx, sc, u32, has)
    vb, _ = surface(ctx, sc, ref[:, 0], has)
    h = hausdorff(va, vb)
    log(f"fp64 gpu == oracle bit for bit: {bit_exact}; fp32 linf u {d.max():.3g}; Hausdorff {h / leaf:.3g} leaf")
    return {"sigma": sigma, "tau": tau, "gap_tol": tol, "launch": launch or "default",
            "fp32_state_equal_to_default_launch": same32,
            "oracle_threads": threads, "oracle_s": round(t_or, 1),
            "oracle_iters_per_level": its.tolist(), "gpu_fp64_iters_per_level": gpu[FP64][1],
            "gpu_fp32_iters_per_level": gpu[FP32][1], "gpu_fp64_ms": gpu[FP64][2], "gpu_fp32_ms": gpu[FP32][2],
            "fp64_state_bit_exact_with_oracle": bit_exact,
            "fp64_iters_equal": gpu[FP64][1] == its.tolist(),
            "fp32_linf_u": float(d.max()), "fp32_linf_u_voted": float(d[has > 0].max()), "fp32_mean_abs_u": float(d.mean()),
            "fp32_rms_u": float(np.sqrt(np.mean(d * d))), "fp32_l2_u": float(np.linalg.norm(d)),
            "hausdorff": h, "leaf_edge": leaf, "hausdorff_leaf": h / leaf,
            "surface_vertices": [int(len(va)), int(len(vb))]}


def config3(ctx, sc, hist, threads, cap=6000, tol=1e-4):
    """Storage precisions on the sphere workload against the FP64 CPU oracle:
    each precision is solved to relative gap tol (at most cap iterations per
    level) with the per-depth steps, the oracle in FP64 to the same gap with
    the same steps; L_inf / RMS of u and the surfaces' Hausdorff distance
    against the oracle's solution."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from tgvreconclaude_amd import step_schedule
    has = (hist.sum(axis=1) > 0).astype(np.uint8)
    d = int(sc.cubes["depth"].max())
    leaf = 2.0 * sc.r_root / 2 ** d
    ratios = [256 if k <= d - 4 else (1024 if k < d else 4096) for k in range(d + 1)]
    sd, td = step_schedule(ratios)
    steps = dict(sigma=8.0, tau=1.0 / 128, sigma_depth=sd, tau_depth=td)
    t0 = time.time()
    ref, _, its, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=tol, chunk=50,
                                                             max_iters=cap, threads=threads, sigma=8.0, tau=1.0 / 128,
                                                             sigma_depth=sd, tau_depth=td), f64=True)
    log(f"oracle fp64 to gap: {time.time() - t0:.1f} s, iterations {its.tolist()}")
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    vr, _ = surface(ctx, sc, ref[:, 0], has)
    out = {"gap_tol": tol, "cap_per_level": cap, "oracle_iters_per_level": its.tolist(), "leaf_edge": leaf,
           "results": []}
    for prec in (FP32, BF16_DUAL, BF16):
        ctx.load(sc.cubes, hist, precision=prec, op_mode=OP_ADJ)
        rep = ctx.run(_params(prec, gap_tol=tol, max_iters=cap, **steps))
        lv = rep.levels()
        u = u_of(ctx.state())
        e = np.abs(u - ref[:, 0])
        # the storage error alone: FP64 (the oracle's iteration, bit for bit)
        # run for the same per-level iteration counts
        its_p = [l["iters"] for l in lv]
        ctx.load(sc.cubes, hist, precision=FP64, op_mode=OP_ADJ)
        ctx.run_levels(_params(FP64, **steps), its_p)
        e_same = np.abs(u - u_of(ctx.state()))
        ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
        va, _ = surface(ctx, sc, u, has)
        h = hausdorff(va, vr)
        r = {"precision": NAMES[prec], "reached": all(l["rel_gap"] <= tol for l in lv),
             "iters_per_level": its_p, "final_rel_gap": [l["rel_gap"] for l in lv],
             "linf_u": float(e.max()), "rms_u": float(np.sqrt(np.mean(e * e))), "hausdorff_leaf": h / leaf,
             "linf_u_same_iters": float(e_same.max()), "rms_u_same_iters": float(np.sqrt(np.mean(e_same ** 2))),
             "bytes_per_leaf_iter": {FP32: 168, BF16_DUAL: 132, BF16: 100}[prec]}
        log(f"{NAMES[prec]}: reached {r['reached']}, iterations {r['iters_per_level']}, max gap "
            f"{max(r['final_rel_gap']):.3g}, linf u {r['linf_u']:.3g} (same iterations in FP64: "
            f"{r['linf_u_same_iters']:.3g}), Hausdorff {r['hausdorff_leaf']:.3g} leaf")
        out["results"].append(r)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--bf16-max-iters", type=int, default=6000, help="per-level cap of the BF16 to-gap attempt")
    ap.add_argument("--oracle", action="store_true", help="FP64 reference = the CPU oracle (steps sigma/tau = 1024)")
    ap.add_argument("--config3", action="store_true",
                    help="storage precisions on the sphere workload against the FP64 oracle solved to the gap")
    ap.add_argument("--threads", type=int, default=0, help="oracle host threads (0: the cores available, at most 16)")
    ap.add_argument("--fused", type=int, default=None, help="with --oracle: the GPU's launch variant (0-6)")
    ap.add_argument("--block", type=int, default=None, help="with --oracle: the GPU's workgroup / tile size")
    args = ap.parse_args()
    from tgvreconclaude_amd import Context, scene
    if args.config3:
        sc = scene.sphere_scene()
        ctx = Context(0)
        hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
        res = config3(ctx, sc, hist, args.threads or min(16, len(os.sched_getaffinity(0))))
        res.update({"workload": "sphere_depth7", "cubes": int(len(sc.cubes))})
        with open(args.out, "w") as f:
            json.dump(res, f, indent=1)
        ctx.close()
        return
    t0 = time.time()
    sc = scene.torus_box_scene()
    print(f"[accuracy] scene: {len(sc.cubes)} cubes ({time.time() - t0:.0f} s)", flush=True)
    ctx = Context(0)
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    if args.oracle:
        threads = args.threads or min(16, len(os.sched_getaffinity(0)))
        res = against_oracle(ctx, sc, hist, args.tol, 8.0, 1.0 / 128, threads, args.fused, args.block)
        with open(args.out, "w") as f:
            json.dump({"workload": "torus_box_depth10", "cubes": int(len(sc.cubes)), "oracle": res}, f, indent=1)
        ctx.close()
        return
    out = {"workload": "torus_box_depth10", "cubes": int(len(sc.cubes)), "results": []}
    # FP32 to the gap; BF16 for the FP32 solve's iterations; BF16 to the gap
    r32 = compare(ctx, sc, hist, FP32, args.tol)
    runs = [r32, compare(ctx, sc, hist, BF16, args.tol, level_iters=r32["iters_per_level"]),
            compare(ctx, sc, hist, BF16, args.tol, max_iters=args.bf16_max_iters)]
    for r in runs:
        out["results"].append(r)
        print(f"[accuracy] {r['precision']} (gap_tol {r['gap_tol']}): iters {r['iters_per_level']} "
              f"reached {r['reached']} max rel gap {max(r['final_rel_gap']):.3g} linf_u {r['linf_u']:.3g} "
              f"hausdorff {r['hausdorff_leaf']:.3f} leaf ({r['wall_s']} s)", flush=True)
    out["fp64_to_gap"] = fp64_to_gap(ctx, sc, hist, args.tol)
    print(f"[accuracy] fp64 to gap: iters {out['fp64_to_gap']['iters_per_level']}", flush=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    main()
