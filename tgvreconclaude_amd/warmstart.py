"""Learned warm start: training-data generation, training and evaluation.

Training pairs come from this solver itself: seeded CSG scenes (24 ring views
of 320x240, octree depth <= 9, as the small sphere workload) are solved to relative gap 1e-4 on the GPU; for
every level, the features of each cube (computed with the parent level's final
state, exactly as at inference) are paired with the cube's converged (u, v)
minus its cold initialisation. The network is trained on the GPU with the
library's own forward / backward / Adam kernels. Evaluation solves held-out
depth-10 scenes (24 ring views of 1280x960) to the same gap: the plain cold
loop with the reference's steps (sigma = tau = 0.25, p = q = 0), and with steps
sigma / tau = 1024: cold, cold with the dual recovered from the initial
(u, v), and warm (network + recovered dual, with the per-level fallback to
cold); each must reach the gap.

    python -m tgvreconclaude_amd.warmstart train --scenes 150 --out weights.npy
    python -m tgvreconclaude_amd.warmstart eval --scenes 50 --weights weights.npy
    python -m tgvreconclaude_amd.warmstart bound --eval-scenes 5   # what any predictor could gain
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import FP32, OP_ADJ, Context, TGVParams, scene

ALPHA0, ALPHA1 = 2.0, 1.0
GAP = 1e-4
DUAL_EPS = 1e-4
# primal-dual steps: sigma / tau = 1024 with sigma * tau = 1/16 (the reference's
# 0.25 * 0.25); at the reference's sigma = tau = 0.25 the same solves take
# 5-8x longer
SIGMA, TAU = 8.0, 1.0 / 128
REFERENCE_STEPS = dict(sigma=0.25, tau=0.25)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "warmstart_mlp.npy")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


USE_GRAPH = 1  # 0 (--no-graph): plain launches, e.g. under a counter-collecting profiler


def params(**kw):
    # launches: levels of at most 1024 cubes in one workgroup, the others two
    # iterations per launch with tiles of 512
    base = dict(alpha0=ALPHA0, alpha1=ALPHA1, precision=FP32, op_mode=OP_ADJ, gap_tol=GAP, chunk=50,
                max_iters=40000, sigma=SIGMA, tau=TAU, use_graph=USE_GRAPH, fused=6)
    base.update(kw)
    return TGVParams(**base)


def train_scene(seed, max_depth=9, W=320, H=240):
    return scene.csg_scene(seed, views=24, W=W, H=H, max_depth=max_depth)


def eval_scene(seed, max_depth=10, W=1280, H=960):
    return scene.csg_scene(seed, views=24, W=W, H=H, max_depth=max_depth)


def harvest(ctx, sc, cap, rng):
    """(features, residual targets) of every level of one scene, solved to the gap."""
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    rep = ctx.run(params(dual_eps=DUAL_EPS))
    xs, ts = [], []
    for depth, n in ctx.levels():
        x, uv, _ = ctx.features(depth)
        if n > cap:
            keep = rng.choice(n, cap, replace=False)
            x, uv = x[keep], uv[keep]
        xs.append(x)
        ts.append(uv - x[:, 20:24])
    return np.concatenate(xs), np.concatenate(ts), rep


def make_training_set(ctx, n_scenes, seed0=1000, cap=4000):
    rng = np.random.default_rng(seed0)
    X, T = [], []
    t0 = time.time()
    for k in range(n_scenes):
        sc = train_scene(seed0 + k)
        x, t, rep = harvest(ctx, sc, cap, rng)
        X.append(x)
        T.append(t)
        if k % 10 == 0:
            log(f"[warmstart] scene {k}/{n_scenes}: {len(sc.cubes)} cubes, {len(x)} pairs, "
                f"solve {rep.solve_ms:.0f} ms ({time.time() - t0:.0f} s)")
    return np.concatenate(X), np.concatenate(T)


def train(ctx, X, T, epochs=12, batch=4096, lr=1e-3, seed=0):
    ctx.mlp_set(initial_weights(seed))
    curve = ctx.mlp_train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return ctx.mlp_get(), curve


def initial_weights(seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(6404, dtype=np.float32)
    w[:1920] = rng.normal(0, np.sqrt(2 / 30), 1920)
    w[1984:6080] = rng.normal(0, np.sqrt(2 / 64), 4096)
    w[6144:6400] = rng.normal(0, 0.01, 256)
    return w


def depth_schedule(max_depth):
    """Per-depth steps: sigma / tau = 256 up to four levels above the finest,
    1024 for the next three, 4096 at the finest (sigma * tau = 1/16)."""
    from . import step_schedule
    sd, td = step_schedule([256 if d <= max_depth - 4 else (1024 if d < max_depth else 4096)
                            for d in range(max_depth + 1)])
    return dict(sigma_depth=sd, tau_depth=td)


def solve_variants(ctx, sc, weights, diagnostics=False):
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    out = {"cubes": int(len(sc.cubes))}
    sched = depth_schedule(int(sc.cubes["depth"].max()))
    runs = {"cold_reference_steps": params(**REFERENCE_STEPS), "cold": params(),
            "cold_dual": params(dual_eps=DUAL_EPS),
            "warm": params(dual_eps=DUAL_EPS, warm_start=1, warm_fallback=1),
            "cold_sched": params(**sched), "cold_dual_sched": params(dual_eps=DUAL_EPS, **sched),
            "warm_sched": params(dual_eps=DUAL_EPS, warm_start=1, warm_fallback=1, **sched)}
    if diagnostics:
        # the network without the fallback rule, and the converged solution
        # itself as the initial (u, v): what a perfect predictor could gain
        runs["warm_no_fallback"] = params(dual_eps=DUAL_EPS, warm_start=1, warm_fallback=0)
        runs["warm_cold_dual"] = params(dual_eps=DUAL_EPS, warm_start=1, warm_fallback=1, warm_dual=1)
        runs["warm_cold_dual_no_fallback"] = params(dual_eps=DUAL_EPS, warm_start=1, warm_fallback=0, warm_dual=1)
        runs["solution_init"] = params(dual_eps=DUAL_EPS, warm_fallback=0)
    ctx.mlp_set(weights)
    ctx.run(params(max_iters=50, **REFERENCE_STEPS))  # untimed: graphs of the reference steps
    ctx.run(runs["cold"])  # untimed: graphs of the tuned steps, and the solution
    st = ctx.state()
    sol = np.ascontiguousarray(np.concatenate([st["u"][:, None], st["v"]], axis=1))
    for name, p in runs.items():
        rep = ctx.run(p, init_uv=sol if name == "solution_init" else None)
        lv = rep.levels()
        out[name] = {"ms": rep.solve_ms, "iters": [l["iters"] for l in lv], "leaf_iters": rep.leaf_iters,
                     "reached": bool(all(l["rel_gap"] <= GAP for l in lv)),
                     "warm_levels": [l["warm"] for l in lv]}
    return out


def evaluate(ctx, weights, n_scenes, seed0=900000, diagnostics=False):
    res = []
    for k in range(n_scenes):
        sc = eval_scene(seed0 + k)
        r = solve_variants(ctx, sc, weights, diagnostics)
        r["seed"] = seed0 + k
        res.append(r)
        log(f"[warmstart] eval {k}: {r['cubes']} cubes  cold (reference steps) "
            f"{r['cold_reference_steps']['ms']:.0f} ms  cold {r['cold']['ms']:.0f} ms  "
            f"cold+dual {r['cold_dual']['ms']:.0f} ms  warm {r['warm']['ms']:.0f} ms")
    ref = np.array([r["cold_reference_steps"]["ms"] for r in res])
    cold = np.array([r["cold"]["ms"] for r in res])
    cd = np.array([r["cold_dual"]["ms"] for r in res])
    warm = np.array([r["warm"]["ms"] for r in res])
    summary = {"scenes": n_scenes, "gap": GAP, "sigma": SIGMA, "tau": TAU,
               "all_reached": bool(all(r[v]["reached"] for r in res
                                       for v in ("cold_reference_steps", "cold", "cold_dual", "warm"))),
               "reached_per_variant": {v: int(sum(r[v]["reached"] for r in res))
                                       for v in ("cold_reference_steps", "cold", "cold_dual", "warm")},
               "speedup_warm_vs_cold_reference_steps_total": float(ref.sum() / warm.sum()),
               "speedup_warm_vs_cold_reference_steps_median": float(np.median(ref / warm)),
               "speedup_cold_vs_cold_reference_steps_total": float(ref.sum() / cold.sum()),
               "cold_reference_steps_ms_total": float(ref.sum()),
               "speedup_warm_vs_cold_total": float(cold.sum() / warm.sum()),
               "speedup_warm_vs_cold_median": float(np.median(cold / warm)),
               "speedup_cold_dual_vs_cold_total": float(cold.sum() / cd.sum()),
               "speedup_warm_vs_cold_dual_total": float(cd.sum() / warm.sum()),
               "cold_ms_total": float(cold.sum()), "cold_dual_ms_total": float(cd.sum()),
               "warm_ms_total": float(warm.sum()),
               "leaf_iters_total": {v: float(sum(r[v]["leaf_iters"] for r in res))
                                    for v in ("cold", "cold_dual", "warm", "cold_sched", "cold_dual_sched",
                                              "warm_sched")},
               "scenes_warm_fewer_leaf_iters_than_cold_dual": int(sum(
                   r["warm"]["leaf_iters"] <= r["cold_dual"]["leaf_iters"] for r in res))}
    for extra in ("warm_no_fallback", "warm_cold_dual", "warm_cold_dual_no_fallback", "solution_init"):
        if extra in res[0]:
            x = np.array([r[extra]["ms"] for r in res])
            summary[f"speedup_{extra}_vs_cold_total"] = float(cold.sum() / x.sum())
    # per-depth step schedule (cold, cold + dual, warm), against the same references
    cs = np.array([r["cold_sched"]["ms"] for r in res])
    cds = np.array([r["cold_dual_sched"]["ms"] for r in res])
    ws = np.array([r["warm_sched"]["ms"] for r in res])
    summary["schedule"] = {
        "reached_per_variant": {v: int(sum(r[v]["reached"] for r in res))
                                for v in ("cold_sched", "cold_dual_sched", "warm_sched")},
        "cold_ms_total": float(cs.sum()), "cold_dual_ms_total": float(cds.sum()), "warm_ms_total": float(ws.sum()),
        "speedup_warm_vs_cold_reference_steps_total": float(ref.sum() / ws.sum()),
        "speedup_warm_vs_cold_reference_steps_median": float(np.median(ref / ws)),
        "speedup_cold_dual_vs_cold_reference_steps_total": float(ref.sum() / cds.sum()),
        "speedup_warm_vs_cold_same_steps_total": float(cs.sum() / ws.sum()),
        "speedup_warm_vs_cold_dual_same_steps_total": float(cds.sum() / ws.sum())}
    return summary, res


def perturbed(st, e_uv, e_pq, rng):
    """The state st (u, v, p, q) with Gaussian noise of RMS e_uv on u, v and
    e_pq times the ball radius on p, q; p, q then projected back onto their
    balls (alpha1, alpha0), as the dual step would."""
    x = np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1).astype(np.float64)
    x[:, 0:4] += rng.normal(0.0, e_uv, (len(x), 4)) if e_uv > 0 else 0.0
    if e_pq > 0:
        x[:, 4:7] += rng.normal(0.0, e_pq * ALPHA1, (len(x), 3))
        x[:, 7:13] += rng.normal(0.0, e_pq * ALPHA0, (len(x), 6))
    for sl, r in ((slice(4, 7), ALPHA1), (slice(7, 13), ALPHA0)):
        n = np.sqrt((x[:, sl] ** 2).sum(1))
        x[:, sl] *= np.where(n > r, r / np.maximum(n, 1e-30), 1.0)[:, None]
    return np.ascontiguousarray(x, dtype=np.float32)


def bound(ctx, n_scenes, seed0=900000, e_uv_list=(0.0, 0.018), e_pq_list=(0.0, 0.01, 0.03, 0.1, 0.3),
          weights=None):
    """What any per-cube predictor of the initial state could gain: each
    held-out scene is solved cold to the gap (per-depth steps); then every
    level is started from that converged (u, v, p, q) with noise added (e_uv:
    RMS on u and v, the trained network's residual is about 0.018; e_pq: RMS
    on p and q as a fraction of their ball radius) and solved to the same gap
    again. The converged primal with the dual recovered from it (what a
    perfect (u, v) network gives) is run beside it, and, with weights, the
    trained network's (u, v) (from the features of the cold solve's states)
    with the converged dual and with the recovered one."""
    rng = np.random.default_rng(5)
    rows = []
    for k in range(n_scenes):
        sc = eval_scene(seed0 + k)
        hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
        ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
        sched = depth_schedule(int(sc.cubes["depth"].max()))
        pc = params(**sched)
        ctx.run(pc)  # untimed: graphs
        cold = ctx.run(pc)
        st = ctx.state()
        uv = np.ascontiguousarray(np.concatenate([st["u"][:, None], st["v"]], axis=1))
        row = {"seed": seed0 + k, "cubes": int(len(sc.cubes)), "cold_ms": cold.solve_ms,
               "cold_iters": [l["iters"] for l in cold.levels()]}
        r = ctx.run(params(dual_eps=DUAL_EPS, warm_fallback=0, **sched), init_uv=uv)
        row["primal_solution_dual_recovered"] = {"ms": r.solve_ms, "iters": [l["iters"] for l in r.levels()]}
        if weights is not None:
            ctx.mlp_set(weights)
            pred = uv.copy()
            for depth, _ in ctx.levels():
                x, _, idx = ctx.features(depth)
                pred[idx] = x[:, 20:24] + ctx.mlp_forward(x)
            row["network_uv_rms"] = float(np.sqrt(np.mean((pred - uv) ** 2)))
            full = perturbed(st, 0.0, 0.0, rng)
            full[:, 0:4] = pred
            r = ctx.run(params(warm_fallback=0, **sched), init_state=full)
            row["network_uv_exact_dual"] = {"ms": r.solve_ms, "iters": [l["iters"] for l in r.levels()]}
            r = ctx.run(params(dual_eps=DUAL_EPS, warm_fallback=0, **sched), init_uv=pred)
            row["network_uv_dual_recovered"] = {"ms": r.solve_ms, "iters": [l["iters"] for l in r.levels()]}
        for e_uv in e_uv_list:
            for e_pq in e_pq_list:
                r = ctx.run(params(warm_fallback=0, **sched), init_state=perturbed(st, e_uv, e_pq, rng))
                row[f"uv{e_uv}_pq{e_pq}"] = {"ms": r.solve_ms, "iters": [l["iters"] for l in r.levels()],
                                             "reached": bool(all(l["rel_gap"] <= GAP for l in r.levels()))}
        rows.append(row)
        log(f"[warmstart] bound scene {k}: cold {row['cold_ms']:.0f} ms; " + ", ".join(
            f"{key} {v['ms']:.0f}" for key, v in row.items() if isinstance(v, dict)))
    cold = sum(r["cold_ms"] for r in rows)
    keys = [k for k in rows[0] if isinstance(rows[0][k], dict)]
    summary = {"scenes": n_scenes, "cold_ms_total": cold,
               "speedup_vs_cold_total": {k: cold / sum(r[k]["ms"] for r in rows) for k in keys}}
    return summary, rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval", "both", "bound"])
    ap.add_argument("--scenes", type=int, default=2000)
    ap.add_argument("--eval-scenes", type=int, default=50)
    ap.add_argument("--epochs", type=int, default=12)
    ap.add_argument("--cap", type=int, default=4000, help="training pairs kept per level and scene")
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--out", default=None, help="JSON results file")
    ap.add_argument("--diagnostics", action="store_true", help="also run without fallback and from the solution")
    ap.add_argument("--no-graph", action="store_true", help="launch iteration chunks without HIP graphs")
    args = ap.parse_args()
    global USE_GRAPH
    USE_GRAPH = 0 if args.no_graph else 1
    ctx = Context(0)
    result = {}
    if args.cmd in ("train", "both"):
        t0 = time.time()
        X, T = make_training_set(ctx, args.scenes, cap=args.cap)
        t_data = time.time() - t0
        t0 = time.time()
        w, curve = train(ctx, X, T, epochs=args.epochs)
        t_train = time.time() - t0
        base = float(np.mean(T ** 2))
        log(f"[warmstart] {len(X)} pairs from {args.scenes} scenes in {t_data:.0f} s; training {t_train:.1f} s; "
            f"loss {curve.tolist()} (zero-correction loss {base:.4g})")
        os.makedirs(os.path.dirname(os.path.abspath(args.weights)), exist_ok=True)
        np.save(args.weights, w)
        result["train"] = {"scenes": args.scenes, "pairs": int(len(X)), "epochs": args.epochs,
                           "loss_curve": curve.tolist(), "zero_correction_loss": base,
                           "data_s": t_data, "train_s": t_train}
    if args.cmd == "bound":
        summary, rows = bound(ctx, args.eval_scenes,
                              weights=np.load(args.weights) if os.path.exists(args.weights) else None)
        log(f"[warmstart] {json.dumps(summary)}")
        result["bound"] = summary
        result["bound_scenes"] = rows
    if args.cmd in ("eval", "both"):
        w = np.load(args.weights)
        summary, res = evaluate(ctx, w, args.eval_scenes, diagnostics=args.diagnostics)
        log(f"[warmstart] {json.dumps(summary)}")
        result["eval"] = summary
        result["eval_scenes"] = res
    if args.out:
        with open(args.out, "w") as f:
            json.dump(result, f, indent=1)
    print(json.dumps({k: v for k, v in result.items() if k not in ("eval_scenes", "bound_scenes")}))


if __name__ == "__main__":
    main()
