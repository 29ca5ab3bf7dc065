"""The depth-10 workload itself: torus (R = 1, r = 0.3) union a rotated box,
96 seeded views of 1280x960 on the radius-3 sphere, depth noise 0.003, 3 %
outliers, octree clamped at depth 10 (about 6 M cubes), 10 bins.

  * FP32 solved to relative duality gap 1e-4 against FP64 run for the same
    per-level iterations: u within 1e-4, surface Hausdorff within half a
    finest leaf (the workload's accuracy bar);
  * one 50-iteration chunk per level on the GPU against the CPU oracle at
    full size, bit for bit: FP64 in the exact-adjoint mode, and FP32 with the
    reference's own update rules (tgv.cpp:137-232, the drop-in path);
  * the two-iteration kernel at full size against the two-kernel iteration,
    bit for bit, every storage precision, both rule sets, three tile sizes.
"""
import os

import numpy as np
import pytest

import oracle as O
from accuracy import compare
from tgvreconclaude_amd import BF16, BF16_DUAL, FP32, FP64, OP_ADJ, OP_REF, TGVParams, scene, step_schedule

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def torus_box(ctx):
    sc = scene.torus_box_scene()
    print(f"torus + box scene: {len(sc.cubes)} cubes", flush=True)
    assert 4_000_000 <= len(sc.cubes) <= 8_000_000
    return sc, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)


def _steps(sc):
    d = int(sc.cubes["depth"].max())
    ratios = [256] * (d - 3) + [1024] * 3 + [4096]
    sd, td = step_schedule(ratios)
    return dict(sigma=8.0, tau=1.0 / 128, sigma_depth=sd, tau_depth=td)


def test_fp32_to_gap_against_fp64(ctx, torus_box):
    sc, hist = torus_box
    r = compare(ctx, sc, hist, FP32, tol=1e-4, **_steps(sc))
    print({k: r[k] for k in ("iters_per_level", "linf_u", "rms_u", "hausdorff_leaf", "solve_ms")}, flush=True)
    assert r["reached"]
    assert r["linf_u"] <= 1e-4
    assert r["hausdorff_leaf"] <= 0.5


def _flat(st):
    if st.dtype.names:
        return np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1)
    return np.ascontiguousarray(st[:, :13])


@pytest.mark.parametrize("prec,mode", [(FP64, OP_ADJ), (FP32, OP_REF)])
def test_one_chunk_per_level_equals_oracle(ctx, torus_box, prec, mode):
    sc, hist = torus_box
    # the benchmark's launches: two iterations per launch (the smallest levels
    # in one workgroup in the exact-adjoint mode)
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=prec, op_mode=mode, fused=6)
    ctx.load(sc.cubes, hist, precision=prec, op_mode=mode)
    rep = ctx.run(p)
    got = _flat(ctx.state())
    print(f"gpu {rep.solve_ms:.0f} ms; oracle on {THREADS} threads", flush=True)
    ref = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=50, op_mode=mode, threads=THREADS),
                         f64=prec == FP64)[0]
    word = np.uint64 if prec == FP64 else np.uint32
    same = got.view(word) == np.ascontiguousarray(ref).view(word)
    assert same.all(), f"{(~same).any(axis=1).sum()} cubes differ"
    assert np.abs(ref[:, 4:]).max() > 0


@pytest.mark.parametrize("prec", [FP32, FP64, BF16, BF16_DUAL])
@pytest.mark.parametrize("mode", [OP_ADJ, OP_REF])
def test_two_iteration_tiles_at_full_size(ctx, torus_box, prec, mode):
    """The two-iteration kernel at the workload's own sizes, where its finest
    levels split into two LDS classes with tables in launch order, equals the
    two-kernel iteration bit for bit for tiles of 256, 512 and 1024 (head
    launch, two-iteration launches and the single last steps: 7 iterations)."""
    sc, hist = torus_box
    ctx.load(sc.cubes, hist, precision=prec, op_mode=mode)
    base = dict(alpha0=2.0, alpha1=1.0, iters=7, precision=prec, op_mode=mode, no_energy=1, use_graph=0)
    ctx.run(TGVParams(fused=0, block=256, **base))
    want = np.ascontiguousarray(ctx.state()).view(np.uint8)
    for block in (1024, 512, 256):
        ctx.run(TGVParams(fused=5, block=block, **base))
        got = np.ascontiguousarray(ctx.state()).view(np.uint8)
        assert np.array_equal(got, want), block
