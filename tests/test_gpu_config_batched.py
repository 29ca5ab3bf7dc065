"""The batched workload itself: 256 independent blocks, each a seeded union of
1-6 spheres and boxes, cube counts log-uniform in [4096, 524288] (about 33 M
cubes in all), 10 bins, solved as one resident problem with the autotuned
launches.

  * with 50 iterations per level, every block's state equals the state of
    solving that block alone, bit for bit;
  * a seeded sample of blocks equals the CPU oracle, bit for bit.
"""
import os

import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import FP32, OP_ADJ, TGVParams, scene

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def blocks(ctx):
    out = []
    for b, tgt in enumerate(scene.batch_targets(256)):
        sc = scene.batch_block(4000 + b, tgt)
        out.append((sc.cubes, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)))
        if b % 32 == 31:
            print(f"{b + 1}/256 blocks generated", flush=True)
    counts = np.array([len(c) for c, _ in out])
    print(f"{counts.sum()} cubes, per block {counts.min()} - {counts.max()}", flush=True)
    return out


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8)


@pytest.fixture(scope="module")
def batched_state(ctx, blocks):
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=FP32, op_mode=OP_ADJ, fused=5, block=1024)
    ctx.load_blocks(blocks, precision=FP32, op_mode=OP_ADJ)
    ctx.autotune(p, iters=50)
    p.tune = 1
    rep = ctx.run(p)
    print(f"batched sweep: {rep.leaf_iters:.3g} leaf-iterations in {rep.solve_ms:.1f} ms", flush=True)
    return p, ctx.split_blocks(ctx.state())


def test_counts_span_the_workload(blocks):
    counts = np.array([len(c) for c, _ in blocks])
    assert len(counts) == 256 and counts.min() < 8192 and counts.max() > 262144
    assert 2.5e7 < counts.sum() < 4.5e7


def test_every_block_equals_its_own_solve(ctx, blocks, batched_state):
    p, parts = batched_state
    for b, ((cubes, hist), got) in enumerate(zip(blocks, parts)):
        ctx.load(cubes, hist, precision=FP32, op_mode=OP_ADJ)
        ctx.run(p)
        assert np.array_equal(_raw(got), _raw(ctx.state())), f"block {b}"


def test_sampled_blocks_equal_oracle(blocks, batched_state):
    _, parts = batched_state
    counts = np.array([len(c) for c, _ in blocks])
    rng = np.random.default_rng(7)
    sample = sorted({int(np.argmin(counts)), int(np.argmax(counts))} | set(rng.choice(256, 3, replace=False).tolist()))
    for b in sample:
        cubes, hist = blocks[b]
        ref = O.tgv_minimize(cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=50, op_mode=1, threads=THREADS))[0]
        g = parts[b]
        got = np.concatenate([g["u"][:, None], g["v"], g["p"], g["q"]], axis=1)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), f"block {b} ({counts[b]} cubes)"
