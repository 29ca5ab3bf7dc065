"""A context keeps the device blocks its calls release and hands them out to
its later allocations. Loads of problems of different sizes one after the
other in one context (blocks reused, some of them larger than asked for)
give every problem the state a fresh context gives it, bit for bit, in both
operator modes; so does the out-of-core solve, whose slots run on streams of
their own."""
import numpy as np
import pytest

from tgvreconclaude_amd import FP32, OP_ADJ, OP_REF, Context, TGVParams, scene

pytestmark = pytest.mark.gpu


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8).reshape(len(st), -1)


@pytest.fixture(scope="module")
def scenes(ctx):
    out = []
    for seed, depth in ((3, 7), (4, 8), (5, 6)):
        sc = scene.csg_scene(seed, views=12, W=320, H=240, max_depth=depth)
        out.append((sc, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 8)))
    return out


@pytest.mark.parametrize("mode", [OP_ADJ, OP_REF])
def test_alternating_sizes_match_a_fresh_context(ctx, scenes, mode):
    p = TGVParams(alpha0=1.0, alpha1=0.5, iters=15, precision=FP32, op_mode=mode, fused=6)
    fresh = []
    with Context(0) as c2:
        for sc, hist in scenes:
            fresh.append(_raw(c2.solve(sc.cubes, hist, p)[0]))
    for rnd in range(2):
        for k in (1, 0, 2, 1):
            sc, hist = scenes[k]
            got = _raw(ctx.solve(sc.cubes, hist, p)[0])
            assert np.array_equal(got, fresh[k]), (rnd, k)


def test_out_of_core_between_loads(ctx, scenes):
    sc, hist = scenes[1]
    leaves = scene.treetop(sc.cubes, 4000)
    p = TGVParams(alpha0=1.0, alpha1=0.5, iters=10)
    want, _ = ctx.solve_streaming(sc.cubes, hist, leaves, p)
    for k in (0, 2):
        ctx.solve(scenes[k][0].cubes, scenes[k][1], TGVParams(iters=5, fused=6))
        st, _, _ = ctx.solve_streaming_ooc(sc.cubes, hist, leaves, p, device_bytes=2 * 113 * 20000)
        assert np.array_equal(_raw(st), _raw(want))
