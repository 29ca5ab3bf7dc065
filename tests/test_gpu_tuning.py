"""Launch configurations and the autotuner.

Every launch configuration (kernel variant, workgroup / tile size, tile
order) runs the same arithmetic per cube, so results must be bit-identical;
the autotuner must pick, per level, a configuration at least as fast as the
parameters' own (which is among its candidates) and solves that use its
table must match untuned solves bit for bit.
"""
import numpy as np
import pytest

from tgvreconclaude_amd import BF16, FP32, FP64, OP_ADJ, OP_REF, TGVParams

pytestmark = pytest.mark.gpu


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8)


@pytest.mark.parametrize("fused,block", [(0, 64), (0, 128), (1, 128), (1, 512), (1, 1024), (2, 64), (2, 128),
                                         (3, 128), (3, 256), (3, 1024), (4, 128), (4, 256), (4, 512), (4, 1024),
                                         (5, 128), (5, 256), (5, 512), (5, 1024), (6, 0)])
def test_launch_shapes_identical(ctx, sphere, fused, block):
    sc, hist = sphere
    base = dict(alpha0=2.0, alpha1=1.0, iters=60, precision=FP32, op_mode=OP_ADJ)
    ref, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=0, block=256, **base))
    got, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=fused, block=block, **base))
    assert np.array_equal(_raw(got), _raw(ref))


@pytest.mark.parametrize("prec,mode", [(FP32, OP_ADJ), (BF16, OP_ADJ), (FP32, OP_REF), (FP64, OP_ADJ)])
def test_autotune_choice_and_identity(ctx, sphere, prec, mode):
    sc, hist = sphere
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=60, precision=prec, op_mode=mode)
    ctx.load(sc.cubes, hist, precision=prec, op_mode=mode)
    iters = 10
    levels = ctx.autotune(p, iters=iters)
    assert [(l["depth"], l["leaves"]) for l in levels] == [tuple(x) for x in ctx.levels()]
    for l in levels:
        assert l["variant"] in ((0, 1, 2, 3, 4, 5, 6, 7, 8) if mode == OP_ADJ else (0, 5, 6, 8))
        assert l["block"] in (64, 128, 256, 512, 1024) and l["xcd"] in (0, 1)
        if l["variant"] in (1, 3, 4):
            assert l["block"] in (128, 256, 512, 1024)
        elif l["variant"] == 5:
            assert l["block"] in (128, 256, 512, 1024)
        elif l["variant"] == 6:
            assert l["block"] == 1024 and l["leaves"] <= 1024
        elif l["variant"] == 7:
            assert l["block"] in (256, 1024)
        elif l["variant"] == 8:
            assert l["block"] in (128, 256)
        else:
            assert l["block"] <= 256
        # candidates are ranked by time + 1 us per kernel launch and iteration
        def launches(v, n):
            if v == 0:
                return 2.0
            if v == 5 or (v == 6 and n > 1024):
                return (2.0 + (iters - 1) / 2.0) / iters
            if v == 7:
                return 2.0 / iters
            if v == 8:
                return (2.0 + (iters - 1) / 4.0) / iters
            return 1.0 / iters if v == 6 else (iters + 1.0) / iters
        own = 0 if mode == OP_REF else p.fused
        assert 0 < l["best_us"] and 0 < l["default_us"]
        assert l["best_us"] + launches(l["variant"], l["leaves"]) <= l["default_us"] + launches(own, l["leaves"]) + 1e-9
    ctx.run(p)
    plain = ctx.state()
    p.tune = 1
    ctx.run(p)
    tuned = ctx.state()
    assert np.array_equal(_raw(tuned), _raw(plain))


@pytest.mark.parametrize("prec,tile,graph", [(FP32, 128, 1), (FP32, 256, 1), (FP32, 512, 1), (FP32, 1024, 1),
                                             (FP32, 256, 0),
                                             (BF16, 256, 1), (FP64, 128, 1)])
def test_two_iteration_kernel(ctx, sphere, prec, tile, graph):
    # fused = 5: two iterations per launch over the tabulated tile sets; the
    # finest level's sets fit (so the kernel runs there), and the state equals
    # the two-kernel iteration's bit for bit, for odd and even pass counts
    sc, hist = sphere
    ctx.load(sc.cubes, hist, precision=prec, op_mode=OP_ADJ)
    for depth, n in ctx.levels():
        ok, (sa, sd, sp), lds = ctx.tile_sets(depth, tile)
        if ok:
            assert sa >= sd >= sp >= tile and sa <= 4 * tile and 0 < lds <= 160 * 1024
    assert ctx.tile_sets(ctx.levels()[-1][0], tile)[0]
    for iters in (61, 60):
        base = dict(alpha0=2.0, alpha1=1.0, iters=iters, precision=prec, op_mode=OP_ADJ, use_graph=graph)
        ref, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=0, block=256, **base))
        got, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=5, block=tile, **base))
        assert np.array_equal(_raw(got), _raw(ref))


def test_two_iteration_level_bench(ctx, sphere):
    # the level bench times the two-iteration launch per iteration and leaves
    # the state where the same number of iterations of any variant leaves it
    sc, hist = sphere
    depth = None
    outs = []
    for fused in (5, 0):
        ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
        p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=FP32, op_mode=OP_ADJ, fused=fused, block=256)
        ctx.run(p)
        depth = ctx.levels()[-1][0]
        total, dual, primal, fz = ctx.level_bench(depth, p, 20)
        assert total > 0 and dual > 0 and primal > 0 and fz > 0
        outs.append(ctx.state())
    assert np.array_equal(_raw(outs[0]), _raw(outs[1]))


@pytest.mark.parametrize("prec", [FP32, BF16, FP64])
def test_whole_level_kernel(ctx, sphere, prec):
    # fused = 6: the levels of at most 1024 cubes run all their iterations in
    # one workgroup (the larger ones as fused = 5); the state equals the
    # two-kernel iteration's bit for bit, with and without graphs
    sc, hist = sphere
    for graph in (1, 0):
        base = dict(alpha0=2.0, alpha1=1.0, iters=57, precision=prec, op_mode=OP_ADJ, use_graph=graph)
        ref, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=0, block=256, **base))
        got, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=6, **base))
        assert np.array_equal(_raw(got), _raw(ref))
    assert min(n for _, n in ctx.levels()) <= 1024  # some level ran in one workgroup
