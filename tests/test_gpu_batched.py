"""Several independent octrees solved as one resident problem.

Every block's cubes only see cubes of the same block, so with a fixed
iteration count a batched solve must give each block exactly the state that
solving it alone gives (bit for bit). With a gap target, a block stops being
updated at the chunk boundary where its own gap first reaches the target, so
its iteration counts and final state equal those of its own solve.
"""
import numpy as np
import pytest

from tgvreconclaude_amd import FP32, FP64, OP_ADJ, OP_REF, TGVParams, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def blocks(ctx):
    out = []
    for seed, depth in ((11, 7), (12, 6), (13, 7)):
        sc = scene.csg_scene(seed, views=12, W=320, H=240, max_depth=depth)
        out.append((sc.cubes, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)))
    return out


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8)


@pytest.mark.parametrize("mode,prec", [(OP_ADJ, FP32), (OP_REF, FP32), (OP_ADJ, FP64)])
def test_batched_fixed_iterations_bit_exact(ctx, blocks, mode, prec):
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=120, precision=prec, op_mode=mode)
    ctx.load_blocks(blocks, precision=prec, op_mode=mode)
    ctx.run(p)
    parts = ctx.split_blocks(ctx.state())
    for (cubes, hist), got in zip(blocks, parts):
        alone, _ = ctx.solve(cubes, hist, p)
        assert np.array_equal(_raw(got), _raw(alone))


def test_solve_batched_entry_matches_resident(ctx, blocks):
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=80, precision=FP32, op_mode=OP_ADJ)
    states, _ = ctx.solve_batched(blocks, p)
    for (cubes, hist), got in zip(blocks, states):
        alone, _ = ctx.solve(cubes, hist, p)
        assert np.array_equal(_raw(got), _raw(alone))


@pytest.mark.parametrize("fused", [1, 4, 5, 6])
def test_batched_to_gap_per_block(ctx, blocks, fused):
    p = TGVParams(alpha0=2.0, alpha1=1.0, precision=FP32, op_mode=OP_ADJ, gap_tol=1e-3, chunk=50,
                  max_iters=20000, fused=fused)
    ctx.load_blocks(blocks, precision=FP32, op_mode=OP_ADJ)
    rep = ctx.run(p)
    it, gap = ctx.block_report()
    parts = ctx.split_blocks(ctx.state())
    depths = [d for d, _ in ctx.levels()]
    assert it.shape == (len(depths), len(blocks))
    work = 0
    for b, (cubes, hist) in enumerate(blocks):
        have = set(np.unique(cubes["depth"]).tolist())
        alone_state, alone = ctx.solve(cubes, hist, p)
        own = {l["depth"]: l["iters"] for l in alone.levels()}
        for li, d in enumerate(depths):
            if d not in have:
                assert it[li, b] == -1
                continue
            assert 0 <= it[li, b] <= 20000
            assert gap[li, b] <= 1e-3
            # a block that reached the gap stops where its own solve stops
            assert it[li, b] == own[d]
            work += int((cubes["depth"] == d).sum()) * int(it[li, b])
        assert np.array_equal(_raw(parts[b]), _raw(alone_state))
    # a level runs until its slowest block is done; the work counts each
    # block's own iterations
    lv = rep.levels()
    for li in range(len(depths)):
        assert lv[li]["iters"] == it[li][it[li] >= 0].max()
    assert rep.leaf_iters == work
