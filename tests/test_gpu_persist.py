"""The persistent level kernel (fused = 7): a level's iterations in one
cooperative launch, two grid barriers per iteration.

Each thread keeps its cube's state in registers and exchanges u_bar, v_bar
and p, q through the main set between barriers, with k_dual's and
primal_adj_at's expressions, so the state must equal the two-kernel
iteration's bit for bit: on one problem (both workgroup sizes, every
precision), on several blocks iterated to a gap (where converged blocks are
skipped) and against the FP64 oracle. A grid barrier whose wait gives up
(forced here from the host with a spin limit of 0) must fail the solve and
leave the library usable.
"""
import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import BF16, BF16_DUAL, FP32, FP64, OP_ADJ, TGVParams, scene

pytestmark = pytest.mark.gpu


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8)


@pytest.mark.parametrize("prec", [FP32, FP64, BF16, BF16_DUAL])
@pytest.mark.parametrize("block", [1024, 256])
def test_persistent_level_kernel_bit_exact(ctx, sphere, prec, block):
    sc, hist = sphere
    base = dict(alpha0=2.0, alpha1=1.0, iters=57, precision=prec, op_mode=OP_ADJ)
    ref, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=0, block=256, **base))
    got, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=7, block=block, **base))
    assert np.array_equal(_raw(got), _raw(ref))
    # every level of the sphere fits one cooperative launch
    assert max(n for _, n in ctx.levels()) <= 256 * 1024


def test_persistent_level_kernel_fp64_oracle(ctx, sphere):
    sc, hist = sphere
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=FP64, op_mode=OP_ADJ, fused=7)
    st, _ = ctx.solve(sc.cubes, hist, p)
    ref, _, _, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=40, op_mode=1), f64=True)
    assert np.array_equal(st.view(np.uint64), ref.view(np.uint64))


def test_persistent_level_kernel_blocks_to_gap(ctx):
    blocks = []
    for seed, depth in ((11, 7), (12, 6), (13, 7)):
        sc = scene.csg_scene(seed, views=12, W=320, H=240, max_depth=depth)
        blocks.append((sc.cubes, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)))
    out = []
    for fused in (0, 7):
        p = TGVParams(alpha0=2.0, alpha1=1.0, precision=FP32, op_mode=OP_ADJ, gap_tol=1e-3, chunk=50,
                      max_iters=20000, fused=fused, block=1024 if fused else 256)
        ctx.load_blocks(blocks, precision=FP32, op_mode=OP_ADJ)
        ctx.run(p)
        it, _ = ctx.block_report()
        out.append((it.copy(), ctx.state()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(_raw(out[0][1]), _raw(out[1][1]))


def test_grid_barrier_gives_up_and_recovers(ctx, sphere, monkeypatch):
    sc, hist = sphere
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=FP32, op_mode=OP_ADJ, fused=7, block=256)
    monkeypatch.setenv("TGVX_GRID_SPIN", "0")
    with pytest.raises(RuntimeError, match="grid barrier"):
        ctx.solve(sc.cubes, hist, p)
    monkeypatch.delenv("TGVX_GRID_SPIN")
    ref, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=FP32,
                                                 op_mode=OP_ADJ, fused=0, block=256))
    got, _ = ctx.solve(sc.cubes, hist, p)
    assert np.array_equal(_raw(got), _raw(ref))


def test_grid_barrier_bench(ctx):
    us = [ctx.grid_barrier_us(w, 1024, barriers=200) for w in (1, 16, 64)]
    assert all(0.0 < u < 1000.0 for u in us)
    with pytest.raises(RuntimeError, match="workgroups must be"):
        ctx.grid_barrier_us(1 << 20, 1024, barriers=10)  # refused on the host, nothing launched
