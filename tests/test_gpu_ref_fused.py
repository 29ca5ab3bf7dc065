"""The reference's update rules (tgv.cpp:137-232, the drop-in tgv_minimize)
with two iterations per launch over tabulated tile sets, and small levels in
one workgroup: bit for bit the two-kernel iteration and the CPU oracle, for
every tile size, odd and even iteration counts, FP32 and FP64, inputs where
octant-0 children share their parents' codes, and several blocks in one
problem."""
import numpy as np
import pytest

import oracle as O
import test_ref_rules_pins as pins
from tgvreconclaude_amd import BF16, BF16_DUAL, FP32, FP64, OP_REF, TGVParams, scene

pytestmark = pytest.mark.gpu


def _flat(st):
    if st.dtype.names:
        return np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1)
    return np.ascontiguousarray(st)


def _bits(a):
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


@pytest.mark.parametrize("fused,block,iters", [(5, 128, 60), (5, 256, 61), (5, 512, 60), (5, 1024, 61),
                                               (6, 0, 60), (6, 0, 1), (6, 0, 2), (6, 0, 3)])
def test_ref_two_iterations_per_launch_equal_oracle(ctx, sphere, fused, block, iters):
    sc, hist = sphere
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=iters, precision=FP32, op_mode=OP_REF, fused=fused, block=block)
    st, _ = ctx.solve(sc.cubes, hist, p)
    ref = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=iters, op_mode=0))[0]
    bad = np.flatnonzero(np.any(_bits(_flat(st)) != _bits(ref), axis=1))
    assert bad.size == 0, f"{bad.size} cubes differ"


@pytest.mark.parametrize("prec", [FP64, BF16, BF16_DUAL])
@pytest.mark.parametrize("iters", [1, 2, 7, 30])
def test_ref_small_levels_in_one_workgroup(ctx, sphere, prec, iters):
    """Levels of at most 1024 cubes in one workgroup (fused = 6) against the
    two kernels, bit for bit, in every other storage precision."""
    sc, hist = sphere
    base = dict(alpha0=2.0, alpha1=1.0, iters=iters, precision=prec, op_mode=OP_REF)
    want, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=0, block=256, **base))
    got, _ = ctx.solve(sc.cubes, hist, TGVParams(fused=6, **base))
    assert np.array_equal(np.ascontiguousarray(got).view(np.uint8), np.ascontiguousarray(want).view(np.uint8))


def test_ref_fp64_equals_oracle(ctx, sphere):
    sc, hist = sphere
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=100, precision=FP64, op_mode=OP_REF, fused=6)
    st, _ = ctx.solve(sc.cubes, hist, p)
    ref = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=100, op_mode=0), f64=True)[0]
    assert np.array_equal(_bits(_flat(st)), _bits(ref))


@pytest.mark.parametrize("seed", pins.SEEDS)
@pytest.mark.parametrize("fused,block", [(5, 128), (5, 512), (6, 0)])
def test_ref_code_collisions(ctx, seed, fused, block):
    # octant-0 children sharing codes with coarser cubes, in both orders,
    # coarser fallbacks and neighbours of other depths (read frozen or stale);
    # fused = 6: every level of these small problems in one workgroup
    cs, h = pins.scene(seed)
    p = TGVParams(alpha0=0.5, alpha1=0.3, iters=20, precision=FP32, op_mode=OP_REF, fused=fused, block=block)
    st, _ = ctx.solve(cs, h, p)
    ref = O.tgv_minimize(cs, h, O.params(alpha0=0.5, alpha1=0.3, iters=20, op_mode=0))[0]
    assert np.array_equal(_bits(_flat(st)), _bits(ref))


def test_ref_csg_scene_tuned(ctx):
    sc = scene.csg_scene(21, views=24, W=640, H=480, max_depth=8)
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=50, precision=FP32, op_mode=OP_REF, fused=6)
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_REF)
    tuned = ctx.autotune(p, iters=50)
    assert all(l["variant"] in (0, 5, 6, 8) for l in tuned)  # the reference rules' variants
    p.tune = 1
    ctx.run(p)
    got = _flat(ctx.state())
    ref = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=50, op_mode=0, threads=8))[0]
    assert np.array_equal(_bits(got), _bits(ref))


def test_ref_batched_blocks(ctx):
    blocks = []
    for seed, depth in ((31, 7), (32, 6), (33, 7)):
        sc = scene.csg_scene(seed, views=12, W=320, H=240, max_depth=depth)
        blocks.append((sc.cubes, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)))
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=FP32, op_mode=OP_REF, fused=6)
    ctx.load_blocks(blocks, precision=FP32, op_mode=OP_REF)
    ctx.run(p)
    for (cubes, hist), got in zip(blocks, ctx.split_blocks(ctx.state())):
        ref = O.tgv_minimize(cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=40, op_mode=0))[0]
        assert np.array_equal(_bits(_flat(got)), _bits(ref))


def _leaves(n, parts):
    from tgvreconclaude_amd import LEAF_DT
    cuts = np.linspace(0, n, parts + 1).astype(np.int64)
    lv = np.zeros(parts, dtype=LEAF_DT)
    lv["first_idx"] = cuts[:-1]
    lv["last_idx"] = cuts[1:] - 1
    return lv


@pytest.mark.parametrize("fused,block,iters,graph", [(5, 128, 60, 1), (5, 512, 61, 1), (5, 1024, 60, 1),
                                                     (6, 0, 61, 1), (6, 0, 60, 0), (6, 0, 3, 1), (0, 256, 60, 1)])
def test_ref_frozen_leaves_fused_equal_oracle(ctx, sphere, fused, block, iters, graph):
    # tgv_minimize_oc (tgv.cpp:423-511) with each leaf's part of a level as
    # the two-iteration kernel's level: out-of-leaf u_bar, v_bar from the
    # level-start snapshot, out-of-leaf p, q live, leaves one after another
    sc, hist = sphere
    leaves = scene.treetop(sc.cubes, 8000)
    assert len(leaves) > 2
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=iters, precision=FP32, op_mode=OP_REF, fused=fused, block=block,
                  use_graph=graph)
    st, _ = ctx.solve_oc(sc.cubes, hist, leaves, p)
    ref = O.tgv_minimize_oc(sc.cubes, hist, leaves, O.params(alpha0=2.0, alpha1=1.0, iters=iters))
    bad = np.flatnonzero(np.any(_bits(_flat(st)) != _bits(ref), axis=1))
    assert bad.size == 0, f"{bad.size} cubes differ"


def test_ref_frozen_leaves_fp64_equal_oracle(ctx, sphere):
    sc, hist = sphere
    leaves = scene.treetop(sc.cubes, 8000)
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=FP64, op_mode=OP_REF, fused=6)
    st, _ = ctx.solve_oc(sc.cubes, hist, leaves, p)
    ref = O.tgv_minimize_oc(sc.cubes, hist, leaves, O.params(alpha0=2.0, alpha1=1.0, iters=40), f64=True)
    assert np.array_equal(_bits(_flat(st)), _bits(ref))


@pytest.mark.parametrize("seed", pins.SEEDS)
def test_ref_frozen_leaves_code_collisions(ctx, seed):
    # leaves cut across the index range of scenes whose octant-0 children
    # share their parents' codes; the leaves in reverse order as well
    cs, h = pins.scene(seed)
    lv = _leaves(len(cs), 3)
    for order in (lv, lv[::-1].copy()):
        p = TGVParams(alpha0=0.5, alpha1=0.3, iters=21, precision=FP32, op_mode=OP_REF, fused=5, block=128)
        st, _ = ctx.solve_oc(cs, h, order, p)
        ref = O.tgv_minimize_oc(cs, h, order, O.params(alpha0=0.5, alpha1=0.3, iters=21))
        assert np.array_equal(_bits(_flat(st)), _bits(ref))
