"""The device octree build (build + merge + balance) against the host
generator's stages 1-3, record for record (code, depth and r_c bits).

The balance runs as rounds of parallel checks that must land on the FIFO
queue's own result; the random inputs here change their result when the
queue is reversed, so a balance that computed some other 2:1 closure, or
ignored the queue order, fails them.
"""
import numpy as np
import pytest

import balance_bfs as B
from tgvreconclaude_amd import CUBE_DT, scene

pytestmark = pytest.mark.gpu


def _raw(a):
    return np.ascontiguousarray(a).view(np.uint8)


@pytest.mark.parametrize("seed", [0, 1, 2, 4, 5])
def test_balance_random_cells_both_queue_orders(ctx, seed):
    cells = B.random_cells(seed, 300)
    for order in (cells, cells[::-1]):
        inp = B.to_cubes(order)
        got, rep = ctx.octree_balance(inp, 1.5)
        want = scene.balance(inp, 1.5)
        assert np.array_equal(_raw(got), _raw(want))
        assert rep["bridges"] == len(want) - len(inp) and rep["rounds"] >= 1


def test_balance_known_cases(ctx):
    step5, step6 = 1 << 16, 1 << 15
    two = B.to_cubes([(B.interleave(0, 0, 0), 5, np.float32(100.0 / 32)),
                      (B.interleave(step5, 0, 0), 7, np.float32(100.0 / 128))])
    got, rep = ctx.octree_balance(two, 100.0)
    assert np.array_equal(_raw(got), _raw(scene.balance(two, 100.0)))
    assert len(got) == 3 and rep["bridges"] == 1
    assert int(B.codes(got[got["depth"] == 6])[0]) == B.interleave(step5 - step6, 0, 0)
    one = B.to_cubes([(B.interleave(0, 0, 0), 5, np.float32(0.3))])
    got, rep = ctx.octree_balance(one, 10.0)
    assert np.array_equal(_raw(got), _raw(one)) and rep["bridges"] == 0
    got, rep = ctx.octree_balance(np.zeros(0, dtype=CUBE_DT), 1.5)
    assert len(got) == 0


def test_balance_repeated_cells(ctx):
    # a cell given twice is queued twice and written twice, in input order
    cells = B.random_cells(7, 200)
    cells = cells + cells[:50]
    inp = B.to_cubes(cells)
    inp["r_c"][200:] = np.float32(0.5)
    got, _ = ctx.octree_balance(inp, 1.5)
    assert np.array_equal(_raw(got), _raw(scene.balance(inp, 1.5)))


def test_balance_coarse_and_unaligned_inputs(ctx):
    # depth-0 / depth-1 cells only (the code sort has no bits to sort on)
    inp = B.to_cubes([(0, 0, np.float32(1.5)), (1 << 60, 1, np.float32(0.75)), (0, 1, np.float32(0.75))])
    got, _ = ctx.octree_balance(inp, 1.5)
    assert np.array_equal(_raw(got), _raw(scene.balance(inp, 1.5)))
    # a code with bits below its cell size is not a cube the builder writes
    bad = B.to_cubes([(B.interleave(1, 0, 0), 19, np.float32(0.1))])
    with pytest.raises(RuntimeError, match="below its cell size"):
        ctx.octree_balance(bad, 1.5)


def test_balance_thinned_scene(ctx):
    # a real octree with 40 % of its cubes dropped: many violations at once
    c = scene.sphere_scene().cubes
    keep = np.random.default_rng(3).random(len(c)) >= 0.4
    inp = c[keep]
    got, rep = ctx.octree_balance(inp, 1.5)
    want = scene.balance(inp, 1.5)
    assert rep["bridges"] > 0
    assert np.array_equal(_raw(got), _raw(want))


@pytest.mark.parametrize("which", ["sphere", "csg"])
def test_build_equals_host(ctx, which):
    if which == "sphere":
        sc = scene.sphere_scene()
        md = 7
    else:
        sc = scene.csg_scene(21, views=12, W=320, H=240, max_depth=8)
        md = 8
    got, rep = ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, md)
    assert np.array_equal(_raw(got), _raw(sc.cubes))
    assert rep["pixels"] == sc.depths.size
    assert rep["raw"] >= rep["merged"] and rep["merged"] + rep["bridges"] == len(got)


@pytest.mark.parametrize("stage", [1, 2])
def test_build_stages_equal_host(ctx, stage):
    sc = scene.sphere_scene()
    got, rep = ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, 7, last_stage=stage)
    want = scene.build_octree(sc.depths, sc.cams, 7, last_stage=stage)
    assert np.array_equal(_raw(got), _raw(want))
    assert len(got) == (rep["raw"] if stage == 1 else rep["merged"])


def test_reference_build_and_merge_cases(ctx):
    from test_octree_pins import _build, reference_build_case, reference_merge_case
    for case, stages in ((reference_build_case, (1, 3)), (reference_merge_case, (1, 2, 3))):
        depths, cams, origin, r_root = case()
        for stage in stages:
            got, _ = ctx.octree_build(depths, cams, origin, r_root, 21, last_stage=stage)
            assert np.array_equal(_raw(got), _raw(_build(depths, cams, origin, r_root, 21, stage)))


def test_merge_equals_host(ctx):
    rng = np.random.default_rng(9)
    cells = B.random_cells(12, 300)
    inp = B.to_cubes([cells[i] for i in rng.integers(0, len(cells), 3000)])
    inp["r_c"] = rng.uniform(0.01, 0.2, len(inp)).astype(np.float32)
    got, rep = ctx.octree_merge(inp)
    assert np.array_equal(_raw(got), _raw(scene.merge(inp)))
    assert rep["raw"] == len(inp) and rep["merged"] == len(got)
    got, _ = ctx.octree_merge(inp[:0])
    assert len(got) == 0


def test_build_rejects_bad_arguments(ctx):
    sc = scene.sphere_scene(views=2)
    with pytest.raises(RuntimeError, match="max_depth"):
        ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, 22)
    with pytest.raises(RuntimeError, match="r_root"):
        ctx.octree_build(sc.depths, sc.cams, sc.origin, 0.0, 7)
    with pytest.raises(RuntimeError, match="last_stage"):
        ctx.octree_build(sc.depths, sc.cams, sc.origin, sc.r_root, 7, last_stage=4)
