"""Problem load of inputs large enough for the host's chunked passes (the
level-major counting sort and the input-order test run on several host
threads): one level per depth with all of its cubes, and input in any order
gives each cube the same state, bit for bit (exact-adjoint rules, no repeated
(code, depth) pair), whether the load searches the input as it is (sorted by
(code, depth)) or sorts it first."""
import numpy as np
import pytest

from tgvreconclaude_amd import FP32, OP_ADJ, TGVParams, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csg9(ctx):
    sc = scene.csg_scene(21, views=24, W=640, H=480, max_depth=9)
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    assert len(sc.cubes) >= 4 << 16  # four host chunks at least
    return sc, hist


def _raw(st):
    return np.ascontiguousarray(st).view(np.uint8).reshape(len(st), -1)


def _params():
    return TGVParams(alpha0=2.0, alpha1=1.0, iters=20, precision=FP32, op_mode=OP_ADJ, fused=6)


def test_levels_by_depth(ctx, csg9):
    sc, hist = csg9
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    got = [(d, n) for d, n in ctx.levels()]
    depths, counts = np.unique(sc.cubes["depth"], return_counts=True)
    assert got == [(int(d), int(n)) for d, n in zip(depths, counts)]


def test_shuffled_input_same_state(ctx, csg9):
    sc, hist = csg9
    want, _ = ctx.solve(sc.cubes, hist, _params())
    perm = np.random.default_rng(5).permutation(len(sc.cubes))
    got, _ = ctx.solve(sc.cubes[perm], hist[perm], _params())
    assert np.array_equal(_raw(got), _raw(want)[perm])


@pytest.mark.parametrize("chunks", [2, 3, 4, 8])
def test_order_broken_at_a_chunk_boundary(ctx, csg9, chunks):
    """Two neighbouring cubes swapped where the host's order test splits the
    input: the load must see the input as unsorted and sort it."""
    sc, hist = csg9
    n = len(sc.cubes)
    want, _ = ctx.solve(sc.cubes, hist, _params())
    k = (n + chunks - 1) // chunks
    perm = np.arange(n)
    perm[[k - 1, k]] = perm[[k, k - 1]]
    got, _ = ctx.solve(sc.cubes[perm], hist[perm], _params())
    assert np.array_equal(_raw(got), _raw(want)[perm])
