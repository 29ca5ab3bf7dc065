"""FP32 and BF16 storage against FP64 run for the same iterations (a depth-8
scene; the depth-10 report is tests/accuracy.py)."""
import pytest

from accuracy import compare
from tgvreconclaude_amd import BF16, FP32, scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def csg8(ctx):
    sc = scene.csg_scene(7, views=24, W=640, H=480, max_depth=8)
    return sc, ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)


def test_fp32_against_fp64(ctx, csg8):
    sc, hist = csg8
    r = compare(ctx, sc, hist, FP32, tol=1e-4)
    print(r)
    assert r["reached"]
    assert r["linf_u"] <= 1e-4
    assert r["hausdorff_leaf"] <= 0.5


def test_bf16_against_fp64(ctx, csg8):
    # BF16 storage (8 significand bits) stalls the iteration at a relative gap
    # of a few 1e-2, far above 1e-4: it is compared at the iteration counts
    # of the FP32 to-gap solve, and its error is bounded loosely
    sc, hist = csg8
    r32 = compare(ctx, sc, hist, FP32, tol=1e-4)
    r = compare(ctx, sc, hist, BF16, level_iters=r32["iters_per_level"])
    print(r)
    assert r["iters_per_level"] == r32["iters_per_level"]
    assert r["linf_u"] <= 0.05
    assert r["hausdorff_leaf"] <= 4.0
