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


def test_precisions_on_the_sphere_workload(ctx, sphere):
    # the sphere workload's precision report (FP32, BF16 p, q with FP32 u, v,
    # BF16 everything) against the FP64 oracle solved to the same gap
    from accuracy import config3
    sc, hist = sphere
    res = config3(ctx, sc, hist, threads=8, cap=3000)
    r = {x["precision"]: x for x in res["results"]}
    print({k: (v["reached"], max(v["final_rel_gap"]), v["linf_u_same_iters"], v["hausdorff_leaf"])
           for k, v in r.items()})
    assert r["fp32"]["reached"]
    assert r["fp32"]["linf_u_same_iters"] <= 1e-4
    assert r["fp32"]["hausdorff_leaf"] <= 0.5
    # 8 significand bits stall the gap: p, q in BF16 well above 1e-4, all
    # fields in BF16 higher still; the storage error grows in the same order
    assert not r["bf16-dual"]["reached"] and not r["bf16"]["reached"]
    assert max(r["bf16-dual"]["final_rel_gap"]) < max(r["bf16"]["final_rel_gap"])
    assert r["fp32"]["linf_u_same_iters"] < r["bf16-dual"]["linf_u_same_iters"] < r["bf16"]["linf_u_same_iters"]
