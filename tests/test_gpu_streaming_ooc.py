"""The streaming solve out of core: the state in pinned host memory, at most a
given number of bytes of per-cube data on the device, treetop leaves run in
batches on two alternating slots. Bit for bit the resident streaming solve
(which the oracle pins) for every budget, from many small batches per level
to one."""
import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import TGVParams, scene

pytestmark = pytest.mark.gpu

PER_CUBE = 4 * 28 + 1  # bytes of slot data per local cube


def _bits(st):
    st = np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1) if st.dtype.names else st
    return st.view(np.uint32)


@pytest.mark.parametrize("slot_cubes", [3000, 20000, 10 ** 6])
def test_sphere_out_of_core_equals_resident(ctx, sphere, slot_cubes):
    sc, hist = sphere
    leaves = scene.treetop(sc.cubes, 3000)
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40)
    ref, _ = ctx.solve_streaming(sc.cubes, hist, leaves, p)
    budget = 2 * PER_CUBE * slot_cubes
    st, rep, orep = ctx.solve_streaming_ooc(sc.cubes, hist, leaves, p, device_bytes=budget)
    print(orep)
    assert np.array_equal(_bits(st), _bits(ref))
    assert orep["device_bytes"] <= budget
    nlev = len(np.unique(sc.cubes["depth"]))
    if slot_cubes <= 3000:
        assert orep["batches"] > 2 * nlev  # several batches per level, alternating slots
    assert len(rep.levels()) == nlev


def test_csg_out_of_core_equals_oracle(ctx):
    sc = scene.csg_scene(41, views=12, W=320, H=240, max_depth=7)
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 8)
    leaves = scene.treetop(sc.cubes, 2000)
    p = TGVParams(alpha0=0.5, alpha1=0.3, iters=25)
    st, _, orep = ctx.solve_streaming_ooc(sc.cubes, hist, leaves, p, device_bytes=2 * PER_CUBE * 4000)
    ref = O.tgv_minimize_streaming(sc.cubes, hist, leaves, O.params(alpha0=0.5, alpha1=0.3, iters=25))
    assert np.array_equal(_bits(st), ref.view(np.uint32))
    assert orep["batches"] > len(np.unique(sc.cubes["depth"]))


def test_budget_below_a_leaf_is_refused(ctx, sphere):
    sc, hist = sphere
    leaves = scene.treetop(sc.cubes, 40000)
    with pytest.raises(RuntimeError, match="too small|budget"):
        ctx.solve_streaming_ooc(sc.cubes, hist, leaves, TGVParams(iters=2), device_bytes=2 * PER_CUBE * 100)


@pytest.fixture(scope="module")
def csg9(ctx):
    sc = scene.csg_scene(21, views=24, W=640, H=480, max_depth=9)
    hist = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    assert len(sc.cubes) >= 4 << 16  # the host passes split it over several threads
    return sc, hist


@pytest.mark.parametrize("leaf_max,slot_cubes", [(1 << 16, 150000), (1 << 17, 300000), (1 << 19, 300000)])
def test_large_out_of_core_equals_resident(ctx, csg9, leaf_max, slot_cubes):
    """Several host threads share each pass at this size (input check, parent
    copy, snapshot, gathering and scatter; with leaves or batches of more than
    2^17 cubes of one depth also the batch cut and the local numbering)."""
    sc, hist = csg9
    leaves = scene.treetop(sc.cubes, leaf_max)
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=10)
    ref, _ = ctx.solve_streaming(sc.cubes, hist, leaves, p)
    st, rep, orep = ctx.solve_streaming_ooc(sc.cubes, hist, leaves, p, device_bytes=2 * PER_CUBE * slot_cubes)
    assert np.array_equal(_bits(st), _bits(ref))
    if leaf_max == 1 << 16:
        assert orep["batches"] > len(np.unique(sc.cubes["depth"]))


@pytest.mark.parametrize("parts", [2, 3, 4, 8])
def test_unsorted_input_refused_at_a_range_boundary(ctx, csg9, parts):
    sc, hist = csg9
    cubes = sc.cubes.copy()
    n = len(cubes)
    k = (n + parts - 1) // parts
    # cube k's code below every other (neighbouring codes can be equal)
    cubes["hi"][k] = cubes["lo"][k] = 0
    assert cubes["hi"][k - 1] or cubes["lo"][k - 1]
    with pytest.raises(RuntimeError, match="sorted"):
        ctx.solve_streaming_ooc(cubes, hist, scene.treetop(sc.cubes, 1 << 16), TGVParams(iters=1))


def test_first_bad_cube_in_input_order_is_reported(ctx, csg9):
    sc, hist = csg9
    n = len(sc.cubes)
    leaves = scene.treetop(sc.cubes, 1 << 16)
    early, late = n // 8, n - 2
    code = (sc.cubes["hi"].astype(np.uint64) << 32) | sc.cubes["lo"]
    assert code[early] < code[early + 1] and code[late] < code[late + 1]
    # an unsorted pair before a cube of impossible depth
    cubes = sc.cubes.copy()
    cubes[[early, early + 1]] = cubes[[early + 1, early]]
    cubes["depth"][late] = 30
    with pytest.raises(RuntimeError, match="sorted"):
        ctx.solve_streaming_ooc(cubes, hist, leaves, TGVParams(iters=1))
    # and the other way round
    cubes = sc.cubes.copy()
    cubes["depth"][early] = 30
    cubes[[late, late + 1]] = cubes[[late + 1, late]]
    with pytest.raises(RuntimeError, match="depth"):
        ctx.solve_streaming_ooc(cubes, hist, leaves, TGVParams(iters=1))
