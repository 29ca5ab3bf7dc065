"""Pins of the host octree-input generator's stages 1-3 (build, merge,
balance), the CPU statement the device build is compared with.

Reference tests followed: tests/test_stage1.cpp:84-117 (build: sorted,
cell size rule), tests/test_stages2_4.cpp:147-281 (balance: a single cube is
unchanged; two cubes two levels apart get a bridge and no violation
remains). The balance is also compared, record for record, with a plain
Python FIFO statement (tests/balance_bfs.py) on random inputs, and the
inputs are chosen so that the queue order changes the result (the
reference's balance is order dependent, so the device must reproduce the
queue, not just some 2:1 closure).
"""
import ctypes as C

import numpy as np
import pytest

import balance_bfs as B
from tgvreconclaude_amd import _native as N, scene


def _raw(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _sorted_unique(c):
    k = B.codes(c)
    return bool(np.all((k[1:] > k[:-1]) | ((k[1:] == k[:-1]) & (c["depth"][1:] > c["depth"][:-1]))))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_balance_equals_python_fifo(seed):
    cells = B.random_cells(seed, 250)
    want = B.to_cubes(B.balance(cells, 1.5))
    got = scene.balance(B.to_cubes(cells), 1.5)
    assert np.array_equal(_raw(got), _raw(want))
    assert not B.violations(got)


def test_balance_depends_on_queue_order():
    # seeds 1, 2, 4, 5 of random_cells give different cell sets for the
    # reversed queue (measured); the generator follows the queue it is given
    differs = 0
    for seed in (1, 2):
        cells = B.random_cells(seed, 300)
        fwd = scene.balance(B.to_cubes(cells), 1.5)
        rev = scene.balance(B.to_cubes(cells[::-1]), 1.5)
        assert np.array_equal(_raw(rev), _raw(B.to_cubes(B.balance(cells[::-1], 1.5))))
        a = set(zip(B.codes(fwd).tolist(), fwd["depth"].tolist()))
        b = set(zip(B.codes(rev).tolist(), rev["depth"].tolist()))
        differs += a != b
    assert differs == 2


def test_balance_single_cube_unchanged():
    one = B.to_cubes([(B.interleave(0, 0, 0), 5, np.float32(0.3))])
    got = scene.balance(one, 10.0)
    assert np.array_equal(_raw(got), _raw(one))


def test_balance_two_cubes_get_one_bridge():
    # depth 5 at the origin, depth 7 right of it in +x: the bridge is the
    # depth-6 cell left of the fine cube (test_stages2_4.cpp:191-281)
    step5, step6 = 1 << 16, 1 << 15
    r_root = 100.0
    cells = [(B.interleave(0, 0, 0), 5, np.float32(r_root) / np.float32(32)),
             (B.interleave(step5, 0, 0), 7, np.float32(r_root) / np.float32(128))]
    got = scene.balance(B.to_cubes(cells), r_root)
    assert len(got) == 3 and _sorted_unique(got) and not B.violations(got)
    bridge = got[got["depth"] == 6]
    assert len(bridge) == 1
    assert int(B.codes(bridge)[0]) == B.interleave(step5 - step6, 0, 0)
    assert bridge["r_c"][0] == np.float32(r_root) / np.float32(64)


def test_balance_empty():
    assert len(scene.balance(np.zeros(0, dtype=N.CUBE_DT), 1.5)) == 0


def _identity_camera(fx, cx, cy):
    cam = N.Camera()
    cam.fx = cam.fy = fx
    cam.cx, cam.cy = cx, cy
    for i in range(12):
        cam.cam_to_world[i] = cam.world_to_cam[i] = 1.0 if i in (0, 5, 10) else 0.0
    return cam


def _build(depths, cams, origin, r_root, max_depth, last_stage):
    S = N.scene_lib()
    p = C.c_void_p()
    o = np.asarray(origin, dtype=np.float32)
    V, H, W = depths.shape
    n = S.tgvs_build_octree(N.ptr(depths), cams, V, W, H, N.ptr(o), r_root, max_depth, last_stage, C.byref(p))
    assert n >= 0
    c = np.empty(n, dtype=N.CUBE_DT)
    if n:
        C.memmove(c.ctypes.data, p, n * N.CUBE_DT.itemsize)
    S.tgvs_free(p)
    return c


def reference_build_case():
    """tests/test_stage1.cpp:84-117: one 10 x 10 map at depth 5, K = (50, 50, 5, 5),
    Rt = I, origin (-1, -1, 4), r_root 2.5, max depth 21."""
    cams = (N.Camera * 1)(_identity_camera(50.0, 5.0, 5.0))
    return np.full((1, 10, 10), 5.0, dtype=np.float32), cams, (-1.0, -1.0, 4.0), 2.5


def reference_merge_case():
    """tests/test_stages2_4.cpp:29-44,104-133: two 8 x 8 maps at depths 5 and 6,
    K = (50, 50, 4, 4), Rt = I, origin (-1.5, -1.5, 4), r_root 3."""
    cams = (N.Camera * 2)(_identity_camera(50.0, 4.0, 4.0), _identity_camera(50.0, 4.0, 4.0))
    d = np.stack([np.full((8, 8), 5.0, np.float32), np.full((8, 8), 6.0, np.float32)])
    return d, cams, (-1.5, -1.5, 4.0), 3.0


def test_build_reference_case():
    # the builder's own output (stage 1): non-empty, sorted, every cube obeys the
    # cell size rule 0.75 r <= r_root / 2^d < 1.5 r
    depths, cams, origin, r_root = reference_build_case()
    for stage in (1, 3):
        c = _build(depths, cams, origin, r_root, 21, stage)
        assert len(c) > 0
        k = B.codes(c)
        assert np.all(k[1:] >= k[:-1])
        cell = np.float32(2.5) / (np.float32(2.0) ** c["depth"].astype(np.float32))
        assert np.all(cell >= np.float32(0.75) * c["r_c"]) and np.all(cell < np.float32(1.5) * c["r_c"])


def test_merge_reference_case():
    # two disjoint planes: the merged list is sorted and holds both maps' cubes
    depths, cams, origin, r_root = reference_merge_case()
    a = _build(depths[:1], cams, origin, r_root, 21, 1)
    b = _build(depths[1:], (N.Camera * 1)(cams[1]), origin, r_root, 21, 1)
    m = _build(depths, cams, origin, r_root, 21, 2)
    assert len(a) and len(b) and len(m) == len(a) + len(b)
    assert _sorted_unique(m)


def test_merge_equals_python_restatement():
    # runs of equal (code, depth) in stable (code, depth) order, mean r_c summed
    # in double in that order (merge.cpp:57-92 with the tie order fixed)
    rng = np.random.default_rng(5)
    cells = B.random_cells(11, 120)
    pick = rng.integers(0, len(cells), 600)
    inp = B.to_cubes([cells[i] for i in pick])
    inp["r_c"] = rng.uniform(0.01, 0.2, len(inp)).astype(np.float32)
    got = scene.merge(inp)
    order = sorted(range(len(inp)), key=lambda i: (int(B.codes(inp[i:i + 1])[0]), int(inp["depth"][i])))
    runs = {}
    for i in order:
        runs.setdefault((int(B.codes(inp[i:i + 1])[0]), int(inp["depth"][i])), []).append(float(inp["r_c"][i]))
    want = B.to_cubes([(k[0], k[1], np.float32(sum(v) / len(v))) for k, v in runs.items()])
    assert np.array_equal(_raw(got), _raw(want))


def test_build_sphere_properties():
    c = scene.sphere_scene().cubes
    assert _sorted_unique(c)
    cell = np.float32(1.5) / (np.float32(2.0) ** c["depth"].astype(np.float32))
    assert np.all(cell >= np.float32(0.75) * c["r_c"])
    # the upper bound holds except where max_depth (7) clamps the depth
    assert np.all((cell < np.float32(1.5) * c["r_c"]) | (c["depth"] == 7))
    assert not B.violations(c)
