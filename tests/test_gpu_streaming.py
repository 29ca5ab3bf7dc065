"""tgv_minimize_streaming on the GPU against the oracle's restatement.

Bit-exact in FP32 and FP64 with many treetop leaves (the leaf-local lookups,
the cube file's bisection, level-start snapshots, zero p, q across leaf
borders and the leaf-by-leaf parent copy all show up in the sphere scene,
whose octant-0 children share their parents' codes); the hand-derived parent
copy of the oracle pins holds on the GPU too.
"""
import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import CUBE_DT, FP32, FP64, LEAF_DT, OP_ADJ, TGVParams, scene

pytestmark = pytest.mark.gpu


def _cube(ix, iy, iz, depth, r_c=1.0):
    s = 21 - depth
    c = O.morton_encode(ix << s, iy << s, iz << s)
    return (c >> 32, c & 0xFFFFFFFF, r_c, depth)


def _leaves(*ranges):
    lv = np.zeros(len(ranges), dtype=LEAF_DT)
    for k, (a, b) in enumerate(ranges):
        lv[k]["first_idx"], lv[k]["last_idx"] = a, b
    return lv


def _bits(st):
    st = np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1) if st.dtype.names else st
    return st.view(np.uint32 if st.dtype == np.float32 else np.uint64)


def test_single_leaf_chain_equals_in_memory(ctx):
    cubes = np.array([_cube(0, 0, 0, 5), _cube(1, 0, 0, 5), _cube(2, 0, 0, 5)], dtype=CUBE_DT)
    h = np.zeros((3, 8), np.uint32)
    h[0, 6] = 50
    h[2, 1] = 50
    p = TGVParams(alpha0=0.1, alpha1=0.2, iters=50)
    a, _ = ctx.solve(cubes, h, p)
    b, _ = ctx.solve_streaming(cubes, h, _leaves((0, 2)), p)
    assert np.array_equal(_bits(a), _bits(b))


def test_parent_copy_follows_the_bisection(ctx):
    cubes = np.array([_cube(0, 0, 0, 1), _cube(0, 0, 0, 2), _cube(0, 0, 0, 3), _cube(0, 0, 0, 4),
                      _cube(1, 0, 0, 3)], dtype=CUBE_DT)
    h = np.zeros((5, 8), np.uint32)
    for i, k in enumerate((0, 2, 4, 6, 7)):
        h[i, k] = 5
    m = [O.median_indicator(x) for x in h]
    st, _ = ctx.solve_streaming(cubes, h, _leaves((0, 4)), TGVParams(iters=0))
    assert st["u"].tolist() == [m[0], m[2], m[2], m[2], m[2]]


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_sphere_many_leaves_bit_exact(ctx, sphere, prec):
    sc, hist8 = sphere
    hist = hist8
    leaves = scene.treetop(sc.cubes, 3000)
    assert len(leaves) > 5
    p = TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=prec)
    st, rep = ctx.solve_streaming(sc.cubes, hist, leaves, p)
    ref = O.tgv_minimize_streaming(sc.cubes, hist, leaves, O.params(alpha0=2.0, alpha1=1.0, iters=40),
                                   f64=prec == FP64)
    assert np.array_equal(_bits(st), ref.view(np.uint32 if prec == FP32 else np.uint64))
    assert len(rep.levels()) == len(np.unique(sc.cubes["depth"]))


def test_rejects_unsorted_and_adjoint_mode(ctx):
    cubes = np.array([_cube(1, 0, 0, 5), _cube(0, 0, 0, 5)], dtype=CUBE_DT)
    h = np.ones((2, 8), np.uint32)
    with pytest.raises(RuntimeError, match="sorted|reference rules"):
        ctx.solve_streaming(cubes, h, _leaves((0, 1)), TGVParams(iters=2))
    with pytest.raises(RuntimeError, match="sorted|reference rules"):
        ctx.solve_streaming(cubes[::-1].copy(), h, _leaves((0, 1)), TGVParams(iters=2, op_mode=OP_ADJ))
