"""Iso-surface extraction on the GPU against the oracle's restatement of
marching_cubes.cpp:63-181: the same vertices (bit for bit, same numbering)
and the same triangles."""
import itertools

import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import CUBE_DT, FP32, FP64, OP_ADJ, OP_REF, TGVParams

pytestmark = pytest.mark.gpu


def _cube(ix, iy, iz, depth, r_c=1.0):
    s = 21 - depth
    c = O.morton_encode(ix << s, iy << s, iz << s)
    return (c >> 32, c & 0xFFFFFFFF, r_c, depth)


def _same_mesh(a, b):
    (va, fa), (vb, fb) = a, b
    assert va.shape == vb.shape and fa.shape == fb.shape
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(fa, fb)


def test_closed_box(ctx):
    # tests/test_stage7.cpp:245-273: a negative cube inside positive ones gives
    # a closed box of 12 triangles
    cs, u = [], []
    for iz, iy, ix in itertools.product(range(3), repeat=3):
        cs.append(_cube(ix, iy, iz, 3, 24.0 / 8))
        u.append(-1.0 if (ix, iy, iz) == (1, 1, 1) else 1.0)
    cubes = np.array(cs, dtype=CUBE_DT)
    u = np.array(u, np.float32)
    ctx.load(cubes, np.zeros((len(cubes), 8), np.uint32), precision=FP32, op_mode=OP_REF)
    g = ctx.surface((0, 0, 0), 24.0, u=u)
    assert len(g[1]) == 12 and len(g[0]) == 8
    _same_mesh(g, O.extract_surface(cubes, u, (0, 0, 0), 24.0))


@pytest.mark.parametrize("with_data", [False, True])
def test_sphere_solution_matches_oracle(ctx, sphere, with_data):
    sc, hist = sphere
    st, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=100, op_mode=OP_ADJ))
    u = st["u"].copy()
    has = (hist.sum(axis=1) > 0).astype(np.uint8) if with_data else None
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    g = ctx.surface(sc.origin, sc.r_root, u=u, has_data=has)
    o = O.extract_surface(sc.cubes, u, sc.origin, sc.r_root, has)
    assert len(o[1]) > 1000
    _same_mesh(g, o)


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_resident_state_surface(ctx, sphere, prec):
    # u = None: the problem's own state after a solve (FP64 rounded to float)
    sc, hist = sphere
    ctx.load(sc.cubes, hist, precision=prec, op_mode=OP_ADJ)
    ctx.run(TGVParams(alpha0=2.0, alpha1=1.0, iters=80, precision=prec, op_mode=OP_ADJ))
    st = ctx.state()
    u = (st["u"] if st.dtype.names else st[:, 0]).astype(np.float32)
    has = (hist.sum(axis=1) > 0).astype(np.uint8)
    g = ctx.surface(sc.origin, sc.r_root, has_data=has)
    _same_mesh(g, O.extract_surface(sc.cubes, u, sc.origin, sc.r_root, has))


def test_random_signs_dedup(ctx):
    # a 2:1 mix of depths with random signs: coarse fallbacks, pairs met from
    # both sides and from two faces of one cube
    rng = np.random.default_rng(3)
    cells = set()
    for _ in range(3000):
        d = int(rng.integers(3, 6))
        x, y, z = (int(v) for v in rng.integers(0, 1 << d, 3))
        cells.add((d, x, y, z))
    lst = sorted(cells, key=lambda c: (O.morton_encode(c[1] << (21 - c[0]), c[2] << (21 - c[0]),
                                                       c[3] << (21 - c[0])), c[0]))
    cubes = np.array([_cube(x, y, z, d) for d, x, y, z in lst], dtype=CUBE_DT)
    u = rng.uniform(-1, 1, len(cubes)).astype(np.float32)
    ctx.load(cubes, np.zeros((len(cubes), 8), np.uint32), precision=FP32, op_mode=OP_REF)
    g = ctx.surface((-1.5, -1.5, -1.5), 1.5, u=u)
    o = O.extract_surface(cubes, u, (-1.5, -1.5, -1.5), 1.5)
    assert len(o[1]) > 1000
    _same_mesh(g, o)
