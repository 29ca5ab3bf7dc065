"""Pins of the CPU oracle against the reference's own known answers and
against what the mathematics fixes (CPU only).

Known answers cited from the reference's tests (paths relative to its root);
the rest are closed forms, exact fixed points, brute force on tiny inputs and
an independent dense-matrix statement of the level operator.
"""
import itertools

import numpy as np
import pytest

import oracle as O

CUBE = O.CUBE_DT


def cube(ix, iy, iz, depth, r_c=1.0):
    s = 21 - depth
    c = O.morton_encode(ix << s, iy << s, iz << s)
    return (c >> 32, c & 0xFFFFFFFF, r_c, depth)


def cubes_of(lst):
    return np.array(lst, dtype=CUBE)


# ---------------------------------------------------------------- Morton

def _py_encode(x, y, z):
    r = 0
    for b in range(21):
        r |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return r


def test_morton_roundtrip_known_values():
    # tests/test_stage1.cpp:20-33
    for ix in (0, 1, 127, 1023, 2097151):
        for iy in (0, 5, 255):
            for iz in (0, 3, 100):
                c = O.morton_encode(ix, iy, iz)
                assert O.morton_decode(c) == (ix, iy, iz)
                assert c == _py_encode(ix, iy, iz)


def test_morton_order_is_z_curve():
    # tests/test_stage1.cpp:36-42
    for i in range(99):
        assert O.morton_encode(i, 7, 3) < O.morton_encode(i + 1, 7, 3)
    # bit layout (types.hpp:10): bit0 = x0, bit1 = y0, bit2 = z0
    assert O.morton_encode(1, 0, 0) == 1 and O.morton_encode(0, 1, 0) == 2 and O.morton_encode(0, 0, 1) == 4


def test_parent_code_known_values():
    # tests/test_stages2_4.cpp:47-67
    code = O.morton_encode(5 << 11, 0, 0)
    p9 = O.parent_code(code, 10)
    assert O.morton_decode(p9) == (2 << 12, 0, 0)
    assert p9 != code
    p7 = O.parent_code(O.parent_code(p9, 9), 8)
    assert O.morton_decode(p7) == (0, 0, 0)


def test_face_neighbor_known_values():
    # tests/test_stages2_4.cpp:71-98
    step = 1 << 17
    c = O.morton_encode(step, step, step)
    assert O.morton_decode(O.face_neighbor(c, 0, 4)) == (2 * step, step, step)
    assert O.morton_decode(O.face_neighbor(c, 1, 4))[0] == 0
    none = 0xFFFFFFFFFFFFFFFF
    assert O.face_neighbor(O.morton_encode(0, 0, 0), 1, 4) == none
    assert O.face_neighbor(O.morton_encode(15 * step, 0, 0), 0, 4) == none


def test_face_neighbor_brute_force():
    # every face of every cell of a depth-3 grid against coordinate arithmetic
    d, n = 3, 8
    s = 1 << (21 - d)
    for x, y, z in itertools.product(range(n), repeat=3):
        c = O.morton_encode(x * s, y * s, z * s)
        for f in range(6):
            a, sign = f // 2, (1 if f % 2 == 0 else -1)
            p = [x, y, z]
            p[a] += sign
            got = O.face_neighbor(c, f, d)
            if 0 <= p[a] < n:
                assert got == O.morton_encode(p[0] * s, p[1] * s, p[2] * s)
            else:
                assert got == 0xFFFFFFFFFFFFFFFF


# ---------------------------------------------------------------- histograms

def test_median_mean_known_values():
    # tests/test_stage5.cpp:84-101
    h = np.zeros(8, np.uint32)
    h[4] = 10
    assert O.mean_indicator(h) == pytest.approx(0.125, abs=0.01)
    assert O.median_indicator(h) == np.float32(0.125)
    h2 = np.zeros(8, np.uint32)
    h2[3] = 5
    h2[4] = 5
    assert O.median_indicator(h2) == np.float32(-0.125)
    z = np.zeros(8, np.uint32)
    assert O.median_indicator(z) == 0.0 and O.mean_indicator(z) == 0.0


def test_median_brute_force():
    # definition: centre of the first bin whose cumulative count reaches half the total
    rng = np.random.default_rng(0)
    for nb in (8, 10):
        for _ in range(300):
            h = rng.integers(0, 4, nb).astype(np.uint32) * (rng.random(nb) < 0.5)
            tot = int(h.sum())
            if tot == 0:
                assert O.median_indicator(h) == 0.0
                continue
            k = int(np.argmax(2 * np.cumsum(h) >= tot))
            assert O.median_indicator(h, f64=True) == -1.0 + (k + 0.5) / (nb / 2)


def _plane_view(w, h, depth, fx, cx, cy):
    m = np.eye(3, 4, dtype=np.float32)  # Rt = identity: camera at the origin, z forward
    return O.view(np.full((h, w), depth, np.float32), fx, fx, cx, cy, m)


def test_mip_build_constant_map():
    # tests/test_stage5.cpp:21-41
    v, keep = _plane_view(8, 8, 10.0, 100, 4, 4)
    lv = O.mip_build(v)
    assert len(lv) >= 4 and [x.shape[1] for x in lv[:3]] == [8, 4, 2]
    for x in lv:
        assert np.all(x == 10.0)


def test_mip_build_brute_force():
    rng = np.random.default_rng(1)
    d = rng.uniform(1, 5, (13, 11)).astype(np.float32)
    d[rng.random(d.shape) < 0.3] = np.nan
    d[0, 0] = -1.0
    v, keep = O.view(d, 10, 10, 5, 5, np.eye(3, 4))
    lv = O.mip_build(v)
    prev = d
    for nxt in lv[1:]:
        H, W = nxt.shape
        assert (H, W) == (max(1, prev.shape[0] // 2), max(1, prev.shape[1] // 2))
        for y in range(H):
            for x in range(W):
                s, c = np.float32(0), 0
                for dv in range(2):
                    for du in range(2):
                        py, px = 2 * y + dv, 2 * x + du
                        if py < prev.shape[0] and px < prev.shape[1]:
                            q = prev[py, px]
                            if np.isfinite(q) and q > 0:
                                s = np.float32(s + q)
                                c += 1
                want = np.float32(s / np.float32(c)) if c else np.nan
                got = nxt[y, x]
                assert (np.isnan(want) and np.isnan(got)) or want == got
        prev = nxt


def test_project_known_cases():
    # tests/test_stage5.cpp:45-80
    v, keep = _plane_view(32, 32, 5.0, 100, 16, 16)
    ok, lvl, u, w, cz = O.project(v, (0, 0, 5), 0.04)
    assert ok and lvl == 0 and cz == pytest.approx(5.0) and (u, w) == (16, 16)
    assert not O.project(v, (0, 0, -1), 0.1)[0]
    assert not O.project(v, (1000, 0, 5), 0.1)[0]
    ok, lvl, u, w, cz = O.project(v, (0, 0, 1), 2.0)
    assert ok and lvl >= 1


def test_single_cube_single_view_votes_bin3():
    # tests/test_stage5.cpp:105-146: a = 5 - 5.0625, delta = 0.375,
    # (a/delta + 1) * 4 = 3.33 -> exactly one vote, in bin 3 (bin 4 with 10 bins).
    v, keep = _plane_view(16, 16, 5.0, 50, 8, 8)
    c = cubes_of([cube(4, 4, 8, 4, r_c=1.0 / 16)])
    h8 = O.histograms(c, [v], (-0.5, -0.5, 4.0), 1.0, 8)
    assert h8.tolist() == [[0, 0, 0, 1, 0, 0, 0, 0]]
    h10 = O.histograms(c, [v], (-0.5, -0.5, 4.0), 1.0, 10)
    assert h10[0, 4] == 1 and h10.sum() == 1


def test_histogram_brute_force_fp64():
    # per (cube, view) pair recomputed in double: the float path may only
    # disagree where the double value sits on a bin edge
    rng = np.random.default_rng(2)
    v, keep = _plane_view(64, 48, 3.0, 60, 32, 24)
    cs = [cube(int(x), int(y), int(z), 5, r_c=float(rng.uniform(0.01, 0.1)))
          for x, y, z in rng.integers(8, 24, (200, 3))]
    c = cubes_of(cs)
    origin = (-1.0, -1.0, 1.5)
    h = O.histograms(c, [v], origin, 1.0, 10)
    for i, (hi, lo, r, d) in enumerate(cs):
        x, y, z = O.morton_decode((hi << 32) | lo)
        sh = 21 - d
        cs_ = 2.0 / (1 << d)
        P = [origin[0] + ((x >> sh) + 0.5) * cs_, origin[1] + ((y >> sh) + 0.5) * cs_,
             origin[2] + ((z >> sh) + 0.5) * cs_]
        uf, vf = P[0] / P[2] * 60 + 32, P[1] / P[2] * 60 + 24
        if not (-1 < uf < 64 and -1 < vf < 48) or P[2] <= 0:
            assert h[i].sum() == 0
            continue
        a = 3.0 - P[2]
        if a < -18 * r:
            assert h[i].sum() == 0
            continue
        t = min(max(a / (6 * r), -1.0), 1.0)
        b = min(max(int((t + 1) * 5), 0), 9)
        assert h[i].sum() == 1
        if abs((t + 1) * 5 - round((t + 1) * 5)) > 1e-4:
            assert h[i, b] == 1


def test_log2_level_rule_exhaustive():
    # The device computes trunc(log2f(x)) for x >= 1 as e + (x >= t[e]);
    # t[e] found by scanning down from 2^(e+1) with the C library's log2f,
    # then every float in [1, 8192) is checked against it.
    thr = []
    for e in range(13):
        top = np.float32(2.0 ** (e + 1))
        t = top
        x = np.nextafter(top, np.float32(0))
        while x >= 2.0 ** e and O.lib().orc_log2_level(x) >= e + 1:
            t = x
            x = np.nextafter(x, np.float32(0))
        thr.append(t)
    assert O.log2_rule_mismatches(1.0, 8192.0, np.array(thr, np.float32)) == 0


# ---------------------------------------------------------------- TGV invariants

def _hist(nb, **bins):
    h = np.zeros(nb, np.uint32)
    for k, v in bins.items():
        h[int(k[1:])] = v
    return h


def chain3():
    cs = cubes_of([cube(0, 0, 0, 5), cube(1, 0, 0, 5), cube(2, 0, 0, 5)])
    h = np.zeros((3, 8), np.uint32)
    h[0, 6] = 50
    h[2, 1] = 50
    return cs, h


@pytest.mark.parametrize("mode", [0, 1])
def test_three_cube_chain_polarity(mode):
    # tests/test_stage6.cpp:25-66
    cs, h = chain3()
    st, _, _, _ = O.tgv_minimize(cs, h, O.params(alpha0=0.1, alpha1=0.2, iters=50, op_mode=mode))
    u = st[:, 0]
    assert u[0] > 0 and u[2] < 0 and u[2] < u[1] < u[0]


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("mode", [0, 1])
def test_constant_field_is_exact_fixed_point(f64, mode):
    # tests/test_stage6.cpp:68-94 allows 0.625 +- 0.15; with every histogram
    # equal, u = m, v = p = q = 0 is an exact fixed point of every pass.
    cs = cubes_of([cube(ix, 0, 0, 4) for ix in range(4)] + [cube(ix, 1, 0, 4) for ix in range(4)])
    h = np.zeros((8, 8), np.uint32)
    h[:, 6] = 10
    st, _, _, _ = O.tgv_minimize(cs, h, O.params(iters=20, op_mode=mode), f64=f64)
    assert np.all(st[:, 0] == 0.625) and np.all(st[:, 1:] == 0)


def test_isolated_leaf_snaps_to_median_and_empty_stays_zero():
    cs = cubes_of([cube(3, 3, 3, 4)])
    h = np.zeros((1, 10), np.uint32)
    h[0, 7] = 4
    st, _, _, _ = O.tgv_minimize(cs, h, O.params(iters=1))
    assert st[0, 0] == np.float32(-1 + 7.5 / 5)
    cs2 = cubes_of([cube(i, 0, 0, 4) for i in range(5)])
    st2, _, _, _ = O.tgv_minimize(cs2, np.zeros((5, 10), np.uint32), O.params(iters=30))
    assert np.all(st2 == 0)


def test_empty_input():
    # tests/test_stage6.cpp:178-184
    st, _, it, _ = O.tgv_minimize(np.zeros(0, CUBE), np.zeros((0, 8), np.uint32))
    assert st.shape == (0, 13) and len(it) == 0
    leaves = np.zeros(0, dtype=[("first_idx", "<u4"), ("last_idx", "<u4")])
    assert O.tgv_minimize_oc(np.zeros(0, CUBE), np.zeros((0, 8), np.uint32), leaves).shape == (0, 13)


def test_oc_single_leaf_equals_in_memory_bitwise():
    # tests/test_stage6.cpp:133-153 allows 1e-5; the frozen snapshot of a
    # single leaf is never read, so the results are identical.
    cs, h = chain3()
    leaves = np.array([(0, 2)], dtype=[("first_idx", "<u4"), ("last_idx", "<u4")])
    prm = O.params(alpha0=0.1, alpha1=0.2, iters=50)
    a = O.tgv_minimize(cs, h, prm)[0]
    b = O.tgv_minimize_oc(cs, h, leaves, prm)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_oc_two_leaf_polarity():
    # tests/test_stage6.cpp:157-175
    cs, h = chain3()
    leaves = np.array([(0, 1), (2, 2)], dtype=[("first_idx", "<u4"), ("last_idx", "<u4")])
    u = O.tgv_minimize_oc(cs, h, leaves, O.params(alpha0=0.1, alpha1=0.2, iters=50))[:, 0]
    assert u[0] > 0 and u[2] < 0 and u[2] < u[1] < u[0]


def _leaves(*ranges):
    return np.array(list(ranges), dtype=[("first_idx", "<u4"), ("last_idx", "<u4")])


def test_streaming_single_leaf_equals_in_memory_bitwise():
    # tests/test_stage6.cpp:241-275 allows 1e-4; one leaf, one depth: no
    # snapshot and no border is ever read, so the results are identical
    cs, h = chain3()
    prm = O.params(alpha0=0.1, alpha1=0.2, iters=50)
    a = O.tgv_minimize(cs, h, prm)[0]
    b = O.tgv_minimize_streaming(cs, h, _leaves((0, 2)), prm)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_streaming_two_leaf_polarity():
    # tests/test_stage6.cpp:277-310
    cs, h = chain3()
    st = O.tgv_minimize_streaming(cs, h, _leaves((0, 1), (2, 2)), O.params(alpha0=0.1, alpha1=0.2, iters=50))
    u = st[:, 0]
    assert u[0] > 0 and u[2] < 0 and u[2] < u[1] < u[0]
    assert np.all(np.isfinite(st))


def test_streaming_parent_copy_follows_the_bisection():
    # A chain of octant-0 descendants shares code 0 at depths 1-4, then a
    # cube B at depth 3 with a larger code. The cube file's bisection for
    # code 0 over five cubes probes index 2 first and stops there (the depth-3
    # cube), so with no iterations every parent copy reads index 2's median:
    # u = [m0, m2, m2, m2, m2]. The in-memory solve takes the last cube with
    # the code (index 3) instead.
    cs = cubes_of([cube(0, 0, 0, 1), cube(0, 0, 0, 2), cube(0, 0, 0, 3), cube(0, 0, 0, 4), cube(1, 0, 0, 3)])
    codes = [(int(c["hi"]) << 32) | int(c["lo"]) for c in cs]
    assert codes[:4] == [0, 0, 0, 0] and codes[4] > 0
    h = np.zeros((5, 8), np.uint32)
    for i, k in enumerate((0, 2, 4, 6, 7)):
        h[i, k] = 5
    m = [O.median_indicator(x) for x in h]
    st = O.tgv_minimize_streaming(cs, h, _leaves((0, 4)), O.params(iters=0))
    assert st[:, 0].tolist() == [m[0], m[2], m[2], m[2], m[2]]
    assert np.all(st[:, 1:] == 0)
    mem = O.tgv_minimize(cs, h, O.params(iters=0))[0]
    assert mem[1, 0] == m[3]


def test_threads_do_not_change_bits(tmp_path):
    # Every pass is a Jacobi sweep: splitting it over threads changes nothing.
    rng = np.random.default_rng(3)
    pts = {tuple(p) for p in rng.integers(0, 16, (900, 3))}
    cs = cubes_of([cube(*p, 4) for p in sorted(pts)])
    h = rng.integers(0, 3, (len(cs), 10)).astype(np.uint32)
    for mode in (0, 1):
        a = O.tgv_minimize(cs, h, O.params(alpha0=2, alpha1=1, iters=40, op_mode=mode, threads=1))[0]
        b = O.tgv_minimize(cs, h, O.params(alpha0=2, alpha1=1, iters=40, op_mode=mode, threads=4))[0]
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------- dense operator

def dense_level(cs, level):
    """Independent statement of one level's linear map K(u, v) = (grad u - v, eps(v))
    as a matrix, from the neighbour table alone (all cubes on one level)."""
    act, nb = O.neighbour_table(cs, level, mode=1)
    n = len(act)
    pos = {int(a): i for i, a in enumerate(act)}
    # variables x = [u (n), vx (n), vy (n), vz (n)]; rows: p (3n) then q (6n)
    K = np.zeros((9 * n, 4 * n))
    U = lambda i: i
    V = lambda c, i: (1 + c) * n + i
    for i in range(n):
        for a in range(3):
            j = nb[i, 2 * a]
            r = a * n + i
            K[r, V(a, i)] -= 1.0
            if j >= 0:
                K[r, U(pos[j])] += 1.0
                K[r, U(i)] -= 1.0
        # eps rows
        def D(c, a):  # row vector of D_a v^c at i
            row = np.zeros(4 * n)
            j = nb[i, 2 * a]
            if j >= 0:
                row[V(c, pos[j])] += 1.0
                row[V(c, i)] -= 1.0
            return row
        K[3 * n + 0 * n + i] = D(0, 0)
        K[3 * n + 1 * n + i] = D(1, 1)
        K[3 * n + 2 * n + i] = D(2, 2)
        K[3 * n + 3 * n + i] = 0.5 * (D(0, 1) + D(1, 0))
        K[3 * n + 4 * n + i] = 0.5 * (D(0, 2) + D(2, 0))
        K[3 * n + 5 * n + i] = 0.5 * (D(1, 2) + D(2, 1))
    return act, K


def _random_level(seed, n_pts=60, grid=6, depth=3):
    rng = np.random.default_rng(seed)
    pts = sorted({tuple(p) for p in rng.integers(0, grid, (n_pts, 3))})
    cs = cubes_of([cube(*p, depth) for p in pts])
    h = rng.integers(0, 3, (len(cs), 10)).astype(np.uint32) * (rng.random((len(cs), 1)) < 0.8)
    return cs, h.astype(np.uint32)


def _unpack(st, n):
    u, v, p, q = st[:, 0], st[:, 1:4], st[:, 4:7], st[:, 7:13]
    x = np.concatenate([u, v[:, 0], v[:, 1], v[:, 2]])
    y = np.concatenate([p[:, 0], p[:, 1], p[:, 2]] + [q[:, k] for k in range(6)])
    return x, y


def _clip_rows(y, n, a1, a0):
    P = y[:3 * n].reshape(3, n)
    Q = y[3 * n:].reshape(6, n)
    nP = np.sqrt((P ** 2).sum(0))
    nQ = np.sqrt((Q ** 2).sum(0))
    P = P * np.where(nP > a1, a1 / np.maximum(nP, 1e-300), 1.0)
    Q = Q * np.where(nQ > a0, a0 / np.maximum(nQ, 1e-300), 1.0)
    return np.concatenate([P.reshape(-1), Q.reshape(-1)])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_dense_dual_step_and_exact_adjoint(seed):
    # one iteration from a random (u, v): p, q = clip(sigma K x), then the
    # exact-adjoint primal step x' = prox(x - tau K^T y) with K^T from numpy
    cs, h = _random_level(seed)
    n = len(cs)
    act, K = dense_level(cs, 3)
    assert np.array_equal(act, np.arange(n))
    rng = np.random.default_rng(100 + seed)
    init = rng.uniform(-0.5, 0.5, (n, 4)).astype(np.float32)
    a0, a1 = 2.0, 1.0
    st = O.tgv_minimize(cs, h, O.params(alpha0=a0, alpha1=a1, iters=1, op_mode=1, warm_fallback=0), f64=True,
                        init_uv=init)[0]
    x0 = np.concatenate([init[:, 0], init[:, 1], init[:, 2], init[:, 3]]).astype(np.float64)
    y1 = _clip_rows(0.25 * (K @ x0), n, a1, a0)
    x1, y1_or = _unpack(st, n)
    np.testing.assert_allclose(y1_or, y1, rtol=1e-12, atol=1e-13)
    g = x0 - 0.25 * (K.T @ y1)
    tot = h.sum(1).astype(np.float64)
    med = np.array([O.median_indicator(r, f64=True) for r in h])
    z = g[:n]
    u_prox = np.where(tot == 0, z, np.where(z - med > 0.25 * tot, z - 0.25 * tot,
                                            np.where(z - med < -0.25 * tot, z + 0.25 * tot, med)))
    np.testing.assert_allclose(x1[:n], u_prox, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(x1[n:], g[n:], rtol=1e-12, atol=1e-13)


def dense_gap(K, x, y, W, m, lam, a0, a1, vbox, n):
    kx = K @ x
    P3 = kx[:3 * n].reshape(3, n)
    Q6 = kx[3 * n:].reshape(6, n)
    u = x[:n]
    P = (lam * W * np.abs(u - m)).sum() + a1 * np.sqrt((P3 ** 2).sum(0)).sum() + a0 * np.sqrt((Q6 ** 2).sum(0)).sum()
    c = K.T @ y
    cu, cv = c[:n], c[n:]
    cand = np.stack([lam * W * np.abs(t - m) + cu * t for t in (-1.0, 1.0)] + [lam * W * 0 + cu * m])
    mu = np.where(W == 0, -np.abs(cu), cand.min(0))
    D = mu.sum() - vbox * np.abs(cv).sum()
    return P, D


@pytest.mark.parametrize("seed", [0, 1])
def test_dense_energy_and_gap(seed):
    cs, h = _random_level(seed, n_pts=50)
    n = len(cs)
    _, K = dense_level(cs, 3)
    prm = O.params(alpha0=2.0, alpha1=1.0, iters=7, op_mode=1)
    st, bar, _, pd = O.tgv_minimize(cs, h, prm, f64=True, with_bar=True)
    x, y = _unpack(st, n)
    W = h.sum(1).astype(np.float64)
    m = np.array([O.median_indicator(r, f64=True) for r in h])
    P, D = dense_gap(K, x, y, W, m, 1.0, 2.0, 1.0, 2.0, n)
    Po, Do = O.level_gap(cs, h, prm, 3, st, bar)
    assert Po == pytest.approx(P, rel=1e-12) and Do == pytest.approx(D, rel=1e-12, abs=1e-9)
    assert pd[0][0] == pytest.approx(P, rel=1e-12)


def test_gap_nonnegative_and_closes_in_adjoint_mode():
    cs, h = _random_level(5, n_pts=80)
    gaps = []
    for iters in (10, 100, 1000, 10000):
        _, _, _, pd = O.tgv_minimize(cs, h, O.params(alpha0=2.0, alpha1=1.0, iters=iters, op_mode=1), f64=True)
        P, D = pd[0]
        assert P - D >= -1e-9 * max(1.0, abs(P))
        gaps.append((P - D) / max(1.0, abs(P)))
    assert all(b < a for a, b in zip(gaps, gaps[1:])) and gaps[-1] < 1e-4


def test_gap_stopping_rule():
    cs, h = _random_level(6, n_pts=80)
    _, _, it, pd = O.tgv_minimize(cs, h, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-3, chunk=25,
                                                  max_iters=5000), f64=True)
    P, D = pd[0]
    assert (P - D) / max(1.0, abs(P)) <= 1e-3 and it[0] % 25 == 0 and it[0] < 5000


# ---------------------------------------------------------------- surface

def _edge_valence(faces):
    from collections import Counter
    e = Counter()
    for a, b, c in faces:
        for x, y in ((a, b), (b, c), (c, a)):
            e[(min(x, y), max(x, y))] += 1
    vals = np.array(list(e.values()))
    return (vals == 1).sum(), (vals == 2).sum(), (vals > 2).sum()


def test_closed_box_surface():
    # tests/test_stage7.cpp:245-273: 12 triangles, no boundary edges
    cs, u = [], []
    for iz, iy, ix in itertools.product(range(3), repeat=3):
        cs.append(cube(ix, iy, iz, 3, 24.0 / 8))
        u.append(-1.0 if (ix, iy, iz) == (1, 1, 1) else 1.0)
    v, f = O.extract_surface(cubes_of(cs), np.array(u, np.float32), (0, 0, 0), 24.0)
    assert len(f) == 12
    b, i, irr = _edge_valence(f)
    assert b == 0 and irr == 0


def test_plane_patch_surface():
    # tests/test_stage7.cpp:62-111
    cs, u = [], []
    for iz, iy, ix in itertools.product(range(2), repeat=3):
        cs.append(cube(ix, iy, iz, 4, 1.0))
        u.append(1.0 if ix == 0 else -1.0)
    v, f = O.extract_surface(cubes_of(cs), np.array(u, np.float32), (0, 0, 0), 16.0)
    assert len(f) >= 4
    assert np.all(np.abs(v[:, 0] - 2.0) < 0.01 * 2.0 + 1e-4)
    assert len({tuple(r) for r in np.round(v * 1e6).astype(np.int64)}) == len(v)
    b, i, irr = _edge_valence(f)
    assert irr == 0 and i > 0


# ---------------------------------------------------------------- depth-aware table

def test_depth_aware_table_brute_force():
    # exact-adjoint mode: same-depth face neighbour, else the cube one level
    # coarser covering that position, else none; one-to-one within a level
    rng = np.random.default_rng(11)
    cells = set()
    for _ in range(400):
        d = int(rng.integers(2, 5))
        x, y, z = rng.integers(0, 1 << d, 3)
        cells.add((d, int(x), int(y), int(z)))
    lst = sorted(cells, key=lambda c: (O.morton_encode(c[1] << (21 - c[0]), c[2] << (21 - c[0]),
                                                       c[3] << (21 - c[0])), c[0]))
    cs = cubes_of([cube(x, y, z, d) for d, x, y, z in lst])
    index = {c: i for i, c in enumerate(lst)}
    for level in (2, 3, 4):
        act, nb = O.neighbour_table(cs, level, mode=1)
        seen = set()
        for row, i in zip(nb, act):
            d, x, y, z = lst[i]
            assert d == level
            for f in range(6):
                a, sgn = f // 2, (1 if f % 2 == 0 else -1)
                p = [x, y, z]
                p[a] += sgn
                want = -1
                if 0 <= p[a] < (1 << d):
                    want = index.get((d, *p), -1)
                    if want < 0:
                        want = index.get((d - 1, p[0] >> 1, p[1] >> 1, p[2] >> 1), -1)
                assert row[f] == want, (lst[i], f, row[f], want)
                if f % 2 == 0 and want >= 0 and lst[want][0] == level:
                    assert (want, f) not in seen
                    seen.add((want, f))


# ---------------------------------------------------------------- network

def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(4)
    w = rng.normal(0, 0.3, O.MLP_NP)
    x = rng.normal(0, 1, (7, 30))
    t = rng.normal(0, 1, (7, 4))
    loss, g = O.mlp_loss_grad(w, x, t)
    for i in list(range(0, O.MLP_NP, 97)) + [O.MLP_NP - 1, 1920, 1983]:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (O.mlp_loss_grad(wp, x, t)[0] - O.mlp_loss_grad(wm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_forward_matches_numpy():
    rng = np.random.default_rng(5)
    w = rng.normal(0, 0.3, O.MLP_NP)
    x = rng.normal(0, 1, (9, 30))
    W1, b1 = w[:1920].reshape(30, 64), w[1920:1984]
    W2, b2 = w[1984:6080].reshape(64, 64), w[6080:6144]
    W3, b3 = w[6144:6400].reshape(64, 4), w[6400:]
    h = np.maximum(x @ W1 + b1, 0)
    h = np.maximum(h @ W2 + b2, 0)
    np.testing.assert_allclose(O.mlp_forward(w, x), h @ W3 + b3, rtol=1e-12, atol=1e-12)


def test_adam_step_closed_form():
    rng = np.random.default_rng(6)
    w = rng.normal(size=O.MLP_NP)
    g = rng.normal(size=O.MLP_NP)
    m, v = np.zeros(O.MLP_NP), np.zeros(O.MLP_NP)
    w0 = w.copy()
    O.mlp_adam(w, m, v, g, 1, 1e-3)
    # first step with zero moments: w -= lr * sign(g) * |g| / (|g| + eps)
    np.testing.assert_allclose(w, w0 - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)


# ---------------------------------------------------------------- per-depth steps

def _two_level_octree(seed):
    """A random level at depth 3 and its children-level neighbours at depth 4."""
    rng = np.random.default_rng(seed)
    coarse = sorted({tuple(p) for p in rng.integers(0, 6, (40, 3))})
    fine = sorted({tuple(p) for p in rng.integers(0, 12, (120, 3))} - {(2 * x + i, 2 * y + j, 2 * z + k)
                                                                         for x, y, z in coarse for i in (0, 1)
                                                                         for j in (0, 1) for k in (0, 1)})
    cs = cubes_of([cube(*p, 3) for p in coarse] + [cube(*p, 4) for p in fine])
    order = np.lexsort((cs["depth"], cs["lo"], cs["hi"]))
    cs = cs[order]
    h = (rng.integers(0, 3, (len(cs), 10)) * (rng.random((len(cs), 1)) < 0.8)).astype(np.uint32)
    return cs, h


def test_depth_steps_equal_to_scalar_steps_bitwise():
    cs, h = _two_level_octree(3)
    base = dict(alpha0=2.0, alpha1=1.0, iters=120, op_mode=1)
    a, _, _, _ = O.tgv_minimize(cs, h, O.params(sigma=2.0, tau=1.0 / 32, **base), f64=True)
    b, _, _, _ = O.tgv_minimize(cs, h, O.params(sigma_depth=[2.0] * 22, tau_depth=[1.0 / 32] * 22, **base), f64=True)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))


def test_depth_steps_act_on_their_level_only():
    # steps given for depth 4 change the depth-4 cubes and leave the depth-3
    # level (iterated before, never written afterwards) bit for bit as it was
    cs, h = _two_level_octree(4)
    base = dict(alpha0=2.0, alpha1=1.0, iters=120, op_mode=1)
    sd, td = [0.0] * 22, [0.0] * 22
    sd[4], td[4] = 2.0, 1.0 / 32
    a, _, _, _ = O.tgv_minimize(cs, h, O.params(**base), f64=True)
    b, _, _, _ = O.tgv_minimize(cs, h, O.params(sigma_depth=sd, tau_depth=td, **base), f64=True)
    c3 = cs["depth"] == 3
    assert np.array_equal(a[c3].view(np.uint64), b[c3].view(np.uint64))
    assert not np.array_equal(a[~c3].view(np.uint64), b[~c3].view(np.uint64))


def test_depth_steps_gap_closes():
    # sigma * tau = 1/16 at every depth (ratios 64 and 4096): exact PDHG, the gap closes
    cs, h = _two_level_octree(5)
    sd, td = [0.0] * 22, [0.0] * 22
    sd[3], td[3] = 2.0, 1.0 / 32
    sd[4], td[4] = 16.0, 1.0 / 256
    _, _, it, pd = O.tgv_minimize(cs, h, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-6, chunk=50,
                                                  max_iters=40000, sigma_depth=sd, tau_depth=td), f64=True)
    for P, D in pd:
        assert -1e-9 <= (P - D) / max(1.0, abs(P)) <= 1e-6
    assert all(0 < k < 40000 for k in it)
