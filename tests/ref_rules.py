"""A second, independent statement of the reference's TGV update rules in NumPy.

TEST INFRASTRUCTURE ONLY. It shares no code with oracle/tgv_oracle.cpp and is
written from the reference's source (paths relative to its root), structured
differently on purpose so that a slip in the oracle is not repeated here:

  * the code-only neighbour table (src/tgv.cpp:50-54 `key` / CodeMap, :97-109
    `neighbor_idx_ctf`, :115-133 `NeighbourTable`) by brute force: a full scan
    of the code array for the largest index holding a code (the
    unordered_map's last writer, :367-370), neighbour codes from integer
    coordinates (src/balance.cpp:22-49, src/types.hpp:61-73);
  * one level's iteration (src/tgv.cpp:137-232 `run_level`) with 0/1
    selection matrices: S[a] picks the +a neighbour (or the cube itself, the
    Neumann rule of :163), M[a] the -a neighbour (a zero row where there is
    none, :196, :209); the self term of the primal passes is always present
    and the v divergence has no 1/2 (:198-216);
  * tgv_minimize (:356-419), tgv_minimize_oc with run_level_frozen
    (:241-352, :423-511: out-of-leaf u_bar / v_bar from the level-start
    snapshot, out-of-leaf p / q live) and tgv_minimize_streaming (:802-1070:
    per treetop leaf a local index space of the leaf's cubes and the border
    cubes found by the cube file's bisection, border p / q zero, border bars
    from the level-start (u, v)).

Every expression is evaluated in the reference's order, one NumPy operation at
a time on arrays of the chosen dtype (no fused multiply-add, IEEE rounding per
operation), so float32 results can be compared with the oracle bit for bit.
"""
import numpy as np

NONE = None
FINE = 21


# ---------------------------------------------------------------- geometry
def encode(x, y, z):
    """Morton code: bit b of x, y, z at bits 3b, 3b+1, 3b+2 (src/types.hpp:10-11, 61-66)."""
    r = 0
    for b in range(FINE):
        r |= ((x >> b) & 1) << (3 * b)
        r |= ((y >> b) & 1) << (3 * b + 1)
        r |= ((z >> b) & 1) << (3 * b + 2)
    return r


def decode(c):
    x = y = z = 0
    for b in range(FINE):
        x |= ((c >> (3 * b)) & 1) << b
        y |= ((c >> (3 * b + 1)) & 1) << b
        z |= ((c >> (3 * b + 2)) & 1) << b
    return x, y, z


def face_code(code, face, depth):
    """The same-size neighbour's code across face 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z,
    or None outside the root cube (src/balance.cpp:32-49)."""
    p = list(decode(code))
    step = 1 << (FINE - depth)
    a = face // 2
    p[a] += step if face % 2 == 0 else -step
    if p[a] < 0 or p[a] >= (1 << FINE):
        return NONE
    return encode(*p)


def coarser_code(code, depth):
    """Lower corner of the cell one level coarser (depth - 1) that contains the
    cell (src/balance.cpp:22-30), from coordinates; depth 0 gives 0."""
    if depth <= 0:
        return 0
    size = 1 << (FINE - (depth - 1))
    x, y, z = decode(code)
    return encode(x // size * size, y // size * size, z // size * size)


def codes_of(cubes):
    return [(int(c["hi"]) << 32) | int(c["lo"]) for c in cubes]


def last_with_code(codes, target):
    """Largest index whose code equals target, or -1 (the CodeMap filled in index
    order keeps the last writer, src/tgv.cpp:367-370)."""
    hit = -1
    for j, c in enumerate(codes):
        if c == target:
            hit = j
    return hit


def ref_neighbours(codes, depths, i):
    """neighbor_idx_ctf for the six faces of cube i (src/tgv.cpp:97-109)."""
    out = []
    for f in range(6):
        t = face_code(codes[i], f, depths[i])
        if t is NONE:
            out.append(-1)
            continue
        j = last_with_code(codes, t)
        if j < 0 and depths[i] > 0:
            j = last_with_code(codes, coarser_code(t, depths[i]))
        out.append(j)
    return out


# ---------------------------------------------------------------- histograms
def median8(bins, dt):
    """CubeHistogram::median_indicator generalised to nb bins
    (src/histogram.cpp:21-31: centre -1 + (k + 0.5) / (nb / 2))."""
    nb = len(bins)
    tot = int(np.sum(bins, dtype=np.int64))
    if tot == 0:
        return dt(0)
    cum = 0
    for k in range(nb):
        cum += int(bins[k])
        if 2 * cum >= tot:
            return dt(dt(-1) + dt(dt(k) + dt(0.5)) / dt(dt(0.5) * dt(nb)))
    return dt(1)


def prox(z, tl, bins, dt):
    """prox_data (src/tgv.cpp:75-94): the median, or z shrunk by tau * total."""
    tot = int(np.sum(bins, dtype=np.int64))
    if tot == 0:
        return z
    med = median8(bins, dt)
    diff = dt(z - med)
    shrink = dt(tl * dt(tot))
    if diff > shrink:
        return dt(z - shrink)
    if diff < -shrink:
        return dt(z + shrink)
    return med


# ---------------------------------------------------------------- one level
def _clip_rows(v, r):
    """clip_ball3 / clip_ball6 row-wise: squared norm summed left to right."""
    n2 = v[:, 0] * v[:, 0]
    for k in range(1, v.shape[1]):
        n2 = n2 + v[:, k] * v[:, k]
    out = v.copy()
    big = n2 > r * r
    if np.any(big):
        s = r / np.sqrt(n2[big])
        out[big] = v[big] * s[:, None]
    return out


def selection(n, rows, cols, default_self):
    """n x n 0/1 matrix with row rows[k] picking column cols[k]; a negative
    column picks the row itself (default_self) or nothing."""
    S = np.zeros((n, n))
    for r, c in zip(rows, cols):
        if c >= 0:
            S[r, c] = 1.0
        elif default_self:
            S[r, r] = 1.0
    return S


def iterate_level(active, nb, state, bar, bins, sigma, tau, tl, a0, a1, iters, dt,
                  read_bar=None, read_pq=None):
    """run_level (src/tgv.cpp:137-232) on the whole arrays.

    active: indices iterated; nb: {i: six neighbour indices};
    state: N x 13 (u, v[3], p[3], q[6]); bar: N x 4 (u_bar, v_bar[3]).
    read_bar(bar) -> the N x 4 values the dual passes read at neighbours
    (identity for run_level; the frozen snapshot outside the leaf for
    run_level_frozen); read_pq(state) -> the N x 9 p, q the primal passes read
    at -a neighbours (identity for run_level).
    """
    n = state.shape[0]
    act = np.asarray(active, dtype=np.int64)
    S = [selection(n, act, [nb[i][2 * a] for i in act], False) for a in range(3)]
    M = [selection(n, act, [nb[i][2 * a + 1] for i in act], False) for a in range(3)]
    # rows whose +a neighbour is missing read the cube's own live bars (Neumann, :163)
    own = [np.isin(np.arange(n), [i for i in act if nb[i][2 * a] < 0]) for a in range(3)]
    sg, ta = dt(sigma), dt(tau)
    for _ in range(iters):
        rb = bar if read_bar is None else read_bar(bar)
        # neighbour values picked by the selection matrices (exact: one 1 per row)
        nbv = []
        for a in range(3):
            g = (S[a] @ rb.astype(np.float64)).astype(dt)
            g[own[a]] = bar[own[a]]
            nbv.append(g)
        ub, vx, vy, vz = bar[:, 0], bar[:, 1], bar[:, 2], bar[:, 3]
        # ---- dual p (:159-169)
        p = state[:, 4:7].copy()
        p[:, 0] = p[:, 0] + sg * ((nbv[0][:, 0] - ub) - vx)
        p[:, 1] = p[:, 1] + sg * ((nbv[1][:, 0] - ub) - vy)
        p[:, 2] = p[:, 2] + sg * ((nbv[2][:, 0] - ub) - vz)
        # ---- dual q (:172-189)
        q = state[:, 7:13].copy()
        e = np.stack([nbv[0][:, 1] - vx,
                      nbv[1][:, 2] - vy,
                      nbv[2][:, 3] - vz,
                      dt(0.5) * (((nbv[1][:, 1] - vx) + nbv[0][:, 2]) - vy),
                      dt(0.5) * (((nbv[2][:, 1] - vx) + nbv[0][:, 3]) - vz),
                      dt(0.5) * (((nbv[2][:, 2] - vy) + nbv[1][:, 3]) - vz)], axis=1)
        q = q + sg * e
        new = state.copy()
        # TGVParams holds floats (src/tgv.hpp:24-28): the radii are float values
        new[act, 4:7] = _clip_rows(p[act], dt(np.float32(a1)))
        new[act, 7:13] = _clip_rows(q[act], dt(np.float32(a0)))
        # ---- primal u, v (:192-217): -a neighbours' p, q (0 where none)
        pq = new[:, 4:13] if read_pq is None else read_pq(new)
        mp = [(M[a] @ pq.astype(np.float64)).astype(dt) for a in range(3)]
        P, Q = new[:, 4:7], new[:, 7:13]
        div_p = ((P[:, 0] - mp[0][:, 0]) + (P[:, 1] - mp[1][:, 1])) + (P[:, 2] - mp[2][:, 2])
        z = state[:, 0] + ta * div_p
        un = np.array([prox(z[i], dt(tl), bins[i], dt) for i in range(n)], dtype=dt)
        # q columns: 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz -> pq columns 3 + k
        dqx = ((Q[:, 0] - mp[0][:, 3]) + (Q[:, 3] - mp[1][:, 6])) + (Q[:, 4] - mp[2][:, 7])
        dqy = ((Q[:, 3] - mp[0][:, 6]) + (Q[:, 1] - mp[1][:, 4])) + (Q[:, 5] - mp[2][:, 8])
        dqz = ((Q[:, 4] - mp[0][:, 7]) + (Q[:, 5] - mp[1][:, 8])) + (Q[:, 2] - mp[2][:, 5])
        vn = np.stack([state[:, 1] + ta * (P[:, 0] + dqx),
                       state[:, 2] + ta * (P[:, 1] + dqy),
                       state[:, 3] + ta * (P[:, 2] + dqz)], axis=1)
        # ---- extrapolation and commit (:220-230), active cubes only
        nbar = bar.copy()
        nbar[act, 0] = dt(2) * un[act] - state[act, 0]
        nbar[act, 1:4] = dt(2) * vn[act] - state[act, 1:4]
        new[act, 0] = un[act]
        new[act, 1:4] = vn[act]
        state[:] = new
        bar[:] = nbar


# ---------------------------------------------------------------- drivers
def _setup(cubes, hists, dt):
    codes = codes_of(cubes)
    depths = [int(c["depth"]) for c in cubes]
    N = len(codes)
    state = np.zeros((N, 13), dtype=dt)
    for i in range(N):
        state[i, 0] = median8(hists[i], dt)
    bar = np.zeros((N, 4), dtype=dt)
    bar[:, 0] = state[:, 0]
    return codes, depths, state, bar


def _parent_copy(codes, depths, state, bar, active, L):
    """tgv.cpp:391-409, cube by cube in index order."""
    for i in active:
        pi = last_with_code(codes, coarser_code(codes[i], L))
        if pi >= 0:
            state[i, 0:4] = state[pi, 0:4]
        state[i, 4:13] = 0
        bar[i] = state[i, 0:4]


def _steps(sigma, tau, lam, dt):
    s = dt(np.float32(sigma)) if sigma > 0 else dt(0.25)
    t = dt(np.float32(tau)) if tau > 0 else dt(0.25)
    return s, t, dt(t * dt(np.float32(lam)))


def tgv_minimize(cubes, hists, alpha0, alpha1, iters, dt=np.float32, sigma=0.0, tau=0.0, lam=1.0):
    """src/tgv.cpp:356-419; returns the N x 13 state."""
    codes, depths, state, bar = _setup(cubes, hists, dt)
    if not codes:
        return state
    sg, ta, tl = _steps(sigma, tau, lam, dt)
    lo, hi = min(depths), max(depths)
    for L in range(lo, hi + 1):
        active = [i for i in range(len(codes)) if depths[i] == L]
        if not active:
            continue
        if L > lo:
            _parent_copy(codes, depths, state, bar, active, L)
        nb = {i: ref_neighbours(codes, depths, i) for i in active}
        iterate_level(active, nb, state, bar, hists, sg, ta, tl, alpha0, alpha1, iters, dt)
    return state


def tgv_minimize_oc(cubes, hists, leaves, alpha0, alpha1, iters, dt=np.float32):
    """src/tgv.cpp:423-511 with run_level_frozen (:241-352): leaves in order;
    a neighbour index inside [first, last] of the leaf being iterated reads the
    live bars, any other the level-start snapshot; p, q are read live."""
    codes, depths, state, bar = _setup(cubes, hists, dt)
    if not codes:
        return state
    sg, ta, tl = _steps(0, 0, 1.0, dt)
    lo, hi = min(depths), max(depths)
    for L in range(lo, hi + 1):
        level = [i for i in range(len(codes)) if depths[i] == L]
        if not level:
            continue
        if L > lo:
            _parent_copy(codes, depths, state, bar, level, L)
        snap = bar.copy()
        for first, last in leaves:
            act = [i for i in level if first <= i <= last]
            if not act:
                continue
            nb = {i: ref_neighbours(codes, depths, i) for i in act}
            inleaf = np.array([first <= j <= last for j in range(len(codes))])

            def read_bar(b, inleaf=inleaf):
                return np.where(inleaf[:, None], b, snap)

            iterate_level(act, nb, state, bar, hists, sg, ta, tl, alpha0, alpha1, iters, dt, read_bar=read_bar)
    return state


def bisect_code(codes, target):
    """CubeFileReader::lookup (src/tgv.cpp:673-683): the first probe of a binary
    search over the code-sorted file whose code matches, or -1."""
    lo, hi = 0, len(codes)
    while lo < hi:
        mid = lo + (hi - lo) // 2
        if codes[mid] == target:
            return mid
        if codes[mid] < target:
            lo = mid + 1
        else:
            hi = mid
    return -1


def tgv_minimize_streaming(cubes, hists, leaves, alpha0, alpha1, iters, dt=np.float32):
    """src/tgv.cpp:802-1070 with the files replaced by arrays (cubes sorted by code)."""
    codes, depths, state, _ = _setup(cubes, hists, dt)
    N = len(codes)
    if N == 0:
        return state
    sg, ta, tl = _steps(0, 0, 1.0, dt)
    lo, hi = min(depths), max(depths)
    leaves = [(f, min(l, N - 1)) for f, l in leaves]
    for L in range(lo, hi + 1):
        if L > lo:  # :872-934, leaf after leaf
            for first, last in leaves:
                lvl = [i for i in range(first, last + 1) if depths[i] == L]
                if not lvl:
                    continue
                par = [bisect_code(codes, coarser_code(codes[i], L)) for i in lvl]
                pst = {g: state[g].copy() for g in set(par) if g >= 0}
                for i, g in zip(lvl, par):
                    if g >= 0:
                        state[i, 0:4] = pst[g][0:4]
                    state[i, 4:13] = 0
        frozen = state[:, 0:4].copy()  # build_bars_frozen (:606-635)
        for first, last in leaves:
            n_leaf = last - first + 1
            if n_leaf <= 0:
                continue
            loc = list(range(first, last + 1))
            act = [k for k in range(n_leaf) if depths[loc[k]] == L]
            if not act:
                continue
            leaf_map = {}
            for k in range(n_leaf):
                leaf_map[codes[loc[k]]] = k
            border = []
            for k in act:  # :960-975
                c, d = codes[loc[k]], depths[loc[k]]
                for f in range(6):
                    t = face_code(c, f, d)
                    if t is NONE or t in leaf_map:
                        continue
                    border.append(t)
                    if d > 0 and coarser_code(t, d) not in leaf_map:
                        border.append(coarser_code(t, d))
            bglob = []
            for t in sorted(set(border)):  # :981-991
                g = bisect_code(codes, t)
                if g < 0 or first <= g <= last:
                    continue
                bglob.append(g)
            local_map = dict(leaf_map)
            for b, g in enumerate(bglob):
                local_map[codes[g]] = n_leaf + b
            idx = loc + bglob
            nb = {}
            for k in act:  # :1000-1014
                c, d = codes[loc[k]], depths[loc[k]]
                row = [-1] * 6
                for f in range(6):
                    t = face_code(c, f, d)
                    if t is NONE:
                        continue
                    if t in local_map:
                        row[f] = local_map[t]
                    elif d > 0 and coarser_code(t, d) in local_map:
                        row[f] = local_map[coarser_code(t, d)]
                nb[k] = row
            st = state[idx].copy()
            st[n_leaf:, 4:13] = 0  # border p, q suppressed (:1030-1034)
            bar = np.concatenate([st[:n_leaf, 0:4], frozen[bglob].reshape(-1, 4)]).astype(dt)
            hl = [hists[j] for j in idx]
            iterate_level(act, nb, st, bar, hl, sg, ta, tl, alpha0, alpha1, iters, dt)
            state[first:last + 1] = st[:n_leaf]
    return state
