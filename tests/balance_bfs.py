"""A second, independent statement of balance_octree (reference
src/balance.cpp:60-150) in plain Python, for small inputs: Morton codes as
Python ints, one set per depth, a deque for the FIFO queue. Used to pin the
host generator's balance and, through it, the device balance.

Cells are keyed by the full 63-bit code (the reference's dc_key,
balance.cpp:56-58, keeps 60 bits; see the generator's note).
"""
from collections import deque

import numpy as np

from tgvreconclaude_amd import CUBE_DT


def interleave(x, y, z):
    c = 0
    for b in range(21):
        c |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return c


def deinterleave(c):
    x = y = z = 0
    for b in range(21):
        x |= ((c >> (3 * b)) & 1) << b
        y |= ((c >> (3 * b + 1)) & 1) << b
        z |= ((c >> (3 * b + 2)) & 1) << b
    return x, y, z


def ancestor(c, d):
    s = 3 * (21 - d)
    return 0 if s >= 64 else c & ~((1 << s) - 1)


def face_neighbor(c, f, d):
    """faces 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z; None outside the grid (balance.cpp:32-49)"""
    p = list(deinterleave(c))
    step = 1 << (21 - d)
    a = f // 2
    if f % 2 == 0:
        if p[a] + step >= 1 << 21:
            return None
        p[a] += step
    else:
        if p[a] < step:
            return None
        p[a] -= step
    return interleave(*p)


def balance(cells, r_root):
    """cells: list of (code, depth, r_c) in queue order -> list sorted by
    (code, depth), input order among equal pairs, bridges appended before the sort."""
    by_depth = [set() for _ in range(22)]
    for c, d, _ in cells:
        by_depth[d].add(c)
    out = list(cells)
    q = deque((d, c) for c, d, _ in cells)
    while q:
        d, c = q.popleft()
        if d < 2:
            continue
        for f in range(6):
            nb = face_neighbor(c, f, d)
            if nb is None or nb in by_depth[d]:
                continue
            par = ancestor(nb, d - 1)
            if par in by_depth[d - 1]:
                continue
            if not any(ancestor(nb, d2) in by_depth[d2] for d2 in range(d - 2, -1, -1)):
                continue
            by_depth[d - 1].add(par)
            out.append((par, d - 1, np.float32(r_root) / np.float32(1 << (d - 1))))
            q.append((d - 1, par))
    return sorted(out, key=lambda t: (t[0], t[1]))  # stable


def violations(cubes):
    """The reference test's predicate (tests/test_stages2_4.cpp:253-281): a
    cube whose face neighbour has neither a same-depth cell nor a parent cell
    one level coarser, while a cell two or more levels coarser covers it."""
    cs = set((int(c), int(d)) for c, d in zip(codes(cubes), cubes["depth"]))
    bad = []
    for c, d in cs:
        if d < 2:
            continue
        for f in range(6):
            nb = face_neighbor(c, f, d)
            if nb is None:
                continue
            if (nb, d) in cs or (ancestor(nb, d - 1), d - 1) in cs:
                continue
            if any((ancestor(nb, d2), d2) in cs for d2 in range(d - 2, -1, -1)):
                bad.append((c, d, f))
    return bad


def codes(cubes):
    return (cubes["hi"].astype(np.uint64) << np.uint64(32)) | cubes["lo"].astype(np.uint64)


def to_cubes(cells):
    out = np.zeros(len(cells), dtype=CUBE_DT)
    for i, (c, d, r) in enumerate(cells):
        out[i]["hi"] = c >> 32
        out[i]["lo"] = c & 0xFFFFFFFF
        out[i]["depth"] = d
        out[i]["r_c"] = r
    return out


def random_cells(seed, n, lo=2, hi=9, box=3):
    """n distinct random cells of depths lo..hi inside a small corner region,
    so that cells of very different depths touch (many violations)."""
    rng = np.random.default_rng(seed)
    seen, cells = set(), []
    while len(cells) < n:
        d = int(rng.integers(lo, hi + 1))
        g = 1 << d
        lim = max(1, (g * box) // 16)
        x, y, z = (int(v) for v in rng.integers(0, min(g, lim), 3))
        sh = 21 - d
        c = interleave(x << sh, y << sh, z << sh)
        if (c, d) in seen:
            continue
        seen.add((c, d))
        cells.append((c, d, np.float32(1.5) / np.float32(g)))
    return cells
