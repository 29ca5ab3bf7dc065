"""Pins of the oracle's reference-rule paths against tests/ref_rules.py, a second
NumPy statement written from the reference's source alone (CPU only).

Covered (paths relative to the reference root):
  * the code-only neighbour table with last-writer-wins lookups, including
    octant-0 children sharing their parents' codes in either order
    (src/tgv.cpp:50-54, :97-133, :367-370);
  * tgv_minimize with the reference's update rules (src/tgv.cpp:137-232,
    :356-419), FP32 and FP64, bit for bit;
  * tgv_minimize_oc / run_level_frozen over several treetop leaves
    (src/tgv.cpp:241-352, :423-511);
  * tgv_minimize_streaming's per-leaf tables, bisection lookups, border bars
    and zeroed border p / q (src/tgv.cpp:802-1070).
The last test builds copies of the oracle with one plausible one-line slip
each and requires that these pins reject every one of them.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import ref_rules as R

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CUBE = O.CUBE_DT
LEAF = [("first_idx", "<u4"), ("last_idx", "<u4")]


def scene(seed, n_coarse=26, grid=5):
    """Cubes of depths 2-4 over a small region: depth-3 cells, some split into
    depth-4 children (octant 0 among them, sharing the parent's code), some
    depth-2 cells, sorted by code with ties in random order (the reference's
    unstable sort, src/balance.cpp:138-139)."""
    rng = np.random.default_rng(seed)
    cells = set()
    for x, y, z in rng.integers(0, grid, (n_coarse, 3)):
        cells.add((int(x), int(y), int(z), 3))
    for (x, y, z, d) in list(cells):
        if rng.random() < 0.45:
            kids = [k for k in range(8) if k == 0 or rng.random() < 0.5]
            for k in kids:
                cells.add((2 * x + (k & 1), 2 * y + ((k >> 1) & 1), 2 * z + (k >> 2), 4))
            if rng.random() < 0.5:
                cells.discard((x, y, z, 3))
    for x, y, z in rng.integers(grid // 2, grid // 2 + 3, (4, 3)):
        cells.add((int(x), int(y), int(z), 2))
    recs = []
    for (x, y, z, d) in cells:
        s = 21 - d
        c = R.encode(x << s, y << s, z << s)
        recs.append((c, rng.random(), d))
    recs.sort()
    cs = np.array([(c >> 32, c & 0xFFFFFFFF, 0.01, d) for c, _, d in recs], dtype=CUBE)
    h = (rng.integers(0, 6, (len(cs), 8)) * (rng.random((len(cs), 8)) < 0.4)).astype(np.uint32)
    h[rng.random(len(cs)) < 0.15] = 0
    return cs, h


def _collisions(cs):
    codes = R.codes_of(cs)
    d = [int(c["depth"]) for c in cs]
    return sum(1 for i in range(1, len(codes)) if codes[i] == codes[i - 1] and d[i] != d[i - 1])


def bits(a):
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


SEEDS = [0, 1, 2, 3]


@pytest.mark.parametrize("seed", SEEDS)
def test_scene_has_collisions_and_fallbacks(seed):
    cs, h = scene(seed)
    assert _collisions(cs) >= 2
    codes, depths = R.codes_of(cs), [int(c["depth"]) for c in cs]
    fallback = 0
    for i in range(len(cs)):
        for f in range(6):
            t = R.face_code(codes[i], f, depths[i])
            if t is not None and R.last_with_code(codes, t) < 0 and \
                    R.last_with_code(codes, R.coarser_code(t, depths[i])) >= 0:
                fallback += 1
    assert fallback >= 3


@pytest.mark.parametrize("seed", SEEDS)
def test_reference_table_brute_force(seed):
    cs, _ = scene(seed)
    codes, depths = R.codes_of(cs), [int(c["depth"]) for c in cs]
    for L in sorted(set(depths)):
        act, nb = O.neighbour_table(cs, L, mode=0)
        want_act = [i for i in range(len(cs)) if depths[i] == L]
        assert act.tolist() == want_act
        want = np.array([R.ref_neighbours(codes, depths, i) for i in want_act], dtype=np.int32)
        assert np.array_equal(nb, want)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("f64", [False, True])
def test_reference_rules_minimize(seed, f64):
    cs, h = scene(seed)
    a0, a1, it = (0.02, 0.04, 12) if seed % 2 else (2.0, 1.0, 12)
    got = O.tgv_minimize(cs, h, O.params(alpha0=a0, alpha1=a1, iters=it, op_mode=0), f64=f64)[0]
    want = R.tgv_minimize(cs, h, a0, a1, it, dt=np.float64 if f64 else np.float32)
    assert np.array_equal(bits(got), bits(want))
    assert np.abs(got[:, 4:]).max() > 0


@pytest.mark.parametrize("seed", SEEDS)
def test_reference_rules_one_iteration(seed):
    # one iteration per level from the median / parent-copy start: the
    # smallest case, where every rule of the five passes acts once
    cs, h = scene(seed)
    got = O.tgv_minimize(cs, h, O.params(alpha0=0.05, alpha1=0.1, iters=1, op_mode=0), f64=True)[0]
    want = R.tgv_minimize(cs, h, 0.05, 0.1, 1, dt=np.float64)
    assert np.array_equal(bits(got), bits(want))


def _leaves_of(n, cuts):
    edges = [0] + sorted(cuts) + [n]
    return [(edges[k], edges[k + 1] - 1) for k in range(len(edges) - 1) if edges[k] < edges[k + 1]]


@pytest.mark.parametrize("seed", SEEDS)
def test_frozen_border_oc(seed):
    # The leaves run in the order given. With code-sorted cubes every +a
    # neighbour has a larger code than its cube (Morton order grows with each
    # coordinate), so in ascending leaf order an out-of-leaf +a neighbour has
    # not moved yet when it is read; in descending order it has, and only the
    # level-start snapshot gives the reference's values.
    cs, h = scene(seed)
    n = len(cs)
    rng = np.random.default_rng(50 + seed)
    lv = _leaves_of(n, rng.choice(np.arange(3, n - 3), 3, replace=False).tolist())
    for order in (lv, lv[::-1]):
        arr = np.array(order, dtype=LEAF)
        got = O.tgv_minimize_oc(cs, h, arr, O.params(alpha0=0.5, alpha1=0.3, iters=10))
        want = R.tgv_minimize_oc(cs, h, order, 0.5, 0.3, 10)
        assert np.array_equal(bits(got), bits(want))
    # the snapshot matters: the same leaves solved as one differ
    one = R.tgv_minimize_oc(cs, h, [(0, n - 1)], 0.5, 0.3, 10)
    assert not np.array_equal(bits(one), bits(want))


@pytest.mark.parametrize("seed", SEEDS)
def test_streaming_rules(seed):
    cs, h = scene(seed)
    n = len(cs)
    rng = np.random.default_rng(70 + seed)
    lv = _leaves_of(n, rng.choice(np.arange(3, n - 3), 3, replace=False).tolist())
    arr = np.array(lv, dtype=LEAF)
    got = O.tgv_minimize_streaming(cs, h, arr, O.params(alpha0=0.5, alpha1=0.3, iters=10))
    want = R.tgv_minimize_streaming(cs, h, lv, 0.5, 0.3, 10)
    assert np.array_equal(bits(got), bits(want))


# ---------------------------------------------------------------- slips the pins must catch
# (file, text, replacement): each a one-line change to the oracle's source.
MUTANTS = {
    "first_writer_wins": ("cm[code_of(cubes[i])] = (uint32_t)i;  // last writer wins",
                          "cm.emplace(code_of(cubes[i]), (uint32_t)i);"),
    "no_coarser_fallback": ("        it = cm.find(parent(nb, depth));\n        if (it != cm.end()) return (int32_t)it->second;",
                            "        it = cm.end();\n        if (it != cm.end()) return (int32_t)it->second;"),
    "primal_plus_neighbour": (
        "return j >= 0 ? st[j].p[c] : T(0);\n                    };\n                    T div_p = (st[i].p[0] - gp(1, 0))",
        "return j >= 0 ? st[j].p[c] : T(0);\n                    };\n                    T div_p = (st[i].p[0] - gp(0, 0))"),
    "self_term_only_with_neighbour": (
        "return j >= 0 ? st[j].p[c] : T(0);\n                    };\n                    T div_p = (st[i].p[0] - gp(1, 0))",
        "return j >= 0 ? st[j].p[c] : T(0);\n                    };\n"
        "                    T div_p = ((L.nb[ai * 6 + 0] >= 0 ? st[i].p[0] : T(0)) - gp(1, 0))"),
    "half_in_v_divergence": (
        "return j >= 0 ? st[j].q[c] : T(0);\n                    };\n                    const State<T>& s = st[i];\n"
        "                    T dqx = (s.q[0] - gq(1, 0)) + (s.q[3] - gq(3, 3))",
        "return j >= 0 ? st[j].q[c] : T(0);\n                    };\n                    const State<T>& s = st[i];\n"
        "                    T dqx = (s.q[0] - gq(1, 0)) + T(0.5) * (s.q[3] - gq(3, 3))"),
    "frozen_read_live": ("                    if (fu && !inleaf(j)) return fu[j];",
                         "                    if (fu && !inleaf(j)) return ub[j];"),
    "oc_border_pq_zero": ("return j >= 0 ? st[j].p[c] : T(0);",
                          "return (j >= 0 && (!fu || inleaf(j))) ? st[j].p[c] : T(0);"),
    "streaming_border_pq_live": ("return (j >= 0 && inside(i, j)) ? st[(size_t)j].p[c] : T(0);",
                                 "return (j >= 0) ? st[(size_t)j].p[c] : T(0);"),
    "streaming_primal_plus_neighbour": (
        "st[(size_t)j].p[c] : T(0);\n                    };\n                    T div_p = (st[i].p[0] - gp(1, 0))",
        "st[(size_t)j].p[c] : T(0);\n                    };\n                    T div_p = (st[i].p[0] - gp(0, 0))"),
    "streaming_border_bars_live": ("return live(i, j) ? ub[(size_t)j] : su[(size_t)j];",
                                   "return ub[(size_t)j];"),
    "streaming_last_writer": ("int64_t j = it != leaf_map.end() ? (int64_t)it->second : bisect(k);",
                              "int64_t j = it != leaf_map.end() ? (int64_t)it->second : (int64_t)(-1);"),
}

PIN_SCRIPT = r"""
import sys
sys.path[:0] = [{oracle!r}, {tests!r}]
import numpy as np
import test_ref_rules_pins as P
fails = 0
checks = [P.test_reference_table_brute_force, P.test_reference_rules_minimize, P.test_frozen_border_oc,
          P.test_streaming_rules]
for fn in checks:
    for seed in P.SEEDS:
        try:
            if fn is P.test_reference_rules_minimize:
                fn(seed, False)
            else:
                fn(seed)
        except AssertionError:
            fails += 1
print("FAILS", fails)
"""


def _build_mutant(name, out_dir):
    src = open(os.path.join(ROOT, "oracle", "tgv_oracle.cpp")).read()
    old, new = MUTANTS[name]
    assert src.count(old) == 1, name
    d = os.path.join(out_dir, name)
    os.makedirs(d)
    with open(os.path.join(d, "tgv_oracle.cpp"), "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(d, "liboracle.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp", "-o", so,
                    os.path.join(d, "tgv_oracle.cpp"), os.path.join(ROOT, "oracle", "mlp_oracle.cpp")],
                   check=True, capture_output=True)
    return so


def _run_pins(so):
    env = dict(os.environ, ORACLE_LIB=so)
    script = PIN_SCRIPT.format(oracle=os.path.join(ROOT, "oracle"), tests=HERE)
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return int(r.stdout.split("FAILS")[-1])


def test_pins_reject_oracle_slips():
    with tempfile.TemporaryDirectory() as tmp:
        with ThreadPoolExecutor(max_workers=4) as ex:
            libs = dict(zip(MUTANTS, ex.map(lambda m: _build_mutant(m, tmp), MUTANTS)))
            fails = dict(zip(libs, ex.map(_run_pins, libs.values())))
        caught = {m: f for m, f in fails.items() if f > 0}
        assert caught == {m: fails[m] for m in MUTANTS}, f"slips not caught: {set(MUTANTS) - set(caught)}"
