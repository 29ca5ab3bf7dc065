"""The warm-start network on the GPU (MFMA kernels) against the FP64 CPU oracle."""
import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import FP32, OP_ADJ, TGVParams

pytestmark = pytest.mark.gpu


def _rand(seed, n):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.2, O.MLP_NP).astype(np.float32)
    x = rng.normal(0, 1, (n, 30)).astype(np.float32)
    t = rng.normal(0, 0.5, (n, 4)).astype(np.float32)
    return w, x, t


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_forward_matches_oracle(ctx, n):
    w, x, _ = _rand(n, n)
    ctx.mlp_set(w)
    y = ctx.mlp_forward(x)
    ref = O.mlp_forward(w.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(y, ref, rtol=2e-4, atol=2e-4 * np.abs(ref).max())


@pytest.mark.parametrize("n", [17, 64, 300])
def test_gradient_matches_oracle(ctx, n):
    w, x, t = _rand(100 + n, n)
    ctx.mlp_set(w)
    loss, g = ctx.mlp_loss_grad(x, t)
    rl, rg = O.mlp_loss_grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert loss == pytest.approx(rl, rel=1e-4)
    err = np.abs(g - rg).max() / np.abs(rg).max()
    assert err < 1e-4, err


def test_training_fits_a_teacher(ctx):
    # targets from a random teacher network: the loss must drop by 10x
    teacher, x, _ = _rand(7, 20000)
    t = O.mlp_forward(teacher.astype(np.float64), x.astype(np.float64)).astype(np.float32)
    w0, _, _ = _rand(8, 1)
    ctx.mlp_set(w0 * 0.5)
    loss0, _ = ctx.mlp_loss_grad(x, t)
    curve = ctx.mlp_train(x, t, epochs=30, batch=1024, lr=2e-3, seed=1)
    assert np.all(np.isfinite(curve))
    assert curve[-1] < 0.1 * loss0, (loss0, curve)


def test_one_adam_step_matches_oracle(ctx):
    w, x, t = _rand(9, 256)
    ctx.mlp_set(w)
    _, g = ctx.mlp_loss_grad(x, t)
    ctx.mlp_train(x, t, epochs=1, batch=256, lr=1e-3, seed=0)
    w1 = ctx.mlp_get()
    wr = w.astype(np.float64).copy()
    m, v = np.zeros(O.MLP_NP), np.zeros(O.MLP_NP)
    O.mlp_adam(wr, m, v, g.astype(np.float64), 1, 1e-3)
    np.testing.assert_allclose(w1, wr, rtol=1e-5, atol=2e-6)


def test_features_match_numpy(ctx, sphere):
    sc, hist = sphere
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    ctx.run(TGVParams(alpha0=2.0, alpha1=1.0, iters=30, op_mode=OP_ADJ))
    st = ctx.state()
    u = st["u"]
    v = st["v"]
    d = int(sc.cubes["depth"].max())
    x, uv, idx = ctx.features(d)
    assert np.array_equal(uv[:, 0], u[idx]) and np.array_equal(uv[:, 1:], v[idx])
    act, nb = O.neighbour_table(sc.cubes, d, mode=1)
    assert np.array_equal(act, idx)
    W = hist.sum(1).astype(np.float64)
    norm = np.where(W[:, None] > 0, hist / np.maximum(W, 1)[:, None], 0.0)
    med = np.array([O.median_indicator(h) for h in hist], dtype=np.float64)
    codes = (sc.cubes["hi"].astype(np.uint64) << 32) | sc.cubes["lo"]
    key = {(int(c), int(dd)): i for i, (c, dd) in enumerate(zip(codes, sc.cubes["depth"]))}
    for r in range(0, len(idx), 97):
        i = idx[r]
        f = np.zeros(30)
        f[:10] = norm[i]
        js = [j for j in nb[r] if j >= 0]
        if js:
            f[10:20] = norm[js].mean(0)
        pc = O.parent_code(int(codes[i]), d)
        p = key.get((int(pc), d - 1), -1)
        f[20] = u[p] if p >= 0 else med[i]
        f[21:24] = v[p] if p >= 0 else 0
        for a in range(3):
            jp, jm = nb[r][2 * a], nb[r][2 * a + 1]
            f[24 + a] = 0.5 * ((med[jp] if jp >= 0 else med[i]) - (med[jm] if jm >= 0 else med[i]))
        f[27] = np.log1p(W[i]) / 8
        f[28] = len(js) / 6
        f[29] = med[i]
        np.testing.assert_allclose(x[r], f, rtol=1e-5, atol=1e-6)


def test_warm_start_from_network_runs_and_falls_back(ctx, sphere):
    # an untrained network: every level must stay valid, and the per-level
    # fallback must keep the cold start wherever the network starts worse
    sc, hist = sphere
    w, _, _ = _rand(11, 1)
    ctx.mlp_set(w)
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    rep = ctx.run(TGVParams(alpha0=2.0, alpha1=1.0, op_mode=OP_ADJ, warm_start=1, warm_fallback=1, dual_eps=1e-4,
                            gap_tol=1e-3, chunk=50, max_iters=4000))
    for l in rep.levels():
        assert l["rel_gap"] <= 1e-3
