"""GPU (libtgvx.so through its C ABI) against the CPU oracle on the same inputs.

Bars: histogram bins bit-exact; FP32 and FP64 TGV states bit-exact in the
reference update rules (REF) and in the exact-adjoint rules (ADJ) -- the FP64
bar of the workload description is |du| <= 1e-10, bit equality is stronger;
FP64 energies / gaps within 1e-9 relative (reduction order differs).
"""
import numpy as np
import pytest

import oracle as O
from tgvreconclaude_amd import BF16, FP32, FP64, OP_ADJ, OP_REF, TGVParams, scene

pytestmark = pytest.mark.gpu


def _state13(st):
    if st.dtype.names:
        return np.concatenate([st["u"][:, None], st["v"], st["p"], st["q"]], axis=1)
    return st


def test_histograms_bit_exact_nb10(ctx, sphere):
    sc, hist = sphere
    g = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 10)
    assert g.shape == hist.shape
    assert np.array_equal(g, hist)
    assert hist.sum() > 0


def test_histograms_bit_exact_nb8(ctx, sphere):
    sc, _ = sphere
    vs, keep = O.views_from_scene(sc)
    ref = O.histograms(sc.cubes, vs, sc.origin, sc.r_root, 8)
    g = ctx.histograms(sc.cubes, sc.views(), sc.origin, sc.r_root, 8)
    assert np.array_equal(g, ref)


def test_mip_build_bit_exact(ctx, sphere):
    sc, _ = sphere
    v = sc.views()[3]
    ov, keep = O.view(sc.depths[3], v.fx, v.fy, v.cx, v.cy, list(v.world_to_cam))
    a = ctx.mip_build(v)
    b = O.mip_build(ov)
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert x.shape == y.shape
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_median_bit_exact(ctx, sphere):
    _, hist = sphere
    g = ctx.median_indicator(hist)
    ref = np.array([O.median_indicator(h) for h in hist[:2000]], dtype=np.float32)
    assert np.array_equal(g[:2000].view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("mode", [OP_REF, OP_ADJ])
def test_tgv_fp32_bit_exact(ctx, sphere, mode):
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, iters=300, precision=FP32, op_mode=mode)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=300, op_mode=mode))
    g = _state13(st)
    bad = np.flatnonzero(np.any(g.view(np.uint32) != ref.view(np.uint32), axis=1))
    assert bad.size == 0, f"{bad.size} cubes differ, first {bad[:5]}: {g[bad[:1]]} vs {ref[bad[:1]]}"
    assert rep.n_levels == len(it)
    for l in range(rep.n_levels):
        P, D = rep.level_energy[l], rep.level_dual[l]
        assert P == pytest.approx(pd[l][0], rel=1e-9, abs=1e-9)
        assert D == pytest.approx(pd[l][1], rel=1e-9, abs=1e-9)


@pytest.mark.parametrize("mode", [OP_REF, OP_ADJ])
def test_tgv_fp64_matches_oracle(ctx, sphere, mode):
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, iters=300, precision=FP64, op_mode=mode)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, _, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=300, op_mode=mode), f64=True)
    assert np.max(np.abs(st[:, 0] - ref[:, 0])) <= 1e-10
    assert np.array_equal(st.view(np.uint64), ref.view(np.uint64))


def test_tgv_to_gap_adj(ctx, sphere):
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, precision=FP32, op_mode=OP_ADJ, gap_tol=1e-3, chunk=50, max_iters=3000)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-3,
                                                             chunk=50, max_iters=3000))
    for l in range(rep.n_levels):
        lv = rep.levels()[l]
        assert lv["rel_gap"] <= 1e-3
        assert lv["iters"] == it[l]
    assert np.array_equal(_state13(st).view(np.uint32), ref.view(np.uint32))


def test_bf16_storage_close(ctx, sphere):
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, iters=300, precision=BF16, op_mode=OP_ADJ)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, _, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=300, op_mode=1), f64=True)
    err = np.abs(st["u"].astype(np.float64) - ref[:, 0])
    assert np.all(np.isfinite(st["u"]))
    assert np.quantile(err, 0.99) < 0.05


def test_oc_matches_oracle(ctx, sphere):
    sc, hist = sphere
    leaves = scene.treetop(sc.cubes, 8000)
    assert len(leaves) > 2
    prm = TGVParams(alpha0=2.0, alpha1=1.0, iters=60, precision=FP32, op_mode=OP_REF)
    st, rep = ctx.solve_oc(sc.cubes, hist, leaves, prm)
    ref = O.tgv_minimize_oc(sc.cubes, hist, leaves, O.params(alpha0=2.0, alpha1=1.0, iters=60))
    assert np.array_equal(_state13(st).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("fallback,warm_dual", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_warm_start_and_dual_recovery_bit_exact(ctx, sphere, fallback, warm_dual):
    # an initial (u, v) field for every level, the dual recovered from it, and
    # (fallback=1) the per-level choice between warm and cold by initial gap
    sc, hist = sphere
    rng = np.random.default_rng(7)
    ref_cold, _, _, _ = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=200, op_mode=1))
    init = np.ascontiguousarray(ref_cold[:, :4] + rng.normal(0, 0.02, (len(hist), 4)).astype(np.float32))
    prm = TGVParams(alpha0=2.0, alpha1=1.0, iters=120, precision=FP32, op_mode=OP_ADJ, dual_eps=1e-4,
                    warm_fallback=fallback, warm_dual=warm_dual)
    st, rep = ctx.solve(sc.cubes, hist, prm, init_uv=init)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, iters=120, op_mode=1,
                                                             dual_eps=1e-4, warm_fallback=fallback,
                                                             warm_dual=warm_dual), init_uv=init)
    assert np.array_equal(_state13(st).view(np.uint32), ref.view(np.uint32))
    if fallback:
        assert 0 < sum(rep.level_warm[i] for i in range(rep.n_levels)) <= rep.n_levels


def test_cold_dual_recovery_to_gap(ctx, sphere):
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, precision=FP32, op_mode=OP_ADJ, gap_tol=1e-3, chunk=50, max_iters=4000,
                    dual_eps=1e-4)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-3, chunk=50,
                                                             max_iters=4000, dual_eps=1e-4))
    assert [l["iters"] for l in rep.levels()] == it.tolist()
    assert np.array_equal(_state13(st).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("prec", [FP32, FP64])
def test_fused_iteration_identical_to_two_kernels(ctx, sphere, prec, variant):
    # exact-adjoint mode: one fused kernel per iteration (LDS-tiled or not) ==
    # dual + primal kernels, bit for bit
    sc, hist = sphere
    a, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=75, precision=prec, op_mode=OP_ADJ,
                                               fused=variant))
    b, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=75, precision=prec, op_mode=OP_ADJ,
                                               fused=0))
    a, b = _state13(a), _state13(b)
    assert np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                          b.view(np.uint32 if b.dtype == np.float32 else np.uint64))


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5, 6])
def test_fused_bf16_matches_two_kernels(ctx, sphere, variant):
    sc, hist = sphere
    a, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=BF16, op_mode=OP_ADJ,
                                               fused=variant))
    b, _ = ctx.solve(sc.cubes, hist, TGVParams(alpha0=2.0, alpha1=1.0, iters=40, precision=BF16, op_mode=OP_ADJ,
                                               fused=0))
    assert np.array_equal(_state13(a).view(np.uint32), _state13(b).view(np.uint32))


def test_parameter_change_on_resident_problem(ctx, sphere):
    # graphs captured for one set of weights / steps must not be replayed for
    # another: a second solve with other parameters equals a fresh solve
    sc, hist = sphere
    a = TGVParams(alpha0=2.0, alpha1=1.0, iters=60, op_mode=OP_ADJ)
    b = TGVParams(alpha0=0.5, alpha1=0.3, iters=60, op_mode=OP_ADJ, sigma=0.5, tau=0.125)
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    ctx.run(a)
    ctx.run(b)
    got = ctx.state()
    ref, _ = ctx.solve(sc.cubes, hist, b)
    assert np.array_equal(_state13(got).view(np.uint32), _state13(ref).view(np.uint32))
    ctx.load(sc.cubes, hist, precision=FP32, op_mode=OP_ADJ)
    ctx.run(a)
    first = ctx.state()
    ctx.run(b)
    ctx.run(a)
    assert np.array_equal(_state13(ctx.state()).view(np.uint32), _state13(first).view(np.uint32))


@pytest.mark.parametrize("fused", [0, 1, 4, 5, 6])
def test_to_gap_adj_large_step_ratio(ctx, sphere, fused):
    # sigma / tau = 1024 (sigma = 8, tau = 1/128, sigma * tau = 1/16): the
    # same iteration with other steps, bit-exact with the oracle, counts equal
    sc, hist = sphere
    prm = TGVParams(alpha0=2.0, alpha1=1.0, op_mode=OP_ADJ, gap_tol=1e-4, chunk=50, max_iters=8000,
                    sigma=8.0, tau=1.0 / 128, fused=fused)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-4,
                                                             chunk=50, max_iters=8000, sigma=8.0, tau=1.0 / 128))
    assert [l["iters"] for l in rep.levels()] == it.tolist()
    assert all(l["rel_gap"] <= 1e-4 for l in rep.levels())
    assert np.array_equal(_state13(st).view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("prec", [FP32, FP64])
def test_to_gap_with_depth_schedule(ctx, sphere, prec):
    # per-depth steps (sigma / tau = 256 up to depth 5, 1024 at 6, 4096 at 7;
    # sigma * tau = 1/16): bit-exact with the oracle under the same schedule,
    # iteration counts equal, every level reaches the gap
    from tgvreconclaude_amd import step_schedule
    sc, hist = sphere
    sd, td = step_schedule([256] * 6 + [1024, 4096])
    prm = TGVParams(alpha0=2.0, alpha1=1.0, op_mode=OP_ADJ, gap_tol=1e-4, chunk=50, max_iters=8000,
                    sigma=8.0, tau=1.0 / 128, precision=prec, fused=5, block=512, sigma_depth=sd, tau_depth=td)
    st, rep = ctx.solve(sc.cubes, hist, prm)
    ref, _, it, pd = O.tgv_minimize(sc.cubes, hist, O.params(alpha0=2.0, alpha1=1.0, op_mode=1, gap_tol=1e-4,
                                                             chunk=50, max_iters=8000, sigma=8.0, tau=1.0 / 128,
                                                             sigma_depth=sd, tau_depth=td), f64=prec == FP64)
    assert [l["iters"] for l in rep.levels()] == it.tolist()
    assert all(l["rel_gap"] <= 1e-4 for l in rep.levels())
    w = np.uint64 if prec == FP64 else np.uint32
    assert np.array_equal(np.ascontiguousarray(_state13(st)).view(w), ref.view(w))
