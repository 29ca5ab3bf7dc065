"""tgvx: the global-TGV hot path of TGVReconClaude on AMD Instinct MI355X (gfx950).

Python front end of the C ABI in include/tgvx.h. The functions mirror the
reference's C++ interface (src/tgv.hpp, src/histogram.hpp):

    compute_histograms(cubes, views, origin, r_root, nb)      -> uint32 [n, nb]
    tgv_minimize(cubes, hists, origin, r_root, params)        -> state records
    tgv_minimize_oc(cubes, hists, leaves, origin, r_root, params)
    median_indicator(hists)                                   -> float32 [n]

Every call runs on the GPU through libtgvx.so; there is no CPU fallback. A
missing or unloadable library raises RuntimeError, as does any failure reported
by the library (the reference signals errors with std::runtime_error).
"""
import ctypes as C

import numpy as np

from . import _native as N
from ._native import BF16, BF16_DUAL, CUBE_DT, FP32, FP64, LEAF_DT, OP_ADJ, OP_REF, STATE_DT, Params, Report

__all__ = [
    "Context", "TGVParams", "compute_histograms", "tgv_minimize", "tgv_minimize_oc", "median_indicator",
    "CUBE_DT", "STATE_DT", "LEAF_DT", "FP64", "FP32", "BF16", "BF16_DUAL", "OP_REF", "OP_ADJ", "make_view", "step_schedule",
]


def TGVParams(**kw):
    """TGVParams (src/tgv.hpp:24-28) with the extra controls of tgvx_params.

    Defaults reproduce the reference: alpha0=0.02, alpha1=0.04, iters=200,
    FP32, reference update rules, fixed iterations.
    """
    p = Params()
    N.lib().tgvx_default_params(C.byref(p))
    for k, v in kw.items():
        if k == "lambda":
            k = "lambda_"
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        if k in ("sigma_depth", "tau_depth"):  # per octree depth, 22 entries (0 = sigma / tau)
            v = (C.c_float * 22)(*[float(x) for x in v])
        setattr(p, k, v)
    return p


def step_schedule(ratios, product=1.0 / 16):
    """Per-depth steps (sigma_depth, tau_depth) with sigma / tau = ratios[depth] and
    sigma * tau = product, for TGVParams(sigma_depth=..., tau_depth=...); depths
    without a ratio (None or 0) keep the scalar steps."""
    sd, td = [0.0] * 22, [0.0] * 22
    for d, r in enumerate(ratios):
        if r:
            sd[d], td[d] = (product * r) ** 0.5, (product / r) ** 0.5
    return sd, td


def make_view(depths, fx, fy, cx, cy, world_to_cam, lidar=False):
    """One depth map (DepthMap + DepthMipmap::build inputs, src/depth_map.hpp:7-12).

    depths: float32 [H, W], NaN = invalid; world_to_cam: 3x4 (rows 0-2 of Rt^-1).
    Returns (View, keep-alive array).
    """
    d = np.ascontiguousarray(depths, dtype=np.float32)
    v = N.View()
    v.depths = d.ctypes.data
    v.height, v.width = d.shape
    v.fx, v.fy, v.cx, v.cy = fx, fy, cx, cy
    m = np.asarray(world_to_cam, dtype=np.float32).reshape(12)
    for i in range(12):
        v.world_to_cam[i] = float(m[i])
    v.lidar = 1 if lidar else 0
    return v, d


class Context:
    """One tgvx context (device buffers + stream) on one GPU."""

    def __init__(self, device=0):
        self._L = N.lib()
        h = C.c_void_p()
        rc = self._L.tgvx_create(C.byref(h), device)
        if rc != 0:
            raise RuntimeError(f"tgvx_create failed with code {rc} (no GPU visible?)")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self._L.tgvx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"tgvx error {rc}: {self._L.tgvx_last_error(self.h).decode()}")

    def device_info(self):
        name = C.create_string_buffer(64)
        cus = C.c_int32()
        self._check(self._L.tgvx_device_info(self.h, name, 64, C.byref(cus)))
        return name.value.decode(), cus.value

    # ---- histograms
    def histograms(self, cubes, views, origin, r_root, nb=8):
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        arr = (N.View * max(1, len(views)))(*[v for v in views])
        o = np.asarray(origin, dtype=np.float32)
        out = np.zeros((len(cubes), nb), dtype=np.uint32)
        self._check(self._L.tgvx_histograms(self.h, N.ptr(cubes), len(cubes), arr, len(views), N.ptr(o),
                                            float(r_root), nb, N.ptr(out)))
        return out

    def mip_build(self, view):
        w, h = view.width, view.height
        out = np.zeros(w * h * 4 // 3 + 64, dtype=np.float32)
        dims = np.zeros(2 * 16, dtype=np.int32)
        nl = C.c_int32()
        self._check(self._L.tgvx_mip_build(self.h, C.byref(view), N.ptr(out), N.ptr(dims), C.byref(nl)))
        levels, off = [], 0
        for l in range(nl.value):
            lw, lh = int(dims[2 * l]), int(dims[2 * l + 1])
            levels.append(out[off:off + lw * lh].reshape(lh, lw).copy())
            off += lw * lh
        return levels

    def median_indicator(self, hists):
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        out = np.zeros(len(h), dtype=np.float32)
        self._check(self._L.tgvx_median_init(self.h, N.ptr(h), len(h), h.shape[1], N.ptr(out)))
        return out

    # ---- solves
    @staticmethod
    def _state_array(n, precision):
        if precision == FP64:
            return np.zeros((n, 13), dtype=np.float64)
        return np.zeros(n, dtype=STATE_DT)

    def solve(self, cubes, hists, params=None, init_uv=None):
        """tgv_minimize; returns (state, Report)."""
        params = params or TGVParams()
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        st = self._state_array(len(cubes), params.precision)
        rep = Report()
        init = None if init_uv is None else np.ascontiguousarray(init_uv, dtype=np.float32)
        self._check(self._L.tgvx_solve(self.h, N.ptr(cubes), len(cubes), N.ptr(h), h.shape[1] if h.ndim == 2 else 8,
                                       C.byref(params), N.ptr(init), N.ptr(st), C.byref(rep)))
        return st, rep

    def solve_streaming(self, cubes, hists, leaves, params=None):
        """tgv_minimize_streaming on arrays; returns (state, Report)."""
        params = params or TGVParams()
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        lv = np.ascontiguousarray(leaves, dtype=LEAF_DT)
        st = self._state_array(len(cubes), params.precision)
        rep = Report()
        self._check(self._L.tgvx_solve_streaming(self.h, N.ptr(cubes), len(cubes), N.ptr(h), h.shape[1], N.ptr(lv),
                                                 len(lv), C.byref(params), N.ptr(st), C.byref(rep)))
        return st, rep

    def solve_streaming_ooc(self, cubes, hists, leaves, params=None, device_bytes=256 << 20):
        """tgv_minimize_streaming with at most device_bytes of per-cube data on the
        device (FP32); returns (state, Report, out-of-core report dict)."""
        params = params or TGVParams()
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        lv = np.ascontiguousarray(leaves, dtype=LEAF_DT)
        st = self._state_array(len(cubes), FP32)
        rep, orep = Report(), N.OocReport()
        self._check(self._L.tgvx_solve_streaming_ooc(self.h, N.ptr(cubes), len(cubes), N.ptr(h), h.shape[1],
                                                     N.ptr(lv), len(lv), C.byref(params), int(device_bytes),
                                                     N.ptr(st), C.byref(rep), C.byref(orep)))
        return st, rep, orep.as_dict()

    def solve_oc(self, cubes, hists, leaves, params=None):
        params = params or TGVParams()
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        lv = np.ascontiguousarray(leaves, dtype=LEAF_DT)
        st = self._state_array(len(cubes), params.precision)
        rep = Report()
        self._check(self._L.tgvx_solve_oc(self.h, N.ptr(cubes), len(cubes), N.ptr(h),
                                          h.shape[1] if h.ndim == 2 else 8, N.ptr(lv), len(lv),
                                          C.byref(params), N.ptr(st), C.byref(rep)))
        return st, rep

    # ---- resident problem
    def load(self, cubes, hists, precision=FP32, op_mode=OP_REF):
        cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        h = np.ascontiguousarray(hists, dtype=np.uint32)
        self._n, self._prec = len(cubes), precision
        self._check(self._L.tgvx_problem_load(self.h, N.ptr(cubes), len(cubes), N.ptr(h), h.shape[1], precision,
                                              op_mode))

    def load_blocks(self, blocks, precision=FP32, op_mode=OP_ADJ):
        """Several independent octrees [(cubes, hists), ...] as one resident problem."""
        keep = [(np.ascontiguousarray(c, dtype=CUBE_DT), np.ascontiguousarray(h, dtype=np.uint32)) for c, h in blocks]
        arr = (N.Block * len(keep))()
        for i, (c, h) in enumerate(keep):
            arr[i].cubes, arr[i].n, arr[i].bins = c.ctypes.data, len(c), h.ctypes.data
        nb = keep[0][1].shape[1]
        self._n, self._prec = sum(len(c) for c, _ in keep), precision
        self._block_sizes = [len(c) for c, _ in keep]
        self._check(self._L.tgvx_problem_load_blocks(self.h, arr, len(keep), nb, precision, op_mode))

    def autotune(self, params, iters=20):
        """Choose each level's launch configuration for the loaded problem (see tgvx_autotune)."""
        rep = N.TuneReport()
        self._check(self._L.tgvx_autotune(self.h, C.byref(params), iters, C.byref(rep)))
        return rep.levels()

    def solve_batched(self, blocks, params=None):
        """tgv_minimize of several independent octrees in one resident problem; (states per block, Report)."""
        params = params or TGVParams()
        keep = [(np.ascontiguousarray(c, dtype=CUBE_DT), np.ascontiguousarray(h, dtype=np.uint32)) for c, h in blocks]
        arr = (N.Block * len(keep))()
        states = [self._state_array(len(c), params.precision) for c, _ in keep]
        outs = (C.c_void_p * len(keep))()
        for i, (c, h) in enumerate(keep):
            arr[i].cubes, arr[i].n, arr[i].bins = c.ctypes.data, len(c), h.ctypes.data
            outs[i] = states[i].ctypes.data
        rep = Report()
        self._check(self._L.tgvx_solve_batched(self.h, arr, len(keep), keep[0][1].shape[1], C.byref(params), outs,
                                               C.byref(rep)))
        self._n, self._prec = sum(len(c) for c, _ in keep), params.precision
        self._block_sizes = [len(c) for c, _ in keep]
        return states, rep

    def block_report(self):
        """(iterations [levels, blocks], relative gaps [levels, blocks]) of the last solve."""
        nl, nb = len(self.levels()), len(self._block_sizes)
        it = np.zeros((nl, nb), dtype=np.int32)
        g = np.zeros((nl, nb), dtype=np.float64)
        self._check(self._L.tgvx_problem_block_report(self.h, N.ptr(it), N.ptr(g)))
        return it, g

    def split_blocks(self, state):
        out, o = [], 0
        for n in self._block_sizes:
            out.append(state[o:o + n])
            o += n
        return out

    def run(self, params, init_uv=None, init_state=None):
        """Solve the resident problem; init_uv (n x 4: u, v) or init_state (n
        records of 13 FP32 values: u, v, p, q, as STATE_DT) replaces each
        level's cold initialisation."""
        rep = Report()
        if init_state is not None:
            st = np.ascontiguousarray(init_state)
            st = st.view(np.float32).reshape(-1, 13) if st.dtype.names else np.ascontiguousarray(st, np.float32)
            assert st.shape == (self._n, 13)
            self._check(self._L.tgvx_problem_solve_from(self.h, C.byref(params), N.ptr(st), C.byref(rep)))
            return rep
        init = None if init_uv is None else np.ascontiguousarray(init_uv, dtype=np.float32)
        self._check(self._L.tgvx_problem_solve(self.h, C.byref(params), N.ptr(init), C.byref(rep)))
        return rep

    def surface(self, origin, r_root, u=None, has_data=None):
        """extract_surface over the resident problem: (verts [nv, 3] float32, tris [nt, 3] uint32)."""
        uu = None if u is None else np.ascontiguousarray(u, dtype=np.float32)
        hd = None if has_data is None else np.ascontiguousarray(has_data, dtype=np.uint8)
        o = np.ascontiguousarray(origin, dtype=np.float32)
        nv, nt = C.c_uint64(), C.c_uint64()
        self._check(self._L.tgvx_surface_extract(self.h, N.ptr(uu), N.ptr(hd), N.ptr(o), float(r_root),
                                                 C.byref(nv), C.byref(nt)))
        verts = np.zeros((nv.value, 3), dtype=np.float32)
        tris = np.zeros((nt.value, 3), dtype=np.uint32)
        self._check(self._L.tgvx_surface_copy(self.h, N.ptr(verts), N.ptr(tris)))
        return verts, tris

    def octree_build(self, depths, cams, origin, r_root, max_depth, last_stage=3):
        """Stages 1..last_stage on the device from depth maps [V, H, W] and
        cameras (tgvs_camera records); returns (cubes sorted by (code, depth),
        report dict)."""
        d = np.ascontiguousarray(depths, dtype=np.float32)
        V, H, W = d.shape
        views = (N.DepthView * max(1, V))()
        for k in range(V):
            v = views[k]
            v.depths = d[k].ctypes.data
            v.width, v.height = W, H
            v.fx, v.fy, v.cx, v.cy = cams[k].fx, cams[k].fy, cams[k].cx, cams[k].cy
            for i in range(12):
                v.cam_to_world[i] = cams[k].cam_to_world[i]
        o = np.ascontiguousarray(origin, dtype=np.float32)
        n, rep = C.c_uint64(), N.OctreeReport()
        self._check(self._L.tgvx_octree_build(self.h, views, V, N.ptr(o), float(r_root), int(max_depth),
                                              int(last_stage), C.byref(n), C.byref(rep)))
        return self._octree_out(n.value), rep.as_dict()

    def octree_merge(self, cubes):
        """kway_merge on the device over the concatenated cube lists."""
        c = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        n, rep = C.c_uint64(), N.OctreeReport()
        self._check(self._L.tgvx_octree_merge(self.h, N.ptr(c), len(c), C.byref(n), C.byref(rep)))
        return self._octree_out(n.value), rep.as_dict()

    def octree_balance(self, cubes, r_root):
        """balance_octree on the device; returns (cubes, report dict)."""
        c = np.ascontiguousarray(cubes, dtype=CUBE_DT)
        n, rep = C.c_uint64(), N.OctreeReport()
        self._check(self._L.tgvx_octree_balance(self.h, N.ptr(c), len(c), float(r_root), C.byref(n),
                                                C.byref(rep)))
        return self._octree_out(n.value), rep.as_dict()

    def _octree_out(self, n):
        out = np.zeros(n, dtype=CUBE_DT)
        self._check(self._L.tgvx_octree_copy(self.h, N.ptr(out)))
        return out

    def run_levels(self, params, level_iters):
        """Solve the resident problem with a fixed iteration count per level."""
        it = np.ascontiguousarray(level_iters, dtype=np.int32)
        rep = Report()
        self._check(self._L.tgvx_problem_solve_levels(self.h, C.byref(params), N.ptr(it), len(it), C.byref(rep)))
        return rep

    def state(self, with_bar=False):
        st = self._state_array(self._n, self._prec)
        bar = None
        if with_bar:
            bar = np.zeros((self._n, 4), dtype=np.float64 if self._prec == FP64 else np.float32)
        self._check(self._L.tgvx_problem_state(self.h, N.ptr(st), N.ptr(bar)))
        return (st, bar) if with_bar else st

    def gap(self, depth, params):
        P, D = C.c_double(), C.c_double()
        self._check(self._L.tgvx_problem_gap(self.h, depth, C.byref(params), C.byref(P), C.byref(D)))
        return P.value, D.value

    def levels(self):
        d = np.zeros(22, dtype=np.int32)
        c = np.zeros(22, dtype=np.uint32)
        n = C.c_int32()
        self._check(self._L.tgvx_problem_levels(self.h, N.ptr(d), N.ptr(c), C.byref(n)))
        return [(int(d[i]), int(c[i])) for i in range(n.value)]

    # ---- warm-start network
    def mlp_set(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == N.MLP_NPARAMS
        self._check(self._L.tgvx_mlp_set(self.h, N.ptr(w)))

    def mlp_get(self):
        w = np.zeros(N.MLP_NPARAMS, dtype=np.float32)
        self._check(self._L.tgvx_mlp_get(self.h, N.ptr(w)))
        return w

    def mlp_forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 4), dtype=np.float32)
        self._check(self._L.tgvx_mlp_forward(self.h, N.ptr(x), len(x), N.ptr(y)))
        return y

    def mlp_loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(N.MLP_NPARAMS, dtype=np.float32)
        loss = C.c_double()
        self._check(self._L.tgvx_mlp_loss_grad(self.h, N.ptr(x), N.ptr(t), len(x), N.ptr(g), C.byref(loss)))
        return loss.value, g

    def mlp_train(self, x, t, epochs=10, batch=4096, lr=1e-3, seed=0):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        out = np.zeros(epochs, dtype=np.float32)
        self._check(self._L.tgvx_mlp_train(self.h, N.ptr(x), N.ptr(t), len(x), epochs, batch, lr, seed, N.ptr(out)))
        return out

    def features(self, depth):
        """(features [n_L, 30], current (u, v) [n_L, 4], input indices [n_L]) of one level."""
        n = dict(self.levels())[depth]
        x = np.zeros((n, 30), dtype=np.float32)
        uv = np.zeros((n, 4), dtype=np.float32)
        idx = np.zeros(n, dtype=np.uint32)
        self._check(self._L.tgvx_problem_features(self.h, depth, N.ptr(x), N.ptr(uv), N.ptr(idx)))
        return x, uv, idx

    def level_bench(self, depth, params, iters):
        """(total ms of `iters` iterations, avg ms of one dual / primal / fused launch)."""
        t, d, p, f = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self._check(self._L.tgvx_level_bench(self.h, depth, C.byref(params), iters, C.byref(t), C.byref(d),
                                             C.byref(p), C.byref(f)))
        return t.value, d.value, p.value, f.value

    def grid_barrier_us(self, workgroups, block=1024, barriers=2000):
        """Microseconds per grid barrier of the persistent level kernel in a
        cooperative launch of `workgroups` workgroups of `block` threads."""
        us = C.c_double()
        self._check(self._L.tgvx_grid_barrier_bench(self.h, workgroups, block, barriers, C.byref(us)))
        return us.value

    def tile_sets4(self, depth, tile):
        """The four-iteration kernel's tile sets at `depth` (exact-adjoint problem):
        (usable, largest slot count, LDS bytes per workgroup)."""
        ok, ns, lds = C.c_int32(), C.c_int32(), C.c_int64()
        self._check(self._L.tgvx_tile_sets4(self.h, depth, tile, C.byref(ok), C.byref(ns), C.byref(lds)))
        return bool(ok.value), ns.value, lds.value

    def tile_sets(self, depth, tile):
        """The two-iteration kernel's tile sets at `depth` (exact-adjoint problem):
        (usable, (largest PA, DB, PC), LDS bytes per workgroup)."""
        ok, lds = C.c_int32(), C.c_int64()
        sets = (C.c_int32 * 3)()
        self._check(self._L.tgvx_tile_sets(self.h, depth, tile, C.byref(ok), sets, C.byref(lds)))
        return bool(ok.value), tuple(sets), lds.value


_default_ctx = None


def _ctx():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def compute_histograms(cubes, views, origin, r_root, nb=8):
    """compute_histograms (src/histogram.hpp:44-48) with nb bins."""
    return _ctx().histograms(cubes, views, origin, r_root, nb)


def median_indicator(hists):
    """CubeHistogram::median_indicator (src/histogram.hpp:18) of every row."""
    return _ctx().median_indicator(hists)


def tgv_minimize(cubes, hists, origin=None, r_root=1.0, params=None, init_uv=None):
    """tgv_minimize (src/tgv.hpp:31-36). origin / r_root are unused, as in the reference."""
    return _ctx().solve(cubes, hists, params, init_uv)[0]


def tgv_minimize_oc(cubes, hists, leaves, origin=None, r_root=1.0, params=None):
    """tgv_minimize_oc (src/tgv.hpp:42-48)."""
    return _ctx().solve_oc(cubes, hists, leaves, params)[0]
