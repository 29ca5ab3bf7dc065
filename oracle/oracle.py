"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker or the timed CPU
reference, never as part of the product path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

CUBE_DT = np.dtype([("hi", "<u4"), ("lo", "<u4"), ("r_c", "<f4"), ("depth", "<i4")])


class OView(C.Structure):
    _fields_ = [("depths", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("m", C.c_float * 12), ("lidar", C.c_int32)]


class OParams(C.Structure):
    _fields_ = [("alpha0", C.c_float), ("alpha1", C.c_float), ("lambda_", C.c_float),
                ("iters", C.c_int32), ("op_mode", C.c_int32), ("chunk", C.c_int32),
                ("max_iters", C.c_int32), ("gap_tol", C.c_double), ("threads", C.c_int32),
                ("warm_fallback", C.c_int32), ("dual_eps", C.c_float), ("warm_dual", C.c_int32),
                ("sigma", C.c_float), ("tau", C.c_float), ("sigma_depth", C.c_float * 22),
                ("tau_depth", C.c_float * 22)]


_L = None


def lib():
    global _L
    if _L is None:
        # ORACLE_LIB: another build of the same source (the tests that check
        # the pins against deliberately altered copies of the oracle use it)
        p = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
        if not os.path.exists(p):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(p)
        vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int32
        L.orc_morton_encode.argtypes = [C.c_uint32] * 3 + [C.POINTER(C.c_uint32)] * 2
        L.orc_morton_decode.argtypes = [C.c_uint32, C.c_uint32] + [C.POINTER(C.c_uint32)] * 3
        L.orc_parent_code.argtypes = [u64, i32]
        L.orc_parent_code.restype = u64
        L.orc_face_neighbor.argtypes = [u64, i32, i32]
        L.orc_face_neighbor.restype = u64
        L.orc_median_indicator.argtypes = [vp, i32]
        L.orc_median_indicator.restype = C.c_float
        L.orc_median_indicator_f64.argtypes = [vp, i32]
        L.orc_median_indicator_f64.restype = C.c_double
        L.orc_mean_indicator.argtypes = [vp, i32]
        L.orc_mean_indicator.restype = C.c_float
        L.orc_mip_build.argtypes = [C.POINTER(OView), vp, vp]
        L.orc_project.argtypes = [C.POINTER(OView), vp, C.c_float, vp, C.POINTER(C.c_float)]
        L.orc_histograms.argtypes = [vp, u64, vp, i32, vp, C.c_float, i32, vp]
        L.orc_tgv_minimize.argtypes = [vp, u64, vp, i32, C.POINTER(OParams), i32, vp, vp, vp, vp, vp, i32]
        L.orc_tgv_minimize_oc.argtypes = [vp, u64, vp, i32, vp, vp, u64, C.POINTER(OParams), i32, vp]
        L.orc_tgv_minimize_streaming.argtypes = [vp, u64, vp, i32, vp, vp, u64, C.POINTER(OParams), i32, vp]
        L.orc_level_gap.argtypes = [vp, u64, vp, i32, C.POINTER(OParams), i32, vp, vp,
                                    C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_neighbour_table.argtypes = [vp, u64, i32, i32, vp, vp]
        L.orc_neighbour_table.restype = u64
        L.orc_extract_surface.argtypes = [vp, u64, vp, vp, vp, C.c_float, C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(u64)]
        L.orc_extract_surface.restype = u64
        L.orc_free.argtypes = [vp]
        L.orc_free.restype = None
        L.orc_time_level.argtypes = [vp, u64, vp, i32, C.POINTER(OParams), i32, i32, C.POINTER(u64)]
        L.orc_time_level.restype = C.c_double
        L.orc_log2_level.argtypes = [C.c_float]
        L.orc_log2_rule_mismatches.argtypes = [C.c_float, C.c_float, vp, i32]
        L.orc_log2_rule_mismatches.restype = u64
        _L = L
    return _L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def params(alpha0=0.02, alpha1=0.04, lam=1.0, iters=200, op_mode=0, chunk=50, max_iters=20000, gap_tol=0.0,
           threads=1, warm_fallback=1, dual_eps=0.0, warm_dual=0, sigma=0.0, tau=0.0, sigma_depth=None,
           tau_depth=None):
    """sigma_depth / tau_depth: per octree depth (22 entries, > 0 replaces sigma / tau at that depth)."""
    sd = (C.c_float * 22)(*(sigma_depth or [0.0] * 22))
    td = (C.c_float * 22)(*(tau_depth or [0.0] * 22))
    return OParams(alpha0, alpha1, lam, iters, op_mode, chunk, max_iters, gap_tol, threads, warm_fallback, dual_eps,
                   warm_dual, sigma, tau, sd, td)


def view(depths, fx, fy, cx, cy, m, lidar=False):
    d = np.ascontiguousarray(depths, dtype=np.float32)
    v = OView()
    v.depths = d.ctypes.data
    v.height, v.width = d.shape
    v.fx, v.fy, v.cx, v.cy = fx, fy, cx, cy
    mm = np.asarray(m, dtype=np.float32).reshape(12)
    for i in range(12):
        v.m[i] = float(mm[i])
    v.lidar = int(bool(lidar))
    return v, d


def views_from_scene(scene):
    out, keep = [], []
    for k in range(len(scene.depths)):
        c = scene.cams[k]
        v, d = view(scene.depths[k], c.fx, c.fy, c.cx, c.cy, list(c.world_to_cam))
        out.append(v)
        keep.append(d)
    return out, keep


def morton_encode(x, y, z):
    hi, lo = C.c_uint32(), C.c_uint32()
    lib().orc_morton_encode(x, y, z, C.byref(hi), C.byref(lo))
    return (hi.value << 32) | lo.value


def morton_decode(code):
    x, y, z = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().orc_morton_decode(code >> 32, code & 0xffffffff, C.byref(x), C.byref(y), C.byref(z))
    return x.value, y.value, z.value


def parent_code(code, depth):
    return lib().orc_parent_code(code, depth)


def face_neighbor(code, face, depth):
    return lib().orc_face_neighbor(code, face, depth)


def median_indicator(bins, f64=False):
    b = np.ascontiguousarray(bins, dtype=np.uint32)
    return (lib().orc_median_indicator_f64 if f64 else lib().orc_median_indicator)(_p(b), len(b))


def mean_indicator(bins):
    b = np.ascontiguousarray(bins, dtype=np.uint32)
    return lib().orc_mean_indicator(_p(b), len(b))


def mip_build(v):
    out = np.zeros(v.width * v.height * 4 // 3 + 64, dtype=np.float32)
    dims = np.zeros(32, dtype=np.int32)
    nl = lib().orc_mip_build(C.byref(v), _p(out), _p(dims))
    levels, off = [], 0
    for l in range(nl):
        w, h = int(dims[2 * l]), int(dims[2 * l + 1])
        levels.append(out[off:off + w * h].reshape(h, w).copy())
        off += w * h
    return levels


def project(v, P, r_c):
    P = np.asarray(P, dtype=np.float32)
    o = np.zeros(3, dtype=np.int32)
    cz = C.c_float()
    ok = lib().orc_project(C.byref(v), _p(P), r_c, _p(o), C.byref(cz))
    return (bool(ok), int(o[0]), int(o[1]), int(o[2]), cz.value)


def histograms(cubes, views, origin, r_root, nb):
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    arr = (OView * max(1, len(views)))(*views)
    o = np.asarray(origin, dtype=np.float32)
    out = np.zeros((len(cubes), nb), dtype=np.uint32)
    lib().orc_histograms(_p(cubes), len(cubes), arr, len(views), _p(o), r_root, nb, _p(out))
    return out


def tgv_minimize(cubes, hists, prm=None, f64=False, init_uv=None, with_bar=False):
    """Returns (state [n,13], bar [n,4] or None, level_iters, level (P, D))."""
    prm = prm or params()
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    dt = np.float64 if f64 else np.float32
    st = np.zeros((len(cubes), 13), dtype=dt)
    bar = np.zeros((len(cubes), 4), dtype=dt) if with_bar else None
    it = np.zeros(22, dtype=np.int32)
    pd = np.zeros(44, dtype=np.float64)
    init = None if init_uv is None else np.ascontiguousarray(init_uv, dtype=np.float32)
    nl = lib().orc_tgv_minimize(_p(cubes), len(cubes), _p(h), h.shape[1], C.byref(prm), int(f64), _p(init), _p(st),
                                _p(bar), _p(it), _p(pd), 22)
    return st, bar, it[:nl].copy(), pd[:2 * nl].reshape(-1, 2).copy()


def tgv_minimize_oc(cubes, hists, leaves, prm=None, f64=False):
    prm = prm or params()
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    first = np.ascontiguousarray(leaves["first_idx"], dtype=np.uint32)
    last = np.ascontiguousarray(leaves["last_idx"], dtype=np.uint32)
    st = np.zeros((len(cubes), 13), dtype=np.float64 if f64 else np.float32)
    lib().orc_tgv_minimize_oc(_p(cubes), len(cubes), _p(h), h.shape[1], _p(first), _p(last), len(first),
                              C.byref(prm), int(f64), _p(st))
    return st


def level_gap(cubes, hists, prm, level, state, bar):
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    s = np.ascontiguousarray(state, dtype=np.float64)
    b = np.ascontiguousarray(bar, dtype=np.float64)
    P, D = C.c_double(), C.c_double()
    lib().orc_level_gap(_p(cubes), len(cubes), _p(h), h.shape[1], C.byref(prm), level, _p(s), _p(b), C.byref(P),
                        C.byref(D))
    return P.value, D.value


def neighbour_table(cubes, level, mode=0):
    """mode 0: the reference's table; mode 1: the depth-aware table of the exact-adjoint mode."""
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    nb = np.zeros((len(cubes), 6), dtype=np.int32)
    act = np.zeros(len(cubes), dtype=np.uint32)
    na = lib().orc_neighbour_table(_p(cubes), len(cubes), level, mode, _p(nb), _p(act))
    return act[:na].copy(), nb[:na].copy()


def tgv_minimize_streaming(cubes, hists, leaves, prm=None, f64=False):
    """tgv_minimize_streaming on arrays: state [n, 13]."""
    prm = prm or params()
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    first = np.ascontiguousarray(leaves["first_idx"], dtype=np.uint32)
    last = np.ascontiguousarray(leaves["last_idx"], dtype=np.uint32)
    st = np.zeros((len(cubes), 13), dtype=np.float64 if f64 else np.float32)
    lib().orc_tgv_minimize_streaming(_p(cubes), len(cubes), _p(h), h.shape[1], _p(first), _p(last), len(first),
                                     C.byref(prm), int(f64), _p(st))
    return st


def extract_surface(cubes, u, origin, r_root, has_data=None):
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    uu = np.ascontiguousarray(u, dtype=np.float32)
    hd = None if has_data is None else np.ascontiguousarray(has_data, dtype=np.uint8)
    o = np.asarray(origin, dtype=np.float32)
    vp, fp = C.c_void_p(), C.c_void_p()
    nf = C.c_uint64()
    nv = lib().orc_extract_surface(_p(cubes), len(cubes), _p(uu), _p(hd), _p(o), r_root, C.byref(vp), C.byref(fp),
                                   C.byref(nf))
    verts = np.ctypeslib.as_array(C.cast(vp, C.POINTER(C.c_float)), shape=(max(1, nv) * 3,))[:nv * 3].copy()
    faces = np.ctypeslib.as_array(C.cast(fp, C.POINTER(C.c_uint32)), shape=(max(1, nf.value) * 3,))[
        :nf.value * 3].copy()
    lib().orc_free(vp)
    lib().orc_free(fp)
    return verts.reshape(-1, 3), faces.reshape(-1, 3)


def log2_rule_mismatches(lo, hi, thr):
    t = np.ascontiguousarray(thr, dtype=np.float32)
    return lib().orc_log2_rule_mismatches(lo, hi, _p(t), len(t))


def host_cpu():
    """(model name, has the AVX-512 subsets of x86-64-v4) from /proc/cpuinfo."""
    model, flags = "unknown", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "unknown":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("flags") and not flags:
                    flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags


_T = {}


def timing_lib():
    """(library, -march) for timing on this host: the x86-64-v4 build where the
    host has AVX-512, else the x86-64-v3 one (same source, same results)."""
    if "lib" not in _T:
        v4 = os.path.join(HERE, "liboracle_v4.so")
        if host_cpu()[1] and os.path.exists(v4):
            L, march = C.CDLL(v4), "x86-64-v4"
        else:
            L, march = lib(), "x86-64-v3"
        L.orc_time_level.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32, C.POINTER(OParams), C.c_int32,
                                     C.c_int32, C.POINTER(C.c_uint64)]
        L.orc_time_level.restype = C.c_double
        _T["lib"], _T["march"] = L, march
    return _T["lib"], _T["march"]


def time_level(cubes, hists, prm, depth, iters):
    """Seconds for `iters` iterations of one level of the CPU restatement (the
    reference's AoS loops; prm.op_mode 0 = its own update rules), and the
    level's cube count."""
    cubes = np.ascontiguousarray(cubes, dtype=CUBE_DT)
    h = np.ascontiguousarray(hists, dtype=np.uint32)
    na = C.c_uint64()
    L, _ = timing_lib()
    t = L.orc_time_level(_p(cubes), len(cubes), _p(h), h.shape[1], C.byref(prm), depth, iters, C.byref(na))
    return t, na.value


# ---------------------------------------------------------------- warm-start network
MLP_NP = 6404


def _mlp_lib():
    L = lib()
    if not hasattr(L, "_mlp_ready"):
        vp = C.c_void_p
        L.orc_mlp_forward.argtypes = [vp, vp, C.c_int32, vp]
        L.orc_mlp_loss_grad.argtypes = [vp, vp, vp, C.c_int32, vp]
        L.orc_mlp_loss_grad.restype = C.c_double
        L.orc_mlp_adam.argtypes = [vp, vp, vp, vp, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double]
        L._mlp_ready = True
    return L


def mlp_forward(w, x):
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros((len(x), 4))
    _mlp_lib().orc_mlp_forward(_p(w), _p(x), len(x), _p(y))
    return y


def mlp_loss_grad(w, x, t):
    w = np.ascontiguousarray(w, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    g = np.zeros(MLP_NP)
    loss = _mlp_lib().orc_mlp_loss_grad(_p(w), _p(x), _p(t), len(x), _p(g))
    return loss, g


def mlp_adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    for a in (w, m, v):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    g = np.ascontiguousarray(g, dtype=np.float64)
    _mlp_lib().orc_mlp_adam(_p(w), _p(m), _p(v), _p(g), step, lr, b1, b2, eps)
