"""ctypes bindings of libtgvx.so (include/tgvx.h) and libtgvx_scene.so (include/tgvx_scene.h).

The structs mirror the C headers byte for byte; the numpy dtypes mirror the
reference PODs (OctreeCube 16 B, TGVCube 52 B, TreetopLeaf 20 B).
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))

CUBE_DT = np.dtype([("hi", "<u4"), ("lo", "<u4"), ("r_c", "<f4"), ("depth", "<i4")])
STATE_DT = np.dtype([("u", "<f4"), ("v", "<f4", 3), ("p", "<f4", 3), ("q", "<f4", 6)])
LEAF_DT = np.dtype([("hi", "<u4"), ("lo", "<u4"), ("depth", "<i4"), ("first_idx", "<u4"), ("last_idx", "<u4")])
assert CUBE_DT.itemsize == 16 and STATE_DT.itemsize == 52 and LEAF_DT.itemsize == 20

FP64, FP32, BF16, BF16_DUAL = 0, 1, 2, 3
OP_REF, OP_ADJ = 0, 1
MAX_LEVELS = 22


class View(C.Structure):
    _fields_ = [("depths", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("world_to_cam", C.c_float * 12), ("lidar", C.c_int32)]


class Params(C.Structure):
    _fields_ = [("alpha0", C.c_float), ("alpha1", C.c_float), ("lambda_", C.c_float),
                ("iters", C.c_int32), ("precision", C.c_int32), ("op_mode", C.c_int32),
                ("chunk", C.c_int32), ("max_iters", C.c_int32), ("gap_tol", C.c_double),
                ("warm_fallback", C.c_int32), ("use_graph", C.c_int32), ("block", C.c_int32),
                ("dual_eps", C.c_float), ("warm_start", C.c_int32), ("warm_dual", C.c_int32),
                ("fused", C.c_int32), ("tune", C.c_int32),
                ("sigma", C.c_float), ("tau", C.c_float),
                ("sigma_depth", C.c_float * 22), ("tau_depth", C.c_float * 22), ("no_energy", C.c_int32)]


class Report(C.Structure):
    _fields_ = [("n_levels", C.c_int32),
                ("level_depth", C.c_int32 * MAX_LEVELS), ("level_leaves", C.c_uint32 * MAX_LEVELS),
                ("level_iters", C.c_int32 * MAX_LEVELS), ("level_warm", C.c_int32 * MAX_LEVELS),
                ("level_energy", C.c_double * MAX_LEVELS), ("level_dual", C.c_double * MAX_LEVELS),
                ("leaf_iters", C.c_double), ("solve_ms", C.c_double), ("iterate_ms", C.c_double),
                ("gap_ms", C.c_double), ("setup_ms", C.c_double), ("level_ms", C.c_double * MAX_LEVELS)]

    def levels(self):
        out = []
        for i in range(self.n_levels):
            P, D = self.level_energy[i], self.level_dual[i]
            out.append(dict(depth=self.level_depth[i], leaves=self.level_leaves[i],
                            iters=self.level_iters[i], warm=self.level_warm[i], P=P, D=D,
                            rel_gap=(P - D) / max(1.0, abs(P)), ms=self.level_ms[i]))
        return out


ML = MAX_LEVELS


class TuneReport(C.Structure):
    _fields_ = [("n_levels", C.c_int32), ("level_depth", C.c_int32 * ML), ("level_leaves", C.c_uint32 * ML),
                ("variant", C.c_int32 * ML), ("block", C.c_int32 * ML), ("xcd", C.c_int32 * ML),
                ("best_us", C.c_double * ML), ("default_us", C.c_double * ML)]

    def levels(self):
        return [{"depth": self.level_depth[i], "leaves": self.level_leaves[i], "variant": self.variant[i],
                 "block": self.block[i], "xcd": self.xcd[i], "best_us": self.best_us[i],
                 "default_us": self.default_us[i]} for i in range(self.n_levels)]


class OocReport(C.Structure):
    _fields_ = [("batches", C.c_int32), ("slot_cubes", C.c_uint64), ("device_bytes", C.c_uint64),
                ("index_bytes", C.c_uint64), ("moved_bytes", C.c_uint64), ("tables_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Block(C.Structure):
    _fields_ = [("cubes", C.c_void_p), ("n", C.c_uint64), ("bins", C.c_void_p)]


class Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("center", C.c_float * 3), ("size", C.c_float * 3), ("rot", C.c_float * 9)]


class Camera(C.Structure):
    _fields_ = [("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("cam_to_world", C.c_float * 12), ("world_to_cam", C.c_float * 12)]


class DepthView(C.Structure):
    _fields_ = [("depths", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_float), ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("cam_to_world", C.c_float * 12)]


class OctreeReport(C.Structure):
    _fields_ = [("pixels", C.c_uint64), ("raw", C.c_uint64), ("merged", C.c_uint64), ("bridges", C.c_uint64),
                ("rounds", C.c_int32), ("upload_ms", C.c_double), ("cubes_ms", C.c_double),
                ("merge_ms", C.c_double), ("balance_ms", C.c_double), ("total_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# exported symbols of each library, as declared in include/*.h
TGVX_SYMBOLS = [
    "tgvx_default_params", "tgvx_create", "tgvx_destroy", "tgvx_last_error", "tgvx_device_info",
    "tgvx_histograms", "tgvx_mip_build", "tgvx_median_init", "tgvx_solve", "tgvx_solve_oc",
    "tgvx_problem_load", "tgvx_problem_solve", "tgvx_problem_state", "tgvx_problem_gap",
    "tgvx_problem_levels", "tgvx_level_bench", "tgvx_mlp_set", "tgvx_mlp_get", "tgvx_mlp_forward",
    "tgvx_mlp_loss_grad", "tgvx_mlp_train", "tgvx_problem_features", "tgvx_problem_load_blocks",
    "tgvx_problem_block_report", "tgvx_solve_batched", "tgvx_autotune", "tgvx_problem_solve_levels",
    "tgvx_surface_extract", "tgvx_surface_copy", "tgvx_solve_streaming", "tgvx_octree_build",
    "tgvx_octree_balance", "tgvx_octree_copy", "tgvx_octree_merge", "tgvx_tile_sets", "tgvx_histograms_noalloc",
    "tgvx_problem_solve_from", "tgvx_solve_streaming_ooc", "tgvx_grid_barrier_bench", "tgvx_tile_sets4",
]
MLP_NPARAMS = 6404
SCENE_SYMBOLS = ["tgvs_ring_cameras", "tgvs_random_cameras", "tgvs_render", "tgvs_build_octree",
                 "tgvs_merge", "tgvs_balance", "tgvs_treetop", "tgvs_free"]

_lib = None
_scene = None


def _path(name):
    p = os.path.join(PKG, name)
    if not os.path.exists(p):
        raise RuntimeError(f"{name} is not built: run __graft_entry__.build() (or python tgvreconclaude_amd/build.py)")
    return p


def lib():
    """libtgvx.so (loaded once). Raises if it has not been built."""
    global _lib
    if _lib is None:
        L = C.CDLL(_path("libtgvx.so"))
        vp, u64, i32 = C.c_void_p, C.c_uint64, C.c_int32
        L.tgvx_default_params.argtypes = [C.POINTER(Params)]
        L.tgvx_default_params.restype = None
        L.tgvx_create.argtypes = [C.POINTER(vp), i32]
        L.tgvx_destroy.argtypes = [vp]
        L.tgvx_destroy.restype = None
        L.tgvx_last_error.argtypes = [vp]
        L.tgvx_last_error.restype = C.c_char_p
        L.tgvx_device_info.argtypes = [vp, C.c_char_p, i32, C.POINTER(i32)]
        L.tgvx_histograms.argtypes = [vp, vp, u64, C.POINTER(View), i32, vp, C.c_float, i32, vp]
        L.tgvx_histograms_noalloc.argtypes = [vp, vp, u64, C.POINTER(View), i32, vp, C.c_float, i32, vp]
        L.tgvx_mip_build.argtypes = [vp, C.POINTER(View), vp, vp, C.POINTER(i32)]
        L.tgvx_median_init.argtypes = [vp, vp, u64, i32, vp]
        L.tgvx_solve.argtypes = [vp, vp, u64, vp, i32, C.POINTER(Params), vp, vp, C.POINTER(Report)]
        L.tgvx_solve_oc.argtypes = [vp, vp, u64, vp, i32, vp, u64, C.POINTER(Params), vp, C.POINTER(Report)]
        L.tgvx_problem_load.argtypes = [vp, vp, u64, vp, i32, i32, i32]
        L.tgvx_problem_solve.argtypes = [vp, C.POINTER(Params), vp, C.POINTER(Report)]
        L.tgvx_problem_solve_from.argtypes = [vp, C.POINTER(Params), vp, C.POINTER(Report)]
        L.tgvx_problem_state.argtypes = [vp, vp, vp]
        L.tgvx_problem_gap.argtypes = [vp, i32, C.POINTER(Params), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.tgvx_problem_levels.argtypes = [vp, vp, vp, C.POINTER(i32)]
        L.tgvx_level_bench.argtypes = [vp, i32, C.POINTER(Params), i32, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.tgvx_tile_sets.argtypes = [vp, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_int64)]
        L.tgvx_tile_sets4.argtypes = [vp, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(C.c_int64)]
        L.tgvx_grid_barrier_bench.argtypes = [vp, i32, i32, i32, C.POINTER(C.c_double)]
        L.tgvx_mlp_set.argtypes = [vp, vp]
        L.tgvx_mlp_get.argtypes = [vp, vp]
        L.tgvx_mlp_forward.argtypes = [vp, vp, u64, vp]
        L.tgvx_mlp_loss_grad.argtypes = [vp, vp, vp, u64, vp, C.POINTER(C.c_double)]
        L.tgvx_mlp_train.argtypes = [vp, vp, vp, u64, i32, i32, C.c_float, C.c_uint64, vp]
        L.tgvx_problem_features.argtypes = [vp, i32, vp, vp, vp]
        L.tgvx_problem_load_blocks.argtypes = [vp, vp, C.c_uint32, i32, i32, i32]
        L.tgvx_autotune.argtypes = [vp, C.POINTER(Params), i32, C.POINTER(TuneReport)]
        L.tgvx_problem_solve_levels.argtypes = [vp, C.POINTER(Params), vp, i32, C.POINTER(Report)]
        L.tgvx_surface_extract.argtypes = [vp, vp, vp, vp, C.c_float, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.tgvx_surface_copy.argtypes = [vp, vp, vp]
        L.tgvx_solve_streaming.argtypes = [vp, vp, C.c_uint64, vp, i32, vp, C.c_uint64, C.POINTER(Params), vp,
                                           C.POINTER(Report)]
        L.tgvx_problem_block_report.argtypes = [vp, vp, vp]
        L.tgvx_solve_streaming_ooc.argtypes = [vp, vp, C.c_uint64, vp, i32, vp, C.c_uint64, C.POINTER(Params),
                                               C.c_uint64, vp, C.POINTER(Report), C.POINTER(OocReport)]
        L.tgvx_solve_batched.argtypes = [vp, vp, C.c_uint32, i32, C.POINTER(Params), vp, C.POINTER(Report)]
        L.tgvx_octree_build.argtypes = [vp, C.POINTER(DepthView), i32, vp, C.c_float, i32, i32, C.POINTER(u64),
                                        C.POINTER(OctreeReport)]
        L.tgvx_octree_merge.argtypes = [vp, vp, u64, C.POINTER(u64), C.POINTER(OctreeReport)]
        L.tgvx_octree_balance.argtypes = [vp, vp, u64, C.c_float, C.POINTER(u64), C.POINTER(OctreeReport)]
        L.tgvx_octree_copy.argtypes = [vp, vp]
        for name in TGVX_SYMBOLS:
            f = getattr(L, name)
            if f.restype is C.c_int:  # default
                f.restype = C.c_int
        _lib = L
    return _lib


def scene_lib():
    global _scene
    if _scene is None:
        L = C.CDLL(_path("libtgvx_scene.so"))
        vp, i32, f32 = C.c_void_p, C.c_int32, C.c_float
        L.tgvs_ring_cameras.argtypes = [i32, f32, i32, i32, f32, vp]
        L.tgvs_random_cameras.argtypes = [i32, f32, i32, i32, f32, C.c_uint64, vp]
        L.tgvs_render.argtypes = [vp, i32, vp, i32, i32, i32, f32, f32, f32, f32, C.c_uint64, vp]
        L.tgvs_build_octree.argtypes = [vp, vp, i32, i32, i32, vp, f32, i32, i32, C.POINTER(vp)]
        L.tgvs_merge.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
        L.tgvs_merge.restype = C.c_int64
        L.tgvs_build_octree.restype = C.c_int64
        L.tgvs_balance.argtypes = [vp, C.c_uint64, f32, C.POINTER(vp)]
        L.tgvs_balance.restype = C.c_int64
        L.tgvs_treetop.argtypes = [vp, C.c_uint64, C.c_uint32, C.POINTER(vp)]
        L.tgvs_treetop.restype = C.c_int64
        L.tgvs_free.argtypes = [vp]
        L.tgvs_free.restype = None
        _scene = L
    return _scene


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
