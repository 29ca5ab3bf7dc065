"""Seeded synthetic scenes and octree inputs (wraps libtgvx_scene.so).

Scene choices not fixed by the workload descriptions are made here and
recorded in Scene.meta: root cube [-1.5, 1.5]^3 (r_root 1.5), 60 degree
horizontal field of view, outliers uniform in [0.5, 6] along the pixel ray.
"""
import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

SPHERE, BOX, TORUS = 0, 1, 2
R_ROOT = 1.5
ORIGIN = (-1.5, -1.5, -1.5)
HFOV = 60.0


def rot_axis_angle(axis, deg):
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    t = math.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + math.sin(t) * K + (1 - math.cos(t)) * (K @ K)


def prim(kind, center, size, rot=None):
    p = N.Prim()
    p.kind = kind
    for i in range(3):
        p.center[i] = float(center[i])
        p.size[i] = float(size[i]) if i < len(size) else 0.0
    R = np.eye(3) if rot is None else np.asarray(rot, dtype=np.float64)
    for i in range(9):
        p.rot[i] = float(R.reshape(9)[i])
    return p


@dataclass
class Scene:
    depths: np.ndarray          # float32 [V, H, W]
    cams: object                # ctypes array of Camera
    cubes: np.ndarray           # CUBE_DT, sorted by (code, depth)
    origin: tuple = ORIGIN
    r_root: float = R_ROOT
    meta: dict = field(default_factory=dict)

    def views(self):
        """tgvx View structs (keep self alive while they are used)."""
        out = []
        for k in range(len(self.depths)):
            c = self.cams[k]
            v = N.View()
            v.depths = self.depths[k].ctypes.data
            v.height, v.width = self.depths[k].shape
            v.fx, v.fy, v.cx, v.cy = c.fx, c.fy, c.cx, c.cy
            for i in range(12):
                v.world_to_cam[i] = c.world_to_cam[i]
            v.lidar = 0
            out.append(v)
        return out


def cameras_ring(n, radius, W, H, hfov=HFOV):
    cams = (N.Camera * n)()
    if N.scene_lib().tgvs_ring_cameras(n, radius, W, H, hfov, cams) != 0:
        raise RuntimeError("tgvs_ring_cameras failed")
    return cams


def cameras_random(n, radius, W, H, seed, hfov=HFOV):
    cams = (N.Camera * n)()
    if N.scene_lib().tgvs_random_cameras(n, radius, W, H, hfov, seed, cams) != 0:
        raise RuntimeError("tgvs_random_cameras failed")
    return cams


def render(prims, cams, W, H, sigma, outlier_frac, seed, lo=0.5, hi=6.0):
    pa = (N.Prim * max(1, len(prims)))(*prims)
    out = np.empty((len(cams), H, W), dtype=np.float32)
    if N.scene_lib().tgvs_render(pa, len(prims), cams, len(cams), W, H, sigma, outlier_frac, lo, hi, seed,
                                 N.ptr(out)) != 0:
        raise RuntimeError("tgvs_render failed")
    return out


def build_octree(depths, cams, max_depth, origin=ORIGIN, r_root=R_ROOT, last_stage=3):
    S = N.scene_lib()
    o = np.asarray(origin, dtype=np.float32)
    p = C.c_void_p()
    V, H, W = depths.shape
    n = S.tgvs_build_octree(N.ptr(depths), cams, V, W, H, N.ptr(o), r_root, max_depth, last_stage, C.byref(p))
    if n < 0:
        raise RuntimeError("tgvs_build_octree failed")
    cubes = np.empty(n, dtype=N.CUBE_DT)
    if n:
        C.memmove(cubes.ctypes.data, p, n * N.CUBE_DT.itemsize)
    S.tgvs_free(p)
    return cubes


def merge(cubes):
    S = N.scene_lib()
    c = np.ascontiguousarray(cubes, dtype=N.CUBE_DT)
    p = C.c_void_p()
    n = S.tgvs_merge(N.ptr(c), len(c), C.byref(p))
    if n < 0:
        raise RuntimeError("tgvs_merge failed")
    out = np.empty(n, dtype=N.CUBE_DT)
    if n:
        C.memmove(out.ctypes.data, p, n * N.CUBE_DT.itemsize)
    S.tgvs_free(p)
    return out


def balance(cubes, r_root=R_ROOT):
    S = N.scene_lib()
    c = np.ascontiguousarray(cubes, dtype=N.CUBE_DT)
    p = C.c_void_p()
    n = S.tgvs_balance(N.ptr(c), len(c), r_root, C.byref(p))
    out = np.empty(n, dtype=N.CUBE_DT)
    if n:
        C.memmove(out.ctypes.data, p, n * N.CUBE_DT.itemsize)
    S.tgvs_free(p)
    return out


def treetop(cubes, max_per_leaf):
    S = N.scene_lib()
    c = np.ascontiguousarray(cubes, dtype=N.CUBE_DT)
    p = C.c_void_p()
    n = S.tgvs_treetop(N.ptr(c), len(c), max_per_leaf, C.byref(p))
    out = np.empty(n, dtype=N.LEAF_DT)
    if n:
        C.memmove(out.ctypes.data, p, n * N.LEAF_DT.itemsize)
    S.tgvs_free(p)
    return out


def make_scene(prims, cams, W, H, sigma, outlier_frac, seed, max_depth, **meta):
    depths = render(prims, cams, W, H, sigma, outlier_frac, seed)
    cubes = build_octree(depths, cams, max_depth)
    m = dict(W=W, H=H, views=len(cams), sigma=sigma, outliers=outlier_frac, seed=seed, max_depth=max_depth,
             r_root=R_ROOT, origin=ORIGIN, hfov=HFOV)
    m.update(meta)
    return Scene(depths, cams, cubes, meta=m)


def sphere_scene(views=24, W=320, H=240, sigma=0.005, outliers=0.02, seed=0, max_depth=7):
    """Unit sphere, cameras on the ring of radius 3 (the CPU-sized workload)."""
    cams = cameras_ring(views, 3.0, W, H)
    return make_scene([prim(SPHERE, (0, 0, 0), (1.0,))], cams, W, H, sigma, outliers, seed, max_depth,
                      name="sphere")


def torus_box_prims():
    """Torus (R=1, r=0.3) union a box of half extent 0.45 rotated 30 degrees about (1,1,0)."""
    return [prim(TORUS, (0, 0, 0), (1.0, 0.3)),
            prim(BOX, (0, 0, 0), (0.45, 0.45, 0.45), rot_axis_angle((1, 1, 0), 30.0))]


def torus_box_scene(views=96, W=1280, H=960, sigma=0.003, outliers=0.03, seed=1, max_depth=10):
    """The depth-10 workload: torus union rotated box, seeded random poses on the radius-3 sphere."""
    cams = cameras_random(views, 3.0, W, H, seed)
    return make_scene(torus_box_prims(), cams, W, H, sigma, outliers, seed, max_depth, name="torus_box")


def csg_prims(seed, scale=1.0):
    """Seeded union of 1-6 spheres and boxes inside the root cube (scaled about the origin)."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 7))
    out = []
    for _ in range(k):
        c = rng.uniform(-0.6, 0.6, 3) * scale
        if rng.random() < 0.5:
            out.append(prim(SPHERE, c, (float(rng.uniform(0.2, 0.55)) * scale,)))
        else:
            ax = rng.normal(size=3)
            out.append(prim(BOX, c, tuple(rng.uniform(0.15, 0.45, 3) * scale),
                            rot_axis_angle(ax, float(rng.uniform(0, 90)))))
    return out


def csg_scene(seed, views=24, W=320, H=240, sigma=0.005, outliers=0.02, max_depth=9):
    cams = cameras_ring(views, 3.0, W, H)
    return make_scene(csg_prims(seed), cams, W, H, sigma, outliers, seed, max_depth, name=f"csg{seed}")


def batch_block(seed, target, views=12, sigma=0.005, outliers=0.02):
    """One block of the batched workload: a seeded union of 1-6 spheres and
    boxes whose octree has about `target` cubes. The depth clamp is the
    smallest whose estimated count (the count at clamp 7 times 4 per extra
    depth) reaches the target; the primitives are then scaled by
    sqrt(target / estimate), so the surface, and with it the count, shrinks
    to the target, corrected from the measured count where leaves reach the
    pixel footprint. Images are 640x480, or 1280x960 for clamps above 8."""
    def make(scale, depth):
        W, H = (1280, 960) if depth > 8 else (640, 480)
        cams = cameras_ring(views, 3.0, W, H)
        return make_scene(csg_prims(seed, scale), cams, W, H, sigma, outliers, seed, depth, name=f"block{seed}",
                          target=int(target), scale=scale)

    probe = make(1.0, 7)
    c7 = max(len(probe.cubes), 1)
    depth, est = 10, c7 * 4.0 ** 3
    for d in range(4, 11):
        if c7 * 4.0 ** (d - 7) >= target:
            depth, est = d, c7 * 4.0 ** (d - 7)
            break
    scale = float(np.clip(np.sqrt(target / est), 0.3, 1.0))
    sc = probe if (depth == 7 and scale == 1.0) else make(scale, depth)
    # counts stop growing fourfold once leaves approach the pixel footprint:
    # correct the scale (then the clamp) from the measured count, twice at most
    for _ in range(2):
        n = len(sc.cubes)
        if n >= 0.8 * target or (scale >= 1.0 and depth >= 10):
            break
        if scale < 1.0:
            scale = float(min(1.0, scale * np.sqrt(target / max(n, 1))))
        else:
            depth += 1
        sc = make(scale, depth)
    return sc


def batch_targets(n_blocks=256, seed=4, lo=4096, hi=524288):
    """Log-uniform target cube counts of the batched workload's blocks."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n_blocks))
