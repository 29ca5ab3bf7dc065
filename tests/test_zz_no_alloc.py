"""tgvx_histograms_noalloc through the C ABI: device inputs and output, no
allocation inside the call, results equal to tgvx_histograms bit for bit on
two calls in a row (the second smaller), and oversized inputs refused."""
import ctypes as C

import numpy as np
import pytest

from tgvreconclaude_amd import _native as N
from tgvreconclaude_amd import scene


class Hip:
    """The few runtime calls the test needs to own device buffers."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), C.c_size_t(max(1, nbytes))) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def fill(self, p, byte, nbytes):
        assert self.L.hipMemset(p, C.c_int(byte), C.c_size_t(nbytes)) == 0

    def free_all(self):
        self.L.hipDeviceSynchronize()
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


def device_views(hip, sc, nv):
    out = []
    for k, v in enumerate(sc.views()[:nv]):
        d = N.View()
        C.pointer(d)[0] = v
        d.depths = hip.upload(sc.depths[k].astype(np.float32)).value
        out.append(d)
    return (N.View * max(1, nv))(*out)


@pytest.mark.gpu
def test_noalloc_histograms_equal_the_allocating_call(ctx):
    sc = scene.sphere_scene()  # 24 ring views of 320 x 240, depth <= 7
    L = N.lib()
    hip = Hip()
    origin = np.asarray(sc.origin, dtype=np.float32)
    try:
        n_all = len(sc.cubes)
        # first call: the whole workload; second: fewer cubes, views and bins
        for nv, n, nb in ((24, n_all, 10), (9, n_all // 3, 8)):
            cubes = np.ascontiguousarray(sc.cubes[:n])
            want = ctx.histograms(cubes, sc.views()[:nv], sc.origin, sc.r_root, nb)
            assert want.sum() > 0
            d_cubes = hip.upload(cubes)
            dviews = device_views(hip, sc, nv)
            d_bins = hip.alloc(n * nb * 4)
            hip.fill(d_bins, 0xA5, n * nb * 4)  # every count must be written
            hip.L.hipDeviceSynchronize()
            print(f"noalloc call: {n} cubes, {nv} views, {nb} bins", flush=True)
            rc = L.tgvx_histograms_noalloc(ctx.h, d_cubes, n, dviews, nv, N.ptr(origin), float(sc.r_root), nb,
                                           d_bins)
            assert rc == 0, L.tgvx_last_error(ctx.h).decode()
            got = hip.download(d_bins, (n, nb), np.uint32)
            assert np.array_equal(got, want)
    finally:
        hip.free_all()


@pytest.mark.gpu
def test_noalloc_refuses_what_does_not_fit(ctx):
    L = N.lib()
    hip = Hip()
    origin = np.zeros(3, np.float32)
    try:
        cubes = np.zeros(4, N.CUBE_DT)
        d_cubes = hip.upload(cubes)
        d_bins = hip.alloc(4 * 8 * 4)
        base = hip.alloc(1280 * 960 * 4)
        hip.fill(base, 0, 1280 * 960 * 4)

        def views(k):
            vs = (N.View * k)()
            for v in vs:
                v.depths, v.width, v.height, v.fx, v.fy, v.cx, v.cy = base.value, 1280, 960, 1000.0, 1000.0, 640.0, 480.0
                v.world_to_cam[0] = v.world_to_cam[5] = v.world_to_cam[10] = 1.0
            return vs

        # more views than the view table holds
        assert L.tgvx_histograms_noalloc(ctx.h, d_cubes, 4, views(1025), 1025, N.ptr(origin), 1.0, 8, d_bins) == 1
        assert b"at most" in L.tgvx_last_error(ctx.h)
        # 64 views of 1280 x 960: their pyramid levels exceed the scratch array
        assert L.tgvx_histograms_noalloc(ctx.h, d_cubes, 4, views(64), 64, N.ptr(origin), 1.0, 8, d_bins) == 1
        assert b"pyramid" in L.tgvx_last_error(ctx.h)
        # host memory where device memory is required
        assert L.tgvx_histograms_noalloc(ctx.h, N.ptr(cubes), 4, views(1), 1, N.ptr(origin), 1.0, 8, d_bins) == 1
        assert b"cubes is not" in L.tgvx_last_error(ctx.h)
        # within the limits the same call works
        assert L.tgvx_histograms_noalloc(ctx.h, d_cubes, 4, views(1), 1, N.ptr(origin), 1.0, 8, d_bins) == 0
    finally:
        hip.free_all()
