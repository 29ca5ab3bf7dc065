"""Builds the native libraries in-tree.

libtgvx.so        HIP kernels + C ABI (gfx950), the product
libtgvx_scene.so  host-only synthetic-scene / octree-input generator
oracle/liboracle.so  CPU oracle (test infrastructure; built by oracle/Makefile)
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -ffp-contract=off: no fused multiply-add anywhere, so every expression rounds
# as written (bit parity with the CPU definition). The packed-FP32 feature is
# switched off for the device code. Division and square root stay correctly
# rounded (HIP's default, stated explicitly).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + INCLUDE, "-I" + CSRC,
]
HIP_SOURCES = ["api.hip", "hist.hip", "solve.hip", "mlp.hip", "surface.hip", "octree.hip", "xfer.hip"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build_tgvx(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]

    def one(i):
        _run([HIPCC] + HIP_FLAGS + ["-c", "-o", objs[i], os.path.join(CSRC, srcs[i])])

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(one, range(len(srcs))))
    out = os.path.join(PKG, "libtgvx.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
    return out


def build_scene():
    out = os.path.join(PKG, "libtgvx_scene.so")
    # -ffp-contract=off: every float expression rounds as written, so the
    # device octree build reproduces the generator's cubes bit for bit
    _run(["g++", "-std=c++17", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared", "-fopenmp",
          "-I" + INCLUDE, "-o", out, os.path.join(CSRC, "scene.cpp")])
    return out


def build_host():
    """libtgvx_host.so: the reference's C++ interface over the C ABI, and the
    C++ drop-in test program linked against it."""
    host = os.path.join(PKG, "host")
    out = os.path.join(PKG, "libtgvx_host.so")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-I" + INCLUDE, "-I" + host, "-o", out,
          os.path.join(host, "tgvx_host.cpp"), "-L" + PKG, "-ltgvx", "-Wl,-rpath,$ORIGIN"])
    test = os.path.join(ROOT, "tests", "cpp", "test_dropin")
    _run(["g++", "-std=c++17", "-O2", "-I" + INCLUDE, "-I" + host, "-o", test,
          os.path.join(ROOT, "tests", "cpp", "test_dropin.cpp"), "-L" + PKG, "-ltgvx_host", "-ltgvx",
          "-Wl,-rpath," + PKG])
    return out


def build_oracle():
    _run(["make", "-s", "-B", "-j2", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all():
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_tgvx), ex.submit(build_scene), ex.submit(build_oracle)]
        out = [f.result() for f in futs]
    out.append(build_host())
    return out


if __name__ == "__main__":
    for p in build_all():
        print(p)
