// Histogram voting on the device: depth pyramids and per-cube signed-distance
// histograms (reference src/histogram.cpp:35-163, the paper's Algorithm 1).
//
// One thread per cube walks every view; its nb counters live in LDS, one
// private column per thread (no atomics), and the block's counters leave LDS
// in one coalesced cube-major write. Counts are integers, so the result does
// not depend on order and is bit-exact with the CPU definition as long as each
// per-(cube, view) decision is: every float expression below is evaluated in
// the reference's order with no contraction (-ffp-contract=off), and the
// pyramid level uses an exact threshold form of trunc(log2f(x)).
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "tgvx_device.h"
#include "tgvx_internal.h"

namespace tgvx {

constexpr int kMaxLevels = 13;  // base + at most 12 (histogram.cpp:42)
constexpr int kMaxBins = 32;

struct ViewMeta {
    float m[12];
    float fx, fy, cx, cy;
    int32_t lidar, nlev, pad0, pad1;
    int32_t w[kMaxLevels], h[kMaxLevels];
    const float* lv[kMaxLevels];  // device address of each pyramid level
};

struct Log2Thr {
    float t[kMaxLevels];  // smallest x in [2^e, 2^(e+1)) whose trunc(log2f(x)) is e+1
};

// For e = 0..12, the smallest float in [2^e, 2^(e+1)) for which the host C
// library's log2f truncates to e+1 (2^(e+1) when none does). The device then
// computes trunc(log2f(x)) for x >= 1 as e + (x >= t[e]) with e the binary
// exponent, which reproduces the reference's (int)std::log2(float)
// (histogram.cpp:98-99) exactly, independent of any device log2 rounding.
static Log2Thr log2_thresholds() {
    Log2Thr r;
    for (int e = 0; e < kMaxLevels; ++e) {
        const float top = std::ldexp(1.0f, e + 1), bot = std::ldexp(1.0f, e);
        float t = top;
        for (float x = std::nextafter(top, 0.0f); x >= bot; x = std::nextafter(x, 0.0f)) {
            if ((int)std::log2(x) >= e + 1) t = x;
            else break;
        }
        r.t[e] = t;
    }
    return r;
}

// histogram.cpp:58-71: average of the finite, positive pixels of each 2x2
// block (du inner, dv outer), NaN when there are none.
__global__ void k_mip(const float* __restrict__ prev, int pw, int ph, float* __restrict__ next, int nw, int nh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw * nh) return;
    const int x = i % nw, y = i / nw;
    float sum = 0.0f;
    int cnt = 0;
    for (int dv = 0; dv < 2; ++dv)
        for (int du = 0; du < 2; ++du) {
            const int px = 2 * x + du, py = 2 * y + dv;
            if (px < pw && py < ph) {
                const float d = prev[(long)py * pw + px];
                if (isfinite(d) && d > 0.0f) {
                    sum += d;
                    ++cnt;
                }
            }
        }
    next[i] = cnt > 0 ? sum / (float)cnt : __int_as_float(0x7fc00000);
}

struct VoteArgs {
    const tgvx_cube* cubes;
    uint64_t n;
    const ViewMeta* views;
    int32_t nv, nb;
    float ox, oy, oz, r_root;
    Log2Thr thr;
    uint32_t* bins;
};

// compute_histograms (histogram.cpp:127-163) for one cube per thread.
__global__ void __launch_bounds__(256) k_vote(VoteArgs a) {
    extern __shared__ uint32_t cnt[];  // [nb][blockDim.x]
    const int T = blockDim.x, t = threadIdx.x;
    const uint64_t base = (uint64_t)xcd_tile() * T;
    const uint64_t ci = base + t;
    for (int k = 0; k < a.nb; ++k) cnt[k * T + t] = 0u;
    if (ci < a.n) {
        const tgvx_cube c = a.cubes[ci];
        const float r_c = c.r_c;
        if (r_c > 0.0f) {
            // cube_center (histogram.cpp:118-125)
            const uint64_t code = ((uint64_t)c.hi << 32) | c.lo;
            const int shift = 21 - c.depth;
            const uint32_t ix = compact3(code) >> shift, iy = compact3(code >> 1) >> shift,
                           iz = compact3(code >> 2) >> shift;
            const float cs = 2.0f * a.r_root / (float)(1u << c.depth);
            const float P0 = a.ox + (ix + 0.5f) * cs;
            const float P1 = a.oy + (iy + 0.5f) * cs;
            const float P2 = a.oz + (iz + 0.5f) * cs;
            const float delta = 6.0f * r_c;
            const float eta = 18.0f * r_c;
            const float scale = 0.5f * (float)a.nb;
            for (int vi = 0; vi < a.nv; ++vi) {
                const ViewMeta& V = a.views[vi];
                // DepthMipmap::project (histogram.cpp:76-106)
                const float camx = ((V.m[0] * P0 + V.m[1] * P1) + V.m[2] * P2) + V.m[3];
                const float camy = ((V.m[4] * P0 + V.m[5] * P1) + V.m[6] * P2) + V.m[7];
                const float z = ((V.m[8] * P0 + V.m[9] * P1) + V.m[10] * P2) + V.m[11];
                if (!(z > 0.0f)) continue;
                const float uf = camx / z * V.fx + V.cx;
                const float vf = camy / z * V.fy + V.cy;
                if (!(uf > -1.0f && uf < (float)V.w[0])) continue;
                if (!(vf > -1.0f && vf < (float)V.h[0])) continue;
                const int u0 = (int)uf, v0 = (int)vf;
                float fp = fabsf(V.fx) * r_c / z;
                fp = fmaxf(1.0f, fp);
                int lvl;
                if (!(fp <= FLT_MAX)) {
                    lvl = kMaxLevels;  // infinite footprint: coarsest level
                } else {
                    const int e = (int)((__float_as_uint(fp) >> 23) & 0xffu) - 127;
                    lvl = (e < kMaxLevels && fp >= a.thr.t[e]) ? e + 1 : e;
                }
                lvl = min(lvl, V.nlev - 1);
                const int uu = min(max(u0 >> lvl, 0), V.w[lvl] - 1);
                const int vv = min(max(v0 >> lvl, 0), V.h[lvl] - 1);
                // depth_at + vote (histogram.cpp:148-158)
                const float d = V.lv[lvl][(uint64_t)vv * V.w[lvl] + uu];
                if (!isfinite(d) || d <= 0.0f) continue;
                float s = d - z;
                if (s < -eta) continue;
                s = s / delta;
                s = s < -1.0f ? -1.0f : (1.0f < s ? 1.0f : s);
                int bin = (int)((s + 1.0f) * scale);
                bin = min(max(bin, 0), a.nb - 1);
                cnt[bin * T + t] += V.lidar ? 5u : 1u;
            }
        }
    }
    __syncthreads();
    // cube-major, coalesced: element i of the block's output is bin i % nb of cube i / nb
    const uint64_t nvalid = (a.n > base) ? min((uint64_t)T, a.n - base) : 0;
    const uint64_t total = nvalid * (uint64_t)a.nb;
    for (uint64_t i = t; i < total; i += T) {
        const int k = (int)(i % a.nb), cc = (int)(i / a.nb);
        a.bins[base * a.nb + i] = cnt[k * T + cc];
    }
}

// median_indicator (histogram.cpp:21-31) with nb bins, FP32.
__global__ void k_median(const uint32_t* __restrict__ bins, uint64_t n, int nb, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* b = bins + i * nb;
    uint32_t tot = 0;
    for (int k = 0; k < nb; ++k) tot += b[k];
    float r = 0.0f;
    if (tot != 0) {
        r = 1.0f;
        uint32_t cum = 0;
        const float half = 0.5f * (float)nb;
        for (int k = 0; k < nb; ++k) {
            cum += b[k];
            if (cum * 2u >= tot) {
                r = -1.0f + ((float)k + 0.5f) / half;
                break;
            }
        }
    }
    out[i] = r;
}

namespace {
// Level sizes of DepthMipmap::build (histogram.cpp:42-50).
std::vector<std::pair<int, int>> mip_dims(int w, int h) {
    std::vector<std::pair<int, int>> d{{w, h}};
    for (int l = 0; l < 12; ++l) {
        const auto [pw, ph] = d.back();
        if (pw <= 1 || ph <= 1) break;
        d.push_back({std::max(1, pw / 2), std::max(1, ph / 2)});
    }
    return d;
}

// Fills the view table (pyramid levels at base + off) and returns the floats
// all levels need.
uint64_t view_meta(const tgvx_view& in, ViewMeta& m, uint64_t off) {
    if (in.width <= 0 || in.height <= 0 || !in.depths) throw Error(TGVX_E_ARG, "view without depths");
    std::memset(&m, 0, sizeof(m));
    std::memcpy(m.m, in.world_to_cam, sizeof(m.m));
    m.fx = in.fx; m.fy = in.fy; m.cx = in.cx; m.cy = in.cy;
    m.lidar = in.lidar;
    const auto d = mip_dims(in.width, in.height);
    m.nlev = (int)d.size();
    for (int l = 0; l < m.nlev; ++l) {
        m.w[l] = d[l].first;
        m.h[l] = d[l].second;
        m.lv[l] = reinterpret_cast<const float*>(off * sizeof(float));  // offset until placed
        off += (uint64_t)d[l].first * d[l].second;
    }
    return off;
}

void launch_mips(const ViewMeta& m, hipStream_t s) {
    for (int l = 1; l < m.nlev; ++l) {
        const int cnt = m.w[l] * m.h[l];
        k_mip<<<(cnt + 255) / 256, 256, 0, s>>>(m.lv[l - 1], m.w[l - 1], m.h[l - 1], const_cast<float*>(m.lv[l]),
                                               m.w[l], m.h[l]);
        TGVX_HIP(hipGetLastError());
    }
}

// Uploads the base levels and builds every pyramid on the device.
void build_pyramids(tgvx_ctx* c, const tgvx_view* views, int nv, std::vector<ViewMeta>& meta) {
    meta.assign(nv, ViewMeta{});
    uint64_t total = 0;
    for (int v = 0; v < nv; ++v) view_meta(views[v], meta[v], 0);
    // every view's base level first, one after another (one gathered upload),
    // then the views' coarser levels
    std::vector<HostPiece> base(nv);
    for (int v = 0; v < nv; ++v) {
        ViewMeta& m = meta[v];
        base[v] = HostPiece{views[v].depths, sizeof(float) * (size_t)m.w[0] * m.h[0]};
        m.lv[0] = reinterpret_cast<const float*>(total * sizeof(float));
        total += (uint64_t)m.w[0] * m.h[0];
    }
    for (int v = 0; v < nv; ++v) {
        ViewMeta& m = meta[v];
        for (int l = 1; l < m.nlev; ++l) {
            m.lv[l] = reinterpret_cast<const float*>(total * sizeof(float));
            total += (uint64_t)m.w[l] * m.h[l];
        }
    }
    c->h_pyr.ensure(total * sizeof(float));
    char* pyr = c->h_pyr.as<char>();
    for (int v = 0; v < nv; ++v)
        for (int l = 0; l < meta[v].nlev; ++l)
            meta[v].lv[l] = reinterpret_cast<const float*>(pyr + (uintptr_t)meta[v].lv[l]);
    h2dv(c, pyr, base.data(), nv);
    for (int v = 0; v < nv; ++v) launch_mips(meta[v], c->stream);
}
}  // namespace

void histograms(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views, int nv,
                const float origin[3], float r_root, int nb, uint32_t* bins_out) {
    if (nb <= 0 || nb > kMaxBins) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (nv < 0 || (nv > 0 && !views)) throw Error(TGVX_E_ARG, "bad views");
    if (n == 0) return;
    if (!cubes || !bins_out || !origin) throw Error(TGVX_E_ARG, "null pointer");
    static const Log2Thr thr = log2_thresholds();
    std::vector<ViewMeta> meta;
    if (nv > 0) build_pyramids(c, views, nv, meta);
    c->h_meta.ensure(sizeof(ViewMeta) * std::max(1, nv));
    if (nv > 0)
        TGVX_HIP(hipMemcpyAsync(c->h_meta.p, meta.data(), sizeof(ViewMeta) * nv, hipMemcpyHostToDevice, c->stream));
    c->h_cubes.ensure(sizeof(tgvx_cube) * n);
    c->h_bins.ensure(sizeof(uint32_t) * n * nb);
    h2d(c, c->h_cubes.p, cubes, sizeof(tgvx_cube) * n);
    VoteArgs a;
    a.cubes = c->h_cubes.as<tgvx_cube>();
    a.n = n;
    a.views = c->h_meta.as<ViewMeta>();
    a.nv = nv;
    a.nb = nb;
    a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
    a.r_root = r_root;
    a.thr = thr;
    a.bins = c->h_bins.as<uint32_t>();
    const int T = 256;
    const uint64_t blocks = (n + T - 1) / T;
    if (blocks > 0x7fffffffull) throw Error(TGVX_E_ARG, "too many cubes");
    k_vote<<<(unsigned)blocks, T, sizeof(uint32_t) * nb * T, c->stream>>>(a);
    TGVX_HIP(hipGetLastError());
    d2h(c, bins_out, c->h_bins.p, sizeof(uint32_t) * n * nb);
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

void mip_build(tgvx_ctx* c, const tgvx_view* v, float* out, int32_t* dims, int32_t* nlevels) {
    if (!v || !out || !dims || !nlevels) throw Error(TGVX_E_ARG, "null pointer");
    std::vector<ViewMeta> meta;
    build_pyramids(c, v, 1, meta);
    const ViewMeta& m = meta[0];
    uint64_t total = 0;
    for (int l = 0; l < m.nlev; ++l) {
        dims[2 * l] = m.w[l];
        dims[2 * l + 1] = m.h[l];
        total += (uint64_t)m.w[l] * m.h[l];
    }
    *nlevels = m.nlev;
    TGVX_HIP(hipMemcpyAsync(out, m.lv[0], sizeof(float) * total, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

void median_init(tgvx_ctx* c, const uint32_t* bins, uint64_t n, int nb, float* out) {
    if (nb <= 0 || nb > kMaxBins) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (n == 0) return;
    c->h_bins.ensure(sizeof(uint32_t) * n * nb);
    c->h_pyr.ensure(sizeof(float) * n);
    h2d(c, c->h_bins.p, bins, sizeof(uint32_t) * n * nb);
    k_median<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->h_bins.as<uint32_t>(), n, nb,
                                                                   c->h_pyr.as<float>());
    TGVX_HIP(hipGetLastError());
    d2h(c, out, c->h_pyr.p, sizeof(float) * n);
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

// ---- allocation-free histogram voting (no reference counterpart)
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): inputs and output are the caller's device buffers, and
// everything else the vote needs -- the view table and the pyramid levels above
// each caller-held base level -- lives in one file-scope device array. Every
// device holds its own copy until the process ends; host code never uses the
// array's name, only the address hipGetSymbolAddress gives on each device.
constexpr uint64_t kScratchBytes = (uint64_t)TGVX_NOALLOC_SCRATCH_BYTES;
constexpr uint64_t kMetaBytes = (sizeof(ViewMeta) * TGVX_NOALLOC_MAX_VIEWS + 255) / 256 * 256;
static_assert(kMetaBytes < kScratchBytes, "scratch too small for the view table");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

void* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) throw Error(TGVX_E_STATE, "device index beyond the scratch table");
    if (!g_scratch[device])
        throw Error(TGVX_E_STATE, std::string("allocation-free scratch has no address on this device: ") +
                                      hipGetErrorString(g_scratch_err[device]));
    return g_scratch[device];
}

// The caller's buffer must be device memory of the context's device.
void require_device(const void* p, int device, const char* what) {
    hipPointerAttribute_t at;
    if (!p || hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        throw Error(TGVX_E_ARG, std::string(what) + " is not device memory");
    }
    if ((at.type != hipMemoryTypeDevice && at.type != hipMemoryTypeManaged) || at.device != device)
        throw Error(TGVX_E_ARG, std::string(what) + " is not memory of the context's device");
}
}  // namespace

// Once per device, with that device current (tgvx_create): the address of the
// device's copy of scratch. A failure is kept and reported by the entry point.
void noalloc_setup(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || g_scratch[device]) return;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e == hipSuccess && ptr) {
        g_scratch[device] = ptr;
    } else {
        g_scratch_err[device] = e;
        (void)hipGetLastError();
    }
}

void histograms_noalloc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views, int nv,
                        const float origin[3], float r_root, int nb, uint32_t* bins) {
    if (nb <= 0 || nb > kMaxBins) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (nv < 0 || (nv > 0 && !views)) throw Error(TGVX_E_ARG, "bad views");
    if (nv > TGVX_NOALLOC_MAX_VIEWS)
        throw Error(TGVX_E_ARG, "the allocation-free call takes at most " + std::to_string(TGVX_NOALLOC_MAX_VIEWS) +
                                    " views");
    if (n == 0) return;
    if (!origin) throw Error(TGVX_E_ARG, "null pointer");
    const uint64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffull) throw Error(TGVX_E_ARG, "too many cubes");
    require_device(cubes, c->device, "cubes");
    require_device(bins, c->device, "bins");
    // pyramid levels above the base, as DepthMipmap::build makes them
    uint64_t need = 0;
    for (int v = 0; v < nv; ++v) {
        ViewMeta m;
        need += view_meta(views[v], m, 0) - (uint64_t)m.w[0] * m.h[0];
        require_device(views[v].depths, c->device, "view depths");
    }
    if (need * sizeof(float) > kScratchBytes - kMetaBytes)
        throw Error(TGVX_E_ARG, "the views' pyramid levels need " + std::to_string(need * sizeof(float)) +
                                    " bytes; the allocation-free call holds at most " +
                                    std::to_string(kScratchBytes - kMetaBytes));
    char* base = static_cast<char*>(scratch_of(c->device));
    ViewMeta* meta_dev = reinterpret_cast<ViewMeta*>(base);
    float* mip = reinterpret_cast<float*>(base + kMetaBytes);
    static const Log2Thr thr = log2_thresholds();
    // the view table in chunks from the stack: each chunk is copied and its
    // pyramids launched, then the stream drains before the chunk is reused
    constexpr int kChunk = 16;
    ViewMeta chunk[kChunk];
    uint64_t off = 0;
    for (int v0 = 0; v0 < nv; v0 += kChunk) {
        const int k = std::min(kChunk, nv - v0);
        for (int j = 0; j < k; ++j) {
            ViewMeta& m = chunk[j];
            view_meta(views[v0 + j], m, 0);
            m.lv[0] = views[v0 + j].depths;
            for (int l = 1; l < m.nlev; ++l) {
                m.lv[l] = mip + off;
                off += (uint64_t)m.w[l] * m.h[l];
            }
            launch_mips(m, c->stream);
        }
        TGVX_HIP(hipMemcpyAsync(meta_dev + v0, chunk, sizeof(ViewMeta) * k, hipMemcpyHostToDevice, c->stream));
        TGVX_HIP(hipStreamSynchronize(c->stream));
    }
    VoteArgs a;
    a.cubes = cubes;
    a.n = n;
    a.views = meta_dev;
    a.nv = nv;
    a.nb = nb;
    a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
    a.r_root = r_root;
    a.thr = thr;
    a.bins = bins;
    k_vote<<<(unsigned)blocks, 256, sizeof(uint32_t) * nb * 256, c->stream>>>(a);
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace tgvx
