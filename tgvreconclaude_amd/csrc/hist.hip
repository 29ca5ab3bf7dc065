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
    uint64_t off[kMaxLevels];
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
    const float* pyr;
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
                const float d = a.pyr[V.off[lvl] + (uint64_t)vv * V.w[lvl] + uu];
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

// Uploads the base levels and builds every pyramid on the device.
void build_pyramids(tgvx_ctx* c, const tgvx_view* views, int nv, std::vector<ViewMeta>& meta) {
    meta.assign(nv, ViewMeta{});
    uint64_t total = 0;
    for (int v = 0; v < nv; ++v) {
        const tgvx_view& in = views[v];
        if (in.width <= 0 || in.height <= 0 || !in.depths) throw Error(TGVX_E_ARG, "view without depths");
        ViewMeta& m = meta[v];
        std::memcpy(m.m, in.world_to_cam, sizeof(m.m));
        m.fx = in.fx; m.fy = in.fy; m.cx = in.cx; m.cy = in.cy;
        m.lidar = in.lidar;
        const auto d = mip_dims(in.width, in.height);
        m.nlev = (int)d.size();
        for (int l = 0; l < m.nlev; ++l) {
            m.w[l] = d[l].first;
            m.h[l] = d[l].second;
            m.off[l] = total;
            total += (uint64_t)d[l].first * d[l].second;
        }
    }
    c->h_pyr.ensure(total * sizeof(float));
    float* pyr = c->h_pyr.as<float>();
    for (int v = 0; v < nv; ++v) {
        const ViewMeta& m = meta[v];
        TGVX_HIP(hipMemcpyAsync(pyr + m.off[0], views[v].depths, sizeof(float) * (size_t)m.w[0] * m.h[0],
                                hipMemcpyHostToDevice, c->stream));
        for (int l = 1; l < m.nlev; ++l) {
            const int cnt = m.w[l] * m.h[l];
            k_mip<<<(cnt + 255) / 256, 256, 0, c->stream>>>(pyr + m.off[l - 1], m.w[l - 1], m.h[l - 1],
                                                            pyr + m.off[l], m.w[l], m.h[l]);
            TGVX_HIP(hipGetLastError());
        }
    }
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
    TGVX_HIP(hipMemcpyAsync(c->h_cubes.p, cubes, sizeof(tgvx_cube) * n, hipMemcpyHostToDevice, c->stream));
    VoteArgs a;
    a.cubes = c->h_cubes.as<tgvx_cube>();
    a.n = n;
    a.views = c->h_meta.as<ViewMeta>();
    a.nv = nv;
    a.nb = nb;
    a.pyr = c->h_pyr.as<float>();
    a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2];
    a.r_root = r_root;
    a.thr = thr;
    a.bins = c->h_bins.as<uint32_t>();
    const int T = 256;
    const uint64_t blocks = (n + T - 1) / T;
    if (blocks > 0x7fffffffull) throw Error(TGVX_E_ARG, "too many cubes");
    k_vote<<<(unsigned)blocks, T, sizeof(uint32_t) * nb * T, c->stream>>>(a);
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipMemcpyAsync(bins_out, c->h_bins.p, sizeof(uint32_t) * n * nb, hipMemcpyDeviceToHost, c->stream));
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
    TGVX_HIP(hipMemcpyAsync(out, c->h_pyr.p, sizeof(float) * total, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

void median_init(tgvx_ctx* c, const uint32_t* bins, uint64_t n, int nb, float* out) {
    if (nb <= 0 || nb > kMaxBins) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (n == 0) return;
    c->h_bins.ensure(sizeof(uint32_t) * n * nb);
    c->h_pyr.ensure(sizeof(float) * n);
    TGVX_HIP(hipMemcpyAsync(c->h_bins.p, bins, sizeof(uint32_t) * n * nb, hipMemcpyHostToDevice, c->stream));
    k_median<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(c->h_bins.as<uint32_t>(), n, nb,
                                                                   c->h_pyr.as<float>());
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipMemcpyAsync(out, c->h_pyr.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace tgvx
