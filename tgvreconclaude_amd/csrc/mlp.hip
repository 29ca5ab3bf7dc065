// Warm-start network on the device (no reference counterpart): per-cube
// features, a 30 -> 64 -> 64 -> 4 ReLU MLP on FP32 matrix cores
// (v_mfma_f32_16x16x4f32), its backward pass and an Adam optimizer.
//
// Parameters, packed as W1[30][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]
// (6404 floats), live in the context. A workgroup of 4 wavefronts handles 64
// rows (cubes) at a time, each wavefront 16 of them; every product is a set
// of 16x16 output tiles accumulated by MFMA over K in steps of 4.
//
// MFMA 16x16x4 operand layout (wave64): lane l supplies A[l%16][l/16] and
// B[l/16][l%16]; the four results of lane l are D[4(l/16)+r][l%16], r = 0..3.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "tgvx_device.h"
#include "tgvx_internal.h"

namespace tgvx {

constexpr int kIn = 30, kInPad = 32, kHid = 64, kOut = 4;
constexpr int kOW1 = 0, kOB1 = kOW1 + kIn * kHid, kOW2 = kOB1 + kHid, kOB2 = kOW2 + kHid * kHid,
              kOW3 = kOB2 + kHid, kOB3 = kOW3 + kHid * kOut, kNP = kOB3 + kOut;
static_assert(kNP == 6404, "parameter count");
constexpr int kRows = 64;          // rows per workgroup step
constexpr int kSX = kInPad + 1;    // LDS row strides (padded against bank conflicts)
constexpr int kSH = kHid + 1;

typedef float f4 __attribute__((ext_vector_type(4)));

// acc += A[16 x K] * B[K x 16]; A(i, k) and B(k, j) are LDS / global reads.
template <int K, class FA, class FB>
__device__ inline f4 mma16(FA A, FB B, f4 acc) {
    const int lane = threadIdx.x & 63, i = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 4) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A(i, k0 + kk), B(k0 + kk, i), acc, 0, 0, 0);
    return acc;
}
__device__ inline int out_row(int r) { return 4 * ((threadIdx.x & 63) >> 4) + r; }
__device__ inline int out_col() { return threadIdx.x & 15; }

struct MlpW {  // packed parameters in LDS, rows of W padded to 16-column tiles
    float w1[kInPad * kHid];
    float w2[kHid * kHid];
    float w3[kHid * 16];
    float b1[kHid], b2[kHid], b3[16];
};

__device__ inline void stage_weights(const float* __restrict__ w, MlpW& s) {
    for (int i = threadIdx.x; i < kInPad * kHid; i += blockDim.x) s.w1[i] = i < kIn * kHid ? w[kOW1 + i] : 0.0f;
    for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) s.w2[i] = w[kOW2 + i];
    for (int i = threadIdx.x; i < kHid * 16; i += blockDim.x) {
        const int r = i / 16, cc = i % 16;
        s.w3[i] = cc < kOut ? w[kOW3 + r * kOut + cc] : 0.0f;
    }
    for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
        s.b1[i] = w[kOB1 + i];
        s.b2[i] = w[kOB2 + i];
    }
    for (int i = threadIdx.x; i < 16; i += blockDim.x) s.b3[i] = i < kOut ? w[kOB3 + i] : 0.0f;
}

// Forward of the 16 rows owned by this wavefront: x -> h1 -> h2 (h2
// overwrites the wavefront's own rows of x) -> y[16][4] in registers. x and
// h use the same row stride, so the rows of different wavefronts never overlap.
__device__ inline f4 forward16(const MlpW& W, float* xb, float* hb, int r0) {
    const int c = out_col();
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma16<kInPad>([&](int i, int k) { return xb[(r0 + i) * kSH + k]; },
                            [&](int k, int j) { return W.w1[k * kHid + ct * 16 + j]; }, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) hb[(r0 + out_row(r)) * kSH + ct * 16 + c] = fmaxf(acc[r] + W.b1[ct * 16 + c], 0.0f);
    }
    f4 h2[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        h2[ct] = mma16<kHid>([&](int i, int k) { return hb[(r0 + i) * kSH + k]; },
                             [&](int k, int j) { return W.w2[k * kHid + ct * 16 + j]; }, acc);
    }
    // x is no longer needed by this wavefront: h2 goes into its rows of xb
    // (stride kSH, the buffer is sized for it)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) xb[(r0 + out_row(r)) * kSH + ct * 16 + c] = fmaxf(h2[ct][r] + W.b2[ct * 16 + c], 0.0f);
    f4 y = {0.0f, 0.0f, 0.0f, 0.0f};
    y = mma16<kHid>([&](int i, int k) { return xb[(r0 + i) * kSH + k]; },
                    [&](int k, int j) { return W.w3[k * 16 + j]; }, y);
#pragma unroll
    for (int r = 0; r < 4; ++r) y[r] += W.b3[c];
    return y;
}

// y[n][4] = network(x[n][32]) (columns 30, 31 of x are zero)
__global__ void __launch_bounds__(256) k_mlp_fwd(const float* __restrict__ w, const float* __restrict__ x, uint32_t n,
                                                 float* __restrict__ y) {
    __shared__ MlpW W;
    __shared__ float xb[kRows * kSH];
    __shared__ float hb[kRows * kSH];
    stage_weights(w, W);
    __syncthreads();
    const int r0 = 16 * (threadIdx.x >> 6);
    for (uint32_t base = blockIdx.x * kRows; base < n; base += gridDim.x * kRows) {
        // each wavefront loads and processes its own 16 rows
        for (int i = threadIdx.x & 63; i < 16 * kInPad; i += 64) {
            const int rr = r0 + i / kInPad, k = i % kInPad;
            xb[rr * kSH + k] = base + rr < n ? x[(size_t)(base + rr) * kInPad + k] : 0.0f;
        }
        const f4 out = forward16(W, xb, hb, r0);
        const int c = out_col();
        if (c < kOut)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const uint32_t row = base + r0 + out_row(r);
                if (row < n) y[(size_t)row * kOut + c] = out[r];
            }
    }
}

// Loss and gradient of one tile of kRows rows (rows >= nvalid contribute
// nothing); partial gradients (kNP) and loss go to part[blockIdx.x].
// rows[] maps the tile's rows to data rows (a shuffled epoch); scale =
// 1 / (4 * rows in the whole batch).
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ w, const float* __restrict__ x,
                                                  const float* __restrict__ t, const uint32_t* __restrict__ rows,
                                                  uint32_t nvalid, float scale, float* __restrict__ part) {
    __shared__ MlpW W;
    __shared__ float xb[kRows * kSH];   // x, later h2
    __shared__ float hb[kRows * kSH];   // h1, later dz1
    __shared__ float xs[kRows * kSX];   // x kept for dW1
    __shared__ float dy[kRows * 16];
    __shared__ float dz2[kRows * kSH];
    __shared__ float red[256];
    stage_weights(w, W);
    const uint32_t base = blockIdx.x * kRows;
    for (int i = threadIdx.x; i < kRows * kInPad; i += blockDim.x) {
        const int rr = i / kInPad, k = i % kInPad;
        const uint32_t row = base + rr;
        const float v = row < nvalid ? x[(size_t)rows[row] * kInPad + k] : 0.0f;
        xb[rr * kSH + k] = v;
        xs[rr * kSX + k] = v;
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, r0 = 16 * wv, c = out_col();
    const f4 y = forward16(W, xb, hb, r0);  // h1 in hb, h2 in xb
    // dy = 2 (y - t) scale on valid rows; per-thread loss
    float lsum = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int rr = r0 + out_row(r);
        const uint32_t row = base + rr;
        float d = 0.0f;
        if (c < kOut && row < nvalid) {
            const float e = y[r] - t[(size_t)rows[row] * kOut + c];
            lsum += e * e * scale;
            d = 2.0f * e * scale;
        }
        dy[rr * 16 + c] = d;
    }
    red[threadIdx.x] = lsum;
    __syncthreads();
    float* g = part + (size_t)blockIdx.x * (kNP + 1);
    // dW3 [64 x 4] = h2^T dy: wave wv computes hidden rows 16 wv .. 16 wv + 15
    {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma16<kRows>([&](int i, int k) { return xb[k * kSH + r0 + i]; },
                           [&](int k, int j) { return dy[k * 16 + j]; }, acc);
        if (c < kOut)
#pragma unroll
            for (int r = 0; r < 4; ++r) g[kOW3 + (r0 + out_row(r)) * kOut + c] = acc[r];
    }
    // dz2 = (dy W3^T) * [h2 > 0] for this wavefront's rows
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma16<16>([&](int i, int k) { return dy[(r0 + i) * 16 + k]; },
                        [&](int k, int j) { return W.w3[(ct * 16 + j) * 16 + k]; }, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rr = r0 + out_row(r), cc = ct * 16 + c;
            dz2[rr * kSH + cc] = xb[rr * kSH + cc] > 0.0f ? acc[r] : 0.0f;
        }
    }
    __syncthreads();
    // dW2 [64 x 64] = h1^T dz2: wave wv computes hidden rows 16 wv ..
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma16<kRows>([&](int i, int k) { return hb[k * kSH + r0 + i]; },
                           [&](int k, int j) { return dz2[k * kSH + ct * 16 + j]; }, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) g[kOW2 + (r0 + out_row(r)) * kHid + ct * 16 + c] = acc[r];
    }
    // dz1 = (dz2 W2^T) * [h1 > 0] for this wavefront's rows, kept in registers
    f4 dz1[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        dz1[ct] = mma16<kHid>([&](int i, int k) { return dz2[(r0 + i) * kSH + k]; },
                              [&](int k, int j) { return W.w2[(ct * 16 + j) * kHid + k]; }, acc);
    }
    __syncthreads();  // every wavefront is done reading h1 (hb)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int rr = r0 + out_row(r), cc = ct * 16 + c;
            hb[rr * kSH + cc] = hb[rr * kSH + cc] > 0.0f ? dz1[ct][r] : 0.0f;
        }
    __syncthreads();
    // dW1 [30 x 64] = x^T dz1: 2 x 4 tiles, two per wavefront
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        const int tile = wv * 2 + tt, rt = tile & 1, ct = tile >> 1;
        f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma16<kRows>([&](int i, int k) { return xs[k * kSX + rt * 16 + i]; },
                           [&](int k, int j) { return hb[k * kSH + ct * 16 + j]; }, acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int fi = rt * 16 + out_row(r);
            if (fi < kIn) g[kOW1 + fi * kHid + ct * 16 + c] = acc[r];
        }
    }
    // biases: column sums of dz1, dz2, dy; loss
    if (threadIdx.x < kHid) {
        float s1 = 0.0f, s2 = 0.0f;
        for (int rr = 0; rr < kRows; ++rr) {
            s1 += hb[rr * kSH + threadIdx.x];
            s2 += dz2[rr * kSH + threadIdx.x];
        }
        g[kOB1 + threadIdx.x] = s1;
        g[kOB2 + threadIdx.x] = s2;
    } else if (threadIdx.x < kHid + kOut) {
        const int j = threadIdx.x - kHid;
        float s = 0.0f;
        for (int rr = 0; rr < kRows; ++rr) s += dy[rr * 16 + j];
        g[kOB3 + j] = s;
    } else if (threadIdx.x == 255) {
        float s = 0.0f;
        for (int i = 0; i < 256; ++i) s += red[i];
        g[kNP] = s;
    }
}

// Sum the per-tile partials in tile order; optionally one Adam step.
__global__ void k_mlp_reduce(const float* __restrict__ part, uint32_t nparts, float* __restrict__ grad,
                             float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, int adam, float lr,
                             float c1, float c2, float* __restrict__ loss_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kNP) return;
    float s = 0.0f;
    for (uint32_t b = 0; b < nparts; ++b) s += part[(size_t)b * (kNP + 1) + i];
    if (i == kNP) {
        if (loss_out) *loss_out = s;
        return;
    }
    if (grad) grad[i] = s;
    if (adam) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mi = b1 * m[i] + (1.0f - b1) * s;
        const float vi = b2 * v[i] + (1.0f - b2) * s * s;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// ---- features of the level's cubes (30 per cube, padded to 32)
//  0-9   own histogram / W (0 for an empty one)
//  10-19 mean over the existing depth-aware face neighbours of theirs
//  20-23 (u, v) of the cube one level coarser at the parent code (the cold
//        initialisation), else (median, 0, 0, 0)
//  24-26 per axis (m(+a) - m(-a)) / 2, m the neighbour's median (own if none)
//  27    log(1 + W) / 8     28 existing face neighbours / 6     29 own median
struct FeatArgs {
    const uint32_t* bins;
    const uint32_t* info;
    const int32_t* nbp[3];
    const int32_t* nbm[3];
    const int32_t* parent;
    const float* u;
    const float* v[3];
    uint32_t off, n;
    float* x;
};

__device__ inline float median10(uint32_t w) {
    if (InfoWord::W(w) == 0u) return 0.0f;
    return -1.0f + ((float)((w >> 24) & 31u) + 0.5f) / 5.0f;
}

__global__ void k_features(FeatArgs a) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const uint32_t g = a.off + t;
    float f[kInPad];
    const uint32_t w = a.info[g];
    const uint32_t W = InfoWord::W(w);
    const float m = median10(w);
    for (int k = 0; k < 10; ++k) f[k] = W ? (float)a.bins[(size_t)g * 10 + k] / (float)W : 0.0f;
    float acc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    int cnt = 0;
    float mp[3], mm[3];
    for (int ax = 0; ax < 3; ++ax) {
        const int jp = a.nbp[ax][g], jm = a.nbm[ax][g];
        mp[ax] = m;
        mm[ax] = m;
        if (jp >= 0) {
            const uint32_t wj = a.info[jp], Wj = InfoWord::W(wj);
            for (int k = 0; k < 10; ++k) acc[k] += Wj ? (float)a.bins[(size_t)jp * 10 + k] / (float)Wj : 0.0f;
            mp[ax] = median10(wj);
            ++cnt;
        }
        if (jm >= 0) {
            const uint32_t wj = a.info[jm], Wj = InfoWord::W(wj);
            for (int k = 0; k < 10; ++k) acc[k] += Wj ? (float)a.bins[(size_t)jm * 10 + k] / (float)Wj : 0.0f;
            mm[ax] = median10(wj);
            ++cnt;
        }
    }
    for (int k = 0; k < 10; ++k) f[10 + k] = cnt ? acc[k] / (float)cnt : 0.0f;
    const int pj = a.parent[g];
    f[20] = pj >= 0 ? a.u[pj] : m;
    for (int cc = 0; cc < 3; ++cc) f[21 + cc] = pj >= 0 ? a.v[cc][pj] : 0.0f;
    for (int ax = 0; ax < 3; ++ax) f[24 + ax] = 0.5f * (mp[ax] - mm[ax]);
    f[27] = log1pf((float)W) / 8.0f;
    f[28] = (float)cnt / 6.0f;
    f[29] = m;
    f[30] = 0.0f;
    f[31] = 0.0f;
    for (int k = 0; k < kInPad; ++k) a.x[(size_t)t * kInPad + k] = f[k];
}

// init (u, v) of the level = cold initialisation (features 20-23) + network output
__global__ void k_warm_apply(const float* __restrict__ x, const float* __restrict__ y, uint32_t off, uint32_t n,
                             float* __restrict__ init) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    for (int c = 0; c < 4; ++c) init[(size_t)(off + t) * 4 + c] = x[(size_t)t * kInPad + 20 + c] + y[(size_t)t * 4 + c];
}

// ---------------------------------------------------------------- host side
namespace {
void ensure_mlp(tgvx_ctx* c) {
    if (!c->mlp_w.p) {
        c->mlp_w.ensure(sizeof(float) * kNP);
        c->mlp_m.ensure(sizeof(float) * kNP);
        c->mlp_v.ensure(sizeof(float) * kNP);
        TGVX_HIP(hipMemsetAsync(c->mlp_m.p, 0, sizeof(float) * kNP, c->stream));
        TGVX_HIP(hipMemsetAsync(c->mlp_v.p, 0, sizeof(float) * kNP, c->stream));
        // He-normal initial weights, zero biases, fixed seed
        std::vector<float> w(kNP, 0.0f);
        std::mt19937 rng(12345u);
        std::normal_distribution<float> n1(0.0f, std::sqrt(2.0f / kIn)), n2(0.0f, std::sqrt(2.0f / kHid));
        for (int i = 0; i < kIn * kHid; ++i) w[kOW1 + i] = n1(rng);
        for (int i = 0; i < kHid * kHid; ++i) w[kOW2 + i] = n2(rng);
        for (int i = 0; i < kHid * kOut; ++i) w[kOW3 + i] = 0.1f * n2(rng);
        TGVX_HIP(hipMemcpyAsync(c->mlp_w.p, w.data(), sizeof(float) * kNP, hipMemcpyHostToDevice, c->stream));
        c->mlp_step = 0;
    }
}

void upload_x(tgvx_ctx* c, DevBuf& d, const float* x, uint64_t n) {
    std::vector<float> xp((size_t)std::max<uint64_t>(n, 1) * kInPad, 0.0f);
    for (uint64_t r = 0; r < n; ++r) std::memcpy(&xp[r * kInPad], x + r * kIn, sizeof(float) * kIn);
    d.ensure(sizeof(float) * xp.size());
    TGVX_HIP(hipMemcpyAsync(d.p, xp.data(), sizeof(float) * xp.size(), hipMemcpyHostToDevice, c->stream));
}

unsigned fwd_grid(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + kRows - 1) / kRows, 2048)); }
}  // namespace

void mlp_set(tgvx_ctx* c, const float* w) {
    ensure_mlp(c);
    TGVX_HIP(hipMemcpyAsync(c->mlp_w.p, w, sizeof(float) * kNP, hipMemcpyHostToDevice, c->stream));
    TGVX_HIP(hipMemsetAsync(c->mlp_m.p, 0, sizeof(float) * kNP, c->stream));
    TGVX_HIP(hipMemsetAsync(c->mlp_v.p, 0, sizeof(float) * kNP, c->stream));
    c->mlp_step = 0;
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

void mlp_get(tgvx_ctx* c, float* w) {
    ensure_mlp(c);
    TGVX_HIP(hipMemcpyAsync(w, c->mlp_w.p, sizeof(float) * kNP, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

void mlp_forward(tgvx_ctx* c, const float* x, uint64_t n, float* y) {
    ensure_mlp(c);
    if (n == 0) return;
    if (n >= (1ull << 31)) throw Error(TGVX_E_ARG, "too many rows");
    upload_x(c, c->mlp_x, x, n);
    c->mlp_y.ensure(sizeof(float) * n * kOut);
    k_mlp_fwd<<<fwd_grid(n), 256, 0, c->stream>>>(c->mlp_w.as<float>(), c->mlp_x.as<float>(), (uint32_t)n,
                                                   c->mlp_y.as<float>());
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipMemcpyAsync(y, c->mlp_y.p, sizeof(float) * n * kOut, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

namespace {
// gradient of the rows rows[0..nrows) of the resident data; adam: update too
void grad_step(tgvx_ctx* c, const uint32_t* d_rows, uint32_t nrows, bool adam, float lr, float* d_grad,
               float* d_loss) {
    const uint32_t tiles = (nrows + kRows - 1) / kRows;
    c->mlp_part.ensure(sizeof(float) * (size_t)tiles * (kNP + 1));
    k_mlp_grad<<<tiles, 256, 0, c->stream>>>(c->mlp_w.as<float>(), c->mlp_x.as<float>(), c->mlp_t.as<float>(), d_rows,
                                             nrows, 1.0f / (4.0f * (float)nrows), c->mlp_part.as<float>());
    float c1 = 1.0f, c2 = 1.0f;
    if (adam) {
        ++c->mlp_step;
        c1 = 1.0f - std::pow(0.9f, (float)c->mlp_step);
        c2 = 1.0f - std::pow(0.999f, (float)c->mlp_step);
    }
    k_mlp_reduce<<<(kNP + 1 + 255) / 256, 256, 0, c->stream>>>(c->mlp_part.as<float>(), tiles, d_grad,
                                                               c->mlp_w.as<float>(), c->mlp_m.as<float>(),
                                                               c->mlp_v.as<float>(), adam ? 1 : 0, lr, c1, c2, d_loss);
    TGVX_HIP(hipGetLastError());
}
}  // namespace

void mlp_loss_grad(tgvx_ctx* c, const float* x, const float* t, uint64_t n, float* grad, double* loss) {
    ensure_mlp(c);
    if (n == 0 || n >= (1ull << 31)) throw Error(TGVX_E_ARG, "bad row count");
    upload_x(c, c->mlp_x, x, n);
    c->mlp_t.ensure(sizeof(float) * n * kOut);
    TGVX_HIP(hipMemcpyAsync(c->mlp_t.p, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice, c->stream));
    std::vector<uint32_t> rows(n);
    for (uint64_t i = 0; i < n; ++i) rows[i] = (uint32_t)i;
    c->mlp_rows.ensure(sizeof(uint32_t) * n);
    TGVX_HIP(hipMemcpyAsync(c->mlp_rows.p, rows.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
    c->mlp_y.ensure(sizeof(float) * (kNP + 1));
    float* d_grad = c->mlp_y.as<float>();
    grad_step(c, c->mlp_rows.as<uint32_t>(), (uint32_t)n, false, 0.0f, d_grad, d_grad + kNP);
    std::vector<float> out(kNP + 1);
    TGVX_HIP(hipMemcpyAsync(out.data(), d_grad, sizeof(float) * (kNP + 1), hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    if (grad) std::memcpy(grad, out.data(), sizeof(float) * kNP);
    if (loss) *loss = out[kNP];
}

void mlp_train(tgvx_ctx* c, const float* x, const float* t, uint64_t n, int epochs, int batch, float lr,
               uint64_t seed, float* epoch_loss) {
    ensure_mlp(c);
    if (n == 0 || n >= (1ull << 31) || epochs <= 0 || batch <= 0) throw Error(TGVX_E_ARG, "bad training arguments");
    upload_x(c, c->mlp_x, x, n);
    c->mlp_t.ensure(sizeof(float) * n * kOut);
    TGVX_HIP(hipMemcpyAsync(c->mlp_t.p, t, sizeof(float) * n * kOut, hipMemcpyHostToDevice, c->stream));
    c->mlp_rows.ensure(sizeof(uint32_t) * n);
    const uint64_t nb = (n + batch - 1) / batch;
    c->mlp_y.ensure(sizeof(float) * nb);
    std::vector<uint32_t> perm(n);
    std::mt19937_64 rng(seed);
    std::vector<float> losses(nb);
    for (int e = 0; e < epochs; ++e) {
        for (uint64_t i = 0; i < n; ++i) perm[i] = (uint32_t)i;
        std::shuffle(perm.begin(), perm.end(), rng);
        TGVX_HIP(hipMemcpyAsync(c->mlp_rows.p, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
        for (uint64_t b = 0; b < nb; ++b) {
            const uint64_t r0 = b * (uint64_t)batch, rn = std::min<uint64_t>(batch, n - r0);
            grad_step(c, c->mlp_rows.as<uint32_t>() + r0, (uint32_t)rn, true, lr, nullptr, c->mlp_y.as<float>() + b);
        }
        TGVX_HIP(hipMemcpyAsync(losses.data(), c->mlp_y.p, sizeof(float) * nb, hipMemcpyDeviceToHost, c->stream));
        TGVX_HIP(hipStreamSynchronize(c->stream));
        double s = 0.0, wsum = 0.0;
        for (uint64_t b = 0; b < nb; ++b) {
            const double rn = (double)std::min<uint64_t>(batch, n - b * (uint64_t)batch);
            s += losses[b] * rn;
            wsum += rn;
        }
        if (epoch_loss) epoch_loss[e] = (float)(s / wsum);
    }
}

// Features of the level at `depth` into d_x (n_L x 32); used by the warm
// start (then network + k_warm_apply) and to export training data.
void level_features(tgvx_ctx* c, uint32_t off, uint32_t n) {
    Problem& P = c->prob;
    if (P.nb != 10) throw Error(TGVX_E_ARG, "the warm-start network is defined for 10 histogram bins");
    if (P.precision != TGVX_FP32) throw Error(TGVX_E_ARG, "the warm start runs on FP32 state");
    FeatArgs a;
    a.bins = P.d_bins.as<uint32_t>();
    a.info = P.d_info.as<uint32_t>();
    for (int ax = 0; ax < 3; ++ax) {
        a.nbp[ax] = P.d_nba[ax].as<int32_t>();
        a.nbm[ax] = P.d_nbam[ax].as<int32_t>();
    }
    a.parent = P.d_parent_a.as<int32_t>();
    a.u = P.d_state[0].as<float>();
    for (int cc = 0; cc < 3; ++cc) a.v[cc] = P.d_state[1 + cc].as<float>();
    a.off = off;
    a.n = n;
    c->mlp_x.ensure(sizeof(float) * std::max<uint32_t>(n, 1) * kInPad);
    a.x = c->mlp_x.as<float>();
    if (n) k_features<<<(n + 255) / 256, 256, 0, c->stream>>>(a);
    TGVX_HIP(hipGetLastError());
}

// Learned initial (u, v) of one level into P.d_init (positions off .. off+n).
void warm_init(tgvx_ctx* c, uint32_t off, uint32_t n) {
    ensure_mlp(c);
    Problem& P = c->prob;
    level_features(c, off, n);
    if (n == 0) return;
    c->mlp_y.ensure(sizeof(float) * n * kOut);
    k_mlp_fwd<<<fwd_grid(n), 256, 0, c->stream>>>(c->mlp_w.as<float>(), c->mlp_x.as<float>(), n, c->mlp_y.as<float>());
    P.d_init.ensure(sizeof(float) * 4 * std::max<uint64_t>(P.n, 1));
    k_warm_apply<<<(n + 255) / 256, 256, 0, c->stream>>>(c->mlp_x.as<float>(), c->mlp_y.as<float>(), off, n,
                                                         P.d_init.as<float>());
    TGVX_HIP(hipGetLastError());
}

void problem_features(tgvx_ctx* c, int depth, float* x_out, float* uv_out, uint32_t* index_out) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    const Level* L = nullptr;
    for (const Level& l : P.levels)
        if (l.depth == depth) L = &l;
    if (!L) throw Error(TGVX_E_ARG, "no cubes at depth " + std::to_string(depth));
    level_features(c, L->off, L->n);
    std::vector<float> xp((size_t)L->n * kInPad);
    TGVX_HIP(hipMemcpyAsync(xp.data(), c->mlp_x.p, sizeof(float) * xp.size(), hipMemcpyDeviceToHost, c->stream));
    std::vector<float> col[4];
    for (int cc = 0; cc < 4; ++cc) {
        col[cc].resize(L->n);
        TGVX_HIP(hipMemcpyAsync(col[cc].data(), P.d_state[cc].as<float>() + L->off, sizeof(float) * L->n,
                                hipMemcpyDeviceToHost, c->stream));
    }
    TGVX_HIP(hipStreamSynchronize(c->stream));
    for (uint32_t t = 0; t < L->n; ++t) {
        if (x_out) std::memcpy(x_out + (size_t)t * kIn, &xp[(size_t)t * kInPad], sizeof(float) * kIn);
        if (uv_out)
            for (int cc = 0; cc < 4; ++cc) uv_out[(size_t)t * 4 + cc] = col[cc][t];
        if (index_out) index_out[t] = P.perm[L->off + t];
    }
}

}  // namespace tgvx
