// Octree input build on the device: one cube per depth-map pixel
// (reference src/octree_builder.cpp:11-97), the merge of equal (code, depth)
// cells (src/merge.cpp:57-92) and the 2:1 balance with bridge cubes
// (src/balance.cpp:60-150).
//
// Stage 1, per pixel: DepthMap::unproject / sample_radius (src/depth_map.cpp:
// 4-30) in float, the depth rule of cube_depth_for_radius, the fine-grid code.
// Stage 2: the valid pixels, sorted by (code, depth, view, pixel) with two
// stable radix sorts, and per run of equal (code, depth) the mean r_c, summed
// in double in that order.
// Stage 3: the balance is a FIFO queue: the seeds in input order, then every
// bridge in the order it was inserted; cube k's check of face f sees the
// cells inserted before it. The device runs it as rounds of independent
// checks: a round evaluates every queued cube's six checks against the
// bridges of the previous round, each bridge time-stamped with the (queue
// position, face) of the check that inserted it, so a check at time t sees
// exactly the bridges stamped before t. A bridge's stamp is the earliest check
// that inserts it; its queue position is the seed count plus its rank by
// stamp. The round's bridges become the next round's; the rounds stop when a
// round reproduces its input. At that fixed point every check agrees with the
// state its predecessors produce, which by induction over the queue order is
// the sequential run's state; and each round fixes at least the earliest check
// that was still wrong, so the rounds terminate (8-14 rounds on the scenes
// measured).
#include <hip/hip_runtime.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "tgvx_device.h"
#include "tgvx_internal.h"

namespace tgvx {
namespace {

double ms_now() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------- stage 1
struct PixView {
    const float* d;
    int W, H;
    float fx, fy, cx, cy;
    float M[12];
};

__device__ inline bool pix_valid(const PixView& v, int u, int w) {
    if (u < 0 || w < 0 || u >= v.W || w >= v.H) return false;
    const float x = v.d[(long)w * v.W + u];
    return isfinite(x) && x > 0.0f;
}
// DepthMap::unproject (depth_map.cpp:4-11): camera point ((u - cx) z / fx,
// (v - cy) z / fy, z), then Rt.
__device__ inline void pix_at(const PixView& v, int u, int w, float out[3]) {
    const float z = v.d[(long)w * v.W + u];
    const float cx = ((float)u - v.cx) * z / v.fx, cy = ((float)w - v.cy) * z / v.fy;
#pragma unroll
    for (int r = 0; r < 3; ++r) out[r] = ((v.M[r * 4] * cx + v.M[r * 4 + 1] * cy) + v.M[r * 4 + 2] * z) + v.M[r * 4 + 3];
}
// DepthMap::sample_radius (depth_map.cpp:13-30): half the mean distance to
// the valid 4-neighbours, in the order +u, -u, +v, -v.
__device__ inline float pix_radius(const PixView& v, int u, int w) {
    if (!pix_valid(v, u, w)) return 0.0f;
    float c[3];
    pix_at(v, u, w, c);
    const int du[4] = {1, -1, 0, 0}, dv[4] = {0, 0, 1, -1};
    float sum = 0.0f;
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int nu = u + du[i], nv = w + dv[i];
        if (!pix_valid(v, nu, nv)) continue;
        float p[3];
        pix_at(v, nu, nv, p);
        const float dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
        sum += sqrtf(dx * dx + dy * dy + dz * dz);
        ++cnt;
    }
    if (cnt == 0) return 0.0f;
    return 0.5f * sum / (float)cnt;
}
// cube_depth_for_radius (octree_builder.cpp:11-28): the estimate rounds
// log2(r_root / (1.125 r)) and is then moved to the one depth whose cell
// size lies in [0.75 r, 1.5 r) (or a clamp); that depth does not depend on
// the last bit of the logarithm.
__device__ inline int depth_for_radius(float r_x, float r_root, int max_depth) {
    if (r_x <= 0.0f) return max_depth;
    int d = (int)roundf(log2f(r_root / (r_x * 1.125f)));
    if (d < 0) d = 0;
    if (d > max_depth) d = max_depth;
    for (int tries = 0; tries < 3; ++tries) {
        const float cell = r_root / (float)(1u << d);
        if (cell >= 0.75f * r_x && cell < 1.5f * r_x) break;
        if (cell < 0.75f * r_x) {
            if (d > 0) --d;
            else break;
        } else {
            if (d < max_depth) ++d;
            else break;
        }
    }
    return d;
}

// One thread per pixel of one view: the pixel's cube (code, depth, r) at
// slot base + pixel, flag 1 if it yields one.
__global__ void k_pix_cubes(PixView v, long base, float o0, float o1, float o2, float inv_2r, float r_root,
                            int max_depth, uint64_t* __restrict__ code, uint32_t* __restrict__ depth,
                            float* __restrict__ rad, uint8_t* __restrict__ flag) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)v.W * v.H) return;
    const int u = (int)(i % v.W), w = (int)(i / v.W);
    uint8_t ok = 0;
    if (pix_valid(v, u, w)) {
        const float r_x = pix_radius(v, u, w);
        if (r_x > 0.0f) {
            float p[3];
            pix_at(v, u, w, p);
            const float n0 = (p[0] - o0) * inv_2r, n1 = (p[1] - o1) * inv_2r, n2 = (p[2] - o2) * inv_2r;
            if (!(n0 < 0 || n0 >= 1 || n1 < 0 || n1 >= 1 || n2 < 0 || n2 >= 1)) {
                const int d = depth_for_radius(r_x, r_root, max_depth);
                const uint32_t g = 1u << d;
                uint32_t ix = (uint32_t)(n0 * (float)g), iy = (uint32_t)(n1 * (float)g), iz = (uint32_t)(n2 * (float)g);
                if (ix >= g) ix = g - 1;
                if (iy >= g) iy = g - 1;
                if (iz >= g) iz = g - 1;
                const int sh = 21 - d;
                code[base + i] = morton(ix << sh, iy << sh, iz << sh);
                depth[base + i] = (uint32_t)d;
                rad[base + i] = r_x;
                ok = 1;
            }
        }
    }
    flag[base + i] = ok;
}

// ---------------------------------------------------------------- stage 2
template <class T>
__global__ void k_gather(const T* __restrict__ src, const uint32_t* __restrict__ idx, uint64_t n, T* __restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}
// 1 where a run of equal (code, depth) starts (code sorted, depth by idx)
__global__ void k_run_heads(const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth,
                            const uint32_t* __restrict__ idx, uint64_t n, uint8_t* __restrict__ head) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = i == 0 || code[i] != code[i - 1] || depth[idx[i]] != depth[idx[i - 1]];
}
// merge.cpp:57-92 with the sum in double over the run, in sorted order.
// rs = the radii already in sorted order (contiguous per run). Runs of the
// coarsest cells collect up to millions of pixels (outliers), so the loads
// of a run are issued eight at a time; the additions keep their order.
__global__ void k_merge_runs(const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth,
                             const float* __restrict__ rs, const uint32_t* __restrict__ idx,
                             const uint32_t* __restrict__ start, uint64_t nruns, uint64_t nraw,
                             uint64_t* __restrict__ mcode, uint32_t* __restrict__ mdepth, float* __restrict__ mrc) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nruns) return;
    const uint64_t a = start[j], b = j + 1 < nruns ? start[j + 1] : nraw;
    double s = 0.0;
    uint64_t i = a;
    for (; i + 8 <= b; i += 8) {
        float x[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) x[k] = rs[i + k];
#pragma unroll
        for (int k = 0; k < 8; ++k) s += (double)x[k];
    }
    for (; i < b; ++i) s += (double)rs[i];
    mcode[j] = code[a];
    mdepth[j] = depth[idx[a]];
    mrc[j] = (float)(s / (double)(b - a));
}

// ---------------------------------------------------------------- stage 3
// A cell (depth d, code c) as one nonzero 64-bit key: the code's bits below
// the cell's size are zero, so the highest of them marks the depth (d < 21);
// a depth-21 cell sets bit 63 instead (codes use bits 0-62).
__host__ __device__ inline uint64_t cell_key(int d, uint64_t c) {
    return d >= 21 ? (c | (1ull << 63)) : (c | (1ull << (3 * (21 - d) - 1)));
}
__host__ __device__ inline void cell_of(uint64_t k, int& d, uint64_t& c) {
    if (k >> 63) {
        d = 21;
        c = k & ~(1ull << 63);
        return;
    }
    const int b = __builtin_ctzll(k);
    d = 21 - (b + 1) / 3;
    c = k & ~(1ull << b);
}
__host__ __device__ inline uint64_t key_hash(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// Open-addressing tables (linear probing, key 0 = empty): the seeds, and the
// current bridges with their time stamps.
struct Table {
    uint64_t* key;
    uint64_t* ts;  // null for the seed table
    uint64_t mask;
};
__global__ void k_table_insert(Table t, const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth,
                               const uint64_t* __restrict__ keys, const uint64_t* __restrict__ ts, uint64_t n,
                               int min_depth) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!keys && (int)depth[i] < min_depth) return;
    const uint64_t k = keys ? keys[i] : cell_key((int)depth[i], code[i]);
    for (uint64_t h = key_hash(k) & t.mask;; h = (h + 1) & t.mask) {
        const unsigned long long prev = atomicCAS((unsigned long long*)&t.key[h], 0ull, (unsigned long long)k);
        if (prev == 0ull || prev == k) {
            if (t.ts) t.ts[h] = ts[i];  // keys are distinct in the bridge table
            return;
        }
    }
}
__device__ inline bool table_find(const Table& t, uint64_t k, uint64_t* ts) {
    for (uint64_t h = key_hash(k) & t.mask;; h = (h + 1) & t.mask) {
        const uint64_t x = t.key[h];
        if (x == k) {
            if (ts) *ts = t.ts[h];
            return true;
        }
        if (x == 0ull) return false;
    }
}
// Cells of depth <= kBitDepth are also kept as bits, one per cell of their
// level (Morton index), levels packed coarse to fine (19 MB): the check's
// walk over the coarser ancestors reads bits instead of probing the tables,
// and the seed table holds only the deeper cells.
constexpr int kBitDepth = 9;
constexpr uint64_t kBitWords = ((((1ull << (3 * (kBitDepth + 1))) - 1) / 7) + 31) / 32;
__host__ __device__ inline uint64_t bit_of(int d, uint64_t c) {
    return ((1ull << (3 * d)) - 1) / 7 + (c >> (3 * (21 - d)));
}
__device__ inline bool bit_set(const uint32_t* bits, uint64_t b) { return (bits[b >> 5] >> (b & 31)) & 1u; }
// bits of the n cells (code, depth) or keys with depth <= kBitDepth
__global__ void k_bits_set(uint32_t* __restrict__ bits, const uint64_t* __restrict__ code,
                           const uint32_t* __restrict__ depth, const uint64_t* __restrict__ keys, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int d;
    uint64_t c;
    if (keys) {
        cell_of(keys[i], d, c);
    } else {
        d = (int)depth[i];
        c = code[i];
    }
    if (d > kBitDepth) return;
    const uint64_t b = bit_of(d, c);
    atomicOr(&bits[b >> 5], 1u << (b & 31));
}
struct Cells {
    Table s0, br;               // seeds deeper than kBitDepth; all bridges with their stamps
    const uint32_t* s0bits;     // seeds of depth <= kBitDepth
    const uint32_t* brbits;     // bridges of depth <= kBitDepth
};
// the cell is a seed
__device__ inline bool seed_has(const Cells& t, int d, uint64_t c) {
    if (d <= kBitDepth) return bit_set(t.s0bits, bit_of(d, c));
    return table_find(t.s0, cell_key(d, c), nullptr);
}
// the cell is a bridge stamped before tau
__device__ inline bool bridge_has(const Cells& t, int d, uint64_t c, uint64_t tau) {
    if (t.br.mask == 0) return false;
    if (d <= kBitDepth && !bit_set(t.brbits, bit_of(d, c))) return false;
    uint64_t ts;
    return table_find(t.br, cell_key(d, c), &ts) && ts < tau;
}

// One thread per queued cube k (seeds, then bridges by position): the six
// checks of balance.cpp:93-123 at times 1 + 6k + f; an insertion is recorded
// as (cell key, time). A check inserts iff neither the same-level neighbour
// (A) nor its parent one level coarser (B) exists and a cell two or more
// levels coarser covers the neighbour (C), each "exists" meaning a seed or an
// earlier bridge. The seeds' part never changes: the first round stores it
// per seed and face in seed_bits (4 bits per face: no neighbour, A, B, C by
// seeds) and later rounds only look the bridges up.
__global__ void __launch_bounds__(256) k_balance_checks(
    const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth, uint64_t nseed,
    const uint64_t* __restrict__ bkey, uint64_t nbr, Cells cells, uint32_t* __restrict__ seed_bits, int first_round,
    uint64_t* __restrict__ dkey, uint64_t* __restrict__ dts, unsigned long long* __restrict__ count, uint64_t cap) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int d = 0;
    uint64_t cv = 0;
    const bool seed = k < nseed;
    if (seed) {
        d = (int)depth[k];
        cv = code[k];
    } else if (k < nseed + nbr) {
        cell_of(bkey[k - nseed], d, cv);
    }
    // this cube's insertions (at most one per face), appended below with one
    // atomic per wavefront: a single counter hit by every lane serialises
    uint64_t ins_key[6], ins_ts[6];
    int nins = 0;
    if (d >= 2) {
        const bool known = seed && !first_round;
        const uint32_t old_bits = known ? seed_bits[k] : 0u;
        uint32_t bits = 0;
        for (int f = 0; f < 6; ++f) {
            uint32_t m;  // bit 0 no neighbour, 1 A, 2 B, 3 C, by seeds
            const uint64_t nb = face_nb(cv, f, d);
            const uint64_t par = nb == kNone ? 0 : ancestor(nb, d - 1);
            if (known) {
                m = (old_bits >> (4 * f)) & 15u;
            } else {
                m = nb == kNone ? 1u : 0u;
                if (!m && seed_has(cells, d, nb)) m |= 2u;
                if (!m && seed_has(cells, d - 1, par)) m |= 4u;
                if (!m)
                    for (int d2 = d - 2; d2 >= 0; --d2)
                        if (seed_has(cells, d2, ancestor(nb, d2))) {
                            m |= 8u;
                            break;
                        }
            }
            bits |= m << (4 * f);
            if (m & 7u) continue;  // no neighbour, or A / B by a seed: never an insertion
            const uint64_t tau = 1 + 6 * k + (uint64_t)f;
            if (bridge_has(cells, d, nb, tau) || bridge_has(cells, d - 1, par, tau)) continue;
            bool viol = m & 8u;
            for (int d2 = d - 2; d2 >= 0 && !viol; --d2) viol = bridge_has(cells, d2, ancestor(nb, d2), tau);
            if (!viol) continue;
            ins_key[nins] = cell_key(d - 1, par);
            ins_ts[nins] = tau;
            ++nins;
        }
        if (seed && first_round) seed_bits[k] = bits;
    }
    // wavefront prefix sum of the counts, one atomic per wavefront
    const int lane = (int)(threadIdx.x & 63u);
    int incl = nins;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    const int total = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 63 && total > 0) base = atomicAdd(count, (unsigned long long)total);
    base = __shfl(base, 63, 64);
    const unsigned long long first = base + (unsigned long long)(incl - nins);
    for (int i = 0; i < nins; ++i)
        if (first + i < cap) {
            dkey[first + i] = ins_key[i];
            dts[first + i] = ins_ts[i];
        }
}
// after sorting by (key, time): 1 at the first (earliest) record of a key
__global__ void k_key_heads(const uint64_t* __restrict__ key, uint64_t n, uint8_t* __restrict__ head) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) head[i] = i == 0 || key[i] != key[i - 1];
}
__global__ void k_differs(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, const uint64_t* __restrict__ c,
                          const uint64_t* __restrict__ e, uint64_t n, uint32_t* __restrict__ flag) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (a[i] != b[i] || c[i] != e[i])) flag[0] = 1u;
}
// final cells: seeds (their r_c), then bridges (r_root / 2^d, balance.cpp:118-119)
__global__ void k_cells(const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth,
                        const float* __restrict__ rc, uint64_t nseed, const uint64_t* __restrict__ bkey, uint64_t nbr,
                        float r_root, uint64_t* __restrict__ ocode, uint32_t* __restrict__ odepth,
                        float* __restrict__ orc) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nseed + nbr) return;
    if (i < nseed) {
        ocode[i] = code[i];
        odepth[i] = depth[i];
        orc[i] = rc[i];
    } else {
        int d;
        uint64_t cv;
        cell_of(bkey[i - nseed], d, cv);
        ocode[i] = cv;
        odepth[i] = (uint32_t)d;
        orc[i] = r_root / (float)(1u << d);
    }
}
__global__ void k_iota(uint32_t* p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}
__global__ void k_records(const uint64_t* __restrict__ code, const uint32_t* __restrict__ depth,
                          const float* __restrict__ rc, const uint32_t* __restrict__ idx, uint64_t n,
                          tgvx_cube* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = idx[i];
    tgvx_cube c;
    c.hi = (uint32_t)(code[j] >> 32);
    c.lo = (uint32_t)code[j];
    c.r_c = rc[j];
    c.depth = (int32_t)depth[j];
    out[i] = c;
}

unsigned grid_of(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

uint64_t pow2_at_least(uint64_t x) {
    uint64_t p = 1024;
    while (p < x) p <<= 1;
    return p;
}

// Scratch buffers of the octree calls, by role.
enum {
    B_CODE = 0, B_DEPTH, B_RAD, B_FLAG, B_IDX, B_IDX2, B_KEY32, B_KEY32S, B_CODES, B_HEAD, B_START, B_COUNT,
    B_TMP, B_S0, B_BRK, B_BRT, B_DKEY, B_DTS, B_DKEY2, B_DTS2, B_NKEY, B_NTS, B_PKEY, B_PTS, B_S0BITS, B_BRBITS,
    B_SEEDBITS
};

// stable sort of (keys, idx) pairs in place through the alternate buffers
// (bits [begin_bit, end_bit) of the keys)
template <class K>
void sort_pairs(tgvx_ctx* c, DevBuf& keys, DevBuf& keys_alt, DevBuf& vals, DevBuf& vals_alt, uint64_t n,
                int begin_bit, int end_bit) {
    if (n == 0) return;
    if (begin_bit >= end_bit) {  // nothing to sort by: the order stays
        TGVX_HIP(hipMemcpyAsync(keys_alt.p, keys.p, sizeof(K) * n, hipMemcpyDeviceToDevice, c->stream));
        TGVX_HIP(hipMemcpyAsync(vals_alt.p, vals.p, 4 * n, hipMemcpyDeviceToDevice, c->stream));
        return;
    }
    DevBuf& tmp = c->oct_tmp[B_TMP];
    size_t bytes = 0;
    TGVX_HIP(rocprim::radix_sort_pairs(nullptr, bytes, keys.as<K>(), keys_alt.as<K>(), vals.as<uint32_t>(),
                                       vals_alt.as<uint32_t>(), (size_t)n, begin_bit, end_bit, c->stream));
    tmp.ensure(bytes);
    TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, keys.as<K>(), keys_alt.as<K>(), vals.as<uint32_t>(),
                                       vals_alt.as<uint32_t>(), (size_t)n, begin_bit, end_bit, c->stream));
}
// stable compaction of the positions i < n with flag[i] != 0 into out; returns the count
uint64_t select_flagged(tgvx_ctx* c, const uint8_t* flag, uint64_t n, uint32_t* out) {
    if (n == 0) return 0;
    DevBuf& tmp = c->oct_tmp[B_TMP];
    DevBuf& cnt = c->oct_tmp[B_COUNT];
    cnt.ensure(16);
    rocprim::counting_iterator<uint32_t> it(0u);
    size_t bytes = 0;
    TGVX_HIP(rocprim::select(nullptr, bytes, it, flag, out, cnt.as<uint64_t>(), (size_t)n, c->stream));
    tmp.ensure(bytes);
    TGVX_HIP(rocprim::select(tmp.p, bytes, it, flag, out, cnt.as<uint64_t>(), (size_t)n, c->stream));
    uint64_t m = 0;
    TGVX_HIP(hipMemcpyAsync(&m, cnt.p, sizeof(m), hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    return m;
}

// Balance of the nseed cells (code, depth, r_c on the device, queue order)
// into c->oct_out, sorted by (code, depth), stable. Every code (seeds and
// bridges) has its bits below low_bit zero. Returns the bridge count.
uint64_t balance_device(tgvx_ctx* c, DevBuf& code, DevBuf& depth, DevBuf& rc, uint64_t nseed, float r_root,
                        int low_bit, int32_t* rounds) {
    DevBuf* T = c->oct_tmp;
    hipStream_t st = c->stream;
    // seed table
    const uint64_t cap0 = pow2_at_least(2 * std::max<uint64_t>(nseed, 1));
    T[B_S0].ensure(8 * cap0);
    TGVX_HIP(hipMemsetAsync(T[B_S0].p, 0, 8 * cap0, st));
    Table s0{T[B_S0].as<uint64_t>(), nullptr, cap0 - 1};
    T[B_SEEDBITS].ensure(4 * std::max<uint64_t>(nseed, 1));
    T[B_S0BITS].ensure(4 * kBitWords);
    T[B_BRBITS].ensure(4 * kBitWords);
    TGVX_HIP(hipMemsetAsync(T[B_S0BITS].p, 0, 4 * kBitWords, st));
    if (nseed) {
        k_table_insert<<<grid_of(nseed), 256, 0, st>>>(s0, code.as<uint64_t>(), depth.as<uint32_t>(), nullptr, nullptr,
                                                        nseed, kBitDepth + 1);
        k_bits_set<<<grid_of(nseed), 256, 0, st>>>(T[B_S0BITS].as<uint32_t>(), code.as<uint64_t>(),
                                                   depth.as<uint32_t>(), nullptr, nseed);
    }
    uint64_t nbr = 0;       // bridges of the previous round
    uint64_t dcap = std::max<uint64_t>(4096, nseed / 2);
    T[B_COUNT].ensure(16);
    T[B_HEAD].ensure(16);
    int r = 0;
    for (;;) {
        if (++r > 100000) throw Error(TGVX_E_HIP, "balance rounds did not settle");
        Table br{T[B_BRK].as<uint64_t>(), T[B_BRT].as<uint64_t>(), 0};
        TGVX_HIP(hipMemsetAsync(T[B_BRBITS].p, 0, 4 * kBitWords, st));
        if (nbr) {
            const uint64_t capb = pow2_at_least(2 * nbr);
            T[B_BRK].ensure(8 * capb);
            T[B_BRT].ensure(8 * capb);
            TGVX_HIP(hipMemsetAsync(T[B_BRK].p, 0, 8 * capb, st));
            br = Table{T[B_BRK].as<uint64_t>(), T[B_BRT].as<uint64_t>(), capb - 1};
            k_table_insert<<<grid_of(nbr), 256, 0, st>>>(br, nullptr, nullptr, T[B_PKEY].as<uint64_t>(),
                                                          T[B_PTS].as<uint64_t>(), nbr, 0);
            k_bits_set<<<grid_of(nbr), 256, 0, st>>>(T[B_BRBITS].as<uint32_t>(), nullptr, nullptr,
                                                     T[B_PKEY].as<uint64_t>(), nbr);
        }
        const Cells cells{s0, br, T[B_S0BITS].as<uint32_t>(), T[B_BRBITS].as<uint32_t>()};
        // this round's checks (a repeat of the first round stores the seed bits again)
        uint64_t nd = 0;
        const int first_round = r == 1;
        for (;;) {
            T[B_DKEY].ensure(8 * dcap);
            T[B_DTS].ensure(8 * dcap);
            TGVX_HIP(hipMemsetAsync(T[B_COUNT].p, 0, 8, st));
            const uint64_t items = nseed + nbr;
            if (items)
                k_balance_checks<<<grid_of(items), 256, 0, st>>>(
                    code.as<uint64_t>(), depth.as<uint32_t>(), nseed, T[B_PKEY].as<uint64_t>(), nbr, cells,
                    T[B_SEEDBITS].as<uint32_t>(), first_round,
                    T[B_DKEY].as<uint64_t>(), T[B_DTS].as<uint64_t>(), T[B_COUNT].as<unsigned long long>(), dcap);
            TGVX_HIP(hipGetLastError());
            TGVX_HIP(hipMemcpyAsync(&nd, T[B_COUNT].p, 8, hipMemcpyDeviceToHost, st));
            TGVX_HIP(hipStreamSynchronize(st));
            if (nd <= dcap) break;
            dcap = nd + nd / 4;  // too many records for the buffer: again with room
        }
        // the new bridges: the earliest record per cell, ordered by its time
        uint64_t nnew = 0;
        if (nd) {
            T[B_DKEY2].ensure(8 * nd);
            T[B_DTS2].ensure(8 * nd);
            DevBuf& tmp = T[B_TMP];
            size_t bytes = 0;
            // times are below 6 (nseed + nbr) + 7: only their low bits are sorted
            int ts_bits = 1;
            while (ts_bits < 64 && (1ull << ts_bits) <= 6 * (nseed + nbr) + 7) ++ts_bits;
            // by time, then (stably) by key: each key's records in time order
            TGVX_HIP(rocprim::radix_sort_pairs(nullptr, bytes, T[B_DTS].as<uint64_t>(), T[B_DTS2].as<uint64_t>(),
                                               T[B_DKEY].as<uint64_t>(), T[B_DKEY2].as<uint64_t>(), (size_t)nd, 0,
                                               ts_bits, st));
            tmp.ensure(bytes);
            TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, T[B_DTS].as<uint64_t>(), T[B_DTS2].as<uint64_t>(),
                                               T[B_DKEY].as<uint64_t>(), T[B_DKEY2].as<uint64_t>(), (size_t)nd, 0,
                                               ts_bits, st));
            bytes = 0;
            TGVX_HIP(rocprim::radix_sort_pairs(nullptr, bytes, T[B_DKEY2].as<uint64_t>(), T[B_DKEY].as<uint64_t>(),
                                               T[B_DTS2].as<uint64_t>(), T[B_DTS].as<uint64_t>(), (size_t)nd, 0,
                                               64, st));
            tmp.ensure(bytes);
            TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, T[B_DKEY2].as<uint64_t>(), T[B_DKEY].as<uint64_t>(),
                                               T[B_DTS2].as<uint64_t>(), T[B_DTS].as<uint64_t>(), (size_t)nd, 0,
                                               64, st));
            T[B_HEAD].ensure(nd);
            k_key_heads<<<grid_of(nd), 256, 0, st>>>(T[B_DKEY].as<uint64_t>(), nd, T[B_HEAD].as<uint8_t>());
            T[B_IDX2].ensure(4 * nd);
            nnew = select_flagged(c, T[B_HEAD].as<uint8_t>(), nd, T[B_IDX2].as<uint32_t>());
            T[B_NKEY].ensure(8 * nnew);
            T[B_NTS].ensure(8 * nnew);
            T[B_DKEY2].ensure(8 * nnew);
            T[B_DTS2].ensure(8 * nnew);
            k_gather<uint64_t><<<grid_of(nnew), 256, 0, st>>>(T[B_DKEY].as<uint64_t>(), T[B_IDX2].as<uint32_t>(), nnew,
                                                              T[B_DKEY2].as<uint64_t>());
            k_gather<uint64_t><<<grid_of(nnew), 256, 0, st>>>(T[B_DTS].as<uint64_t>(), T[B_IDX2].as<uint32_t>(), nnew,
                                                              T[B_DTS2].as<uint64_t>());
            bytes = 0;
            TGVX_HIP(rocprim::radix_sort_pairs(nullptr, bytes, T[B_DTS2].as<uint64_t>(), T[B_NTS].as<uint64_t>(),
                                               T[B_DKEY2].as<uint64_t>(), T[B_NKEY].as<uint64_t>(), (size_t)nnew, 0,
                                               ts_bits, st));
            tmp.ensure(bytes);
            TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, T[B_DTS2].as<uint64_t>(), T[B_NTS].as<uint64_t>(),
                                               T[B_DKEY2].as<uint64_t>(), T[B_NKEY].as<uint64_t>(), (size_t)nnew, 0,
                                               ts_bits, st));
        }
        // settled when the round reproduced the bridges it started from
        bool same = nnew == nbr;
        if (same && nnew) {
            TGVX_HIP(hipMemsetAsync(T[B_COUNT].p, 0, 4, st));
            k_differs<<<grid_of(nnew), 256, 0, st>>>(T[B_NKEY].as<uint64_t>(), T[B_PKEY].as<uint64_t>(),
                                                     T[B_NTS].as<uint64_t>(), T[B_PTS].as<uint64_t>(), nnew,
                                                     T[B_COUNT].as<uint32_t>());
            uint32_t diff = 0;
            TGVX_HIP(hipMemcpyAsync(&diff, T[B_COUNT].p, 4, hipMemcpyDeviceToHost, st));
            TGVX_HIP(hipStreamSynchronize(st));
            same = diff == 0;
        }
        if (same) break;
        std::swap(T[B_NKEY].p, T[B_PKEY].p);
        std::swap(T[B_NKEY].bytes, T[B_PKEY].bytes);
        std::swap(T[B_NTS].p, T[B_PTS].p);
        std::swap(T[B_NTS].bytes, T[B_PTS].bytes);
        nbr = nnew;
    }
    if (rounds) *rounds = r;
    // the cells, sorted by depth and then by code (both stable)
    const uint64_t n = nseed + nbr;
    T[B_CODES].ensure(8 * std::max<uint64_t>(n, 1));
    T[B_KEY32].ensure(4 * std::max<uint64_t>(n, 1));
    DevBuf cells_rc;
    cells_rc.ensure(4 * std::max<uint64_t>(n, 1));
    if (n)
        k_cells<<<grid_of(n), 256, 0, st>>>(code.as<uint64_t>(), depth.as<uint32_t>(), rc.as<float>(), nseed,
                                            T[B_PKEY].as<uint64_t>(), nbr, r_root, T[B_CODES].as<uint64_t>(),
                                            T[B_KEY32].as<uint32_t>(), cells_rc.as<float>());
    DevBuf dk, dk2, cc2, cc3, idx, idx2;
    dk.ensure(4 * std::max<uint64_t>(n, 1));
    dk2.ensure(4 * std::max<uint64_t>(n, 1));
    cc2.ensure(8 * std::max<uint64_t>(n, 1));
    idx.ensure(4 * std::max<uint64_t>(n, 1));
    idx2.ensure(4 * std::max<uint64_t>(n, 1));
    if (n) {
        k_iota<<<grid_of(n), 256, 0, st>>>(idx.as<uint32_t>(), n);
        TGVX_HIP(hipMemcpyAsync(dk.p, T[B_KEY32].p, 4 * n, hipMemcpyDeviceToDevice, st));
        sort_pairs<uint32_t>(c, dk, dk2, idx, idx2, n, 0, 5);  // idx2: by depth
        k_gather<uint64_t><<<grid_of(n), 256, 0, st>>>(T[B_CODES].as<uint64_t>(), idx2.as<uint32_t>(), n,
                                                       cc2.as<uint64_t>());
        cc3.ensure(8 * n);
        sort_pairs<uint64_t>(c, cc2, cc3, idx2, idx, n, low_bit, 63);  // idx: by (code, depth)
        c->oct_out.ensure(sizeof(tgvx_cube) * n);
        k_records<<<grid_of(n), 256, 0, st>>>(T[B_CODES].as<uint64_t>(), T[B_KEY32].as<uint32_t>(),
                                              cells_rc.as<float>(), idx.as<uint32_t>(), n,
                                              c->oct_out.as<tgvx_cube>());
        TGVX_HIP(hipGetLastError());
    }
    TGVX_HIP(hipStreamSynchronize(st));
    c->oct_n = n;
    return nbr;
}

// Stage 2 over n entries: idx lists them in tie order (ascending pixel or
// input index) over the arrays code / depth / rad. Sorts idx by (code,
// depth), stably; with merge, writes one cell per run of equal (code, depth)
// to mcode / mdepth / mrc (mean r_c, double sum in sorted order) and returns
// the run count, else returns n with idx sorted. Codes have their bits
// below low_bit zero.
uint64_t sort_merge(tgvx_ctx* c, const uint64_t* code, const uint32_t* depth, const float* rad, DevBuf& idx, uint64_t n,
                    int low_bit, bool merge, DevBuf& mcode, DevBuf& mdepth, DevBuf& mrc) {
    DevBuf* T = c->oct_tmp;
    hipStream_t st = c->stream;
    if (n == 0) return 0;
    DevBuf dk, dk2, idx2, cs, cs2, rs;  // rs: radii in sorted order
    dk.ensure(4 * n);
    dk2.ensure(4 * n);
    idx2.ensure(4 * n);
    cs.ensure(8 * n);
    cs2.ensure(8 * n);
    // by depth, then stably by code
    k_gather<uint32_t><<<grid_of(n), 256, 0, st>>>(depth, idx.as<uint32_t>(), n, dk.as<uint32_t>());
    sort_pairs<uint32_t>(c, dk, dk2, idx, idx2, n, 0, 5);
    k_gather<uint64_t><<<grid_of(n), 256, 0, st>>>(code, idx2.as<uint32_t>(), n, cs.as<uint64_t>());
    sort_pairs<uint64_t>(c, cs, cs2, idx2, idx, n, low_bit, 63);
    uint64_t nm = n;
    if (merge) {
        T[B_HEAD].ensure(n);
        k_run_heads<<<grid_of(n), 256, 0, st>>>(cs2.as<uint64_t>(), depth, idx.as<uint32_t>(), n,
                                                T[B_HEAD].as<uint8_t>());
        T[B_START].ensure(4 * n);
        nm = select_flagged(c, T[B_HEAD].as<uint8_t>(), n, T[B_START].as<uint32_t>());
        mcode.ensure(8 * nm);
        mdepth.ensure(4 * nm);
        mrc.ensure(4 * nm);
        rs.ensure(4 * n);
        k_gather<float><<<grid_of(n), 256, 0, st>>>(rad, idx.as<uint32_t>(), n, rs.as<float>());
        k_merge_runs<<<grid_of(nm), 256, 0, st>>>(cs2.as<uint64_t>(), depth, rs.as<float>(), idx.as<uint32_t>(),
                                                  T[B_START].as<uint32_t>(), nm, n, mcode.as<uint64_t>(),
                                                  mdepth.as<uint32_t>(), mrc.as<float>());
    }
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(st));  // before the scratch buffers go
    return nm;
}

// cells (code, depth, r_c) listed by idx -> c->oct_out
void write_records(tgvx_ctx* c, const uint64_t* code, const uint32_t* depth, const float* rc, const uint32_t* idx,
                   uint64_t n) {
    c->oct_n = n;
    if (n == 0) return;
    c->oct_out.ensure(sizeof(tgvx_cube) * n);
    k_records<<<grid_of(n), 256, 0, c->stream>>>(code, depth, rc, idx, n, c->oct_out.as<tgvx_cube>());
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(c->stream));
}

// host cubes -> device code / depth / r_c arrays; returns the lowest set bit
// over the codes (63 when all are zero)
int upload_cubes(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, DevBuf& dcode, DevBuf& ddepth, DevBuf& drc) {
    std::vector<uint64_t> code(n);
    std::vector<uint32_t> depth(n);
    std::vector<float> rc(n);
    int low_bit = 63;
    for (uint64_t i = 0; i < n; ++i) {
        code[i] = ((uint64_t)cubes[i].hi << 32) | cubes[i].lo;
        depth[i] = (uint32_t)cubes[i].depth;
        rc[i] = cubes[i].r_c;
        if (code[i]) low_bit = std::min(low_bit, __builtin_ctzll(code[i]));
    }
    dcode.ensure(8 * std::max<uint64_t>(n, 1));
    ddepth.ensure(4 * std::max<uint64_t>(n, 1));
    drc.ensure(4 * std::max<uint64_t>(n, 1));
    if (n) {
        TGVX_HIP(hipMemcpyAsync(dcode.p, code.data(), 8 * n, hipMemcpyHostToDevice, c->stream));
        TGVX_HIP(hipMemcpyAsync(ddepth.p, depth.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
        TGVX_HIP(hipMemcpyAsync(drc.p, rc.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
    }
    TGVX_HIP(hipStreamSynchronize(c->stream));
    return low_bit;
}

}  // namespace

void octree_build(tgvx_ctx* c, const tgvx_depth_view* views, int32_t nviews, const float origin[3], float r_root,
                  int32_t max_depth, int32_t last_stage, uint64_t* n_out, tgvx_octree_report* rep) {
    if (nviews < 0 || (nviews > 0 && !views) || !origin) throw Error(TGVX_E_ARG, "null pointer");
    if (max_depth < 0 || max_depth > 21) throw Error(TGVX_E_ARG, "max_depth must be in 0..21");
    if (last_stage < 1 || last_stage > 3) throw Error(TGVX_E_ARG, "last_stage must be 1, 2 or 3");
    if (!(r_root > 0.0f)) throw Error(TGVX_E_ARG, "r_root must be positive");
    DevBuf* T = c->oct_tmp;
    hipStream_t st = c->stream;
    tgvx_octree_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = ms_now();
    uint64_t total = 0, maxpix = 0;
    for (int k = 0; k < nviews; ++k) {
        if (!views[k].depths || views[k].width <= 0 || views[k].height <= 0)
            throw Error(TGVX_E_ARG, "view " + std::to_string(k) + ": no depth map");
        const uint64_t p = (uint64_t)views[k].width * views[k].height;
        total += p;
        maxpix = std::max(maxpix, p);
    }
    if (total >= (1ull << 32)) throw Error(TGVX_E_ARG, "at most 2^32 - 1 pixels");
    r.pixels = total;
    // depth maps in, one cube per pixel
    DevBuf dmaps;
    dmaps.ensure(4 * std::max<uint64_t>(total, 1));
    std::vector<HostPiece> pieces(nviews);
    for (int k = 0; k < nviews; ++k)
        pieces[k] = HostPiece{views[k].depths, 4 * (size_t)views[k].width * views[k].height};
    h2dv(c, dmaps.p, pieces.data(), nviews, st);
    uint64_t off = 0;
    TGVX_HIP(hipStreamSynchronize(st));
    const double t1 = ms_now();
    T[B_CODE].ensure(8 * std::max<uint64_t>(total, 1));
    T[B_DEPTH].ensure(4 * std::max<uint64_t>(total, 1));
    T[B_RAD].ensure(4 * std::max<uint64_t>(total, 1));
    T[B_FLAG].ensure(std::max<uint64_t>(total, 1));
    const float inv_2r = 1.0f / (2.0f * r_root);
    off = 0;
    for (int k = 0; k < nviews; ++k) {
        PixView v;
        v.d = dmaps.as<float>() + off;
        v.W = views[k].width;
        v.H = views[k].height;
        v.fx = views[k].fx;
        v.fy = views[k].fy;
        v.cx = views[k].cx;
        v.cy = views[k].cy;
        std::memcpy(v.M, views[k].cam_to_world, sizeof(v.M));
        const uint64_t p = (uint64_t)v.W * v.H;
        k_pix_cubes<<<grid_of(p), 256, 0, st>>>(v, (long)off, origin[0], origin[1], origin[2], inv_2r, r_root,
                                                max_depth, T[B_CODE].as<uint64_t>(), T[B_DEPTH].as<uint32_t>(),
                                                T[B_RAD].as<float>(), T[B_FLAG].as<uint8_t>());
        off += p;
    }
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(st));  // the maps are read until here
    dmaps.release();
    T[B_IDX].ensure(4 * std::max<uint64_t>(total, 1));
    const uint64_t nraw = select_flagged(c, T[B_FLAG].as<uint8_t>(), total, T[B_IDX].as<uint32_t>());
    r.raw = nraw;
    const double t2 = ms_now();
    // per-pixel codes are aligned to their depth (<= max_depth); so is a
    // bridge, an ancestor of a neighbour at a coarser depth
    const int low_bit = 3 * (21 - max_depth);
    DevBuf mcode, mdepth, mrc;
    const uint64_t nm = sort_merge(c, T[B_CODE].as<uint64_t>(), T[B_DEPTH].as<uint32_t>(), T[B_RAD].as<float>(),
                                   T[B_IDX], nraw, low_bit, last_stage >= 2, mcode, mdepth, mrc);
    if (last_stage == 1)
        write_records(c, T[B_CODE].as<uint64_t>(), T[B_DEPTH].as<uint32_t>(), T[B_RAD].as<float>(),
                      T[B_IDX].as<uint32_t>(), nraw);
    r.merged = last_stage >= 2 ? nm : 0;
    const double t3 = ms_now();
    if (last_stage == 2) {
        DevBuf id;
        id.ensure(4 * std::max<uint64_t>(nm, 1));
        if (nm) k_iota<<<grid_of(nm), 256, 0, st>>>(id.as<uint32_t>(), nm);
        write_records(c, mcode.as<uint64_t>(), mdepth.as<uint32_t>(), mrc.as<float>(), id.as<uint32_t>(), nm);
    } else if (last_stage == 3) {
        mcode.ensure(8);
        mdepth.ensure(4);
        mrc.ensure(4);
        r.bridges = balance_device(c, mcode, mdepth, mrc, nm, r_root, low_bit, &r.rounds);
    }
    const double t4 = ms_now();
    r.upload_ms = t1 - t0;
    r.cubes_ms = t2 - t1;
    r.merge_ms = t3 - t2;
    r.balance_ms = t4 - t3;
    r.total_ms = t4 - t0;
    if (n_out) *n_out = c->oct_n;
    if (rep) *rep = r;
}

void octree_balance(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, float r_root, uint64_t* n_out,
                    tgvx_octree_report* rep) {
    if (n && !cubes) throw Error(TGVX_E_ARG, "null pointer");
    if (n >= (1ull << 32)) throw Error(TGVX_E_ARG, "at most 2^32 - 1 cubes");
    for (uint64_t i = 0; i < n; ++i) {
        if (cubes[i].depth < 0 || cubes[i].depth > 21) throw Error(TGVX_E_ARG, "cube depth outside 0..21");
        // a cube's code is its cell's lower corner (octree_builder.cpp:70-71):
        // the bits below the cell size are zero, which the cell keys rely on
        const uint64_t cv = ((uint64_t)cubes[i].hi << 32) | cubes[i].lo;
        if (cv & ((1ull << (3 * (21 - cubes[i].depth))) - 1ull))
            throw Error(TGVX_E_ARG, "cube " + std::to_string(i) + ": code has bits below its cell size");
    }
    int max_depth = 0;
    for (uint64_t i = 0; i < n; ++i) max_depth = std::max(max_depth, (int)cubes[i].depth);
    tgvx_octree_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = ms_now();
    DevBuf dcode, ddepth, drc;
    // the sort's lowest key bit: every seed code's lowest set bit, and the
    // alignment of a bridge (one level above the deepest seed at most)
    int low_bit = upload_cubes(c, cubes, n, dcode, ddepth, drc);
    const double t1 = ms_now();
    r.merged = n;
    low_bit = std::min(low_bit, 3 * (21 - max_depth));
    r.bridges = balance_device(c, dcode, ddepth, drc, n, r_root, low_bit, &r.rounds);
    const double t2 = ms_now();
    r.upload_ms = t1 - t0;
    r.balance_ms = t2 - t1;
    r.total_ms = t2 - t0;
    if (n_out) *n_out = c->oct_n;
    if (rep) *rep = r;
}

void octree_merge(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, uint64_t* n_out, tgvx_octree_report* rep) {
    if (n && !cubes) throw Error(TGVX_E_ARG, "null pointer");
    if (n >= (1ull << 32)) throw Error(TGVX_E_ARG, "at most 2^32 - 1 cubes");
    for (uint64_t i = 0; i < n; ++i)
        if (cubes[i].depth < 0 || cubes[i].depth > 21) throw Error(TGVX_E_ARG, "cube depth outside 0..21");
    tgvx_octree_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = ms_now();
    DevBuf dcode, ddepth, drc, idx, mcode, mdepth, mrc, id;
    const int low_bit = upload_cubes(c, cubes, n, dcode, ddepth, drc);
    const double t1 = ms_now();
    idx.ensure(4 * std::max<uint64_t>(n, 1));
    if (n) k_iota<<<grid_of(n), 256, 0, c->stream>>>(idx.as<uint32_t>(), n);  // ties: input order
    const uint64_t nm = sort_merge(c, dcode.as<uint64_t>(), ddepth.as<uint32_t>(), drc.as<float>(), idx, n, low_bit,
                                   true, mcode, mdepth, mrc);
    id.ensure(4 * std::max<uint64_t>(nm, 1));
    if (nm) k_iota<<<grid_of(nm), 256, 0, c->stream>>>(id.as<uint32_t>(), nm);
    write_records(c, mcode.as<uint64_t>(), mdepth.as<uint32_t>(), mrc.as<float>(), id.as<uint32_t>(), nm);
    const double t2 = ms_now();
    r.raw = n;
    r.merged = nm;
    r.upload_ms = t1 - t0;
    r.merge_ms = t2 - t1;
    r.total_ms = t2 - t0;
    if (n_out) *n_out = c->oct_n;
    if (rep) *rep = r;
}

void octree_copy(tgvx_ctx* c, tgvx_cube* out) {
    if (c->oct_n && !out) throw Error(TGVX_E_ARG, "null pointer");
    if (c->oct_n)
        d2h(c, out, c->oct_out.p, sizeof(tgvx_cube) * c->oct_n);
}

}  // namespace tgvx
