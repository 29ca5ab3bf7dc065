// Device-side helpers shared by the tgvx kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tgvx {

// ---------------------------------------------------------------- Morton
// 21 bits per axis, bit b of x/y/z at bit 3b/3b+1/3b+2 of the 63-bit code
// (reference src/types.hpp:10-11,39-73).
__host__ __device__ inline uint64_t dilate3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
__host__ __device__ inline uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}
__host__ __device__ inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return dilate3(x) | (dilate3(y) << 1) | (dilate3(z) << 2);
}
// Ancestor at target depth: clear the low 3*(21-d) bits (balance.cpp:22-26).
__host__ __device__ inline uint64_t ancestor(uint64_t c, int d) {
    const unsigned s = 3u * (21u - (unsigned)d);
    return s >= 64u ? 0ull : (c & ~((1ull << s) - 1ull));
}
constexpr uint64_t kNone = ~0ull;
// Same-depth face neighbour, faces 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z; kNone
// outside [0, 2^21) (balance.cpp:32-49).
__host__ __device__ inline uint64_t face_nb(uint64_t c, int f, int d) {
    uint32_t a[3] = {compact3(c), compact3(c >> 1), compact3(c >> 2)};
    const uint32_t s = 1u << (21 - d);
    uint32_t& x = a[f >> 1];
    if (f & 1) {
        if (x < s) return kNone;
        x -= s;
    } else {
        if (x + s >= (1u << 21)) return kNone;
        x += s;
    }
    return morton(a[0], a[1], a[2]);
}

// ---------------------------------------------------------------- storage
// State arrays are stored as S and computed in T. BF16 storage rounds to
// nearest-even on every store.
struct bf16s {
    uint16_t x;
};

template <class S>
struct Io;
template <>
struct Io<float> {
    using T = float;
    __device__ static inline T ld(const float* p, long i) { return p[i]; }
    __device__ static inline void st(float* p, long i, T v) { p[i] = v; }
    __device__ static inline void stn(float* p, long i, T v) { __builtin_nontemporal_store(v, p + i); }
    // record r of an array of 4-value records (16-byte aligned)
    __device__ static inline void ld4(const float* p, long r, T o[4]) {
        const float4 x = reinterpret_cast<const float4*>(p)[r];
        o[0] = x.x, o[1] = x.y, o[2] = x.z, o[3] = x.w;
    }
    __device__ static inline T rt(T v) { return v; }  // value after a store / load round trip
    __device__ static inline T rtd(T v) { return v; }  // the same for a dual field (p, q)
};
template <>
struct Io<double> {
    using T = double;
    __device__ static inline T ld(const double* p, long i) { return p[i]; }
    __device__ static inline void st(double* p, long i, T v) { p[i] = v; }
    __device__ static inline void stn(double* p, long i, T v) { __builtin_nontemporal_store(v, p + i); }
    __device__ static inline void ld4(const double* p, long r, T o[4]) {
        const double2 a = reinterpret_cast<const double2*>(p)[2 * r], b = reinterpret_cast<const double2*>(p)[2 * r + 1];
        o[0] = a.x, o[1] = a.y, o[2] = b.x, o[3] = b.y;
    }
    __device__ static inline T rt(T v) { return v; }
    __device__ static inline T rtd(T v) { return v; }
};
template <>
struct Io<bf16s> {
    using T = float;
    __device__ static inline T ld(const bf16s* p, long i) {
        return __uint_as_float((uint32_t)p[i].x << 16);
    }
    __device__ static inline void st(bf16s* p, long i, T v) {
        const uint32_t u = __float_as_uint(v);
        uint16_t h;
        if ((u & 0x7fffffffu) > 0x7f800000u) h = (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
        else h = (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
        p[i].x = h;
    }
    __device__ static inline void stn(bf16s* p, long i, T v) { st(p, i, v); }
    __device__ static inline void ld4(const bf16s* p, long r, T o[4]) {
        const uint2 x = reinterpret_cast<const uint2*>(p)[r];
        o[0] = __uint_as_float(x.x << 16), o[1] = __uint_as_float(x.x & 0xffff0000u);
        o[2] = __uint_as_float(x.y << 16), o[3] = __uint_as_float(x.y & 0xffff0000u);
    }
    __device__ static inline T rt(T v) {
        bf16s b;
        st(&b, 0, v);
        return ld(&b, 0);
    }
    __device__ static inline T rtd(T v) { return rt(v); }
};

// Mixed storage: u, v and u_bar, v_bar in FP32, p, q in BF16, arithmetic in
// FP32. The arrays of the dual fields are marked by the low bit of their
// address (host side: dual_ptr), which every access tests; the pointers are
// kernel arguments, so the branch is uniform across a wave.
struct mixs {
    float x;
};
template <>
struct Io<mixs> {
    using T = float;
    __device__ static inline bool dual(const mixs* p) { return ((uintptr_t)p & 1u) != 0u; }
    __device__ static inline bf16s* half(const mixs* p) { return (bf16s*)((uintptr_t)p & ~(uintptr_t)1); }
    __device__ static inline T ld(const mixs* p, long i) {
        return dual(p) ? Io<bf16s>::ld(half(p), i) : reinterpret_cast<const float*>(p)[i];
    }
    __device__ static inline void st(mixs* p, long i, T v) {
        if (dual(p)) Io<bf16s>::st(half(p), i, v);
        else reinterpret_cast<float*>(p)[i] = v;
    }
    __device__ static inline void stn(mixs* p, long i, T v) {
        if (dual(p)) Io<bf16s>::st(half(p), i, v);
        else __builtin_nontemporal_store(v, reinterpret_cast<float*>(p) + i);
    }
    // records of the primal bars only (FP32)
    __device__ static inline void ld4(const mixs* p, long r, T o[4]) {
        const float4 x = reinterpret_cast<const float4*>(p)[r];
        o[0] = x.x, o[1] = x.y, o[2] = x.z, o[3] = x.w;
    }
    __device__ static inline T rt(T v) { return v; }
    __device__ static inline T rtd(T v) { return Io<bf16s>::rt(v); }
};

// Per-cube constant word: the vote total W (bits 0-23), the weighted-median
// bin k (bits 24-28) and "the depth-aware +x / +y / +z neighbour exists"
// (bits 29-31).
struct InfoWord {
    static constexpr uint32_t kWMax = (1u << 24) - 1;
    __host__ __device__ static uint32_t W(uint32_t w) { return w & kWMax; }
    __host__ __device__ static int bin(uint32_t w) { return (int)((w >> 24) & 31u); }
    __host__ __device__ static bool has(uint32_t w, int ax) { return (w >> (29 + ax)) & 1u; }
};

// median_indicator (histogram.cpp:21-31) with nb bins, from the median bin:
// the centre -1 + (k + 0.5) / (nb / 2) of bin k, 0 for an empty histogram.
template <class T>
__host__ __device__ inline T median_of(uint32_t w, int nb) {
    if (InfoWord::W(w) == 0u) return T(0);
    return T(-1) + (T(InfoWord::bin(w)) + T(0.5)) / (T(0.5) * T(nb));
}

// ---------------------------------------------------------------- launch
// Workgroups are dispatched to the 8 XCDs round-robin by id; remap so that
// each XCD gets one contiguous run of tiles (neighbouring leaves share its L2).
__device__ inline int xcd_tile() {
    const int nwg = (int)gridDim.x, b = (int)blockIdx.x;
    const int per = nwg >> 3, rem = nwg & 7, x = b & 7, k = b >> 3;
    return x < rem ? x * (per + 1) + k : rem * (per + 1) + (x - rem) * per + k;
}

}  // namespace tgvx
