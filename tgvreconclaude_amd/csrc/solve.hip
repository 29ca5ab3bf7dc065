// Primal-dual TGV iteration over the level-major permuted octree
// (reference src/tgv.cpp:137-232 run_level, :241-352 run_level_frozen,
// :356-511 tgv_minimize / tgv_minimize_oc), energy / duality gap in FP64, and
// the host orchestration of the coarse-to-fine level loop.
//
// Layout in device memory (n cubes, positions g = 0..n-1, level-major):
//   17 state arrays of S (u, v[3], p[3], q[6], u_bar, v_bar[3]), SoA;
//   6 int32 neighbour arrays (face f of position g -> position or -1);
//   W (uint32 vote total), med (T median indicator), parent (int32);
//   incoming-edge CSR of each level (for the exact-adjoint mode and the gap).
// One reference iteration = five Jacobi passes; they fuse into two kernels
// with one device-wide dependency between them:
//   k_dual   = passes 1-2 (p, q ascent + ball projections), reads u_bar, v_bar;
//   k_primal = passes 3-5 (u prox, v step, extrapolation), reads the new p, q.
// Every expression keeps the reference's evaluation order; the library is
// compiled with -ffp-contract=off, and HIP's division and square root are
// correctly rounded, so FP32 / FP64 results equal the CPU definition bit for bit.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <numeric>
#include <set>
#include <vector>

#include "tgvx_device.h"
#include "tgvx_internal.h"

namespace tgvx {

enum { F_U = 0, F_VX, F_VY, F_VZ, F_PX, F_PY, F_PZ, F_Q0, F_Q1, F_Q2, F_Q3, F_Q4, F_Q5, F_UB, F_VBX, F_VBY, F_VBZ };

template <class S>
struct KArgs {
    using T = typename Io<S>::T;
    S* s[17];
    const int32_t* nbp[3];    // +x, +y, +z neighbours read by the dual pass
    const int32_t* nbm[3];    // -x, -y, -z neighbours (reference primal pass)
    const int32_t* nba[3];    // depth-aware +x, +y, +z neighbours (exact-adjoint mode, gap)
    const int32_t* nbt[3];    // the level's cube whose depth-aware +a neighbour this is, or -1
    const uint32_t* info;     // packed per-cube word, see InfoWord
    int32_t nbins;
    const int32_t* parent;
    uint32_t off, n;          // positions iterated by this launch
    uint32_t lvl_off, lvl_n;  // the level's positions (active set for the gap)
    T sigma, tau, a0, a1, tl;
    // frozen-border mode (run_level_frozen): out-of-leaf u_bar / v_bar come from fz[]
    const S* fz[4];
    const uint32_t* orig;  // position -> input index
    uint32_t leaf_first, leaf_last;
    const float* init;     // initial field: init_stride values per position (4: u, v; 13: u, v, p, q)
    int32_t init_stride;
    S* save[4];
    // several blocks iterated to a gap: a cube is skipped while its block's
    // byte in bact is 0 (the block has converged); null: every cube runs
    const int32_t* blk;
    const uint8_t* bact;
    // streaming mode (tgv_minimize_streaming): per cube, bit f of face f set
    // for the +a faces (0, 2, 4) when that neighbour is iterated in the same
    // treetop leaf (else its level-start snapshot in fz is read), and for the
    // -a faces (1, 3, 5) when that neighbour lies outside the leaf (its p, q
    // read as 0); bit 6: the cube belongs to a leaf. Null otherwise.
    const uint8_t* sflag;
    // the previous level's u_bar, v_bar as records of 4 (positions
    // [fz_off, fz_off + fz_n)), for the fused kernel's frozen reads; or null
    const S* fzaos;
    uint32_t fz_off, fz_n;
    // reference rules, two-iteration kernel: per level cube t (position
    // lvl_off + t) and axis ax, record ax * lvl_n + t holds, where the
    // reference table's +a neighbour lies outside the level, its u_bar, v_bar
    // (rfz_bar) and, where the -a neighbour does, the p, q the primal step
    // reads of it (rfz_dual: kF[ax]); both constant while the level iterates
    const S* rfz_bar;
    const S* rfz_dual;
};

template <class S>
__device__ inline bool masked(const KArgs<S>& a, long g) {
    return a.bact && !a.bact[a.blk[g]];
}

// Workgroup -> tile: XCD-contiguous runs of tiles (xcd_tile) or plain order.
template <bool XCD>
__device__ inline int tile_id() {
    return XCD ? xcd_tile() : (int)blockIdx.x;
}

template <class T>
__device__ inline void clip3(T& a, T& b, T& c, T r) {
    const T n2 = a * a + b * b + c * c;
    if (n2 > r * r) {
        const T s = r / sqrt(n2);
        a *= s;
        b *= s;
        c *= s;
    }
}
template <class T>
__device__ inline void clip6(T* q, T r) {
    T n2 = T(0);
    for (int i = 0; i < 6; ++i) n2 += q[i] * q[i];
    if (n2 > r * r) {
        const T s = r / sqrt(n2);
        for (int i = 0; i < 6; ++i) q[i] *= s;
    }
}
// prox_data (tgv.cpp:75-94): shrink towards the weighted median by
// tau*lambda*W; the median depends only on the histogram, so it comes from
// the packed word instead of a scan of the bins every iteration.
// The same with the weighted median given (median_of's value for w).
template <class T>
__device__ inline T prox_m(T z, uint32_t w, T med, T tl) {
    const uint32_t W = InfoWord::W(w);
    if (W == 0u) return z;
    const T diff = z - med;
    const T shrink = tl * (T)W;
    if (diff > shrink) return z - shrink;
    if (diff < -shrink) return z + shrink;
    return med;
}
template <class T>
__device__ inline T prox(T z, uint32_t w, int nb, T tl) {
    const uint32_t W = InfoWord::W(w);
    if (W == 0u) return z;
    const T med = median_of<T>(w, nb);
    const T diff = z - med;
    const T shrink = tl * (T)W;
    if (diff > shrink) return z - shrink;
    if (diff < -shrink) return z + shrink;
    return med;
}

// Passes 1-2 of one cube (tgv.cpp:159-189): p += sigma (grad u_bar - v_bar),
// q += sigma eps(v_bar), ball projections. bar[0] = this cube's (u_bar,
// v_bar), bar[1 + a] = its +a neighbour's (or its own when there is none).
template <class T>
__device__ inline void dual_step(T p[3], T q[6], const T bar[4][4], T sigma, T a0, T a1) {
    const T ubi = bar[0][0], vxi = bar[0][1], vyi = bar[0][2], vzi = bar[0][3];
    p[0] += sigma * (bar[1][0] - ubi - vxi);
    p[1] += sigma * (bar[2][0] - ubi - vyi);
    p[2] += sigma * (bar[3][0] - ubi - vzi);
    clip3(p[0], p[1], p[2], a1);
    const T x0 = bar[1][1], y0 = bar[1][2], z0 = bar[1][3];
    const T x2 = bar[2][1], y2 = bar[2][2], z2 = bar[2][3];
    const T x4 = bar[3][1], y4 = bar[3][2], z4 = bar[3][3];
    const T eps[6] = {x0 - vxi,
                      y2 - vyi,
                      z4 - vzi,
                      T(0.5) * (x2 - vxi + y0 - vyi),
                      T(0.5) * (x4 - vxi + z0 - vzi),
                      T(0.5) * (y4 - vyi + z2 - vzi)};
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = q[k] + sigma * eps[k];
    clip6(q, a0);
}

// ---- phase D: dual p and q (tgv.cpp:158-189; frozen variant :274-309;
// streaming variant :738-767). MODE 0: every neighbour live; 1 (frozen
// borders): live inside the treetop leaf [leaf_first, leaf_last]; 2
// (streaming): live as sflag says. A neighbour that is not live gives its
// level-start snapshot fz.
template <class S, int MODE, bool XCD = true>
__global__ void __launch_bounds__(256) k_dual(KArgs<S> a) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = (uint32_t)tile_id<XCD>() * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    if (masked(a, g)) return;
    const uint32_t sf = MODE == 2 ? a.sflag[g] : 0u;
    if (MODE == 2 && !(sf & 64u)) return;  // in no leaf: never iterated
    const T ubi = IO::ld(a.s[F_UB], g), vxi = IO::ld(a.s[F_VBX], g), vyi = IO::ld(a.s[F_VBY], g),
            vzi = IO::ld(a.s[F_VBZ], g);
    const int j0 = a.nbp[0][g], j2 = a.nbp[1][g], j4 = a.nbp[2][g];
    auto live = [&](int j, int ax) -> bool {
        if (MODE == 0) return true;
        if (MODE == 2) return (sf >> (2 * ax)) & 1u;
        const uint32_t o = a.orig[j];
        return o >= a.leaf_first && o <= a.leaf_last;
    };
    // A missing neighbour reads the cube's own value (Neumann, tgv.cpp:163):
    // loading at g itself gives exactly that, so every load is unconditional.
    const long s0 = j0 >= 0 ? j0 : g, s2 = j2 >= 0 ? j2 : g, s4 = j4 >= 0 ? j4 : g;
    auto rd = [&](int j, long sj, int f, int fzi, T self, int ax) -> T {
        if (MODE != 0) {
            if (j < 0) return self;
            if (!live(j, ax)) return IO::ld(a.fz[fzi], j);
        }
        return IO::ld(a.s[f], sj);
    };
    // bars: [0] this cube, [1..3] its +x, +y, +z neighbours; each (u, vx, vy, vz)
    T bar[4][4] = {{ubi, vxi, vyi, vzi}};
    const int js[3] = {j0, j2, j4};
    const long ss[3] = {s0, s2, s4};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int c = 0; c < 4; ++c) bar[1 + ax][c] = rd(js[ax], ss[ax], F_UB + c, c, bar[0][c], ax);
    T p[3] = {IO::ld(a.s[F_PX], g), IO::ld(a.s[F_PY], g), IO::ld(a.s[F_PZ], g)};
    T q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = IO::ld(a.s[F_Q0 + k], g);
    dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::st(a.s[F_PX + c], g, p[c]);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::st(a.s[F_Q0 + k], g, q[k]);
}

// ---- phase P, reference rules: primal u, primal v, extrapolation
// (tgv.cpp:191-230).
template <class S, bool XCD = true>
__global__ void __launch_bounds__(256) k_primal_ref(KArgs<S> a) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = (uint32_t)tile_id<XCD>() * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    if (masked(a, g)) return;
    int m0 = a.nbm[0][g], m2 = a.nbm[1][g], m4 = a.nbm[2][g];
    if (a.sflag) {
        // streaming: p, q of a cube outside the leaf count as 0, as a missing one
        const uint32_t sf = a.sflag[g];
        if (!(sf & 64u)) return;
        m0 = (sf & 2u) ? -1 : m0;
        m2 = (sf & 8u) ? -1 : m2;
        m4 = (sf & 32u) ? -1 : m4;
    }
    // A missing neighbour contributes 0 (tgv.cpp:196,209). Loads go to g
    // instead and the value is replaced by 0 afterwards, so that every load
    // is unconditional and they can all be in flight together.
    const long s0 = m0 >= 0 ? m0 : g, s2 = m2 >= 0 ? m2 : g, s4 = m4 >= 0 ? m4 : g;
    auto nx = [&](int f) -> T { const T x = IO::ld(a.s[f], s0); return m0 >= 0 ? x : T(0); };
    auto ny = [&](int f) -> T { const T x = IO::ld(a.s[f], s2); return m2 >= 0 ? x : T(0); };
    auto nz = [&](int f) -> T { const T x = IO::ld(a.s[f], s4); return m4 >= 0 ? x : T(0); };
    const T px = IO::ld(a.s[F_PX], g), py = IO::ld(a.s[F_PY], g), pz = IO::ld(a.s[F_PZ], g);
    const T div_p = (px - nx(F_PX)) + (py - ny(F_PY)) + (pz - nz(F_PZ));
    const T u = IO::ld(a.s[F_U], g);
    const T un = prox(u + a.tau * div_p, a.info[g], a.nbins, a.tl);

    const T q0 = IO::ld(a.s[F_Q0], g), q1 = IO::ld(a.s[F_Q1], g), q2 = IO::ld(a.s[F_Q2], g),
            q3 = IO::ld(a.s[F_Q3], g), q4 = IO::ld(a.s[F_Q4], g), q5 = IO::ld(a.s[F_Q5], g);
    const T dqx = (q0 - nx(F_Q0)) + (q3 - ny(F_Q3)) + (q4 - nz(F_Q4));
    const T dqy = (q3 - nx(F_Q3)) + (q1 - ny(F_Q1)) + (q5 - nz(F_Q5));
    const T dqz = (q4 - nx(F_Q4)) + (q5 - ny(F_Q5)) + (q2 - nz(F_Q2));
    const T vx = IO::ld(a.s[F_VX], g), vy = IO::ld(a.s[F_VY], g), vz = IO::ld(a.s[F_VZ], g);
    const T vxn = vx + a.tau * (px + dqx);
    const T vyn = vy + a.tau * (py + dqy);
    const T vzn = vz + a.tau * (pz + dqz);
    IO::st(a.s[F_UB], g, T(2) * un - u);
    IO::st(a.s[F_VBX], g, T(2) * vxn - vx);
    IO::st(a.s[F_VBY], g, T(2) * vyn - vy);
    IO::st(a.s[F_VBZ], g, T(2) * vzn - vz);
    IO::st(a.s[F_U], g, un);
    IO::st(a.s[F_VX], g, vxn);
    IO::st(a.s[F_VY], g, vyn);
    IO::st(a.s[F_VZ], g, vzn);
}

// q component paired with (v component c, axis a) in eps: the diagonal entry
// for c == a, else xy / xz / yz for {c, a} = {0,1} / {0,2} / {1,2}.
__host__ __device__ constexpr int qidx(int c, int ax) {
    return c == ax ? F_Q0 + c : (c + ax == 1 ? F_Q3 : (c + ax == 2 ? F_Q4 : F_Q5));
}
// Field f + ax for a run-time axis, without indexing the argument array.
template <class S>
__device__ inline const S* field3(const KArgs<S>& a, int f0, int f1, int f2, int ax) {
    return ax == 0 ? a.s[f0] : (ax == 1 ? a.s[f1] : a.s[f2]);
}

// ---- phase P, exact adjoint (no reference counterpart): the primal step
// uses minus the transpose of the dual step's linear map: self terms only
// where the +a neighbour exists, incoming terms from every active cube whose
// +a neighbour this one is, and 1/2 on the off-diagonal eps entries.
// The 13 iterated values (u, v[3], p[3], q[6]) of every cube, as one array
// each; the fused iteration ping-pongs between two such sets.
template <class S>
struct Set13 {
    S* f[13];
};

// Passes 3-5 with the reference's rules (as k_primal_ref) reading (u, v, p, q)
// from src -- a cube of another depth from the main set, which the level
// never writes -- and writing the new u, v to dst and u_bar, v_bar to the main
// set; copy_dual also moves p, q from src to dst. Ends a run of
// reference-rule iterations that alternate between two sets. Same
// expressions, in the same order, as k_primal_ref.
template <class S, bool XCD = true>
__global__ void __launch_bounds__(256) k_primal_ref_set(KArgs<S> a, Set13<S> src, Set13<S> dst, int copy_dual) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = (uint32_t)tile_id<XCD>() * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    if (masked(a, g)) return;
    const int m0 = a.nbm[0][g], m2 = a.nbm[1][g], m4 = a.nbm[2][g];
    auto rd = [&](int f, int m) -> T {
        if (m < 0) return T(0);
        const bool live = (uint32_t)m - a.lvl_off < a.lvl_n;
        return live ? IO::ld(src.f[f], m) : IO::ld(a.s[f], m);
    };
    const T px = IO::ld(src.f[F_PX], g), py = IO::ld(src.f[F_PY], g), pz = IO::ld(src.f[F_PZ], g);
    const T div_p = (px - rd(F_PX, m0)) + (py - rd(F_PY, m2)) + (pz - rd(F_PZ, m4));
    const T u = IO::ld(src.f[F_U], g);
    const T un = prox(u + a.tau * div_p, a.info[g], a.nbins, a.tl);
    const T q0 = IO::ld(src.f[F_Q0], g), q1 = IO::ld(src.f[F_Q1], g), q2 = IO::ld(src.f[F_Q2], g),
            q3 = IO::ld(src.f[F_Q3], g), q4 = IO::ld(src.f[F_Q4], g), q5 = IO::ld(src.f[F_Q5], g);
    const T dqx = (q0 - rd(F_Q0, m0)) + (q3 - rd(F_Q3, m2)) + (q4 - rd(F_Q4, m4));
    const T dqy = (q3 - rd(F_Q3, m0)) + (q1 - rd(F_Q1, m2)) + (q5 - rd(F_Q5, m4));
    const T dqz = (q4 - rd(F_Q4, m0)) + (q5 - rd(F_Q5, m2)) + (q2 - rd(F_Q2, m4));
    const T vx = IO::ld(src.f[F_VX], g), vy = IO::ld(src.f[F_VY], g), vz = IO::ld(src.f[F_VZ], g);
    const T vxn = vx + a.tau * (px + dqx);
    const T vyn = vy + a.tau * (py + dqy);
    const T vzn = vz + a.tau * (pz + dqz);
    IO::st(a.s[F_UB], g, T(2) * un - u);
    IO::st(a.s[F_VBX], g, T(2) * vxn - vx);
    IO::st(a.s[F_VBY], g, T(2) * vyn - vy);
    IO::st(a.s[F_VBZ], g, T(2) * vzn - vz);
    IO::st(dst.f[F_U], g, un);
    IO::st(dst.f[F_VX], g, vxn);
    IO::st(dst.f[F_VY], g, vyn);
    IO::st(dst.f[F_VZ], g, vzn);
    if (copy_dual) {
        IO::st(dst.f[F_PX], g, px);
        IO::st(dst.f[F_PY], g, py);
        IO::st(dst.f[F_PZ], g, pz);
        IO::st(dst.f[F_Q0], g, q0);
        IO::st(dst.f[F_Q1], g, q1);
        IO::st(dst.f[F_Q2], g, q2);
        IO::st(dst.f[F_Q3], g, q3);
        IO::st(dst.f[F_Q4], g, q4);
        IO::st(dst.f[F_Q5], g, q5);
    }
}

// Passes 3-5 of cube g in the exact-adjoint mode, from the values in src:
// the new (u, v) in nw, the old ones in od.
// p_own / q_own receive the cube's own p, q from src.
template <class S>
__device__ inline void primal_adj_at(const KArgs<S>& a, const Set13<S>& src, long g, typename Io<S>::T nw[4],
                                     typename Io<S>::T od[4], typename Io<S>::T p_own[3],
                                     typename Io<S>::T q_own[6]) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t w = a.info[g];
    T p[3] = {IO::ld(src.f[F_PX], g), IO::ld(src.f[F_PY], g), IO::ld(src.f[F_PZ], g)};
    T dp = T(0);
    T dq[3] = {T(0), T(0), T(0)};
    const T qs[6] = {IO::ld(src.f[F_Q0], g), IO::ld(src.f[F_Q1], g), IO::ld(src.f[F_Q2], g),
                     IO::ld(src.f[F_Q3], g), IO::ld(src.f[F_Q4], g), IO::ld(src.f[F_Q5], g)};
    // self terms where the +a neighbour exists; a select keeps each sum
    // exactly as if the term had been skipped
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const bool h = InfoWord::has(w, ax);
        dp = h ? dp + p[ax] : dp;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T w2 = (c == ax) ? T(1) : T(0.5);
            dq[c] = h ? dq[c] + w2 * qs[qidx(c, ax) - F_Q0] : dq[c];
        }
    }
    // incoming edges, one per axis at most (the depth-aware table is one-to-one)
    const int k0 = a.nbt[0][g], k1 = a.nbt[1][g], k2 = a.nbt[2][g];
    const long s0 = k0 >= 0 ? k0 : g, s1 = k1 >= 0 ? k1 : g, s2 = k2 >= 0 ? k2 : g;
    const T p0 = IO::ld(src.f[F_PX], s0), a0 = IO::ld(src.f[F_Q0], s0), a3 = IO::ld(src.f[F_Q3], s0),
            a4 = IO::ld(src.f[F_Q4], s0);
    const T p1 = IO::ld(src.f[F_PY], s1), b3 = IO::ld(src.f[F_Q3], s1), b1 = IO::ld(src.f[F_Q1], s1),
            b5 = IO::ld(src.f[F_Q5], s1);
    const T p2 = IO::ld(src.f[F_PZ], s2), c4 = IO::ld(src.f[F_Q4], s2), c5 = IO::ld(src.f[F_Q5], s2),
            c2 = IO::ld(src.f[F_Q2], s2);
    if (k0 >= 0) {
        dp -= p0;
        dq[0] -= T(1) * a0;
        dq[1] -= T(0.5) * a3;
        dq[2] -= T(0.5) * a4;
    }
    if (k1 >= 0) {
        dp -= p1;
        dq[0] -= T(0.5) * b3;
        dq[1] -= T(1) * b1;
        dq[2] -= T(0.5) * b5;
    }
    if (k2 >= 0) {
        dp -= p2;
        dq[0] -= T(0.5) * c4;
        dq[1] -= T(0.5) * c5;
        dq[2] -= T(1) * c2;
    }
    const T u = IO::ld(src.f[F_U], g);
    od[0] = u;
    nw[0] = prox(u + a.tau * dp, w, a.nbins, a.tl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        od[1 + c] = IO::ld(src.f[F_VX + c], g);
        nw[1 + c] = od[1 + c] + a.tau * (p[c] + dq[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) p_own[c] = p[c];
#pragma unroll
    for (int k = 0; k < 6; ++k) q_own[k] = qs[k];
}
template <class S>
__device__ inline void primal_adj_at(const KArgs<S>& a, const Set13<S>& src, long g, typename Io<S>::T nw[4],
                                     typename Io<S>::T od[4]) {
    typename Io<S>::T p[3], q[6];
    primal_adj_at(a, src, g, nw, od, p, q);
}

// ---- phase P, exact adjoint (no reference counterpart): the primal step
// uses minus the transpose of the dual step's linear map: self terms only
// where the +a neighbour exists, incoming terms from the cube whose +a
// neighbour this one is, and 1/2 on the off-diagonal eps entries. Reads src,
// writes (u, v) to dst and u_bar, v_bar to the main set; copy_dual also
// moves p, q from src to dst.
template <class S, bool XCD = true>
__global__ void __launch_bounds__(256) k_primal_adj(KArgs<S> a, Set13<S> src, Set13<S> dst, int copy_dual) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = (uint32_t)tile_id<XCD>() * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    if (masked(a, g)) return;
    T nw[4], od[4];
    primal_adj_at(a, src, g, nw, od);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_UB + c], g, T(2) * nw[c] - od[c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(dst.f[F_U + c], g, nw[c]);
    if (copy_dual) {
#pragma unroll
        for (int f = F_PX; f <= F_Q5; ++f) IO::st(dst.f[f], g, IO::ld(src.f[f], g));
    }
}

// ---- one fused iteration, exact-adjoint mode (no reference counterpart):
// the primal passes of iteration n for this cube and, redundantly, for its
// +x, +y, +z neighbours, then the dual passes of iteration n + 1 for this
// cube. u_bar, v_bar never leave registers; src holds (u^n, v^n, p^n+1,
// q^n+1), dst receives (u^n+1, v^n+1, p^n+2, q^n+2). Every value is computed
// by the same expressions as k_primal_adj followed by k_dual, so the results
// are identical to the two-kernel iteration.
template <class S, bool XCD = true>
__global__ void __launch_bounds__(256) k_fused_adj(KArgs<S> a, Set13<S> src, Set13<S> dst) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = (uint32_t)tile_id<XCD>() * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    if (masked(a, g)) return;
    T nw[4], od[4];
    primal_adj_at(a, src, g, nw, od);
    T bar[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bar[0][c] = IO::rt(T(2) * nw[c] - od[c]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int j = a.nbp[ax][g];
        if (j < 0) {
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[1 + ax][c] = bar[0][c];
        } else if ((uint32_t)j >= a.lvl_off && (uint32_t)j < a.lvl_off + a.lvl_n) {
            T n2[4], o2[4];
            primal_adj_at(a, src, j, n2, o2);
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[1 + ax][c] = IO::rt(T(2) * n2[c] - o2[c]);
        } else {
            // a cube of another level: its frozen u_bar, v_bar
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[1 + ax][c] = IO::ld(a.s[F_UB + c], j);
        }
    }
    T p[3] = {IO::ld(src.f[F_PX], g), IO::ld(src.f[F_PY], g), IO::ld(src.f[F_PZ], g)};
    T q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = IO::ld(src.f[F_Q0 + k], g);
    dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(dst.f[F_U + c], g, nw[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::st(dst.f[F_PX + c], g, p[c]);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::st(dst.f[F_Q0 + k], g, q[k]);
}

// ---- one fused iteration over a tile of TILE consecutive positions
// (no reference counterpart). Same values as k_fused_adj, but the tile's
// p, q and its primal results are shared through LDS: the primal step reads
// the -a neighbours' p, q from LDS when they lie in the tile, the dual step
// reads the +a neighbours' u_bar, v_bar from LDS, and only the +a neighbours
// outside the tile (a few tenths per cube in Morton order) have their primal
// values recomputed, once per reference, by the whole workgroup.
constexpr int kFuseTile = 256;  // default tile of the tiled fused kernel

// Passes 3-5 of cube g (as primal_adj_at) with its own p, q given and the
// -a neighbours' p, q taken from the tile's copy in LDS where they are in
// [tile_lo, tile_hi). Same operations in the same order as primal_adj_at.
template <class S, int TILE>
__device__ inline void primal_tile(const KArgs<S>& a, const Set13<S>& src, long g, const typename Io<S>::T p[3],
                                   const typename Io<S>::T qs[6], const typename Io<S>::T (*lpq)[TILE], long tile_lo,
                                   long tile_hi, typename Io<S>::T nw[4], typename Io<S>::T od[4]) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t w = a.info[g];
    T dp = T(0);
    T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const bool h = InfoWord::has(w, ax);
        dp = h ? dp + p[ax] : dp;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const T w2 = (c == ax) ? T(1) : T(0.5);
            dq[c] = h ? dq[c] + w2 * qs[qidx(c, ax) - F_Q0] : dq[c];
        }
    }
    // incoming edge of axis ax: p_ax and the q entries of row ax of eps
    constexpr int kF[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
    const int ks[3] = {a.nbt[0][g], a.nbt[1][g], a.nbt[2][g]};
    T in[3][4];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const long k = ks[ax];
        if (k >= tile_lo && k < tile_hi) {
#pragma unroll
            for (int e = 0; e < 4; ++e) in[ax][e] = lpq[kF[ax][e] - F_PX][k - tile_lo];
        } else {
            const long sk = k >= 0 ? k : g;
#pragma unroll
            for (int e = 0; e < 4; ++e) in[ax][e] = IO::ld(src.f[kF[ax][e]], sk);
        }
    }
    if (ks[0] >= 0) {
        dp -= in[0][0];
        dq[0] -= T(1) * in[0][1];
        dq[1] -= T(0.5) * in[0][2];
        dq[2] -= T(0.5) * in[0][3];
    }
    if (ks[1] >= 0) {
        dp -= in[1][0];
        dq[0] -= T(0.5) * in[1][1];
        dq[1] -= T(1) * in[1][2];
        dq[2] -= T(0.5) * in[1][3];
    }
    if (ks[2] >= 0) {
        dp -= in[2][0];
        dq[0] -= T(0.5) * in[2][1];
        dq[1] -= T(0.5) * in[2][2];
        dq[2] -= T(1) * in[2][3];
    }
    const T u = IO::ld(src.f[F_U], g);
    od[0] = u;
    nw[0] = prox(u + a.tau * dp, w, a.nbins, a.tl);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        od[1 + c] = IO::ld(src.f[F_VX + c], g);
        nw[1 + c] = od[1 + c] + a.tau * (p[c] + dq[c]);
    }
}

// STAGE: the tile's p, q go through LDS first, so the primal step's incoming
// edges inside the tile are read from LDS (one more barrier); otherwise every
// cube's primal step reads them from memory. Which is faster depends on the
// level and the device: the autotuner decides.
template <class S, int TILE = kFuseTile, bool XCD = true, bool STAGE = false>
__global__ void __launch_bounds__(TILE) k_fused_tile(KArgs<S> a, Set13<S> src, Set13<S> dst) {
    using IO = Io<S>;
    using T = typename IO::T;
    // LDS: the tile's p, q (STAGE, phases 0-1), reused for the halo's bars
    // (phases 2-3)
    constexpr int kPQ = 9 * TILE, kHalo = 4 * 3 * TILE;
    __shared__ T shared_pq_halo[kPQ > kHalo ? kPQ : kHalo];
    __shared__ T own_bar[4][TILE];
    __shared__ int32_t halo_j[3 * TILE];
    __shared__ int32_t n_halo;
    T(*own_pq)[TILE] = reinterpret_cast<T(*)[TILE]>(shared_pq_halo);
    T(*halo_bar)[3 * TILE] = reinterpret_cast<T(*)[3 * TILE]>(shared_pq_halo);
    const uint32_t t0 = (uint32_t)tile_id<XCD>() * TILE;
    const uint32_t t = t0 + threadIdx.x;
    const long g = (long)a.off + t;
    const bool run = t < a.n && !masked(a, g);
    const long tile_lo = (long)a.off + t0;
    const long tile_hi = (long)a.off + min(t0 + (uint32_t)TILE, a.n);
    if (threadIdx.x == 0) n_halo = 0;
    T p[3], q[6], nw[4];
    if (STAGE) {
        // phase 0: the tile's p, q into LDS
        if (run) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                p[c] = IO::ld(src.f[F_PX + c], g);
                own_pq[c][threadIdx.x] = p[c];
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                q[k] = IO::ld(src.f[F_Q0 + k], g);
                own_pq[3 + k][threadIdx.x] = q[k];
            }
        }
        __syncthreads();
    }
    // phase 1: own primal values; where each +a neighbour's bar comes from:
    // >= 0 an LDS slot (own tile below TILE, halo above), -1 this cube (no
    // neighbour), -2 a cube of another level (its frozen u_bar, v_bar, read
    // here already)
    int32_t src_slot[3] = {-1, -1, -1};
    T frozen[3][4];
    if (run) {
        T od[4];
        if (STAGE) primal_tile<S, TILE>(a, src, g, p, q, own_pq, tile_lo, tile_hi, nw, od);
        else primal_adj_at(a, src, g, nw, od, p, q);
#pragma unroll
        for (int c = 0; c < 4; ++c) own_bar[c][threadIdx.x] = IO::rt(T(2) * nw[c] - od[c]);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int j = __builtin_nontemporal_load(&a.nbp[ax][g]);  // read once
            if (j < 0) {
                src_slot[ax] = -1;
            } else if (j >= tile_lo && j < tile_hi) {
                src_slot[ax] = (int32_t)(j - tile_lo);
            } else if ((uint32_t)j >= a.lvl_off && (uint32_t)j < a.lvl_off + a.lvl_n) {
                const int h = atomicAdd(&n_halo, 1);
                halo_j[h] = j;
                src_slot[ax] = TILE + h;
            } else {
                src_slot[ax] = -2;
                if (a.fzaos && (uint32_t)j - a.fz_off < a.fz_n) {
                    IO::ld4(a.fzaos, (long)((uint32_t)j - a.fz_off), frozen[ax]);
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c) frozen[ax][c] = IO::ld(a.s[F_UB + c], j);
                }
            }
        }
    }
    __syncthreads();
    // phase 2: primal values of the out-of-tile neighbours (LDS p, q are dead)
    const int nh = n_halo;
    for (int h = threadIdx.x; h < nh; h += TILE) {
        T n2[4], o2[4];
        primal_adj_at(a, src, (long)halo_j[h], n2, o2);
#pragma unroll
        for (int c = 0; c < 4; ++c) halo_bar[c][h] = IO::rt(T(2) * n2[c] - o2[c]);
    }
    __syncthreads();
    // phase 3: dual step
    if (!run) return;
    T bar[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bar[0][c] = own_bar[c][threadIdx.x];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int sl = src_slot[ax];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            T v;
            if (sl == -1) v = bar[0][c];
            else if (sl == -2) v = frozen[ax][c];
            else if (sl < TILE) v = own_bar[c][sl];
            else v = halo_bar[c][sl - TILE];
            bar[1 + ax][c] = v;
        }
    }
    dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
    // streaming stores: nothing reads dst in this launch, so keep L2 for the
    // gathered neighbour values
    for (int c = 0; c < 4; ++c) IO::stn(dst.f[F_U + c], g, nw[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::stn(dst.f[F_PX + c], g, p[c]);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::stn(dst.f[F_Q0 + k], g, q[k]);
}

// ---- fused iteration with a precomputed halo (no reference counterpart).
// The neighbour structure of a level never changes during its iterations, so
// where each cube's +a neighbour's u_bar, v_bar come from, and which cubes of
// the level outside a tile the tile needs, are tabulated once per launch range
// and tile size. The kernel then computes the tile's own primal values and
// the halo's in the same phase (no discovery pass, no atomics, one barrier).
struct HaloTab {
    const int32_t* slot[3];  // per position and axis: -1 the cube itself; 0..TILE-1 a
                             // cube of the tile; TILE + k halo entry k; <= -2 the
                             // cube -2 - v of another level (frozen u_bar, v_bar)
    const int32_t* cnt;      // per tile: number of halo entries
    const int32_t* list;     // per tile, 3 * TILE entries: halo positions, in order
                             // of the referencing cube, then axis
};

// One workgroup per tile (plain order). nb = the dual step's +a neighbour
// tables; the halo entries of a tile are numbered by an exclusive scan of the
// per-cube counts, so the table does not depend on scheduling.
template <int TILE>
__global__ void __launch_bounds__(TILE) k_halo_table(const int32_t* __restrict__ nb0, const int32_t* __restrict__ nb1,
                                                     const int32_t* __restrict__ nb2, uint32_t off, uint32_t n,
                                                     uint32_t lvl_off, uint32_t lvl_n, int32_t* __restrict__ s0,
                                                     int32_t* __restrict__ s1, int32_t* __restrict__ s2,
                                                     int32_t* __restrict__ cnt, int32_t* __restrict__ list) {
    __shared__ int32_t scan[2][TILE];
    const uint32_t tile = blockIdx.x;
    const uint32_t t = tile * TILE + threadIdx.x;
    const long tile_lo = (long)off + (long)tile * TILE;
    const long tile_hi = (long)off + min(tile * TILE + (uint32_t)TILE, n);
    int32_t j[3] = {-1, -1, -1};
    int nh = 0;
    if (t < n) {
        j[0] = nb0[off + t];
        j[1] = nb1[off + t];
        j[2] = nb2[off + t];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const long v = j[ax];
            if (v >= 0 && !(v >= tile_lo && v < tile_hi) && (uint32_t)v >= lvl_off && (uint32_t)v < lvl_off + lvl_n)
                ++nh;
        }
    }
    int cur = 0;
    scan[0][threadIdx.x] = nh;
    __syncthreads();
    for (int d = 1; d < TILE; d <<= 1) {
        const int v = scan[cur][threadIdx.x] + (threadIdx.x >= (uint32_t)d ? scan[cur][threadIdx.x - d] : 0);
        scan[cur ^ 1][threadIdx.x] = v;
        cur ^= 1;
        __syncthreads();
    }
    const int incl = scan[cur][threadIdx.x];
    if (threadIdx.x == TILE - 1) cnt[tile] = incl;
    if (t >= n) return;
    int k = incl - nh;
    int32_t* const out[3] = {s0, s1, s2};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const long v = j[ax];
        int32_t code;
        if (v < 0) {
            code = -1;
        } else if (v >= tile_lo && v < tile_hi) {
            code = (int32_t)(v - tile_lo);
        } else if ((uint32_t)v >= lvl_off && (uint32_t)v < lvl_off + lvl_n) {
            list[(long)tile * 3 * TILE + k] = (int32_t)v;
            code = TILE + k;
            ++k;
        } else {
            code = -2 - (int32_t)v;
        }
        out[ax][t] = code;
    }
}

// Same values as k_fused_adj / k_fused_tile (identical expressions).
template <class S, int TILE, bool XCD = true>
__global__ void __launch_bounds__(TILE) k_fused_halo(KArgs<S> a, Set13<S> src, Set13<S> dst, HaloTab h) {
    using IO = Io<S>;
    using T = typename IO::T;
    __shared__ T own_bar[4][TILE];
    __shared__ T halo_bar[4][3 * TILE];
    const uint32_t tile = (uint32_t)tile_id<XCD>();
    const uint32_t t = tile * TILE + threadIdx.x;
    const long g = (long)a.off + t;
    const bool run = t < a.n && !masked(a, g);
    T p[3], q[6], nw[4];
    int32_t sl[3] = {-1, -1, -1};
    T frozen[3][4];
    if (run) {
        T od[4];
        primal_adj_at(a, src, g, nw, od, p, q);
#pragma unroll
        for (int c = 0; c < 4; ++c) own_bar[c][threadIdx.x] = IO::rt(T(2) * nw[c] - od[c]);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            sl[ax] = __builtin_nontemporal_load(&h.slot[ax][t]);  // read once
            if (sl[ax] <= -2) {
                const long j = -2L - sl[ax];
                if (a.fzaos && (uint32_t)j - a.fz_off < a.fz_n) {
                    IO::ld4(a.fzaos, (long)((uint32_t)j - a.fz_off), frozen[ax]);
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c) frozen[ax][c] = IO::ld(a.s[F_UB + c], j);
                }
            }
        }
    }
    // the halo's primal values (a cube of a converged block is never read)
    const int nh = h.cnt[tile];
    const int32_t* hl = h.list + (long)tile * 3 * TILE;
    for (int k = threadIdx.x; k < nh; k += TILE) {
        const long j = hl[k];
        if (masked(a, j)) continue;
        T n2[4], o2[4];
        primal_adj_at(a, src, j, n2, o2);
#pragma unroll
        for (int c = 0; c < 4; ++c) halo_bar[c][k] = IO::rt(T(2) * n2[c] - o2[c]);
    }
    __syncthreads();
    if (!run) return;
    T bar[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bar[0][c] = own_bar[c][threadIdx.x];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int v = sl[ax];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            T x;
            if (v == -1) x = bar[0][c];
            else if (v <= -2) x = frozen[ax][c];
            else if (v < TILE) x = own_bar[c][v];
            else x = halo_bar[c][v - TILE];
            bar[1 + ax][c] = x;
        }
    }
    dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::stn(dst.f[F_U + c], g, nw[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::stn(dst.f[F_PX + c], g, p[c]);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::stn(dst.f[F_Q0 + k], g, q[k]);
}

// ---- two iterations per launch (no reference counterpart). A tile T of
// consecutive positions advances by two exact-adjoint iterations in one
// launch: primal n+1 over PA, dual n+2 over DB, primal n+2 over PC and dual
// n+3 over T, where, within the level,
//   PC = T and the +a neighbours of T (the u_bar, v_bar the last dual reads),
//   DB = PC and the cubes whose +a neighbour lies in PC (the p, q that PC's
//        primal step reads),
//   PA = DB and the +a neighbours of DB.
// Every intermediate value stays in LDS, so a cube's own state is read and
// written once per two iterations. The sets are tabulated once per level and
// tile size (k_halo2_table): slot s < TILE is the tile's cube s, slots from
// TILE on are the other members, numbered so that PC, DB and PA are the
// prefixes [0, nPC), [0, nDB) and [0, nPA) of the slots. Each value is
// computed by the expressions of k_primal_adj and k_dual, so the results are
// bit-identical to every other launch variant.
constexpr int kHalo2Cap = 4;  // slots per tile: at most kHalo2Cap * TILE
constexpr int kHalo2Lds = 160 * 1024;  // LDS budget of k_fused2 per workgroup (gfx950: 160 KiB)

__host__ __device__ constexpr int log2_ceil(int x) {
    int l = 0;
    while ((1 << l) < x) ++l;
    return l;
}

// Built by k_halo2_table in tile order; halo2_tab then reorders the rows into
// launch order (LDS class by class) and replaces each count record's last
// entry by its tile id.
struct Halo2Tab {
    const int32_t* cnt;    // per tile: nPC, nDB, nPA, failed (the tile's sets exceed the cap)
    const int32_t* list;   // per tile, (kHalo2Cap - 1) * TILE: position of slot TILE + k
    const int32_t* bslot;  // per tile, 3 x kHalo2Cap * TILE: the +a source of slot s < nDB:
                           // -1 the cube itself, >= 0 a slot, <= -2 the cube -2 - v of
                           // another level (its frozen u_bar, v_bar)
    const int32_t* cslot;  // per tile, 3 x kHalo2Cap * TILE: the cube whose +a neighbour
                           // slot s < nPA is (reference rules: slot s's -a neighbour):
                           // -1 none, >= 0 its slot, <= -2 the position -2 - v outside
                           // the tile's sets (for s < nPC only outside the level)
};

// One workgroup per tile. nba = depth-aware +a neighbours, nbt = their
// transpose within the level; [off, off + n) must be the whole level. Members
// outside the tile are collected in an LDS hash set, then the members of
// PC \ T, DB \ PC and PA \ DB are each numbered in position order, so that
// neighbouring lanes of k_fused2 read neighbouring addresses of the halo.
// REF (the reference's rules): nba = the reference table's +a neighbours and
// nbt its -a neighbours (tgv.cpp:97-133), whose p, q the primal step reads;
// both may point outside the level (cubes of other depths, constant during
// the level): such a -a source is stored as -2 - position.
template <int TILE, bool REF = false>
__global__ void __launch_bounds__(TILE) k_halo2_table(const int32_t* __restrict__ nba0,
                                                      const int32_t* __restrict__ nba1,
                                                      const int32_t* __restrict__ nba2,
                                                      const int32_t* __restrict__ nbt0,
                                                      const int32_t* __restrict__ nbt1,
                                                      const int32_t* __restrict__ nbt2, uint32_t off, uint32_t n,
                                                      int32_t* __restrict__ cnt, int32_t* __restrict__ list,
                                                      int32_t* __restrict__ bslot, int32_t* __restrict__ cslot) {
    constexpr int CA = kHalo2Cap * TILE, NH = CA - TILE;
    constexpr int LH = log2_ceil(2 * NH), H = 1 << LH;
    __shared__ int32_t hkey[H], hval[H], lst[NH], rk[NH];
    __shared__ int32_t nh, failed, cP, cD;
    const int32_t* const nba[3] = {nba0, nba1, nba2};
    const int32_t* const nbt[3] = {nbt0, nbt1, nbt2};
    const uint32_t tile = blockIdx.x;
    const int tid = (int)threadIdx.x;
    const long lo = (long)off + (long)tile * TILE;
    const int cntT = (int)min((uint32_t)TILE, n - tile * TILE);
    const long hi = lo + cntT, lvl_hi = (long)off + n;
    for (int i = tid; i < H; i += TILE) hkey[i] = -1;
    if (tid == 0) {
        nh = 0;
        failed = 0;
    }
    __syncthreads();
    auto in_tile = [&](long j) { return j >= lo && j < hi; };
    auto in_level = [&](long j) { return j >= (long)off && j < lvl_hi; };
    auto hash = [](int32_t j) { return ((uint32_t)j * 0x9E3779B1u) >> (32 - LH); };
    auto insert = [&](int32_t j) {
        uint32_t hp = hash(j);
        for (int it = 0; it < H; ++it) {
            const int32_t prev = atomicCAS(&hkey[hp], -1, j);
            if (prev == -1) {
                const int k = atomicAdd(&nh, 1);
                if (k < NH) {
                    lst[k] = j;
                    hval[hp] = TILE + k;
                } else {
                    hval[hp] = -1;
                    atomicOr(&failed, 1);
                }
                return;
            }
            if (prev == j) return;
            hp = (hp + 1u) & (uint32_t)(H - 1);
        }
        atomicOr(&failed, 1);
    };
    auto find = [&](int32_t j) -> int32_t {
        uint32_t hp = hash(j);
        for (int it = 0; it < H; ++it) {
            const int32_t k = hkey[hp];
            if (k == j) return hval[hp];
            if (k == -1) break;
            hp = (hp + 1u) & (uint32_t)(H - 1);
        }
        return -1;
    };
    auto lookup = [&](int32_t j) -> int32_t {
        const int32_t v = find(j);
        if (v < 0) atomicOr(&failed, 1);
        return v;
    };
    auto pos = [&](int s) -> long { return s < TILE ? lo + s : (long)lst[s - TILE]; };
    auto valid = [&](int s) { return !(s >= cntT && s < TILE); };
    // PC: the tile's +a neighbours in the level
    if (tid < cntT) {
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int32_t j = nba[ax][lo + tid];
            if (j >= 0 && in_level(j) && !in_tile(j)) insert(j);
        }
    }
    __syncthreads();
    if (tid == 0) cP = TILE + min((int)nh, NH);
    __syncthreads();
    const int nPC = cP;
    // DB: the cubes whose +a neighbour is in PC (always in the level)
    for (int s = tid; s < nPC; s += TILE) {
        if (!valid(s)) continue;
        const long g = pos(s);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int32_t k = nbt[ax][g];
            if (REF) {
                if (k >= 0 && in_level(k) && !in_tile(k)) insert(k);
            } else if (k >= 0 && !in_level(k)) {
                atomicOr(&failed, 1);
            } else if (k >= 0 && !in_tile(k)) {
                insert(k);
            }
        }
    }
    __syncthreads();
    if (tid == 0) cD = TILE + min((int)nh, NH);
    __syncthreads();
    const int nDB = cD;
    // PA: the +a neighbours of DB in the level
    for (int s = tid; s < nDB; s += TILE) {
        if (!valid(s)) continue;
        const long g = pos(s);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int32_t j = nba[ax][g];
            if (j >= 0 && in_level(j) && !in_tile(j)) insert(j);
        }
    }
    __syncthreads();
    const int nPA = TILE + min((int)nh, NH);
    // position order within each class: rank among the class's members
    const int nX = nPA - TILE;
    for (int k = tid; k < nX; k += TILE) {
        const int sl = TILE + k;
        const int k0 = (sl < nPC ? TILE : sl < nDB ? nPC : nDB) - TILE;
        const int k1 = (sl < nPC ? nPC : sl < nDB ? nDB : nPA) - TILE;
        const int32_t v = lst[k];
        int r = k0;
        for (int m = k0; m < k1; ++m) r += lst[m] < v ? 1 : 0;
        rk[k] = r;
    }
    int32_t mv[kHalo2Cap - 1];
#pragma unroll
    for (int i = 0; i < kHalo2Cap - 1; ++i) {
        const int k = tid + i * TILE;
        mv[i] = k < nX ? lst[k] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kHalo2Cap - 1; ++i) {
        const int k = tid + i * TILE;
        if (k < nX) lst[rk[k]] = mv[i];
    }
    for (int i = tid; i < H; i += TILE) {
        const int32_t sv = hval[i];
        if (hkey[i] != -1 && sv >= TILE) hval[i] = TILE + rk[sv - TILE];
    }
    __syncthreads();
    int32_t* const L = list + (long)tile * NH;
    for (int k = tid; k < nPA - TILE; k += TILE) L[k] = lst[k];
    int32_t* const B = bslot + (long)tile * 3 * CA;
    int32_t* const C = cslot + (long)tile * 3 * CA;
    for (int s = tid; s < nPA; s += TILE) {
        if (!valid(s)) continue;
        const long g = pos(s);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (s < nDB) {
                const int32_t j = nba[ax][g];
                int32_t v;
                if (j < 0) v = -1;
                else if (in_tile(j)) v = (int32_t)(j - lo);
                else if (in_level(j)) v = lookup(j);
                else v = -2 - j;
                B[ax * CA + s] = v;
            }
            const int32_t k = nbt[ax][g];
            int32_t v;
            if (k < 0) v = -1;
            else if (in_tile(k)) v = (int32_t)(k - lo);
            else if (REF && !in_level(k)) v = -2 - k;
            else if (s < nPC) v = lookup(k);
            else {
                v = find(k);
                if (v < 0) v = -2 - k;
            }
            C[ax * CA + s] = v;
        }
    }
    __syncthreads();
    if (tid == 0) {
        cnt[4 * tile] = nPC;
        cnt[4 * tile + 1] = nDB;
        cnt[4 * tile + 2] = nPA;
        cnt[4 * tile + 3] = (failed || nh > NH) ? 1 : 0;
    }
}

// One 4-byte value per lane from global memory straight into LDS: lane l of
// the wave writes l_base[l] (l_base wave-uniform); inactive lanes write
// nothing. The LDS write completes with the load (vmcnt), before a barrier.
__device__ inline void lds_dma4(const float* g, float* l_base) {
    __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1)))*)g,
                                     (void __attribute__((address_space(3)))*)l_base, 4, 0, 0);
}

// The dynamic LDS of k_fused2: u_bar, v_bar and p, q of PA (sa slots), and
// u, v of the PC slots from 2 * TILE on (sx slots; the first two slots of
// each thread keep theirs in registers, as they keep their table entries,
// info words and frozen neighbours).
extern __shared__ uint4 k_fused2_lds[];

// h: the tables of the tiles this launch runs (one LDS class), in launch
// order: workgroup k (XCD-contiguous or plain) runs the tile of record k,
// whose count record holds nPC, nDB, nPA and the tile id, so that every table
// read of the staging depends on the workgroup id alone.
// REF: the reference's update rules (k_dual and k_primal_ref's expressions):
// the primal step reads p, q of the -a neighbours of the reference table and
// always its own; a -a neighbour of another depth gives its p, q from the
// main set, which the level never writes.
// HEAD: the first launch of a level that starts from its initial state
// (u_bar = u, v_bar = v, p = q = 0, as k_reset and k_parent_init leave it):
// dual 1 over DB, primal 1 over PC, dual 2 over T, so src holds (u^0, v^0)
// and dst receives (u^1, v^1, p^2, q^2); src's p, q are not read, the
// launch's first dual step starts from zero.
template <class S, int TILE, bool XCD = true, bool REF = false, bool HEAD = false>
__global__ void __launch_bounds__(TILE) k_fused2(KArgs<S> a, Set13<S> src, Set13<S> dst, Halo2Tab h, int sa,
                                                 int sx) {
    using IO = Io<S>;
    using T = typename IO::T;
    constexpr int CA = kHalo2Cap * TILE, NH = CA - TILE;
    T* const ub = reinterpret_cast<T*>(k_fused2_lds);  // [4][sa]
    T* const pq = ub + 4 * sa;                         // [9][sa]: p[3], q[6]
    T* const uvx = pq + 9 * sa;                        // [4][sx]
    // the weighted median of every bin (median_of's expression, once per
    // workgroup instead of a division per prox step); read after the
    // staging barrier
    __shared__ T s_med[32];
    const long kp = tile_id<XCD>();
    const int32_t* const hd = h.cnt + 4 * kp;
    const int nP = hd[0], nD = hd[1], nA = hd[2];
    const uint32_t tile = (uint32_t)hd[3];
    const int tid = (int)threadIdx.x;
    if (tid < 32) s_med[tid] = median_of<T>(((uint32_t)tid << 24) | 1u, a.nbins);
    const long lo = (long)a.off + (long)tile * TILE;
    const int cntT = (int)min((uint32_t)TILE, a.n - tile * TILE);
    const int32_t* const hl = h.list + kp * NH;
    const int32_t* const bs = h.bslot + kp * 3 * CA;
    const int32_t* const cs = h.cslot + kp * 3 * CA;
    auto live = [&](int s, long& g) {
        if (s >= cntT && s < TILE) return false;
        g = s < TILE ? lo + s : (long)hl[s - TILE];
        return !masked(a, g);
    };
    // frozen u_bar, v_bar of the +a neighbour (axis ax) of the cube at g,
    // encoded v <= -2
    auto frozen = [&](int v, T o[4], int ax, long g) {
        const long j = -2L - v;
        if (REF && a.rfz_bar) {
            IO::ld4(a.rfz_bar, (long)ax * a.lvl_n + (g - (long)a.lvl_off), o);
        } else if (a.fzaos && (uint32_t)j - a.fz_off < a.fz_n) {
            IO::ld4(a.fzaos, (long)((uint32_t)j - a.fz_off), o);
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) o[c] = IO::ld(a.s[F_UB + c], j);
        }
    };
    // own p, q of slot s into LDS
    auto stage = [&](int s, long g) {
#pragma unroll
        for (int c = 0; c < 3; ++c) pq[c * sa + s] = IO::ld(src.f[F_PX + c], g);
#pragma unroll
        for (int k = 0; k < 6; ++k) pq[(3 + k) * sa + s] = IO::ld(src.f[F_Q0 + k], g);
    };
    // own p, q of the slots base + lane of this wave (g: this lane's cube)
    // into LDS without passing through registers (single precision, tiles
    // of 1024: 0.9 % faster there, 0.8 % slower at 512)
    auto stage_w = [&](int s, long g, int base) {
        if constexpr (std::is_same<S, float>::value && TILE >= 1024) {
#pragma unroll
            for (int c = 0; c < 3; ++c) lds_dma4(src.f[F_PX + c] + g, pq + c * sa + base);
#pragma unroll
            for (int k = 0; k < 6; ++k) lds_dma4(src.f[F_Q0 + k] + g, pq + (3 + k) * sa + base);
        } else {
            stage(s, g);
        }
    };
    // incoming edge of axis ax: p_ax and the q entries of row ax of eps
    constexpr int kF[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
    // field f of the cube at position j outside the tile's sets: in the
    // level from src, else (reference rules only) from the main set
    auto ld_out = [&](int f, long j) -> T {
        if (REF && !((uint32_t)j - a.lvl_off < a.lvl_n)) return IO::ld(a.s[f], j);
        return IO::ld(src.f[f], j);
    };
    // the incoming values (kF[ax]) of the cube at position j, the -a
    // neighbour of the cube at g, outside the tile's sets
    auto gather_in = [&](int ax, long j, long g, T o[4]) {
        if (REF && a.rfz_dual && !((uint32_t)j - a.lvl_off < a.lvl_n)) {
            IO::ld4(a.rfz_dual, (long)ax * a.lvl_n + (g - (long)a.lvl_off), o);
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) o[e] = ld_out(kF[ax][e], j);
        }
    };
    // passes 3-5 of slot s (the expressions of primal_adj_at): own p, q and
    // the incoming ones from LDS (cv >= 0) or from src (cv <= -2, p, q of the
    // first iteration only; pin holds them if have_in); u_bar, v_bar to LDS,
    // the new u, v in nw
    auto primal = [&](int s, long g, uint32_t w, const T od[4], const int cv[3], T nw[4], bool have_in,
                      const T (&pin)[3][4]) {
        T p[3], qs[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = pq[c * sa + s];
#pragma unroll
        for (int k = 0; k < 6; ++k) qs[k] = pq[(3 + k) * sa + s];
        T dp = T(0);
        T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const bool hs = InfoWord::has(w, ax);
            dp = hs ? dp + p[ax] : dp;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const T w2 = (c == ax) ? T(1) : T(0.5);
                dq[c] = hs ? dq[c] + w2 * qs[qidx(c, ax) - F_Q0] : dq[c];
            }
        }
        T in[3][4];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int v = cv[ax];
            if (v >= 0) {
#pragma unroll
                for (int e = 0; e < 4; ++e) in[ax][e] = pq[(kF[ax][e] - F_PX) * sa + v];
            } else if (have_in) {
#pragma unroll
                for (int e = 0; e < 4; ++e) in[ax][e] = v <= -2 ? pin[ax][e] : T(0);
            } else if (v <= -2) {
                gather_in(ax, -2L - v, g, in[ax]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) in[ax][e] = T(0);
            }
        }
        if constexpr (REF) {
            // tgv.cpp:192-217 (k_primal_ref): a missing -a neighbour gives 0
            const T div_p = (p[0] - in[0][0]) + (p[1] - in[1][0]) + (p[2] - in[2][0]);
            nw[0] = prox_m(od[0] + a.tau * div_p, w, s_med[InfoWord::bin(w)], a.tl);
            const T dqx = (qs[0] - in[0][1]) + (qs[3] - in[1][1]) + (qs[4] - in[2][1]);
            const T dqy = (qs[3] - in[0][2]) + (qs[1] - in[1][2]) + (qs[5] - in[2][2]);
            const T dqz = (qs[4] - in[0][3]) + (qs[5] - in[1][3]) + (qs[2] - in[2][3]);
            nw[1] = od[1] + a.tau * (p[0] + dqx);
            nw[2] = od[2] + a.tau * (p[1] + dqy);
            nw[3] = od[3] + a.tau * (p[2] + dqz);
#pragma unroll
            for (int c = 0; c < 4; ++c) ub[c * sa + s] = IO::rt(T(2) * nw[c] - od[c]);
            return;
        }
        if (cv[0] != -1) {
            dp -= in[0][0];
            dq[0] -= T(1) * in[0][1];
            dq[1] -= T(0.5) * in[0][2];
            dq[2] -= T(0.5) * in[0][3];
        }
        if (cv[1] != -1) {
            dp -= in[1][0];
            dq[0] -= T(0.5) * in[1][1];
            dq[1] -= T(1) * in[1][2];
            dq[2] -= T(0.5) * in[1][3];
        }
        if (cv[2] != -1) {
            dp -= in[2][0];
            dq[0] -= T(0.5) * in[2][1];
            dq[1] -= T(0.5) * in[2][2];
            dq[2] -= T(1) * in[2][3];
        }
        nw[0] = prox_m(od[0] + a.tau * dp, w, s_med[InfoWord::bin(w)], a.tl);
#pragma unroll
        for (int c = 0; c < 3; ++c) nw[1 + c] = od[1 + c] + a.tau * (p[c] + dq[c]);
#pragma unroll
        for (int c = 0; c < 4; ++c) ub[c * sa + s] = IO::rt(T(2) * nw[c] - od[c]);
    };
    // dual n+2 of slot s; fz: the frozen neighbours already read (have_fz)
    auto dual2 = [&](int s, long g, const int bv[3], bool have_fz, const T (&fz)[3][4]) {
        T bar[4][4];
#pragma unroll
        for (int c = 0; c < 4; ++c) bar[0][c] = ub[c * sa + s];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int v = bv[ax];
            if (v <= -2) {
                if (have_fz) {
#pragma unroll
                    for (int c = 0; c < 4; ++c) bar[1 + ax][c] = fz[ax][c];
                } else {
                    frozen(v, bar[1 + ax], ax, g);
                }
            } else {
#pragma unroll
                for (int c = 0; c < 4; ++c) bar[1 + ax][c] = v == -1 ? bar[0][c] : ub[c * sa + v];
            }
        }
        T p[3], q[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = pq[c * sa + s];
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = pq[(3 + k) * sa + s];
        dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
        for (int c = 0; c < 3; ++c) pq[c * sa + s] = IO::rtd(p[c]);
#pragma unroll
        for (int k = 0; k < 6; ++k) pq[(3 + k) * sa + s] = IO::rtd(q[k]);
    };
    auto load_uv = [&](long g, T o[4]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) o[c] = IO::ld(src.f[F_U + c], g);
    };
    // HEAD: slot s starts from u_bar = (u, v) = o and p = q = 0
    auto seed = [&](int s, const T o[4]) {
#pragma unroll
        for (int c = 0; c < 4; ++c) ub[c * sa + s] = o[c];
#pragma unroll
        for (int k = 0; k < 9; ++k) pq[k * sa + s] = T(0);
    };
    // ---- stage p, q of PA. Slots tid (0) and tid + TILE (1) keep their
    // table entries, info words, u, v and (slot 0) frozen neighbours in
    // registers. With tiles of 1024 a CU holds one workgroup whatever the
    // VGPRs up to 128, so in single precision slot 1 also keeps its frozen
    // neighbours and its incoming p, q from outside the sets (the global
    // reads of the first primal and the dual step), which leaves those phases
    // LDS-only (in double precision these registers would spill).
    constexpr bool PF = TILE >= 1024 && sizeof(T) == 4;
    const int s1 = tid + TILE;
    // the table entries of slot 0, read at once (within the tile's rows,
    // whatever its size), and those of slot 1 with its position where the
    // tile has that slot: the values they address are then one dependent
    // read away from the count record
    int32_t h1 = 0;
    int t_b0[3], t_c0[3], t_b1[3] = {-1, -1, -1}, t_c1[3] = {-1, -1, -1};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        t_b0[ax] = bs[ax * CA + tid];
        t_c0[ax] = cs[ax * CA + tid];
    }
    if (s1 < nA) {
        h1 = hl[tid];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            t_b1[ax] = bs[ax * CA + s1];
            t_c1[ax] = cs[ax * CA + s1];
        }
    }
    long g0 = 0, g1 = 0;
    const bool l0 = live(tid, g0);
    bool l1 = s1 < nA;
    if (l1) {
        g1 = (long)h1;
        l1 = !masked(a, g1);
    }
    int b0[3], c0[3], b1[3], c1[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        b0[ax] = l0 ? t_b0[ax] : -1;
        c0[ax] = l0 ? t_c0[ax] : -1;
        b1[ax] = l1 && s1 < nD ? t_b1[ax] : -1;
        c1[ax] = l1 ? t_c1[ax] : -1;
    }
    uint32_t w0 = 0u, w1 = 0u;
    T od0[4] = {T(0), T(0), T(0), T(0)}, od1[4] = {T(0), T(0), T(0), T(0)};
    T fz0[3][4], fz1[3][4], in1[3][4];
    // reference rules (tiles of 1024, single precision): slot 0's -a
    // neighbours of other depths, whose p, q both primal steps read, are
    // gathered while staging too (the exact-adjoint sources are all in the
    // tile's sets)
    constexpr bool PF0 = PF && REF;
    T in0[3][4];
    if (l0) {
        w0 = a.info[g0];
        load_uv(g0, od0);
        if constexpr (HEAD) seed(tid, od0);
        else stage_w(tid, g0, tid & ~63);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
            if (b0[ax] <= -2) frozen(b0[ax], fz0[ax], ax, g0);
        if constexpr (PF0) {
#pragma unroll
            for (int ax = 0; ax < 3; ++ax)
                if (c0[ax] <= -2) gather_in(ax, -2L - c0[ax], g0, in0[ax]);
        }
    }
    if (l1) {
        w1 = a.info[g1];
        load_uv(g1, od1);
        if constexpr (HEAD) seed(s1, od1);
        else stage_w(s1, g1, TILE + (tid & ~63));
        if constexpr (PF) {
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                if (b1[ax] <= -2) frozen(b1[ax], fz1[ax], ax, g1);
                // reference rules: for a slot of PC, a -a neighbour outside
                // the sets lies outside the level, so its p, q are constant
                // and serve the second primal step too (the head launch's only)
                if ((!HEAD || (REF && s1 < nP)) && c1[ax] <= -2) gather_in(ax, -2L - c1[ax], g1, in1[ax]);
            }
        }
    }
    for (int s = tid + 2 * TILE; s < nA; s += TILE) {
        long g;
        if (!live(s, g)) continue;
        if constexpr (HEAD) {
            T od[4];
            load_uv(g, od);
            seed(s, od);
        } else {
            stage(s, g);
        }
    }
    __syncthreads();
    // ---- primal n+1 over PA (HEAD: none, the launch starts with the dual)
    T nw0[4] = {T(0), T(0), T(0), T(0)}, nw1[4] = {T(0), T(0), T(0), T(0)};
    if constexpr (!HEAD) {
        if (l0) primal(tid, g0, w0, od0, c0, nw0, PF0, in0);
        if (l1) primal(s1, g1, w1, od1, c1, nw1, PF, in1);
        for (int s = tid + 2 * TILE; s < nA; s += TILE) {
            long g;
            if (!live(s, g)) continue;
            const int cv[3] = {cs[s], cs[CA + s], cs[2 * CA + s]};
            T od[4], nw[4];
            load_uv(g, od);
            primal(s, g, a.info[g], od, cv, nw, false, in1);
            if (s < nP) {
#pragma unroll
                for (int c = 0; c < 4; ++c) uvx[c * sx + s - 2 * TILE] = IO::rt(nw[c]);
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            nw0[c] = IO::rt(nw0[c]);  // the stored u, v of the first iteration
            nw1[c] = IO::rt(nw1[c]);
        }
        __syncthreads();
    }
    // ---- dual n+2 over DB (HEAD: dual 1)
    if (l0) dual2(tid, g0, b0, true, fz0);
    if (l1 && s1 < nD) dual2(s1, g1, b1, PF, fz1);
    for (int s = tid + 2 * TILE; s < nD; s += TILE) {
        long g;
        if (!live(s, g)) continue;
        const int bv[3] = {bs[s], bs[CA + s], bs[2 * CA + s]};
        dual2(s, g, bv, false, fz0);
    }
    __syncthreads();
    // ---- primal n+2 over PC (HEAD: primal 1, from the staged (u, v); a
    // slot's own u_bar is read before its primal step replaces it)
    T own[4] = {T(0), T(0), T(0), T(0)};
    if (l0) primal(tid, g0, w0, HEAD ? od0 : nw0, c0, own, PF0, in0);
    if (l1 && s1 < nP) {
        T x[4];
        primal(s1, g1, w1, HEAD ? od1 : nw1, c1, x, PF && REF, in1);
    }
    for (int s = tid + 2 * TILE; s < nP; s += TILE) {
        long g;
        if (!live(s, g)) continue;
        const int cv[3] = {cs[s], cs[CA + s], cs[2 * CA + s]};
        T od[4], x[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) od[c] = HEAD ? ub[c * sa + s] : uvx[c * sx + s - 2 * TILE];
        primal(s, g, a.info[g], od, cv, x, false, in1);
    }
    __syncthreads();
    // ---- dual n+3 over the tile (HEAD: dual 2)
    if (!l0) return;
    T bar[4][4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bar[0][c] = ub[c * sa + tid];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int v = b0[ax];
#pragma unroll
        for (int c = 0; c < 4; ++c) bar[1 + ax][c] = v <= -2 ? fz0[ax][c] : (v == -1 ? bar[0][c] : ub[c * sa + v]);
    }
    T p[3], q[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) p[c] = pq[c * sa + tid];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = pq[(3 + k) * sa + tid];
    dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::stn(dst.f[F_U + c], g0, own[c]);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::stn(dst.f[F_PX + c], g0, p[c]);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::stn(dst.f[F_Q0 + k], g0, q[k]);
}

// ---- four iterations per launch (no reference counterpart). Where a level
// is small, its two-iteration launches are bound by latency, not bandwidth
// (one workgroup runs its tile through staging and four dependent phases
// while most CUs idle), so these levels take four iterations per launch over
// deeper tile sets, at the price of more redundant work at the sets' edges.
// For a tile T the nested sets are G_0 = T and G_k = G_{k-1} plus its +a
// neighbours (k odd) or its -a sources (k even) within the level, k = 1..7;
// the launch runs primal n+1 over G_7, dual n+2 over G_6, primal n+2 over
// G_5, ..., primal n+4 over G_1 and dual n+5 over G_0, every value in LDS.
// Neighbours of other depths (constant while the level iterates) become
// constant slots after G_7, holding their u_bar, v_bar and, with the
// reference's rules, the p, q the primal step reads of them. Each value is
// computed by the expressions of k_fused2, so the results are bit-identical
// to every other launch variant.
constexpr int kHalo4Cap = 8;   // slots per tile: at most kHalo4Cap * TILE
constexpr int kHalo4Sets = 8;  // G_0 .. G_7
constexpr int kHalo4Cnt = 10;  // per tile: ends of G_0 .. G_7, end of the constant slots, failed

struct Halo4Tab {
    const int32_t* cnt;    // per tile, kHalo4Cnt
    const int32_t* list;   // per tile, (kHalo4Cap - 1) * TILE: position of slot TILE + k
    const int32_t* bslot;  // per tile, 3 x kHalo4Cap * TILE: the +a source of slot s < end(G_6):
                           // -1 the cube itself, >= 0 a slot (a constant slot for another depth)
    const int32_t* cslot;  // per tile, 3 x kHalo4Cap * TILE: for slot s < end(G_7) the cube whose
                           // +a neighbour it is (reference rules: its -a neighbour): -1 none,
                           // >= 0 a slot, <= -2 the position -2 - v in the level outside the
                           // sets (slots beyond G_5 only, read from the launch's input)
};

// One workgroup per tile; [off, off + n) is the whole level. As
// k_halo2_table: members outside the tile go into an LDS hash set stage by
// stage, then each class (G_k \ G_{k-1}, the constant slots) is numbered in
// position order.
template <int TILE, bool REF = false>
__global__ void __launch_bounds__(TILE) k_halo4_table(const int32_t* __restrict__ nba0,
                                                      const int32_t* __restrict__ nba1,
                                                      const int32_t* __restrict__ nba2,
                                                      const int32_t* __restrict__ nbt0,
                                                      const int32_t* __restrict__ nbt1,
                                                      const int32_t* __restrict__ nbt2, uint32_t off, uint32_t n,
                                                      int32_t* __restrict__ cnt, int32_t* __restrict__ list,
                                                      int32_t* __restrict__ bslot, int32_t* __restrict__ cslot) {
    constexpr int CA = kHalo4Cap * TILE, NH = CA - TILE;
    constexpr int LH = log2_ceil(2 * NH), H = 1 << LH;
    __shared__ int32_t hkey[H], hval[H], lst[NH], rk[NH];
    __shared__ int32_t nh, failed, ends[kHalo4Sets + 1];
    const int32_t* const nba[3] = {nba0, nba1, nba2};
    const int32_t* const nbt[3] = {nbt0, nbt1, nbt2};
    const uint32_t tile = blockIdx.x;
    const int tid = (int)threadIdx.x;
    const long lo = (long)off + (long)tile * TILE;
    const int cntT = (int)min((uint32_t)TILE, n - tile * TILE);
    const long hi = lo + cntT, lvl_hi = (long)off + n;
    for (int i = tid; i < H; i += TILE) hkey[i] = -1;
    if (tid == 0) {
        nh = 0;
        failed = 0;
        ends[0] = TILE;
    }
    __syncthreads();
    auto in_tile = [&](long j) { return j >= lo && j < hi; };
    auto in_level = [&](long j) { return j >= (long)off && j < lvl_hi; };
    auto hash = [](int32_t j) { return ((uint32_t)j * 0x9E3779B1u) >> (32 - LH); };
    auto insert = [&](int32_t j) {
        uint32_t hp = hash(j);
        for (int it = 0; it < H; ++it) {
            const int32_t prev = atomicCAS(&hkey[hp], -1, j);
            if (prev == -1) {
                const int k = atomicAdd(&nh, 1);
                if (k < NH) {
                    lst[k] = j;
                    hval[hp] = TILE + k;
                } else {
                    hval[hp] = -1;
                    atomicOr(&failed, 1);
                }
                return;
            }
            if (prev == j) return;
            hp = (hp + 1u) & (uint32_t)(H - 1);
        }
        atomicOr(&failed, 1);
    };
    auto find = [&](int32_t j) -> int32_t {
        uint32_t hp = hash(j);
        for (int it = 0; it < H; ++it) {
            const int32_t k = hkey[hp];
            if (k == j) return hval[hp];
            if (k == -1) break;
            hp = (hp + 1u) & (uint32_t)(H - 1);
        }
        return -1;
    };
    auto lookup = [&](int32_t j) -> int32_t {
        const int32_t v = find(j);
        if (v < 0) atomicOr(&failed, 1);
        return v;
    };
    auto pos = [&](int s) -> long { return s < TILE ? lo + s : (long)lst[s - TILE]; };
    auto valid = [&](int s) { return !(s >= cntT && s < TILE); };
    // G_1 .. G_7: +a neighbours for odd k, -a sources for even k, within the
    // level (the exact-adjoint sources always are)
    for (int k = 1; k < kHalo4Sets; ++k) {
        const int prev = ends[k - 1];
        const bool plus = (k & 1) != 0;
        for (int s = tid; s < prev; s += TILE) {
            if (!valid(s)) continue;
            const long g = pos(s);
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const int32_t j = plus ? nba[ax][g] : nbt[ax][g];
                if (j < 0) continue;
                if (!in_level(j)) {
                    if (!plus && !REF) atomicOr(&failed, 1);
                    continue;
                }
                if (!in_tile(j)) insert(j);
            }
        }
        __syncthreads();
        if (tid == 0) ends[k] = TILE + min((int)nh, NH);
        __syncthreads();
    }
    // constant slots: the other-depth +a neighbours of the dual slots and,
    // with the reference's rules, the other-depth -a neighbours of the
    // primal slots
    const int e6 = ends[6], e7 = ends[7];
    for (int s = tid; s < e7; s += TILE) {
        if (!valid(s)) continue;
        const long g = pos(s);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (s < e6) {
                const int32_t j = nba[ax][g];
                if (j >= 0 && !in_level(j)) insert(j);
            }
            if (REF) {
                const int32_t k = nbt[ax][g];
                if (k >= 0 && !in_level(k)) insert(k);
            }
        }
    }
    __syncthreads();
    if (tid == 0) ends[kHalo4Sets] = TILE + min((int)nh, NH);
    __syncthreads();
    const int eF = ends[kHalo4Sets];
    // position order within each class
    const int nX = eF - TILE;
    for (int k = tid; k < nX; k += TILE) {
        const int sl = TILE + k;
        int c = 1;
        while (c < kHalo4Sets && sl >= ends[c]) ++c;
        const int k0 = ends[c - 1] - TILE, k1 = ends[c] - TILE;
        const int32_t v = lst[k];
        int r = k0;
        for (int m = k0; m < k1; ++m) r += lst[m] < v ? 1 : 0;
        rk[k] = r;
    }
    int32_t mv[kHalo4Cap - 1];
#pragma unroll
    for (int i = 0; i < kHalo4Cap - 1; ++i) {
        const int k = tid + i * TILE;
        mv[i] = k < nX ? lst[k] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kHalo4Cap - 1; ++i) {
        const int k = tid + i * TILE;
        if (k < nX) lst[rk[k]] = mv[i];
    }
    for (int i = tid; i < H; i += TILE) {
        const int32_t sv = hval[i];
        if (hkey[i] != -1 && sv >= TILE) hval[i] = TILE + rk[sv - TILE];
    }
    __syncthreads();
    int32_t* const L = list + (long)tile * NH;
    for (int k = tid; k < nX; k += TILE) L[k] = lst[k];
    int32_t* const B = bslot + (long)tile * 3 * CA;
    int32_t* const C = cslot + (long)tile * 3 * CA;
    const int e5 = ends[5];
    for (int s = tid; s < e7; s += TILE) {
        if (!valid(s)) continue;
        const long g = pos(s);
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            if (s < e6) {
                const int32_t j = nba[ax][g];
                B[ax * CA + s] = j < 0 ? -1 : (in_tile(j) ? (int32_t)(j - lo) : lookup(j));
            }
            const int32_t k = nbt[ax][g];
            int32_t v;
            if (k < 0) v = -1;
            else if (in_tile(k)) v = (int32_t)(k - lo);
            else if (!in_level(k)) v = lookup(k);  // a constant slot (reference rules)
            else if (s < e5) v = lookup(k);        // in G_6
            else {
                v = find(k);
                if (v < 0) v = -2 - k;
            }
            C[ax * CA + s] = v;
        }
    }
    __syncthreads();
    if (tid == 0) {
        for (int k = 0; k <= kHalo4Sets; ++k) cnt[kHalo4Cnt * tile + k] = ends[k];
        cnt[kHalo4Cnt * tile + kHalo4Sets + 1] = (failed || nh > NH) ? 1 : 0;
    }
}

// The dynamic LDS of k_fused4: u_bar, v_bar (4), p, q (9) and the stored
// u, v (4) of every slot, each as an array of sa values.
extern __shared__ uint4 k_fused4_lds[];

// Threads per workgroup: more than the tile, so that each phase has at most
// a few slots per thread and sixteen waves per CU hide the LDS and divide
// latencies (with one thread per tile cube the launch is latency-bound).
constexpr int kFused4Threads = 1024;

// One workgroup per tile, in XCD-contiguous order. src holds (u^n, v^n,
// p^n+1, q^n+1), dst receives (u^n+4, v^n+4, p^n+5, q^n+5) of the tiles'
// cubes; sa = the slots of the largest tile's sets and constant slots.
template <class S, int TILE, bool REF = false>
__global__ void __launch_bounds__(kFused4Threads) k_fused4(KArgs<S> a, Set13<S> src, Set13<S> dst, Halo4Tab h,
                                                           int sa) {
    constexpr int NT = kFused4Threads;
    using IO = Io<S>;
    using T = typename IO::T;
    constexpr int CA = kHalo4Cap * TILE, NH = CA - TILE;
    // iterated slots per thread: slot tid + r * NT is thread tid's r-th; its
    // position, info word and slot tables stay in registers for the launch
    constexpr int R = (CA + NT - 1) / NT;
    static_assert(R <= 2, "tiles of at most 256");
    T* const ub = reinterpret_cast<T*>(k_fused4_lds);  // [4][sa]
    T* const pq = ub + 4 * sa;                         // [9][sa]: p[3], q[6]
    T* const uv = pq + 9 * sa;                         // [4][sa]: u, v as stored
    __shared__ T s_med[32];
    const uint32_t tile = (uint32_t)tile_id<true>();
    const int tid = (int)threadIdx.x;
    if (tid < 32) s_med[tid] = median_of<T>(((uint32_t)tid << 24) | 1u, a.nbins);
    const long lo = (long)a.off + (long)tile * TILE;
    const int cntT = (int)min((uint32_t)TILE, a.n - tile * TILE);
    const int32_t* const e = h.cnt + (long)tile * kHalo4Cnt;
    int ends[kHalo4Sets + 1];
#pragma unroll
    for (int k = 0; k <= kHalo4Sets; ++k) ends[k] = e[k];
    const int e6 = ends[6], e7 = ends[7], eF = ends[kHalo4Sets];
    const int32_t* const hl = h.list + (long)tile * NH;
    const int32_t* const bs = h.bslot + (long)tile * 3 * CA;
    const int32_t* const cs = h.cslot + (long)tile * 3 * CA;
    auto live = [&](int s, long& g) {
        if (s >= cntT && s < TILE) return false;
        g = s < TILE ? lo + s : (long)hl[s - TILE];
        return !masked(a, g);
    };
    constexpr int kF[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
    // ---- staging: the thread's iterated slots (records, u, v, p, q from
    // src); u_bar, v_bar (and with the reference's rules p, q) of the
    // constant slots from the main set, which the level never writes
    bool lv[R];
    long gr[R];
    uint32_t wr[R];
    int br[R][3], cr[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int s = tid + r * NT;
        lv[r] = s < e7 && live(s, gr[r]);
        wr[r] = 0u;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) br[r][ax] = cr[r][ax] = -1;
        if (!lv[r]) continue;
        const long g = gr[r];
        wr[r] = a.info[g];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            cr[r][ax] = cs[ax * CA + s];
            if (s < e6) br[r][ax] = bs[ax * CA + s];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) uv[c * sa + s] = IO::ld(src.f[F_U + c], g);
#pragma unroll
        for (int k = 0; k < 9; ++k) pq[k * sa + s] = IO::ld(src.f[F_PX + k], g);
    }
    for (int s = e7 + tid; s < eF; s += NT) {
        const long g = (long)hl[s - TILE];
#pragma unroll
        for (int c = 0; c < 4; ++c) ub[c * sa + s] = IO::ld(a.s[F_UB + c], g);
        if (REF) {
#pragma unroll
            for (int k = 0; k < 9; ++k) pq[k * sa + s] = IO::ld(a.s[F_PX + k], g);
        }
    }
    __syncthreads();
    // passes 3-5 of the thread's slot r (k_fused2's primal expressions): own
    // p, q and the incoming ones from LDS, or (first phase only) from src
    auto primal = [&](int r) {
        const int s = tid + r * NT;
        const uint32_t w = wr[r];
        T od[4], nw[4], p[3], qs[6];
#pragma unroll
        for (int c = 0; c < 4; ++c) od[c] = uv[c * sa + s];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = pq[c * sa + s];
#pragma unroll
        for (int k = 0; k < 6; ++k) qs[k] = pq[(3 + k) * sa + s];
        T in[3][4];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int v = cr[r][ax];
            if (v >= 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) in[ax][q] = pq[(kF[ax][q] - F_PX) * sa + v];
            } else if (v <= -2) {
#pragma unroll
                for (int q = 0; q < 4; ++q) in[ax][q] = IO::ld(src.f[kF[ax][q]], -2L - v);
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) in[ax][q] = T(0);
            }
        }
        const T med = s_med[InfoWord::bin(w)];
        if constexpr (REF) {
            const T div_p = (p[0] - in[0][0]) + (p[1] - in[1][0]) + (p[2] - in[2][0]);
            nw[0] = prox_m(od[0] + a.tau * div_p, w, med, a.tl);
            const T dqx = (qs[0] - in[0][1]) + (qs[3] - in[1][1]) + (qs[4] - in[2][1]);
            const T dqy = (qs[3] - in[0][2]) + (qs[1] - in[1][2]) + (qs[5] - in[2][2]);
            const T dqz = (qs[4] - in[0][3]) + (qs[5] - in[1][3]) + (qs[2] - in[2][3]);
            nw[1] = od[1] + a.tau * (p[0] + dqx);
            nw[2] = od[2] + a.tau * (p[1] + dqy);
            nw[3] = od[3] + a.tau * (p[2] + dqz);
        } else {
            T dp = T(0);
            T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const bool hs = InfoWord::has(w, ax);
                dp = hs ? dp + p[ax] : dp;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T w2 = (c == ax) ? T(1) : T(0.5);
                    dq[c] = hs ? dq[c] + w2 * qs[qidx(c, ax) - F_Q0] : dq[c];
                }
            }
            if (cr[r][0] != -1) {
                dp -= in[0][0];
                dq[0] -= T(1) * in[0][1];
                dq[1] -= T(0.5) * in[0][2];
                dq[2] -= T(0.5) * in[0][3];
            }
            if (cr[r][1] != -1) {
                dp -= in[1][0];
                dq[0] -= T(0.5) * in[1][1];
                dq[1] -= T(1) * in[1][2];
                dq[2] -= T(0.5) * in[1][3];
            }
            if (cr[r][2] != -1) {
                dp -= in[2][0];
                dq[0] -= T(0.5) * in[2][1];
                dq[1] -= T(0.5) * in[2][2];
                dq[2] -= T(1) * in[2][3];
            }
            nw[0] = prox_m(od[0] + a.tau * dp, w, med, a.tl);
#pragma unroll
            for (int c = 0; c < 3; ++c) nw[1 + c] = od[1 + c] + a.tau * (p[c] + dq[c]);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            ub[c * sa + s] = IO::rt(T(2) * nw[c] - od[c]);
            uv[c * sa + s] = IO::rt(nw[c]);
        }
    };
    // passes 1-2 of the thread's slot r (k_dual's expressions)
    auto dual = [&](int r) {
        const int s = tid + r * NT;
        T bar[4][4];
#pragma unroll
        for (int c = 0; c < 4; ++c) bar[0][c] = ub[c * sa + s];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            const int v = br[r][ax];
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[1 + ax][c] = v < 0 ? bar[0][c] : ub[c * sa + v];
        }
        T p[3], q[6];
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = pq[c * sa + s];
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = pq[(3 + k) * sa + s];
        dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
        for (int c = 0; c < 3; ++c) pq[c * sa + s] = IO::rtd(p[c]);
#pragma unroll
        for (int k = 0; k < 6; ++k) pq[(3 + k) * sa + s] = IO::rtd(q[k]);
    };
    // primal over G_7, dual over G_6, ..., primal over G_1, dual over G_0
#pragma unroll 1
    for (int k = kHalo4Sets - 1; k >= 1; k -= 2) {
        const int np = ends[k], nd = ends[k - 1];
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lv[r] && tid + r * NT < np) primal(r);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (lv[r] && tid + r * NT < nd) dual(r);
        __syncthreads();
    }
    // ---- the tile's cubes (slots 0 .. TILE - 1: the thread's slot 0)
    if (tid < TILE && lv[0]) {
        const long g = gr[0];
#pragma unroll
        for (int c = 0; c < 4; ++c) IO::stn(dst.f[F_U + c], g, uv[c * sa + tid]);
#pragma unroll
        for (int k = 0; k < 9; ++k) IO::stn(dst.f[F_PX + k], g, pq[k * sa + tid]);
    }
}

// ---- a whole level in one workgroup (no reference counterpart), for levels
// of at most kLevel1 cubes: `iters` iterations in one launch. Thread t owns
// the level's cube t and keeps its u, v, p, q, u_bar, v_bar, table entries and
// frozen neighbours in registers; each iteration publishes u_bar, v_bar, then
// p, q through LDS (two barriers). The dual and primal steps are k_dual's and
// primal_adj_at's expressions on the main set (REF: k_primal_ref's, the
// reference's rules, tgv.cpp:192-217, with the -a neighbours of the
// reference's table; one of another depth gives its p, q from the main set,
// constant while the level iterates), so the state after the launch is that
// of `iters` two-kernel iterations, bit for bit.
// NT threads, the smallest of 64, 256, 1024 that covers the level, so a small
// level's barriers wait for as few waves as it needs.
constexpr int kLevel1 = 1024;

template <class S, int NT = kLevel1, bool REF = false>
__global__ void __launch_bounds__(NT) k_level1(KArgs<S> a, int iters) {
    using IO = Io<S>;
    using T = typename IO::T;
    __shared__ T lub[4][NT];
    __shared__ T lpq[9][NT];
    const int t = (int)threadIdx.x;
    const long g = (long)a.off + (t < (int)a.n ? t : 0);
    const bool run = t < (int)a.n && !masked(a, g);
    auto in_level = [&](int j) { return j >= 0 && (uint32_t)j >= a.off && (uint32_t)j < a.off + a.n; };
    T u = IO::ld(a.s[F_U], g), v[3], p[3], q[6], ub[4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        v[c] = IO::ld(a.s[F_VX + c], g);
        p[c] = IO::ld(a.s[F_PX + c], g);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = IO::ld(a.s[F_Q0 + k], g);
#pragma unroll
    for (int c = 0; c < 4; ++c) ub[c] = IO::ld(a.s[F_UB + c], g);
    const uint32_t w = a.info[g];
    const T med = median_of<T>(w, a.nbins);  // constant over the iterations
    // +a sources: >= 0 an LDS slot, -1 this cube, -2 frozen (values in fz);
    // incoming edges: the LDS slot or -1 (REF: -2 a -a neighbour of another
    // depth, its p_ax and row ax of q in fin)
    constexpr int kFf[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
    int src[3], inc[3];
    T fz[3][4], fin[3][4];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int j = a.nbp[ax][g];
        src[ax] = j < 0 ? -1 : (in_level(j) ? j - (int)a.off : -2);
#pragma unroll
        for (int c = 0; c < 4; ++c) fz[ax][c] = IO::ld(a.s[F_UB + c], j >= 0 ? (long)j : g);
        if constexpr (REF) {
            const int m = a.nbm[ax][g];
            inc[ax] = m < 0 ? -1 : (in_level(m) ? m - (int)a.off : -2);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const T x = IO::ld(a.s[kFf[ax][e]], m >= 0 ? (long)m : g);
                fin[ax][e] = inc[ax] == -2 ? x : T(0);
            }
        } else {
            const int k = a.nbt[ax][g];
            inc[ax] = k < 0 ? -1 : k - (int)a.off;
        }
    }
    // the +a neighbour's slot (this cube's own where there is none: its
    // u_bar, v_bar are published too), read unconditionally, the frozen
    // values selected after the read (no branch per component)
    int nsl[3];
    bool frz[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        nsl[ax] = src[ax] >= 0 ? src[ax] : t;
        frz[ax] = src[ax] == -2;
    }
    for (int it = 0; it < iters; ++it) {
        // passes 1-2 (k_dual)
#pragma unroll
        for (int c = 0; c < 4; ++c) lub[c][t] = ub[c];
        __syncthreads();
        if (run) {
            T bar[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[0][c] = ub[c];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const T x = lub[c][nsl[ax]];
                    bar[1 + ax][c] = frz[ax] ? fz[ax][c] : x;
                }
            }
            dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
            for (int c = 0; c < 3; ++c) p[c] = IO::rtd(p[c]);
#pragma unroll
            for (int k = 0; k < 6; ++k) q[k] = IO::rtd(q[k]);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) lpq[c][t] = p[c];
#pragma unroll
        for (int k = 0; k < 6; ++k) lpq[3 + k][t] = q[k];
        __syncthreads();
        // passes 3-5 (primal_adj_at, then the extrapolation; REF:
        // k_primal_ref's, a missing -a neighbour giving 0)
        if (REF && run) {
            constexpr int kF[3][4] = {{0, 3, 6, 7}, {1, 6, 4, 8}, {2, 7, 8, 5}};  // p_ax, q row ax in lpq
            T in[3][4];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const int k = inc[ax] >= 0 ? inc[ax] : t;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const T x = lpq[kF[ax][e]][k];
                    in[ax][e] = inc[ax] >= 0 ? x : fin[ax][e];
                }
            }
            const T div_p = (p[0] - in[0][0]) + (p[1] - in[1][0]) + (p[2] - in[2][0]);
            T nw[4];
            nw[0] = prox_m(u + a.tau * div_p, w, med, a.tl);
            const T dqx = (q[0] - in[0][1]) + (q[3] - in[1][1]) + (q[4] - in[2][1]);
            const T dqy = (q[3] - in[0][2]) + (q[1] - in[1][2]) + (q[5] - in[2][2]);
            const T dqz = (q[4] - in[0][3]) + (q[5] - in[1][3]) + (q[2] - in[2][3]);
            nw[1] = v[0] + a.tau * (p[0] + dqx);
            nw[2] = v[1] + a.tau * (p[1] + dqy);
            nw[3] = v[2] + a.tau * (p[2] + dqz);
            ub[0] = IO::rt(T(2) * nw[0] - u);
#pragma unroll
            for (int c = 0; c < 3; ++c) ub[1 + c] = IO::rt(T(2) * nw[1 + c] - v[c]);
            u = IO::rt(nw[0]);
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = IO::rt(nw[1 + c]);
        }
        if (!REF && run) {
            T dp = T(0);
            T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const bool hs = InfoWord::has(w, ax);
                dp = hs ? dp + p[ax] : dp;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T w2 = (c == ax) ? T(1) : T(0.5);
                    dq[c] = hs ? dq[c] + w2 * q[qidx(c, ax) - F_Q0] : dq[c];
                }
            }
            constexpr int kF[3][4] = {{0, 3, 6, 7}, {1, 6, 4, 8}, {2, 7, 8, 5}};  // p_ax, q row ax in lpq
            T in[3][4];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const int k = inc[ax] >= 0 ? inc[ax] : t;
#pragma unroll
                for (int e = 0; e < 4; ++e) in[ax][e] = lpq[kF[ax][e]][k];
            }
            if (inc[0] >= 0) {
                dp -= in[0][0];
                dq[0] -= T(1) * in[0][1];
                dq[1] -= T(0.5) * in[0][2];
                dq[2] -= T(0.5) * in[0][3];
            }
            if (inc[1] >= 0) {
                dp -= in[1][0];
                dq[0] -= T(0.5) * in[1][1];
                dq[1] -= T(1) * in[1][2];
                dq[2] -= T(0.5) * in[1][3];
            }
            if (inc[2] >= 0) {
                dp -= in[2][0];
                dq[0] -= T(0.5) * in[2][1];
                dq[1] -= T(0.5) * in[2][2];
                dq[2] -= T(1) * in[2][3];
            }
            T nw[4];
            nw[0] = prox_m(u + a.tau * dp, w, med, a.tl);
#pragma unroll
            for (int c = 0; c < 3; ++c) nw[1 + c] = v[c] + a.tau * (p[c] + dq[c]);
            ub[0] = IO::rt(T(2) * nw[0] - u);
#pragma unroll
            for (int c = 0; c < 3; ++c) ub[1 + c] = IO::rt(T(2) * nw[1 + c] - v[c]);
            u = IO::rt(nw[0]);
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = IO::rt(nw[1 + c]);
        }
    }
    if (!run) return;
    IO::st(a.s[F_U], g, u);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        IO::st(a.s[F_VX + c], g, v[c]);
        IO::st(a.s[F_PX + c], g, p[c]);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::st(a.s[F_Q0 + k], g, q[k]);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_UB + c], g, ub[c]);
}

// ---- a whole level in one cooperative launch (no reference counterpart):
// one cube per thread over as many workgroups as the level needs, all
// co-resident, `iters` iterations with two grid-wide barriers each. Thread
// state (u, v, p, q, u_bar, v_bar, table entries) stays in registers; the
// exchange goes through the main set: u_bar, v_bar are published, barrier,
// the dual step reads the +a neighbours' (k_dual's expressions), p, q are
// published, barrier, the primal step reads the incoming ones
// (primal_adj_at's expressions). A value is never overwritten before the
// barrier that follows its last read, so the state after the launch is that
// of `iters` two-kernel iterations, bit for bit.
//
// Barrier: one arrival counter per launch (cleared by the host before it);
// barrier k is complete once the counter reaches (k + 1) * gridDim.x. The
// arrival releases and the wait acquires at agent scope: the XCDs' L2 caches
// are not coherent with each other, so the release writes the arriving L2
// back and the acquire invalidates the waiting one. The wait is bounded:
// after spin_max polls, or once another workgroup has given up, it sets
// *err and every workgroup returns, so the launch always ends; the host
// reads *err and fails the solve.
struct GridBar {
    uint32_t* cnt;
    int32_t* err;
    uint32_t spin_max;
};

__device__ inline bool grid_barrier(const GridBar& b, uint32_t k) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t target = (k + 1u) * gridDim.x;
        __hip_atomic_fetch_add(b.cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        int ok = 1;
        uint32_t polls = 0;
        while (__hip_atomic_load(b.cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            ++polls;
            if (polls > b.spin_max ||
                ((polls & 63u) == 0u && __hip_atomic_load(b.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                __hip_atomic_store(b.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

template <class S, int B>
__global__ void __launch_bounds__(B) k_persist(KArgs<S> a, int iters, GridBar gb) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = blockIdx.x * (uint32_t)B + threadIdx.x;
    const bool inr = t < a.n;
    const long g = (long)a.off + (inr ? t : 0u);
    const bool run = inr && !masked(a, g);
    T u = T(0), v[3] = {T(0), T(0), T(0)}, p[3] = {T(0), T(0), T(0)}, ub[4] = {T(0), T(0), T(0), T(0)};
    T q[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    uint32_t w = 0u;
    int nb[3] = {-1, -1, -1}, inc[3] = {-1, -1, -1};
    if (run) {
        u = IO::ld(a.s[F_U], g);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            v[c] = IO::ld(a.s[F_VX + c], g);
            p[c] = IO::ld(a.s[F_PX + c], g);
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) q[k] = IO::ld(a.s[F_Q0 + k], g);
#pragma unroll
        for (int c = 0; c < 4; ++c) ub[c] = IO::ld(a.s[F_UB + c], g);
        w = a.info[g];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            nb[ax] = a.nbp[ax][g];
            inc[ax] = a.nbt[ax][g];
        }
    }
    uint32_t k = 0;  // barriers passed
    for (int it = 0; it < iters; ++it) {
        // passes 1-2 (k_dual): a missing +a neighbour gives the cube's own
        // value; one outside the level gives its u_bar, v_bar, which no
        // launch of this level writes
        if (run) {
#pragma unroll
            for (int c = 0; c < 4; ++c) IO::st(a.s[F_UB + c], g, ub[c]);
        }
        if (!grid_barrier(gb, k++)) return;
        if (run) {
            T bar[4][4];
#pragma unroll
            for (int c = 0; c < 4; ++c) bar[0][c] = ub[c];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
                for (int c = 0; c < 4; ++c) bar[1 + ax][c] = nb[ax] < 0 ? ub[c] : IO::ld(a.s[F_UB + c], nb[ax]);
            }
            dual_step(p, q, bar, a.sigma, a.a0, a.a1);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                p[c] = IO::rtd(p[c]);
                IO::st(a.s[F_PX + c], g, p[c]);
            }
#pragma unroll
            for (int m = 0; m < 6; ++m) {
                q[m] = IO::rtd(q[m]);
                IO::st(a.s[F_Q0 + m], g, q[m]);
            }
        }
        if (!grid_barrier(gb, k++)) return;
        // passes 3-5 (primal_adj_at, then the extrapolation)
        if (run) {
            T dp = T(0);
            T dq[3] = {T(0), T(0), T(0)};
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const bool hs = InfoWord::has(w, ax);
                dp = hs ? dp + p[ax] : dp;
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const T w2 = (c == ax) ? T(1) : T(0.5);
                    dq[c] = hs ? dq[c] + w2 * q[qidx(c, ax) - F_Q0] : dq[c];
                }
            }
            constexpr int kF[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
            T in[3][4];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
                const long s = inc[ax] >= 0 ? (long)inc[ax] : g;
#pragma unroll
                for (int e = 0; e < 4; ++e) in[ax][e] = IO::ld(a.s[kF[ax][e]], s);
            }
            if (inc[0] >= 0) {
                dp -= in[0][0];
                dq[0] -= T(1) * in[0][1];
                dq[1] -= T(0.5) * in[0][2];
                dq[2] -= T(0.5) * in[0][3];
            }
            if (inc[1] >= 0) {
                dp -= in[1][0];
                dq[0] -= T(0.5) * in[1][1];
                dq[1] -= T(1) * in[1][2];
                dq[2] -= T(0.5) * in[1][3];
            }
            if (inc[2] >= 0) {
                dp -= in[2][0];
                dq[0] -= T(0.5) * in[2][1];
                dq[1] -= T(0.5) * in[2][2];
                dq[2] -= T(1) * in[2][3];
            }
            T nw[4];
            nw[0] = prox(u + a.tau * dp, w, a.nbins, a.tl);
#pragma unroll
            for (int c = 0; c < 3; ++c) nw[1 + c] = v[c] + a.tau * (p[c] + dq[c]);
            ub[0] = IO::rt(T(2) * nw[0] - u);
#pragma unroll
            for (int c = 0; c < 3; ++c) ub[1 + c] = IO::rt(T(2) * nw[1 + c] - v[c]);
            u = IO::rt(nw[0]);
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = IO::rt(nw[1 + c]);
        }
    }
    if (!run) return;
    IO::st(a.s[F_U], g, u);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::st(a.s[F_VX + c], g, v[c]);
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_UB + c], g, ub[c]);
}

// `nbar` grid barriers and nothing else: the cost of one barrier of k_persist
// at a given grid (no reference counterpart).
template <int B>
__global__ void __launch_bounds__(B) k_barrier_bench(GridBar gb, int nbar) {
    for (int k = 0; k < nbar; ++k)
        if (!grid_barrier(gb, (uint32_t)k)) return;
}

// ---- streaming mode (tgv_minimize_streaming, tgv.cpp:515-1070)
// The cube file's bisection (CubeFileReader::lookup, tgv.cpp:674-683): the
// first probe whose code matches, not necessarily the last cube with it.
__device__ inline int64_t bisect_code(const uint64_t* code, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        const uint64_t k = code[mid];
        if (k == key) return mid;
        if (k < key) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}
// The leaf map (tgv.cpp:916-920): the last cube of [f0, f1] with this code;
// the cube array is sorted by code.
__device__ inline int64_t leaf_code(const uint64_t* code, uint32_t f0, uint32_t f1, uint64_t key) {
    uint32_t lo = f0, hi = f1 + 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (code[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return (lo > f0 && code[lo - 1] == key) ? (int64_t)(lo - 1) : -1;
}

// Per input cube i (codes and depths in input order): its six neighbours as
// its leaf's table has them (tgv.cpp:925-975), the flags of KArgs::sflag and
// the bisection's parent (tgv.cpp:877-887), all stored at i's position.
__global__ void k_stream_tables(const uint64_t* __restrict__ code, const int32_t* __restrict__ depth, uint32_t n,
                                const int32_t* __restrict__ leaf_of, const uint32_t* __restrict__ lfirst,
                                const uint32_t* __restrict__ llast, const uint32_t* __restrict__ inv, int32_t* nb0,
                                int32_t* nb1, int32_t* nb2, int32_t* nb3, int32_t* nb4, int32_t* nb5,
                                uint8_t* __restrict__ flag, int32_t* __restrict__ par) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t g = inv[i];
    int32_t* out[6] = {nb0, nb1, nb2, nb3, nb4, nb5};
    const int l = leaf_of[i];
    if (l < 0) {
        for (int f = 0; f < 6; ++f) out[f][g] = -1;
        flag[g] = 0;
        par[g] = -1;
        return;
    }
    const uint32_t f0 = lfirst[l], f1 = llast[l];
    const uint64_t c = code[i];
    const int d = depth[i];
    uint32_t fl = 64u;
    for (int f = 0; f < 6; ++f) {
        const uint64_t k = face_nb(c, f, d);
        int64_t j = -1;
        if (k != kNone) {
            j = leaf_code(code, f0, f1, k);
            if (j < 0) j = bisect_code(code, n, k);
            if (j < 0 && d > 0) {
                const uint64_t pk = ancestor(k, d - 1);
                j = leaf_code(code, f0, f1, pk);
                if (j < 0) j = bisect_code(code, n, pk);
            }
        }
        out[f][g] = j >= 0 ? (int32_t)inv[j] : -1;
        if (j >= 0) {
            const bool in = (uint64_t)j >= f0 && (uint64_t)j <= f1;
            if ((f & 1) == 0 && in && depth[j] == d) fl |= 1u << f;
            if ((f & 1) == 1 && !in) fl |= 1u << f;
        }
    }
    flag[g] = (uint8_t)fl;
    const int64_t pj = d > 0 ? bisect_code(code, n, ancestor(c, d - 1)) : -1;
    par[g] = pj >= 0 ? (int32_t)inv[pj] : -1;
}

__global__ void k_stream_inputs(const uint64_t* __restrict__ pcode, const int32_t* __restrict__ pdepth,
                                const uint32_t* __restrict__ inv, uint32_t n, uint64_t* __restrict__ code,
                                int32_t* __restrict__ depth) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    code[i] = pcode[inv[i]];
    depth[i] = pdepth[inv[i]];
}

// The streaming tables of input cubes [i0, i0 + cn) in input order (global
// input indices), for the out-of-core solve: the same lookups as
// k_stream_tables.
__global__ void k_stream_tables_ooc(const uint64_t* __restrict__ code, const int32_t* __restrict__ depth, uint32_t n,
                                    const int32_t* __restrict__ leaf_of, const uint32_t* __restrict__ lfirst,
                                    const uint32_t* __restrict__ llast, uint32_t i0, uint32_t cn,
                                    int32_t* __restrict__ nb, uint8_t* __restrict__ flag, int32_t* __restrict__ par) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cn) return;
    const uint32_t i = i0 + t;
    const int l = leaf_of[i];
    if (l < 0) {
        for (int f = 0; f < 6; ++f) nb[6 * (size_t)t + f] = -1;
        flag[t] = 0;
        par[t] = -1;
        return;
    }
    const uint32_t f0 = lfirst[l], f1 = llast[l];
    const uint64_t c = code[i];
    const int d = depth[i];
    uint32_t fl = 64u;
    for (int f = 0; f < 6; ++f) {
        const uint64_t k = face_nb(c, f, d);
        int64_t j = -1;
        if (k != kNone) {
            j = leaf_code(code, f0, f1, k);
            if (j < 0) j = bisect_code(code, n, k);
            if (j < 0 && d > 0) {
                const uint64_t pk = ancestor(k, d - 1);
                j = leaf_code(code, f0, f1, pk);
                if (j < 0) j = bisect_code(code, n, pk);
            }
        }
        nb[6 * (size_t)t + f] = (int32_t)j;
        if (j >= 0) {
            const bool in = (uint64_t)j >= f0 && (uint64_t)j <= f1;
            if ((f & 1) == 0 && in && depth[j] == d) fl |= 1u << f;
            if ((f & 1) == 1 && !in) fl |= 1u << f;
        }
    }
    flag[t] = (uint8_t)fl;
    const int64_t pj = d > 0 ? bisect_code(code, n, ancestor(c, d - 1)) : -1;
    par[t] = (int32_t)pj;
}

// Parent copy of one leaf's part of a level (tgv.cpp:895-911): read every
// parent first, then write.
template <class S>
__global__ void k_stream_parent_read(KArgs<S> a, const int32_t* __restrict__ par, typename Io<S>::T* __restrict__ tmp) {
    using IO = Io<S>;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    const int j = par[g];
#pragma unroll
    for (int c = 0; c < 4; ++c) tmp[4 * t + c] = j >= 0 ? IO::ld(a.s[F_U + c], j) : IO::ld(a.s[F_U + c], g);
}
template <class S>
__global__ void k_stream_parent_write(KArgs<S> a, const typename Io<S>::T* __restrict__ tmp) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_U + c], g, tmp[4 * t + c]);
#pragma unroll
    for (int f = F_PX; f <= F_Q5; ++f) IO::st(a.s[f], g, T(0));
}
// Level start (tgv.cpp:914, 1041-1056): every cube's bars and snapshot are
// its (u, v).
template <class S>
__global__ void k_stream_snapshot(KArgs<S> a, uint32_t n) {
    using IO = Io<S>;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const typename IO::T x = IO::ld(a.s[F_U + c], g);
        IO::st(a.s[F_UB + c], g, x);
        IO::st(const_cast<S*>(a.fz[c]), g, x);
    }
}

template <class S>
__global__ void k_fz_aos(KArgs<S> a, S* __restrict__ rec) {
    using IO = Io<S>;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.fz_n) return;
    const long g = (long)a.fz_off + t;
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(rec, 4l * t + c, IO::ld(a.s[F_UB + c], g));
}

// The reference rules' out-of-level neighbour values of a level (rfz_bar,
// rfz_dual of KArgs), gathered once at level start: the two-iteration kernel
// then reads one record per such neighbour, next to its own cube's, instead
// of four scattered fields per iteration. count (when not null): the level's
// out-of-level references are counted into *count (one atomic per
// workgroup) and nothing is gathered. SNAP (frozen-border leaves, where the
// range [lvl_off, lvl_off + lvl_n) is one treetop leaf's part of a level):
// u_bar, v_bar come from the level-start snapshot fz (run_level_frozen,
// tgv.cpp:274-309), p, q from the main set (live, constant while the leaf
// iterates).
template <class S, bool SNAP = false>
__global__ void __launch_bounds__(256) k_ref_frozen(KArgs<S> a, S* __restrict__ bar, S* __restrict__ dual,
                                                    uint32_t* __restrict__ count) {
    using IO = Io<S>;
    __shared__ uint32_t s_cnt;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = t < a.lvl_n;
    const long g = (long)a.lvl_off + (in ? t : 0u);
    auto out = [&](int j) { return j >= 0 && !((uint32_t)j - a.lvl_off < a.lvl_n); };
    constexpr int kF[3][4] = {{F_PX, F_Q0, F_Q3, F_Q4}, {F_PY, F_Q3, F_Q1, F_Q5}, {F_PZ, F_Q4, F_Q5, F_Q2}};
    if (count) {
        if (threadIdx.x == 0) s_cnt = 0u;
        __syncthreads();
        uint32_t k = 0;
        if (in) {
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) k += (out(a.nbp[ax][g]) ? 1u : 0u) + (out(a.nbm[ax][g]) ? 1u : 0u);
        }
        if (k) atomicAdd(&s_cnt, k);
        __syncthreads();
        if (threadIdx.x == 0 && s_cnt) atomicAdd(count, s_cnt);
        return;
    }
    if (!in) return;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const long r = (long)ax * a.lvl_n + t;
        const int jp = a.nbp[ax][g], jm = a.nbm[ax][g];
        if (out(jp)) {
#pragma unroll
            for (int c = 0; c < 4; ++c) IO::st(bar, 4 * r + c, IO::ld(SNAP ? a.fz[c] : a.s[F_UB + c], jp));
        }
        if (out(jm)) {
#pragma unroll
            for (int e = 0; e < 4; ++e) IO::st(dual, 4 * r + e, IO::ld(a.s[kF[ax][e]], jm));
        }
    }
}

// ---- initialisation
// tgv.cpp:372-385: u = median, everything else 0, u_bar = u.
template <class S>
// p, q are cleared at the first npq positions only: a caller passes fewer
// than n when every later level sets its own p, q before anything reads them
// (exact-adjoint solves, whose other-depth reads are u_bar, v_bar only).
__global__ void k_reset(KArgs<S> a, uint32_t n, uint32_t npq) {
    using IO = Io<S>;
    using T = typename IO::T;
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const T m = median_of<T>(a.info[g], a.nbins);
    IO::st(a.s[F_U], g, m);
    IO::st(a.s[F_UB], g, m);
#pragma unroll
    for (int f = F_VX; f <= F_VZ; ++f) IO::st(a.s[f], g, T(0));
    if (g < (long)npq) {
#pragma unroll
        for (int f = F_PX; f <= F_Q5; ++f) IO::st(a.s[f], g, T(0));
    }
#pragma unroll
    for (int f = F_VBX; f <= F_VBZ; ++f) IO::st(a.s[f], g, T(0));
}

// tgv.cpp:391-409: u, v from the cube holding the parent code (if any), p = q
// = 0, u_bar = u, v_bar = v. A parent found at the same depth holds the code
// of its own parent and copies itself, so reading it while it is written
// returns the same value: the gather order does not matter.
// FULL = false (the level opens with the head launch of the two-iteration
// kernel, which reads u, v only): u, v alone. rec (when not null): the
// previous level's u_bar, v_bar records of k_fz_aos, in the same launch.
template <class S, bool FULL = true>
__global__ void k_parent_init(KArgs<S> a, S* __restrict__ rec) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (rec && t < a.fz_n) {
        const long f = (long)a.fz_off + t;
#pragma unroll
        for (int c = 0; c < 4; ++c) IO::st(rec, 4l * t + c, IO::ld(a.s[F_UB + c], f));
    }
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    const int pi = a.parent[g];
    const long src = pi >= 0 ? pi : g;
    const T u = IO::ld(a.s[F_U], src), vx = IO::ld(a.s[F_VX], src), vy = IO::ld(a.s[F_VY], src),
            vz = IO::ld(a.s[F_VZ], src);
    IO::st(a.s[F_U], g, u);
    IO::st(a.s[F_VX], g, vx);
    IO::st(a.s[F_VY], g, vy);
    IO::st(a.s[F_VZ], g, vz);
    if (!FULL) return;
#pragma unroll
    for (int f = F_PX; f <= F_Q5; ++f) IO::st(a.s[f], g, T(0));
    IO::st(a.s[F_UB], g, u);
    IO::st(a.s[F_VBX], g, vx);
    IO::st(a.s[F_VBY], g, vy);
    IO::st(a.s[F_VBZ], g, vz);
}

// Warm start: u, v from a given field; mode 1 also saves the current u, v of
// the level first, mode 2 restores the saved u, v instead; p, q are zeroed,
// or left as they are with keep_dual.
template <class S>
__global__ void k_field_init(KArgs<S> a, int mode, int keep_dual) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    const int st = a.init_stride;
    T x[4];
    if (mode == 2) {
#pragma unroll
        for (int c = 0; c < 4; ++c) x[c] = IO::ld(a.save[c], t);
    } else {
        if (mode == 1) {
#pragma unroll
            for (int c = 0; c < 4; ++c) IO::st(a.save[c], t, IO::ld(a.s[F_U + c], g));
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) x[c] = (T)a.init[g * st + c];
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_U + c], g, x[c]);
    if (mode != 2 && st == 13) {
        // a whole state (u, v, p, q): its dual too
#pragma unroll
        for (int f = F_PX; f <= F_Q5; ++f) IO::st(a.s[f], g, (T)a.init[g * 13 + (f - F_U)]);
    } else if (!keep_dual) {
#pragma unroll
        for (int f = F_PX; f <= F_Q5; ++f) IO::st(a.s[f], g, T(0));
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) IO::st(a.s[F_UB + c], g, x[c]);
}

// Dual initialisation from the level's initial (u, v) (no reference
// counterpart): p = alpha1 K1 / max(|K1|, eps), q = alpha0 e / max(|e|, eps)
// with K1 = grad u - v and e = eps(v) of the depth-aware operators, frozen
// values outside the level. These are the optimality conditions wherever the
// residuals are non-zero, so a good primal guess yields a good dual one.
template <class S>
__global__ void k_dual_recover(KArgs<S> a, typename Io<S>::T eps) {
    using IO = Io<S>;
    using T = typename IO::T;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.n) return;
    const long g = (long)a.off + t;
    auto in_lvl = [&](int j) { return (uint32_t)j >= a.lvl_off && (uint32_t)j < a.lvl_off + a.lvl_n; };
    const T u = IO::ld(a.s[F_U], g);
    const T v[3] = {IO::ld(a.s[F_VX], g), IO::ld(a.s[F_VY], g), IO::ld(a.s[F_VZ], g)};
    T gr[3], dv[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int j = a.nba[ax][g];
        const bool in = j >= 0 && in_lvl(j);
        const long sj = j >= 0 ? j : g;
        gr[ax] = j < 0 ? T(0) : (in ? IO::ld(a.s[F_U], sj) : IO::ld(a.s[F_UB], sj)) - u;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            dv[c][ax] = j < 0 ? T(0) : (in ? IO::ld(a.s[F_VX + c], sj) : IO::ld(a.s[F_VBX + c], sj)) - v[c];
    }
    const T k1[3] = {gr[0] - v[0], gr[1] - v[1], gr[2] - v[2]};
    const T e[6] = {dv[0][0], dv[1][1], dv[2][2], T(0.5) * (dv[0][1] + dv[1][0]), T(0.5) * (dv[0][2] + dv[2][0]),
                    T(0.5) * (dv[1][2] + dv[2][1])};
    const T n1 = sqrt(k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
    T n0 = T(0);
#pragma unroll
    for (int k = 0; k < 6; ++k) n0 += e[k] * e[k];
    n0 = sqrt(n0);
    const T s1 = a.a1 / (n1 < eps ? eps : n1), s0 = a.a0 / (n0 < eps ? eps : n0);
#pragma unroll
    for (int c = 0; c < 3; ++c) IO::st(a.s[F_PX + c], g, k1[c] * s1);
#pragma unroll
    for (int k = 0; k < 6; ++k) IO::st(a.s[F_Q0 + k], g, e[k] * s0);
}

// ---- energy and duality gap (no reference counterpart)
// Per active cube i (neighbours outside the level frozen at u_bar / v_bar):
//   P_i = lambda W |u - m| + alpha1 |K1(u, v)| + alpha0 |eps(v)|
//   D_i = <frozen part, (p, q)> + min_{u in [-1,1]} (lambda W |u - m| + c_u u)
//         - V (|c_vx| + |c_vy| + |c_vz|)
// with c_u, c_v the coefficients of u_i, v_i in the Lagrangian. Summed in FP64
// per block, then over blocks in a fixed order.
constexpr int kGapBlock = 256;

// Contribution of cube g to the level's energy P and dual value D.
template <class S>
__device__ inline void gap_cube(const KArgs<S>& a, long g, double lam, double a0, double a1, double vbox, double& P,
                                double& D) {
    using IO = Io<S>;
    using T = typename IO::T;
    auto in_lvl = [&](int j) { return (uint32_t)j >= a.lvl_off && (uint32_t)j < a.lvl_off + a.lvl_n; };
    auto L = [&](int f, long j) -> double { return (double)IO::ld(a.s[f], j); };
    const uint32_t wd = a.info[g];
    const uint32_t Wc = InfoWord::W(wd);
    const double W = (double)Wc, m = (double)median_of<T>(wd, a.nbins), ui = L(F_U, g);
    const double v[3] = {L(F_VX, g), L(F_VY, g), L(F_VZ, g)};
    const double p[3] = {L(F_PX, g), L(F_PY, g), L(F_PZ, g)};
    double q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) q[k] = L(F_Q0 + k, g);
    // every load unconditional (the cube's own address where a neighbour is
    // missing, the value then unused), so that they are all in flight at
    // once; the sums below skip the same terms in the same order
    int jn[3], kn[3];
    bool in[3];
    double nu[3], nv[3][3], ip[3], iq[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        jn[ax] = a.nba[ax][g];
        kn[ax] = a.nbt[ax][g];
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const int j = jn[ax], k = kn[ax];
        const long jj = j >= 0 ? j : g, kk = k >= 0 ? k : g;
        // a neighbour in the level gives its u, v, one outside it (frozen)
        // its u_bar, v_bar
        in[ax] = j >= 0 && in_lvl(j);
        nu[ax] = (double)IO::ld(in[ax] ? a.s[F_U] : a.s[F_UB], jj);
#pragma unroll
        for (int c = 0; c < 3; ++c) nv[ax][c] = (double)IO::ld(in[ax] ? a.s[F_VX + c] : a.s[F_VBX + c], jj);
        ip[ax] = L(F_PX + ax, kk);
#pragma unroll
        for (int c = 0; c < 3; ++c) iq[ax][c] = L(qidx(c, ax), kk);
    }
    double gr[3], dv[3][3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const bool h = jn[ax] >= 0;
        gr[ax] = h ? nu[ax] - ui : 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) dv[c][ax] = h ? nv[ax][c] - v[c] : 0.0;
    }
    const double k1[3] = {gr[0] - v[0], gr[1] - v[1], gr[2] - v[2]};
    const double e[6] = {dv[0][0], dv[1][1], dv[2][2], 0.5 * (dv[0][1] + dv[1][0]), 0.5 * (dv[0][2] + dv[2][0]),
                         0.5 * (dv[1][2] + dv[2][1])};
    const double nk = sqrt(k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
    double ne = 0.0;
    for (int k = 0; k < 6; ++k) ne += e[k] * e[k];
    ne = sqrt(ne);
    P = (Wc ? lam * W * fabs(ui - m) : 0.0) + a1 * nk + a0 * ne;

    double cst = 0.0, cu = 0.0, cv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (jn[ax] < 0) continue;
        cu -= p[ax];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double w = (c == ax) ? 1.0 : 0.5;
            cv[c] -= w * q[qidx(c, ax) - F_Q0];
        }
        if (!in[ax]) {
            cst += p[ax] * nu[ax];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double w = (c == ax) ? 1.0 : 0.5;
                cst += w * q[qidx(c, ax) - F_Q0] * nv[ax][c];
            }
        }
    }
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        if (kn[ax] < 0) continue;
        cu += ip[ax];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double w = (c == ax) ? 1.0 : 0.5;
            cv[c] += w * iq[ax][c];
        }
    }
    for (int c = 0; c < 3; ++c) cv[c] -= p[c];
    double mu;
    if (Wc == 0u) {
        mu = -fabs(cu);
    } else {
        const double g0 = lam * W * fabs(-1.0 - m) + cu * -1.0;
        const double g1 = lam * W * fabs(m - m) + cu * m;
        const double g2 = lam * W * fabs(1.0 - m) + cu * 1.0;
        mu = fmin(g0, fmin(g1, g2));
    }
    D = cst + mu - vbox * (fabs(cv[0]) + fabs(cv[1]) + fabs(cv[2]));
}

template <class S>
__global__ void __launch_bounds__(kGapBlock) k_gap(KArgs<S> a, double lam, double a0, double a1, double vbox,
                                                    double* part) {
    __shared__ double sp[kGapBlock], sd[kGapBlock];
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    double P = 0.0, D = 0.0;
    if (t < a.n) gap_cube(a, (long)a.off + t, lam, a0, a1, vbox, P, D);
    sp[threadIdx.x] = P;
    sd[threadIdx.x] = D;
    __syncthreads();
    for (int s = kGapBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sp[threadIdx.x] += sp[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sp[0];
        part[2 * blockIdx.x + 1] = sd[0];
    }
}

// Per-block energy and dual value of a level shared by several blocks:
// workgroup b sums its block's segment [seg[b], seg[b+1]) of the level.
template <class S>
__global__ void __launch_bounds__(kGapBlock) k_gap_seg(KArgs<S> a, const uint32_t* __restrict__ seg, double lam,
                                                        double a0, double a1, double vbox, double* out) {
    __shared__ double sp[kGapBlock], sd[kGapBlock];
    const uint32_t b = blockIdx.x;
    double P = 0.0, D = 0.0;
    for (uint32_t t = seg[b] + threadIdx.x; t < seg[b + 1]; t += kGapBlock) {
        double p1, d1;
        gap_cube(a, (long)a.off + t, lam, a0, a1, vbox, p1, d1);
        P += p1;
        D += d1;
    }
    sp[threadIdx.x] = P;
    sd[threadIdx.x] = D;
    __syncthreads();
    for (int s = kGapBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sp[threadIdx.x] += sp[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * b] = sp[0];
        out[2 * b + 1] = sd[0];
    }
}

__global__ void __launch_bounds__(kGapBlock) k_sum_pairs(const double* part, uint32_t nparts, double* out) {
    __shared__ double sp[kGapBlock], sd[kGapBlock];
    double P = 0.0, D = 0.0;
    // eight pairs' loads in flight at a time; the sums keep their order
    constexpr int kBatch = 8;
    uint32_t i = threadIdx.x;
    for (; i + (kBatch - 1) * kGapBlock < nparts; i += kBatch * kGapBlock) {
        double x[2 * kBatch];
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            x[2 * k] = part[2 * (i + k * kGapBlock)];
            x[2 * k + 1] = part[2 * (i + k * kGapBlock) + 1];
        }
#pragma unroll
        for (int k = 0; k < kBatch; ++k) {
            P += x[2 * k];
            D += x[2 * k + 1];
        }
    }
    for (; i < nparts; i += kGapBlock) {
        P += part[2 * i];
        D += part[2 * i + 1];
    }
    sp[threadIdx.x] = P;
    sd[threadIdx.x] = D;
    __syncthreads();
    for (int s = kGapBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sp[threadIdx.x] += sp[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sp[0];
        out[1] = sd[0];
    }
}

// ---- neighbour tables on the device (tgv.cpp:97-133)
// Lookup in the code map: ucode ascending, uidx = position of the last input
// cube with that code (the reference's unordered_map "last writer wins").
__device__ inline int code_lookup(const uint64_t* uc, const int32_t* ui, uint32_t U, uint64_t key) {
    // the last entry with the key (the arrays may repeat a code: the last
    // writer wins, as in the reference's map)
    uint32_t lo = 0, hi = U;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (uc[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return (lo > 0 && uc[lo - 1] == key) ? ui[lo - 1] : -1;
}

// Several independent octrees (blocks) may share one problem: each position's
// lookups are confined to its block's segment [uoff[b], uoff[b+1]) of the
// sorted code array (blk == nullptr: one block, the whole array).
__global__ void k_tables(const uint64_t* __restrict__ codes, const int32_t* __restrict__ depth, uint32_t n,
                         const uint64_t* __restrict__ uc0, const int32_t* __restrict__ ui0, uint32_t U0,
                         const int32_t* __restrict__ blk, const uint32_t* __restrict__ uoff,
                         int32_t* nb0, int32_t* nb1, int32_t* nb2, int32_t* nb3, int32_t* nb4, int32_t* nb5,
                         int32_t* parent) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint64_t c = codes[g];
    const int d = depth[g];
    const uint32_t lo = blk ? uoff[blk[g]] : 0u, U = blk ? uoff[blk[g] + 1] - lo : U0;
    const uint64_t* uc = uc0 + lo;
    const int32_t* ui = ui0 + lo;
    int32_t* out[6] = {nb0, nb1, nb2, nb3, nb4, nb5};
    for (int f = 0; f < 6; ++f) {
        const uint64_t k = face_nb(c, f, d);
        int j = -1;
        if (k != kNone) {
            j = code_lookup(uc, ui, U, k);
            if (j < 0 && d > 0) j = code_lookup(uc, ui, U, ancestor(k, d - 1));
        }
        out[f][g] = j;
    }
    parent[g] = d > 0 ? code_lookup(uc, ui, U, ancestor(c, d - 1)) : -1;
}

// Depth-aware table of the exact-adjoint mode (no reference counterpart):
// dc / dd ascending by (code, depth), dp = position of the last input cube
// with that pair. Same-depth neighbour, else the cube one level coarser,
// else -1; duplicates of a (code, depth) pair other than the last get none.
__device__ inline int pair_lookup(const uint64_t* dc, const int32_t* dd, const int32_t* dp, uint32_t U, uint64_t key,
                                  int depth) {
    uint32_t lo = 0, hi = U;  // the last entry with the pair
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (dc[mid] < key || (dc[mid] == key && dd[mid] <= depth)) lo = mid + 1;
        else hi = mid;
    }
    return (lo > 0 && dc[lo - 1] == key && dd[lo - 1] == depth) ? dp[lo - 1] : -1;
}

__global__ void k_tables_geo(const uint64_t* __restrict__ codes, const int32_t* __restrict__ depth, uint32_t n,
                             const uint64_t* __restrict__ dc0, const int32_t* __restrict__ dd0,
                             const int32_t* __restrict__ dp0, uint32_t U0, const int32_t* __restrict__ blk,
                             const uint32_t* __restrict__ doff, int32_t* nba0, int32_t* nba1, int32_t* nba2,
                             int32_t* nbm0, int32_t* nbm1, int32_t* nbm2, int32_t* parent, uint32_t* info) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint64_t c = codes[g];
    const int d = depth[g];
    const uint32_t lo = blk ? doff[blk[g]] : 0u, U = blk ? doff[blk[g] + 1] - lo : U0;
    const uint64_t* dc = dc0 + lo;
    const int32_t* dd = dd0 + lo;
    const int32_t* dp = dp0 + lo;
    const bool shadow = pair_lookup(dc, dd, dp, U, c, d) != (int)g;
    int32_t* out[3] = {nba0, nba1, nba2};
    uint32_t flags = 0;
    for (int ax = 0; ax < 3; ++ax) {
        int j = -1;
        const uint64_t k = shadow ? kNone : face_nb(c, 2 * ax, d);
        if (k != kNone) {
            j = pair_lookup(dc, dd, dp, U, k, d);
            if (j < 0 && d > 0) j = pair_lookup(dc, dd, dp, U, ancestor(k, d - 1), d - 1);
        }
        out[ax][g] = j;
        if (j >= 0) flags |= 1u << (29 + ax);
    }
    // -x, -y, -z faces (read only by the warm-start features)
    int32_t* outm[3] = {nbm0, nbm1, nbm2};
    for (int ax = 0; ax < 3; ++ax) {
        int j = -1;
        const uint64_t k = shadow ? kNone : face_nb(c, 2 * ax + 1, d);
        if (k != kNone) {
            j = pair_lookup(dc, dd, dp, U, k, d);
            if (j < 0 && d > 0) j = pair_lookup(dc, dd, dp, U, ancestor(k, d - 1), d - 1);
        }
        outm[ax][g] = j;
    }
    info[g] |= flags;
    parent[g] = d > 0 ? pair_lookup(dc, dd, dp, U, ancestor(c, d - 1), d - 1) : -1;
}

// The state of position g as the records of input cube perm[g]: (u, v, p, q)
// in st (13 values), (u_bar, v_bar) in bar (4), widened to the arithmetic type.
template <class S>
__global__ void k_state_out(KArgs<S> a, const uint32_t* __restrict__ perm, uint32_t n,
                            typename Io<S>::T* __restrict__ st, typename Io<S>::T* __restrict__ bar) {
    using IO = Io<S>;
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const size_t i = perm[g];
    if (st) {
#pragma unroll
        for (int f = 0; f < 13; ++f) st[13 * i + f] = IO::ld(a.s[f], g);
    }
    if (bar) {
#pragma unroll
        for (int k = 0; k < 4; ++k) bar[4 * i + k] = IO::ld(a.s[F_UB + k], g);
    }
}

__global__ void k_fill_i32(int32_t* p, uint32_t n, int32_t v) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) p[g] = v;
}

// The per-position inputs of a load, from the cubes and histograms in input
// order and the position -> input permutation: code, depth, the histogram
// (nb bins) and the packed word of vote total and median bin
// (histogram.cpp:21-31: the first k with 2 * cumulative >= total). A total
// above the word's 24 bits sets *bad. icode, idepth (when not null): code and
// depth in input order.
__global__ void k_load_positions(const tgvx_cube* __restrict__ cubes, const uint32_t* __restrict__ bins,
                                 const uint32_t* __restrict__ perm, uint32_t n, int nb, uint64_t* __restrict__ code,
                                 int32_t* __restrict__ depth, uint32_t* __restrict__ pbins,
                                 uint32_t* __restrict__ info, int32_t* __restrict__ bad,
                                 uint64_t* __restrict__ icode, int32_t* __restrict__ idepth) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint32_t i = perm[g];
    const tgvx_cube cb = cubes[i];
    const uint64_t cc = ((uint64_t)cb.hi << 32) | cb.lo;
    code[g] = cc;
    depth[g] = cb.depth;
    if (icode) {  // the same in input order
        icode[i] = cc;
        idepth[i] = cb.depth;
    }
    const uint32_t* b = bins + (size_t)i * nb;
    uint32_t* o = pbins + (size_t)g * nb;
    uint64_t tot = 0;
    for (int k = 0; k < nb; ++k) {
        const uint32_t x = b[k];
        o[k] = x;
        tot += x;
    }
    if (tot > InfoWord::kWMax) {
        atomicOr(bad, 1);
        tot = 0;
    }
    uint32_t kmed = 0, cum = 0;
    for (int k = 0; k < nb; ++k) {
        cum += b[k];
        if (cum * 2u >= (uint32_t)tot) {
            kmed = (uint32_t)k;
            break;
        }
    }
    info[g] = (uint32_t)tot | (kmed << 24);
}

// The depth-aware table is one-to-one within a level, so each target has at
// most one writer here.
__global__ void k_transpose(const int32_t* __restrict__ nba0, const int32_t* __restrict__ nba1,
                            const int32_t* __restrict__ nba2, const int32_t* __restrict__ depth, uint32_t n,
                            int32_t* nbt0, int32_t* nbt1, int32_t* nbt2) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t* in[3] = {nba0, nba1, nba2};
    int32_t* out[3] = {nbt0, nbt1, nbt2};
    for (int ax = 0; ax < 3; ++ax) {
        const int j = in[ax][k];
        if (j >= 0 && depth[j] == depth[k]) out[ax][j] = (int32_t)k;
    }
}

// ---------------------------------------------------------------- host side
namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Address of state field f of storage S: with mixed storage the dual fields
// (p, q) hold BF16 and their address carries the mark Io<mixs> tests.
template <class S>
S* field_ptr(DevBuf& b, int f) {
    S* ptr = b.as<S>();
    if (std::is_same<S, mixs>::value && f >= F_PX && f <= F_Q5)
        ptr = reinterpret_cast<S*>(reinterpret_cast<uintptr_t>(ptr) | 1u);
    return ptr;
}

template <class S>
KArgs<S> make_args(tgvx_ctx* c, const tgvx_params* prm) {
    using T = typename Io<S>::T;
    Problem& P = c->prob;
    KArgs<S> a;
    std::memset(&a, 0, sizeof(a));
    for (int i = 0; i < 17; ++i) a.s[i] = field_ptr<S>(P.d_state[i], i);
    const bool adj = prm && prm->op_mode == TGVX_OP_ADJ;
    for (int ax = 0; ax < 3; ++ax) {
        a.nba[ax] = P.d_nba[ax].as<int32_t>();
        a.nbt[ax] = P.d_nbt[ax].as<int32_t>();
        a.nbm[ax] = P.d_nb[2 * ax + 1].as<int32_t>();
        a.nbp[ax] = adj ? a.nba[ax] : P.d_nb[2 * ax].as<int32_t>();
    }
    for (int ax = 0; ax < 3; ++ax) a.nbt[ax] = P.d_nbt[ax].as<int32_t>();
    a.info = P.d_info.as<uint32_t>();
    a.nbins = P.nb;
    a.parent = adj ? P.d_parent_a.as<int32_t>() : P.d_parent.as<int32_t>();
    for (int i = 0; i < 4; ++i) a.save[i] = P.d_save[i].as<S>();
    a.init = P.d_init.as<float>();
    a.init_stride = 4;
    if (prm) {
        a.sigma = prm->sigma > 0.f ? (T)prm->sigma : T(0.25);
        a.tau = prm->tau > 0.f ? (T)prm->tau : T(0.25);
        a.a0 = (T)prm->alpha0;
        a.a1 = (T)prm->alpha1;
        a.tl = a.tau * (T)prm->lambda;
    }
    return a;
}

// The level of octree depth `depth` iterates with the per-depth steps where
// the parameters give them (sigma_depth, tau_depth > 0), else sigma, tau.
template <class S>
void set_steps(KArgs<S>& a, const tgvx_params* p, int depth) {
    using T = typename Io<S>::T;
    const bool in = depth >= 0 && depth < TGVX_MAX_LEVELS;
    const float sg = in && p->sigma_depth[depth] > 0.f ? p->sigma_depth[depth] : p->sigma;
    const float ta = in && p->tau_depth[depth] > 0.f ? p->tau_depth[depth] : p->tau;
    a.sigma = sg > 0.f ? (T)sg : T(0.25);
    a.tau = ta > 0.f ? (T)ta : T(0.25);
    a.tl = a.tau * (T)p->lambda;
}

// Before level li iterates in the exact-adjoint mode: the previous level's
// bars (the only frozen values it reads) as 4-value records.
// The records of the reference rules' out-of-level neighbours (u_bar, v_bar,
// then p, q; 4 values per level cube and axis each), sized for the largest
// level at the first use, so that the buffer never moves under a captured
// graph of an earlier level.
template <class S>
void rfz_bufs(tgvx_ctx* c, S** bar, S** dual) {
    Problem& P = c->prob;
    uint32_t max_n = 1;
    for (const Level& l : P.levels) max_n = std::max(max_n, l.n);
    P.d_rfz.ensure(2 * 4 * 3 * (size_t)max_n * P.elem);
    const size_t rec = 4 * 3 * (size_t)max_n;
    *bar = P.d_rfz.as<S>();
    *dual = reinterpret_cast<S*>(static_cast<char*>(P.d_rfz.p) + rec * P.elem);
}

template <class S>
// launch_fz = false: the exact-adjoint records of the previous level are
// left to the caller's parent initialisation (same launch).
void prepare_level(tgvx_ctx* c, KArgs<S>& a, const tgvx_params* p, size_t li, bool launch_fz = true) {
    Problem& P = c->prob;
    a.fzaos = nullptr;
    a.fz_off = a.fz_n = 0;
    a.rfz_bar = a.rfz_dual = nullptr;
    if (p->op_mode == TGVX_OP_REF && !(p->fused == 0 && !p->tune)) {
        // the reference rules' out-of-level neighbours (two-iteration kernel)
        const Level& L = P.levels[li];
        S *bar, *dual;
        rfz_bufs<S>(c, &bar, &dual);
        KArgs<S> g = a;
        g.lvl_off = L.off;
        g.lvl_n = L.n;
        if (L.n == 0) return;
        // Records pay where many references leave the level: measured on the
        // depth-10 torus + box scene, levels with 40-60 % of their six
        // references outside (depths 4-9) run the two-iteration kernel 18-47 %
        // faster with them, the finest level (18 %) 8 % slower, its sparse
        // record reads fetching whole lines. Decided once per level and
        // problem from a count (one synchronisation).
        if (P.rfz_use.size() != P.levels.size()) P.rfz_use.assign(P.levels.size(), -1);
        if (P.rfz_use[li] < 0) {
            P.d_rfz_cnt.ensure(sizeof(uint32_t));
            uint32_t* cnt = P.d_rfz_cnt.as<uint32_t>();
            k_fill_i32<<<1, 64, 0, c->stream>>>(reinterpret_cast<int32_t*>(cnt), 1u, 0);
            k_ref_frozen<S><<<(L.n + 255) / 256, 256, 0, c->stream>>>(g, bar, dual, cnt);
            uint32_t h = 0;
            TGVX_HIP(hipMemcpyAsync(&h, cnt, sizeof(h), hipMemcpyDeviceToHost, c->stream));
            TGVX_HIP(hipStreamSynchronize(c->stream));
            P.rfz_use[li] = (uint64_t)h * 10u >= (uint64_t)L.n * 6u * 3u ? 1 : 0;  // at least 30 %
        }
        if (!P.rfz_use[li]) return;  // out-of-level reads from the main set
        k_ref_frozen<S><<<(L.n + 255) / 256, 256, 0, c->stream>>>(g, bar, dual, nullptr);
        a.rfz_bar = bar;
        a.rfz_dual = dual;
        return;
    }
    if (p->op_mode != TGVX_OP_ADJ || li == 0) return;
    const Level& prev = P.levels[li - 1];
    a.fz_off = prev.off;
    a.fz_n = prev.n;
    if (launch_fz) k_fz_aos<S><<<(prev.n + 255) / 256, 256, 0, c->stream>>>(a, P.d_fzaos.as<S>());
    a.fzaos = P.d_fzaos.as<S>();
}

int block_of(const tgvx_params* p) {
    int b = p && p->block > 0 ? p->block : 256;
    if (b != 64 && b != 128 && b != 256 && b != 512 && b != 1024)
        throw Error(TGVX_E_ARG, "block must be 64, 128, 256, 512 or 1024");
    const bool tiled = p && ((p->op_mode == TGVX_OP_ADJ && (p->fused == 1 || p->fused == 3 || p->fused == 4)) ||
                             p->fused == 5 || p->fused == 6 || p->fused == 7 || p->fused == 8);
    if (b > 256 && !tiled) throw Error(TGVX_E_ARG, "blocks above 256 exist only for the tiled fused kernels");
    return b;
}

// How one level's iterations are launched: variant 0 = dual + primal
// kernels, 1 = LDS-tiled fused kernel, 2 = untiled fused kernel, 3 = tiled
// with the tile's p, q staged in LDS, 4 = tiled with a precomputed halo,
// 5 = two iterations per launch over tabulated tile sets, 6 = a level of at
// most kLevel1 cubes in one workgroup for all its iterations (larger levels
// run variant 5 with tiles of 512), 7 = a level in one cooperative launch
// with grid barriers (workgroups of 256 or 1024; a level whose workgroups
// cannot all be co-resident runs as variant 6 does), 8 = four iterations per
// launch over tabulated tile sets where four remain and the sets fit (tiles
// of 128 or 256), otherwise as variant 5;
// block = the workgroup size (the tile size for variants 1, 3, 4, 5, 8); xcd =
// XCD-aware tile order.
struct Launch {
    int variant, block, xcd;
};

int size_bucket(uint32_t n) {
    int b = 0;
    while (n > 1u) {
        n >>= 1;
        ++b;
    }
    return b;
}

Launch launch_for(const tgvx_ctx* c, const tgvx_params* p, uint32_t n, bool frozen) {
    const bool adj = p->op_mode == TGVX_OP_ADJ;
    if (p->fused < 0 || p->fused > 8) throw Error(TGVX_E_ARG, "fused must be 0 to 8");
    // the reference's rules have the two-kernel iteration, the two- and
    // four-iteration kernels (5, 8), the whole-level kernel for small levels
    // (6; larger levels as 5) and run 7 as 6 does beyond its small levels,
    // also for frozen-border leaves (with 5); the exact-adjoint frozen-border
    // solve has the two kernels only
    const int var = !adj ? (p->fused >= 5 ? p->fused : 0) : frozen ? 0 : p->fused;
    Launch L{var, block_of(p), 1};
    const bool tiled = L.variant == 1 || L.variant == 3 || L.variant == 4 || L.variant == 5 || L.variant == 8;
    if (tiled && p->block <= 0) L.block = kFuseTile;
    if (L.variant == 6) L = Launch{6, kLevel1, 0};
    if (L.variant == 7) L = Launch{7, L.block == 256 ? 256 : 1024, 0};
    if (tiled && L.block == 64)
        throw Error(TGVX_E_ARG, "the tiled fused kernel takes block 128 or more");
    if (p->tune && !frozen) {
        auto it = c->tuned.find(std::make_tuple(c->prob.precision, p->op_mode, size_bucket(n)));
        if (it != c->tuned.end()) L = Launch{it->second.variant, it->second.block, it->second.xcd};
    }
    return L;
}

template <class S>
Set13<S> main_set(const KArgs<S>& a) {
    Set13<S> m;
    for (int f = 0; f < 13; ++f) m.f[f] = a.s[f];
    return m;
}

template <class S>
Set13<S> alt_set(tgvx_ctx* c) {
    Problem& P = c->prob;
    Set13<S> m;
    for (int f = 0; f < 13; ++f) {
        P.d_alt[f].ensure(P.elem * std::max<uint64_t>(P.n, 1));
        m.f[f] = field_ptr<S>(P.d_alt[f], f);
    }
    return m;
}

template <class S, bool X>
void launch_iter_x(tgvx_ctx* c, KArgs<S>& a, int op_mode, int block, bool frozen) {
    if (block > 256) block = 256;  // the dual / primal kernels take at most 256
    const unsigned grid = (a.n + block - 1) / block;
    if (frozen) k_dual<S, 1, X><<<grid, block, 0, c->stream>>>(a);
    else k_dual<S, 0, X><<<grid, block, 0, c->stream>>>(a);
    if (op_mode == TGVX_OP_ADJ) {
        const Set13<S> m = main_set(a);
        k_primal_adj<S, X><<<grid, block, 0, c->stream>>>(a, m, m, 0);
    } else {
        k_primal_ref<S, X><<<grid, block, 0, c->stream>>>(a);
    }
}
template <class S>
void launch_iter(tgvx_ctx* c, KArgs<S>& a, int op_mode, const Launch& L, bool frozen) {
    if (L.xcd) launch_iter_x<S, true>(c, a, op_mode, L.block, frozen);
    else launch_iter_x<S, false>(c, a, op_mode, L.block, frozen);
}

// The halo table of the fused tiled kernel for a's launch range and a tile
// size, built on first use. Callers build it before any graph capture starts.
template <int TILE>
void halo_table_launch(tgvx_ctx* c, const int32_t* const nb[3], uint32_t off, uint32_t n, uint32_t lvl_off,
                       uint32_t lvl_n, HaloBufs& hb) {
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    if (ntiles == 0) return;
    k_halo_table<TILE><<<ntiles, TILE, 0, c->stream>>>(nb[0], nb[1], nb[2], off, n, lvl_off, lvl_n,
                                                        hb.slot[0].as<int32_t>(), hb.slot[1].as<int32_t>(),
                                                        hb.slot[2].as<int32_t>(), hb.cnt.as<int32_t>(),
                                                        hb.list.as<int32_t>());
}
template <class S>
HaloTab halo_tab(tgvx_ctx* c, const KArgs<S>& a, int tile) {
    Problem& P = c->prob;
    const auto key = std::make_tuple(a.off, a.n, tile);
    auto it = P.halo.find(key);
    if (it == P.halo.end()) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        TGVX_HIP(hipStreamIsCapturing(c->stream, &cs));
        if (cs != hipStreamCaptureStatusNone) throw Error(TGVX_E_STATE, "halo table built inside a graph capture");
        auto hb = std::make_unique<HaloBufs>();
        const uint32_t ntiles = (a.n + tile - 1) / tile;
        for (int ax = 0; ax < 3; ++ax) hb->slot[ax].ensure(sizeof(int32_t) * std::max<uint32_t>(a.n, 1));
        hb->cnt.ensure(sizeof(int32_t) * std::max<uint32_t>(ntiles, 1));
        hb->list.ensure(sizeof(int32_t) * 3 * (size_t)tile * std::max<uint32_t>(ntiles, 1));
        switch (tile) {
            case 128: halo_table_launch<128>(c, a.nbp, a.off, a.n, a.lvl_off, a.lvl_n, *hb); break;
            case 256: halo_table_launch<256>(c, a.nbp, a.off, a.n, a.lvl_off, a.lvl_n, *hb); break;
            case 512: halo_table_launch<512>(c, a.nbp, a.off, a.n, a.lvl_off, a.lvl_n, *hb); break;
            case 1024: halo_table_launch<1024>(c, a.nbp, a.off, a.n, a.lvl_off, a.lvl_n, *hb); break;
            default: throw Error(TGVX_E_ARG, "the tiled fused kernel takes tiles of 128, 256, 512 or 1024");
        }
        TGVX_HIP(hipGetLastError());
        it = P.halo.emplace(key, std::move(hb)).first;
    }
    const HaloBufs& hb = *it->second;
    HaloTab h;
    for (int ax = 0; ax < 3; ++ax) h.slot[ax] = static_cast<const int32_t*>(hb.slot[ax].p);
    h.cnt = static_cast<const int32_t*>(hb.cnt.p);
    h.list = static_cast<const int32_t*>(hb.list.p);
    return h;
}

// LDS per workgroup the two-iteration kernel may use: the device's limit
// (160 KiB on gfx950), at most kHalo2Lds.
inline size_t lds_budget(const tgvx_ctx* c) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess || v <= 0)
        v = 64 * 1024;
    return std::min<size_t>((size_t)v, (size_t)kHalo2Lds);
}

// LDS of one CU (160 KiB on gfx950).
inline size_t lds_per_cu(const tgvx_ctx* c) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) != hipSuccess ||
        v <= 0)
        v = 64 * 1024;
    return (size_t)v;
}
// Workgroups of k_fused2 (tile threads) that its VGPRs let a CU hold: four
// SIMDs of 512 VGPRs per lane, allocated in granules of 8, at most 8 waves each.
template <class S, bool REF>
size_t fused2_wg_per_cu(int tile) {
    hipFuncAttributes fa;
    hipError_t e;
    switch (tile) {
        case 128: e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_fused2<S, 128, true, REF>)); break;
        case 256: e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_fused2<S, 256, true, REF>)); break;
        case 1024: e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_fused2<S, 1024, true, REF>)); break;
        default: e = hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(&k_fused2<S, 512, true, REF>)); break;
    }
    if (e != hipSuccess || fa.numRegs <= 0) return 1;
    const int waves = std::min(8, 512 / ((fa.numRegs + 7) / 8 * 8));
    return (size_t)std::max(1, 4 * waves / std::max(1, tile / 64));
}

// Row r of dst = row order[r] of src, rows of len int32 (one workgroup a row).
__global__ void k_permute_rows(const int32_t* __restrict__ src, int32_t* __restrict__ dst,
                               const int32_t* __restrict__ order, uint32_t len) {
    const int32_t* const s = src + (size_t)order[blockIdx.x] * len;
    int32_t* const d = dst + (size_t)blockIdx.x * len;
    for (uint32_t i = threadIdx.x; i < len; i += blockDim.x) d[i] = s[i];
}

// The two-iteration kernel's tile sets for a's level and a tile size, built
// on first use (callers build them before any graph capture starts).
template <int TILE, bool REF>
void halo2_table_launch(tgvx_ctx* c, const int32_t* const nba[3], const int32_t* const nbt[3], uint32_t off,
                        uint32_t n, Halo2Bufs& hb) {
    const uint32_t ntiles = (n + TILE - 1) / TILE;
    if (ntiles == 0) return;
    k_halo2_table<TILE, REF><<<ntiles, TILE, 0, c->stream>>>(nba[0], nba[1], nba[2], nbt[0], nbt[1], nbt[2], off, n,
                                                              hb.cnt.as<int32_t>(), hb.list.as<int32_t>(),
                                                              hb.bslot.as<int32_t>(), hb.cslot.as<int32_t>());
}
// REF: the tile sets of the reference's rules (its own +a / -a table).
template <class S, bool REF = false>
const Halo2Bufs& halo2_tab(tgvx_ctx* c, const KArgs<S>& a, int tile) {
    Problem& P = c->prob;
    const auto key = std::make_tuple(a.off, a.n, tile, (int)sizeof(S) + (REF ? 1000 : 0));
    auto it = P.halo2.find(key);
    if (it != P.halo2.end()) return *it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    TGVX_HIP(hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) throw Error(TGVX_E_STATE, "tile sets built inside a graph capture");
    auto hb = std::make_unique<Halo2Bufs>();
    const uint32_t ntiles = (a.n + tile - 1) / tile;
    const size_t ca = (size_t)kHalo2Cap * tile;
    // the kernel needs the whole level as its launch range (the sets are
    // closed under the level's neighbour relation)
    const bool whole = a.off == a.lvl_off && a.n == a.lvl_n;
    if (whole && ntiles > 0) {
        hb->cnt.ensure(sizeof(int32_t) * 4 * ntiles);
        hb->list.ensure(sizeof(int32_t) * (ca - tile) * ntiles);
        hb->bslot.ensure(sizeof(int32_t) * 3 * ca * ntiles);
        hb->cslot.ensure(sizeof(int32_t) * 3 * ca * ntiles);
        const int32_t* const* nb_plus = REF ? a.nbp : a.nba;
        const int32_t* const* nb_in = REF ? a.nbm : a.nbt;
        switch (tile) {
            case 128: halo2_table_launch<128, REF>(c, nb_plus, nb_in, a.off, a.n, *hb); break;
            case 256: halo2_table_launch<256, REF>(c, nb_plus, nb_in, a.off, a.n, *hb); break;
            case 512: halo2_table_launch<512, REF>(c, nb_plus, nb_in, a.off, a.n, *hb); break;
            case 1024: halo2_table_launch<1024, REF>(c, nb_plus, nb_in, a.off, a.n, *hb); break;
            default: throw Error(TGVX_E_ARG, "the two-iteration kernel takes tiles of 128, 256, 512 or 1024");
        }
        TGVX_HIP(hipGetLastError());
        std::vector<int32_t> cnt(4 * (size_t)ntiles);
        TGVX_HIP(hipMemcpyAsync(cnt.data(), hb->cnt.p, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost,
                                c->stream));
        TGVX_HIP(hipStreamSynchronize(c->stream));
        bool failed = false;
        const size_t el = sizeof(typename Io<S>::T);
        // the kernel's static LDS (its median table) comes on top of the
        // dynamic part a class is launched with
        const size_t st = 32 * el;
        auto need = [&](uint32_t t) {
            return st + el * (13 * (size_t)cnt[4 * t + 2] + 4 * (size_t)std::max(0, cnt[4 * t] - 2 * tile));
        };
        std::vector<size_t> needs(ntiles);
        for (uint32_t t = 0; t < ntiles; ++t) {
            hb->sp = std::max(hb->sp, cnt[4 * t]);
            hb->sd = std::max(hb->sd, cnt[4 * t + 1]);
            hb->sa = std::max(hb->sa, cnt[4 * t + 2]);
            failed = failed || cnt[4 * t + 3] != 0;
            needs[t] = need(t);
        }
        // the main launch takes every tile that keeps the occupancy (workgroups
        // per CU, from the kernel's VGPRs and the CU's LDS) of the tile at the
        // 90th percentile of LDS need; the rest run in a second launch
        std::vector<size_t> sorted = needs;
        std::nth_element(sorted.begin(), sorted.begin() + (ntiles * 9) / 10, sorted.end());
        const size_t p90 = std::max<size_t>(sorted[(ntiles * 9) / 10], 1);
        const size_t cu_lds = lds_per_cu(c);
        const size_t wg = std::max<size_t>(1, std::min(fused2_wg_per_cu<S, REF>(tile), cu_lds / p90));
        const size_t thr = std::max(p90, cu_lds / wg);
        std::vector<int32_t> order;
        order.reserve(ntiles);
        for (int k = 0; k < 2; ++k) {
            Halo2Bufs::Class& cl = hb->cls[k];
            cl.first = (uint32_t)order.size();
            for (uint32_t t = 0; t < ntiles; ++t) {
                if ((needs[t] <= thr) != (k == 0)) continue;
                order.push_back((int32_t)t);
                cl.sa = std::max(cl.sa, cnt[4 * t + 2]);
                cl.sd = std::max(cl.sd, cnt[4 * t + 1]);
                cl.sx = std::max(cl.sx, cnt[4 * t] - 2 * tile);
            }
            cl.count = (uint32_t)order.size() - cl.first;
            cl.lds = el * (13 * (size_t)cl.sa + 4 * (size_t)cl.sx);
            hb->lds = std::max(hb->lds, st + cl.lds);
        }
        hb->ok = !failed && hb->lds <= lds_budget(c);
        if (hb->ok) {
            // the tables in launch order, each count record ending in its
            // tile id (the kernel's table reads then depend on its workgroup
            // id alone)
            std::vector<int32_t> rec(4 * order.size());
            for (size_t k = 0; k < order.size(); ++k) {
                const int32_t t = order[k];
                rec[4 * k] = cnt[4 * (size_t)t];
                rec[4 * k + 1] = cnt[4 * (size_t)t + 1];
                rec[4 * k + 2] = cnt[4 * (size_t)t + 2];
                rec[4 * k + 3] = t;
            }
            DevBuf ord;
            ord.ensure(sizeof(int32_t) * order.size());
            TGVX_HIP(hipMemcpyAsync(ord.p, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice,
                                    c->stream));
            TGVX_HIP(hipMemcpyAsync(hb->cnt.p, rec.data(), sizeof(int32_t) * rec.size(), hipMemcpyHostToDevice,
                                    c->stream));
            auto permute = [&](DevBuf& b, size_t len) {
                DevBuf o;
                o.ensure(sizeof(int32_t) * len * ntiles);
                k_permute_rows<<<ntiles, 256, 0, c->stream>>>(b.as<int32_t>(), o.as<int32_t>(), ord.as<int32_t>(),
                                                              (uint32_t)len);
                TGVX_HIP(hipGetLastError());
                b.swap(o);
                TGVX_HIP(hipStreamSynchronize(c->stream));  // before o (the old rows) is freed
            };
            permute(hb->list, ca - tile);
            permute(hb->bslot, 3 * ca);
            permute(hb->cslot, 3 * ca);
        } else {  // the level falls back to one iteration per launch
            hb->cnt.release();
            hb->list.release();
            hb->bslot.release();
            hb->cslot.release();
        }
    }
    return *P.halo2.emplace(key, std::move(hb)).first->second;
}

// two iterations (primal n+1 ... dual n+3) from src into dst, or with HEAD
// the level's first one and a half (dual 1, primal 1, dual 2); false when
// the level's tile sets do not fit (nothing launched)
template <class S, bool X, bool REF = false, bool HEAD = false>
bool launch_fused2_x(tgvx_ctx* c, KArgs<S>& a, const Set13<S>& src, const Set13<S>& dst, int tile) {
    const Halo2Bufs& hb = halo2_tab<S, REF>(c, a, tile);
    if (!hb.ok) return false;
    Halo2Tab h;
    const size_t ca = (size_t)kHalo2Cap * tile;
    for (const Halo2Bufs::Class& cl : hb.cls) {
        if (cl.count == 0) continue;
        // the class's rows (the tables are in launch order)
        h.cnt = hb.cnt.as<int32_t>() + 4 * (size_t)cl.first;
        h.list = hb.list.as<int32_t>() + (ca - tile) * cl.first;
        h.bslot = hb.bslot.as<int32_t>() + 3 * ca * cl.first;
        h.cslot = hb.cslot.as<int32_t>() + 3 * ca * cl.first;
        switch (tile) {
            case 128:
                k_fused2<S, 128, X, REF, HEAD><<<cl.count, 128, cl.lds, c->stream>>>(a, src, dst, h, cl.sa, cl.sx);
                break;
            case 256:
                k_fused2<S, 256, X, REF, HEAD><<<cl.count, 256, cl.lds, c->stream>>>(a, src, dst, h, cl.sa, cl.sx);
                break;
            case 1024:
                k_fused2<S, 1024, X, REF, HEAD><<<cl.count, 1024, cl.lds, c->stream>>>(a, src, dst, h, cl.sa, cl.sx);
                break;
            default:
                k_fused2<S, 512, X, REF, HEAD><<<cl.count, 512, cl.lds, c->stream>>>(a, src, dst, h, cl.sa, cl.sx);
                break;
        }
    }
    return true;
}

// The four-iteration kernel's tile sets for a's level (the whole level) and a
// tile of 128 or 256, built on first use outside any graph capture.
// ok = no tile's sets exceeded the cap and the largest tile's LDS fits.
template <class S, bool REF = false>
const Halo4Bufs& halo4_tab(tgvx_ctx* c, const KArgs<S>& a, int tile) {
    Problem& P = c->prob;
    const auto key = std::make_tuple(a.off, a.n, tile, (int)sizeof(S) + (REF ? 1000 : 0));
    auto it = P.halo4.find(key);
    if (it != P.halo4.end()) return *it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    TGVX_HIP(hipStreamIsCapturing(c->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) throw Error(TGVX_E_STATE, "tile sets built inside a graph capture");
    auto hb = std::make_unique<Halo4Bufs>();
    const uint32_t ntiles = (a.n + tile - 1) / tile;
    const size_t ca = (size_t)kHalo4Cap * tile;
    const bool whole = a.off == a.lvl_off && a.n == a.lvl_n;
    if (whole && ntiles > 0 && (tile == 128 || tile == 256)) {
        hb->cnt.ensure(sizeof(int32_t) * kHalo4Cnt * ntiles);
        hb->list.ensure(sizeof(int32_t) * (ca - tile) * ntiles);
        hb->bslot.ensure(sizeof(int32_t) * 3 * ca * ntiles);
        hb->cslot.ensure(sizeof(int32_t) * 3 * ca * ntiles);
        const int32_t* const* nb_plus = REF ? a.nbp : a.nba;
        const int32_t* const* nb_in = REF ? a.nbm : a.nbt;
        int32_t* const o[4] = {hb->cnt.as<int32_t>(), hb->list.as<int32_t>(), hb->bslot.as<int32_t>(),
                               hb->cslot.as<int32_t>()};
        switch (tile) {
            case 128:
                k_halo4_table<128, REF><<<ntiles, 128, 0, c->stream>>>(nb_plus[0], nb_plus[1], nb_plus[2], nb_in[0],
                                                                        nb_in[1], nb_in[2], a.off, a.n, o[0], o[1],
                                                                        o[2], o[3]);
                break;
            default:
                k_halo4_table<256, REF><<<ntiles, 256, 0, c->stream>>>(nb_plus[0], nb_plus[1], nb_plus[2], nb_in[0],
                                                                        nb_in[1], nb_in[2], a.off, a.n, o[0], o[1],
                                                                        o[2], o[3]);
                break;
        }
        TGVX_HIP(hipGetLastError());
        std::vector<int32_t> cnt((size_t)kHalo4Cnt * ntiles);
        TGVX_HIP(hipMemcpyAsync(cnt.data(), hb->cnt.p, sizeof(int32_t) * cnt.size(), hipMemcpyDeviceToHost,
                                c->stream));
        TGVX_HIP(hipStreamSynchronize(c->stream));
        bool failed = false;
        for (uint32_t t = 0; t < ntiles; ++t) {
            hb->sa = std::max(hb->sa, cnt[(size_t)kHalo4Cnt * t + kHalo4Sets]);
            failed = failed || cnt[(size_t)kHalo4Cnt * t + kHalo4Sets + 1] != 0;
        }
        const size_t el = sizeof(typename Io<S>::T);
        hb->lds = el * (32 + 17 * (size_t)hb->sa);  // the median table and 17 arrays of sa
        hb->ok = !failed && hb->lds <= lds_budget(c);
        if (!hb->ok) {
            hb->cnt.release();
            hb->list.release();
            hb->bslot.release();
            hb->cslot.release();
        }
    }
    return *P.halo4.emplace(key, std::move(hb)).first->second;
}

// four iterations (primal n+1 ... dual n+5) from src into dst; false when
// the level's four-iteration sets do not fit (nothing launched)
template <class S, bool REF = false>
bool launch_fused4(tgvx_ctx* c, KArgs<S>& a, const Set13<S>& src, const Set13<S>& dst, int tile) {
    const Halo4Bufs& hb = halo4_tab<S, REF>(c, a, tile);
    if (!hb.ok) return false;
    Halo4Tab h;
    h.cnt = hb.cnt.as<int32_t>();
    h.list = hb.list.as<int32_t>();
    h.bslot = hb.bslot.as<int32_t>();
    h.cslot = hb.cslot.as<int32_t>();
    const uint32_t ntiles = (a.n + tile - 1) / tile;
    const size_t dyn = hb.lds - 32 * sizeof(typename Io<S>::T);
    switch (tile) {
        case 128: k_fused4<S, 128, REF><<<ntiles, kFused4Threads, dyn, c->stream>>>(a, src, dst, h, hb.sa); break;
        default: k_fused4<S, 256, REF><<<ntiles, kFused4Threads, dyn, c->stream>>>(a, src, dst, h, hb.sa); break;
    }
    return true;
}

// one fused pass (primal n, dual n+1) from src into dst
template <class S, bool X>
void launch_fused_pass_x(tgvx_ctx* c, KArgs<S>& a, const Set13<S>& src, const Set13<S>& dst, const Launch& L0) {
    // variants 5 and 8 advance one iteration with the tiled kernel of the same tile
    const Launch L = L0.variant == 5 || L0.variant == 8 ? Launch{1, L0.block, L0.xcd} : L0;
    if (L.variant == 4) {
        const HaloTab h = halo_tab<S>(c, a, L.block);
        switch (L.block) {
            case 128: k_fused_halo<S, 128, X><<<(a.n + 127) / 128, 128, 0, c->stream>>>(a, src, dst, h); break;
            case 256: k_fused_halo<S, 256, X><<<(a.n + 255) / 256, 256, 0, c->stream>>>(a, src, dst, h); break;
            case 1024: k_fused_halo<S, 1024, X><<<(a.n + 1023) / 1024, 1024, 0, c->stream>>>(a, src, dst, h); break;
            default: k_fused_halo<S, 512, X><<<(a.n + 511) / 512, 512, 0, c->stream>>>(a, src, dst, h); break;
        }
    } else if (L.variant == 1 || L.variant == 3) {
        const bool stage = L.variant == 3;
        switch (L.block) {
            case 128:
                if (stage) k_fused_tile<S, 128, X, true><<<(a.n + 127) / 128, 128, 0, c->stream>>>(a, src, dst);
                else k_fused_tile<S, 128, X, false><<<(a.n + 127) / 128, 128, 0, c->stream>>>(a, src, dst);
                break;
            case 256:
                if (stage) k_fused_tile<S, 256, X, true><<<(a.n + 255) / 256, 256, 0, c->stream>>>(a, src, dst);
                else k_fused_tile<S, 256, X, false><<<(a.n + 255) / 256, 256, 0, c->stream>>>(a, src, dst);
                break;
            case 1024:
                if (stage) k_fused_tile<S, 1024, X, true><<<(a.n + 1023) / 1024, 1024, 0, c->stream>>>(a, src, dst);
                else k_fused_tile<S, 1024, X, false><<<(a.n + 1023) / 1024, 1024, 0, c->stream>>>(a, src, dst);
                break;
            default:
                if (stage) k_fused_tile<S, 512, X, true><<<(a.n + 511) / 512, 512, 0, c->stream>>>(a, src, dst);
                else k_fused_tile<S, 512, X, false><<<(a.n + 511) / 512, 512, 0, c->stream>>>(a, src, dst);
                break;
        }
    } else {
        k_fused_adj<S, X><<<(a.n + L.block - 1) / L.block, L.block, 0, c->stream>>>(a, src, dst);
    }
}
template <class S>
void launch_fused_pass(tgvx_ctx* c, KArgs<S>& a, const Set13<S>& src, const Set13<S>& dst, const Launch& L) {
    if (L.xcd) launch_fused_pass_x<S, true>(c, a, src, dst, L);
    else launch_fused_pass_x<S, false>(c, a, src, dst, L);
}

// u, v from the set uv, p, q from the set pq
template <class S>
Set13<S> mix_sets(const Set13<S>& uv, const Set13<S>& pq) {
    Set13<S> m = pq;
    for (int f = F_U; f <= F_VZ; ++f) m.f[f] = uv.f[f];
    return m;
}

// `iters` >= 2 iterations as: dual of the first, iters - 1 fused (primal n,
// dual n+1) passes alternating between the main and the alternate set,
// primal of the last back into the main set. A level that starts from its
// initial state (cold: u_bar = u, v_bar = v, p = q = 0) with the
// two-iteration kernel opens with its head launch instead (dual 1, primal 1,
// dual 2), which writes u, v to the alternate set and p, q to the main set;
// the sets then alternate field group by field group, so after an even
// number of two-iteration launches p, q are back in the main set and the
// last primal step leaves them in place.
template <class S, bool X>
void launch_fused_x(tgvx_ctx* c, KArgs<S>& a, int iters, const Launch& L, int op_mode, bool frozen, bool cold,
                    bool need_head) {
    const int wg = L.block > 256 ? 256 : L.block;  // the dual / primal kernels take at most 256
    const unsigned grid = (a.n + wg - 1) / wg;
    Set13<S> cur = main_set(a), oth = alt_set<S>(c);
    const Set13<S> mainset = cur, altset = oth;
    // frozen-border leaves (reference rules only): out-of-leaf u_bar, v_bar
    // from the snapshot in the dual kernel, from the records in k_fused2
    auto dual = [&] {
        if (frozen) k_dual<S, 1, X><<<grid, wg, 0, c->stream>>>(a);
        else k_dual<S, 0, X><<<grid, wg, 0, c->stream>>>(a);
    };
    const bool ref = op_mode == TGVX_OP_REF;
    int i = 1;
    bool head = false;
    const bool two = L.variant == 5 || L.variant == 8;
    // variant 8: four iterations per launch while four remain (whole levels)
    const bool four = L.variant == 8 && !frozen;
    if (cold && two) {
        const Set13<S> hd = mix_sets(altset, mainset);
        head = ref ? launch_fused2_x<S, X, true, true>(c, a, mainset, hd, L.block)
                   : launch_fused2_x<S, X, false, true>(c, a, mainset, hd, L.block);
        if (head) {
            cur = hd;
            oth = mix_sets(mainset, altset);
            i = 2;
        }
    }
    if (need_head && !head) throw Error(TGVX_E_STATE, "a level initialised for the head launch runs without it");
    if (!head) dual();
    // the last primal step copies p, q into the main set unless they are there
    auto copy = [&] { return cur.f[F_PX] != mainset.f[F_PX] ? 1 : 0; };
    if (ref) {
        // the reference's rules: four (variant 8) or two iterations per
        // launch while as many remain, then primal and dual kernels on the
        // main set
        for (; four && i + 3 < iters; i += 4) {
            if (!launch_fused4<S, true>(c, a, cur, oth, L.block)) break;
            std::swap(cur, oth);
        }
        for (; i + 1 < iters; i += 2) {
            if (!launch_fused2_x<S, X, true>(c, a, cur, oth, L.block)) break;
            std::swap(cur, oth);
        }
        for (; i < iters; ++i) {
            k_primal_ref_set<S, X><<<grid, wg, 0, c->stream>>>(a, cur, mainset, copy());
            cur = mainset;
            oth = altset;
            dual();
        }
        k_primal_ref_set<S, X><<<grid, wg, 0, c->stream>>>(a, cur, mainset, copy());
        return;
    }
    for (; four && i + 3 < iters; i += 4) {
        if (!launch_fused4<S>(c, a, cur, oth, L.block)) break;
        std::swap(cur, oth);
    }
    if (two) {
        for (; i + 1 < iters; i += 2) {
            if (!launch_fused2_x<S, X>(c, a, cur, oth, L.block)) break;
            std::swap(cur, oth);
        }
    }
    for (; i < iters; ++i) {
        launch_fused_pass_x<S, X>(c, a, cur, oth, L);
        std::swap(cur, oth);
    }
    k_primal_adj<S, X><<<grid, wg, 0, c->stream>>>(a, cur, mainset, copy());
}
template <class S>
void launch_fused(tgvx_ctx* c, KArgs<S>& a, int iters, const Launch& L, int op_mode, bool frozen = false,
                  bool cold = false, bool need_head = false) {
    if (L.xcd) launch_fused_x<S, true>(c, a, iters, L, op_mode, frozen, cold, need_head);
    else launch_fused_x<S, false>(c, a, iters, L, op_mode, frozen, cold, need_head);
}

// Cached graph of `iters` iterations of one level.
// Everything a captured graph bakes in: the launch shape and the kernel
// arguments' scalars (steps, weights), so a solve with other parameters on
// the same problem never replays a stale graph.
// A frozen-border leaf's graph also bakes in the leaf's input index range
// (its in-leaf test); the snapshot it reads lives as long as the problem.
struct GraphKey {
    uint32_t off, n, lvl_off, lvl_n;
    int iters, op, block, prec, variant, xcd, mask;
    double sigma, tau, a0, a1, tl;
    int frozen;
    uint32_t leaf_first, leaf_last;
    int cold;
    bool operator<(const GraphKey& o) const {
        return std::tie(off, n, lvl_off, lvl_n, iters, op, block, prec, variant, xcd, mask, sigma, tau, a0, a1, tl,
                        frozen, leaf_first, leaf_last, cold) <
               std::tie(o.off, o.n, o.lvl_off, o.lvl_n, o.iters, o.op, o.block, o.prec, o.variant, o.xcd, o.mask,
                        o.sigma, o.tau, o.a0, o.a1, o.tl, o.frozen, o.leaf_first, o.leaf_last, o.cold);
    }
};
struct GraphCache {
    std::map<GraphKey, hipGraphExec_t> m;
    // the whole fixed-iteration sweep as one graph, for the parameters (and
    // per-level iteration counts) in sweep_key; seen_key: the parameters of
    // the last sweep run without one (a sweep is captured on its second run,
    // when every table and buffer its launches read exists)
    hipGraphExec_t sweep = nullptr;
    std::vector<char> sweep_key, seen_key;
    ~GraphCache() { clear(); }
    void clear() {
        for (auto& kv : m) (void)hipGraphExecDestroy(kv.second);
        m.clear();
        if (sweep) (void)hipGraphExecDestroy(sweep);
        sweep = nullptr;
        sweep_key.clear();
        seen_key.clear();
    }
};
std::map<const tgvx_ctx*, GraphCache> g_graphs;

// Polls of a grid-barrier wait before it gives up (each poll sleeps about
// 64 cycles and reads the counter from memory: about a second in all, far
// above any legitimate wait of a few microseconds). The environment variable
// TGVX_GRID_SPIN replaces it (0 makes every wait that has to poll give up at
// once, which is how the tests drive the failure path).
constexpr uint32_t kGridSpin = 1u << 20;
inline uint32_t grid_spin() {
    const char* e = std::getenv("TGVX_GRID_SPIN");
    return e && *e ? (uint32_t)std::strtoul(e, nullptr, 10) : kGridSpin;
}

// Whether a level of n cubes fits one cooperative launch of k_persist with
// workgroups of `block` threads (all workgroups co-resident).
template <class S>
bool persist_fits(tgvx_ctx* c, uint32_t n, int block) {
    const void* f = block == 1024 ? (const void*)k_persist<S, 1024> : (const void*)k_persist<S, 256>;
    int per_cu = 0, cus = 0;
    TGVX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, block, 0));
    TGVX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    const uint64_t grid = ((uint64_t)n + block - 1) / block;
    return n > 0 && per_cu > 0 && grid <= (uint64_t)per_cu * (uint64_t)cus;
}

// Variant 7: `iters` iterations of a's range in one cooperative launch; false
// (nothing launched) when its workgroups cannot all be co-resident.
template <class S>
bool launch_persist(tgvx_ctx* c, KArgs<S>& a, int iters, int block) {
    if (!persist_fits<S>(c, a.n, block)) return false;
    const bool fresh = c->grid_bar.p == nullptr;
    c->grid_bar.ensure(512);
    int32_t* const w = c->grid_bar.as<int32_t>();
    // the counter restarts at 0 for every launch; the error flag is cleared
    // once, and again by the host after it reports a failure
    k_fill_i32<<<1, 128, 0, c->stream>>>(w, fresh ? 128u : 1u, 0);
    GridBar gb{reinterpret_cast<uint32_t*>(w), w + 64, grid_spin()};
    const uint32_t grid = (a.n + block - 1) / block;
    void* args[] = {&a, &iters, &gb};
    const void* f = block == 1024 ? (const void*)k_persist<S, 1024> : (const void*)k_persist<S, 256>;
    TGVX_HIP(hipLaunchCooperativeKernel(f, dim3(grid), dim3(block), args, 0, c->stream));
    c->grid_used = true;
    return true;
}

// Microseconds per grid barrier of a cooperative launch of `wgs` workgroups
// of `block` threads: two launches of 10 and 10 + nbar barriers, on events.
void grid_barrier_bench_run(tgvx_ctx* c, int wgs, int block, int nbar, double* us) {
    if (block != 256 && block != 1024) throw Error(TGVX_E_ARG, "block must be 256 or 1024");
    if (nbar < 1 || nbar > 100000) throw Error(TGVX_E_ARG, "barriers must be in 1..100000");
    const void* f = block == 1024 ? (const void*)k_barrier_bench<1024> : (const void*)k_barrier_bench<256>;
    int per_cu = 0, cus = 0;
    TGVX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, block, 0));
    TGVX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    if (wgs < 1 || (int64_t)wgs > (int64_t)per_cu * cus)
        throw Error(TGVX_E_ARG, "workgroups must be in 1.." + std::to_string((int64_t)per_cu * cus));
    const bool fresh = c->grid_bar.p == nullptr;
    c->grid_bar.ensure(512);
    int32_t* const w = c->grid_bar.as<int32_t>();
    GridBar gb{reinterpret_cast<uint32_t*>(w), w + 64, grid_spin()};
    float ms[2] = {0.f, 0.f};
    const int counts[2] = {10, 10 + nbar};
    for (int r = 0; r < 2; ++r) {
        k_fill_i32<<<1, 128, 0, c->stream>>>(w, fresh && r == 0 ? 128u : 1u, 0);
        int k = counts[r];
        void* args[] = {&gb, &k};
        TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
        TGVX_HIP(hipLaunchCooperativeKernel(f, dim3(wgs), dim3(block), args, 0, c->stream));
        TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
        c->grid_used = true;
        TGVX_HIP(hipEventSynchronize(c->ev[1]));
        TGVX_HIP(hipEventElapsedTime(&ms[r], c->ev[0], c->ev[1]));
    }
    *us = 1e3 * (double)(ms[1] - ms[0]) / nbar;
}

// After the stream has drained past persistent launches: fail if one of
// their grid barriers gave up (the state is then not the iteration's).
void check_grid(tgvx_ctx* c) {
    if (!c->grid_used) return;
    c->grid_used = false;
    int32_t e = 0;
    int32_t* const w = c->grid_bar.as<int32_t>();
    TGVX_HIP(hipMemcpyAsync(&e, w + 64, sizeof(e), hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    if (e != 0) {
        k_fill_i32<<<1, 128, 0, c->stream>>>(w, 128u, 0);
        TGVX_HIP(hipStreamSynchronize(c->stream));
        throw Error(TGVX_E_STATE, "persistent level kernel: a grid barrier gave up waiting");
    }
}

// Variant 6 on a level larger than kLevel1: the two-iteration kernel, tiles of
// 1024 once the level has one per CU (2^18 cubes; measured 2.6-6 % faster than
// 512 at 0.38-4 M cubes, 12 % slower at 97 k), else 512.
inline Launch level1_fallback(uint32_t n) { return Launch{5, n >= (1u << 18) ? 1024 : 512, 1}; }

// Whether variant 6 runs a's range as one k_level1 launch: a whole level of
// at most kLevel1 cubes, outside the frozen-border and streaming solves (both
// operator modes).
template <class S>
bool level1_runs(const KArgs<S>& a, bool frozen) {
    return !frozen && !a.sflag && a.n <= (uint32_t)kLevel1 && a.off == a.lvl_off && a.n == a.lvl_n;
}

template <class S>
void run_launch(tgvx_ctx* c, KArgs<S>& a, const tgvx_params* p, int iters, bool frozen, const Launch& L0,
                bool cold = false, bool need_head = false) {
    if (iters <= 0 || a.n == 0) return;
    Launch L = L0;
    if (L.variant == 6) {
        if (level1_runs(a, frozen)) {
            const bool ref = p->op_mode == TGVX_OP_REF;
            if (a.n <= 64) {
                if (ref) k_level1<S, 64, true><<<1, 64, 0, c->stream>>>(a, iters);
                else k_level1<S, 64><<<1, 64, 0, c->stream>>>(a, iters);
            } else if (a.n <= 256) {
                if (ref) k_level1<S, 256, true><<<1, 256, 0, c->stream>>>(a, iters);
                else k_level1<S, 256><<<1, 256, 0, c->stream>>>(a, iters);
            } else {
                if (ref) k_level1<S, kLevel1, true><<<1, kLevel1, 0, c->stream>>>(a, iters);
                else k_level1<S, kLevel1><<<1, kLevel1, 0, c->stream>>>(a, iters);
            }
            TGVX_HIP(hipGetLastError());
            return;
        }
        L = level1_fallback(a.n);
    }
    if (L.variant == 7) {
        if (!frozen && p->op_mode == TGVX_OP_ADJ && !a.sflag && launch_persist<S>(c, a, iters, L.block)) {
            TGVX_HIP(hipGetLastError());
            return;
        }
        L = level1_fallback(a.n);
    }
    // a frozen-border leaf's fused iterations read its out-of-leaf neighbours
    // from the records; without them, the two kernels
    if (frozen && !a.rfz_bar) L.variant = 0;
    const bool fused = L.variant != 0 && iters >= 2;
    // a level initialised for the head launch (u, v only) must take it
    if (need_head && !(fused && (L.variant == 5 || L.variant == 8)))
        throw Error(TGVX_E_STATE, "a level initialised for the head launch runs without it");
    if (fused) alt_set<S>(c);  // allocate outside any capture
    if (fused && L.variant == 4) halo_tab<S>(c, a, L.block);
    if (fused && (L.variant == 5 || L.variant == 8)) {
        if (p->op_mode == TGVX_OP_REF) halo2_tab<S, true>(c, a, L.block);
        else halo2_tab<S>(c, a, L.block);
    }
    if (fused && L.variant == 8 && !frozen && iters >= 4) {
        if (p->op_mode == TGVX_OP_REF) halo4_tab<S, true>(c, a, L.block);
        else halo4_tab<S>(c, a, L.block);
    }
    if (p->use_graph && iters > 1 && !c->in_capture) {
        GraphKey key{a.off,
                     a.n,
                     a.lvl_off,
                     a.lvl_n,
                     iters,
                     p->op_mode,
                     L.block,
                     c->prob.precision,
                     fused ? L.variant : 0,
                     L.xcd,
                     a.bact ? 1 : 0,
                     (double)a.sigma,
                     (double)a.tau,
                     (double)a.a0,
                     (double)a.a1,
                     (double)a.tl,
                     frozen ? 1 : 0,
                     frozen ? a.leaf_first : 0u,
                     frozen ? a.leaf_last : 0u,
                     fused && cold ? 1 : 0};
        GraphCache& gc = g_graphs[c];
        auto it = gc.m.find(key);
        if (it == gc.m.end()) {
            hipGraph_t graph;
            TGVX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            if (fused) launch_fused<S>(c, a, iters, L, p->op_mode, frozen, cold, need_head);
            else
                for (int i = 0; i < iters; ++i) launch_iter<S>(c, a, p->op_mode, L, frozen);
            TGVX_HIP(hipStreamEndCapture(c->stream, &graph));
            hipGraphExec_t exec;
            TGVX_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            TGVX_HIP(hipGraphDestroy(graph));
            it = gc.m.emplace(key, exec).first;
        }
        TGVX_HIP(hipGraphLaunch(it->second, c->stream));
        return;
    }
    if (fused) launch_fused<S>(c, a, iters, L, p->op_mode, frozen, cold, need_head);
    else
        for (int i = 0; i < iters; ++i) launch_iter<S>(c, a, p->op_mode, L, frozen);
    TGVX_HIP(hipGetLastError());
}

// cold: the level starts from its initial state (u_bar = u, v_bar = v,
// p = q = 0), which lets the two-iteration kernel open with its head launch
template <class S>
void run_iters(tgvx_ctx* c, KArgs<S>& a, const tgvx_params* p, int iters, bool frozen, bool cold = false,
               bool need_head = false) {
    run_launch<S>(c, a, p, iters, frozen, launch_for(c, p, a.n, frozen), cold, need_head);
}

// Whether run_iters(cold) opens this whole level with the head launch: the
// choice run_launch and launch_fused_x make (variant 5 or 8, directly or as
// variant 6's fallback above kLevel1 cubes, and the level's tile sets fit).
template <class S>
bool opens_with_head(tgvx_ctx* c, const KArgs<S>& a, const tgvx_params* p, int iters) {
    if (iters < 2 || a.n == 0 || a.off != a.lvl_off || a.n != a.lvl_n) return false;
    Launch L = launch_for(c, p, a.n, false);
    if (L.variant == 6) {
        if (level1_runs(a, false)) return false;
        L = level1_fallback(a.n);
    }
    if (L.variant != 5 && L.variant != 8) return false;
    return p->op_mode == TGVX_OP_REF ? halo2_tab<S, true>(c, a, L.block).ok : halo2_tab<S>(c, a, L.block).ok;
}

// The cold initialisation of level li > 0 (tgv.cpp:391-409), with the
// previous level's u_bar, v_bar records (exact-adjoint mode, left to it by
// prepare_level(..., false)) in the same launch. uv_only: the level opens
// with the head launch, which reads u, v alone.
template <class S>
void parent_init_cold(tgvx_ctx* c, const KArgs<S>& a, bool uv_only) {
    S* rec = a.fzaos ? c->prob.d_fzaos.as<S>() : nullptr;
    const uint32_t m = std::max(a.n, rec ? a.fz_n : 0u);
    if (m == 0) return;
    if (uv_only) k_parent_init<S, false><<<(m + 255) / 256, 256, 0, c->stream>>>(a, rec);
    else k_parent_init<S, true><<<(m + 255) / 256, 256, 0, c->stream>>>(a, rec);
}

template <class S>
void eval_gap(tgvx_ctx* c, KArgs<S> a, const tgvx_params* p, double* Pout, double* Dout) {
    Problem& P = c->prob;
    a.off = a.lvl_off;
    a.n = a.lvl_n;
    const uint32_t blocks = (a.n + kGapBlock - 1) / kGapBlock;
    P.d_part.ensure(sizeof(double) * 2 * std::max<uint32_t>(blocks, 1));
    P.d_pd.ensure(sizeof(double) * 2);
    if (blocks > 0)
        k_gap<S><<<blocks, kGapBlock, 0, c->stream>>>(a, (double)p->lambda, (double)p->alpha0, (double)p->alpha1,
                                                      2.0, P.d_part.as<double>());
    k_sum_pairs<<<1, kGapBlock, 0, c->stream>>>(P.d_part.as<double>(), blocks, P.d_pd.as<double>());
    TGVX_HIP(hipGetLastError());
    double pd[2];
    TGVX_HIP(hipMemcpyAsync(pd, P.d_pd.p, sizeof(pd), hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    *Pout = pd[0];
    *Dout = pd[1];
}

// P, D of the level into dst (device), without waiting for them
template <class S>
void eval_gap_to(tgvx_ctx* c, KArgs<S> a, const tgvx_params* p, double* dst) {
    Problem& P = c->prob;
    a.off = a.lvl_off;
    a.n = a.lvl_n;
    const uint32_t blocks = (a.n + kGapBlock - 1) / kGapBlock;
    P.d_part.ensure(sizeof(double) * 2 * std::max<uint32_t>(blocks, 1));
    if (blocks > 0)
        k_gap<S><<<blocks, kGapBlock, 0, c->stream>>>(a, (double)p->lambda, (double)p->alpha0, (double)p->alpha1,
                                                      2.0, P.d_part.as<double>());
    k_sum_pairs<<<1, kGapBlock, 0, c->stream>>>(P.d_part.as<double>(), blocks, dst);
    TGVX_HIP(hipGetLastError());
}

double rel_gap(double P, double D) { return (P - D) / std::max(1.0, std::fabs(P)); }

// Per-block relative gaps of level li of a problem holding several blocks.
template <class S>
void eval_gap_blocks(tgvx_ctx* c, KArgs<S> a, const tgvx_params* p, size_t li, std::vector<double>& rg) {
    Problem& P = c->prob;
    const uint32_t nb = P.nblocks;
    P.d_seg.ensure(sizeof(uint32_t) * (nb + 1));
    P.d_pd.ensure(sizeof(double) * 2 * std::max<uint32_t>(nb, 1));
    TGVX_HIP(hipMemcpyAsync(P.d_seg.p, P.seg[li].data(), sizeof(uint32_t) * (nb + 1), hipMemcpyHostToDevice,
                            c->stream));
    a.off = a.lvl_off;
    a.n = a.lvl_n;
    k_gap_seg<S><<<nb, kGapBlock, 0, c->stream>>>(a, P.d_seg.as<uint32_t>(), (double)p->lambda, (double)p->alpha0,
                                                   (double)p->alpha1, 2.0, P.d_pd.as<double>());
    TGVX_HIP(hipGetLastError());
    std::vector<double> pd(2 * nb);
    TGVX_HIP(hipMemcpyAsync(pd.data(), P.d_pd.p, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, c->stream));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    rg.resize(nb);
    for (uint32_t b = 0; b < nb; ++b) rg[b] = rel_gap(pd[2 * b], pd[2 * b + 1]);
}

// A fixed-iteration solve of one block from the initial state as one graph
// (no reference counterpart): k_reset, then per level the previous level's
// u_bar, v_bar records, the parent initialisation, the iterations and,
// unless no_energy, the FP64 energy, all nodes of one graph, so nothing
// waits on the host between levels. The launches are the ones solve_t makes
// for such a solve, in the same order. A sweep is captured on its second run
// with the same parameters, when every table and buffer it reads exists
// (they are never reallocated while the problem stays loaded); the per-level
// times are not taken (use_graph = 2 keeps the per-level graphs and takes
// them). Returns false when the solve is not of this kind or has not been
// seen before: the caller runs it level by level.
template <class S>
void enqueue_sweep(tgvx_ctx* c, const tgvx_params* p, KArgs<S>& a) {
    Problem& P = c->prob;
    // exact-adjoint levels set their own p, q before anything reads them
    const uint32_t npq = p->op_mode == TGVX_OP_ADJ ? P.levels[0].n : (uint32_t)P.n;
    k_reset<S><<<(unsigned)((P.n + 255) / 256), 256, 0, c->stream>>>(a, (uint32_t)P.n, npq);
    for (size_t li = 0; li < P.levels.size(); ++li) {
        const Level& L = P.levels[li];
        a.off = a.lvl_off = L.off;
        a.n = a.lvl_n = L.n;
        set_steps<S>(a, p, L.depth);
        prepare_level<S>(c, a, p, li, false);
        const int it = P.level_iters.empty() ? p->iters : P.level_iters[li];
        const bool head = li > 0 && opens_with_head<S>(c, a, p, it);
        if (li > 0) parent_init_cold<S>(c, a, head);
        run_iters<S>(c, a, p, it, false, true, head);
        if (!p->no_energy) eval_gap_to<S>(c, a, p, P.d_lvlpd.as<double>() + 2 * li);
    }
    TGVX_HIP(hipGetLastError());
}

template <class S>
bool sweep_graph(tgvx_ctx* c, const tgvx_params* p, tgvx_report* rep) {
    Problem& P = c->prob;
    if (p->use_graph != 1 || p->gap_tol > 0.0 || P.nblocks != 1 || p->warm_start || p->dual_eps > 0.0f ||
        P.levels.size() > (size_t)TGVX_MAX_LEVELS || P.n == 0)
        return false;
    for (const Level& L : P.levels)
        if (launch_for(c, p, L.n, false).variant == 7) return false;  // cooperative launches stay eager
    // the parameters' bytes up to the last field (the struct has no inner
    // padding; its tail padding may hold anything) and the per-level counts
    constexpr size_t kp = offsetof(tgvx_params, no_energy) + sizeof(int32_t);
    std::vector<char> key(kp + sizeof(int32_t) * P.level_iters.size());
    std::memcpy(key.data(), p, kp);
    if (!P.level_iters.empty())
        std::memcpy(key.data() + kp, P.level_iters.data(), sizeof(int32_t) * P.level_iters.size());
    GraphCache& gc = g_graphs[c];
    if (!(gc.sweep && gc.sweep_key == key)) {
        if (gc.seen_key != key) {
            gc.seen_key = key;
            return false;
        }
        KArgs<S> a = make_args<S>(c, p);
        a.init_stride = 4;
        TGVX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->in_capture = true;
        try {
            enqueue_sweep<S>(c, p, a);
        } catch (...) {
            // not capturable after all (a table missing): run level by level
            c->in_capture = false;
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            (void)hipGetLastError();
            gc.seen_key.clear();
            return false;
        }
        c->in_capture = false;
        hipGraph_t graph = nullptr;
        TGVX_HIP(hipStreamEndCapture(c->stream, &graph));
        hipGraphExec_t exec = nullptr;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        TGVX_HIP(e);
        if (gc.sweep) (void)hipGraphExecDestroy(gc.sweep);
        gc.sweep = exec;
        gc.sweep_key = key;
    }
    tgvx_report r;
    std::memset(&r, 0, sizeof(r));
    P.block_iters.assign(P.levels.size(), -1);
    P.block_gap.assign(P.levels.size(), 0.0);
    const double t0 = now_ms();
    TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
    TGVX_HIP(hipGraphLaunch(gc.sweep, c->stream));
    TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
    TGVX_HIP(hipEventSynchronize(c->ev[1]));
    check_grid(c);
    float ms = 0.f;
    TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    std::vector<double> pd(2 * P.levels.size(), 0.0);
    if (!p->no_energy && !pd.empty())
        TGVX_HIP(hipMemcpy(pd.data(), P.d_lvlpd.p, sizeof(double) * pd.size(), hipMemcpyDeviceToHost));
    for (size_t li = 0; li < P.levels.size(); ++li) {
        const Level& L = P.levels[li];
        const int it = P.level_iters.empty() ? p->iters : P.level_iters[li];
        r.level_depth[li] = L.depth;
        r.level_leaves[li] = L.n;
        r.level_iters[li] = it;
        r.level_energy[li] = pd[2 * li];
        r.level_dual[li] = pd[2 * li + 1];
        r.leaf_iters += (double)L.n * it;
    }
    r.n_levels = (int)P.levels.size();
    r.solve_ms = now_ms() - t0;
    r.iterate_ms = ms;
    if (rep) {
        const double setup = rep->setup_ms;
        *rep = r;
        rep->setup_ms = setup;
    }
    return true;
}

template <class S>
void solve_t(tgvx_ctx* c, const tgvx_params* p, const float* init_uv, tgvx_report* rep,
             const tgvx_treetop_leaf* leaves, uint64_t nleaves, int init_stride) {
    Problem& P = c->prob;
    KArgs<S> a = make_args<S>(c, p);
    const bool to_gap = p->gap_tol > 0.0;
    const int chunk = p->chunk > 0 ? p->chunk : 50;
    const int max_iters = p->max_iters > 0 ? p->max_iters : 100000;
    tgvx_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = now_ms();
    if (p->warm_start && (init_uv || leaves))
        throw Error(TGVX_E_ARG, "warm_start excludes an explicit initial field and the frozen-border solve");
    if (P.nblocks > 1 && (p->warm_start || init_uv || leaves))
        throw Error(TGVX_E_ARG, "a problem of several blocks takes neither warm starts nor treetop leaves");
    if (!init_uv && !leaves && sweep_graph<S>(c, p, rep)) return;
    P.block_iters.assign(P.levels.size() * P.nblocks, -1);
    P.block_gap.assign(P.levels.size() * P.nblocks, 0.0);
    a.init_stride = 4;
    const bool full = init_uv && init_stride == 13;  // a whole state: no dual recovery after it
    if (init_uv) {
        const size_t w = (size_t)init_stride;
        P.d_init.ensure(sizeof(float) * w * P.n);
        std::vector<float> perm_init(w * P.n);
        for (uint64_t g = 0; g < P.n; ++g)
            std::memcpy(&perm_init[w * g], init_uv + w * (size_t)P.perm[g], sizeof(float) * w);
        h2d(c, P.d_init.p, perm_init.data(), sizeof(float) * w * P.n);
        TGVX_HIP(hipStreamSynchronize(c->stream));
        a.init = P.d_init.as<float>();
        a.init_stride = init_stride;
    }
    const double t_solve0 = now_ms();
    // an exact-adjoint solve from the cold initialisation sets every later
    // level's p, q before anything reads them
    const bool lean = p->op_mode == TGVX_OP_ADJ && !init_uv && !p->warm_start && !leaves && P.nblocks == 1;
    if (P.n > 0) {
        const uint32_t npq = lean && !P.levels.empty() ? P.levels[0].n : (uint32_t)P.n;
        k_reset<S><<<(unsigned)((P.n + 255) / 256), 256, 0, c->stream>>>(a, (uint32_t)P.n, npq);
        TGVX_HIP(hipGetLastError());
    }
    double it_ms = 0.0, gap_ms = 0.0;
    std::vector<S*> fz_store;
    std::vector<int> deferred(std::min<size_t>(P.levels.size(), TGVX_MAX_LEVELS), 0);
    P.d_lvlpd.ensure(sizeof(double) * 2 * TGVX_MAX_LEVELS);
    for (size_t li = 0; li < P.levels.size(); ++li) {
        const Level& L = P.levels[li];
        a.off = L.off;
        a.n = L.n;
        a.lvl_off = L.off;
        a.lvl_n = L.n;
        set_steps<S>(a, p, L.depth);
        const unsigned grid = (L.n + 255) / 256;
        int warm_kept = 0;
        const bool field = init_uv || p->warm_start;
        // a cold level's parent initialisation takes the previous level's
        // records along (prepare_level leaves them to it)
        if (!leaves) prepare_level<S>(c, a, p, li, field);
        // the level's initialisation; with dual_eps the dual follows from (u, v)
        const bool rec = p->dual_eps > 0.0f && !leaves;
        auto recover = [&] {
            if (rec) k_dual_recover<S><<<grid, 256, 0, c->stream>>>(a, (typename Io<S>::T)p->dual_eps);
        };
        // fixed iterations from the cold initialisation, level by level with
        // nothing read between: a level that opens with the head launch
        // needs u, v only
        const bool fixed_deferred = !to_gap && P.nblocks == 1 && li < (size_t)TGVX_MAX_LEVELS && !leaves;
        const int fixed_iters = P.level_iters.empty() ? p->iters : P.level_iters[li];
        bool need_head = false;
        if (p->warm_start) {
            // the network needs the parent level's final state: features now
            warm_init(c, L.off, L.n);
            a.init = P.d_init.as<float>();
        }
        if (field) {
            // warm_dual: keep the dual recovered from the cold initialisation
            const int keep = rec && p->warm_dual ? 1 : 0;
            if (p->warm_fallback) {
                // keep the warm start only where it does not start from a larger gap
                if (li > 0) k_parent_init<S><<<grid, 256, 0, c->stream>>>(a, nullptr);
                recover();
                double Pc, Dc, Pw, Dw;
                eval_gap<S>(c, a, p, &Pc, &Dc);
                k_field_init<S><<<grid, 256, 0, c->stream>>>(a, 1, keep);
                if (!keep && !full) recover();
                eval_gap<S>(c, a, p, &Pw, &Dw);
                if (Pw - Dw > Pc - Dc) {
                    k_field_init<S><<<grid, 256, 0, c->stream>>>(a, 2, keep);
                    if (!keep) recover();
                } else {
                    warm_kept = 1;
                }
            } else {
                if (keep) {
                    if (li > 0) k_parent_init<S><<<grid, 256, 0, c->stream>>>(a, nullptr);
                    recover();
                }
                k_field_init<S><<<grid, 256, 0, c->stream>>>(a, 0, keep);
                if (!keep && !full) recover();
                warm_kept = 1;
            }
        } else {
            if (li > 0) {
                need_head = fixed_deferred && !rec && opens_with_head<S>(c, a, p, fixed_iters);
                if (leaves) k_parent_init<S><<<grid, 256, 0, c->stream>>>(a, nullptr);
                else parent_init_cold<S>(c, a, need_head);
            }
            recover();
        }
        TGVX_HIP(hipGetLastError());
        // the level's first iterations start from its initial state
        const bool cold = !field && !rec;
        int done = 0;
        double work = -1.0;  // cube-iterations actually run, when not n * done
        double Pv = 0.0, Dv = 0.0;
        if (leaves) {
            // tgv_minimize_oc: snapshot u_bar / v_bar, then each treetop leaf's
            // part of the level in turn (tgv.cpp:482-507).
            for (int i = 0; i < 4; ++i) {
                TGVX_HIP(hipMemcpyAsync(P.d_save[i].p, P.d_state[F_UB + i].p, P.elem * P.n,
                                        hipMemcpyDeviceToDevice, c->stream));
                a.fz[i] = P.d_save[i].as<S>();
            }
            a.orig = P.d_orig.as<uint32_t>();
            const uint32_t* pb = P.perm.data() + L.off;
            // reference rules with a fused variant: each leaf's part of the
            // level is the two-iteration kernel's level, its out-of-leaf
            // neighbours gathered into records at the leaf's start (u_bar,
            // v_bar from the snapshot, p, q live)
            const bool ref_fused = p->op_mode == TGVX_OP_REF && launch_for(c, p, L.n, true).variant != 0;
            S *rbar = nullptr, *rdual = nullptr;
            if (ref_fused) rfz_bufs<S>(c, &rbar, &rdual);
            TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
            for (uint64_t l = 0; l < nleaves; ++l) {
                const uint32_t* lo = std::lower_bound(pb, pb + L.n, leaves[l].first_idx);
                const uint32_t* hi = std::upper_bound(pb, pb + L.n, leaves[l].last_idx);
                if (lo == hi) continue;
                a.off = L.off + (uint32_t)(lo - pb);
                a.n = (uint32_t)(hi - lo);
                a.leaf_first = leaves[l].first_idx;
                a.leaf_last = leaves[l].last_idx;
                if (ref_fused) {
                    a.lvl_off = a.off;
                    a.lvl_n = a.n;
                    k_ref_frozen<S, true><<<(a.n + 255) / 256, 256, 0, c->stream>>>(a, rbar, rdual, nullptr);
                    a.rfz_bar = rbar;
                    a.rfz_dual = rdual;
                }
                run_iters<S>(c, a, p, p->iters, true, cold);
            }
            a.lvl_off = L.off;
            a.lvl_n = L.n;
            a.rfz_bar = a.rfz_dual = nullptr;
            TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
            TGVX_HIP(hipEventSynchronize(c->ev[1]));
            float ms = 0.f;
            TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            it_ms += ms;
            if (li < (size_t)TGVX_MAX_LEVELS) r.level_ms[li] = ms;
            a.off = L.off;
            a.n = L.n;
            done = p->iters;
        } else if (!to_gap && P.nblocks == 1 && li < (size_t)TGVX_MAX_LEVELS) {
            // fixed iterations: nothing waits for the level; its time and
            // P, D are read after the last level
            done = fixed_iters;
            TGVX_HIP(hipEventRecord(c->ev[4 + 2 * li], c->stream));
            run_iters<S>(c, a, p, done, false, cold, need_head);
            TGVX_HIP(hipEventRecord(c->ev[5 + 2 * li], c->stream));
            if (!p->no_energy) eval_gap_to<S>(c, a, p, P.d_lvlpd.as<double>() + 2 * li);
            deferred[li] = p->no_energy ? 2 : 1;
        } else if (!to_gap) {
            const int fixed = P.level_iters.empty() ? p->iters : P.level_iters[li];
            TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
            run_iters<S>(c, a, p, fixed, false, cold);
            TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
            TGVX_HIP(hipEventSynchronize(c->ev[1]));
            check_grid(c);
            float ms = 0.f;
            TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
            it_ms += ms;
            done = fixed;
            const double g0 = now_ms();
            if (!p->no_energy) eval_gap<S>(c, a, p, &Pv, &Dv);
            if (P.nblocks > 1) {
                std::vector<double> rg(P.nblocks, 0.0);
                if (!p->no_energy) eval_gap_blocks<S>(c, a, p, li, rg);
                for (uint32_t b = 0; b < P.nblocks; ++b) {
                    const bool has = P.seg[li][b + 1] > P.seg[li][b];
                    P.block_iters[li * P.nblocks + b] = has ? done : -1;
                    P.block_gap[li * P.nblocks + b] = has ? rg[b] : 0.0;
                }
            }
            gap_ms += now_ms() - g0;
        } else if (P.nblocks > 1) {
            // several blocks: iterate the whole level until every block has
            // reached the gap; each block's count is where it first did
            std::vector<double> rg;
            int* bi = &P.block_iters[li * P.nblocks];
            std::vector<uint8_t> act(P.nblocks);
            auto upload = [&] {
                for (uint32_t b = 0; b < P.nblocks; ++b) act[b] = bi[b] < 0 && P.seg[li][b + 1] > P.seg[li][b];
                TGVX_HIP(hipMemcpyAsync(P.d_bact.p, act.data(), P.nblocks, hipMemcpyHostToDevice, c->stream));
            };
            a.blk = P.d_blk.as<int32_t>();
            a.bact = P.d_bact.as<uint8_t>();
            double g0 = now_ms();
            eval_gap_blocks<S>(c, a, p, li, rg);
            gap_ms += now_ms() - g0;
            auto mark = [&] {
                bool all = true;
                for (uint32_t b = 0; b < P.nblocks; ++b) {
                    if (!std::isfinite(rg[b]))
                        throw Error(TGVX_E_NUMERIC, "non-finite gap at depth " + std::to_string(L.depth));
                    if (P.seg[li][b + 1] == P.seg[li][b]) continue;
                    if (rg[b] <= p->gap_tol) {
                        if (bi[b] < 0) {
                            bi[b] = done;
                            P.block_gap[li * P.nblocks + b] = rg[b];
                        }
                    } else {
                        all = false;
                    }
                }
                return all;
            };
            while (!mark() && done < max_iters) {
                const int k = std::min(chunk, max_iters - done);
                upload();  // blocks that reached the gap keep their state
                TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
                run_iters<S>(c, a, p, k, false, cold && done == 0);
                TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
                TGVX_HIP(hipEventSynchronize(c->ev[1]));
                check_grid(c);
                done += k;
                g0 = now_ms();
                eval_gap_blocks<S>(c, a, p, li, rg);
                gap_ms += now_ms() - g0;
                float ms = 0.f;
                TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                it_ms += ms;
            }
            work = 0.0;
            for (uint32_t b = 0; b < P.nblocks; ++b) {
                if (bi[b] < 0) P.block_gap[li * P.nblocks + b] = rg[b];  // not reached (or no cube there)
                work += (double)(P.seg[li][b + 1] - P.seg[li][b]) * (bi[b] < 0 ? done : bi[b]);
            }
            a.blk = nullptr;
            a.bact = nullptr;
            eval_gap<S>(c, a, p, &Pv, &Dv);
        } else {
            double g0 = now_ms();
            eval_gap<S>(c, a, p, &Pv, &Dv);
            gap_ms += now_ms() - g0;
            while (done < max_iters && rel_gap(Pv, Dv) > p->gap_tol) {
                if (!std::isfinite(Pv) || !std::isfinite(Dv))
                    throw Error(TGVX_E_NUMERIC, "non-finite energy or gap at depth " + std::to_string(L.depth));
                const int k = std::min(chunk, max_iters - done);
                TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
                run_iters<S>(c, a, p, k, false, cold && done == 0);
                TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
                TGVX_HIP(hipEventSynchronize(c->ev[1]));
                check_grid(c);
                done += k;
                g0 = now_ms();
                eval_gap<S>(c, a, p, &Pv, &Dv);
                gap_ms += now_ms() - g0;
                float ms = 0.f;
                TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                it_ms += ms;
            }
        }
        if (li < (size_t)TGVX_MAX_LEVELS) {
            r.level_depth[li] = L.depth;
            r.level_leaves[li] = L.n;
            r.level_iters[li] = done;
            r.level_warm[li] = warm_kept;
            r.level_energy[li] = Pv;
            r.level_dual[li] = Dv;
        }
        r.leaf_iters += work >= 0.0 ? work : (double)L.n * done;
    }
    TGVX_HIP(hipStreamSynchronize(c->stream));
    check_grid(c);
    if (std::find_if(deferred.begin(), deferred.end(), [](int d) { return d != 0; }) != deferred.end()) {
        // deferred 1: P, D evaluated on the device; 2: iterations only
        std::vector<double> pd(2 * deferred.size(), 0.0);
        if (std::find(deferred.begin(), deferred.end(), 1) != deferred.end())
            TGVX_HIP(hipMemcpy(pd.data(), P.d_lvlpd.p, sizeof(double) * pd.size(), hipMemcpyDeviceToHost));
        for (size_t li = 0; li < deferred.size(); ++li) {
            if (!deferred[li]) continue;
            r.level_energy[li] = deferred[li] == 1 ? pd[2 * li] : 0.0;
            r.level_dual[li] = deferred[li] == 1 ? pd[2 * li + 1] : 0.0;
            float ms = 0.f;
            TGVX_HIP(hipEventElapsedTime(&ms, c->ev[4 + 2 * li], c->ev[5 + 2 * li]));
            it_ms += ms;
            r.level_ms[li] = ms;
        }
    }
    r.n_levels = (int)std::min<size_t>(P.levels.size(), TGVX_MAX_LEVELS);
    r.solve_ms = now_ms() - t_solve0;
    r.iterate_ms = it_ms;
    r.gap_ms = gap_ms;
    (void)t0;
    if (rep) {
        const double setup = rep->setup_ms;
        *rep = r;
        rep->setup_ms = setup;
    }
}

template <class S>
KArgs<S> level_args(tgvx_ctx* c, const tgvx_params* p, const Level& L) {
    KArgs<S> a = make_args<S>(c, p);
    a.off = a.lvl_off = L.off;
    a.n = a.lvl_n = L.n;
    return a;
}

template <class S>
void bench_t(tgvx_ctx* c, int depth, const tgvx_params* p, int iters, double* total_ms, double* dual_ms,
             double* primal_ms, double* fused_ms) {
    Problem& P = c->prob;
    const Level* L = nullptr;
    for (const Level& l : P.levels)
        if (l.depth == depth) L = &l;
    if (!L) throw Error(TGVX_E_ARG, "no cubes at depth " + std::to_string(depth));
    KArgs<S> a = level_args<S>(c, p, *L);
    prepare_level<S>(c, a, p, (size_t)(L - P.levels.data()));
    const Launch cfg = launch_for(c, p, a.n, false);
    // whole iterations (graph or plain launches as configured)
    TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
    run_launch<S>(c, a, p, iters, false, cfg);
    TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
    TGVX_HIP(hipEventSynchronize(c->ev[1]));
    check_grid(c);
    float ms = 0.f;
    TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    *total_ms = ms;
    // dual and primal kernels, each bracketed by its own events (the
    // configuration's workgroup size and tile order)
    const int k = std::max(2, std::min(iters, 20) & ~1);  // even: the fused passes below end on the main set
    Launch split = cfg;
    split.variant = 0;
    if (split.block > 256) split.block = 256;
    const unsigned grid = (a.n + split.block - 1) / split.block;
    double sd = 0.0, sp = 0.0;
    for (int i = 0; i < k; ++i) {
        TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
        if (split.xcd) k_dual<S, 0, true><<<grid, split.block, 0, c->stream>>>(a);
        else k_dual<S, 0, false><<<grid, split.block, 0, c->stream>>>(a);
        TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
        if (p->op_mode == TGVX_OP_ADJ) {
            const Set13<S> m = main_set(a);
            if (split.xcd) k_primal_adj<S, true><<<grid, split.block, 0, c->stream>>>(a, m, m, 0);
            else k_primal_adj<S, false><<<grid, split.block, 0, c->stream>>>(a, m, m, 0);
        } else {
            if (split.xcd) k_primal_ref<S, true><<<grid, split.block, 0, c->stream>>>(a);
            else k_primal_ref<S, false><<<grid, split.block, 0, c->stream>>>(a);
        }
        TGVX_HIP(hipEventRecord(c->ev[2], c->stream));
        TGVX_HIP(hipEventSynchronize(c->ev[2]));
        float m1 = 0.f, m2 = 0.f;
        TGVX_HIP(hipEventElapsedTime(&m1, c->ev[0], c->ev[1]));
        TGVX_HIP(hipEventElapsedTime(&m2, c->ev[1], c->ev[2]));
        sd += m1;
        sp += m2;
    }
    *dual_ms = sd / k;
    *primal_ms = sp / k;
    // the fused pass alone (the configuration's variant, the tiled kernel if
    // it has none; an even count, so the main set ends current)
    double sf = 0.0;
    const bool ref = p->op_mode != TGVX_OP_ADJ;
    // the reference's rules: the two-iteration kernel when configured and the
    // level's tile sets fit (they have no one-iteration fused kernel)
    const bool ref2 = ref && cfg.variant >= 5 &&
                      halo2_tab<S, true>(c, a, cfg.variant >= 6 ? level1_fallback(a.n).block : cfg.block).ok;
    if (!ref || ref2) {
        Launch fz = cfg;
        if (fz.variant == 0) fz = Launch{1, kFuseTile, cfg.xcd};
        if (fz.variant >= 6) fz = level1_fallback(a.n);
        // the two-iteration kernel: per launch / 2 (one iteration per
        // launch where the level's tile sets do not fit)
        int per = 1;
        if (fz.variant == 5) {
            if (ref2 || halo2_tab<S>(c, a, fz.block).ok) per = 2;
            else fz.variant = 1;
        }
        Set13<S> cur = main_set(a), oth = alt_set<S>(c);
        const int kf = 2 * k / per;  // 2k iterations whatever the variant
        for (int i = 0; i < kf; ++i) {
            TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
            if (per == 2 && ref) {
                if (fz.xcd) launch_fused2_x<S, true, true>(c, a, cur, oth, fz.block);
                else launch_fused2_x<S, false, true>(c, a, cur, oth, fz.block);
            } else if (per == 2) {
                if (fz.xcd) launch_fused2_x<S, true>(c, a, cur, oth, fz.block);
                else launch_fused2_x<S, false>(c, a, cur, oth, fz.block);
            } else {
                launch_fused_pass<S>(c, a, cur, oth, fz);
            }
            TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
            TGVX_HIP(hipEventSynchronize(c->ev[1]));
            check_grid(c);
            float m1 = 0.f;
            TGVX_HIP(hipEventElapsedTime(&m1, c->ev[0], c->ev[1]));
            sf += m1;
            std::swap(cur, oth);
        }
        sf /= (double)kf * per;
    }
    if (fused_ms) *fused_ms = sf;
}

// Autotuner: for every level of the loaded problem, time `iters` iterations
// of each launch configuration (as a captured graph when use_graph is set)
// and keep the fastest per (precision, operator mode, log2 size bucket).
template <class S>
void autotune_t(tgvx_ctx* c, const tgvx_params* p, int iters, tgvx_tune_report* rep) {
    Problem& P = c->prob;
    std::vector<Launch> cand;
    for (int x = 1; x >= 0; --x) {
        for (int b : {256, 128, 64}) cand.push_back(Launch{0, b, x});
        if (p->op_mode == TGVX_OP_ADJ) {
            for (int v : {1, 3, 4})
                for (int b : {256, 128, 512, 1024}) cand.push_back(Launch{v, b, x});
            for (int b : {256, 128}) cand.push_back(Launch{2, b, x});
        }
        for (int b : {256, 128, 512, 1024}) cand.push_back(Launch{5, b, x});
    }
    for (int b : {256, 128}) cand.push_back(Launch{8, b, 1});
    cand.push_back(Launch{6, kLevel1, 0});
    if (p->op_mode == TGVX_OP_ADJ)
        for (int b : {1024, 256}) cand.push_back(Launch{7, b, 0});
    // Kernel launches per iteration of a candidate over `iters` iterations of
    // a level of n cubes. Inside a sweep every launch costs about 1 us more
    // than when one level's graph is timed alone (measured at depths 4-6 of
    // the depth-10 workload: two kernels per iteration 7.5-7.8 us in the
    // sweep against 5.0-5.5 us alone), so candidates are compared by their
    // time plus kSweepLaunchUs per launch and iteration.
    constexpr double kSweepLaunchUs = 1.0;
    auto launches = [&](const Launch& cf, uint32_t n) -> double {
        const double k = (double)std::max(iters, 1);
        switch (cf.variant) {
            case 0: return 2.0;
            case 5: return (2.0 + (k - 1.0) / 2.0) / k;
            case 6: return n <= (uint32_t)kLevel1 ? 1.0 / k : (2.0 + (k - 1.0) / 2.0) / k;
            case 7: return 2.0 / k;  // the counter's clear and the launch
            case 8: return (2.0 + (k - 1.0) / 4.0) / k;
            default: return (k + 1.0) / k;
        }
    };
    tgvx_params q = *p;
    q.tune = 0;
    const Launch base = launch_for(c, &q, 1u << 20, false);
    if (rep) std::memset(rep, 0, sizeof(*rep));
    int li = 0;
    for (const Level& L : P.levels) {
        KArgs<S> a = level_args<S>(c, p, L);
        prepare_level<S>(c, a, p, (size_t)(&L - P.levels.data()));
        double best = 1e300, def = 0.0, best_score = 1e300, def_score = 0.0;
        Launch win = base;
        for (const Launch& cf : cand) {
            if (cf.variant == 6 && L.n > (uint32_t)kLevel1) continue;  // would run as 5
            if (cf.variant == 8 && L.n > (1u << 20)) continue;         // bandwidth-bound: two per launch
            if (cf.variant == 7 && !persist_fits<S>(c, L.n, cf.block)) continue;
            run_launch<S>(c, a, &q, iters, false, cf);  // warm-up (and graph build)
            double us = 1e300;
            for (int rep = 0; rep < 2; ++rep) {  // the better of two timings
                TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
                run_launch<S>(c, a, &q, iters, false, cf);
                TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
                TGVX_HIP(hipEventSynchronize(c->ev[1]));
                check_grid(c);
                float ms = 0.f;
                TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
                us = std::min(us, 1e3 * ms / iters);
            }
            const double score = us + kSweepLaunchUs * launches(cf, L.n);
            const bool is_base = cf.variant == base.variant && cf.block == base.block && cf.xcd == base.xcd;
            if (is_base) {
                def = us;
                def_score = score;
            }
            if (score < best_score) {
                best_score = score;
                best = us;
                win = cf;
            }
        }
        // keep the caller's configuration unless a candidate beats it by 3 %
        if (def > 0.0 && best_score > 0.97 * def_score) {
            best = def;
            win = base;
        }
        c->tuned[std::make_tuple(P.precision, p->op_mode, size_bucket(L.n))] = TuneEntry{win.variant, win.block,
                                                                                         win.xcd, best};
        if (rep && li < TGVX_MAX_LEVELS) {
            rep->level_depth[li] = L.depth;
            rep->level_leaves[li] = L.n;
            rep->variant[li] = win.variant;
            rep->block[li] = win.block;
            rep->xcd[li] = win.xcd;
            rep->best_us[li] = best;
            rep->default_us[li] = def;
        }
        ++li;
    }
    if (rep) rep->n_levels = std::min(li, TGVX_MAX_LEVELS);
    release_graphs(c);  // the candidates' graphs
    // keep only the tables the chosen configurations use (the candidates'
    // tables of other tile sizes hold about 100 B per cube each)
    std::set<std::tuple<uint32_t, uint32_t, int>> keep4, keep5, keep8;
    for (const Level& L : P.levels) {
        auto it = c->tuned.find(std::make_tuple(P.precision, p->op_mode, size_bucket(L.n)));
        if (it == c->tuned.end()) continue;
        const TuneEntry& e = it->second;
        if (e.variant == 4) keep4.insert(std::make_tuple(L.off, L.n, e.block));
        if (e.variant == 5 || e.variant == 8) keep5.insert(std::make_tuple(L.off, L.n, e.block));
        if (e.variant == 8) keep8.insert(std::make_tuple(L.off, L.n, e.block));
        if (e.variant == 6 && L.n > (uint32_t)kLevel1)
            keep5.insert(std::make_tuple(L.off, L.n, level1_fallback(L.n).block));
    }
    for (auto it = P.halo.begin(); it != P.halo.end();)
        it = keep4.count(it->first) ? std::next(it) : P.halo.erase(it);
    for (auto it = P.halo2.begin(); it != P.halo2.end();) {
        const auto k = std::make_tuple(std::get<0>(it->first), std::get<1>(it->first), std::get<2>(it->first));
        it = keep5.count(k) ? std::next(it) : P.halo2.erase(it);
    }
    for (auto it = P.halo4.begin(); it != P.halo4.end();) {
        const auto k = std::make_tuple(std::get<0>(it->first), std::get<1>(it->first), std::get<2>(it->first));
        it = keep8.count(k) ? std::next(it) : P.halo4.erase(it);
    }
}

// tgv_minimize_streaming (tgv.cpp:802-1070) with every treetop leaf of a
// level iterated in the same launch: within a level the leaves are
// independent (border values come from the level-start snapshot, border
// p, q are zero), so this equals the reference's leaf-by-leaf order. The
// parent copy still runs leaf after leaf, as there.
template <class S>
void solve_stream_t(tgvx_ctx* c, const tgvx_params* p, const tgvx_treetop_leaf* leaves, uint64_t nleaves,
                    tgvx_report* rep) {
    using T = typename Io<S>::T;
    Problem& P = c->prob;
    const uint32_t n = (uint32_t)P.n;
    tgvx_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = now_ms();
    // the leaf of every cube (first leaf wins) and the leaf ranges
    std::vector<int32_t> leaf_of(n, -1);
    std::vector<uint32_t> lf(std::max<uint64_t>(nleaves, 1), 1), ll(std::max<uint64_t>(nleaves, 1), 0);
    for (uint64_t l = 0; l < nleaves; ++l) {
        if (leaves[l].first_idx > leaves[l].last_idx || leaves[l].first_idx >= n) continue;
        lf[l] = leaves[l].first_idx;
        ll[l] = std::min<uint32_t>(leaves[l].last_idx, n - 1);
        for (uint32_t i = lf[l]; i <= ll[l]; ++i)
            if (leaf_of[i] < 0) leaf_of[i] = (int32_t)l;
    }
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        b.ensure(bytes);
        h2d(c, b.p, src, bytes);
    };
    DevBuf d_inv, d_leaf, d_lf, d_ll, d_code, d_depth;
    up(d_inv, P.inv.data(), 4ull * std::max<uint32_t>(n, 1));
    up(d_leaf, leaf_of.data(), 4ull * std::max<uint32_t>(n, 1));
    up(d_lf, lf.data(), 4 * lf.size());
    up(d_ll, ll.data(), 4 * ll.size());
    d_code.ensure(8ull * std::max<uint32_t>(n, 1));
    d_depth.ensure(4ull * std::max<uint32_t>(n, 1));
    for (int f = 0; f < 6; ++f) P.d_snb[f].ensure(4ull * std::max<uint32_t>(n, 1));
    P.d_sflag.ensure(std::max<uint32_t>(n, 1));
    P.d_spar.ensure(4ull * std::max<uint32_t>(n, 1));
    const unsigned grid_all = (n + 255) / 256;
    if (n) {
        k_stream_inputs<<<grid_all, 256, 0, c->stream>>>(P.d_code.as<uint64_t>(), P.d_depth.as<int32_t>(),
                                                         d_inv.as<uint32_t>(), n, d_code.as<uint64_t>(),
                                                         d_depth.as<int32_t>());
        k_stream_tables<<<grid_all, 256, 0, c->stream>>>(
            d_code.as<uint64_t>(), d_depth.as<int32_t>(), n, d_leaf.as<int32_t>(), d_lf.as<uint32_t>(),
            d_ll.as<uint32_t>(), d_inv.as<uint32_t>(), P.d_snb[0].as<int32_t>(), P.d_snb[1].as<int32_t>(),
            P.d_snb[2].as<int32_t>(), P.d_snb[3].as<int32_t>(), P.d_snb[4].as<int32_t>(), P.d_snb[5].as<int32_t>(),
            P.d_sflag.as<uint8_t>(), P.d_spar.as<int32_t>());
    }
    KArgs<S> a = make_args<S>(c, p);
    for (int ax = 0; ax < 3; ++ax) {
        a.nbp[ax] = P.d_snb[2 * ax].as<int32_t>();
        a.nbm[ax] = P.d_snb[2 * ax + 1].as<int32_t>();
    }
    a.sflag = P.d_sflag.as<uint8_t>();
    for (int i = 0; i < 4; ++i) a.fz[i] = P.d_save[i].as<S>();
    // tgv.cpp:821-845: u = median, everything else 0
    if (n) k_reset<S><<<grid_all, 256, 0, c->stream>>>(a, n, n);
    uint32_t max_level = 1;
    for (const Level& L : P.levels) max_level = std::max(max_level, L.n);
    DevBuf tmp;
    tmp.ensure(4 * sizeof(T) * (size_t)max_level);
    double it_ms = 0.0;
    for (size_t li = 0; li < P.levels.size(); ++li) {
        const Level& L = P.levels[li];
        if (li > 0) {
            const uint32_t* pb = P.perm.data() + L.off;
            for (uint64_t l = 0; l < nleaves; ++l) {
                if (lf[l] > ll[l]) continue;
                const uint32_t* lo = std::lower_bound(pb, pb + L.n, lf[l]);
                const uint32_t* hi = std::upper_bound(pb, pb + L.n, ll[l]);
                if (lo == hi) continue;
                KArgs<S> b = a;
                b.off = L.off + (uint32_t)(lo - pb);
                b.n = (uint32_t)(hi - lo);
                const unsigned gl = (b.n + 255) / 256;
                k_stream_parent_read<S><<<gl, 256, 0, c->stream>>>(b, P.d_spar.as<int32_t>(), tmp.as<T>());
                k_stream_parent_write<S><<<gl, 256, 0, c->stream>>>(b, tmp.as<T>());
            }
        }
        k_stream_snapshot<S><<<grid_all, 256, 0, c->stream>>>(a, n);
        a.off = a.lvl_off = L.off;
        a.n = a.lvl_n = L.n;
        const unsigned grid = (L.n + 255) / 256;
        TGVX_HIP(hipEventRecord(c->ev[0], c->stream));
        for (int it = 0; it < p->iters; ++it) {
            k_dual<S, 2, true><<<grid, 256, 0, c->stream>>>(a);
            k_primal_ref<S, true><<<grid, 256, 0, c->stream>>>(a);
        }
        TGVX_HIP(hipEventRecord(c->ev[1], c->stream));
        TGVX_HIP(hipEventSynchronize(c->ev[1]));
        float ms = 0.f;
        TGVX_HIP(hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
        it_ms += ms;
        if (li < (size_t)TGVX_MAX_LEVELS) {
            r.level_depth[li] = L.depth;
            r.level_leaves[li] = L.n;
            r.level_iters[li] = p->iters;
        }
        r.leaf_iters += (double)L.n * p->iters;
    }
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(c->stream));
    r.n_levels = (int)std::min<size_t>(P.levels.size(), TGVX_MAX_LEVELS);
    r.solve_ms = now_ms() - t0;
    r.iterate_ms = it_ms;
    if (rep) {
        const double setup = rep->setup_ms;
        *rep = r;
        rep->setup_ms = setup;
    }
}

}  // namespace

void problem_load(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int nb, int precision,
                  int op_mode, const uint64_t* blk_start, uint32_t nblocks) {
    if (nb <= 0 || nb > 32) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (n >= (1ull << 30)) throw Error(TGVX_E_ARG, "at most 2^30 - 1 cubes");
    if (precision < TGVX_FP64 || precision > TGVX_BF16_DUAL) throw Error(TGVX_E_ARG, "bad precision");
    if (n > 0 && (!cubes || !bins)) throw Error(TGVX_E_ARG, "null pointer");
    const double t0 = now_ms();
    g_graphs[c].clear();
    Problem& P = c->prob;
    P.loaded = false;
    P.halo.clear();
    P.halo2.clear();
    P.halo4.clear();
    P.rfz_use.clear();
    P.n = n;
    P.nb = nb;
    P.precision = precision;
    P.op_mode = op_mode;
    // bytes per element of the state arrays (mixed storage: FP32-sized
    // arrays, the dual fields using the first half as BF16)
    P.elem = precision == TGVX_FP64 ? 8 : (precision == TGVX_BF16 ? 2 : 4);
    // blocks: contiguous input ranges [bs[b], bs[b+1]) (one block by default)
    std::vector<uint64_t> bs;
    if (blk_start && nblocks > 0) {
        bs.assign(blk_start, blk_start + nblocks + 1);
        if (bs[0] != 0 || bs[nblocks] != n) throw Error(TGVX_E_ARG, "block ranges must cover the cubes");
        for (uint32_t b = 0; b < nblocks; ++b)
            if (bs[b + 1] < bs[b]) throw Error(TGVX_E_ARG, "block ranges must be ascending");
    } else {
        bs = {0, n};
        nblocks = 1;
    }
    P.nblocks = nblocks;
    P.block_start = bs;
    // level-major permutation: a stable counting sort by depth, in parallel
    // over contiguous chunks of the input (per chunk and depth a count, then
    // every chunk scatters from its own offsets, so each level keeps input
    // order), with the depths checked and the input's (code, depth) order
    // within each block tested in the same pass
    const int nt = (int)std::min<uint64_t>((uint64_t)host_threads(c), std::max<uint64_t>(1, n >> 16));
    const uint64_t per = (n + (uint64_t)nt - 1) / (uint64_t)nt;
    std::vector<std::array<uint64_t, 22>> ccnt(nt);
    std::vector<uint8_t> bad(nt, 0), unordered(nt, 0);
    parallel_for(c, nt, [&](int k) {
        std::array<uint64_t, 22>& cc = ccnt[k];
        cc.fill(0);
        const uint64_t lo = std::min(n, per * (uint64_t)k), hi = std::min(n, lo + per);
        if (lo >= hi) return;
        size_t b = (size_t)(std::upper_bound(bs.begin(), bs.end(), lo) - bs.begin()) - 1;  // lo's block
        bool ord = true;
        for (uint64_t i = lo; i < hi; ++i) {
            const int d = cubes[i].depth;
            if (d < 0 || d > 21) {
                bad[k] = 1;
                return;
            }
            cc[d]++;
            while (i >= bs[b + 1]) ++b;
            if (ord && i > bs[b]) {
                const uint64_t x = ((uint64_t)cubes[i - 1].hi << 32) | cubes[i - 1].lo,
                               y = ((uint64_t)cubes[i].hi << 32) | cubes[i].lo;
                ord = !(y < x || (y == x && d < cubes[i - 1].depth));
            }
        }
        unordered[k] = ord ? 0 : 1;
    });
    for (int k = 0; k < nt; ++k)
        if (bad[k]) throw Error(TGVX_E_ARG, "cube depth outside 0..21");
    bool in_order = true;
    for (int k = 0; k < nt; ++k) in_order = in_order && !unordered[k];
    P.levels.clear();
    std::vector<std::array<uint64_t, 22>> coff(nt);
    uint64_t acc = 0;
    for (int d = 0; d < 22; ++d) {
        const uint64_t first = acc;
        for (int k = 0; k < nt; ++k) {
            coff[k][d] = acc;
            acc += ccnt[k][d];
        }
        if (acc > first) P.levels.push_back({d, (uint32_t)first, (uint32_t)(acc - first)});
    }
    P.perm.resize(n);
    P.inv.resize(n);
    parallel_for(c, nt, [&](int k) {
        std::array<uint64_t, 22> o = coff[k];
        const uint64_t lo = std::min(n, per * (uint64_t)k), hi = std::min(n, lo + per);
        for (uint64_t i = lo; i < hi; ++i) {
            const uint64_t g = o[cubes[i].depth]++;
            P.perm[g] = (uint32_t)i;
            P.inv[i] = (uint32_t)g;
        }
    });
    // per-level, per-block segments of the positions (a level's positions are
    // in input order, hence grouped by block); the block of each position
    // when there are several
    P.seg.assign(P.levels.size(), std::vector<uint32_t>(nblocks + 1, 0));
    std::vector<int32_t> pblk(nblocks > 1 ? n : 0);
    for (size_t li = 0; li < P.levels.size(); ++li) {
        const Level& L = P.levels[li];
        std::vector<uint32_t>& sg = P.seg[li];
        if (nblocks == 1) {
            sg[1] = L.n;
            continue;
        }
        uint32_t b = 0;
        for (uint32_t t = 0; t < L.n; ++t) {
            const uint64_t i = P.perm[L.off + t];
            while (i >= bs[b + 1]) ++b;
            pblk[L.off + t] = (int32_t)b;
            sg[b + 1]++;
        }
        for (uint32_t k = 0; k < nblocks; ++k) sg[k + 1] += sg[k];
    }
    // The code maps the lookups search: per block, codes ascending, the last
    // input cube of each code (code-only map, the reference's last writer)
    // and of each (code, depth) pair (depth-aware map). Input in (code,
    // depth) order within each block -- the octree builder's output -- is
    // searched as it is (the lookups take the last of equal entries); other
    // input is sorted and deduplicated here.
    std::vector<uint64_t> ucode, dcode;
    std::vector<int32_t> uidx, ddepth, dpos;
    std::vector<uint32_t> uoff(nblocks + 1, 0), doff(nblocks + 1, 0);
    if (in_order) {
        for (uint32_t b = 0; b <= nblocks; ++b) uoff[b] = doff[b] = (uint32_t)bs[b];
    } else {
        for (uint32_t b = 0; b < nblocks; ++b) {
            std::vector<std::pair<uint64_t, uint32_t>> kv;
            kv.reserve(bs[b + 1] - bs[b]);
            for (uint64_t i = bs[b]; i < bs[b + 1]; ++i)
                kv.push_back({((uint64_t)cubes[i].hi << 32) | cubes[i].lo, (uint32_t)i});
            std::stable_sort(kv.begin(), kv.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t i = 0; i < kv.size(); ++i) {
                if (i + 1 < kv.size() && kv[i + 1].first == kv[i].first) continue;
                ucode.push_back(kv[i].first);
                uidx.push_back((int32_t)P.inv[kv[i].second]);
            }
            uoff[b + 1] = (uint32_t)ucode.size();
            std::stable_sort(kv.begin(), kv.end(), [&](const auto& x, const auto& y) {
                return x.first != y.first ? x.first < y.first : cubes[x.second].depth < cubes[y.second].depth;
            });
            for (size_t i = 0; i < kv.size(); ++i) {
                if (i + 1 < kv.size() && kv[i + 1].first == kv[i].first &&
                    cubes[kv[i + 1].second].depth == cubes[kv[i].second].depth)
                    continue;
                dcode.push_back(kv[i].first);
                ddepth.push_back(cubes[kv[i].second].depth);
                dpos.push_back((int32_t)P.inv[kv[i].second]);
            }
            doff[b + 1] = (uint32_t)dcode.size();
        }
    }
    auto up = [&](DevBuf& b, const void* src, size_t bytes) {
        b.ensure(bytes);
        h2d(c, b.p, src, bytes);
    };
    // per-position code, depth, histogram, vote total and median bin, on the
    // device from the inputs in input order and the permutation (the position
    // -> input map stays resident: the frozen-border mode's in-leaf test
    // reads it). The host buffers are read by the copies, which complete
    // before the synchronisation below.
    DevBuf& d_code = P.d_code;
    DevBuf& d_depth = P.d_depth;
    DevBuf d_uc, d_ui, d_uoff, d_doff, d_cubes, d_rbins, d_bad, d_icode, d_idepth, d_inv, d_dc, d_dd, d_dp;
    up(P.d_orig, P.perm.data(), 4 * std::max<uint64_t>(n, 1));
    if (in_order) {
        d_icode.ensure(8 * std::max<uint64_t>(n, 1));
        d_idepth.ensure(4 * std::max<uint64_t>(n, 1));
        up(d_inv, P.inv.data(), 4 * std::max<uint64_t>(n, 1));
    }
    d_code.ensure(8 * std::max<uint64_t>(n, 1));
    d_depth.ensure(4 * std::max<uint64_t>(n, 1));
    P.d_info.ensure(4 * std::max<uint64_t>(n, 1));
    P.d_bins.ensure(4 * (size_t)std::max<uint64_t>(n, 1) * nb);
    if (n > 0) {
        up(d_cubes, cubes, sizeof(tgvx_cube) * n);
        up(d_rbins, bins, 4 * (size_t)n * nb);
        d_bad.ensure(sizeof(int32_t));
        k_fill_i32<<<1, 64, 0, c->stream>>>(d_bad.as<int32_t>(), 1u, 0);
        k_load_positions<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
            d_cubes.as<tgvx_cube>(), d_rbins.as<uint32_t>(), P.d_orig.as<uint32_t>(), (uint32_t)n, nb,
            d_code.as<uint64_t>(), d_depth.as<int32_t>(), P.d_bins.as<uint32_t>(), P.d_info.as<uint32_t>(),
            d_bad.as<int32_t>(), in_order ? d_icode.as<uint64_t>() : nullptr,
            in_order ? d_idepth.as<int32_t>() : nullptr);
        TGVX_HIP(hipGetLastError());
        int32_t bad = 0;
        TGVX_HIP(hipMemcpyAsync(&bad, d_bad.p, sizeof(bad), hipMemcpyDeviceToHost, c->stream));
        TGVX_HIP(hipStreamSynchronize(c->stream));
        if (bad) throw Error(TGVX_E_ARG, "vote total of a cube exceeds 2^23 - 1");
    }
    if (!in_order) {
        up(d_uc, ucode.data(), 8 * std::max<size_t>(ucode.size(), 1));
        up(d_ui, uidx.data(), 4 * std::max<size_t>(uidx.size(), 1));
        up(d_dc, dcode.data(), 8 * std::max<size_t>(dcode.size(), 1));
        up(d_dd, ddepth.data(), 4 * std::max<size_t>(ddepth.size(), 1));
        up(d_dp, dpos.data(), 4 * std::max<size_t>(dpos.size(), 1));
    }
    const uint64_t* const map_code = in_order ? d_icode.as<uint64_t>() : d_uc.as<uint64_t>();
    const int32_t* const map_pos = in_order ? d_inv.as<int32_t>() : d_ui.as<int32_t>();
    const uint32_t map_n = in_order ? (uint32_t)n : (uint32_t)ucode.size();
    const uint64_t* const pair_code = in_order ? d_icode.as<uint64_t>() : d_dc.as<uint64_t>();
    const int32_t* const pair_depth = in_order ? d_idepth.as<int32_t>() : d_dd.as<int32_t>();
    const int32_t* const pair_pos = in_order ? d_inv.as<int32_t>() : d_dp.as<int32_t>();
    const uint32_t pair_n = in_order ? (uint32_t)n : (uint32_t)dcode.size();
    const bool multi = nblocks > 1;
    if (multi) {
        up(P.d_blk, pblk.data(), 4 * n);
        up(d_uoff, uoff.data(), 4 * uoff.size());
        up(d_doff, doff.data(), 4 * doff.size());
        P.d_bact.ensure(nblocks);
    }
    const int32_t* kblk = multi ? P.d_blk.as<int32_t>() : nullptr;
    for (int f = 0; f < 6; ++f) P.d_nb[f].ensure(4 * std::max<uint64_t>(n, 1));
    P.d_parent.ensure(4 * std::max<uint64_t>(n, 1));
    if (n > 0) {
        k_tables<<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
            d_code.as<uint64_t>(), d_depth.as<int32_t>(), (uint32_t)n, map_code, map_pos, map_n, kblk,
            d_uoff.as<uint32_t>(), P.d_nb[0].as<int32_t>(), P.d_nb[1].as<int32_t>(), P.d_nb[2].as<int32_t>(),
            P.d_nb[3].as<int32_t>(), P.d_nb[4].as<int32_t>(), P.d_nb[5].as<int32_t>(), P.d_parent.as<int32_t>());
        TGVX_HIP(hipGetLastError());
    }
    for (int ax = 0; ax < 3; ++ax) {
        P.d_nba[ax].ensure(4 * std::max<uint64_t>(n, 1));
        P.d_nbam[ax].ensure(4 * std::max<uint64_t>(n, 1));
        P.d_nbt[ax].ensure(4 * std::max<uint64_t>(n, 1));
    }
    P.d_parent_a.ensure(4 * std::max<uint64_t>(n, 1));
    if (n > 0) {
        const unsigned grid = (unsigned)((n + 255) / 256);
        k_tables_geo<<<grid, 256, 0, c->stream>>>(d_code.as<uint64_t>(), d_depth.as<int32_t>(), (uint32_t)n,
                                                  pair_code, pair_depth, pair_pos, pair_n, kblk, d_doff.as<uint32_t>(),
                                                  P.d_nba[0].as<int32_t>(),
                                                  P.d_nba[1].as<int32_t>(), P.d_nba[2].as<int32_t>(),
                                                  P.d_nbam[0].as<int32_t>(), P.d_nbam[1].as<int32_t>(),
                                                  P.d_nbam[2].as<int32_t>(),
                                                  P.d_parent_a.as<int32_t>(), P.d_info.as<uint32_t>());
        for (int ax = 0; ax < 3; ++ax) k_fill_i32<<<grid, 256, 0, c->stream>>>(P.d_nbt[ax].as<int32_t>(), (uint32_t)n, -1);
        k_transpose<<<grid, 256, 0, c->stream>>>(P.d_nba[0].as<int32_t>(), P.d_nba[1].as<int32_t>(),
                                                 P.d_nba[2].as<int32_t>(), d_depth.as<int32_t>(), (uint32_t)n,
                                                 P.d_nbt[0].as<int32_t>(), P.d_nbt[1].as<int32_t>(),
                                                 P.d_nbt[2].as<int32_t>());
        TGVX_HIP(hipGetLastError());
    }
    for (int i = 0; i < 17; ++i) P.d_state[i].ensure(P.elem * std::max<uint64_t>(n, 1));
    // the frozen-bar records hold one level; sized once per problem, as
    // captured graphs keep the pointer
    uint32_t max_level = 1;
    for (const Level& L : P.levels) max_level = std::max(max_level, L.n);
    P.d_fzaos.ensure(4 * P.elem * (size_t)max_level);
    for (int i = 0; i < 4; ++i) P.d_save[i].ensure(P.elem * std::max<uint64_t>(n, 1));
    TGVX_HIP(hipStreamSynchronize(c->stream));
    P.loaded = true;
    (void)t0;
}

void problem_solve(tgvx_ctx* c, const tgvx_params* p, const float* init_uv, tgvx_report* rep,
                   const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_cube* cubes, int init_stride) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (!p) throw Error(TGVX_E_ARG, "null params");
    if (p->op_mode != P.op_mode && p->op_mode != TGVX_OP_REF && p->op_mode != TGVX_OP_ADJ)
        throw Error(TGVX_E_ARG, "bad op_mode");
    if (leaves && init_uv) throw Error(TGVX_E_ARG, "the frozen-border solve takes no initial field");
    (void)cubes;
    if (init_stride != 4 && init_stride != 13) throw Error(TGVX_E_ARG, "bad initial field layout");
    if (init_stride == 13 && p->warm_start) throw Error(TGVX_E_ARG, "warm_start excludes an initial state");
    switch (P.precision) {
        case TGVX_FP64: solve_t<double>(c, p, init_uv, rep, leaves, nleaves, init_stride); break;
        case TGVX_FP32: solve_t<float>(c, p, init_uv, rep, leaves, nleaves, init_stride); break;
        case TGVX_BF16_DUAL: solve_t<mixs>(c, p, init_uv, rep, leaves, nleaves, init_stride); break;
        default: solve_t<bf16s>(c, p, init_uv, rep, leaves, nleaves, init_stride); break;
    }
}

void problem_solve_levels(tgvx_ctx* c, const tgvx_params* p, const int32_t* level_iters, int32_t n_levels,
                          tgvx_report* rep) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (!p || !level_iters) throw Error(TGVX_E_ARG, "null pointer");
    if (n_levels != (int32_t)P.levels.size())
        throw Error(TGVX_E_ARG, "level_iters needs one count per level (" + std::to_string(P.levels.size()) + ")");
    if (p->gap_tol > 0.0) throw Error(TGVX_E_ARG, "per-level iteration counts exclude gap_tol");
    for (int32_t i = 0; i < n_levels; ++i)
        if (level_iters[i] < 0) throw Error(TGVX_E_ARG, "negative iteration count");
    P.level_iters.assign(level_iters, level_iters + n_levels);
    struct Clear {
        Problem& P;
        ~Clear() { P.level_iters.clear(); }
    } clear{P};
    problem_solve(c, p, nullptr, rep, nullptr, 0, nullptr);
}

void problem_solve_streaming(tgvx_ctx* c, const tgvx_params* p, const tgvx_treetop_leaf* leaves, uint64_t nleaves,
                             tgvx_report* rep) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (!p || (nleaves && !leaves)) throw Error(TGVX_E_ARG, "null pointer");
    if (p->op_mode != TGVX_OP_REF) throw Error(TGVX_E_ARG, "the streaming solve runs the reference rules (op_mode 0)");
    if (p->iters < 0) throw Error(TGVX_E_ARG, "negative iteration count");
    if (P.nblocks > 1) throw Error(TGVX_E_ARG, "a problem of several blocks has no streaming solve");
    switch (P.precision) {
        case TGVX_FP64: solve_stream_t<double>(c, p, leaves, nleaves, rep); break;
        case TGVX_FP32: solve_stream_t<float>(c, p, leaves, nleaves, rep); break;
        case TGVX_BF16_DUAL: solve_stream_t<mixs>(c, p, leaves, nleaves, rep); break;
        default: solve_stream_t<bf16s>(c, p, leaves, nleaves, rep); break;
    }
}

namespace {
// Pinned host memory owned by one solve.
struct PinnedBuf {
    void* p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void alloc(size_t b) {
        TGVX_HIP(hipHostMalloc(&p, std::max<size_t>(b, 16), hipHostMallocDefault));
        bytes = b;
    }
    template <class X>
    X* as() const { return static_cast<X*>(p); }
};

// One of the two alternating device slots of the out-of-core solve: SoA
// arrays of capacity C cubes in one allocation (state[13], bars[4],
// snapshot[4] as float; info, tables[6] as 32-bit words; flags), the pinned
// staging of the same layout and the pinned copy of the state block back.
struct OocSlot {
    size_t cap = 0;
    DevBuf dev;
    PinnedBuf up, down;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<uint32_t> act;  // global indices of the batch's iterated cubes
    bool busy = false;
    size_t w = 0;  // the current batch's array stride: its local cubes
    static size_t words(size_t cap) { return (13 + 4 + 4 + 1 + 6) * cap + (cap + 3) / 4; }
    float* st(float* b) const { return b; }
    float* bar(float* b) const { return b + 13 * w; }
    float* fz(float* b) const { return b + 17 * w; }
    uint32_t* info(float* b) const { return reinterpret_cast<uint32_t*>(b + 21 * w); }
    int32_t* nb(float* b) const { return reinterpret_cast<int32_t*>(b + 22 * w); }
    uint8_t* flag(float* b) const { return reinterpret_cast<uint8_t*>(b + 28 * w); }
    ~OocSlot() {
        // the slot's work has ended before its device block goes back to the
        // context's cache (a call that threw may leave some in flight)
        if (stream) (void)hipStreamSynchronize(stream);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
    }
};
}  // namespace

// tgv_minimize_streaming (tgv.cpp:802-1070) out of core: the state of every
// cube lives in pinned host memory (the reference keeps it in its state file)
// and at most `budget` bytes of per-cube data are on the device at a time.
// Per level, on the host: the parent copy leaf after leaf (bisection parent,
// tgv.cpp:872-934) and the level-start snapshot of (u, v) (build_bars_frozen).
// Then the treetop leaves holding cubes of the level run in batches: a
// batch's iterated cubes, the cubes they read (their leaf's other cubes, other
// leaves' and border cubes: snapshot values) and its tables go to a device
// slot, the level's iterations run there with the resident solve's kernels
// (streaming mode), and the iterated cubes' state comes back. The two slots
// alternate on two streams, so one batch's transfers overlap the other's
// iterations; within a level the leaves are independent (borders read the
// snapshot, border p, q are zero), so the result is the resident solve's, bit
// for bit. The sorted codes, depths and leaf of every cube (16 B per cube)
// stay on the device for the neighbour lookups (CubeFileReader::lookup's
// bisection, tgv.cpp:673-683), like the reference's open cube file. The
// host's passes over the cubes run on the context's host threads; each gives
// the single-thread result (the same batches and the same slot layout).
void solve_stream_ooc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n64, const uint32_t* bins, int nbins,
                      const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* p, uint64_t budget,
                      float* state_out, tgvx_report* rep, tgvx_ooc_report* orep) {
    if (!p || !state_out || (n64 && (!cubes || !bins)) || (nleaves && !leaves)) throw Error(TGVX_E_ARG, "null pointer");
    if (nbins <= 0 || nbins > 32) throw Error(TGVX_E_ARG, "nb must be in 1..32");
    if (n64 >= (1ull << 30)) throw Error(TGVX_E_ARG, "at most 2^30 - 1 cubes");
    if (p->op_mode != TGVX_OP_REF) throw Error(TGVX_E_ARG, "the streaming solve runs the reference rules (op_mode 0)");
    if (p->iters < 0) throw Error(TGVX_E_ARG, "negative iteration count");
    const uint32_t n = (uint32_t)n64;
    tgvx_report r;
    std::memset(&r, 0, sizeof(r));
    const double t0 = now_ms();
    // ---- inputs: codes (sorted), depths, per-cube constants, leaves
    std::vector<uint64_t> code(n);
    std::vector<int32_t> depth(n);
    std::vector<uint32_t> info(n);
    // host passes over the cubes run on the context's host threads in ranges
    // of at least kGrain cubes; each writes only its own cubes' entries
    constexpr size_t kGrain = 1u << 16;
    int min_d = 21, max_d = 0;
    bool has_depth[22] = {};
    // the first bad cube in input order and what is wrong with it
    uint64_t bad_at = ~0ull;
    int bad_why = 0;
    std::mutex mu;
    auto cube_code = [&](uint32_t i) { return ((uint64_t)cubes[i].hi << 32) | cubes[i].lo; };
    parallel_range(c, n, kGrain, [&](size_t lo, size_t hi) {
        int dmin = 21, dmax = 0, why = 0;
        bool seen[22] = {};
        size_t i = lo;
        for (; i < hi; ++i) {
            code[i] = cube_code((uint32_t)i);
            depth[i] = cubes[i].depth;
            if (depth[i] < 0 || depth[i] > 21) {
                why = 1;
                break;
            }
            if (i && code[i] < cube_code((uint32_t)i - 1)) {
                why = 2;
                break;
            }
            dmin = std::min(dmin, depth[i]);
            dmax = std::max(dmax, depth[i]);
            seen[depth[i]] = true;
            const uint32_t* b = bins + i * nbins;
            uint64_t tot = 0;
            for (int k = 0; k < nbins; ++k) tot += b[k];
            if (tot > InfoWord::kWMax) {
                why = 3;
                break;
            }
            uint32_t kmed = 0, cum = 0;
            for (int k = 0; k < nbins; ++k) {
                cum += b[k];
                if (cum * 2u >= (uint32_t)tot) {
                    kmed = (uint32_t)k;
                    break;
                }
            }
            info[i] = (uint32_t)tot | (kmed << 24);
        }
        std::lock_guard<std::mutex> lk(mu);
        if (why && i < bad_at) {
            bad_at = i;
            bad_why = why;
        }
        min_d = std::min(min_d, dmin);
        max_d = std::max(max_d, dmax);
        for (int d = 0; d < 22; ++d) has_depth[d] = has_depth[d] || seen[d];
    });
    if (bad_why == 1) throw Error(TGVX_E_ARG, "cube depth outside 0..21");
    if (bad_why == 2) throw Error(TGVX_E_ARG, "cubes must be sorted by Morton code");
    if (bad_why == 3) throw Error(TGVX_E_ARG, "vote total of a cube exceeds 2^23 - 1");
    std::vector<int32_t> leaf_of(n, -1);
    std::vector<uint32_t> lf(std::max<uint64_t>(nleaves, 1), 1), ll(std::max<uint64_t>(nleaves, 1), 0);
    for (uint64_t l = 0; l < nleaves; ++l) {
        if (leaves[l].first_idx > leaves[l].last_idx || leaves[l].first_idx >= n) continue;
        lf[l] = leaves[l].first_idx;
        ll[l] = std::min<uint32_t>(leaves[l].last_idx, n - 1);
        // a cube in several leaves' ranges belongs to the first
        const uint32_t f0 = lf[l];
        parallel_range(c, (size_t)ll[l] - f0 + 1, kGrain, [&](size_t lo, size_t hi) {
            for (size_t i = f0 + lo; i < f0 + hi; ++i)
                if (leaf_of[i] < 0) leaf_of[i] = (int32_t)l;
        });
    }
    // state: u = median (tgv.cpp:821-845), everything else 0
    PinnedBuf st, snap;
    st.alloc(sizeof(float) * 13 * (size_t)n);
    snap.alloc(sizeof(float) * 4 * (size_t)n);
    float* S = st.as<float>();
    float* Z = snap.as<float>();
    parallel_range(c, n, kGrain, [&](size_t lo, size_t hi) {
        std::memset(S + 13 * lo, 0, sizeof(float) * 13 * (hi - lo));
        for (size_t i = lo; i < hi; ++i) S[13 * i] = median_of<float>(info[i], nbins);
    });
    if (n == 0) {
        if (rep) *rep = r;
        return;
    }
    // ---- per-cube tables on the device, in chunks: neighbours as each
    // cube's leaf sees them (global indices), flags, bisection parent
    std::vector<int32_t> nbg(6 * (size_t)n), par(n);
    std::vector<uint8_t> sflag(n);
    {
        DevBuf d_code, d_depth, d_leaf, d_lf, d_ll, d_nb, d_fl, d_par;
        d_code.ensure(8ull * n);
        d_depth.ensure(4ull * n);
        d_leaf.ensure(4ull * n);
        d_lf.ensure(4 * lf.size());
        d_ll.ensure(4 * ll.size());
        h2d(c, d_code.p, code.data(), 8ull * n);
        h2d(c, d_depth.p, depth.data(), 4ull * n);
        h2d(c, d_leaf.p, leaf_of.data(), 4ull * n);
        TGVX_HIP(hipMemcpyAsync(d_lf.p, lf.data(), 4 * lf.size(), hipMemcpyHostToDevice, c->stream));
        TGVX_HIP(hipMemcpyAsync(d_ll.p, ll.data(), 4 * ll.size(), hipMemcpyHostToDevice, c->stream));
        const uint32_t chunk = 1u << 20;
        d_nb.ensure(24ull * chunk);
        d_fl.ensure(chunk);
        d_par.ensure(4ull * chunk);
        for (uint32_t i0 = 0; i0 < n; i0 += chunk) {
            const uint32_t cn = std::min(chunk, n - i0);
            k_stream_tables_ooc<<<(cn + 255) / 256, 256, 0, c->stream>>>(
                d_code.as<uint64_t>(), d_depth.as<int32_t>(), n, d_leaf.as<int32_t>(), d_lf.as<uint32_t>(),
                d_ll.as<uint32_t>(), i0, cn, d_nb.as<int32_t>(), d_fl.as<uint8_t>(), d_par.as<int32_t>());
            TGVX_HIP(hipGetLastError());
            d2h(c, nbg.data() + 6 * (size_t)i0, d_nb.p, 24ull * cn);
            d2h(c, sflag.data() + i0, d_fl.p, cn);
            d2h(c, par.data() + i0, d_par.p, 4ull * cn);
        }
    }
    r.setup_ms = now_ms() - t0;
    // ---- device slots: two, each a capacity of C local cubes
    const size_t per_cube = 4 * (13 + 4 + 4 + 1 + 6) + 1;
    const size_t cap = budget / (2 * per_cube);
    if (cap < 64) throw Error(TGVX_E_ARG, "device budget below 2 x 64 cubes of slot data");
    OocSlot slot[2];
    for (OocSlot& sl : slot) {
        sl.cap = cap;
        sl.dev.ensure(4 * OocSlot::words(cap));
        sl.up.alloc(4 * OocSlot::words(cap));
        sl.down.alloc(sizeof(float) * 13 * cap);
        TGVX_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        TGVX_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    const float sg = p->sigma > 0.f ? p->sigma : 0.25f, ta = p->tau > 0.f ? p->tau : 0.25f;
    auto finish = [&](OocSlot& sl) {
        if (!sl.busy) return;
        TGVX_HIP(hipEventSynchronize(sl.done));
        const float* D = sl.down.as<float>();
        parallel_range(c, sl.act.size(), kGrain, [&](size_t lo, size_t hi) {
            for (size_t a = lo; a < hi; ++a)
                for (int f = 0; f < 13; ++f) S[13 * (size_t)sl.act[a] + f] = D[f * sl.w + a];
        });
        sl.busy = false;
    };
    std::vector<int32_t> loc(n, -1);
    std::vector<uint32_t> mark(n, 0u);
    uint32_t stamp = 0;
    double it_ms = 0.0;
    int li = 0, nb_batches = 0;
    uint64_t moved = 0;
    for (int L = min_d; L <= max_d; ++L) {
        // the leaves' cubes of depth L, leaf after leaf
        if (!has_depth[L]) continue;
        std::vector<std::vector<uint32_t>> by_leaf(std::max<uint64_t>(nleaves, 1));
        const int nlt = (int)std::min<uint64_t>((uint64_t)host_threads(c), nleaves);
        parallel_for(c, nlt, [&](int k) {
            for (uint64_t l = (uint64_t)k; l < nleaves; l += (uint64_t)nlt) {
                if (lf[l] > ll[l]) continue;
                for (uint32_t i = lf[l]; i <= ll[l]; ++i)
                    if (depth[i] == L && leaf_of[i] == (int32_t)l) by_leaf[l].push_back(i);
            }
        });
        uint64_t nL = 0;
        for (uint64_t l = 0; l < nleaves; ++l) nL += by_leaf[l].size();
        // parent copy, leaf after leaf: each leaf reads all its parents, then
        // writes its cubes
        if (L > min_d) {
            std::vector<float> tmp;
            for (uint64_t l = 0; l < nleaves; ++l) {
                const std::vector<uint32_t>& v = by_leaf[l];
                tmp.assign(4 * v.size(), 0.f);
                parallel_range(c, v.size(), kGrain, [&](size_t lo, size_t hi) {
                    for (size_t k = lo; k < hi; ++k)
                        if (par[v[k]] >= 0) std::memcpy(&tmp[4 * k], S + 13 * (size_t)par[v[k]], 16);
                });
                parallel_range(c, v.size(), kGrain, [&](size_t lo, size_t hi) {
                    for (size_t k = lo; k < hi; ++k) {
                        float* s = S + 13 * (size_t)v[k];
                        if (par[v[k]] >= 0) std::memcpy(s, &tmp[4 * k], 16);
                        std::memset(s + 4, 0, 9 * sizeof(float));
                    }
                });
            }
        }
        // level-start snapshot of every cube's (u, v)
        parallel_range(c, n, kGrain, [&](size_t lo, size_t hi) {
            for (size_t i = lo; i < hi; ++i) std::memcpy(Z + 4 * i, S + 13 * i, 16);
        });
        const double tl0 = now_ms();
        // batches of whole leaves; a leaf's cubes read at most 7 local slots each
        uint64_t l = 0, used_slot = 0;
        while (l < nleaves) {
            if (by_leaf[l].empty()) {
                ++l;
                continue;
            }
            OocSlot& sl = slot[used_slot & 1];
            finish(sl);
            float* U = sl.up.as<float>();
            const size_t C = sl.cap;
            sl.act.clear();
            std::vector<uint32_t> others;
            // the batch: whole leaves while the distinct cubes they iterate
            // and read fit the slot, counted with a stamp per batch. A
            // leaf's cubes are stamped on the host threads: the thread whose
            // exchange first sets a cube's stamp counts it, so the count is
            // the leaf's new distinct cubes in any thread order
            ++stamp;
            uint64_t l1 = l;
            size_t count = 0;
            std::vector<std::vector<uint32_t>> fresh(host_threads(c));
            while (l1 < nleaves) {
                const std::vector<uint32_t>& v = by_leaf[l1];
                for (auto& fr : fresh) fr.clear();
                const int parts = parallel_split(c, v.size(), kGrain, [&](int k, size_t lo, size_t hi) {
                    std::vector<uint32_t>& fr = fresh[k];
                    auto see = [&](int32_t j) {
                        if (j >= 0 && __atomic_load_n(&mark[j], __ATOMIC_RELAXED) != stamp &&
                            __atomic_exchange_n(&mark[j], stamp, __ATOMIC_RELAXED) != stamp)
                            fr.push_back((uint32_t)j);
                    };
                    for (size_t x = lo; x < hi; ++x) {
                        const uint32_t i = v[x];
                        see((int32_t)i);
                        for (int f = 0; f < 6; ++f) see(nbg[6 * (size_t)i + f]);
                    }
                });
                size_t nf = 0;
                for (int k = 0; k < parts; ++k) nf += fresh[k].size();
                if (count + nf > C) {
                    for (int k = 0; k < parts; ++k)
                        for (uint32_t j : fresh[k]) mark[j] = 0;
                    break;
                }
                count += nf;
                ++l1;
            }
            if (l1 == l) throw Error(TGVX_E_ARG, "device budget too small for treetop leaf " + std::to_string(l));
            for (uint64_t k = l; k < l1; ++k) sl.act.insert(sl.act.end(), by_leaf[k].begin(), by_leaf[k].end());
            const size_t na = sl.act.size();
            parallel_range(c, na, kGrain, [&](size_t lo, size_t hi) {
                for (size_t a = lo; a < hi; ++a) loc[sl.act[a]] = (int32_t)a;
            });
            // the other cubes the batch reads get local indices na, na + 1,
            // ... in the order of their first reading (iterated cube a, then
            // face f). On the host threads: a cube's claim -(2 + a) keeps
            // the smallest reader a (an atomic max), then the thread holding
            // that reader lists the cube at its first face and marks it
            // listed (INT32_MIN); the threads' lists in range order are that
            // first-reading order
            std::vector<std::vector<uint32_t>>& mine = fresh;
            for (auto& m : mine) m.clear();
            parallel_split(c, na, kGrain, [&](int, size_t lo, size_t hi) {
                for (size_t a = lo; a < hi; ++a) {
                    const int32_t want = -2 - (int32_t)a;
                    for (int f = 0; f < 6; ++f) {
                        const int32_t j = nbg[6 * (size_t)sl.act[a] + f];
                        if (j < 0) continue;
                        int32_t cur = __atomic_load_n(&loc[j], __ATOMIC_RELAXED);
                        while ((cur == -1 || cur < want) &&
                               !__atomic_compare_exchange_n(&loc[j], &cur, want, false, __ATOMIC_RELAXED,
                                                            __ATOMIC_RELAXED)) {
                        }
                    }
                }
            });
            const int parts = parallel_split(c, na, kGrain, [&](int k, size_t lo, size_t hi) {
                std::vector<uint32_t>& m = mine[k];
                for (size_t a = lo; a < hi; ++a) {
                    const int32_t want = -2 - (int32_t)a;
                    for (int f = 0; f < 6; ++f) {
                        const int32_t j = nbg[6 * (size_t)sl.act[a] + f];
                        if (j >= 0 && __atomic_load_n(&loc[j], __ATOMIC_RELAXED) == want) {
                            __atomic_store_n(&loc[j], INT32_MIN, __ATOMIC_RELAXED);
                            m.push_back((uint32_t)j);
                        }
                    }
                }
            });
            for (int k = 0; k < parts; ++k) others.insert(others.end(), mine[k].begin(), mine[k].end());
            parallel_range(c, others.size(), kGrain, [&](size_t lo, size_t hi) {
                for (size_t k = lo; k < hi; ++k) loc[others[k]] = (int32_t)(na + k);
            });
            const size_t W = na + others.size();
            sl.w = W;
            int32_t* NB = sl.nb(U);
            auto put = [&](size_t a, uint32_t i) {
                for (int f = 0; f < 13; ++f) sl.st(U)[f * W + a] = S[13 * (size_t)i + f];
                for (int k = 0; k < 4; ++k) sl.fz(U)[k * W + a] = Z[4 * (size_t)i + k];
            };
            parallel_range(c, na, kGrain, [&](size_t lo, size_t hi) {
                for (size_t a = lo; a < hi; ++a) {
                    const uint32_t i = sl.act[a];
                    for (int f = 0; f < 6; ++f) {
                        const int32_t j = nbg[6 * (size_t)i + f];
                        NB[f * W + a] = j < 0 ? -1 : loc[j];
                    }
                    sl.info(U)[a] = info[i];
                    sl.flag(U)[a] = sflag[i];
                    put(a, i);
                    for (int k = 0; k < 4; ++k) sl.bar(U)[k * W + a] = Z[4 * (size_t)i + k];
                }
            });
            parallel_range(c, others.size(), kGrain, [&](size_t lo, size_t hi) {
                for (size_t k = lo; k < hi; ++k) put(na + k, others[k]);
            });
            parallel_range(c, na, kGrain, [&](size_t lo, size_t hi) {
                for (size_t a = lo; a < hi; ++a) loc[sl.act[a]] = -1;
            });
            parallel_range(c, others.size(), kGrain, [&](size_t lo, size_t hi) {
                for (size_t k = lo; k < hi; ++k) loc[others[k]] = -1;
            });
            // up, iterate, down
            if (W > C) throw Error(TGVX_E_STATE, "out-of-core batch exceeds its slot");
            TGVX_HIP(hipMemcpyAsync(sl.dev.p, U, 4 * OocSlot::words(W), hipMemcpyHostToDevice, sl.stream));
            float* dv = sl.dev.as<float>();
            KArgs<float> a;
            std::memset(&a, 0, sizeof(a));
            for (int f = 0; f < 13; ++f) a.s[f] = dv + f * W;
            for (int k = 0; k < 4; ++k) {
                a.s[F_UB + k] = dv + (13 + k) * W;
                a.fz[k] = dv + (17 + k) * W;
            }
            a.info = reinterpret_cast<const uint32_t*>(dv + 21 * W);
            for (int ax = 0; ax < 3; ++ax) {
                a.nbp[ax] = reinterpret_cast<const int32_t*>(dv + (22 + 2 * ax) * W);
                a.nbm[ax] = reinterpret_cast<const int32_t*>(dv + (23 + 2 * ax) * W);
            }
            a.sflag = reinterpret_cast<const uint8_t*>(dv + 28 * W);
            a.nbins = nbins;
            a.off = 0;
            a.n = (uint32_t)na;
            a.lvl_off = 0;
            a.lvl_n = (uint32_t)na;
            a.sigma = sg;
            a.tau = ta;
            a.a0 = p->alpha0;
            a.a1 = p->alpha1;
            a.tl = ta * p->lambda;
            a.init_stride = 4;
            const unsigned grid = (unsigned)((na + 255) / 256);
            for (int it = 0; it < p->iters; ++it) {
                k_dual<float, 2, true><<<grid, 256, 0, sl.stream>>>(a);
                k_primal_ref<float, true><<<grid, 256, 0, sl.stream>>>(a);
            }
            TGVX_HIP(hipGetLastError());
            TGVX_HIP(hipMemcpyAsync(sl.down.p, dv, sizeof(float) * 13 * W, hipMemcpyDeviceToHost, sl.stream));
            TGVX_HIP(hipEventRecord(sl.done, sl.stream));
            sl.busy = true;
            moved += 4 * OocSlot::words(W) + sizeof(float) * 13 * W;
            ++used_slot;
            ++nb_batches;
            l = l1;
        }
        finish(slot[0]);
        finish(slot[1]);
        it_ms += now_ms() - tl0;
        if (li < TGVX_MAX_LEVELS) {
            r.level_depth[li] = L;
            r.level_leaves[li] = (uint32_t)nL;
            r.level_iters[li] = p->iters;
        }
        r.leaf_iters += (double)nL * p->iters;
        ++li;
    }
    parallel_range(c, n, kGrain, [&](size_t lo, size_t hi) {
        std::memcpy(state_out + 13 * lo, S + 13 * lo, sizeof(float) * 13 * (hi - lo));
    });
    r.n_levels = std::min(li, TGVX_MAX_LEVELS);
    r.solve_ms = now_ms() - t0;
    r.iterate_ms = it_ms;
    if (rep) *rep = r;
    if (orep) {
        orep->batches = nb_batches;
        orep->slot_cubes = cap;
        orep->device_bytes = 2 * 4 * OocSlot::words(cap);
        orep->index_bytes = 16ull * n;
        orep->moved_bytes = moved;
        orep->tables_ms = r.setup_ms;
    }
}

void problem_state(tgvx_ctx* c, void* state_out, void* ubar_out) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    const size_t n = P.n;
    if (n == 0 || (!state_out && !ubar_out)) return;
    // records in input order on the device (one scattered write per cube),
    // then one copy each
    auto go = [&](auto tag) {
        using S = decltype(tag);
        using T = typename Io<S>::T;
        tgvx_params q;
        tgvx_default_params(&q);
        const KArgs<S> a = make_args<S>(c, &q);
        DevBuf st, bar;
        if (state_out) st.ensure(13 * sizeof(T) * n);
        if (ubar_out) bar.ensure(4 * sizeof(T) * n);
        k_state_out<S><<<(unsigned)((n + 255) / 256), 256, 0, c->stream>>>(
            a, P.d_orig.as<uint32_t>(), (uint32_t)n, state_out ? st.as<T>() : nullptr, ubar_out ? bar.as<T>() : nullptr);
        TGVX_HIP(hipGetLastError());
        if (state_out)
            d2h(c, state_out, st.p, 13 * sizeof(T) * n);
        if (ubar_out) d2h(c, ubar_out, bar.p, 4 * sizeof(T) * n);
        TGVX_HIP(hipStreamSynchronize(c->stream));
    };
    switch (P.precision) {
        case TGVX_FP64: go(double{}); break;
        case TGVX_FP32: go(float{}); break;
        case TGVX_BF16_DUAL: go(mixs{}); break;
        default: go(bf16s{}); break;
    }
}

void problem_gap(tgvx_ctx* c, int depth, const tgvx_params* p, double* Pv, double* Dv) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    const Level* L = nullptr;
    for (const Level& l : P.levels)
        if (l.depth == depth) L = &l;
    if (!L) throw Error(TGVX_E_ARG, "no cubes at depth " + std::to_string(depth));
    auto go = [&](auto tag) {
        using S = decltype(tag);
        KArgs<S> a = make_args<S>(c, p);
        a.off = a.lvl_off = L->off;
        a.n = a.lvl_n = L->n;
        eval_gap<S>(c, a, p, Pv, Dv);
    };
    switch (P.precision) {
        case TGVX_FP64: go(double{}); break;
        case TGVX_FP32: go(float{}); break;
        case TGVX_BF16_DUAL: go(mixs{}); break;
        default: go(bf16s{}); break;
    }
}

void autotune(tgvx_ctx* c, const tgvx_params* p, int iters, tgvx_tune_report* rep) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (!p) throw Error(TGVX_E_ARG, "null params");
    if (iters < 2 || iters > 10000) throw Error(TGVX_E_ARG, "iters must be in 2..10000");
    switch (P.precision) {
        case TGVX_FP64: autotune_t<double>(c, p, iters, rep); break;
        case TGVX_FP32: autotune_t<float>(c, p, iters, rep); break;
        case TGVX_BF16_DUAL: autotune_t<mixs>(c, p, iters, rep); break;
        default: autotune_t<bf16s>(c, p, iters, rep); break;
    }
}

void grid_barrier_bench(tgvx_ctx* c, int wgs, int block, int nbar, double* us) {
    grid_barrier_bench_run(c, wgs, block, nbar, us);
    check_grid(c);
}

void level_bench(tgvx_ctx* c, int depth, const tgvx_params* p, int iters, double* total_ms, double* dual_ms,
                 double* primal_ms, double* fused_ms) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    switch (P.precision) {
        case TGVX_FP64: bench_t<double>(c, depth, p, iters, total_ms, dual_ms, primal_ms, fused_ms); break;
        case TGVX_FP32: bench_t<float>(c, depth, p, iters, total_ms, dual_ms, primal_ms, fused_ms); break;
        case TGVX_BF16_DUAL: bench_t<mixs>(c, depth, p, iters, total_ms, dual_ms, primal_ms, fused_ms); break;
        default: bench_t<bf16s>(c, depth, p, iters, total_ms, dual_ms, primal_ms, fused_ms); break;
    }
}

template <class S>
void tile_sets_t(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_sets, int64_t* lds_bytes) {
    Problem& P = c->prob;
    const Level* L = nullptr;
    for (const Level& l : P.levels)
        if (l.depth == depth) L = &l;
    if (!L) throw Error(TGVX_E_ARG, "no cubes at depth " + std::to_string(depth));
    if (tile != 128 && tile != 256 && tile != 512 && tile != 1024)
        throw Error(TGVX_E_ARG, "the two-iteration kernel takes tiles of 128, 256, 512 or 1024");
    tgvx_params q;
    tgvx_default_params(&q);
    q.op_mode = TGVX_OP_ADJ;
    const KArgs<S> a = level_args<S>(c, &q, *L);
    const Halo2Bufs& hb = halo2_tab<S>(c, a, tile);
    *usable = hb.ok ? 1 : 0;
    max_sets[0] = hb.sa;
    max_sets[1] = hb.sd;
    max_sets[2] = hb.sp;
    *lds_bytes = (int64_t)hb.lds;
}

void tile_sets(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_sets, int64_t* lds_bytes) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (P.op_mode != TGVX_OP_ADJ) throw Error(TGVX_E_STATE, "the two-iteration kernel needs the exact-adjoint tables");
    switch (P.precision) {
        case TGVX_FP64: tile_sets_t<double>(c, depth, tile, usable, max_sets, lds_bytes); break;
        case TGVX_FP32: tile_sets_t<float>(c, depth, tile, usable, max_sets, lds_bytes); break;
        case TGVX_BF16_DUAL: tile_sets_t<mixs>(c, depth, tile, usable, max_sets, lds_bytes); break;
        default: tile_sets_t<bf16s>(c, depth, tile, usable, max_sets, lds_bytes); break;
    }
}

template <class S>
void tile_sets4_t(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_slots, int64_t* lds_bytes) {
    Problem& P = c->prob;
    const Level* L = nullptr;
    for (const Level& l : P.levels)
        if (l.depth == depth) L = &l;
    if (!L) throw Error(TGVX_E_ARG, "no cubes at depth " + std::to_string(depth));
    if (tile != 128 && tile != 256) throw Error(TGVX_E_ARG, "the four-iteration kernel takes tiles of 128 or 256");
    tgvx_params q;
    tgvx_default_params(&q);
    q.op_mode = TGVX_OP_ADJ;
    const KArgs<S> a = level_args<S>(c, &q, *L);
    const Halo4Bufs& hb = halo4_tab<S>(c, a, tile);
    *usable = hb.ok ? 1 : 0;
    *max_slots = hb.sa;
    *lds_bytes = (int64_t)hb.lds;
}

void tile_sets4(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_slots, int64_t* lds_bytes) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (P.op_mode != TGVX_OP_ADJ) throw Error(TGVX_E_STATE, "the four-iteration kernel needs the exact-adjoint tables");
    switch (P.precision) {
        case TGVX_FP64: tile_sets4_t<double>(c, depth, tile, usable, max_slots, lds_bytes); break;
        case TGVX_FP32: tile_sets4_t<float>(c, depth, tile, usable, max_slots, lds_bytes); break;
        case TGVX_BF16_DUAL: tile_sets4_t<mixs>(c, depth, tile, usable, max_slots, lds_bytes); break;
        default: tile_sets4_t<bf16s>(c, depth, tile, usable, max_slots, lds_bytes); break;
    }
}

void release_graphs(const tgvx_ctx* c) { g_graphs.erase(c); }

}  // namespace tgvx
