// Iso-surface extraction on the device over the resident problem's cubes
// (reference src/marching_cubes.cpp:16-181, extract_surface).
//
// For every face-adjacent pair of cubes (the reference's neighbour rule: the
// same-depth face neighbour by code, else the cube at its parent code; the
// problem's reference-mode table holds exactly this) whose u changes sign, one
// quad is emitted at the interpolated zero crossing between the cube centres,
// spanning the smaller cube's face. A pair is emitted once, by its first
// visitor in (input index, face) order; vertices are merged on the 20-bit per
// axis quantised key of their position, each merged vertex keeping the
// position of its first occurrence and vertices numbered in order of first
// occurrence, so the mesh is the reference's, vertex for vertex and triangle
// for triangle.
//
// Device steps: per input cube a face mask and count -> exclusive scan -> the
// quads' corners and keys -> radix sort of (key, occurrence) -> run heads ->
// first occurrence per key -> sort of first occurrences (numbering) ->
// vertex ids per occurrence -> triangles.
#include <hip/hip_runtime.h>
#include <string.h>

#include <cstring>
#include <vector>

#include <rocprim/rocprim.hpp>

#include "tgvx_device.h"
#include "tgvx_internal.h"

namespace tgvx {
namespace {

struct SurfArgs {
    const uint64_t* code;  // per position
    const int32_t* depth;
    const int32_t* nb[6];   // reference-mode face table, positions
    const uint32_t* orig;   // position -> input index
    const uint32_t* inv;    // input index -> position
    const float* u;         // per position
    const uint8_t* has;     // per position, or null
    uint32_t n;
    float o[3], r_root, inv_q;
};

__device__ inline void corner(uint64_t code, int depth, const SurfArgs& a, float out[3]) {
    const uint32_t f[3] = {compact3(code), compact3(code >> 1), compact3(code >> 2)};
    const int shift = 21 - depth;
    const float cs = 2.0f * a.r_root / (float)(1u << depth);
#pragma unroll
    for (int k = 0; k < 3; ++k) out[k] = a.o[k] + (float)(f[k] >> shift) * cs;
}

__device__ inline uint32_t quant(float v, float o, float inv) {
    int q = (int)((v - o) * inv + 0.5f);
    q = q < 0 ? 0 : q;
    q = q > (1 << 20) - 1 ? (1 << 20) - 1 : q;
    return (uint32_t)q;
}

// faces of input cube i that emit a quad
__device__ inline uint32_t face_mask(const SurfArgs& a, uint32_t i) {
    const uint32_t p = a.inv[i];
    uint32_t m = 0;
    int32_t seen[6];
    for (int f = 0; f < 6; ++f) {
        const int32_t q = a.nb[f][p];
        seen[f] = q;
        if (q < 0 || (uint32_t)q == p) continue;
        bool dup = false;
        for (int g = 0; g < f; ++g) dup = dup || seen[g] == q;
        if (dup) continue;  // this cube met the pair on an earlier face
        if (a.orig[q] < i) {  // the other cube comes first if its table names this one
            for (int g = 0; g < 6; ++g) dup = dup || a.nb[g][q] == (int32_t)p;
            if (dup) continue;
        }
        if (a.has && !(a.has[p] && a.has[q])) continue;
        const float ui = a.u[p], uj = a.u[q];
        if ((ui >= 0.0f) == (uj >= 0.0f)) continue;
        m |= 1u << f;
    }
    return m;
}

__global__ void k_srf_inv(const uint32_t* __restrict__ orig, uint32_t n, uint32_t* __restrict__ inv) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) inv[orig[p]] = p;
}

template <class S>
__global__ void k_srf_u_state(const S* __restrict__ us, uint32_t n, float* __restrict__ uf) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) uf[p] = (float)Io<S>::ld(us, p);
}

__global__ void k_srf_gather(const float* __restrict__ u_in, const uint8_t* __restrict__ has_in,
                             const uint32_t* __restrict__ orig, uint32_t n, float* __restrict__ uf,
                             uint8_t* __restrict__ hd) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (u_in) uf[p] = u_in[orig[p]];
    if (has_in) hd[p] = has_in[orig[p]] ? 1 : 0;
}

__global__ void k_srf_mark(SurfArgs a, uint8_t* __restrict__ mask, uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t m = face_mask(a, i);
    mask[i] = (uint8_t)m;
    cnt[i] = (uint32_t)__popc(m);
}

__global__ void k_srf_quads(SurfArgs a, const uint8_t* __restrict__ mask, const uint32_t* __restrict__ qoff,
                            uint64_t* __restrict__ key, float* __restrict__ px, float* __restrict__ py,
                            float* __restrict__ pz, uint8_t* __restrict__ flip) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    uint32_t m = mask[i];
    if (!m) return;
    const uint32_t p = a.inv[i];
    uint32_t qi = qoff[i];
    const int di = a.depth[p];
    float pi[3];
    corner(a.code[p], di, a, pi);
    const float csi = 2.0f * a.r_root / (float)(1u << di);
    const float ui = a.u[p];
    for (int f = 0; f < 6; ++f) {
        if (!((m >> f) & 1u)) continue;
        const int32_t q = a.nb[f][p];
        const int dj = a.depth[q];
        float pj[3];
        corner(a.code[q], dj, a, pj);
        const float csj = 2.0f * a.r_root / (float)(1u << dj);
        const float csm = csi < csj ? csi : csj;
        const float h = csm * 0.5f;
        const float uj = a.u[q];
        const float t = ui / (ui - uj);
        const int ax = f / 2;
        float s[3];
        const float ci = pi[ax] + csi * 0.5f;
        s[ax] = ci + t * ((pj[ax] + csj * 0.5f) - ci);
        const float* sc = (csi <= csj) ? pi : pj;
        float a1[3] = {0.f, 0.f, 0.f}, a2[3] = {0.f, 0.f, 0.f};
        if (ax == 0) {
            s[1] = sc[1] + h;
            s[2] = sc[2] + h;
            a1[1] = h;
            a2[2] = h;
        } else if (ax == 1) {
            s[0] = sc[0] + h;
            s[2] = sc[2] + h;
            a1[2] = h;
            a2[0] = h;
        } else {
            s[0] = sc[0] + h;
            s[1] = sc[1] + h;
            a1[0] = h;
            a2[1] = h;
        }
        float v[4][3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            v[0][k] = s[k] - a1[k] - a2[k];
            v[1][k] = s[k] + a1[k] - a2[k];
            v[2][k] = s[k] + a1[k] + a2[k];
            v[3][k] = s[k] - a1[k] + a2[k];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const uint64_t o = 4ull * qi + c;
            key[o] = (uint64_t)quant(v[c][0], a.o[0], a.inv_q) | ((uint64_t)quant(v[c][1], a.o[1], a.inv_q) << 20) |
                     ((uint64_t)quant(v[c][2], a.o[2], a.inv_q) << 40);
            px[o] = v[c][0];
            py[o] = v[c][1];
            pz[o] = v[c][2];
        }
        flip[qi] = (f % 2 == 0) ? (ui > 0.0f) : (ui < 0.0f);
        ++qi;
    }
}

__global__ void k_srf_iota(uint32_t* __restrict__ x, uint64_t n) {
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) x[k] = (uint32_t)k;
}

__global__ void k_srf_heads(const uint64_t* __restrict__ ks, uint64_t m, uint32_t* __restrict__ head) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) head[s] = (s == 0 || ks[s] != ks[s - 1]) ? 1u : 0u;
}

// hs = inclusive scan of the heads: run id of sorted occurrence s is hs[s] - 1
__global__ void k_srf_first(const uint32_t* __restrict__ head, const uint32_t* __restrict__ hs,
                            const uint32_t* __restrict__ occ_s, uint64_t m, uint32_t* __restrict__ first,
                            uint32_t* __restrict__ uid) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= m || !head[s]) return;
    const uint32_t u = hs[s] - 1;
    first[u] = occ_s[s];
    uid[u] = u;
}

// numbering by first occurrence: rank r holds run uid_s[r], first seen at fo_s[r]
__global__ void k_srf_rank(const uint32_t* __restrict__ fo_s, const uint32_t* __restrict__ uid_s, uint64_t nu,
                           const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
                           uint32_t* __restrict__ rank, float* __restrict__ verts) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nu) return;
    rank[uid_s[r]] = (uint32_t)r;
    const uint32_t o = fo_s[r];
    verts[3 * r] = px[o];
    verts[3 * r + 1] = py[o];
    verts[3 * r + 2] = pz[o];
}

__global__ void k_srf_vid(const uint32_t* __restrict__ occ_s, const uint32_t* __restrict__ hs,
                          const uint32_t* __restrict__ rank, uint64_t m, uint32_t* __restrict__ vid) {
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s < m) vid[occ_s[s]] = rank[hs[s] - 1];
}

__global__ void k_srf_tris(const uint32_t* __restrict__ vid, const uint8_t* __restrict__ flip, uint64_t nq,
                           uint32_t* __restrict__ tris) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const uint32_t i0 = vid[4 * q], i1 = vid[4 * q + 1], i2 = vid[4 * q + 2], i3 = vid[4 * q + 3];
    uint32_t* t = tris + 6 * q;
    if (!flip[q]) {
        t[0] = i0, t[1] = i1, t[2] = i2, t[3] = i0, t[4] = i2, t[5] = i3;
    } else {
        t[0] = i0, t[1] = i2, t[2] = i1, t[3] = i0, t[4] = i3, t[5] = i2;
    }
}

unsigned blocks_for(uint64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void surface_extract(tgvx_ctx* c, const float* u_in, const uint8_t* has_in, const float origin[3], float r_root,
                     uint64_t* nv, uint64_t* nt) {
    Problem& P = c->prob;
    if (!P.loaded) throw Error(TGVX_E_STATE, "no problem loaded");
    if (!origin || !nv || !nt) throw Error(TGVX_E_ARG, "null pointer");
    if (!(r_root > 0.0f)) throw Error(TGVX_E_ARG, "r_root must be positive");
    const uint32_t n = (uint32_t)P.n;
    hipStream_t st = c->stream;
    DevBuf* T = c->srf_tmp;
    // per position: u as float, votes flag; input index -> position
    DevBuf& inv = T[0];
    DevBuf& uf = T[1];
    DevBuf& hd = T[2];
    inv.ensure(4ull * n);
    uf.ensure(4ull * n);
    k_srf_inv<<<blocks_for(n), 256, 0, st>>>(P.d_orig.as<uint32_t>(), n, inv.as<uint32_t>());
    DevBuf d_uin, d_hin;
    if (u_in) {
        d_uin.ensure(4ull * n);
        h2d(c, d_uin.p, u_in, 4ull * n, st);
    } else {
        switch (P.precision) {
            case TGVX_FP64:
                k_srf_u_state<double><<<blocks_for(n), 256, 0, st>>>(P.d_state[0].as<double>(), n, uf.as<float>());
                break;
            case TGVX_FP32:
            case TGVX_BF16_DUAL:  // u is FP32 in both
                k_srf_u_state<float><<<blocks_for(n), 256, 0, st>>>(P.d_state[0].as<float>(), n, uf.as<float>());
                break;
            default:
                k_srf_u_state<bf16s><<<blocks_for(n), 256, 0, st>>>(P.d_state[0].as<bf16s>(), n, uf.as<float>());
                break;
        }
    }
    if (has_in) {
        d_hin.ensure(n);
        hd.ensure(n);
        h2d(c, d_hin.p, has_in, n, st);
    }
    if (u_in || has_in)
        k_srf_gather<<<blocks_for(n), 256, 0, st>>>(u_in ? d_uin.as<float>() : nullptr,
                                                    has_in ? d_hin.as<uint8_t>() : nullptr, P.d_orig.as<uint32_t>(),
                                                    n, uf.as<float>(), hd.as<uint8_t>());
    SurfArgs a;
    std::memset(&a, 0, sizeof(a));
    a.code = P.d_code.as<uint64_t>();
    a.depth = P.d_depth.as<int32_t>();
    for (int f = 0; f < 6; ++f) a.nb[f] = P.d_nb[f].as<int32_t>();
    a.orig = P.d_orig.as<uint32_t>();
    a.inv = inv.as<uint32_t>();
    a.u = uf.as<float>();
    a.has = has_in ? hd.as<uint8_t>() : nullptr;
    a.n = n;
    for (int k = 0; k < 3; ++k) a.o[k] = origin[k];
    a.r_root = r_root;
    a.inv_q = (float)(1u << 20) / (2.0f * r_root);

    // faces and quad offsets in input order
    DevBuf& mask = T[3];
    DevBuf& cnt = T[4];
    DevBuf& qoff = T[5];
    mask.ensure(n);
    cnt.ensure(4ull * n + 4);
    qoff.ensure(4ull * n + 4);
    k_srf_mark<<<blocks_for(n), 256, 0, st>>>(a, mask.as<uint8_t>(), cnt.as<uint32_t>());
    size_t tmp_bytes = 0;
    TGVX_HIP(rocprim::exclusive_scan(nullptr, tmp_bytes, cnt.as<uint32_t>(), qoff.as<uint32_t>(), 0u, (size_t)n + 1,
                                     rocprim::plus<uint32_t>(), st));
    DevBuf tmp;
    tmp.ensure(tmp_bytes);
    // cnt[n] = 0, so qoff[n] is the total
    TGVX_HIP(hipMemsetAsync(cnt.as<uint32_t>() + n, 0, 4, st));
    TGVX_HIP(rocprim::exclusive_scan(tmp.p, tmp_bytes, cnt.as<uint32_t>(), qoff.as<uint32_t>(), 0u, (size_t)n + 1,
                                     rocprim::plus<uint32_t>(), st));
    uint32_t nq32 = 0;
    TGVX_HIP(hipMemcpyAsync(&nq32, qoff.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, st));
    TGVX_HIP(hipStreamSynchronize(st));
    const uint64_t nq = nq32, m = 4 * nq;
    if (m >= (1ull << 32)) throw Error(TGVX_E_ARG, "surface too large (2^30 quads at most)");
    c->srf_nv = 0;
    c->srf_nt = 2 * nq;
    if (nq == 0) {
        *nv = 0;
        *nt = 0;
        return;
    }
    // corners of every quad, in (input index, face, corner) occurrence order
    DevBuf keys, keys_s, occ, occ_s, px, py, pz, flip;
    keys.ensure(8 * m);
    keys_s.ensure(8 * m);
    occ.ensure(4 * m);
    occ_s.ensure(4 * m);
    px.ensure(4 * m);
    py.ensure(4 * m);
    pz.ensure(4 * m);
    flip.ensure(nq);
    k_srf_quads<<<blocks_for(n), 256, 0, st>>>(a, mask.as<uint8_t>(), qoff.as<uint32_t>(), keys.as<uint64_t>(),
                                               px.as<float>(), py.as<float>(), pz.as<float>(), flip.as<uint8_t>());
    k_srf_iota<<<blocks_for(m), 256, 0, st>>>(occ.as<uint32_t>(), m);
    // stable sort by key keeps each key's occurrences in order
    tmp_bytes = 0;
    TGVX_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys.as<uint64_t>(), keys_s.as<uint64_t>(),
                                       occ.as<uint32_t>(), occ_s.as<uint32_t>(), (size_t)m, 0u, 60u, st));
    tmp.ensure(tmp_bytes);
    TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, keys.as<uint64_t>(), keys_s.as<uint64_t>(),
                                       occ.as<uint32_t>(), occ_s.as<uint32_t>(), (size_t)m, 0u, 60u, st));
    // runs of equal keys: one vertex each
    DevBuf head, hs;
    head.ensure(4 * m);
    hs.ensure(4 * m);
    k_srf_heads<<<blocks_for(m), 256, 0, st>>>(keys_s.as<uint64_t>(), m, head.as<uint32_t>());
    tmp_bytes = 0;
    TGVX_HIP(rocprim::inclusive_scan(nullptr, tmp_bytes, head.as<uint32_t>(), hs.as<uint32_t>(), (size_t)m,
                                     rocprim::plus<uint32_t>(), st));
    tmp.ensure(tmp_bytes);
    TGVX_HIP(rocprim::inclusive_scan(tmp.p, tmp_bytes, head.as<uint32_t>(), hs.as<uint32_t>(), (size_t)m,
                                     rocprim::plus<uint32_t>(), st));
    uint32_t nu32 = 0;
    TGVX_HIP(hipMemcpyAsync(&nu32, hs.as<uint32_t>() + (m - 1), 4, hipMemcpyDeviceToHost, st));
    TGVX_HIP(hipStreamSynchronize(st));
    const uint64_t nu = nu32;
    // first occurrence of each run, then number the runs in that order
    DevBuf first, uid, fo_s, uid_s, rank, vid;
    first.ensure(4 * nu);
    uid.ensure(4 * nu);
    fo_s.ensure(4 * nu);
    uid_s.ensure(4 * nu);
    rank.ensure(4 * nu);
    vid.ensure(4 * m);
    k_srf_first<<<blocks_for(m), 256, 0, st>>>(head.as<uint32_t>(), hs.as<uint32_t>(), occ_s.as<uint32_t>(), m,
                                               first.as<uint32_t>(), uid.as<uint32_t>());
    tmp_bytes = 0;
    TGVX_HIP(rocprim::radix_sort_pairs(nullptr, tmp_bytes, first.as<uint32_t>(), fo_s.as<uint32_t>(),
                                       uid.as<uint32_t>(), uid_s.as<uint32_t>(), (size_t)nu, 0u, 32u, st));
    tmp.ensure(tmp_bytes);
    TGVX_HIP(rocprim::radix_sort_pairs(tmp.p, tmp_bytes, first.as<uint32_t>(), fo_s.as<uint32_t>(),
                                       uid.as<uint32_t>(), uid_s.as<uint32_t>(), (size_t)nu, 0u, 32u, st));
    c->srf_verts.ensure(12 * nu);
    c->srf_tris.ensure(4 * 6 * nq);
    k_srf_rank<<<blocks_for(nu), 256, 0, st>>>(fo_s.as<uint32_t>(), uid_s.as<uint32_t>(), nu, px.as<float>(),
                                               py.as<float>(), pz.as<float>(), rank.as<uint32_t>(),
                                               c->srf_verts.as<float>());
    k_srf_vid<<<blocks_for(m), 256, 0, st>>>(occ_s.as<uint32_t>(), hs.as<uint32_t>(), rank.as<uint32_t>(), m,
                                             vid.as<uint32_t>());
    k_srf_tris<<<blocks_for(nq), 256, 0, st>>>(vid.as<uint32_t>(), flip.as<uint8_t>(), nq,
                                               c->srf_tris.as<uint32_t>());
    TGVX_HIP(hipGetLastError());
    TGVX_HIP(hipStreamSynchronize(st));
    c->srf_nv = nu;
    *nv = nu;
    *nt = 2 * nq;
}

void surface_copy(tgvx_ctx* c, float* verts, uint32_t* tris) {
    if (c->srf_nv && verts) d2h(c, verts, c->srf_verts.p, 12 * c->srf_nv);
    if (c->srf_nt && tris) d2h(c, tris, c->srf_tris.p, 12 * c->srf_nt);
}

}  // namespace tgvx
