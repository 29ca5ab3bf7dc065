// CPU oracle for the global-TGV hot path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, the reference's
// algorithm for the hot path so that the HIP implementation can be checked
// against it. Only tests/, the smoke check and the CPU-baseline leg of the
// benchmark load it; the product library never links or calls it.
//
// Everything here is templated on the arithmetic type T (float or double) and
// built with -ffp-contract=off so that no multiply-add is fused: the result of
// every expression is the IEEE-rounded result of the same expression written
// in the reference.
//
// Reference files restated (paths relative to the reference root):
//   src/types.hpp:39-73          Morton dilate/compact/encode/decode
//   src/balance.cpp:22-49        ancestor64, parent_code, face_neighbor
//   src/histogram.cpp:12-31      CubeHistogram::mean_indicator / median_indicator
//   src/histogram.cpp:35-113     DepthMipmap::build / project / depth_at
//   src/histogram.cpp:118-163    cube_center, compute_histograms
//   src/tgv.cpp:57-133           clip_ball3/6, prox_data, neighbor_idx_ctf, NeighbourTable
//   src/tgv.cpp:137-232          run_level (five Jacobi passes per iteration)
//   src/tgv.cpp:241-352          run_level_frozen
//   src/tgv.cpp:356-511          tgv_minimize, tgv_minimize_oc
//   src/marching_cubes.cpp:16-181  extract_surface (used only to measure surfaces)
// Functions with no reference counterpart (energy / duality gap, the exact-adjoint
// operator mode, the MLP) are marked as such where they are defined.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------- Morton
// Straight bit loops (not the reference's magic-constant dilation) so that
// this file is an independent statement of the same interleaving:
// bit b of x goes to bit 3b, of y to 3b+1, of z to 3b+2 (types.hpp:10,62).
static inline uint64_t spread3(uint32_t v) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= (uint64_t)((v >> b) & 1u) << (3 * b);
    return r;
}
static inline uint32_t gather3(uint64_t v) {
    uint32_t r = 0;
    for (int b = 0; b < 21; ++b) r |= (uint32_t)((v >> (3 * b)) & 1u) << b;
    return r;
}
static inline uint64_t encode(uint32_t ix, uint32_t iy, uint32_t iz) {
    return spread3(ix) | (spread3(iy) << 1) | (spread3(iz) << 2);
}
static inline void decode(uint64_t c, uint32_t& ix, uint32_t& iy, uint32_t& iz) {
    ix = gather3(c);
    iy = gather3(c >> 1);
    iz = gather3(c >> 2);
}
// balance.cpp:22-26 (target_depth may be -1: the shift then exceeds 63 and the
// ancestor is 0, as the unsigned arithmetic there yields).
static inline uint64_t ancestor(uint64_t c, int target_depth) {
    unsigned shift = 3u * (21u - (unsigned)target_depth);
    if (shift >= 64u) return 0;
    return c & ~((1ull << shift) - 1ull);
}
static inline uint64_t parent(uint64_t c, int depth) { return ancestor(c, depth - 1); }

static const uint64_t NONE = 0xffffffffffffffffull;  // {UINT32_MAX, UINT32_MAX}

// balance.cpp:32-49; faces 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z
static inline uint64_t face_nb(uint64_t c, int face, int depth) {
    uint32_t x, y, z;
    decode(c, x, y, z);
    const uint32_t step = 1u << (21 - depth), bound = 1u << 21;
    switch (face) {
        case 0: if (x + step >= bound) return NONE; x += step; break;
        case 1: if (x < step) return NONE; x -= step; break;
        case 2: if (y + step >= bound) return NONE; y += step; break;
        case 3: if (y < step) return NONE; y -= step; break;
        case 4: if (z + step >= bound) return NONE; z += step; break;
        case 5: if (z < step) return NONE; z -= step; break;
        default: return NONE;
    }
    return encode(x, y, z);
}

// ---------------------------------------------------------------- inputs
struct Cube {  // == OctreeCube (types.hpp:31-35)
    uint32_t hi, lo;
    float r_c;
    int32_t depth;
};
static inline uint64_t code_of(const Cube& c) { return ((uint64_t)c.hi << 32) | c.lo; }

struct View {  // one depth map: base level + intrinsics + world->camera 3x4 rows
    const float* depths;
    int32_t width, height;
    float fx, fy, cx, cy;
    float m[12];
    int32_t lidar;
};

// ---------------------------------------------------------------- histograms
// histogram.cpp:21-31 generalised to nb bins; bin centre -1 + (k+0.5)/(nb/2).
template <class T>
static T median_indicator(const uint32_t* bins, int nb) {
    uint32_t tot = 0;
    for (int k = 0; k < nb; ++k) tot += bins[k];
    if (tot == 0) return T(0);
    uint32_t cum = 0;
    const T half = T(0.5) * T(nb);
    for (int k = 0; k < nb; ++k) {
        cum += bins[k];
        if (cum * 2u >= tot) return T(-1) + (T(k) + T(0.5)) / half;
    }
    return T(1);
}
// histogram.cpp:12-19
template <class T>
static T mean_indicator(const uint32_t* bins, int nb) {
    uint32_t tot = 0;
    for (int k = 0; k < nb; ++k) tot += bins[k];
    if (tot == 0) return T(0);
    const T half = T(0.5) * T(nb);
    T sum = T(0);
    for (int k = 0; k < nb; ++k) sum += T(bins[k]) * (T(-1) + (T(k) + T(0.5)) / half);
    return sum / T(tot);
}

struct Mip {  // DepthMipmap restated (histogram.hpp:22-38)
    std::vector<std::vector<float>> lv;
    std::vector<int> w, h;
    float fx, fy, cx, cy, m[12];
    int lidar;
};

// histogram.cpp:35-74: each level averages the finite, positive pixels of the
// 2x2 parent block (du inner, dv outer), NaN when none; at most 12 extra levels.
static Mip build_mip(const View& v) {
    Mip mp;
    mp.fx = v.fx; mp.fy = v.fy; mp.cx = v.cx; mp.cy = v.cy;
    std::memcpy(mp.m, v.m, sizeof(mp.m));
    mp.lidar = v.lidar;
    mp.lv.emplace_back(v.depths, v.depths + (size_t)v.width * v.height);
    mp.w.push_back(v.width);
    mp.h.push_back(v.height);
    for (int l = 0; l < 12; ++l) {
        const int pw = mp.w.back(), ph = mp.h.back();
        if (pw <= 1 || ph <= 1) break;
        const int nw = std::max(1, pw / 2), nh = std::max(1, ph / 2);
        std::vector<float> nx((size_t)nw * nh, std::numeric_limits<float>::quiet_NaN());
        const std::vector<float>& prev = mp.lv.back();
        for (int y = 0; y < nh; ++y)
            for (int x = 0; x < nw; ++x) {
                float sum = 0.0f;
                int cnt = 0;
                for (int dv = 0; dv < 2; ++dv)
                    for (int du = 0; du < 2; ++du) {
                        const int px = 2 * x + du, py = 2 * y + dv;
                        if (px < pw && py < ph) {
                            const float d = prev[(size_t)py * pw + px];
                            if (std::isfinite(d) && d > 0.0f) { sum += d; ++cnt; }
                        }
                    }
                if (cnt > 0) nx[(size_t)y * nw + x] = sum / (float)cnt;
            }
        mp.lv.push_back(std::move(nx));
        mp.w.push_back(nw);
        mp.h.push_back(nh);
    }
    return mp;
}

// histogram.cpp:76-106. The camera point is the 3x4 world->camera matrix times
// [P,1], each row summed x, y, z, then the translation. The level is the
// truncated float log2 of the footprint, as there.
static bool project(const Mip& mp, const float P[3], float r_c, int& level, int& u, int& v, float& cz) {
    const float* m = mp.m;
    const float camx = ((m[0] * P[0] + m[1] * P[1]) + m[2] * P[2]) + m[3];
    const float camy = ((m[4] * P[0] + m[5] * P[1]) + m[6] * P[2]) + m[7];
    const float z = ((m[8] * P[0] + m[9] * P[1]) + m[10] * P[2]) + m[11];
    if (z <= 0.0f) return false;
    const float uf = camx / z * mp.fx + mp.cx;
    const float vf = camy / z * mp.fy + mp.cy;
    // (int) truncation; anything that truncates outside [0, size) is rejected,
    // which is also what x86 does for NaN / out-of-range conversions.
    if (!(uf > -1.0f && uf < (float)mp.w[0])) return false;
    if (!(vf > -1.0f && vf < (float)mp.h[0])) return false;
    const int u0 = (int)uf, v0 = (int)vf;
    cz = z;
    const float footprint = std::fabs(mp.fx) * r_c / z;
    const float fp = std::max(1.0f, footprint);
    // An infinite footprint (z underflowing) makes the float->int conversion
    // undefined in the reference; it is taken as the coarsest level here.
    int lvl = fp <= std::numeric_limits<float>::max() ? std::max(0, (int)std::log2(fp)) : (int)mp.lv.size();
    lvl = std::min(lvl, (int)mp.lv.size() - 1);
    level = lvl;
    u = std::clamp(u0 >> lvl, 0, mp.w[lvl] - 1);
    v = std::clamp(v0 >> lvl, 0, mp.h[lvl] - 1);
    return true;
}

// histogram.cpp:118-125
static void cube_center(const Cube& c, const float origin[3], float r_root, float out[3]) {
    uint32_t fx, fy, fz;
    decode(code_of(c), fx, fy, fz);
    const int shift = 21 - c.depth;
    const uint32_t ix = fx >> shift, iy = fy >> shift, iz = fz >> shift;
    const float cs = 2.0f * r_root / (float)(1u << c.depth);
    out[0] = origin[0] + (ix + 0.5f) * cs;
    out[1] = origin[1] + (iy + 0.5f) * cs;
    out[2] = origin[2] + (iz + 0.5f) * cs;
}

// histogram.cpp:127-163 with nb bins: bin = clamp(int((a+1) * nb/2), 0, nb-1).
static void compute_histograms(const Cube* cubes, size_t n, const std::vector<Mip>& mips,
                               const float origin[3], float r_root, int nb, uint32_t* bins) {
    std::memset(bins, 0, sizeof(uint32_t) * n * nb);
    const float scale = 0.5f * (float)nb;
    for (size_t ci = 0; ci < n; ++ci) {
        const Cube& c = cubes[ci];
        float ctr[3];
        cube_center(c, origin, r_root, ctr);
        const float r_c = c.r_c;
        const float delta = 6.0f * r_c;
        const float eta = 18.0f * r_c;
        if (r_c <= 0.0f) continue;
        for (const Mip& mp : mips) {
            int level, u, v;
            float cz;
            if (!project(mp, ctr, r_c, level, u, v, cz)) continue;
            const float d = mp.lv[level][(size_t)v * mp.w[level] + u];
            if (!std::isfinite(d) || d <= 0.0f) continue;
            float a = d - cz;
            if (a < -eta) continue;
            a = std::clamp(a / delta, -1.0f, 1.0f);
            int bin = (int)((a + 1.0f) * scale);
            bin = std::clamp(bin, 0, nb - 1);
            bins[ci * nb + bin] += mp.lidar ? 5u : 1u;
        }
    }
}

// ---------------------------------------------------------------- TGV
template <class T>
struct State {  // TGVCube (tgv.hpp:12-17) with arithmetic type T
    T u;
    T v[3];
    T p[3];
    T q[6];
};

struct Params {
    float alpha0, alpha1, lambda;
    int32_t iters;        // fixed iterations per level (when gap_tol <= 0)
    int32_t op_mode;      // 0 = reference rules, 1 = exact adjoint
    int32_t chunk;        // iterations between gap evaluations
    int32_t max_iters;    // cap per level when iterating to a gap
    double gap_tol;       // relative gap target; <= 0 means fixed iterations
    int32_t threads;      // OpenMP threads for the passes (1 = reference)
    int32_t warm_fallback;  // with an init field: keep cold init where it has the lower gap
    float dual_eps;         // with an init field: > 0 recovers p, q from (u, v), see dual_recover
    int32_t warm_dual;      // with an init field: 1 keeps the dual recovered from the cold init
    float sigma, tau;       // primal-dual steps; 0 = the reference's 0.25 (tgv.cpp:149-150)
    // per octree depth: > 0 replaces sigma, tau for that depth's level (project-defined)
    float sigma_depth[22], tau_depth[22];
};

using CodeMap = std::unordered_map<uint64_t, uint32_t>;

// (code, depth) -> last input index with that pair. Used only by the
// exact-adjoint mode, whose neighbour table is depth-aware (no reference
// counterpart; the reference keys its map on the code alone, tgv.cpp:50-54).
struct PairHash {
    size_t operator()(const std::pair<uint64_t, int>& k) const {
        return std::hash<uint64_t>()(k.first * 31u + (uint64_t)k.second);
    }
};
using DepthCodeMap = std::unordered_map<std::pair<uint64_t, int>, uint32_t, PairHash>;

// tgv.cpp:57-72
template <class T>
static inline void clip3(T p[3], T r) {
    T n2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    if (n2 > r * r) {
        T s = r / std::sqrt(n2);
        p[0] *= s; p[1] *= s; p[2] *= s;
    }
}
template <class T>
static inline void clip6(T q[6], T r) {
    T n2 = 0;
    for (int i = 0; i < 6; ++i) n2 += q[i] * q[i];
    if (n2 > r * r) {
        T s = r / std::sqrt(n2);
        for (int i = 0; i < 6; ++i) q[i] *= s;
    }
}

// tgv.cpp:75-94 with the data weight lambda (the reference's is 1) and nb bins.
// The weighted median is recomputed every call, as there.
template <class T>
static inline T prox_data(T z, T tau_lambda, const uint32_t* bins, int nb) {
    uint32_t tot = 0;
    for (int k = 0; k < nb; ++k) tot += bins[k];
    if (tot == 0) return z;
    uint32_t cum = 0;
    T med = z;
    const T half = T(0.5) * T(nb);
    for (int k = 0; k < nb; ++k) {
        cum += bins[k];
        if (cum * 2u >= tot) { med = T(-1) + (T(k) + T(0.5)) / half; break; }
    }
    const T diff = z - med;
    const T shrink = tau_lambda * (T)tot;
    if (diff > shrink) return z - shrink;
    if (diff < -shrink) return z + shrink;
    return med;
}

// tgv.cpp:97-109
static inline int32_t nb_ctf(const CodeMap& cm, uint64_t code, int face, int depth) {
    const uint64_t nb = face_nb(code, face, depth);
    if (nb == NONE) return -1;
    auto it = cm.find(nb);
    if (it != cm.end()) return (int32_t)it->second;
    if (depth > 0) {
        it = cm.find(parent(nb, depth));
        if (it != cm.end()) return (int32_t)it->second;
    }
    return -1;
}

// Geometric neighbour of the exact-adjoint mode: the cube of the same depth
// at the face-neighbour code, else the cube one level coarser covering it,
// else none. A cube that is not the last one with its own (code, depth) is a
// duplicate and gets no neighbours (the lookups find the last one).
static inline int32_t nb_geo(const DepthCodeMap& dm, uint64_t code, int face, int depth, bool shadow) {
    if (shadow) return -1;
    const uint64_t nb = face_nb(code, face, depth);
    if (nb == NONE) return -1;
    auto it = dm.find({nb, depth});
    if (it != dm.end()) return (int32_t)it->second;
    if (depth > 0) {
        it = dm.find({parent(nb, depth), depth - 1});
        if (it != dm.end()) return (int32_t)it->second;
    }
    return -1;
}

// Incoming-edge lists for the exact-adjoint operator and the gap (no reference
// counterpart): for active position ai, every (active position k, axis a) with
// nb_{+a}(active[k]) == active[ai], ordered by (a, k).
struct Incoming {
    std::vector<uint32_t> start;               // size na+1
    std::vector<std::pair<uint32_t, int>> ent;  // (k, a)
};

struct Level {
    std::vector<uint32_t> active;  // global indices, ascending
    std::vector<int32_t> nb;       // reference table nb[ai*6+f] (global index or -1), tgv.cpp:115-133
    std::vector<int32_t> nba;      // exact-adjoint (depth-aware) table, same layout
    std::vector<int32_t> pos;      // global index -> active position or -1 (size N)
    Incoming in;                   // incoming +a edges of nba within the level
};

static void build_level(Level& L, const Cube* cubes, const CodeMap& cm, const DepthCodeMap& dm, size_t N) {
    const size_t na = L.active.size();
    L.nb.resize(na * 6);
    L.nba.resize(na * 6);
    for (size_t ai = 0; ai < na; ++ai) {
        const uint32_t i = L.active[ai];
        const Cube& c = cubes[i];
        const bool shadow = dm.at({code_of(c), c.depth}) != i;
        for (int f = 0; f < 6; ++f) {
            L.nb[ai * 6 + f] = nb_ctf(cm, code_of(c), f, c.depth);
            L.nba[ai * 6 + f] = nb_geo(dm, code_of(c), f, c.depth, shadow);
        }
    }
    L.pos.assign(N, -1);
    for (size_t ai = 0; ai < na; ++ai) L.pos[L.active[ai]] = (int32_t)ai;
    std::vector<uint32_t> cnt(na + 1, 0);
    for (size_t k = 0; k < na; ++k)
        for (int a = 0; a < 3; ++a) {
            const int32_t j = L.nba[k * 6 + 2 * a];
            if (j >= 0 && L.pos[j] >= 0) cnt[L.pos[j] + 1]++;
        }
    for (size_t i = 0; i < na; ++i) cnt[i + 1] += cnt[i];
    L.in.start = cnt;
    L.in.ent.resize(cnt[na]);
    std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int a = 0; a < 3; ++a)
        for (size_t k = 0; k < na; ++k) {
            const int32_t j = L.nba[k * 6 + 2 * a];
            if (j >= 0 && L.pos[j] >= 0) L.in.ent[fill[L.pos[j]]++] = {(uint32_t)k, a};
        }
}

// Energy and duality gap of one level (no reference counterpart; the reference
// computes neither). Level problem, with frozen values for neighbours outside
// the active set A (their u_bar / v_bar):
//   P(u,v) = sum_i lambda W_i |u_i - m_i| + alpha1 |grad u - v|_i + alpha0 |eps(v)|_i
//   D(p,q) = min over u in [-1,1], v in [-V,V]^3 of the Lagrangian
// grad / eps are the forward differences of the dual passes (missing neighbour:
// zero difference), eps with 1/2 on the off-diagonals, |.| the Euclidean norm
// of the 3- resp. 6-vector. Returns P and D accumulated in double.
template <class T>
static void level_gap(const Level& L, const std::vector<State<T>>& st, const std::vector<T>& ub,
                      const std::vector<T>* vb, const uint32_t* bins, int nb, const Params& prm,
                      double vbox, double& P_out, double& D_out) {
    const size_t na = L.active.size();
    const double a0 = prm.alpha0, a1 = prm.alpha1, lam = prm.lambda;
    double P = 0, D = 0;
    // value of u / v component c at global index j, for the energy at (u, v)
    auto U = [&](int32_t j) -> double { return L.pos[j] >= 0 ? (double)st[j].u : (double)ub[j]; };
    auto Vc = [&](int32_t j, int c) -> double {
        return L.pos[j] >= 0 ? (double)st[j].v[c] : (double)vb[c][j];
    };
    // q index for (component c, axis a) and its weight in eps
    static const int QI[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
    for (size_t ai = 0; ai < na; ++ai) {
        const uint32_t i = L.active[ai];
        const State<T>& s = st[i];
        const uint32_t* b = bins + (size_t)i * nb;
        uint32_t tot = 0;
        for (int k = 0; k < nb; ++k) tot += b[k];
        const double W = (double)tot;
        const double m = (double)median_indicator<T>(b, nb);
        const double ui = (double)s.u;
        // primal
        double g[3], dv[3][3];  // dv[c][a] = D_a v^c
        bool has[3];
        for (int a = 0; a < 3; ++a) {
            const int32_t j = L.nba[ai * 6 + 2 * a];
            has[a] = j >= 0;
            g[a] = has[a] ? U(j) - ui : 0.0;
            for (int c = 0; c < 3; ++c) dv[c][a] = has[a] ? Vc(j, c) - (double)s.v[c] : 0.0;
        }
        double k1[3] = {g[0] - (double)s.v[0], g[1] - (double)s.v[1], g[2] - (double)s.v[2]};
        double e[6] = {dv[0][0], dv[1][1], dv[2][2], 0.5 * (dv[0][1] + dv[1][0]),
                       0.5 * (dv[0][2] + dv[2][0]), 0.5 * (dv[1][2] + dv[2][1])};
        double nk = std::sqrt(k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
        double ne = 0;
        for (int k = 0; k < 6; ++k) ne += e[k] * e[k];
        ne = std::sqrt(ne);
        P += (tot ? lam * W * std::fabs(ui - m) : 0.0) + a1 * nk + a0 * ne;
        // dual: constant part from frozen neighbours, and the u / v coefficients
        double cst = 0;
        double cu = 0;         // coefficient of u_i
        double cv[3] = {0, 0, 0};  // coefficient of v_i
        for (int a = 0; a < 3; ++a) {
            const int32_t j = L.nba[ai * 6 + 2 * a];
            if (j < 0) continue;
            cu -= (double)s.p[a];
            for (int c = 0; c < 3; ++c) {
                const double w = (c == a) ? 1.0 : 0.5;
                cv[c] -= w * (double)s.q[QI[c][a]];
            }
            if (L.pos[j] < 0) {
                cst += (double)s.p[a] * (double)ub[j];
                for (int c = 0; c < 3; ++c) {
                    const double w = (c == a) ? 1.0 : 0.5;
                    cst += w * (double)s.q[QI[c][a]] * (double)vb[c][j];
                }
            }
        }
        for (uint32_t e2 = L.in.start[ai]; e2 < L.in.start[ai + 1]; ++e2) {
            const uint32_t k = L.active[L.in.ent[e2].first];
            const int a = L.in.ent[e2].second;
            cu += (double)st[k].p[a];
            for (int c = 0; c < 3; ++c) {
                const double w = (c == a) ? 1.0 : 0.5;
                cv[c] += w * (double)st[k].q[QI[c][a]];
            }
        }
        for (int c = 0; c < 3; ++c) cv[c] -= (double)s.p[c];
        // min over u in [-1, 1] of lam W |u - m| + cu u
        double mu;
        if (tot == 0) {
            mu = -std::fabs(cu);
        } else {
            auto gfun = [&](double x) { return lam * W * std::fabs(x - m) + cu * x; };
            mu = std::min(gfun(-1.0), std::min(gfun(m), gfun(1.0)));
        }
        D += cst + mu - vbox * (std::fabs(cv[0]) + std::fabs(cv[1]) + std::fabs(cv[2]));
    }
    P_out = P;
    D_out = D;
}

template <class T>
struct Solver {
    const Cube* cubes;
    size_t N;
    const uint32_t* bins;
    int nb;
    Params prm;
    std::vector<State<T>> st;
    std::vector<T> ub, vb[3];
    CodeMap cm;
    DepthCodeMap dm;
    std::vector<std::vector<uint32_t>> by_depth;
    int min_d = 21, max_d = 0;
    // per-level report
    std::vector<int> level_iters;
    std::vector<double> level_P, level_D;

    void setup() {
        st.assign(N, State<T>{});
        cm.reserve(N);
        for (size_t i = 0; i < N; ++i) cm[code_of(cubes[i])] = (uint32_t)i;  // last writer wins
        dm.reserve(N);
        for (size_t i = 0; i < N; ++i) dm[{code_of(cubes[i]), (int)cubes[i].depth}] = (uint32_t)i;
        for (size_t i = 0; i < N; ++i) st[i].u = median_indicator<T>(bins + i * nb, nb);
        for (size_t i = 0; i < N; ++i) {
            min_d = std::min(min_d, (int)cubes[i].depth);
            max_d = std::max(max_d, (int)cubes[i].depth);
        }
        by_depth.assign(max_d + 1, {});
        for (size_t i = 0; i < N; ++i) by_depth[cubes[i].depth].push_back((uint32_t)i);
        ub.assign(N, T(0));
        for (int c = 0; c < 3; ++c) vb[c].assign(N, T(0));
        for (size_t i = 0; i < N; ++i) ub[i] = st[i].u;
    }

    // tgv.cpp:391-409: copy u, v from the cube holding the parent code (if any),
    // zero p, q, reset the extrapolated values.
    // The exact-adjoint mode copies from the cube exactly one level coarser.
    void parent_init(const std::vector<uint32_t>& active, int L) {
        for (uint32_t i : active) {
            const uint64_t pc = parent(code_of(cubes[i]), L);
            int64_t pi = -1;
            if (prm.op_mode == 0) {
                auto it = cm.find(pc);
                if (it != cm.end()) pi = it->second;
            } else {
                auto it = dm.find({pc, L - 1});
                if (it != dm.end()) pi = it->second;
            }
            if (pi >= 0) {
                st[i].u = st[pi].u;
                for (int c = 0; c < 3; ++c) st[i].v[c] = st[pi].v[c];
            }
            for (int c = 0; c < 3; ++c) st[i].p[c] = T(0);
            for (int k = 0; k < 6; ++k) st[i].q[k] = T(0);
            ub[i] = st[i].u;
            for (int c = 0; c < 3; ++c) vb[c][i] = st[i].v[c];
        }
    }
    void set_init(const std::vector<uint32_t>& active, const float* init_uv, bool keep_dual = false) {
        for (uint32_t i : active) {
            st[i].u = (T)init_uv[(size_t)i * 4];
            for (int c = 0; c < 3; ++c) st[i].v[c] = (T)init_uv[(size_t)i * 4 + 1 + c];
            if (!keep_dual) {
                for (int c = 0; c < 3; ++c) st[i].p[c] = T(0);
                for (int k = 0; k < 6; ++k) st[i].q[k] = T(0);
            }
            ub[i] = st[i].u;
            for (int c = 0; c < 3; ++c) vb[c][i] = st[i].v[c];
        }
    }

    // Dual initialisation from a primal guess (no reference counterpart):
    // the optimality conditions give p = alpha1 K1/|K1| and q = alpha0 e/|e|
    // wherever K1 = grad u - v resp. e = eps(v) is non-zero; with a guess the
    // norms are clamped below by eps, so small residuals give small duals.
    // Operators as in the gap (depth-aware table, frozen values outside A).
    void dual_recover(const Level& L, T eps) {
        const T a0 = (T)prm.alpha0, a1 = (T)prm.alpha1;
        const long na = (long)L.active.size();
        for (long ai = 0; ai < na; ++ai) {
            const uint32_t i = L.active[ai];
            State<T>& s = st[i];
            T g[3], dv[3][3];
            for (int a = 0; a < 3; ++a) {
                const int32_t j = L.nba[ai * 6 + 2 * a];
                const bool in = j >= 0 && L.pos[j] >= 0;
                g[a] = j < 0 ? T(0) : (in ? st[j].u : ub[j]) - s.u;
                for (int c = 0; c < 3; ++c)
                    dv[c][a] = j < 0 ? T(0) : (in ? st[j].v[c] : vb[c][j]) - s.v[c];
            }
            const T k1[3] = {g[0] - s.v[0], g[1] - s.v[1], g[2] - s.v[2]};
            const T e[6] = {dv[0][0], dv[1][1], dv[2][2], T(0.5) * (dv[0][1] + dv[1][0]),
                            T(0.5) * (dv[0][2] + dv[2][0]), T(0.5) * (dv[1][2] + dv[2][1])};
            const T n1 = std::sqrt(k1[0] * k1[0] + k1[1] * k1[1] + k1[2] * k1[2]);
            T n0 = T(0);
            for (int k = 0; k < 6; ++k) n0 += e[k] * e[k];
            n0 = std::sqrt(n0);
            const T s1 = a1 / std::max(n1, eps), s0 = a0 / std::max(n0, eps);
            for (int c = 0; c < 3; ++c) s.p[c] = k1[c] * s1;
            for (int k = 0; k < 6; ++k) s.q[k] = e[k] * s0;
        }
    }

    // One iteration = the five passes of tgv.cpp:158-230. frozen_* are the
    // level-start snapshots of run_level_frozen (tgv.cpp:263-298), or null.
    void iterate(const Level& L, int iters, std::vector<T>& un, std::vector<T>* vn, const T* fu,
                 const T* const* fv, uint32_t leaf_first, uint32_t leaf_last) {
        const int dl = L.active.empty() ? 0 : (int)cubes[L.active[0]].depth;
        const float sg = prm.sigma_depth[dl] > 0.f ? prm.sigma_depth[dl] : prm.sigma;
        const float ta = prm.tau_depth[dl] > 0.f ? prm.tau_depth[dl] : prm.tau;
        const T sigma = sg > 0.f ? (T)sg : T(0.25), tau = ta > 0.f ? (T)ta : T(0.25);
        const T a0 = (T)prm.alpha0, a1 = (T)prm.alpha1;
        const T tl = tau * (T)prm.lambda;
        const long na = (long)L.active.size();
        const int mode = prm.op_mode;
        auto inleaf = [&](int32_t j) { return (uint32_t)j >= leaf_first && (uint32_t)j <= leaf_last; };
        static const int QI[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
        // dual passes: the reference table, or the depth-aware one (exact adjoint)
        const std::vector<int32_t>& NB = mode == 0 ? L.nb : L.nba;
#ifdef _OPENMP
        const int nt = prm.threads > 0 ? prm.threads : 1;
#endif
        for (int it = 0; it < iters; ++it) {
            // ---- dual p (tgv.cpp:159-169)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
            for (long ai = 0; ai < na; ++ai) {
                const uint32_t i = L.active[ai];
                auto ubf = [&](int f) -> T {
                    const int32_t j = NB[ai * 6 + f];
                    if (j < 0) return ub[i];
                    if (fu && !inleaf(j)) return fu[j];
                    return ub[j];
                };
                st[i].p[0] += sigma * (ubf(0) - ub[i] - vb[0][i]);
                st[i].p[1] += sigma * (ubf(2) - ub[i] - vb[1][i]);
                st[i].p[2] += sigma * (ubf(4) - ub[i] - vb[2][i]);
                clip3(st[i].p, a1);
            }
            // ---- dual q (tgv.cpp:172-189)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
            for (long ai = 0; ai < na; ++ai) {
                const uint32_t i = L.active[ai];
                auto vbf = [&](int f, int c) -> T {
                    const int32_t j = NB[ai * 6 + f];
                    if (j < 0) return vb[c][i];
                    if (fv && !inleaf(j)) return fv[c][j];
                    return vb[c][j];
                };
                T eps[6] = {vbf(0, 0) - vb[0][i],
                            vbf(2, 1) - vb[1][i],
                            vbf(4, 2) - vb[2][i],
                            T(0.5) * (vbf(2, 0) - vb[0][i] + vbf(0, 1) - vb[1][i]),
                            T(0.5) * (vbf(4, 0) - vb[0][i] + vbf(0, 2) - vb[2][i]),
                            T(0.5) * (vbf(4, 1) - vb[1][i] + vbf(2, 2) - vb[2][i])};
                for (int k = 0; k < 6; ++k) st[i].q[k] += sigma * eps[k];
                clip6(st[i].q, a0);
            }
            if (mode == 0) {
                // ---- primal u (tgv.cpp:192-202)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
                for (long ai = 0; ai < na; ++ai) {
                    const uint32_t i = L.active[ai];
                    auto gp = [&](int f, int c) -> T {
                        const int32_t j = L.nb[ai * 6 + f];
                        return j >= 0 ? st[j].p[c] : T(0);
                    };
                    T div_p = (st[i].p[0] - gp(1, 0)) + (st[i].p[1] - gp(3, 1)) + (st[i].p[2] - gp(5, 2));
                    un[ai] = prox_data(st[i].u + tau * div_p, tl, bins + (size_t)i * nb, nb);
                }
                // ---- primal v (tgv.cpp:205-217)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
                for (long ai = 0; ai < na; ++ai) {
                    const uint32_t i = L.active[ai];
                    auto gq = [&](int f, int c) -> T {
                        const int32_t j = L.nb[ai * 6 + f];
                        return j >= 0 ? st[j].q[c] : T(0);
                    };
                    const State<T>& s = st[i];
                    T dqx = (s.q[0] - gq(1, 0)) + (s.q[3] - gq(3, 3)) + (s.q[4] - gq(5, 4));
                    T dqy = (s.q[3] - gq(1, 3)) + (s.q[1] - gq(3, 1)) + (s.q[5] - gq(5, 5));
                    T dqz = (s.q[4] - gq(1, 4)) + (s.q[5] - gq(3, 5)) + (s.q[2] - gq(5, 2));
                    vn[0][ai] = s.v[0] + tau * (s.p[0] + dqx);
                    vn[1][ai] = s.v[1] + tau * (s.p[1] + dqy);
                    vn[2][ai] = s.v[2] + tau * (s.p[2] + dqz);
                }
            } else {
                // ---- exact-adjoint primal u and v (no reference counterpart):
                // self terms only where the +a neighbour exists, incoming terms
                // from every active cube whose +a neighbour this one is, and the
                // 1/2 of the off-diagonal eps entries in the v divergence.
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
                for (long ai = 0; ai < na; ++ai) {
                    const uint32_t i = L.active[ai];
                    const State<T>& s = st[i];
                    T dp = T(0);
                    T dq[3] = {T(0), T(0), T(0)};
                    for (int a = 0; a < 3; ++a) {
                        if (L.nba[ai * 6 + 2 * a] < 0) continue;
                        dp += s.p[a];
                        for (int c = 0; c < 3; ++c) {
                            const T w = (c == a) ? T(1) : T(0.5);
                            dq[c] += w * s.q[QI[c][a]];
                        }
                    }
                    for (uint32_t e = L.in.start[ai]; e < L.in.start[ai + 1]; ++e) {
                        const State<T>& o = st[L.active[L.in.ent[e].first]];
                        const int a = L.in.ent[e].second;
                        dp -= o.p[a];
                        for (int c = 0; c < 3; ++c) {
                            const T w = (c == a) ? T(1) : T(0.5);
                            dq[c] -= w * o.q[QI[c][a]];
                        }
                    }
                    un[ai] = prox_data(s.u + tau * dp, tl, bins + (size_t)i * nb, nb);
                    for (int c = 0; c < 3; ++c) vn[c][ai] = s.v[c] + tau * (s.p[c] + dq[c]);
                }
            }
            // ---- extrapolation + commit (tgv.cpp:220-230)
#pragma omp parallel for schedule(static) num_threads(nt) if (nt > 1)
            for (long ai = 0; ai < na; ++ai) {
                const uint32_t i = L.active[ai];
                ub[i] = T(2) * un[ai] - st[i].u;
                for (int c = 0; c < 3; ++c) vb[c][i] = T(2) * vn[c][ai] - st[i].v[c];
                st[i].u = un[ai];
                for (int c = 0; c < 3; ++c) st[i].v[c] = vn[c][ai];
            }
        }
    }

    void gap(const Level& L, double& P, double& D) {
        level_gap<T>(L, st, ub, vb, bins, nb, prm, 2.0, P, D);
    }

    // Iterate one level: fixed prm.iters, or in chunks until the relative gap
    // (P - D) / max(1, |P|) reaches prm.gap_tol or prm.max_iters is reached.
    void run(const Level& L, const T* fu, const T* const* fv, uint32_t lf, uint32_t ll) {
        const size_t na = L.active.size();
        std::vector<T> un(na), vn[3];
        for (int c = 0; c < 3; ++c) vn[c].resize(na);
        int done = 0;
        double P = 0, D = 0;
        if (prm.gap_tol <= 0) {
            iterate(L, prm.iters, un, vn, fu, fv, lf, ll);
            done = prm.iters;
            gap(L, P, D);
        } else {
            gap(L, P, D);
            while (done < prm.max_iters && (P - D) / std::max(1.0, std::fabs(P)) > prm.gap_tol) {
                const int k = std::min(prm.chunk, prm.max_iters - done);
                iterate(L, k, un, vn, fu, fv, lf, ll);
                done += k;
                gap(L, P, D);
            }
        }
        level_iters.push_back(done);
        level_P.push_back(P);
        level_D.push_back(D);
    }

    // tgv.cpp:356-419, with an optional per-cube initial (u, v) in place of the
    // median / parent-copy initialisation (the warm start).
    void minimize(const float* init_uv) {
        setup();
        for (int L = min_d; L <= max_d; ++L) {
            Level lv;
            lv.active = by_depth[L];
            if (lv.active.empty()) continue;
            const bool rec = prm.dual_eps > 0;
            const bool keep = rec && prm.warm_dual;  // dual from the cold init
            if (init_uv) {
                // warm start, kept only where it does not start from a larger gap
                if (prm.warm_fallback) {
                    if (L > min_d) parent_init(lv.active, L);
                    build_level(lv, cubes, cm, dm, N);
                    if (rec) dual_recover(lv, (T)prm.dual_eps);
                    double Pc, Dc;
                    gap(lv, Pc, Dc);
                    std::vector<State<T>> keepst(lv.active.size());
                    for (size_t k = 0; k < lv.active.size(); ++k) keepst[k] = st[lv.active[k]];
                    set_init(lv.active, init_uv, keep);
                    if (rec && !keep) dual_recover(lv, (T)prm.dual_eps);
                    double Pw, Dw;
                    gap(lv, Pw, Dw);
                    if (Pw - Dw > Pc - Dc) {  // back to the cold init
                        for (size_t k = 0; k < lv.active.size(); ++k) {
                            const uint32_t i = lv.active[k];
                            st[i] = keepst[k];
                            ub[i] = st[i].u;
                            for (int c = 0; c < 3; ++c) vb[c][i] = st[i].v[c];
                        }
                    }
                } else {
                    if (keep) {
                        if (L > min_d) parent_init(lv.active, L);
                        build_level(lv, cubes, cm, dm, N);
                        dual_recover(lv, (T)prm.dual_eps);
                    }
                    set_init(lv.active, init_uv, keep);
                    build_level(lv, cubes, cm, dm, N);
                    if (rec && !keep) dual_recover(lv, (T)prm.dual_eps);
                }
            } else {
                if (L > min_d) parent_init(lv.active, L);
                build_level(lv, cubes, cm, dm, N);
                if (rec) dual_recover(lv, (T)prm.dual_eps);
            }
            run(lv, nullptr, nullptr, 0, 0);
        }
    }

    // tgv.cpp:423-511: per treetop leaf, out-of-leaf u_bar / v_bar read from the
    // level-start snapshot.
    void minimize_oc(const uint32_t* leaf_first, const uint32_t* leaf_last, size_t nleaves) {
        setup();
        if (N == 0) return;
        for (int L = min_d; L <= max_d; ++L) {
            const std::vector<uint32_t>& all = by_depth[L];
            if (all.empty()) continue;
            if (L > min_d) parent_init(all, L);
            const std::vector<T> fu = ub;
            const std::vector<T> fvx = vb[0], fvy = vb[1], fvz = vb[2];
            const T* fv[3] = {fvx.data(), fvy.data(), fvz.data()};
            for (size_t l = 0; l < nleaves; ++l) {
                auto lo = std::lower_bound(all.begin(), all.end(), leaf_first[l]);
                auto hi = std::upper_bound(all.begin(), all.end(), leaf_last[l]);
                if (lo == hi) continue;
                Level lv;
                lv.active.assign(lo, hi);
                build_level(lv, cubes, cm, dm, N);
                run(lv, fu.data(), fv, leaf_first[l], leaf_last[l]);
            }
        }
    }
};

// ---------------------------------------------------------------- streaming
// tgv_minimize_streaming (tgv.cpp:515-1070) held in memory instead of files.
// Per treetop leaf: the leaf's own cubes are found by code with the last
// writer winning (the leaf map, tgv.cpp:916-920), every other cube by the
// cube file's bisection (CubeFileReader::lookup, tgv.cpp:674-683), which
// returns the first probe that hits, not necessarily the last cube with the
// code. Neighbour values: a cube iterated in the same leaf reads the live
// u_bar, v_bar; every other cube the (u, v) snapshot taken at level start;
// p, q of cubes outside the leaf read as zero (tgv.cpp:1023-1039). Parents
// are copied leaf by leaf from the bisection's parent (tgv.cpp:865-912).
template <class T>
struct Streaming {
    const Cube* cubes;
    size_t N;
    const uint32_t* bins;
    int nb;
    Params prm;
    std::vector<State<T>> st;

    int64_t bisect(uint64_t key) const {
        uint64_t lo = 0, hi = N;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2;
            const uint64_t k = code_of(cubes[mid]);
            if (k == key) return (int64_t)mid;
            if (k < key) lo = mid + 1;
            else hi = mid;
        }
        return -1;
    }

    void run(const uint32_t* leaf_first, const uint32_t* leaf_last, size_t nleaves) {
        st.assign(N, State<T>{});
        for (size_t i = 0; i < N; ++i) st[i].u = median_indicator<T>(bins + i * nb, nb);
        if (N == 0) return;
        int min_d = 21, max_d = 0;
        for (size_t i = 0; i < N; ++i) {
            min_d = std::min(min_d, cubes[i].depth);
            max_d = std::max(max_d, cubes[i].depth);
        }
        // per cube: its leaf, and its six neighbours as the leaf's table has them
        std::vector<int64_t> leaf_of(N, -1);
        std::vector<int64_t> nbr(N * 6, -1);
        for (size_t l = 0; l < nleaves; ++l) {
            const uint32_t f0 = leaf_first[l], f1 = std::min<uint64_t>(leaf_last[l], N - 1);
            if (f0 > f1) continue;
            CodeMap leaf_map;
            for (uint32_t i = f0; i <= f1; ++i) leaf_map[code_of(cubes[i])] = i;
            for (uint32_t i = f0; i <= f1; ++i) {
                leaf_of[i] = (int64_t)l;
                const uint64_t c = code_of(cubes[i]);
                const int d = cubes[i].depth;
                for (int f = 0; f < 6; ++f) {
                    const uint64_t k = face_nb(c, f, d);
                    if (k == NONE) continue;
                    auto it = leaf_map.find(k);
                    int64_t j = it != leaf_map.end() ? (int64_t)it->second : bisect(k);
                    if (j < 0 && d > 0) {
                        const uint64_t pk = parent(k, d);
                        auto pt = leaf_map.find(pk);
                        j = pt != leaf_map.end() ? (int64_t)pt->second : bisect(pk);
                    }
                    nbr[i * 6 + f] = j;
                }
            }
        }
        const T sigma = prm.sigma > 0.f ? (T)prm.sigma : T(0.25), tau = prm.tau > 0.f ? (T)prm.tau : T(0.25);
        const T a0 = (T)prm.alpha0, a1 = (T)prm.alpha1, tl = tau * (T)prm.lambda;
        std::vector<T> ub(N), vb[3] = {std::vector<T>(N), std::vector<T>(N), std::vector<T>(N)};
        std::vector<T> su(N), sv[3] = {std::vector<T>(N), std::vector<T>(N), std::vector<T>(N)};
        for (int L = min_d; L <= max_d; ++L) {
            // parent copy, leaf after leaf: each leaf reads its parents first
            if (L > min_d) {
                for (size_t l = 0; l < nleaves; ++l) {
                    const uint32_t f0 = leaf_first[l], f1 = std::min<uint64_t>(leaf_last[l], N - 1);
                    std::vector<uint32_t> lvl;
                    std::vector<int64_t> par;
                    for (uint32_t i = f0; i <= f1 && f0 <= f1; ++i)
                        if (cubes[i].depth == L) {
                            lvl.push_back(i);
                            par.push_back(bisect(parent(code_of(cubes[i]), L)));
                        }
                    std::vector<State<T>> ps(lvl.size());
                    for (size_t k = 0; k < lvl.size(); ++k)
                        if (par[k] >= 0) ps[k] = st[(size_t)par[k]];
                    for (size_t k = 0; k < lvl.size(); ++k) {
                        State<T>& s = st[lvl[k]];
                        if (par[k] >= 0) {
                            s.u = ps[k].u;
                            for (int c = 0; c < 3; ++c) s.v[c] = ps[k].v[c];
                        }
                        for (int c = 0; c < 3; ++c) s.p[c] = T(0);
                        for (int c = 0; c < 6; ++c) s.q[c] = T(0);
                    }
                }
            }
            // level start: snapshot and bars of every cube
            std::vector<uint32_t> act;
            for (size_t i = 0; i < N; ++i) {
                su[i] = ub[i] = st[i].u;
                for (int c = 0; c < 3; ++c) sv[c][i] = vb[c][i] = st[i].v[c];
                if (cubes[i].depth == L && leaf_of[i] >= 0) act.push_back((uint32_t)i);
            }
            if (act.empty()) continue;
            auto live = [&](uint32_t i, int64_t j) {
                return leaf_of[(size_t)j] == leaf_of[i] && cubes[(size_t)j].depth == L;
            };
            auto inside = [&](uint32_t i, int64_t j) { return leaf_of[(size_t)j] == leaf_of[i]; };
            const size_t na = act.size();
            std::vector<T> un(na), vn[3] = {std::vector<T>(na), std::vector<T>(na), std::vector<T>(na)};
            for (int it = 0; it < prm.iters; ++it) {
                for (size_t ai = 0; ai < na; ++ai) {  // dual p (tgv.cpp:738-748)
                    const uint32_t i = act[ai];
                    auto ubf = [&](int f) -> T {
                        const int64_t j = nbr[i * 6 + f];
                        if (j < 0) return ub[i];
                        return live(i, j) ? ub[(size_t)j] : su[(size_t)j];
                    };
                    st[i].p[0] += sigma * (ubf(0) - ub[i] - vb[0][i]);
                    st[i].p[1] += sigma * (ubf(2) - ub[i] - vb[1][i]);
                    st[i].p[2] += sigma * (ubf(4) - ub[i] - vb[2][i]);
                    clip3(st[i].p, a1);
                }
                for (size_t ai = 0; ai < na; ++ai) {  // dual q (tgv.cpp:751-767)
                    const uint32_t i = act[ai];
                    auto vbf = [&](int f, int c) -> T {
                        const int64_t j = nbr[i * 6 + f];
                        if (j < 0) return vb[c][i];
                        return live(i, j) ? vb[c][(size_t)j] : sv[c][(size_t)j];
                    };
                    T eps[6] = {vbf(0, 0) - vb[0][i],
                                vbf(2, 1) - vb[1][i],
                                vbf(4, 2) - vb[2][i],
                                T(0.5) * (vbf(2, 0) - vb[0][i] + vbf(0, 1) - vb[1][i]),
                                T(0.5) * (vbf(4, 0) - vb[0][i] + vbf(0, 2) - vb[2][i]),
                                T(0.5) * (vbf(4, 1) - vb[1][i] + vbf(2, 2) - vb[2][i])};
                    for (int k = 0; k < 6; ++k) st[i].q[k] += sigma * eps[k];
                    clip6(st[i].q, a0);
                }
                for (size_t ai = 0; ai < na; ++ai) {  // primal u (tgv.cpp:770-779)
                    const uint32_t i = act[ai];
                    auto gp = [&](int f, int c) -> T {
                        const int64_t j = nbr[i * 6 + f];
                        return (j >= 0 && inside(i, j)) ? st[(size_t)j].p[c] : T(0);
                    };
                    T div_p = (st[i].p[0] - gp(1, 0)) + (st[i].p[1] - gp(3, 1)) + (st[i].p[2] - gp(5, 2));
                    un[ai] = prox_data(st[i].u + tau * div_p, tl, bins + (size_t)i * nb, nb);
                }
                for (size_t ai = 0; ai < na; ++ai) {  // primal v (tgv.cpp:782-795)
                    const uint32_t i = act[ai];
                    auto gq = [&](int f, int c) -> T {
                        const int64_t j = nbr[i * 6 + f];
                        return (j >= 0 && inside(i, j)) ? st[(size_t)j].q[c] : T(0);
                    };
                    const State<T>& s = st[i];
                    T dqx = (s.q[0] - gq(1, 0)) + (s.q[3] - gq(3, 3)) + (s.q[4] - gq(5, 4));
                    T dqy = (s.q[3] - gq(1, 3)) + (s.q[1] - gq(3, 1)) + (s.q[5] - gq(5, 5));
                    T dqz = (s.q[4] - gq(1, 4)) + (s.q[5] - gq(3, 5)) + (s.q[2] - gq(5, 2));
                    vn[0][ai] = s.v[0] + tau * (s.p[0] + dqx);
                    vn[1][ai] = s.v[1] + tau * (s.p[1] + dqy);
                    vn[2][ai] = s.v[2] + tau * (s.p[2] + dqz);
                }
                for (size_t ai = 0; ai < na; ++ai) {  // extrapolation (tgv.cpp:798-810)
                    const uint32_t i = act[ai];
                    ub[i] = T(2) * un[ai] - st[i].u;
                    for (int c = 0; c < 3; ++c) vb[c][i] = T(2) * vn[c][ai] - st[i].v[c];
                    st[i].u = un[ai];
                    for (int c = 0; c < 3; ++c) st[i].v[c] = vn[c][ai];
                }
            }
        }
    }
};

// ---------------------------------------------------------------- surface
// marching_cubes.cpp:16-181: one quad per face-adjacent pair with a sign
// change in u, at the interpolated zero crossing between the cube centres,
// vertices merged on a 20-bit-per-axis quantised key.
struct Mesh {
    std::vector<float> verts;
    std::vector<uint32_t> faces;
};

static void corner(const Cube& c, const float o[3], float r_root, float out[3]) {
    uint32_t fx, fy, fz;
    decode(code_of(c), fx, fy, fz);
    const int shift = 21 - c.depth;
    const float cs = 2.0f * r_root / (float)(1u << c.depth);
    out[0] = o[0] + (float)(fx >> shift) * cs;
    out[1] = o[1] + (float)(fy >> shift) * cs;
    out[2] = o[2] + (float)(fz >> shift) * cs;
}
static uint32_t quant(float v, float o, float inv) {
    int q = (int)((v - o) * inv + 0.5f);
    if (q < 0) q = 0;
    if (q > (1 << 20) - 1) q = (1 << 20) - 1;
    return (uint32_t)q;
}

static Mesh extract_surface(const Cube* cubes, size_t N, const float* u, const uint8_t* has_data,
                            const float o[3], float r_root) {
    const float inv = (float)(1u << 20) / (2.0f * r_root);
    CodeMap cm;
    cm.reserve(N * 2);
    for (size_t i = 0; i < N; ++i) cm[code_of(cubes[i])] = (uint32_t)i;
    std::unordered_set<uint64_t> seen;
    std::unordered_map<uint64_t, uint32_t> vmap;
    Mesh m;
    auto vid = [&](const float p[3]) -> uint32_t {
        const uint64_t k = (uint64_t)quant(p[0], o[0], inv) | ((uint64_t)quant(p[1], o[1], inv) << 20) |
                           ((uint64_t)quant(p[2], o[2], inv) << 40);
        auto r = vmap.emplace(k, (uint32_t)(m.verts.size() / 3));
        if (r.second) m.verts.insert(m.verts.end(), p, p + 3);
        return r.first->second;
    };
    for (uint32_t i = 0; i < N; ++i) {
        const Cube& ci = cubes[i];
        for (int f = 0; f < 6; ++f) {
            const uint64_t nb = face_nb(code_of(ci), f, ci.depth);
            if (nb == NONE) continue;
            int32_t j = -1;
            auto it = cm.find(nb);
            if (it != cm.end()) j = (int32_t)it->second;
            if (j < 0 && ci.depth > 0) {
                it = cm.find(parent(nb, ci.depth));
                if (it != cm.end()) j = (int32_t)it->second;
            }
            if (j < 0) continue;
            const uint32_t ju = (uint32_t)j;
            if (ju == i) continue;
            const uint32_t lo = std::min(i, ju), hi = std::max(i, ju);
            if (!seen.insert(((uint64_t)lo << 32) | hi).second) continue;
            const float ui = u[i], uj = u[ju];
            if (has_data && !(has_data[i] && has_data[ju])) continue;
            if ((ui >= 0.0f) == (uj >= 0.0f)) continue;
            const float csi = 2.0f * r_root / (float)(1u << ci.depth);
            const float csj = 2.0f * r_root / (float)(1u << cubes[ju].depth);
            const float csm = csi < csj ? csi : csj;
            const float h = csm * 0.5f;
            float pi[3], pj[3];
            corner(ci, o, r_root, pi);
            corner(cubes[ju], o, r_root, pj);
            const float t = ui / (ui - uj);
            const int ax = f / 2;
            float s[3];
            s[ax] = (pi[ax] + csi * 0.5f) + t * ((pj[ax] + csj * 0.5f) - (pi[ax] + csi * 0.5f));
            const float* sc = (csi <= csj) ? pi : pj;
            float a1[3] = {0, 0, 0}, a2[3] = {0, 0, 0};
            switch (ax) {
                case 0: s[1] = sc[1] + h; s[2] = sc[2] + h; a1[1] = h; a2[2] = h; break;
                case 1: s[0] = sc[0] + h; s[2] = sc[2] + h; a1[2] = h; a2[0] = h; break;
                default: s[0] = sc[0] + h; s[1] = sc[1] + h; a1[0] = h; a2[1] = h; break;
            }
            float v0[3], v1[3], v2[3], v3[3];
            for (int k = 0; k < 3; ++k) {
                v0[k] = s[k] - a1[k] - a2[k];
                v1[k] = s[k] + a1[k] - a2[k];
                v2[k] = s[k] + a1[k] + a2[k];
                v3[k] = s[k] - a1[k] + a2[k];
            }
            const uint32_t i0 = vid(v0), i1 = vid(v1), i2 = vid(v2), i3 = vid(v3);
            const bool flip = (f % 2 == 0) ? (ui > 0.0f) : (ui < 0.0f);
            if (!flip) {
                const uint32_t t6[6] = {i0, i1, i2, i0, i2, i3};
                m.faces.insert(m.faces.end(), t6, t6 + 6);
            } else {
                const uint32_t t6[6] = {i0, i2, i1, i0, i3, i2};
                m.faces.insert(m.faces.end(), t6, t6 + 6);
            }
        }
    }
    return m;
}

}  // namespace orc

// ====================================================================== C API
template <class T>
static void copy_out(orc::Solver<T>& s, void* state_out, void* ubar_out, int32_t* level_iters,
                     double* level_pd, int32_t max_levels) {
    T* so = (T*)state_out;
    for (size_t i = 0; i < s.N; ++i) {
        const orc::State<T>& x = s.st[i];
        T* r = so + i * 13;
        r[0] = x.u;
        for (int c = 0; c < 3; ++c) r[1 + c] = x.v[c];
        for (int c = 0; c < 3; ++c) r[4 + c] = x.p[c];
        for (int k = 0; k < 6; ++k) r[7 + k] = x.q[k];
    }
    if (ubar_out) {
        T* b = (T*)ubar_out;
        for (size_t i = 0; i < s.N; ++i) {
            b[i * 4] = s.ub[i];
            for (int c = 0; c < 3; ++c) b[i * 4 + 1 + c] = s.vb[c][i];
        }
    }
    for (size_t l = 0; l < s.level_iters.size() && (int32_t)l < max_levels; ++l) {
        if (level_iters) level_iters[l] = s.level_iters[l];
        if (level_pd) { level_pd[2 * l] = s.level_P[l]; level_pd[2 * l + 1] = s.level_D[l]; }
    }
}

extern "C" {

typedef struct {
    float alpha0, alpha1, lambda;
    int32_t iters, op_mode, chunk, max_iters;
    double gap_tol;
    int32_t threads, warm_fallback;
    float dual_eps;
    int32_t warm_dual;
    float sigma, tau;
    float sigma_depth[22], tau_depth[22];
} orc_params;

void orc_morton_encode(uint32_t x, uint32_t y, uint32_t z, uint32_t* hi, uint32_t* lo) {
    const uint64_t c = orc::encode(x, y, z);
    *hi = (uint32_t)(c >> 32);
    *lo = (uint32_t)c;
}
void orc_morton_decode(uint32_t hi, uint32_t lo, uint32_t* x, uint32_t* y, uint32_t* z) {
    orc::decode(((uint64_t)hi << 32) | lo, *x, *y, *z);
}
uint64_t orc_parent_code(uint64_t code, int32_t depth) { return orc::parent(code, depth); }
uint64_t orc_face_neighbor(uint64_t code, int32_t face, int32_t depth) {
    return orc::face_nb(code, face, depth);
}
float orc_median_indicator(const uint32_t* bins, int32_t nb) { return orc::median_indicator<float>(bins, nb); }
double orc_median_indicator_f64(const uint32_t* bins, int32_t nb) {
    return orc::median_indicator<double>(bins, nb);
}
float orc_mean_indicator(const uint32_t* bins, int32_t nb) { return orc::mean_indicator<float>(bins, nb); }

// Mip pyramid of one view: writes every level, concatenated, into out (caller
// sizes it as width*height*4/3+levels); dims gets (w,h) per level. Returns the
// number of levels.
int32_t orc_mip_build(const orc::View* v, float* out, int32_t* dims) {
    orc::Mip mp = orc::build_mip(*v);
    size_t off = 0;
    for (size_t l = 0; l < mp.lv.size(); ++l) {
        std::memcpy(out + off, mp.lv[l].data(), mp.lv[l].size() * sizeof(float));
        off += mp.lv[l].size();
        dims[2 * l] = mp.w[l];
        dims[2 * l + 1] = mp.h[l];
    }
    return (int32_t)mp.lv.size();
}

int32_t orc_project(const orc::View* v, const float* P, float r_c, int32_t* out4, float* cz) {
    orc::Mip mp = orc::build_mip(*v);
    int l, u, w;
    float z;
    if (!orc::project(mp, P, r_c, l, u, w, z)) return 0;
    out4[0] = l; out4[1] = u; out4[2] = w;
    *cz = z;
    return 1;
}

void orc_histograms(const orc::Cube* cubes, uint64_t n, const orc::View* views, int32_t nv,
                    const float* origin, float r_root, int32_t nb, uint32_t* bins) {
    std::vector<orc::Mip> mips;
    for (int i = 0; i < nv; ++i) mips.push_back(orc::build_mip(views[i]));
    orc::compute_histograms(cubes, n, mips, origin, r_root, nb, bins);
}

static orc::Params to_params(const orc_params* p) {
    orc::Params q;
    q.alpha0 = p->alpha0; q.alpha1 = p->alpha1; q.lambda = p->lambda;
    q.iters = p->iters; q.op_mode = p->op_mode; q.chunk = p->chunk > 0 ? p->chunk : 50;
    q.max_iters = p->max_iters; q.gap_tol = p->gap_tol;
    q.threads = p->threads; q.warm_fallback = p->warm_fallback; q.dual_eps = p->dual_eps; q.warm_dual = p->warm_dual;
    q.sigma = p->sigma; q.tau = p->tau;
    for (int d = 0; d < 22; ++d) {
        q.sigma_depth[d] = p->sigma_depth[d];
        q.tau_depth[d] = p->tau_depth[d];
    }
    return q;
}

// tgv_minimize in FP32 (is_f64 = 0) or FP64 (is_f64 = 1). state_out holds 13
// values per cube in TGVCube order (u, v[3], p[3], q[6]) of the chosen type;
// ubar_out (optional) 4 per cube (u_bar, v_bar[3]). Returns the number of levels.
int32_t orc_tgv_minimize(const orc::Cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                         const orc_params* p, int32_t is_f64, const float* init_uv, void* state_out,
                         void* ubar_out, int32_t* level_iters, double* level_pd, int32_t max_levels) {
    if (is_f64) {
        orc::Solver<double> s{cubes, n, bins, nb, to_params(p)};
        s.minimize(init_uv);
        copy_out(s, state_out, ubar_out, level_iters, level_pd, max_levels);
        return (int32_t)s.level_iters.size();
    }
    orc::Solver<float> s{cubes, n, bins, nb, to_params(p)};
    s.minimize(init_uv);
    copy_out(s, state_out, ubar_out, level_iters, level_pd, max_levels);
    return (int32_t)s.level_iters.size();
}

// tgv_minimize_oc; leaves given as inclusive [first, last] index ranges.
int32_t orc_tgv_minimize_oc(const orc::Cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                            const uint32_t* leaf_first, const uint32_t* leaf_last, uint64_t nleaves,
                            const orc_params* p, int32_t is_f64, void* state_out) {
    if (is_f64) {
        orc::Solver<double> s{cubes, n, bins, nb, to_params(p)};
        s.minimize_oc(leaf_first, leaf_last, nleaves);
        copy_out(s, state_out, nullptr, nullptr, nullptr, 0);
        return (int32_t)s.level_iters.size();
    }
    orc::Solver<float> s{cubes, n, bins, nb, to_params(p)};
    s.minimize_oc(leaf_first, leaf_last, nleaves);
    copy_out(s, state_out, nullptr, nullptr, nullptr, 0);
    return (int32_t)s.level_iters.size();
}

// tgv_minimize_streaming on arrays (cubes sorted by code, as the cube file);
// leaves as inclusive [first, last] index ranges. state_out: 13 values per
// cube of the chosen type.
int32_t orc_tgv_minimize_streaming(const orc::Cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                                   const uint32_t* leaf_first, const uint32_t* leaf_last, uint64_t nleaves,
                                   const orc_params* p, int32_t is_f64, void* state_out) {
    auto out = [&](auto& s, auto* dst) {
        for (size_t i = 0; i < s.N; ++i) {
            const auto& x = s.st[i];
            auto* r = dst + i * 13;
            r[0] = x.u;
            for (int c = 0; c < 3; ++c) r[1 + c] = x.v[c];
            for (int c = 0; c < 3; ++c) r[4 + c] = x.p[c];
            for (int c = 0; c < 6; ++c) r[7 + c] = x.q[c];
        }
    };
    if (is_f64) {
        orc::Streaming<double> s{cubes, n, bins, nb, to_params(p), {}};
        s.run(leaf_first, leaf_last, nleaves);
        out(s, (double*)state_out);
    } else {
        orc::Streaming<float> s{cubes, n, bins, nb, to_params(p), {}};
        s.run(leaf_first, leaf_last, nleaves);
        out(s, (float*)state_out);
    }
    return 0;
}

// Energy / gap of one level for a given full state (FP64 in, FP64 out): the
// level's active set is every cube with depth == level. state: 13 doubles per
// cube; ubar: 4 doubles per cube (frozen values of inactive neighbours).
void orc_level_gap(const orc::Cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                   const orc_params* p, int32_t level, const double* state, const double* ubar,
                   double* P, double* D) {
    orc::CodeMap cm;
    orc::DepthCodeMap dm;
    for (size_t i = 0; i < n; ++i) {
        cm[orc::code_of(cubes[i])] = (uint32_t)i;
        dm[{orc::code_of(cubes[i]), (int)cubes[i].depth}] = (uint32_t)i;
    }
    orc::Level L;
    for (size_t i = 0; i < n; ++i)
        if (cubes[i].depth == level) L.active.push_back((uint32_t)i);
    orc::build_level(L, cubes, cm, dm, n);
    std::vector<orc::State<double>> st(n);
    std::vector<double> ub(n), vb[3];
    for (int c = 0; c < 3; ++c) vb[c].resize(n);
    for (size_t i = 0; i < n; ++i) {
        const double* r = state + i * 13;
        st[i].u = r[0];
        for (int c = 0; c < 3; ++c) st[i].v[c] = r[1 + c];
        for (int c = 0; c < 3; ++c) st[i].p[c] = r[4 + c];
        for (int k = 0; k < 6; ++k) st[i].q[k] = r[7 + k];
        ub[i] = ubar[i * 4];
        for (int c = 0; c < 3; ++c) vb[c][i] = ubar[i * 4 + 1 + c];
    }
    orc::level_gap<double>(L, st, ub, vb, bins, nb, to_params(p), 2.0, *P, *D);
}

// Neighbour table of the level's active set (6 per active cube, global
// indices or -1), in ascending global index order: the reference's table
// (mode 0) or the depth-aware table of the exact-adjoint mode (mode 1).
// Returns the active count.
uint64_t orc_neighbour_table(const orc::Cube* cubes, uint64_t n, int32_t level, int32_t mode, int32_t* nb_out,
                             uint32_t* active_out) {
    orc::CodeMap cm;
    orc::DepthCodeMap dm;
    for (size_t i = 0; i < n; ++i) {
        cm[orc::code_of(cubes[i])] = (uint32_t)i;
        dm[{orc::code_of(cubes[i]), (int)cubes[i].depth}] = (uint32_t)i;
    }
    orc::Level L;
    for (size_t i = 0; i < n; ++i)
        if (cubes[i].depth == level) L.active.push_back((uint32_t)i);
    orc::build_level(L, cubes, cm, dm, n);
    const std::vector<int32_t>& t = mode == 0 ? L.nb : L.nba;
    for (size_t k = 0; k < L.active.size(); ++k) {
        if (active_out) active_out[k] = L.active[k];
        for (int f = 0; f < 6; ++f) nb_out[k * 6 + f] = t[k * 6 + f];
    }
    return L.active.size();
}

// Surface of the u = 0 level set. Returns the vertex count; verts / faces are
// allocated with malloc and must be released with orc_free.
uint64_t orc_extract_surface(const orc::Cube* cubes, uint64_t n, const float* u, const uint8_t* has_data,
                             const float* origin, float r_root, float** verts, uint32_t** faces,
                             uint64_t* nfaces) {
    orc::Mesh m = orc::extract_surface(cubes, n, u, has_data, origin, r_root);
    *verts = (float*)std::malloc(std::max<size_t>(1, m.verts.size()) * sizeof(float));
    *faces = (uint32_t*)std::malloc(std::max<size_t>(1, m.faces.size()) * sizeof(uint32_t));
    std::memcpy(*verts, m.verts.data(), m.verts.size() * sizeof(float));
    std::memcpy(*faces, m.faces.data(), m.faces.size() * sizeof(uint32_t));
    *nfaces = m.faces.size() / 3;
    return m.verts.size() / 3;
}
void orc_free(void* p) { std::free(p); }

// Truncated float log2 as the histogram's level choice computes it
// (histogram.cpp:98-100); exposed so the exact-threshold rule of the GPU path
// can be compared against the C library over every float in a range.
int32_t orc_log2_level(float footprint) { return std::max(0, (int)std::log2(std::max(1.0f, footprint))); }

// Count of floats in [lo, hi) where trunc(log2f(x)) differs from the rule
// "e = exponent(x); e + (x >= thr[e])". Used by a test to pin the rule.
uint64_t orc_log2_rule_mismatches(float lo, float hi, const float* thr, int32_t nthr) {
    uint64_t bad = 0;
    for (float x = lo; x < hi; x = std::nextafter(x, hi)) {
        int e;
        std::frexp(x, &e);
        e -= 1;
        int rule = e;
        if (e < nthr && x >= thr[e]) rule = e + 1;
        if (rule != (int)std::log2(x)) ++bad;
    }
    return bad;
}

// CPU baseline: times `iters` iterations of the level at `depth` of the
// reference restatement (AoS state, active[] indirection, five passes, the
// median recomputed inside prox_data) with `threads` OpenMP threads.
// Level set-up (code map, neighbour table) is excluded. Returns seconds and
// the number of active cubes.
double orc_time_level(const orc::Cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb, const orc_params* p,
                      int32_t depth, int32_t iters, uint64_t* n_active) {
    orc::Solver<float> s{cubes, n, bins, nb, to_params(p)};
    s.setup();
    orc::Level lv;
    if (depth <= s.max_d) lv.active = s.by_depth[depth];
    *n_active = lv.active.size();
    if (lv.active.empty()) return 0.0;
    orc::build_level(lv, cubes, s.cm, s.dm, n);
    const size_t na = lv.active.size();
    std::vector<float> un(na), vn[3];
    for (int c = 0; c < 3; ++c) vn[c].resize(na);
    const auto t0 = std::chrono::steady_clock::now();
    s.iterate(lv, iters, un, vn, nullptr, nullptr, 0, 0);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

int32_t orc_sizeof_cube() { return (int32_t)sizeof(orc::Cube); }
int32_t orc_sizeof_view() { return (int32_t)sizeof(orc::View); }
int32_t orc_sizeof_params() { return (int32_t)sizeof(orc_params); }

}  // extern "C"
