// Synthetic-scene and octree-input generator (host C++, OpenMP).
//
// Renders seeded CSG unions into noisy pinhole depth maps and builds the
// balanced Morton octree from them the way the reference's stages 1-3 do
// (src/octree_builder.cpp:11-97, src/merge.cpp:7-97, src/balance.cpp:60-150),
// plus the treetop split of stage 4 (src/treetop.cpp:23-70). Ties the
// reference leaves to run order (its unstable sorts, heap pop order) are fixed
// here: cubes sort by (code, depth), and merged r_c averages are summed in
// (view, pixel) order in double.
#include "tgvx_scene.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <limits>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- Morton
// Magic-constant dilation: bit b of v moves to bit 3b (21 input bits).
inline uint64_t dilate3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
inline uint32_t compact3(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}
inline uint64_t morton(uint32_t x, uint32_t y, uint32_t z) {
    return dilate3(x) | (dilate3(y) << 1) | (dilate3(z) << 2);
}
inline void unmorton(uint64_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    x = compact3(c);
    y = compact3(c >> 1);
    z = compact3(c >> 2);
}
inline uint64_t ancestor(uint64_t c, int d) {
    const unsigned s = 3u * (21u - (unsigned)d);
    return s >= 64u ? 0 : (c & ~((1ull << s) - 1ull));
}
constexpr uint64_t NONE = ~0ull;
inline uint64_t face_nb(uint64_t c, int f, int d) {
    uint32_t x, y, z;
    unmorton(c, x, y, z);
    const uint32_t s = 1u << (21 - d), B = 1u << 21;
    uint32_t* a = f < 2 ? &x : (f < 4 ? &y : &z);
    if (f & 1) {
        if (*a < s) return NONE;
        *a -= s;
    } else {
        if (*a + s >= B) return NONE;
        *a += s;
    }
    return morton(x, y, z);
}
inline uint64_t code_of(const tgvx_cube& c) { return ((uint64_t)c.hi << 32) | c.lo; }

// ---------------------------------------------------------------- RNG
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline double u01(uint64_t h) { return ((h >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

// ---------------------------------------------------------------- geometry
struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 a) { return a * (1.0 / std::sqrt(dot(a, a))); }

// world -> local: local = R^T (p - c)
inline V3 to_local(const tgvs_prim& p, V3 w) {
    const V3 d = w - V3{p.center[0], p.center[1], p.center[2]};
    const float* R = p.rot;
    return {R[0] * d.x + R[3] * d.y + R[6] * d.z, R[1] * d.x + R[4] * d.y + R[7] * d.z,
            R[2] * d.x + R[5] * d.y + R[8] * d.z};
}
inline V3 dir_local(const tgvs_prim& p, V3 w) {
    const float* R = p.rot;
    return {R[0] * w.x + R[3] * w.y + R[6] * w.z, R[1] * w.x + R[4] * w.y + R[7] * w.z,
            R[2] * w.x + R[5] * w.y + R[8] * w.z};
}

// First hit t > 0 of the ray o + t d (|d| = 1) with one solid, or +inf.
double hit(const tgvs_prim& p, V3 o, V3 d) {
    const double INF = std::numeric_limits<double>::infinity();
    const V3 lo = to_local(p, o), ld = dir_local(p, d);
    if (p.kind == TGVS_SPHERE) {
        const double r = p.size[0];
        const double b = dot(lo, ld), c = dot(lo, lo) - r * r;
        const double disc = b * b - c;
        if (disc < 0) return INF;
        const double s = std::sqrt(disc);
        double t = -b - s;
        if (t <= 1e-9) t = -b + s;
        return t > 1e-9 ? t : INF;
    }
    if (p.kind == TGVS_BOX) {
        double t0 = -INF, t1 = INF;
        const double oo[3] = {lo.x, lo.y, lo.z}, dd[3] = {ld.x, ld.y, ld.z};
        for (int a = 0; a < 3; ++a) {
            const double h = p.size[a];
            if (std::fabs(dd[a]) < 1e-300) {
                if (oo[a] < -h || oo[a] > h) return INF;
                continue;
            }
            double ta = (-h - oo[a]) / dd[a], tb = (h - oo[a]) / dd[a];
            if (ta > tb) std::swap(ta, tb);
            t0 = std::max(t0, ta);
            t1 = std::min(t1, tb);
        }
        if (t0 > t1 || t1 <= 1e-9) return INF;
        return t0 > 1e-9 ? t0 : t1;
    }
    // torus (axis = local y): sphere tracing on the exact distance, entered at
    // the bounding sphere, then bisection on the sign of the implicit function.
    const double R = p.size[0], r = p.size[1];
    const double br = R + r;
    const double b = dot(lo, ld), c = dot(lo, lo) - br * br;
    const double disc = b * b - c;
    if (disc < 0) return INF;
    const double s = std::sqrt(disc);
    double t = std::max(0.0, -b - s);
    const double tend = -b + s;
    auto sdf = [&](double tt) {
        const V3 q = lo + ld * tt;
        const double a = std::sqrt(q.x * q.x + q.z * q.z) - R;
        return std::sqrt(a * a + q.y * q.y) - r;
    };
    for (int it = 0; it < 400 && t <= tend; ++it) {
        const double dist = sdf(t);
        if (dist < 1e-7) {
            // refine: step back until outside, then bisect
            double a = std::max(0.0, t - 1e-4), bb = t;
            if (sdf(a) <= 0) return t;
            for (int k = 0; k < 40; ++k) {
                const double m = 0.5 * (a + bb);
                if (sdf(m) > 0) a = m; else bb = m;
            }
            return bb;
        }
        t += dist;
    }
    return INF;
}

void make_camera(V3 C, V3 target, V3 up, int W, int H, float hfov_deg, tgvs_camera* out) {
    const V3 f = normalize(target - C);
    if (std::fabs(dot(f, normalize(up))) > 0.999) up = V3{0, 0, 1};
    V3 dn = normalize(up * -1.0 - f * dot(up * -1.0, f));  // camera +y (down)
    V3 rt = cross(dn, f);                                   // camera +x: y cross z
    const double fx = 0.5 * W / std::tan(0.5 * hfov_deg * M_PI / 180.0);
    out->fx = (float)fx;
    out->fy = (float)fx;
    out->cx = 0.5f * W;
    out->cy = 0.5f * H;
    // cam_to_world: columns rt, dn, f; translation C
    const double Rc[3][3] = {{rt.x, dn.x, f.x}, {rt.y, dn.y, f.y}, {rt.z, dn.z, f.z}};
    const double Cc[3] = {C.x, C.y, C.z};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) out->cam_to_world[r * 4 + c] = (float)Rc[r][c];
        out->cam_to_world[r * 4 + 3] = (float)Cc[r];
    }
    // world_to_cam = [R^T | -R^T C], from the rounded cam_to_world in double
    double Rf[3][3], tf[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Rf[r][c] = out->cam_to_world[r * 4 + c];
        tf[r] = out->cam_to_world[r * 4 + 3];
    }
    // general 3x3 inverse (rounded rotation is not exactly orthonormal)
    const double det = Rf[0][0] * (Rf[1][1] * Rf[2][2] - Rf[1][2] * Rf[2][1]) -
                       Rf[0][1] * (Rf[1][0] * Rf[2][2] - Rf[1][2] * Rf[2][0]) +
                       Rf[0][2] * (Rf[1][0] * Rf[2][1] - Rf[1][1] * Rf[2][0]);
    double Ri[3][3];
    Ri[0][0] = (Rf[1][1] * Rf[2][2] - Rf[1][2] * Rf[2][1]) / det;
    Ri[0][1] = (Rf[0][2] * Rf[2][1] - Rf[0][1] * Rf[2][2]) / det;
    Ri[0][2] = (Rf[0][1] * Rf[1][2] - Rf[0][2] * Rf[1][1]) / det;
    Ri[1][0] = (Rf[1][2] * Rf[2][0] - Rf[1][0] * Rf[2][2]) / det;
    Ri[1][1] = (Rf[0][0] * Rf[2][2] - Rf[0][2] * Rf[2][0]) / det;
    Ri[1][2] = (Rf[0][2] * Rf[1][0] - Rf[0][0] * Rf[1][2]) / det;
    Ri[2][0] = (Rf[1][0] * Rf[2][1] - Rf[1][1] * Rf[2][0]) / det;
    Ri[2][1] = (Rf[0][1] * Rf[2][0] - Rf[0][0] * Rf[2][1]) / det;
    Ri[2][2] = (Rf[0][0] * Rf[1][1] - Rf[0][1] * Rf[1][0]) / det;
    for (int r = 0; r < 3; ++r) {
        double t = 0;
        for (int c = 0; c < 3; ++c) {
            out->world_to_cam[r * 4 + c] = (float)Ri[r][c];
            t -= Ri[r][c] * tf[c];
        }
        out->world_to_cam[r * 4 + 3] = (float)t;
    }
}

// ---------------------------------------------------------------- octree
// DepthMap::unproject (depth_map.cpp:4-11) with the 3x4 cam_to_world, float.
struct Unproj {
    const float* d;
    int W, H;
    const tgvs_camera* c;
    bool valid(int u, int v) const {
        if (u < 0 || v < 0 || u >= W || v >= H) return false;
        const float x = d[(size_t)v * W + u];
        return std::isfinite(x) && x > 0.0f;
    }
    void at(int u, int v, float out[3]) const {
        const float z = d[(size_t)v * W + u];
        const float cx = (u - c->cx) * z / c->fx, cy = (v - c->cy) * z / c->fy;
        const float* M = c->cam_to_world;
        for (int r = 0; r < 3; ++r) out[r] = ((M[r * 4] * cx + M[r * 4 + 1] * cy) + M[r * 4 + 2] * z) + M[r * 4 + 3];
    }
    // depth_map.cpp:13-30
    float radius(int u, int v) const {
        if (!valid(u, v)) return 0.0f;
        float c[3];
        at(u, v, c);
        static const int du[4] = {1, -1, 0, 0}, dv[4] = {0, 0, 1, -1};
        float sum = 0.0f;
        int cnt = 0;
        for (int i = 0; i < 4; ++i) {
            const int nu = u + du[i], nv = v + dv[i];
            if (!valid(nu, nv)) continue;
            float p[3];
            at(nu, nv, p);
            const float dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
            sum += std::sqrt(dx * dx + dy * dy + dz * dz);
            ++cnt;
        }
        if (cnt == 0) return 0.0f;
        return 0.5f * sum / cnt;
    }
};

// octree_builder.cpp:11-28
int depth_for_radius(float r_x, float r_root, int max_depth) {
    if (r_x <= 0.0f) return max_depth;
    int d = (int)std::round(std::log2(r_root / (r_x * 1.125f)));
    if (d < 0) d = 0;
    if (d > max_depth) d = max_depth;
    for (int tries = 0; tries < 3; ++tries) {
        const float cell = r_root / (float)(1u << d);
        if (cell >= 0.75f * r_x && cell < 1.5f * r_x) break;
        if (cell < 0.75f * r_x) {
            if (d > 0) --d; else break;
        } else {
            if (d < max_depth) ++d; else break;
        }
    }
    return d;
}

struct Raw {
    uint64_t key;  // code
    int32_t depth;
    float r;
};

// Open-addressing set of 64-bit keys (key ~0 reserved).
struct KeySet {
    std::vector<uint64_t> slot;
    size_t mask = 0, count = 0;
    explicit KeySet(size_t expect = 16) {
        size_t cap = 16;
        while (cap < expect * 2) cap <<= 1;
        slot.assign(cap, NONE);
        mask = cap - 1;
    }
    static size_t h(uint64_t k) { return (size_t)mix64(k); }
    bool has(uint64_t k) const {
        for (size_t i = h(k) & mask;; i = (i + 1) & mask) {
            if (slot[i] == k) return true;
            if (slot[i] == NONE) return false;
        }
    }
    bool add(uint64_t k) {
        if ((count + 1) * 2 > slot.size()) grow();
        for (size_t i = h(k) & mask;; i = (i + 1) & mask) {
            if (slot[i] == k) return false;
            if (slot[i] == NONE) { slot[i] = k; ++count; return true; }
        }
    }
    void grow() {
        std::vector<uint64_t> old;
        old.swap(slot);
        slot.assign(old.size() * 2, NONE);
        mask = slot.size() - 1;
        count = 0;
        for (uint64_t k : old)
            if (k != NONE) add(k);
    }
};

std::vector<tgvx_cube> balance(const std::vector<tgvx_cube>& in, float r_root) {
    // balance.cpp:60-150, FIFO BFS seeded with the input cubes in order. Cells
    // are keyed by the full 63-bit code: the reference's dc_key
    // (balance.cpp:56-58) keeps code bits 0-59 only, which would merge cells
    // of different root octants and clear bits 60-62 of the output codes.
    std::vector<KeySet> by_depth;
    by_depth.reserve(22);
    std::vector<size_t> per(22, 0);
    for (const auto& c : in) per[c.depth]++;
    for (int d = 0; d < 22; ++d) by_depth.emplace_back(per[d] + 16);
    std::vector<tgvx_cube> out = in;
    for (const auto& c : in) by_depth[c.depth].add(code_of(c));
    std::deque<std::pair<int, uint64_t>> q;
    for (const auto& c : in) q.emplace_back(c.depth, code_of(c));
    while (!q.empty()) {
        const auto [d, cv] = q.front();
        q.pop_front();
        if (d < 2) continue;
        for (int f = 0; f < 6; ++f) {
            const uint64_t nb = face_nb(cv, f, d);
            if (nb == NONE) continue;
            if (by_depth[d].has(nb)) continue;
            const uint64_t par = ancestor(nb, d - 1);
            if (by_depth[d - 1].has(par)) continue;
            bool viol = false;
            for (int d2 = d - 2; d2 >= 0; --d2)
                if (by_depth[d2].has(ancestor(nb, d2))) { viol = true; break; }
            if (!viol) continue;
            if (by_depth[d - 1].add(par)) {
                tgvx_cube b;
                b.hi = (uint32_t)(par >> 32);
                b.lo = (uint32_t)par;
                b.depth = d - 1;
                b.r_c = r_root / (float)(1u << (d - 1));
                out.push_back(b);
                q.emplace_back(d - 1, par);
            }
        }
    }
    // by (code, depth); equal pairs (repeated input cells only: a bridge is
    // never an existing cell) keep their input order
    std::stable_sort(out.begin(), out.end(), [](const tgvx_cube& a, const tgvx_cube& b) {
        const uint64_t ka = code_of(a), kb = code_of(b);
        return ka != kb ? ka < kb : a.depth < b.depth;
    });
    return out;
}

tgvx_cube cube_of(const Raw& r) {
    tgvx_cube c;
    c.hi = (uint32_t)(r.key >> 32);
    c.lo = (uint32_t)r.key;
    c.depth = r.depth;
    c.r_c = r.r;
    return c;
}

// positions of `all` sorted by (code, depth), position order among equal pairs
std::vector<uint32_t> sorted_order(const std::vector<Raw>& all) {
    std::vector<uint32_t> ord(all.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (uint32_t)i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
        const Raw &x = all[a], &y = all[b];
        if (x.key != y.key) return x.key < y.key;
        if (x.depth != y.depth) return x.depth < y.depth;
        return a < b;
    });
    return ord;
}

// merge.cpp:57-92: one cell per run of equal (code, depth) in sorted order,
// r_c the mean of the run (double sum in that order)
std::vector<tgvx_cube> merge_sorted(const std::vector<Raw>& all, const std::vector<uint32_t>& ord) {
    std::vector<tgvx_cube> merged;
    merged.reserve(all.size() / 2 + 16);
    for (size_t i = 0; i < ord.size();) {
        size_t j = i;
        double s = 0;
        const Raw& r0 = all[ord[i]];
        while (j < ord.size() && all[ord[j]].key == r0.key && all[ord[j]].depth == r0.depth) {
            s += all[ord[j]].r;
            ++j;
        }
        tgvx_cube c = cube_of(r0);
        c.r_c = (float)(s / (double)(j - i));
        merged.push_back(c);
        i = j;
    }
    return merged;
}

void treetop_rec(const tgvx_cube* c, uint32_t first, uint32_t last, uint64_t node, int depth, uint32_t cap,
                 std::vector<tgvx_treetop_leaf>& out) {
    const uint32_t count = last - first;
    if (count == 0) return;
    if (count <= cap || depth >= 21) {
        out.push_back({(uint32_t)(node >> 32), (uint32_t)node, depth, first, last - 1});
        return;
    }
    const unsigned sh = depth <= 20 ? 3u * (20u - (unsigned)depth) : 0u;
    uint32_t cf = first;
    for (int q = 0; q < 8; ++q) {
        const uint64_t child = node | ((uint64_t)q << sh);
        const uint64_t cmax = child | (sh > 0 ? ((1ull << sh) - 1ull) : 0ull);
        uint32_t lo = cf, hi = last;
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (!(cmax < code_of(c[mid]))) lo = mid + 1; else hi = mid;
        }
        treetop_rec(c, cf, lo, child, depth + 1, cap, out);
        cf = lo;
        if (cf >= last) break;
    }
}

}  // namespace

extern "C" {

int32_t tgvs_ring_cameras(int32_t n, float radius, int32_t W, int32_t H, float hfov, tgvs_camera* out) {
    if (n <= 0 || !out) return -1;
    for (int k = 0; k < n; ++k) {
        const double t = 2.0 * M_PI * k / n;
        make_camera({radius * std::cos(t), 0.0, radius * std::sin(t)}, {0, 0, 0}, {0, 1, 0}, W, H, hfov, out + k);
    }
    return 0;
}

int32_t tgvs_random_cameras(int32_t n, float radius, int32_t W, int32_t H, float hfov, uint64_t seed,
                            tgvs_camera* out) {
    if (n <= 0 || !out) return -1;
    for (int k = 0; k < n; ++k) {
        const double a = u01(mix64(seed * 0x100000001b3ull + 2 * k));
        const double b = u01(mix64(seed * 0x100000001b3ull + 2 * k + 1));
        const double z = 2.0 * a - 1.0, phi = 2.0 * M_PI * b, s = std::sqrt(std::max(0.0, 1 - z * z));
        make_camera({radius * s * std::cos(phi), radius * z, radius * s * std::sin(phi)}, {0, 0, 0}, {0, 1, 0}, W, H,
                    hfov, out + k);
    }
    return 0;
}

int32_t tgvs_render(const tgvs_prim* prims, int32_t np, const tgvs_camera* cams, int32_t nc, int32_t W, int32_t H,
                    float sigma, float ofrac, float olo, float ohi, uint64_t seed, float* out) {
    if (!cams || !out || W <= 0 || H <= 0 || nc < 0 || np < 0) return -1;
    const float nanv = std::numeric_limits<float>::quiet_NaN();
#pragma omp parallel for schedule(dynamic, 4) collapse(2)
    for (int c = 0; c < nc; ++c)
        for (int v = 0; v < H; ++v) {
            const tgvs_camera& cam = cams[c];
            const float* M = cam.cam_to_world;
            const V3 o{M[3], M[7], M[11]};
            for (int u = 0; u < W; ++u) {
                // pixel (u, v) unprojects through ((u - cx)/fx, (v - cy)/fy, 1)
                const double dx = (u - cam.cx) / cam.fx, dy = (v - cam.cy) / cam.fy;
                const V3 dc{dx, dy, 1.0};
                const V3 dw{M[0] * dc.x + M[1] * dc.y + M[2] * dc.z, M[4] * dc.x + M[5] * dc.y + M[6] * dc.z,
                            M[8] * dc.x + M[9] * dc.y + M[10] * dc.z};
                const double len = std::sqrt(dot(dw, dw));
                const V3 dn = dw * (1.0 / len);
                double t = std::numeric_limits<double>::infinity();
                for (int k = 0; k < np; ++k) t = std::min(t, hit(prims[k], o, dn));
                const size_t pix = ((size_t)c * H + v) * W + u;
                const uint64_t h0 = mix64(seed ^ mix64(pix * 4 + 1));
                const uint64_t h1 = mix64(seed ^ mix64(pix * 4 + 2));
                const uint64_t h2 = mix64(seed ^ mix64(pix * 4 + 3));
                float depth = nanv;
                if (std::isfinite(t)) {
                    const double z = t / len;  // camera-frame z of the hit
                    const double g = std::sqrt(-2.0 * std::log(u01(h0))) * std::cos(2.0 * M_PI * u01(h1));
                    depth = (float)(z + sigma * g);
                }
                if (u01(h2) < ofrac) {
                    const uint64_t h3 = mix64(seed ^ mix64(pix * 4 + 4));
                    depth = (float)(olo + (ohi - olo) * u01(h3));
                }
                out[pix] = depth;
            }
        }
    return 0;
}

int64_t tgvs_build_octree(const float* depths, const tgvs_camera* cams, int32_t nc, int32_t W, int32_t H,
                          const float origin[3], float r_root, int32_t max_depth, int32_t last_stage,
                          tgvx_cube** out) {
    if (!depths || !cams || !out || max_depth < 0 || max_depth > 21 || last_stage < 1 || last_stage > 3) return -1;
    const float inv_2r = 1.0f / (2.0f * r_root);
    // stage 1 per view (octree_builder.cpp:30-97), then stage 2: one global
    // merge of equal (code, depth) cells.
    std::vector<std::vector<Raw>> per(nc);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nc; ++c) {
        Unproj U{depths + (size_t)c * W * H, W, H, cams + c};
        std::vector<Raw>& loc = per[c];
        for (int v = 0; v < H; ++v)
            for (int u = 0; u < W; ++u) {
                if (!U.valid(u, v)) continue;
                const float r_x = U.radius(u, v);
                if (r_x <= 0.0f) continue;
                float p[3];
                U.at(u, v, p);
                const float n0 = (p[0] - origin[0]) * inv_2r, n1 = (p[1] - origin[1]) * inv_2r,
                            n2 = (p[2] - origin[2]) * inv_2r;
                if (n0 < 0 || n0 >= 1 || n1 < 0 || n1 >= 1 || n2 < 0 || n2 >= 1) continue;
                const int d = depth_for_radius(r_x, r_root, max_depth);
                const uint32_t g = 1u << d;
                uint32_t ix = (uint32_t)(n0 * g), iy = (uint32_t)(n1 * g), iz = (uint32_t)(n2 * g);
                if (ix >= g) ix = g - 1;
                if (iy >= g) iy = g - 1;
                if (iz >= g) iz = g - 1;
                const int sh = 21 - d;
                loc.push_back({morton(ix << sh, iy << sh, iz << sh), d, r_x});
            }
        std::stable_sort(loc.begin(), loc.end(), [](const Raw& a, const Raw& b) {
            return a.key != b.key ? a.key < b.key : a.depth < b.depth;
        });
    }
    // merge: k-way over the sorted per-view runs, equal (code, depth) averaged
    // with a double sum in (view, pixel) order.
    size_t total = 0;
    for (auto& v : per) total += v.size();
    std::vector<Raw> all;
    all.reserve(total);
    std::vector<uint32_t> src;
    src.reserve(total);
    for (int c = 0; c < nc; ++c) {
        all.insert(all.end(), per[c].begin(), per[c].end());
        src.insert(src.end(), per[c].size(), (uint32_t)c);
        std::vector<Raw>().swap(per[c]);
    }
    const std::vector<uint32_t> ord = sorted_order(all);
    std::vector<tgvx_cube> res;
    if (last_stage == 1) {
        res.resize(ord.size());
        for (size_t i = 0; i < ord.size(); ++i) res[i] = cube_of(all[ord[i]]);
    } else {
        std::vector<tgvx_cube> merged = merge_sorted(all, ord);
        std::vector<Raw>().swap(all);
        res = last_stage == 2 ? std::move(merged) : balance(merged, r_root);
    }
    const std::vector<tgvx_cube>& bal = res;
    *out = (tgvx_cube*)std::malloc(std::max<size_t>(1, bal.size()) * sizeof(tgvx_cube));
    if (!*out) return -1;
    std::memcpy(*out, bal.data(), bal.size() * sizeof(tgvx_cube));
    return (int64_t)bal.size();
}

int64_t tgvs_merge(const tgvx_cube* cubes, uint64_t n, tgvx_cube** out) {
    if ((n && !cubes) || !out) return -1;
    std::vector<Raw> all(n);
    for (uint64_t i = 0; i < n; ++i) all[i] = {code_of(cubes[i]), cubes[i].depth, cubes[i].r_c};
    const std::vector<tgvx_cube> m = merge_sorted(all, sorted_order(all));
    *out = (tgvx_cube*)std::malloc(std::max<size_t>(1, m.size()) * sizeof(tgvx_cube));
    if (!*out) return -1;
    std::memcpy(*out, m.data(), m.size() * sizeof(tgvx_cube));
    return (int64_t)m.size();
}

int64_t tgvs_balance(const tgvx_cube* cubes, uint64_t n, float r_root, tgvx_cube** out) {
    std::vector<tgvx_cube> in(cubes, cubes + n);
    std::vector<tgvx_cube> bal = balance(in, r_root);
    *out = (tgvx_cube*)std::malloc(std::max<size_t>(1, bal.size()) * sizeof(tgvx_cube));
    if (!*out) return -1;
    std::memcpy(*out, bal.data(), bal.size() * sizeof(tgvx_cube));
    return (int64_t)bal.size();
}

int64_t tgvs_treetop(const tgvx_cube* cubes, uint64_t n, uint32_t cap, tgvx_treetop_leaf** out) {
    std::vector<tgvx_treetop_leaf> leaves;
    if (cap == 0) cap = 1u << 24;
    treetop_rec(cubes, 0, (uint32_t)n, 0ull, 0, cap, leaves);
    *out = (tgvx_treetop_leaf*)std::malloc(std::max<size_t>(1, leaves.size()) * sizeof(tgvx_treetop_leaf));
    if (!*out) return -1;
    std::memcpy(*out, leaves.data(), leaves.size() * sizeof(tgvx_treetop_leaf));
    return (int64_t)leaves.size();
}

void tgvs_free(void* p) { std::free(p); }

}  // extern "C"
