// The reference's C++ hot-path interface (src/tgv.hpp, src/histogram.hpp)
// implemented over the tgvx C ABI: histograms and TGV solves run on the GPU.
// Errors reported by the library become std::runtime_error, as the
// reference's only error channel; the small per-histogram accessors and the
// file formats are plain host code.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <mutex>
#include <stdexcept>
#include <string>

#include "tgv.hpp"
#include "tgvx.h"

namespace {

tgvx_ctx* ctx() {
    static std::once_flag once;
    static tgvx_ctx* c = nullptr;
    static int rc = 0;
    std::call_once(once, [] { rc = tgvx_create(&c, 0); });
    if (rc != 0 || !c) throw std::runtime_error("tgvx: no usable GPU (tgvx_create failed with code " + std::to_string(rc) + ")");
    return c;
}

void check(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("tgvx: ") + tgvx_last_error(ctx()));
}

static_assert(sizeof(OctreeCube) == sizeof(tgvx_cube), "OctreeCube layout");
static_assert(sizeof(TGVCube) == sizeof(tgvx_state), "TGVCube layout");
static_assert(sizeof(TreetopLeaf) == sizeof(tgvx_treetop_leaf), "TreetopLeaf layout");
static_assert(sizeof(CubeHistogram) == 8 * sizeof(uint32_t), "CubeHistogram layout");

struct FileHeader {
    char magic[4];
    uint32_t version;
    uint32_t count;
};

template <class T>
void save_records(const std::string& path, const char* magic, const std::vector<T>& v) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open: " + path);
    FileHeader h{{magic[0], magic[1], magic[2], magic[3]}, 1u, (uint32_t)v.size()};
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    if (!v.empty()) out.write(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}

template <class T>
std::vector<T> load_records(const std::string& path, const char* magic) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open: " + path);
    FileHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (std::memcmp(h.magic, magic, 4) != 0) throw std::runtime_error("Bad magic in " + path);
    std::vector<T> v(h.count);
    if (h.count) in.read(reinterpret_cast<char*>(v.data()), h.count * sizeof(T));
    return v;
}

tgvx_params solve_params(const TGVParams& p) {
    tgvx_params q;
    tgvx_default_params(&q);
    q.alpha0 = p.alpha0;
    q.alpha1 = p.alpha1;
    q.iters = p.iters;
    q.precision = TGVX_FP32;
    q.op_mode = TGVX_OP_REF;
    return q;
}

}  // namespace

// ---- CubeHistogram (histogram.cpp:12-31)
float CubeHistogram::mean_indicator() const {
    const uint32_t tot = total();
    if (tot == 0) return 0.0f;
    float sum = 0.0f;
    for (int k = 0; k < 8; ++k) sum += bins[k] * (-1.0f + (k + 0.5f) / 4.0f);
    return sum / tot;
}

float CubeHistogram::median_indicator() const {
    const uint32_t tot = total();
    if (tot == 0) return 0.0f;
    uint32_t cum = 0;
    for (int k = 0; k < 8; ++k) {
        cum += bins[k];
        if (cum * 2 >= tot) return -1.0f + (k + 0.5f) / 4.0f;
    }
    return 1.0f;
}

// ---- DepthMipmap (histogram.cpp:35-113); the pyramid is built on the GPU
void DepthMipmap::build(const DepthMap& base, bool is_lidar) {
    lidar = is_lidar;
    // Rt is world <- camera; invert once in double (affine: R^-1, -R^-1 t)
    double R[3][3], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) R[r][c] = base.Rt(r, c);
        t[r] = base.Rt(r, 3);
    }
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                       R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                       R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
    if (det == 0.0) throw std::runtime_error("DepthMipmap::build: singular Rt");
    double I[3][3];
    I[0][0] = (R[1][1] * R[2][2] - R[1][2] * R[2][1]) / det;
    I[0][1] = (R[0][2] * R[2][1] - R[0][1] * R[2][2]) / det;
    I[0][2] = (R[0][1] * R[1][2] - R[0][2] * R[1][1]) / det;
    I[1][0] = (R[1][2] * R[2][0] - R[1][0] * R[2][2]) / det;
    I[1][1] = (R[0][0] * R[2][2] - R[0][2] * R[2][0]) / det;
    I[1][2] = (R[0][2] * R[1][0] - R[0][0] * R[1][2]) / det;
    I[2][0] = (R[1][0] * R[2][1] - R[1][1] * R[2][0]) / det;
    I[2][1] = (R[0][1] * R[2][0] - R[0][0] * R[2][1]) / det;
    I[2][2] = (R[0][0] * R[1][1] - R[0][1] * R[1][0]) / det;
    Rt_inv = Matrix4f::Identity();
    for (int r = 0; r < 3; ++r) {
        double tr = 0.0;
        for (int c = 0; c < 3; ++c) {
            Rt_inv(r, c) = (float)I[r][c];
            tr -= I[r][c] * t[c];
        }
        Rt_inv(r, 3) = (float)tr;
    }
    tgvx_view v{};
    v.depths = base.depths.data();
    v.width = base.width;
    v.height = base.height;
    v.fx = base.K(0, 0);
    v.fy = base.K(1, 1);
    v.cx = base.K(0, 2);
    v.cy = base.K(1, 2);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) v.world_to_cam[r * 4 + c] = Rt_inv(r, c);
    std::vector<float> pyr((size_t)base.width * base.height * 4 / 3 + 64);
    int32_t dims[2 * 16];
    int32_t nl = 0;
    check(tgvx_mip_build(ctx(), &v, pyr.data(), dims, &nl));
    levels.clear();
    levels.push_back(base);
    size_t off = (size_t)base.width * base.height;
    for (int l = 1; l < nl; ++l) {
        const DepthMap& prev = levels.back();
        DepthMap next;
        next.width = dims[2 * l];
        next.height = dims[2 * l + 1];
        next.K = prev.K;
        next.K(0, 0) *= 0.5f;
        next.K(1, 1) *= 0.5f;
        next.K(0, 2) = prev.K(0, 2) * 0.5f;
        next.K(1, 2) = prev.K(1, 2) * 0.5f;
        next.Rt = prev.Rt;
        next.depths.assign(pyr.begin() + off, pyr.begin() + off + (size_t)next.width * next.height);
        off += (size_t)next.width * next.height;
        levels.push_back(std::move(next));
    }
}

bool DepthMipmap::project(const Vec3f& P, float r_c, int& level_out, int& u_out, int& v_out, float& cam_z_out) const {
    if (levels.empty()) return false;
    const DepthMap& base = levels[0];
    const Matrix4f& m = Rt_inv;
    const float cx_ = ((m(0, 0) * P.x() + m(0, 1) * P.y()) + m(0, 2) * P.z()) + m(0, 3);
    const float cy_ = ((m(1, 0) * P.x() + m(1, 1) * P.y()) + m(1, 2) * P.z()) + m(1, 3);
    const float z = ((m(2, 0) * P.x() + m(2, 1) * P.y()) + m(2, 2) * P.z()) + m(2, 3);
    if (z <= 0.0f) return false;
    const float fx = base.K(0, 0), fy = base.K(1, 1), cx = base.K(0, 2), cy = base.K(1, 2);
    const float uf = cx_ / z * fx + cx, vf = cy_ / z * fy + cy;
    if (!(uf > -1.0f && uf < (float)base.width) || !(vf > -1.0f && vf < (float)base.height)) return false;
    const int u0 = (int)uf, v0 = (int)vf;
    cam_z_out = z;
    const float fp = std::max(1.0f, std::abs(fx) * r_c / z);
    int level = fp <= std::numeric_limits<float>::max() ? std::max(0, (int)std::log2(fp)) : (int)levels.size();
    level = std::min(level, (int)levels.size() - 1);
    level_out = level;
    u_out = std::clamp(u0 >> level, 0, levels[level].width - 1);
    v_out = std::clamp(v0 >> level, 0, levels[level].height - 1);
    return true;
}

float DepthMipmap::depth_at(int level, int u, int v) const {
    if (level < 0 || level >= (int)levels.size()) return std::numeric_limits<float>::quiet_NaN();
    const DepthMap& lm = levels[level];
    if (u < 0 || u >= lm.width || v < 0 || v >= lm.height) return std::numeric_limits<float>::quiet_NaN();
    return lm.depths[v * lm.width + u];
}

std::vector<CubeHistogram> compute_histograms(const std::vector<OctreeCube>& cubes,
                                              const std::vector<DepthMipmap>& depth_maps, const Vec3f& origin,
                                              float r_root) {
    std::vector<CubeHistogram> hists(cubes.size());
    if (cubes.empty()) return hists;
    std::vector<tgvx_view> views;
    for (const DepthMipmap& dm : depth_maps) {
        if (dm.levels.empty()) continue;
        const DepthMap& b = dm.levels[0];
        tgvx_view v{};
        v.depths = b.depths.data();
        v.width = b.width;
        v.height = b.height;
        v.fx = b.K(0, 0);
        v.fy = b.K(1, 1);
        v.cx = b.K(0, 2);
        v.cy = b.K(1, 2);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) v.world_to_cam[r * 4 + c] = dm.Rt_inv(r, c);
        v.lidar = dm.lidar ? 1 : 0;
        views.push_back(v);
    }
    const float o[3] = {origin.x(), origin.y(), origin.z()};
    check(tgvx_histograms(ctx(), reinterpret_cast<const tgvx_cube*>(cubes.data()), cubes.size(), views.data(),
                          (int32_t)views.size(), o, r_root, 8, reinterpret_cast<uint32_t*>(hists.data())));
    return hists;
}

void save_histograms(const std::string& path, const std::vector<CubeHistogram>& h) { save_records(path, "TGVH", h); }
std::vector<CubeHistogram> load_histograms(const std::string& path) { return load_records<CubeHistogram>(path, "TGVH"); }
void save_tgv(const std::string& path, const std::vector<TGVCube>& s) { save_records(path, "TGVS", s); }
std::vector<TGVCube> load_tgv(const std::string& path) { return load_records<TGVCube>(path, "TGVS"); }

std::vector<TGVCube> tgv_minimize(const std::vector<OctreeCube>& cubes, const std::vector<CubeHistogram>& hists,
                                  const Vec3f& /*origin*/, float /*r_root*/, const TGVParams& params) {
    std::vector<TGVCube> state(cubes.size());
    if (cubes.empty()) return state;
    if (hists.size() != cubes.size()) throw std::runtime_error("tgv_minimize: one histogram per cube expected");
    const tgvx_params p = solve_params(params);
    check(tgvx_solve(ctx(), reinterpret_cast<const tgvx_cube*>(cubes.data()), cubes.size(),
                     reinterpret_cast<const uint32_t*>(hists.data()), 8, &p, nullptr, state.data(), nullptr));
    return state;
}

std::vector<TGVCube> tgv_minimize_oc(const std::vector<OctreeCube>& cubes, const std::vector<CubeHistogram>& hists,
                                     const std::vector<TreetopLeaf>& leaves, const Vec3f& /*origin*/, float /*r_root*/,
                                     const TGVParams& params) {
    std::vector<TGVCube> state(cubes.size());
    if (cubes.empty()) return state;
    if (hists.size() != cubes.size()) throw std::runtime_error("tgv_minimize_oc: one histogram per cube expected");
    const tgvx_params p = solve_params(params);
    check(tgvx_solve_oc(ctx(), reinterpret_cast<const tgvx_cube*>(cubes.data()), cubes.size(),
                        reinterpret_cast<const uint32_t*>(hists.data()), 8,
                        reinterpret_cast<const tgvx_treetop_leaf*>(leaves.data()), leaves.size(), &p, state.data(),
                        nullptr));
    return state;
}
