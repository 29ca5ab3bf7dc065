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

#include "balance.hpp"
#include "marching_cubes.hpp"
#include "merge.hpp"
#include "octree_builder.hpp"
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
    // two iterations per launch over tabulated tile sets, the frozen-border
    // solve's leaves included; results are the same bits
    q.fused = 6;
    // the reference evaluates no energy
    q.no_energy = 1;
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

// ---- file readers of the cube and treetop files (octree_builder.cpp:99-140,
// treetop.cpp:93-108)
std::vector<OctreeCube> load_cubes_range(const std::string& path, uint32_t first_idx, uint32_t last_idx,
                                         float& r_root_out) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open cube file: " + path);
    CubeFileHeader h;
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (std::memcmp(h.magic, "TGV1", 4) != 0) throw std::runtime_error("Bad magic in cube file: " + path);
    r_root_out = h.r_root;
    if (first_idx >= h.cube_count || first_idx > last_idx) return {};
    const uint32_t end = std::min<uint64_t>((uint64_t)last_idx + 1, h.cube_count);
    std::vector<OctreeCube> v(end - first_idx);
    in.seekg(sizeof(CubeFileHeader) + (std::streamoff)first_idx * sizeof(OctreeCube));
    in.read(reinterpret_cast<char*>(v.data()), v.size() * sizeof(OctreeCube));
    return v;
}

std::vector<OctreeCube> load_cubes(const std::string& path, float& r_root_out) {
    return load_cubes_range(path, 0, std::numeric_limits<uint32_t>::max() - 1, r_root_out);
}

void save_cubes(const std::string& path, const std::vector<OctreeCube>& cubes, float r_root) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open output: " + path);
    CubeFileHeader h;
    h.r_root = r_root;
    h.cube_count = (uint32_t)cubes.size();
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    if (!cubes.empty()) out.write(reinterpret_cast<const char*>(cubes.data()), cubes.size() * sizeof(OctreeCube));
}

namespace {
std::vector<OctreeCube> octree_result(uint64_t n) {
    std::vector<OctreeCube> v(n);
    check(tgvx_octree_copy(ctx(), reinterpret_cast<tgvx_cube*>(v.data())));
    return v;
}
}  // namespace

// ---- stages 1-3 (octree_builder.cpp:11-97, merge.cpp:7-97, balance.cpp:22-150)
int cube_depth_for_radius(float r_x, float r_root, int max_depth) {
    if (r_x <= 0.0f) return max_depth;
    int d = (int)std::round(std::log2(r_root / (r_x * 1.125f)));
    d = std::min(std::max(d, 0), max_depth);
    // the one depth whose cell size lies in [0.75 r_x, 1.5 r_x), three steps at most
    for (int step = 0; step < 3; ++step) {
        const float cell = r_root / (float)(1u << d);
        if (cell < 0.75f * r_x && d > 0) --d;
        else if (cell >= 1.5f * r_x && d < max_depth) ++d;
        else break;
    }
    return d;
}

void build_cubes_from_depthmap(const DepthMap& dm, const Vec3f& scene_origin, float r_root,
                               const std::string& out_path) {
    if ((int64_t)dm.depths.size() < (int64_t)dm.width * dm.height) throw std::runtime_error("depth map too small");
    tgvx_depth_view v;
    v.depths = dm.depths.data();
    v.width = dm.width;
    v.height = dm.height;
    v.fx = dm.K(0, 0);
    v.fy = dm.K(1, 1);
    v.cx = dm.K(0, 2);
    v.cy = dm.K(1, 2);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 4; ++c) v.cam_to_world[r * 4 + c] = dm.Rt(r, c);
    const float origin[3] = {scene_origin.x(), scene_origin.y(), scene_origin.z()};
    uint64_t n = 0;
    check(tgvx_octree_build(ctx(), &v, 1, origin, r_root, 21, 1, &n, nullptr));
    save_cubes(out_path, octree_result(n), r_root);
}

void kway_merge(const std::vector<std::string>& in_paths, const std::string& out_path) {
    std::vector<OctreeCube> all;
    float r_root = 0.0f;
    for (size_t i = 0; i < in_paths.size(); ++i) {
        float r = 0.0f;
        std::vector<OctreeCube> v = load_cubes(in_paths[i], r);
        if (i == 0) r_root = r;
        all.insert(all.end(), v.begin(), v.end());
    }
    uint64_t n = 0;
    check(tgvx_octree_merge(ctx(), reinterpret_cast<const tgvx_cube*>(all.data()), all.size(), &n, nullptr));
    save_cubes(out_path, octree_result(n), r_root);
}

void balance_octree(const std::string& in_path, const std::string& out_path) {
    float r_root = 0.0f;
    const std::vector<OctreeCube> in = load_cubes(in_path, r_root);
    uint64_t n = 0;
    check(tgvx_octree_balance(ctx(), reinterpret_cast<const tgvx_cube*>(in.data()), in.size(), r_root, &n, nullptr));
    save_cubes(out_path, octree_result(n), r_root);
}

MortonCode parent_code(const MortonCode& code, int current_depth) {
    const uint64_t c = ((uint64_t)code.hi << 32) | code.lo;
    const unsigned s = 3u * (21u - (unsigned)(current_depth - 1));
    const uint64_t p = s >= 64u ? 0ull : c & ~((1ull << s) - 1ull);
    return {(uint32_t)(p >> 32), (uint32_t)p};
}

MortonCode face_neighbor(const MortonCode& code, int face, int depth) {
    const MortonCode none{UINT32_MAX, UINT32_MAX};
    if (face < 0 || face > 5) return none;
    uint32_t a[3];
    morton_decode(code, a[0], a[1], a[2]);
    const uint32_t step = 1u << (21 - depth);
    uint32_t& x = a[face / 2];
    if (face % 2 == 0) {
        if (x + step >= (1u << 21)) return none;
        x += step;
    } else {
        if (x < step) return none;
        x -= step;
    }
    return morton_encode(a[0], a[1], a[2]);
}

std::vector<TreetopLeaf> load_treetop(const std::string& path, float& r_root_out) {
    struct Header {
        char magic[4];
        uint32_t version;
        float r_root;
        uint32_t leaf_count;
    } h{};
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open treetop: " + path);
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (std::memcmp(h.magic, "TGVT", 4) != 0) throw std::runtime_error("Bad treetop magic: " + path);
    r_root_out = h.r_root;
    std::vector<TreetopLeaf> v(h.leaf_count);
    if (h.leaf_count) in.read(reinterpret_cast<char*>(v.data()), h.leaf_count * sizeof(TreetopLeaf));
    return v;
}

std::vector<CubeHistogram> load_histograms_range(const std::string& path, uint32_t first_idx, uint32_t last_idx) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open: " + path);
    FileHeader h{};
    in.read(reinterpret_cast<char*>(&h), sizeof(h));
    if (std::memcmp(h.magic, "TGVH", 4) != 0) throw std::runtime_error("Bad histogram magic");
    if (first_idx >= h.count || first_idx > last_idx) return {};
    const uint32_t end = std::min<uint64_t>((uint64_t)last_idx + 1, h.count);
    std::vector<CubeHistogram> v(end - first_idx);
    in.seekg(sizeof(FileHeader) + (std::streamoff)first_idx * sizeof(CubeHistogram));
    in.read(reinterpret_cast<char*>(v.data()), v.size() * sizeof(CubeHistogram));
    return v;
}

// ---- histogram.cpp:222-256: leaf after leaf, appended in leaf order
void compute_histograms_oc(const std::string& balanced_cube_path, const std::vector<TreetopLeaf>& leaves,
                           const std::vector<DepthMipmap>& depth_maps, const Vec3f& scene_origin, float r_root,
                           const std::string& out_hist_path) {
    std::ofstream out(out_hist_path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open: " + out_hist_path);
    FileHeader h{{'T', 'G', 'V', 'H'}, 1u, leaves.empty() ? 0u : leaves.back().last_idx + 1};
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    float r_tmp = 0.0f;
    for (const TreetopLeaf& leaf : leaves) {
        const auto cubes = load_cubes_range(balanced_cube_path, leaf.first_idx, leaf.last_idx, r_tmp);
        const auto hists = compute_histograms(cubes, depth_maps, scene_origin, r_root);
        if (!hists.empty())
            out.write(reinterpret_cast<const char*>(hists.data()), hists.size() * sizeof(CubeHistogram));
    }
}

// ---- tgv.cpp:802-1070
void tgv_minimize_streaming(const std::string& balanced_cube_path, const std::string& hist_path,
                            const std::string& treetop_path, const Vec3f& /*origin*/, float /*r_root*/,
                            const std::string& out_state_path, const TGVParams& params) {
    float r_tt = 0.0f, r_c = 0.0f;
    const auto leaves = load_treetop(treetop_path, r_tt);
    const auto cubes = load_cubes(balanced_cube_path, r_c);
    std::vector<TGVCube> state(cubes.size());
    if (!cubes.empty()) {
        const auto hists = load_histograms_range(hist_path, 0, (uint32_t)cubes.size() - 1);
        if (hists.size() != cubes.size())
            throw std::runtime_error("tgv_minimize_streaming: the histogram file holds fewer records than cubes");
        const tgvx_params p = solve_params(params);
        check(tgvx_solve_streaming(ctx(), reinterpret_cast<const tgvx_cube*>(cubes.data()), cubes.size(),
                                   reinterpret_cast<const uint32_t*>(hists.data()), 8,
                                   reinterpret_cast<const tgvx_treetop_leaf*>(leaves.data()), leaves.size(), &p,
                                   state.data(), nullptr));
    }
    save_tgv(out_state_path, state);
}

// ---- marching_cubes.cpp:63-181 on the GPU
Mesh extract_surface(const std::vector<OctreeCube>& cubes, const std::vector<TGVCube>& state, const Vec3f& origin,
                     float r_root, const std::vector<bool>& has_data) {
    Mesh m;
    if (cubes.empty()) return m;
    if (state.size() != cubes.size()) throw std::runtime_error("extract_surface: one state per cube expected");
    std::vector<uint32_t> nobins(cubes.size() * 8, 0u);
    check(tgvx_problem_load(ctx(), reinterpret_cast<const tgvx_cube*>(cubes.data()), cubes.size(), nobins.data(), 8,
                            TGVX_FP32, TGVX_OP_REF));
    std::vector<float> u(cubes.size());
    for (size_t i = 0; i < cubes.size(); ++i) u[i] = state[i].u;
    std::vector<uint8_t> hd;
    if (!has_data.empty()) {
        if (has_data.size() != cubes.size()) throw std::runtime_error("extract_surface: has_data size mismatch");
        hd.resize(cubes.size());
        for (size_t i = 0; i < cubes.size(); ++i) hd[i] = has_data[i] ? 1 : 0;
    }
    const float o[3] = {origin.x(), origin.y(), origin.z()};
    uint64_t nv = 0, nt = 0;
    check(tgvx_surface_extract(ctx(), u.data(), hd.empty() ? nullptr : hd.data(), o, r_root, &nv, &nt));
    std::vector<float> v(3 * nv);
    m.faces.resize(3 * nt);
    check(tgvx_surface_copy(ctx(), v.data(), m.faces.data()));
    m.verts.resize(nv);
    for (uint64_t i = 0; i < nv; ++i) m.verts[i] = Vec3f(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    return m;
}
