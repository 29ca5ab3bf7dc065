// Histogram interface of the reference (src/histogram.hpp:10-71), running on
// the GPU through libtgvx. Signatures and member names are the reference's.
#pragma once
#include <string>
#include <vector>

#include "depth_map.hpp"
#include "treetop.hpp"
#include "types.hpp"

struct CubeHistogram {
    uint32_t bins[8] = {};
    uint32_t total() const {
        uint32_t s = 0;
        for (auto b : bins) s += b;
        return s;
    }
    float mean_indicator() const;
    float median_indicator() const;
};

struct DepthMipmap {
    std::vector<DepthMap> levels;  // levels[0] = original
    Matrix4f Rt_inv;               // camera <- world (inverted once, in double)
    bool lidar = false;
    void build(const DepthMap& base, bool is_lidar = false);
    bool project(const Vec3f& P, float r_c, int& level_out, int& u_out, int& v_out, float& cam_z_out) const;
    float depth_at(int level, int u, int v) const;
};

std::vector<CubeHistogram> compute_histograms(const std::vector<OctreeCube>& cubes,
                                              const std::vector<DepthMipmap>& depth_maps,
                                              const Vec3f& scene_origin, float r_root);

// histogram.hpp:55-61: histograms of the cubes of each treetop leaf of the
// cube file, written leaf after leaf to a 'TGVH' file
void compute_histograms_oc(const std::string& balanced_cube_path, const std::vector<TreetopLeaf>& leaves,
                           const std::vector<DepthMipmap>& depth_maps, const Vec3f& scene_origin, float r_root,
                           const std::string& out_hist_path);

void save_histograms(const std::string& path, const std::vector<CubeHistogram>& hists);
std::vector<CubeHistogram> load_histograms(const std::string& path);
std::vector<CubeHistogram> load_histograms_range(const std::string& path, uint32_t first_idx, uint32_t last_idx);
