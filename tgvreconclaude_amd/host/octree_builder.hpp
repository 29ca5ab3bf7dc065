// Cube file interface of the reference (src/octree_builder.hpp:14-47): the
// Morton-sorted OctreeCube file ('TGV1'), its readers, and the builder, which
// runs on the GPU (tgvx_octree_build, stage 1).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "depth_map.hpp"
#include "types.hpp"

struct CubeFileHeader {
    char magic[4] = {'T', 'G', 'V', '1'};
    uint32_t version = 1;
    float r_root = 0.0f;
    uint32_t cube_count = 0;
};

std::vector<OctreeCube> load_cubes(const std::string& path, float& r_root_out);
std::vector<OctreeCube> load_cubes_range(const std::string& path, uint32_t first_idx, uint32_t last_idx,
                                         float& r_root_out);

// One cube per valid pixel (depth from its sample radius, max depth 21),
// sorted by (code, depth), pixel order among equal pairs, written to out_path.
void build_cubes_from_depthmap(const DepthMap& dm, const Vec3f& scene_origin, float r_root,
                               const std::string& out_path);

// The depth d with 0.75 r_x <= r_root / 2^d < 1.5 r_x, clamped to [0, max_depth].
int cube_depth_for_radius(float r_x, float r_root, int max_depth = 21);

// Writes a cube file (header + records), as the builder, merge and balance do.
void save_cubes(const std::string& path, const std::vector<OctreeCube>& cubes, float r_root);
