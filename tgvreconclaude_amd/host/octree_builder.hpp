// Cube file interface of the reference (src/octree_builder.hpp:14-39): the
// balanced, Morton-sorted OctreeCube file ('TGV1') and its readers. The
// builder itself is not part of this library.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

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
