// TreetopLeaf of the reference interface (src/treetop.hpp:8-13).
#pragma once
#include <cstdint>

#include "types.hpp"

struct TreetopLeaf {
    MortonCode code;
    int32_t depth;
    uint32_t first_idx;
    uint32_t last_idx;
};
static constexpr uint32_t N_CUBES_PER_TREETOP_LEAF = 1u << 24;

// Treetop file ('TGVT', src/treetop.hpp:15-28): leaves of the balanced cube file.
#include <string>
#include <vector>
std::vector<TreetopLeaf> load_treetop(const std::string& path, float& r_root_out);
