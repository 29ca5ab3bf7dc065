// 2:1 balance of a cube file (reference src/balance.hpp:1-18), on the GPU
// (tgvx_octree_balance): the bridge cubes the reference's FIFO queue adds,
// seeded in file order; output sorted by (code, depth).
#pragma once
#include <string>

#include "types.hpp"

// Code of the ancestor one level up (the cell's lower corner at depth - 1).
MortonCode parent_code(const MortonCode& code, int current_depth);

// Same-depth face neighbour, faces 0:+x 1:-x 2:+y 3:-y 4:+z 5:-z;
// {UINT32_MAX, UINT32_MAX} outside the grid.
MortonCode face_neighbor(const MortonCode& code, int face, int depth);

void balance_octree(const std::string& in_path, const std::string& out_path);
