// K-way merge of sorted cube files (reference src/merge.hpp:7), on the GPU
// (tgvx_octree_merge): the files' cubes in file order, equal (code, depth)
// cells merged with r_c averaged (double sum in that order); r_root from the
// first file.
#pragma once
#include <string>
#include <vector>

void kway_merge(const std::vector<std::string>& in_paths, const std::string& out_path);
