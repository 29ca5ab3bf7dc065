// Surface interface of the reference (src/marching_cubes.hpp:8-33): the u = 0
// iso-surface of a TGV state, extracted on the GPU. Same mesh as the
// reference, vertex for vertex.
#pragma once
#include <cstdint>
#include <vector>

#include "octree_builder.hpp"
#include "tgv.hpp"
#include "types.hpp"

struct Mesh {
    std::vector<Vec3f> verts;
    std::vector<uint32_t> faces;  // 3 indices per triangle
    std::vector<Vec3f> normals;   // not filled by extract_surface
    bool empty() const { return faces.empty(); }
    size_t tri_count() const { return faces.size() / 3; }
};

Mesh extract_surface(const std::vector<OctreeCube>& cubes, const std::vector<TGVCube>& state, const Vec3f& origin,
                     float r_root, const std::vector<bool>& has_data = {});
