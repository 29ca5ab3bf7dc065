// DepthMap of the reference interface (src/depth_map.hpp:7-25).
#pragma once
#include <cmath>
#include <vector>

#include "types.hpp"

struct DepthMap {
    int width = 0, height = 0;
    std::vector<float> depths;  // row-major; NaN = invalid pixel
    Matrix3f K;                 // camera intrinsics
    Matrix4f Rt;                // world <- camera
    bool valid(int u, int v) const {
        if (u < 0 || v < 0 || u >= width || v >= height) return false;
        const float d = depths[v * width + u];
        return std::isfinite(d) && d > 0.0f;
    }
};
