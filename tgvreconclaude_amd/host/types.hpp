// Geometry types of the reference interface (src/types.hpp:1-73), without Eigen.
// Vec3f / Matrix3f / Matrix4f keep the accessors the reference's call sites use
// (x(), y(), z(), operator(), Zero(), Identity()); MortonCode / OctreeCube are
// byte-identical to the reference PODs and to tgvx_cube.
#pragma once
#include <cstdint>

struct Vec3f {
    float v[3] = {0.0f, 0.0f, 0.0f};
    Vec3f() = default;
    Vec3f(float a, float b, float c) : v{a, b, c} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float x() const { return v[0]; }
    float y() const { return v[1]; }
    float z() const { return v[2]; }
    static Vec3f Zero() { return {}; }
    Vec3f operator+(const Vec3f& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    Vec3f operator-(const Vec3f& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
};

template <int R, int C>
struct MatrixNf {
    float a[R][C] = {};
    float& operator()(int r, int c) { return a[r][c]; }
    float operator()(int r, int c) const { return a[r][c]; }
    static MatrixNf Identity() {
        MatrixNf m;
        for (int i = 0; i < R && i < C; ++i) m.a[i][i] = 1.0f;
        return m;
    }
};
using Matrix3f = MatrixNf<3, 3>;
using Matrix4f = MatrixNf<4, 4>;

struct MortonCode {
    uint32_t hi = 0;
    uint32_t lo = 0;
    bool operator<(const MortonCode& o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
    bool operator==(const MortonCode& o) const { return hi == o.hi && lo == o.lo; }
};

struct OctreeCube {
    MortonCode code;
    float r_c;
    int32_t depth;
};

// 21 bits per axis, interleaved x, y, z from bit 0 (types.hpp:10-11).
inline uint64_t dilate3_64(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}
inline uint32_t compact3_64(uint64_t x) {
    x &= 0x1249249249249249ull;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ull;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00full;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffull;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffull;
    x = (x ^ (x >> 32)) & 0x1fffffull;
    return (uint32_t)x;
}
inline MortonCode morton_encode(uint32_t ix, uint32_t iy, uint32_t iz) {
    const uint64_t c = dilate3_64(ix) | (dilate3_64(iy) << 1) | (dilate3_64(iz) << 2);
    return {(uint32_t)(c >> 32), (uint32_t)c};
}
inline void morton_decode(const MortonCode& m, uint32_t& ix, uint32_t& iy, uint32_t& iz) {
    const uint64_t c = ((uint64_t)m.hi << 32) | m.lo;
    ix = compact3_64(c);
    iy = compact3_64(c >> 1);
    iz = compact3_64(c >> 2);
}
