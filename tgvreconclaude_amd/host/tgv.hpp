// TGV interface of the reference (src/tgv.hpp:12-64), running on the GPU
// through libtgvx. Signatures and member names are the reference's.
#pragma once
#include <string>
#include <vector>

#include "histogram.hpp"
#include "treetop.hpp"
#include "types.hpp"

struct TGVCube {
    float u = 0.0f;
    float v[3] = {};
    float p[3] = {};
    float q[6] = {};
};

void save_tgv(const std::string& path, const std::vector<TGVCube>& state);
std::vector<TGVCube> load_tgv(const std::string& path);

struct TGVParams {
    float alpha0 = 0.02f;
    float alpha1 = 0.04f;
    int iters = 200;
};

std::vector<TGVCube> tgv_minimize(const std::vector<OctreeCube>& cubes, const std::vector<CubeHistogram>& hists,
                                  const Vec3f& origin, float r_root, const TGVParams& params = TGVParams{});

std::vector<TGVCube> tgv_minimize_oc(const std::vector<OctreeCube>& cubes, const std::vector<CubeHistogram>& hists,
                                     const std::vector<TreetopLeaf>& leaves, const Vec3f& origin, float r_root,
                                     const TGVParams& params = TGVParams{});

// tgv.hpp:57-64: inputs and result in files ('TGV1' cubes, 'TGVH'
// histograms, 'TGVT' treetop; result 'TGVS'). The streaming semantics of the
// reference (per-leaf lookups, level-start snapshots, zero p, q across leaf
// borders) are kept; the whole problem is solved resident on the GPU.
void tgv_minimize_streaming(const std::string& balanced_cube_path, const std::string& hist_path,
                            const std::string& treetop_path, const Vec3f& origin, float r_root,
                            const std::string& out_state_path, const TGVParams& params = TGVParams{});
