/* Synthetic-scene and octree-input generator (host C++, no GPU).
 *
 * Produces the inputs of the TGV hot path: seeded CSG scenes rendered into
 * noisy pinhole depth maps, and the Morton octree built from them the way the
 * reference's stages 1-4 build it. These stages are outside the GPU hot path;
 * they exist so that every input can be generated offline and reproducibly.
 *
 * Reference interfaces restated here (paths relative to the reference root):
 *   tgvs_build_octree  <- build_cubes_from_depthmap (src/octree_builder.hpp:25-29)
 *                         + kway_merge (src/merge.hpp) + balance_octree (src/balance.hpp:14)
 *   tgvs_treetop       <- build_treetop (src/treetop.hpp:24), with the cube cap as
 *                         a parameter instead of N_CUBES_PER_TREETOP_LEAF (treetop.hpp:22)
 *   tgvs_camera        <- DepthMap::K / DepthMap::Rt (src/depth_map.hpp:7-12)
 * The renderer and camera rigs have no reference counterpart.
 */
#ifndef TGVX_SCENE_H
#define TGVX_SCENE_H

#include <stdint.h>

#include "tgvx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { TGVS_SPHERE = 0, TGVS_BOX = 1, TGVS_TORUS = 2 };

/* One solid of a CSG union. size: sphere {radius}, box {half extents},
 * torus {R, r} (axis = local y). rot: row-major local->world rotation. */
typedef struct {
    int32_t kind;
    float center[3];
    float size[3];
    float rot[9];
} tgvs_prim;

/* Pinhole camera. cam_to_world is the reference's DepthMap::Rt (world <- camera)
 * as its top 3x4 rows; world_to_cam its inverse, computed in double and rounded
 * once, shared by every consumer of the view. Camera frame: x right, y down,
 * z forward; depth is the camera-frame z of the surface point. */
typedef struct {
    float fx, fy, cx, cy;
    float cam_to_world[12];
    float world_to_cam[12];
} tgvs_camera;

/* n cameras on the ring (r cos t, 0, r sin t), t = 2 pi k / n, looking at 0, up +y. */
int32_t tgvs_ring_cameras(int32_t n, float radius, int32_t width, int32_t height, float hfov_deg,
                          tgvs_camera* out);
/* n cameras at seeded uniform directions on the sphere of the given radius, looking at 0. */
int32_t tgvs_random_cameras(int32_t n, float radius, int32_t width, int32_t height, float hfov_deg,
                            uint64_t seed, tgvs_camera* out);

/* Renders the union of prims into ncams depth maps of width x height (NaN =
 * no surface), adds N(0, sigma^2) depth noise and replaces outlier_frac of the
 * pixels by uniform depths in [outlier_lo, outlier_hi]. Seeded, deterministic,
 * independent of the thread count. */
int32_t tgvs_render(const tgvs_prim* prims, int32_t nprims, const tgvs_camera* cams, int32_t ncams,
                    int32_t width, int32_t height, float sigma, float outlier_frac, float outlier_lo,
                    float outlier_hi, uint64_t seed, float* depths_out);

/* Stages 1..last_stage: per-pixel cubes (octree_builder.cpp:30-97, with max_depth as a
 * parameter; last_stage 1 returns them sorted by (code, depth), view then pixel order
 * among equal pairs), merge of equal (code, depth) cells averaging r_c in double in that
 * order (merge.cpp:57-92; 2), 2:1 balance with bridge cubes (balance.cpp:60-150; 3).
 * Output sorted by (code, depth), malloc'ed into *out (release with tgvs_free). Returns
 * the cube count, or -1 on error. */
int64_t tgvs_build_octree(const float* depths, const tgvs_camera* cams, int32_t ncams, int32_t width,
                          int32_t height, const float origin[3], float r_root, int32_t max_depth,
                          int32_t last_stage, tgvx_cube** out);

/* Merge alone (merge.cpp:57-92) on an arbitrary cube list: sorted by (code, depth),
 * input order among equal pairs, each run averaged; malloc'ed. */
int64_t tgvs_merge(const tgvx_cube* cubes, uint64_t n, tgvx_cube** out);

/* Balance alone (balance.cpp:60-150) on an arbitrary cube list; output sorted by
 * (code, depth), malloc'ed. */
int64_t tgvs_balance(const tgvx_cube* cubes, uint64_t n, float r_root, tgvx_cube** out);

/* Stage 4 (treetop.cpp:23-70): split the Morton-sorted cubes into contiguous
 * leaves of at most max_per_leaf cubes. malloc'ed into *out. Returns the count. */
int64_t tgvs_treetop(const tgvx_cube* cubes, uint64_t n, uint32_t max_per_leaf, tgvx_treetop_leaf** out);

void tgvs_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* TGVX_SCENE_H */
