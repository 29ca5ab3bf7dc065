/* tgvx: MI355X (gfx950) implementation of the global-TGV hot path.
 *
 * C ABI of libtgvx.so. Plain structs, plain pointers and sizes; every call is
 * synchronous at return; the caller owns all host memory and it is never
 * retained; a context owns its device buffers and stream and is not
 * thread-safe (one per thread). All functions return 0 on success and a
 * non-zero TGVX_E_* code on failure, with a message in tgvx_last_error().
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference root):
 *   tgvx_histograms    <- compute_histograms        src/histogram.hpp:44-48
 *   tgvx_mip_build     <- DepthMipmap::build        src/histogram.hpp:27
 *   tgvx_solve         <- tgv_minimize              src/tgv.hpp:31-36
 *   tgvx_solve_oc      <- tgv_minimize_oc           src/tgv.hpp:42-48
 *   tgvx_median_init   <- CubeHistogram::median_indicator, src/histogram.hpp:18
 *   tgvx_problem_*     <- the level loop of tgv_minimize (src/tgv.cpp:387-416)
 *                         split into load / solve / read-back so that inputs can
 *                         stay resident in device memory between solves
 *   tgvx_level_bench   <- run_level                 src/tgv.cpp:137-232
 * Struct layouts equal the reference PODs byte for byte:
 *   tgvx_cube         == OctreeCube   (src/types.hpp:31-35, 16 B)
 *   tgvx_state        == TGVCube      (src/tgv.hpp:12-17, 52 B)
 *   tgvx_treetop_leaf == TreetopLeaf  (src/treetop.hpp:8-13, 20 B)
 */
#ifndef TGVX_H
#define TGVX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tgvx_ctx tgvx_ctx;

typedef struct {
    uint32_t hi, lo; /* 63-bit Morton code of the lower corner on the 2^21 grid */
    float r_c;       /* sample radius */
    int32_t depth;
} tgvx_cube;

typedef struct {
    float u;
    float v[3];
    float p[3];
    float q[6]; /* Voigt xx yy zz xy xz yz */
} tgvx_state;

typedef struct {
    uint32_t hi, lo;
    int32_t depth;
    uint32_t first_idx, last_idx; /* inclusive cube index range */
} tgvx_treetop_leaf;

/* One depth map. Only the base level is passed: the pyramid is built on the
 * device. world_to_cam is rows 0..2 of the inverse of the reference's
 * DepthMap::Rt (world <- camera), computed once on the host. */
typedef struct {
    const float* depths; /* width*height, row-major, NaN = invalid */
    int32_t width, height;
    float fx, fy, cx, cy;
    float world_to_cam[12];
    int32_t lidar; /* vote 5 instead of 1 (histogram.cpp:158) */
} tgvx_view;

/* storage of the iterated state (arithmetic is FP32 except for TGVX_FP64):
 * TGVX_BF16 stores u, v, p, q (and the extrapolations) in BF16;
 * TGVX_BF16_DUAL stores p, q in BF16 and u, v, u_bar, v_bar in FP32 */
enum { TGVX_FP64 = 0, TGVX_FP32 = 1, TGVX_BF16 = 2, TGVX_BF16_DUAL = 3 };
enum { TGVX_OP_REF = 0, TGVX_OP_ADJ = 1 };
enum {
    TGVX_OK = 0,
    TGVX_E_ARG = 1,    /* bad argument */
    TGVX_E_HIP = 2,    /* HIP runtime error */
    TGVX_E_STATE = 3,  /* call out of order (e.g. solve before load) */
    TGVX_E_NUMERIC = 4 /* non-finite energy or gap */
};

#define TGVX_MAX_LEVELS 22

/* TGVParams (src/tgv.hpp:24-28) plus the controls the reference hard-codes or
 * lacks. Defaults (tgvx_default_params) reproduce the reference. */
typedef struct {
    float alpha0;      /* weight on |eps(v)|  (reference default 0.02) */
    float alpha1;      /* weight on |grad u - v| (reference default 0.04) */
    float lambda;      /* data weight (reference: 1) */
    int32_t iters;     /* iterations per level when gap_tol <= 0 (reference default 200) */
    int32_t precision; /* TGVX_FP64 / TGVX_FP32 / TGVX_BF16 / TGVX_BF16_DUAL (state storage; the BF16 ones
                          compute in FP32) */
    int32_t op_mode;   /* TGVX_OP_REF: the reference's update rules bit for bit;
                          TGVX_OP_ADJ: primal step with the exact adjoint of the dual step */
    int32_t chunk;     /* iterations between gap evaluations (default 50) */
    int32_t max_iters; /* cap per level when iterating to a gap */
    double gap_tol;    /* relative duality gap target; <= 0 selects fixed iterations */
    int32_t warm_fallback; /* with an initial field: per level, keep the cold init when its gap is lower */
    int32_t use_graph;     /* launch iteration chunks as captured HIP graphs */
    int32_t block;         /* threads per workgroup (0 = tuned default) */
    float dual_eps;        /* > 0: after each level's initialisation set p = alpha1 K1/max(|K1|, dual_eps),
                              q = alpha0 e/max(|e|, dual_eps) from the initial (u, v) (K1 = grad u - v,
                              e = eps(v), depth-aware operators); 0: p = q = 0 as the reference */
    int32_t warm_start;    /* 1: every level starts from the context's network (tgvx_mlp_*): cold
                              initialisation + network correction, with warm_fallback as above */
    int32_t warm_dual;     /* with an initial field and dual_eps > 0: 0 recovers p, q from the field,
                              1 keeps the p, q recovered from the cold initialisation */
    int32_t fused;         /* results identical in all cases: 0 a dual and a
                              primal kernel per iteration; 1 (default) one fused kernel per
                              iteration sharing a tile's primal values through LDS; 2 one fused
                              kernel recomputing every neighbour's primal values; 3 as 1 with the
                              tile's p, q staged in LDS first; 4 as 1 with the out-of-tile
                              neighbours tabulated once per level and tile (1-4 exact-adjoint
                              mode only; the reference's rules run them as 0); 5 two iterations
                              per launch over tile sets tabulated once per level and tile (a
                              level whose sets do not fit runs as 1, with the reference's rules
                              as 0); 6 a level of at most 1024 cubes in one workgroup for all
                              its iterations, both operator modes (larger levels run as 5,
                              tiles of 1024 from 2^18 cubes on, else 512); 7 a level in
                              one cooperative launch for all its iterations, two grid barriers
                              per iteration, workgroups of block = 256 or 1024 (default; a level
                              whose workgroups cannot all be resident at once runs as 6 does;
                              exact-adjoint mode, the reference's rules run it as 5);
                              8 four iterations per launch over tile sets tabulated once per level
                              and tile, where four remain and the sets fit (tiles of 128, 256 =
                              default), otherwise as 5 (both operator modes; for small levels,
                              whose launches are bound by latency).
                              For 1, 3, 4, 5 and 8, block is the tile (128, 256 = default, 512,
                              1024) */
    int32_t tune;          /* 1: take each level's launch configuration from tgvx_autotune's table
                              (levels without an entry use fused / block as given) */
    float sigma, tau;      /* primal-dual steps; 0 = the reference's 0.25 (tgv.cpp:149-150).
                              Exact-adjoint mode converges for sigma * tau <= 1/16 */
    /* per octree depth: > 0 replaces sigma (tau) for the level of that depth in tgvx_solve,
       tgvx_solve_oc and the resident / batched solves (not the streaming solve); 0 = sigma
       (tau). No reference counterpart. */
    float sigma_depth[TGVX_MAX_LEVELS];
    float tau_depth[TGVX_MAX_LEVELS];
    int32_t no_energy;     /* fixed-iteration solves (gap_tol <= 0): 1 skips the evaluation of each
                              level's P and D after its iterations (level_energy / level_dual are
                              then 0); 0 (default) evaluates them. The reference computes neither.
                              No reference counterpart. */
} tgvx_params;

typedef struct {
    int32_t n_levels;
    int32_t level_depth[TGVX_MAX_LEVELS];
    uint32_t level_leaves[TGVX_MAX_LEVELS];
    int32_t level_iters[TGVX_MAX_LEVELS];
    int32_t level_warm[TGVX_MAX_LEVELS]; /* 1: warm init kept, 0: cold */
    double level_energy[TGVX_MAX_LEVELS]; /* primal energy P at the last evaluation */
    double level_dual[TGVX_MAX_LEVELS];   /* dual value D at the last evaluation */
    double leaf_iters;    /* sum over levels of leaves x iterations */
    double solve_ms;      /* wall time of the level loop, inputs resident on the device */
    double iterate_ms;    /* time in the iteration kernels alone (events around chunks) */
    double gap_ms;        /* time in the energy / gap evaluations */
    double setup_ms;      /* neighbour tables, permutation and per-leaf constants */
    double level_ms[TGVX_MAX_LEVELS]; /* iteration time of each level (fixed-iteration solves
                                         of one block: events around the level's launches) */
} tgvx_report;

void tgvx_default_params(tgvx_params* p);

int tgvx_create(tgvx_ctx** ctx, int32_t device);
void tgvx_destroy(tgvx_ctx* ctx);
const char* tgvx_last_error(const tgvx_ctx* ctx);
/* Device name and compute units, for reports. */
int tgvx_device_info(tgvx_ctx* ctx, char* name, int32_t name_len, int32_t* cus);

/* compute_histograms with nb bins (the reference has 8): bins_out holds n*nb
 * counts, cube-major. Bit-exact with the CPU definition. */
int tgvx_histograms(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views,
                    int32_t nviews, const float origin[3], float r_root, int32_t nb, uint32_t* bins_out);

/* Allocation-free compute_histograms. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). cubes,
 * bins (n*nb counts, as tgvx_histograms) and every view's depths are DEVICE
 * memory of the context's device; the views array itself is host memory.
 * Everything else -- the view table and the pyramid levels above each base
 * level -- lives in one fixed device array of TGVX_NOALLOC_SCRATCH_BYTES per
 * device that the library holds until the process ends, so the call makes no
 * allocation. Limits: at most TGVX_NOALLOC_MAX_VIEWS views, and their
 * pyramid levels above the base (about a third of the base pixels) within the
 * array less the view table (e.g. 40 views of 1280x960); larger inputs are
 * rejected with TGVX_E_ARG. Results equal tgvx_histograms' bit for bit. Calls
 * on one device share the array: calls that overlap on one device race. */
#define TGVX_NOALLOC_SCRATCH_BYTES (64u << 20)
#define TGVX_NOALLOC_MAX_VIEWS 1024
int tgvx_histograms_noalloc(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views,
                            int32_t nviews, const float origin[3], float r_root, int32_t nb, uint32_t* bins);

/* DepthMipmap::build of one view: all levels concatenated into out (size
 * >= width*height*4/3 + 16), (w, h) per level into dims; level count in *nlevels. */
int tgvx_mip_build(tgvx_ctx* ctx, const tgvx_view* view, float* out, int32_t* dims, int32_t* nlevels);

/* median_indicator of n histograms of nb bins, FP32. */
int tgvx_median_init(tgvx_ctx* ctx, const uint32_t* bins, uint64_t n, int32_t nb, float* out);

/* tgv_minimize. state_out: n records of 13 values (u, v[3], p[3], q[6]) --
 * tgvx_state for FP32/BF16, 13 doubles for FP64. init_uv (optional, n x 4
 * floats: u, vx, vy, vz) replaces the median / parent-copy initialisation of
 * every level (the learned warm start). report may be NULL. */
int tgvx_solve(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
               const tgvx_params* params, const float* init_uv, void* state_out, tgvx_report* report);

/* tgv_minimize_oc: per treetop leaf, out-of-leaf u_bar / v_bar read from the
 * level-start snapshot. The leaves of a level run one after another, in the
 * order given, as in the reference (a later leaf reads the final p, q of the
 * earlier ones, tgv.cpp:489-507). */
int tgvx_solve_oc(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                  const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* params,
                  void* state_out, tgvx_report* report);

/* tgv_minimize_streaming (tgv.hpp:57-64, tgv.cpp:802-1070) on arrays instead of
 * files: cubes sorted by code (as the balanced cube file), leaves as from
 * load_treetop. Per treetop leaf, the leaf's cubes are looked up by code
 * (last one wins), all others by the cube file's bisection; neighbours not
 * iterated in the leaf give their level-start (u, v), p, q outside the leaf
 * count as zero; parents come from the bisection. The reference rules
 * (op_mode 0) with params.iters per level. state_out as tgvx_solve. */
int tgvx_solve_streaming(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                         const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* params,
                         void* state_out, tgvx_report* report);

/* tgvx_solve_streaming out of core (no reference counterpart for the
 * budget; the semantics are tgv_minimize_streaming's, tgv.cpp:802-1070): the
 * state of every cube stays in pinned host memory, as the reference keeps it
 * in its state file, and at most device_bytes of per-cube data are on the
 * device at a time, in two slots that alternate so that one batch of treetop
 * leaves moves over PCIe while the other iterates. Per level the parent copy
 * and the level-start snapshot run on the host; the leaves holding cubes of
 * the level run in batches of whole leaves. The sorted codes, depths and leaf
 * of every cube (16 B per cube, reported as index_bytes) stay on the device
 * for the neighbour lookups. FP32; results equal tgvx_solve_streaming's bit for
 * bit. A leaf whose cubes of one level need more than a slot is an error. */
typedef struct {
    int32_t batches;        /* leaf batches run over all levels */
    uint64_t slot_cubes;    /* capacity of a slot in cubes */
    uint64_t device_bytes;  /* both slots */
    uint64_t index_bytes;   /* codes, depths and leaf of every cube */
    uint64_t moved_bytes;   /* host <-> device */
    double tables_ms;       /* neighbour tables (device lookups) */
} tgvx_ooc_report;
int tgvx_solve_streaming_ooc(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                             const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* params,
                             uint64_t device_bytes, void* state_out, tgvx_report* report, tgvx_ooc_report* ooc);

/* Resident problem: upload cubes + histograms once, build tables on the device. */
int tgvx_problem_load(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins,
                      int32_t nb, int32_t precision, int32_t op_mode);
/* Solve the resident problem from its cold (or given) initialisation. */
int tgvx_problem_solve(tgvx_ctx* ctx, const tgvx_params* params, const float* init_uv,
                       tgvx_report* report);
/* tgvx_problem_solve starting every level from a given state (no reference
 * counterpart): n TGVCube records (save_tgv's payload, tgv.cpp:17-46) in
 * input order; each level's u, v, p, q are taken from it and u_bar = u,
 * v_bar = v, instead of the median / parent-copy initialisation (resuming a
 * saved solve, or a warm start that includes the dual). With warm_fallback a
 * level keeps its cold start where that starts from a smaller gap. */
int tgvx_problem_solve_from(tgvx_ctx* ctx, const tgvx_params* params, const tgvx_state* init, tgvx_report* report);
/* tgvx_problem_solve with a fixed iteration count per level (n_levels = the
 * number of depths present, as tgvx_problem_levels; gap_tol must be 0): runs
 * another precision for exactly the iterations a to-gap solve took */
int tgvx_problem_solve_levels(tgvx_ctx* ctx, const tgvx_params* params, const int32_t* level_iters, int32_t n_levels,
                              tgvx_report* report);
/* Read back state (layout as tgvx_solve) and, optionally, u_bar/v_bar (n x 4 of the state type). */
int tgvx_problem_state(tgvx_ctx* ctx, void* state_out, void* ubar_out);
/* Energy P and dual value D of one level of the resident state (FP64). */
int tgvx_problem_gap(tgvx_ctx* ctx, int32_t depth, const tgvx_params* params, double* P, double* D);
/* Per-level sizes of the resident problem: depth and leaf count, returns count. */
int tgvx_problem_levels(tgvx_ctx* ctx, int32_t* depths, uint32_t* counts, int32_t* nlevels);

/* run_level timing: iters iterations of the level at `depth` of the resident
 * problem (state continues from where it is). Times measured with HIP events
 * on the library's stream: total (the configured path), and the average
 * duration of one launch of each iteration kernel (dual, primal and, in the
 * exact-adjoint mode, the fused pass; fused_ms may be NULL) from separately
 * bracketed launches. */
int tgvx_level_bench(tgvx_ctx* ctx, int32_t depth, const tgvx_params* params, int32_t iters,
                     double* total_ms, double* dual_ms, double* primal_ms, double* fused_ms);

/* Grid-barrier cost (no reference counterpart; measurement for the choice
 * of params.fused = 7): microseconds per barrier of one cooperative launch of
 * `workgroups` workgroups of `block` (256 or 1024) threads that does nothing
 * but `barriers` (1..100000) of the persistent level kernel's grid barriers
 * (difference of two launches, HIP events on the library's stream). Fails
 * when the workgroups cannot all be resident at once. */
int tgvx_grid_barrier_bench(tgvx_ctx* ctx, int32_t workgroups, int32_t block, int32_t barriers,
                            double* us_per_barrier);

/* Tile sets of the two-iteration kernel (params.fused = 5; no reference
 * counterpart) for the level at `depth` of the resident problem (exact-adjoint
 * mode) and a tile of 128, 256 or 512 cubes, built if not yet built: usable = 1
 * when every tile's sets fit (otherwise that level runs one iteration per
 * launch), max_sets = the largest PA, DB, PC over the tiles (the cubes the
 * launch's first primal step, its inner dual step and its second primal step
 * cover), lds_bytes = the kernel's LDS per workgroup. With params.fused = 5,
 * fused_ms of tgvx_level_bench is per iteration (half a launch). */
int tgvx_tile_sets(tgvx_ctx* ctx, int32_t depth, int32_t tile, int32_t* usable, int32_t* max_sets,
                   int64_t* lds_bytes);
/* The same for the four-iteration kernel (params.fused = 8; tiles of 128, 256): usable,
 * max_slots = the largest tile's slots (its nested sets and the constant slots of other-depth
 * neighbours), lds_bytes = that tile's LDS per workgroup. No reference counterpart. */
int tgvx_tile_sets4(tgvx_ctx* ctx, int32_t depth, int32_t tile, int32_t* usable, int32_t* max_slots,
                    int64_t* lds_bytes);

/* ---- iso-surface extraction (marching_cubes.cpp:63-181 extract_surface), on the device
 * Over the resident problem's cubes with u = the given field (n floats in input
 * order) or, when u is NULL, the problem's current state; has_data NULL or n
 * bytes (a pair emits only if both cubes have data). One quad (two triangles)
 * per face-adjacent pair with a sign change, at the interpolated zero crossing;
 * vertices merged on the quantised position key and numbered by first
 * occurrence, as the reference. The mesh stays in the context until the next
 * extraction; tgvx_surface_copy reads it (verts n_verts x 3, tris n_tris x 3). */
int tgvx_surface_extract(tgvx_ctx* ctx, const float* u, const uint8_t* has_data, const float origin[3],
                         float r_root, uint64_t* n_verts, uint64_t* n_tris);
int tgvx_surface_copy(tgvx_ctx* ctx, float* verts, uint32_t* tris);

/* ---- octree input build on the device (the reference's stages 1-3)
 * tgvx_octree_build runs, in memory instead of through cube files,
 *   last_stage 1: build_cubes_from_depthmap (src/octree_builder.hpp:25-29,
 *                 octree_builder.cpp:30-97) over every view, one list sorted by (code,
 *                 depth), view then pixel order among equal pairs;
 *   last_stage 2: + kway_merge (src/merge.hpp:7, merge.cpp:7-97): equal (code, depth)
 *                 cells merged, r_c averaged (a double sum in that order);
 *   last_stage 3: + balance_octree (src/balance.hpp:14, balance.cpp:60-150): the bridge
 *                 cubes the reference's FIFO queue adds, seeded in (code, depth) order.
 * Each pixel gives one cube, its depth from the sample radius (max_depth instead of the
 * reference's 21). tgvx_octree_merge and tgvx_octree_balance run stage 2 / stage 3 alone
 * on given cubes (kway_merge over the concatenated files; balance_octree seeded in input
 * order, codes aligned to their depth). Results are sorted by (code, depth), input
 * order among equal pairs. Cells are keyed by the full 63-bit code (balance.cpp:56-58
 * keeps 60 bits). The cubes stay in the context until the next call; tgvx_octree_copy
 * reads them (n_out records). */
typedef struct {
    const float* depths;    /* host, height x width row-major; valid = finite and > 0 */
    int32_t width, height;
    float fx, fy, cx, cy;   /* DepthMap::K (src/depth_map.hpp:9) */
    float cam_to_world[12]; /* DepthMap::Rt (world <- camera, depth_map.hpp:10), top 3x4 rows */
} tgvx_depth_view;
typedef struct {
    uint64_t pixels;   /* pixels scanned */
    uint64_t raw;      /* one cube per valid pixel inside the root cube */
    uint64_t merged;   /* distinct (code, depth) cells: the balance's seeds */
    uint64_t bridges;  /* cubes the balance added */
    int32_t rounds;    /* parallel balance rounds until the queue's result was reached */
    double upload_ms, cubes_ms, merge_ms, balance_ms, total_ms;
} tgvx_octree_report;
int tgvx_octree_build(tgvx_ctx* ctx, const tgvx_depth_view* views, int32_t nviews, const float origin[3],
                      float r_root, int32_t max_depth, int32_t last_stage, uint64_t* n_out,
                      tgvx_octree_report* report);
int tgvx_octree_merge(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, uint64_t* n_out,
                      tgvx_octree_report* report);
int tgvx_octree_balance(tgvx_ctx* ctx, const tgvx_cube* cubes, uint64_t n, float r_root, uint64_t* n_out,
                        tgvx_octree_report* report);
int tgvx_octree_copy(tgvx_ctx* ctx, tgvx_cube* out);

/* ---- launch autotuner (no reference counterpart)
 * For every level of the loaded problem, times `iters` iterations of each
 * launch configuration (the better of two timings, after a warm-up; pass the
 * solve's chunk, e.g. 50) -- kernel variant (dual + primal kernels, the fused
 * variants 1-7 of tgvx_params.fused; exact-adjoint mode only for the fused
 * ones but 5), workgroup / tile size (64 to 1024) and tile order (XCD-contiguous or
 * plain) -- and stores the fastest in the context, per (precision, operator
 * mode, floor(log2 leaves)), counting 1 us per kernel launch and iteration on
 * top of the measured time (what a launch costs more inside a coarse-to-fine
 * sweep than alone), unless it beats the params' own configuration by less
 * than 3 % (then that one); solves with params.tune = 1 use it. Every
 * configuration gives identical results. The problem's state is overwritten
 * (a later solve re-initialises it). */
typedef struct {
    int32_t n_levels;
    int32_t level_depth[TGVX_MAX_LEVELS];
    uint32_t level_leaves[TGVX_MAX_LEVELS];
    int32_t variant[TGVX_MAX_LEVELS]; /* 0 two kernels, 1 tiled fused, 2 untiled fused, 3 tiled + LDS p, q,
                                         4 tiled + tabulated halo, 5 two iterations per launch,
                                         6 the whole level in one workgroup,
                                         7 the whole level in one cooperative launch */
    int32_t block[TGVX_MAX_LEVELS];
    int32_t xcd[TGVX_MAX_LEVELS];
    double best_us[TGVX_MAX_LEVELS];    /* per iteration, chosen configuration */
    double default_us[TGVX_MAX_LEVELS]; /* per iteration, params' own configuration */
} tgvx_tune_report;
int tgvx_autotune(tgvx_ctx* ctx, const tgvx_params* params, int32_t iters, tgvx_tune_report* report);

/* ---- several independent octrees in one problem (no reference counterpart)
 * Blocks are concatenated level-major: every iteration of a depth is one
 * launch over that depth's cubes of all blocks; neighbour lookups stay inside
 * each block. With gap_tol > 0 a depth is iterated until every block has
 * reached it (blocks that reached it keep iterating with the rest). */
typedef struct {
    const tgvx_cube* cubes;
    uint64_t n;
    const uint32_t* bins; /* n x nb */
} tgvx_block;
int tgvx_problem_load_blocks(tgvx_ctx* ctx, const tgvx_block* blocks, uint32_t nblocks, int32_t nb,
                             int32_t precision, int32_t op_mode);
/* per level (as tgvx_problem_levels) x block, row-major: the iterations
 * after which the block's relative gap first was <= gap_tol and that gap
 * (-1 and 0: the block has no cube at that depth; -1 and the final gap: not
 * reached within max_iters); with a fixed iteration count, that count and
 * the final gap. Either pointer may be NULL. */
int tgvx_problem_block_report(tgvx_ctx* ctx, int32_t* iters, double* rel_gap);
/* load + solve + per-block read-back; state_out[b] as in tgvx_solve */
int tgvx_solve_batched(tgvx_ctx* ctx, const tgvx_block* blocks, uint32_t nblocks, int32_t nb,
                       const tgvx_params* params, void* const* state_out, tgvx_report* report);

/* ---- warm-start network (no reference counterpart)
 * 30 -> 64 -> 64 -> 4 ReLU MLP on FP32 matrix cores; parameters packed as
 * W1[30][64] b1[64] W2[64][64] b2[64] W3[64][4] b3[4]. A context starts with
 * seeded He-normal weights. Inputs are the 30 per-cube features of
 * tgvx_problem_features, outputs the correction (du, dvx, dvy, dvz) to the
 * cold initialisation. */
#define TGVX_MLP_NPARAMS 6404
int tgvx_mlp_set(tgvx_ctx* ctx, const float* params);
int tgvx_mlp_get(tgvx_ctx* ctx, float* params);
/* y[n][4] = network(x[n][30]) */
int tgvx_mlp_forward(tgvx_ctx* ctx, const float* x, uint64_t n, float* y);
/* mean squared error over n rows and 4 outputs, and its gradient (TGVX_MLP_NPARAMS) */
int tgvx_mlp_loss_grad(tgvx_ctx* ctx, const float* x, const float* t, uint64_t n, float* grad, double* loss);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled mini-batches; mean loss per epoch */
int tgvx_mlp_train(tgvx_ctx* ctx, const float* x, const float* t, uint64_t n, int32_t epochs, int32_t batch,
                   float lr, uint64_t seed, float* epoch_loss);
/* Features (n_L x 30) of the resident problem's level at `depth` from its
 * current state, with that level's current (u, v) (n_L x 4) and the input
 * index of each row: after a solve, the training pairs of the warm start. */
int tgvx_problem_features(tgvx_ctx* ctx, int32_t depth, float* x_out, float* uv_out, uint32_t* index_out);

#ifdef __cplusplus
}
#endif

#endif /* TGVX_H */
