// Host-side internals of libtgvx: context, device buffers, resident problem.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "tgvx.h"

namespace tgvx {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TGVX_HIP(call)                                                                             \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::tgvx::Error(TGVX_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_));    \
    } while (0)

// Device blocks released by one context's buffers, kept for that context's
// later allocations: all of a context's device work is in order on its one
// stream, so a block released there can be handed out again at once, without
// the device-wide wait of hipFree (repeated loads of problems of one size
// reuse every block). At most kMaxHeld bytes are kept (the largest blocks
// are freed first beyond that), and an allocation that finds the device full
// frees every kept block and tries again. Freed with the last buffer that
// refers to it.
class BlockCache : public std::enable_shared_from_this<BlockCache> {
  public:
    static constexpr size_t kMaxHeld = 32ull << 30;
    BlockCache() = default;
    BlockCache(const BlockCache&) = delete;
    BlockCache& operator=(const BlockCache&) = delete;
    ~BlockCache();
    // a kept block of b .. b + b / 8 + 2 MiB bytes (its size in *got), else a
    // new one of b bytes
    void* get(size_t b, size_t* got);
    void put(void* p, size_t bytes);
    void trim(size_t keep);  // free kept blocks, largest first, down to keep bytes

  private:
    std::mutex mu_;
    std::multimap<size_t, void*> free_;
    size_t held_ = 0;
};
// the cache of the context whose call runs on this thread (set by every C-ABI
// entry point for the call's duration), or null
BlockCache* current_cache();
struct CacheScope {
    BlockCache* prev;
    explicit CacheScope(BlockCache* c);
    ~CacheScope();
};

// Device allocation owned by the context; released on reset / destruction,
// into the cache of the context it was made for (else freed).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    std::shared_ptr<BlockCache> owner;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            if (owner)
                owner->put(p, bytes);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
        owner.reset();
    }
    void ensure(size_t b) {
        if (b <= bytes && p) return;
        release();
        if (b == 0) b = 16;
        if (BlockCache* bc = current_cache()) {
            owner = bc->shared_from_this();
            p = bc->get(b, &bytes);
            return;
        }
        TGVX_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    template <class X>
    X* as() const { return static_cast<X*>(p); }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(bytes, o.bytes);
        std::swap(owner, o.owner);
    }
};

// Two pinned host buffers that large host <-> device copies of caller buffers
// alternate through, with the host-thread pool that fills and drains them
// (xfer.hip). done[i]: recorded after the last DMA that used buf[i].
class CopyPool;
struct Staging {
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
    int next = 0;
    std::shared_ptr<CopyPool> pool;
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging();
};

// Per launch range and tile size: the fused tiled kernel's precomputed
// neighbour slots (three int32 per position) and per-tile lists of the
// out-of-tile neighbours of the same level (3 * tile entries per tile).
struct HaloBufs {
    DevBuf slot[3], cnt, list;
};

// Per level and tile size: the two-iteration kernel's tile sets (per tile:
// sizes, out-of-tile members, +a source slots, incoming-edge slots). ok is
// false when a tile's sets exceed the cap or its LDS need exceeds the budget;
// the level then runs one-iteration passes. sa, sd, sp: the largest PA, DB,
// PC over the tiles; lds: the largest dynamic LDS of a launch. The tiles run
// in at most two launches (LDS classes): those whose LDS need keeps the
// occupancy of the tile at the 90th percentile, then the rest, each launch
// sized for its largest tile; the tables' rows are in that launch order,
// class by class.
struct Halo2Bufs {
    DevBuf cnt, list, bslot, cslot;
    bool ok = false;
    int sa = 0, sd = 0, sp = 0;
    size_t lds = 0;
    struct Class {
        uint32_t first = 0, count = 0;
        int sa = 0, sd = 0, sx = 0;
        size_t lds = 0;
    } cls[2];
};

// The four-iteration kernel's tile sets of one level and tile size: per
// tile the ends of its nested sets and constant slots, the halo positions and
// the slot tables; lds = the largest tile's LDS in bytes.
struct Halo4Bufs {
    DevBuf cnt, list, bslot, cslot;
    bool ok = false;
    int sa = 0;
    size_t lds = 0;
};

struct Level {
    int depth;
    uint32_t off, n;  // contiguous range of permuted positions
};

// A problem resident on the device: cubes permuted level-major (ascending
// depth; within a depth ascending input index, i.e. the reference's by_depth
// order), neighbour tables in permuted positions, per-leaf data constants and
// the state arrays.
struct Problem {
    bool loaded = false;
    uint64_t n = 0;
    int nb = 0, precision = TGVX_FP32, op_mode = TGVX_OP_REF;
    size_t elem = 4;  // bytes per stored state value
    std::vector<Level> levels;
    uint32_t nblocks = 1;                    // independent octrees sharing the problem
    std::vector<uint64_t> block_start;       // input index range of each block
    std::vector<std::vector<uint32_t>> seg;  // per level: block b owns positions off + [seg[b], seg[b+1])
    std::vector<int32_t> block_iters;        // per level x block: iterations to reach the gap
    std::vector<double> block_gap;           // per level x block: final relative gap
    std::vector<int32_t> level_iters;        // fixed iterations per level for the next solve (empty: params)
    std::vector<uint32_t> perm;      // position -> input index
    std::vector<uint32_t> inv;       // input index -> position
    DevBuf d_code;                   // uint64 Morton code per position
    DevBuf d_depth;                  // int32 depth per position
    DevBuf d_nb[6];                  // int32 per position, faces 0..5
    DevBuf d_parent;                 // int32: position of the cube holding the parent code, or -1
    DevBuf d_info;                   // uint32 packed vote total / median bin / flags
    DevBuf d_nba[3];                 // int32 depth-aware +x, +y, +z neighbours (exact-adjoint mode)
    DevBuf d_nbam[3];                // int32 depth-aware -x, -y, -z neighbours (warm-start features)
    DevBuf d_nbt[3];                 // int32 the level's cube whose depth-aware +a neighbour this is, or -1
    DevBuf d_bins;                   // uint32 histograms, position order
    DevBuf d_blk;                    // int32 block of each position (several blocks only)
    DevBuf d_seg;                    // uint32 one level's block segments
    DevBuf d_bact;                   // uint8 per block: still iterating (to-gap solves of several blocks)
    DevBuf d_parent_a;               // int32 cube exactly one level coarser at the parent code, or -1
    DevBuf d_state[17];              // u vx vy vz px py pz q0..q5 ub vbx vby vbz
    DevBuf d_save[4];                // scratch copy of u, v of one level
    DevBuf d_alt[13];                // second (u, v, p, q) set of the fused iteration
    DevBuf d_snb[6], d_sflag, d_spar;  // streaming mode: per-leaf faces, flags, bisection parent
    DevBuf d_rfz;                    // reference rules: out-of-level neighbour records of a level
    DevBuf d_rfz_cnt;                // uint32: out-of-level references of a level (count)
    std::vector<int8_t> rfz_use;     // per level: records used (1), not (0), not yet decided (-1)
    DevBuf d_fzaos;                  // the previous level's u_bar, v_bar as 4-value records
    DevBuf d_init;                   // float x 4 per position (warm start field), optional
    DevBuf d_orig;                   // uint32 position -> input index
    DevBuf d_part;                   // gap partial sums
    DevBuf d_pd;                     // final P, D
    DevBuf d_lvlpd;                  // per level P, D of a fixed-iteration solve, read at its end
    // halo tables of the fused tiled kernel, keyed by (first position, count, tile)
    std::map<std::tuple<uint32_t, uint32_t, int>, std::unique_ptr<HaloBufs>> halo;
    // tile sets of the two-iteration kernel, keyed likewise (and by the state's element size)
    std::map<std::tuple<uint32_t, uint32_t, int, int>, std::unique_ptr<Halo2Bufs>> halo2;
    std::map<std::tuple<uint32_t, uint32_t, int, int>, std::unique_ptr<Halo4Bufs>> halo4;
};

}  // namespace tgvx

namespace tgvx {
// the autotuner's choice for one (precision, operator mode, log2 size) bucket
struct TuneEntry {
    int variant, block, xcd;
    double us;
};
}  // namespace tgvx

struct tgvx_ctx {
    std::shared_ptr<tgvx::BlockCache> cache = std::make_shared<tgvx::BlockCache>();
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4 + 2 * TGVX_MAX_LEVELS] = {};  // 4 general, then a pair per level
    std::string err;
    tgvx::Problem prob;
    // scratch for histogram calls
    tgvx::DevBuf h_cubes, h_bins, h_pyr, h_meta;
    // warm-start network: weights, Adam moments, activations and data
    tgvx::DevBuf mlp_w, mlp_m, mlp_v, mlp_x, mlp_t, mlp_y, mlp_rows, mlp_part;
    int mlp_step = 0;
    // last extracted surface: vertices (x, y, z) and triangles
    tgvx::DevBuf srf_verts, srf_tris, srf_tmp[8];
    uint64_t srf_nv = 0, srf_nt = 0;
    // last octree build / balance: the cubes (tgvx_cube records) and scratch
    tgvx::DevBuf oct_out, oct_tmp[32];
    uint64_t oct_n = 0;
    // launch configurations chosen by tgvx_autotune
    std::map<std::tuple<int, int, int>, tgvx::TuneEntry> tuned;
    // the persistent level kernel's grid barrier: [0] arrival count, [64]
    // error flag; grid_used: a launch since the flag was last read
    tgvx::DevBuf grid_bar;
    bool grid_used = false;
    // pinned staging of large copies of caller buffers (made on first use)
    std::unique_ptr<tgvx::Staging> stage;
    // a whole fixed-iteration sweep is being captured into one graph: the
    // level launches go onto the stream as kernels, not as graphs of their own
    bool in_capture = false;
};

namespace tgvx {
// xfer.hip: copies of caller (pageable) host buffers on the context's stream
// (or st): h2d returns once src may be reused, its DMAs ordered on the
// stream; d2h returns with dst filled.
void h2d(tgvx_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st = nullptr);
// the pieces' bytes one after another into dst
struct HostPiece {
    const void* src;
    size_t bytes;
};
void h2dv(tgvx_ctx* c, void* dst, const HostPiece* pieces, int count, hipStream_t st = nullptr);
void d2h(tgvx_ctx* c, void* dst, const void* src, size_t bytes, hipStream_t st = nullptr);
// the same pool's host threads: f(0) .. f(parts - 1) with parts at most
// host_threads(c), part 0 on the calling thread
int host_threads(tgvx_ctx* c);
void parallel_for(tgvx_ctx* c, int parts, const std::function<void(int)>& f);
// f(lo, hi) over consecutive ranges covering [0, count), one per part, with at
// least `grain` items per part
void parallel_range(tgvx_ctx* c, size_t count, size_t grain, const std::function<void(size_t, size_t)>& f);
// the same with the part's number k (ranges ascend with k); returns the
// number of parts, at most host_threads(c)
int parallel_split(tgvx_ctx* c, size_t count, size_t grain, const std::function<void(int, size_t, size_t)>& f);
// solve.hip
void problem_load(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int nb, int precision,
                  int op_mode, const uint64_t* blk_start = nullptr, uint32_t nblocks = 1);
void problem_solve(tgvx_ctx* c, const tgvx_params* p, const float* init_uv, tgvx_report* rep,
                   const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_cube* cubes, int init_stride = 4);
void problem_solve_levels(tgvx_ctx* c, const tgvx_params* p, const int32_t* level_iters, int32_t n_levels,
                          tgvx_report* rep);
void problem_state(tgvx_ctx* c, void* state_out, void* ubar_out);
void solve_stream_ooc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int nb,
                      const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* p, uint64_t budget,
                      float* state_out, tgvx_report* rep, tgvx_ooc_report* orep);
void problem_solve_streaming(tgvx_ctx* c, const tgvx_params* p, const tgvx_treetop_leaf* leaves, uint64_t nleaves,
                             tgvx_report* rep);
void problem_gap(tgvx_ctx* c, int depth, const tgvx_params* p, double* P, double* D);
void level_bench(tgvx_ctx* c, int depth, const tgvx_params* p, int iters, double* total_ms, double* dual_ms,
                 double* primal_ms, double* fused_ms);
void autotune(tgvx_ctx* c, const tgvx_params* p, int iters, tgvx_tune_report* rep);
void grid_barrier_bench(tgvx_ctx* c, int wgs, int block, int nbar, double* us);
void tile_sets(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_sets, int64_t* lds_bytes);
void tile_sets4(tgvx_ctx* c, int depth, int tile, int32_t* usable, int32_t* max_slots, int64_t* lds_bytes);
// octree.hip
void octree_build(tgvx_ctx* c, const tgvx_depth_view* views, int32_t nviews, const float origin[3], float r_root,
                  int32_t max_depth, int32_t last_stage, uint64_t* n_out, tgvx_octree_report* rep);
void octree_merge(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, uint64_t* n_out, tgvx_octree_report* rep);
void octree_balance(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, float r_root, uint64_t* n_out,
                    tgvx_octree_report* rep);
void octree_copy(tgvx_ctx* c, tgvx_cube* out);
// surface.hip
void surface_extract(tgvx_ctx* c, const float* u, const uint8_t* has_data, const float origin[3], float r_root,
                     uint64_t* nv, uint64_t* nt);
void surface_copy(tgvx_ctx* c, float* verts, uint32_t* tris);
void release_graphs(const tgvx_ctx* c);
// mlp.hip
void mlp_set(tgvx_ctx* c, const float* w);
void mlp_get(tgvx_ctx* c, float* w);
void mlp_forward(tgvx_ctx* c, const float* x, uint64_t n, float* y);
void mlp_loss_grad(tgvx_ctx* c, const float* x, const float* t, uint64_t n, float* grad, double* loss);
void mlp_train(tgvx_ctx* c, const float* x, const float* t, uint64_t n, int epochs, int batch, float lr,
               uint64_t seed, float* epoch_loss);
void warm_init(tgvx_ctx* c, uint32_t off, uint32_t n);
void problem_features(tgvx_ctx* c, int depth, float* x_out, float* uv_out, uint32_t* index_out);
// hist.hip
void histograms(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views, int nv,
                const float origin[3], float r_root, int nb, uint32_t* bins_out);
void mip_build(tgvx_ctx* c, const tgvx_view* v, float* out, int32_t* dims, int32_t* nlevels);
void median_init(tgvx_ctx* c, const uint32_t* bins, uint64_t n, int nb, float* out);
void noalloc_setup(int device);
void histograms_noalloc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views, int nv,
                        const float origin[3], float r_root, int nb, uint32_t* bins);
}  // namespace tgvx
