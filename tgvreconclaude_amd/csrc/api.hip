// extern "C" entry points of libtgvx.so: argument checks, error capture,
// context lifetime. Compute lives in hist.hip and solve.hip.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "tgvx_internal.h"

namespace tgvx {
void release_graphs(const tgvx_ctx* c);
}

namespace {
template <class F>
int guarded(tgvx_ctx* c, F&& f) {
    if (!c) return TGVX_E_ARG;
    const tgvx::CacheScope scope(c->cache.get());
    try {
        TGVX_HIP(hipSetDevice(c->device));
        f();
        c->err.clear();
        return TGVX_OK;
    } catch (const tgvx::Error& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return TGVX_E_HIP;
    }
}
}  // namespace

extern "C" {

void tgvx_default_params(tgvx_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->alpha0 = 0.02f;  // tgv.hpp:25
    p->alpha1 = 0.04f;  // tgv.hpp:26
    p->lambda = 1.0f;
    p->iters = 200;     // tgv.hpp:27
    p->precision = TGVX_FP32;
    p->op_mode = TGVX_OP_REF;
    p->chunk = 50;
    p->max_iters = 20000;
    p->gap_tol = 0.0;
    p->warm_fallback = 1;
    p->use_graph = 1;
    p->block = 256;
    p->fused = 1;
}

int tgvx_create(tgvx_ctx** out, int32_t device) {
    if (!out) return TGVX_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return TGVX_E_HIP;
    if (device < 0 || device >= n) return TGVX_E_ARG;
    tgvx_ctx* c = new tgvx_ctx;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return TGVX_E_HIP;
    }
    for (auto& e : c->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            delete c;
            return TGVX_E_HIP;
        }
    }
    tgvx::noalloc_setup(device);
    *out = c;
    return TGVX_OK;
}

void tgvx_destroy(tgvx_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->stage.reset();
    tgvx::release_graphs(c);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* tgvx_last_error(const tgvx_ctx* c) { return c ? c->err.c_str() : "null context"; }

int tgvx_device_info(tgvx_ctx* c, char* name, int32_t name_len, int32_t* cus) {
    return guarded(c, [&] {
        hipDeviceProp_t prop;
        TGVX_HIP(hipGetDeviceProperties(&prop, c->device));
        if (name && name_len > 0) {
            std::strncpy(name, prop.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cus) *cus = prop.multiProcessorCount;
    });
}

int tgvx_histograms(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views, int32_t nviews,
                    const float origin[3], float r_root, int32_t nb, uint32_t* bins_out) {
    return guarded(c, [&] { tgvx::histograms(c, cubes, n, views, nviews, origin, r_root, nb, bins_out); });
}

int tgvx_histograms_noalloc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const tgvx_view* views,
                            int32_t nviews, const float origin[3], float r_root, int32_t nb, uint32_t* bins) {
    return guarded(c, [&] { tgvx::histograms_noalloc(c, cubes, n, views, nviews, origin, r_root, nb, bins); });
}

int tgvx_mip_build(tgvx_ctx* c, const tgvx_view* view, float* out, int32_t* dims, int32_t* nlevels) {
    return guarded(c, [&] { tgvx::mip_build(c, view, out, dims, nlevels); });
}

int tgvx_median_init(tgvx_ctx* c, const uint32_t* bins, uint64_t n, int32_t nb, float* out) {
    return guarded(c, [&] { tgvx::median_init(c, bins, n, nb, out); });
}

int tgvx_problem_load(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                      int32_t precision, int32_t op_mode) {
    return guarded(c, [&] { tgvx::problem_load(c, cubes, n, bins, nb, precision, op_mode); });
}

int tgvx_problem_solve(tgvx_ctx* c, const tgvx_params* p, const float* init_uv, tgvx_report* rep) {
    return guarded(c, [&] { tgvx::problem_solve(c, p, init_uv, rep, nullptr, 0, nullptr); });
}

int tgvx_problem_solve_from(tgvx_ctx* c, const tgvx_params* p, const tgvx_state* init, tgvx_report* rep) {
    return guarded(c, [&] {
        if (!init) throw tgvx::Error(TGVX_E_ARG, "null initial state");
        tgvx::problem_solve(c, p, reinterpret_cast<const float*>(init), rep, nullptr, 0, nullptr, 13);
    });
}

int tgvx_problem_solve_levels(tgvx_ctx* c, const tgvx_params* p, const int32_t* level_iters, int32_t n_levels,
                              tgvx_report* rep) {
    return guarded(c, [&] { tgvx::problem_solve_levels(c, p, level_iters, n_levels, rep); });
}

int tgvx_problem_state(tgvx_ctx* c, void* state_out, void* ubar_out) {
    return guarded(c, [&] { tgvx::problem_state(c, state_out, ubar_out); });
}

int tgvx_problem_gap(tgvx_ctx* c, int32_t depth, const tgvx_params* p, double* P, double* D) {
    return guarded(c, [&] {
        if (!p || !P || !D) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::problem_gap(c, depth, p, P, D);
    });
}

int tgvx_problem_levels(tgvx_ctx* c, int32_t* depths, uint32_t* counts, int32_t* nlevels) {
    return guarded(c, [&] {
        const auto& P = c->prob;
        if (!P.loaded) throw tgvx::Error(TGVX_E_STATE, "no problem loaded");
        for (size_t i = 0; i < P.levels.size(); ++i) {
            if (depths) depths[i] = P.levels[i].depth;
            if (counts) counts[i] = P.levels[i].n;
        }
        if (nlevels) *nlevels = (int32_t)P.levels.size();
    });
}

int tgvx_level_bench(tgvx_ctx* c, int32_t depth, const tgvx_params* p, int32_t iters, double* total_ms,
                     double* dual_ms, double* primal_ms, double* fused_ms) {
    return guarded(c, [&] {
        if (!p || !total_ms || !dual_ms || !primal_ms || iters <= 0) throw tgvx::Error(TGVX_E_ARG, "bad argument");
        tgvx::level_bench(c, depth, p, iters, total_ms, dual_ms, primal_ms, fused_ms);
    });
}

int tgvx_grid_barrier_bench(tgvx_ctx* c, int32_t workgroups, int32_t block, int32_t barriers,
                            double* us_per_barrier) {
    return guarded(c, [&] {
        if (!us_per_barrier) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::grid_barrier_bench(c, workgroups, block, barriers, us_per_barrier);
    });
}

int tgvx_tile_sets(tgvx_ctx* c, int32_t depth, int32_t tile, int32_t* usable, int32_t* max_sets,
                   int64_t* lds_bytes) {
    return guarded(c, [&] {
        if (!usable || !max_sets || !lds_bytes) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::tile_sets(c, depth, tile, usable, max_sets, lds_bytes);
    });
}

int tgvx_tile_sets4(tgvx_ctx* c, int32_t depth, int32_t tile, int32_t* usable, int32_t* max_slots,
                    int64_t* lds_bytes) {
    return guarded(c, [&] {
        if (!usable || !max_slots || !lds_bytes) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::tile_sets4(c, depth, tile, usable, max_slots, lds_bytes);
    });
}

int tgvx_mlp_set(tgvx_ctx* c, const float* w) {
    return guarded(c, [&] {
        if (!w) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::mlp_set(c, w);
    });
}

int tgvx_mlp_get(tgvx_ctx* c, float* w) {
    return guarded(c, [&] {
        if (!w) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::mlp_get(c, w);
    });
}

int tgvx_mlp_forward(tgvx_ctx* c, const float* x, uint64_t n, float* y) {
    return guarded(c, [&] {
        if (n && (!x || !y)) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::mlp_forward(c, x, n, y);
    });
}

int tgvx_mlp_loss_grad(tgvx_ctx* c, const float* x, const float* t, uint64_t n, float* grad, double* loss) {
    return guarded(c, [&] {
        if (!x || !t) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::mlp_loss_grad(c, x, t, n, grad, loss);
    });
}

int tgvx_mlp_train(tgvx_ctx* c, const float* x, const float* t, uint64_t n, int32_t epochs, int32_t batch, float lr,
                   uint64_t seed, float* epoch_loss) {
    return guarded(c, [&] {
        if (!x || !t) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        tgvx::mlp_train(c, x, t, n, epochs, batch, lr, seed, epoch_loss);
    });
}

int tgvx_problem_features(tgvx_ctx* c, int32_t depth, float* x_out, float* uv_out, uint32_t* index_out) {
    return guarded(c, [&] { tgvx::problem_features(c, depth, x_out, uv_out, index_out); });
}

namespace {
// concatenated cubes / bins of several blocks and their input ranges
struct Concat {
    std::vector<tgvx_cube> cubes;
    std::vector<uint32_t> bins;
    std::vector<uint64_t> start;
};
Concat concat(const tgvx_block* blocks, uint32_t nblocks, int32_t nb) {
    if (!blocks || nblocks == 0) throw tgvx::Error(TGVX_E_ARG, "no blocks");
    if (nb <= 0 || nb > 32) throw tgvx::Error(TGVX_E_ARG, "nb must be in 1..32");
    Concat c;
    c.start.push_back(0);
    uint64_t tot = 0;
    for (uint32_t b = 0; b < nblocks; ++b) tot += blocks[b].n;
    c.cubes.reserve(tot);
    c.bins.reserve(tot * nb);
    for (uint32_t b = 0; b < nblocks; ++b) {
        const tgvx_block& k = blocks[b];
        if (k.n && (!k.cubes || !k.bins)) throw tgvx::Error(TGVX_E_ARG, "null block arrays");
        c.cubes.insert(c.cubes.end(), k.cubes, k.cubes + k.n);
        c.bins.insert(c.bins.end(), k.bins, k.bins + k.n * nb);
        c.start.push_back(c.start.back() + k.n);
    }
    return c;
}
}  // namespace

int tgvx_surface_extract(tgvx_ctx* c, const float* u, const uint8_t* has_data, const float origin[3], float r_root,
                         uint64_t* nv, uint64_t* nt) {
    return guarded(c, [&] { tgvx::surface_extract(c, u, has_data, origin, r_root, nv, nt); });
}

int tgvx_surface_copy(tgvx_ctx* c, float* verts, uint32_t* tris) {
    return guarded(c, [&] { tgvx::surface_copy(c, verts, tris); });
}

int tgvx_octree_build(tgvx_ctx* c, const tgvx_depth_view* views, int32_t nviews, const float origin[3],
                      float r_root, int32_t max_depth, int32_t last_stage, uint64_t* n_out, tgvx_octree_report* rep) {
    return guarded(c, [&] {
        tgvx::octree_build(c, views, nviews, origin, r_root, max_depth, last_stage, n_out, rep);
    });
}

int tgvx_octree_merge(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, uint64_t* n_out, tgvx_octree_report* rep) {
    return guarded(c, [&] { tgvx::octree_merge(c, cubes, n, n_out, rep); });
}

int tgvx_octree_balance(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, float r_root, uint64_t* n_out,
                        tgvx_octree_report* rep) {
    return guarded(c, [&] { tgvx::octree_balance(c, cubes, n, r_root, n_out, rep); });
}

int tgvx_octree_copy(tgvx_ctx* c, tgvx_cube* out) {
    return guarded(c, [&] { tgvx::octree_copy(c, out); });
}

int tgvx_solve_streaming(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                         const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* p, void* state_out,
                         tgvx_report* rep) {
    return guarded(c, [&] {
        if (!p || !state_out || (n && (!cubes || !bins))) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        // the cube file is sorted by code; its bisection relies on it
        for (uint64_t i = 1; i < n; ++i) {
            const uint64_t a = ((uint64_t)cubes[i - 1].hi << 32) | cubes[i - 1].lo;
            const uint64_t b = ((uint64_t)cubes[i].hi << 32) | cubes[i].lo;
            if (b < a) throw tgvx::Error(TGVX_E_ARG, "cubes must be sorted by Morton code");
        }
        tgvx_report r;
        std::memset(&r, 0, sizeof(r));
        const auto t0 = std::chrono::steady_clock::now();
        tgvx::problem_load(c, cubes, n, bins, nb, p->precision, TGVX_OP_REF);
        r.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        tgvx::problem_solve_streaming(c, p, leaves, nleaves, &r);
        tgvx::problem_state(c, state_out, nullptr);
        if (rep) *rep = r;
    });
}

int tgvx_solve_streaming_ooc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                             const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* p,
                             uint64_t device_bytes, void* state_out, tgvx_report* rep, tgvx_ooc_report* ooc) {
    return guarded(c, [&] {
        if (p && p->precision != TGVX_FP32) throw tgvx::Error(TGVX_E_ARG, "the out-of-core solve stores FP32");
        tgvx::solve_stream_ooc(c, cubes, n, bins, nb, leaves, nleaves, p, device_bytes, static_cast<float*>(state_out),
                               rep, ooc);
    });
}

int tgvx_autotune(tgvx_ctx* c, const tgvx_params* p, int32_t iters, tgvx_tune_report* rep) {
    return guarded(c, [&] { tgvx::autotune(c, p, iters, rep); });
}

int tgvx_problem_load_blocks(tgvx_ctx* c, const tgvx_block* blocks, uint32_t nblocks, int32_t nb, int32_t precision,
                             int32_t op_mode) {
    return guarded(c, [&] {
        Concat k = concat(blocks, nblocks, nb);
        tgvx::problem_load(c, k.cubes.data(), k.cubes.size(), k.bins.data(), nb, precision, op_mode, k.start.data(),
                           nblocks);
    });
}

int tgvx_problem_block_report(tgvx_ctx* c, int32_t* iters, double* rel_gap) {
    return guarded(c, [&] {
        const auto& P = c->prob;
        if (!P.loaded) throw tgvx::Error(TGVX_E_STATE, "no problem loaded");
        if (iters) std::copy(P.block_iters.begin(), P.block_iters.end(), iters);
        if (rel_gap) std::copy(P.block_gap.begin(), P.block_gap.end(), rel_gap);
    });
}

int tgvx_solve_batched(tgvx_ctx* c, const tgvx_block* blocks, uint32_t nblocks, int32_t nb, const tgvx_params* p,
                       void* const* state_out, tgvx_report* rep) {
    return guarded(c, [&] {
        if (!p || !state_out) throw tgvx::Error(TGVX_E_ARG, "null pointer");
        Concat k = concat(blocks, nblocks, nb);
        tgvx_report r;
        std::memset(&r, 0, sizeof(r));
        const auto t0 = std::chrono::steady_clock::now();
        tgvx::problem_load(c, k.cubes.data(), k.cubes.size(), k.bins.data(), nb, p->precision, p->op_mode,
                           k.start.data(), nblocks);
        r.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        tgvx::problem_solve(c, p, nullptr, &r, nullptr, 0, nullptr);
        const size_t rec = p->precision == TGVX_FP64 ? 13 * sizeof(double) : sizeof(tgvx_state);
        std::vector<uint8_t> all(rec * std::max<size_t>(k.cubes.size(), 1));
        tgvx::problem_state(c, all.data(), nullptr);
        for (uint32_t b = 0; b < nblocks; ++b)
            if (blocks[b].n) std::memcpy(state_out[b], all.data() + rec * k.start[b], rec * blocks[b].n);
        if (rep) *rep = r;
    });
}

static int solve_common(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                        const tgvx_params* p, const float* init_uv, const tgvx_treetop_leaf* leaves,
                        uint64_t nleaves, void* state_out, tgvx_report* rep) {
    return guarded(c, [&] {
        if (!p) throw tgvx::Error(TGVX_E_ARG, "null params");
        if (n > 0 && !state_out) throw tgvx::Error(TGVX_E_ARG, "null state_out");
        tgvx_report r;
        std::memset(&r, 0, sizeof(r));
        const auto t0 = std::chrono::steady_clock::now();
        tgvx::problem_load(c, cubes, n, bins, nb, p->precision, p->op_mode);
        r.setup_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        tgvx::problem_solve(c, p, init_uv, &r, leaves, nleaves, cubes);
        tgvx::problem_state(c, state_out, nullptr);
        if (rep) *rep = r;
    });
}

int tgvx_solve(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
               const tgvx_params* params, const float* init_uv, void* state_out, tgvx_report* report) {
    return solve_common(c, cubes, n, bins, nb, params, init_uv, nullptr, 0, state_out, report);
}

int tgvx_solve_oc(tgvx_ctx* c, const tgvx_cube* cubes, uint64_t n, const uint32_t* bins, int32_t nb,
                  const tgvx_treetop_leaf* leaves, uint64_t nleaves, const tgvx_params* params, void* state_out,
                  tgvx_report* report) {
    if (n > 0 && (!leaves || nleaves == 0)) {
        if (c) c->err = "tgvx_solve_oc needs at least one treetop leaf";
        return TGVX_E_ARG;
    }
    return solve_common(c, cubes, n, bins, nb, params, nullptr, leaves, nleaves, state_out, report);
}

}  // extern "C"
