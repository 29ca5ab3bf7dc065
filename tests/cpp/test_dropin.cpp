// The reference's hot-path tests (tests/test_stage5.cpp, tests/test_stage6.cpp,
// tests/test_stage7.cpp of the reference), written against the same C++ interface, running on the
// GPU through libtgvx. Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <string>
#include <vector>

#include <fstream>

#include "balance.hpp"
#include "histogram.hpp"
#include "marching_cubes.hpp"
#include "merge.hpp"
#include "octree_builder.hpp"
#include "tgv.hpp"

static int n_pass = 0, n_fail = 0;
#define EXPECT(c)                                                      \
    do {                                                               \
        if (c) ++n_pass;                                               \
        else {                                                         \
            ++n_fail;                                                  \
            std::printf("FAIL %s:%d  %s\n", __FILE__, __LINE__, #c);  \
        }                                                              \
    } while (0)
#define EXPECT_NEAR(a, b, e) EXPECT(std::fabs((float)(a) - (float)(b)) <= (float)(e))

static DepthMap plane(int w, int h, float depth, float f, float c) {
    DepthMap dm;
    dm.width = w;
    dm.height = h;
    dm.depths.assign((size_t)w * h, depth);
    dm.K = Matrix3f::Identity();
    dm.K(0, 0) = f;
    dm.K(1, 1) = f;
    dm.K(0, 2) = c;
    dm.K(1, 2) = c;
    dm.Rt = Matrix4f::Identity();
    return dm;
}

// test_stage5.cpp:21-41
static void mip_levels_of_constant_map() {
    DepthMipmap mip;
    mip.build(plane(8, 8, 10.0f, 100, 4), false);
    EXPECT(mip.levels.size() >= 4u);
    EXPECT(mip.levels[0].width == 8 && mip.levels[1].width == 4 && mip.levels[2].width == 2);
    EXPECT(mip.depth_at(0, 3, 3) == 10.0f);
    EXPECT(mip.depth_at(1, 2, 2) == 10.0f);
    EXPECT(mip.depth_at(2, 1, 1) == 10.0f);
}

// test_stage5.cpp:45-80
static void projection_cases() {
    DepthMipmap mip;
    mip.build(plane(32, 32, 5.0f, 100, 16), false);
    int l, u, v;
    float z;
    EXPECT(mip.project(Vec3f{0, 0, 5}, 0.04f, l, u, v, z));
    EXPECT_NEAR(z, 5.0f, 1e-4f);
    EXPECT(l == 0);
    EXPECT(!mip.project(Vec3f{0, 0, -1}, 0.1f, l, u, v, z));
    EXPECT(!mip.project(Vec3f{1000, 0, 5}, 0.1f, l, u, v, z));
    EXPECT(mip.project(Vec3f{0, 0, 1}, 2.0f, l, u, v, z));
    EXPECT(l >= 1);
}

// test_stage5.cpp:84-101
static void histogram_statistics() {
    CubeHistogram h;
    h.bins[4] = 10;
    EXPECT(h.median_indicator() == 0.125f);
    EXPECT_NEAR(h.mean_indicator(), 0.125f, 1e-6f);
    CubeHistogram h2;
    h2.bins[3] = 5;
    h2.bins[4] = 5;
    EXPECT(h2.median_indicator() == -0.125f);
    CubeHistogram e;
    EXPECT(e.median_indicator() == 0.0f && e.mean_indicator() == 0.0f);
}

// test_stage5.cpp:105-156: one cube just behind a plane, one vote in bin 3
static void one_cube_one_view() {
    DepthMipmap mip;
    mip.build(plane(16, 16, 5.0f, 50, 8), false);
    OctreeCube c;
    c.depth = 4;
    c.r_c = 1.0f / 16;
    c.code = morton_encode(4u << 17, 4u << 17, 8u << 17);
    auto hists = compute_histograms({c}, {mip}, Vec3f{-0.5f, -0.5f, 4.0f}, 1.0f);
    EXPECT(hists.size() == 1u);
    EXPECT(hists[0].total() == 1u);
    EXPECT(hists[0].bins[3] == 1u);
    const std::string path = std::filesystem::temp_directory_path().string() + "/tgvx_dropin_hist.h";
    save_histograms(path, hists);
    auto back = load_histograms(path);
    EXPECT(back.size() == 1u && back[0].bins[3] == 1u);
    std::filesystem::remove(path);
}

static void chain(std::vector<OctreeCube>& cubes, std::vector<CubeHistogram>& hists, float r_root) {
    cubes.clear();
    for (uint32_t ix = 0; ix < 3; ++ix) {
        OctreeCube c;
        c.depth = 5;
        c.r_c = r_root / 32;
        c.code = morton_encode(ix << 16, 0u, 0u);
        cubes.push_back(c);
    }
    hists.assign(3, CubeHistogram{});
    hists[0].bins[6] = 50;
    hists[2].bins[1] = 50;
}

// test_stage6.cpp:25-66
static void smoothing_chain() {
    std::vector<OctreeCube> cubes;
    std::vector<CubeHistogram> hists;
    chain(cubes, hists, 32.0f);
    TGVParams p;
    p.iters = 50;
    p.alpha0 = 0.1f;
    p.alpha1 = 0.2f;
    auto s = tgv_minimize(cubes, hists, Vec3f::Zero(), 32.0f, p);
    EXPECT(s.size() == 3u);
    EXPECT(s[0].u > 0.0f && s[2].u < 0.0f);
    EXPECT(s[1].u > s[2].u && s[1].u < s[0].u);
}

// test_stage6.cpp:68-94 (exact: a constant histogram field is a fixed point)
static void constant_field() {
    std::vector<OctreeCube> cubes;
    std::vector<CubeHistogram> hists;
    for (uint32_t ix = 0; ix < 4; ++ix) {
        OctreeCube c;
        c.depth = 4;
        c.r_c = 1.0f;
        c.code = morton_encode(ix << 17, 0u, 0u);
        cubes.push_back(c);
        CubeHistogram h;
        h.bins[6] = 10;
        hists.push_back(h);
    }
    TGVParams p;
    p.iters = 20;
    for (auto& s : tgv_minimize(cubes, hists, Vec3f::Zero(), 16.0f, p)) EXPECT(s.u == 0.625f);
}

// test_stage6.cpp:96-109
static void state_file_roundtrip() {
    std::vector<TGVCube> st(3);
    st[0].u = 1.0f;
    st[0].v[0] = 0.1f;
    st[2].u = -1.0f;
    st[2].p[2] = 0.5f;
    const std::string path = std::filesystem::temp_directory_path().string() + "/tgvx_dropin_state.s";
    save_tgv(path, st);
    auto b = load_tgv(path);
    EXPECT(b.size() == 3u && b[0].u == 1.0f && b[2].p[2] == 0.5f);
    std::filesystem::remove(path);
}

// test_stage6.cpp:133-184
static void out_of_core() {
    std::vector<OctreeCube> cubes;
    std::vector<CubeHistogram> hists;
    chain(cubes, hists, 32.0f);
    TGVParams p;
    p.iters = 50;
    p.alpha0 = 0.1f;
    p.alpha1 = 0.2f;
    TreetopLeaf all{cubes[0].code, 5, 0, 2};
    auto a = tgv_minimize(cubes, hists, Vec3f::Zero(), 32.0f, p);
    auto b = tgv_minimize_oc(cubes, hists, {all}, Vec3f::Zero(), 32.0f, p);
    for (int i = 0; i < 3; ++i) EXPECT(a[i].u == b[i].u);
    TreetopLeaf l0{cubes[0].code, 5, 0, 1}, l1{cubes[2].code, 5, 2, 2};
    auto s = tgv_minimize_oc(cubes, hists, {l0, l1}, Vec3f::Zero(), 32.0f, p);
    EXPECT(s[0].u > 0.0f && s[2].u < 0.0f && s[1].u > s[2].u && s[1].u < s[0].u);
    EXPECT(tgv_minimize_oc({}, {}, {}, Vec3f::Zero(), 1.0f).empty());
    EXPECT(tgv_minimize({}, {}, Vec3f::Zero(), 1.0f).empty());
}

// the reference tests' own file writers (test_stage6.cpp:190-238)
static std::string tmp_path(const char* name) {
    return std::filesystem::temp_directory_path().string() + "/" + name;
}
static void write_cubes(const std::string& path, float r_root, const std::vector<OctreeCube>& cubes) {
    std::ofstream out(path, std::ios::binary);
    CubeFileHeader h;
    h.r_root = r_root;
    h.cube_count = (uint32_t)cubes.size();
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    out.write(reinterpret_cast<const char*>(cubes.data()), cubes.size() * sizeof(OctreeCube));
}
static void write_treetop(const std::string& path, float r_root, const std::vector<TreetopLeaf>& leaves) {
    struct {
        char magic[4] = {'T', 'G', 'V', 'T'};
        uint32_t version = 1;
        float r_root = 0.0f;
        uint32_t leaf_count = 0;
    } h;
    h.r_root = r_root;
    h.leaf_count = (uint32_t)leaves.size();
    std::ofstream out(path, std::ios::binary);
    out.write(reinterpret_cast<const char*>(&h), sizeof(h));
    out.write(reinterpret_cast<const char*>(leaves.data()), leaves.size() * sizeof(TreetopLeaf));
}

// test_stage6.cpp:240-344
static void streaming() {
    std::vector<OctreeCube> cubes;
    std::vector<CubeHistogram> hists;
    chain(cubes, hists, 32.0f);
    const std::string cp = tmp_path("tgvx_s_cubes.bin"), hp = tmp_path("tgvx_s_hists.bin"),
                      tp = tmp_path("tgvx_s_tt.bin"), sp = tmp_path("tgvx_s_state.bin");
    write_cubes(cp, 32.0f, cubes);
    save_histograms(hp, hists);
    TGVParams p;
    p.iters = 50;
    p.alpha0 = 0.1f;
    p.alpha1 = 0.2f;
    // one leaf: as in memory
    write_treetop(tp, 32.0f, {TreetopLeaf{cubes[0].code, 5, 0, 2}});
    auto ref = tgv_minimize(cubes, hists, Vec3f::Zero(), 32.0f, p);
    tgv_minimize_streaming(cp, hp, tp, Vec3f::Zero(), 32.0f, sp, p);
    auto s1 = load_tgv(sp);
    EXPECT(s1.size() == 3u);
    for (size_t i = 0; i < s1.size() && i < 3; ++i) EXPECT_NEAR(ref[i].u, s1[i].u, 1e-4f);
    // two leaves: polarity across the leaf border
    write_treetop(tp, 32.0f, {TreetopLeaf{cubes[0].code, 5, 0, 1}, TreetopLeaf{cubes[2].code, 5, 2, 2}});
    tgv_minimize_streaming(cp, hp, tp, Vec3f::Zero(), 32.0f, sp, p);
    auto s2 = load_tgv(sp);
    EXPECT(s2.size() == 3u);
    if (s2.size() == 3u) EXPECT(s2[0].u > 0.0f && s2[2].u < 0.0f && s2[1].u > s2[2].u && s2[1].u < s2[0].u);
    // the state file reads back finite
    TGVParams q;
    q.iters = 20;
    write_treetop(tp, 32.0f, {TreetopLeaf{cubes[0].code, 5, 0, 2}});
    tgv_minimize_streaming(cp, hp, tp, Vec3f::Zero(), 32.0f, sp, q);
    for (const auto& c : load_tgv(sp)) EXPECT(std::isfinite(c.u));
    float r = 0.0f;
    EXPECT(load_cubes(cp, r).size() == 3u && r == 32.0f);
    for (const auto& f : {cp, hp, tp, sp}) std::filesystem::remove(f);
}

// histogram.hpp:55-61: leaf by leaf equals all at once
static void histograms_out_of_core() {
    std::vector<OctreeCube> cubes;
    for (uint32_t ix = 0; ix < 8; ++ix) {
        OctreeCube c;
        c.depth = 3;
        c.r_c = 0.1f;
        c.code = morton_encode(ix << 18, 3u << 18, 4u << 18);
        cubes.push_back(c);
    }
    DepthMipmap dm;
    dm.build(plane(64, 64, 1.0f, 32.0f, 32.0f));
    const std::string cp = tmp_path("tgvx_h_cubes.bin"), hp = tmp_path("tgvx_h_hists.bin");
    write_cubes(cp, 1.0f, cubes);
    const Vec3f o(-1.0f, -1.0f, -0.5f);
    compute_histograms_oc(cp, {TreetopLeaf{cubes[0].code, 1, 0, 4}, TreetopLeaf{cubes[5].code, 1, 5, 7}}, {dm}, o,
                          1.0f, hp);
    auto a = load_histograms(hp);
    auto b = compute_histograms(cubes, {dm}, o, 1.0f);
    EXPECT(a.size() == b.size());
    for (size_t i = 0; i < a.size() && i < b.size(); ++i)
        for (int k = 0; k < 8; ++k) EXPECT(a[i].bins[k] == b[i].bins[k]);
    for (const auto& f : {cp, hp}) std::filesystem::remove(f);
}

// test_stage7.cpp:245-273: a negative cube inside positive ones is a closed
// box of 12 triangles
static void closed_box() {
    std::vector<OctreeCube> cubes;
    std::vector<TGVCube> st;
    for (uint32_t iz = 0; iz < 3; ++iz)
        for (uint32_t iy = 0; iy < 3; ++iy)
            for (uint32_t ix = 0; ix < 3; ++ix) {
                OctreeCube c;
                c.depth = 3;
                c.r_c = 3.0f;
                c.code = morton_encode(ix << 18, iy << 18, iz << 18);
                cubes.push_back(c);
                TGVCube t;
                t.u = (ix == 1 && iy == 1 && iz == 1) ? -1.0f : 1.0f;
                st.push_back(t);
            }
    // the reference's extraction wants code order
    std::vector<size_t> order(cubes.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cubes[a].code < cubes[b].code; });
    std::vector<OctreeCube> cs;
    std::vector<TGVCube> ss;
    for (size_t i : order) {
        cs.push_back(cubes[i]);
        ss.push_back(st[i]);
    }
    Mesh m = extract_surface(cs, ss, Vec3f::Zero(), 24.0f);
    EXPECT(m.tri_count() == 12u);
    EXPECT(m.verts.size() == 8u);
}

// ---- stages 1-3 (tests/test_stage1.cpp:84-117, tests/test_stages2_4.cpp:47-281)

static void build_and_load() {
    DepthMap dm = plane(10, 10, 5.0f, 50, 5);
    const std::string path = tmp_path("build.cubes");
    build_cubes_from_depthmap(dm, Vec3f{-1.0f, -1.0f, 4.0f}, 2.5f, path);
    float rr = 0.0f;
    const auto cubes = load_cubes(path, rr);
    EXPECT_NEAR(rr, 2.5f, 1e-5f);
    EXPECT(!cubes.empty());
    bool sorted = true, rule = true;
    for (size_t i = 1; i < cubes.size(); ++i) sorted = sorted && !(cubes[i].code < cubes[i - 1].code);
    for (const auto& c : cubes) {
        const float cell = 2.5f / (float)(1u << c.depth);
        rule = rule && cell >= 0.75f * c.r_c && cell < 1.5f * c.r_c;
        rule = rule && cube_depth_for_radius(c.r_c, 2.5f) == c.depth;
    }
    EXPECT(sorted);
    EXPECT(rule);
    std::filesystem::remove(path);
}

static void parent_and_neighbour_codes() {
    const MortonCode code = morton_encode(5u << 11, 0, 0);  // ix 5 at depth 10
    const MortonCode p9 = parent_code(code, 10);
    uint32_t x, y, z;
    morton_decode(p9, x, y, z);
    EXPECT(x == (2u << 12) && y == 0 && z == 0);
    EXPECT(!(code == p9));
    morton_decode(parent_code(parent_code(p9, 9), 8), x, y, z);
    EXPECT(x == 0 && y == 0 && z == 0);
    const uint32_t step = 1u << 17;
    const MortonCode c = morton_encode(step, step, step);
    morton_decode(face_neighbor(c, 0, 4), x, y, z);
    EXPECT(x == 2 * step && y == step && z == step);
    morton_decode(face_neighbor(c, 1, 4), x, y, z);
    EXPECT(x == 0);
    EXPECT(face_neighbor(morton_encode(0, 0, 0), 1, 4).hi == UINT32_MAX);
    EXPECT(face_neighbor(morton_encode(15 * step, 0, 0), 0, 4).hi == UINT32_MAX);
}

static void merge_two_planes() {
    const std::string a = tmp_path("merge_a.cubes"), b = tmp_path("merge_b.cubes"), m = tmp_path("merged.cubes");
    const Vec3f origin{-1.5f, -1.5f, 4.0f};
    build_cubes_from_depthmap(plane(8, 8, 5.0f, 50, 4), origin, 3.0f, a);
    build_cubes_from_depthmap(plane(8, 8, 6.0f, 50, 4), origin, 3.0f, b);
    kway_merge({a, b}, m);
    float rr = 0.0f, ra = 0.0f, rb = 0.0f;
    const auto cubes = load_cubes(m, rr);
    EXPECT(rr == 3.0f);
    EXPECT(!cubes.empty());
    bool sorted = true;
    for (size_t i = 1; i < cubes.size(); ++i) sorted = sorted && !(cubes[i].code < cubes[i - 1].code);
    EXPECT(sorted);
    EXPECT(cubes.size() == load_cubes(a, ra).size() + load_cubes(b, rb).size());
    for (const auto& p : {a, b, m}) std::filesystem::remove(p);
}

static void balance_cases() {
    const std::string in = tmp_path("bal_in.cubes"), out = tmp_path("bal_out.cubes");
    OctreeCube one;
    one.code = morton_encode(0, 0, 0);
    one.depth = 5;
    one.r_c = 0.3f;
    save_cubes(in, {one}, 10.0f);
    balance_octree(in, out);
    float rr = 0.0f;
    EXPECT(load_cubes(out, rr).size() == 1u);
    // depths 5 and 7 sharing a face: one bridge at depth 6
    OctreeCube a, b;
    a.depth = 5;
    a.code = morton_encode(0, 0, 0);
    a.r_c = 100.0f / 32;
    b.depth = 7;
    b.code = morton_encode(1u << 16, 0, 0);
    b.r_c = 100.0f / 128;
    save_cubes(in, {a, b}, 100.0f);
    balance_octree(in, out);
    const auto cubes = load_cubes(out, rr);
    EXPECT(cubes.size() == 3u);
    bool sorted = true, bridge = false;
    for (size_t i = 1; i < cubes.size(); ++i) sorted = sorted && !(cubes[i].code < cubes[i - 1].code);
    for (const auto& c : cubes)
        bridge = bridge || (c.depth == 6 && c.code == morton_encode((1u << 16) - (1u << 15), 0, 0));
    EXPECT(sorted);
    EXPECT(bridge);
    // no cube keeps a face neighbour two or more levels coarser
    bool violation = false;
    for (const auto& c : cubes) {
        if (c.depth < 2) continue;
        for (int f = 0; f < 6; ++f) {
            const MortonCode nb = face_neighbor(c.code, f, c.depth);
            if (nb.hi == UINT32_MAX) continue;
            auto has = [&](int d, const MortonCode& k) {
                for (const auto& e : cubes)
                    if (e.depth == d && e.code == k) return true;
                return false;
            };
            if (has(c.depth, nb) || has(c.depth - 1, parent_code(nb, c.depth))) continue;
            MortonCode anc = parent_code(nb, c.depth);
            for (int d2 = c.depth - 2; d2 >= 0; --d2) {
                anc = parent_code(anc, d2 + 1);
                violation = violation || has(d2, anc);
            }
        }
    }
    EXPECT(!violation);
    std::filesystem::remove(in);
    std::filesystem::remove(out);
}

int main() {
    try {
        build_and_load();
        parent_and_neighbour_codes();
        merge_two_planes();
        balance_cases();
        mip_levels_of_constant_map();
        projection_cases();
        histogram_statistics();
        one_cube_one_view();
        smoothing_chain();
        constant_field();
        state_file_roundtrip();
        out_of_core();
        streaming();
        histograms_out_of_core();
        closed_box();
    } catch (const std::exception& e) {
        std::printf("exception: %s\n", e.what());
        return 2;
    }
    std::printf("Results: %d passed, %d failed\n", n_pass, n_fail);
    return n_fail > 0 ? 1 : 0;
}
