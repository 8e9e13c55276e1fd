// sanitize_test.cpp -- host-only checks of the tree builder and the scene-file
// reader, built with AddressSanitizer + UndefinedBehaviorSanitizer and with
// ThreadSanitizer (host/Makefile.sanitize).  No device code: it exercises the
// threaded reference-rule builder (subtree threads, parallel split dims), the
// threaded fast-pass hierarchy build and its level-parallel refit, the objects
// of a whole scene configuration with its shapes, motion interpolation and the
// PFM writer, and checks that thread counts change nothing.
//
// Usage: sanitize_test <scene config.xml> <scratch directory>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <random>
#include <string>
#include <vector>

#include "../../include/walkray.h"
#include "../csrc/wr_host.h"
#include "../csrc/wr_sceneio.h"

// The scene-file reader's error path and its scene hand-over are resolved here:
// this driver builds no wr_scene.
namespace wr {
int fail(wr_ctx*, int code, const std::string& msg) {
    std::fprintf(stderr, "  (reported) %s\n", msg.c_str());
    return code;
}
}  // namespace wr
extern "C" int wr_scene_add_mesh_corners(wr_scene*, const float*, uint32_t, const uint32_t*, uint32_t, const float*,
                                         const wr_material*, uint32_t*) {
    return WR_EINVAL;
}
extern "C" int wr_scene_set_motion(wr_scene*, uint32_t, float, float, const float*, const float*) { return WR_EINVAL; }

static int failures = 0;
#define CHECK(cond, what)                                  \
    do {                                                   \
        if (!(cond)) {                                     \
            std::fprintf(stderr, "FAILED: %s\n", what);   \
            failures++;                                    \
        }                                                  \
    } while (0)

template <class T>
static bool same(const std::vector<T>& a, const std::vector<T>& b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
}

static bool same_tree(const wr::FlatTree& a, const wr::FlatTree& b) {
    return same(a.nodes, b.nodes) && same(a.tris, b.tris) && same(a.wnodes, b.wnodes) && same(a.rec_prim, b.rec_prim) &&
           same(a.rec_first, b.rec_first) && same(a.leaf_up, b.leaf_up) && a.hdr.wide_root == b.hdr.wide_root &&
           a.hdr.root_ref == b.hdr.root_ref;
}

// Every record exactly once under the fast-pass hierarchy.
static bool covers_once(const wr::FlatTree& F) {
    std::vector<uint32_t> seen(F.tris.size(), 0u);
    std::vector<uint32_t> stack;
    if (F.hdr.wide_root == WR_NONE) return F.tris.empty();
    stack.push_back(F.hdr.wide_root);
    while (!stack.empty()) {
        const uint32_t r = stack.back();
        stack.pop_back();
        if (r & WR_LEAF_BIT) {
            const uint32_t rec = r & WR_FIRST_MASK;
            if (rec >= seen.size()) return false;
            seen[rec]++;
            continue;
        }
        for (int c = 0; c < 4; c++)
            if (F.wnodes[r].ref[c] != WR_NONE) stack.push_back(F.wnodes[r].ref[c]);
    }
    for (uint32_t s : seen)
        if (s != 1) return false;
    return true;
}

static void heightfield(int n, std::vector<float>& V, std::vector<uint32_t>& I, float phase) {
    V.clear();
    I.clear();
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) {
            const float u = x / float(n - 1), v = y / float(n - 1);
            V.push_back(u);
            V.push_back(v);
            V.push_back(0.1f * std::sin(9.f * u + phase) * std::cos(7.f * v));
        }
    for (int y = 0; y + 1 < n; y++)
        for (int x = 0; x + 1 < n; x++) {
            const uint32_t a = y * n + x, b = a + 1, c = a + n, d = c + 1;
            I.insert(I.end(), {a, b, d, a, d, c});
        }
}

static void soup(std::mt19937& rng, int nv, int nt, std::vector<float>& V, std::vector<uint32_t>& I) {
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    std::uniform_int_distribution<int> pick(0, nv - 1);
    V.resize(3 * (size_t)nv);
    for (float& x : V) x = std::round(U(rng) * 10.f) / 10.f;  // many coincident vertices
    I.resize(3 * (size_t)nt);
    for (uint32_t& i : I) i = (uint32_t)pick(rng);  // degenerate triangles included
}

static void check_builds(const char* name, const std::vector<float>& V, const std::vector<uint32_t>& I,
                         const std::vector<float>& V2) {
    const uint32_t nv = (uint32_t)(V.size() / 3), nt = (uint32_t)(I.size() / 3);
    const wr::FlatTree a = wr::build_tree(V.data(), nv, I.data(), nt, 1);
    const wr::FlatTree b = wr::build_tree(V.data(), nv, I.data(), nt, 8);
    CHECK(same_tree(a, b), name);
    CHECK(covers_once(a), name);
    const wr::FlatTree ra = wr::build_tree(V2.data(), nv, I.data(), nt, 1, &a);
    const wr::FlatTree rb = wr::build_tree(V2.data(), nv, I.data(), nt, 8, &a);
    CHECK(same_tree(ra, rb), name);
    CHECK(covers_once(ra), name);
    const wr::FlatTree fresh = wr::build_tree(V2.data(), nv, I.data(), nt, 8);
    CHECK(same(fresh.nodes, ra.nodes) && same(fresh.tris, ra.tris), name);  // the reference tree is rebuilt exactly
    std::printf("%s: %u triangles, %zu nodes, %zu wide nodes\n", name, nt, a.nodes.size(), a.wnodes.size());
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <scene config.xml> <scratch directory>\n", argv[0]);
        return 2;
    }
    std::vector<float> V, V2;
    std::vector<uint32_t> I, I2;
    heightfield(201, V, I, 0.f);  // 80,000 triangles: parallel subtrees, split dims and wide build
    heightfield(201, V2, I2, 0.7f);
    check_builds("heightfield", V, I, V2);
    std::mt19937 rng(5);
    soup(rng, 3000, 70000, V, I);
    V2 = V;
    for (size_t k = 0; k < V2.size(); k += 7) V2[k] += 0.05f;
    check_builds("soup", V, I, V2);

    wr::io::Scene sc;
    wr::io::load_config(argv[1], sc);
    CHECK(!sc.objects.empty(), "config objects");
    for (const wr::io::SceneObject& o : sc.objects) {
        const wr::FlatTree t = wr::build_tree(o.verts.data(), (uint32_t)(o.verts.size() / 3), o.tris.data(),
                                              (uint32_t)(o.tris.size() / 3), 4);
        CHECK(covers_once(t), o.name.c_str());
    }
    std::printf("config: %zu objects, camera %ux%u\n", sc.objects.size(), sc.cam.width, sc.cam.height);

    std::uniform_real_distribution<float> U(-2.f, 2.f);
    for (int k = 0; k < 200; k++) {
        wr::io::Mat4 a = wr::io::Mat4::identity(), b = wr::io::Mat4::identity();
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) {
                a.e[r][c] += 0.3f * U(rng);
                b.e[r][c] += 0.3f * U(rng);
            }
        const wr::io::Motion m = wr::io::make_motion(0.f, 1.f, a, b);
        const wr::io::Mat4 x = wr::io::interpolate(m, 0.5f * (U(rng) + 2.f) / 2.f);
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) CHECK(x.e[r][c] == x.e[r][c], "interpolate NaN");
    }
    std::vector<float> rgb(3 * 7 * 5, 0.25f);
    wr::io::write_pfm(std::string(argv[2]) + "/sanitize_test.pfm", rgb.data(), 7, 5);

    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("sanitize_test: all checks passed\n");
    return 0;
}
