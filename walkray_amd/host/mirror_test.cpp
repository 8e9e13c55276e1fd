// mirror_test.cpp -- drives the C++ mirror (walkray.hpp) the way reference
// code would drive KDTree / Scene / PathTracer, for tests/test_gpu_mirror.py.
//
// usage: mirror_test <in.bin> <out.bin>
//        mirror_test --config <config.xml> <out.pfm> <spp>
//   (Config::load, Scene::buildKDTree, PathTracer::renderPixels, savePFM)
// in : u32 nv, nt, nr, W, H, spp, depth; f32 verts[3nv]; u32 tris[3nt];
//      f32 rays[6nr]; f32 cam[10] (pos, focus, up, sight_dist)
// out: u32 prim[nr]; f32 t[nr]            KDTree batch intersect
//      u32 prim1[nr]; f32 t1[nr]          KDTree::intersect one ray at a time
//      u32 obj[nr]; u32 tri[nr]; f32 ts[nr]  Scene::intersect (ObjSourceInfo)
//      u32 inside[nr]                      Scene::findInsideObject
//      f32 rgb[3 W H]                      PathTracer::renderPixels
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "walkray.hpp"

using namespace walkray;

template <class T>
static void rd(FILE* f, T* p, size_t n) {
    if (n && fread(p, sizeof(T), n, f) != n) {
        fprintf(stderr, "short read\n");
        exit(2);
    }
}
template <class T>
static void wr(FILE* f, const T* p, size_t n) {
    if (n) fwrite(p, sizeof(T), n, f);
}

static int render_config(const char* cfg_path, const char* pfm_path, unsigned spp) {
    try {
        Context ctx(0);
        Config cfg(cfg_path);
        Scene scene(ctx);
        cfg.load(scene);
        scene.buildKDTree();
        PathTracer pt(scene, spp, (unsigned)cfg.maxPathLen(), 7);
        const wr_camera& cam = cfg.camera();
        std::vector<vec3f> px = pt.renderPixels(cam);
        savePFM(pfm_path, px, cam.width, cam.height);
        printf("mirror_test: %s (%s, %s), %ux%u, %u spp, maxPathLen %d\n", cfg_path, cfg.sceneName().c_str(),
               cfg.renderer().c_str(), cam.width, cam.height, spp, cfg.maxPathLen());
    } catch (const std::exception& e) {
        fprintf(stderr, "mirror_test: %s\n", e.what());
        return 1;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 5 && std::string(argv[1]) == "--config") return render_config(argv[2], argv[3], (unsigned)atoi(argv[4]));
    if (argc != 3) {
        fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    uint32_t hdr[7];
    rd(in, hdr, 7);
    const uint32_t nv = hdr[0], nt = hdr[1], nr = hdr[2], W = hdr[3], H = hdr[4], spp = hdr[5], depth = hdr[6];
    std::vector<float> verts(3 * nv), rays(6 * nr), cam(10);
    std::vector<uint32_t> tris(3 * nt);
    rd(in, verts.data(), verts.size());
    rd(in, tris.data(), tris.size());
    rd(in, rays.data(), rays.size());
    rd(in, cam.data(), cam.size());
    fclose(in);
    try {
        Context ctx(0);
        KDTree tree(ctx);
        for (uint32_t i = 0; i < nv; i++) tree.vertexPositionList.push_back(vec3f(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]));
        for (uint32_t i = 0; i < nt; i++) {
            KDTree::Triangle t;
            for (int k = 0; k < 3; k++) t.vertexIndices[k] = tris[3 * i + k];
            tree.triangleList.push_back(t);
        }
        tree.build();
        std::vector<KDTree::Ray> rs(nr);
        for (uint32_t i = 0; i < nr; i++) {
            rs[i].origin = vec3f(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
            rs[i].direction = vec3f(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        }
        std::vector<wr_hit> hb = tree.intersect(rs);
        std::vector<uint32_t> prim(nr), prim1(nr, WR_NONE), obj(nr, WR_NONE), tri(nr, WR_NONE), inside(nr);
        std::vector<float> t(nr), t1(nr), ts(nr);
        for (uint32_t i = 0; i < nr; i++) {
            prim[i] = hb[i].prim;
            t[i] = hb[i].t;
            unsigned id = WR_NONE;
            t1[i] = tree.intersect(rs[i], id);
            prim1[i] = id;
        }
        // the same soup as one diffuse object plus an emitting quad above it
        Scene scene(ctx);
        Scene::Object o;
        o.vertices = tree.vertexPositionList;
        o.indices = tris;
        o.material = wr_material{WR_DIFFUSE, {0.7f, 0.6f, 0.5f}, 0.f, 1.f};
        scene.addObject(o);
        Scene::Object light;
        light.vertices = {vec3f(-1, -1, 2), vec3f(2, -1, 2), vec3f(2, 2, 2), vec3f(-1, 2, 2)};
        light.indices = {0, 2, 1, 0, 3, 2};  // face normal -z
        light.material = wr_material{WR_EMISSIVE, {4.f, 4.f, 4.f}, 0.f, 1.f};
        scene.addObject(light);
        scene.buildKDTree();
        for (uint32_t i = 0; i < nr; i++) {
            Scene::ObjSourceInfo src{WR_NONE, WR_NONE};
            ts[i] = scene.intersect(rs[i], src);
            obj[i] = src.objID;
            tri[i] = src.triID;
            inside[i] = scene.findInsideObject(rs[i]);
        }
        wr_camera c{{cam[0], cam[1], cam[2]}, {cam[3], cam[4], cam[5]}, {cam[6], cam[7], cam[8]}, W, H, cam[9]};
        PathTracer pt(scene, spp, depth, 42);
        std::vector<vec3f> px = pt.renderPixels(c);
        FILE* out = fopen(argv[2], "wb");
        if (!out) return 2;
        wr(out, prim.data(), nr);
        wr(out, t.data(), nr);
        wr(out, prim1.data(), nr);
        wr(out, t1.data(), nr);
        wr(out, obj.data(), nr);
        wr(out, tri.data(), nr);
        wr(out, ts.data(), nr);
        wr(out, inside.data(), nr);
        wr(out, &px[0].x, 3 * px.size());
        fclose(out);
    } catch (const std::exception& e) {
        fprintf(stderr, "mirror_test: %s\n", e.what());
        return 1;
    }
    printf("mirror_test: %u rays, %ux%u image\n", nr, W, H);
    return 0;
}
