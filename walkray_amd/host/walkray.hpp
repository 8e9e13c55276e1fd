// walkray.hpp -- C++ host mirror of the reference's hot-path interfaces
// (KDTree, Scene, PathTracer) over libwalkray's C ABI (include/walkray.h).
//
// Names and argument meanings follow the reference:
//   KDTree      kdtree.h:9-48   (Ray, Triangle, Strategy, vertexPositionList,
//                                triangleList, build, destroy, intersect)
//   Scene       scene.h:10-45   (ObjSourceInfo, buildKDTree, intersect,
//                                findInsideObject)
//   PathTracer  pathTracer.h:7-18, mcRenderer.h:14-36 (spp, maxLen, renderPixels)
//   Config      config.h        (load: scene objects, camera, renderer settings)
//   savePFM     film.h:100-110
// Differences, all forced by the GPU boundary:
//   * a Condition cannot run arbitrary host code inside the GPU traversal:
//     KDTree::intersect accepts only a NULL condition, and the reference's one
//     condition on this path, NoSelfCondition, is Scene::intersectNoSelf;
//   * errors: the reference reports none (misses return -1, unknown materials
//     become NULL); here a failing ABI call throws std::runtime_error with the
//     library's message, and a miss still returns -1 with the id untouched.
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/walkray.h"

namespace walkray {

struct vec3f {
    float x = 0, y = 0, z = 0;
    vec3f() = default;
    vec3f(float a, float b, float c) : x(a), y(b), z(c) {}
};

inline void check(int rc, const wr_ctx* ctx, const char* what) {
    if (rc != WR_OK) throw std::runtime_error(std::string(what) + ": " + wr_last_error(ctx));
}

// One device + stream.
class Context {
  public:
    explicit Context(int device = 0) { check(wr_ctx_create(device, &ctx_), nullptr, "wr_ctx_create"); }
    ~Context() { wr_ctx_destroy(ctx_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    wr_ctx* get() const { return ctx_; }

  private:
    wr_ctx* ctx_ = nullptr;
};

template <class T>
class DeviceArray {
  public:
    DeviceArray(wr_ctx* ctx, size_t n) : ctx_(ctx), n_(n) {
        check(wr_device_alloc(ctx, sizeof(T) * (n ? n : 1), &p_), ctx, "wr_device_alloc");
    }
    ~DeviceArray() { wr_device_free(ctx_, p_); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void upload(const T* src) { check(wr_memcpy_htod(ctx_, p_, src, sizeof(T) * n_), ctx_, "wr_memcpy_htod"); }
    void download(T* dst) const { check(wr_memcpy_dtoh(ctx_, dst, p_, sizeof(T) * n_), ctx_, "wr_memcpy_dtoh"); }
    T* get() const { return (T*)p_; }

  private:
    wr_ctx* ctx_;
    size_t n_;
    void* p_ = nullptr;
};

// KDTree (kdtree.h:9-48): fill vertexPositionList / triangleList, build(),
// then intersect -- one ray at a time as in the reference, or a whole batch.
class KDTree {
  public:
    struct Ray {
        vec3f origin;
        vec3f direction;
    };
    struct Triangle {
        unsigned vertexIndices[3];
        void* sourceInformation = nullptr;
    };
    enum Strategy { LOOP, BEST };
    class Condition {
      public:
        virtual ~Condition() = default;
        virtual bool legal(const Ray&, const Triangle&, const float) const { return true; }
    };

    std::vector<vec3f> vertexPositionList;
    std::vector<Triangle> triangleList;

    explicit KDTree(const Context& ctx) : ctx_(ctx.get()) {}
    ~KDTree() { destroy(); }

    void build(Strategy strategy = BEST) {
        wr_scene_destroy(scene_);
        scene_ = nullptr;
        check(wr_scene_create(ctx_, &scene_), ctx_, "wr_scene_create");
        std::vector<uint32_t> idx;
        idx.reserve(3 * triangleList.size());
        for (const Triangle& t : triangleList)
            for (int k = 0; k < 3; k++) idx.push_back(t.vertexIndices[k]);
        wr_material m{WR_DIFFUSE, {0.5f, 0.5f, 0.5f}, 0.f, 1.f};
        check(wr_scene_add_mesh(scene_, &vertexPositionList[0].x, (uint32_t)vertexPositionList.size(), idx.data(),
                                (uint32_t)triangleList.size(), nullptr, &m, nullptr),
              ctx_, "wr_scene_add_mesh");
        check(wr_scene_build(scene_, strategy == BEST ? WR_BEST : WR_LOOP), ctx_, "wr_scene_build");
    }

    void destroy() {
        wr_scene_destroy(scene_);
        scene_ = nullptr;
    }

    // kdtree.cpp:349-353: -1 on a miss, triangleID untouched.
    float intersect(const Ray& ray, unsigned& triangleID, const Condition* condition = nullptr) const {
        if (condition) throw std::runtime_error("KDTree::intersect: GPU traversal takes no host Condition");
        std::vector<Ray> one(1, ray);
        std::vector<wr_hit> h = intersect(one);
        if (h[0].prim == WR_NONE) return -1.f;
        triangleID = h[0].prim;
        return h[0].t;
    }

    // Batch form: one wr_hit per ray (prim, t, barycentrics).
    std::vector<wr_hit> intersect(const std::vector<Ray>& rays) const {
        if (!scene_) throw std::runtime_error("KDTree::intersect: build() first");
        std::vector<wr_hit> out(rays.size());
        if (rays.empty()) return out;
        DeviceArray<wr_ray> dr(ctx_, rays.size());
        DeviceArray<wr_hit> dh(ctx_, rays.size());
        dr.upload(reinterpret_cast<const wr_ray*>(rays.data()));
        check(wr_closest_hit(ctx_, scene_, dr.get(), rays.size(), dh.get(), nullptr), ctx_, "wr_closest_hit");
        dh.download(out.data());
        return out;
    }

  private:
    wr_ctx* ctx_;
    wr_scene* scene_ = nullptr;
};
static_assert(sizeof(KDTree::Ray) == sizeof(wr_ray), "Ray must match wr_ray");

// Scene (scene.h:10-45) with objects in world space.
class Scene {
  public:
    struct ObjSourceInfo {
        unsigned objID;
        unsigned triID;
    };
    struct Object {
        std::vector<vec3f> vertices;
        std::vector<unsigned> indices;  // 3 per triangle
        std::vector<vec3f> normals;     // per vertex, or empty (flat)
        wr_material material;
    };

    explicit Scene(const Context& ctx) : ctx_(ctx.get()) { check(wr_scene_create(ctx_, &scene_), ctx_, "wr_scene_create"); }
    ~Scene() { wr_scene_destroy(scene_); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    unsigned addObject(const Object& o) {
        uint32_t id = 0;
        check(wr_scene_add_mesh(scene_, &o.vertices[0].x, (uint32_t)o.vertices.size(), o.indices.data(),
                                (uint32_t)(o.indices.size() / 3), o.normals.empty() ? nullptr : &o.normals[0].x,
                                &o.material, &id),
              ctx_, "wr_scene_add_mesh");
        return id;
    }
    // Scene::buildKDTree + buildObjKDTrees (scene.cpp:20-65).
    void buildKDTree() { check(wr_scene_build(scene_, WR_BEST), ctx_, "wr_scene_build"); }

    // Scene::intersect without a condition (scene.cpp:102-113).
    float intersect(const KDTree::Ray& ray, ObjSourceInfo& src) const {
        std::vector<KDTree::Ray> one(1, ray);
        DeviceArray<wr_ray> dr(ctx_, 1);
        DeviceArray<wr_hit> dh(ctx_, 1);
        dr.upload(reinterpret_cast<const wr_ray*>(one.data()));
        check(wr_closest_hit(ctx_, scene_, dr.get(), 1, dh.get(), nullptr), ctx_, "wr_closest_hit");
        wr_hit h;
        dh.download(&h);
        return finish(h, src);
    }

    // Scene::intersect with NoSelfCondition(scene, ray) (mcRenderer.cpp:42-44):
    // insideObj / contactObj are object ids or WR_NONE.
    float intersectNoSelf(const KDTree::Ray& ray, unsigned insideObj, unsigned contactObj, ObjSourceInfo& src) const {
        wr_path_ray pr{{ray.origin.x, ray.origin.y, ray.origin.z},
                       {ray.direction.x, ray.direction.y, ray.direction.z}, insideObj, contactObj};
        DeviceArray<wr_path_ray> dr(ctx_, 1);
        DeviceArray<wr_hit> dh(ctx_, 1);
        dr.upload(&pr);
        check(wr_closest_hit_noself(ctx_, scene_, dr.get(), 1, dh.get()), ctx_, "wr_closest_hit_noself");
        wr_hit h;
        dh.download(&h);
        return finish(h, src);
    }

    // Scene::findInsideObject(ray, curObj) (scene.cpp:115-144); WR_NONE = NULL.
    unsigned findInsideObject(const KDTree::Ray& ray, unsigned curObj = WR_NONE) const {
        wr_path_ray pr{{ray.origin.x, ray.origin.y, ray.origin.z},
                       {ray.direction.x, ray.direction.y, ray.direction.z}, WR_NONE, curObj};
        DeviceArray<wr_path_ray> dr(ctx_, 1);
        DeviceArray<uint32_t> dobj(ctx_, 1);
        dr.upload(&pr);
        check(wr_inside_object(ctx_, scene_, dr.get(), 1, dobj.get()), ctx_, "wr_inside_object");
        uint32_t o;
        dobj.download(&o);
        return o;
    }

    wr_scene* get() const { return scene_; }
    wr_ctx* ctx() const { return ctx_; }

  private:
    float finish(const wr_hit& h, ObjSourceInfo& src) const {
        if (h.prim == WR_NONE) return -1.f;
        check(wr_scene_prim_source(scene_, h.prim, &src.objID, &src.triID), ctx_, "wr_scene_prim_source");
        return h.t;
    }
    wr_ctx* ctx_;
    wr_scene* scene_ = nullptr;
};

// PathTracer (pathTracer.h:7-18): renderPixels returns the per-pixel mean
// radiance, row-major, top row first (pathTracer.cpp:10,75).
// useNextEventEstimation starts false as in the reference's constructor
// (pathTracer.h:14); the setter is this mirror's addition.
class PathTracer {
  public:
    PathTracer(const Scene& scene, unsigned spp, unsigned maxLen, uint64_t seed = 1,
               uint32_t mode = WR_WAVEFRONT)
        : scene_(scene), spp_(spp), maxLen_(maxLen), seed_(seed), mode_(mode) {}

    void setNextEventEstimation(bool on) { useNextEventEstimation_ = on; }

    std::vector<vec3f> renderPixels(const wr_camera& cam) const {
        wr_film* film = nullptr;
        check(wr_film_create(scene_.ctx(), cam.width, cam.height, &film), scene_.ctx(), "wr_film_create");
        wr_render_params p{spp_, maxLen_, seed_, mode_, 0, 0, useNextEventEstimation_ ? (uint32_t)WR_RENDER_NEE : 0u};
        int rc = wr_render(scene_.ctx(), scene_.get(), &cam, &p, film, nullptr);
        std::vector<vec3f> px((size_t)cam.width * cam.height);
        if (rc == WR_OK) rc = wr_film_read(film, &px[0].x, nullptr, nullptr, nullptr, nullptr);
        wr_film_destroy(film);
        check(rc, scene_.ctx(), "PathTracer::renderPixels");
        return px;
    }

  private:
    const Scene& scene_;
    unsigned spp_, maxLen_;
    uint64_t seed_;
    uint32_t mode_;
    bool useNextEventEstimation_ = false;
};

// Config (config.h): load reads a reference scene configuration into a Scene
// (objects in Scene::objects order, config.cpp:432-575) and keeps the camera
// and the renderer settings it names.
class Config {
  public:
    explicit Config(const std::string& path) {
        check(wr_config_load(path.c_str(), &cfg_), nullptr, "Config::load");
        check(wr_config_get_info(cfg_, &info_), nullptr, "Config::load");
    }
    ~Config() { wr_config_free(cfg_); }
    Config(const Config&) = delete;
    Config& operator=(const Config&) = delete;

    // adds every object to scene; scene.buildKDTree() follows, as in Config::load
    void load(Scene& scene) const { check(wr_config_add_to_scene(cfg_, scene.get()), scene.ctx(), "Config::load"); }
    const wr_camera& camera() const { return info_.camera; }
    int spp() const { return info_.spp; }
    int maxPathLen() const { return info_.max_len; }
    std::string renderer() const { return info_.renderer; }
    std::string sceneName() const { return info_.scene_name; }

  private:
    wr_config* cfg_ = nullptr;
    wr_config_info info_{};
};

// Film::setBuffer + Film::savePFM (film.h:36-42,100-110) of renderPixels' output.
inline void savePFM(const std::string& filename, const std::vector<vec3f>& pixels, unsigned width, unsigned height) {
    if (pixels.size() != (size_t)width * height) throw std::runtime_error("savePFM: pixel count != width * height");
    check(wr_pfm_write(filename.c_str(), pixels.empty() ? nullptr : &pixels[0].x, width, height), nullptr, "savePFM");
}

}  // namespace walkray
