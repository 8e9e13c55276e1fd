// wr_capi_internal.h -- definitions of the opaque C ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/walkray.h"
#include "wr_host.h"
#include "wr_kernels.h"
#include "wr_pt.h"
#include "wr_sceneio.h"

struct wr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    unsigned long long* d_work = nullptr;  // work counters of the persistent closest-hit kernel (self-resetting)
    wr::Engine engine;  // wavefront engine of wr_render, reused across calls
    uint32_t tuned_slots = 0, tuned_grid = 0;  // wr_autotune_render's choice (0 = defaults)
    uint32_t tuned_lds = 8;                    // wr_autotune_render's traversal stack levels in LDS
    wr::StageStats last_stage;                 // the last WR_RENDER_STAGE_STATS render's per-stage figures
};

namespace wr {
struct HostObject {
    std::vector<float> verts, normals;  // normals: 9 per triangle (corners), empty for flat shading
    std::vector<uint32_t> idx;
    wr_material mat;
    // a moving object (wr_scene_set_motion): verts / normals are placed by the
    // transform of the moment, local_* keep the mesh in its own space
    bool moving = false;
    io::Motion motion;
    std::vector<float> local_verts, local_normals;
};
int fail(wr_ctx* ctx, int code, const std::string& msg);
int hip_fail(wr_ctx* ctx, hipError_t e, const char* what);
void scene_free_device(wr_scene* s);
// Places a moving object with transform M: vertices M (v, 1), corner normals
// transpose(inverse(M)) (n, 0) (mesh.h:47-77,106).
void place_object(HostObject& o, const io::Mat4& M);
// Uniform in [0, 1): the time of object obj in pass `pass` (AbstractObject::
// updateTransform, abstractObj.cpp:70-76), from the counter RNG at vertex
// 0xFFFFFFFE with the object id in the pixel slot.
float motion_time(uint64_t seed, uint32_t obj, uint32_t pass);
// Looks up, on `device` (current), the address of the fixed work list of
// wr_hashgrid_query_noalloc; once per device, from wr_ctx_create.
hipError_t noalloc_scratch_init(int device);
// What a scene keeps for its per-pass rebuilds (moving objects).  Pinned host
// buffers: every pass stages the concatenated scene arrays in the same pages,
// so they are neither faulted in again nor copied through pageable memory;
// slot k grows (with a quarter of headroom) when a pass needs more.  The
// device refit of the fast-pass hierarchies: per tree of the built scene its
// wide nodes deepest level first (made on the first pass that rebuilds the
// tree), and the refit's scratch.
struct RebuildCache {
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
    };
    std::vector<Buf> bufs;
    std::vector<uint32_t*> d_order;                // per tree, nullptr until needed
    std::vector<std::vector<uint32_t>> level_start;
    uint32_t* d_rec_of = nullptr;
    size_t rec_cap = 0;
    float* d_nbox = nullptr;
    size_t box_cap = 0;
    std::mutex mu;  // held by the rebuild that uses all this
    void* get(size_t slot, size_t bytes);  // nullptr when hipHostMalloc fails
    ~RebuildCache();
};
// wr_scene_build; the trees of objects that do not move are copied from
// `reuse` (nullable) when it is the same scene, built, and the host arrays
// are staged in reuse's pinned buffers.
int scene_build(wr_scene* s, const wr_scene* reuse);
// Scene::updateSceneForMotionBlur (scene.cpp:67-80) for pass `pass` of a
// render with `seed`: out (empty) receives s's objects, the moving ones placed
// at their time for the pass, built with scene_build(out, s).
int scene_pass(const wr_scene* s, uint64_t seed, uint32_t pass, wr_scene* out);

// VCMTracer (bdpt.hip): the merging radius of a pass (vcmTracer.cpp:30), and
// passes [pass0, pass0 + npass) side by side [sub-paths, light points, one
// hash grid per pass, connections and merges, film] on the context stream.
// verts / pix / q as for launch_bdpt_passes; vb: the light-point buffers (cap
// points of 3 floats and 2 words), the per-lane counts and offsets (lanes + 1
// each), the per-pass starts (npass + 1), the scan's scratch, and the device
// copy of the per-pass grids (npass records of vcm_grid_bytes()).
struct VcmBufs {
    uint32_t* pc;
    uint32_t* offs;
    uint32_t* starts;
    void* scan;
    size_t scan_bytes;
    float* pos;
    uint32_t* src;
    uint64_t cap;
    void* grids;
    uint32_t* mq;      // VCM merges queued for k_vcm_merge: 2 words each, then the count
    uint32_t mq_cap;
};
float vcm_radius(float base, uint32_t pass);
size_t vcm_scan_bytes(uint32_t lanes);
size_t vcm_grid_bytes();
int launch_vcm_passes(wr_ctx* ctx, const SceneDev& S, const wr_tree_hdr& H, const CamHost& ch, uint32_t max_depth,
                      uint64_t seed, uint32_t pass0, uint32_t npass, float base_radius, uint32_t maxv, void* verts,
                      float* pix, const BdptQueue& q, const VcmBufs& vb, const FilmBufs& film);
// PhotonMap (bdpt.hip): passes [pass0, pass0 + npass) side by side [light
// sub-paths, light points, one hash grid per pass, eye merge paths and density
// estimates, film]; buffers as for launch_vcm_passes (counts: 2 per lane).
int launch_ppm_passes(wr_ctx* ctx, const SceneDev& S, const CamHost& ch, uint32_t max_depth, uint64_t seed,
                      uint32_t pass0, uint32_t npass, float base_radius, uint32_t maxv, void* verts, uint32_t* counts,
                      float* pix, const VcmBufs& vb, const FilmBufs& film);
}  // namespace wr

struct wr_scene {
    wr_ctx* ctx = nullptr;
    std::vector<wr::HostObject> objs;
    bool built = false;
    std::vector<wr::FlatTree> trees;       // [0] global, [1 + i] object i
    std::vector<uint32_t> prim_base;       // first global prim of each object (+ total)
    std::vector<wr_tree_hdr> hdrs;         // host copy of the device headers
    std::vector<wr_object_gpu> host_objs;  // host copy of the device object table
    std::vector<void*> dbufs;
    SceneDev dev{};
    wr::ChTune ch_tune;                    // closest-hit launch configuration (wr_autotune_closest_hit)
    uint64_t device_bytes = 0;
    uint32_t max_obj_depth = 0;
    double build_seconds = 0;
    mutable std::unique_ptr<wr::RebuildCache> rebuild = std::make_unique<wr::RebuildCache>();  // for the per-pass rebuilds
};
