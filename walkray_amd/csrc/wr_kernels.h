// wr_kernels.h -- device scene view and kernel launchers (host <-> .hip glue).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "wr_layout.h"

// Device pointers of one uploaded scene.  Record arrays cover the global tree
// followed by every object's own tree (offsets in the tree headers).
struct SceneDev {
    const wr_node* nodes;
    const wr_wnode* wnodes;          // fast-pass hierarchy of every tree (offsets in the headers)
    const wr_tri* tris;
    const uint32_t* leaf_cnt;
    const wr_chain* chains;
    const float* top_boxes;          // chain top records: 6 floats per box (wr_chain::top)
    const float* rec_nrm;            // record -> 9 floats: vertex normals (objects with normals)
    const wr_tree_hdr* hdrs;         // [0] = global tree, [1 + i] = object i
    const wr_object_gpu* objs;
    const uint32_t* obj_has_normals;
    uint32_t n_objs;
    uint32_t max_obj_depth;
    // object-candidate grid for findInsideObject: cell c lists, in descending
    // object order, every object whose root box overlaps the cell
    const uint32_t* cell_start;  // ncells + 1
    const uint32_t* cell_objs;
    float grid_min[3], grid_max[3], grid_inv[3];
    uint32_t grid_dim[3];
    // light sampler (next-event estimation): emissive objects in object order
    const wr_light* lights;
    const float* light_cum;          // SurfaceSampler::areaThreshold (n_lights)
    const float* light_tri_cum;      // per light triangle: AbstractObject::areaThreshold entry
    const uint32_t* light_tri_rec;   // per light triangle: its global-tree record
    uint32_t n_lights;
    float light_total;               // SurfaceSampler::boundArea
};

// Cell coordinate of x on one grid axis: monotone in x, so a point inside a
// closed box maps into the box's cell range (same arithmetic on host and GPU).
#if defined(__HIPCC__)
__host__ __device__
#endif
inline uint32_t wr_grid_cell(float x, float gmin, float inv, uint32_t dim) {
    const float f = (x - gmin) * inv;
    int c = (int)floorf(f);
    if (c < 0) c = 0;
    if (c > (int)dim - 1) c = (int)dim - 1;
    return (uint32_t)c;
}

namespace wr {
// Launch configuration of the persistent closest-hit kernel.  Scheduling
// only: results never depend on it.
struct ChTune {
    uint32_t refill_min = 16;      // idle lanes of a wave before it takes new rays (>= 1)
    uint32_t node_phase_min = 32;  // lanes at wide nodes for a node-only step (0..64)
    uint32_t blocks_per_cu = 0;    // resident workgroups per CU, 0 = as many as fit
    uint32_t lds_levels = 8;       // traversal stack levels kept in LDS: 4, 8 or 16 (deeper ones spill)
    uint32_t batch = 256;          // rays a wave takes from its XCD's work counter at a time (64..4096)
};
// The persistent closest-hit kernel's work counters: one per XCD, each on its
// own 128 B line, and the count of finished waves after them.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kWorkStride = 16;                    // 64-bit words between counters
constexpr uint32_t kWorkDone = kXcds * kWorkStride;     // the finished-wave count
constexpr uint32_t kWorkWords = kWorkDone + kWorkStride;
// work: kWorkWords zeroed 64-bit words owned by the caller's stream; the
// kernel leaves them zeroed again when it ends.
hipError_t launch_closest_hit(const SceneDev& S, const wr_tree_hdr& H, const float* rays, uint64_t n, void* hits,
                              unsigned long long* counters, unsigned long long* work, const ChTune& tune,
                              hipStream_t st);
hipError_t launch_closest_hit_path(const SceneDev& S, const wr_tree_hdr& H, const uint32_t* rays, uint64_t n,
                                   void* hits, hipStream_t st);
hipError_t launch_hit_sources(const SceneDev& S, uint32_t n_tris, const void* hits, uint64_t n, uint32_t* obj,
                              uint32_t* tri, hipStream_t st);
hipError_t launch_inside_object(const SceneDev& S, uint32_t max_obj_depth, const uint32_t* rays, uint64_t n,
                                uint32_t* out, hipStream_t st);

// The fast-pass hierarchy of one tree refit on the device (refit.hip) to the
// records of a pass: wnodes holds the built scene's nodes of this tree (its
// topology and leaf references) and is overwritten; old_tris / new_tris are
// the tree's records in the built scene and in this pass (the same
// triangles, n_recs of them); order / level_start its node indices deepest
// level first (wide_levels); rec_of (n_recs) and nbox (6 floats per node) are
// scratch.
struct WideRefit {
    wr_wnode* wnodes;
    const wr_tri* old_tris;
    const wr_tri* new_tris;
    uint32_t n_recs;
    const uint32_t* order;
    std::vector<uint32_t> level_start;
    uint32_t* rec_of;
    float* nbox;
};
hipError_t launch_wide_refit(const WideRefit& R, hipStream_t st);
}  // namespace wr
