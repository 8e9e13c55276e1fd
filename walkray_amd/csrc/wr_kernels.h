// wr_kernels.h -- device scene view and kernel launchers (host <-> .hip glue).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_layout.h"

// Device pointers of one uploaded scene.  Record arrays cover the global tree
// followed by every object's own tree (offsets in the tree headers).
struct SceneDev {
    const wr_node* nodes;
    const wr_tri* tris;
    const uint32_t* leaf_up;
    const uint32_t* leaf_cnt;
    const uint32_t* node_up;
    const uint32_t* rec_prim;        // record -> triangle id (global id for the global tree, local for object trees)
    const uint32_t* rec_obj;         // record -> object id
    const float* rec_nrm;            // record -> 9 floats: vertex normals (objects with normals)
    const wr_tree_hdr* hdrs;         // [0] = global tree, [1 + i] = object i
    const wr_object_gpu* objs;
    const uint32_t* obj_has_normals;
    uint32_t n_objs;
    uint32_t max_obj_depth;
};

namespace wr {
hipError_t launch_closest_hit(const SceneDev& S, const wr_tree_hdr& H, const float* rays, uint64_t n, void* hits,
                              unsigned long long* counters, hipStream_t st);
hipError_t launch_closest_hit_path(const SceneDev& S, const wr_tree_hdr& H, const uint32_t* rays, uint64_t n,
                                   void* hits, hipStream_t st);
hipError_t launch_inside_object(const SceneDev& S, uint32_t max_obj_depth, const uint32_t* rays, uint64_t n,
                                uint32_t* out, hipStream_t st);
}  // namespace wr
