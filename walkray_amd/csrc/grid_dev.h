// grid_dev.h -- the photon hash grid (HashGrid, hashgrid.hxx:15-203) as the
// device sees it: its parameters, the cell hash, the query's eight cells and
// the distance test, shared by the grid kernels (photon.hip) and the vertex
// merging of VCMTracer (bdpt.hip); and the library's grid handle.
#pragma once
#include <hip/hip_runtime.h>

#include "pt_dev.h"
#include "wr_capi_internal.h"

#pragma clang fp contract(off)

namespace wrd {

struct GridDev {
    float bmin[3], bmax[3];
    float inv;     // mInvCellSize
    float r2;      // mRadius * mRadius
    uint32_t n_cells;
};

// GetCellIndex(Vec3i) (hashgrid.hxx:170-178)
__device__ __forceinline__ uint32_t grid_hash(const GridDev& g, int cx, int cy, int cz) {
    const uint32_t x = (uint32_t)cx, y = (uint32_t)cy, z = (uint32_t)cz;
    return ((x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u)) % g.n_cells;
}

// The distance test of HashGrid::Process (hashgrid.hxx:142-147): |q - p|^2 <= r^2.
__device__ __forceinline__ bool grid_near(const GridDev& g, const float* pos, f3 qp, uint32_t pi) {
    const f3 dv = mk(qp.x - pos[3 * pi], qp.y - pos[3 * pi + 1], qp.z - pos[3 * pi + 2]);
    const float dist = len(dv);
    return dist * dist <= g.r2;
}

// The query's cell corner and its nearer neighbours (HashGrid::Process); false
// outside the grid's box.
__device__ __forceinline__ bool query_corner(const GridDev& g, f3 qp, int p[6]) {
    const f3 dmin = mk(qp.x - g.bmin[0], qp.y - g.bmin[1], qp.z - g.bmin[2]);
    const f3 dmax = mk(g.bmax[0] - qp.x, g.bmax[1] - qp.y, g.bmax[2] - qp.z);
    if (dmin.x < 0.f || dmin.y < 0.f || dmin.z < 0.f || dmax.x < 0.f || dmax.y < 0.f || dmax.z < 0.f) return false;
    const f3 cp = mk(g.inv * dmin.x, g.inv * dmin.y, g.inv * dmin.z);
    const f3 cf = mk(floorf(cp.x), floorf(cp.y), floorf(cp.z));
    p[0] = (int)cf.x;
    p[1] = (int)cf.y;
    p[2] = (int)cf.z;
    p[3] = p[0] + (cp.x - cf.x < 0.5f ? -1 : +1);
    p[4] = p[1] + (cp.y - cf.y < 0.5f ? -1 : +1);
    p[5] = p[2] + (cp.z - cf.z < 0.5f ? -1 : +1);
    return true;
}

// Cell j of the query's 8, in the reference's order.
__device__ __forceinline__ uint32_t query_cell(const GridDev& g, const int p[6], int j) {
    return grid_hash(g, (j & 4) ? p[3] : p[0], (j & 2) ? p[4] : p[1], (j & 1) ? p[5] : p[2]);
}

}  // namespace wrd

// A built grid: mCellEnds (inclusive ends per cell) and mIndices over the
// caller's points.
struct wr_hashgrid {
    wr_ctx* ctx = nullptr;
    wrd::GridDev g{};
    const float* pos = nullptr;  // the caller's points
    uint64_t n = 0;
    int* ends = nullptr;         // n_cells
    uint32_t* idx = nullptr;     // n
    uint32_t* defer = nullptr;   // query calls: deferred-query counter and list (grown on demand)
    uint64_t defer_cap = 0;      // entries of defer
};
