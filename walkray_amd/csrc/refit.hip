// refit.hip -- the fast-pass hierarchy of a moving object's tree refit on the
// device, once per pass of a render with moving objects.
//
// The reference rebuilds every tree before every pass (Scene::
// updateSceneForMotionBlur, scene.cpp:67-80).  Here the reference trees are
// rebuilt exactly on the host; the 4-wide hierarchy that only speeds up the
// fast pass keeps the built scene's topology: every leaf reference renamed to
// the record that now holds its triangle, every box recomputed bottom-up from
// the pass's triangle records and rounded outward onto its node's grid, one
// launch per level from the deepest.  The arithmetic is the host refit's
// (builder.cpp: refit_wide, encode_wnode, quant_lo / quant_hi) operation for
// operation, so both give the same nodes bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_kernels.h"

namespace wrd {

__device__ __forceinline__ float min_ref(float a, float b) { return b < a ? b : a; }  // std::min
__device__ __forceinline__ float max_ref(float a, float b) { return a < b ? b : a; }  // std::max

// One bound rounded outward onto origin + q * scale (the decode is one fmaf).
__device__ __forceinline__ uint32_t quant_lo(float v, float origin, float scale) {
    const double q = floor(((double)v - (double)origin) / (double)scale);
    int qi = (int)fmax(0.0, fmin(255.0, q));
    while (qi > 0 && fmaf((float)qi, scale, origin) > v) qi--;
    return (uint32_t)qi;
}
__device__ __forceinline__ uint32_t quant_hi(float v, float origin, float scale) {
    const double q = ceil(((double)v - (double)origin) / (double)scale);
    int qi = (int)fmax(0.0, fmin(255.0, q));
    while (qi < 255 && fmaf((float)qi, scale, origin) < v) qi++;
    return (uint32_t)qi;
}

// Per axis the children's minimum as the grid origin and the smallest power
// of two reaching their maximum within 255 steps.
__device__ __forceinline__ void encode(wr_wnode& W, int n, const float (*cb)[6], const uint32_t* ref) {
    float origin[3], scale[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float lo = cb[0][a], hi = cb[0][3 + a];
        for (int c = 1; c < n; c++) {
            lo = min_ref(lo, cb[c][a]);
            hi = max_ref(hi, cb[c][3 + a]);
        }
        origin[a] = lo;
        const double ext = (double)hi - (double)lo;
        int ex = -100;
        if (ext > 0) frexp(ext / 255.0, &ex);
        float sc = ldexpf(1.0f, ex > -126 ? ex : -126);
        while (fmaf(255.0f, sc, lo) < hi) sc *= 2.0f;
        scale[a] = sc;
    }
    uint32_t q[6] = {0u, 0u, 0u, 0u, 0u, 0u};  // byte c of word k: bound k of child c
#pragma unroll
    for (int c = 0; c < 4; c++) {
        if (c >= n) continue;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            q[a] |= quant_lo(cb[c][a], origin[a], scale[a]) << (8 * c);
            q[3 + a] |= quant_hi(cb[c][3 + a], origin[a], scale[a]) << (8 * c);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        W.origin[a] = origin[a];
        W.scale[a] = scale[a];
    }
#pragma unroll
    for (int k = 0; k < 6; k++)
#pragma unroll
        for (int c = 0; c < 4; c++) W.q[k][c] = (uint8_t)(q[k] >> (8 * c));
#pragma unroll
    for (int c = 0; c < 4; c++) W.ref[c] = c < n ? ref[c] : WR_NONE;
}

// rec_of[prim] = the record of this pass that holds triangle prim.
__global__ __launch_bounds__(256) void k_refit_rec_of(const wr_tri* __restrict__ tris, uint32_t n,
                                                      uint32_t* __restrict__ rec_of) {
    const uint32_t r = blockIdx.x * 256 + threadIdx.x;
    if (r < n) rec_of[tris[r].prim] = r;
}

// One level of nodes (no node of it is another's child): each child's box is
// its triangle's (a leaf reference, renamed from the built scene's record to
// this pass's) or the box its node left in nbox one level deeper.
__global__ __launch_bounds__(256) void k_refit_level(const uint32_t* __restrict__ order, uint32_t n,
                                                     wr_wnode* __restrict__ wn, float* __restrict__ nbox,
                                                     const wr_tri* __restrict__ old_tris,
                                                     const wr_tri* __restrict__ new_tris,
                                                     const uint32_t* __restrict__ rec_of) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t node = order[k];
    wr_wnode W = wn[node];
    float cb[4][6];
    uint32_t ref[4];
    int nc = 0;
    for (int c = 0; c < 4 && W.ref[c] != WR_NONE; c++, nc++) {
        uint32_t r = W.ref[c];
        if (r & WR_LEAF_BIT) {
            const uint32_t rec = rec_of[old_tris[r & WR_FIRST_MASK].prim];
            r = (r & ~WR_FIRST_MASK) | rec;
            const wr_tri& T = new_tris[rec];
#pragma unroll
            for (int a = 0; a < 3; a++) {
                cb[c][a] = min_ref(min_ref(T.v0[a], T.v1[a]), T.v2[a]);
                cb[c][3 + a] = max_ref(max_ref(T.v0[a], T.v1[a]), T.v2[a]);
            }
        } else {
#pragma unroll
            for (int a = 0; a < 6; a++) cb[c][a] = nbox[6 * (size_t)r + a];
        }
        ref[c] = r;
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float lo = cb[0][a], hi = cb[0][3 + a];
        for (int c = 1; c < nc; c++) {
            lo = min_ref(lo, cb[c][a]);
            hi = max_ref(hi, cb[c][3 + a]);
        }
        nbox[6 * (size_t)node + a] = lo;
        nbox[6 * (size_t)node + 3 + a] = hi;
    }
    encode(W, nc, cb, ref);
    wn[node] = W;
}

}  // namespace wrd

namespace wr {

hipError_t launch_wide_refit(const WideRefit& R, hipStream_t st) {
    if (R.n_recs) {
        wrd::k_refit_rec_of<<<(R.n_recs + 255) / 256, 256, 0, st>>>(R.new_tris, R.n_recs, R.rec_of);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (size_t k = 0; k + 1 < R.level_start.size(); k++) {
        const uint32_t a = R.level_start[k], n = R.level_start[k + 1] - a;
        if (n == 0) continue;
        wrd::k_refit_level<<<(n + 255) / 256, 256, 0, st>>>(R.order + a, n, R.wnodes, R.nbox, R.old_tris,
                                                            R.new_tris, R.rec_of);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace wr
