// trace.hip -- batch closest-hit and inside-object kernels (gfx950).
//
//  k_closest_hit        KDTree::intersect without Condition (kdtree.cpp:349-353),
//                       as used by Scene::intersect for camera rays (camera.cpp:52)
//  k_closest_hit_path   the same query with NoSelfCondition (mcRenderer.cpp:42-44,
//                       noself.cpp:13-20)
//  k_inside_object      Scene::findInsideObject (scene.cpp:115-144)
//
// One lane per ray, wave64, 256-lane workgroups; per-lane node stack in LDS;
// workgroup ids are remapped so each XCD walks a contiguous range of rays
// (neighbouring rays share tree nodes, and each XCD has its own L2).
#include <hip/hip_runtime.h>

#include "trace_dev.h"
#include "wr_kernels.h"

#pragma clang fp contract(off)

namespace wrd {

using wr::kWorkDone;
using wr::kWorkStride;
using wr::kXcds;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t G) {
    const uint32_t xcd = b & 7u, local = b >> 3;
    const uint32_t per = G >> 3, rem = G & 7u;
    return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

// Gram-system barycentrics of Mesh::getWorldNormal (mesh.h:67-72) at o + d t.
__device__ __forceinline__ void bary(const wr_tri& T, const f3& o, const f3& d, float t, float& u1, float& u2) {
    const f3 v0 = mk(T.v0[0], T.v0[1], T.v0[2]);
    const f3 b1 = sub(mk(T.v1[0], T.v1[1], T.v1[2]), v0);
    const f3 b2 = sub(mk(T.v2[0], T.v2[1], T.v2[2]), v0);
    const f3 v = sub(add(o, muls(d, t)), v0);
    const float d12 = dot(b1, b2);
    const float l1 = len(b1);
    const float l2 = len(b2);
    const float vb1 = dot(v, b1), vb2 = dot(v, b2);
    u2 = (vb1 * d12 - vb2 * l1 * l1) / (d12 * d12 - l1 * l1 * l2 * l2);
    u1 = (vb1 - u2 * d12) / (l1 * l1);
}

__device__ __forceinline__ void add_counters(unsigned long long* out, const Counters& c) {
    unsigned long long a = c.nodes, b = c.tris, f = c.fallback;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        f += __shfl_xor(f, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(out + 0, a);
        atomicAdd(out + 1, b);
        atomicAdd(out + 2, f);
    }
}

// Fast pass of the batch closest-hit query as a persistent kernel: each wave
// takes rays in batches of `batch` from a work counter and refills a lane with
// the next ray as soon as its traversal ends, so a wave never waits for its
// slowest ray.  The result of the fast pass (record, t, first record of the
// leaf, overflow flag) goes to hits[ray]; k_closest_hit_finish completes it.
// Work queues: the rays are cut into kXcds contiguous ranges, range q handed
// out by the counter work[q * kWorkStride] to the workgroups of XCD q
// (workgroups are dispatched to the XCDs round-robin), so neighbouring rays
// share one XCD's L2 and each counter takes an eighth of the atomics.  A wave
// whose range is used up moves on to the next range (and so on round the
// ranges); once it has found all of them used up it is done.  work[kWorkDone]
// counts finished waves; the last wave out zeroes every counter for the next
// launch.  Compiled for 8 waves per SIMD (64
// VGPRs): more waves in flight hide more of the node-fetch latency.
// Scheduling knobs (ChTune, chosen by the autotuner): idle lanes are refilled
// once at least refill_min are idle; a wave runs node-only steps while at
// least node_phase_min lanes sit at wide nodes (a mixed step pays for both the
// node code and the triangle code).
// D: stack levels in LDS (the autotuner's stack-depth knob).
template <bool COUNT, int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_closest_hit(
    SceneDev S, wr_tree_hdr H, const float* __restrict__ rays, uint64_t n, uint4* __restrict__ hits,
    unsigned long long* counters, unsigned long long* work, uint32_t refill_min, uint32_t node_phase_min,
    uint32_t batch) {
    const Tree tr = make_tree(H, S);
    StackSpillT<D> spill;
    LaneStackT<D> st(&spill);
    Counters cnt{0, 0, 0};
    const NoSelf ns{};
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint64_t bnext = 0, bend = 0;  // the wave's current batch [bnext, bend)
    bool exhausted = false;
    bool active = false, overflow = false;
    uint32_t q = blockIdx.x % kXcds;  // the work range of this workgroup's XCD
    uint32_t used_up = 0;             // ranges this wave has found used up (each once: they stay so)
    uint64_t rid = 0;
    uint32_t ref = 0;
    RayPre r{};
    HitState h{};
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!exhausted && (uint32_t)__popcll(idle) >= refill_min) {
            while (idle != 0) {
                if (bnext == bend) {
                    // next batch: from range q, or the next range once q is used up
                    bool got = false;
                    while (used_up < kXcds) {
                        const uint64_t lo = n * q / kXcds, hi = n * (q + 1) / kXcds;
                        unsigned long long base = 0;
                        if (lane == 0) base = atomicAdd(work + kWorkStride * q, (unsigned long long)batch);
                        base = __shfl(base, 0);
                        if (lo + base < hi) {
                            bnext = lo + base;
                            bend = lo + base + batch < hi ? lo + base + batch : hi;
                            got = true;
                            break;
                        }
                        q = q + 1 < kXcds ? q + 1 : 0;
                        used_up++;
                    }
                    if (!got) {
                        exhausted = true;
                        break;
                    }
                }
                const uint64_t avail = bend - bnext;
                const uint64_t rank = (uint64_t)__popcll(idle & below);
                if (!active && rank < avail) {
                    rid = bnext + rank;
                    const float* rp = rays + 6 * rid;
                    r = make_ray(mk(rp[0], rp[1], rp[2]), mk(rp[3], rp[4], rp[5]), tr.scale);
                    overflow = false;
                    if (wide_begin(tr, r, st, h, ref)) active = true;
                    else hits[rid] = make_uint4(WR_NONE, 0u, 0u, 0u);  // misses the root box
                }
                const uint64_t took = (uint64_t)__popcll(idle);
                bnext += took < avail ? took : avail;
                idle = __ballot(!active);
            }
        }
        if (__ballot(active) == 0) break;  // only once the counter is exhausted
        // node phase: while at least node_phase_min lanes sit at wide nodes, only
        // those lanes step (no triangle code runs); then one step for every lane
        for (;;) {
            const bool at_node = active && !(ref & WR_LEAF_BIT);
            if ((uint32_t)__popcll(__ballot(at_node)) < node_phase_min) break;
            if (at_node && !wide_step<false, COUNT>(tr, r, ns, st, h, ref, overflow, cnt)) {
                hits[rid] = make_uint4(h.rec, __float_as_uint(h.t), h.leaf, overflow ? 1u : 0u);
                active = false;
            }
        }
        if (active && !wide_step<false, COUNT>(tr, r, ns, st, h, ref, overflow, cnt)) {
            hits[rid] = make_uint4(h.rec, __float_as_uint(h.t), h.leaf, overflow ? 1u : 0u);
            active = false;
        }
    }
    if (COUNT) add_counters(counters, cnt);
    if (lane == 0) {
        const unsigned long long waves = (unsigned long long)gridDim.x * (blockDim.x / 64);
        if (atomicAdd(work + kWorkDone, 1ull) == waves - 1) {
            for (uint32_t k = 0; k < kXcds; k++) atomicExch(work + kWorkStride * k, 0ull);
            atomicExch(work + kWorkDone, 0ull);
        }
    }
}

// Completion of k_closest_hit's records: binary fast pass after a wide-stack
// overflow, chain verification, exact fallback, barycentrics, output.  One
// lane per ray, so the verification walks of a wave run together.
template <bool COUNT, int D>
__global__ __launch_bounds__(256) void k_closest_hit_finish(SceneDev S, wr_tree_hdr H,
                                                            const float* __restrict__ rays, uint64_t n,
                                                            uint4* __restrict__ hits, unsigned long long* counters) {
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const uint64_t i = (uint64_t)b * 256 + threadIdx.x;
    Counters cnt{0, 0, 0};
    StackSpillT<D> spill;
    LaneStackT<D> st(&spill);
    if (i < n) {
        const uint4 f = hits[i];
        HitState h{__uint_as_float(f.y), f.x, f.z};
        const bool overflow = f.w != 0;
        uint4 out = make_uint4(WR_NONE, __float_as_uint(-1.f), 0u, 0u);
        if (h.rec != WR_NONE || overflow) {
            const float* rp = rays + 6 * i;
            const f3 o = mk(rp[0], rp[1], rp[2]);
            const f3 d = mk(rp[3], rp[4], rp[5]);
            const Tree tr = make_tree(H, S);
            const RayPre r = make_ray(o, d, tr.scale);
            const NoSelf ns{};
            bool failed;
            // the winner's record is fetched alongside the verification's first fetch
            const uint32_t rec0 = h.rec;
            wr_tri T{};
            if (rec0 != WR_NONE) T = tr.tris[rec0];
            finish_hit<false, COUNT>(tr, r, ns, st, cnt, h, overflow, failed);
            if (h.rec != WR_NONE) {
                if (h.rec != rec0) T = tr.tris[h.rec];
                float u1, u2;
                bary(T, o, d, h.t, u1, u2);
                out = make_uint4(T.prim, __float_as_uint(h.t), __float_as_uint(u1), __float_as_uint(u2));
            }
            if (failed) out.x = 0xFFFFFFFEu;  // stack overflow: reported, never silently wrong
        }
        hits[i] = out;
    }
    if (COUNT) add_counters(counters, cnt);
}

__global__ __launch_bounds__(256) void k_closest_hit_path(SceneDev S, wr_tree_hdr H, const uint32_t* __restrict__ rays,
                                                          uint64_t n, uint4* __restrict__ hits) {
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    const uint64_t i = (uint64_t)b * 256 + threadIdx.x;
    if (i >= n) return;
    StackSpill spill;
    LaneStack st(&spill);
    const uint32_t* rp = rays + 8 * i;
    const f3 o = mk(__uint_as_float(rp[0]), __uint_as_float(rp[1]), __uint_as_float(rp[2]));
    const f3 d = mk(__uint_as_float(rp[3]), __uint_as_float(rp[4]), __uint_as_float(rp[5]));
    const Tree tr = make_tree(H, S);
    const RayPre r = make_ray(o, d, tr.scale);
    NoSelf ns{rp[6], rp[7], S.rec_nrm + 9 * (size_t)H.rec_base, S.obj_has_normals};
    Counters cnt{0, 0, 0};
    bool failed;
    const HitState h = closest_hit<true, false>(tr, r, ns, st, cnt, failed);
    uint4 out;
    if (h.rec != WR_NONE) {
        float u1, u2;
        bary(tr.tris[h.rec], o, d, h.t, u1, u2);
        out = make_uint4(tr.tris[h.rec].prim, __float_as_uint(h.t), __float_as_uint(u1),
                         __float_as_uint(u2));
    } else {
        out = make_uint4(WR_NONE, __float_as_uint(-1.f), 0u, 0u);
    }
    if (failed) out.x = 0xFFFFFFFEu;
    hits[i] = out;
}

__global__ __launch_bounds__(256) void k_inside_object(SceneDev S, const uint32_t* __restrict__ rays, uint64_t n,
                                                       uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    StackSpill spill;
    LaneStack st(&spill);
    const uint32_t* rp = rays + 8 * i;
    const f3 o = mk(__uint_as_float(rp[0]), __uint_as_float(rp[1]), __uint_as_float(rp[2]));
    const f3 d = mk(__uint_as_float(rp[3]), __uint_as_float(rp[4]), __uint_as_float(rp[5]));
    out[i] = find_inside(S, o, d, rp[7], st);
}

// Global triangle id -> (object, triangle of the object) for a batch of hits
// (ObjSourceInfo, scene.cpp:34-36,110): the last object whose first global
// triangle is <= prim, found by binary search over the object table.
__global__ __launch_bounds__(256) void k_hit_sources(const wr_object_gpu* __restrict__ objs, uint32_t n_objs,
                                                     uint32_t n_tris, const uint4* __restrict__ hits, uint64_t n,
                                                     uint32_t* __restrict__ obj, uint32_t* __restrict__ tri) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t prim = hits[i].x;
    uint32_t o = WR_NONE, t = WR_NONE;
    if (prim < n_tris) {
        uint32_t lo = 0, hi = n_objs;  // first object with prim_base > prim lies in [lo, hi]
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (objs[mid].prim_base <= prim) lo = mid + 1;
            else hi = mid;
        }
        o = lo - 1;
        t = prim - objs[o].prim_base;
    }
    obj[i] = o;
    tri[i] = t;
}

}  // namespace wrd

namespace wr {

hipError_t launch_hit_sources(const SceneDev& S, uint32_t n_tris, const void* hits, uint64_t n, uint32_t* obj,
                              uint32_t* tri, hipStream_t st) {
    if (n == 0) return hipSuccess;
    wrd::k_hit_sources<<<(uint32_t)((n + 255) / 256), 256, 0, st>>>(S.objs, S.n_objs, n_tris, (const uint4*)hits, n,
                                                                     obj, tri);
    return hipGetLastError();
}

// Workgroups of the persistent kernel with D LDS stack levels: per_cu per CU
// (0: as many as fit).
template <int D>
static uint32_t persistent_blocks(uint64_t n, uint32_t per_cu_want) {
    static int cus = 0, per_cu_max = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_max, wrd::k_closest_hit<false, D>, 256,
                                                         wrd::stack_lds_bytes<D>()) != hipSuccess ||
            cus <= 0 || per_cu_max <= 0) {
            cus = 256;
            per_cu_max = 1;
        }
    }
    const uint32_t per_cu = per_cu_want && per_cu_want < (uint32_t)per_cu_max ? per_cu_want : (uint32_t)per_cu_max;
    const uint64_t resident = (uint64_t)cus * per_cu;
    const uint64_t need = (n + 255) / 256;
    return (uint32_t)(need < resident ? need : resident);
}

template <int D>
static hipError_t launch_closest_hit_d(const SceneDev& S, const wr_tree_hdr& H, const float* rays, uint64_t n,
                                       void* hits, unsigned long long* counters, unsigned long long* work,
                                       uint32_t per_cu, uint32_t rmin, uint32_t nmin, uint32_t batch,
                                       hipStream_t st) {
    const uint32_t pgrid = persistent_blocks<D>(n, per_cu);
    const uint32_t grid = (uint32_t)((n + 255) / 256);
    const size_t lds = wrd::stack_lds_bytes<D>();
    if (counters) {
        wrd::k_closest_hit<true, D><<<pgrid, 256, lds, st>>>(S, H, rays, n, (uint4*)hits, counters, work, rmin, nmin,
                                                                 batch);
        wrd::k_closest_hit_finish<true, D><<<grid, 256, lds, st>>>(S, H, rays, n, (uint4*)hits, counters);
    } else {
        wrd::k_closest_hit<false, D><<<pgrid, 256, lds, st>>>(S, H, rays, n, (uint4*)hits, nullptr, work, rmin, nmin,
                                                                  batch);
        wrd::k_closest_hit_finish<false, D><<<grid, 256, lds, st>>>(S, H, rays, n, (uint4*)hits, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_closest_hit(const SceneDev& S, const wr_tree_hdr& H, const float* rays, uint64_t n, void* hits,
                              unsigned long long* counters, unsigned long long* work, const ChTune& tune,
                              hipStream_t st) {
    if (n == 0) return hipSuccess;
    // refill_min in [1, 64]; node_phase_min in [1, 65] (0 would never leave the node phase; 65 disables it)
    const uint32_t rmin = tune.refill_min < 1 ? 1 : (tune.refill_min > 64 ? 64 : tune.refill_min);
    const uint32_t nmin = tune.node_phase_min < 1 ? 1 : (tune.node_phase_min > 65 ? 65 : tune.node_phase_min);
    // rays per work-counter atomic in [64, 4096]
    const uint32_t batch = tune.batch < 64 ? 64 : (tune.batch > 4096 ? 4096 : tune.batch);
    const uint32_t pc = tune.blocks_per_cu;
    switch (tune.lds_levels) {
        case 4:
            return launch_closest_hit_d<4>(S, H, rays, n, hits, counters, work, pc, rmin, nmin, batch, st);
        case 16:
            return launch_closest_hit_d<16>(S, H, rays, n, hits, counters, work, pc, rmin, nmin, batch, st);
        default:
            return launch_closest_hit_d<8>(S, H, rays, n, hits, counters, work, pc, rmin, nmin, batch, st);
    }
}

hipError_t launch_closest_hit_path(const SceneDev& S, const wr_tree_hdr& H, const uint32_t* rays, uint64_t n,
                                   void* hits, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)((n + 255) / 256);
    wrd::k_closest_hit_path<<<grid, 256, wrd::kStackLdsBytes, st>>>(S, H, rays, n, (uint4*)hits);
    return hipGetLastError();
}

hipError_t launch_inside_object(const SceneDev& S, uint32_t max_obj_depth, const uint32_t* rays, uint64_t n,
                                uint32_t* out, hipStream_t st) {
    (void)max_obj_depth;
    if (n == 0) return hipSuccess;
    const uint32_t grid = (uint32_t)((n + 255) / 256);
    wrd::k_inside_object<<<grid, 256, wrd::kStackLdsBytes, st>>>(S, rays, n, out);
    return hipGetLastError();
}

}  // namespace wr
