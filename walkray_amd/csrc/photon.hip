// photon.hip -- light sub-paths and the photon hash grid (gfx950), with their
// C ABI entry points.
//
//  k_light_paths    one lane per light path: Scene::genLightSample as
//                   AbstractObject::genSurfaceSample (abstractObj.cpp:18-52)
//                   and MCRenderer::samplePath (mcRenderer.cpp:16-57) from it,
//                   every Ray of the Path written as a 64 B record.
//  photon collect   count, scan and write the light points VCM / PPM keep
//                   (vcmTracer.cpp:42-55): deterministic, path-major order.
//  hash grid        HashGrid::Build (hashgrid.hxx:24-75) as bbox reduction,
//                   per-cell counts, an inclusive scan (mCellEnds) and a stable
//                   radix sort of (cell, point) pairs (mIndices: the serial
//                   scatter keeps input order inside a cell, so does a stable
//                   sort); HashGrid::Process (hashgrid.hxx:77-158) as a range
//                   query, one lane per query point.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "grid_dev.h"
#include "pt_dev.h"
#include "wr_capi_internal.h"

#pragma clang fp contract(off)

namespace wrd {

// Vertex counters of a light path's uniforms: kLightVertex (light pick and
// point), kLightVertex + 1 (direction), kLightVertex + 2 + d (scatter at depth d).
constexpr uint32_t kLightVertex = 0x80000000u;

__device__ __forceinline__ void put_vertex(wr_path_vertex* V, uint32_t& nv, uint32_t maxv, f3 o, f3 d, f3 rad,
                                           float op, float dp, uint32_t contact, uint32_t ctri, uint32_t inside,
                                           bool delta) {
    if (nv < maxv) {
        wr_path_vertex& v = V[nv];
        v.origin[0] = o.x; v.origin[1] = o.y; v.origin[2] = o.z;
        v.direction[0] = d.x; v.direction[1] = d.y; v.direction[2] = d.z;
        v.radiance[0] = rad.x; v.radiance[1] = rad.y; v.radiance[2] = rad.z;
        v.origin_prob = op;
        v.direction_prob = dp;
        v.contact_obj = contact;
        v.contact_tri = contact == WR_NONE ? 0u : ctri;
        v.inside_obj = inside;
        v.flags = delta ? (uint32_t)WR_VERTEX_DELTA : 0u;
        v.pad = 0u;
    }
    nv++;
}

// terminateRay (mcRenderer.cpp:20-27): the origin of the ray passed in.
__device__ __forceinline__ void put_terminate(wr_path_vertex* V, uint32_t& nv, uint32_t maxv, f3 o) {
    const f3 z = mk(0.f, 0.f, 0.f);
    put_vertex(V, nv, maxv, o, z, z, 1.f, 1.f, WR_NONE, 0u, WR_NONE, true);
}

// 4 waves per SIMD (128 VGPRs, the rest spilled): 12.8 -> 9.3 ms for 2M light paths on the field.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_light_paths(SceneDev S, uint32_t max_depth, uint64_t seed, uint32_t pass,
                                                     uint32_t path0, uint32_t n, uint32_t maxv, wr_path_vertex* out,
                                                     uint32_t* counts) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    const uint32_t path = path0 + k;
    wr_path_vertex* V = out + (size_t)k * maxv;
    uint32_t nv = 0;
    float rn[4];
    rng4(seed, path, pass, kLightVertex, rn);
    const LightSample L = sample_light(sc, rn);
    f3 o = mk(0.f, 0.f, 0.f), d = o;
    uint32_t inside = WR_NONE;
    PHit hit;
    hit.rec = WR_NONE;
    hit.obj = WR_NONE;
    hit.t = -1.f;
    if (!L.ok) {  // the reference's empty-mesh light ray: no direction, radiance or contact
        put_vertex(V, nv, maxv, o, d, o, 1.f, 1.f, WR_NONE, 0u, WR_NONE, false);
    } else {
        rng4(seed, path, pass, kLightVertex + 1, rn);
        // UniformSphereSampler::sample (uniformSphereSampler.h:6-20)
        const float term1 = (float)(2.0 * kPi * (double)rn[0]);
        const float term2 = 2.f * sqrtf(rn[1] - rn[1] * rn[1]);
        d = mk(cosf(term1) * term2, 1.f - 2.f * rn[1], sinf(term1) * term2);
        float dp = (float)(kInvPi * 0.25f);
        if (dot(d, L.n) < 0.f) d = mk(-d.x, -d.y, -d.z);
        o = L.o;
        inside = find_inside(S, o, d, L.obj, st);
        dp *= 2.f;
        const PHit h = trace_scene<true>(sc, o, d, inside, L.obj, st);
        if (h.rec != WR_NONE && h.t > 0.f) hit = h;  // intersectObj is set only for dist > 0
        put_vertex(V, nv, maxv, o, d, L.rad, L.op, dp, L.obj, L.tri, inside, false);
    }
    for (uint32_t depth = 0;; depth++) {
        const uint32_t self = inside != WR_NONE ? inside : hit.obj;
        if (self == WR_NONE || hit.rec == WR_NONE) {  // nothing to scatter (an open inside object as well)
            put_terminate(V, nv, maxv, o);
            break;
        }
        rng4(seed, path, pass, kLightVertex + 2 + depth, rn);
        const Out r = scatter(sc, self, o, d, hit, inside, rn, st);
        const uint32_t ctri = S.tris[hit.rec].prim - S.objs[hit.obj].prim_base;  // intersectTriangleID
        if ((double)len(r.d) < 0.5) {
            put_vertex(V, nv, maxv, r.o, r.d, r.rad, r.op, r.dp, r.contact, ctri, r.inside, r.delta);
            break;
        }
        if (depth + 1 > max_depth) {
            put_terminate(V, nv, maxv, o);
            break;
        }
        const PHit nh = trace_scene<true>(sc, r.o, r.d, r.inside, r.contact, st);
        put_vertex(V, nv, maxv, r.o, r.d, r.rad, r.op, r.dp, r.contact, ctri, r.inside, r.delta);
        if (nh.rec == WR_NONE) {
            put_terminate(V, nv, maxv, o);
            break;
        }
        o = r.o;
        d = r.d;
        inside = r.inside;
        hit = nh;
    }
    counts[k] = nv;
}

// Light points of one path: vertices 1.. before the first emissive contact,
// without the delta ones.
__device__ __forceinline__ bool photon_at(const SceneDev& S, const wr_path_vertex& v, bool& stop) {
    stop = v.contact_obj != WR_NONE && S.objs[v.contact_obj].mat == WR_MAT_EMISSIVE;
    return !stop && !(v.flags & WR_VERTEX_DELTA);
}

__global__ __launch_bounds__(256) void k_photon_count(SceneDev S, const wr_path_vertex* V, const uint32_t* counts,
                                                      uint32_t n, uint32_t maxv, uint32_t* pc) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t c = min(counts[k], maxv);
    uint32_t m = 0;
    for (uint32_t i = 1; i < c; i++) {
        bool stop;
        const bool keep = photon_at(S, V[(size_t)k * maxv + i], stop);
        if (stop) break;
        m += keep;
    }
    pc[k] = m;
}

__global__ __launch_bounds__(256) void k_photon_write(SceneDev S, const wr_path_vertex* V, const uint32_t* counts,
                                                      uint32_t n, uint32_t maxv, const uint64_t* offs,
                                                      uint64_t base, uint64_t cap, float* pos, uint32_t* src) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint32_t c = min(counts[k], maxv);
    uint64_t j = base + offs[k];
    for (uint32_t i = 1; i < c; i++) {
        const wr_path_vertex& v = V[(size_t)k * maxv + i];
        bool stop;
        const bool keep = photon_at(S, v, stop);
        if (stop) break;
        if (!keep) continue;
        if (j < cap) {
            pos[3 * j] = v.origin[0];
            pos[3 * j + 1] = v.origin[1];
            pos[3 * j + 2] = v.origin[2];
            src[2 * j] = k;
            src[2 * j + 1] = i;
        }
        j++;
    }
}

__global__ __launch_bounds__(256) void k_fill_leading(uint64_t n, uint64_t cap, float* pos, uint32_t* src) {
    for (uint64_t j = blockIdx.x * 256ull + threadIdx.x; j < n && j < cap; j += (uint64_t)gridDim.x * 256) {
        pos[3 * j] = 0.f;
        pos[3 * j + 1] = 0.f;
        pos[3 * j + 2] = 0.f;
        src[2 * j] = WR_NONE;
        src[2 * j + 1] = WR_NONE;
    }
}

// ---- hash grid ------------------------------------------------------------
// Order-preserving map of floats onto signed ints, for atomic min / max.
__device__ __forceinline__ int fkey(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
static inline int fkey_host(float f) {
    int i;
    std::memcpy(&i, &f, 4);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
static inline float fkey_inv(int k) {
    const int i = k >= 0 ? k : k ^ 0x7FFFFFFF;
    float f;
    std::memcpy(&f, &i, 4);
    return f;
}

__global__ __launch_bounds__(256) void k_grid_bbox(const float* pos, uint64_t n, int* keys) {
    float mn[3] = {1e36f, 1e36f, 1e36f}, mx[3] = {-1e36f, -1e36f, -1e36f};
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float v = pos[3 * i + a];
            mn[a] = fminf(mn[a], v);
            mx[a] = fmaxf(mx[a], v);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
        }
    }
    // the block's four waves meet in LDS: one set of atomics per block
    __shared__ float part[4][6];
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; a++) {
            part[w][a] = mn[a];
            part[w][3 + a] = mx[a];
        }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float lo = part[0][a], hi = part[0][3 + a];
            for (int k = 1; k < 4; k++) {
                lo = fminf(lo, part[k][a]);
                hi = fmaxf(hi, part[k][3 + a]);
            }
            atomicMin(keys + a, fkey(lo));
            atomicMax(keys + 3 + a, fkey(hi));
        }
    }
}

__global__ __launch_bounds__(256) void k_grid_keys(GridDev g, const float* pos, uint32_t n, uint32_t* keys,
                                                   uint32_t* vals, uint32_t* cell_count) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    // GetCellIndex(Vec3f) (hashgrid.hxx:180-190)
    const float dx = pos[3 * i] - g.bmin[0], dy = pos[3 * i + 1] - g.bmin[1], dz = pos[3 * i + 2] - g.bmin[2];
    const uint32_t c = grid_hash(g, (int)floorf(g.inv * dx), (int)floorf(g.inv * dy), (int)floorf(g.inv * dz));
    keys[i] = c;
    vals[i] = i;
    // one atomic for a wave whose points share a cell (many coincident points,
    // such as VCM's leading ones), else one per point
    const uint64_t act = __ballot(1);
    const uint32_t leader = __ffsll((unsigned long long)act) - 1;
    const uint64_t same = __ballot(c == __shfl(c, leader));
    if (same == act) {
        if ((threadIdx.x & 63) == leader) atomicAdd(cell_count + c, (uint32_t)__popcll(act));
    } else {
        atomicAdd(cell_count + c, 1u);
    }
}

// Ranges longer than this are scanned by the whole wave, 64 points a step, one
// query at a time (a crowded cell, or one that shares a hash bucket with it).
constexpr int kLongRange = 64;

// Ranges longer than this hand their whole query to k_grid_query_wide (one
// workgroup per query), so that no wave walks such a range alone.
constexpr int kHugeRange = 2048;

// defer: [0] the number of deferred queries, then their indices.
__global__ __launch_bounds__(256) void k_grid_query(GridDev g, const float* pos, const int* ends, const uint32_t* idx,
                                                    const float* q, uint32_t nq, uint32_t cap, uint32_t* counts,
                                                    uint32_t* out, uint32_t* defer) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    // every lane stays to the end: the long ranges are scanned by the whole wave
    bool valid = k < nq;
    f3 qp = mk(0.f, 0.f, 0.f);
    int p[6] = {0, 0, 0, 0, 0, 0};
    if (valid) {
        qp = mk(q[3 * k], q[3 * k + 1], q[3 * k + 2]);
        if (!query_corner(g, qp, p)) {
            counts[k] = 0;
            valid = false;
        } else {
            bool huge = false;
            for (int j = 0; j < 8; j++) {
                const uint32_t c = query_cell(g, p, j);
                huge = huge || ends[c] - (c == 0 ? 0 : ends[c - 1]) > kHugeRange;
            }
            if (huge) {  // k_grid_query_wide writes this query's list and count
                defer[1 + atomicAdd(defer, 1u)] = k;
                valid = false;
            }
        }
    }
    uint32_t m = 0;
    for (int j = 0; j < 8; j++) {  // the reference's cell order
        int a = 0, b = 0;
        if (valid) {
            const uint32_t c = query_cell(g, p, j);
            a = c == 0 ? 0 : ends[c - 1];
            b = ends[c];
        }
        const bool long_range = b - a > kLongRange;
        if (!long_range) {
            for (; a < b; a++) {
                const uint32_t pi = idx[a];
                if (grid_near(g, pos, qp, pi)) {
                    if (m < cap) out[(size_t)k * cap + m] = pi;
                    m++;
                }
            }
        }
        // long ranges: the wave scans them one query at a time, in order, so the
        // points keep the serial loop's order in the output
        for (uint64_t pend = __ballot(long_range); pend != 0; pend &= pend - 1) {
            const uint32_t L = __ffsll((unsigned long long)pend) - 1;
            const f3 lq = mk(__shfl(qp.x, L), __shfl(qp.y, L), __shfl(qp.z, L));
            const int la = __shfl(a, L), lb = __shfl(b, L);
            const uint32_t lk = __shfl(k, L);
            uint32_t lm = __shfl(m, L);
            // sixteen 64-point steps per iteration: their loads are in flight together
            // (a single wave walks the whole range, so its loads in flight set the pace)
            for (int base = la; base < lb; base += 16 * 64) {
                uint32_t pi[16];
                bool hit[16];
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const int ai = base + 64 * u + (int)lane;
                    pi[u] = ai < lb ? idx[ai] : 0u;
                }
#pragma unroll
                for (int u = 0; u < 16; u++)
                    hit[u] = base + 64 * u + (int)lane < lb && grid_near(g, pos, lq, pi[u]);
#pragma unroll
                for (int u = 0; u < 16; u++) {
                    const uint64_t hm = __ballot(hit[u]);
                    if (hit[u]) {
                        const uint32_t slot = lm + (uint32_t)__popcll(hm & below);
                        if (slot < cap) out[(size_t)lk * cap + slot] = pi[u];
                    }
                    lm += (uint32_t)__popcll(hm);
                }
            }
            if (lane == L) m = lm;
        }
    }
    if (valid) counts[k] = m;
}

// The deferred queries, one workgroup per query: the workgroup walks the
// query's 8 cells in the reference's order, 256 points a step and kWideSteps
// steps per iteration.  Step u, wave w, lane l holds point base + 256u + 64w + l,
// so a hit's slot is the hits before it in that order: the query's count so
// far, the wave counts of the earlier steps and waves (LDS), and its rank in
// its wave's ballot.  Every loop bound is the same for the whole workgroup.
constexpr int kWideSteps = 16;

__global__ __launch_bounds__(256) void k_grid_query_wide(GridDev g, const float* pos, const int* ends,
                                                         const uint32_t* idx, const float* q, uint32_t cap,
                                                         uint32_t* counts, uint32_t* out, const uint32_t* defer) {
    __shared__ uint32_t wcnt[kWideSteps][4];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t nd = defer[0];
    for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const uint32_t k = defer[1 + d];
        const f3 qp = mk(q[3 * k], q[3 * k + 1], q[3 * k + 2]);
        int p[6];
        (void)query_corner(g, qp, p);  // inside the box: only such queries are deferred
        uint32_t m = 0;
        for (int j = 0; j < 8; j++) {
            const uint32_t c = query_cell(g, p, j);
            const int a = c == 0 ? 0 : ends[c - 1];
            const int b = ends[c];
            for (int base = a; base < b; base += kWideSteps * 256) {
                uint32_t pi[kWideSteps];
                bool hit[kWideSteps];
#pragma unroll
                for (int u = 0; u < kWideSteps; u++) {
                    const int ai = base + 256 * u + (int)threadIdx.x;
                    pi[u] = ai < b ? idx[ai] : 0u;
                }
#pragma unroll
                for (int u = 0; u < kWideSteps; u++)
                    hit[u] = base + 256 * u + (int)threadIdx.x < b && grid_near(g, pos, qp, pi[u]);
                uint64_t hm[kWideSteps];
#pragma unroll
                for (int u = 0; u < kWideSteps; u++) {
                    hm[u] = __ballot(hit[u]);
                    if (lane == 0) wcnt[u][wave] = (uint32_t)__popcll(hm[u]);
                }
                __syncthreads();
                uint32_t off = m;
#pragma unroll
                for (int u = 0; u < kWideSteps; u++) {
                    uint32_t before = 0, all = 0;
                    for (uint32_t w = 0; w < 4; w++) {
                        before += w < wave ? wcnt[u][w] : 0u;
                        all += wcnt[u][w];
                    }
                    if (hit[u]) {
                        const uint32_t slot = off + before + (uint32_t)__popcll(hm[u] & below);
                        if (slot < cap) out[(size_t)k * cap + slot] = pi[u];
                    }
                    off += all;
                }
                m = off;
                __syncthreads();  // wcnt is rewritten by the next iteration
            }
        }
        if (threadIdx.x == 0) counts[k] = m;
    }
}

// Clears one word with a kernel rather than a memset, so a caller may record
// the query sequence into a graph.
__global__ void k_clear_word(uint32_t* p) { *p = 0u; }

}  // namespace wrd

// Deferred-query work list of wr_hashgrid_query_noalloc: the counter, then up
// to WR_HASHGRID_NOALLOC_MAX_QUERIES query indices (16 MiB per GPU, held until
// the process ends).  Host code never passes this name to a kernel: it uses
// the per-device address from hipGetSymbolAddress (noalloc_ptr below).
static __device__ __align__(256) uint32_t scratch[WR_HASHGRID_NOALLOC_MAX_QUERIES + 1];

namespace {
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kMaxScratchDevices] = {};  // device address of scratch, per device
}  // namespace

namespace wr {
// Looks up scratch's address on `device`, which must be current.  Called once
// per device, from wr_ctx_create.
hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = (uint32_t*)p;
    return hipSuccess;
}

static uint32_t* noalloc_ptr(int device) {
    if (device < 0 || device >= kMaxScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace wr


using namespace wrd;
using wr::fail;
using wr::hip_fail;

#define PH_CHECK(call, what)                         \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

namespace {
// Device scratch owned by one call, freed on every exit.
struct Scratch {
    std::vector<void*> p;
    ~Scratch() {
        for (void* q : p) (void)hipFree(q);
    }
    template <class T>
    hipError_t get(T** out, size_t bytes) {
        void* q = nullptr;
        hipError_t e = hipMalloc(&q, std::max<size_t>(bytes, 4));
        if (e == hipSuccess) p.push_back(q);
        *out = (T*)q;
        return e;
    }
};
}  // namespace

extern "C" {

int wr_light_paths(wr_ctx* ctx, const wr_scene* s, uint32_t max_depth, uint64_t seed, uint32_t pass, uint32_t path0,
                   uint32_t n_paths, uint32_t max_vertices, wr_path_vertex* d_vertices, uint32_t* d_counts) {
    if (!ctx || !s || (n_paths && (!d_vertices || !d_counts))) return fail(ctx, WR_EINVAL, "wr_light_paths: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_light_paths: scene not built");
    if ((uint64_t)max_vertices < (uint64_t)max_depth + 3)
        return fail(ctx, WR_EINVAL, "wr_light_paths: max_vertices must be at least max_depth + 3");
    if ((uint64_t)path0 + n_paths > 0x100000000ull) return fail(ctx, WR_EINVAL, "wr_light_paths: path ids overflow");
    if (n_paths == 0) return WR_OK;
    (void)hipSetDevice(ctx->device);
    k_light_paths<<<(n_paths + 255) / 256, 256, kStackLdsBytes, ctx->stream>>>(s->dev, max_depth, seed, pass, path0,
                                                                               n_paths, max_vertices, d_vertices,
                                                                               d_counts);
    PH_CHECK(hipGetLastError(), "wr_light_paths");
    PH_CHECK(hipStreamSynchronize(ctx->stream), "wr_light_paths");
    return WR_OK;
}

int wr_photons_collect(wr_ctx* ctx, const wr_scene* s, const wr_path_vertex* d_vertices, const uint32_t* d_counts,
                       uint32_t n_paths, uint32_t max_vertices, uint64_t n_leading, float* d_pos, uint32_t* d_src,
                       uint64_t cap, uint64_t* n_out) {
    if (!ctx || !s || !n_out || (cap && (!d_pos || !d_src)) || (n_paths && (!d_vertices || !d_counts)))
        return fail(ctx, WR_EINVAL, "wr_photons_collect: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_photons_collect: scene not built");
    if (n_paths >= (1u << 31)) return fail(ctx, WR_EINVAL, "wr_photons_collect: too many paths");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    Scratch sc;
    uint32_t* pc = nullptr;
    uint64_t* offs = nullptr;
    PH_CHECK(sc.get(&pc, sizeof(uint32_t) * (n_paths + 1)), "hipMalloc(photons)");
    PH_CHECK(sc.get(&offs, sizeof(uint64_t) * (n_paths + 1)), "hipMalloc(photons)");
    uint64_t total = 0;
    if (n_paths) {
        const uint32_t blocks = (n_paths + 255) / 256;
        k_photon_count<<<blocks, 256, 0, st>>>(s->dev, d_vertices, d_counts, n_paths, max_vertices, pc);
        PH_CHECK(hipGetLastError(), "wr_photons_collect");
        // exclusive offsets over n_paths + 1 entries (the last one is the total)
        PH_CHECK(hipMemsetAsync(pc + n_paths, 0, sizeof(uint32_t), st), "wr_photons_collect");
        size_t tb = 0;
        PH_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, pc, offs, (int)n_paths + 1, st), "hipcub scan");
        void* tmp = nullptr;
        PH_CHECK(sc.get(&tmp, tb), "hipMalloc(photons)");
        PH_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tb, pc, offs, (int)n_paths + 1, st), "hipcub scan");
        PH_CHECK(hipMemcpyAsync(&total, offs + n_paths, sizeof(uint64_t), hipMemcpyDeviceToHost, st),
                 "wr_photons_collect");
        k_photon_write<<<blocks, 256, 0, st>>>(s->dev, d_vertices, d_counts, n_paths, max_vertices, offs, n_leading,
                                                cap, d_pos, d_src);
        PH_CHECK(hipGetLastError(), "wr_photons_collect");
    }
    if (n_leading && cap) {
        const uint64_t m = std::min(n_leading, cap);
        k_fill_leading<<<(unsigned)std::min<uint64_t>((m + 255) / 256, 65535), 256, 0, st>>>(n_leading, cap, d_pos,
                                                                                             d_src);
        PH_CHECK(hipGetLastError(), "wr_photons_collect");
    }
    PH_CHECK(hipStreamSynchronize(st), "wr_photons_collect");
    *n_out = n_leading + total;
    return WR_OK;
}

int wr_hashgrid_build(wr_ctx* ctx, const float* d_pos, uint64_t n, float radius, uint32_t n_cells,
                      wr_hashgrid** out) {
    if (!ctx || !out || (n && !d_pos) || n_cells == 0) return fail(ctx, WR_EINVAL, "wr_hashgrid_build: bad argument");
    *out = nullptr;
    if (n >= (1ull << 31)) return fail(ctx, WR_EINVAL, "wr_hashgrid_build: too many points");
    if (!(radius > 0.f) || !std::isfinite(radius)) return fail(ctx, WR_EINVAL, "wr_hashgrid_build: radius must be > 0");
    (void)hipSetDevice(ctx->device);
    hipStream_t st = ctx->stream;
    wr_hashgrid* g = new wr_hashgrid;
    g->ctx = ctx;
    g->pos = d_pos;
    g->n = n;
    // Build with aMaxDist = 0: mCellSize = 2 * radius (hashgrid.hxx:28-35)
    float max_dist = 0.f;
    if (std::fabs(max_dist) < 1e-5) max_dist = radius;
    const float cell = max_dist * 2.f;
    g->g.inv = 1.f / cell;
    g->g.r2 = radius * radius;
    g->g.n_cells = n_cells;
    auto bail = [&](hipError_t e, const char* what) {
        wr_hashgrid_destroy(g);
        return hip_fail(ctx, e, what);
    };
    hipError_t e = hipMalloc((void**)&g->ends, sizeof(int) * n_cells);
    if (e == hipSuccess) e = hipMalloc((void**)&g->idx, sizeof(uint32_t) * std::max<uint64_t>(n, 1));
    if (e != hipSuccess) return bail(e, "hipMalloc(hashgrid)");
    Scratch sc;
    int* keys6 = nullptr;
    uint32_t *ckeys = nullptr, *ckeys2 = nullptr, *vals = nullptr, *cnt = nullptr;
    if ((e = sc.get(&keys6, 6 * sizeof(int))) != hipSuccess) return bail(e, "hipMalloc(hashgrid)");
    const int init[6] = {fkey_host(1e36f), fkey_host(1e36f), fkey_host(1e36f),
                         fkey_host(-1e36f), fkey_host(-1e36f), fkey_host(-1e36f)};
    if ((e = hipMemcpyAsync(keys6, init, sizeof(init), hipMemcpyHostToDevice, st)) != hipSuccess)
        return bail(e, "wr_hashgrid_build");
    if (n) {
        k_grid_bbox<<<(unsigned)std::min<uint64_t>((n + 255) / 256, 4096), 256, 0, st>>>(d_pos, n, keys6);
        if ((e = hipGetLastError()) != hipSuccess) return bail(e, "wr_hashgrid_build");
    }
    int hk[6];
    if ((e = hipMemcpyAsync(hk, keys6, sizeof(hk), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return bail(e, "wr_hashgrid_build");
    for (int a = 0; a < 3; a++) {
        g->g.bmin[a] = fkey_inv(hk[a]);
        g->g.bmax[a] = fkey_inv(hk[3 + a]);
        // the integer cell coordinates must stay representable
        if (n && !((g->g.bmax[a] - g->g.bmin[a]) * g->g.inv < 2147483520.f)) {
            wr_hashgrid_destroy(g);
            return fail(ctx, WR_EINVAL, "wr_hashgrid_build: too many cells across the points' box for this radius");
        }
    }
    if ((e = hipMemsetAsync(g->ends, 0, sizeof(int) * n_cells, st)) != hipSuccess) return bail(e, "wr_hashgrid_build");
    if (n) {
        if ((e = sc.get(&ckeys, 4 * n)) != hipSuccess || (e = sc.get(&ckeys2, 4 * n)) != hipSuccess ||
            (e = sc.get(&vals, 4 * n)) != hipSuccess || (e = sc.get(&cnt, 4ull * n_cells)) != hipSuccess)
            return bail(e, "hipMalloc(hashgrid)");
        if ((e = hipMemsetAsync(cnt, 0, 4ull * n_cells, st)) != hipSuccess) return bail(e, "wr_hashgrid_build");
        k_grid_keys<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(g->g, d_pos, (uint32_t)n, ckeys, vals, cnt);
        if ((e = hipGetLastError()) != hipSuccess) return bail(e, "wr_hashgrid_build");
        // mCellEnds = inclusive prefix sum of the counts
        size_t tb = 0, tb2 = 0;
        if ((e = hipcub::DeviceScan::InclusiveSum(nullptr, tb, cnt, g->ends, (int)n_cells, st)) != hipSuccess)
            return bail(e, "hipcub scan");
        int end_bit = 1;
        while (end_bit < 32 && (n_cells - 1) >> end_bit) end_bit++;
        if ((e = hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, ckeys, ckeys2, vals, g->idx, (int)n, 0, end_bit,
                                                    st)) != hipSuccess)
            return bail(e, "hipcub sort");
        void* tmp = nullptr;
        if ((e = sc.get(&tmp, std::max(tb, tb2))) != hipSuccess) return bail(e, "hipMalloc(hashgrid)");
        if ((e = hipcub::DeviceScan::InclusiveSum(tmp, tb, cnt, g->ends, (int)n_cells, st)) != hipSuccess)
            return bail(e, "hipcub scan");
        // stable: points of one cell keep their input order, as the serial scatter does
        if ((e = hipcub::DeviceRadixSort::SortPairs(tmp, tb2, ckeys, ckeys2, vals, g->idx, (int)n, 0, end_bit,
                                                    st)) != hipSuccess)
            return bail(e, "hipcub sort");
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return bail(e, "wr_hashgrid_build");
    *out = g;
    return WR_OK;
}

int wr_hashgrid_export(const wr_hashgrid* g, float* bbox, int32_t* cell_ends, int32_t* indices) {
    if (!g) return fail(nullptr, WR_EINVAL, "wr_hashgrid_export: grid is NULL");
    wr_ctx* ctx = g->ctx;
    (void)hipSetDevice(ctx->device);
    if (bbox)
        for (int a = 0; a < 3; a++) {
            bbox[a] = g->g.bmin[a];
            bbox[3 + a] = g->g.bmax[a];
        }
    if (cell_ends)
        PH_CHECK(hipMemcpy(cell_ends, g->ends, sizeof(int) * g->g.n_cells, hipMemcpyDeviceToHost), "wr_hashgrid_export");
    if (indices && g->n)
        PH_CHECK(hipMemcpy(indices, g->idx, sizeof(uint32_t) * g->n, hipMemcpyDeviceToHost), "wr_hashgrid_export");
    return WR_OK;
}

int wr_hashgrid_query(wr_hashgrid* g, const float* d_queries, uint64_t nq, uint32_t cap, uint32_t* d_counts,
                      uint32_t* d_idx) {
    if (!g || (nq && (!d_queries || !d_counts)) || (cap && nq && !d_idx))
        return fail(g ? g->ctx : nullptr, WR_EINVAL, "wr_hashgrid_query: bad argument");
    wr_ctx* ctx = g->ctx;
    if (nq >= (1ull << 32)) return fail(ctx, WR_EINVAL, "wr_hashgrid_query: too many queries");
    if (nq == 0) return WR_OK;
    (void)hipSetDevice(ctx->device);
    if (g->defer_cap < nq + 1) {
        if (g->defer) PH_CHECK(hipFree(g->defer), "wr_hashgrid_query");
        g->defer = nullptr;
        g->defer_cap = 0;
        PH_CHECK(hipMalloc((void**)&g->defer, sizeof(uint32_t) * (nq + 1)), "hipMalloc(hashgrid query)");
        g->defer_cap = nq + 1;
    }
    PH_CHECK(hipMemsetAsync(g->defer, 0, sizeof(uint32_t), ctx->stream), "wr_hashgrid_query");
    k_grid_query<<<(unsigned)((nq + 255) / 256), 256, 0, ctx->stream>>>(g->g, g->pos, g->ends, g->idx, d_queries,
                                                                         (uint32_t)nq, cap, d_counts, d_idx, g->defer);
    PH_CHECK(hipGetLastError(), "wr_hashgrid_query");
    // the deferred queries: a workgroup each, as many workgroups as fill the device twice
    const unsigned wide_grid = (unsigned)std::min<uint64_t>(nq, 2048);
    k_grid_query_wide<<<wide_grid, 256, 0, ctx->stream>>>(g->g, g->pos, g->ends, g->idx, d_queries, cap, d_counts,
                                                          d_idx, g->defer);
    PH_CHECK(hipGetLastError(), "wr_hashgrid_query");
    PH_CHECK(hipStreamSynchronize(ctx->stream), "wr_hashgrid_query");
    return WR_OK;
}

int wr_hashgrid_query_noalloc(wr_hashgrid* g, const float* d_queries, uint64_t nq, uint32_t cap, uint32_t* d_counts,
                              uint32_t* d_idx) {
    if (!g || (nq && (!d_queries || !d_counts)) || (cap && nq && !d_idx))
        return fail(g ? g->ctx : nullptr, WR_EINVAL, "wr_hashgrid_query_noalloc: bad argument");
    wr_ctx* ctx = g->ctx;
    if (nq > WR_HASHGRID_NOALLOC_MAX_QUERIES)
        return fail(ctx, WR_EINVAL, "wr_hashgrid_query_noalloc: more than WR_HASHGRID_NOALLOC_MAX_QUERIES (4,194,304) "
                                    "queries; use wr_hashgrid_query or split the call");
    if (nq == 0) return WR_OK;
    PH_CHECK(hipSetDevice(ctx->device), "wr_hashgrid_query_noalloc");
    uint32_t* defer = wr::noalloc_ptr(ctx->device);
    if (!defer) return fail(ctx, WR_ESTATE, "wr_hashgrid_query_noalloc: work list not set up on this device");
    // the list keeps the previous call's contents: only its counter must start at zero
    k_clear_word<<<1, 1, 0, ctx->stream>>>(defer);
    PH_CHECK(hipGetLastError(), "wr_hashgrid_query_noalloc");
    k_grid_query<<<(unsigned)((nq + 255) / 256), 256, 0, ctx->stream>>>(g->g, g->pos, g->ends, g->idx, d_queries,
                                                                         (uint32_t)nq, cap, d_counts, d_idx, defer);
    PH_CHECK(hipGetLastError(), "wr_hashgrid_query_noalloc");
    const unsigned wide_grid = (unsigned)std::min<uint64_t>(nq, 2048);
    k_grid_query_wide<<<wide_grid, 256, 0, ctx->stream>>>(g->g, g->pos, g->ends, g->idx, d_queries, cap, d_counts,
                                                          d_idx, defer);
    PH_CHECK(hipGetLastError(), "wr_hashgrid_query_noalloc");
    return WR_OK;
}

void wr_hashgrid_destroy(wr_hashgrid* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    if (g->ends) (void)hipFree(g->ends);
    if (g->idx) (void)hipFree(g->idx);
    if (g->defer) (void)hipFree(g->defer);
    delete g;
}

}  // extern "C"
