// pt.hip -- path-tracing kernels (gfx950) and their host-side drivers.
//
//  k_pt_mega        the plain one-thread-per-pixel kernel: one launch per
//                   pass, each lane runs its pixel's whole path
//                   (pathTracer.cpp:20-70 with mcRenderer.cpp:16-57) and adds
//                   the sample to the Kahan film it owns.
//  wavefront        the same estimator split into stages (extend = trace +
//                   finish, shade, plan, generate, fold, advance) over a
//                   shared pool of path slots
//                   that any number of jobs (render calls, served requests)
//                   fill; finished paths write their sample to a per-(job,
//                   pass, pixel) slot and complete passes are folded into the
//                   job's Kahan film in pass order, so every film is
//                   bit-identical to the megakernel's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pt_dev.h"
#include "wr_pt.h"

#pragma clang fp contract(off)

namespace wrd {

using wr::kWorkDone;
using wr::kWorkStride;
using wr::kXcds;

template <bool NEE>
__global__ __launch_bounds__(256) void k_pt_mega(SceneDev S, CamDev cam, uint32_t max_depth, uint64_t seed,
                                                 uint32_t pass, FilmDev film, uint32_t npix) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    const f3 v = sample_path<NEE>(sc, cam, max_depth, seed, p, pass, st);
    film_add(film, p, v);
}

// ---- wavefront engine ---------------------------------------------------------
// Live paths of every admitted job share one pool of dense path buffers
// (ping-pong, SoA, coalesced): extend traces every live ray (camera rays
// without NoSelf plus findInsideObject, later rays with NoSelf), shade scatters
// and appends survivors to the other buffer with one atomic per wave, plan
// hands the free slots to the running jobs round-robin (each job keeps at most
// R passes in flight), generate appends their new camera paths, fold adds
// every complete pass of every job to that job's Kahan film in pass order, and
// advance retires finished jobs.  Per-path randomness is keyed by the job's
// seed and the (pixel, pass, vertex) of the path, so a job's film does not
// depend on which other jobs share the pool.
//
// Jobs with next-event estimation (pathTracer.cpp:39-66) keep three more
// fields per path: U and P, the branch-order products with and without the
// path's current ray, and crec, the triangle under the ray's origin (kFront
// holds the incoming side there).  A path whose ray escapes (k_wf_finish) or
// that reaches the depth limit (k_wf_shade) connects that origin to a light
// sample: it becomes a connection entry (kShadow) of the next iteration's
// buffer, traced by the same extend stage with the same NoSelf condition, and
// k_wf_finish ends it with the visibility test.
constexpr uint32_t kPrimary = 1u;
constexpr uint32_t kDead = 2u;
constexpr uint32_t kShadow = 8u;   // connection entry: T = value if visible, ux = distance to the light point
constexpr uint32_t kFront = 16u;   // (incoming direction . N) <= 0 at the ray's origin

struct PathBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tx, *ty, *tz, *ht;
    float *ux, *uy, *uz, *px, *py, *pz;
    uint32_t *pixel, *pass, *depth, *inside, *contact, *hrec, *hobj, *flags, *job, *crec;
};

struct EngineDev {
    PathBuf buf[2];
    wr::EngineCounters* cnt;
    const wr::JobDesc* jobs;
    wr::JobCtl* ctl;
    const SceneDev* scenes;
    uint32_t* exit_list;
    uint32_t cap;
};

__device__ __forceinline__ CamDev job_cam(const wr::JobDesc& J) {
    CamDev c;
    c.pos = mk(J.pos[0], J.pos[1], J.pos[2]);
    c.front = mk(J.front[0], J.front[1], J.front[2]);
    c.right = mk(J.right[0], J.right[1], J.right[2]);
    c.top = mk(J.top[0], J.top[1], J.top[2]);
    c.dist = J.dist;
    c.W = J.W;
    c.H = J.H;
    return c;
}

// One slot per lane that wants one, for the whole 256-lane workgroup with a
// single atomic; every thread of the workgroup must make the call.  Slots go
// out in thread order.
__device__ __forceinline__ uint32_t block_append(uint32_t* counter, bool want) {
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_base;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint64_t mask = __ballot(want);
    if (lane == 0) s_cnt[w] = (uint32_t)__popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        s_base = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    uint32_t off = s_base;
    for (uint32_t v = 0; v < w; v++) off += s_cnt[v];
    off += (uint32_t)__popcll(mask & ((1ull << lane) - 1));
    __syncthreads();  // s_cnt and s_base serve the next call
    return off;
}

// *addr += 1 for every active lane, one atomic per distinct address in the wave
// (lanes of a wave almost always finish paths of the same job and pass).
__device__ __forceinline__ void count_done(uint32_t* addr) {
    const uint64_t a = (uint64_t)addr;
    for (;;) {
        const uint64_t act = __ballot(1);
        const uint32_t leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t lo = __shfl((uint32_t)a, leader), hi = __shfl((uint32_t)(a >> 32), leader);
        const bool mine = (uint32_t)a == lo && (uint32_t)(a >> 32) == hi;
        const uint64_t same = __ballot(mine);
        if (mine) {
            if ((threadIdx.x & 63) == leader) atomicAdd(addr, (uint32_t)__popcll(same));
            return;
        }
    }
}

// A connection entry (MCRenderer::visibilityTest of the NEE branch) at slot j
// of the next iteration's buffer.
__device__ __forceinline__ void write_connection(const PathBuf& B, uint32_t j, f3 o, const NeeConn& c,
                                                 uint32_t inside, uint32_t contact, uint32_t pixel, uint32_t pass,
                                                 uint32_t job) {
    B.ox[j] = o.x; B.oy[j] = o.y; B.oz[j] = o.z;
    B.dx[j] = c.dir.x; B.dy[j] = c.dir.y; B.dz[j] = c.dir.z;
    B.tx[j] = c.C.x; B.ty[j] = c.C.y; B.tz[j] = c.C.z;
    B.ux[j] = c.dist;
    B.pixel[j] = pixel;
    B.pass[j] = pass;
    B.depth[j] = 0;
    B.inside[j] = inside;
    B.contact[j] = contact;
    B.flags[j] = kShadow;
    B.job[j] = job;
}

// The value of a finished path into its pass slot; returns the pass's
// completion counter, which the caller increments.
__device__ __forceinline__ uint32_t* finish_value(const wr::JobDesc& J, const CamDev& cam, uint32_t p, uint32_t ps,
                                                  f3 color) {
    if (!legal(color)) color = mk(0.f, 0.f, 0.f);
    const f3 v = fix_vignetting(cam, color, p);
    const uint32_t key = ps % J.R;
    float* c = J.contrib + 3 * ((size_t)key * J.npix + p);
    c[0] = v.x;
    c[1] = v.y;
    c[2] = v.z;
    return J.pass_done + key;
}

// *addr += 1 for every lane of the workgroup with want set; every thread must
// make the call.  The lanes whose counter is the first wanting lane's (in a
// single-job render, almost all of them) add with one atomic for the
// workgroup, the others per wave as count_done does.
__device__ __forceinline__ void block_count_done(uint32_t* addr, bool want) {
    __shared__ unsigned long long s_addr;
    __shared__ uint32_t s_n;
    if (threadIdx.x == 0) {
        s_addr = 0ull;
        s_n = 0u;
    }
    __syncthreads();
    if (want) atomicCAS(&s_addr, 0ull, (unsigned long long)addr);
    __syncthreads();
    const bool same = want && (unsigned long long)addr == s_addr;
    const uint64_t m = __ballot(same);
    if ((threadIdx.x & 63u) == 0 && m) atomicAdd(&s_n, (uint32_t)__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0 && s_n) atomicAdd((uint32_t*)s_addr, s_n);
    if (want && !same) count_done(addr);
    __syncthreads();  // s_addr and s_n serve the next call
}

// extend, part 1: the fast pass of every live ray (Scene::intersect; camera
// rays without a condition, later rays with NoSelf) as a persistent kernel
// that refills a lane with the next path as soon as its traversal ends, and
// runs node-only steps while most lanes sit at wide nodes (as
// k_closest_hit).  The fast-pass result goes to ht / hrec / hobj (first record
// of the leaf) and the overflow flag; k_wf_finish completes it.
constexpr uint32_t kOverflow = 4u;
constexpr uint32_t kTraceRefillMin = 16;
// Work ranges and batches: the live paths are cut into kXcds contiguous
// ranges, range q handed out by its own counter to the workgroups of XCD q
// (workgroups are dispatched to the XCDs round-robin), so neighbouring paths
// share one XCD's L2 and each counter takes an eighth of the atomics; a wave
// whose range is used up moves on round the others.  Paths a wave takes at a
// time: guided self-scheduling (Polychronopoulos and Kuck, 1987), the paths
// left in the range over sixteen times the XCD's waves, clamped to
// [kTraceBatchMin, kTraceBatchMax] and rounded up to whole waves.  Large
// batches early keep the waves off the counters while most paths remain;
// small ones at the end keep the tail short.  A fixed 64 left a single
// counter contended on large iterations, a fixed 256 stretched the tail of
// small ones; with a counter per XCD a share of 16 beat 8, 4 and 2 (on the
// 512-object field's iterations of about 8M paths it takes 64-128 paths).
constexpr uint32_t kTraceBatchMin = 64;
constexpr uint32_t kTraceBatchMax = 1024;
constexpr uint32_t kTraceBatchShare = 16;
constexpr uint32_t kTraceNodePhaseMin = 32;

// D: traversal stack levels in LDS (4, 8 or 16; the wavefront autotuner's stack-depth knob).
// COUNT: add the fast pass's node and triangle records fetched to the engine's
// totals (the per-stage statistics launches only).
template <int D, bool COUNT = false>
__global__ __launch_bounds__(256) void k_wf_trace(EngineDev E, uint32_t par) {
    const PathBuf& A = E.buf[par];
    const uint32_t n = E.cnt->n[par];
    unsigned long long* work = E.cnt->trace_work;
    // workgroups beyond what n paths need leave at once (and are not counted below)
    const uint32_t used = (uint32_t)(((uint64_t)n + 255) / 256 < gridDim.x ? ((uint64_t)n + 255) / 256 : gridDim.x);
    if (blockIdx.x >= used) return;
    StackSpillT<D> spill;
    LaneStackT<D> st(&spill);
    Counters cnt{0, 0, 0};
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t below = (1ull << lane) - 1ull;
    uint32_t bnext = 0, bend = 0;  // the wave's current batch of paths [bnext, bend)
    uint32_t q = blockIdx.x % kXcds;  // the work range of this workgroup's XCD
    uint32_t used_up = 0;             // ranges this wave has found used up (each once: they stay so)
    const uint32_t xcd_waves = used * 4u / kXcds > 0u ? used * 4u / kXcds : 1u;
    bool exhausted = false;
    bool active = false, overflow = false;
    uint32_t pid = 0, ref = 0;
    RayPre r{};
    HitState h{};
    Tree tr{};
    NoSelf ns{};
    for (;;) {
        uint64_t idle = __ballot(!active);
        if (!exhausted && (uint32_t)__popcll(idle) >= kTraceRefillMin) {
            while (idle != 0) {
                if (bnext == bend) {
                    // next batch: from range q, or the next range once q is used up
                    bool got = false;
                    while (used_up < kXcds) {
                        const uint32_t lo = (uint32_t)((uint64_t)n * q / kXcds);
                        const uint32_t hi = (uint32_t)((uint64_t)n * (q + 1) / kXcds);
                        const uint32_t from = bend >= lo && bend < hi ? bend : lo;  // progress estimate
                        uint32_t take = (hi - from) / (xcd_waves * kTraceBatchShare);
                        take = take < kTraceBatchMin ? kTraceBatchMin : (take > kTraceBatchMax ? kTraceBatchMax : take);
                        take = (take + 63u) & ~63u;
                        unsigned long long base = 0;
                        if (lane == 0) base = atomicAdd(work + kWorkStride * q, (unsigned long long)take);
                        base = __shfl(base, 0);
                        if (lo + base < hi) {
                            bnext = lo + (uint32_t)base;
                            bend = lo + base + take < hi ? lo + (uint32_t)base + take : hi;
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
                const uint32_t avail = bend - bnext;
                const uint32_t rank = (uint32_t)__popcll(idle & below);
                if (!active && rank < avail) {
                    pid = bnext + rank;
                    const SceneDev& S = E.scenes[E.jobs[A.job[pid]].scene];
                    const wr_tree_hdr H = S.hdrs[0];
                    tr = make_tree(H, S);
                    r = make_ray(mk(A.ox[pid], A.oy[pid], A.oz[pid]), mk(A.dx[pid], A.dy[pid], A.dz[pid]), tr.scale);
                    const bool primary = A.flags[pid] & kPrimary;
                    ns = NoSelf{primary ? WR_NONE : A.inside[pid], primary ? WR_NONE : A.contact[pid],
                                S.rec_nrm + 9 * (size_t)H.rec_base, S.obj_has_normals};
                    overflow = false;
                    if (wide_begin(tr, r, st, h, ref)) {
                        active = true;
                    } else {
                        A.hrec[pid] = WR_NONE;  // misses the root box
                    }
                }
                const uint32_t took = (uint32_t)__popcll(idle);
                bnext += took < avail ? took : avail;
                idle = __ballot(!active);
            }
        }
        if (__ballot(active) == 0) break;  // only once the counter is exhausted
        for (;;) {
            const bool at_node = active && !(ref & WR_LEAF_BIT);
            if ((uint32_t)__popcll(__ballot(at_node)) < kTraceNodePhaseMin) break;
            if (at_node && !wide_step<true, COUNT>(tr, r, ns, st, h, ref, overflow, cnt)) {
                active = false;
                A.ht[pid] = h.t;
                A.hrec[pid] = h.rec;
                A.hobj[pid] = h.leaf;
                if (overflow) A.flags[pid] |= kOverflow;
            }
        }
        if (active && !wide_step<true, COUNT>(tr, r, ns, st, h, ref, overflow, cnt)) {
            active = false;
            A.ht[pid] = h.t;
            A.hrec[pid] = h.rec;
            A.hobj[pid] = h.leaf;
            if (overflow) A.flags[pid] |= kOverflow;
        }
    }
    if (COUNT) {
        unsigned long long a = cnt.nodes, b = cnt.tris;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off);
            b += __shfl_xor(b, off);
        }
        if (lane == 0) {
            atomicAdd(&E.cnt->trace_nodes, a);
            atomicAdd(&E.cnt->trace_tris, b);
        }
    }
    if (lane == 0) {
        const unsigned long long waves = (unsigned long long)used * (blockDim.x / 64);
        if (atomicAdd(work + kWorkDone, 1ull) == waves - 1) {
            for (uint32_t k = 0; k < kXcds; k++) atomicExch(work + kWorkStride * k, 0ull);
            atomicExch(work + kWorkDone, 0ull);
        }
    }
}

// extend, part 2 (one lane per path): binary fast pass after an overflow,
// chain verification and exact fallback (trace_dev.h: finish_hit),
// findInsideObject for camera rays (camera.cpp:52-56), the end of paths that
// escape (or their next-event connection), and the visibility test of
// connection entries.
// 4 waves per SIMD (128 VGPRs, the rest spilled) instead of 2 at 191 VGPRs: 2% on the field.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_wf_finish(EngineDev E, uint32_t par) {
    const PathBuf& A = E.buf[par];
    const PathBuf& B = E.buf[par ^ 1];
    const uint32_t n = E.cnt->n[par];
    StackSpill spill;
    LaneStack st(&spill);
    // the workgroup steps through the slots together: the paths that end here
    // are counted, and the connection entries take their slots, with one
    // atomic per workgroup
    for (uint32_t i0 = blockIdx.x * 256; i0 < n; i0 += gridDim.x * 256) {
        const uint32_t i = i0 + threadIdx.x;
        uint32_t* done = nullptr;  // the completion counter of a path that ends here
        bool conn = false;         // a connection entry from an escaped path (NEE)
        NeeConn nc;
        f3 co;
        if (i < n) {
            if (E.ctl[A.job[i]].state != wr::kJobRunning) {  // a cancelled job: drop its path, write nothing
                A.flags[i] = kDead;
            } else {
                const wr::JobDesc& J = E.jobs[A.job[i]];
                const SceneDev& S = E.scenes[J.scene];
                const SceneCtx sc = make_scene_ctx(S);
                const f3 o = mk(A.ox[i], A.oy[i], A.oz[i]);
                const f3 d = mk(A.dx[i], A.dy[i], A.dz[i]);
                const uint32_t fl = A.flags[i];
                const bool primary = fl & kPrimary;
                HitState h{A.ht[i], A.hrec[i], A.hobj[i]};
                if (h.rec != WR_NONE || (fl & kOverflow)) {
                    const RayPre r = make_ray(o, d, sc.tr.scale);
                    const NoSelf ns{primary ? WR_NONE : A.inside[i], primary ? WR_NONE : A.contact[i], sc.gnrm,
                                    S.obj_has_normals};
                    Counters cnt{0, 0, 0};
                    bool failed;
                    finish_hit<true, false>(sc.tr, r, ns, st, cnt, h, (fl & kOverflow) != 0, failed);
                }
                if (fl & kShadow) {  // MCRenderer::visibilityTest (mcRenderer.cpp:59-65)
                    const bool vis = nee_visible(h.rec, h.t, A.ux[i]);
                    done = finish_value(J, job_cam(J), A.pixel[i], A.pass[i],
                                        vis ? mk(A.tx[i], A.ty[i], A.tz[i]) : mk(0.f, 0.f, 0.f));
                    A.flags[i] = kDead;
                } else {
                    if (primary) A.inside[i] = find_inside(S, o, d, WR_NONE, st);
                    if (h.rec == WR_NONE) {  // escaped (no environment light) or nothing to scatter: sample 0
                        if ((J.flags & wr::kJobNee) && !primary) {  // endIndex = this ray: connect from its origin
                            float rn[4];
                            rng4(J.seed, A.pixel[i], J.pass0 + A.pass[i], kNeeVertex, rn);
                            nc = nee_connect(sc, o, A.contact[i], A.crec[i], (fl & kFront) != 0,
                                             mk(A.px[i], A.py[i], A.pz[i]), rn);
                            conn = nc.ok;
                            co = o;
                        }
                        if (!conn) done = finish_value(J, job_cam(J), A.pixel[i], A.pass[i], mk(0.f, 0.f, 0.f));
                        A.flags[i] = kDead;
                    } else {
                        A.ht[i] = h.t;
                        A.hrec[i] = h.rec;
                        A.hobj[i] = sc.tr.tris[h.rec].obj;
                        A.flags[i] = fl & kFront;
                    }
                }
            }
        }
        block_count_done(done, done != nullptr);
        const uint32_t j = block_append(&E.cnt->n[par ^ 1], conn);
        if (conn) write_connection(B, j, co, nc, A.inside[i], A.contact[i], A.pixel[i], A.pass[i], A.job[i]);
    }
}

// shade: scatter + Russian roulette of every path whose ray hit (MCRenderer::
// samplePath, mcRenderer.cpp:30-41), survivors appended to the next buffer.
// EXITS = false shades every live path except refraction exits, whose scatter
// needs findInsideObject (two closest-hit queries per candidate object): those
// are listed and shaded by the EXITS = true instance, so the traversal code
// and its registers stay out of the common case and the queries run on
// converged waves.  The list order only changes which slots the survivors take.
// NEE = false: the instance for pools in which no job asks for next-event
// estimation; without that code the common-case kernel keeps its registers
// (4 spilled instead of 62 at 128 VGPRs).
template <bool EXITS, bool NEE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(EXITS ? 1 : 4))) void k_wf_shade(EngineDev E,
                                                                                                     uint32_t par) {
    const PathBuf& A = E.buf[par];
    const PathBuf& B = E.buf[par ^ 1];
    const uint32_t n = EXITS ? E.cnt->n_exit : E.cnt->n[par];
    StackSpill spill;
    LaneStack st(&spill);
    // the workgroup steps through the slots together: each step's survivors,
    // connection entries and refraction exits take their slots with one
    // atomic per workgroup (block_append) instead of one per wave
    for (uint32_t k0 = blockIdx.x * 256; k0 < n; k0 += gridDim.x * 256) {
        const uint32_t k = k0 + threadIdx.x;
        enum { kNone, kSurvive, kConnect, kExit } fate = kNone;
        uint32_t i = 0, jb = 0, p = 0, ps = 0, depth = 0, inside = 0, contact = 0, crec = 0, flags = 0;
        uint32_t* done = nullptr;  // the completion counter of a path that ends here
        f3 o0, d0, t0, u0, u1;  // the next entry: origin, direction, throughput, NEE products
        NeeConn nc;             // a connection entry's direction, value and distance
        if (k < n) {
            i = EXITS ? E.exit_list[k] : k;
            if (EXITS || !(A.flags[i] & kDead)) {
                jb = A.job[i];
                const wr::JobDesc& J = E.jobs[jb];
                const SceneCtx sc = make_scene_ctx(E.scenes[J.scene]);
                const f3 o = mk(A.ox[i], A.oy[i], A.oz[i]);
                const f3 d = mk(A.dx[i], A.dy[i], A.dz[i]);
                PHit h;
                h.t = A.ht[i];
                h.rec = A.hrec[i];
                h.obj = A.hobj[i];
                inside = A.inside[i];
                depth = A.depth[i];
                p = A.pixel[i];
                ps = A.pass[i];
                const uint32_t self = inside != WR_NONE ? inside : h.obj;
                if (!EXITS && scatter_queries_inside(sc, self, h, inside)) {
                    fate = kExit;
                } else {
                    float rn[4];
                    rng4(J.seed, p, J.pass0 + ps, depth + 1, rn);
                    const Out out = scatter<EXITS>(sc, self, o, d, h, inside, rn, st);
                    f3 T = mk(A.tx[i], A.ty[i], A.tz[i]);
                    const bool nee = NEE && (J.flags & wr::kJobNee) != 0;
                    if ((double)len(out.d) < 0.5) {  // emission / absorption ends the path
                        done = finish_value(J, job_cam(J), p, ps, mulv(T, divs(divs(out.rad, out.op), out.dp)));
                    } else if (depth + 1 > J.max_depth) {
                        NeeConn c;
                        c.ok = false;
                        if (nee && depth >= 1) {  // endIndex = this ray: connect from its origin
                            rng4(J.seed, p, J.pass0 + ps, kNeeVertex, rn);
                            c = nee_connect(sc, o, A.contact[i], A.crec[i], (A.flags[i] & kFront) != 0,
                                            mk(A.px[i], A.py[i], A.pz[i]), rn);
                        }
                        if (c.ok) {
                            fate = kConnect;
                            o0 = o;
                            nc = c;
                            contact = A.contact[i];
                        } else {
                            done = finish_value(J, job_cam(J), p, ps, mk(0.f, 0.f, 0.f));
                        }
                    } else {
                        // the weight of this vertex only matters if the next ray hits; applying
                        // it now keeps the per-pixel loop's multiplication order
                        T = muls(T, cosine_term(out));
                        T = mulv(T, divs(divs(out.rad, out.op), out.dp));
                        fate = kSurvive;
                        o0 = out.o;
                        d0 = out.d;
                        t0 = T;
                        inside = out.inside;
                        contact = out.contact;
                        flags = nee && dot(d, out.n) <= 0.f ? kFront : 0u;
                        if (nee) {  // the next ray's products (branch order) and its origin's triangle
                            u0 = mk(A.ux[i], A.uy[i], A.uz[i]);
                            u1 = mulv(u0, divs(divs(out.rad, out.op), out.dp));
                            u1 = muls(u1, cosine_term(out));
                            crec = h.rec;
                            flags |= 0x80000000u;  // marks the NEE fields for the write below
                        }
                    }
                }
            }
        }
        block_count_done(done, done != nullptr);
        const uint32_t j = block_append(&E.cnt->n[par ^ 1], fate == kSurvive || fate == kConnect);
        if (!EXITS) {
            const uint32_t x = block_append(&E.cnt->n_exit, fate == kExit);
            if (fate == kExit) E.exit_list[x] = i;
        }
        if (fate == kConnect) {
            write_connection(B, j, o0, nc, inside, contact, p, ps, jb);
        } else if (fate == kSurvive) {
            if (flags & 0x80000000u) {
                B.px[j] = u0.x; B.py[j] = u0.y; B.pz[j] = u0.z;
                B.ux[j] = u1.x; B.uy[j] = u1.y; B.uz[j] = u1.z;
                B.crec[j] = crec;
            }
            B.ox[j] = o0.x; B.oy[j] = o0.y; B.oz[j] = o0.z;
            B.dx[j] = d0.x; B.dy[j] = d0.y; B.dz[j] = d0.z;
            B.tx[j] = t0.x; B.ty[j] = t0.y; B.tz[j] = t0.z;
            B.pixel[j] = p;
            B.pass[j] = ps;
            B.depth[j] = depth + 1;
            B.inside[j] = inside;
            B.contact[j] = contact;
            B.flags[j] = flags & ~0x80000000u;
            B.job[j] = jb;
        }
    }
}

// Number of consecutive complete passes of a job starting at its fold cursor.
__device__ __forceinline__ uint32_t foldable(const wr::JobDesc& J, const wr::JobCtl& C) {
    uint32_t k = 0;
    while (k < J.R && C.cursor + k < J.npass && J.pass_done[(C.cursor + k) % J.R] == J.npix) k++;
    return k;
}

// One wave (lane k = job k, then lane k = the k-th job of the round-robin
// order).  First the fold plan: per running job, the complete passes from its
// fold cursor (shade has finished this iteration's paths, and nothing before
// k_wf_advance changes the counts).  Then the free slots behind the survivors
// go to the running jobs round-robin from a rotating start, each within its
// window of free pass slots: the greedy serial allocation in that order, job
// i taking min(avail_i, room - sum of the takes before it), is a prefix sum
// over the wave, clamp(room - sum_{<i} avail, 0, avail_i).
__global__ __launch_bounds__(64) void k_wf_plan(EngineDev E, uint32_t par) {
    static_assert(wr::kMaxJobs == 64, "one lane per job");
    wr::EngineCounters& c = *E.cnt;
    const uint32_t lane = threadIdx.x;
    const uint64_t below = (1ull << lane) - 1ull;
    {
        const wr::JobCtl& C = E.ctl[lane];
        const uint32_t k = C.state == wr::kJobRunning ? foldable(E.jobs[lane], C) : 0u;
        const uint64_t m = __ballot(k > 0);
        c.fold_k[lane] = k;
        if (k > 0) c.fold_job[__popcll(m & below)] = lane;
        if (lane == 0) c.nfold = (uint32_t)__popcll(m);
    }
    const uint32_t base0 = c.n[par ^ 1];
    const uint32_t room = E.cap - base0;
    const uint32_t rr = c.rr;
    const uint32_t j = (rr + lane) % wr::kMaxJobs;
    wr::JobCtl& C = E.ctl[j];
    uint64_t avail = 0, next = 0;
    if (C.state == wr::kJobRunning) {
        const wr::JobDesc& J = E.jobs[j];
        const uint64_t total = (uint64_t)J.npass * J.npix;
        const uint64_t lim_w = (uint64_t)(C.cursor + J.R) * J.npix;
        const uint64_t limit = lim_w < total ? lim_w : total;
        next = C.next_work;
        avail = limit > next ? limit - next : 0;
    }
    if (avail > room) avail = room;  // a take never exceeds room; keeps the sums below 2^39
    // exclusive prefix sum of avail over the lanes
    uint64_t incl = avail;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t)incl, off), hi = __shfl_up((uint32_t)(incl >> 32), off);
        if (lane >= (uint32_t)off) incl += ((uint64_t)hi << 32) | lo;
    }
    const uint64_t before = incl - avail;
    const uint32_t used = (uint32_t)(before < room ? before : room);
    const uint32_t take = (uint32_t)(avail < (uint64_t)(room - used) ? avail : (uint64_t)(room - used));
    const uint64_t m = __ballot(take > 0);
    if (take > 0) {
        const uint32_t s = (uint32_t)__popcll(m & below);
        c.seg_job[s] = j;
        c.seg_base[s] = base0 + used;
        c.seg_take[s] = take;
        c.seg_work[s] = next;
        C.next_work = next + take;
    }
    const uint64_t total_take = __shfl(incl, 63);
    if (lane == 0) {
        const uint32_t all = (uint32_t)(total_take < room ? total_take : room);
        c.traced += c.n[par];
        c.exits += c.n_exit;
        c.generated += all;
        c.nseg = (uint32_t)__popcll(m);
        c.seg_start = base0;
        c.seg_end = base0 + all;
        c.n[par ^ 1] = base0 + all;
        c.rr = (rr + 1) % wr::kMaxJobs;
        c.n_exit = 0;  // the exit list of this iteration is consumed
    }
}

// The pixel of the r-th camera path of a pass: pixels are issued in kCamTile x
// kCamTile tiles (tiles row-major within bands kCamTile rows high, pixels
// row-major within a tile; the last band and the last tile of a band are cut
// to the image), so the paths of one wave and one work range look at one patch
// of the scene.  A bijection on [0, W*H): each (pixel, pass) is generated once,
// and since a path's samples depend only on its pixel and pass, the order
// changes nothing in the film.
constexpr uint32_t kCamTile = 8;
__device__ __forceinline__ uint32_t tile_pixel(uint32_t r, uint32_t W, uint32_t H) {
    const uint32_t band = r / (W * kCamTile);
    const uint32_t hb = H - band * kCamTile < kCamTile ? H - band * kCamTile : kCamTile;
    const uint32_t o = r - band * W * kCamTile;  // index within the band
    const uint32_t nfull = W / kCamTile;         // full-width tiles of the band
    uint32_t x, y;
    if (o < nfull * kCamTile * hb) {
        const uint32_t tc = o / (kCamTile * hb), oo = o - tc * kCamTile * hb;
        x = tc * kCamTile + oo % kCamTile;
        y = band * kCamTile + oo / kCamTile;
    } else {  // the band's last tile, narrower than kCamTile
        const uint32_t oo = o - nfull * kCamTile * hb, wt = W - nfull * kCamTile;
        x = nfull * kCamTile + oo % wt;
        y = band * kCamTile + oo / wt;
    }
    return y * W + x;
}

__global__ __launch_bounds__(256) void k_wf_generate(EngineDev E, uint32_t par) {
    const PathBuf& B = E.buf[par ^ 1];
    const wr::EngineCounters& c = *E.cnt;
    const uint32_t start = c.seg_start, end = c.seg_end, nseg = c.nseg;
    for (uint32_t j = start + blockIdx.x * 256 + threadIdx.x; j < end; j += gridDim.x * 256) {
        uint32_t s = 0;
        while (s + 1 < nseg && c.seg_base[s + 1] <= j) s++;
        const uint32_t jb = c.seg_job[s];
        const wr::JobDesc& J = E.jobs[jb];
        const CamDev cam = job_cam(J);
        const uint64_t w = c.seg_work[s] + (j - c.seg_base[s]);
        const uint32_t ps = (uint32_t)(w / J.npix), p = tile_pixel((uint32_t)(w % J.npix), J.W, J.H);
        float rn[4];
        rng4(J.seed, p, J.pass0 + ps, 0, rn);
        const f3 d = cam_dir(cam, p, rn[0], rn[1]);
        float c0 = fabsf(dot(cam.front, d));  // camera cosineTerm (ray.cpp:52-58, camera.cpp:182-186)
        c0 = (double)c0 > kEps ? c0 : (float)kEps;
        B.ox[j] = cam.pos.x; B.oy[j] = cam.pos.y; B.oz[j] = cam.pos.z;
        B.dx[j] = d.x; B.dy[j] = d.y; B.dz[j] = d.z;
        B.tx[j] = c0; B.ty[j] = c0; B.tz[j] = c0;
        if (J.flags & wr::kJobNee) {  // branch-order product of the camera ray: (1,1,1) * 1/1/1 * cos
            B.ux[j] = c0; B.uy[j] = c0; B.uz[j] = c0;
        }
        B.pixel[j] = p;
        B.pass[j] = ps;
        B.depth[j] = 0;
        B.inside[j] = WR_NONE;
        B.contact[j] = WR_NONE;
        B.flags[j] = kPrimary;
        B.job[j] = jb;
    }
}

// Adds every complete pass of every running job to its film, in pass order
// (the jobs and pass counts of k_wf_plan's fold plan).
__global__ __launch_bounds__(256) void k_wf_fold(EngineDev E) {
    const wr::EngineCounters& c = *E.cnt;
    const uint32_t nf = c.nfold;
    for (uint32_t f = 0; f < nf; f++) {
        const uint32_t jb = c.fold_job[f];
        const uint32_t k = c.fold_k[jb];
        const wr::JobCtl& C = E.ctl[jb];
        const wr::JobDesc& J = E.jobs[jb];
        const FilmDev film{J.fsum, J.fcomp, J.fsumsq};
        for (uint32_t p = blockIdx.x * 256 + threadIdx.x; p < J.npix; p += gridDim.x * 256) {
            for (uint32_t q = 0; q < k; q++) {
                const float* cc = J.contrib + 3 * ((size_t)((C.cursor + q) % J.R) * J.npix + p);
                film_add(film, p, mk(cc[0], cc[1], cc[2]));
            }
        }
    }
}

// One wave, lane k = job k: advance fold cursors by the fold plan, re-arm the
// folded pass slots, retire finished jobs; then empty the consumed path buffer.
__global__ __launch_bounds__(64) void k_wf_advance(EngineDev E, uint32_t par) {
    wr::EngineCounters& c = *E.cnt;
    const uint32_t jb = threadIdx.x;
    const uint32_t it = c.iteration;
    wr::JobCtl& C = E.ctl[jb];
    if (C.state == wr::kJobRunning) {
        const wr::JobDesc& J = E.jobs[jb];
        const uint32_t k = c.fold_k[jb];
        for (uint32_t q = 0; q < k; q++) J.pass_done[(C.cursor + q) % J.R] = 0;
        C.cursor += k;
        if (C.cursor >= J.npass) {
            C.state = wr::kJobDone;
            C.done_iter = it;
        }
    }
    if (jb == 0) {
        c.n[par] = 0;
        c.iteration = it + 1;
    }
}

// run_wavefront's admission in one launch: scene slot 0, job slot 0 (running,
// with the engine iteration it starts at) and the job's pass counters zeroed.
// Nothing else runs on the engine meanwhile (the previous call ended with a
// synchronisation).
__global__ __launch_bounds__(256) void k_wf_admit(SceneDev* scenes, wr::JobDesc* jobs, wr::JobCtl* ctl,
                                                  const wr::EngineCounters* cnt, SceneDev S, wr::JobDesc J) {
    for (uint32_t k = threadIdx.x; k < J.R; k += 256) J.pass_done[k] = 0;
    if (threadIdx.x == 0) {
        scenes[0] = S;
        jobs[0] = J;
        wr::JobCtl c{};
        c.state = wr::kJobRunning;
        c.start_iter = cnt->iteration;
        ctl[0] = c;
    }
}

}  // namespace wrd

namespace wr {

using namespace wrd;

static CamDev to_cam(const CamHost& c) {
    CamDev d;
    d.pos = f3{c.pos[0], c.pos[1], c.pos[2]};
    d.front = f3{c.front[0], c.front[1], c.front[2]};
    d.right = f3{c.right[0], c.right[1], c.right[2]};
    d.top = f3{c.top[0], c.top[1], c.top[2]};
    d.dist = c.dist;
    d.W = c.W;
    d.H = c.H;
    return d;
}

hipError_t launch_pt_mega(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass,
                          bool nee, const FilmBufs& film, hipStream_t st) {
    const uint32_t npix = cam.W * cam.H;
    const FilmDev f{film.sum, film.comp, film.sumsq};
    if (nee)
        k_pt_mega<true><<<(npix + 255) / 256, 256, kStackLdsBytes, st>>>(S, to_cam(cam), max_depth, seed, pass, f, npix);
    else
        k_pt_mega<false><<<(npix + 255) / 256, 256, kStackLdsBytes, st>>>(S, to_cam(cam), max_depth, seed, pass, f, npix);
    return hipGetLastError();
}

#define WR_CHECK(x)                      \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

Engine::~Engine() { release(); }

void Engine::release() {
    for (int m = 0; m < 2; m++) {
        if (exec[m]) (void)hipGraphExecDestroy(exec[m]);
        if (graph[m]) (void)hipGraphDestroy(graph[m]);
        if (tail_exec[m]) (void)hipGraphExecDestroy(tail_exec[m]);
        if (tail_graph[m]) (void)hipGraphDestroy(tail_graph[m]);
        exec[m] = tail_exec[m] = nullptr;
        graph[m] = tail_graph[m] = nullptr;
    }
    ready = false;
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    bufs.clear();
    if (contrib) (void)hipFree(contrib);
    if (pass_done) (void)hipFree(pass_done);
    if (ctl_back) (void)hipHostFree(ctl_back);
    contrib = nullptr;
    pass_done = nullptr;
    contrib_floats = pass_slots = 0;
    ctl_back = nullptr;
    cap = 0;
}

static PathBuf to_pathbuf(void* const* w) {
    PathBuf b;
    float** f[] = {&b.ox, &b.oy, &b.oz, &b.dx, &b.dy, &b.dz, &b.tx, &b.ty, &b.tz, &b.ht,
                   &b.ux, &b.uy, &b.uz, &b.px, &b.py, &b.pz};
    uint32_t** u[] = {&b.pixel, &b.pass, &b.depth, &b.inside, &b.contact, &b.hrec, &b.hobj, &b.flags, &b.job, &b.crec};
    static_assert(sizeof(f) / sizeof(f[0]) + sizeof(u) / sizeof(u[0]) == kPathWords, "path state words");
    int k = 0;
    for (float** p : f) *p = (float*)w[k++];
    for (uint32_t** p : u) *p = (uint32_t*)w[k++];
    return b;
}

hipError_t Engine::init(uint32_t want_cap, uint32_t want_grid, hipStream_t st, uint32_t want_lds) {
    if (want_grid == 0) want_grid = 2048;
    if (want_lds != 4 && want_lds != 16) want_lds = 8;
    want_cap = std::max<uint32_t>(256, (want_cap + 255) / 256 * 256);
    if (cap == want_cap && grid == want_grid && lds_levels == want_lds && ready) return hipSuccess;
    release();
    cap = want_cap;
    grid = want_grid;
    lds_levels = want_lds;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    for (int b = 0; b < 2; b++)
        for (int k = 0; k < kPathWords; k++)
            if (!(words[b][k] = alloc(4 * (size_t)cap))) return hipErrorOutOfMemory;
    if (!(cnt = (EngineCounters*)alloc(sizeof(EngineCounters)))) return hipErrorOutOfMemory;
    if (!(jobs = (JobDesc*)alloc(sizeof(JobDesc) * kMaxJobs))) return hipErrorOutOfMemory;
    if (!(ctl = (JobCtl*)alloc(sizeof(JobCtl) * kMaxJobs))) return hipErrorOutOfMemory;
    if (!(scenes = (SceneDev*)alloc(sizeof(SceneDev) * kMaxScenes))) return hipErrorOutOfMemory;
    if (!(exit_list = (uint32_t*)alloc(sizeof(uint32_t) * (size_t)cap))) return hipErrorOutOfMemory;
    WR_CHECK(hipMemsetAsync(cnt, 0, sizeof(EngineCounters), st));
    WR_CHECK(hipMemsetAsync(jobs, 0, sizeof(JobDesc) * kMaxJobs, st));
    WR_CHECK(hipMemsetAsync(ctl, 0, sizeof(JobCtl) * kMaxJobs, st));
    WR_CHECK(hipMemsetAsync(scenes, 0, sizeof(SceneDev) * kMaxScenes, st));
    WR_CHECK(hipStreamSynchronize(st));
    // the persistent trace kernel runs as many workgroups as fit at once
    int dev = 0, cus = 0, per_cu = 0;
    WR_CHECK(hipGetDevice(&dev));
    WR_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const void* trace_fn = lds_levels == 4    ? (const void*)k_wf_trace<4>
                           : lds_levels == 16 ? (const void*)k_wf_trace<16>
                                              : (const void*)k_wf_trace<8>;
    const size_t trace_lds = lds_levels == 4 ? stack_lds_bytes<4>() : lds_levels == 16 ? stack_lds_bytes<16>()
                                                                                       : stack_lds_bytes<8>();
    WR_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trace_fn, 256, trace_lds));
    trace_grid = (uint32_t)std::max(1, std::min<int>(cus * std::max(per_cu, 1), (int)(cap / 256)));
    ready = true;
    return hipSuccess;
}

// The iteration graphs (kGraphIters and kTailIters iterations: even, the buffer
// parity alternates), captured from the direct launch sequence the first time a
// replay needs them, so direct launches never create a graph.
hipError_t Engine::ensure_graphs(hipStream_t st, bool nee) {
    const int m = nee ? 1 : 0;
    if (exec[m] && tail_exec[m]) return hipSuccess;
    auto capture = [&](int iters, hipGraph_t* g, hipGraphExec_t* x) -> hipError_t {
        WR_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const hipError_t e = enqueue(iters, st, nullptr, nee);
        hipGraph_t cg = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &cg);
        if (e != hipSuccess || ec != hipSuccess) {
            if (cg) (void)hipGraphDestroy(cg);
            return e != hipSuccess ? e : ec;
        }
        *g = cg;
        return hipGraphInstantiate(x, *g, nullptr, nullptr, 0);
    };
    static_assert(kGraphIters % 2 == 0 && kTailIters % 2 == 0, "replays end on the parity they start on");
    if (!exec[m]) WR_CHECK(capture(kGraphIters, &graph[m], &exec[m]));
    if (!tail_exec[m]) WR_CHECK(capture(kTailIters, &tail_graph[m], &tail_exec[m]));
    return hipSuccess;
}

static EngineDev engine_dev(const Engine& En) {
    EngineDev E;
    E.buf[0] = to_pathbuf(En.words[0]);
    E.buf[1] = to_pathbuf(En.words[1]);
    E.cnt = En.cnt;
    E.jobs = En.jobs;
    E.ctl = En.ctl;
    E.scenes = En.scenes;
    E.exit_list = En.exit_list;
    E.cap = En.cap;
    return E;
}

template <bool COUNT>
static void launch_trace(const EngineDev& E, uint32_t lds_levels, uint32_t trace_grid, uint32_t par, hipStream_t st) {
    if (lds_levels == 4)
        k_wf_trace<4, COUNT><<<trace_grid, 256, stack_lds_bytes<4>(), st>>>(E, par);
    else if (lds_levels == 16)
        k_wf_trace<16, COUNT><<<trace_grid, 256, stack_lds_bytes<16>(), st>>>(E, par);
    else
        k_wf_trace<8, COUNT><<<trace_grid, 256, stack_lds_bytes<8>(), st>>>(E, par);
}

hipError_t Engine::enqueue(int iters, hipStream_t st, hipEvent_t* ev, bool nee) {
    const EngineDev E = engine_dev(*this);
    int k = 0;
    auto mark = [&]() -> hipError_t { return ev ? hipEventRecord(ev[k++], st) : hipSuccess; };
    WR_CHECK(mark());
    for (int it = 0; it < iters; it++) {
        const uint32_t par = (uint32_t)(it & 1);
        if (ev) launch_trace<true>(E, lds_levels, trace_grid, par, st);
        else launch_trace<false>(E, lds_levels, trace_grid, par, st);
        WR_CHECK(mark());
        k_wf_finish<<<grid, 256, kStackLdsBytes, st>>>(E, par);
        WR_CHECK(mark());
        if (nee) {
            k_wf_shade<false, true><<<grid, 256, kStackLdsBytes, st>>>(E, par);
            WR_CHECK(mark());
            k_wf_shade<true, true><<<grid, 256, kStackLdsBytes, st>>>(E, par);
        } else {
            k_wf_shade<false, false><<<grid, 256, kStackLdsBytes, st>>>(E, par);
            WR_CHECK(mark());
            k_wf_shade<true, false><<<grid, 256, kStackLdsBytes, st>>>(E, par);
        }
        WR_CHECK(mark());
        k_wf_plan<<<1, 64, 0, st>>>(E, par);
        WR_CHECK(mark());
        k_wf_generate<<<grid, 256, 0, st>>>(E, par);
        WR_CHECK(mark());
        k_wf_fold<<<grid, 256, 0, st>>>(E);
        WR_CHECK(mark());
        k_wf_advance<<<1, 64, 0, st>>>(E, par);
        WR_CHECK(mark());
    }
    return hipGetLastError();
}

hipError_t Engine::set_scene(uint32_t slot, const SceneDev& S, hipStream_t st) {
    WR_CHECK(hipMemcpyAsync(scenes + slot, &S, sizeof(SceneDev), hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);
}

// Admits a job into a free slot (between replays, on the engine's stream).
hipError_t Engine::admit(uint32_t slot, const JobDesc& J, hipStream_t st) {
    WR_CHECK(hipMemcpyAsync(jobs + slot, &J, sizeof(JobDesc), hipMemcpyHostToDevice, st));
    JobCtl c{};
    c.state = kJobRunning;
    WR_CHECK(hipMemcpyAsync(ctl + slot, &c, sizeof(JobCtl), hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);  // the sources are host stack copies
}

hipError_t Engine::cancel_slot(uint32_t slot, hipStream_t st) {
    const uint32_t v = kJobCancelled;
    WR_CHECK(hipMemcpyAsync(&ctl[slot].state, &v, sizeof(v), hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);
}

hipError_t Engine::release_slot(uint32_t slot, hipStream_t st) {
    JobCtl c{};
    WR_CHECK(hipMemcpyAsync(ctl + slot, &c, sizeof(JobCtl), hipMemcpyHostToDevice, st));
    return hipStreamSynchronize(st);
}

// One replay of the iteration graph, then a read-back of the job states.
hipError_t Engine::step(hipStream_t st, JobCtl* host_ctl, EngineCounters* host_cnt, bool nee, bool tail,
                        bool direct) {
    if (direct) {
        WR_CHECK(enqueue(tail ? kTailIters : kGraphIters, st, nullptr, nee));
    } else {
        WR_CHECK(ensure_graphs(st, nee));
        const int m = nee ? 1 : 0;
        WR_CHECK(hipGraphLaunch(tail ? tail_exec[m] : exec[m], st));
    }
    if (host_ctl) WR_CHECK(hipMemcpyAsync(host_ctl, ctl, sizeof(JobCtl) * kMaxJobs, hipMemcpyDeviceToHost, st));
    if (host_cnt) WR_CHECK(hipMemcpyAsync(host_cnt, cnt, sizeof(EngineCounters), hipMemcpyDeviceToHost, st));
    return hipStreamSynchronize(st);
}

// Fills the camera part of a job record.
void fill_job_camera(JobDesc& J, const CamHost& c) {
    for (int k = 0; k < 3; k++) {
        J.pos[k] = c.pos[k];
        J.front[k] = c.front[k];
        J.right[k] = c.right[k];
        J.top[k] = c.top[k];
    }
    J.dist = c.dist;
    J.W = c.W;
    J.H = c.H;
}

static hipError_t run_wavefront_once(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed,
                                     uint32_t pass0, uint32_t npass, bool nee, const FilmBufs& film, Engine& E,
                                     uint32_t cap, uint32_t grid, WaveStats* stats, hipStream_t st,
                                     uint32_t lds_levels, int launch, StageStats* stage, hipEvent_t* ev);

// Renders passes [pass0, pass0 + npass) of one film on the engine, alone.  On
// any error the engine is released, so the next call starts from a fresh pool
// rather than from the live paths and counters of the failed one.
hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, bool nee, const FilmBufs& film, Engine& E, uint32_t cap, uint32_t grid,
                         WaveStats* stats, hipStream_t st, uint32_t lds_levels, int launch, StageStats* stage) {
    // per-stage statistics: one event before the first kernel of a replay and one after each kernel
    std::vector<hipEvent_t> ev;
    hipError_t e = hipSuccess;
    if (launch == kLaunchStageStats) {
        ev.assign(kStages * kGraphIters + 1, nullptr);
        for (hipEvent_t& x : ev)
            if (e == hipSuccess) e = hipEventCreate(&x);
    }
    if (e == hipSuccess)
        e = run_wavefront_once(S, cam, max_depth, seed, pass0, npass, nee, film, E, cap, grid, stats, st, lds_levels,
                               launch, stage, ev.empty() ? nullptr : ev.data());
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        E.release();
    }
    return e;
}

static hipError_t run_wavefront_once(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed,
                                     uint32_t pass0, uint32_t npass, bool nee, const FilmBufs& film, Engine& E,
                                     uint32_t cap, uint32_t grid, WaveStats* stats, hipStream_t st,
                                     uint32_t lds_levels, int launch, StageStats* stage, hipEvent_t* ev) {
    const uint32_t npix = cam.W * cam.H;
    if (npass == 0 || npix == 0) return hipSuccess;
    const uint64_t total = (uint64_t)npass * npix;
    if (cap == 0) cap = kDefaultPathSlots;
    WR_CHECK(E.init(cap, grid, st, lds_levels));
    if (launch == kLaunchGraph) WR_CHECK(E.ensure_graphs(st, nee));
    JobDesc J{};
    fill_job_camera(J, cam);
    J.scene = 0;
    J.max_depth = max_depth;
    J.flags = nee ? kJobNee : 0u;
    J.seed = seed;
    J.pass0 = pass0;
    J.npass = npass;
    J.npix = npix;
    J.R = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(npass, E.cap / npix + 2), kMaxPassWindow);
    // the job's pass window, the engine's buffers grown to fit it
    const size_t want_floats = 3 * (size_t)J.R * npix;
    if (E.contrib_floats < want_floats) {
        if (E.contrib) WR_CHECK(hipFree(E.contrib));
        E.contrib = nullptr;
        E.contrib_floats = 0;
        WR_CHECK(hipMalloc((void**)&E.contrib, sizeof(float) * want_floats));
        E.contrib_floats = want_floats;
    }
    if (E.pass_slots < J.R) {
        if (E.pass_done) WR_CHECK(hipFree(E.pass_done));
        E.pass_done = nullptr;
        E.pass_slots = 0;
        WR_CHECK(hipMalloc((void**)&E.pass_done, sizeof(uint32_t) * J.R));
        E.pass_slots = J.R;
    }
    if (!E.ctl_back) WR_CHECK(hipHostMalloc((void**)&E.ctl_back, sizeof(JobCtl), hipHostMallocDefault));
    J.contrib = E.contrib;
    J.pass_done = E.pass_done;
    J.fsum = film.sum;
    J.fcomp = film.comp;
    J.fsumsq = film.sumsq;
    EngineCounters c0{};
    if (ev) WR_CHECK(hipMemcpy(&c0, E.cnt, sizeof(EngineCounters), hipMemcpyDeviceToHost));
    k_wf_admit<<<1, 256, 0, st>>>(E.scenes, E.jobs, E.ctl, E.cnt, S, J);
    WR_CHECK(hipGetLastError());
    // every pass is generated within npass window turns and drains in at most
    // max_depth + 2 iterations after its last path is issued
    const uint64_t it_bound = 4 * (uint64_t)(max_depth + 4) * (total / E.cap + npass + 2) + 64;
    // every camera path fits the pool and the window: all are issued in the
    // first iteration, so the first replay is followed by a tail replay at once
    // (with per-stage events every replay is read back on its own)
    bool chain_tail = total <= E.cap && npass <= J.R && !ev;
    uint64_t iters = 0;
    bool tail = false;
    JobCtl& back = *E.ctl_back;
    hipError_t e = hipSuccess;
    auto replay = [&](bool short_one) -> hipError_t {
        const int n = short_one ? kTailIters : kGraphIters;
        iters += n;
        const int m = nee ? 1 : 0;
        if (launch == kLaunchGraph) return hipGraphLaunch(short_one ? E.tail_exec[m] : E.exec[m], st);
        return E.enqueue(n, st, ev, nee);
    };
    for (;;) {
        e = replay(tail);
        if (e == hipSuccess && chain_tail) {
            e = replay(true);
            chain_tail = false;
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&back, E.ctl, sizeof(JobCtl), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess && ev && stage) {
            const int n = tail ? kTailIters : kGraphIters;
            for (int k = 0; k < kStages * n && e == hipSuccess; k++) {
                float ms = 0.f;
                e = hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
                stage->ms[k % kStages] += ms;
            }
        }
        if (e != hipSuccess || back.state == kJobDone) break;
        tail = back.next_work >= total;  // all camera paths issued: only draining is left
        if (iters > it_bound) {  // cannot happen
            e = hipErrorUnknown;
            break;
        }
    }
    // job slot 0 stays done until the next call's admission overwrites it
    if (e != hipSuccess) return e;
    if (ev && stage) {
        EngineCounters c1{};
        WR_CHECK(hipMemcpy(&c1, E.cnt, sizeof(EngineCounters), hipMemcpyDeviceToHost));
        stage->traced += c1.traced - c0.traced;
        stage->exits += c1.exits - c0.exits;
        stage->generated += c1.generated - c0.generated;
        stage->nodes += c1.trace_nodes - c0.trace_nodes;
        stage->tris += c1.trace_tris - c0.trace_tris;
    }
    if (stats) {
        stats->iterations = back.done_iter + 1 - back.start_iter;
        stats->paths = total;
    }
    return hipSuccess;
}

}  // namespace wr
