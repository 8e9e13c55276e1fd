// pt.hip -- path-tracing kernels (gfx950) and their host-side drivers.
//
//  k_pt_mega        the plain one-thread-per-pixel kernel: one launch per
//                   pass, each lane runs its pixel's whole path
//                   (pathTracer.cpp:20-70 with mcRenderer.cpp:16-57) and adds
//                   the sample to the Kahan film it owns.
//  wavefront        the same estimator split into stages over a pool of path
//                   slots: generate (camera ray + primary hit + inside object),
//                   shade (material scatter + Russian roulette), extend
//                   (NoSelf closest hit); finished paths write their sample to
//                   a per-(pass, pixel) slot; complete passes are folded into
//                   the Kahan film in pass order, so the film is bit-identical
//                   to the megakernel's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pt_dev.h"
#include "wr_pt.h"

#pragma clang fp contract(off)

namespace wrd {

__global__ __launch_bounds__(256) void k_pt_mega(SceneDev S, CamDev cam, uint32_t max_depth, uint64_t seed,
                                                 uint32_t pass, FilmDev film, uint32_t npix) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    const f3 v = sample_path(sc, cam, max_depth, seed, p, pass, st);
    film_add(film, p, v);
}

// ---- wavefront --------------------------------------------------------------
struct WaveDev {
    // per-slot path state (SoA); pass is relative to the render call's first pass
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tx, *ty, *tz;
    uint32_t *pixel, *pass, *depth, *inside, *contact;
    float* ht;
    uint32_t *hrec, *hobj;
    // queues
    uint32_t* q_shade[2];
    uint32_t* q_extend;
    uint32_t* q_free;  // ring of capacity nslots
    wr::WaveCounters* cnt;
    // per-(pass window slot, pixel) samples
    float* contrib;   // R * npix * 3
    uint32_t nslots, npix, R;
};

__device__ __forceinline__ uint32_t wave_append(uint32_t* counter) {
    // one atomic per wave: lanes that append get consecutive indices
    const uint64_t mask = __ballot(1);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t rank = __popcll(mask & ((1ull << lane) - 1));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + rank;
}

__device__ __forceinline__ void finish_path(const WaveDev& W, const CamDev& cam, uint32_t slot, f3 color) {
    if (!legal(color)) color = mk(0.f, 0.f, 0.f);
    const uint32_t p = W.pixel[slot], ps = W.pass[slot];
    const f3 v = fix_vignetting(cam, color, p);
    float* c = W.contrib + 3 * ((size_t)(ps % W.R) * W.npix + p);
    c[0] = v.x;
    c[1] = v.y;
    c[2] = v.z;
    atomicAdd(&W.cnt->pass_done[ps % W.R], 1u);
    atomicAdd(&W.cnt->finished, 1u);
    const uint32_t at = wave_append(&W.cnt->free_tail);
    W.q_free[at % W.nslots] = slot;
}

// Iteration prologue (one lane): reset the queues the iteration appends to and
// decide how many new paths the generate stage starts.
__global__ void k_wf_plan(WaveDev W, uint32_t parity, uint64_t work_limit) {
    wr::WaveCounters& c = *W.cnt;
    c.n_extend = 0;
    c.n_shade[parity ^ 1] = 0;
    const uint32_t free_n = c.free_tail - c.free_head;
    const uint64_t remaining = work_limit > c.next_work ? work_limit - c.next_work : 0;
    const uint32_t take = (uint32_t)(remaining < free_n ? remaining : free_n);
    c.take = take;
    c.take_head = c.free_head;
    c.take_work = c.next_work;
    c.free_head += take;
    c.next_work += take;
}

__global__ __launch_bounds__(256) void k_wf_generate(SceneDev S, WaveDev W, CamDev cam, uint64_t seed,
                                                     uint32_t pass0, uint32_t parity) {
    const uint32_t take = W.cnt->take, head = W.cnt->take_head;
    const uint64_t work0 = W.cnt->take_work;
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < take; i += gridDim.x * 256) {
        const uint32_t slot = W.q_free[(head + i) % W.nslots];
        const uint64_t w = work0 + i;
        const uint32_t ps = (uint32_t)(w / W.npix), p = (uint32_t)(w % W.npix);
        W.pixel[slot] = p;
        W.pass[slot] = ps;
        float rn[4];
        rng4(seed, p, pass0 + ps, 0, rn);
        const f3 o = cam.pos;
        const f3 d = cam_dir(cam, p, rn[0], rn[1]);
        const PHit hit = trace_scene<false>(sc, o, d, WR_NONE, WR_NONE, st);
        const uint32_t inside = find_inside(S, o, d, WR_NONE, st);
        float c0 = fabsf(dot(cam.front, d));
        c0 = (double)c0 > kEps ? c0 : (float)kEps;
        const uint32_t self = inside != WR_NONE ? inside : hit.obj;
        if (self == WR_NONE || hit.rec == WR_NONE) {
            finish_path(W, cam, slot, mk(0.f, 0.f, 0.f));
            continue;
        }
        W.ox[slot] = o.x; W.oy[slot] = o.y; W.oz[slot] = o.z;
        W.dx[slot] = d.x; W.dy[slot] = d.y; W.dz[slot] = d.z;
        W.tx[slot] = c0; W.ty[slot] = c0; W.tz[slot] = c0;
        W.depth[slot] = 0;
        W.inside[slot] = inside;
        W.ht[slot] = hit.t;
        W.hrec[slot] = hit.rec;
        W.hobj[slot] = hit.obj;
        const uint32_t at = wave_append(&W.cnt->n_shade[parity]);
        W.q_shade[parity][at] = slot;
    }
}

__global__ __launch_bounds__(256) void k_wf_shade(SceneDev S, WaveDev W, CamDev cam, uint32_t max_depth,
                                                  uint64_t seed, uint32_t pass0, uint32_t parity) {
    const uint32_t n = W.cnt->n_shade[parity];
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t slot = W.q_shade[parity][i];
        const f3 o = mk(W.ox[slot], W.oy[slot], W.oz[slot]);
        const f3 d = mk(W.dx[slot], W.dy[slot], W.dz[slot]);
        PHit h;
        h.t = W.ht[slot];
        h.rec = W.hrec[slot];
        h.obj = W.hobj[slot];
        const uint32_t inside = W.inside[slot];
        const uint32_t depth = W.depth[slot];
        const uint32_t self = inside != WR_NONE ? inside : h.obj;
        float rn[4];
        rng4(seed, W.pixel[slot], pass0 + W.pass[slot], depth + 1, rn);
        const Out out = scatter(sc, self, o, d, h, inside, rn, st);
        f3 T = mk(W.tx[slot], W.ty[slot], W.tz[slot]);
        if ((double)len(out.d) < 0.5) {
            finish_path(W, cam, slot, mulv(T, divs(divs(out.rad, out.op), out.dp)));
            continue;
        }
        if (depth + 1 > max_depth) {
            finish_path(W, cam, slot, mk(0.f, 0.f, 0.f));
            continue;
        }
        // the weight of this vertex only matters if the next ray hits; applying it
        // now keeps the same multiplication order as the per-pixel loop
        T = muls(T, cosine_term(out));
        T = mulv(T, divs(divs(out.rad, out.op), out.dp));
        W.tx[slot] = T.x; W.ty[slot] = T.y; W.tz[slot] = T.z;
        W.ox[slot] = out.o.x; W.oy[slot] = out.o.y; W.oz[slot] = out.o.z;
        W.dx[slot] = out.d.x; W.dy[slot] = out.d.y; W.dz[slot] = out.d.z;
        W.inside[slot] = out.inside;
        W.contact[slot] = out.contact;
        const uint32_t at = wave_append(&W.cnt->n_extend);
        W.q_extend[at] = slot;
    }
}

__global__ __launch_bounds__(256) void k_wf_extend(SceneDev S, WaveDev W, CamDev cam, uint32_t parity) {
    const uint32_t n = W.cnt->n_extend;
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t slot = W.q_extend[i];
        const f3 o = mk(W.ox[slot], W.oy[slot], W.oz[slot]);
        const f3 d = mk(W.dx[slot], W.dy[slot], W.dz[slot]);
        const PHit h = trace_scene<true>(sc, o, d, W.inside[slot], W.contact[slot], st);
        if (h.rec == WR_NONE) {
            finish_path(W, cam, slot, mk(0.f, 0.f, 0.f));
            continue;
        }
        W.ht[slot] = h.t;
        W.hrec[slot] = h.rec;
        W.hobj[slot] = h.obj;
        W.depth[slot] = W.depth[slot] + 1;
        const uint32_t at = wave_append(&W.cnt->n_shade[parity ^ 1]);
        W.q_shade[parity ^ 1][at] = slot;
    }
}

// Adds one complete pass to the film (pixel-parallel, Kahan), then re-arms its window slot.
__global__ __launch_bounds__(256) void k_wf_fold(WaveDev W, FilmDev film, uint32_t pass) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= W.npix) return;
    const float* c = W.contrib + 3 * ((size_t)(pass % W.R) * W.npix + p);
    film_add(film, p, mk(c[0], c[1], c[2]));
}

__global__ void k_wf_rearm(WaveDev W, uint32_t pass) { W.cnt->pass_done[pass % W.R] = 0; }

__global__ void k_wf_init_free(WaveDev W) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < W.nslots) W.q_free[i] = i;
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

static FilmDev to_film(const FilmBufs& f) { return FilmDev{f.sum, f.comp, f.sumsq}; }

hipError_t launch_pt_mega(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass,
                          const FilmBufs& film, hipStream_t st) {
    const uint32_t npix = cam.W * cam.H;
    k_pt_mega<<<(npix + 255) / 256, 256, kStackLdsBytes, st>>>(S, to_cam(cam), max_depth, seed, pass,
                                                                to_film(film), npix);
    return hipGetLastError();
}

#define WR_CHECK(x)                          \
    do {                                     \
        hipError_t e_ = (x);                 \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

WavefrontPool::~WavefrontPool() { release(); }

void WavefrontPool::release() {
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    bufs.clear();
    nslots = npix = R = 0;
}

hipError_t WavefrontPool::ensure(uint32_t want_slots, uint32_t want_npix, uint32_t want_R) {
    if (nslots >= want_slots && npix == want_npix && R == want_R && !bufs.empty()) return hipSuccess;
    release();
    nslots = want_slots;
    npix = want_npix;
    R = want_R;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    const size_t f4 = sizeof(float) * (size_t)nslots, u4 = sizeof(uint32_t) * (size_t)nslots;
    float** fp[] = {&ox, &oy, &oz, &dx, &dy, &dz, &tx, &ty, &tz, &ht};
    for (float** p : fp)
        if (!(*p = (float*)alloc(f4))) return hipErrorOutOfMemory;
    uint32_t** up[] = {&pixel, &pass, &depth, &inside, &contact, &hrec, &hobj, &q_shade0, &q_shade1, &q_extend, &q_free};
    for (uint32_t** p : up)
        if (!(*p = (uint32_t*)alloc(u4))) return hipErrorOutOfMemory;
    if (!(cnt = (WaveCounters*)alloc(sizeof(WaveCounters)))) return hipErrorOutOfMemory;
    if (!(contrib = (float*)alloc(sizeof(float) * 3 * (size_t)R * npix))) return hipErrorOutOfMemory;
    return hipSuccess;
}

static WaveDev to_wave(const WavefrontPool& P) {
    WaveDev W;
    W.ox = P.ox; W.oy = P.oy; W.oz = P.oz;
    W.dx = P.dx; W.dy = P.dy; W.dz = P.dz;
    W.tx = P.tx; W.ty = P.ty; W.tz = P.tz;
    W.pixel = P.pixel; W.pass = P.pass; W.depth = P.depth; W.inside = P.inside; W.contact = P.contact;
    W.ht = P.ht; W.hrec = P.hrec; W.hobj = P.hobj;
    W.q_shade[0] = P.q_shade0; W.q_shade[1] = P.q_shade1;
    W.q_extend = P.q_extend; W.q_free = P.q_free;
    W.cnt = P.cnt; W.contrib = P.contrib;
    W.nslots = P.nslots; W.npix = P.npix; W.R = P.R;
    return W;
}

// Renders passes [pass0, pass0 + npass) with the wavefront stages.
hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, const FilmBufs& film, WavefrontPool& P, uint32_t slots, uint32_t grid,
                         WaveStats* stats, hipStream_t st) {
    const uint32_t npix = cam.W * cam.H;
    const uint32_t R = 4;
    if (npass == 0 || npix == 0) return hipSuccess;
    const uint64_t total = (uint64_t)npass * npix;
    if (slots == 0) slots = (uint32_t)std::min<uint64_t>(total, 1u << 21);
    WR_CHECK(P.ensure(slots, npix, R));
    WaveDev W = to_wave(P);
    W.nslots = slots;
    CamDev C = to_cam(cam);
    WaveCounters init{};
    init.free_head = 0;
    init.free_tail = slots;
    init.next_work = 0;
    WR_CHECK(hipMemcpyAsync(P.cnt, &init, sizeof(init), hipMemcpyHostToDevice, st));
    k_wf_init_free<<<(slots + 255) / 256, 256, 0, st>>>(W);
    WR_CHECK(hipGetLastError());
    // work items are numbered pass-major from 0; passes are relative to pass0
    uint32_t folded = 0;  // passes (relative) already folded into the film
    WaveCounters host{};
    uint32_t parity = 0;
    uint64_t iters = 0;
    if (grid == 0) grid = 2048;
    while (folded < npass) {
        const uint64_t limit = std::min<uint64_t>(total, (uint64_t)(folded + R) * npix);
        k_wf_plan<<<1, 1, 0, st>>>(W, parity, limit);
        k_wf_generate<<<grid, 256, kStackLdsBytes, st>>>(S, W, C, seed, pass0, parity);
        k_wf_shade<<<grid, 256, kStackLdsBytes, st>>>(S, W, C, max_depth, seed, pass0, parity);
        k_wf_extend<<<grid, 256, kStackLdsBytes, st>>>(S, W, C, parity);
        WR_CHECK(hipGetLastError());
        WR_CHECK(hipMemcpyAsync(&host, P.cnt, sizeof(host), hipMemcpyDeviceToHost, st));
        WR_CHECK(hipStreamSynchronize(st));
        iters++;
        parity ^= 1;
        // fold every pass whose samples are all written, in pass order
        while (folded < npass && host.pass_done[folded % R] == npix) {
            k_wf_fold<<<(npix + 255) / 256, 256, 0, st>>>(W, FilmDev{film.sum, film.comp, film.sumsq}, folded);
            k_wf_rearm<<<1, 1, 0, st>>>(W, folded);
            WR_CHECK(hipGetLastError());
            host.pass_done[folded % R] = 0;
            folded++;
        }
    }
    if (stats) {
        stats->iterations = iters;
        stats->paths = host.finished;
    }
    return hipSuccess;
}

}  // namespace wr
