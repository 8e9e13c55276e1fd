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
// Live paths are kept dense: every stage reads path i of the current buffer
// (coalesced SoA loads) and the shade stage appends survivors, then the
// generate stage appends new camera paths, to the other buffer with one atomic
// per wave.  Primary (camera) rays are traced by the extend stage like any
// other ray.  Finished paths write their sample to a per-(pass, pixel) slot;
// the fold stage adds complete passes to the Kahan film in pass order.
constexpr uint32_t kPrimary = 1u;
constexpr uint32_t kDead = 2u;

struct PathBuf {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    float *tx, *ty, *tz, *ht;
    uint32_t *pixel, *pass, *depth, *inside, *contact, *hrec, *hobj, *flags;
};

struct WaveDev {
    PathBuf buf[2];
    wr::WaveCounters* cnt;
    float* contrib;  // R * npix * 3: samples of the passes in flight
    uint32_t cap, npix, R, npass;
    uint64_t total;
};

__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, uint32_t amount = 1) {
    // one atomic per wave; lanes get consecutive slots in lane order
    const uint64_t mask = __ballot(1);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t rank = __popcll(mask & ((1ull << lane) - 1));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask) * amount);
    base = __shfl(base, leader);
    return base + rank * amount;
}

// pass_done[key] += 1 for every active lane, with one atomic per distinct key in
// the wave (the lanes of a wave almost always finish paths of one pass).
__device__ __forceinline__ void count_done(uint32_t* pass_done, uint32_t key) {
    for (;;) {
        const uint64_t act = __ballot(1);
        const uint32_t leader = __ffsll((unsigned long long)act) - 1;
        const uint32_t k = __shfl(key, leader);
        const uint64_t same = __ballot(key == k);
        if (key == k) {
            if ((threadIdx.x & 63) == leader) atomicAdd(&pass_done[k], (uint32_t)__popcll(same));
            return;
        }
    }
}

__device__ __forceinline__ void finish(const WaveDev& W, const CamDev& cam, uint32_t p, uint32_t ps, f3 color) {
    if (!legal(color)) color = mk(0.f, 0.f, 0.f);
    const f3 v = fix_vignetting(cam, color, p);
    const uint32_t key = ps % W.R;
    float* c = W.contrib + 3 * ((size_t)key * W.npix + p);
    c[0] = v.x;
    c[1] = v.y;
    c[2] = v.z;
    count_done(W.cnt->pass_done, key);
}

__global__ __launch_bounds__(256) void k_wf_extend(SceneDev S, WaveDev W, CamDev cam, uint32_t par) {
    const PathBuf& A = W.buf[par];
    const uint32_t n = W.cnt->n[par];
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const f3 o = mk(A.ox[i], A.oy[i], A.oz[i]);
        const f3 d = mk(A.dx[i], A.dy[i], A.dz[i]);
        const uint32_t fl = A.flags[i];
        PHit h;
        if (fl & kPrimary) {  // camera ray: Scene::intersect without Condition + findInsideObject (camera.cpp:52-56)
            h = trace_scene<false>(sc, o, d, WR_NONE, WR_NONE, st);
            A.inside[i] = find_inside(S, o, d, WR_NONE, st);
        } else {
            h = trace_scene<true>(sc, o, d, A.inside[i], A.contact[i], st);
        }
        if (h.rec == WR_NONE) {  // escaped (no environment light) or no object to scatter: sample 0
            finish(W, cam, A.pixel[i], A.pass[i], mk(0.f, 0.f, 0.f));
            A.flags[i] = kDead;
            continue;
        }
        A.ht[i] = h.t;
        A.hrec[i] = h.rec;
        A.hobj[i] = h.obj;
        A.flags[i] = 0;
    }
}

__global__ __launch_bounds__(256) void k_wf_shade(SceneDev S, WaveDev W, CamDev cam, uint32_t max_depth,
                                                  uint64_t seed, uint32_t pass0, uint32_t par) {
    const PathBuf& A = W.buf[par];
    const PathBuf& B = W.buf[par ^ 1];
    const uint32_t n = W.cnt->n[par];
    LaneStack st;
    st.lane = threadIdx.x;
    const SceneCtx sc = make_scene_ctx(S);
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (A.flags[i] & kDead) continue;
        const f3 o = mk(A.ox[i], A.oy[i], A.oz[i]);
        const f3 d = mk(A.dx[i], A.dy[i], A.dz[i]);
        PHit h;
        h.t = A.ht[i];
        h.rec = A.hrec[i];
        h.obj = A.hobj[i];
        const uint32_t inside = A.inside[i], depth = A.depth[i], p = A.pixel[i], ps = A.pass[i];
        const uint32_t self = inside != WR_NONE ? inside : h.obj;
        float rn[4];
        rng4(seed, p, pass0 + ps, depth + 1, rn);
        const Out out = scatter(sc, self, o, d, h, inside, rn, st);
        f3 T = mk(A.tx[i], A.ty[i], A.tz[i]);
        if ((double)len(out.d) < 0.5) {  // emission / absorption ends the path
            finish(W, cam, p, ps, mulv(T, divs(divs(out.rad, out.op), out.dp)));
            continue;
        }
        if (depth + 1 > max_depth) {
            finish(W, cam, p, ps, mk(0.f, 0.f, 0.f));
            continue;
        }
        // the weight of this vertex only matters if the next ray hits; applying it
        // now keeps the per-pixel loop's multiplication order
        T = muls(T, cosine_term(out));
        T = mulv(T, divs(divs(out.rad, out.op), out.dp));
        const uint32_t j = wave_append(&W.cnt->n[par ^ 1]);
        B.ox[j] = out.o.x; B.oy[j] = out.o.y; B.oz[j] = out.o.z;
        B.dx[j] = out.d.x; B.dy[j] = out.d.y; B.dz[j] = out.d.z;
        B.tx[j] = T.x; B.ty[j] = T.y; B.tz[j] = T.z;
        B.pixel[j] = p;
        B.pass[j] = ps;
        B.depth[j] = depth + 1;
        B.inside[j] = out.inside;
        B.contact[j] = out.contact;
        B.flags[j] = 0;
    }
}

// One lane: how many new camera paths fit behind the survivors, inside the
// window of passes whose sample slots are free.
__global__ void k_wf_plan(WaveDev W, uint32_t par) {
    wr::WaveCounters& c = *W.cnt;
    const uint32_t nb = c.n[par ^ 1];
    const uint64_t lim_w = (uint64_t)(c.cursor + W.R) * W.npix;
    const uint64_t limit = lim_w < W.total ? lim_w : W.total;
    const uint64_t remaining = limit > c.next_work ? limit - c.next_work : 0;
    const uint32_t room = W.cap - nb;
    const uint32_t take = (uint32_t)(remaining < room ? remaining : room);
    c.take = take;
    c.take_base = nb;
    c.take_work = c.next_work;
    c.n[par ^ 1] = nb + take;
    c.next_work += take;
}

__global__ __launch_bounds__(256) void k_wf_generate(WaveDev W, CamDev cam, uint64_t seed, uint32_t pass0,
                                                     uint32_t par) {
    const PathBuf& B = W.buf[par ^ 1];
    const uint32_t take = W.cnt->take, base = W.cnt->take_base;
    const uint64_t work0 = W.cnt->take_work;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < take; i += gridDim.x * 256) {
        const uint64_t w = work0 + i;
        const uint32_t ps = (uint32_t)(w / W.npix), p = (uint32_t)(w % W.npix);
        float rn[4];
        rng4(seed, p, pass0 + ps, 0, rn);
        const f3 d = cam_dir(cam, p, rn[0], rn[1]);
        float c0 = fabsf(dot(cam.front, d));  // camera cosineTerm (ray.cpp:52-58, camera.cpp:182-186)
        c0 = (double)c0 > kEps ? c0 : (float)kEps;
        const uint32_t j = base + i;
        B.ox[j] = cam.pos.x; B.oy[j] = cam.pos.y; B.oz[j] = cam.pos.z;
        B.dx[j] = d.x; B.dy[j] = d.y; B.dz[j] = d.z;
        B.tx[j] = c0; B.ty[j] = c0; B.tz[j] = c0;
        B.pixel[j] = p;
        B.pass[j] = ps;
        B.depth[j] = 0;
        B.inside[j] = WR_NONE;
        B.contact[j] = WR_NONE;
        B.flags[j] = kPrimary;
    }
}

// Number of consecutive complete passes starting at the fold cursor.
__device__ __forceinline__ uint32_t foldable(const WaveDev& W) {
    const wr::WaveCounters& c = *W.cnt;
    uint32_t k = 0;
    while (k < W.R && c.cursor + k < W.npass && c.pass_done[(c.cursor + k) % W.R] == W.npix) k++;
    return k;
}

// Adds every complete pass (in pass order) to the film; pixel-parallel.
__global__ __launch_bounds__(256) void k_wf_fold(WaveDev W, FilmDev film) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    const uint32_t k = foldable(W);
    if (p >= W.npix) return;
    const uint32_t cur = W.cnt->cursor;
    for (uint32_t j = 0; j < k; j++) {
        const float* c = W.contrib + 3 * ((size_t)((cur + j) % W.R) * W.npix + p);
        film_add(film, p, mk(c[0], c[1], c[2]));
    }
}

// One lane: advance the fold cursor, re-arm folded pass slots, empty the consumed buffer.
__global__ void k_wf_advance(WaveDev W, uint32_t par) {
    const uint32_t k = foldable(W);
    wr::WaveCounters& c = *W.cnt;
    for (uint32_t j = 0; j < k; j++) c.pass_done[(c.cursor + j) % W.R] = 0;
    c.cursor += k;
    c.n[par] = 0;
    c.iterations++;
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

#define WR_CHECK(x)                      \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

WavefrontPool::~WavefrontPool() { release(); }

void WavefrontPool::release() {
    for (void* p : bufs)
        if (p) (void)hipFree(p);
    bufs.clear();
    cap = contrib_floats = 0;
}

hipError_t WavefrontPool::ensure(uint32_t want_cap, size_t want_contrib_floats) {
    if (cap >= want_cap && contrib_floats >= want_contrib_floats && !bufs.empty()) return hipSuccess;
    release();
    cap = want_cap;
    contrib_floats = want_contrib_floats;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(bytes, 4)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return p;
    };
    const size_t w4 = 4 * (size_t)cap;
    for (int b = 0; b < 2; b++) {
        for (int k = 0; k < kPathWords; k++)
            if (!(words[b][k] = alloc(w4))) return hipErrorOutOfMemory;
    }
    if (!(cnt = (WaveCounters*)alloc(sizeof(WaveCounters)))) return hipErrorOutOfMemory;
    if (!(contrib = (float*)alloc(sizeof(float) * contrib_floats))) return hipErrorOutOfMemory;
    return hipSuccess;
}

static PathBuf to_pathbuf(void* const* w) {
    PathBuf b;
    float** f[] = {&b.ox, &b.oy, &b.oz, &b.dx, &b.dy, &b.dz, &b.tx, &b.ty, &b.tz, &b.ht};
    uint32_t** u[] = {&b.pixel, &b.pass, &b.depth, &b.inside, &b.contact, &b.hrec, &b.hobj, &b.flags};
    int k = 0;
    for (float** p : f) *p = (float*)w[k++];
    for (uint32_t** p : u) *p = (uint32_t*)w[k++];
    return b;
}

// Renders passes [pass0, pass0 + npass) of the film with the wavefront stages.
// Iterations run as replays of a captured graph of `kGraphIters` iterations;
// the host reads one counter per replay to detect completion.
hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, const FilmBufs& film, WavefrontPool& P, uint32_t cap, uint32_t grid,
                         WaveStats* stats, hipStream_t st) {
    const uint32_t npix = cam.W * cam.H;
    if (npass == 0 || npix == 0) return hipSuccess;
    const uint64_t total = (uint64_t)npass * npix;
    if (cap == 0) cap = (uint32_t)std::min<uint64_t>(total, 1u << 22);
    cap = std::max<uint32_t>(cap, 256);
    uint32_t R = (uint32_t)std::min<uint64_t>(npass, cap / npix + 2);
    R = std::min<uint32_t>(R, kMaxPassWindow);
    WR_CHECK(P.ensure(cap, (size_t)3 * R * npix));
    WaveDev W;
    W.buf[0] = to_pathbuf(P.words[0]);
    W.buf[1] = to_pathbuf(P.words[1]);
    W.cnt = P.cnt;
    W.contrib = P.contrib;
    W.cap = cap;
    W.npix = npix;
    W.R = R;
    W.npass = npass;
    W.total = total;
    const CamDev C = to_cam(cam);
    const FilmDev F = to_film(film);
    if (grid == 0) grid = 2048;
    WaveCounters init{};
    WR_CHECK(hipMemcpyAsync(P.cnt, &init, sizeof(init), hipMemcpyHostToDevice, st));
    // capture kGraphIters iterations (an even number: buffer parity alternates)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    WR_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    for (int it = 0; it < kGraphIters; it++) {
        const uint32_t par = (uint32_t)(it & 1);
        k_wf_extend<<<grid, 256, kStackLdsBytes, st>>>(S, W, C, par);
        k_wf_shade<<<grid, 256, kStackLdsBytes, st>>>(S, W, C, max_depth, seed, pass0, par);
        k_wf_plan<<<1, 1, 0, st>>>(W, par);
        k_wf_generate<<<grid, 256, 0, st>>>(W, C, seed, pass0, par);
        k_wf_fold<<<(npix + 255) / 256, 256, 0, st>>>(W, F);
        k_wf_advance<<<1, 1, 0, st>>>(W, par);
    }
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return e;
    }
    WaveCounters host{};
    uint64_t replays = 0;
    for (;;) {
        e = hipGraphLaunch(exec, st);
        if (e != hipSuccess) break;
        replays++;
        e = hipMemcpyAsync(&host, P.cnt, sizeof(host), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess || host.cursor >= npass) break;
        if (replays > 4 * (uint64_t)(max_depth + 3) * (total / cap + 2)) {  // cannot happen: every iteration advances every path
            e = hipErrorUnknown;
            break;
        }
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return e;
    if (stats) {
        stats->iterations = host.iterations;
        stats->paths = total;
    }
    return hipSuccess;
}

}  // namespace wr
