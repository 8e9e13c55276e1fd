// capi.cpp -- extern "C" entry points of libwalkray: contexts, device memory,
// scenes (build + upload) and ray queries.  See include/walkray.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/walkray.h"
#include "wr_capi_internal.h"
#include "wr_host.h"
#include "wr_kernels.h"

static thread_local std::string g_err;

namespace wr {

int fail(wr_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    g_err = msg;
    return code;
}

int hip_fail(wr_ctx* ctx, hipError_t e, const char* what) {
    return fail(ctx, WR_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class T>
static int upload(wr_ctx* ctx, const T* v, size_t n, T** d, uint64_t& bytes) {
    *d = nullptr;
    const size_t b = std::max<size_t>(n, 1) * sizeof(T);
    hipError_t e = hipMalloc((void**)d, b);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc(scene)");
    bytes += b;
    if (n) {
        e = hipMemcpy(*d, v, n * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(scene)");
    }
    return WR_OK;
}

template <class T>
static int upload(wr_ctx* ctx, const std::vector<T>& v, T** d, uint64_t& bytes) {
    return upload(ctx, v.data(), v.size(), d, bytes);
}

void scene_free_device(wr_scene* s) {
    for (void* p : s->dbufs)
        if (p) (void)hipFree(p);
    s->dbufs.clear();
    s->dev = SceneDev{};
    s->device_bytes = 0;
}

void place_object(HostObject& o, const io::Mat4& M) {
    o.verts.resize(o.local_verts.size());
    for (size_t i = 0; i + 2 < o.local_verts.size(); i += 3) M.apply(&o.local_verts[i], 1.f, &o.verts[i]);
    o.normals.resize(o.local_normals.size());
    if (!o.local_normals.empty()) {
        const io::Mat4 nm = io::transpose(io::inverse(M));
        for (size_t i = 0; i + 2 < o.local_normals.size(); i += 3) nm.apply(&o.local_normals[i], 0.f, &o.normals[i]);
    }
}

// Philox-4x32-10 on the host (Salmon, Moraes, Dror, Shaw, SC'11), the counter
// RNG of the kernels: the first of its four 24-bit uniforms.
float motion_time(uint64_t seed, uint32_t obj, uint32_t pass) {
    uint32_t c[4] = {obj, pass, 0xFFFFFFFEu, 0x57414C4Bu};
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return (float)(c[0] >> 8) * (1.0f / 16777216.0f);
}

}  // namespace wr

using namespace wr;

extern "C" {

const char* wr_last_error(const wr_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int wr_ctx_create(int device, wr_ctx** out) {
    if (!out) return fail(nullptr, WR_EINVAL, "wr_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(nullptr, WR_EINVAL, "wr_ctx_create: no such device");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
    std::unique_ptr<wr_ctx> c(new wr_ctx);
    c->device = device;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(nullptr, e, "hipStreamCreate");
    if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return hip_fail(nullptr, e, "hipEventCreate");
    if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return hip_fail(nullptr, e, "hipEventCreate");
    if ((e = hipMalloc(&c->d_work, wr::kWorkWords * sizeof(unsigned long long))) != hipSuccess) return hip_fail(nullptr, e, "hipMalloc");
    if ((e = hipMemset(c->d_work, 0, wr::kWorkWords * sizeof(unsigned long long))) != hipSuccess) return hip_fail(nullptr, e, "hipMemset");
    if ((e = wr::noalloc_scratch_init(device)) != hipSuccess) return hip_fail(nullptr, e, "hipGetSymbolAddress(scratch)");
    *out = c.release();
    return WR_OK;
}

void wr_ctx_destroy(wr_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipEventDestroy(ctx->ev0);
    (void)hipEventDestroy(ctx->ev1);
    (void)hipFree(ctx->d_work);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int wr_synchronize(wr_ctx* ctx) {
    if (!ctx) return fail(nullptr, WR_EINVAL, "wr_synchronize: ctx is NULL");
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
    return WR_OK;
}

int wr_device_alloc(wr_ctx* ctx, size_t bytes, void** d_out) {
    if (!ctx || !d_out) return fail(ctx, WR_EINVAL, "wr_device_alloc: bad argument");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(d_out, std::max<size_t>(bytes, 1));
    if (e != hipSuccess) return fail(ctx, WR_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return WR_OK;
}

int wr_device_free(wr_ctx* ctx, void* d_ptr) {
    if (!ctx) return fail(nullptr, WR_EINVAL, "wr_device_free: ctx is NULL");
    if (!d_ptr) return WR_OK;
    hipError_t e = hipFree(d_ptr);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    return WR_OK;
}

int wr_memcpy_htod(wr_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
    if (!ctx || (!d_dst && bytes) || (!src && bytes)) return fail(ctx, WR_EINVAL, "wr_memcpy_htod: bad argument");
    hipError_t e = hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(htod)");
    return WR_OK;
}

int wr_memcpy_dtoh(wr_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
    if (!ctx || (!d_src && bytes) || (!dst && bytes)) return fail(ctx, WR_EINVAL, "wr_memcpy_dtoh: bad argument");
    hipError_t e = hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(dtoh)");
    return WR_OK;
}

int wr_memcpy_dtod(wr_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx || (!d_src && bytes) || (!d_dst && bytes)) return fail(ctx, WR_EINVAL, "wr_memcpy_dtod: bad argument");
    hipError_t e = hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMemcpy(dtod)");
    return WR_OK;
}

int wr_timer_start(wr_ctx* ctx) {
    if (!ctx) return fail(nullptr, WR_EINVAL, "wr_timer_start: ctx is NULL");
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventRecord");
    return WR_OK;
}

int wr_timer_stop(wr_ctx* ctx, float* ms) {
    if (!ctx || !ms) return fail(ctx, WR_EINVAL, "wr_timer_stop: bad argument");
    hipError_t e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventElapsedTime");
    return WR_OK;
}

int wr_scene_create(wr_ctx* ctx, wr_scene** out) {
    if (!ctx || !out) return fail(ctx, WR_EINVAL, "wr_scene_create: bad argument");
    *out = new wr_scene;
    (*out)->ctx = ctx;
    return WR_OK;
}

void wr_scene_destroy(wr_scene* s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    scene_free_device(s);
    delete s;
}

int wr_scene_add_mesh(wr_scene* s, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                      const float* normals, const wr_material* mat, uint32_t* obj_id) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_scene_add_mesh: scene is NULL");
    wr_ctx* ctx = s->ctx;
    if ((!verts && nv) || (!idx && nt) || !mat) return fail(ctx, WR_EINVAL, "wr_scene_add_mesh: NULL array");
    if (mat->type < WR_DIFFUSE || mat->type > WR_EMISSIVE) return fail(ctx, WR_EINVAL, "wr_scene_add_mesh: unknown material");
    for (size_t i = 0; i < 3 * (size_t)nt; i++)
        if (idx[i] >= nv) return fail(ctx, WR_EINVAL, "wr_scene_add_mesh: vertex index out of range");
    HostObject o;
    o.verts.assign(verts, verts + 3 * (size_t)nv);
    o.idx.assign(idx, idx + 3 * (size_t)nt);
    if (normals) {  // kept per triangle corner
        o.normals.resize(9 * (size_t)nt);
        for (size_t c = 0; c < 3 * (size_t)nt; c++)
            for (int a = 0; a < 3; a++) o.normals[3 * c + a] = normals[3 * (size_t)idx[c] + a];
    }
    o.mat = *mat;
    s->objs.push_back(std::move(o));
    s->built = false;
    if (obj_id) *obj_id = (uint32_t)(s->objs.size() - 1);
    return WR_OK;
}

int wr_scene_set_motion(wr_scene* s, uint32_t obj, float start_time, float end_time, const float* start,
                        const float* end) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_scene_set_motion: scene is NULL");
    wr_ctx* ctx = s->ctx;
    if (obj >= s->objs.size() || !start || !end) return fail(ctx, WR_EINVAL, "wr_scene_set_motion: bad argument");
    io::Mat4 a, b;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a.e[i][j] = start[4 * i + j];
            b.e[i][j] = end[4 * i + j];
        }
    HostObject& o = s->objs[obj];
    if (!o.moving) {
        o.local_verts = o.verts;
        o.local_normals = o.normals;
        o.moving = true;
    }
    o.motion = io::make_motion(start_time, end_time, a, b);
    place_object(o, io::interpolate(o.motion, start_time));  // as built: the start transform
    s->built = false;
    return WR_OK;
}

int wr_scene_add_mesh_corners(wr_scene* s, const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt,
                              const float* corner_normals, const wr_material* mat, uint32_t* obj_id) {
    const int rc = wr_scene_add_mesh(s, verts, nv, idx, nt, nullptr, mat, obj_id);
    if (rc != WR_OK || !corner_normals) return rc;
    s->objs.back().normals.assign(corner_normals, corner_normals + 9 * (size_t)nt);
    return WR_OK;
}

// Mesh::getTriangleArea (mesh.h:80-88) in float: |(v1 - v0) x (v2 - v1)| / 2 with
// nvVector's cross and length (nvVector.h:490-525).
static float triangle_area(const HostObject& o, uint32_t t) {
    float v[3][3];
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) v[k][a] = o.verts[3 * (size_t)o.idx[3 * (size_t)t + k] + a];
    float l0[3], l1[3];
    for (int a = 0; a < 3; a++) {
        l0[a] = v[1][a] - v[0][a];
        l1[a] = v[2][a] - v[1][a];
    }
    const float c[3] = {l0[1] * l1[2] - l0[2] * l1[1], l0[2] * l1[0] - l0[0] * l1[2], l0[0] * l1[1] - l0[1] * l1[0]};
    float r = 0.f;
    r += c[0] * c[0];
    r += c[1] * c[1];
    r += c[2] * c[2];
    return std::sqrt(r) / 2;
}

int wr_scene_build(wr_scene* s, int strategy) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_scene_build: scene is NULL");
    if (strategy != WR_BEST)
        return fail(s->ctx, WR_EINVAL, "wr_scene_build: only BEST is supported (Scene always builds BEST, scene.cpp:41)");
    return wr::scene_build(s, nullptr);
}

}  // extern "C"

namespace wr {

void* RebuildCache::get(size_t slot, size_t bytes) {
    if (bufs.size() <= slot) bufs.resize(slot + 1);
    Buf& b = bufs[slot];
    if (b.cap < bytes) {
        if (b.p) (void)hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        const size_t cap = bytes + bytes / 4;
        if (hipHostMalloc(&b.p, cap, hipHostMallocDefault) != hipSuccess) {
            b.p = nullptr;
            return nullptr;
        }
        b.cap = cap;
    }
    return b.p;
}

RebuildCache::~RebuildCache() {
    for (Buf& b : bufs)
        if (b.p) (void)hipHostFree(b.p);
    for (uint32_t* p : d_order)
        if (p) (void)hipFree(p);
    if (d_rec_of) (void)hipFree(d_rec_of);
    if (d_nbox) (void)hipFree(d_nbox);
}

// The device refit of tree ti's fast-pass hierarchy for a pass (see
// WideRefit): its level order from the built scene's host tree on first use,
// scratch grown as needed.  d_wnodes: the pass's copy of the built scene's
// nodes; d_tris: the pass's records.
static hipError_t refit_on_device(RebuildCache& C, const wr_scene* built, size_t ti, wr_wnode* d_wnodes,
                                  const wr_tree_hdr& h, const wr_tri* d_tris, hipStream_t st) {
    const FlatTree& prev = built->trees[ti];
    if (C.d_order.size() < built->trees.size()) {
        C.d_order.resize(built->trees.size(), nullptr);
        C.level_start.resize(built->trees.size());
    }
    hipError_t e = hipSuccess;
    if (!C.d_order[ti]) {
        const std::vector<uint32_t> order = wide_levels(prev, C.level_start[ti]);
        e = hipMalloc((void**)&C.d_order[ti], std::max<size_t>(order.size(), 1) * sizeof(uint32_t));
        if (e != hipSuccess) return e;
        if (!order.empty())
            e = hipMemcpy(C.d_order[ti], order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
        if (e != hipSuccess) return e;
    }
    const size_t nr = prev.tris.size(), nw = prev.wnodes.size();
    if (C.rec_cap < nr) {
        if (C.d_rec_of) (void)hipFree(C.d_rec_of);
        C.d_rec_of = nullptr;
        C.rec_cap = 0;
        if ((e = hipMalloc((void**)&C.d_rec_of, nr * sizeof(uint32_t))) != hipSuccess) return e;
        C.rec_cap = nr;
    }
    if (C.box_cap < 6 * nw) {
        if (C.d_nbox) (void)hipFree(C.d_nbox);
        C.d_nbox = nullptr;
        C.box_cap = 0;
        if ((e = hipMalloc((void**)&C.d_nbox, 6 * nw * sizeof(float))) != hipSuccess) return e;
        C.box_cap = 6 * nw;
    }
    WideRefit R;
    R.wnodes = d_wnodes + h.wide_base;
    R.old_tris = built->dev.tris + built->hdrs[ti].rec_base;
    R.new_tris = d_tris + h.rec_base;
    R.n_recs = (uint32_t)nr;
    R.order = C.d_order[ti];
    R.level_start = C.level_start[ti];
    R.rec_of = C.d_rec_of;
    R.nbox = C.d_nbox;
    return launch_wide_refit(R, st);
}

int scene_pass(const wr_scene* s, uint64_t seed, uint32_t pass, wr_scene* out) {
    out->ctx = s->ctx;
    out->objs = s->objs;
    out->ch_tune = s->ch_tune;
    for (uint32_t oi = 0; oi < out->objs.size(); oi++) {
        HostObject& o = out->objs[oi];
        if (o.moving) place_object(o, io::interpolate(o.motion, motion_time(seed, oi, pass)));
    }
    return scene_build(out, s);  // static objects keep their trees; s stages the arrays
}

int scene_build(wr_scene* s, const wr_scene* reuse) {
    wr_ctx* ctx = s->ctx;
    try {
        (void)hipSetDevice(ctx->device);
        scene_free_device(s);
        const int nth = (int)std::max(1u, std::thread::hardware_concurrency());
        // global tree over the concatenated objects (scene.cpp:20-44)
        std::vector<float> V;
        std::vector<uint32_t> I;
        s->prim_base.assign(s->objs.size() + 1, 0);
        uint32_t voff = 0;
        for (size_t oi = 0; oi < s->objs.size(); oi++) {
            const HostObject& o = s->objs[oi];
            V.insert(V.end(), o.verts.begin(), o.verts.end());
            for (uint32_t x : o.idx) I.push_back(x + voff);
            voff += (uint32_t)(o.verts.size() / 3);
            s->prim_base[oi + 1] = s->prim_base[oi] + (uint32_t)(o.idx.size() / 3);
        }
        const uint32_t nt = (uint32_t)(I.size() / 3);
        const bool can_reuse = reuse && reuse->built && reuse->trees.size() == 1 + s->objs.size();
        s->trees.clear();
        // with `reuse`, the global tree's fast-pass hierarchy is refit rather than built
        // a pass of a render with moving objects: the rebuilt trees' fast-pass
        // hierarchies are refit on the device from the built scene's
        const bool gpu_refit = can_reuse;
        s->trees.push_back(build_tree(V.data(), (uint32_t)(V.size() / 3), I.data(), nt, nth,
                                      can_reuse ? &reuse->trees[0] : nullptr, gpu_refit));
        // per-object trees (scene.cpp:46-65).  A static object's tree from `reuse`
        // (the same scene before its moving objects were placed) is the tree this
        // deterministic builder would make again.  The others: large objects one
        // after another with every thread, the rest one object per thread.
        const size_t no = s->objs.size();
        s->trees.resize(1 + no);
        std::vector<size_t> big, small;
        for (size_t oi = 0; oi < no; oi++) {
            if (can_reuse && !s->objs[oi].moving) {
                s->trees[1 + oi] = reuse->trees[1 + oi];
                s->trees[1 + oi].build_seconds = 0;  // not built here
                continue;
            }
            (s->objs[oi].idx.size() / 3 >= (1u << 18) ? big : small).push_back(oi);
        }
        auto build_obj = [&](size_t oi, int threads) {
            const HostObject& o = s->objs[oi];
            s->trees[1 + oi] = build_tree(o.verts.data(), (uint32_t)(o.verts.size() / 3), o.idx.data(),
                                          (uint32_t)(o.idx.size() / 3), threads,
                                          can_reuse ? &reuse->trees[1 + oi] : nullptr, gpu_refit);
        };
        for (size_t oi : big) build_obj(oi, nth);
        if (!small.empty()) {
            std::atomic<size_t> next{0};
            std::exception_ptr err;
            std::mutex err_mu;
            auto worker = [&]() {
                for (size_t k; (k = next.fetch_add(1)) < small.size();) {
                    try {
                        build_obj(small[k], 1);
                    } catch (...) {
                        std::lock_guard<std::mutex> lk(err_mu);
                        if (!err) err = std::current_exception();
                    }
                }
            };
            const int nw = (int)std::min<size_t>((size_t)nth, small.size());
            std::vector<std::thread> pool;
            for (int w = 1; w < nw; w++) pool.emplace_back(worker);
            worker();
            for (std::thread& t : pool) t.join();
            if (err) std::rethrow_exception(err);
        }
        if (s->trees[0].hdr.depth > 64) return fail(ctx, WR_EINVAL, "wr_scene_build: tree deeper than 64 levels");
        uint32_t max_obj_depth = 0;
        for (size_t i = 1; i < s->trees.size(); i++) {
            if (s->trees[i].hdr.depth > 64) return fail(ctx, WR_EINVAL, "wr_scene_build: object tree deeper than 64 levels");
            max_obj_depth = std::max(max_obj_depth, s->trees[i].hdr.depth);
        }
        // concatenate into device arrays: the bases of every tree first, then
        // each tree's part filled in place, its records in chunks over threads
        const size_t ntree = s->trees.size();
        std::vector<wr_tree_hdr> hdrs(ntree);
        std::vector<uint64_t> chain_base(ntree + 1, 0), top_base(ntree + 1, 0);
        uint64_t n_nodes = 0, n_wide = 0, n_recs = 0;
        for (size_t ti = 0; ti < ntree; ti++) {
            const FlatTree& F = s->trees[ti];
            wr_tree_hdr& h = hdrs[ti];
            h = F.hdr;
            h.node_base = (uint32_t)n_nodes;
            h.wide_base = (uint32_t)n_wide;
            h.rec_base = (uint32_t)n_recs;
            n_nodes += F.nodes.size();
            n_wide += F.hdr.n_wide;
            n_recs += F.tris.size();
            chain_base[ti + 1] = chain_base[ti] + F.chains.size();
            top_base[ti + 1] = top_base[ti] + F.top_boxes.size() / 6;
        }
        if (std::max({n_nodes, n_wide, n_recs, chain_base[ntree], top_base[ntree]}) > 0xFFFFFFFFull)
            return fail(ctx, WR_EINVAL, "wr_scene_build: scene too large");
        bool any_normals = false;
        for (const HostObject& o : s->objs) any_normals |= !o.normals.empty();
        // host arrays, written in full below: a rebuild's in `reuse`'s pinned
        // staging, else fresh (uninitialised, first touched by the fill threads)
        std::vector<std::unique_ptr<char[]>> own;
        RebuildCache* stg = reuse ? reuse->rebuild.get() : nullptr;
        std::unique_lock<std::mutex> rebuild_lock;
        if (stg) rebuild_lock = std::unique_lock<std::mutex>(stg->mu);
        auto host_array = [&](auto* type, size_t slot, size_t n) {
            using T = std::remove_pointer_t<decltype(type)>;
            const size_t b = std::max<size_t>(n, 1) * sizeof(T);
            void* p = stg ? stg->get(slot, b) : nullptr;
            if (!p) {  // no staging, or no pinned memory for it
                own.emplace_back(new char[b]);
                p = own.back().get();
            }
            return static_cast<T*>(p);
        };
        const size_t n_chains = chain_base[ntree], n_top = 6 * (top_base[ntree] + 5);  // five boxes' room past every record
        const size_t n_nrm = any_normals ? 9 * n_recs : 0;
        wr_node* nodes = host_array((wr_node*)nullptr, 0, n_nodes);
        wr_wnode* wnodes = gpu_refit ? nullptr : host_array((wr_wnode*)nullptr, 1, n_wide);
        wr_tri* tris = host_array((wr_tri*)nullptr, 2, n_recs);
        uint32_t* leaf_cnt = host_array((uint32_t*)nullptr, 3, n_recs);
        uint32_t* rec_obj = host_array((uint32_t*)nullptr, 4, n_recs);
        wr_chain* chains = host_array((wr_chain*)nullptr, 5, n_chains);
        float* top_boxes = host_array((float*)nullptr, 6, n_top);
        float* rec_nrm = host_array((float*)nullptr, 7, n_nrm);
        std::fill(top_boxes + 6 * top_base[ntree], top_boxes + n_top, 0.f);
        {
            struct Job {
                uint32_t ti, r0, r1;
            };
            constexpr uint32_t kJobRecs = 1u << 16;
            std::vector<Job> jobs;
            for (uint32_t ti = 0; ti < ntree; ti++) {
                const uint32_t nr = (uint32_t)s->trees[ti].tris.size();
                for (uint32_t r0 = 0; r0 == 0 || r0 < nr; r0 += kJobRecs) jobs.push_back({ti, r0, std::min(nr, r0 + kJobRecs)});
            }
            auto run_job = [&](const Job& j) {
                const FlatTree& F = s->trees[j.ti];
                const wr_tree_hdr& h = hdrs[j.ti];
                if (j.r0 == 0) {  // the tree's nodes, wide nodes, chains and top records
                    std::copy(F.nodes.begin(), F.nodes.end(), nodes + h.node_base);
                    if (wnodes) std::copy(F.wnodes.begin(), F.wnodes.end(), wnodes + h.wide_base);
                    std::copy(F.top_boxes.begin(), F.top_boxes.end(), top_boxes + 6 * top_base[j.ti]);
                    for (size_t c = 0; c < F.chains.size(); c++) {
                        wr_chain ch = F.chains[c];
                        if (ch.next != WR_NONE) ch.top += (uint32_t)top_base[j.ti];
                        chains[chain_base[j.ti] + c] = ch;
                    }
                }
                for (uint32_t r = j.r0; r < j.r1; r++) {
                    const size_t g = (size_t)h.rec_base + r;
                    uint32_t oi, local;
                    if (j.ti == 0) {
                        const uint32_t p = F.rec_prim[r];
                        oi = (uint32_t)(std::upper_bound(s->prim_base.begin(), s->prim_base.end(), p) -
                                        s->prim_base.begin() - 1);
                        local = p - s->prim_base[oi];
                    } else {
                        oi = j.ti - 1;
                        local = F.rec_prim[r];
                    }
                    wr_tri T = F.tris[r];
                    // the leaf's chain record: tie key and verification
                    T.leaf = F.rec_chain[r] == WR_NONE ? WR_NONE : F.rec_chain[r] + (uint32_t)chain_base[j.ti];
                    T.prim = F.rec_prim[r];
                    T.obj = oi;
                    T.pad = 0;
                    tris[g] = T;
                    leaf_cnt[g] = F.leaf_cnt[r];
                    rec_obj[g] = oi;
                    if (any_normals) {
                        const HostObject& o = s->objs[oi];
                        for (int k = 0; k < 9; k++)
                            rec_nrm[9 * g + k] = o.normals.empty() ? 0.f : o.normals[9 * (size_t)local + k];
                    }
                }
            };
            std::atomic<size_t> next{0};
            auto worker = [&]() {
                for (size_t k; (k = next.fetch_add(1)) < jobs.size();) run_job(jobs[k]);
            };
            const int nw = (int)std::min<size_t>((size_t)nth, jobs.size());
            std::vector<std::thread> pool;
            for (int w = 1; w < nw; w++) pool.emplace_back(worker);
            worker();
            for (std::thread& t : pool) t.join();
        }
        std::vector<wr_object_gpu> objs;
        std::vector<uint32_t> has_n;
        for (size_t oi = 0; oi < s->objs.size(); oi++) {
            const HostObject& o = s->objs[oi];
            wr_object_gpu g{};
            const wr_tree_hdr& h = s->trees[oi + 1].hdr;
            for (int a = 0; a < 3; a++) {
                g.aabb_min[a] = h.root_min[a];
                g.aabb_max[a] = h.root_max[a];
                g.color[a] = o.mat.color[a];
            }
            if (h.root_ref == WR_NONE) {  // no triangles: an empty box nothing is inside
                for (int a = 0; a < 3; a++) {
                    g.aabb_min[a] = 1.f;
                    g.aabb_max[a] = -1.f;
                }
            }
            g.power = o.mat.power;
            g.ior = o.mat.type == WR_REFRACTIVE ? o.mat.ior : 1.f;
            float pm = o.mat.color[0];
            if (o.mat.color[1] > pm) pm = o.mat.color[1];
            if (o.mat.color[2] > pm) pm = o.mat.color[2];
            g.pmax = pm;
            g.mat = o.mat.type;
            g.has_normals = o.normals.empty() ? 0u : 1u;
            g.tree = (uint32_t)(oi + 1);
            g.prim_base = s->prim_base[oi];
            objs.push_back(g);
            has_n.push_back(g.has_normals);
        }
        // object-candidate grid for findInsideObject (see SceneDev)
        float gmin[3] = {INFINITY, INFINITY, INFINITY}, gmax[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (size_t oi = 0; oi < objs.size(); oi++) {
            if (s->trees[oi + 1].hdr.root_ref == WR_NONE) continue;
            for (int a = 0; a < 3; a++) {
                gmin[a] = std::min(gmin[a], objs[oi].aabb_min[a]);
                gmax[a] = std::max(gmax[a], objs[oi].aabb_max[a]);
            }
        }
        uint32_t gdim[3] = {1, 1, 1};
        float ginv[3] = {0.f, 0.f, 0.f};
        const bool any_obj = gmin[0] <= gmax[0];
        if (any_obj) {
            const uint32_t want = (uint32_t)std::min(64.0, std::max(1.0, std::round(2.0 * std::cbrt((double)objs.size()))));
            for (int a = 0; a < 3; a++) {
                const float ext = gmax[a] - gmin[a];
                if (ext > 0.f) {
                    gdim[a] = want;
                    ginv[a] = (float)want / ext;
                }
            }
        }
        const uint32_t ncells = gdim[0] * gdim[1] * gdim[2];
        std::vector<std::vector<uint32_t>> cell_lists(ncells);
        for (int oi = (int)objs.size() - 1; oi >= 0 && any_obj; oi--) {
            if (s->trees[oi + 1].hdr.root_ref == WR_NONE) continue;
            uint32_t lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                lo[a] = wr_grid_cell(objs[oi].aabb_min[a], gmin[a], ginv[a], gdim[a]);
                hi[a] = wr_grid_cell(objs[oi].aabb_max[a], gmin[a], ginv[a], gdim[a]);
            }
            for (uint32_t z = lo[2]; z <= hi[2]; z++)
                for (uint32_t y = lo[1]; y <= hi[1]; y++)
                    for (uint32_t x = lo[0]; x <= hi[0]; x++)
                        cell_lists[(z * gdim[1] + y) * gdim[0] + x].push_back((uint32_t)oi);
        }
        std::vector<uint32_t> cell_start(ncells + 1, 0), cell_objs;
        for (uint32_t c = 0; c < ncells; c++) {
            cell_start[c] = (uint32_t)cell_objs.size();
            cell_objs.insert(cell_objs.end(), cell_lists[c].begin(), cell_lists[c].end());
        }
        cell_start[ncells] = (uint32_t)cell_objs.size();

        // light sampler (Scene::prepareForLightSampling, scene.cpp:5-14;
        // SurfaceSampler::prepareForSampling, surfaceSampler.cpp:7-15;
        // AbstractObject::prepareForSurfaceSampling, abstractObj.cpp:9-16): running
        // float sums of Mesh::getTriangleArea in object and triangle order
        std::vector<wr_light> lights;
        std::vector<float> light_cum, light_tri_cum;
        std::vector<uint32_t> light_of(s->objs.size(), WR_NONE);
        float light_total = 0.f;
        for (size_t oi = 0; oi < s->objs.size(); oi++) {
            const HostObject& o = s->objs[oi];
            if (o.mat.type != WR_EMISSIVE) continue;
            wr_light L{};
            L.obj = (uint32_t)oi;
            L.tri_base = (uint32_t)light_tri_cum.size();
            L.n_tris = (uint32_t)(o.idx.size() / 3);
            float area = 0.f;
            for (uint32_t t = 0; t < L.n_tris; t++) {
                area += triangle_area(o, t);
                light_tri_cum.push_back(area);
            }
            L.area = area;
            light_total += area;
            light_cum.push_back(light_total);
            light_of[oi] = (uint32_t)lights.size();
            lights.push_back(L);
        }
        for (wr_light& L : lights) {
            const float pick = L.area / light_total;  // pickProb (surfaceSampler.cpp:13-14)
            L.oprob = 1.f / L.area * pick;
        }
        std::vector<uint32_t> light_tri_rec(light_tri_cum.size(), WR_NONE);
        {
            const FlatTree& G = s->trees[0];  // its records come first in the record arrays
            for (size_t r = 0; r < G.rec_prim.size(); r++) {
                const uint32_t oi = rec_obj[r];
                if (light_of[oi] == WR_NONE) continue;
                light_tri_rec[lights[light_of[oi]].tri_base + (G.rec_prim[r] - s->prim_base[oi])] = (uint32_t)r;
            }
        }
        for (uint32_t r : light_tri_rec)
            if (r == WR_NONE) return fail(ctx, WR_EINVAL, "wr_scene_build: an emissive triangle has no tree record");

        uint64_t bytes = 0;
        uint32_t *d_cs, *d_co;
        wr_node* d_nodes;
        wr_wnode* d_wnodes;
        wr_tri* d_tris;
        uint32_t *d_lc, *d_hn;
        wr_chain* d_ch;
        float* d_rn;
        wr_tree_hdr* d_h;
        wr_object_gpu* d_o;
        int rc;
        if ((rc = upload(ctx, nodes, n_nodes, &d_nodes, bytes))) return rc;
        s->dbufs.push_back(d_nodes);
        if ((rc = upload(ctx, tris, n_recs, &d_tris, bytes))) return rc;
        s->dbufs.push_back(d_tris);
        if (!gpu_refit) {
            if ((rc = upload(ctx, wnodes, n_wide, &d_wnodes, bytes))) return rc;
            s->dbufs.push_back(d_wnodes);
        } else {
            // the built scene's nodes (the same count per tree), then the
            // rebuilt trees refit in place; synchronised so that the refit's
            // scratch is free when the rebuild lock is released
            uint64_t built_wide = 0;
            for (const wr_tree_hdr& bh : reuse->hdrs) built_wide += bh.n_wide;
            if (built_wide != n_wide) return fail(ctx, WR_EINVAL, "wr_scene_build: fast-pass hierarchy size changed");
            hipError_t e = hipMalloc((void**)&d_wnodes, std::max<uint64_t>(n_wide, 1) * sizeof(wr_wnode));
            if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc(scene)");
            s->dbufs.push_back(d_wnodes);
            bytes += std::max<uint64_t>(n_wide, 1) * sizeof(wr_wnode);
            if (n_wide)
                e = hipMemcpyAsync(d_wnodes, reuse->dev.wnodes, n_wide * sizeof(wr_wnode), hipMemcpyDeviceToDevice,
                                   ctx->stream);
            for (size_t ti = 0; ti < ntree && e == hipSuccess; ti++)
                if (ti == 0 || s->objs[ti - 1].moving)
                    e = refit_on_device(*stg, reuse, ti, d_wnodes, hdrs[ti], d_tris, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) return hip_fail(ctx, e, "wr_scene_build: device refit");
        }
        if ((rc = upload(ctx, leaf_cnt, n_recs, &d_lc, bytes))) return rc;
        s->dbufs.push_back(d_lc);
        if ((rc = upload(ctx, chains, n_chains, &d_ch, bytes))) return rc;
        s->dbufs.push_back(d_ch);
        float* d_tb;
        if ((rc = upload(ctx, top_boxes, n_top, &d_tb, bytes))) return rc;
        s->dbufs.push_back(d_tb);
        if ((rc = upload(ctx, rec_nrm, n_nrm, &d_rn, bytes))) return rc;
        s->dbufs.push_back(d_rn);
        if ((rc = upload(ctx, hdrs, &d_h, bytes))) return rc;
        s->dbufs.push_back(d_h);
        if ((rc = upload(ctx, objs, &d_o, bytes))) return rc;
        s->dbufs.push_back(d_o);
        if ((rc = upload(ctx, has_n, &d_hn, bytes))) return rc;
        s->dbufs.push_back(d_hn);
        if ((rc = upload(ctx, cell_start, &d_cs, bytes))) return rc;
        s->dbufs.push_back(d_cs);
        if ((rc = upload(ctx, cell_objs, &d_co, bytes))) return rc;
        s->dbufs.push_back(d_co);
        wr_light* d_li;
        float *d_lcum, *d_ltcum;
        uint32_t* d_ltrec;
        if ((rc = upload(ctx, lights, &d_li, bytes))) return rc;
        s->dbufs.push_back(d_li);
        if ((rc = upload(ctx, light_cum, &d_lcum, bytes))) return rc;
        s->dbufs.push_back(d_lcum);
        if ((rc = upload(ctx, light_tri_cum, &d_ltcum, bytes))) return rc;
        s->dbufs.push_back(d_ltcum);
        if ((rc = upload(ctx, light_tri_rec, &d_ltrec, bytes))) return rc;
        s->dbufs.push_back(d_ltrec);
        SceneDev& D = s->dev;
        D.nodes = d_nodes;
        D.wnodes = d_wnodes;
        D.tris = d_tris;
        D.leaf_cnt = d_lc;
        D.chains = d_ch;
        D.top_boxes = d_tb;
        D.rec_nrm = d_rn;
        D.hdrs = d_h;
        D.objs = d_o;
        D.obj_has_normals = d_hn;
        D.n_objs = (uint32_t)s->objs.size();
        D.max_obj_depth = max_obj_depth;
        D.cell_start = d_cs;
        D.cell_objs = d_co;
        D.lights = d_li;
        D.light_cum = d_lcum;
        D.light_tri_cum = d_ltcum;
        D.light_tri_rec = d_ltrec;
        D.n_lights = (uint32_t)lights.size();
        D.light_total = light_total;
        for (int a = 0; a < 3; a++) {
            D.grid_min[a] = gmin[a];
            D.grid_max[a] = gmax[a];
            D.grid_inv[a] = ginv[a];
            D.grid_dim[a] = gdim[a];
        }
        s->hdrs = hdrs;
        s->host_objs = objs;
        s->device_bytes = bytes;
        s->max_obj_depth = max_obj_depth;
        s->build_seconds = 0;
        for (auto& F : s->trees) s->build_seconds += F.build_seconds;
        s->built = true;
        return WR_OK;
    } catch (const std::exception& ex) {
        scene_free_device(s);
        return fail(ctx, WR_EINVAL, std::string("wr_scene_build: ") + ex.what());
    }
}

}  // namespace wr

extern "C" {

int wr_scene_get_info(const wr_scene* s, wr_scene_info* info) {
    if (!s || !info) return fail(nullptr, WR_EINVAL, "wr_scene_get_info: bad argument");
    if (!s->built) return fail(s->ctx, WR_ESTATE, "wr_scene_get_info: scene not built");
    info->n_objects = (uint32_t)s->objs.size();
    info->n_triangles = s->prim_base.back();
    info->n_nodes = s->trees[0].hdr.n_nodes;
    info->n_leaves = s->trees[0].n_leaves;
    info->depth = s->trees[0].hdr.depth;
    info->max_obj_depth = s->max_obj_depth;
    info->device_bytes = s->device_bytes;
    info->build_seconds = s->build_seconds;
    return WR_OK;
}

int wr_scene_prim_source(const wr_scene* s, uint32_t prim, uint32_t* obj, uint32_t* tri) {
    if (!s || !obj || !tri) return fail(nullptr, WR_EINVAL, "wr_scene_prim_source: bad argument");
    if (s->prim_base.empty() || prim >= s->prim_base.back())
        return fail(s->ctx, WR_EINVAL, "wr_scene_prim_source: prim out of range");
    const uint32_t oi =
        (uint32_t)(std::upper_bound(s->prim_base.begin(), s->prim_base.end(), prim) - s->prim_base.begin() - 1);
    *obj = oi;
    *tri = prim - s->prim_base[oi];
    return WR_OK;
}

int wr_scene_motion_pass(const wr_scene* s, uint64_t seed, uint32_t pass, wr_scene** out) {
    if (!s || !out) return fail(nullptr, WR_EINVAL, "wr_scene_motion_pass: bad argument");
    *out = nullptr;
    if (!s->built) return fail(s->ctx, WR_ESTATE, "wr_scene_motion_pass: scene not built");
    (void)hipSetDevice(s->ctx->device);
    std::unique_ptr<wr_scene> t(new wr_scene);
    const int rc = scene_pass(s, seed, pass, t.get());
    if (rc != WR_OK) {
        scene_free_device(t.get());
        return rc;
    }
    for (HostObject& o : t->objs) o.moving = false;  // a snapshot: placed once
    *out = t.release();
    return WR_OK;
}

int wr_scene_export_wide(const wr_scene* s, uint32_t obj, uint32_t* n_wide, uint32_t* wide_root, void* nodes) {
    if (!s || !n_wide || !wide_root) return fail(nullptr, WR_EINVAL, "wr_scene_export_wide: bad argument");
    if (!s->built) return fail(s->ctx, WR_ESTATE, "wr_scene_export_wide: scene not built");
    const size_t ti = obj == WR_NONE ? 0 : (size_t)obj + 1;
    if (ti >= s->hdrs.size()) return fail(s->ctx, WR_EINVAL, "wr_scene_export_wide: no such object");
    const wr_tree_hdr& h = s->hdrs[ti];
    *n_wide = h.n_wide;
    *wide_root = h.wide_root;
    if (nodes && h.n_wide) {
        (void)hipSetDevice(s->ctx->device);
        hipError_t e = hipStreamSynchronize(s->ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpy(nodes, s->dev.wnodes + h.wide_base, (size_t)h.n_wide * sizeof(wr_wnode), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_scene_export_wide");
    }
    return WR_OK;
}

int wr_scene_export_tree(const wr_scene* s, uint32_t obj, uint32_t* n_nodes, uint32_t* n_leaf_tris, float* box,
                         uint32_t* right, uint32_t* leaf_first, uint32_t* leaf_count, uint32_t* tri_list) {
    if (!s || !n_nodes || !n_leaf_tris) return fail(nullptr, WR_EINVAL, "wr_scene_export_tree: bad argument");
    if (!s->built) return fail(s->ctx, WR_ESTATE, "wr_scene_export_tree: scene not built");
    const size_t ti = obj == WR_NONE ? 0 : (size_t)obj + 1;
    if (ti >= s->trees.size()) return fail(s->ctx, WR_EINVAL, "wr_scene_export_tree: no such object");
    TreeExport X = export_tree(s->trees[ti]);
    *n_nodes = (uint32_t)X.right.size();
    *n_leaf_tris = (uint32_t)X.tri_list.size();
    if (box) std::memcpy(box, X.box.data(), X.box.size() * sizeof(float));
    if (right) std::memcpy(right, X.right.data(), X.right.size() * sizeof(uint32_t));
    if (leaf_first) std::memcpy(leaf_first, X.leaf_first.data(), X.leaf_first.size() * sizeof(uint32_t));
    if (leaf_count) std::memcpy(leaf_count, X.leaf_count.data(), X.leaf_count.size() * sizeof(uint32_t));
    if (tri_list) std::memcpy(tri_list, X.tri_list.data(), X.tri_list.size() * sizeof(uint32_t));
    return WR_OK;
}

struct wr_tree {
    wr::FlatTree t;
};

int wr_tree_build(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads, wr_tree** out) {
    if (!out || (!verts && nv) || (!idx && nt)) return fail(nullptr, WR_EINVAL, "wr_tree_build: bad argument");
    *out = nullptr;
    try {
        std::unique_ptr<wr_tree> t(new wr_tree);
        t->t = build_tree(verts, nv, idx, nt, nthreads);
        *out = t.release();
        return WR_OK;
    } catch (const std::exception& ex) {
        return fail(nullptr, WR_EINVAL, std::string("wr_tree_build: ") + ex.what());
    }
}

int wr_tree_build_refit(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads,
                        const wr_tree* prev, wr_tree** out) {
    if (!out || !prev || (!verts && nv) || (!idx && nt)) return fail(nullptr, WR_EINVAL, "wr_tree_build_refit: bad argument");
    *out = nullptr;
    if (prev->t.hdr.n_recs != nt) return fail(nullptr, WR_EINVAL, "wr_tree_build_refit: prev has another triangle count");
    try {
        std::unique_ptr<wr_tree> t(new wr_tree);
        t->t = build_tree(verts, nv, idx, nt, nthreads, &prev->t);
        *out = t.release();
        return WR_OK;
    } catch (const std::exception& ex) {
        return fail(nullptr, WR_EINVAL, std::string("wr_tree_build_refit: ") + ex.what());
    }
}

int wr_tree_export(const wr_tree* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* depth, float* box,
                   uint32_t* right, uint32_t* leaf_first, uint32_t* leaf_count, uint32_t* tri_list) {
    if (!tree || !n_nodes || !n_leaf_tris) return fail(nullptr, WR_EINVAL, "wr_tree_export: bad argument");
    TreeExport X = export_tree(tree->t);
    *n_nodes = (uint32_t)X.right.size();
    *n_leaf_tris = (uint32_t)X.tri_list.size();
    if (depth) *depth = tree->t.hdr.depth;
    if (box) std::memcpy(box, X.box.data(), X.box.size() * sizeof(float));
    if (right) std::memcpy(right, X.right.data(), X.right.size() * sizeof(uint32_t));
    if (leaf_first) std::memcpy(leaf_first, X.leaf_first.data(), X.leaf_first.size() * sizeof(uint32_t));
    if (leaf_count) std::memcpy(leaf_count, X.leaf_count.data(), X.leaf_count.size() * sizeof(uint32_t));
    if (tri_list) std::memcpy(tri_list, X.tri_list.data(), X.tri_list.size() * sizeof(uint32_t));
    return WR_OK;
}

void wr_tree_free(wr_tree* tree) { delete tree; }

int wr_tree_wide(const wr_tree* tree, uint32_t* n_wide, uint32_t* n_records, uint32_t* wide_root, void* wide_nodes,
                 uint32_t* record_first, uint32_t* record_tri) {
    if (!tree || !n_wide || !n_records || !wide_root) return fail(nullptr, WR_EINVAL, "wr_tree_wide: bad argument");
    const FlatTree& F = tree->t;
    *n_wide = (uint32_t)F.wnodes.size();
    *n_records = (uint32_t)F.tris.size();
    *wide_root = F.hdr.wide_root;
    if (wide_nodes) std::memcpy(wide_nodes, F.wnodes.data(), F.wnodes.size() * sizeof(wr_wnode));
    if (record_first) std::memcpy(record_first, F.rec_first.data(), F.rec_first.size() * sizeof(uint32_t));
    if (record_tri) std::memcpy(record_tri, F.rec_prim.data(), F.rec_prim.size() * sizeof(uint32_t));
    return WR_OK;
}

int wr_closest_hit(wr_ctx* ctx, const wr_scene* s, const wr_ray* d_rays, uint64_t n, wr_hit* d_hits,
                   uint64_t* d_counters) {
    if (!ctx || !s || (n && (!d_rays || !d_hits))) return fail(ctx, WR_EINVAL, "wr_closest_hit: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_closest_hit: scene not built");
    hipError_t e = launch_closest_hit(s->dev, s->hdrs[0], (const float*)d_rays, n, d_hits,
                                      (unsigned long long*)d_counters, ctx->d_work, s->ch_tune, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "k_closest_hit launch");
    return WR_OK;
}

int wr_hit_sources(wr_ctx* ctx, const wr_scene* s, const wr_hit* d_hits, uint64_t n, uint32_t* d_obj,
                   uint32_t* d_tri) {
    if (!ctx || !s || (n && (!d_hits || !d_obj || !d_tri))) return fail(ctx, WR_EINVAL, "wr_hit_sources: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_hit_sources: scene not built");
    hipError_t e = launch_hit_sources(s->dev, s->prim_base.back(), d_hits, n, d_obj, d_tri, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "k_hit_sources launch");
    return WR_OK;
}

int wr_autotune_closest_hit(wr_ctx* ctx, wr_scene* s, const wr_ray* d_rays, uint64_t n,
                            wr_closest_hit_tuning* best) {
    if (!ctx || !s || !d_rays || n == 0) return fail(ctx, WR_EINVAL, "wr_autotune_closest_hit: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_autotune_closest_hit: scene not built");
    (void)hipSetDevice(ctx->device);
    void* d_hits = nullptr;
    hipError_t e = hipMalloc(&d_hits, 16 * (size_t)n);
    if (e != hipSuccess) return fail(ctx, WR_ENOMEM, std::string("wr_autotune_closest_hit: ") + hipGetErrorString(e));
    // time of one call under t: best of three after one warm-up call
    auto time_of = [&](const ChTune& t, float& ms) -> hipError_t {
        ms = 1e30f;
        for (int rep = 0; rep < 4; rep++) {
            hipError_t r = hipEventRecord(ctx->ev0, ctx->stream);
            if (r == hipSuccess)
                r = launch_closest_hit(s->dev, s->hdrs[0], (const float*)d_rays, n, d_hits, nullptr, ctx->d_work, t,
                                       ctx->stream);
            if (r == hipSuccess) r = hipEventRecord(ctx->ev1, ctx->stream);
            if (r == hipSuccess) r = hipEventSynchronize(ctx->ev1);
            float m = 0.f;
            if (r == hipSuccess) r = hipEventElapsedTime(&m, ctx->ev0, ctx->ev1);
            if (r != hipSuccess) return r;
            if (rep > 0) ms = std::min(ms, m);
        }
        return hipSuccess;
    };
    // coordinate descent over the knobs, starting from the current choice;
    // a candidate must be 1% faster to replace it (timing noise stays put)
    ChTune cur = s->ch_tune;
    float cur_ms = 0.f;
    e = time_of(cur, cur_ms);
    const uint32_t nodes[] = {16, 24, 32, 40, 48, 65};
    const uint32_t refills[] = {4, 8, 16, 24, 32};
    const uint32_t blocks[] = {0, 7, 6, 5};
    const uint32_t levels[] = {4, 8, 16};
    const uint32_t batches[] = {128, 256, 512, 1024};
    for (int pass = 0; pass < 2 && e == hipSuccess; pass++) {
        for (uint32_t v : nodes) {
            ChTune t = cur;
            t.node_phase_min = v;
            float m;
            if ((e = time_of(t, m)) != hipSuccess) break;
            if (m < 0.99f * cur_ms) cur = t, cur_ms = m;
        }
        for (uint32_t v : refills) {
            if (e != hipSuccess) break;
            ChTune t = cur;
            t.refill_min = v;
            float m;
            if ((e = time_of(t, m)) != hipSuccess) break;
            if (m < 0.99f * cur_ms) cur = t, cur_ms = m;
        }
        for (uint32_t v : blocks) {
            if (e != hipSuccess) break;
            ChTune t = cur;
            t.blocks_per_cu = v;
            float m;
            if ((e = time_of(t, m)) != hipSuccess) break;
            if (m < 0.99f * cur_ms) cur = t, cur_ms = m;
        }
        for (uint32_t v : levels) {
            if (e != hipSuccess) break;
            ChTune t = cur;
            t.lds_levels = v;
            float m;
            if ((e = time_of(t, m)) != hipSuccess) break;
            if (m < 0.99f * cur_ms) cur = t, cur_ms = m;
        }
        for (uint32_t v : batches) {
            if (e != hipSuccess) break;
            ChTune t = cur;
            t.batch = v;
            float m;
            if ((e = time_of(t, m)) != hipSuccess) break;
            if (m < 0.99f * cur_ms) cur = t, cur_ms = m;
        }
    }
    (void)hipFree(d_hits);
    if (e != hipSuccess) return hip_fail(ctx, e, "wr_autotune_closest_hit");
    s->ch_tune = cur;
    if (best)
        *best = wr_closest_hit_tuning{cur.refill_min, cur.node_phase_min, cur.blocks_per_cu, cur.lds_levels, cur.batch,
                                      cur_ms};
    return WR_OK;
}

int wr_scene_set_closest_hit_tuning(wr_scene* s, const wr_closest_hit_tuning* t) {
    if (!s || !t) return fail(nullptr, WR_EINVAL, "wr_scene_set_closest_hit_tuning: bad argument");
    const uint32_t lv = t->lds_levels ? t->lds_levels : 8u;
    if (lv != 4 && lv != 8 && lv != 16)
        return fail(s->ctx, WR_EINVAL, "wr_scene_set_closest_hit_tuning: lds_levels must be 4, 8 or 16");
    s->ch_tune.lds_levels = lv;
    s->ch_tune.refill_min = t->refill_min;
    s->ch_tune.node_phase_min = t->node_phase_min;
    s->ch_tune.blocks_per_cu = t->blocks_per_cu;
    s->ch_tune.batch = t->batch ? t->batch : 256u;
    return WR_OK;
}

int wr_closest_hit_noself(wr_ctx* ctx, const wr_scene* s, const wr_path_ray* d_rays, uint64_t n, wr_hit* d_hits) {
    if (!ctx || !s || (n && (!d_rays || !d_hits))) return fail(ctx, WR_EINVAL, "wr_closest_hit_noself: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_closest_hit_noself: scene not built");
    hipError_t e = launch_closest_hit_path(s->dev, s->hdrs[0], (const uint32_t*)d_rays, n, d_hits, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "k_closest_hit_path launch");
    return WR_OK;
}

int wr_inside_object(wr_ctx* ctx, const wr_scene* s, const wr_path_ray* d_rays, uint64_t n, uint32_t* d_obj) {
    if (!ctx || !s || (n && (!d_rays || !d_obj))) return fail(ctx, WR_EINVAL, "wr_inside_object: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_inside_object: scene not built");
    hipError_t e = launch_inside_object(s->dev, s->max_obj_depth, (const uint32_t*)d_rays, n, d_obj, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "k_inside_object launch");
    return WR_OK;
}

}  // extern "C"
