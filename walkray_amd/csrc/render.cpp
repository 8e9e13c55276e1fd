// render.cpp -- extern "C" film and render entry points (include/walkray.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/walkray.h"
#include "wr_capi_internal.h"
#include "wr_pt.h"

struct wr_film {
    wr_ctx* ctx = nullptr;
    uint32_t W = 0, H = 0;
    float* sum = nullptr;
    float* comp = nullptr;
    float* sumsq = nullptr;
    uint32_t passes = 0;
    // set when a render failed part way: the sums no longer match `passes`, so
    // the film is refused until wr_film_clear or wr_film_write resets it
    bool broken = false;
};

namespace {

// nvVector semantics (nvVector.h:490-525): sequential dot from 0, sqrt, r > 0
// guard, true division.
void vnormalize(float v[3]) {
    float r2 = 0.f;
    r2 += v[0] * v[0];
    r2 += v[1] * v[1];
    r2 += v[2] * v[2];
    const float r = std::sqrt(r2);
    if (r > 0)
        for (int k = 0; k < 3; k++) v[k] = v[k] / r;
}
void vcross(const float a[3], const float b[3], float o[3]) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// Camera basis of camera.cpp:7-11.
wr::CamHost make_cam(const wr_camera& c) {
    wr::CamHost h;
    for (int k = 0; k < 3; k++) {
        h.pos[k] = c.pos[k];
        h.front[k] = c.focus[k] - c.pos[k];
    }
    vnormalize(h.front);
    vcross(h.front, c.up, h.right);
    vnormalize(h.right);
    vcross(h.right, h.front, h.top);
    h.dist = c.sight_dist;
    h.W = c.width;
    h.H = c.height;
    return h;
}

}  // namespace

using namespace wr;

extern "C" {

int wr_film_create(wr_ctx* ctx, uint32_t width, uint32_t height, wr_film** out) {
    if (!ctx || !out || width == 0 || height == 0) return fail(ctx, WR_EINVAL, "wr_film_create: bad argument");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    wr_film* f = new wr_film;
    f->ctx = ctx;
    f->W = width;
    f->H = height;
    const size_t b = sizeof(float) * 3 * (size_t)width * height;
    if (hipMalloc((void**)&f->sum, b) != hipSuccess || hipMalloc((void**)&f->comp, b) != hipSuccess ||
        hipMalloc((void**)&f->sumsq, b) != hipSuccess) {
        wr_film_destroy(f);
        return fail(ctx, WR_ENOMEM, "wr_film_create: out of device memory");
    }
    int rc = wr_film_clear(f);
    if (rc) {
        wr_film_destroy(f);
        return rc;
    }
    *out = f;
    return WR_OK;
}

void wr_film_destroy(wr_film* f) {
    if (!f) return;
    (void)hipStreamSynchronize(f->ctx->stream);
    if (f->sum) (void)hipFree(f->sum);
    if (f->comp) (void)hipFree(f->comp);
    if (f->sumsq) (void)hipFree(f->sumsq);
    delete f;
}

int wr_film_clear(wr_film* f) {
    if (!f) return fail(nullptr, WR_EINVAL, "wr_film_clear: film is NULL");
    const size_t b = sizeof(float) * 3 * (size_t)f->W * f->H;
    hipError_t e = hipMemsetAsync(f->sum, 0, b, f->ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(f->comp, 0, b, f->ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(f->sumsq, 0, b, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    if (e != hipSuccess) return hip_fail(f->ctx, e, "wr_film_clear");
    f->passes = 0;
    f->broken = false;
    return WR_OK;
}

static int film_broken(wr_film* f, const char* who) {
    return fail(f->ctx, WR_ESTATE,
                std::string(who) + ": a render into this film failed part way; clear it or write a checkpoint into it");
}

int wr_film_read(wr_film* f, float* mean_rgb, float* sum, float* comp, float* sumsq, uint32_t* passes) {
    if (!f) return fail(nullptr, WR_EINVAL, "wr_film_read: film is NULL");
    if (f->broken) return film_broken(f, "wr_film_read");
    const size_t n = 3 * (size_t)f->W * f->H;
    std::vector<float> tmp;
    float* s = sum;
    if (!s && mean_rgb) {
        tmp.resize(n);
        s = tmp.data();
    }
    hipError_t e = hipSuccess;
    if (s) e = hipMemcpyAsync(s, f->sum, n * sizeof(float), hipMemcpyDeviceToHost, f->ctx->stream);
    if (e == hipSuccess && comp) e = hipMemcpyAsync(comp, f->comp, n * sizeof(float), hipMemcpyDeviceToHost, f->ctx->stream);
    if (e == hipSuccess && sumsq) e = hipMemcpyAsync(sumsq, f->sumsq, n * sizeof(float), hipMemcpyDeviceToHost, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    if (e != hipSuccess) return hip_fail(f->ctx, e, "wr_film_read");
    if (mean_rgb) {
        const float inv = f->passes ? 1.f / (float)f->passes : 0.f;
        for (size_t i = 0; i < n; i++) mean_rgb[i] = f->passes ? s[i] * inv : 0.f;
    }
    if (passes) *passes = f->passes;
    return WR_OK;
}

int wr_film_write(wr_film* f, const float* sum, const float* comp, const float* sumsq, uint32_t passes) {
    if (!f || !sum || !comp || !sumsq) return fail(f ? f->ctx : nullptr, WR_EINVAL, "wr_film_write: bad argument");
    const size_t b = sizeof(float) * 3 * (size_t)f->W * f->H;
    hipError_t e = hipMemcpyAsync(f->sum, sum, b, hipMemcpyHostToDevice, f->ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(f->comp, comp, b, hipMemcpyHostToDevice, f->ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(f->sumsq, sumsq, b, hipMemcpyHostToDevice, f->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(f->ctx->stream);
    if (e != hipSuccess) return hip_fail(f->ctx, e, "wr_film_write");
    f->passes = passes;
    f->broken = false;
    return WR_OK;
}

int wr_render(wr_ctx* ctx, const wr_scene* s, const wr_camera* cam, const wr_render_params* prm, wr_film* film,
              wr_render_stats* stats) {
    if (!ctx || !s || !cam || !prm || !film) return fail(ctx, WR_EINVAL, "wr_render: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_render: scene not built");
    if (film->W != cam->width || film->H != cam->height)
        return fail(ctx, WR_EINVAL, "wr_render: film size differs from the camera resolution");
    if (prm->mode != WR_MEGAKERNEL && prm->mode != WR_WAVEFRONT) return fail(ctx, WR_EINVAL, "wr_render: unknown mode");
    if (prm->flags & ~(uint32_t)(WR_RENDER_NEE | WR_RENDER_DIRECT | WR_RENDER_STAGE_STATS))
        return fail(ctx, WR_EINVAL, "wr_render: unknown flags");
    const bool nee = (prm->flags & WR_RENDER_NEE) != 0;
    const int launch = (prm->flags & WR_RENDER_STAGE_STATS) ? kLaunchStageStats
                       : (prm->flags & WR_RENDER_DIRECT)   ? kLaunchDirect
                                                           : kLaunchGraph;
    if (launch != kLaunchGraph && prm->mode != WR_WAVEFRONT)
        return fail(ctx, WR_EINVAL, "wr_render: launch-mode flags are for the wavefront engine");
    StageStats stage;
    double build_ms = 0;  // host time of the per-pass rebuilds
    if ((uint64_t)cam->width * cam->height >= (1ull << 31)) return fail(ctx, WR_EINVAL, "wr_render: image too large");
    if (film->broken) return film_broken(film, "wr_render");
    (void)hipSetDevice(ctx->device);
    const CamHost ch = make_cam(*cam);
    FilmBufs fb{film->sum, film->comp, film->sumsq};
    bool moving = false;
    for (const HostObject& o : s->objs) moving = moving || o.moving;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    WaveStats ws;
    if (e == hipSuccess && moving) {
        // Scene::updateSceneForMotionBlur before every pass (pathTracer.cpp:18,
        // scene.cpp:67-80): each moving object placed at its time for the pass,
        // then the trees rebuilt, so every pass sees the reference's trees
        for (uint32_t k = 0; k < prm->spp && e == hipSuccess; k++) {
            const uint32_t pass = film->passes + k;
            wr_scene tmp;
            const auto b0 = std::chrono::steady_clock::now();
            const int rc = scene_pass(s, prm->seed, pass, &tmp);
            build_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - b0).count();
            if (rc != WR_OK) {
                scene_free_device(&tmp);
                film->passes += k;  // the passes before this one are in the film
                return rc;
            }
            if (prm->mode == WR_MEGAKERNEL)
                e = launch_pt_mega(tmp.dev, ch, prm->max_depth, prm->seed, pass, nee, fb, ctx->stream);
            else
                e = run_wavefront(tmp.dev, ch, prm->max_depth, prm->seed, pass, 1, nee, fb, ctx->engine,
                                  prm->wave_slots ? prm->wave_slots : ctx->tuned_slots,
                                  prm->wave_grid ? prm->wave_grid : ctx->tuned_grid, &ws, ctx->stream,
                                  ctx->tuned_lds, launch, &stage);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            scene_free_device(&tmp);
        }
    } else if (e == hipSuccess) {
        if (prm->mode == WR_MEGAKERNEL) {
            for (uint32_t k = 0; k < prm->spp && e == hipSuccess; k++)
                e = launch_pt_mega(s->dev, ch, prm->max_depth, prm->seed, film->passes + k, nee, fb, ctx->stream);
        } else {
            e = run_wavefront(s->dev, ch, prm->max_depth, prm->seed, film->passes, prm->spp, nee, fb, ctx->engine,
                              prm->wave_slots ? prm->wave_slots : ctx->tuned_slots,
                              prm->wave_grid ? prm->wave_grid : ctx->tuned_grid, &ws, ctx->stream,
                              ctx->tuned_lds, launch, &stage);
        }
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    if (e != hipSuccess) {
        film->broken = true;
        return hip_fail(ctx, e, "wr_render");
    }
    film->passes += prm->spp;
    if (launch == kLaunchStageStats) ctx->last_stage = stage;
    if (stats) {
        stats->paths = (uint64_t)prm->spp * cam->width * cam->height;
        stats->iterations = ws.iterations;
        stats->kernel_ms = ms;
        stats->build_ms = (float)build_ms;
    }
    return WR_OK;
}

int wr_render_stage_stats(wr_ctx* ctx, wr_stage_stats* out) {
    if (!ctx || !out) return fail(ctx, WR_EINVAL, "wr_render_stage_stats: bad argument");
    const StageStats& s = ctx->last_stage;
    for (int k = 0; k < kStages; k++) out->ms[k] = (float)s.ms[k];
    out->segments = s.traced;
    out->exits = s.exits;
    out->generated = s.generated;
    out->box_records = s.nodes;
    out->tri_records = s.tris;
    return WR_OK;
}

}  // extern "C"

extern "C" int wr_render_bdpt(wr_ctx* ctx, const wr_scene* s, const wr_camera* cam, const wr_render_params* prm,
                              wr_film* film, wr_render_stats* stats) {
    if (!ctx || !s || !cam || !prm || !film) return fail(ctx, WR_EINVAL, "wr_render_bdpt: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_render_bdpt: scene not built");
    if (film->W != cam->width || film->H != cam->height)
        return fail(ctx, WR_EINVAL, "wr_render_bdpt: film size differs from the camera resolution");
    if (prm->mode != 0 || prm->wave_grid || prm->flags)
        return fail(ctx, WR_EINVAL, "wr_render_bdpt: mode, wave_grid and flags must be 0");
    if (prm->max_depth > kBdptMaxDepth) return fail(ctx, WR_EINVAL, "wr_render_bdpt: max_depth above 32");
    for (const HostObject& o : s->objs)
        if (o.moving) return fail(ctx, WR_EINVAL, "wr_render_bdpt: moving objects are not supported");
    const uint64_t npix = (uint64_t)cam->width * cam->height;
    if (npix == 0 || npix >= (1ull << 31)) return fail(ctx, WR_EINVAL, "wr_render_bdpt: bad image size");
    if (film->broken) return film_broken(film, "wr_render_bdpt");
    (void)hipSetDevice(ctx->device);
    const CamHost ch = make_cam(*cam);
    const uint32_t maxv = prm->max_depth + 3;  // the longest sub-path: first ray, max_depth scatters, the end
    // passes per launch: as many as fill the device, within 4 GiB of sub-path records
    const uint64_t per_pass = bdpt_vertex_bytes() * 2 * maxv * npix;
    uint64_t batch = std::max<uint64_t>(1, (4ull << 30) / per_pass);
    batch = std::min<uint64_t>(std::min<uint64_t>(batch, prm->spp ? prm->spp : 1), std::max<uint64_t>(1, (1ull << 22) / npix));
    void* verts = nullptr;
    float* pix = nullptr;
    // the visibility queue: 32 entries per lane, at most 24M (1.6 GB); a full queue
    // makes the remaining connections trace in place
    BdptQueue q{};
    q.cap = prm->wave_slots ? prm->wave_slots : (uint32_t)std::min<uint64_t>(32ull * npix * batch, 24ull << 20);
    void* qbuf = nullptr;
    hipError_t e = hipMalloc(&verts, per_pass * batch);
    if (e == hipSuccess) e = hipMalloc((void**)&pix, sizeof(float) * 3 * npix * batch);
    if (e == hipSuccess) e = hipMalloc(&qbuf, kBdptQueueEntryBytes * (size_t)q.cap + 256);
    if (e == hipSuccess) e = hipMalloc((void**)&q.counts, sizeof(uint32_t) * 2 * npix * batch);
    if (e == hipSuccess) {
        char* b = (char*)qbuf;
        q.rays = (uint32_t*)b;
        q.hits = b + 32 * (size_t)q.cap;
        q.pay = (float*)(b + 48 * (size_t)q.cap);
        q.dst = (uint32_t*)(b + 64 * (size_t)q.cap);
        q.count = (uint32_t*)(b + 68 * (size_t)q.cap + 128);
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    const FilmBufs fb{film->sum, film->comp, film->sumsq};
    const bool launched = e == hipSuccess && prm->spp > 0;  // from here on a failure leaves the film part-written
    for (uint32_t k = 0; k < prm->spp && e == hipSuccess; k += (uint32_t)batch) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, prm->spp - k);
        e = launch_bdpt_passes(s->dev, s->hdrs[0], ch, prm->max_depth, prm->seed, film->passes + k, nb, maxv, verts,
                               pix, q, fb, ctx->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(verts);
    (void)hipFree(pix);
    (void)hipFree(qbuf);
    (void)hipFree(q.counts);
    if (e != hipSuccess) {
        film->broken = film->broken || launched;
        return hip_fail(ctx, e, "wr_render_bdpt");
    }
    film->passes += prm->spp;
    if (stats) {
        stats->paths = npix * prm->spp;
        stats->iterations = 0;
        stats->kernel_ms = ms;
        stats->build_ms = 0.f;
    }
    return WR_OK;
}

extern "C" float wr_vcm_radius(float base_radius, uint32_t pass) { return vcm_radius(base_radius, pass); }

extern "C" int wr_render_vcm(wr_ctx* ctx, const wr_scene* s, const wr_camera* cam, const wr_render_params* prm,
                             float base_radius, wr_film* film, wr_render_stats* stats) {
    if (!ctx || !s || !cam || !prm || !film) return fail(ctx, WR_EINVAL, "wr_render_vcm: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_render_vcm: scene not built");
    if (film->W != cam->width || film->H != cam->height)
        return fail(ctx, WR_EINVAL, "wr_render_vcm: film size differs from the camera resolution");
    if (prm->mode != 0 || prm->wave_grid || prm->flags)
        return fail(ctx, WR_EINVAL, "wr_render_vcm: mode, wave_grid and flags must be 0");
    if (prm->max_depth > kBdptMaxDepth) return fail(ctx, WR_EINVAL, "wr_render_vcm: max_depth above 32");
    if (!(base_radius > 0.f) || !std::isfinite(base_radius))
        return fail(ctx, WR_EINVAL, "wr_render_vcm: base_radius must be finite and > 0");
    for (const HostObject& o : s->objs)
        if (o.moving) return fail(ctx, WR_EINVAL, "wr_render_vcm: moving objects are not supported");
    const uint64_t npix = (uint64_t)cam->width * cam->height;
    const uint32_t maxv = prm->max_depth + 3;
    // light points of a pass: npix default ones plus at most maxv - 1 per light path
    if (npix == 0 || npix * maxv >= (1ull << 31)) return fail(ctx, WR_EINVAL, "wr_render_vcm: bad image size");
    if (film->broken) return film_broken(film, "wr_render_vcm");
    (void)hipSetDevice(ctx->device);
    const CamHost ch = make_cam(*cam);
    // passes per launch: as many as fill the device, within 4 GiB of sub-path records and 2^31 points
    const uint64_t per_pass = bdpt_vertex_bytes() * 2 * maxv * npix;
    uint64_t batch = std::max<uint64_t>(1, (4ull << 30) / per_pass);
    batch = std::min<uint64_t>(std::min<uint64_t>(batch, prm->spp ? prm->spp : 1), std::max<uint64_t>(1, (1ull << 22) / npix));
    batch = std::min<uint64_t>(batch, std::max<uint64_t>(1, ((1ull << 31) - 1) / (npix * maxv)));
    const uint64_t lanes = batch * npix;
    BdptQueue q{};
    q.cap = prm->wave_slots ? prm->wave_slots : (uint32_t)std::min<uint64_t>(32ull * lanes, 24ull << 20);
    VcmBufs vb{};
    vb.cap = lanes * maxv;
    vb.scan_bytes = vcm_scan_bytes((uint32_t)lanes);
    vb.mq_cap = (uint32_t)std::min<uint64_t>(64ull * lanes, 48ull << 20);  // beyond it a lane merges in place
    void *verts = nullptr, *qbuf = nullptr;
    float* pix = nullptr;
    hipError_t e = hipMalloc(&verts, per_pass * batch);
    if (e == hipSuccess) e = hipMalloc((void**)&pix, sizeof(float) * 3 * lanes);
    if (e == hipSuccess) e = hipMalloc((void**)&vb.mq, sizeof(uint32_t) * (2 * (size_t)vb.mq_cap + 1));
    if (e == hipSuccess) e = hipMalloc(&qbuf, kBdptQueueEntryBytes * (size_t)q.cap + 256);
    if (e == hipSuccess) e = hipMalloc((void**)&q.counts, sizeof(uint32_t) * 2 * lanes);
    if (e == hipSuccess) e = hipMalloc((void**)&vb.pc, sizeof(uint32_t) * (lanes + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.offs, sizeof(uint32_t) * (lanes + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.starts, sizeof(uint32_t) * (batch + 1));
    if (e == hipSuccess) e = hipMalloc(&vb.scan, std::max<size_t>(vb.scan_bytes, 4));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.pos, sizeof(float) * 3 * vb.cap);
    if (e == hipSuccess) e = hipMalloc((void**)&vb.src, sizeof(uint32_t) * 2 * vb.cap);
    if (e == hipSuccess) e = hipMalloc(&vb.grids, vcm_grid_bytes() * batch);
    if (e == hipSuccess) {
        char* b = (char*)qbuf;
        q.rays = (uint32_t*)b;
        q.hits = b + 32 * (size_t)q.cap;
        q.pay = (float*)(b + 48 * (size_t)q.cap);
        q.dst = (uint32_t*)(b + 64 * (size_t)q.cap);
        q.count = (uint32_t*)(b + 68 * (size_t)q.cap + 128);
    }
    auto release = [&]() {
        for (void* ptr : {verts, (void*)pix, qbuf, (void*)q.counts, (void*)vb.pc, (void*)vb.offs, (void*)vb.starts,
                          vb.scan, (void*)vb.pos, (void*)vb.src, vb.grids, (void*)vb.mq})
            (void)hipFree(ptr);
    };
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) {
        release();
        return hip_fail(ctx, e, "wr_render_vcm");
    }
    const FilmBufs fb{film->sum, film->comp, film->sumsq};
    for (uint32_t k = 0; k < prm->spp; k += (uint32_t)batch) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, prm->spp - k);
        const int rc = launch_vcm_passes(ctx, s->dev, s->hdrs[0], ch, prm->max_depth, prm->seed, film->passes + k, nb,
                                         base_radius, maxv, verts, pix, q, vb, fb);
        if (rc != WR_OK) {  // the film may hold part of this call's passes
            (void)hipStreamSynchronize(ctx->stream);
            release();
            film->broken = true;
            return rc;
        }
    }
    e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    release();
    if (e != hipSuccess) {
        film->broken = film->broken || prm->spp > 0;
        return hip_fail(ctx, e, "wr_render_vcm");
    }
    film->passes += prm->spp;
    if (stats) {
        stats->paths = npix * prm->spp;
        stats->iterations = 0;
        stats->kernel_ms = ms;
        stats->build_ms = 0.f;
    }
    return WR_OK;
}

extern "C" int wr_render_ppm(wr_ctx* ctx, const wr_scene* s, const wr_camera* cam, const wr_render_params* prm,
                             float base_radius, wr_film* film, wr_render_stats* stats) {
    if (!ctx || !s || !cam || !prm || !film) return fail(ctx, WR_EINVAL, "wr_render_ppm: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_render_ppm: scene not built");
    if (film->W != cam->width || film->H != cam->height)
        return fail(ctx, WR_EINVAL, "wr_render_ppm: film size differs from the camera resolution");
    if (prm->mode != 0 || prm->wave_grid || prm->flags || prm->wave_slots)
        return fail(ctx, WR_EINVAL, "wr_render_ppm: mode, wave_slots, wave_grid and flags must be 0");
    if (prm->max_depth > kBdptMaxDepth) return fail(ctx, WR_EINVAL, "wr_render_ppm: max_depth above 32");
    if (!(base_radius > 0.f) || !std::isfinite(base_radius))
        return fail(ctx, WR_EINVAL, "wr_render_ppm: base_radius must be finite and > 0");
    for (const HostObject& o : s->objs)
        if (o.moving) return fail(ctx, WR_EINVAL, "wr_render_ppm: moving objects are not supported");
    const uint64_t npix = (uint64_t)cam->width * cam->height;
    const uint32_t maxv = prm->max_depth + 3;
    if (npix == 0 || npix * maxv >= (1ull << 31)) return fail(ctx, WR_EINVAL, "wr_render_ppm: bad image size");
    if (film->broken) return film_broken(film, "wr_render_ppm");
    (void)hipSetDevice(ctx->device);
    const CamHost ch = make_cam(*cam);
    const uint64_t per_pass = bdpt_vertex_bytes() * 2 * maxv * npix;
    uint64_t batch = std::max<uint64_t>(1, (4ull << 30) / per_pass);
    batch = std::min<uint64_t>(std::min<uint64_t>(batch, prm->spp ? prm->spp : 1), std::max<uint64_t>(1, (1ull << 22) / npix));
    batch = std::min<uint64_t>(batch, std::max<uint64_t>(1, ((1ull << 31) - 1) / (npix * maxv)));
    const uint64_t lanes = batch * npix;
    VcmBufs vb{};
    vb.cap = lanes * maxv;
    vb.scan_bytes = vcm_scan_bytes((uint32_t)lanes);
    void* verts = nullptr;
    float* pix = nullptr;
    uint32_t* counts = nullptr;
    hipError_t e = hipMalloc(&verts, per_pass * batch);
    if (e == hipSuccess) e = hipMalloc((void**)&pix, sizeof(float) * 3 * lanes);
    if (e == hipSuccess) e = hipMalloc((void**)&counts, sizeof(uint32_t) * 2 * lanes);
    if (e == hipSuccess) e = hipMalloc((void**)&vb.pc, sizeof(uint32_t) * (lanes + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.offs, sizeof(uint32_t) * (lanes + 1));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.starts, sizeof(uint32_t) * (batch + 1));
    if (e == hipSuccess) e = hipMalloc(&vb.scan, std::max<size_t>(vb.scan_bytes, 4));
    if (e == hipSuccess) e = hipMalloc((void**)&vb.pos, sizeof(float) * 3 * vb.cap);
    if (e == hipSuccess) e = hipMalloc((void**)&vb.src, sizeof(uint32_t) * 2 * vb.cap);
    if (e == hipSuccess) e = hipMalloc(&vb.grids, vcm_grid_bytes() * batch);
    auto release = [&]() {
        for (void* ptr : {verts, (void*)pix, (void*)counts, (void*)vb.pc, (void*)vb.offs, (void*)vb.starts, vb.scan,
                          (void*)vb.pos, (void*)vb.src, vb.grids})
            (void)hipFree(ptr);
    };
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e != hipSuccess) {
        release();
        return hip_fail(ctx, e, "wr_render_ppm");
    }
    const FilmBufs fb{film->sum, film->comp, film->sumsq};
    for (uint32_t k = 0; k < prm->spp; k += (uint32_t)batch) {
        const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, prm->spp - k);
        const int rc = launch_ppm_passes(ctx, s->dev, ch, prm->max_depth, prm->seed, film->passes + k, nb, base_radius,
                                         maxv, verts, counts, pix, vb, fb);
        if (rc != WR_OK) {  // the film may hold part of this call's passes
            (void)hipStreamSynchronize(ctx->stream);
            release();
            film->broken = true;
            return rc;
        }
    }
    e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    release();
    if (e != hipSuccess) {
        film->broken = film->broken || prm->spp > 0;
        return hip_fail(ctx, e, "wr_render_ppm");
    }
    film->passes += prm->spp;
    if (stats) {
        stats->paths = npix * prm->spp;
        stats->iterations = 0;
        stats->kernel_ms = ms;
        stats->build_ms = 0.f;
    }
    return WR_OK;
}

extern "C" int wr_set_render_tuning(wr_ctx* ctx, const wr_render_tuning* t) {
    if (!ctx || !t) return fail(ctx, WR_EINVAL, "wr_set_render_tuning: bad argument");
    const uint32_t lv = t->lds_levels ? t->lds_levels : 8u;
    if (lv != 4 && lv != 8 && lv != 16) return fail(ctx, WR_EINVAL, "wr_set_render_tuning: lds_levels must be 4, 8 or 16");
    ctx->tuned_slots = t->wave_slots;
    ctx->tuned_grid = t->wave_grid;
    ctx->tuned_lds = lv;
    return WR_OK;
}

extern "C" int wr_autotune_render(wr_ctx* ctx, const wr_scene* s, const wr_camera* cam, uint32_t max_depth,
                                  uint32_t spp, wr_render_tuning* best) {
    if (!ctx || !s || !cam || spp == 0) return fail(ctx, WR_EINVAL, "wr_autotune_render: bad argument");
    if (!s->built) return fail(ctx, WR_ESTATE, "wr_autotune_render: scene not built");
    wr_film* film = nullptr;
    int rc = wr_film_create(ctx, cam->width, cam->height, &film);
    if (rc != WR_OK) return rc;
    const uint32_t slots[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23};
    const uint32_t grids[] = {1024, 2048, 4096};
    uint32_t best_slots = 0, best_grid = 0;
    float best_ms = 1e30f;
    int last_rc = WR_OK;
    const uint32_t lds0 = ctx->tuned_lds;
    ctx->tuned_lds = 8;  // the pool and grid are chosen at 8 LDS levels
    for (uint32_t sl : slots) {
        for (uint32_t g : grids) {
            // the first render of a configuration also builds the engine: time the second;
            // a configuration that cannot be set up (device memory) is skipped
            wr_render_params p{spp, max_depth, 1, WR_WAVEFRONT, sl, g, 0};
            wr_render_stats st{};
            int r = WR_OK;
            for (int rep = 0; rep < 2 && r == WR_OK; rep++) r = wr_render(ctx, s, cam, &p, film, &st);
            if (r != WR_OK) {
                last_rc = r;
                if (r != WR_ENOMEM && r != WR_EHIP) break;
                continue;
            }
            if (st.kernel_ms < best_ms) {
                best_ms = st.kernel_ms;
                best_slots = sl;
                best_grid = g;
            }
        }
        if (last_rc != WR_OK && last_rc != WR_ENOMEM && last_rc != WR_EHIP) break;
    }
    // then the trace stage's stack levels in LDS at that pool and grid (a move needs a 1% gain)
    uint32_t best_lds = 8;
    if (best_slots != 0) {
        for (uint32_t lv : {4u, 16u}) {
            ctx->tuned_lds = lv;
            wr_render_params p{spp, max_depth, 1, WR_WAVEFRONT, best_slots, best_grid, 0};
            wr_render_stats st{};
            int r = WR_OK;
            for (int rep = 0; rep < 2 && r == WR_OK; rep++) r = wr_render(ctx, s, cam, &p, film, &st);
            if (r != WR_OK) {
                last_rc = r;
                if (r != WR_ENOMEM && r != WR_EHIP) break;
                continue;
            }
            if (st.kernel_ms < 0.99f * best_ms) {
                best_ms = st.kernel_ms;
                best_lds = lv;
            }
        }
    }
    wr_film_destroy(film);
    if (best_slots == 0) {
        ctx->tuned_lds = lds0;
        return last_rc != WR_OK ? last_rc : fail(ctx, WR_EINVAL, "wr_autotune_render: no configuration ran");
    }
    ctx->tuned_slots = best_slots;
    ctx->tuned_grid = best_grid;
    ctx->tuned_lds = best_lds;
    if (best) *best = wr_render_tuning{best_slots, best_grid, best_lds, best_ms};
    return WR_OK;
}
