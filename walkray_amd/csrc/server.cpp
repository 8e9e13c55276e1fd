// server.cpp -- render-job server over the shared wavefront engine
// (include/walkray.h, "serving").  Continuous batching: jobs are admitted
// between engine steps whenever a job slot and enough of the device-memory
// budget are free, and their paths join the same path pool as the jobs
// already running.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "../../include/walkray.h"
#include "wr_capi_internal.h"
#include "wr_pt.h"

namespace {

void vnormalize(float v[3]) {  // nvVector.h:520-525 semantics
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

struct Job {
    wr_job_desc desc{};
    wr_job_status st{};
    int slot = -1;
    uint32_t R = 0;
    float *sum = nullptr, *comp = nullptr, *sumsq = nullptr, *contrib = nullptr;
    uint32_t* pass_done = nullptr;
    uint64_t film_bytes = 0, window_bytes = 0;
    uint64_t total = 0;  // camera paths of the job (spp x pixels)
};

}  // namespace

struct wr_server {
    wr_ctx* ctx = nullptr;
    wr::Engine engine;
    uint64_t budget = 0, used = 0;
    uint32_t max_jobs = 0;
    std::vector<const wr_scene*> scenes;
    std::map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;
    std::vector<int64_t> slot_job;  // engine job slot -> job id (-1 free, kDraining: a cancelled job's paths)
    std::vector<uint8_t> slot_nee;  // the slot's job (running or draining) asks for next-event estimation
    uint32_t inject_admit_failures = 0;  // test hook: admissions left to fail as if allocation had failed
    uint64_t next_id = 1;
    bool tail = false;  // every running job had issued all its paths at the last read-back
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

using namespace wr;

// slot_job value of a slot whose job was cancelled while running: the slot is
// reused once the next replay has dropped the job's paths from the pool.
static constexpr int64_t kDraining = -2;

static void free_job_window(Job& j) {
    if (j.contrib) (void)hipFree(j.contrib);
    if (j.pass_done) (void)hipFree(j.pass_done);
    j.contrib = nullptr;
    j.pass_done = nullptr;
}

static void free_job_film(Job& j) {
    if (j.sum) (void)hipFree(j.sum);
    if (j.comp) (void)hipFree(j.comp);
    if (j.sumsq) (void)hipFree(j.sumsq);
    j.sum = j.comp = j.sumsq = nullptr;
}

extern "C" {

int wr_server_create(wr_ctx* ctx, const wr_server_config* cfg, wr_server** out) {
    if (!ctx || !out) return fail(ctx, WR_EINVAL, "wr_server_create: bad argument");
    *out = nullptr;
    wr_server* s = new wr_server;
    s->ctx = ctx;
    s->t0 = std::chrono::steady_clock::now();
    s->budget = cfg && cfg->job_memory_budget ? cfg->job_memory_budget : (8ull << 30);
    s->max_jobs = cfg && cfg->max_jobs ? std::min<uint32_t>(cfg->max_jobs, kMaxJobs) : kMaxJobs;
    const uint32_t slots = cfg && cfg->path_slots ? cfg->path_slots : kDefaultPathSlots;
    s->slot_job.assign(kMaxJobs, -1);
    s->slot_nee.assign(kMaxJobs, 0);
    (void)hipSetDevice(ctx->device);
    hipError_t e = s->engine.init(slots, 0, ctx->stream);
    if (e != hipSuccess) {
        delete s;
        return hip_fail(ctx, e, "wr_server_create: engine");
    }
    *out = s;
    return WR_OK;
}

void wr_server_destroy(wr_server* s) {
    if (!s) return;
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto& kv : s->jobs) {
        free_job_window(kv.second);
        free_job_film(kv.second);
    }
    delete s;
}

double wr_server_clock(const wr_server* s) { return s ? s->now() : 0.0; }

int wr_server_add_scene(wr_server* s, const wr_scene* scene, uint32_t* scene_id) {
    if (!s || !scene || !scene_id) return fail(s ? s->ctx : nullptr, WR_EINVAL, "wr_server_add_scene: bad argument");
    if (!scene->built) return fail(s->ctx, WR_ESTATE, "wr_server_add_scene: scene not built");
    for (const HostObject& o : scene->objs)
        if (o.moving) return fail(s->ctx, WR_EINVAL, "wr_server_add_scene: moving objects are rendered by wr_render only");
    if (s->scenes.size() >= kMaxScenes) return fail(s->ctx, WR_EBUSY, "wr_server_add_scene: too many scenes");
    const uint32_t id = (uint32_t)s->scenes.size();
    hipError_t e = s->engine.set_scene(id, scene->dev, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_server_add_scene");
    s->scenes.push_back(scene);
    *scene_id = id;
    return WR_OK;
}

int wr_job_submit(wr_server* s, const wr_job_desc* d, uint64_t* job_id) {
    if (!s || !d || !job_id) return fail(s ? s->ctx : nullptr, WR_EINVAL, "wr_job_submit: bad argument");
    if (d->scene >= s->scenes.size()) return fail(s->ctx, WR_EINVAL, "wr_job_submit: unknown scene");
    if (d->cam.width == 0 || d->cam.height == 0 || (uint64_t)d->cam.width * d->cam.height >= (1ull << 31))
        return fail(s->ctx, WR_EINVAL, "wr_job_submit: bad resolution");
    Job j;
    j.desc = *d;
    j.st.state = WR_JOB_QUEUED;
    j.st.submit_s = s->now();
    const uint64_t npix = (uint64_t)d->cam.width * d->cam.height;
    j.st.samples = npix * d->spp;
    j.R = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(std::max<uint32_t>(d->spp, 1), s->engine.cap / npix + 2),
                                       kMaxPassWindow);
    j.film_bytes = 3 * sizeof(float) * 3 * npix;
    j.window_bytes = 3 * sizeof(float) * j.R * npix + sizeof(uint32_t) * j.R;
    j.st.device_bytes = j.film_bytes + j.window_bytes;
    const uint64_t id = s->next_id++;
    if (j.st.device_bytes > s->budget || d->spp == 0) {
        j.st.state = d->spp == 0 ? WR_JOB_DONE : WR_JOB_FAILED;
        s->jobs[id] = j;
        *job_id = id;
        if (j.st.state == WR_JOB_FAILED) return fail(s->ctx, WR_ENOMEM, "wr_job_submit: job exceeds the memory budget");
        return WR_OK;
    }
    s->jobs[id] = j;
    s->queue.push_back(id);
    *job_id = id;
    return WR_OK;
}

static hipError_t admit_one(wr_server* s, Job& j, int slot) {
    hipStream_t st = s->ctx->stream;
    const uint32_t npix = j.desc.cam.width * j.desc.cam.height;
    const size_t fb = sizeof(float) * 3 * (size_t)npix;
    if (s->inject_admit_failures) {
        s->inject_admit_failures--;
        return hipErrorOutOfMemory;
    }
    hipError_t e = hipMalloc((void**)&j.sum, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&j.comp, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&j.sumsq, fb);
    if (e == hipSuccess) e = hipMalloc((void**)&j.contrib, sizeof(float) * 3 * (size_t)j.R * npix);
    if (e == hipSuccess) e = hipMalloc((void**)&j.pass_done, sizeof(uint32_t) * j.R);
    if (e == hipSuccess) e = hipMemsetAsync(j.sum, 0, fb, st);
    if (e == hipSuccess) e = hipMemsetAsync(j.comp, 0, fb, st);
    if (e == hipSuccess) e = hipMemsetAsync(j.sumsq, 0, fb, st);
    if (e == hipSuccess) e = hipMemsetAsync(j.pass_done, 0, sizeof(uint32_t) * j.R, st);
    if (e != hipSuccess) {
        free_job_window(j);
        free_job_film(j);
        return e;
    }
    wr_camera c = j.desc.cam;
    CamHost ch;
    for (int k = 0; k < 3; k++) {
        ch.pos[k] = c.pos[k];
        ch.front[k] = c.focus[k] - c.pos[k];
    }
    vnormalize(ch.front);
    vcross(ch.front, c.up, ch.right);
    vnormalize(ch.right);
    vcross(ch.right, ch.front, ch.top);
    ch.dist = c.sight_dist;
    ch.W = c.width;
    ch.H = c.height;
    JobDesc J{};
    fill_job_camera(J, ch);
    J.scene = j.desc.scene;
    J.max_depth = j.desc.max_depth;
    J.flags = (j.desc.flags & WR_RENDER_NEE) ? kJobNee : 0u;
    J.seed = j.desc.seed;
    J.pass0 = 0;
    J.npass = j.desc.spp;
    J.npix = npix;
    j.total = (uint64_t)j.desc.spp * npix;
    J.R = j.R;
    J.contrib = j.contrib;
    J.pass_done = j.pass_done;
    J.fsum = j.sum;
    J.fcomp = j.comp;
    J.fsumsq = j.sumsq;
    e = s->engine.admit((uint32_t)slot, J, st);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        free_job_window(j);
        free_job_film(j);
        return e;
    }
    j.slot = slot;
    j.st.state = WR_JOB_RUNNING;
    j.st.admit_s = s->now();
    s->used += j.film_bytes + j.window_bytes;
    return hipSuccess;
}

int wr_server_step(wr_server* s, uint32_t* n_active) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_server_step: server is NULL");
    hipStream_t st = s->ctx->stream;
    bool admitted = false;
    // admission (first come, first served) against the slot limit and the budget
    while (!s->queue.empty()) {
        Job& j = s->jobs[s->queue.front()];
        int slot = -1;
        uint32_t running = 0;
        for (uint32_t k = 0; k < kMaxJobs; k++) {
            if (s->slot_job[k] >= 0) running++;
            else if (s->slot_job[k] == -1 && slot < 0) slot = (int)k;  // a draining slot is not free yet
        }
        if (slot < 0 || running >= s->max_jobs || s->used + j.film_bytes + j.window_bytes > s->budget) break;
        hipError_t e = admit_one(s, j, slot);
        if (e != hipSuccess) {
            // the failure stays with this job: its buffers are freed, it leaves the
            // queue as failed, and admission goes on with the next one
            (void)hipGetLastError();  // a failed allocation must not surface at a later launch check
            j.st.state = WR_JOB_FAILED;
            j.st.done_s = s->now();
            s->queue.pop_front();
            continue;
        }
        s->slot_job[slot] = (int64_t)s->queue.front();
        s->slot_nee[slot] = (j.desc.flags & WR_RENDER_NEE) != 0;
        s->queue.pop_front();
        admitted = true;
    }
    uint32_t running = 0;  // slots with paths in the pool: admitted jobs and draining cancelled ones
    for (uint32_t k = 0; k < kMaxJobs; k++) running += s->slot_job[k] >= 0 || s->slot_job[k] == kDraining;
    if (running) {
        std::vector<JobCtl> hc(kMaxJobs);
        // the shading instance with next-event estimation while any job in the pool asks for it
        bool nee = false;
        for (uint32_t k = 0; k < kMaxJobs; k++) nee = nee || s->slot_nee[k];
        // short replays once only draining is left (nothing queued or newly admitted)
        hipError_t e = s->engine.step(st, hc.data(), nullptr, nee, s->tail && !admitted && s->queue.empty());
        if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_server_step: engine");
        const double t = s->now();
        s->tail = true;
        for (uint32_t k = 0; k < kMaxJobs; k++)
            if (s->slot_job[k] >= 0 && hc[k].state != kJobDone &&
                hc[k].next_work < s->jobs[(uint64_t)s->slot_job[k]].total)
                s->tail = false;
        for (uint32_t k = 0; k < kMaxJobs; k++) {
            if (s->slot_job[k] != kDraining) continue;  // the replay dropped the cancelled job's paths
            e = s->engine.release_slot(k, st);
            if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_server_step: release");
            s->slot_job[k] = -1;
            s->slot_nee[k] = 0;
        }
        for (uint32_t k = 0; k < kMaxJobs; k++) {
            if (s->slot_job[k] < 0 || hc[k].state != kJobDone) continue;
            Job& j = s->jobs[(uint64_t)s->slot_job[k]];
            j.st.state = WR_JOB_DONE;
            j.st.done_s = t;
            free_job_window(j);
            s->used -= j.window_bytes;
            e = s->engine.release_slot(k, st);
            if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_server_step: release");
            s->slot_job[k] = -1;
            s->slot_nee[k] = 0;
            j.slot = -1;
        }
        running = 0;
        for (uint32_t k = 0; k < kMaxJobs; k++) running += s->slot_job[k] >= 0 || s->slot_job[k] == kDraining;
    }
    if (n_active) *n_active = running + (uint32_t)s->queue.size();
    return WR_OK;
}

int wr_job_cancel(wr_server* s, uint64_t id) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_job_cancel: server is NULL");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s->ctx, WR_EINVAL, "wr_job_cancel: unknown job");
    Job& j = it->second;
    if (j.st.state == WR_JOB_CANCELLED) return WR_OK;
    if (j.st.state == WR_JOB_DONE || j.st.state == WR_JOB_FAILED)
        return fail(s->ctx, WR_ESTATE, "wr_job_cancel: job already finished");
    if (j.st.state == WR_JOB_QUEUED) {
        s->queue.erase(std::find(s->queue.begin(), s->queue.end(), id));
    } else {
        // running: once the state is on the device no kernel touches the job's
        // window or film (k_wf_finish drops its paths first), so both are freed now;
        // the slot waits for the next replay
        hipError_t e = s->engine.cancel_slot((uint32_t)j.slot, s->ctx->stream);
        if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_job_cancel");
        free_job_window(j);
        free_job_film(j);
        s->used -= j.film_bytes + j.window_bytes;
        s->slot_job[j.slot] = kDraining;
        j.slot = -1;
    }
    j.st.state = WR_JOB_CANCELLED;
    j.st.done_s = s->now();
    return WR_OK;
}

int wr_server_inject_admit_failures(wr_server* s, uint32_t n) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_server_inject_admit_failures: server is NULL");
    s->inject_admit_failures = n;
    return WR_OK;
}

int wr_job_get_status(wr_server* s, uint64_t id, wr_job_status* out) {
    if (!s || !out) return fail(s ? s->ctx : nullptr, WR_EINVAL, "wr_job_get_status: bad argument");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s->ctx, WR_EINVAL, "wr_job_get_status: unknown job");
    *out = it->second.st;
    return WR_OK;
}

int wr_job_read(wr_server* s, uint64_t id, float* mean_rgb, float* sum, float* comp, float* sumsq, uint32_t* passes) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_job_read: server is NULL");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s->ctx, WR_EINVAL, "wr_job_read: unknown job");
    Job& j = it->second;
    if (j.st.state != WR_JOB_DONE || !j.sum) return fail(s->ctx, WR_ESTATE, "wr_job_read: job not finished");
    const size_t n = 3 * (size_t)j.desc.cam.width * j.desc.cam.height;
    hipStream_t st = s->ctx->stream;
    std::vector<float> tmp;
    float* sp = sum;
    if (!sp && mean_rgb) {
        tmp.resize(n);
        sp = tmp.data();
    }
    hipError_t e = hipSuccess;
    if (sp) e = hipMemcpyAsync(sp, j.sum, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && comp) e = hipMemcpyAsync(comp, j.comp, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && sumsq) e = hipMemcpyAsync(sumsq, j.sumsq, n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "wr_job_read");
    if (mean_rgb) {
        const float inv = 1.f / (float)j.desc.spp;
        for (size_t i = 0; i < n; i++) mean_rgb[i] = sp[i] * inv;
    }
    if (passes) *passes = j.desc.spp;
    return WR_OK;
}

int wr_job_release(wr_server* s, uint64_t id) {
    if (!s) return fail(nullptr, WR_EINVAL, "wr_job_release: server is NULL");
    auto it = s->jobs.find(id);
    if (it == s->jobs.end()) return fail(s->ctx, WR_EINVAL, "wr_job_release: unknown job");
    Job& j = it->second;
    if (j.st.state == WR_JOB_RUNNING || j.st.state == WR_JOB_QUEUED)
        return fail(s->ctx, WR_EBUSY, "wr_job_release: job still queued or running");
    (void)hipStreamSynchronize(s->ctx->stream);
    if (j.sum) s->used -= j.film_bytes;
    free_job_film(j);
    s->jobs.erase(it);
    return WR_OK;
}

}  // extern "C"
