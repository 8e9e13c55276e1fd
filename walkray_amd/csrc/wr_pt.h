// wr_pt.h -- host-side declarations of the path-tracing drivers (pt.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "wr_kernels.h"

namespace wr {

// Camera basis computed on the host with the reference's float semantics
// (camera.cpp:7-11): front = normalize(focus - pos), right =
// normalize(front x up), top = right x front.
struct CamHost {
    float pos[3], front[3], right[3], top[3];
    float dist;
    uint32_t W, H;
};

struct FilmBufs {
    float* sum;
    float* comp;
    float* sumsq;
};

// Device counters of one wavefront run (all 32-bit but next_work / take_work).
struct WaveCounters {
    uint32_t n_shade[2];
    uint32_t n_extend;
    uint32_t free_head, free_tail;
    uint32_t take, take_head;
    uint32_t finished;
    uint64_t next_work, take_work;
    uint32_t pass_done[4];
};

struct WaveStats {
    uint64_t iterations = 0;
    uint64_t paths = 0;
};

// Path-slot pool, reused across render calls of one context.
struct WavefrontPool {
    float *ox = nullptr, *oy = nullptr, *oz = nullptr, *dx = nullptr, *dy = nullptr, *dz = nullptr;
    float *tx = nullptr, *ty = nullptr, *tz = nullptr, *ht = nullptr;
    uint32_t *pixel = nullptr, *pass = nullptr, *depth = nullptr, *inside = nullptr, *contact = nullptr;
    uint32_t *hrec = nullptr, *hobj = nullptr, *q_shade0 = nullptr, *q_shade1 = nullptr, *q_extend = nullptr,
             *q_free = nullptr;
    WaveCounters* cnt = nullptr;
    float* contrib = nullptr;
    uint32_t nslots = 0, npix = 0, R = 0;
    std::vector<void*> bufs;
    ~WavefrontPool();
    void release();
    hipError_t ensure(uint32_t slots, uint32_t npix, uint32_t R);
};

hipError_t launch_pt_mega(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass,
                          const FilmBufs& film, hipStream_t st);
hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, const FilmBufs& film, WavefrontPool& pool, uint32_t slots, uint32_t grid,
                         WaveStats* stats, hipStream_t st);

}  // namespace wr
