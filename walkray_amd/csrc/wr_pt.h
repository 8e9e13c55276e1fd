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

constexpr int kPathWords = 18;       // 32-bit words of path state per slot (SoA arrays)
constexpr uint32_t kMaxPassWindow = 64;  // passes whose samples may be in flight at once
constexpr int kGraphIters = 8;       // wavefront iterations per captured graph (even)

// Device counters of one wavefront run.
struct WaveCounters {
    uint32_t n[2];       // live paths in path buffer 0 / 1
    uint32_t take, take_base;
    uint64_t next_work, take_work;
    uint32_t cursor;     // passes folded into the film
    uint32_t iterations;
    uint32_t pass_done[kMaxPassWindow];
};

struct WaveStats {
    uint64_t iterations = 0;
    uint64_t paths = 0;
};

// Two path buffers (ping-pong), counters and the per-(pass, pixel) sample
// window; reused across render calls of one context.
struct WavefrontPool {
    void* words[2][kPathWords] = {};
    WaveCounters* cnt = nullptr;
    float* contrib = nullptr;
    uint32_t cap = 0;
    size_t contrib_floats = 0;
    std::vector<void*> bufs;
    ~WavefrontPool();
    void release();
    hipError_t ensure(uint32_t cap, size_t contrib_floats);
};

hipError_t launch_pt_mega(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass,
                          const FilmBufs& film, hipStream_t st);
hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, const FilmBufs& film, WavefrontPool& pool, uint32_t slots, uint32_t grid,
                         WaveStats* stats, hipStream_t st);

}  // namespace wr
