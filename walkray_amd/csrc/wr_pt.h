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

constexpr int kPathWords = 26;           // 32-bit words of path state per slot (SoA arrays)
constexpr uint32_t kMaxPassWindow = 64;  // passes of one job whose samples may be in flight
constexpr uint32_t kMaxJobs = 64;        // concurrently admitted jobs
constexpr uint32_t kMaxScenes = 64;      // resident scenes of one engine
constexpr int kGraphIters = 8;           // wavefront iterations per graph replay (even)
constexpr int kTailIters = 2;            // iterations per replay once every job has issued all its paths (even)
constexpr uint32_t kDefaultPathSlots = 1u << 22;  // 4M paths in flight: 2 x 416 MB of path state
constexpr uint32_t kBdptMaxDepth = 32;           // BidirectionalPathTracer: maxPathLen limit of wr_render_bdpt

enum : uint32_t { kJobFree = 0, kJobRunning = 1, kJobDone = 2, kJobCancelled = 3 };
// A cancelled job's paths still in the pool are dropped by the next extend
// stage (k_wf_finish reads the job state first), so no kernel touches its
// sample window or film after the state is written; its slot is reused only
// after one more replay, once those paths are gone.
enum : uint32_t { kJobNee = 1 };  // JobDesc::flags: PathTracer's next-event branch

// Device record of one admitted job.
struct JobDesc {
    float pos[3], front[3], right[3], top[3];
    float dist;
    uint32_t W, H;
    uint32_t scene, max_depth;
    uint64_t seed;
    uint32_t pass0, npass, npix, R;  // R: passes in flight (sample window)
    uint32_t flags, pad;             // kJobNee
    float* contrib;                  // R * npix * 3 samples of the passes in flight
    uint32_t* pass_done;             // R counters: finished paths per window slot
    float *fsum, *fcomp, *fsumsq;    // the job's film
};

struct JobCtl {
    uint64_t next_work;  // camera paths issued (pass-major)
    uint32_t cursor;     // passes folded into the film
    uint32_t state;      // kJobFree / kJobRunning / kJobDone
    uint32_t done_iter;  // engine iteration that folded the last pass
    uint32_t start_iter; // engine iteration at admission (set by run_wavefront's admission kernel)
};

struct EngineCounters {
    uint32_t n[2];  // live paths in path buffer 0 / 1
    uint32_t iteration;
    uint32_t rr;    // round-robin start of the planner
    uint32_t nseg, seg_start, seg_end, pad;
    uint32_t seg_job[kMaxJobs], seg_base[kMaxJobs], seg_take[kMaxJobs];
    uint64_t seg_work[kMaxJobs];
    unsigned long long trace_work[kWorkWords];  // k_wf_trace: one path counter per XCD, then finished waves (self-resetting)
    uint32_t n_exit;                   // paths whose shading waits for findInsideObject (k_wf_shade)
    uint32_t nfold;                    // jobs with complete passes to fold this iteration
    uint32_t fold_job[kMaxJobs];       // those jobs, in job order
    uint32_t fold_k[kMaxJobs];         // per job: complete passes from its fold cursor
    // running totals since the engine was set up (k_wf_plan adds each iteration's)
    unsigned long long traced;     // path segments traced (live paths of every iteration)
    unsigned long long exits;      // of those, refraction exits shaded by the findInsideObject instance
    unsigned long long generated;  // camera paths generated
    unsigned long long trace_nodes, trace_tris;  // fast-pass records fetched by the counting trace kernel
};

// The eight kernels of one wavefront iteration, in launch order.
enum : int { kStTrace = 0, kStFinish, kStShade, kStShadeExits, kStPlan, kStGenerate, kStFold, kStAdvance, kStages };

// How run_wavefront issues its iterations: replays of the captured graphs, the
// same kernels launched one by one (identical results; for tools that cannot
// follow graph launches), or launched one by one with a HIP event around every
// kernel and the counting trace kernel (per-stage times and traffic counts).
enum : int { kLaunchGraph = 0, kLaunchDirect = 1, kLaunchStageStats = 2 };

struct StageStats {
    double ms[kStages] = {};
    uint64_t traced = 0, exits = 0, generated = 0, nodes = 0, tris = 0;
};

struct WaveStats {
    uint64_t iterations = 0;
    uint64_t paths = 0;
};

// The shared wavefront: path pool, job and scene tables, and the captured
// iteration graph.  One per context (wr_render) or per server.
struct Engine {
    uint32_t cap = 0, grid = 0, lds_levels = 0;  // lds_levels: k_wf_trace's traversal stack levels in LDS
    EngineCounters* cnt = nullptr;
    JobDesc* jobs = nullptr;
    JobCtl* ctl = nullptr;
    SceneDev* scenes = nullptr;
    uint32_t* exit_list = nullptr;  // path slots of the deferred refraction exits (cap entries)
    // the iteration graphs, per shading instance: [0] without next-event
    // estimation (no running job asks for it), [1] with it
    hipGraph_t graph[2] = {}, tail_graph[2] = {};
    hipGraphExec_t exec[2] = {}, tail_exec[2] = {};
    std::vector<void*> bufs;
    // run_wavefront's per-render buffers, kept between renders (grow-only):
    // the per-pass contributions and completion counts of its one job
    float* contrib = nullptr;
    uint32_t* pass_done = nullptr;
    size_t contrib_floats = 0, pass_slots = 0;
    // pinned host buffer of run_wavefront's job-state read-back
    JobCtl* ctl_back = nullptr;
    // what the iteration kernels are launched with, kept for direct launches
    void* words[2][kPathWords] = {};
    uint32_t trace_grid = 0;
    // Launches `iters` iterations one kernel at a time (the graphs' sequence).
    // With ev (kStages * iters + 1 events) an event is recorded before the first
    // kernel and after every kernel, and the trace stage counts its fetches.
    // nee: the shading instance with next-event estimation, needed as soon as
    // one job in the pool asks for it (the other instance leaves its code and
    // registers out).
    hipError_t enqueue(int iters, hipStream_t st, hipEvent_t* ev, bool nee);
    ~Engine();
    void release();
    bool ready = false;  // buffers allocated and zeroed for (cap, grid, lds_levels)
    hipError_t init(uint32_t cap, uint32_t grid, hipStream_t st, uint32_t lds_levels = 8);
    hipError_t ensure_graphs(hipStream_t st, bool nee);
    hipError_t set_scene(uint32_t slot, const SceneDev& S, hipStream_t st);
    hipError_t admit(uint32_t slot, const JobDesc& J, hipStream_t st);
    hipError_t release_slot(uint32_t slot, hipStream_t st);
    hipError_t cancel_slot(uint32_t slot, hipStream_t st);
    // One replay: kGraphIters iterations, or kTailIters when `tail` (every
    // running job has issued all its camera paths, so only draining is left).
    // nee: some job in the pool (running or draining) asks for next-event
    // estimation.
    hipError_t step(hipStream_t st, JobCtl* host_ctl, EngineCounters* host_cnt, bool nee, bool tail = false,
                    bool direct = false);
};

void fill_job_camera(JobDesc& J, const CamHost& c);

hipError_t launch_pt_mega(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass,
                          bool nee, const FilmBufs& film, hipStream_t st);
// BidirectionalPathTracer passes [pass0, pass0 + npass) side by side (bdpt.hip):
// verts holds 2 * maxv vertex records (bdpt_vertex_bytes() each) per pixel and
// pass, pix 3 floats per pixel and pass.
// q: the visibility queue (cap entries: rays 32 B, pay 16 B, dst 4 B, hits
// 16 B each) and its counter.
struct BdptQueue {
    uint32_t* rays;
    float* pay;
    uint32_t* dst;
    void* hits;
    uint32_t* count;
    uint32_t cap;
    uint32_t* counts;  // 2 per pixel and pass: the eye and light sub-path record counts
};
constexpr size_t kBdptQueueEntryBytes = 32 + 16 + 4 + 16;
size_t bdpt_vertex_bytes();
hipError_t launch_bdpt_passes(const SceneDev& S, const wr_tree_hdr& H, const CamHost& cam, uint32_t max_depth,
                              uint64_t seed, uint32_t pass0, uint32_t npass, uint32_t maxv, void* verts, float* pix,
                              const BdptQueue& q, const FilmBufs& film, hipStream_t st);

hipError_t run_wavefront(const SceneDev& S, const CamHost& cam, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                         uint32_t npass, bool nee, const FilmBufs& film, Engine& engine, uint32_t cap, uint32_t grid,
                         WaveStats* stats, hipStream_t st, uint32_t lds_levels = 8, int launch = kLaunchGraph,
                         StageStats* stage = nullptr);

}  // namespace wr
