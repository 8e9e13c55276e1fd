// bdpt.hip -- BidirectionalPathTracer (bidirPathTracer.cpp:10-313) on gfx950.
//
// One lane per pixel and pass (the reference's per-pixel loop), megakernel
// style:
//   1. the eye sub-path: the camera ray (camera.cpp:6-60) and
//      MCRenderer::samplePath (mcRenderer.cpp:16-57) from it;
//   2. the light sub-path: Scene::genLightSample as
//      AbstractObject::genSurfaceSample (abstractObj.cpp:18-52) and samplePath
//      from it, with the light-path vertex counters;
//   3. every (light length, eye length) split of every joined length
//      (bidirPathTracer.cpp:53-140): connectEdge (294-324), the visibility
//      test (mcRenderer.cpp:59-65), connectingColorProb (266-291) and
//      connectingMisWeight (142-221) with linkVertex (228-243), in one sweep;
//   4. contributions with an eye side of two or more vertices go to the
//      lane's pixel, light-tracing ones (the camera vertex alone on the eye
//      side) are splatted to the pixel Camera::convToRaster names
//      (camera.cpp:148-170), scaled by 1 / pixel solid angle / pixel count.
// The visibility tests of the connections are not traced by the lane that
// forms them: each connection that passes every other test appends its
// visibility ray (and its contribution and target pixel) to a queue, traced
// afterwards one ray per lane by the batch NoSelf query (k_closest_hit_path:
// the same traversal, so the same answers) and folded in by k_bdpt_accum; a
// lane whose queue is full traces in place.
// The sub-paths live in device memory (one block of records per lane).  A
// launch runs several passes side by side (lane = pass x pixel: one pass of a
// 256^2 image is only 1,024 waves of long, serial lanes).  The pixel values
// of each pass are summed with float atomics in a buffer of their own (the
// splats of other lanes land on any pixel), then k_bdpt_fold applies
// isLegalColor and fixVignetting and adds them to the Kahan film in pass
// order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <vector>

#include "grid_dev.h"
#include "pt_dev.h"
#include "wr_pt.h"

#pragma clang fp contract(off)

namespace wrd {

constexpr uint32_t kBdptLightVertex = 0x80000000u;  // as the light sub-paths of photon.hip
constexpr uint32_t kCameraObj = 0xFFFFFFFEu;        // the camera as a NoSelf contact: not an object, not NULL
constexpr uint32_t kVDelta = 1u, kVCamera = 2u;
constexpr int kBdptMaxJoined = 2 * (wr::kBdptMaxDepth + 3);

// One Ray of a Path as the estimator reads it (ray.h:8-37).
struct BVert {
    f3 o, d, rad, n;  // n: the contact normal at o (the camera's: front)
    float op, dp;
    uint32_t inside, contact, flags, pad;
};

__device__ __forceinline__ void put(BVert* V, uint32_t& nv, uint32_t maxv, const BVert& v) {
    if (nv < maxv) V[nv] = v;
    nv++;
}

// terminateRay (mcRenderer.cpp:20-27) after a ray from o.
__device__ __forceinline__ BVert terminate_at(f3 o) {
    BVert t;
    t.o = o;
    t.d = mk(0.f, 0.f, 0.f);
    t.rad = t.d;
    t.n = t.d;
    t.op = 1.f;
    t.dp = 1.f;
    t.inside = WR_NONE;
    t.contact = WR_NONE;
    t.flags = kVDelta;
    t.pad = 0;
    return t;
}

// samplePath from v0 (already traced: hit) with vertex counters vbase + depth + 1.
__device__ void walk(const SceneCtx& sc, const BVert& v0, PHit hit, uint32_t max_depth, uint64_t seed, uint32_t p,
                     uint32_t pass, uint32_t vbase, BVert* V, uint32_t& nv, uint32_t maxv, LaneStack& st) {
    put(V, nv, maxv, v0);
    f3 o = v0.o, d = v0.d;
    uint32_t inside = v0.inside;
    for (uint32_t depth = 0;; depth++) {
        const uint32_t self = inside != WR_NONE ? inside : hit.obj;
        if (self == WR_NONE || hit.rec == WR_NONE) {
            put(V, nv, maxv, terminate_at(o));
            return;
        }
        float rn[4];
        rng4(seed, p, pass, vbase + depth + 1, rn);
        const Out r = scatter(sc, self, o, d, hit, inside, rn, st);
        BVert nx;
        nx.o = r.o;
        nx.d = r.d;
        nx.rad = r.rad;
        nx.n = r.n;
        nx.op = r.op;
        nx.dp = r.dp;
        nx.inside = r.inside;
        nx.contact = r.contact;
        nx.flags = r.delta ? kVDelta : 0u;
        nx.pad = 0;
        if ((double)len(r.d) < 0.5) {
            put(V, nv, maxv, nx);
            return;
        }
        if (depth + 1 > max_depth) {
            put(V, nv, maxv, terminate_at(o));
            return;
        }
        const PHit nh = trace_scene<true>(sc, r.o, r.d, r.inside, r.contact, st);
        put(V, nv, maxv, nx);
        if (nh.rec == WR_NONE) {
            put(V, nv, maxv, terminate_at(o));
            return;
        }
        o = r.o;
        d = r.d;
        inside = r.inside;
        hit = nh;
    }
}

// Camera::getPixelSolidAngle (camera.cpp:124-131).
__device__ __forceinline__ float pixel_solid_angle(const CamDev& c, uint32_t p) {
    const uint32_t x = p % c.W, y = p / c.W;
    const float xc = (float)((double)x + 0.5 - (c.W / 2.0)), yc = (float)((c.H / 2.0) - 0.5 - (double)y);
    const float pd = sqrtf(xc * xc + yc * yc + c.dist * c.dist);
    return c.dist / powf(pd, 3.f);
}

// Ray::cosineTerm (ray.cpp:52-58).
__device__ __forceinline__ float cos_term_v(const BVert& r) {
    float c = 1.f;
    if (((r.flags & kVCamera) || r.contact != WR_NONE) && (double)len(r.d) > 0.5) c = fabsf(dot(r.n, r.d));
    return (double)c > kEps ? c : (float)kEps;
}

// in.evalBSDF(out) (ray.cpp:15-22): the object of out's contact (else inside).
__device__ f3 eval_bsdf(const SceneDev& S, const BVert& in, const BVert& out) {
    if (out.flags & kVCamera) return mk(1.f, 1.f, 1.f);  // AbstractObject::evaluateBSDF
    const uint32_t obj = out.contact != WR_NONE ? out.contact : out.inside;
    if (obj == WR_NONE) return mk(0.f, 0.f, 0.f);
    const wr_object_gpu& ob = S.objs[obj];
    const f3 color = mk(ob.color[0], ob.color[1], ob.color[2]);
    if (ob.mat == WR_MAT_REFRACTIVE) return mk(0.f, 0.f, 0.f);  // refractiveObj.h:13-15
    if (ob.mat == WR_MAT_EMISSIVE) return color;                // emissiveObj.h:13
    if (out.contact == WR_NONE) return mk(0.f, 0.f, 0.f);       // diffuseObj.h:18, glossyObj.h:21
    const f3 n = out.n;
    if (!(dot(in.d, n) <= 0.f && dot(out.d, n) >= 0.f)) return mk(0.f, 0.f, 0.f);
    if (ob.mat == WR_MAT_DIFFUSE) return muls(color, (float)kInvPi);  // diffuseBsdf.h:29-34
    const f3 refl = normalize(add(muls(muls(n, -dot(n, in.d)), 2.f), in.d));  // isotropicPhong.h:41-48
    const f3 rho = muls(muls(muls(color, (float)((double)ob.power + 2.0)), 0.5f), (float)kInvPi);
    const float dRW = dot(refl, out.d);
    return muls(rho, powf(dRW > 0.f ? dRW : 0.f, ob.power));
}

// in.evalDirectionProbability(out) (ray.cpp:31-38); same: the emitted branch of
// AbstractObject::evalDirectionProbability (in and out one object).
__device__ float eval_dir_prob(const SceneDev& S, const CamDev& cam, uint32_t pixel, const BVert& in,
                               const BVert& out, bool same) {
    (void)in;
    if (out.flags & kVCamera) return (float)(1.0 / (double)pixel_solid_angle(cam, pixel));  // camera.cpp:144-146
    const uint32_t obj = out.contact != WR_NONE ? out.contact : out.inside;
    if (obj == WR_NONE) return 0.f;
    const wr_object_gpu& ob = S.objs[obj];
    if (ob.mat == WR_MAT_REFRACTIVE || ob.mat == WR_MAT_EMISSIVE)  // abstractObj.cpp:61-68
        return same ? (float)(kInvPi * 0.25f) * 2.f : 0.f;
    if (out.contact == WR_NONE) return 0.f;
    const float p = ob.pmax;
    const float c = dot(out.n, out.d);
    if (ob.mat == WR_MAT_DIFFUSE) return p * (c > 0.f ? (float)((double)c * kInvPi) : 0.f);  // diffuseBsdf.h:12-14
    const float cos_theta = (double)c > kEps ? c : (float)kEps;                                // isotropicPhong.h:13-18
    return p * (float)(((double)ob.power + 1.0) * (double)powf(cos_theta, ob.power) * (kInvPi * 0.5));
}

// The joined path light[0..nl) + reversed eye[0..ne), with the two connection
// records (indices cl, cl + 1) replaced by their edited copies once connected.
struct Joined {
    const BVert* L;
    const BVert* E;
    int nl, ne, n, cl;
    bool edited;
    BVert lc, ec;
    __device__ __forceinline__ BVert at(int i) const {
        if (edited && i == cl) return lc;
        if (edited && i == cl + 1) return ec;
        return i < nl ? L[i] : E[ne - 1 - (i - nl)];
    }
    __device__ __forceinline__ bool delta(int i) const { return (at(i).flags & kVDelta) != 0; }
};

// pathCanNotConnect (bidirPathTracer.cpp:251-257)
__device__ __forceinline__ bool cannot_connect(const Joined& P) {
    for (int i = 0; i + 1 < P.n; i++)
        if (!P.delta(i) && !P.delta(i + 1)) return false;
    return true;
}

// connectEdge (bidirPathTracer.cpp:294-324): edits the copies lc / ec; the
// previous records are read after the edits (for cl = 0, and for an eye side
// of one record, the previous record is the connection record itself).
__device__ bool connect_edge(const SceneDev& S, Joined& P) {
    P.lc = P.at(P.cl);
    P.ec = P.at(P.cl + 1);
    if ((P.lc.flags & kVDelta) || (P.ec.flags & kVDelta)) return false;
    P.edited = true;
    const int last = P.n - 1;
    const int il = P.cl - 1 > 0 ? P.cl - 1 : 0;
    const int ie = P.cl + 2 < last ? P.cl + 2 : last;
    P.lc.d = normalize(sub(P.ec.o, P.lc.o));
    P.ec.d = mk(-P.lc.d.x, -P.lc.d.y, -P.lc.d.z);
    P.lc.dp = 1.f;
    P.lc.flags &= ~kVDelta;
    {
        const BVert pre = P.at(il);  // lc itself when il == cl
        P.lc.rad = eval_bsdf(S, pre, P.lc);
        if (P.lc.contact != WR_NONE && P.cl > 0)
            P.lc.inside = dot(P.lc.n, P.lc.d) > 0.f ? pre.inside : P.lc.contact;
    }
    P.ec.dp = 1.f;
    P.ec.flags &= ~kVDelta;
    {
        const BVert pre = P.at(ie);  // ec itself when ie == cl + 1
        P.ec.rad = eval_bsdf(S, pre, P.ec);
        if (P.ec.contact != WR_NONE && P.cl < last - 1)
            P.ec.inside = dot(P.ec.n, P.ec.d) > 0.f ? pre.inside : P.ec.contact;
    }
    return true;
}

// connectingColorProb (bidirPathTracer.cpp:266-291; no participating media)
// and connectingMisWeight (142-221, expTerm 1; every evalOriginProbability
// there compares two temporaries, so it is 1) in one sweep over the joined
// path, each record read once.  linkVertex(j, reverse = false) is record j
// re-aimed at j + 1 when j > cl, linkVertex(j, true) record j re-aimed at
// j - 1 when j <= cl; the weight reads only the cosine term and the
// direction probability of those links (evalDirectionProbability ignores its
// incoming ray), so the sweep keeps those per record and the two recurrences
// run over them afterwards, in the reference's order.
// With vt (VCMTracer) the colour and probability are connectingColorProb of
// vcmTracer.cpp:329-362 (a merge also takes the merging disc, the area
// density of the merged vertex and, inside a medium, the sphere's extra length)
// and the weight is vcmMisWeight (vcmTracer.cpp:221-320, expTerm 1, nVC = 1):
// the same two recurrences, with every merging strategy beside every
// connection in the sum.
struct VcmTerms {
    float r;     // the pass's merging radius
    double nvm;  // light sub-paths of the pass (one per pixel)
    bool vm;     // a merge (VM), else a connection (VC)
};

// insideObj && !contactObj: a point inside a medium (the camera is a contact)
__device__ __forceinline__ bool in_volume(const BVert& v) {
    return v.inside != WR_NONE && v.contact == WR_NONE && !(v.flags & kVCamera);
}

__device__ float color_prob_weight(const SceneDev& S, const CamDev& cam, uint32_t pixel, const Joined& P, f3& color,
                                   float& prob, const VcmTerms* vt = nullptr) {
    float cos_f[kBdptMaxJoined], cos_b[kBdptMaxJoined], dp_f[kBdptMaxJoined], dp_b[kBdptMaxJoined];
    float dist[kBdptMaxJoined], op[kBdptMaxJoined];
    uint8_t delta[kBdptMaxJoined], vol[kBdptMaxJoined];
    const int n = P.n, cl = P.cl;
    color = mk(1.f, 1.f, 1.f);
    prob = 1.f;
    BVert prev = P.at(0), cur = prev, next = prev;
    for (int j = 0; j < n; j++) {
        if (j + 1 < n) next = P.at(j + 1);
        BVert f = cur, b = cur;
        if (j > cl) {  // linkVertex(j, false)
            const BVert& e = j + 1 < n ? next : cur;
            f.d = normalize(sub(e.o, cur.o));
            f.inside = e.inside;
        }
        if (j <= cl) {  // linkVertex(j, true)
            const BVert& e = j > 0 ? prev : cur;
            b.d = normalize(sub(e.o, cur.o));
            b.inside = e.inside;
        }
        cos_f[j] = cos_term_v(f);
        cos_b[j] = cos_term_v(b);
        dp_f[j] = eval_dir_prob(S, cam, pixel, f, f, j == 0);
        dp_b[j] = eval_dir_prob(S, cam, pixel, b, b, j == n - 1);
        delta[j] = (cur.flags & kVDelta) ? 1 : 0;
        vol[j] = in_volume(cur) ? 1 : 0;
        op[j] = cur.op;
        if (j > 0) {
            const float l = len(sub(cur.o, prev.o));
            dist[j] = (double)l > kEps ? l : (float)kEps;
        }
        // connectingColorProb
        color = mulv(color, cur.rad);
        if (j == cl && j < n - 1) {
            const float l = len(sub(cur.o, next.o));
            const float d = (double)l > kEps ? l : (float)kEps;
            color = muls(color, cos_term_v(cur) * cos_term_v(next) / (d * d));
        }
        if (j != cl && j != cl + 1) color = muls(color, cos_term_v(cur));
        prob *= cur.dp * cur.op;
        if (vt && vt->vm && j == cl) {  // a merge: cl + 1 (the eye-side vertex) always exists
            const float l = len(sub(cur.o, next.o));
            const float d = (double)l > kEps ? l : (float)kEps;
            prob = (float)((double)prob * (kPi * (double)vt->r * (double)vt->r));
            prob *= 1.0f * fabsf(cos_term_v(next)) / (d * d);  // PdfWtoA at the eye-side vertex
            if (in_volume(next)) prob = (float)((double)prob * (4.0 / 3 * (double)vt->r));
        }
        prev = cur;
        cur = next;
    }
    if (prob == 0.f || len(color) == 0.f) return 0.f;  // the caller skips these
    double direct[kBdptMaxJoined], reverse[kBdptMaxJoined];
    direct[0] = op[0];
    for (int i = 1; i < n; i++) {
        const float d = dist[i];
        direct[i] = direct[i - 1] * (double)(1.0f * fabsf(cos_b[i]) / (d * d));  // PdfWtoA
        if (delta[i - 1]) continue;
        direct[i] *= dp_f[i - 1] * 1.f;
    }
    reverse[n - 1] = op[n - 1];
    for (int i = n - 2; i >= 0; i--) {
        const float d = dist[i + 1];
        reverse[i] = reverse[i + 1] * (double)(1.0f * fabsf(cos_f[i]) / (d * d));
        if (delta[i + 1]) continue;
        reverse[i] *= dp_b[i + 1] * 1.f;
    }
    double all = 0.0;
    bool cannot = true;  // pathCanNotConnect (bidirPathTracer.cpp:251-257)
    if (!vt) {
        for (int i = 0; i < n - 1; i++) {
            if (!delta[i] && !delta[i + 1]) {
                all += direct[i] * reverse[i + 1];
                cannot = false;
            }
        }
        if (cannot) all += reverse[0];
        const double self = cl == -1 ? reverse[0] : direct[cl] * reverse[cl + 1];
        const double w = self / all;
        return (float)(w > 0.0 ? w : 0.0);
    }
    double self = cl == -1 ? reverse[0] : direct[cl] * reverse[cl + 1];
    for (int i = 0; i < n - 1; i++) {
        if (!delta[i] && !delta[i + 1]) {  // VC
            all += direct[i] * reverse[i + 1] * 1.0;
            cannot = false;
        }
        if (i < n - 2 && !delta[i + 1]) {  // VM at vertex i + 1
            double p = direct[i] * reverse[i + 1];
            if (!delta[i]) p *= dp_f[i];
            p *= 1.f;  // evalOriginProbability of two temporaries
            p *= (double)(1.0f * fabsf(cos_b[i + 1]) / (dist[i + 1] * dist[i + 1]));
            p *= kPi * (double)vt->r * (double)vt->r;
            if (vol[i + 1]) p *= 4.0 / 3 * (double)vt->r;
            if (vt->vm && i == cl) self = p;
            all += p * vt->nvm;
        }
    }
    if (cannot) all += reverse[0];
    self *= vt->vm ? vt->nvm : 1.0;
    const double w = self / all;
    if (!(w >= 0.0)) return 0.f;
    return (float)w;
}

// connectEdge for a merge (vcmTracer.cpp:366-398 with VM): both directions are
// re-aimed, only the eye-side record is edited (radiance, probability, delta,
// inside); its previous record is an eye record (the merged eye vertex is at
// least the second of its path).
__device__ bool connect_edge_vm(const SceneDev& S, Joined& P) {
    P.lc = P.at(P.cl);
    P.ec = P.at(P.cl + 1);
    if (P.ec.flags & kVDelta) return false;
    P.edited = true;
    const int last = P.n - 1;
    const int ie = P.cl + 2 < last ? P.cl + 2 : last;
    P.lc.d = normalize(sub(P.ec.o, P.lc.o));
    P.ec.d = mk(-P.lc.d.x, -P.lc.d.y, -P.lc.d.z);
    P.ec.dp = 1.f;
    P.ec.flags &= ~kVDelta;
    const BVert pre = P.at(ie);
    P.ec.rad = eval_bsdf(S, pre, P.ec);
    if (P.ec.contact != WR_NONE && P.cl < last - 1)
        P.ec.inside = dot(P.ec.n, P.ec.d) > 0.f ? pre.inside : P.ec.contact;
    return true;
}

__device__ __forceinline__ bool emissive(const SceneDev& S, uint32_t obj) {
    return obj != WR_NONE && S.objs[obj].mat == WR_MAT_EMISSIVE;
}

// One slot of a queue per active lane, one atomic per wave.
__device__ __forceinline__ uint32_t wave_slot(uint32_t* counter) {
    const uint64_t mask = __ballot(1);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t leader = __ffsll((unsigned long long)mask) - 1;
    const uint32_t rank = __popcll(mask & ((1ull << lane) - 1));
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + rank;
}

// The visibility queue: rays in the batch NoSelf layout (o, d, inside,
// contact), per entry the contribution and the distance to the light-side
// record, and the target (pass buffer offset + pixel).
struct VisQueue {
    uint32_t* rays;   // 8 words per entry
    float* pay;       // 4 floats per entry: contribution, distance
    uint32_t* dst;    // pixel index into the pass buffers (k * npix + q)
    uint32_t* count;
    uint32_t cap;
};

__device__ __forceinline__ void atomic_add3(float* buf, uint32_t p, f3 v) {
    atomicAdd(buf + 3 * p + 0, v.x);
    atomicAdd(buf + 3 * p + 1, v.y);
    atomicAdd(buf + 3 * p + 2, v.z);
}

// The two sub-paths of one pixel and pass (kept apart from the connections:
// this half traverses at a register count that lets several waves share a
// SIMD); counts[2 * lane] = eye records, counts[2 * lane + 1] = light records.
// 4 waves per SIMD (128 VGPRs, the rest spilled) for both kernels: 80 -> 43 ms on the
// Cornell box against the unconstrained 264 / 190 registers.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bdpt_subpaths(SceneDev S, CamDev cam, uint32_t max_depth, uint64_t seed,
                                                       uint32_t pass0, uint32_t npass, uint32_t npix, uint32_t maxv,
                                                       BVert* verts, uint32_t* counts, bool eye = true) {
    const uint64_t lane = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (lane >= (uint64_t)npass * npix) return;
    const uint32_t k = (uint32_t)(lane / npix), p = (uint32_t)(lane % npix);
    const uint32_t pass = pass0 + k;
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    BVert* E = verts + lane * 2 * maxv;
    BVert* L = E + maxv;
    uint32_t ne = 0, nl = 0;
    float rn[4];
    // eye sub-path: the camera ray (camera.cpp:6-60) then samplePath (PhotonMap forms its own)
    if (eye) {
        rng4(seed, p, pass, 0, rn);
        BVert c;
        c.o = cam.pos;
        c.d = cam_dir(cam, p, rn[0], rn[1]);
        c.rad = mk(1.f, 1.f, 1.f);
        c.n = cam.front;
        c.op = 1.f;
        c.dp = 1.f;
        c.contact = WR_NONE;
        c.flags = kVCamera;
        c.pad = 0;
        const PHit h = trace_scene<false>(sc, c.o, c.d, WR_NONE, WR_NONE, st);
        c.inside = find_inside(S, c.o, c.d, WR_NONE, st);
        walk(sc, c, h, max_depth, seed, p, pass, 0u, E, ne, maxv, st);
    }
    // light sub-path (as k_light_paths)
    {
        rng4(seed, p, pass, kBdptLightVertex, rn);
        const LightSample ls = sample_light(sc, rn);
        BVert r;
        r.o = mk(0.f, 0.f, 0.f);
        r.d = r.o;
        r.rad = r.o;
        r.n = r.o;
        r.op = 1.f;
        r.dp = 1.f;
        r.inside = WR_NONE;
        r.contact = WR_NONE;
        r.flags = 0;
        r.pad = 0;
        PHit hit;
        hit.t = -1.f;
        hit.rec = WR_NONE;
        hit.obj = WR_NONE;
        if (ls.ok) {
            rng4(seed, p, pass, kBdptLightVertex + 1, rn);
            const float term1 = (float)(2.0 * kPi * (double)rn[0]);  // uniformSphereSampler.h:6-20
            const float term2 = 2.f * sqrtf(rn[1] - rn[1] * rn[1]);
            f3 d = mk(cosf(term1) * term2, 1.f - 2.f * rn[1], sinf(term1) * term2);
            r.dp = (float)(kInvPi * 0.25f);
            if (dot(d, ls.n) < 0.f) d = mk(-d.x, -d.y, -d.z);
            r.o = ls.o;
            r.n = ls.n;
            r.d = d;
            r.rad = ls.rad;
            r.contact = ls.obj;
            r.inside = find_inside(S, r.o, r.d, ls.obj, st);
            r.dp *= 2.f;
            r.op = ls.op;
            const PHit h = trace_scene<true>(sc, r.o, r.d, r.inside, r.contact, st);
            if (h.rec != WR_NONE && h.t > 0.f) hit = h;
        }
        walk(sc, r, hit, max_depth, seed, p, pass, kBdptLightVertex + 1, L, nl, maxv, st);
    }
    counts[2 * lane] = ne;
    counts[2 * lane + 1] = nl;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bdpt(SceneDev S, CamDev cam, uint32_t max_depth, uint64_t seed,
                                              uint32_t pass0, uint32_t npass, uint32_t npix, uint32_t maxv,
                                              BVert* verts, const uint32_t* counts, float* pix_all, VisQueue Q) {
    (void)max_depth;
    (void)seed;
    (void)pass0;
    const uint64_t lane = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (lane >= (uint64_t)npass * npix) return;
    const uint32_t k = (uint32_t)(lane / npix), p = (uint32_t)(lane % npix);
    float* pix = pix_all + 3 * (size_t)k * npix;
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    BVert* E = verts + lane * 2 * maxv;
    BVert* L = E + maxv;
    const uint32_t ne = counts[2 * lane], nl = counts[2 * lane + 1];
    const int max_eye = (int)(ne < maxv ? ne : maxv), max_light = (int)(nl < maxv ? nl : maxv);
    f3 own = mk(0.f, 0.f, 0.f);
    for (int whole = 2; whole <= max_eye + max_light; whole++) {
        const int min_l = whole - max_eye > 0 ? whole - max_eye : 0;
        const int max_l = whole - 1 >= max_light ? max_light : whole - 1;
        for (int light_len = min_l; light_len <= max_l; light_len++) {
            Joined P;
            P.L = L;
            P.E = E;
            P.nl = light_len;
            P.ne = whole - light_len;
            P.n = whole;
            P.cl = light_len - 1;
            P.edited = false;
            if (!emissive(S, P.at(0).contact)) continue;
            if (light_len == 0 && !cannot_connect(P)) continue;
            const bool connected = P.ne > 0 && light_len > 0;
            if (connected) {
                if (light_len > 1 && emissive(S, P.at(P.cl).contact)) continue;
                if (emissive(S, P.at(P.cl + 1).contact)) continue;
                if (!connect_edge(S, P)) continue;
            }
            f3 color;
            float prob;
            const float w = color_prob_weight(S, cam, p, P, color, prob);
            if (prob == 0.f) continue;
            if (len(color) == 0.f) continue;
            f3 contrib = muls(divs(color, prob), w);
            uint32_t target = p;
            if (P.ne <= 1) {  // light tracing: splat to the pixel the camera sees the point in
                const f3 v = sub(P.at(whole - 2).o, cam.pos);
                const float vfd = dot(v, cam.front);
                const float scale = cam.dist / vfd;
                const float rx = (float)(cam.W / 2.0 + (double)(dot(v, cam.right) * scale));
                const float ry = (float)(cam.H / 2.0 - (double)(dot(v, cam.top) * scale));
                const int x = (int)rx, y = (int)ry;  // checkPixelBound truncates (camera.cpp:164-170)
                if (!(x >= 0 && y >= 0 && x < (int)cam.W && y < (int)cam.H)) continue;
                target = (uint32_t)(y * (int)cam.W + x);
                const float factor = (float)(1.0 / (double)pixel_solid_angle(cam, target));
                contrib = muls(contrib, factor / (float)npix);
            }
            if (!connected) {  // no visibility test (a path found by the eye side alone)
                if (target == p) own = add(own, contrib);
                else atomic_add3(pix, target, contrib);
                continue;
            }
            // MCRenderer::visibilityTest from the eye-side connection record
            const uint32_t contact = (P.ec.flags & kVCamera) ? kCameraObj : P.ec.contact;
            const float dist = len(sub(P.ec.o, P.lc.o));
            const uint32_t slot = wave_slot(Q.count);
            if (slot < Q.cap) {
                uint32_t* r = Q.rays + 8 * (size_t)slot;
                r[0] = __float_as_uint(P.ec.o.x); r[1] = __float_as_uint(P.ec.o.y); r[2] = __float_as_uint(P.ec.o.z);
                r[3] = __float_as_uint(P.ec.d.x); r[4] = __float_as_uint(P.ec.d.y); r[5] = __float_as_uint(P.ec.d.z);
                r[6] = P.ec.inside;
                r[7] = contact;
                float* y = Q.pay + 4 * (size_t)slot;
                y[0] = contrib.x; y[1] = contrib.y; y[2] = contrib.z; y[3] = dist;
                Q.dst[slot] = k * npix + target;
            } else {  // queue full: trace here
                const PHit h = trace_scene<true>(sc, P.ec.o, P.ec.d, P.ec.inside, contact, st);
                if (!nee_visible(h.rec, h.t, dist)) continue;
                if (target == p) own = add(own, contrib);
                else atomic_add3(pix, target, contrib);
            }
        }
    }
    atomic_add3(pix, p, own);
}

// The queued visibility tests, answered by the batch NoSelf query (hits:
// prim WR_NONE on a miss, t): visible connections add their contribution.
__global__ __launch_bounds__(256) void k_bdpt_accum(uint32_t n, const float* pay, const uint32_t* dst,
                                                    const uint4* hits, float* pix_all) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint4 h = hits[i];
    const float* y = pay + 4 * (size_t)i;
    if (!nee_visible(h.x == WR_NONE ? WR_NONE : 0u, __uint_as_float(h.y), y[3])) return;
    atomic_add3(pix_all, dst[i], mk(y[0], y[1], y[2]));
}

// Each pass's pixel values, in pass order: isLegalColor, fixVignetting, the Kahan film.
__global__ __launch_bounds__(256) void k_bdpt_fold(CamDev cam, uint32_t npass, uint32_t npix, const float* pix,
                                                   FilmDev film) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    for (uint32_t k = 0; k < npass; k++) {
        const float* c = pix + 3 * ((size_t)k * npix + p);
        f3 v = mk(c[0], c[1], c[2]);
        if (!legal(v)) v = mk(0.f, 0.f, 0.f);
        film_add(film, p, fix_vignetting(cam, v, p));
    }
}

// ---- VCMTracer (vcmTracer.cpp:6-446) ---------------------------------------
// Several passes side by side (lane = pass x pixel, as k_bdpt), each with its
// own light points, radius and grid: the sub-paths of k_bdpt_subpaths (one
// light path per pixel); per pass the light points (vcmTracer.cpp:42-55: the
// pixel count of default-constructed ones, then every non-delta vertex past
// the first of each light path, up to an emitter, path-major) in a block of
// their own; HashGrid::Build over each block; then k_vcm, one lane per pixel
// and pass: the connections of k_bdpt with vcmMisWeight, and the merges.

// true: vertex i of a light path is a light point; stop: the path's points end there
__device__ __forceinline__ bool light_point(const SceneDev& S, const BVert& v, bool& stop) {
    stop = emissive(S, v.contact);
    return !stop && !(v.flags & kVDelta);
}

__global__ __launch_bounds__(256) void k_vcm_point_count(SceneDev S, const BVert* verts, const uint32_t* counts,
                                                         uint32_t lanes, uint32_t maxv, uint32_t* pc) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= lanes) return;
    const BVert* L = verts + (size_t)k * 2 * maxv + maxv;
    const uint32_t nl = min(counts[2 * k + 1], maxv);
    uint32_t m = 0;
    for (uint32_t i = 1; i < nl; i++) {
        bool stop;
        const bool keep = light_point(S, L[i], stop);
        if (stop) break;
        m += keep;
    }
    pc[k] = m;
}

// Pass b's block of points starts at b * npix + offs[b * npix] (offs: exclusive
// sums of the counts over all lanes): the npix default points (origin, no
// path), then the points of the pass's paths; src holds the lane in the batch
// and the vertex.
__global__ __launch_bounds__(256) void k_vcm_point_write(SceneDev S, const BVert* verts, const uint32_t* counts,
                                                         uint32_t npix, uint32_t lanes, uint32_t maxv,
                                                         const uint32_t* offs, float* pos, uint32_t* src) {
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= lanes) return;
    const uint32_t b = k / npix, p = k % npix;
    const size_t block = (size_t)b * npix + offs[(size_t)b * npix];
    pos[3 * (block + p)] = 0.f;
    pos[3 * (block + p) + 1] = 0.f;
    pos[3 * (block + p) + 2] = 0.f;
    src[2 * (block + p)] = WR_NONE;
    src[2 * (block + p) + 1] = WR_NONE;
    const BVert* L = verts + (size_t)k * 2 * maxv + maxv;
    const uint32_t nl = min(counts[2 * k + 1], maxv);
    size_t j = block + npix + (offs[k] - offs[(size_t)b * npix]);
    for (uint32_t i = 1; i < nl; i++) {
        bool stop;
        const bool keep = light_point(S, L[i], stop);
        if (stop) break;
        if (!keep) continue;
        pos[3 * j] = L[i].o.x;
        pos[3 * j + 1] = L[i].o.y;
        pos[3 * j + 2] = L[i].o.z;
        src[2 * j] = k;
        src[2 * j + 1] = i;
        j++;
    }
}

// The first offset of every pass: starts[b] = offs[b * npix], b = 0..npass.
__global__ void k_vcm_pass_starts(const uint32_t* offs, uint32_t npix, uint32_t npass, uint32_t* starts) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b <= npass) starts[b] = offs[(size_t)b * npix];
}

// One pass's grid: pos / src point at the pass's block, idx indexes it.
struct VcmGrid {
    GridDev g;
    const int* ends;
    const uint32_t* idx;
    const float* pos;
    const uint32_t* src;
    float r;
};

// The merges of the VCM lanes, one entry per light point in reach of an eye
// vertex: word 0 = lane | eye vertex << 24 (lanes < 2^22), word 1 = the light
// point; count at ent[2 * cap].  A separate kernel takes one merge per lane, so
// the uneven number of merges per pixel does not leave most lanes of a wave idle.
struct MergeQueue {
    uint32_t* ent;
    uint32_t cap;
};

// One merge (throughputByMerging, vcmTracer.cpp:196-213): the eye vertex i of
// lane `lane` with light point pi of the lane's pass.  Returns the contribution
// (before the division by the pixel count); zero when the merge gives nothing.
__device__ __forceinline__ f3 vcm_merge(const SceneDev& S, const CamDev& cam, uint32_t npix, uint32_t maxv,
                                        const BVert* verts, const VcmGrid& G, uint64_t lane, int i, uint32_t pi) {
    const uint32_t p = (uint32_t)(lane % npix);
    const uint32_t k = G.src[2 * (size_t)pi], j = G.src[2 * (size_t)pi + 1];
    Joined P;
    P.L = verts + (size_t)k * 2 * maxv + maxv;
    P.E = verts + lane * 2 * maxv;
    P.nl = (int)j;
    P.ne = i + 1;
    P.n = (int)j + i + 1;
    P.cl = (int)j - 1;
    P.edited = false;
    if (!connect_edge_vm(S, P)) return mk(0.f, 0.f, 0.f);
    const VcmTerms vm{G.r, (double)npix, true};
    f3 color;
    float prob;
    const float w = color_prob_weight(S, cam, p, P, color, prob, &vm);
    if (len(color) == 0.f || prob == 0.f) return mk(0.f, 0.f, 0.f);
    return divs(muls(color, w), prob);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_vcm(SceneDev S, CamDev cam, uint32_t npix, uint32_t npass, uint32_t maxv,
                                             const BVert* verts, const uint32_t* counts, float* pix_all, VisQueue Q,
                                             const VcmGrid* grids, MergeQueue M) {
    // every lane of a wave stays to the end (the merge reservations below are
    // wave-wide scans); a lane past the last one stands in for lane 0 and does
    // nothing
    const uint64_t lane_id = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool valid = lane_id < (uint64_t)npass * npix;
    const uint64_t lane = valid ? lane_id : 0;
    const uint32_t kb = (uint32_t)(lane / npix), p = (uint32_t)(lane % npix);
    const VcmGrid G = grids[kb];
    float* pix = pix_all + 3 * (size_t)kb * npix;
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    const BVert* E = verts + lane * 2 * maxv;
    const BVert* L = E + maxv;
    const uint32_t ne = counts[2 * lane], nl = counts[2 * lane + 1];
    const int max_eye = (int)(ne < maxv ? ne : maxv), max_light = (int)(nl < maxv ? nl : maxv);
    const VcmTerms vc{G.r, (double)npix, false};
    f3 own = mk(0.f, 0.f, 0.f);
    // throughputByConnecting (vcmTracer.cpp:88-173): k_bdpt's loop with vcmMisWeight
    for (int whole = 2; valid && whole <= max_eye + max_light; whole++) {
        const int min_l = whole - max_eye > 0 ? whole - max_eye : 0;
        const int max_l = whole - 1 >= max_light ? max_light : whole - 1;
        for (int light_len = min_l; light_len <= max_l; light_len++) {
            Joined P;
            P.L = L;
            P.E = E;
            P.nl = light_len;
            P.ne = whole - light_len;
            P.n = whole;
            P.cl = light_len - 1;
            P.edited = false;
            if (!emissive(S, P.at(0).contact)) continue;
            if (light_len == 0 && !cannot_connect(P)) continue;
            const bool connected = P.ne > 0 && light_len > 0;
            if (connected) {
                if (light_len > 1 && emissive(S, P.at(P.cl).contact)) continue;
                if (emissive(S, P.at(P.cl + 1).contact)) continue;
                if (!connect_edge(S, P)) continue;
            }
            f3 color;
            float prob;
            const float w = color_prob_weight(S, cam, p, P, color, prob, &vc);
            if (prob == 0.f) continue;
            if (len(color) == 0.f) continue;
            f3 contrib = muls(divs(color, prob), w);
            uint32_t target = p;
            if (P.ne <= 1) {  // light tracing: splat to the pixel the camera sees the point in
                const f3 v = sub(P.at(whole - 2).o, cam.pos);
                const float vfd = dot(v, cam.front);
                const float scale = cam.dist / vfd;
                const float rx = (float)(cam.W / 2.0 + (double)(dot(v, cam.right) * scale));
                const float ry = (float)(cam.H / 2.0 - (double)(dot(v, cam.top) * scale));
                const int x = (int)rx, y = (int)ry;  // checkPixelBound truncates (camera.cpp:164-170)
                if (!(x >= 0 && y >= 0 && x < (int)cam.W && y < (int)cam.H)) continue;
                target = (uint32_t)(y * (int)cam.W + x);
                const float factor = (float)(1.0 / (double)pixel_solid_angle(cam, target));
                contrib = muls(contrib, factor / (float)npix);
            }
            if (!connected) {
                if (target == p) own = add(own, contrib);
                else atomic_add3(pix, target, contrib);
                continue;
            }
            const uint32_t contact = (P.ec.flags & kVCamera) ? kCameraObj : P.ec.contact;
            const float dist = len(sub(P.ec.o, P.lc.o));
            const uint32_t slot = wave_slot(Q.count);
            if (slot < Q.cap) {
                uint32_t* r = Q.rays + 8 * (size_t)slot;
                r[0] = __float_as_uint(P.ec.o.x); r[1] = __float_as_uint(P.ec.o.y); r[2] = __float_as_uint(P.ec.o.z);
                r[3] = __float_as_uint(P.ec.d.x); r[4] = __float_as_uint(P.ec.d.y); r[5] = __float_as_uint(P.ec.d.z);
                r[6] = P.ec.inside;
                r[7] = contact;
                float* y = Q.pay + 4 * (size_t)slot;
                y[0] = contrib.x; y[1] = contrib.y; y[2] = contrib.z; y[3] = dist;
                Q.dst[slot] = kb * npix + target;
            } else {  // queue full: trace here
                const PHit h = trace_scene<true>(sc, P.ec.o, P.ec.d, P.ec.inside, contact, st);
                if (!nee_visible(h.rec, h.t, dist)) continue;
                if (target == p) own = add(own, contrib);
                else atomic_add3(pix, target, contrib);
            }
        }
    }
    // throughputByMerging (vcmTracer.cpp:175-217): the light points within the
    // radius of every non-delta eye vertex before an emitter, in the grid's
    // order (HashGrid::Process); the default points have no path and give
    // nothing.  Each merge is queued for k_vcm_merge: per eye vertex the lane
    // counts its points in reach, the wave reserves their slots with one
    // atomic (an exclusive scan of the counts), and the lane writes them.
    const uint32_t wl = threadIdx.x & 63u;
    uint32_t* mcount = M.ent + 2 * (size_t)M.cap;
    bool alive = valid;
    for (int i = 1; i < (int)maxv; i++) {
        bool act = alive && i < max_eye;
        BVert ev;
        if (act) {
            ev = E[i];
            if (emissive(S, ev.contact)) {
                alive = false;
                act = false;
            } else if (ev.flags & kVDelta) {
                act = false;
            }
        }
        if (i + 1 >= max_eye) alive = false;
        int pc[6];
        const bool q = act && query_corner(G.g, ev.o, pc);
        uint32_t cnt = 0;
        for (int pass2 = 0; pass2 < 2; pass2++) {
            uint32_t k = 0;
            if (pass2) {  // the slots of this lane's merges: the wave's scan, one atomic
                uint32_t v = cnt;
                for (int off = 1; off < 64; off <<= 1) {
                    const uint32_t t = __shfl_up(v, off, 64);
                    if (wl >= (uint32_t)off) v += t;
                }
                const uint32_t total = __shfl(v, 63, 64);
                uint32_t base = 0;
                if (wl == 0 && total) base = atomicAdd(mcount, total);
                k = __shfl(base, 0, 64) + (v - cnt);
                if (cnt == 0) break;
            }
            f3 contrib = mk(0.f, 0.f, 0.f);
            for (int c8 = 0; q && c8 < 8; c8++) {
                const uint32_t cell = query_cell(G.g, pc, c8);
                int a = cell == 0 ? 0 : G.ends[cell - 1];
                const int b = G.ends[cell];
                // the default points lead their cell's range (the sort keeps input order in a cell)
                int hi = b;
                while (a < hi) {
                    const int mid = (a + hi) >> 1;
                    if (G.idx[mid] < npix) a = mid + 1;
                    else hi = mid;
                }
                for (; a < b; a++) {
                    const uint32_t pi = G.idx[a];
                    if (!grid_near(G.g, G.pos, ev.o, pi)) continue;
                    if (!pass2) {
                        cnt++;
                    } else if (k < M.cap) {
                        M.ent[2 * (size_t)k] = (uint32_t)lane | ((uint32_t)i << 24);
                        M.ent[2 * (size_t)k + 1] = pi;
                        k++;
                    } else {  // queue full: merge here
                        contrib = add(contrib, vcm_merge(S, cam, npix, maxv, verts, G, lane, i, pi));
                        k++;
                    }
                }
            }
            if (pass2) own = add(own, divs(contrib, (float)npix));
        }
        if (__ballot(alive) == 0) break;
    }
    if (valid) atomic_add3(pix, p, own);
}

// The queued merges of k_vcm, one per lane, each added to its pixel of its pass.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_vcm_merge(SceneDev S, CamDev cam, uint32_t npix, uint32_t maxv,
                                                   const BVert* verts, float* pix_all, MergeQueue M,
                                                   const VcmGrid* grids) {
    const uint32_t n = min(M.ent[2 * (size_t)M.cap], M.cap);
    for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < n; e += gridDim.x * 256) {
        const uint32_t w0 = M.ent[2 * (size_t)e], pi = M.ent[2 * (size_t)e + 1];
        const uint32_t lane = w0 & 0xFFFFFFu;
        const int i = (int)(w0 >> 24);
        const uint32_t kb = lane / npix, p = lane % npix;
        const f3 c = vcm_merge(S, cam, npix, maxv, verts, grids[kb], lane, i, pi);
        if (c.x != 0.f || c.y != 0.f || c.z != 0.f) atomic_add3(pix_all + 3 * (size_t)kb * npix, p, divs(c, (float)npix));
    }
}

// ---- PhotonMap (photonMap.cpp:21-680), progressive photon mapping ----------
// The light points and per-pass grids of VCMTracer (the same light paths);
// k_ppm, one lane per pixel and pass: the eye "merge path" (samplePath with
// maxPathLen MERGE_LEN = 10, photonMap.cpp:10,150-190) up to its first
// emitter (its emitted radiance) or first non-specular, i.e. diffuse, vertex
// (the kernel density estimate of the photons around it, photonMap.cpp:
// 198-248,656-668), times the camera vertex's factor only (photonMap.cpp:
// 670-674).  No object here is a volume or an anisotropic glossy object, so
// every photon is a surface photon and every estimate the plain range query.
constexpr uint32_t kMergeLen = 10;

// Query::Process for the photons around outRay (the eye vertex, its direction
// towards the previous eye vertex).
__device__ f3 ppm_density(const SceneDev& S, const VcmGrid& G, const BVert* verts, uint32_t maxv, uint32_t npix,
                          const BVert& outRay) {
    f3 contrib = mk(0.f, 0.f, 0.f);
    int pc[6];
    if (!query_corner(G.g, outRay.o, pc)) return contrib;
    const float r = G.r;
    for (int c8 = 0; c8 < 8; c8++) {
        const uint32_t cell = query_cell(G.g, pc, c8);
        int a = cell == 0 ? 0 : G.ends[cell - 1];
        const int b = G.ends[cell];
        int hi = b;  // the default points lead their cell's range
        while (a < hi) {
            const int mid = (a + hi) >> 1;
            if (G.idx[mid] < npix) a = mid + 1;
            else hi = mid;
        }
        for (; a < b; a++) {
            const uint32_t pi = G.idx[a];
            if (!grid_near(G.g, G.pos, outRay.o, pi)) continue;
            const uint32_t k = G.src[2 * (size_t)pi], j = G.src[2 * (size_t)pi + 1];
            const BVert* L = verts + (size_t)k * 2 * maxv + maxv;
            f3 thr = mk(1.f, 1.f, 1.f);
            for (uint32_t m = 0; m < j; m++) {
                const BVert lm = L[m];
                thr = mulv(thr, divs(divs(lm.rad, lm.dp), lm.op));
                thr = muls(thr, cos_term_v(lm));
            }
            BVert ph = L[j];
            thr = divs(thr, ph.op);
            ph.d = L[j - 1].d;
            const f3 col = mulv(thr, eval_bsdf(S, ph, outRay));
            const float l = len(sub(outRay.o, ph.o));
            const float dist_sqr = l * l;  // powf(l, 2): the correctly rounded square
            if (!legal(col)) continue;
            const float q = 1.f - dist_sqr / (r * r);
            const float sk = 0.f > q ? 0.f : q;
            const float kernel = (float)((double)(3.f * sk * sk) / kPi);  // Kernel (photonMap.cpp:209-212)
            const float norm = kernel / ((float)npix * r * r);
            contrib = add(contrib, muls(col, norm));
        }
    }
    return contrib;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_ppm(SceneDev S, CamDev cam, uint32_t npix, uint32_t npass, uint32_t maxv,
                                             const BVert* verts, uint64_t seed, uint32_t pass0, float* pix_all,
                                             const VcmGrid* grids) {
    const uint64_t lane = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (lane >= (uint64_t)npass * npix) return;
    const uint32_t kb = (uint32_t)(lane / npix), p = (uint32_t)(lane % npix);
    const uint32_t pass = pass0 + kb;
    const VcmGrid G = grids[kb];
    StackSpill spill;
    LaneStack st(&spill);
    const SceneCtx sc = make_scene_ctx(S);
    float rn[4];
    rng4(seed, p, pass, 0, rn);
    BVert c;
    c.o = cam.pos;
    c.d = cam_dir(cam, p, rn[0], rn[1]);
    c.rad = mk(1.f, 1.f, 1.f);
    c.n = cam.front;
    c.op = 1.f;
    c.dp = 1.f;
    c.contact = WR_NONE;
    c.flags = kVCamera;
    c.pad = 0;
    PHit hit = trace_scene<false>(sc, c.o, c.d, WR_NONE, WR_NONE, st);
    c.inside = find_inside(S, c.o, c.d, WR_NONE, st);
    f3 surface = mk(0.f, 0.f, 0.f);
    f3 o = c.o, d = c.d;
    uint32_t inside = c.inside;
    for (uint32_t depth = 0;; depth++) {
        const uint32_t self = inside != WR_NONE ? inside : hit.obj;
        if (self == WR_NONE || hit.rec == WR_NONE) break;  // the terminating record: no contact
        rng4(seed, p, pass, depth + 1, rn);
        const Out r = scatter(sc, self, o, d, hit, inside, rn, st);
        const bool absorbed = (double)len(r.d) < 0.5;
        if (!absorbed && depth + 1 > kMergeLen) break;  // the terminating record
        if (emissive(S, r.contact)) {  // the eye path hit a light
            surface = r.rad;
            break;
        }
        if (r.contact != WR_NONE && S.objs[r.contact].mat == WR_MAT_DIFFUSE) {  // nonSpecular()
            BVert out;
            out.o = r.o;
            out.d = mk(-d.x, -d.y, -d.z);
            out.rad = r.rad;
            out.n = r.n;
            out.op = r.op;
            out.dp = r.dp;
            out.inside = r.inside;
            out.contact = r.contact;
            out.flags = 0;
            out.pad = 0;
            surface = ppm_density(S, G, verts, maxv, npix, out);
            break;
        }
        if (absorbed) break;
        const PHit nh = trace_scene<true>(sc, r.o, r.d, r.inside, r.contact, st);
        if (nh.rec == WR_NONE) break;
        o = r.o;
        d = r.d;
        inside = r.inside;
        hit = nh;
    }
    const f3 color = mulv(mulv(mk(1.f, 1.f, 1.f), surface), divs(divs(muls(c.rad, cos_term_v(c)), c.dp), c.op));
    float* px = pix_all + 3 * lane;
    px[0] = color.x;
    px[1] = color.y;
    px[2] = color.z;
}

}  // namespace wrd

namespace wr {

size_t bdpt_vertex_bytes() { return sizeof(wrd::BVert); }

float vcm_radius(float base, uint32_t pass) {
    const float alpha = (float)(2.0 / 3);  // vcmTracer.h:39
    const float r = base * std::sqrt(std::pow((float)(pass + 1), alpha - 1.f));
    return (double)r > 1e-5 ? r : (float)1e-5;
}

static wrd::CamDev make_cam_dev(const CamHost& ch) {
    wrd::CamDev cam;
    cam.pos = wrd::f3{ch.pos[0], ch.pos[1], ch.pos[2]};
    cam.front = wrd::f3{ch.front[0], ch.front[1], ch.front[2]};
    cam.right = wrd::f3{ch.right[0], ch.right[1], ch.right[2]};
    cam.top = wrd::f3{ch.top[0], ch.top[1], ch.top[2]};
    cam.dist = ch.dist;
    cam.W = ch.W;
    cam.H = ch.H;
    return cam;
}

// The light sub-paths of passes [pass0, pass0 + npass) (and their eye sub-paths
// when eye), their light points and one grid per pass, uploaded to vb.grids;
// the grids go to `grids` (the caller destroys them).
static int light_point_grids(wr_ctx* ctx, const SceneDev& S, const wrd::CamDev& cam, uint32_t npix, uint32_t max_depth,
                             uint64_t seed, uint32_t pass0, uint32_t npass, float base_radius, uint32_t maxv,
                             wrd::BVert* V, uint32_t* counts, const VcmBufs& vb, bool eye,
                             std::vector<wr_hashgrid*>& grids, const char* who) {
    hipStream_t st = ctx->stream;
    const uint32_t lanes = npass * npix;  // the caller keeps npass * npix below 2^31
    const uint32_t blocks = (lanes + 255) / 256;
    hipError_t e = hipMemsetAsync(vb.pc + lanes, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) return hip_fail(ctx, e, who);
    wrd::k_bdpt_subpaths<<<blocks, 256, wrd::kStackLdsBytes, st>>>(S, cam, max_depth, seed, pass0, npass, npix, maxv, V,
                                                                  counts, eye);
    wrd::k_vcm_point_count<<<blocks, 256, 0, st>>>(S, V, counts, lanes, maxv, vb.pc);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(ctx, e, who);
    size_t tb = vb.scan_bytes;
    if ((e = hipcub::DeviceScan::ExclusiveSum(vb.scan, tb, vb.pc, vb.offs, (int)lanes + 1, st)) != hipSuccess)
        return hip_fail(ctx, e, who);
    wrd::k_vcm_pass_starts<<<(npass + 1 + 63) / 64, 64, 0, st>>>(vb.offs, npix, npass, vb.starts);
    std::vector<uint32_t> starts(npass + 1);
    if ((e = hipGetLastError()) != hipSuccess ||
        (e = hipMemcpyAsync(starts.data(), vb.starts, sizeof(uint32_t) * (npass + 1), hipMemcpyDeviceToHost, st)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(ctx, e, who);
    if ((uint64_t)lanes + starts[npass] > vb.cap)
        return fail(ctx, WR_EINVAL, std::string(who) + ": more light points than the buffer holds");
    wrd::k_vcm_point_write<<<blocks, 256, 0, st>>>(S, V, counts, npix, lanes, maxv, vb.offs, vb.pos, vb.src);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(ctx, e, who);
    // per pass: HashGrid::Reserve(pixels) + Build(its light points, its radius)
    // (vcmTracer.cpp:30,33,57; photonMap.cpp:48,56,93)
    grids.assign(npass, nullptr);
    std::vector<wrd::VcmGrid> G(npass);
    for (uint32_t b = 0; b < npass; b++) {
        const uint64_t block = (uint64_t)b * npix + starts[b];
        const uint64_t n = (uint64_t)npix + (starts[b + 1] - starts[b]);
        const float r = vcm_radius(base_radius, pass0 + b);
        const int rc = wr_hashgrid_build(ctx, vb.pos + 3 * block, n, r, npix, &grids[b]);
        if (rc != WR_OK) return rc;
        G[b] = wrd::VcmGrid{grids[b]->g, grids[b]->ends, grids[b]->idx, vb.pos + 3 * block, vb.src + 2 * block, r};
    }
    if ((e = hipMemcpyAsync(vb.grids, G.data(), sizeof(wrd::VcmGrid) * npass, hipMemcpyHostToDevice, st)) != hipSuccess)
        return hip_fail(ctx, e, who);
    return WR_OK;
}

int launch_vcm_passes(wr_ctx* ctx, const SceneDev& S, const wr_tree_hdr& H, const CamHost& ch, uint32_t max_depth,
                      uint64_t seed, uint32_t pass0, uint32_t npass, float base_radius, uint32_t maxv, void* verts,
                      float* pix, const BdptQueue& q, const VcmBufs& vb, const FilmBufs& film) {
    hipStream_t st = ctx->stream;
    const wrd::CamDev cam = make_cam_dev(ch);
    const uint32_t npix = ch.W * ch.H;
    const uint32_t lanes = npass * npix;
    const uint32_t blocks = (lanes + 255) / 256;
    wrd::BVert* V = (wrd::BVert*)verts;
    hipError_t e = hipMemsetAsync(pix, 0, sizeof(float) * 3 * (size_t)lanes, st);
    if (e == hipSuccess) e = hipMemsetAsync(q.count, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) return hip_fail(ctx, e, "wr_render_vcm");
    std::vector<wr_hashgrid*> grids;
    auto destroy = [&]() {
        for (wr_hashgrid* g : grids) wr_hashgrid_destroy(g);
    };
    int rc = light_point_grids(ctx, S, cam, npix, max_depth, seed, pass0, npass, base_radius, maxv, V, q.counts, vb,
                               true, grids, "wr_render_vcm");
    if (rc != WR_OK) {
        destroy();
        return rc;
    }
    const wrd::VisQueue Q{q.rays, q.pay, q.dst, q.count, q.cap};
    const wrd::MergeQueue M{vb.mq, vb.mq_cap};
    e = hipMemsetAsync(vb.mq + 2 * (size_t)vb.mq_cap, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) {
        destroy();
        return hip_fail(ctx, e, "wr_render_vcm");
    }
    wrd::k_vcm<<<blocks, 256, wrd::kStackLdsBytes, st>>>(S, cam, npix, npass, maxv, V, q.counts, pix, Q,
                                                        (const wrd::VcmGrid*)vb.grids, M);
    e = hipGetLastError();
    if (e == hipSuccess) {
        wrd::k_vcm_merge<<<4096, 256, 0, st>>>(S, cam, npix, maxv, V, pix, M,
                                                                (const wrd::VcmGrid*)vb.grids);
        e = hipGetLastError();
    }
    uint32_t nq = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nq, q.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    destroy();
    if (e != hipSuccess) return hip_fail(ctx, e, "wr_render_vcm");
    if (nq > q.cap) nq = q.cap;
    if (nq) {
        if ((e = launch_closest_hit_path(S, H, q.rays, nq, q.hits, st)) != hipSuccess)
            return hip_fail(ctx, e, "wr_render_vcm: visibility");
        wrd::k_bdpt_accum<<<(nq + 255) / 256, 256, 0, st>>>(nq, q.pay, q.dst, (const uint4*)q.hits, pix);
    }
    wrd::k_bdpt_fold<<<(npix + 255) / 256, 256, 0, st>>>(cam, npass, npix, pix,
                                                         wrd::FilmDev{film.sum, film.comp, film.sumsq});
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(ctx, e, "wr_render_vcm");
    return WR_OK;
}

int launch_ppm_passes(wr_ctx* ctx, const SceneDev& S, const CamHost& ch, uint32_t max_depth, uint64_t seed,
                      uint32_t pass0, uint32_t npass, float base_radius, uint32_t maxv, void* verts, uint32_t* counts,
                      float* pix, const VcmBufs& vb, const FilmBufs& film) {
    hipStream_t st = ctx->stream;
    const wrd::CamDev cam = make_cam_dev(ch);
    const uint32_t npix = ch.W * ch.H;
    const uint32_t lanes = npass * npix;
    const uint32_t blocks = (lanes + 255) / 256;
    wrd::BVert* V = (wrd::BVert*)verts;
    std::vector<wr_hashgrid*> grids;
    auto destroy = [&]() {
        for (wr_hashgrid* g : grids) wr_hashgrid_destroy(g);
    };
    int rc = light_point_grids(ctx, S, cam, npix, max_depth, seed, pass0, npass, base_radius, maxv, V, counts, vb,
                               false, grids, "wr_render_ppm");
    if (rc != WR_OK) {
        destroy();
        return rc;
    }
    wrd::k_ppm<<<blocks, 256, wrd::kStackLdsBytes, st>>>(S, cam, npix, npass, maxv, V, seed, pass0, pix,
                                                        (const wrd::VcmGrid*)vb.grids);
    wrd::k_bdpt_fold<<<(npix + 255) / 256, 256, 0, st>>>(cam, npass, npix, pix,
                                                         wrd::FilmDev{film.sum, film.comp, film.sumsq});
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the grids are read until k_ppm ends
    destroy();
    if (e != hipSuccess) return hip_fail(ctx, e, "wr_render_ppm");
    return WR_OK;
}

size_t vcm_grid_bytes() { return sizeof(wrd::VcmGrid); }

size_t vcm_scan_bytes(uint32_t lanes) {
    size_t tb = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (uint32_t*)nullptr, (uint32_t*)nullptr, (int)lanes + 1);
    return tb;
}

hipError_t launch_bdpt_passes(const SceneDev& S, const wr_tree_hdr& H, const CamHost& ch, uint32_t max_depth,
                              uint64_t seed, uint32_t pass0, uint32_t npass, uint32_t maxv, void* verts, float* pix,
                              const BdptQueue& q, const FilmBufs& film, hipStream_t st) {
    const wrd::CamDev cam = make_cam_dev(ch);
    const uint32_t npix = ch.W * ch.H;
    const uint64_t lanes = (uint64_t)npass * npix;
    hipError_t e = hipMemsetAsync(pix, 0, sizeof(float) * 3 * (size_t)lanes, st);
    if (e == hipSuccess) e = hipMemsetAsync(q.count, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) return e;
    const wrd::VisQueue Q{q.rays, q.pay, q.dst, q.count, q.cap};
    const uint32_t blocks = (uint32_t)((lanes + 255) / 256);
    wrd::k_bdpt_subpaths<<<blocks, 256, wrd::kStackLdsBytes, st>>>(S, cam, max_depth, seed, pass0, npass, npix, maxv,
                                                                  (wrd::BVert*)verts, q.counts);
    wrd::k_bdpt<<<blocks, 256, wrd::kStackLdsBytes, st>>>(S, cam, max_depth, seed, pass0, npass, npix, maxv,
                                                         (wrd::BVert*)verts, q.counts, pix, Q);
    uint32_t nq = 0;
    e = hipMemcpyAsync(&nq, q.count, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return e;
    if (nq > q.cap) nq = q.cap;
    if (nq) {
        e = launch_closest_hit_path(S, H, q.rays, nq, q.hits, st);
        if (e != hipSuccess) return e;
        wrd::k_bdpt_accum<<<(nq + 255) / 256, 256, 0, st>>>(nq, q.pay, q.dst, (const uint4*)q.hits, pix);
    }
    wrd::k_bdpt_fold<<<(npix + 255) / 256, 256, 0, st>>>(cam, npass, npix, pix,
                                                         wrd::FilmDev{film.sum, film.comp, film.sumsq});
    return hipGetLastError();
}

}  // namespace wr
