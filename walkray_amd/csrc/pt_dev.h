// pt_dev.h -- device-side path tracing for gfx950: counter RNG, camera,
// the four hot materials' scatter, the per-path loop and the Kahan film.
//
// Estimator (kept): MCRenderer::samplePath (mcRenderer.cpp:16-57) and the
// PathTracer pixel loop (pathTracer.cpp:20-70): path weight = product over
// inner vertices of cosineTerm * radiance / originProb / directionProb
// (vec3 / 0 is a no-op, nvVector.h:375-379), times the terminal vertex's
// radiance; NaN/Inf/negative samples are zeroed (util.h:8-18); the camera
// vignetting fix (camera.cpp:107-122) divides by |cos| of the pixel-centre
// direction.  Randomness comes from Philox-4x32-10 keyed by (seed, pixel, pass,
// vertex) instead of the reference's one shared mt19937_64 (rng.h:6-59); the
// draws of one vertex are: camera (u, v); diffuse/glossy (roulette, s.x, s.y);
// refractive (Fresnel choice); the next-event light sample (object pick,
// triangle pick, u1, u2) under vertex kNeeVertex.
#pragma once
#include "trace_dev.h"
#include "wr_kernels.h"

#pragma clang fp contract(off)

namespace wrd {

constexpr double kPi = 3.1415926535897932384626433832795;      // macros.h:5
constexpr double kInvPi = 0.31830988618379067153776752674503;  // macros.h:6
constexpr double kEps = 1e-5;                                  // macros.h:7

// ---- Philox-4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers:
// as easy as 1, 2, 3", SC'11) -------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        const uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Four uniforms in [0, 1) (24-bit) for one path vertex.
__device__ __forceinline__ void rng4(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t vertex, float f[4]) {
    uint32_t c[4] = {pixel, pass, vertex, 0x57414C4Bu};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int i = 0; i < 4; i++) f[i] = (float)(c[i] >> 8) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ bool legal(f3 c) {  // util.h:8-18
    return isfinite(c.x) && isfinite(c.y) && isfinite(c.z) && c.x >= 0.f && c.y >= 0.f && c.z >= 0.f;
}
__device__ __forceinline__ f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }

// ---- camera (camera.cpp:6-60, 90-123) -------------------------------------
struct CamDev {
    f3 pos, front, right, top;
    float dist;
    uint32_t W, H;
};

__device__ __forceinline__ f3 cam_dir(const CamDev& c, uint32_t p, float u, float v) {
    const uint32_t x = p % c.W, y = p / c.W;
    const f3 d = add(add(muls(c.front, c.dist), muls(c.right, ((float)x + u) - (c.W / 2.f))),
                     muls(c.top, ((c.H / 2.f) - (float)y) - v));
    return normalize(d);
}

__device__ __forceinline__ f3 fix_vignetting(const CamDev& c, f3 color, uint32_t p) {
    const uint32_t x = p % c.W, y = p / c.W;
    // x + 0.5 - W/2.0 and H/2.0 - 0.5 - y are exact in float for sizes < 2^23
    f3 d = add(add(muls(c.front, c.dist), muls(c.right, (float)x + 0.5f - c.W * 0.5f)),
               muls(c.top, c.H * 0.5f - 0.5f - (float)y));
    d = normalize(d);
    const float fc = dot(d, c.front);
    if (fc == 0.f) return mk(0.f, 0.f, 0.f);
    const f3 r = divs(color, fabsf(fc));
    return legal(r) ? r : mk(0.f, 0.f, 0.f);
}

// ---- scene access -----------------------------------------------------------
struct SceneCtx {
    SceneDev S;
    Tree tr;               // global tree
    const float* gnrm;     // global record -> 9 vertex-normal floats
};

__device__ __forceinline__ SceneCtx make_scene_ctx(const SceneDev& S) {
    SceneCtx c;
    c.S = S;
    const wr_tree_hdr H = S.hdrs[0];
    c.tr = make_tree(H, S);
    c.gnrm = S.rec_nrm + 9 * (size_t)H.rec_base;
    return c;
}

struct PHit {
    float t;
    uint32_t rec;  // global-tree record, WR_NONE on a miss
    uint32_t obj;
};

// Scene::intersect (scene.cpp:102-113), with or without NoSelfCondition.
template <bool NOSELF>
__device__ __forceinline__ PHit trace_scene(const SceneCtx& sc, f3 o, f3 d, uint32_t inside, uint32_t contact,
                                            LaneStack& st) {
    const RayPre r = make_ray(o, d, sc.tr.scale);
    NoSelf ns{inside, contact, sc.gnrm, sc.S.obj_has_normals};
    Counters cnt{0, 0, 0};
    bool failed;
    const HitState h = closest_hit<NOSELF, false>(sc.tr, r, ns, st, cnt, failed);
    PHit ph;
    ph.t = h.t;
    ph.rec = h.rec;
    ph.obj = h.rec == WR_NONE ? WR_NONE : sc.tr.tris[h.rec].obj;
    return ph;
}

// Mesh::getWorldNormal of a global record at p.
__device__ __forceinline__ f3 rec_normal(const SceneCtx& sc, uint32_t rec, uint32_t obj, f3 p) {
    const float* nrm = sc.S.obj_has_normals[obj] ? sc.gnrm + 9 * (size_t)rec : nullptr;
    float u1, u2;
    return world_normal(sc.tr.tris[rec], nrm, p, u1, u2);
}

// ---- scatter (AbstractObject::scatter of the hot materials) -----------------
struct Out {
    f3 o, d, rad, n;  // n: contactObj->getWorldNormal(contactTriangleID, o)
    float op, dp;
    uint32_t inside, contact;
    bool delta;       // isDeltaDirection (the refractive scatter, refractiveObj.h:85)
};

// Ray::cosineTerm (ray.cpp:52-58) of a scatter vertex.
__device__ __forceinline__ float cosine_term(const Out& r) {
    float c = 1.f;
    if ((double)len(r.d) > 0.5) c = fabsf(dot(r.n, r.d));
    return (double)c > kEps ? c : (float)kEps;
}

// Frame::setFromN + toWorldFrame (frame.h:23-49).
__device__ __forceinline__ f3 to_world(f3 w, f3 l) {
    f3 mX, mZ;
    if ((double)fabsf(w.x) > kEps) {
        mX = normalize(cross(mk(0.f, 1.f, 0.f), w));
        mZ = cross(mX, w);
    } else {
        mZ = normalize(cross(mk(1.f, 0.f, 0.f), w));
        mX = cross(w, mZ);
    }
    return add(add(muls(mX, l.x), muls(w, l.y)), muls(mZ, l.z));
}

__device__ __forceinline__ f3 mirror(f3 N, f3 d) {  // -N.dot(d)*N*2 + d, normalized
    return normalize(add(muls(muls(N, -dot(N, d)), 2.f), d));
}

// The object that scatters a ray from in_inside hitting h (refractiveObj.h:32-33:
// inside a refractive object and hitting another, that other object scatters).
__device__ __forceinline__ uint32_t scatter_object(const SceneCtx& sc, uint32_t self, const PHit& h,
                                                   uint32_t in_inside) {
    if (sc.S.objs[self].mat == WR_MAT_REFRACTIVE && h.obj != self && in_inside == self) return h.obj;
    return self;
}

// Whether scatter needs Scene::findInsideObject: a refractive object's scatter
// of a ray leaving it (refractiveObj.h:43-44).
__device__ __forceinline__ bool scatter_queries_inside(const SceneCtx& sc, uint32_t self, const PHit& h,
                                                       uint32_t in_inside) {
    const uint32_t s = scatter_object(sc, self, h, in_inside);
    return sc.S.objs[s].mat == WR_MAT_REFRACTIVE && in_inside == s;
}

// INSIDE = false compiles the findInsideObject query out, for callers that
// never pass a ray for which scatter_queries_inside holds.
template <bool INSIDE = true>
__device__ Out scatter(const SceneCtx& sc, uint32_t self, f3 in_o, f3 in_d, const PHit& h, uint32_t in_inside,
                       const float* rn, LaneStack& st) {
    const f3 pos = add(in_o, muls(in_d, h.t));
    const f3 N = rec_normal(sc, h.rec, h.obj, pos);
    self = scatter_object(sc, self, h, in_inside);
    const wr_object_gpu& ob = sc.S.objs[self];
    Out out;
    out.o = pos;
    out.n = N;
    out.contact = self;
    out.op = 1.f;
    out.dp = 1.f;
    out.rad = mk(1.f, 1.f, 1.f);
    out.inside = in_inside;
    out.delta = false;
    const f3 color = mk(ob.color[0], ob.color[1], ob.color[2]);
    if (ob.mat == WR_MAT_DIFFUSE) {  // diffuseObj.h:34-57, diffuseBsdf.h:17-26, cosHemiSphereSampler.h:6-18
        const float p = ob.pmax;
        if (rn[0] < p) {
            const float term1 = (float)(2.0 * kPi * (double)rn[1]);
            const float term2 = sqrtf(1.f - rn[2]);
            const f3 local = mk(sinf(term1) * term2, sqrtf(rn[2]), cosf(term1) * term2);
            out.d = to_world(N, local);
            const float c = dot(N, out.d);
            out.dp = c > 0.f ? (float)((double)c * kInvPi) : 0.f;
            out.rad = muls(color, (float)kInvPi);
            out.dp *= p;
        } else {
            out.d = mk(0.f, 0.f, 0.f);
            out.rad = mk(0.f, 0.f, 0.f);
            out.dp = 1.f - p;
        }
    } else if (ob.mat == WR_MAT_GLOSSY) {  // glossyObj.h:39-62, isotropicPhong.h:20-38, powerCos...h:8-21
        const float p = ob.pmax;
        if (rn[0] < p) {
            const float power = ob.power;
            const float term1 = (float)(2.0 * kPi * (double)rn[1]);
            const float term2 = (float)pow((double)rn[2], 1.0 / ((double)power + 1.0));
            const float term3 = (float)sqrt(1.0 - (double)(term2 * term2));
            const f3 local = mk(sinf(term1) * term3, term2, cosf(term1) * term3);
            const f3 w = to_world(N, local);
            const float c = dot(N, w);
            out.dp = c > 0.f ? (float)(((double)power + 1.0) * (double)powf(c, power) * (kInvPi * 0.5)) : 0.f;
            const f3 refl = mirror(N, in_d);
            const f3 rho = muls(muls(muls(color, (float)((double)power + 2.0)), 0.5f), (float)kInvPi);
            const float dRW = dot(refl, w);
            if (dot(w, N) < 0.f) {
                out.d = mk(0.f, 0.f, 0.f);
                out.rad = mk((float)kEps, (float)kEps, (float)kEps);
            } else {
                out.d = w;
                out.rad = muls(rho, powf(dRW > 0.f ? dRW : 0.f, power));
            }
            out.dp *= p;
        } else {
            out.d = mk(0.f, 0.f, 0.f);
            out.rad = mk(0.f, 0.f, 0.f);
            out.dp = 1.f - p;
        }
    } else if (ob.mat == WR_MAT_REFRACTIVE) {  // refractiveObj.h:27-88
        out.d = in_d;
        const f3 refl = mirror(N, in_d);
        const float theta = acosf(dot(in_d, N));
        uint32_t cur_out = self;
        if (INSIDE && in_inside == self) cur_out = find_inside(sc.S, out.o, in_d, self, st);
        const float cn = in_inside != WR_NONE ? sc.S.objs[in_inside].ior : 1.f;
        const float nn = cur_out != WR_NONE ? sc.S.objs[cur_out].ior : 1.f;
        const float sin_phi = cn / nn * sinf(theta);
        if (sin_phi > 1.f) {  // total internal reflection: radiance 1, probabilities 1
            out.d = refl;
            out.inside = in_inside;
        } else {
            float phi = asinf(sin_phi);
            if ((double)theta > kPi / 2) phi = (float)(kPi - (double)phi);
            const f3 ax = normalize(cross(N, in_d));
            // RotateMatrix(axis, phi) * vec4f(N, 0)  (util.h:40-54, nvMatrix.h:98-105,424-433)
            const float c = cosf(phi), s = sinf(phi);
            const float _c = 1.f - c;
            const float x = ax.x, y = ax.y, z = ax.z;
            const float m00 = c + _c * x * x, m01 = _c * x * y - s * z, m02 = _c * x * z + s * y;
            const float m10 = _c * x * y + s * z, m11 = c + _c * y * y, m12 = _c * y * z - s * x;
            const float m20 = _c * x * z - s * y, m21 = _c * y * z + s * x, m22 = c + _c * z * z;
            f3 rd;
            rd.x = N.x * m00 + N.y * m01 + N.z * m02 + 0.f * 0.f;
            rd.y = N.x * m10 + N.y * m11 + N.z * m12 + 0.f * 0.f;
            rd.z = N.x * m20 + N.y * m21 + N.z * m22 + 0.f * 0.f;
            out.d = normalize(rd);
            const float cos_theta = fabsf(cosf(theta));
            const float cos_phi = fabsf(cosf(phi));
            const float a = fabsf(cn * cos_theta - nn * cos_phi) / (cn * cos_theta + nn * cos_phi);
            const float b = fabsf(nn * cos_theta - cn * cos_phi) / (nn * cos_theta + cn * cos_phi);
            const float er = (a * a + b * b) / 2.f;
            if (rn[0] < er) {
                out.d = refl;
                out.rad = muls(out.rad, er / cosine_term(out));
                out.dp = er;
                out.inside = in_inside;
            } else {
                out.rad = muls(out.rad, (1.f - er) / cosine_term(out));
                out.dp = 1.f - er;
                out.inside = cur_out;
            }
        }
        out.d = normalize(out.d);
        out.delta = true;
    } else {  // emissive, emissiveObj.h:15-28
        out.d = mk(0.f, 0.f, 0.f);
        out.inside = WR_NONE;
        out.rad = dot(in_d, N) <= 0.f ? color : mk(0.f, 0.f, 0.f);
    }
    return out;
}

// ---- next-event estimation (PathTracer's useNextEventEstimation branch) -------
// Vertex counter of the light sample's uniforms; path vertices use 0 .. max_depth + 1.
constexpr uint32_t kNeeVertex = 0xFFFFFFFFu;

// std::lower_bound over n ascending floats: the first index whose value is not
// less than v (n when there is none).
__device__ __forceinline__ uint32_t lower_bound_f(const float* a, uint32_t n, float v) {
    uint32_t lo = 0;
    while (n > 0) {
        const uint32_t half = n >> 1;
        if (a[lo + half] < v) {
            lo += half + 1;
            n -= half + 1;
        } else {
            n = half;
        }
    }
    return lo;
}

struct LightSample {
    f3 o, n, rad;
    float op;
    uint32_t obj, tri;  // the light object and its (local) triangle
    bool ok;
};

// Scene::genLightSample -> SurfaceSampler::genSample (surfaceSampler.cpp:17-23)
// -> AbstractObject::genSurfaceSample (abstractObj.cpp:18-52) with
// Mesh::genRandTrianglePosition (mesh.h:89-101).  rn: object pick, triangle
// pick, u1, u2.  The hemisphere direction, findInsideObject and the NoSelf
// intersect of the light ray are not formed: the path tracer overwrites the
// direction and reads neither of the others.  ok == false: no emissive object,
// or the picked one has no triangles (its sample has no contact object).
__device__ __forceinline__ LightSample sample_light(const SceneCtx& sc, const float* rn) {
    const SceneDev& S = sc.S;
    LightSample L;
    L.ok = false;
    if (S.n_lights == 0) return L;
    uint32_t li = lower_bound_f(S.light_cum, S.n_lights, rn[0] * S.light_total);
    if (li == S.n_lights) li--;
    const wr_light lt = S.lights[li];
    if (lt.n_tris == 0) return L;
    uint32_t ti = lower_bound_f(S.light_tri_cum + lt.tri_base, lt.n_tris, rn[1] * lt.area);
    if (ti == lt.n_tris) ti--;
    const uint32_t rec = S.light_tri_rec[lt.tri_base + ti];
    const wr_tri& T = sc.tr.tris[rec];
    const f3 v0 = mk(T.v0[0], T.v0[1], T.v0[2]);
    const f3 b1 = sub(mk(T.v1[0], T.v1[1], T.v1[2]), v0);
    const f3 b2 = sub(mk(T.v2[0], T.v2[1], T.v2[2]), v0);
    float u1 = rn[2], u2 = rn[3];
    if (u1 + u2 > 1.f) {
        u1 = 1.f - u1;
        u2 = 1.f - u2;
    }
    L.o = add(add(v0, muls(b1, u1)), muls(b2, u2));
    L.n = rec_normal(sc, rec, lt.obj, L.o);
    const wr_object_gpu& ob = S.objs[lt.obj];
    L.rad = mk(ob.color[0], ob.color[1], ob.color[2]);  // EmissiveObject::evaluateBSDF
    L.op = lt.oprob;
    L.obj = lt.obj;
    L.tri = ti;
    L.ok = true;
    return L;
}

// Ray::cosineTerm (ray.cpp:52-58) of a ray with a contact object.
__device__ __forceinline__ float cos_term(f3 n, f3 d) {
    float c = 1.f;
    if ((double)len(d) > 0.5) c = fabsf(dot(n, d));
    return (double)c > kEps ? c : (float)kEps;
}

struct NeeConn {
    f3 dir, C;  // connection direction; the sample's value if the light point is visible
    float dist; // |light point - o|
    bool ok;
};

// The branch up to its visibility test (pathTracer.cpp:42-52, 58-65), at the
// origin o of the path's ray endRay: contact = its contactObj, crec = the
// global record of the triangle under o (contactTriangleID), front = (the
// incoming direction . N) <= 0 (the first condition of DiffuseBSDF::evalBSDF,
// diffuseBsdf.h:29-34), P = the product over the rays before endRay in the
// branch's order (radiance / originProb / directionProb, cosine, decay).
// ok == false: the branch adds nothing.
__device__ __forceinline__ NeeConn nee_connect(const SceneCtx& sc, f3 o, uint32_t contact, uint32_t crec, bool front,
                                               f3 P, const float* rn) {
    NeeConn c;
    c.ok = false;
    const wr_object_gpu& ob = sc.S.objs[contact];
    // isEmissive / !nonSpecular: only diffuse vertices connect (glossy is
    // specular without MERGE_GLOSSY, glossyObj.h:32-38; refractiveObj.h:26)
    if (ob.mat != WR_MAT_DIFFUSE) return c;
    const LightSample L = sample_light(sc, rn);
    if (!L.ok) return c;
    const f3 dir = normalize(sub(L.o, o));
    const float dist = len(sub(L.o, o));
    const float cd = (double)dist > kEps ? dist : (float)kEps;  // MAX(dist, EPSILON)
    if (dot(dir, L.n) >= 0.f) return c;
    const f3 nE = rec_normal(sc, crec, sc.tr.tris[crec].obj, o);
    f3 f = (front && dot(dir, nE) >= 0.f) ? muls(mk(ob.color[0], ob.color[1], ob.color[2]), (float)kInvPi)
                                          : mk(0.f, 0.f, 0.f);
    f = mulv(f, L.rad);  // radianceDecay is (1,1,1) here: multiplying by it is exact
    f = muls(f, cos_term(L.n, mk(-dir.x, -dir.y, -dir.z)));
    f = muls(f, cos_term(nE, dir));
    f = divs(f, cd * cd);
    // endRay's originProb is 1 for every scatter here, so the divisor is the light's
    c.C = divs(mulv(P, f), L.op);
    c.dir = dir;
    c.dist = dist;
    c.ok = true;
    return c;
}

// MCRenderer::visibilityTest (mcRenderer.cpp:59-65) from the closest hit of the
// connection ray.
__device__ __forceinline__ bool nee_visible(uint32_t rec, float t, float dist) {
    return rec == WR_NONE || (double)(t - dist) >= -kEps;
}

// ---- Kahan film ---------------------------------------------------------------
struct FilmDev {
    float* sum;    // 3 per pixel
    float* comp;   // 3 per pixel (Kahan compensation)
    float* sumsq;  // 3 per pixel, plain sum of squares (for standard errors)
};

__device__ __forceinline__ void kahan_add(float& s, float& c, float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

__device__ __forceinline__ void film_add(const FilmDev& f, uint32_t p, f3 v) {
    const float vv[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float s = f.sum[3 * p + k], c = f.comp[3 * p + k];
        kahan_add(s, c, vv[k]);
        f.sum[3 * p + k] = s;
        f.comp[3 * p + k] = c;
        f.sumsq[3 * p + k] += vv[k] * vv[k];
    }
}

// The next-event branch for one path, connection and visibility test included.
__device__ __forceinline__ f3 nee_sample(const SceneCtx& sc, f3 o, uint32_t contact, uint32_t crec, bool front, f3 P,
                                         uint32_t inside, uint64_t seed, uint32_t p, uint32_t pass, LaneStack& st) {
    float rn[4];
    rng4(seed, p, pass, kNeeVertex, rn);
    const NeeConn c = nee_connect(sc, o, contact, crec, front, P, rn);
    if (!c.ok) return mk(0.f, 0.f, 0.f);
    const PHit h = trace_scene<true>(sc, o, c.dir, inside, contact, st);
    return nee_visible(h.rec, h.t, c.dist) ? c.C : mk(0.f, 0.f, 0.f);
}

// ---- one eye path, megakernel style (mcRenderer.cpp:16-57, pathTracer.cpp:22-69) ----
// NEE: PathTracer's next-event branch (pathTracer.cpp:39-66).  It only changes
// paths whose last entry is terminateRay after an escape or at the depth limit
// (hasToConnect with endIndex > 0): they connect the origin of their last
// traced ray to a light sample.  Every other path keeps the plain value.
template <bool NEE>
__device__ f3 sample_path(const SceneCtx& sc, const CamDev& cam, uint32_t max_depth, uint64_t seed, uint32_t p,
                          uint32_t pass, LaneStack& st) {
    float rn[4];
    rng4(seed, p, pass, 0, rn);
    f3 o = cam.pos;
    f3 d = cam_dir(cam, p, rn[0], rn[1]);
    PHit hit = trace_scene<false>(sc, o, d, WR_NONE, WR_NONE, st);
    uint32_t inside = find_inside(sc.S, o, d, WR_NONE, st);
    float c0 = fabsf(dot(cam.front, d));
    c0 = (double)c0 > kEps ? c0 : (float)kEps;
    f3 T = mk(c0, c0, c0);
    // NEE state of the current ray (index depth): U = the branch-order product up
    // to it, P = without it; contact object, triangle record and incoming side
    // at its origin (valid for depth >= 1)
    f3 U = T, P = T;
    uint32_t c_contact = WR_NONE, c_rec = WR_NONE;
    bool c_front = false;
    f3 color = mk(0.f, 0.f, 0.f);
    uint32_t depth = 0;
    for (;;) {
        uint32_t self = inside != WR_NONE ? inside : hit.obj;
        if (self == WR_NONE || hit.rec == WR_NONE) break;  // terminateRay: contributes 0
        rng4(seed, p, pass, depth + 1, rn);
        const Out out = scatter(sc, self, o, d, hit, inside, rn, st);
        if ((double)len(out.d) < 0.5) {  // emission / absorption ends the path
            color = mulv(T, divs(divs(out.rad, out.op), out.dp));
            break;
        }
        if (depth + 1 > max_depth) {  // endIndex = depth: connect from this ray's origin
            if (NEE && depth >= 1) color = nee_sample(sc, o, c_contact, c_rec, c_front, P, inside, seed, p, pass, st);
            break;
        }
        const PHit nh = trace_scene<true>(sc, out.o, out.d, out.inside, out.contact, st);
        if (nh.rec == WR_NONE) {  // escaped: endIndex = depth + 1, connect from out's origin
            if (NEE)
                color = nee_sample(sc, out.o, out.contact, hit.rec, dot(d, out.n) <= 0.f, U, out.inside, seed, p,
                                   pass, st);
            break;
        }
        T = muls(T, cosine_term(out));
        T = mulv(T, divs(divs(out.rad, out.op), out.dp));
        if (NEE) {
            P = U;
            U = mulv(U, divs(divs(out.rad, out.op), out.dp));
            U = muls(U, cosine_term(out));
            c_contact = out.contact;
            c_rec = hit.rec;
            c_front = dot(d, out.n) <= 0.f;
        }
        o = out.o;
        d = out.d;
        hit = nh;
        inside = out.inside;
        depth++;
    }
    if (!legal(color)) color = mk(0.f, 0.f, 0.f);
    return fix_vignetting(cam, color, p);
}

}  // namespace wrd
