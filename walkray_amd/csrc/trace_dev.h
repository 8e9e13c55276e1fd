// trace_dev.h -- device-side closest-hit traversal for gfx950 (wave64).
//
// Semantics are those of KDTree::intersect (kdtree.cpp:279-353): the result is
// the minimum-t triangle among those whose every ancestor box passes the
// reference box test (kdtree.cpp:46-64) and whose reference triangle test
// (kdtree.cpp:279-306) accepts, ties going to the right subtree
// (kdtree.cpp:331-346) and, inside a leaf, to the earlier triangle
// (kdtree.cpp:316-324).  t, the triangle test and the box test use the
// reference operation order with correctly rounded division / sqrt and no
// contraction, so they are bit-identical to the reference's FP32 arithmetic.
//
// Execution differs (and is proven equal, see below):
//   1. fast pass: ordered stack traversal (nearer child first, stack of node
//      references per lane in LDS) with a conservative slab test - each box
//      is enlarged by e = 2^-19 * max(scene scale, |origin|), a bound well
//      above the rounding error of the reference's face-point computation,
//      so every box the reference accepts is accepted here - and t-pruning
//      with a relative margin of 2^-10;
//   2. the winner's root-to-leaf chain is re-checked with the exact reference
//      box test; if every box passes, the winner is in the reference's
//      candidate set and is the minimum over a superset of it, hence the
//      reference's answer (ties use the same total order: t, then larger
//      first-record, then earlier slot);
//   3. otherwise the ray is re-traced with the exact reference box test at
//      every node (the fallback; rare: only rays grazing box edges).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wr_kernels.h"
#include "wr_layout.h"

#pragma clang fp contract(off)

namespace wrd {

struct Tree {
    const wr_node* nodes;      // already offset to this tree
    const wr_wnode* wnodes;    // already offset to this tree
    const wr_tri* tris;        // already offset to this tree
    const uint32_t* leaf_cnt;  // offset like tris
    const wr_chain* chains;
    const float* top_boxes;    // the scene's pool (chain records hold global box indices)
    float rmin[3], rmax[3];
    uint32_t root_ref;
    uint32_t wide_root;
    float scale;
};

__device__ __forceinline__ Tree make_tree(const wr_tree_hdr& h, const SceneDev& S) {
    Tree t;
    t.nodes = S.nodes + h.node_base;
    t.wnodes = S.wnodes + h.wide_base;
    t.tris = S.tris + h.rec_base;
    t.leaf_cnt = S.leaf_cnt + h.rec_base;
    t.chains = S.chains;
    t.top_boxes = S.top_boxes;
    for (int a = 0; a < 3; a++) {
        t.rmin[a] = h.root_min[a];
        t.rmax[a] = h.root_max[a];
    }
    t.root_ref = h.root_ref;
    t.wide_root = h.wide_root;
    t.scale = h.scale;
    return t;
}

// ---- reference float helpers (nvVector.h semantics) ------------------------
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 muls(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float dot(f3 a, f3 b) {
    float r = 0.f;  // nvVector.h:501-507 starts from 0 (matters for signed zeros)
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 divs(f3 a, float s) {  // operator/=: zero divisor is a no-op
    if (s == 0.f) return a;
    return mk(a.x / s, a.y / s, a.z / s);
}
__device__ __forceinline__ f3 normalize(f3 a) {
    float r = len(a);
    if (r > 0.f) return divs(a, r);
    return a;
}

// ---- ray pre-computation --------------------------------------------------
struct RayPre {
    f3 o, d;
    float ix, iy, iz;     // 1/d (finite stand-in for d == 0)
    // per axis, the offsets of the enlarged box's near and far planes: with
    // ol = (o + e) / d and oh = (o - e) / d, near = ol and far = oh where
    // 1/d > 0, swapped where 1/d < 0
    float onx, ony, onz;
    float ofx, ofy, ofz;
    float e;              // box enlargement in world units
    float tscale;         // 2^-20 * max(1, |d|_inf): t-proportional part of the face-point error bound
};

__device__ __forceinline__ float safe_inv(float d) {
    return fabsf(d) > 1e-30f ? 1.0f / d : copysignf(1e30f, d);
}

__device__ __forceinline__ RayPre make_ray(f3 o, f3 d, float scale) {
    RayPre r;
    r.o = o;
    r.d = d;
    float m = fmaxf(fmaxf(scale, fabsf(o.x)), fmaxf(fabsf(o.y), fabsf(o.z)));
    r.e = m * 1.9073486328125e-06f;  // 2^-19
    r.ix = safe_inv(d.x);
    r.iy = safe_inv(d.y);
    r.iz = safe_inv(d.z);
    const float olx = (o.x + r.e) * r.ix, ohx = (o.x - r.e) * r.ix;
    const float oly = (o.y + r.e) * r.iy, ohy = (o.y - r.e) * r.iy;
    const float olz = (o.z + r.e) * r.iz, ohz = (o.z - r.e) * r.iz;
    r.onx = r.ix > 0.f ? olx : ohx;
    r.ofx = r.ix > 0.f ? ohx : olx;
    r.ony = r.iy > 0.f ? oly : ohy;
    r.ofy = r.iy > 0.f ? ohy : oly;
    r.onz = r.iz > 0.f ? olz : ohz;
    r.ofz = r.iz > 0.f ? ohz : olz;
    r.tscale = fmaxf(1.f, fmaxf(fmaxf(fabsf(d.x), fabsf(d.y)), fabsf(d.z))) * 9.5367431640625e-07f;  // 2^-20
    return r;
}

// Conservative slab test on the enlarged box, clipped to [0, tcut], from the
// near-plane and far-plane coordinates of each axis (slab_nf) or from the box
// (slab).  The plane distances are t0 = fma(lo, 1/d, -(o + e)/d) and t1 =
// fma(hi, 1/d, -(o - e)/d); rounding is monotone, so t0 <= t1 where 1/d > 0
// and t0 >= t1 where 1/d < 0: picking the near plane by the sign of 1/d gives
// exactly min(t0, t1), and the far plane max(t0, t1).
__device__ __forceinline__ bool slab_nf(const RayPre& r, float nx, float ny, float nz, float fx, float fy, float fz,
                                        float tcut, float& tn_out) {
    const float tnx = __builtin_fmaf(nx, r.ix, -r.onx), tfx = __builtin_fmaf(fx, r.ix, -r.ofx);
    const float tny = __builtin_fmaf(ny, r.iy, -r.ony), tfy = __builtin_fmaf(fy, r.iy, -r.ofy);
    const float tnz = __builtin_fmaf(nz, r.iz, -r.onz), tfz = __builtin_fmaf(fz, r.iz, -r.ofz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, 0.f));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, tcut));
    tn_out = tn;
    return tn <= tf;
}

__device__ __forceinline__ bool slab(const RayPre& r, float lx, float ly, float lz, float hx, float hy,
                                     float hz, float tcut, float& tn_out) {
    const bool px = r.ix > 0.f, py = r.iy > 0.f, pz = r.iz > 0.f;
    return slab_nf(r, px ? lx : hx, py ? ly : hy, pz ? lz : hz, px ? hx : lx, py ? hy : ly, pz ? hz : lz, tcut,
                   tn_out);
}

// Node::BoundingBox::intersect (kdtree.cpp:46-64) reduced to "returns >= 0".
// Arguments by value so the (rare) call keeps everything in VGPRs.
__device__ __noinline__ bool ref_box_exact(float ox, float oy, float oz, float dx, float dy, float dz, float lx,
                                           float ly, float lz, float hx, float hy, float hz) {
    const float oo[3] = {ox, oy, oz};
    const float dd[3] = {dx, dy, dz};
    const float mn[3] = {lx, ly, lz};
    const float mx[3] = {hx, hy, hz};
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int a = i % 3, a1 = (i + 1) % 3, a2 = (i + 2) % 3;
        if (dd[a] == 0.f) continue;
        const float plane = i < 3 ? mn[a] : mx[a];
        const float dist = (plane - oo[a]) / dd[a];
        const float p1 = oo[a1] + dist * dd[a1];
        const float p2 = oo[a2] + dist * dd[a2];
        if (dist >= 0.f && mn[a1] <= p1 && mx[a1] >= p1 && mn[a2] <= p2 && mx[a2] >= p2) return true;
    }
    return false;
}

// One far-plane face of the reference box test decided without division:
// the sign of the reference's dist = (plane - o)/d is exact from the signs of
// (plane - o) and d; the face point on the two other axes is approximated with
// a reciprocal and an fma, whose distance to the reference's rounded point is
// below delta = e + |t| * tscale (e >= 2^-19 max(scale, |o|)).  A point at
// least delta inside the face rectangle is therefore inside for the reference
// too, and the reference box test returns >= 0.
__device__ __forceinline__ bool far_face_accepts(float far_a, float o_a, float d_a, float inv_a, float o_b, float d_b,
                                                 float lo_b, float hi_b, float o_c, float d_c, float lo_c,
                                                 float hi_c, float e, float tscale) {
    const float num = far_a - o_a;
    const bool sign_ok = (d_a != 0.f) & ((num == 0.f) | ((num > 0.f) == (d_a > 0.f)));
    const float t = num * inv_a;
    const float delta = __builtin_fmaf(fabsf(t), tscale, e);
    const float pb = __builtin_fmaf(t, d_b, o_b);
    const float pc = __builtin_fmaf(t, d_c, o_c);
    return sign_ok & (pb >= lo_b + delta) & (pb <= hi_b - delta) & (pc >= lo_c + delta) & (pc <= hi_c - delta);
}

// The reference box test (exact result): a division-free proof on the three
// far-plane faces first, the exact six-face test only when none is conclusive.
__device__ __forceinline__ bool ref_box(const RayPre& r, const float* mn, const float* mx) {
    const float fx = r.d.x > 0.f ? mx[0] : mn[0];
    const float fy = r.d.y > 0.f ? mx[1] : mn[1];
    const float fz = r.d.z > 0.f ? mx[2] : mn[2];
    if (far_face_accepts(fx, r.o.x, r.d.x, r.ix, r.o.y, r.d.y, mn[1], mx[1], r.o.z, r.d.z, mn[2], mx[2], r.e, r.tscale) ||
        far_face_accepts(fy, r.o.y, r.d.y, r.iy, r.o.z, r.d.z, mn[2], mx[2], r.o.x, r.d.x, mn[0], mx[0], r.e, r.tscale) ||
        far_face_accepts(fz, r.o.z, r.d.z, r.iz, r.o.x, r.d.x, mn[0], mx[0], r.o.y, r.d.y, mn[1], mx[1], r.e, r.tscale))
        return true;
    return ref_box_exact(r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]);
}

// Reference triangle test (kdtree.cpp:279-306) on a triangle record; -1 on a miss.
// The test is a pure function, so its three conditions may be evaluated in
// any order: t (exactly the reference's orthoDist / d) comes first, and the
// three edge-sign tests are skipped when t cannot beat the current best
// (t > best_t, or equal with a tie the current best wins) - an exact cut.
__device__ __forceinline__ float ref_tri(const wr_tri& T, const f3& o, const f3& d, float best_t, bool tie_loses) {
    const f3 v0 = mk(T.v0[0], T.v0[1], T.v0[2]);
    const f3 v1 = mk(T.v1[0], T.v1[1], T.v1[2]);
    const f3 v2 = mk(T.v2[0], T.v2[1], T.v2[2]);
    const f3 r2 = sub(v2, o);
    const f3 n = mk(T.n[0], T.n[1], T.n[2]);
    const float od = dot(n, r2);
    const float dn = dot(n, d);
    if (dn == 0.f) return -1.f;
    const float t = od / dn;
    if (!(t >= 0.f) || t > best_t || (t == best_t && tie_loses)) return -1.f;
    const f3 l0 = sub(v1, v0), l1 = sub(v2, v1), l2 = sub(v0, v2);
    const f3 r0 = sub(v0, o), r1 = sub(v1, o);
    const float s0 = dot(d, cross(l0, r0));
    const float s1 = dot(d, cross(l1, r1));
    const float s2 = dot(d, cross(l2, r2));
    const bool pos = (s0 >= 0.f) & (s1 >= 0.f) & (s2 >= 0.f);
    const bool neg = (s0 <= 0.f) & (s1 <= 0.f) & (s2 <= 0.f);
    if (!(pos | neg)) return -1.f;
    return t;
}

// Mesh::getWorldNormal (mesh.h:47-77) for a record, identity transform.
// nrm: 9 floats (vertex normals n0, n1, n2) or nullptr for flat shading.
__device__ __forceinline__ f3 world_normal(const wr_tri& T, const float* nrm, f3 pos, float& u1, float& u2) {
    const f3 v0 = mk(T.v0[0], T.v0[1], T.v0[2]);
    const f3 b1 = sub(mk(T.v1[0], T.v1[1], T.v1[2]), v0);
    const f3 b2 = sub(mk(T.v2[0], T.v2[1], T.v2[2]), v0);
    const f3 v = sub(pos, v0);
    const float d12 = dot(b1, b2);
    const float l1 = len(b1);
    const float l2 = len(b2);
    const float vb1 = dot(v, b1), vb2 = dot(v, b2);
    u2 = (vb1 * d12 - vb2 * l1 * l1) / (d12 * d12 - l1 * l1 * l2 * l2);
    u1 = (vb1 - u2 * d12) / (l1 * l1);
    if (!nrm) return normalize(cross(b1, b2));
    const float w0 = 1.f - u1 - u2;
    const f3 n0 = mk(nrm[0], nrm[1], nrm[2]), n1 = mk(nrm[3], nrm[4], nrm[5]), n2 = mk(nrm[6], nrm[7], nrm[8]);
    return normalize(add(add(muls(n0, w0), muls(n1, u1)), muls(n2, u2)));
}

// NoSelfCondition::legal (noself.cpp:13-20) on object ids (WR_NONE = NULL).
struct NoSelf {
    uint32_t inside, contact;
    const float* rec_nrm;       // record -> 9 normals, already offset
    const uint32_t* obj_flags;  // object -> has_normals
};

__device__ __forceinline__ bool noself_legal(const NoSelf& ns, const wr_tri& T, uint32_t rec, const f3& o,
                                             const f3& d, float t) {
    // both NULL: legal() returns !(in && NULL == obj), true for every object
    if (ns.inside == WR_NONE && ns.contact == WR_NONE) return true;
    const uint32_t obj = T.obj;
    const float* nrm = ns.obj_flags[obj] ? ns.rec_nrm + 9 * (size_t)rec : nullptr;
    float u1, u2;
    const f3 pos = add(o, muls(d, t));
    const f3 N = world_normal(T, nrm, pos, u1, u2);
    const float s = dot(d, N);
    const bool in = s <= 0.f, out = s >= 0.f;
    if (ns.inside == ns.contact) return !(in && ns.contact == obj);
    return !(out && ns.inside != obj);
}

struct Counters {
    uint32_t nodes, tris, fallback;
};

struct HitState {
    float t;
    uint32_t rec;   // record index within the tree, WR_NONE on a miss
    uint32_t leaf;  // the winner's leaf key (wr_tri::leaf): tie order and chain record
};

// Per-lane stack of (reference, entry distance) pairs: the first D levels
// live in LDS (lane-interleaved, stk[level * kBlock + lane]); deeper levels
// spill to a private array (scratch), which only deep traversals touch.  An
// entry whose entry distance exceeds the current cut when it is popped is
// dropped unvisited: that is the slab test it would fail anyway.  D is a
// launch choice of the closest-hit kernels (the autotuner's stack-depth knob:
// 4, 8 or 16 levels, 8 KB to 32 KB of LDS per 256-lane workgroup); the other
// traversal kernels use kLdsDepth = 8, 16 KB, which never holds occupancy
// below 8 waves per SIMD.  Which entries sit in LDS never changes a result.
constexpr int kLdsDepth = 8;
constexpr int kMaxDepth = 64;
constexpr int kBlock = 256;  // every traversal kernel runs 256-lane workgroups

// Dynamic LDS of the launching kernel: 2 * D * kBlock words (references, then
// entry distances).
extern __shared__ uint32_t wr_lds_stack[];

// The spill levels are a separate private array of the launching kernel: kept
// inside LaneStack, the dynamically indexed arrays would put the whole struct,
// sp and lane included, in scratch, and every traversal step would start with
// a scratch load of sp before its node fetch.
template <int D>
struct StackSpillT {
    uint32_t ref[kMaxDepth - D];
    float t[kMaxDepth - D];
};

template <int D>
struct LaneStackT {
    static_assert(D > 0 && (D & (D - 1)) == 0 && D < kMaxDepth, "LDS levels: a power of two below kMaxDepth");
    StackSpillT<D>* spill;
    uint32_t lane;  // threadIdx.x
    int sp;
    __device__ explicit LaneStackT(StackSpillT<D>* s) : spill(s), lane(threadIdx.x), sp(0) {}
    // The LDS accesses are unconditional (a wrapped slot when spilling) so they
    // stay ds_read/ds_write instead of being merged with the scratch access
    // into a generic (flat) one.
    __device__ __forceinline__ void push(uint32_t v, float t) {
        const uint32_t slot = (sp & (D - 1)) * kBlock + lane;
        if (sp >= D) {
            spill->ref[sp - D] = wr_lds_stack[slot];
            spill->t[sp - D] = __uint_as_float(wr_lds_stack[D * kBlock + slot]);
        }
        wr_lds_stack[slot] = v;
        wr_lds_stack[D * kBlock + slot] = __float_as_uint(t);
        sp++;
    }
    __device__ __forceinline__ uint32_t pop(float& t) {
        --sp;
        const uint32_t slot = (sp & (D - 1)) * kBlock + lane;
        const uint32_t v = wr_lds_stack[slot];
        t = __uint_as_float(wr_lds_stack[D * kBlock + slot]);
        if (sp >= D) {
            wr_lds_stack[slot] = spill->ref[sp - D];
            wr_lds_stack[D * kBlock + slot] = __float_as_uint(spill->t[sp - D]);
        }
        return v;
    }
    // Pop the next entry that can still hold a hit at t <= tcut; false when
    // the stack runs empty.
    __device__ __forceinline__ bool next(float tcut, uint32_t& ref) {
        while (sp > 0) {
            float t;
            const uint32_t v = pop(t);
            if (t <= tcut) {
                ref = v;
                return true;
            }
        }
        return false;
    }
};
using StackSpill = StackSpillT<kLdsDepth>;
using LaneStack = LaneStackT<kLdsDepth>;
template <int D>
constexpr size_t stack_lds_bytes() {
    return 2 * sizeof(uint32_t) * D * kBlock;
}
constexpr size_t kStackLdsBytes = stack_lds_bytes<kLdsDepth>();

// Distance cut of the t-pruning: boxes entered beyond it cannot hold a
// triangle whose reference t is <= best t (relative margin 2^-10 plus 4 e).
__device__ __forceinline__ float prune_cut(const RayPre& r, float best_t) {
    return best_t * 1.0009765625f + 4.f * r.e;
}

// Reference triangle tests of one leaf (kdtree.cpp:308-327), folded into h
// with the total order (t, then the larger leaf key, then the earlier slot).
// The leaf key (wr_tri::leaf, the same in every record of a leaf) orders
// leaves like their first records, the way kdtree.cpp:331-346 breaks ties.
template <bool NOSELF, bool COUNT>
__device__ __forceinline__ void visit_leaf(const Tree& tr, const RayPre& r, const NoSelf& ns, uint32_t ref,
                                           HitState& h, Counters& cnt) {
    const uint32_t first = ref & WR_FIRST_MASK;
    uint32_t c = (ref >> WR_CNT_SHIFT) & WR_CNT_MASK;
    if (c == 0) c = tr.leaf_cnt[first];
    for (uint32_t k = 0; k < c; k++) {
        const wr_tri T = tr.tris[first + k];
        if (COUNT) cnt.tris++;
        // an equal t loses to a leaf with a larger key, and to an earlier slot of this leaf
        // (once a slot of it is accepted, h.leaf is this leaf's key)
        const bool tie_loses = h.rec != WR_NONE && T.leaf <= h.leaf;
        const float t = ref_tri(T, r.o, r.d, h.t, tie_loses);
        if (t >= 0.f) {
            if (!NOSELF || noself_legal(ns, T, first + k, r.o, r.d, t)) {
                h.t = t;
                h.rec = first + k;
                h.leaf = T.leaf;
            }
        }
    }
}

// Reference triangle test of one record reached through the wide hierarchy
// (its own box), folded into h with the same total order as visit_leaf: the
// leaf key is the record's reference leaf (carried in the record).
template <bool NOSELF, bool COUNT>
__device__ __forceinline__ void visit_tri_rec(const Tree& tr, const RayPre& r, const NoSelf& ns, uint32_t rec,
                                              const wr_tri& T, HitState& h, Counters& cnt) {
    if (COUNT) cnt.tris++;
    const float t = ref_tri(T, r.o, r.d, h.t, false);
    if (!(t >= 0.f)) return;
    const uint32_t f = T.leaf;
    // t <= h.t here; on equality the larger leaf key wins, then the earlier slot
    const bool better = h.rec == WR_NONE || t < h.t || f > h.leaf || (f == h.leaf && rec < h.rec);
    if (better && (!NOSELF || noself_legal(ns, T, rec, r.o, r.d, t))) {
        h.t = t;
        h.rec = rec;
        h.leaf = f;
    }
}

// Ordered traversal of the reference binary tree.  EXACT = false: the fast
// conservative pass; EXACT = true: every box also passes the reference test.
template <bool NOSELF, bool EXACT, bool COUNT, class Stack>
__device__ __forceinline__ void traverse(const Tree& tr, const RayPre& r, const NoSelf& ns, Stack& st,
                                         HitState& h, bool& overflow, Counters& cnt) {
    h.t = __builtin_inff();
    h.rec = WR_NONE;
    h.leaf = 0;
    overflow = false;
    st.sp = 0;
    if (tr.root_ref == WR_NONE) return;
    {
        float tn;
        if (!slab(r, tr.rmin[0], tr.rmin[1], tr.rmin[2], tr.rmax[0], tr.rmax[1], tr.rmax[2], __builtin_inff(), tn))
            return;
        if (EXACT && !ref_box(r, tr.rmin, tr.rmax)) return;
    }
    uint32_t ref = tr.root_ref;
    for (;;) {
        if (ref & WR_LEAF_BIT) {
            visit_leaf<NOSELF, COUNT>(tr, r, ns, ref, h, cnt);
            if (!st.next(prune_cut(r, h.t), ref)) break;
            continue;
        }
        const wr_node N = tr.nodes[ref];
        if (COUNT) cnt.nodes += 2;
        const float tcut = prune_cut(r, h.t);
        float tl, tr_;
        bool hl = slab(r, N.lmin[0], N.lmin[1], N.lmin[2], N.lmax[0], N.lmax[1], N.lmax[2], tcut, tl);
        bool hr = slab(r, N.rmin[0], N.rmin[1], N.rmin[2], N.rmax[0], N.rmax[1], N.rmax[2], tcut, tr_);
        if (EXACT) {
            hl = hl && ref_box(r, N.lmin, N.lmax);
            hr = hr && ref_box(r, N.rmin, N.rmax);
        }
        if (hl && hr) {
            const bool left_first = tl <= tr_;
            if (st.sp >= kMaxDepth) {
                overflow = true;
                return;
            }
            st.push(left_first ? N.rref : N.lref, left_first ? tr_ : tl);
            ref = left_first ? N.lref : N.rref;
        } else if (hl) {
            ref = N.lref;
        } else if (hr) {
            ref = N.rref;
        } else if (!st.next(tcut, ref)) {
            break;
        }
    }
}

// Conditional swap of two (distance, reference) pairs into ascending order.
__device__ __forceinline__ void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
    const bool s = tb < ta;
    const float t = s ? tb : ta;
    const uint32_t r = s ? rb : ra;
    tb = s ? ta : tb;
    rb = s ? ra : rb;
    ta = t;
    ra = r;
}

// Fast conservative pass over the wide hierarchy.  Every decoded child box
// contains the reference box of the same subtree (outward rounding on the
// host, and the decode below is the host's fmaf), the slab test is monotone
// in the box, and skipping the intermediate level's boxes only adds
// candidates, so the set of triangles tested is a superset of the binary fast
// pass's, with the same leaves and first-record tie keys.
//
// The pass is split into wide_begin / wide_step so a persistent kernel can
// interleave the steps of many rays on one lane.
template <class Stack>
__device__ __forceinline__ bool wide_begin(const Tree& tr, const RayPre& r, Stack& st, HitState& h,
                                           uint32_t& ref) {
    h.t = __builtin_inff();
    h.rec = WR_NONE;
    h.leaf = 0;
    st.sp = 0;
    ref = tr.wide_root;
    if (ref == WR_NONE) return false;
    float tn;
    return slab(r, tr.rmin[0], tr.rmin[1], tr.rmin[2], tr.rmax[0], tr.rmax[1], tr.rmax[2], __builtin_inff(), tn);
}

// Next stack entry that can still hold a hit, WR_NONE when there is none.
template <class Stack>
__device__ __forceinline__ uint32_t next_or_end(Stack& st, float tcut) {
    uint32_t ref;
    return st.next(tcut, ref) ? ref : WR_NONE;
}

// Visit wide node `ref`: push the far children the ray enters, farthest
// first, and return the nearest one, else the next stack entry, else WR_NONE
// (pass ended).  Sets overflow (and returns WR_NONE) when the stack is full:
// the caller then re-runs the ray on the binary tree.
// The node is four 16 B words: origin[3] scale[3] | q rows (one dword per
// bound, child k in byte k) | ref[4].
template <bool COUNT, class Stack>
__device__ __forceinline__ uint32_t wide_node(const RayPre& r, Stack& st, float best_t, const uint4& n0,
                                              const uint4& n1, const uint4& n2, const uint4& n3, bool& overflow,
                                              Counters& cnt) {
    if (COUNT) cnt.nodes += 2;  // node bytes in 32 B units, as for a binary node
    const float tcut = prune_cut(r, best_t);
    const float ox = __uint_as_float(n0.x), oy = __uint_as_float(n0.y), oz = __uint_as_float(n0.z);
    const float sx = __uint_as_float(n0.w), sy = __uint_as_float(n1.x), sz = __uint_as_float(n1.y);
    // rows of the near and far bounds per axis (min row where 1/d > 0, else max)
    const bool px = r.ix > 0.f, py = r.iy > 0.f, pz = r.iz > 0.f;
    const uint32_t qnx = px ? n1.z : n2.y, qfx = px ? n2.y : n1.z;
    const uint32_t qny = py ? n1.w : n2.z, qfy = py ? n2.z : n1.w;
    const uint32_t qnz = pz ? n2.x : n2.w, qfz = pz ? n2.w : n2.x;
    const uint32_t cref[4] = {n3.x, n3.y, n3.z, n3.w};
    float t[4];
    uint32_t c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int sh = 8 * k;
        const float nx = __builtin_fmaf((float)((qnx >> sh) & 0xFFu), sx, ox);
        const float ny = __builtin_fmaf((float)((qny >> sh) & 0xFFu), sy, oy);
        const float nz = __builtin_fmaf((float)((qnz >> sh) & 0xFFu), sz, oz);
        const float fx = __builtin_fmaf((float)((qfx >> sh) & 0xFFu), sx, ox);
        const float fy = __builtin_fmaf((float)((qfy >> sh) & 0xFFu), sy, oy);
        const float fz = __builtin_fmaf((float)((qfz >> sh) & 0xFFu), sz, oz);
        float tn;
        const bool hit = slab_nf(r, nx, ny, nz, fx, fy, fz, tcut, tn) && cref[k] != WR_NONE;
        // a hit's key is finite, a miss's +inf, so misses sort last
        t[k] = hit ? fminf(tn, 3.0e38f) : __builtin_inff();
        c[k] = cref[k];
    }
    cswap(t[0], c[0], t[1], c[1]);
    cswap(t[2], c[2], t[3], c[3]);
    cswap(t[0], c[0], t[2], c[2]);
    cswap(t[1], c[1], t[3], c[3]);
    cswap(t[1], c[1], t[2], c[2]);
    if (st.sp + 3 > kMaxDepth) {
        overflow = true;
        return WR_NONE;
    }
    if (t[3] != __builtin_inff()) st.push(c[3], t[3]);
    if (t[2] != __builtin_inff()) st.push(c[2], t[2]);
    if (t[1] != __builtin_inff()) st.push(c[1], t[1]);
    if (t[0] != __builtin_inff()) return c[0];
    return next_or_end(st, tcut);
}

// Visit `ref` (a wide node or a leaf) and pick the next reference; false when
// the pass has ended (stack exhausted) or the stack would overflow (overflow
// set: the caller re-runs the ray on the binary tree).
//
// The record of the step is fetched before the lanes part ways: the wide node
// (64 B) at a node, the triangle record (64 B) at a leaf, with the same four
// loads.  A step where some lanes sit at nodes and others at triangles then
// waits for memory once instead of once per branch.
template <bool NOSELF, bool COUNT, class Stack>
__device__ __forceinline__ bool wide_step(const Tree& tr, const RayPre& r, const NoSelf& ns, Stack& st,
                                          HitState& h, uint32_t& ref, bool& overflow, Counters& cnt) {
    const bool leaf = (ref & WR_LEAF_BIT) != 0;  // always a single triangle in the wide hierarchy
    const uint32_t rec = ref & WR_FIRST_MASK;
    const uint4* p = leaf ? reinterpret_cast<const uint4*>(tr.tris + rec) : reinterpret_cast<const uint4*>(tr.wnodes + ref);
    const uint4 d0 = p[0], d1 = p[1], d2 = p[2], d3 = p[3];
    // keeps the fourth load ahead of the branch (it would otherwise sink into
    // the node branch and wait on its own)
    asm volatile("" ::"v"(d3.x));
    if (leaf) {
        wr_tri T;
        T.v0[0] = __uint_as_float(d0.x); T.v0[1] = __uint_as_float(d0.y); T.v0[2] = __uint_as_float(d0.z);
        T.v1[0] = __uint_as_float(d0.w); T.v1[1] = __uint_as_float(d1.x); T.v1[2] = __uint_as_float(d1.y);
        T.v2[0] = __uint_as_float(d1.z); T.v2[1] = __uint_as_float(d1.w); T.v2[2] = __uint_as_float(d2.x);
        T.n[0] = __uint_as_float(d2.y); T.n[1] = __uint_as_float(d2.z); T.n[2] = __uint_as_float(d2.w);
        T.leaf = d3.x;
        T.prim = d3.y;
        T.obj = d3.z;
        T.pad = d3.w;
        visit_tri_rec<NOSELF, COUNT>(tr, r, ns, rec, T, h, cnt);
        return st.next(prune_cut(r, h.t), ref);
    }
    ref = wide_node<COUNT>(r, st, h.t, d0, d1, d2, d3, overflow, cnt);
    return ref != WR_NONE;
}

template <bool NOSELF, bool COUNT>
__device__ __forceinline__ void traverse_wide(const Tree& tr, const RayPre& r, const NoSelf& ns, LaneStack& st,
                                              HitState& h, bool& overflow, Counters& cnt) {
    overflow = false;
    uint32_t ref;
    if (!wide_begin(tr, r, st, h, ref)) return;
    while (wide_step<NOSELF, COUNT>(tr, r, ns, st, h, ref, overflow, cnt)) {
    }
}

// Re-check the winner's root-to-leaf chain with the reference box test.  The
// ten levels nearest the leaf come from the leaf's chain record (two lines,
// five levels each; the leaf key is its index, so its fetch does not wait for
// the winner's triangle record); the levels above them from the chain's top
// record (the boxes up to the root, shared by every chain that ends below the
// same link), five at a time with their loads issued together.
struct Box6 {
    float mn[3], mx[3];
};

// The first n (at most five) of the five boxes of 6 floats at p, all five
// loaded before any is tested (every caller's p has five boxes' room).
__device__ __forceinline__ bool chain_five(const RayPre& r, const float* p, uint32_t n) {
    float b[5][6];
#pragma unroll
    for (int k = 0; k < 5; k++)
#pragma unroll
        for (int a = 0; a < 6; a++) b[k][a] = p[6 * k + a];
#pragma unroll
    for (int k = 0; k < 5; k++)
        if ((uint32_t)k < n && !ref_box(r, b[k], b[k] + 3)) return false;
    return true;
}

// leaf: the winner's leaf key, WR_NONE when the tree is a single leaf.
__device__ __forceinline__ bool verify_chain(const Tree& tr, const RayPre& r, uint32_t leaf) {
    if (leaf == WR_NONE) return ref_box(r, tr.rmin, tr.rmax);
    const wr_chain* C = tr.chains + leaf;
    const uint32_t count = C->count;
    const uint32_t up = C->next;
    const uint32_t top = C->top, ntop = C->top_count;
    if (!chain_five(r, &C->box[0][0], count < 5 ? count : 5)) return false;
    if (count > 5 && !chain_five(r, &C->box[5][0], count - 5)) return false;
    if (up != WR_NONE) {
        const float* T = tr.top_boxes + 6 * (size_t)top;
        for (uint32_t k = 0; k < ntop; k += 5) {
            const uint32_t m = ntop - k;
            if (!chain_five(r, T + 6 * (size_t)k, m < 5 ? m : 5)) return false;
        }
    }
    return ref_box(r, tr.rmin, tr.rmax);
}

// Completion of a fast-pass result: the binary fast pass when the wide stack
// overflowed, chain verification, and the exact fallback when it fails.
template <bool NOSELF, bool COUNT, class Stack>
__device__ __forceinline__ void finish_hit(const Tree& tr, const RayPre& r, const NoSelf& ns, Stack& st,
                                           Counters& cnt, HitState& h, bool overflow, bool& failed) {
    failed = false;
    if (overflow) traverse<NOSELF, false, COUNT>(tr, r, ns, st, h, overflow, cnt);
    if (!overflow && (h.rec == WR_NONE || verify_chain(tr, r, h.leaf))) return;
    if (COUNT) cnt.fallback++;
    traverse<NOSELF, true, COUNT>(tr, r, ns, st, h, overflow, cnt);
    failed = overflow;
}

// Full closest-hit query: fast pass (wide hierarchy; the binary tree if the
// wide stack overflows), chain verification, exact fallback.
template <bool NOSELF, bool COUNT>
__device__ __forceinline__ HitState closest_hit(const Tree& tr, const RayPre& r, const NoSelf& ns, LaneStack& st,
                                                Counters& cnt, bool& failed) {
    HitState h;
    bool overflow;
    traverse_wide<NOSELF, COUNT>(tr, r, ns, st, h, overflow, cnt);
    finish_hit<NOSELF, COUNT>(tr, r, ns, st, cnt, h, overflow, failed);
    return h;
}

// Scene::findInsideObject for one ray; objects whose root box does not contain
// the origin are skipped: one of the two opposite half-rays then provably fails
// the reference box test of that object's root (its face points move away from
// the box along the separating axis), so the reference skips them too.
__device__ __forceinline__ uint32_t find_inside(const SceneDev& S, const f3& o, const f3& d, uint32_t cur,
                                                LaneStack& st) {
    NoSelf ns{};
    Counters cnt{0, 0, 0};
    const f3 db = mk(-d.x, -d.y, -d.z);
    if (!(S.grid_min[0] <= o.x && o.x <= S.grid_max[0] && S.grid_min[1] <= o.y && o.y <= S.grid_max[1] &&
          S.grid_min[2] <= o.z && o.z <= S.grid_max[2]))
        return WR_NONE;  // outside every object's box
    const uint32_t cx = wr_grid_cell(o.x, S.grid_min[0], S.grid_inv[0], S.grid_dim[0]);
    const uint32_t cy = wr_grid_cell(o.y, S.grid_min[1], S.grid_inv[1], S.grid_dim[1]);
    const uint32_t cz = wr_grid_cell(o.z, S.grid_min[2], S.grid_inv[2], S.grid_dim[2]);
    const uint32_t cell = (cz * S.grid_dim[1] + cy) * S.grid_dim[0] + cx;
    const uint32_t c_end = S.cell_start[cell + 1];
    for (uint32_t k = S.cell_start[cell]; k < c_end; k++) {  // descending object ids (scene.cpp:121)
        const uint32_t i = S.cell_objs[k];
        if (i == cur) continue;
        const wr_object_gpu& ob = S.objs[i];
        if (!(ob.aabb_min[0] <= o.x && o.x <= ob.aabb_max[0] && ob.aabb_min[1] <= o.y && o.y <= ob.aabb_max[1] &&
              ob.aabb_min[2] <= o.z && o.z <= ob.aabb_max[2]))
            continue;
        const wr_tree_hdr H = S.hdrs[ob.tree];
        const Tree tr = make_tree(H, S);
        bool failed;
        const float* nrm_base = ob.has_normals ? S.rec_nrm + 9 * (size_t)H.rec_base : nullptr;
        {
            const RayPre r = make_ray(o, db, tr.scale);
            const HitState h = closest_hit<false, false>(tr, r, ns, st, cnt, failed);
            if (h.rec == WR_NONE) continue;
            float u1, u2;
            const f3 N = world_normal(tr.tris[h.rec], nrm_base ? nrm_base + 9 * (size_t)h.rec : nullptr,
                                      add(o, muls(db, h.t)), u1, u2);
            if (dot(N, db) < 0.f) continue;
        }
        {
            const RayPre r = make_ray(o, d, tr.scale);
            const HitState h = closest_hit<false, false>(tr, r, ns, st, cnt, failed);
            if (h.rec == WR_NONE) continue;
            float u1, u2;
            const f3 N = world_normal(tr.tris[h.rec], nrm_base ? nrm_base + 9 * (size_t)h.rec : nullptr,
                                      add(o, muls(d, h.t)), u1, u2);
            if (dot(N, d) < 0.f) continue;
        }
        return (uint32_t)i;
    }
    return WR_NONE;
}

}  // namespace wrd
