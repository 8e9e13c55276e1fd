/*
 * oracle.cpp -- CPU oracle for the WalkRay path-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithms in
 * plain C++ (libstdc++ + OpenMP, nothing else) so that the tests can check the
 * HIP product against it.  It is compiled with -ffp-contract=off on x86-64
 * (SSE2 scalar float, IEEE single precision, correctly rounded + - * / sqrt),
 * which is the arithmetic the reference's own operation order implies.
 *
 * What is faithful to the reference and what is not:
 *   - KDTree::build(BEST) / splitBest  (kdtree.cpp:141-258): faithful.  The
 *     reference collects unique vertex indices in an unordered_set; only the
 *     multiset of their coordinates matters to nth_element's result, so a
 *     sorted-unique vector is used instead.
 *   - Node::BoundingBox::intersect (kdtree.cpp:46-64), triangle test
 *     (kdtree.cpp:279-306), recursive unpruned traversal with its tie rules
 *     (kdtree.cpp:308-353): faithful, same operation order.
 *   - NoSelfCondition (noself.cpp:13-20), Mesh::getWorldNormal (mesh.h:47-77),
 *     Scene::intersect / findInsideObject (scene.cpp:102-144): faithful, with
 *     identity object transforms (all scenes here are generated in world space).
 *   - Camera::generateRay / fixVignetting (camera.cpp:6-60, 90-123), materials
 *     (diffuseObj.h, glossyObj.h, isotropicPhong.h, refractiveObj.h,
 *     emissiveObj.h, frame.h, cos/powerCos samplers), samplePath
 *     (mcRenderer.cpp:16-57) and the PathTracer pixel loop
 *     (pathTracer.cpp:20-70): faithful estimator.  The reference's single
 *     shared std::mt19937_64 (rng.h:6-59, racy across OpenMP threads) is
 *     replaced by a counter RNG keyed by (seed, pixel, pass, vertex): this
 *     changes which random numbers are drawn, not the estimator.
 *   - The running-mean film (pathTracer.cpp:24,69) is replaced by FP64 sums of
 *     the per-pass sample values and of their squares, so tests can form the
 *     mean and its standard error.
 *   - The per-pass tree rebuild (pathTracer.cpp:18 -> scene.cpp:67-80) is not
 *     repeated: the builder is deterministic, so every pass would rebuild the
 *     identical tree.
 *   - BidirectionalPathTracer (bidirPathTracer.cpp:10-313: eye and light
 *     sub-paths, every connection with its MIS weight, light-tracing splats):
 *     restated as the checker of a later GPU estimator.  Pinned as a whole by
 *     its limit: at a depth where both estimators have converged it agrees
 *     with the PathTracer image within the combined standard error (the two
 *     differ at small depths because BDPT's maxPathLen bounds each sub-path,
 *     not the joined path).  The individual MIS terms are parity unpinned.
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>

namespace orc {

static const double PI_D = 3.1415926535897932384626433832795;      /* macros.h:5 */
static const double INV_PI_D = 0.31830988618379067153776752674503; /* macros.h:6 */
static const double EPS_D = 1e-5;                                  /* macros.h:7 */

/* ---- vec3 with the reference's float semantics (nvVector.h:363-525) ---- */
struct V {
    float x, y, z;
    V() : x(0), y(0), z(0) {}
    V(float a, float b, float c) : x(a), y(b), z(c) {}
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
static inline V add(V a, V b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V sub(V a, V b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V neg(V a) { return V(-a.x, -a.y, -a.z); }
static inline V mul(V a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline V mulv(V a, V b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
/* operator/=(vec3, T): a zero divisor leaves the vector unchanged (nvVector.h:375-379). */
static inline V divs(V a, float s) {
    if (s == 0) return a;
    return V(a.x / s, a.y / s, a.z / s);
}
static inline V cross(V a, V b) { /* nvVector.h:490-499 */
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float dot(V a, V b) { /* nvVector.h:501-507: r = 0; r += ... */
    float r = 0.f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
static inline float length(V a) { return std::sqrt(dot(a, a)); }
static inline V normalize(V a) { /* nvVector.h:520-525 */
    float r = length(a);
    if (r > 0) return divs(a, r);
    return a;
}
static inline float maxc(V a) { /* std::max_element over x,y,z */
    float m = a.x;
    if (a.y > m) m = a.y;
    if (a.z > m) m = a.z;
    return m;
}
static inline bool legal(V c) { /* util.h:8-18 isLegalColor */
    for (int i = 0; i < 3; i++) {
        if (std::isnan(c[i])) return false;
        if (!std::isfinite(c[i])) return false;
        if (c[i] < 0) return false;
    }
    return true;
}

/* ---- KDTree ---------------------------------------------------------- */
struct Ray { V o, d; };

struct Node {
    Node* left = nullptr;
    Node* right = nullptr;
    std::vector<unsigned> tris;
    float box[6];  /* minCoord xyz, maxCoord xyz: (float*)this order of kdtree.cpp:52 */
};

struct Tree {
    std::vector<V> verts;
    std::vector<unsigned> idx;  /* 3 per triangle */
    Node* root = nullptr;
    unsigned depth = 0;

    ~Tree() { destroy(root); }
    void destroy(Node* n) {
        if (!n) return;
        destroy(n->left);
        destroy(n->right);
        delete n;
    }

    /* kdtree.cpp:141-240 (splitBest).  The two subtrees of a node are built as
     * independent OpenMP tasks when the node is large: each subtree depends only
     * on its own triangle list, so the tree is the serial recursion's. */
    void splitBest(Node* node, unsigned d) {
#pragma omp critical(orc_depth)
        if (d > depth) depth = d;
        float bestRatio = 10;
        unsigned bestDim = 0;
        float bestSplit = 0;
        const size_t n = node->tris.size();
        node->left = node->right = nullptr;
        if (n == 0) return;
        /* Tight box over every vertex of every triangle, in list order
         * (boundvec<1>/<0>, kdtree.cpp:5-21,151-165). */
        V mn = verts[idx[3 * node->tris[0]]], mx = mn;
        std::vector<unsigned> uniq;
        uniq.reserve(3 * n);
        for (size_t i = 0; i < n; i++)
            for (int k = 0; k < 3; k++) {
                unsigned vi = idx[3 * node->tris[i] + k];
                V p = verts[vi];
                mn.x = p.x < mn.x ? p.x : mn.x;
                mn.y = p.y < mn.y ? p.y : mn.y;
                mn.z = p.z < mn.z ? p.z : mn.z;
                mx.x = p.x > mx.x ? p.x : mx.x;
                mx.y = p.y > mx.y ? p.y : mx.y;
                mx.z = p.z > mx.z ? p.z : mx.z;
                uniq.push_back(vi);
            }
        node->box[0] = mn.x; node->box[1] = mn.y; node->box[2] = mn.z;
        node->box[3] = mx.x; node->box[4] = mx.y; node->box[5] = mx.z;
        if (n <= 1) return;  /* maxLeafTriangleNum = 1 (kdtree.cpp:25,170) */
        std::sort(uniq.begin(), uniq.end());
        uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        std::vector<float> coord(uniq.size());
        for (unsigned dim = 0; dim < 3; dim++) {
            for (size_t i = 0; i < uniq.size(); i++) coord[i] = verts[uniq[i]][dim];
            std::nth_element(coord.begin(), coord.begin() + (coord.size() >> 1), coord.end());
            float split = coord[coord.size() >> 1];  /* kdtree.cpp:179-183 */
            unsigned leftCount = 0;
            for (size_t i = 0; i < n; i++) {
                unsigned mask = 0;
                for (int k = 0; k < 3; k++)
                    mask += (verts[idx[3 * node->tris[i] + k]][dim] <= split) << k;
                if (mask != 0) leftCount++;
            }
            float ratio = leftCount / float(n);
            /* abs(ratio-0.5) < abs(bestRatio-0.5) in double (kdtree.cpp:202-208) */
            if (std::fabs((double)ratio - 0.5) < std::fabs((double)bestRatio - 0.5)) {
                bestRatio = ratio;
                bestDim = dim;
                bestSplit = split;
            }
        }
        if (bestRatio == 0 || bestRatio == 1) return;  /* kdtree.cpp:211-215 */
        node->left = new Node;
        node->right = new Node;
        for (size_t i = 0; i < n; i++) {  /* order-preserving partition, kdtree.cpp:220-236 */
            unsigned mask = 0;
            for (int k = 0; k < 3; k++)
                mask += (verts[idx[3 * node->tris[i] + k]][bestDim] <= bestSplit) << k;
            if (mask == 0) node->right->tris.push_back(node->tris[i]);
            else node->left->tris.push_back(node->tris[i]);
        }
        if (n >= 65536) {
            Node *l = node->left, *r = node->right;
#pragma omp task firstprivate(l, d)
            splitBest(l, d + 1);
#pragma omp task firstprivate(r, d)
            splitBest(r, d + 1);
#pragma omp taskwait
        } else {
            splitBest(node->left, d + 1);
            splitBest(node->right, d + 1);
        }
    }

    void build() { /* kdtree.cpp:242-258 */
        destroy(root);
        root = new Node;
        const unsigned nt = (unsigned)(idx.size() / 3);
        for (unsigned i = 0; i < nt; i++) root->tris.push_back(i);
        depth = 0;
        if (nt >= 65536) {
#pragma omp parallel
#pragma omp single
            splitBest(root, 0);
        } else {
            splitBest(root, 0);
        }
    }

    /* kdtree.cpp:279-306; cond is evaluated by the caller-supplied functor. */
    template <class Cond>
    float triIsect(const Ray& r, unsigned ti, const Cond& cond) const {
        V vs[3], lines[3], r2v, n;
        for (int i = 0; i < 3; i++) vs[i] = verts[idx[3 * ti + i]];
        unsigned mask_pos = 0, mask_neg = 0;
        for (int i = 0; i < 3; i++) {
            lines[i] = sub(vs[(i + 1) % 3], vs[i]);
            r2v = sub(vs[i], r.o);
            n = cross(lines[i], r2v);
            mask_pos += (dot(r.d, n) >= 0) << i;
            mask_neg += (dot(r.d, n) <= 0) << i;
        }
        n = cross(lines[0], lines[1]);
        n = normalize(n);
        float orthoDist = dot(n, r2v);
        float d = dot(n, r.d);
        if (d == 0 || orthoDist / d < 0 || (mask_pos != 7 && mask_neg != 7)) return -1;
        float dist = orthoDist / d;
        if (!cond(r, ti, dist)) return -1;
        return dist;
    }

    template <class Cond>
    float isect(const Ray& r, const Node* node, unsigned& tid, const Cond& cond) const {
        if (boxIsect(node->box, r) < 0) return -1;  /* kdtree.cpp:310-312 */
        if (!node->left && !node->right) {
            float minDist = -1;
            for (size_t k = 0; k < node->tris.size(); k++) {
                float dist = triIsect(r, node->tris[k], cond);
                if (dist >= 0 && (dist < minDist || minDist < 0)) {  /* earlier wins ties */
                    minDist = dist;
                    tid = node->tris[k];
                }
            }
            return minDist;
        }
        unsigned ltid = 0, rtid = 0;
        float ld = isect(r, node->left, ltid, cond);
        float rd = isect(r, node->right, rtid, cond);
        if (ld >= 0 && rd >= 0) {  /* equal t goes to the right subtree (kdtree.cpp:331-334) */
            tid = ld < rd ? ltid : rtid;
            return ld < rd ? ld : rd;
        } else if (rd < 0) {
            tid = ltid;
            return ld;
        } else if (ld < 0) {
            tid = rtid;
            return rd;
        }
        return -1;
    }

    template <class Cond>
    float intersect(const Ray& r, unsigned& tid, const Cond& cond) const {
        if (!root || root->tris.empty()) return -1;  /* empty tree: no box was ever set */
        return isect(r, root, tid, cond);
    }

    static float boxIsect(const float* b, const Ray& r) { /* kdtree.cpp:46-64 */
        float minDist = -1;
        for (unsigned i = 0; i < 6; i++) {
            const int a = i % 3;
            if (r.d[a] == 0) continue;
            float dist = (b[i] - r.o[a]) / r.d[a];
            V ip = add(r.o, mul(r.d, dist));
            bool inter = dist >= 0;
            const int a1 = (i + 1) % 3, a2 = (i + 2) % 3;
            inter &= b[a1] <= ip[a1];
            inter &= b[3 + a1] >= ip[a1];
            inter &= b[a2] <= ip[a2];
            inter &= b[3 + a2] >= ip[a2];
            minDist = (inter && (dist < minDist || minDist < 0)) ? dist : minDist;
        }
        return minDist;
    }
};

/* Counts, during an unpruned traversal, the triangles other than `win` that the
 * triangle test accepts at a distance within tol of t, and accepts none, so the
 * traversal visits every candidate (the near-ties of a closest hit). */
struct TieCount {
    float t, tol;
    unsigned win;
    mutable unsigned n = 0;
    bool operator()(const Ray&, unsigned ti, float dist) const {
        if (ti != win && std::fabs((double)dist - (double)t) < (double)tol) n++;
        return false;
    }
};

struct NoCond {
    bool operator()(const Ray&, unsigned, float) const { return true; }
};

/* ---- Scene --------------------------------------------------------------- */
/* ---- 4x4 matrices with nv::matrix4 semantics (nvMatrix.h), element(row, col) ---- */
struct M4 {
    float a[4][4];
    static M4 ident() {
        M4 m;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) m.a[i][j] = i == j ? 1.f : 0.f;
        return m;
    }
};
static M4 m4mul(const M4& l, const M4& r) { /* nvMatrix.h:409-421: r(i,j) = 0; += l(i,c) r(c,j) */
    M4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float x = 0.f;
            for (int c = 0; c < 4; c++) x += l.a[i][c] * r.a[c][j];
            o.a[i][j] = x;
        }
    return o;
}
static M4 m4transpose(const M4& m) {
    M4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o.a[i][j] = m.a[j][i];
    return o;
}
/* inverse (nvMatrix.h:242-378): Gauss-Jordan with scaled partial pivoting; a
 * singular input returns the default-constructed (identity) matrix. */
static M4 m4inverse(const M4& m) {
    float r[4][8];
    float* row[4] = {r[0], r[1], r[2], r[3]};
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            row[i][j] = m.a[i][j];
            row[i][j + 4] = i == j ? 1.f : 0.f;
        }
    float scp[4];
    for (int i = 0; i < 4; i++) {
        scp[i] = std::fabs(row[i][0]);
        for (int j = 1; j < 4; j++)
            if (std::fabs(row[i][j]) > scp[i]) scp[i] = std::fabs(row[i][j]);
        if (scp[i] == 0.f) return M4::ident();
    }
    for (int i = 0; i < 4; i++) {
        int to = i;
        float best = std::fabs(row[i][i] / scp[i]);
        for (int p = i + 1; p < 4; p++)
            if (std::fabs(row[p][i] / scp[p]) > best) {
                best = std::fabs(row[p][i] / scp[p]);
                to = p;
            }
        if (to != i) {
            std::swap(row[i], row[to]);
            std::swap(scp[i], scp[to]);
        }
        for (int j = i + 1; j < 4; j++) {
            const float mji = row[j][i] / row[i][i];
            row[j][i] = 0.f;
            for (int jj = i + 1; jj < 8; jj++) row[j][jj] -= mji * row[i][jj];
        }
    }
    if (row[3][3] == 0.f) return M4::ident();
    for (int i = 3; i > 0; i--)
        for (int j = i - 1; j > -1; j--) {
            const float mij = row[j][i] / row[i][i];
            for (int jj = j + 1; jj < 8; jj++) row[j][jj] -= mij * row[i][jj];
        }
    M4 o;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) o.a[i][j] = row[i][j + 4] / row[i][i];
    return o;
}
/* vec3(M * vec4(v, w)) (nvMatrix.h:424-433) */
static V m4apply(const M4& m, V v, float w) {
    V o;
    for (int i = 0; i < 3; i++) o[i] = v.x * m.a[i][0] + v.y * m.a[i][1] + v.z * m.a[i][2] + w * m.a[i][3];
    return o;
}

/* ---- Quaternion (quaternion.h, quaternion.cpp) ---- */
struct Q {
    V v;
    float w = 1.f;
};
static float qDot(const Q& a, const Q& b) { return dot(a.v, b.v) + a.w * b.w; }
static Q qScale(const Q& q, float f) { return Q{mul(q.v, f), q.w * f}; }
static Q qPlus(const Q& a, const Q& b) { return Q{add(a.v, b.v), a.w + b.w}; }
static Q qMinus(const Q& a, const Q& b) { return Q{sub(a.v, b.v), a.w - b.w}; }
static Q qDivide(const Q& q, float f) { return Q{divs(q.v, f), q.w / f}; }
static Q qNormalize(const Q& q) { return qDivide(q, std::sqrt(qDot(q, q))); }
static M4 qToTransform(const Q& q) { /* quaternion.cpp:6-31 */
    const float xx = q.v.x * q.v.x, yy = q.v.y * q.v.y, zz = q.v.z * q.v.z;
    const float xy = q.v.x * q.v.y, xz = q.v.x * q.v.z, yz = q.v.y * q.v.z;
    const float wx = q.v.x * q.w, wy = q.v.y * q.w, wz = q.v.z * q.w;
    M4 m = M4::ident();
    m.a[0][0] = 1.f - 2.f * (yy + zz);
    m.a[0][1] = 2.f * (xy + wz);
    m.a[0][2] = 2.f * (xz - wy);
    m.a[1][0] = 2.f * (xy - wz);
    m.a[1][1] = 1.f - 2.f * (xx + zz);
    m.a[1][2] = 2.f * (yz + wx);
    m.a[2][0] = 2.f * (xz + wy);
    m.a[2][1] = 2.f * (yz - wx);
    m.a[2][2] = 1.f - 2.f * (xx + yy);
    return m;
}
static Q qFromTransform(const M4& m) { /* quaternion.cpp:34-70 */
    Q q;
    const float trace = m.a[0][0] + m.a[1][1] + m.a[2][2];
    if (trace > 0.f) {
        float s = sqrtf(trace + 1.0);
        q.w = s / 2.0f;
        s = 0.5f / s;
        q.v.x = (m.a[2][1] - m.a[1][2]) * s;
        q.v.y = (m.a[0][2] - m.a[2][0]) * s;
        q.v.z = (m.a[1][0] - m.a[0][1]) * s;
    } else {
        const int nxt[3] = {1, 2, 0};
        float c[3];
        int i = 0;
        if (m.a[1][1] > m.a[0][0]) i = 1;
        if (m.a[2][2] > m.a[i][i]) i = 2;
        const int j = nxt[i], k = nxt[j];
        float s = sqrtf((m.a[i][i] - (m.a[j][j] + m.a[k][k])) + 1.0);
        c[i] = s * 0.5f;
        if (s != 0.f) s = 0.5f / s;
        q.w = (m.a[k][j] - m.a[j][k]) * s;
        c[j] = (m.a[j][i] + m.a[i][j]) * s;
        c[k] = (m.a[k][i] + m.a[i][k]) * s;
        q.v = V(c[0], c[1], c[2]);
    }
    return q;
}
static Q qSlerp(float t, const Q& q1, const Q& q2) { /* quaternion.cpp:73-84 */
    const float cosTheta = qDot(q1, q2);
    if (cosTheta > .9995f) return qNormalize(qPlus(qScale(q1, 1.f - t), qScale(q2, t)));
    const float theta = acosf(cosTheta < -1.f ? -1.f : cosTheta > 1.f ? 1.f : cosTheta);
    const float thetap = theta * t;
    const Q qperp = qNormalize(qMinus(q2, qScale(q1, cosTheta)));
    return qPlus(qScale(q1, cosf(thetap)), qScale(qperp, sinf(thetap)));
}

/* ---- AnimatedTransform (motionBlur.h, motionBlur.cpp) ---- */
struct Animated {
    float t0 = 0, t1 = 1;
    M4 start = M4::ident(), end = M4::ident();
    V T[2];
    Q R[2];
    M4 S[2];
    bool actually = false;
    static void decompose(const M4& m, V& T, Q& Rq, M4& S) { /* motionBlur.cpp:48-87 */
        T = V(m.a[0][3], m.a[1][3], m.a[2][3]);
        M4 M = m;
        for (int i = 0; i < 3; ++i) M.a[i][3] = M.a[3][i] = 0.f;
        M.a[3][3] = 1.f;
        float norm;
        int count = 0;
        M4 R = M;
        do {
            M4 Rnext;
            const M4 Rit = m4inverse(m4transpose(R));
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) Rnext.a[i][j] = 0.5f * (R.a[i][j] + Rit.a[i][j]);
            norm = 0.f;
            for (int i = 0; i < 3; ++i) {
                const float n = fabsf(R.a[i][0] - Rnext.a[i][0]) + fabsf(R.a[i][1] - Rnext.a[i][1]) +
                                fabsf(R.a[i][2] - Rnext.a[i][2]);
                norm = std::max(norm, n);
            }
            R = Rnext;
        } while (++count < 100 && norm > .0001f);
        Rq = qFromTransform(R);
        S = m4mul(m4inverse(R), M);
    }
    Animated() = default;
    Animated(const M4& a, float ta, const M4& b, float tb) : t0(ta), t1(tb), start(a), end(b) {
        decompose(start, T[0], R[0], S[0]);
        decompose(end, T[1], R[1], S[1]);
        bool same = true;
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) same = same && start.a[i][j] == end.a[i][j];
        actually = !same;
    }
    M4 interpolate(float time) const { /* motionBlur.cpp:6-45 */
        if (!actually) return start;
        if (time <= t0) return start;
        if (time >= t1) return end;
        const float dt = (time - t0) / (t1 - t0);
        const V trans = add(mul(T[0], 1.f - dt), mul(T[1], dt));
        const Q rotate = qSlerp(dt, R[0], R[1]);
        M4 scale = M4::ident();
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) scale.a[i][j] = (1.f - dt) * S[0].a[i][j] + dt * S[1].a[i][j];
        M4 I = M4::ident();
        for (int i = 0; i < 3; i++) I.a[i][3] = trans[i];
        return m4mul(m4mul(I, qToTransform(rotate)), scale);
    }
};

struct Object {
    std::vector<V> verts;
    std::vector<unsigned> idx;
    std::vector<V> normals;  /* empty: flat (mesh.h:59-65) */
    int mat;
    V color;      /* albedo or emitted radiance */
    float power;  /* Phong exponent */
    float ior;
    /* motion blur: the mesh in its own space and its AnimatedTransform */
    bool moving = false;
    Animated anim;
    std::vector<V> localVerts, localNormals;
    /* getWorldVertexPosition / getWorldNormal with transform M (mesh.h:47-77,106) */
    void place(const M4& M) {
        for (size_t i = 0; i < localVerts.size(); i++) verts[i] = m4apply(M, localVerts[i], 1.f);
        if (!localNormals.empty()) {
            const M4 nm = m4transpose(m4inverse(M));
            for (size_t i = 0; i < localNormals.size(); i++) normals[i] = m4apply(nm, localNormals[i], 0.f);
        }
    }
};

struct Scene {
    std::vector<Object> objs;
    Tree tree;                         /* global tree, scene.cpp:20-44 */
    std::vector<Tree*> objTrees;       /* per-object trees, scene.cpp:46-65 */
    std::vector<unsigned> srcObj, srcTri;  /* ObjSourceInfo per global triangle */
    ~Scene() { for (auto* t : objTrees) delete t; }

    void build() {
        tree.verts.clear(); tree.idx.clear(); srcObj.clear(); srcTri.clear();
        unsigned off = 0;
        for (unsigned oi = 0; oi < objs.size(); oi++) {
            const Object& o = objs[oi];
            for (const V& v : o.verts) tree.verts.push_back(v);
            const unsigned nt = (unsigned)(o.idx.size() / 3);
            for (unsigned i = 0; i < nt; i++) {
                for (int k = 0; k < 3; k++) tree.idx.push_back(o.idx[3 * i + k] + off);
                srcObj.push_back(oi);
                srcTri.push_back(i);
            }
            off += (unsigned)o.verts.size();
        }
        tree.build();
        prepareLights();
        for (auto* t : objTrees) delete t;
        objTrees.clear();
        for (unsigned oi = 0; oi < objs.size(); oi++) {
            Tree* t = new Tree;
            t->verts = objs[oi].verts;
            t->idx = objs[oi].idx;
            t->build();
            objTrees.push_back(t);
        }
    }

    /* Light sampling tables (Scene::prepareForLightSampling, scene.cpp:5-14;
     * SurfaceSampler::prepareForSampling, surfaceSampler.cpp:7-15;
     * AbstractObject::prepareForSurfaceSampling, abstractObj.cpp:9-16). */
    std::vector<unsigned> lightObj;           /* emissive objects, in object order */
    std::vector<float> lightCum;              /* SurfaceSampler::areaThreshold */
    std::vector<std::vector<float>> triCum;   /* per light: AbstractObject::areaThreshold */
    std::vector<float> lightArea, pickProb;   /* per light: boundArea, pickProb */
    float lightTotal = 0;                     /* SurfaceSampler::boundArea */

    /* Scene::updateSceneForMotionBlur (scene.cpp:67-80) before pass `pass`: every
     * moving object at its time (uniform, counter RNG with the object in the
     * pixel slot and vertex 0xFFFFFFFE), then every tree rebuilt.  Returns
     * false (and changes nothing) when no object moves. */
    bool updateForMotion(uint64_t seed, unsigned pass);

    /* Mesh::getTriangleArea (mesh.h:80-88) */
    float triangleArea(unsigned oi, unsigned ti) const {
        const Object& o = objs[oi];
        V vps[3];
        for (int i = 0; i < 3; i++) vps[i] = o.verts[o.idx[3 * ti + i]];
        V l0 = sub(vps[1], vps[0]);
        V l1 = sub(vps[2], vps[1]);
        float c = length(cross(l0, l1));
        return c / 2;
    }

    void prepareLights() {
        lightObj.clear(); lightCum.clear(); triCum.clear(); lightArea.clear(); pickProb.clear();
        lightTotal = 0;
        for (unsigned oi = 0; oi < objs.size(); oi++) {
            if (objs[oi].mat != ORC_EMISSIVE) continue;
            lightObj.push_back(oi);
            std::vector<float> cum;
            float areaValue = 0;
            const unsigned nt = (unsigned)(objs[oi].idx.size() / 3);
            for (unsigned i = 0; i < nt; i++) {
                areaValue += triangleArea(oi, i);
                cum.push_back(areaValue);
            }
            triCum.push_back(cum);
            lightArea.push_back(areaValue);
            lightTotal += areaValue;
            lightCum.push_back(lightTotal);
        }
        for (size_t i = 0; i < lightObj.size(); i++) pickProb.push_back(lightArea[i] / lightTotal);
    }

    /* Mesh::getWorldNormal (mesh.h:47-77) with an identity transform. */
    V worldNormal(unsigned oi, unsigned fi, V pos, float* u1o = nullptr, float* u2o = nullptr) const {
        const Object& o = objs[oi];
        V vps[3];
        for (int i = 0; i < 3; i++) vps[i] = o.verts[o.idx[3 * fi + i]];
        V b1 = sub(vps[1], vps[0]);
        V b2 = sub(vps[2], vps[0]);
        V face = normalize(cross(b1, b2));
        V v = sub(pos, vps[0]);
        float d12 = dot(b1, b2);
        float l1 = length(b1);
        float l2 = length(b2);
        float u2 = (dot(v, b1) * d12 - dot(v, b2) * l1 * l1) / (d12 * d12 - l1 * l1 * l2 * l2);
        float u1 = (dot(v, b1) - u2 * d12) / (l1 * l1);
        if (u1o) *u1o = u1;
        if (u2o) *u2o = u2;
        if (o.normals.empty()) return face;
        V n0 = o.normals[o.idx[3 * fi + 0]], n1 = o.normals[o.idx[3 * fi + 1]],
          n2 = o.normals[o.idx[3 * fi + 2]];
        float w0 = 1 - u1 - u2;
        V nn = add(add(mul(n0, w0), mul(n1, u1)), mul(n2, u2));
        return normalize(nn);
    }

    /* NoSelfCondition::legal (noself.cpp:13-20) */
    struct NoSelf {
        const Scene* s;
        unsigned inside, contact;
        bool operator()(const Ray& r, unsigned gtri, float dist) const {
            unsigned hitObj = s->srcObj[gtri];
            unsigned fi = s->srcTri[gtri];
            float d = dot(r.d, s->worldNormal(hitObj, fi, add(r.o, mul(r.d, dist))));
            bool in = d <= 0, out = d >= 0;
            if (inside == contact) return !(in && contact == hitObj);
            return !(out && inside != hitObj);
        }
    };

    /* Scene::intersect (scene.cpp:102-113) */
    float intersect(const Ray& r, unsigned& gtri, bool noself, unsigned inside,
                    unsigned contact) const {
        unsigned tid = ORC_NONE;
        float dist;
        if (noself) dist = tree.intersect(r, tid, NoSelf{this, inside, contact});
        else dist = tree.intersect(r, tid, NoCond{});
        if (dist >= 0) gtri = tid;
        return dist;
    }

    /* Scene::findInsideObject (scene.cpp:115-144) */
    unsigned findInside(const Ray& r, unsigned cur) const {
        Ray front = r, back;
        back.o = r.o;
        back.d = neg(r.d);
        for (int i = (int)objTrees.size() - 1; i >= 0; i--) {
            if (cur == (unsigned)i) continue;
            unsigned tid = 0;
            float dist = objTrees[i]->intersect(back, tid, NoCond{});
            if (dist < 0) continue;
            V n = worldNormal(i, tid, add(back.o, mul(back.d, dist)));
            if (dot(n, back.d) < 0) continue;
            dist = objTrees[i]->intersect(front, tid, NoCond{});
            if (dist < 0) continue;
            n = worldNormal(i, tid, add(front.o, mul(front.d, dist)));
            if (dot(n, front.d) < 0) continue;
            return (unsigned)i;
        }
        return ORC_NONE;
    }
};

/* ---- Counter RNG: Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) ---- */
static inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
static inline void rng4(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t vertex, float f[4]) {
    uint32_t c[4] = {pixel, pass, vertex, 0x57414C4Bu};
    uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    philox(c, k, o);
    for (int i = 0; i < 4; i++) f[i] = (float)(o[i] >> 8) * (1.0f / 16777216.0f);
}

/* Back to every moving object's start time, as the scene was before a render. */
static void restoreMotion(Scene* s) {
    bool any = false;
    for (Object& o : s->objs)
        if (o.moving) {
            o.place(o.anim.interpolate(o.anim.t0));
            any = true;
        }
    if (any) s->build();
}

bool Scene::updateForMotion(uint64_t seed, unsigned pass) {
    bool any = false;
    for (unsigned oi = 0; oi < objs.size(); oi++) {
        Object& o = objs[oi];
        if (!o.moving) continue;
        float u[4];
        rng4(seed, oi, pass, 0xFFFFFFFEu, u);
        o.place(o.anim.interpolate(u[0]));
        any = true;
    }
    if (any) build();
    return any;
}

/* Vertex counter of the light sample's uniforms (next-event estimation); path
 * vertices use 0 .. maxPathLen + 1. */
static const uint32_t kNeeVertex = 0xFFFFFFFFu;
/* First vertex counter of light sub-paths (counters kLightVertex + 0, 1, 2 + d). */
static const uint32_t kLightVertex = 0x80000000u;

/* ---- Camera (camera.cpp) -------------------------------------------------- */
struct Cam {
    V pos, focus, up, front, right, top;
    unsigned W, H;
    float dist;
    explicit Cam(const orc_camera* c) {
        pos = V(c->pos[0], c->pos[1], c->pos[2]);
        focus = V(c->focus[0], c->focus[1], c->focus[2]);
        up = V(c->up[0], c->up[1], c->up[2]);
        W = c->width; H = c->height; dist = c->sight_dist;
        front = normalize(sub(focus, pos));   /* camera.cpp:7-11 */
        right = normalize(cross(front, up));
        top = cross(right, front);
    }
    V dir(unsigned p, float u, float v) const { /* camera.cpp:13-23 */
        unsigned x = p % W, y = p / W;
        V d = add(add(mul(front, dist), mul(right, ((float)x + u) - (W / 2.f))),
                  mul(top, ((H / 2.f) - (float)y) - v));
        return normalize(d);
    }
    float pixelSolidAngle(unsigned p) const { /* Camera::getPixelSolidAngle (camera.cpp:124-131) */
        unsigned x = p % W, y = p / W;
        float xc = (float)(x + 0.5 - (W / 2.0)), yc = (float)((H / 2.0) - 0.5 - y);
        float pixelDist = std::sqrt(xc * xc + yc * yc + dist * dist);
        return dist / powf(pixelDist, 3);
    }
    void toRaster(V p, float& rx, float& ry) const { /* Camera::convToRaster (camera.cpp:148-162) */
        V v = sub(p, pos);
        float vFrontDist = dot(v, front);
        float scale = dist / vFrontDist;
        rx = (float)(W / 2.0 + dot(v, right) * scale);
        ry = (float)(H / 2.0 - dot(v, top) * scale);
    }
    V fixVignetting(V color, unsigned p) const { /* camera.cpp:107-122 */
        unsigned x = p % W, y = p / W;
        V d = add(add(mul(front, dist), mul(right, (float)(x + 0.5 - (W / 2.0)))),
                  mul(top, (float)((H / 2.0) - 0.5 - y)));
        d = normalize(d);
        if (dot(d, front) == 0) return V(0, 0, 0);
        V res = divs(color, std::fabs(dot(d, front)));
        return legal(res) ? res : V(0, 0, 0);
    }
};

/* ---- One path vertex produced by AbstractObject::scatter ----------------- */
struct Out {
    V o, d, rad, n;  /* n: contact normal, == contactObj->getWorldNormal(contactTri, o) */
    float op = 1.f, dp = 1.f;
    unsigned inside = ORC_NONE, contact = ORC_NONE, ctri = 0;
    bool delta = false;  /* isDeltaDirection: set by the refractive scatter (refractiveObj.h:85) */
};

/* Ray::cosineTerm (ray.cpp:52-58) for a scatter vertex (contactObj set, hasCosine true). */
static inline float cosineTerm(const Out& r) {
    float c = 1;
    if (length(r.d) > 0.5) c = std::fabs(dot(r.n, r.d));
    return c > EPS_D ? c : (float)EPS_D;
}

/* Frame::setFromN + toWorldFrame (frame.h:23-49) */
static inline V toWorld(V w, V local) {
    V mX, mY = w, mZ;
    if (std::fabs(w.x) > EPS_D) {
        mX = normalize(cross(V(0, 1, 0), w));
        mZ = cross(mX, mY);
    } else {
        mZ = normalize(cross(V(1, 0, 0), w));
        mX = cross(mY, mZ);
    }
    return add(add(mul(mX, local.x), mul(mY, local.y)), mul(mZ, local.z));
}

struct Tracer {
    const Scene* s;
    Cam cam;
    unsigned maxDepth;
    uint64_t seed;

    struct Hit {
        bool valid = false;
        float dist = -1;
        unsigned obj = ORC_NONE, tri = 0;
    };

    /* dispatch of AbstractObject::scatter for the four hot materials */
    Out scatter(unsigned self, const Ray& in, const Hit& h, unsigned inInside, const float* rn) const {
        const Object& ob = s->objs[self];
        Out out;
        V position = add(in.o, mul(in.d, h.dist));
        V N = s->worldNormal(h.obj, h.tri, position);
        if (ob.mat == ORC_REFRACTIVE && h.obj != self && inInside == self)  /* refractiveObj.h:32-33 */
            return scatter(h.obj, in, h, inInside, rn);
        out.o = position;
        out.n = N;
        out.contact = self;
        out.ctri = h.tri;
        switch (ob.mat) {
        case ORC_DIFFUSE: { /* diffuseObj.h:34-57, diffuseBsdf.h:17-26, cosHemiSphereSampler.h:6-18 */
            out.inside = inInside;
            float p = maxc(ob.color);
            if (rn[0] < p) {
                const float term1 = (float)(2.0 * PI_D * (double)rn[1]);
                const float term2 = std::sqrt(1.f - rn[2]);
                V local(std::sin(term1) * term2, std::sqrt(rn[2]), std::cos(term1) * term2);
                out.d = toWorld(N, local);
                float c = dot(N, out.d);
                out.dp = c > 0 ? (float)((double)c * INV_PI_D) : 0.f;
                out.rad = mul(ob.color, (float)INV_PI_D);
                out.dp *= p;
            } else {
                out.d = V(0, 0, 0);
                out.rad = V(0, 0, 0);
                out.dp = 1 - p;
            }
            break;
        }
        case ORC_GLOSSY: { /* glossyObj.h:39-62, isotropicPhong.h:20-38, powerCosHemiSphereSampler.h:8-21 */
            out.inside = inInside;
            float p = maxc(ob.color);
            if (rn[0] < p) {
                const float power = ob.power;
                const float term1 = (float)(2.0 * PI_D * (double)rn[1]);
                const float term2 = (float)std::pow((double)rn[2], 1.0 / (power + 1.0));
                const float term3 = (float)std::sqrt(1.0 - (double)(term2 * term2));
                V local(std::sin(term1) * term3, term2, std::cos(term1) * term3);
                V w = toWorld(N, local);
                float c = dot(N, w);
                out.dp = c > 0 ? (float)((power + 1.0) * (double)std::pow(c, power) * (INV_PI_D * 0.5)) : 0.f;
                V refl = add(mul(mul(N, -dot(N, in.d)), 2.f), in.d);
                refl = normalize(refl);
                V rho = mul(mul(mul(ob.color, (float)(power + 2.0)), 0.5f), (float)INV_PI_D);
                float dRW = dot(refl, w);
                if (dot(w, N) < 0) {
                    out.d = V(0, 0, 0);
                    out.rad = V((float)EPS_D, (float)EPS_D, (float)EPS_D);
                } else {
                    out.d = w;
                    out.rad = mul(rho, std::pow(dRW > 0 ? dRW : 0.f, power));
                }
                out.dp *= p;
            } else {
                out.d = V(0, 0, 0);
                out.rad = V(0, 0, 0);
                out.dp = 1 - p;
            }
            break;
        }
        case ORC_REFRACTIVE: { /* refractiveObj.h:27-88 */
            out.d = in.d;
            out.rad = V(1, 1, 1);
            V refl = add(mul(mul(N, -dot(N, in.d)), 2.f), in.d);
            refl = normalize(refl);
            float theta = std::acos(dot(in.d, N));
            unsigned curIn = inInside, curOut = self;
            if (inInside == self) {
                Ray q;
                q.o = out.o;
                q.d = out.d;
                curOut = s->findInside(q, self);
            }
            float cn = curIn != ORC_NONE ? (s->objs[curIn].mat == ORC_REFRACTIVE ? s->objs[curIn].ior : 1.f) : 1.f;
            float nn = curOut != ORC_NONE ? (s->objs[curOut].mat == ORC_REFRACTIVE ? s->objs[curOut].ior : 1.f) : 1.f;
            float sin_phi = cn / nn * std::sin(theta);
            if (sin_phi > 1) {  /* total internal reflection: radiance 1, prob 1 */
                out.d = refl;
                out.inside = inInside;
            } else {
                float phi = std::asin(sin_phi);
                if (theta > PI_D / 2) phi = (float)(PI_D - phi);
                V axis = normalize(cross(N, in.d));
                /* RotateMatrix(axis, phi) * vec4f(N, 0)  (util.h:40-54, nvMatrix.h:98-105,424-433) */
                float c = std::cos(phi), sn = std::sin(phi);
                float _c = 1 - c;
                float x = axis.x, y = axis.y, z = axis.z;
                float m00 = c + _c * x * x, m01 = _c * x * y - sn * z, m02 = _c * x * z + sn * y;
                float m10 = _c * x * y + sn * z, m11 = c + _c * y * y, m12 = _c * y * z - sn * x;
                float m20 = _c * x * z - sn * y, m21 = _c * y * z + sn * x, m22 = c + _c * z * z;
                V rd;
                rd.x = N.x * m00 + N.y * m01 + N.z * m02 + 0.f * 0.f;
                rd.y = N.x * m10 + N.y * m11 + N.z * m12 + 0.f * 0.f;
                rd.z = N.x * m20 + N.y * m21 + N.z * m22 + 0.f * 0.f;
                out.d = normalize(rd);
                float cos_theta = std::fabs(std::cos(theta));
                float cos_phi = std::fabs(std::cos(phi));
                float a = std::fabs(cn * cos_theta - nn * cos_phi) / (cn * cos_theta + nn * cos_phi);
                float b = std::fabs(nn * cos_theta - cn * cos_phi) / (nn * cos_theta + cn * cos_phi);
                float esr = a * a, epr = b * b;
                float er = (esr + epr) / 2;
                float p = er;
                if (rn[0] < p) {
                    out.d = refl;
                    out.rad = mul(out.rad, er / cosineTerm(out));
                    out.dp = p;
                    out.inside = inInside;
                } else {
                    out.rad = mul(out.rad, (1 - er) / cosineTerm(out));
                    out.dp = 1 - p;
                    out.inside = curOut;
                }
            }
            out.d = normalize(out.d);
            out.delta = true;
            break;
        }
        default: { /* ORC_EMISSIVE, emissiveObj.h:15-28 */
            out.d = V(0, 0, 0);
            out.inside = ORC_NONE;
            out.rad = dot(in.d, N) <= 0 ? ob.color : V(0, 0, 0);
            break;
        }
        }
        return out;
    }

    Hit trace(const Ray& r, bool noself, unsigned inside, unsigned contact) const {
        Hit h;
        unsigned g = 0;
        float d = s->intersect(r, g, noself, inside, contact);
        if (d >= 0) {
            h.valid = true;
            h.dist = d;
            h.obj = s->srcObj[g];
            h.tri = s->srcTri[g];
        }
        return h;
    }

    /* One eye path (mcRenderer.cpp:16-57) and its weight (pathTracer.cpp:22-69). */
    V sample(unsigned p, unsigned pass, unsigned* nv) const {
        float rn[4];
        rng4(seed, p, pass, 0, rn);
        Ray ray;
        ray.o = cam.pos;
        ray.d = cam.dir(p, rn[0], rn[1]);
        Hit hit = trace(ray, false, 0, 0);
        unsigned inside = s->findInside(ray, ORC_NONE);
        /* camera vertex: cosineTerm = |front . d| (camera.cpp:182-186, ray.cpp:52-58) */
        float c0 = std::fabs(dot(normalize(sub(cam.focus, cam.pos)), ray.d));
        c0 = c0 > EPS_D ? c0 : (float)EPS_D;
        V T = mul(V(1, 1, 1), c0);
        T = mulv(T, divs(divs(V(1, 1, 1), 1.f), 1.f));
        V color(0, 0, 0);
        unsigned depth = 0, verts = 1;
        for (;;) {
            unsigned self;
            if (inside != ORC_NONE) self = inside;
            else if (hit.valid) self = hit.obj;
            else { verts++; break; }  /* terminateRay: contributes 0 */
            /* inside an object but no hit (an open mesh): the reference would
             * dereference a NULL intersectObj; treated as terminateRay */
            if (!hit.valid) { verts++; break; }
            rng4(seed, p, pass, depth + 1, rn);
            Out out = scatter(self, ray, hit, inside, rn);
            verts++;
            if (length(out.d) < 0.5) {  /* emission / absorption ends the path */
                color = mulv(T, divs(divs(out.rad, out.op), out.dp));
                break;
            }
            if (depth + 1 > maxDepth) { verts++; break; }
            Ray nr;
            nr.o = out.o;
            nr.d = out.d;
            Hit nh = trace(nr, true, out.inside, out.contact);
            if (!nh.valid) { verts++; break; }
            /* out becomes an inner vertex: cosine term, decay (1 for these materials), weight */
            T = mul(T, cosineTerm(out));
            T = mulv(T, divs(divs(out.rad, out.op), out.dp));
            ray = nr;
            hit = nh;
            inside = out.inside;
            depth++;
        }
        if (nv) *nv = verts;
        if (!legal(color)) color = V(0, 0, 0);
        return cam.fixVignetting(color, p);
    }

    /* ---- The same estimator restated over the explicit Path of Rays ----------
     * samplePath (mcRenderer.cpp:16-57) builds the vector of rays, terminateRay
     * entries included; the pixel loop body (pathTracer.cpp:27-68) then forms
     * the sample with either branch.  This is the form the next-event branch
     * needs (it indexes the path from its end); with nee == false it must
     * equal sample() bit for bit, which the tests check. */
    struct PRay {               /* the Ray fields the estimator reads (ray.h:8-37) */
        V o, d, rad = V(1, 1, 1);
        V n;                    /* contactNormal() at o; for the camera ray normalize(focus - o) */
        float op = 1.f, dp = 1.f;
        unsigned inside = ORC_NONE, contact = ORC_NONE, ctri = 0;
        bool camera = false;    /* contactObj is the Camera (camera.cpp:54) */
        bool delta = false;     /* isDeltaDirection */
        unsigned pixel = 0;     /* pixelID of the camera ray */
        Hit hit;                /* intersectObj / intersectDist / intersectTriangleID */
    };

    /* Ray::cosineTerm (ray.cpp:52-58); every contact object here has a cosine. */
    static float cosTerm(const PRay& r) {
        float c = 1;
        if ((r.camera || r.contact != ORC_NONE) && length(r.d) > 0.5) c = std::fabs(dot(r.n, r.d));
        return c > EPS_D ? c : (float)EPS_D;
    }

    /* vbase: the vertex counter of the scatter at depth d is vbase + d + 1. */
    void samplePath(std::vector<PRay>& path, const PRay& prev, unsigned depth, unsigned p, unsigned pass,
                    uint32_t vbase = 0) const {
        path.push_back(prev);
        PRay term;                                  /* terminateRay, mcRenderer.cpp:20-27 */
        term.o = prev.o;
        term.rad = V(0, 0, 0);
        term.d = V(0, 0, 0);
        term.delta = true;
        unsigned self;
        if (prev.inside != ORC_NONE) self = prev.inside;
        else if (prev.hit.valid) self = prev.hit.obj;
        else { path.push_back(term); return; }
        if (!prev.hit.valid) { path.push_back(term); return; }  /* inside but no hit: as in sample() */
        float rn[4];
        rng4(seed, p, pass, vbase + depth + 1, rn);
        Ray in;
        in.o = prev.o;
        in.d = prev.d;
        const Out out = scatter(self, in, prev.hit, prev.inside, rn);
        PRay next;
        next.o = out.o; next.d = out.d; next.rad = out.rad; next.n = out.n;
        next.op = out.op; next.dp = out.dp;
        next.inside = out.inside; next.contact = out.contact; next.ctri = out.ctri;
        next.delta = out.delta;
        if (length(next.d) < 0.5) { path.push_back(next); return; }
        if (depth + 1 > maxDepth) { path.push_back(term); return; }
        Ray nr;
        nr.o = next.o;
        nr.d = next.d;
        next.hit = trace(nr, true, next.inside, next.contact);
        if (!next.hit.valid) {
            path.push_back(next);
            path.push_back(term);
            return;
        }
        samplePath(path, next, depth + 1, p, pass, vbase);
    }

    /* Scene::genLightSample -> SurfaceSampler::genSample (surfaceSampler.cpp:17-23)
     * -> AbstractObject::genSurfaceSample (abstractObj.cpp:18-52) with
     * Mesh::genRandTrianglePosition (mesh.h:89-101).  The four uniforms are
     * the object pick, the triangle pick and (u1, u2).  The hemisphere
     * direction, findInsideObject and the NoSelf intersect of the light ray
     * are not formed: the path tracer overwrites the direction and reads
     * neither of the others.  valid == false: the scene has no emissive
     * object (the reference indexes an empty vector there) or the picked one
     * has no triangles (contactObj stays NULL). */
    struct LightSample {
        bool valid = false;
        V o, n, rad;
        float op = 1.f;
        unsigned obj = ORC_NONE, tri = 0;
    };
    LightSample genLightSample(const float* rn) const {
        LightSample L;
        if (s->lightObj.empty()) return L;
        const float chosen = rn[0] * s->lightTotal;
        size_t li = std::lower_bound(s->lightCum.begin(), s->lightCum.end(), chosen) - s->lightCum.begin();
        if (li == s->lightCum.size()) li--;
        const std::vector<float>& cum = s->triCum[li];
        if (cum.empty()) return L;
        const unsigned oi = s->lightObj[li];
        const float randArea = rn[1] * s->lightArea[li];
        size_t ti = std::lower_bound(cum.begin(), cum.end(), randArea) - cum.begin();
        if (ti == cum.size()) ti--;
        const Object& o = s->objs[oi];
        V vps[3];
        for (int i = 0; i < 3; i++) vps[i] = o.verts[o.idx[3 * ti + i]];
        V b1 = sub(vps[1], vps[0]);
        V b2 = sub(vps[2], vps[0]);
        float u1 = rn[2], u2 = rn[3];
        if (u1 + u2 > 1) {
            u1 = 1 - u1;
            u2 = 1 - u2;
        }
        L.o = add(add(vps[0], mul(b1, u1)), mul(b2, u2));
        L.n = s->worldNormal(oi, (unsigned)ti, L.o);
        L.rad = o.color;                                  /* EmissiveObject::evaluateBSDF */
        L.op = 1 / s->lightArea[li] * s->pickProb[li];    /* abstractObj.cpp:41 */
        L.obj = oi;
        L.tri = (unsigned)ti;
        L.valid = true;
        return L;
    }

    /* MCRenderer::visibilityTest (mcRenderer.cpp:59-65) */
    bool visible(const PRay& in, const V& lightOrigin) const {
        Ray r;
        r.o = in.o;
        r.d = in.d;
        unsigned g = 0;
        const float dist = s->intersect(r, g, true, in.inside, in.contact);
        return dist < 0 || (dist - length(sub(in.o, lightOrigin)) >= -EPS_D);
    }

    /* Light sub-path of BDPT / VCM / PPM (bidirPathTracer.cpp:31-33,
     * vcmTracer.cpp:37-41, photonMap.cpp:61-65): Ray lightRay = genLightSample();
     * samplePath(lightPath, lightRay).  The light ray is the full
     * genSurfaceSample (abstractObj.cpp:18-52): the light point as in
     * genLightSample, a uniform-sphere direction (uniformSphereSampler.h:6-20)
     * flipped into the normal's hemisphere, directionProb = 2 * 1/(4 pi),
     * findInsideObject, and a NoSelf intersect whose hit counts only at
     * dist > 0.  Uniforms: vertex kLightVertex (object pick, triangle pick,
     * u1, u2), kLightVertex + 1 (the sphere), kLightVertex + 2 + d (the scatter
     * at depth d).  With nothing to sample the reference's empty-mesh ray
     * (direction and radiance 0, no contact) starts the path. */
    std::vector<PRay> lightPath(unsigned path, unsigned pass) const {
        float rn[4];
        rng4(seed, path, pass, kLightVertex, rn);
        const LightSample L = genLightSample(rn);
        PRay r;
        if (!L.valid) {
            r.d = V(0, 0, 0);
            r.rad = V(0, 0, 0);
        } else {
            r.o = L.o;
            r.n = L.n;
            r.contact = L.obj;
            r.ctri = L.tri;
            rng4(seed, path, pass, kLightVertex + 1, rn);
            const float term1 = (float)(2.0 * PI_D * (double)rn[0]);
            const float term2 = 2.f * std::sqrt(rn[1] - rn[1] * rn[1]);
            V dir(std::cos(term1) * term2, 1.f - 2.f * rn[1], std::sin(term1) * term2);
            r.dp = (float)(INV_PI_D * 0.25f);
            if (dot(dir, L.n) < 0) dir = neg(dir);
            r.d = dir;
            r.rad = L.rad;                                /* ray.evalBSDF(ray): the emission */
            Ray q;
            q.o = r.o;
            q.d = r.d;
            r.inside = s->findInside(q, r.contact);
            r.dp *= 2;
            r.op = L.op;
            const Hit h = trace(q, true, r.inside, r.contact);
            if (h.valid && h.dist > 0) r.hit = h;
        }
        std::vector<PRay> out;
        samplePath(out, r, 0, path, pass, kLightVertex + 1);
        return out;
    }

    /* ---- BidirectionalPathTracer (bidirPathTracer.cpp) ------------------------ */
    /* The camera ray's contactObj as a NoSelf contact: not an object, not NULL. */
    static constexpr unsigned kCameraObj = 0xFFFFFFFEu;

    PRay cameraRay(unsigned p, unsigned pass) const {
        float rn[4];
        rng4(seed, p, pass, 0, rn);
        PRay r0;
        r0.o = cam.pos;
        r0.d = cam.dir(p, rn[0], rn[1]);
        r0.camera = true;
        r0.pixel = p;
        r0.n = normalize(sub(cam.focus, r0.o));
        Ray cr;
        cr.o = r0.o;
        cr.d = r0.d;
        r0.hit = trace(cr, false, 0, 0);
        r0.inside = s->findInside(cr, ORC_NONE);
        return r0;
    }

    /* in.evalBSDF(out) (ray.cpp:15-22): out's contact (else inside) object's evaluateBSDF */
    V evalBSDF(const PRay& in, const PRay& out) const {
        if (out.camera) return V(1, 1, 1);  /* AbstractObject::evaluateBSDF */
        const unsigned obj = out.contact != ORC_NONE ? out.contact : out.inside;
        if (obj == ORC_NONE) return V(0, 0, 0);
        const Object& ob = s->objs[obj];
        if (ob.mat == ORC_REFRACTIVE) return V(0, 0, 0);            /* refractiveObj.h:13-15 */
        if (ob.mat == ORC_EMISSIVE) return ob.color;                /* emissiveObj.h:13 */
        if (out.contact == ORC_NONE) return V(0, 0, 0);             /* diffuseObj.h:18, glossyObj.h:21 */
        const V n = out.n;
        if (!(dot(in.d, n) <= 0 && dot(out.d, n) >= 0)) return V(0, 0, 0);
        if (ob.mat == ORC_DIFFUSE) return mul(ob.color, (float)INV_PI_D);  /* diffuseBsdf.h:29-34 */
        V refl = normalize(add(mul(mul(n, -dot(n, in.d)), 2.f), in.d));     /* isotropicPhong.h:41-48 */
        V rho = mul(mul(mul(ob.color, (float)(ob.power + 2.0)), 0.5f), (float)INV_PI_D);
        const float dRW = dot(refl, out.d);
        return mul(rho, std::pow(dRW > 0 ? dRW : 0.f, ob.power));
    }

    /* in.evalDirectionProbability(out) (ray.cpp:31-38); same: `in` and `out` are one
     * object (the "emitted" branch of AbstractObject::evalDirectionProbability) */
    float evalDirProb(const PRay& in, const PRay& out, bool same) const {
        if (out.camera) return (float)(1.0 / cam.pixelSolidAngle(out.pixel));  /* camera.cpp:144-146 */
        const unsigned obj = out.contact != ORC_NONE ? out.contact : out.inside;
        if (obj == ORC_NONE) return 0.f;
        const Object& ob = s->objs[obj];
        if (ob.mat == ORC_REFRACTIVE || ob.mat == ORC_EMISSIVE)  /* abstractObj.cpp:61-68 */
            return same ? (float)(INV_PI_D * 0.25f) * 2 : 0.f;
        if (out.contact == ORC_NONE) return 0.f;
        const float p = maxc(ob.color);
        const float c = dot(out.n, out.d);
        if (ob.mat == ORC_DIFFUSE) return p * (c > 0 ? (float)((double)c * INV_PI_D) : 0.f);  /* diffuseBsdf.h:12-14 */
        const float cosTheta = c > EPS_D ? c : (float)EPS_D;                                   /* isotropicPhong.h:13-18 */
        return p * (float)((ob.power + 1.0) * (double)std::pow(cosTheta, ob.power) * (INV_PI_D * 0.5));
    }

    static float cosTermB(const PRay& r) { return cosTerm(r); }

    /* BidirectionalPathTracer::linkVertex (bidirPathTracer.cpp:228-243) */
    static PRay linkVertex(const std::vector<PRay>& path, int cl, int startID, bool reverse) {
        const int endID = reverse ? startID - 1 : startID + 1;
        const int last = (int)path.size() - 1;
        const PRay& startRay = path[startID < 0 ? 0 : startID > last ? last : startID];
        const PRay& endRay = path[endID < 0 ? 0 : endID > last ? last : endID];
        PRay linkRay = startRay;
        if ((reverse && startID > cl) || (!reverse && startID <= cl)) return linkRay;
        linkRay.d = normalize(sub(endRay.o, startRay.o));
        linkRay.inside = endRay.inside;
        return linkRay;
    }

    static bool pathCanNotConnect(const std::vector<PRay>& path) { /* bidirPathTracer.cpp:251-257 */
        for (size_t i = 0; i + 1 < path.size(); i++)
            if (!path[i].delta && !path[i + 1].delta) return false;
        return true;
    }

    /* connectingColorProb (bidirPathTracer.cpp:266-291): color and probability */
    static void colorProb(const std::vector<PRay>& path, int cl, V& color, float& prob) {
        color = V(1, 1, 1);
        prob = 1;
        const int n = (int)path.size();
        for (int i = 0; i < n; i++) {
            color = mulv(color, path[i].rad);
            float dist = 0.f;
            if (i <= cl) {
                const float l = length(sub(path[i].o, path[i + 1].o));
                dist = l > EPS_D ? l : (float)EPS_D;
                color = mulv(color, V(1, 1, 1));  /* radianceDecay: no volumes */
            } else if (i > cl + 1) {
                const float l = length(sub(path[i - 1].o, path[i].o));
                dist = l > EPS_D ? l : (float)EPS_D;
                color = mulv(color, V(1, 1, 1));
            }
            if (i == cl && i < n - 1) color = mul(color, cosTerm(path[i]) * cosTerm(path[i + 1]) / (dist * dist));
            if (i != cl && i != cl + 1) color = mul(color, cosTerm(path[i]));
            prob *= path[i].dp * path[i].op;
        }
    }

    /* connectingMisWeight (bidirPathTracer.cpp:142-221), expTerm 1; every
     * evalOriginProbability here compares two temporaries, so it is 1 */
    float misWeight(const std::vector<PRay>& path, int cl) const {
        const int n = (int)path.size();
        std::vector<double> direct(n), reverse(n);
        std::vector<float> distRecord;
        direct[0] = path[0].op;
        for (int i = 1; i < n; i++) {
            const float l = length(sub(path[i].o, path[i - 1].o));
            const float dist = l > EPS_D ? l : (float)EPS_D;
            distRecord.push_back(dist);
            const float cosThere = cosTerm(linkVertex(path, cl, i, true));
            direct[i] = direct[i - 1] * (double)(1.0f * std::fabs(cosThere) / (dist * dist));  /* PdfWtoA */
            if (path[i - 1].delta) continue;
            float linkDirProb;
            if (i > 1)
                linkDirProb = evalDirProb(linkVertex(path, cl, i - 2, false), linkVertex(path, cl, i - 1, false), false);
            else {
                const PRay r = linkVertex(path, cl, i - 1, false);
                linkDirProb = evalDirProb(r, r, true);
            }
            direct[i] *= linkDirProb * 1.f;
        }
        reverse[n - 1] = path[n - 1].op;
        for (int i = n - 2; i >= 0; i--) {
            const float dist = distRecord[i];
            const float cosThere = cosTerm(linkVertex(path, cl, i, false));
            reverse[i] = reverse[i + 1] * (double)(1.0f * std::fabs(cosThere) / (dist * dist));
            if (path[i + 1].delta) continue;
            float linkDirProb;
            if (i < n - 2)
                linkDirProb = evalDirProb(linkVertex(path, cl, i + 2, true), linkVertex(path, cl, i + 1, true), false);
            else {
                const PRay r = linkVertex(path, cl, i + 1, true);
                linkDirProb = evalDirProb(r, r, true);
            }
            reverse[i] *= linkDirProb * 1.f;
        }
        double all = 0;
        for (int i = 0; i < n - 1; i++)
            if (!path[i].delta && !path[i + 1].delta) all += std::pow(direct[i] * reverse[i + 1], 1.0);
        if (pathCanNotConnect(path)) all += std::pow(reverse[0], 1.0);
        const double self = cl == -1 ? reverse[0] : direct[cl] * reverse[cl + 1];
        const double w = std::pow(self, 1.0) / all;
        return (float)(w > 0 ? w : 0);
    }

    /* connectEdge (bidirPathTracer.cpp:294-324) */
    bool connectEdge(std::vector<PRay>& path, int cl) const {
        PRay& lightRay = path[cl];
        PRay& eyeRay = path[cl + 1];
        if (lightRay.delta || eyeRay.delta) return false;
        const int last = (int)path.size() - 1;
        /* references, read after the edits: for cl = 0 (and for a camera-only eye
         * side) the previous ray is the connection ray itself */
        const PRay& preLight = path[cl - 1 > 0 ? cl - 1 : 0];
        const PRay& preEye = path[cl + 2 < last ? cl + 2 : last];
        lightRay.d = normalize(sub(eyeRay.o, lightRay.o));
        eyeRay.d = neg(lightRay.d);
        lightRay.dp = 1;
        lightRay.delta = false;
        const V lr = evalBSDF(preLight, lightRay);
        lightRay.rad = lr;
        if (lightRay.contact != ORC_NONE && cl > 0)
            lightRay.inside = dot(lightRay.n, lightRay.d) > 0 ? preLight.inside : lightRay.contact;
        eyeRay.dp = 1;
        eyeRay.delta = false;
        const V er = evalBSDF(preEye, eyeRay);
        eyeRay.rad = er;
        if (eyeRay.contact != ORC_NONE && cl < last - 1)
            eyeRay.inside = dot(eyeRay.n, eyeRay.d) > 0 ? preEye.inside : eyeRay.contact;
        return true;
    }

    /* One pixel's share of a BidirectionalPathTracer pass (bidirPathTracer.cpp:28-36,
     * 53-140): its eye path and its light path, every connection; contributions to
     * its own pixel in `own`, light-tracing splats to other pixels in `splats`. */
    void bdptPixel(unsigned p, unsigned pass, V& own, std::vector<std::pair<unsigned, V>>& splats) const {
        std::vector<PRay> eyePath;
        samplePath(eyePath, cameraRay(p, pass), 0, p, pass);
        const std::vector<PRay> lightP = lightPath(p, pass);
        own = V(0, 0, 0);
        const int maxEye = (int)eyePath.size(), maxLight = (int)lightP.size();
        const unsigned npix = cam.W * cam.H;
        for (int whole = 2; whole <= maxEye + maxLight; whole++) {
            const int minL = whole - maxEye > 0 ? whole - maxEye : 0;
            const int maxL = whole - 1 >= maxLight ? maxLight : whole - 1;
            for (int lightLen = minL; lightLen <= maxL; lightLen++) {
                const int eyeLen = whole - lightLen;
                std::vector<PRay> path(whole);
                for (int i = 0; i < lightLen; i++) path[i] = lightP[i];
                for (int i = 0; i < eyeLen; i++) path[lightLen + i] = eyePath[eyeLen - i - 1];
                const int cl = lightLen - 1;
                const PRay& front = path[0];
                if (!(front.contact != ORC_NONE && s->objs[front.contact].mat == ORC_EMISSIVE)) continue;
                if (lightLen == 0 && !pathCanNotConnect(path)) continue;
                if (eyeLen > 0 && lightLen > 0) {
                    const PRay& lc = path[cl];
                    const PRay& ec = path[cl + 1];
                    if (lightLen > 1 && lc.contact != ORC_NONE && s->objs[lc.contact].mat == ORC_EMISSIVE) continue;
                    if (ec.contact != ORC_NONE && s->objs[ec.contact].mat == ORC_EMISSIVE) continue;
                    if (!connectEdge(path, cl)) continue;
                    PRay vis = path[cl + 1];
                    if (vis.camera) vis.contact = kCameraObj;
                    if (!visible(vis, path[cl].o)) continue;
                }
                V color;
                float prob;
                colorProb(path, cl, color, prob);
                if (prob == 0) continue;
                if (length(color) == 0) continue;
                const float w = misWeight(path, cl);
                const V contrib = mul(divs(color, prob), w);
                if (eyeLen > 1) {
                    own = add(own, contrib);
                } else {
                    float rx, ry;
                    cam.toRaster(path[whole - 2].o, rx, ry);
                    const int x = (int)rx, y = (int)ry;  /* checkPixelBound truncates (camera.cpp:164-170) */
                    if (x >= 0 && y >= 0 && x < (int)cam.W && y < (int)cam.H) {
                        const float sa = cam.pixelSolidAngle((unsigned)(y * (int)cam.W + x));
                        const float factor = (float)(1.0 / sa);
                        splats.push_back({(unsigned)(y * (int)cam.W + x), mul(contrib, factor / (float)npix)});
                    }
                }
            }
        }
    }

    /* One pixel sample of PathTracer::renderPixels (pathTracer.cpp:22-69).
     * The branch's `continue` leaves the pixel's running mean scaled by
     * s/(s+1) with nothing added, i.e. a sample of 0.  The light sample's
     * uniforms come from the vertex counter kNeeVertex. */
    V samplePathVector(unsigned p, unsigned pass, bool nee, unsigned* nv) const {
        float rn[4];
        rng4(seed, p, pass, 0, rn);
        PRay r0;                                          /* Camera::generateRay (camera.cpp:6-60) */
        r0.o = cam.pos;
        r0.d = cam.dir(p, rn[0], rn[1]);
        r0.camera = true;
        r0.pixel = p;
        r0.n = normalize(sub(cam.focus, r0.o));            /* Camera::getWorldNormal (camera.cpp:182-186) */
        Ray cr;
        cr.o = r0.o;
        cr.d = r0.d;
        r0.hit = trace(cr, false, 0, 0);
        r0.inside = s->findInside(cr, ORC_NONE);
        std::vector<PRay> eyePath;
        samplePath(eyePath, r0, 0, p, pass);
        if (nv) *nv = (unsigned)eyePath.size();
        const size_t n = eyePath.size();
        V color(1, 1, 1);
        const bool hasToConnect = length(eyePath[n - 1].rad) <= 0;
        if (!nee || !hasToConnect) {
            for (size_t i = 0; i < n; i++) {
                if (i != n - 1) {
                    color = mul(color, cosTerm(eyePath[i]));
                    color = mulv(color, V(1, 1, 1));      /* radianceDecay: no volumes */
                }
                color = mulv(color, divs(divs(eyePath[i].rad, eyePath[i].op), eyePath[i].dp));
            }
        } else {
            const PRay& back = eyePath[n - 1];
            const int endIndex = (back.contact != ORC_NONE || back.camera || back.inside != ORC_NONE) ? -1 : (int)n - 2;
            if (endIndex <= 0) return V(0, 0, 0);
            PRay endRay = eyePath[endIndex];
            rng4(seed, p, pass, kNeeVertex, rn);
            const LightSample L = genLightSample(rn);
            if (endRay.contact != ORC_NONE && s->objs[endRay.contact].mat == ORC_EMISSIVE) return V(0, 0, 0);
            /* nonSpecular() is false for the glossy (MERGE_GLOSSY unset) and refractive classes */
            if (endRay.contact != ORC_NONE && s->objs[endRay.contact].mat != ORC_DIFFUSE) return V(0, 0, 0);
            if (!L.valid) return V(0, 0, 0);              /* contactNormal() of a NULL contact is 0: dot >= 0 */
            endRay.d = normalize(sub(L.o, endRay.o));
            PRay lightRay;
            lightRay.o = L.o;
            lightRay.d = neg(endRay.d);
            lightRay.n = L.n;
            lightRay.contact = L.obj;
            lightRay.ctri = L.tri;
            lightRay.rad = L.rad;
            lightRay.op = L.op;
            const float len = length(sub(L.o, endRay.o));
            const float connectDist = len > EPS_D ? len : (float)EPS_D;
            if (dot(endRay.d, L.n) >= 0) return V(0, 0, 0);
            if (!visible(endRay, L.o)) return V(0, 0, 0);
            for (int i = 0; i < endIndex; i++) {
                color = mulv(color, divs(divs(eyePath[i].rad, eyePath[i].op), eyePath[i].dp));
                color = mul(color, cosTerm(eyePath[i]));
                color = mulv(color, V(1, 1, 1));
            }
            /* eyePath[endIndex-1].evalBSDF(endRay) -> DiffuseObject::evaluateBSDF
             * (diffuseObj.h:17-21) -> DiffuseBSDF::evalBSDF (diffuseBsdf.h:29-34) */
            const V nE = endRay.n;
            const V inD = eyePath[endIndex - 1].d;
            V f = (dot(inD, nE) <= 0 && dot(endRay.d, nE) >= 0) ? mul(s->objs[endRay.contact].color, (float)INV_PI_D)
                                                                   : V(0, 0, 0);
            f = mulv(f, lightRay.rad);
            f = mulv(f, V(1, 1, 1));
            f = mul(f, cosTerm(lightRay));
            f = mul(f, cosTerm(endRay));
            f = divs(f, connectDist * connectDist);
            color = mulv(color, f);
            color = divs(color, endRay.op * lightRay.op);
        }
        if (!legal(color)) color = V(0, 0, 0);
        return cam.fixVignetting(color, p);
    }
};

/* ---- HashGrid (hashgrid.hxx:15-203) ---------------------------------------- */
struct HashGrid {
    V bmin, bmax;
    std::vector<int> cellEnds, indices;
    float radius = 0, cellSize = 0, invCellSize = 0;

    int cellIndex(int cx, int cy, int cz) const { /* hashgrid.hxx:170-178 */
        const uint32_t x = (uint32_t)cx, y = (uint32_t)cy, z = (uint32_t)cz;
        return (int)(((x * 73856093u) ^ (y * 19349663u) ^ (z * 83492791u)) % (uint32_t)cellEnds.size());
    }
    int cellIndex(V p) const { /* hashgrid.hxx:180-190 */
        const V dm = sub(p, bmin);
        return cellIndex((int)std::floor(invCellSize * dm.x), (int)std::floor(invCellSize * dm.y),
                         (int)std::floor(invCellSize * dm.z));
    }
    /* Reserve(nCells) + Build(particles, radius) with aMaxDist = 0 (hashgrid.hxx:18-75) */
    void build(const std::vector<V>& pts, float r, size_t nCells) {
        cellEnds.assign(nCells, 0);
        float maxDist = 0.f;
        if (std::fabs(maxDist) < EPS_D) maxDist = r;
        radius = r;
        cellSize = maxDist * 2.f;
        invCellSize = 1.f / cellSize;
        bmin = V(1e36f, 1e36f, 1e36f);
        bmax = V(-1e36f, -1e36f, -1e36f);
        for (const V& p : pts)
            for (int j = 0; j < 3; j++) {
                bmax[j] = std::max(bmax[j], p[j]);
                bmin[j] = std::min(bmin[j], p[j]);
            }
        indices.assign(pts.size(), 0);
        for (const V& p : pts) cellEnds[cellIndex(p)]++;
        int sum = 0;
        for (size_t i = 0; i < cellEnds.size(); i++) {
            const int t = cellEnds[i];
            cellEnds[i] = sum;
            sum += t;
        }
        for (size_t i = 0; i < pts.size(); i++) indices[cellEnds[cellIndex(pts[i])]++] = (int)i;
    }
    /* Process(particles, query) (hashgrid.hxx:77-158): the particles it hands to
     * the query, in order (a cell reached twice through a hash collision is
     * visited twice, as there). */
    void query(const std::vector<V>& pts, V q, std::vector<int>& out) const {
        out.clear();
        const V dmin = sub(q, bmin), dmax = sub(bmax, q);
        for (int i = 0; i < 3; i++)
            if (dmin[i] < 0.f || dmax[i] < 0.f) return;
        const V cp = V(invCellSize * dmin.x, invCellSize * dmin.y, invCellSize * dmin.z);
        const V cf = V(std::floor(cp.x), std::floor(cp.y), std::floor(cp.z));
        const int px = (int)cf.x, py = (int)cf.y, pz = (int)cf.z;
        const V fr = sub(cp, cf);
        const int pxo = px + (fr.x < 0.5f ? -1 : +1);
        const int pyo = py + (fr.y < 0.5f ? -1 : +1);
        const int pzo = pz + (fr.z < 0.5f ? -1 : +1);
        const int cells[8][3] = {{px, py, pz},   {px, py, pzo},   {px, pyo, pz},   {px, pyo, pzo},
                                 {pxo, py, pz},  {pxo, py, pzo},  {pxo, pyo, pz},  {pxo, pyo, pzo}};
        for (int j = 0; j < 8; j++) {
            const int c = cellIndex(cells[j][0], cells[j][1], cells[j][2]);
            int a = c == 0 ? 0 : cellEnds[c - 1], b = cellEnds[c];
            for (; a < b; a++) {
                const int pi = indices[a];
                const float dist = length(sub(q, pts[pi]));
                const float d2 = dist * dist;
                if (d2 <= radius * radius) out.push_back(pi);
            }
        }
    }
};

/* ---- VCMTracer (vcmTracer.cpp:6-446) --------------------------------------- */
/* One pass of VCMTracer::renderPixels with the pass's radius: the sub-paths and
 * the vertex-connection half are BidirectionalPathTracer's (the same light and
 * eye sub-paths, the same connectEdge / colour / visibility rules), with
 * vcmMisWeight in place of connectingMisWeight; the vertex-merging half
 * queries the light points around every non-delta eye vertex. */
struct Vcm {
    using PRay = Tracer::PRay;
    const Tracer& tr;
    float radius;
    unsigned npix;

    bool emissiveObj(unsigned o) const { return o != ORC_NONE && tr.s->objs[o].mat == ORC_EMISSIVE; }
    /* insideObj && !contactObj (a point inside a medium; the camera is a contact) */
    static bool inVolume(const PRay& r) { return r.inside != ORC_NONE && r.contact == ORC_NONE && !r.camera; }

    /* connectEdge (vcmTracer.cpp:366-398): both directions re-aimed; for a
     * connection (vm false) the light-side ray is edited as in BDPT, for a
     * merge only the eye-side ray is.  The previous rays are references, read
     * after the edits. */
    bool connectEdge(std::vector<PRay>& path, int cl, bool vm) const {
        PRay& lightRay = path[cl];
        PRay& eyeRay = path[cl + 1];
        if (!vm && (lightRay.delta || eyeRay.delta)) return false;
        if (vm && eyeRay.delta) return false;
        const int last = (int)path.size() - 1;
        const PRay& preLight = path[cl - 1 > 0 ? cl - 1 : 0];
        const PRay& preEye = path[cl + 2 < last ? cl + 2 : last];
        lightRay.d = normalize(sub(eyeRay.o, lightRay.o));
        eyeRay.d = neg(lightRay.d);
        if (!vm) {
            lightRay.dp = 1;
            lightRay.delta = false;
            const V lr = tr.evalBSDF(preLight, lightRay);
            lightRay.rad = lr;
            if (lightRay.contact != ORC_NONE && cl > 0)
                lightRay.inside = dot(lightRay.n, lightRay.d) > 0 ? preLight.inside : lightRay.contact;
        }
        eyeRay.dp = 1;
        eyeRay.delta = false;
        const V er = tr.evalBSDF(preEye, eyeRay);
        eyeRay.rad = er;
        if (eyeRay.contact != ORC_NONE && cl < last - 1)
            eyeRay.inside = dot(eyeRay.n, eyeRay.d) > 0 ? preEye.inside : eyeRay.contact;
        return true;
    }

    /* connectingColorProb (vcmTracer.cpp:329-362); a merge (vm) also takes the
     * merging disc, the area density of the merged vertex and, inside a medium,
     * the sphere's extra length into the probability */
    void colorProb(const std::vector<PRay>& path, int cl, bool vm, V& color, float& prob) const {
        color = V(1, 1, 1);
        prob = 1;
        const int n = (int)path.size();
        for (int i = 0; i < n; i++) {
            color = mulv(color, path[i].rad);
            float dist = 0.f;
            if (i <= cl) {
                const float l = length(sub(path[i].o, path[i + 1].o));
                dist = l > EPS_D ? l : (float)EPS_D;
                color = mulv(color, V(1, 1, 1));  /* radianceDecay: no volumes */
            } else if (i > cl + 1) {
                const float l = length(sub(path[i - 1].o, path[i].o));
                dist = l > EPS_D ? l : (float)EPS_D;
                color = mulv(color, V(1, 1, 1));
            }
            if (i == cl && i < n - 1) color = mul(color, Tracer::cosTerm(path[i]) * Tracer::cosTerm(path[i + 1]) / (dist * dist));
            if (i != cl && i != cl + 1) color = mul(color, Tracer::cosTerm(path[i]));
            prob *= path[i].dp * path[i].op;
            if (vm && i == cl) {
                prob = (float)((double)prob * (PI_D * (double)radius * (double)radius));
                const float cosThere = Tracer::cosTerm(Tracer::linkVertex(path, cl, i + 1, true));
                prob *= 1.0f * std::fabs(cosThere) / (dist * dist);  /* PdfWtoA */
                prob *= 1.f;                                         /* evalOriginProbability of two temporaries */
                if (inVolume(path[i + 1])) prob = (float)((double)prob * (4.0 / 3 * (double)radius));
            }
        }
    }

    /* vcmMisWeight (vcmTracer.cpp:221-320), expTerm 1, nVC = 1, nVM = pixels;
     * every evalOriginProbability there compares two temporaries, so it is 1 */
    float misWeight(const std::vector<PRay>& path, int cl, bool vm) const {
        const int n = (int)path.size();
        std::vector<double> direct(n), reverse(n);
        std::vector<float> distRecord;
        direct[0] = path[0].op;
        for (int i = 1; i < n; i++) {
            const float l = length(sub(path[i].o, path[i - 1].o));
            const float dist = l > EPS_D ? l : (float)EPS_D;
            distRecord.push_back(dist);
            const float cosThere = Tracer::cosTerm(Tracer::linkVertex(path, cl, i, true));
            direct[i] = direct[i - 1] * (double)(1.0f * std::fabs(cosThere) / (dist * dist));
            if (path[i - 1].delta) continue;
            float linkDirProb;
            if (i > 1)
                linkDirProb = tr.evalDirProb(Tracer::linkVertex(path, cl, i - 2, false),
                                             Tracer::linkVertex(path, cl, i - 1, false), false);
            else {
                const PRay r = Tracer::linkVertex(path, cl, i - 1, false);
                linkDirProb = tr.evalDirProb(r, r, true);
            }
            direct[i] *= linkDirProb * 1.f;
        }
        reverse[n - 1] = path[n - 1].op;
        for (int i = n - 2; i >= 0; i--) {
            const float dist = distRecord[i];
            const float cosThere = Tracer::cosTerm(Tracer::linkVertex(path, cl, i, false));
            reverse[i] = reverse[i + 1] * (double)(1.0f * std::fabs(cosThere) / (dist * dist));
            if (path[i + 1].delta) continue;
            float linkDirProb;
            if (i < n - 2)
                linkDirProb = tr.evalDirProb(Tracer::linkVertex(path, cl, i + 2, true),
                                             Tracer::linkVertex(path, cl, i + 1, true), false);
            else {
                const PRay r = Tracer::linkVertex(path, cl, i + 1, true);
                linkDirProb = tr.evalDirProb(r, r, true);
            }
            reverse[i] *= linkDirProb * 1.f;
        }
        const double nVC = 1.0, nVM = (double)npix;
        double all = 0;
        double self = cl == -1 ? reverse[0] : direct[cl] * reverse[cl + 1];
        for (int i = 0; i < n - 1; i++) {
            if (!path[i].delta && !path[i + 1].delta) all += direct[i] * reverse[i + 1] * nVC;
            if (i < n - 2 && !path[i + 1].delta) {
                double p = direct[i] * reverse[i + 1];
                if (!path[i].delta) {
                    float dp;
                    if (i > 0)
                        dp = tr.evalDirProb(Tracer::linkVertex(path, cl, i - 1, false),
                                            Tracer::linkVertex(path, cl, i, false), false);
                    else {
                        const PRay r = Tracer::linkVertex(path, cl, i, false);
                        dp = tr.evalDirProb(r, r, true);
                    }
                    p *= dp;
                }
                p *= 1.f;  /* evalOriginProbability */
                const float cosThere = Tracer::cosTerm(Tracer::linkVertex(path, cl, i + 1, true));
                p *= (double)(1.0f * std::fabs(cosThere) / (distRecord[i] * distRecord[i]));
                p *= PI_D * (double)radius * (double)radius;
                if (inVolume(path[i + 1])) p *= 4.0 / 3 * (double)radius;
                if (vm && i == cl) self = p;
                all += p * nVM;
            }
        }
        if (Tracer::pathCanNotConnect(path)) all += reverse[0];
        self *= vm ? nVM : nVC;
        const double w = self / all;
        if (!(w >= 0)) return 0.f;
        return (float)w;
    }

    /* One pixel's share of a pass: throughputByConnecting (vcmTracer.cpp:88-173),
     * then throughputByMerging (175-217) over the pass's light points (pts, with
     * their source path and index; path -1 for the leading default points). */
    void pixel(unsigned p, unsigned pass, const std::vector<std::vector<PRay>>& lightPaths,
               const std::vector<V>& pts, const std::vector<std::pair<int, int>>& src, const HashGrid& grid,
               V& own, std::vector<std::pair<unsigned, V>>& splats) const {
        std::vector<PRay> eyePath;
        tr.samplePath(eyePath, tr.cameraRay(p, pass), 0, p, pass);
        const std::vector<PRay>& lightP = lightPaths[p];
        own = V(0, 0, 0);
        const int maxEye = (int)eyePath.size(), maxLight = (int)lightP.size();
        for (int whole = 2; whole <= maxEye + maxLight; whole++) {
            const int minL = whole - maxEye > 0 ? whole - maxEye : 0;
            const int maxL = whole - 1 >= maxLight ? maxLight : whole - 1;
            for (int lightLen = minL; lightLen <= maxL; lightLen++) {
                const int eyeLen = whole - lightLen;
                std::vector<PRay> path(whole);
                for (int i = 0; i < lightLen; i++) path[i] = lightP[i];
                for (int i = 0; i < eyeLen; i++) path[lightLen + i] = eyePath[eyeLen - i - 1];
                const int cl = lightLen - 1;
                if (!emissiveObj(path[0].contact)) continue;
                if (lightLen == 0 && !Tracer::pathCanNotConnect(path)) continue;
                if (eyeLen > 0 && lightLen > 0) {
                    if (lightLen > 1 && emissiveObj(path[cl].contact)) continue;
                    if (emissiveObj(path[cl + 1].contact)) continue;
                    if (!connectEdge(path, cl, false)) continue;
                    PRay vis = path[cl + 1];
                    if (vis.camera) vis.contact = Tracer::kCameraObj;
                    if (!tr.visible(vis, path[cl].o)) continue;
                }
                V color;
                float prob;
                colorProb(path, cl, false, color, prob);
                if (prob == 0) continue;
                if (length(color) == 0) continue;
                const float w = misWeight(path, cl, false);
                const V contrib = mul(divs(color, prob), w);
                if (eyeLen > 1) {
                    own = add(own, contrib);
                } else {
                    float rx, ry;
                    tr.cam.toRaster(path[whole - 2].o, rx, ry);
                    const int x = (int)rx, y = (int)ry;
                    if (x >= 0 && y >= 0 && x < (int)tr.cam.W && y < (int)tr.cam.H) {
                        const float sa = tr.cam.pixelSolidAngle((unsigned)(y * (int)tr.cam.W + x));
                        const float factor = (float)(1.0 / sa);
                        splats.push_back({(unsigned)(y * (int)tr.cam.W + x), mul(contrib, factor / (float)npix)});
                    }
                }
            }
        }
        std::vector<int> near;
        for (int i = 1; i < maxEye; i++) {
            if (emissiveObj(eyePath[i].contact)) break;
            if (eyePath[i].delta) continue;
            V contrib(0, 0, 0);
            grid.query(pts, eyePath[i].o, near);
            for (int pi : near) {
                const int k = src[pi].first, j = src[pi].second;
                if (k < 0 || j < 0) continue;  /* a leading default point: no path */
                const std::vector<PRay>& L = lightPaths[k];
                std::vector<PRay> path(L.begin(), L.begin() + j);
                for (int e = 0; e <= i; e++) path.push_back(eyePath[i - e]);
                if (!connectEdge(path, j - 1, true)) continue;
                V color;
                float prob;
                colorProb(path, j - 1, true, color, prob);
                if (length(color) == 0 || prob == 0) continue;
                const float w = misWeight(path, j - 1, true);
                contrib = add(contrib, divs(mul(color, w), prob));
            }
            own = add(own, divs(contrib, (float)npix));
        }
    }
};

/* ---- PhotonMap (photonMap.cpp:21-680) ----------------------------------------
 * Progressive photon mapping on surfaces (photonMap.cpp with PPM defined and
 * NO_RAY_MARCHING, ANI_PM and RESEARCH undefined): no object here is a volume
 * or an anisotropic glossy object, so every light point is a surface photon
 * and every density estimate is the plain range query.  Per pass: the radius
 * of vcmRadius, the light points of VCMTracer (the same light paths), the
 * hash grid; per pixel the eye "merge path" (samplePath with maxPathLen
 * MERGE_LEN = 10, photonMap.cpp:10,150-190), and at its first emitter or
 * non-specular (diffuse; Phong is specular without MERGE_GLOSSY) vertex the
 * emitted radiance or the kernel density estimate of the photons around it
 * (photonMap.cpp:198-248,656-668), times the camera vertex's factor only
 * (photonMap.cpp:670-674). */
struct Ppm {
    using PRay = Tracer::PRay;
    const Tracer& tr;  /* light paths: maxPathLen of the renderer */
    const Tracer& te;  /* eye merge paths: maxPathLen MERGE_LEN */
    float radius;
    unsigned npix;

    /* Query::Process (photonMap.cpp:219-248) for one light point */
    V photon(const std::vector<PRay>& L, int index, const PRay& outRay) const {
        V thr(1, 1, 1);
        for (int i = 0; i < index; i++) {
            thr = mulv(thr, divs(divs(L[i].rad, L[i].dp), L[i].op));
            thr = mul(thr, Tracer::cosTerm(L[i]));
            thr = mulv(thr, V(1, 1, 1));  /* radianceDecay: no volumes */
        }
        thr = divs(thr, L[index].op);
        PRay photonRay = L[index];
        photonRay.d = L[index - 1].d;
        const V color = mulv(thr, tr.evalBSDF(photonRay, outRay));
        const float l = length(sub(outRay.o, L[index].o));
        const float distSqr = l * l;  /* powf(l, 2): the correctly rounded square */
        if (!legal(color)) return V(0, 0, 0);
        const float radiusSqr = radius * radius;
        const float q = 1 - distSqr / radiusSqr;
        const float sk = 0 > q ? 0.f : q;                     /* MAX(0, 1 - distSqr / radiusSqr) */
        const float kernel = (float)(3 * sk * sk / PI_D);     /* Kernel (photonMap.cpp:209-212) */
        const float normalization = kernel / ((float)npix * radius * radius);
        return mul(color, normalization);
    }

    V pixel(unsigned p, unsigned pass, const std::vector<std::vector<PRay>>& lightPaths, const std::vector<V>& pts,
            const std::vector<std::pair<int, int>>& src, const HashGrid& grid) const {
        std::vector<PRay> eye;
        te.samplePath(eye, te.cameraRay(p, pass), 0, p, pass);
        V surface(0, 0, 0);
        std::vector<int> near;
        for (size_t i = 1; i < eye.size(); i++) {
            const unsigned c = eye[i].contact;
            if (c != ORC_NONE && tr.s->objs[c].mat == ORC_EMISSIVE) {  /* the eye path hit a light */
                surface = eye[i].rad;
                break;
            }
            if (c != ORC_NONE && tr.s->objs[c].mat == ORC_DIFFUSE) {  /* nonSpecular() */
                PRay outRay = eye[i];
                outRay.d = neg(eye[i - 1].d);
                grid.query(pts, eye[i].o, near);
                V contrib(0, 0, 0);
                for (int pi : near) {
                    const int k = src[pi].first, j = src[pi].second;
                    if (k < 0 || j < 0) continue;  /* a leading default point: no path */
                    contrib = add(contrib, photon(lightPaths[k], j, outRay));
                }
                surface = contrib;
                break;
            }
        }
        V color = add(mulv(V(1, 1, 1), surface), V(0, 0, 0));  /* Tr * SurfaceColor + VolumeColor */
        const PRay& e0 = eye[0];
        color = mulv(color, divs(divs(mul(e0.rad, Tracer::cosTerm(e0)), e0.dp), e0.op));
        return color;
    }
};

/* VCMTracer's radius of pass s (vcmTracer.cpp:30): MAX(baseRadius *
 * sqrt(powf(s+1, alpha-1)), EPSILON) with alpha = 2/3 (vcmTracer.h:39). */
static float vcmRadius(float base, uint32_t s) {
    const float alpha = (float)(2.0 / 3);
    const float r = base * std::sqrt(std::pow((float)(s + 1), alpha - 1.f));
    return r > EPS_D ? r : (float)EPS_D;
}

}  // namespace orc

using namespace orc;

/* ---- C ABI -------------------------------------------------------------- */
extern "C" {

void* orc_tree_build(const float* verts, uint32_t nv, const uint32_t* tris, uint32_t nt) {
    Tree* t = new Tree;
    t->verts.resize(nv);
    for (uint32_t i = 0; i < nv; i++) t->verts[i] = V(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    t->idx.assign(tris, tris + 3 * (size_t)nt);
    t->build();
    return t;
}

void orc_tree_free(void* tree) { delete (Tree*)tree; }

static void count_nodes(const Node* n, uint32_t& nodes, uint32_t& ltris, uint32_t& leaves) {
    if (!n) return;
    nodes++;
    if (!n->left && !n->right) { ltris += (uint32_t)n->tris.size(); leaves++; return; }
    count_nodes(n->left, nodes, ltris, leaves);
    count_nodes(n->right, nodes, ltris, leaves);
}

void orc_tree_stats(void* tree, uint32_t* n_nodes, uint32_t* n_leaf_tris, uint32_t* depth, uint32_t* n_leaves) {
    Tree* t = (Tree*)tree;
    uint32_t a = 0, b = 0, c = 0;
    if (t->root && !t->root->tris.empty()) count_nodes(t->root, a, b, c);
    *n_nodes = a; *n_leaf_tris = b; *depth = t->depth; *n_leaves = c;
}

static void export_rec(const Node* n, uint32_t& ni, uint32_t& li, float* box, uint32_t* right,
                       uint32_t* lf, uint32_t* lc, uint32_t* tl) {
    uint32_t me = ni++;
    memcpy(box + 6 * (size_t)me, n->box, 6 * sizeof(float));
    if (!n->left && !n->right) {
        right[me] = ORC_NONE;
        lf[me] = li;
        lc[me] = (uint32_t)n->tris.size();
        for (unsigned t : n->tris) tl[li++] = t;
        return;
    }
    lf[me] = 0;
    lc[me] = 0;
    export_rec(n->left, ni, li, box, right, lf, lc, tl);
    right[me] = ni;
    export_rec(n->right, ni, li, box, right, lf, lc, tl);
}

void orc_tree_export(void* tree, float* box, uint32_t* right, uint32_t* leaf_first,
                     uint32_t* leaf_count, uint32_t* tri_list) {
    Tree* t = (Tree*)tree;
    if (!t->root || t->root->tris.empty()) return;
    uint32_t ni = 0, li = 0;
    export_rec(t->root, ni, li, box, right, leaf_first, leaf_count, tri_list);
}

void orc_tree_intersect(void* tree, const float* rays, uint64_t n, uint32_t* prim, float* t, int nthreads) {
    Tree* tr = (Tree*)tree;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Ray r;
        r.o = V(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = V(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        unsigned tid = ORC_NONE;
        float d = tr->intersect(r, tid, NoCond{});
        prim[i] = d >= 0 ? tid : ORC_NONE;
        t[i] = d >= 0 ? d : -1.f;
    }
}

float orc_box_intersect(const float* box6, const float* ray6) {
    Ray r;
    r.o = V(ray6[0], ray6[1], ray6[2]);
    r.d = V(ray6[3], ray6[4], ray6[5]);
    return Tree::boxIsect(box6, r);
}

float orc_tri_intersect(const float* v0, const float* v1, const float* v2, const float* ray6) {
    Tree t;
    t.verts = {V(v0[0], v0[1], v0[2]), V(v1[0], v1[1], v1[2]), V(v2[0], v2[1], v2[2])};
    t.idx = {0, 1, 2};
    Ray r;
    r.o = V(ray6[0], ray6[1], ray6[2]);
    r.d = V(ray6[3], ray6[4], ray6[5]);
    return t.triIsect(r, 0, NoCond{});
}

void* orc_scene_create(void) { return new Scene; }
void orc_scene_free(void* s) { delete (Scene*)s; }

uint32_t orc_scene_add_object(void* sp, const float* verts, uint32_t nv, const uint32_t* tris,
                              uint32_t nt, const float* normals, int mat, const float* params) {
    Scene* s = (Scene*)sp;
    Object o;
    o.verts.resize(nv);
    for (uint32_t i = 0; i < nv; i++) o.verts[i] = V(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    o.idx.assign(tris, tris + 3 * (size_t)nt);
    if (normals) {
        o.normals.resize(nv);
        for (uint32_t i = 0; i < nv; i++) o.normals[i] = V(normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
    }
    o.mat = mat;
    o.power = 0;
    o.ior = 1;
    if (mat == ORC_REFRACTIVE) o.ior = params[0];
    else o.color = V(params[0], params[1], params[2]);
    if (mat == ORC_GLOSSY) o.power = params[3];
    s->objs.push_back(o);
    return (uint32_t)(s->objs.size() - 1);
}

void orc_scene_build(void* s) { ((Scene*)s)->build(); }

void orc_scene_intersect(void* sp, const float* rays, const uint32_t* inside, const uint32_t* contact,
                         uint64_t n, int noself, uint32_t* prim, uint32_t* obj, uint32_t* tri,
                         float* t, float* b1, float* b2, int nthreads) {
    Scene* s = (Scene*)sp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Ray r;
        r.o = V(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = V(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        unsigned g = ORC_NONE;
        unsigned in = inside ? inside[i] : ORC_NONE, ct = contact ? contact[i] : ORC_NONE;
        float d = s->intersect(r, g, noself != 0, in, ct);
        if (d >= 0) {
            prim[i] = g;
            obj[i] = s->srcObj[g];
            tri[i] = s->srcTri[g];
            t[i] = d;
            float u1 = 0, u2 = 0;
            s->worldNormal(s->srcObj[g], s->srcTri[g], add(r.o, mul(r.d, d)), &u1, &u2);
            b1[i] = u1;
            b2[i] = u2;
        } else {
            prim[i] = ORC_NONE; obj[i] = ORC_NONE; tri[i] = ORC_NONE;
            t[i] = -1.f; b1[i] = 0.f; b2[i] = 0.f;
        }
    }
}

void orc_scene_near_ties(void* sp, const float* rays, uint64_t n, const uint32_t* prim, const float* t, float tol,
                         uint32_t* count, int nthreads) {
    Scene* s = (Scene*)sp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        count[i] = 0;
        if (prim[i] == ORC_NONE) continue;
        Ray r;
        r.o = V(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = V(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        TieCount tc{t[i], tol, prim[i]};
        unsigned tid = 0;
        (void)s->tree.intersect(r, tid, tc);
        count[i] = tc.n;
    }
}

void orc_scene_find_inside(void* sp, const float* rays, const uint32_t* cur, uint64_t n,
                           uint32_t* out, int nthreads) {
    Scene* s = (Scene*)sp;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        Ray r;
        r.o = V(rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]);
        r.d = V(rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]);
        out[i] = s->findInside(r, cur ? cur[i] : ORC_NONE);
    }
}

void orc_world_normal(void* sp, uint32_t obj, uint32_t tri, const float* p, float* n_out,
                      float* u1_out, float* u2_out) {
    Scene* s = (Scene*)sp;
    V n = s->worldNormal(obj, tri, V(p[0], p[1], p[2]), u1_out, u2_out);
    n_out[0] = n.x; n_out[1] = n.y; n_out[2] = n.z;
}

void orc_render(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                uint32_t npass, double* sum, double* sumsq, int nthreads) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    const int64_t np = (int64_t)c->width * c->height;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    for (uint32_t pass = pass0; pass < pass0 + npass; pass++) {
        s->updateForMotion(seed, pass);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) if(np > 64)
        for (int64_t p = 0; p < np; p++) {
            V v = tr.sample((unsigned)p, pass, nullptr);
            for (int k = 0; k < 3; k++) {
                sum[3 * p + k] += v[k];
                sumsq[3 * p + k] += (double)v[k] * v[k];
            }
        }
    }
    restoreMotion(s);
}

/* Camera::generateRay's direction (camera.cpp:13-23) for pixel p and jitter
 * (u, v), and the sample weight the camera gives it: the camera vertex's cosine
 * term |front . d| (ray.cpp:52-58 with Camera::getWorldNormal, camera.cpp:182-186)
 * divided by fixVignetting's |front . d_centre| (camera.cpp:107-122). */
void orc_camera_ray(const orc_camera* c, uint32_t p, float u, float v, float* dir, float* weight) {
    Cam cam(c);
    V d = cam.dir(p, u, v);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    float c0 = std::fabs(dot(normalize(sub(cam.focus, cam.pos)), d));
    c0 = c0 > EPS_D ? c0 : (float)EPS_D;
    *weight = cam.fixVignetting(V(c0, c0, c0), p).x;
}

/* AbstractObject::scatter of object `self` (dispatched as mcRenderer.cpp:30-31)
 * for the ray (o, d) whose closest hit is local triangle `tri` of object `obj`
 * at distance dist, with insideObj `inside`, from the uniforms rn[0..2].
 * out: origin (3), direction (3), radiance (3), originProb, directionProb,
 * contact normal (3); ids: insideObj, contactObj.  Returns isDeltaDirection. */
int orc_scatter(void* sp, uint32_t self, const float* ray6, uint32_t obj, uint32_t tri, float dist, uint32_t inside,
                const float* rn, float* out, uint32_t* ids) {
    Scene* s = (Scene*)sp;
    orc_camera c0{};
    c0.focus[2] = 1.f;
    c0.up[1] = 1.f;
    c0.width = c0.height = 1;
    c0.sight_dist = 1.f;
    Tracer tr{s, Cam(&c0), 0, 0};
    Ray r;
    r.o = V(ray6[0], ray6[1], ray6[2]);
    r.d = V(ray6[3], ray6[4], ray6[5]);
    Tracer::Hit h;
    h.valid = true;
    h.dist = dist;
    h.obj = obj;
    h.tri = tri;
    float u[4] = {rn[0], rn[1], rn[2], 0.f};
    Out o = tr.scatter(self, r, h, inside, u);
    const V* vs[3] = {&o.o, &o.d, &o.rad};
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) out[3 * k + a] = (*vs[k])[a];
    out[9] = o.op;
    out[10] = o.dp;
    out[11] = o.n.x; out[12] = o.n.y; out[13] = o.n.z;
    ids[0] = o.inside;
    ids[1] = o.contact;
    return o.delta ? 1 : 0;
}

void orc_render_sample(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed,
                       uint32_t pixel, uint32_t pass, float* rgb, uint32_t* n_vertices) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    V v = tr.sample(pixel, pass, n_vertices);
    rgb[0] = v.x; rgb[1] = v.y; rgb[2] = v.z;
}

void orc_render2(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t pass0,
                 uint32_t npass, uint32_t flags, double* sum, double* sumsq, int nthreads) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    const int64_t np = (int64_t)c->width * c->height;
    const bool nee = (flags & ORC_NEE) != 0, pv = (flags & (ORC_NEE | ORC_PATH_VECTOR)) != 0;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    for (uint32_t pass = pass0; pass < pass0 + npass; pass++) {
        s->updateForMotion(seed, pass);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) if(np > 64)
        for (int64_t p = 0; p < np; p++) {
            V v = pv ? tr.samplePathVector((unsigned)p, pass, nee, nullptr) : tr.sample((unsigned)p, pass, nullptr);
            for (int k = 0; k < 3; k++) {
                sum[3 * p + k] += v[k];
                sumsq[3 * p + k] += (double)v[k] * v[k];
            }
        }
    }
    restoreMotion(s);
}

void orc_render_sample2(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t flags,
                        uint32_t pixel, uint32_t pass, float* rgb, uint32_t* n_vertices) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    const bool nee = (flags & ORC_NEE) != 0, pv = (flags & (ORC_NEE | ORC_PATH_VECTOR)) != 0;
    V v = pv ? tr.samplePathVector(pixel, pass, nee, n_vertices) : tr.sample(pixel, pass, n_vertices);
    rgb[0] = v.x; rgb[1] = v.y; rgb[2] = v.z;
}

int orc_light_sample(void* sp, const float rn[4], float* origin, float* normal, float* radiance, float* origin_prob,
                     uint32_t* obj, uint32_t* tri) {
    Scene* s = (Scene*)sp;
    orc_camera c{};
    c.width = c.height = 1;
    Tracer tr{s, Cam(&c), 0, 0};
    const Tracer::LightSample L = tr.genLightSample(rn);
    if (!L.valid) return 0;
    for (int k = 0; k < 3; k++) {
        origin[k] = L.o[k];
        normal[k] = L.n[k];
        radiance[k] = L.rad[k];
    }
    *origin_prob = L.op;
    *obj = L.obj;
    *tri = L.tri;
    return 1;
}

uint32_t orc_light_paths(void* sp, uint32_t max_depth, uint64_t seed, uint32_t pass, uint32_t path0,
                         uint32_t n_paths, uint32_t max_vertices, float* rec, uint32_t* counts, int nthreads) {
    Scene* s = (Scene*)sp;
    orc_camera c{};
    c.width = c.height = 1;
    Tracer tr{s, Cam(&c), max_depth, seed};
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    uint32_t longest = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads) reduction(max : longest)
    for (int64_t k = 0; k < (int64_t)n_paths; k++) {
        const std::vector<Tracer::PRay> P = tr.lightPath(path0 + (uint32_t)k, pass);
        const uint32_t n = (uint32_t)P.size();
        longest = std::max(longest, n);
        counts[k] = n;
        for (uint32_t i = 0; i < std::min(n, max_vertices); i++) {
            float* o = rec + 16 * ((size_t)k * max_vertices + i);
            const Tracer::PRay& r = P[i];
            for (int a = 0; a < 3; a++) {
                o[a] = r.o[a];
                o[3 + a] = r.d[a];
                o[6 + a] = r.rad[a];
            }
            o[9] = r.op;
            o[10] = r.dp;
            uint32_t* u = (uint32_t*)o;
            u[11] = r.contact;
            u[12] = r.contact == ORC_NONE ? 0u : r.ctri;
            u[13] = r.inside;
            u[14] = r.delta ? 1u : 0u;
            u[15] = 0u;
        }
    }
    return longest;
}

void* orc_hashgrid_build(const float* pos, uint64_t n, float radius, uint32_t n_cells) {
    HashGrid* g = new HashGrid;
    std::vector<V> pts(n);
    for (uint64_t i = 0; i < n; i++) pts[i] = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    g->build(pts, radius, n_cells);
    return g;
}

void orc_hashgrid_export(void* gp, float* bbox, int32_t* cell_ends, int32_t* indices) {
    HashGrid* g = (HashGrid*)gp;
    for (int a = 0; a < 3; a++) {
        bbox[a] = g->bmin[a];
        bbox[3 + a] = g->bmax[a];
    }
    std::copy(g->cellEnds.begin(), g->cellEnds.end(), cell_ends);
    std::copy(g->indices.begin(), g->indices.end(), indices);
}

void orc_hashgrid_query(void* gp, const float* pos, uint64_t n, const float* q, uint64_t nq, uint32_t cap,
                        uint32_t* counts, int32_t* idx) {
    HashGrid* g = (HashGrid*)gp;
    std::vector<V> pts(n);
    for (uint64_t i = 0; i < n; i++) pts[i] = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    std::vector<int> out;
    for (uint64_t k = 0; k < nq; k++) {
        g->query(pts, V(q[3 * k], q[3 * k + 1], q[3 * k + 2]), out);
        counts[k] = (uint32_t)out.size();
        for (uint32_t j = 0; j < cap && j < out.size(); j++) idx[(size_t)k * cap + j] = out[j];
    }
}

void orc_hashgrid_free(void* gp) { delete (HashGrid*)gp; }

void orc_scene_set_motion(void* sp, uint32_t obj, float t0, float t1, const float* start, const float* end) {
    Scene* s = (Scene*)sp;
    Object& o = s->objs[obj];
    M4 a, b;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a.a[i][j] = start[4 * i + j];
            b.a[i][j] = end[4 * i + j];
        }
    if (!o.moving) {
        o.localVerts = o.verts;
        o.localNormals = o.normals;
        o.moving = true;
    }
    o.anim = Animated(a, t0, b, t1);
    o.place(o.anim.interpolate(t0));
    s->build();
}

void orc_motion_interpolate(float t0, float t1, const float* start, const float* end, float time, float* out) {
    M4 a, b;
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a.a[i][j] = start[4 * i + j];
            b.a[i][j] = end[4 * i + j];
        }
    const M4 t = Animated(a, t0, b, t1).interpolate(time);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) out[4 * i + j] = t.a[i][j];
}

void orc_render_bdpt(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t pass0, uint32_t npass,
                     double* sum, double* sumsq, int nthreads) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    const int64_t np = (int64_t)c->width * c->height;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<V> iter(np);
    std::vector<std::vector<std::pair<unsigned, V>>> splats(np);
    for (uint32_t pass = pass0; pass < pass0 + npass; pass++) {
        s->updateForMotion(seed, pass);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
        for (int64_t p = 0; p < np; p++) {
            splats[p].clear();
            tr.bdptPixel((unsigned)p, pass, iter[p], splats[p]);
        }
        for (int64_t p = 0; p < np; p++)
            for (const auto& sp2 : splats[p]) iter[sp2.first] = add(iter[sp2.first], sp2.second);
        for (int64_t p = 0; p < np; p++) {
            V v = iter[p];
            if (!legal(v)) v = V(0, 0, 0);
            v = tr.cam.fixVignetting(v, (unsigned)p);
            for (int k = 0; k < 3; k++) {
                sum[3 * p + k] += v[k];
                sumsq[3 * p + k] += (double)v[k] * v[k];
            }
        }
    }
    restoreMotion(s);
}

void orc_render_vcm(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t pass0, uint32_t npass,
                    float base_radius, double* sum, double* sumsq, int nthreads) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    const int64_t np = (int64_t)c->width * c->height;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<V> iter(np);
    std::vector<std::vector<std::pair<unsigned, V>>> splats(np);
    std::vector<std::vector<Tracer::PRay>> lightPaths(np);
    for (uint32_t pass = pass0; pass < pass0 + npass; pass++) {
        const Vcm vcm{tr, vcmRadius(base_radius, pass), (unsigned)np};
        /* step 1 (vcmTracer.cpp:36-57): one light path per pixel; its light points
         * after the np default-constructed ones, in path order */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
        for (int64_t p = 0; p < np; p++) lightPaths[p] = tr.lightPath((unsigned)p, pass);
        std::vector<V> pts((size_t)np, V(0, 0, 0));
        std::vector<std::pair<int, int>> src((size_t)np, {-1, -1});
        for (int64_t p = 0; p < np; p++) {
            const auto& L = lightPaths[p];
            for (size_t i = 1; i < L.size(); i++) {
                if (vcm.emissiveObj(L[i].contact)) break;
                if (L[i].delta) continue;
                pts.push_back(L[i].o);
                src.push_back({(int)p, (int)i});
            }
        }
        HashGrid grid;
        grid.build(pts, vcm.radius, (size_t)np);
        /* step 2 (vcmTracer.cpp:59-70) */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
        for (int64_t p = 0; p < np; p++) {
            splats[p].clear();
            vcm.pixel((unsigned)p, pass, lightPaths, pts, src, grid, iter[p], splats[p]);
        }
        for (int64_t p = 0; p < np; p++)
            for (const auto& sp2 : splats[p]) iter[sp2.first] = add(iter[sp2.first], sp2.second);
        for (int64_t p = 0; p < np; p++) {
            V v = iter[p];
            if (!legal(v)) v = V(0, 0, 0);
            v = tr.cam.fixVignetting(v, (unsigned)p);
            for (int k = 0; k < 3; k++) {
                sum[3 * p + k] += v[k];
                sumsq[3 * p + k] += (double)v[k] * v[k];
            }
        }
    }
}

void orc_render_ppm(void* sp, const orc_camera* c, uint32_t max_depth, uint64_t seed, uint32_t pass0, uint32_t npass,
                    float base_radius, double* sum, double* sumsq, int nthreads) {
    Scene* s = (Scene*)sp;
    Tracer tr{s, Cam(c), max_depth, seed};
    Tracer te{s, Cam(c), 10, seed};  /* MERGE_LEN (photonMap.cpp:10) */
    const int64_t np = (int64_t)c->width * c->height;
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    std::vector<V> iter(np);
    std::vector<std::vector<Tracer::PRay>> lightPaths(np);
    for (uint32_t pass = pass0; pass < pass0 + npass; pass++) {
        const Ppm ppm{tr, te, vcmRadius(base_radius, pass), (unsigned)np};
        /* step 1 (photonMap.cpp:59-96) */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
        for (int64_t p = 0; p < np; p++) lightPaths[p] = tr.lightPath((unsigned)p, pass);
        std::vector<V> pts((size_t)np, V(0, 0, 0));
        std::vector<std::pair<int, int>> src((size_t)np, {-1, -1});
        for (int64_t p = 0; p < np; p++) {
            const auto& L = lightPaths[p];
            for (size_t i = 1; i < L.size(); i++) {
                if (L[i].contact != ORC_NONE && s->objs[L[i].contact].mat == ORC_EMISSIVE) break;
                if (L[i].delta) continue;
                pts.push_back(L[i].o);
                src.push_back({(int)p, (int)i});
            }
        }
        HashGrid grid;
        grid.build(pts, ppm.radius, (size_t)np);
        /* step 2 (photonMap.cpp:98-116) */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads)
        for (int64_t p = 0; p < np; p++) iter[p] = ppm.pixel((unsigned)p, pass, lightPaths, pts, src, grid);
        for (int64_t p = 0; p < np; p++) {
            V v = iter[p];
            if (!legal(v)) v = V(0, 0, 0);
            v = tr.cam.fixVignetting(v, (unsigned)p);
            for (int k = 0; k < 3; k++) {
                sum[3 * p + k] += v[k];
                sumsq[3 * p + k] += (double)v[k] * v[k];
            }
        }
    }
}

float orc_vcm_radius(float base_radius, uint32_t pass) { return vcmRadius(base_radius, pass); }

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

void orc_rng4(uint64_t seed, uint32_t pixel, uint32_t pass, uint32_t vertex, float out[4]) {
    rng4(seed, pixel, pass, vertex, out);
}

}  // extern "C"
