// builder.cpp -- the reference KD-tree builder, re-implemented for large
// scenes, and its flattening to the GPU layout of wr_layout.h.
//
// Algorithm kept from KDTree::splitBest (kdtree.cpp:141-240); it must be
// identical because tie-breaks between equal-distance hits and the reference's
// box-test rounding both depend on the exact tree:
//   * node box = componentwise min/max over every vertex of every triangle,
//     scanned in list order with `p < m ? p : m` (kdtree.cpp:5-21,151-165);
//   * leaf if the node holds <= 1 triangle (maxLeafTriangleNum = 1,
//     kdtree.cpp:25,170);
//   * for dim 0,1,2: split = element of rank floor(U/2) among the coordinates
//     of the node's U distinct vertex indices (kdtree.cpp:154-183); leftCount =
//     triangles with any vertex coordinate <= split; keep the dim whose
//     |leftCount/n - 0.5| (evaluated in double) is strictly smallest
//     (kdtree.cpp:187-208);
//   * leaf if the best ratio is 0 or 1 (kdtree.cpp:211-215);
//   * order-preserving partition, "any vertex <= split" goes left
//     (kdtree.cpp:220-236).
// What differs is only how it runs: one array of triangle records (vertex
// coordinates and indices) is partitioned in place (stable), distinct vertices
// are found with a per-thread stamp array instead of an unordered_set, the
// median is a selection chosen by size (branch-free rank count, branch-free
// quickselect, radix select) that returns std::nth_element's value, large
// subtrees are built on their own threads and the scans of a large node are
// split over workers, and the recursion leaves the triangles of each leaf
// contiguous and in left-to-right leaf order, which is exactly the record
// order the GPU layout wants.  A NaN vertex coordinate is refused: the
// reference's median (std::nth_element, kdtree.cpp:179) is undefined on it.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <thread>

#include "wr_host.h"

namespace wr {
namespace {

struct BNode {
    uint32_t b = 0, e = 0;
    uint32_t n_int = 0;  // internal nodes in this subtree (0 for a leaf)
    float box[6];
    std::unique_ptr<BNode[]> kids;  // left, right; null for a leaf
};

// One triangle as the split sees it: its vertex coordinates (vertex-major,
// as the reference reads them), its vertex indices and its input index.
// These records are partitioned in place of the triangle order, so every scan
// over a node reads them in sequence and the vertex array is read only once.
struct TriRec {
    float p[9];
    uint32_t v[3];
    uint32_t ti;
};

struct Scratch {
    // per vertex: the id of the last node that met it (atomic so the workers
    // of one large node can share it; relaxed loads and stores elsewhere)
    std::unique_ptr<std::atomic<uint32_t>[]> stamp;
    uint32_t stamp_id = 0;
    std::vector<float> u[3];  // per dim, the coordinates of the node's distinct vertices
};


// Order-preserving map of a non-NaN float to an unsigned key (-0 sorts just
// below +0, which compare equal as floats).
inline uint32_t float_key(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}

inline float key_float(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k ^ 0x80000000u) : ~k;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// The k-th smallest of x (NaN-free, k < x.size()) by a most-significant-digit
// radix select over float_key: three counting passes of 11, 11 and 10 bits,
// each keeping only the bucket that holds rank k.  The value equals the one
// std::nth_element puts at position k, but for the sign of a zero, which no
// comparison against it can see.  Reorders nothing; x is only read.
float select_kth(const std::vector<float>& x, size_t k) {
    uint32_t hist[2048];
    std::vector<uint32_t> keep;
    const size_t n = x.size();
    // first digit: over x directly
    std::fill(hist, hist + 2048, 0u);
    for (size_t i = 0; i < n; i++) hist[float_key(x[i]) >> 21]++;
    uint32_t d = 0;
    while (k >= hist[d]) k -= hist[d++];
    uint32_t prefix = d << 21;
    keep.resize(hist[d] + 1);  // + 1: the branch-free filter writes one past the last kept key
    size_t m = 0;
    for (size_t i = 0; i < n; i++) {
        const uint32_t key = float_key(x[i]);
        keep[m] = key;
        m += (key >> 21) == d;
    }
    // second and third digits: over the kept keys, compacted in place
    const int shift[2] = {10, 0};
    const uint32_t width[2] = {11, 10};
    for (int p = 0; p < 2; p++) {
        const uint32_t mask = (1u << width[p]) - 1u;
        std::fill(hist, hist + (1u << width[p]), 0u);
        for (size_t i = 0; i < m; i++) hist[(keep[i] >> shift[p]) & mask]++;
        d = 0;
        while (k >= hist[d]) k -= hist[d++];
        prefix |= d << shift[p];
        size_t w = 0;
        for (size_t i = 0; i < m; i++) {
            const uint32_t key = keep[i];
            keep[w] = key;
            w += ((key >> shift[p]) & mask) == d;
        }
        m = w;
    }
    return key_float(prefix);
}

// The k-th smallest of a[0, n) (NaN-free, k < n), reordering a: quickselect
// with branch-free partitions (the swap-always form of Lomuto's scheme), one
// pass splitting off the values below the pivot and, when rank k is not among
// them, one splitting the rest at the pivot, so runs of equal values end it
// (std::nth_element takes over after 64 rounds).  Same value as std::nth_element but for the sign of a zero.
float quickselect(float* a, size_t n, size_t k) {
    size_t lo = 0, hi = n;
    for (int budget = 64; hi - lo > 12; budget--) {
        if (budget == 0) {  // bad pivots all along: the library's introselect
            std::nth_element(a + lo, a + k, a + hi);
            return a[k];
        }
        const float x = a[lo], y = a[lo + (hi - lo) / 2], z = a[hi - 1];
        const float p = std::max(std::min(x, y), std::min(std::max(x, y), z));  // median of three
        size_t j = lo;
        for (size_t i = lo; i < hi; i++) {
            const float v = a[i];
            a[i] = a[j];
            a[j] = v;
            j += v < p;
        }
        if (k < j) {
            hi = j;
            continue;
        }
        size_t m = j;
        for (size_t i = j; i < hi; i++) {
            const float v = a[i];
            a[i] = a[m];
            a[m] = v;
            m += !(p < v);
        }
        if (k < m) return p;
        lo = m;
    }
    for (size_t i = lo + 1; i < hi; i++)  // insertion sort of the last few
        for (size_t j = i; j > lo && a[j] < a[j - 1]; j--) std::swap(a[j], a[j - 1]);
    return a[k];
}

// The k-th smallest of a[0, n) for a small NaN-free n, without branches: the
// largest value with at most k values below it.
float select_small(const float* a, size_t n, size_t k) {
    float best = -INFINITY;
    for (size_t i = 0; i < n; i++) {
        uint32_t below = 0;
        for (size_t j = 0; j < n; j++) below += a[j] < a[i];
        best = (below <= k && a[i] > best) ? a[i] : best;
    }
    return best;
}

// Runs fn(0) .. fn(n - 1), each on its own thread but the last.
template <class Fn>
void run_parallel(int n, Fn fn) {
    std::vector<std::thread> pool;
    for (int w = 0; w + 1 < n; w++) pool.emplace_back(fn, w);
    fn(n - 1);
    for (std::thread& t : pool) t.join();
}

// Folds the coordinates of recs[b, e) into a box with the reference's rule,
// `p < m ? p : m` (and `>` for the maximum) over every vertex in list order,
// starting from the first vertex.
inline void fold_box(const TriRec* recs, uint32_t b, uint32_t e, float* mn, float* mx) {
    for (int a = 0; a < 3; a++) mn[a] = mx[a] = recs[b].p[a];
    for (uint32_t i = b; i < e; i++)
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) {
                const float p = recs[i].p[3 * k + a];
                mn[a] = p < mn[a] ? p : mn[a];
                mx[a] = p > mx[a] ? p : mx[a];
            }
}

// "any vertex coordinate <= split" (kdtree.cpp:192-198,224-230)
inline bool any_le(const TriRec& R, int dim, float split) {
    return ((int)(R.p[dim] <= split) | (int)(R.p[3 + dim] <= split) | (int)(R.p[6 + dim] <= split)) != 0;
}

struct Builder {
    const float* V;
    const uint32_t* I;
    uint32_t nv, nt;
    int nthreads = 1;
    std::vector<TriRec> recs, rtmp;
    std::atomic<int> spare{0};
    std::atomic<uint32_t> depth_seen{0};
    static constexpr uint32_t kParallelMin = 1u << 14;
    static constexpr uint32_t kChunk = 1u << 15;  // triangles per worker in one node's scans
    static constexpr int kMaxWorkers = 64;
    static constexpr size_t kSmallSelect = 24;    // distinct vertices up to which the median is select_small
    static constexpr size_t kRadixMin = 1024;     // distinct vertices from which the median is a radix select

    void note_depth(uint32_t d) {
        uint32_t cur = depth_seen.load(std::memory_order_relaxed);
        while (d > cur && !depth_seen.compare_exchange_weak(cur, d)) {
        }
    }

    void init_recs() {
        recs.resize(nt);
        rtmp.resize(nt);
        const int nw = nt >= 4 * kChunk ? std::min(nthreads, kMaxWorkers) : 1;
        char nan[kMaxWorkers] = {};
        run_parallel(nw, [&](int w) {
            const uint32_t r0 = (uint32_t)((uint64_t)nt * w / nw), r1 = (uint32_t)((uint64_t)nt * (w + 1) / nw);
            char f = 0;
            for (uint32_t ti = r0; ti < r1; ti++) {
                TriRec& R = recs[ti];
                for (int k = 0; k < 3; k++) {
                    R.v[k] = I[3 * (size_t)ti + k];
                    for (int a = 0; a < 3; a++) {
                        const float p = V[3 * (size_t)R.v[k] + a];
                        R.p[3 * k + a] = p;
                        f |= (char)(p != p);
                    }
                }
                R.ti = ti;
            }
            nan[w] = f;
        });
        for (int w = 0; w < nw; w++)
            if (nan[w]) throw std::runtime_error("a vertex coordinate is NaN (the split rule's median is undefined)");
    }

    // The upper median of the distinct vertices' coordinates along dim (the
    // order of u does not change it).
    float median(std::vector<float>& u) const {
        const size_t U = u.size();
        if (U <= kSmallSelect) return select_small(u.data(), U, U >> 1);
        if (U < kRadixMin) return quickselect(u.data(), U, U >> 1);
        return select_kth(u, U >> 1);
    }

    void split(BNode* node, Scratch& s, uint32_t depth) {
        const uint32_t b = node->b, e = node->e, n = e - b;
        if (n == 0) return;
        note_depth(depth);
        // workers for this node's scans: about the threads this level of the
        // tree has, and one per kChunk triangles
        int nw = (int)std::min<uint64_t>(std::max(1, nthreads >> std::min(depth, 20u)), n / kChunk);
        nw = std::min(nw, kMaxWorkers);
        if (nw < 2) nw = 1;
        auto lo_of = [&](int w) { return b + (uint32_t)((uint64_t)n * w / nw); };

        if (!s.stamp) {
            s.stamp.reset(new std::atomic<uint32_t>[nv]);
            for (uint32_t i = 0; i < nv; i++) s.stamp[i].store(0u, std::memory_order_relaxed);
        }
        const uint32_t sid = ++s.stamp_id;
        std::atomic<uint32_t>* st = s.stamp.get();
        const TriRec* rc = recs.data();

        // node box and distinct vertices (kdtree.cpp:151-165)
        float mn[3], mx[3];
        if (nw == 1) {
            fold_box(rc, b, e, mn, mx);
            for (int a = 0; a < 3; a++) s.u[a].clear();
            for (uint32_t i = b; i < e; i++)
                for (int k = 0; k < 3; k++) {
                    const uint32_t vi = rc[i].v[k];
                    if (st[vi].load(std::memory_order_relaxed) != sid) {
                        st[vi].store(sid, std::memory_order_relaxed);
                        for (int a = 0; a < 3; a++) s.u[a].push_back(rc[i].p[3 * k + a]);
                    }
                }
        } else {
            // a worker's fold keeps the first of its smallest (largest)
            // values, and folding those in worker order keeps the first
            // overall, as the serial fold does (no NaN: init_recs refuses it)
            float wbox[kMaxWorkers][6];
            std::vector<float> wu[kMaxWorkers][3];
            run_parallel(nw, [&](int w) {
                fold_box(rc, lo_of(w), lo_of(w + 1), wbox[w], wbox[w] + 3);
                for (uint32_t i = lo_of(w); i < lo_of(w + 1); i++)
                    for (int k = 0; k < 3; k++) {
                        const uint32_t vi = rc[i].v[k];
                        if (st[vi].load(std::memory_order_relaxed) != sid &&
                            st[vi].exchange(sid, std::memory_order_relaxed) != sid)
                            for (int a = 0; a < 3; a++) wu[w][a].push_back(rc[i].p[3 * k + a]);
                    }
            });
            for (int a = 0; a < 3; a++) {
                mn[a] = wbox[0][a];
                mx[a] = wbox[0][3 + a];
                size_t tot = 0;
                for (int w = 0; w < nw; w++) {
                    mn[a] = wbox[w][a] < mn[a] ? wbox[w][a] : mn[a];
                    mx[a] = wbox[w][3 + a] > mx[a] ? wbox[w][3 + a] : mx[a];
                    tot += wu[w][a].size();
                }
                s.u[a].resize(tot);
                size_t o = 0;
                for (int w = 0; w < nw; w++) {
                    std::copy(wu[w][a].begin(), wu[w][a].end(), s.u[a].begin() + o);
                    o += wu[w][a].size();
                }
            }
        }
        for (int a = 0; a < 3; a++) {
            node->box[a] = mn[a];
            node->box[3 + a] = mx[a];
        }
        if (n <= 1) return;

        // per dim: the split and the triangles with a vertex at or below it
        // (kdtree.cpp:177-208); a large node's three medians run side by side
        float dsplit[3];
        if (n >= kParallelMin * 4) {
            std::thread t1([&]() { dsplit[1] = median(s.u[1]); });
            std::thread t2([&]() { dsplit[2] = median(s.u[2]); });
            dsplit[0] = median(s.u[0]);
            t1.join();
            t2.join();
        } else {
            for (int dim = 0; dim < 3; dim++) dsplit[dim] = median(s.u[dim]);
        }
        uint32_t wleft[kMaxWorkers][3];
        run_parallel(nw, [&](int w) {
            uint32_t l0 = 0, l1 = 0, l2 = 0;
            for (uint32_t i = lo_of(w); i < lo_of(w + 1); i++) {
                l0 += (uint32_t)any_le(rc[i], 0, dsplit[0]);
                l1 += (uint32_t)any_le(rc[i], 1, dsplit[1]);
                l2 += (uint32_t)any_le(rc[i], 2, dsplit[2]);
            }
            wleft[w][0] = l0;
            wleft[w][1] = l1;
            wleft[w][2] = l2;
        });
        float bestRatio = 10.f, bestSplit = 0.f;
        int bestDim = 0;
        for (int dim = 0; dim < 3; dim++) {
            uint32_t left = 0;
            for (int w = 0; w < nw; w++) left += wleft[w][dim];
            const float ratio = left / float(n);
            if (std::fabs((double)ratio - 0.5) < std::fabs((double)bestRatio - 0.5)) {
                bestRatio = ratio;
                bestDim = dim;
                bestSplit = dsplit[dim];
            }
        }
        if (bestRatio == 0 || bestRatio == 1) return;  // kdtree.cpp:211-215

        // stable partition of recs[b, e): left part first (kdtree.cpp:220-236)
        uint32_t nleft = 0;
        for (int w = 0; w < nw; w++) nleft += wleft[w][bestDim];
        const uint32_t mid = b + nleft;
        if (nw == 1) {
            uint32_t wl = b, wr_ = b;
            for (uint32_t i = b; i < e; i++) {
                if (any_le(recs[i], bestDim, bestSplit)) recs[wl++] = recs[i];
                else rtmp[wr_++] = recs[i];
            }
            std::copy(rtmp.begin() + b, rtmp.begin() + wr_, recs.begin() + wl);
        } else {
            uint32_t lpos[kMaxWorkers], rpos[kMaxWorkers];
            for (uint32_t w = 0, l = b, r = mid; w < (uint32_t)nw; w++) {
                lpos[w] = l;
                rpos[w] = r;
                l += wleft[w][bestDim];
                r += lo_of(w + 1) - lo_of(w) - wleft[w][bestDim];
            }
            run_parallel(nw, [&](int w) {
                uint32_t l = lpos[w], r = rpos[w];
                for (uint32_t i = lo_of(w); i < lo_of(w + 1); i++) {
                    if (any_le(recs[i], bestDim, bestSplit)) rtmp[l++] = recs[i];
                    else rtmp[r++] = recs[i];
                }
            });
            run_parallel(nw, [&](int w) { std::copy(rtmp.begin() + lo_of(w), rtmp.begin() + lo_of(w + 1), recs.begin() + lo_of(w)); });
        }

        node->kids.reset(new BNode[2]);
        BNode* L = &node->kids[0];
        BNode* R = &node->kids[1];
        L->b = b;
        L->e = mid;
        R->b = mid;
        R->e = e;
        if (n >= kParallelMin && spare.fetch_sub(1) > 0) {
            std::thread th([this, L, depth]() {
                Scratch s2;
                split(L, s2, depth + 1);
            });
            split(R, s, depth + 1);
            th.join();
            spare.fetch_add(1);
        } else {
            if (n >= kParallelMin) spare.fetch_add(1);
            split(L, s, depth + 1);
            split(R, s, depth + 1);
        }
        node->n_int = 1 + L->n_int + R->n_int;
    }
};

// normalize((v1 - v0) x (v2 - v1)) with the reference's float semantics
// (kdtree.cpp:288-296; nvVector.h:490-525: sequential dot from 0, sqrt, r > 0
// guard, true division).
void ref_normal(const float* v0, const float* v1, const float* v2, float* out) {
    float l0[3], l1[3];
    for (int k = 0; k < 3; k++) {
        l0[k] = v1[k] - v0[k];
        l1[k] = v2[k] - v1[k];
    }
    float n[3] = {l0[1] * l1[2] - l0[2] * l1[1], l0[2] * l1[0] - l0[0] * l1[2], l0[0] * l1[1] - l0[1] * l1[0]};
    float r2 = 0.f;
    r2 += n[0] * n[0];
    r2 += n[1] * n[1];
    r2 += n[2] * n[2];
    const float r = std::sqrt(r2);
    if (r > 0) {
        for (int k = 0; k < 3; k++) n[k] = n[k] / r;
    }
    for (int k = 0; k < 3; k++) out[k] = n[k];
}

// Writes the internal nodes under n in preorder from index k (a node, then
// its left subtree, then its right one: the right child's index follows from
// the left subtree's node count), the links and counts of the leaves and each
// record's first record; large subtrees run on their own threads.  Every
// node's children are freed once written.  Returns n's reference.
struct Flattener {
    FlatTree& F;
    std::atomic<int> spare{0};
    static constexpr uint32_t kParallelNodes = 1u << 12;
    explicit Flattener(FlatTree& f, int nthreads) : F(f) { spare = nthreads - 1; }

    uint32_t visit(BNode* n, uint32_t k, uint32_t up) {
        if (!n->kids) {
            const uint32_t cnt = n->e - n->b;
            const uint32_t field = cnt <= WR_CNT_MASK ? cnt : 0u;
            if (field == 0) F.leaf_cnt[n->b] = cnt;
            F.leaf_up[n->b] = up;
            for (uint32_t r = n->b; r < n->e; r++) F.rec_first[r] = n->b;
            return WR_LEAF_BIT | (field << WR_CNT_SHIFT) | n->b;
        }
        F.node_up[k] = up;
        BNode* nl = &n->kids[0];
        BNode* nr = &n->kids[1];
        const uint32_t kl = k + 1, kr = k + 1 + nl->n_int;
        uint32_t lr = 0, rr = 0;
        if (n->n_int >= kParallelNodes && spare.fetch_sub(1) > 0) {
            std::thread th([&]() { lr = visit(nl, kl, k << 1); });
            rr = visit(nr, kr, (k << 1) | 1u);
            th.join();
            spare.fetch_add(1);
        } else {
            if (n->n_int >= kParallelNodes) spare.fetch_add(1);
            lr = visit(nl, kl, k << 1);
            rr = visit(nr, kr, (k << 1) | 1u);
        }
        wr_node& w = F.nodes[k];
        for (int a = 0; a < 3; a++) {
            w.lmin[a] = nl->box[a];
            w.lmax[a] = nl->box[3 + a];
            w.rmin[a] = nr->box[a];
            w.rmax[a] = nr->box[3 + a];
        }
        w.lref = lr;
        w.rref = rr;
        n->kids.reset();
        return k;
    }
};

// Outward quantization of one bound on the grid origin + q * scale, checked
// against the GPU's decode (one fused multiply-add, fmaf).
uint8_t quant_lo(float v, float origin, float scale) {
    const double q = std::floor(((double)v - (double)origin) / (double)scale);
    int qi = (int)std::max(0.0, std::min(255.0, q));
    while (qi > 0 && std::fmaf((float)qi, scale, origin) > v) qi--;
    return (uint8_t)qi;
}
uint8_t quant_hi(float v, float origin, float scale) {
    const double q = std::ceil(((double)v - (double)origin) / (double)scale);
    int qi = (int)std::max(0.0, std::min(255.0, q));
    while (qi < 255 && std::fmaf((float)qi, scale, origin) < v) qi++;
    return (uint8_t)qi;
}

// Encodes a wide node over n <= 4 children with boxes cbox and references
// ref: per axis a grid origin (the children's minimum) and a power-of-two
// scale reaching their maximum within 255 steps, each child box rounded
// outward onto the grid.
void encode_wnode(wr_wnode& W, int n, const float (*cbox)[6], const uint32_t* ref) {
    float origin[3], scale[3];
    for (int a = 0; a < 3; a++) {
        float lo = cbox[0][a], hi = cbox[0][3 + a];
        for (int c = 1; c < n; c++) {
            lo = std::min(lo, cbox[c][a]);
            hi = std::max(hi, cbox[c][3 + a]);
        }
        origin[a] = lo;
        const double ext = (double)hi - (double)lo;
        int ex = 0;
        if (ext > 0) {
            std::frexp(ext / 255.0, &ex);  // ext / 255 < 2^ex
        } else {
            ex = -100;
        }
        float sc = std::ldexp(1.0f, std::max(ex, -126));
        while (std::fmaf(255.0f, sc, origin[a]) < hi) sc *= 2.0f;
        scale[a] = sc;
    }
    for (int a = 0; a < 3; a++) {
        W.origin[a] = origin[a];
        W.scale[a] = scale[a];
    }
    for (int c = 0; c < 4; c++) {
        if (c < n) {
            for (int a = 0; a < 3; a++) {
                W.q[a][c] = quant_lo(cbox[c][a], origin[a], scale[a]);
                W.q[3 + a][c] = quant_hi(cbox[c][3 + a], origin[a], scale[a]);
            }
            W.ref[c] = ref[c];
        } else {
            for (int k = 0; k < 6; k++) W.q[k][c] = 0;
            W.ref[c] = WR_NONE;
        }
    }
}

// The fast-pass hierarchy of `prev` (the same triangles at other positions)
// refit to F: every leaf reference renamed to the record that now holds its
// triangle, every box recomputed bottom-up from the current vertices and
// encoded outward again.  The topology is prev's, so its quality is prev's
// for the positions prev was built at; correctness needs only that every box
// contains the triangles below it, which the recomputation gives.
void refit_wide(FlatTree& F, const FlatTree& prev, int nthreads) {
    const size_t nr = F.tris.size();
    std::vector<uint32_t> rec_of(nr);  // input triangle -> record
    for (size_t r = 0; r < nr; r++) rec_of[F.rec_prim[r]] = (uint32_t)r;
    F.wnodes = prev.wnodes;
    const size_t nw = F.wnodes.size();
    auto rename = [&](uint32_t ref) {
        if (ref == WR_NONE || !(ref & WR_LEAF_BIT)) return ref;
        const uint32_t old = ref & WR_FIRST_MASK;
        return (ref & ~WR_FIRST_MASK) | rec_of[prev.rec_prim[old]];
    };
    auto tri_box = [&](uint32_t rec, float* b) {
        const wr_tri& T = F.tris[rec];
        for (int a = 0; a < 3; a++) {
            b[a] = std::min(std::min(T.v0[a], T.v1[a]), T.v2[a]);
            b[3 + a] = std::max(std::max(T.v0[a], T.v1[a]), T.v2[a]);
        }
    };
    // nodes are in preorder (a node's wide children come after it): their depths in
    // one forward sweep, then the levels from the deepest up, each level's nodes
    // independent of one another
    std::vector<uint32_t> depth(nw, 0u);
    uint32_t max_d = 0;
    for (size_t i = 0; i < nw; i++)
        for (int c = 0; c < 4; c++) {
            const uint32_t r = F.wnodes[i].ref[c];
            if (r != WR_NONE && !(r & WR_LEAF_BIT)) {
                depth[r] = depth[i] + 1;
                max_d = std::max(max_d, depth[r]);
            }
        }
    std::vector<std::vector<uint32_t>> level(max_d + 1);
    for (size_t i = 0; i < nw; i++) level[depth[i]].push_back((uint32_t)i);
    std::vector<std::array<float, 6>> nbox(nw);
    auto refit_node = [&](uint32_t i) {
        wr_wnode& W = F.wnodes[i];
        float cbox[4][6];
        uint32_t ref[4];
        int n = 0;
        for (int c = 0; c < 4 && W.ref[c] != WR_NONE; c++, n++) {
            ref[c] = rename(W.ref[c]);
            if (ref[c] & WR_LEAF_BIT) tri_box(ref[c] & WR_FIRST_MASK, cbox[c]);
            else std::copy(nbox[ref[c]].begin(), nbox[ref[c]].end(), cbox[c]);
        }
        auto& nb = nbox[i];
        for (int a = 0; a < 3; a++) {
            nb[a] = cbox[0][a];
            nb[3 + a] = cbox[0][3 + a];
            for (int c = 1; c < n; c++) {
                nb[a] = std::min(nb[a], cbox[c][a]);
                nb[3 + a] = std::max(nb[3 + a], cbox[c][3 + a]);
            }
        }
        encode_wnode(W, n, cbox, ref);
    };
    for (size_t d = level.size(); d-- > 0;) {
        const std::vector<uint32_t>& L = level[d];
        const int nt = (int)std::min<size_t>((size_t)std::max(nthreads, 1), L.size() / 2048 + 1);
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (size_t k; (k = next.fetch_add(256)) < L.size();)
                for (size_t j = k; j < std::min(L.size(), k + 256); j++) refit_node(L[j]);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; t++) pool.emplace_back(worker);
        worker();
        for (std::thread& t : pool) t.join();
    }
    F.hdr.wide_root = rename(prev.hdr.wide_root);
}

// The fast-pass hierarchy: a 4-wide BVH over the triangle records, built
// top-down with a binned surface-area heuristic (two binary splits per wide
// node), one triangle per leaf reference.  It is independent of the reference
// tree: the fast pass only needs every triangle under a box that contains it
// (superset of the reference's candidates) and every record's tie key
// (rec_first); exactness comes from verifying the winner on the reference
// tree.
struct WideBuilder {
    const FlatTree& F;
    std::vector<wr_wnode>& out;
    std::vector<std::array<float, 6>> box;  // per record
    std::vector<std::array<float, 3>> cen;  // per record (box centre)
    std::vector<uint32_t> idx;              // record order, partitioned in place

    static uint32_t tri_ref(uint32_t rec) { return WR_LEAF_BIT | (1u << WR_CNT_SHIFT) | rec; }

    WideBuilder(const FlatTree& f, std::vector<wr_wnode>& o) : F(f), out(o) {
        const size_t n = F.tris.size();
        box.resize(n);
        cen.resize(n);
        idx.resize(n);
        for (size_t r = 0; r < n; r++) {
            const wr_tri& T = F.tris[r];
            for (int a = 0; a < 3; a++) {
                box[r][a] = std::min(std::min(T.v0[a], T.v1[a]), T.v2[a]);
                box[r][3 + a] = std::max(std::max(T.v0[a], T.v1[a]), T.v2[a]);
                cen[r][a] = 0.5f * box[r][a] + 0.5f * box[r][3 + a];
            }
            idx[r] = (uint32_t)r;
        }
    }

    void range_box(uint32_t b, uint32_t e, float* bx) const {
        for (int a = 0; a < 3; a++) {
            bx[a] = INFINITY;
            bx[3 + a] = -INFINITY;
        }
        for (uint32_t i = b; i < e; i++) {
            const auto& t = box[idx[i]];
            for (int a = 0; a < 3; a++) {
                bx[a] = std::min(bx[a], t[a]);
                bx[3 + a] = std::max(bx[3 + a], t[3 + a]);
            }
        }
    }

    static double half_area(const float* lo, const float* hi) {
        const double x = std::max(0.0, (double)hi[0] - lo[0]), y = std::max(0.0, (double)hi[1] - lo[1]),
                     z = std::max(0.0, (double)hi[2] - lo[2]);
        return x * y + y * z + z * x;
    }

    // Splits [b, e) (at least two records) in two by a binned SAH over the
    // box centres; returns the split point.  Deterministic.
    uint32_t split(uint32_t b, uint32_t e) {
        constexpr int kBins = 16;
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = b; i < e; i++)
            for (int a = 0; a < 3; a++) {
                clo[a] = std::min(clo[a], cen[idx[i]][a]);
                chi[a] = std::max(chi[a], cen[idx[i]][a]);
            }
        double best = INFINITY;
        int best_axis = -1, best_bin = 0;
        for (int a = 0; a < 3; a++) {
            const double ext = (double)chi[a] - clo[a];
            if (!(ext > 0)) continue;
            uint32_t cnt[kBins] = {};
            float blo[kBins][3], bhi[kBins][3];
            for (int k = 0; k < kBins; k++)
                for (int c = 0; c < 3; c++) {
                    blo[k][c] = INFINITY;
                    bhi[k][c] = -INFINITY;
                }
            const double sc = kBins / ext;
            for (uint32_t i = b; i < e; i++) {
                const uint32_t r = idx[i];
                const int k = std::min(kBins - 1, (int)(((double)cen[r][a] - clo[a]) * sc));
                cnt[k]++;
                for (int c = 0; c < 3; c++) {
                    blo[k][c] = std::min(blo[k][c], box[r][c]);
                    bhi[k][c] = std::max(bhi[k][c], box[r][3 + c]);
                }
            }
            double lcost[kBins];  // cost of bins [0, k] on the left
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            uint32_t n = 0;
            for (int k = 0; k < kBins - 1; k++) {
                n += cnt[k];
                for (int c = 0; c < 3; c++) {
                    lo[c] = std::min(lo[c], blo[k][c]);
                    hi[c] = std::max(hi[c], bhi[k][c]);
                }
                lcost[k] = n ? half_area(lo, hi) * n : 0.0;
            }
            for (int c = 0; c < 3; c++) {
                lo[c] = INFINITY;
                hi[c] = -INFINITY;
            }
            n = 0;
            for (int k = kBins - 1; k > 0; k--) {
                n += cnt[k];
                for (int c = 0; c < 3; c++) {
                    lo[c] = std::min(lo[c], blo[k][c]);
                    hi[c] = std::max(hi[c], bhi[k][c]);
                }
                const uint32_t nl = (e - b) - n;
                if (n == 0 || nl == 0) continue;
                const double cost = lcost[k - 1] + half_area(lo, hi) * n;
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_bin = k;
                }
            }
        }
        if (best_axis < 0) return b + (e - b) / 2;  // coincident centres: split by order
        const double ext = (double)chi[best_axis] - clo[best_axis];
        const double sc = kBins / ext;
        const int a = best_axis;
        const float c0 = clo[a];
        auto mid = std::stable_partition(idx.begin() + b, idx.begin() + e, [&](uint32_t r) {
            return std::min(kBins - 1, (int)(((double)cen[r][a] - c0) * sc)) < best_bin;
        });
        const uint32_t m = (uint32_t)(mid - idx.begin());
        if (m == b || m == e) return b + (e - b) / 2;
        return m;
    }

    // Child record ranges of a wide node over idx[b, e) (at least two
    // records): single records for four or fewer, else two SAH splits.
    int child_ranges(uint32_t b, uint32_t e, uint32_t* cb, uint32_t* ce) {
        int n = 0;
        if (e - b <= 4) {
            for (uint32_t i = b; i < e; i++) {
                cb[n] = i;
                ce[n] = i + 1;
                n++;
            }
            return n;
        }
        const uint32_t m = split(b, e);
        const uint32_t halves[2][2] = {{b, m}, {m, e}};
        for (const auto& hv : halves) {
            if (hv[1] - hv[0] == 1) {
                cb[n] = hv[0];
                ce[n] = hv[1];
                n++;
            } else {
                const uint32_t mm = split(hv[0], hv[1]);
                cb[n] = hv[0];
                ce[n] = mm;
                n++;
                cb[n] = mm;
                ce[n] = hv[1];
                n++;
            }
        }
        return n;
    }

    // Wide subtree over the records idx[b, e) (at least two), appended to o.
    uint32_t build(std::vector<wr_wnode>& o, uint32_t b, uint32_t e) {
        uint32_t cb[4], ce[4];
        const int n = child_ranges(b, e, cb, ce);
        float cbox[4][6];
        for (int c = 0; c < n; c++) range_box(cb[c], ce[c], cbox[c]);
        return make_node(o, n, cbox, [&](int c) {
            return ce[c] - cb[c] == 1 ? tri_ref(idx[cb[c]]) : build(o, cb[c], ce[c]);
        });
    }

    // One wide node over n <= 4 children, appended to o; child_ref(c) is called
    // after the node's slot is taken (preorder) and returns the child's final
    // reference.
    template <class ChildRef>
    uint32_t make_node(std::vector<wr_wnode>& o, int n, const float (*cbox)[6], ChildRef child_ref) {
        const uint32_t node = (uint32_t)o.size();
        o.push_back(wr_wnode{});
        uint32_t wref[4];
        for (int c = 0; c < n; c++) wref[c] = child_ref(c);
        encode_wnode(o[node], n, cbox, wref);
        return node;
    }

    // Parallel build: the top levels are split here, the subtrees below them
    // are built by worker threads into their own arrays and spliced in
    // preorder, so the result is the sequential build's, node for node.
    struct PlanNode {
        int n = 0;
        uint32_t cb[4], ce[4];
        int kind[4];  // 0: one record, 1: task, 2: plan node
        uint32_t id[4];
    };
    struct Task {
        uint32_t b, e;
        std::vector<wr_wnode> nodes;
        uint32_t root;
    };
    std::vector<PlanNode> plans;
    std::vector<Task> tasks;

    uint32_t plan(uint32_t b, uint32_t e, int level, uint32_t grain) {
        PlanNode P;
        P.n = child_ranges(b, e, P.cb, P.ce);
        for (int c = 0; c < P.n; c++) {
            const uint32_t sz = P.ce[c] - P.cb[c];
            if (sz == 1) {
                P.kind[c] = 0;
            } else if (level >= 3 || sz <= grain) {
                P.kind[c] = 1;
                P.id[c] = (uint32_t)tasks.size();
                tasks.push_back(Task{P.cb[c], P.ce[c], {}, 0});
            } else {
                P.kind[c] = 2;
                P.id[c] = plan(P.cb[c], P.ce[c], level + 1, grain);
            }
        }
        plans.push_back(P);
        return (uint32_t)plans.size() - 1;
    }

    uint32_t emit(uint32_t pid) {
        const PlanNode P = plans[pid];
        float cbox[4][6];
        for (int c = 0; c < P.n; c++) range_box(P.cb[c], P.ce[c], cbox[c]);
        return make_node(out, P.n, cbox, [&](int c) -> uint32_t {
            if (P.kind[c] == 0) return tri_ref(idx[P.cb[c]]);
            if (P.kind[c] == 2) return emit(P.id[c]);
            Task& T = tasks[P.id[c]];
            const uint32_t off = (uint32_t)out.size();
            for (wr_wnode w : T.nodes) {
                for (int k = 0; k < 4; k++)
                    if (w.ref[k] != WR_NONE && !(w.ref[k] & WR_LEAF_BIT)) w.ref[k] += off;
                out.push_back(w);
            }
            std::vector<wr_wnode>().swap(T.nodes);
            return T.root + off;
        });
    }

    uint32_t build_root(int nthreads) {
        const uint32_t n = (uint32_t)idx.size();
        if (n == 0) return WR_NONE;
        if (n == 1) return tri_ref(0);
        if (nthreads <= 1 || n < (1u << 16)) return build(out, 0, n);
        const uint32_t grain = std::max<uint32_t>(4096, n / (uint32_t)(8 * nthreads));
        const uint32_t root_plan = plan(0, n, 0, grain);
        std::atomic<size_t> next{0};
        auto worker = [&]() {
            for (size_t t; (t = next.fetch_add(1)) < tasks.size();) {
                Task& T = tasks[t];
                T.nodes.reserve((T.e - T.b) / 2 + 1);
                T.root = build(T.nodes, T.b, T.e);
            }
        };
        std::vector<std::thread> pool;
        for (int i = 1; i < nthreads; i++) pool.emplace_back(worker);
        worker();
        for (auto& th : pool) th.join();
        return emit(root_plan);
    }
};

// The chain records of F (the layout is wr_chain's): one per leaf that is not
// the root, in record order (so left-to-right leaf order), each holding the
// child boxes of its first kChainLevels ancestors on its path and the link
// where it stops; the levels above come from the top record of that link
// (its boxes up to the root), one per distinct link.  rec_chain: record ->
// its leaf's chain index, WR_NONE under a root leaf.  Leaves and links are
// split over threads; the result does not depend on their number.
void build_chains(FlatTree& F, int nthreads) {
    const size_t nr = F.rec_first.size(), nn = F.nodes.size();
    std::vector<uint32_t> leaves;  // first record of each non-root leaf
    for (size_t r = 0; r < nr; r++)
        if (F.rec_first[r] == r && F.leaf_up[r] != WR_NONE) leaves.push_back((uint32_t)r);
    const size_t nl = leaves.size();
    F.chains.assign(nl, wr_chain{});
    F.rec_chain.assign(nr, WR_NONE);
    F.top_boxes.clear();
    if (nl == 0) return;
    // node depths (nodes are in preorder: a parent comes before its children)
    std::vector<uint32_t> depth(nn, 0u);
    for (size_t i = 1; i < nn; i++) depth[i] = depth[F.node_up[i] >> 1] + 1;
    auto ranges = [&](size_t n, auto fn) {
        const int nw = n >= 4096 ? std::max(1, nthreads) : 1;
        run_parallel(nw, [&](int w) { fn(n * w / nw, n * (w + 1) / nw); });
    };
    auto child_box = [&](uint32_t link) {
        const wr_node& N = F.nodes[link >> 1];
        return (link & 1u) ? N.rmin : N.lmin;  // min then max, contiguous
    };
    ranges(nl, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) {
            wr_chain& c = F.chains[i];
            uint32_t up = F.leaf_up[leaves[i]];
            while (up != WR_NONE && c.count < kChainLevels) {
                const float* bx = child_box(up);
                std::copy(bx, bx + 6, c.box[c.count]);
                c.count++;
                up = F.node_up[up >> 1];
            }
            c.next = up;
        }
    });
    // top records: one per distinct stopping link, in link order (a link is
    // parent << 1 | side, so 2 * nn of them)
    std::vector<uint32_t> first(2 * nn + 1, 0u);  // link -> first box of its top record
    for (const wr_chain& c : F.chains)
        if (c.next != WR_NONE) first[c.next] = 1u;
    std::vector<uint32_t> links;
    uint32_t boxes = 0;
    for (uint32_t l = 0; l < 2 * nn; l++) {
        const bool used = first[l] != 0;
        first[l] = boxes;
        if (used) {
            links.push_back(l);
            boxes += depth[l >> 1] + 1;
        }
    }
    F.top_boxes.resize(6 * (size_t)boxes);
    ranges(links.size(), [&](size_t a, size_t b) {
        for (size_t k = a; k < b; k++) {
            float* out = F.top_boxes.data() + 6 * (size_t)first[links[k]];
            for (uint32_t u = links[k]; u != WR_NONE; u = F.node_up[u >> 1], out += 6) {
                const float* bx = child_box(u);
                std::copy(bx, bx + 6, out);
            }
        }
    });
    ranges(nl, [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++) {
            wr_chain& c = F.chains[i];
            if (c.next != WR_NONE) {
                c.top = first[c.next];
                c.top_count = depth[c.next >> 1] + 1;
            }
            // the leaf's records are contiguous from its first
            const uint32_t f = leaves[i];
            for (uint32_t r = f; r < nr && F.rec_first[r] == f; r++) F.rec_chain[r] = (uint32_t)i;
        }
    });
}

}  // namespace

std::vector<uint32_t> wide_levels(const FlatTree& t, std::vector<uint32_t>& level_start) {
    const size_t nw = t.wnodes.size();
    std::vector<uint32_t> depth(nw, 0u);
    uint32_t max_d = 0;
    for (size_t i = 0; i < nw; i++)  // preorder: a node's wide children come after it
        for (int c = 0; c < 4; c++) {
            const uint32_t r = t.wnodes[i].ref[c];
            if (r != WR_NONE && !(r & WR_LEAF_BIT)) {
                depth[r] = depth[i] + 1;
                max_d = std::max(max_d, depth[r]);
            }
        }
    std::vector<uint32_t> count(max_d + 2, 0u);
    for (size_t i = 0; i < nw; i++) count[max_d - depth[i] + 1]++;
    for (size_t k = 1; k < count.size(); k++) count[k] += count[k - 1];
    level_start.assign(count.begin(), count.end());
    std::vector<uint32_t> order(nw);
    for (size_t i = 0; i < nw; i++) order[count[max_d - depth[i]]++] = (uint32_t)i;
    if (nw == 0) level_start.assign(1, 0u);
    return order;
}

FlatTree build_tree(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads,
                    const FlatTree* refit_from, bool wide_on_device) {
    auto t0 = std::chrono::steady_clock::now();
    if (nt > WR_MAX_RECORDS) throw std::runtime_error("too many triangles for one tree (max 2^27)");
    for (size_t i = 0; i < 3 * (size_t)nt; i++)
        if (idx[i] >= nv) throw std::runtime_error("triangle vertex index out of range");
    FlatTree F;
    F.hdr.root_ref = WR_NONE;
    F.hdr.wide_root = WR_NONE;
    F.hdr.n_recs = nt;
    float scale = 0.f;
    for (size_t i = 0; i < 3 * (size_t)nv; i++) scale = std::max(scale, std::fabs(verts[i]));
    F.hdr.scale = scale;
    if (nt == 0) {
        for (int a = 0; a < 3; a++) {
            F.hdr.root_min[a] = 0.f;
            F.hdr.root_max[a] = 0.f;
        }
        return F;
    }
    Builder B;
    B.V = verts;
    B.I = idx;
    B.nv = nv;
    B.nt = nt;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    B.nthreads = nthreads;
    B.spare = nthreads - 1;
    B.init_recs();
    BNode root;
    root.b = 0;
    root.e = nt;
    {
        Scratch s;
        B.split(&root, s, 0);
    }

    F.tris.resize(nt);
    F.rec_prim.resize(nt);
    F.leaf_up.assign(nt, WR_NONE);
    F.leaf_cnt.assign(nt, 0u);
    auto fill = [&](uint32_t r0, uint32_t r1) {
        for (uint32_t r = r0; r < r1; r++) {
            const uint32_t ti = B.recs[r].ti;
            wr_tri& t = F.tris[r];
            for (int a = 0; a < 3; a++) {
                t.v0[a] = verts[3 * (size_t)idx[3 * (size_t)ti + 0] + a];
                t.v1[a] = verts[3 * (size_t)idx[3 * (size_t)ti + 1] + a];
                t.v2[a] = verts[3 * (size_t)idx[3 * (size_t)ti + 2] + a];
            }
            ref_normal(t.v0, t.v1, t.v2, t.n);
            F.rec_prim[r] = ti;
        }
    };
    {
        const uint32_t nw = nt >= (1u << 16) ? (uint32_t)std::max(1, nthreads) : 1u;
        std::vector<std::thread> pool;
        for (uint32_t w = 1; w < nw; w++)
            pool.emplace_back(fill, (uint32_t)((uint64_t)nt * w / nw), (uint32_t)((uint64_t)nt * (w + 1) / nw));
        fill(0, (uint32_t)((uint64_t)nt / nw));
        for (std::thread& t : pool) t.join();
    }
    F.nodes.resize(root.n_int);
    F.node_up.resize(root.n_int);
    F.rec_first.resize(nt);
    Flattener fl(F, nthreads);
    F.hdr.root_ref = fl.visit(&root, 0, WR_NONE);
    for (int a = 0; a < 3; a++) {
        F.hdr.root_min[a] = root.box[a];
        F.hdr.root_max[a] = root.box[3 + a];
    }
    F.hdr.n_nodes = (uint32_t)F.nodes.size();
    F.hdr.depth = B.depth_seen.load();  // depth of the deepest leaf
    build_chains(F, nthreads);
    const bool refit = refit_from && refit_from->hdr.n_recs == nt && refit_from->rec_prim.size() == nt;
    if (refit && wide_on_device) {
        F.hdr.n_wide = refit_from->hdr.n_wide;
        F.hdr.wide_root = refit_from->hdr.wide_root;
        if (F.hdr.wide_root != WR_NONE && (F.hdr.wide_root & WR_LEAF_BIT)) {  // a one-record tree: rename its leaf
            const uint32_t prim = refit_from->rec_prim[F.hdr.wide_root & WR_FIRST_MASK];
            const uint32_t rec = (uint32_t)(std::find(F.rec_prim.begin(), F.rec_prim.end(), prim) - F.rec_prim.begin());
            F.hdr.wide_root = (F.hdr.wide_root & ~WR_FIRST_MASK) | rec;
        }
        F.n_leaves = root.n_int + 1;
        F.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return F;
    }
    if (refit) {
        refit_wide(F, *refit_from, nthreads);
    } else {
        WideBuilder wb(F, F.wnodes);
        F.wnodes.reserve(F.tris.size() / 2 + 1);
        F.hdr.wide_root = wb.build_root(nthreads);
    }
    F.hdr.n_wide = (uint32_t)F.wnodes.size();
    F.n_leaves = root.n_int + 1;
    F.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return F;
}

TreeExport export_tree(const FlatTree& t) {
    TreeExport X;
    if (t.hdr.root_ref == WR_NONE) return X;
    struct Item {
        uint32_t ref;
        float box[6];
    };
    // explicit-stack preorder: node, left subtree, right subtree
    std::vector<std::pair<Item, uint32_t>> st;  // item, index of the parent slot to patch (WR_NONE if none)
    Item root;
    root.ref = t.hdr.root_ref;
    for (int a = 0; a < 3; a++) {
        root.box[a] = t.hdr.root_min[a];
        root.box[3 + a] = t.hdr.root_max[a];
    }
    st.push_back({root, WR_NONE});
    while (!st.empty()) {
        auto [it, patch] = st.back();
        st.pop_back();
        const uint32_t me = (uint32_t)X.right.size();
        if (patch != WR_NONE) X.right[patch] = me;
        X.box.insert(X.box.end(), it.box, it.box + 6);
        if (it.ref & WR_LEAF_BIT) {
            const uint32_t first = it.ref & WR_FIRST_MASK;
            uint32_t cnt = (it.ref >> WR_CNT_SHIFT) & WR_CNT_MASK;
            if (cnt == 0) cnt = t.leaf_cnt[first];
            X.right.push_back(WR_NONE);
            X.leaf_first.push_back((uint32_t)X.tri_list.size());
            X.leaf_count.push_back(cnt);
            for (uint32_t k = 0; k < cnt; k++) X.tri_list.push_back(t.rec_prim[first + k]);
        } else {
            const wr_node& n = t.nodes[it.ref];
            X.right.push_back(WR_NONE);
            X.leaf_first.push_back(0);
            X.leaf_count.push_back(0);
            Item L, R;
            L.ref = n.lref;
            R.ref = n.rref;
            for (int a = 0; a < 3; a++) {
                L.box[a] = n.lmin[a];
                L.box[3 + a] = n.lmax[a];
                R.box[a] = n.rmin[a];
                R.box[3 + a] = n.rmax[a];
            }
            st.push_back({R, me});        // patched into right[me] when popped
            st.push_back({L, WR_NONE});   // left child is the preorder successor
        }
    }
    return X;
}

}  // namespace wr
