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
// What differs is only how it runs: one shared triangle-index array is
// partitioned in place (stable), distinct vertices are found with a per-thread
// stamp array instead of an unordered_set, large subtrees are built on their
// own threads, and the recursion leaves the triangles of each leaf contiguous
// and in left-to-right leaf order, which is exactly the record order the GPU
// layout wants.
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
    float box[6];
    std::unique_ptr<BNode> left, right;
};

struct Scratch {
    std::vector<uint32_t> stamp;
    uint32_t stamp_id = 0;
    std::vector<uint32_t> uniq;
    std::vector<float> coords;
};

struct Builder {
    const float* V;
    const uint32_t* I;
    uint32_t nv, nt;
    std::vector<uint32_t> order, tmp;
    std::atomic<int> spare{0};
    uint32_t max_depth = 0;
    std::atomic<uint32_t> depth_seen{0};
    static constexpr uint32_t kParallelMin = 1u << 14;

    float c(uint32_t vi, int dim) const { return V[3 * (size_t)vi + dim]; }

    void note_depth(uint32_t d) {
        uint32_t cur = depth_seen.load(std::memory_order_relaxed);
        while (d > cur && !depth_seen.compare_exchange_weak(cur, d)) {
        }
    }

    void split(BNode* node, Scratch& s, uint32_t depth) {
        const uint32_t b = node->b, e = node->e, n = e - b;
        if (n == 0) return;
        note_depth(depth);
        float mn[3], mx[3];
        {
            const uint32_t v0 = I[3 * (size_t)order[b]];
            for (int k = 0; k < 3; k++) mn[k] = mx[k] = c(v0, k);
        }
        if (s.stamp.size() != nv) s.stamp.assign(nv, 0u);
        const uint32_t sid = ++s.stamp_id;
        s.uniq.clear();
        for (uint32_t i = b; i < e; i++) {
            const uint32_t* t = I + 3 * (size_t)order[i];
            for (int k = 0; k < 3; k++) {
                const uint32_t vi = t[k];
                for (int a = 0; a < 3; a++) {
                    const float p = c(vi, a);
                    mn[a] = p < mn[a] ? p : mn[a];
                    mx[a] = p > mx[a] ? p : mx[a];
                }
                if (s.stamp[vi] != sid) {
                    s.stamp[vi] = sid;
                    s.uniq.push_back(vi);
                }
            }
        }
        for (int a = 0; a < 3; a++) {
            node->box[a] = mn[a];
            node->box[3 + a] = mx[a];
        }
        if (n <= 1) return;

        float bestRatio = 10.f, bestSplit = 0.f;
        int bestDim = 0;
        uint32_t bestLeft = 0;
        const size_t U = s.uniq.size();
        // per dim: the upper median of the unique vertices' coordinates and the
        // triangles with a vertex at or below it; a large node's three dims run
        // side by side (each on its own coordinate copy), then are compared in
        // dim order as the serial loop does
        float dsplit[3];
        uint32_t dleft[3];
        auto dim_stats = [&](int dim, std::vector<float>& coords) {
            coords.resize(U);
            for (size_t i = 0; i < U; i++) coords[i] = c(s.uniq[i], dim);
            std::nth_element(coords.begin(), coords.begin() + (U >> 1), coords.end());
            const float split = coords[U >> 1];
            uint32_t left = 0;
            for (uint32_t i = b; i < e; i++) {
                const uint32_t* t = I + 3 * (size_t)order[i];
                left += (uint32_t)((int)(c(t[0], dim) <= split) | (int)(c(t[1], dim) <= split) | (int)(c(t[2], dim) <= split));
            }
            dsplit[dim] = split;
            dleft[dim] = left;
        };
        if (n >= kParallelMin * 4) {
            std::vector<float> c1, c2;
            std::thread t1([&]() { dim_stats(1, c1); });
            std::thread t2([&]() { dim_stats(2, c2); });
            dim_stats(0, s.coords);
            t1.join();
            t2.join();
        } else {
            for (int dim = 0; dim < 3; dim++) dim_stats(dim, s.coords);
        }
        for (int dim = 0; dim < 3; dim++) {
            const float split = dsplit[dim];
            const uint32_t left = dleft[dim];
            const float ratio = left / float(n);
            if (std::fabs((double)ratio - 0.5) < std::fabs((double)bestRatio - 0.5)) {
                bestRatio = ratio;
                bestDim = dim;
                bestSplit = split;
                bestLeft = left;
            }
        }
        if (bestRatio == 0 || bestRatio == 1) return;

        // stable partition of order[b, e): left part first
        uint32_t wl = b, wr_ = b;
        for (uint32_t i = b; i < e; i++) {
            const uint32_t ti = order[i];
            const uint32_t* t = I + 3 * (size_t)ti;
            const bool goes_left = ((int)(c(t[0], bestDim) <= bestSplit) | (int)(c(t[1], bestDim) <= bestSplit) |
                                    (int)(c(t[2], bestDim) <= bestSplit)) != 0;
            if (goes_left) order[wl++] = ti;
            else tmp[wr_++] = ti;
        }
        std::memcpy(order.data() + wl, tmp.data() + b, (size_t)(wr_ - b) * sizeof(uint32_t));
        (void)bestLeft;

        node->left.reset(new BNode);
        node->right.reset(new BNode);
        node->left->b = b;
        node->left->e = wl;
        node->right->b = wl;
        node->right->e = e;

        BNode* L = node->left.get();
        BNode* R = node->right.get();
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

struct Flattener {
    FlatTree& F;
    uint32_t leaves = 0;
    explicit Flattener(FlatTree& f) : F(f) {}

    uint32_t leaf_ref(const BNode* n) {
        const uint32_t cnt = n->e - n->b;
        leaves++;
        uint32_t field = cnt <= WR_CNT_MASK ? cnt : 0u;
        if (field == 0) F.leaf_cnt[n->b] = cnt;
        return WR_LEAF_BIT | (field << WR_CNT_SHIFT) | n->b;
    }

    // Returns the reference to `n`; `up` is (parent << 1 | side) or WR_NONE.
    uint32_t visit(const BNode* n, uint32_t up) {
        if (!n->left) {
            uint32_t r = leaf_ref(n);
            F.leaf_up[n->b] = up;
            return r;
        }
        const uint32_t k = (uint32_t)F.nodes.size();
        F.nodes.push_back(wr_node{});
        F.node_up.push_back(up);
        const uint32_t lr = visit(n->left.get(), k << 1);
        const uint32_t rr = visit(n->right.get(), (k << 1) | 1u);
        wr_node& w = F.nodes[k];
        for (int a = 0; a < 3; a++) {
            w.lmin[a] = n->left->box[a];
            w.lmax[a] = n->left->box[3 + a];
            w.rmin[a] = n->right->box[a];
            w.rmax[a] = n->right->box[3 + a];
        }
        w.lref = lr;
        w.rref = rr;
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

// record -> first record of its reference leaf, for every leaf under ref
void fill_rec_first(FlatTree& F, uint32_t ref) {
    std::vector<uint32_t> stack{ref};
    while (!stack.empty()) {
        const uint32_t r = stack.back();
        stack.pop_back();
        if (r & WR_LEAF_BIT) {
            const uint32_t first = r & WR_FIRST_MASK;
            uint32_t cnt = (r >> WR_CNT_SHIFT) & WR_CNT_MASK;
            if (cnt == 0) cnt = F.leaf_cnt[first];
            for (uint32_t k = 0; k < cnt; k++) F.rec_first[first + k] = first;
        } else {
            stack.push_back(F.nodes[r].lref);
            stack.push_back(F.nodes[r].rref);
        }
    }
}

uint32_t internal_depth(const BNode* n) {
    if (!n->left) return 0;
    return 1 + std::max(internal_depth(n->left.get()), internal_depth(n->right.get()));
}

}  // namespace

FlatTree build_tree(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads,
                    const FlatTree* refit_from) {
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
    B.order.resize(nt);
    B.tmp.resize(nt);
    for (uint32_t i = 0; i < nt; i++) B.order[i] = i;
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    B.spare = nthreads - 1;
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
            const uint32_t ti = B.order[r];
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
    Flattener fl(F);
    F.nodes.reserve(nt);
    const uint32_t rref = fl.visit(&root, WR_NONE);
    F.hdr.root_ref = rref;
    for (int a = 0; a < 3; a++) {
        F.hdr.root_min[a] = root.box[a];
        F.hdr.root_max[a] = root.box[3 + a];
    }
    F.hdr.n_nodes = (uint32_t)F.nodes.size();
    F.hdr.depth = internal_depth(&root);
    F.rec_first.resize(F.tris.size());
    fill_rec_first(F, rref);
    if (refit_from && refit_from->hdr.n_recs == nt && refit_from->rec_prim.size() == nt) {
        refit_wide(F, *refit_from, nthreads);
    } else {
        WideBuilder wb(F, F.wnodes);
        F.wnodes.reserve(F.tris.size() / 2 + 1);
        F.hdr.wide_root = wb.build_root(nthreads);
    }
    F.hdr.n_wide = (uint32_t)F.wnodes.size();
    F.n_leaves = fl.leaves;
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
