// wr_host.h -- host-side internals of libwalkray (not part of the C ABI).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "wr_layout.h"

namespace wr {

// Output of the reference builder flattened to the GPU layout (wr_layout.h).
struct FlatTree {
    std::vector<wr_node> nodes;
    std::vector<wr_tri> tris;          // records, in left-to-right leaf order
    std::vector<uint32_t> rec_prim;    // record -> input triangle index
    std::vector<uint32_t> leaf_up;     // first record of a leaf -> (parent << 1 | side); WR_NONE for a root leaf
    std::vector<uint32_t> leaf_cnt;    // first record of a long leaf -> triangle count
    std::vector<uint32_t> node_up;     // internal node -> (parent << 1 | side); WR_NONE for the root
    std::vector<wr_wnode> wnodes;      // fast-pass hierarchy: four children per node, single triangles at the bottom
    std::vector<uint32_t> rec_first;   // record -> first record of its reference leaf
    std::vector<wr_chain> chains;      // chain record of every non-root leaf, in leaf order
    std::vector<float> top_boxes;      // top records (6 floats a box), indexed by wr_chain::top
    std::vector<uint32_t> rec_chain;   // record -> its leaf's chain (the tie key), WR_NONE under a root leaf
    wr_tree_hdr hdr{};
    uint32_t n_leaves = 0;
    double build_seconds = 0.0;
};

// KDTree::build(BEST) (kdtree.cpp:242-258, splitBest 141-240) over the given
// triangle soup, then flattened.  Deterministic; threads only change speed.
// refit_from (nullable): a tree over the same triangles (at other positions)
// whose fast-pass hierarchy is refit instead of built anew (moving objects);
// with wide_on_device the refit is left to the caller's device pass: wnodes
// stays empty, hdr.n_wide is refit_from's and hdr.wide_root its root renamed.
FlatTree build_tree(const float* verts, uint32_t nv, const uint32_t* idx, uint32_t nt, int nthreads,
                    const FlatTree* refit_from = nullptr, bool wide_on_device = false);

// The fast-pass hierarchy's node indices grouped by depth, deepest level
// first (children before parents); level_start[k] .. level_start[k + 1] is
// level k of that order.
std::vector<uint32_t> wide_levels(const FlatTree& t, std::vector<uint32_t>& level_start);

// Canonical preorder export (same format as the oracle's): used by tests to
// compare this builder with the oracle's restatement node by node.
struct TreeExport {
    std::vector<float> box;        // 6 per node
    std::vector<uint32_t> right;   // preorder index of the right child, WR_NONE for leaves
    std::vector<uint32_t> leaf_first, leaf_count, tri_list;
};
TreeExport export_tree(const FlatTree& t);

}  // namespace wr
