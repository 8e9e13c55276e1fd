// wr_layout.h -- GPU-resident scene layout shared by the host builder and the
// HIP kernels.  Plain PODs, no HIP or STL types, so both compilers see the
// same bytes.
//
// The tree is the reference KDTree (kdtree.cpp:31-42,141-258) flattened for
// traversal on CDNA4:
//   * internal nodes become 64 B "child-pair" records holding BOTH children's
//     boxes, so one 64 B fetch (one half of a 128 B L2 line) feeds two box
//     tests; left child = preorder next, right child index stored;
//   * a child reference with bit 31 set is a leaf: bits 27..30 hold the
//     triangle count (1..15; 0 = long leaf, count in leaf_cnt[first]) and bits
//     0..26 the first triangle record;
//   * triangle records (64 B) are stored in left-to-right leaf order, and
//     leaf keys follow that order, so they order leaves the way
//     kdtree.cpp:331-346 breaks ties (equal t -> right subtree -> larger key).
#pragma once
#include <stdint.h>

#define WR_LEAF_BIT 0x80000000u
#define WR_CNT_SHIFT 27u
#define WR_CNT_MASK 0xFu
#define WR_FIRST_MASK 0x07FFFFFFu
#define WR_NONE 0xFFFFFFFFu
#define WR_MAX_RECORDS (WR_FIRST_MASK + 1u)

// 64 B internal node: boxes of both children + their references.
struct wr_node {
    float lmin[3];
    float lmax[3];
    float rmin[3];
    float rmax[3];
    uint32_t lref, rref;
    uint32_t pad0, pad1;
};
static_assert(sizeof(wr_node) == 64, "node must be 64 bytes");

// 64 B triangle record (half an L2 line, never split across lines): the three
// vertices in reference order, the precomputed reference normal
// normalize((v1-v0) x (v2-v1)) (kdtree.cpp:294-295), and what the queries need
// about the record once its triangle test accepts, so no side array is read on
// the hot path: the key of its reference leaf, the triangle id (global in the
// global tree, the object's own in an object tree) and the object.  The leaf
// key is the index of the leaf's chain record (WR_NONE when the leaf is the
// root); chain records are numbered in left-to-right leaf order, so keys order
// leaves like their first records, the way kdtree.cpp:331-346 breaks ties.
struct wr_tri {
    float v0[3];
    float v1[3];
    float v2[3];
    float n[3];
    uint32_t leaf;
    uint32_t prim;
    uint32_t obj;
    uint32_t pad;
};
static_assert(sizeof(wr_tri) == 64, "triangle record must be 64 bytes");

// 64 B wide node of the fast-pass hierarchy: the reference tree collapsed two
// levels at a time (up to four children: grandchildren, or a child that is a
// leaf).  Child boxes are 8-bit coordinates on the grid origin + q * scale
// (scale a power of two), rounded outward on the host so the decoded box
// contains the reference box.  q[k][c]: k = 0..2 min x,y,z, 3..5 max x,y,z of
// child c.  ref[c]: wide-node index, a leaf reference (as in wr_node), or
// WR_NONE.
struct wr_wnode {
    float origin[3];
    float scale[3];
    uint8_t q[6][4];
    uint32_t ref[4];
};
static_assert(sizeof(wr_wnode) == 64, "wide node must be 64 bytes");
static_assert(__builtin_offsetof(wr_wnode, q) == 24 && __builtin_offsetof(wr_wnode, ref) == 48,
              "wide-node fields are read as four 16 B words");

// 256 B chain record of one non-root reference leaf (two L2 lines), indexed
// by the leaf key: the boxes that the reference tests on the way to the leaf
// at its first up-to-ten ancestors, leaf-side first (box[k] is the box of the
// ancestor-k child on the path, stored in ancestor k + 1), their number, and
// the link (parent << 1 | side) of the next level up, WR_NONE when the chain
// ends at the root.  The first line holds the header and the five levels
// nearest the leaf.  The levels above the chain come from its top record:
// the boxes from `next` up to the root (top_count of them, 6 floats each, at
// box top of the scene's top-box pool), stored once for all the chains that
// end below the same link; the upper levels are few and shared by many
// leaves, so these records stay in L2.
constexpr uint32_t kChainLevels = 10;
struct wr_chain {
    uint32_t count;
    uint32_t next;
    float box[kChainLevels][6];
    uint32_t top;        // first box of the top record in the pool (when next != WR_NONE)
    uint32_t top_count;  // its boxes: the levels from next up to the root
};
static_assert(sizeof(wr_chain) == 256, "chain record must be 256 bytes");
static_assert(__builtin_offsetof(wr_chain, box) + 5 * 24 == 128, "the first line holds five levels");

// Per-tree header.  root_ref may itself be a leaf.
struct wr_tree_hdr {
    float root_min[3];
    float root_max[3];
    uint32_t root_ref;
    uint32_t node_base;   // offset of this tree's nodes in the node array
    uint32_t rec_base;    // offset of this tree's records in the record arrays
    uint32_t n_nodes;
    uint32_t n_recs;
    uint32_t depth;       // max number of internal nodes on a root-to-leaf path
    float scale;          // max |coordinate| over the tree's vertices
    uint32_t wide_root;   // root of the wide hierarchy (wide-node index or leaf reference)
    uint32_t wide_base;   // offset of this tree's wide nodes
    uint32_t n_wide;
};

// Object (material) record; mat follows the reference object classes.
enum { WR_MAT_DIFFUSE = 0, WR_MAT_GLOSSY = 1, WR_MAT_REFRACTIVE = 2, WR_MAT_EMISSIVE = 3 };
struct wr_object_gpu {
    float aabb_min[3];
    float aabb_max[3];
    float color[3];       // albedo (diffuse/glossy) or emitted radiance
    float power;          // Phong exponent
    float ior;
    float pmax;           // max(albedo): the Russian-roulette survival probability
    int32_t mat;
    uint32_t has_normals;
    uint32_t tree;        // index of the object's own tree header (findInsideObject)
    uint32_t prim_base;   // global triangle id of the object's first triangle
    uint32_t pad[2];
};

// One emissive object of the light sampler (Scene::prepareForLightSampling,
// scene.cpp:5-14): its area (AbstractObject::boundArea), originProb = 1 /
// boundArea * pickProb (abstractObj.cpp:41) and its triangles' entries in the
// per-triangle tables (cumulative areas, global-tree records).
struct wr_light {
    float area;
    float oprob;
    uint32_t obj;
    uint32_t tri_base;
    uint32_t n_tris;
    uint32_t pad[3];
};
static_assert(sizeof(wr_light) == 32, "light record must be 32 bytes");
