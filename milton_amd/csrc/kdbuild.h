// Host SAH kd-tree builder of the path-tracing core.
//
// Replaces kdTreeAccel::init / _buildTree / _computeSplitPlaneSAH
// (accel/kdTreeAccel.cpp:403-472, 533-630, 722-877).  The reference sorts the
// primitive extrema of every node (O(N log^2 N), single-threaded) and stores
// 16-byte pointer nodes with per-leaf heap vectors.  This builder bins the
// clipped primitive bounds per node (O(N log N)); the large nodes near the
// root bin and partition on several threads, subtrees go to threads as they
// free up, nodes live in per-thread arenas, and the device layout is written
// by several threads into precomputed ranges.  The result does not depend on
// the thread count:
//   node   = 2 x u32: word0 = (child << 3) | axis for an interior node, i.e.
//            the byte offset of its left child in the node array plus the
//            axis (the right child follows at + 8, and every child pair
//            starts at an even node, so the offsets of the two differ only in
//            bit 3); (count << 2) | 3 for a leaf;
//            word1 = split position (float bits) or the leaf's first record;
//            node 0 is the root, node 1 unused (it aligns the pairs); the
//            top of the tree comes first, breadth-first (every node of the
//            first Params::top_nodes slots is shallower than any node after
//            them), the rest depth-first;
//   record = 12 x f32 per leaf reference, stored in leaf order so a leaf is a
//            contiguous run: (v0, id), (v1, kind), (v2, 0) for a triangle
//            (kind 0 two-sided, 1 one-sided),
//            (centre, id), (r, 0, 0, kind), 0 for a sphere,
//            (centre, id), (ex, kind), (ez, 0) for a quad.
#ifndef MPT_KDBUILD_H_
#define MPT_KDBUILD_H_

#include <stdint.h>

#include <functional>
#include <vector>

namespace mpt_kd {

enum { KIND_TRIANGLE = 0, KIND_TRIANGLE_ONE_SIDED = 1, KIND_SPHERE = 2, KIND_QUAD = 3 };   // triangles: kind < 2

struct Box {
  float lo[3], hi[3];
};

struct Params {
  uint32_t leaf_size = 4;
  uint32_t max_depth = 0;          // 0: 8 + 1.3 log2(N), at most 48
  float cost_traversal = 1.0f;
  float cost_intersect = 1.5f;
  float empty_bias = 1.0f;
  uint32_t num_threads = 0;        // 0: hardware concurrency
  uint32_t bins = 32;
  bool clip = true;                // clip triangles to the cells they are split across
  uint32_t top_nodes = 512;        // node slots laid out breadth-first at the front of the array
};

struct Output {
  std::vector<uint32_t> nodes;     // 2 words per node
  std::vector<float> records;      // 12 floats per reference
  Box box;
  uint64_t n_nodes = 0, n_leaves = 0, n_refs = 0, max_depth = 0, max_leaf = 0;   // n_nodes: root and children
  uint64_t top_slots = 0;          // node slots [0, top_slots) hold the top of the tree, breadth-first
  bool too_large = false;          // node offsets past 2^32 bytes: the tree is unusable
  double build_ms = 0;
};

// prims: bounds per primitive; write_record(prim, out12) fills its record.
// tri_verts (9 floats per triangle, primitives [0, num_tris)) lets the builder
// clip triangles to cells; other primitives are split by their bounds.
void build(const std::vector<Box> &prims, const Params &p,
           const std::function<void(uint32_t, float *)> &write_record, Output &out,
           const float *tri_verts = nullptr, uint32_t num_tris = 0);

}  // namespace mpt_kd

#endif
