// Host SAH kd-tree builder (binned SAH, subtree-parallel).  Layout and the
// reference functions it replaces are described in kdbuild.h.
#include "kdbuild.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <thread>

namespace mpt_kd {
namespace {

struct BNode {
  uint32_t axis = 3;   // 3 == leaf
  float split = 0;
  std::unique_ptr<BNode[]> kids;
  std::vector<uint32_t> refs;
};

inline float area(const float d[3]) { return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]); }

class Builder {
 public:
  Builder(const std::vector<Box> &prims, const Params &p, uint32_t max_depth, int spawn_depth)
      : prims_(prims), p_(p), max_depth_(max_depth), spawn_depth_(spawn_depth) {}

  void build(BNode *node, std::vector<uint32_t> refs, Box cell, uint32_t depth) {
    const size_t n = refs.size();
    if (n <= p_.leaf_size || depth >= max_depth_) {
      make_leaf(node, std::move(refs));
      return;
    }
    uint32_t axis;
    float split;
    if (!find_split(refs, cell, axis, split)) {
      make_leaf(node, std::move(refs));
      return;
    }
    std::vector<uint32_t> left, right;
    left.reserve(n / 2 + 16);
    right.reserve(n / 2 + 16);
    for (uint32_t r : refs) {
      const Box &b = prims_[r];
      // A reference goes to every child whose closed cell its bounds touch.
      if (b.lo[axis] <= split) left.push_back(r);
      if (b.hi[axis] >= split) right.push_back(r);
    }
    if (left.size() == n && right.size() == n) {
      make_leaf(node, std::move(refs));
      return;
    }
    std::vector<uint32_t>().swap(refs);
    node->axis = axis;
    node->split = split;
    node->kids.reset(new BNode[2]);
    Box lcell = cell, rcell = cell;
    lcell.hi[axis] = split;
    rcell.lo[axis] = split;
    if ((int)depth < spawn_depth_ && left.size() + right.size() > 4096) {
      BNode *rk = &node->kids[1];
      std::thread t([this, rk, &right, rcell, depth]() { build(rk, std::move(right), rcell, depth + 1); });
      build(&node->kids[0], std::move(left), lcell, depth + 1);
      t.join();
    } else {
      build(&node->kids[0], std::move(left), lcell, depth + 1);
      build(&node->kids[1], std::move(right), rcell, depth + 1);
    }
  }

 private:
  void make_leaf(BNode *node, std::vector<uint32_t> refs) {
    node->axis = 3;
    node->refs = std::move(refs);
  }

  // Binned SAH over the reference bounds clipped to the cell.
  bool find_split(const std::vector<uint32_t> &refs, const Box &cell, uint32_t &best_axis,
                  float &best_split) const {
    const uint32_t K = std::max<uint32_t>(4, std::min<uint32_t>(p_.bins, 256));
    float d[3];
    for (int a = 0; a < 3; ++a) d[a] = cell.hi[a] - cell.lo[a];
    const float sa = area(d);
    if (!(sa > 0)) return false;
    const double n = (double)refs.size();
    double best = (double)p_.cost_intersect * n;   // leaf cost
    bool found = false;
    uint32_t start[256], end[256];
    for (uint32_t a = 0; a < 3; ++a) {
      const float ext = d[a];
      if (!(ext > 0)) continue;
      std::memset(start, 0, sizeof(uint32_t) * K);
      std::memset(end, 0, sizeof(uint32_t) * K);
      const float scale = (float)K / ext;
      const float lo0 = cell.lo[a];
      for (uint32_t r : refs) {
        const Box &b = prims_[r];
        const float lo = std::max(b.lo[a], cell.lo[a]), hi = std::min(b.hi[a], cell.hi[a]);
        int bl = (int)((lo - lo0) * scale), bh = (int)((hi - lo0) * scale);
        bl = std::min(std::max(bl, 0), (int)K - 1);
        bh = std::min(std::max(bh, 0), (int)K - 1);
        ++start[bl];
        ++end[bh];
      }
      uint32_t nl = 0, nr = (uint32_t)refs.size();
      for (uint32_t i = 1; i < K; ++i) {
        nl += start[i - 1];
        nr -= end[i - 1];
        const float s = lo0 + ext * ((float)i / (float)K);
        if (!(s > cell.lo[a] && s < cell.hi[a])) continue;
        float dl[3] = {d[0], d[1], d[2]}, dr[3] = {d[0], d[1], d[2]};
        dl[a] = s - cell.lo[a];
        dr[a] = cell.hi[a] - s;
        double cost = p_.cost_traversal +
                      p_.cost_intersect * ((double)area(dl) * nl + (double)area(dr) * nr) / sa;
        if (nl == 0 || nr == 0) cost *= p_.empty_bias;
        if (cost < best) {
          best = cost;
          best_axis = a;
          best_split = s;
          found = true;
        }
      }
    }
    return found;
  }

  const std::vector<Box> &prims_;
  const Params &p_;
  uint32_t max_depth_;
  int spawn_depth_;
};

struct Flattener {
  const std::function<void(uint32_t, float *)> &write;
  Output &out;
  void run(const BNode &b, uint64_t idx, uint64_t depth) {
    out.max_depth = std::max(out.max_depth, depth);
    if (b.axis == 3) {
      const uint64_t first = out.records.size() / 12;
      out.nodes[2 * idx] = ((uint32_t)b.refs.size() << 2) | 3u;
      out.nodes[2 * idx + 1] = (uint32_t)first;
      out.records.resize(out.records.size() + 12 * b.refs.size());
      for (size_t i = 0; i < b.refs.size(); ++i) write(b.refs[i], &out.records[12 * (first + i)]);
      ++out.n_leaves;
      out.n_refs += b.refs.size();
      out.max_leaf = std::max<uint64_t>(out.max_leaf, b.refs.size());
      return;
    }
    const uint64_t child = out.nodes.size() / 2;
    out.nodes.resize(out.nodes.size() + 4);
    uint32_t bits;
    std::memcpy(&bits, &b.split, 4);
    out.nodes[2 * idx] = ((uint32_t)child << 2) | b.axis;
    out.nodes[2 * idx + 1] = bits;
    run(b.kids[0], child, depth + 1);
    run(b.kids[1], child + 1, depth + 1);
  }
};

}  // namespace

void build(const std::vector<Box> &prims, const Params &p,
           const std::function<void(uint32_t, float *)> &write_record, Output &out) {
  const auto t0 = std::chrono::steady_clock::now();
  out = Output();
  Box box;
  for (int a = 0; a < 3; ++a) {
    box.lo[a] = INFINITY;
    box.hi[a] = -INFINITY;
  }
  for (const Box &b : prims)
    for (int a = 0; a < 3; ++a) {
      box.lo[a] = std::min(box.lo[a], b.lo[a]);
      box.hi[a] = std::max(box.hi[a], b.hi[a]);
    }
  if (prims.empty())
    for (int a = 0; a < 3; ++a) box.lo[a] = box.hi[a] = 0;
  out.box = box;
  const double n = (double)std::max<size_t>(prims.size(), 1);
  uint32_t max_depth = p.max_depth ? p.max_depth : (uint32_t)(8 + 1.3 * std::log2(n));
  max_depth = std::min<uint32_t>(max_depth, 48);
  unsigned threads = p.num_threads ? p.num_threads : std::max(1u, std::thread::hardware_concurrency());
  threads = std::min(threads, 64u);
  const int spawn = threads > 1 ? (int)std::ceil(std::log2((double)threads)) + 2 : 0;
  std::vector<uint32_t> refs(prims.size());
  for (uint32_t i = 0; i < refs.size(); ++i) refs[i] = i;
  BNode root;
  Builder(prims, p, max_depth, spawn).build(&root, std::move(refs), box, 0);
  out.nodes.assign(2, 0);
  Flattener{write_record, out}.run(root, 0, 0);
  out.n_nodes = out.nodes.size() / 2;
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mpt_kd
