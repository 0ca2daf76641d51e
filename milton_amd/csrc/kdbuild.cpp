// Host SAH kd-tree builder (binned SAH, clipped reference bounds,
// subtree-parallel).  Layout and the reference functions it replaces are
// described in kdbuild.h.
//
// Each reference carries its own bounds, clipped to the cell it lives in.
// When a split cuts through a triangle's bounds, the triangle itself is
// clipped against each child cell (Sutherland-Hodgman in double precision);
// a child the triangle does not reach gets no reference, and the child's
// reference bounds are tight.  This is the "perfect split" refinement of SAH
// kd-trees (Havran; Wald & Havran 2006); the reference builder partitions by
// unclipped primitive bounds (accel/kdTreeAccel.cpp:571-582).
#include "kdbuild.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>

namespace mpt_kd {
namespace {

struct Ref {
  uint32_t prim;
  Box box;
};

struct BNode {
  uint32_t axis = 3;   // 3 == leaf
  float split = 0;
  std::unique_ptr<BNode[]> kids;
  std::vector<uint32_t> prims;
};

inline float area(const float d[3]) { return 2.0f * (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]); }

// Clips triangle v (9 floats) against the cell (widened by a few ulps) and
// returns the bounds of the remaining polygon, rounded outward.  False if
// nothing of the triangle is inside.
bool clip_triangle(const float *v, const Box &cell, Box &out) {
  double poly[2][12][3];
  int n = 3, cur = 0;
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) poly[0][i][k] = v[3 * i + k];
  for (int axis = 0; axis < 3; ++axis) {
    for (int side = 0; side < 2; ++side) {
      const double lim = side == 0 ? (double)cell.lo[axis] - 1e-7 * (1.0 + std::fabs(cell.lo[axis]))
                                   : (double)cell.hi[axis] + 1e-7 * (1.0 + std::fabs(cell.hi[axis]));
      const double sgn = side == 0 ? 1.0 : -1.0;   // inside: sgn * (x - lim) >= 0
      double (*in)[3] = poly[cur];
      double (*o)[3] = poly[cur ^ 1];
      int m = 0;
      for (int i = 0; i < n; ++i) {
        const double *p = in[i], *q = in[(i + 1) % n];
        const double dp = sgn * (p[axis] - lim), dq = sgn * (q[axis] - lim);
        if (dp >= 0) {
          for (int k = 0; k < 3; ++k) o[m][k] = p[k];
          ++m;
        }
        if ((dp >= 0) != (dq >= 0)) {
          const double t = dp / (dp - dq);
          for (int k = 0; k < 3; ++k) o[m][k] = p[k] + t * (q[k] - p[k]);
          o[m][axis] = lim;
          ++m;
        }
      }
      n = m;
      cur ^= 1;
      if (n == 0) return false;
    }
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], poly[cur][i][k]);
      hi[k] = std::max(hi[k], poly[cur][i][k]);
    }
  for (int k = 0; k < 3; ++k) {
    out.lo[k] = std::max(cell.lo[k], std::nextafter((float)lo[k], -INFINITY));
    out.hi[k] = std::min(cell.hi[k], std::nextafter((float)hi[k], INFINITY));
    if (out.lo[k] > out.hi[k]) out.lo[k] = out.hi[k] = (out.lo[k] + out.hi[k]) * 0.5f;
  }
  return true;
}

class Builder {
 public:
  Builder(const Params &p, uint32_t max_depth, int spawn_depth, const float *tri_verts, uint32_t num_tris)
      : p_(p), max_depth_(max_depth), spawn_depth_(spawn_depth), tri_verts_(tri_verts), num_tris_(num_tris) {}

  void build(BNode *node, std::vector<Ref> refs, Box cell, uint32_t depth) {
    const size_t n = refs.size();
    if (n <= p_.leaf_size || depth >= max_depth_) {
      make_leaf(node, refs);
      return;
    }
    uint32_t axis;
    float split;
    if (!find_split(refs, cell, axis, split)) {
      make_leaf(node, refs);
      return;
    }
    Box lcell = cell, rcell = cell;
    lcell.hi[axis] = split;
    rcell.lo[axis] = split;
    std::vector<Ref> left, right;
    left.reserve(n / 2 + 16);
    right.reserve(n / 2 + 16);
    for (const Ref &r : refs) {
      const Box &b = r.box;
      if (b.lo[axis] < split && b.hi[axis] > split) {
        // The split cuts the reference: clip it into each child.
        Ref lr{r.prim, b}, rr{r.prim, b};
        bool in_l = true, in_r = true;
        if (tri_verts_ && r.prim < num_tris_) {
          in_l = clip_triangle(tri_verts_ + 9ull * r.prim, intersect(b, lcell), lr.box);
          in_r = clip_triangle(tri_verts_ + 9ull * r.prim, intersect(b, rcell), rr.box);
          if (!in_l && !in_r) {   // rounding: keep both, unclipped
            in_l = in_r = true;
            lr.box = intersect(b, lcell);
            rr.box = intersect(b, rcell);
          }
        } else {
          lr.box.hi[axis] = split;
          rr.box.lo[axis] = split;
        }
        if (in_l) left.push_back(lr);
        if (in_r) right.push_back(rr);
      } else {
        // A reference goes to every child whose closed cell its bounds touch.
        if (b.lo[axis] <= split) left.push_back(r);
        if (b.hi[axis] >= split) right.push_back(r);
      }
    }
    if (left.size() == n && right.size() == n) {
      make_leaf(node, refs);
      return;
    }
    std::vector<Ref>().swap(refs);
    node->axis = axis;
    node->split = split;
    node->kids.reset(new BNode[2]);
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
  static Box intersect(const Box &a, const Box &b) {
    Box r;
    for (int k = 0; k < 3; ++k) {
      r.lo[k] = std::max(a.lo[k], b.lo[k]);
      r.hi[k] = std::min(a.hi[k], b.hi[k]);
      if (r.lo[k] > r.hi[k]) r.lo[k] = r.hi[k];
    }
    return r;
  }

  void make_leaf(BNode *node, const std::vector<Ref> &refs) {
    node->axis = 3;
    node->prims.resize(refs.size());
    for (size_t i = 0; i < refs.size(); ++i) node->prims[i] = refs[i].prim;
  }

  // Binned SAH over the (cell-clipped) reference bounds.
  bool find_split(const std::vector<Ref> &refs, const Box &cell, uint32_t &best_axis, float &best_split) const {
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
      for (const Ref &r : refs) {
        const float lo = std::max(r.box.lo[a], cell.lo[a]), hi = std::min(r.box.hi[a], cell.hi[a]);
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

  const Params &p_;
  uint32_t max_depth_;
  int spawn_depth_;
  const float *tri_verts_;
  uint32_t num_tris_;
};

struct Flattener {
  const std::function<void(uint32_t, float *)> &write;
  Output &out;
  void run(const BNode &b, uint64_t idx, uint64_t depth) {
    out.max_depth = std::max(out.max_depth, depth);
    if (b.axis == 3) {
      const uint64_t first = out.records.size() / 12;
      out.nodes[2 * idx] = ((uint32_t)b.prims.size() << 2) | 3u;
      out.nodes[2 * idx + 1] = (uint32_t)first;
      out.records.resize(out.records.size() + 12 * b.prims.size());
      for (size_t i = 0; i < b.prims.size(); ++i) write(b.prims[i], &out.records[12 * (first + i)]);
      ++out.n_leaves;
      out.n_refs += b.prims.size();
      out.max_leaf = std::max<uint64_t>(out.max_leaf, b.prims.size());
      return;
    }
    const uint64_t child = out.nodes.size() / 2;   // even: node 1 pads the root
    if (child + 2 > (1ull << 29)) out.too_large = true;
    out.nodes.resize(out.nodes.size() + 4);
    uint32_t bits;
    std::memcpy(&bits, &b.split, 4);
    out.nodes[2 * idx] = ((uint32_t)child << 3) | b.axis;
    out.nodes[2 * idx + 1] = bits;
    run(b.kids[0], child, depth + 1);
    run(b.kids[1], child + 1, depth + 1);
  }
  // The top first, breadth-first while a node's children still fit in
  // `budget` slots, so that the top is a prefix of the array (a kernel can
  // keep it in local memory); every node after that, depth-first.
  void run_top_first(const BNode &root, uint64_t budget) {
    struct Item {
      const BNode *b;
      uint64_t idx, depth;
    };
    std::deque<Item> open{{&root, 0, 0}};
    std::vector<Item> rest;
    while (!open.empty()) {
      const Item it = open.front();
      open.pop_front();
      if (it.b->axis == 3 || out.nodes.size() / 2 + 2 > budget) {
        rest.push_back(it);
        continue;
      }
      out.max_depth = std::max(out.max_depth, it.depth);
      const uint64_t child = out.nodes.size() / 2;
      out.nodes.resize(out.nodes.size() + 4);
      uint32_t bits;
      std::memcpy(&bits, &it.b->split, 4);
      out.nodes[2 * it.idx] = ((uint32_t)child << 3) | it.b->axis;
      out.nodes[2 * it.idx + 1] = bits;
      open.push_back({&it.b->kids[0], child, it.depth + 1});
      open.push_back({&it.b->kids[1], child + 1, it.depth + 1});
    }
    out.top_slots = out.nodes.size() / 2;
    for (const Item &it : rest) run(*it.b, it.idx, it.depth);
  }
};

}  // namespace

void build(const std::vector<Box> &prims, const Params &p,
           const std::function<void(uint32_t, float *)> &write_record, Output &out, const float *tri_verts,
           uint32_t num_tris) {
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
  std::vector<Ref> refs(prims.size());
  for (uint32_t i = 0; i < refs.size(); ++i) refs[i] = Ref{i, prims[i]};
  BNode root;
  Builder(p, max_depth, spawn, p.clip ? tri_verts : nullptr, num_tris).build(&root, std::move(refs), box, 0);
  out.nodes.assign(4, 0);
  Flattener{write_record, out}.run_top_first(root, p.top_nodes);
  out.n_nodes = out.nodes.size() / 2 - 1;
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mpt_kd
