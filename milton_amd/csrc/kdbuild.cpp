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
#include <mutex>
#include <list>
#include <atomic>
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
  uint32_t axis = 3;          // 3 == leaf
  float split = 0;
  BNode *kids = nullptr;      // interior: the child pair, in an arena
  const uint32_t *prims = nullptr;   // leaf: its references, in an arena
  uint32_t count = 0;         // leaf: number of references
  uint64_t sub_slots = 0;     // node slots the subtree below this node takes (2 per interior node)
  uint64_t sub_refs = 0;      // references in the subtree's leaves
};

// Nodes and leaf references of the subtrees one thread builds, allocated in
// chunks and released together (tens of millions of nodes would otherwise
// be as many allocations to make and free).
class Arena {
 public:
  BNode *pair() {
    if (nodes_used_ + 2 > kNodeChunk) {
      nodes_.emplace_back(new BNode[kNodeChunk]);
      nodes_used_ = 0;
    }
    BNode *p = nodes_.back().get() + nodes_used_;
    nodes_used_ += 2;
    return p;
  }
  uint32_t *prims(size_t n) {
    if (n == 0) return nullptr;   // an empty leaf
    if (prims_used_ + n > prims_cap_) {
      prims_cap_ = std::max<size_t>(kPrimChunk, n);
      prims_.emplace_back(new uint32_t[prims_cap_]);
      prims_used_ = 0;
    }
    uint32_t *p = prims_.back().get() + prims_used_;
    prims_used_ += n;
    return p;
  }

 private:
  static constexpr size_t kNodeChunk = 1 << 16, kPrimChunk = 1 << 18;
  std::vector<std::unique_ptr<BNode[]>> nodes_;
  std::vector<std::unique_ptr<uint32_t[]>> prims_;
  size_t nodes_used_ = kNodeChunk, prims_used_ = 0, prims_cap_ = 0;
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
  Builder(const Params &p, uint32_t max_depth, const float *tri_verts, uint32_t num_tris, unsigned threads)
      : p_(p), max_depth_(max_depth), tri_verts_(tri_verts), num_tris_(num_tris), threads_(threads) {}

  // A new arena for a subtree built on its own thread; it lives as long as
  // the builder.
  Arena &new_arena() {
    std::lock_guard<std::mutex> lk(mu_);
    arenas_.emplace_back();
    return arenas_.back();
  }

  void build(BNode *node, std::vector<Ref> refs, Box cell, uint32_t depth, Arena &ar) {
    const size_t n = refs.size();
    if (n <= p_.leaf_size || depth >= max_depth_) {
      make_leaf(node, refs, ar);
      return;
    }
    uint32_t axis;
    float split;
    if (!find_split(refs, cell, axis, split)) {
      make_leaf(node, refs, ar);
      return;
    }
    Box lcell = cell, rcell = cell;
    lcell.hi[axis] = split;
    rcell.lo[axis] = split;
    std::vector<Ref> left, right;
    const unsigned parts = helpers(n);
    if (parts <= 1) {
      left.reserve(n / 2 + 16);
      right.reserve(n / 2 + 16);
      partition(refs, 0, n, axis, split, lcell, rcell, left, right);
    } else {
      // consecutive runs of the references on helper threads, joined in
      // order: the same children as one pass
      std::vector<std::vector<Ref>> pl(parts), pr(parts);
      parallel_parts(n, parts, [&](unsigned t, size_t b, size_t e) {
        partition(refs, b, e, axis, split, lcell, rcell, pl[t], pr[t]);
      });
      size_t nl = 0, nr = 0;
      for (unsigned t = 0; t < parts; ++t) {
        nl += pl[t].size();
        nr += pr[t].size();
      }
      left.reserve(nl);
      right.reserve(nr);
      for (unsigned t = 0; t < parts; ++t) {
        left.insert(left.end(), pl[t].begin(), pl[t].end());
        right.insert(right.end(), pr[t].begin(), pr[t].end());
      }
    }
    if (left.size() == n && right.size() == n) {
      make_leaf(node, refs, ar);
      return;
    }
    std::vector<Ref>().swap(refs);
    node->axis = axis;
    node->split = split;
    node->kids = ar.pair();
    // the right subtree on a thread of its own while fewer than `threads_`
    // build and both sides are large; a split that only cuts empty space off
    // (one side nearly empty) keeps its thread for the other side
    if (left.size() > 4096 && right.size() > 4096 && live_.fetch_add(1) < (int)threads_) {
      BNode *rk = &node->kids[1];
      Arena &rar = new_arena();
      std::thread t([this, rk, &right, rcell, depth, &rar]() {
        build(rk, std::move(right), rcell, depth + 1, rar);
        live_.fetch_sub(1);
      });
      build(&node->kids[0], std::move(left), lcell, depth + 1, ar);
      t.join();
    } else {
      if (left.size() > 4096 && right.size() > 4096) live_.fetch_sub(1);
      build(&node->kids[0], std::move(left), lcell, depth + 1, ar);
      build(&node->kids[1], std::move(right), rcell, depth + 1, ar);
    }
    node->sub_slots = 2 + node->kids[0].sub_slots + node->kids[1].sub_slots;
    node->sub_refs = node->kids[0].sub_refs + node->kids[1].sub_refs;
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

  void make_leaf(BNode *node, const std::vector<Ref> &refs, Arena &ar) {
    node->axis = 3;
    uint32_t *pr = ar.prims(refs.size());
    for (size_t i = 0; i < refs.size(); ++i) pr[i] = refs[i].prim;
    node->prims = pr;
    node->count = (uint32_t)refs.size();
    node->sub_slots = 0;
    node->sub_refs = refs.size();
  }

  // Threads to split n references over: several only for the large nodes
  // near the root, whose subtrees have not yet spread over threads.
  unsigned helpers(size_t n) const {
    if (n < (1u << 17)) return 1;
    return (unsigned)std::min<size_t>(threads_, n >> 16);
  }

  template <class F>
  static void parallel_parts(size_t n, unsigned parts, F &&f) {
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < parts; ++t) pool.emplace_back([&, t]() { f(t, n * t / parts, n * (t + 1) / parts); });
    f(0, 0, n / parts);
    for (auto &x : pool) x.join();
  }

  // References [b, e) into the children of the split (clipped where it
  // cuts them), appended in order.
  void partition(const std::vector<Ref> &refs, size_t b0, size_t e0, uint32_t axis, float split, const Box &lcell,
                 const Box &rcell, std::vector<Ref> &left, std::vector<Ref> &right) const {
    for (size_t i = b0; i < e0; ++i) {
      const Ref &r = refs[i];
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
  }

  // Bin counts of references [b, e) on every axis: first and last bin of
  // their cell-clipped bounds.
  void count_bins(const std::vector<Ref> &refs, size_t b0, size_t e0, const Box &cell, uint32_t K,
                  uint32_t *start, uint32_t *end) const {
    for (uint32_t a = 0; a < 3; ++a) {
      const float ext = cell.hi[a] - cell.lo[a];
      if (!(ext > 0)) continue;
      const float scale = (float)K / ext;
      const float lo0 = cell.lo[a];
      uint32_t *st = start + a * K, *en = end + a * K;
      for (size_t i = b0; i < e0; ++i) {
        const Ref &r = refs[i];
        const float lo = std::max(r.box.lo[a], cell.lo[a]), hi = std::min(r.box.hi[a], cell.hi[a]);
        int bl = (int)((lo - lo0) * scale), bh = (int)((hi - lo0) * scale);
        bl = std::min(std::max(bl, 0), (int)K - 1);
        bh = std::min(std::max(bh, 0), (int)K - 1);
        ++st[bl];
        ++en[bh];
      }
    }
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
    std::vector<uint32_t> start(3 * K, 0), end(3 * K, 0);
    const unsigned parts = helpers(refs.size());
    if (parts <= 1) {
      count_bins(refs, 0, refs.size(), cell, K, start.data(), end.data());
    } else {
      std::vector<std::vector<uint32_t>> ps(parts, std::vector<uint32_t>(3 * K, 0)), pe = ps;
      parallel_parts(refs.size(), parts, [&](unsigned t, size_t b, size_t e) {
        count_bins(refs, b, e, cell, K, ps[t].data(), pe[t].data());
      });
      for (unsigned t = 0; t < parts; ++t)
        for (uint32_t i = 0; i < 3 * K; ++i) {
          start[i] += ps[t][i];
          end[i] += pe[t][i];
        }
    }
    for (uint32_t a = 0; a < 3; ++a) {
      const float ext = d[a];
      if (!(ext > 0)) continue;
      const float lo0 = cell.lo[a];
      const uint32_t *st = start.data() + a * K, *en = end.data() + a * K;
      uint32_t nl = 0, nr = (uint32_t)refs.size();
      for (uint32_t i = 1; i < K; ++i) {
        nl += st[i - 1];
        nr -= en[i - 1];
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
  const float *tri_verts_;
  uint32_t num_tris_;
  unsigned threads_;
  std::atomic<int> live_{1};   // threads building subtrees
  std::mutex mu_;
  std::list<Arena> arenas_;
};

// Writes the device layout.  Every subtree knows how many node slots and
// references it takes, so the array is sized once and the subtrees below
// the breadth-first top are written on several threads, each into its own
// range: the same bytes as writing them one after another.
struct Flattener {
  const std::function<void(uint32_t, float *)> &write;
  Output &out;
  struct Stats {
    uint64_t max_depth = 0, n_leaves = 0, n_refs = 0, max_leaf = 0;
  };
  // the subtree of b at slot idx, its children from slot `slot` on and its
  // references from record `ref` on, depth-first
  void run(const BNode &b, uint64_t idx, uint64_t depth, uint64_t &slot, uint64_t &ref, Stats &st) {
    st.max_depth = std::max(st.max_depth, depth);
    if (b.axis == 3) {
      out.nodes[2 * idx] = (b.count << 2) | 3u;
      out.nodes[2 * idx + 1] = (uint32_t)ref;
      for (uint32_t i = 0; i < b.count; ++i) write(b.prims[i], &out.records[12 * (ref + i)]);
      ref += b.count;
      ++st.n_leaves;
      st.n_refs += b.count;
      st.max_leaf = std::max<uint64_t>(st.max_leaf, b.count);
      return;
    }
    const uint64_t child = slot;
    slot += 2;
    uint32_t bits;
    std::memcpy(&bits, &b.split, 4);
    out.nodes[2 * idx] = ((uint32_t)child << 3) | b.axis;
    out.nodes[2 * idx + 1] = bits;
    run(b.kids[0], child, depth + 1, slot, ref, st);
    run(b.kids[1], child + 1, depth + 1, slot, ref, st);
  }
  // The top first, breadth-first while a node's children still fit in
  // `budget` slots, so that the top is a prefix of the array (a kernel can
  // keep it in local memory); every node after that, depth-first.
  void run_top_first(const BNode &root, uint64_t budget, unsigned threads) {
    const uint64_t slots = 2 + root.sub_slots;   // the root and the padding slot first
    if (slots > (1ull << 29)) {
      out.too_large = true;
      return;
    }
    out.nodes.assign(2 * slots, 0);
    out.records.assign(12 * root.sub_refs, 0.0f);
    struct Item {
      const BNode *b;
      uint64_t idx, depth;
    };
    std::deque<Item> open{{&root, 0, 0}};
    std::vector<Item> rest;
    uint64_t slot = 2;
    Stats top;
    while (!open.empty()) {
      const Item it = open.front();
      open.pop_front();
      if (it.b->axis == 3 || slot + 2 > budget) {
        rest.push_back(it);
        continue;
      }
      top.max_depth = std::max(top.max_depth, it.depth);
      const uint64_t child = slot;
      slot += 2;
      uint32_t bits;
      std::memcpy(&bits, &it.b->split, 4);
      out.nodes[2 * it.idx] = ((uint32_t)child << 3) | it.b->axis;
      out.nodes[2 * it.idx + 1] = bits;
      open.push_back({&it.b->kids[0], child, it.depth + 1});
      open.push_back({&it.b->kids[1], child + 1, it.depth + 1});
    }
    out.top_slots = slot;
    // each remaining subtree's ranges, in the order a sequential pass would
    // give them
    std::vector<uint64_t> slot0(rest.size()), ref0(rest.size());
    uint64_t ref = 0;
    for (size_t i = 0; i < rest.size(); ++i) {
      slot0[i] = slot;
      ref0[i] = ref;
      slot += rest[i].b->sub_slots;
      ref += rest[i].b->sub_refs;
    }
    std::vector<Stats> st(std::max(1u, threads));
    std::atomic<size_t> next(0);
    auto work = [&](unsigned t) {
      for (size_t i; (i = next.fetch_add(1)) < rest.size();) {
        uint64_t sl = slot0[i], rf = ref0[i];
        run(*rest[i].b, rest[i].idx, rest[i].depth, sl, rf, st[t]);
      }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < st.size(); ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &x : pool) x.join();
    out.max_depth = top.max_depth;
    for (const Stats &x : st) {
      out.max_depth = std::max(out.max_depth, x.max_depth);
      out.n_leaves += x.n_leaves;
      out.n_refs += x.n_refs;
      out.max_leaf = std::max(out.max_leaf, x.max_leaf);
    }
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
  std::vector<Ref> refs(prims.size());
  for (uint32_t i = 0; i < refs.size(); ++i) refs[i] = Ref{i, prims[i]};
  BNode root;
  Builder builder(p, max_depth, p.clip ? tri_verts : nullptr, num_tris, threads);
  builder.build(&root, std::move(refs), box, 0, builder.new_arena());
  Flattener{write_record, out}.run_top_first(root, p.top_nodes, threads);
  out.n_nodes = out.nodes.empty() ? 0 : out.nodes.size() / 2 - 1;
  out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace mpt_kd
