/*
 * oracle_pt.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * A double-precision CPU restatement of Milton's unidirectional path tracer
 * and the functions it calls.  It is the checker for the HIP product path:
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it;
 * the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * mlt/src/milton/).  Deliberate differences from the reference, each shared
 * with the GPU path so both compute the same estimator:
 *   - random numbers come from a counter-based Philox4x32-10 stream keyed by
 *     (seed, job, pixel, sample, depth, stream) instead of the process-global,
 *     unsynchronised mt19937 (stats/Random.cpp:29-56), whose per-path stream is
 *     not reproducible in the reference either;
 *   - each pixel draws its own jitter in its stratum instead of cycling 133
 *     shared pre-drawn patterns (renderers/generators/SuperSampleGenerator.cpp:
 *     23-91); the stratum order (row-major, floor(sqrt(spp))^2 strata) is kept;
 *   - a maximum path depth caps the recursion (the reference has none);
 *   - a BSDF sample whose pdf is zero terminates the path instead of
 *     producing a NaN (ModifiedPhongBSDF.cpp:150-160 can return pdf 0 for a
 *     back-side hit);
 *   - the kd node stores its child index instead of a 32-bit-truncated
 *     pointer (accel/kdTreeAccel.cpp:80-85).
 * Parity status of each part is listed in the repository design notes.
 */
#include "oracle_pt.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace {

const double kEps = 1e-7;                                   // common/arch.h:137
const double kInf = std::numeric_limits<double>::infinity(); // common/arch.h:136
const double kAirIor = 1.0003;                              // materials/IndexOfRefraction.cpp:16

struct V3 {
  double x, y, z;
  V3() : x(0), y(0), z(0) {}
  V3(double a, double b, double c) : x(a), y(b), z(c) {}
  double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  double &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return a * s; }
inline V3 operator/(V3 a, double s) { return V3(a.x / s, a.y / s, a.z / s); }
inline V3 mul(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
  return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 normalized(V3 a) { return a / len(a); }
inline bool is_black(V3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }
inline double max_comp(V3 a) { return std::max(a.x, std::max(a.y, a.z)); }

// core/Ray.h:39-43 -- the ray caches the componentwise reciprocal direction.
struct Ray {
  V3 o, d, inv;
  Ray(V3 o_, V3 d_) : o(o_), d(d_), inv(1.0 / d_.x, 1.0 / d_.y, 1.0 / d_.z) {}
};

struct Aabb {
  V3 lo, hi;
  Aabb() : lo(kInf, kInf, kInf), hi(-kInf, -kInf, -kInf) {}
  void add(V3 p) {
    lo = V3(std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z));
    hi = V3(std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z));
  }
  void add(const Aabb &b) { add(b.lo); add(b.hi); }
  bool is_point() const { return lo.x == hi.x && lo.y == hi.y && lo.z == hi.z; }
  // accel/AABB.inl:112-118
  double area() const {
    V3 e = hi - lo;
    return 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
  }
  // accel/AABB.inl:120-136
  void split_areas(int axis, double pos, double &l, double &r) const {
    V3 e = hi - lo;
    e[axis] = pos - lo[axis];
    l = 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
    e[axis] = hi[axis] - pos;
    r = 2 * (e.x * e.y + e.y * e.z + e.z * e.x);
  }
  // accel/AABB.inl:25-56 (Williams et al. slab test)
  bool intersects(const Ray &ray, double &tmin, double &tmax) const {
    const V3 b[2] = {lo, hi};
    const int xs = ray.inv.x < 0, ys = ray.inv.y < 0, zs = ray.inv.z < 0;
    tmin = (b[xs].x - ray.o.x) * ray.inv.x;
    tmax = (b[1 - xs].x - ray.o.x) * ray.inv.x;
    const double tymin = (b[ys].y - ray.o.y) * ray.inv.y;
    const double tymax = (b[1 - ys].y - ray.o.y) * ray.inv.y;
    if (tmin > tymax || tymin > tmax) return false;
    if (tymin > tmin) tmin = tymin;
    if (tymax < tmax) tmax = tymax;
    const double tzmin = (b[zs].z - ray.o.z) * ray.inv.z;
    const double tzmax = (b[1 - zs].z - ray.o.z) * ray.inv.z;
    if (tmin > tzmax || tzmin > tmax) return false;
    if (tzmin > tmin) tmin = tzmin;
    if (tzmax < tmax) tmax = tzmax;
    return tmin <= tmax;
  }
};

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011).
// ---------------------------------------------------------------------------
inline void philox(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Stream layout shared with the device path: counter = (pixel, sample,
// depth << 8 | stream, job), key = (seed, 0x4D505431).
const uint32_t kKeyHi = 0x4D505431u;
const uint32_t kCameraDepth = 255;
struct Uniforms { double u[4]; uint32_t raw[4]; };
inline Uniforms draw(uint32_t seed, uint32_t job, uint32_t pixel, uint32_t sample,
                     uint32_t depth, uint32_t stream) {
  const uint32_t ctr[4] = {pixel, sample, (depth << 8) | stream, job};
  const uint32_t key[2] = {seed, kKeyHi};
  Uniforms r;
  philox(ctr, key, r.raw);
  for (int i = 0; i < 4; ++i) r.u[i] = (double)(r.raw[i] >> 8) * (1.0 / 16777216.0);
  return r;
}

// ---------------------------------------------------------------------------
// kd-tree: accel/kdTreeAccel.cpp restated.
// ---------------------------------------------------------------------------
struct KdParams {
  unsigned min_prims = 3;     // kdTreeAccel.h:97-104 (BuildParams defaults)
  unsigned max_depth = 24;
  double cost_traversal = 1;
  double empty_bias = 0.9;
};

struct BNode {
  int axis = 3;   // 3 == leaf (KD_LEAF_FLAG, kdTreeAccel.cpp:64)
  float split = 0;
  std::unique_ptr<BNode[]> kids;
  std::vector<uint32_t> prims;
};

struct SplitPlane {
  // kdTreeAccel.cpp:279-297: SPLIT_MAX < SPLIT_PLANE < SPLIT_MIN on ties
  enum { MAX = 0, PLANE = 1, MIN = 2 };
  double pos;
  int type;
  bool operator<(const SplitPlane &p) const {
    if (pos < p.pos) return true;
    if (pos > p.pos) return false;
    return type < p.type;
  }
};

struct KdNode {
  uint32_t axis;   // 3 == leaf
  float split;
  uint32_t child;  // left child index; the right child is child + 1
  uint32_t first, count;
};

class KdTree {
 public:
  Aabb box;
  std::vector<KdNode> nodes;
  std::vector<uint32_t> lists;
  std::vector<Aabb> bounds;   // per primitive
  KdParams params;
  bool naive = false;
  uint64_t n_internal = 0, n_leaves = 0, max_depth_seen = 0;

  // kdTreeAccel.cpp:403-472 (init) and SpatialAccel::init (global AABB).
  void build(std::vector<Aabb> prim_bounds, const KdParams &p, bool use_naive, int nthreads) {
    bounds = std::move(prim_bounds);
    params = p;
    naive = use_naive;
    box = Aabb();
    for (const Aabb &b : bounds) box.add(b);
    nodes.clear();
    lists.clear();
    if (naive) return;
    BNode root;
    std::vector<uint32_t> all(bounds.size());
    for (uint32_t i = 0; i < all.size(); ++i) all[i] = i;
    const int spawn_depth = nthreads > 1 ? std::min(6, (int)std::ceil(std::log2((double)nthreads)) + 1) : 0;
    build_helper(&root, std::move(all), 3, 0, box, spawn_depth);
    flatten(root);
  }

  // kdTreeAccel.cpp:893-1016 (getIntersection).  `isect(prim, ray)` returns
  // the primitive's t (INFINITY on a miss) and may set a sub-index.
  template <class F>
  double closest(const Ray &ray, uint32_t &hit_prim, F &&isect) const {
    if (naive) return closest_naive(ray, hit_prim, isect);
    double tmin, tmax;
    if (!box.intersects(ray, tmin, tmax)) return kInf;
    tmin -= kEps;
    tmax += kEps;
    struct Entry { uint32_t node; double tmax; } stack[64];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
      while (nodes[cur].axis != 3) {
        const KdNode &n = nodes[cur];
        const int a = n.axis;
        const double diff = (double)n.split - ray.o[a];
        const double d = diff * ray.inv[a];
        const bool rev = diff <= 0 ? (ray.d[a] <= 0) : (ray.d[a] < 0);
        cur = n.child + (rev ? 1 : 0);
        if (d > tmax) continue;
        const uint32_t far_node = n.child + (rev ? 0 : 1);
        if (d < tmin) {
          cur = far_node;
        } else {
          stack[sp].node = far_node;
          stack[sp].tmax = tmax;
          ++sp;
          tmax = d + kEps;
        }
      }
      const KdNode &leaf = nodes[cur];
      double tcell = kInf;
      uint32_t best = ORC_MISS;
      for (uint32_t i = leaf.count; i--;) {
        const uint32_t prim = lists[leaf.first + i];
        uint32_t sub = prim;
        const double t = isect(prim, ray, sub);
        if (t > kEps && t < tcell) {
          tcell = t;
          best = sub;
        }
      }
      if (tcell <= tmax) {
        hit_prim = best;
        return tcell;
      }
      if (sp == 0) return kInf;
      tmin = tmax - 2 * kEps;
      --sp;
      cur = stack[sp].node;
      tmax = stack[sp].tmax;
    }
  }

  // kdTreeAccel.cpp:1018-1094 (intersects).  `any(prim, ray, clip)` is the
  // primitive's own any-hit test.
  template <class F>
  bool any(const Ray &ray, double clip, F &&anyf) const {
    double tmin, tmax;
    if (naive) {
      if (!box.intersects(ray, tmin, tmax)) return false;
      for (uint32_t i = (uint32_t)bounds.size(); i--;)
        if (anyf(i, ray, clip)) return true;
      return false;
    }
    if (!box.intersects(ray, tmin, tmax) || clip < tmin) return false;
    if (clip < tmax) tmax = clip;
    tmin -= kEps;
    tmax += kEps;
    struct Entry { uint32_t node; double tmax; } stack[64];
    int sp = 0;
    uint32_t cur = 0;
    for (;;) {
      while (nodes[cur].axis != 3) {
        const KdNode &n = nodes[cur];
        const int a = n.axis;
        const double diff = (double)n.split - ray.o[a];
        const double d = diff * ray.inv[a];
        const bool rev = diff <= 0 ? (ray.d[a] <= 0) : (ray.d[a] < 0);
        cur = n.child + (rev ? 1 : 0);
        if (d > tmax) continue;
        const uint32_t far_node = n.child + (rev ? 0 : 1);
        if (d < tmin) {
          cur = far_node;
        } else {
          stack[sp].node = far_node;
          stack[sp].tmax = tmax;
          ++sp;
          tmax = d + kEps;
        }
      }
      const KdNode &leaf = nodes[cur];
      for (uint32_t i = leaf.count; i--;)
        if (anyf(lists[leaf.first + i], ray, clip)) return true;
      if (sp == 0) return false;
      tmin = tmax;
      --sp;
      cur = stack[sp].node;
      tmax = stack[sp].tmax;
    }
  }

  // Prims whose bounds overlap a leaf cell with positive volume but are not
  // referenced by it (the integrity check disabled at kdTreeAccel.cpp:245-266).
  uint64_t check_invariant() const {
    if (naive) return 0;
    uint64_t bad = 0;
    check_rec(0, box, bad);
    return bad;
  }

 private:
  template <class F>
  double closest_naive(const Ray &ray, uint32_t &hit_prim, F &&isect) const {
    // accel/NaiveSpatialAccel.cpp:16-50: the AABB miss returns 0 (a miss).
    double tmin, tmax;
    if (!box.intersects(ray, tmin, tmax)) return 0;
    double best_t = kInf;
    for (uint32_t i = (uint32_t)bounds.size(); i--;) {
      uint32_t sub = i;
      const double t = isect(i, ray, sub);
      if (t > kEps && t < best_t) {
        best_t = t;
        hit_prim = sub;
      }
    }
    return best_t;
  }

  void check_rec(uint32_t idx, const Aabb &cell, uint64_t &bad) const {
    const KdNode &n = nodes[idx];
    if (n.axis != 3) {
      Aabb l = cell, r = cell;
      l.hi[n.axis] = n.split;
      r.lo[n.axis] = n.split;
      check_rec(n.child, l, bad);
      check_rec(n.child + 1, r, bad);
      return;
    }
    std::vector<uint32_t> ref(lists.begin() + n.first, lists.begin() + n.first + n.count);
    std::sort(ref.begin(), ref.end());
    for (uint32_t p = 0; p < bounds.size(); ++p) {
      const Aabb &b = bounds[p];
      bool overlap = true;
      for (int a = 0; a < 3; ++a)
        overlap = overlap && b.lo[a] < cell.hi[a] && b.hi[a] > cell.lo[a];
      if (overlap && !std::binary_search(ref.begin(), ref.end(), p)) ++bad;
    }
  }

  // kdTreeAccel.cpp:533-630 (_buildTree) + 632-660 (_buildTreeHelper).
  void build_helper(BNode *node, std::vector<uint32_t> prims, int parent_axis,
                    unsigned depth, Aabb cell, int spawn_depth) {
    if (prims.empty()) {
      node->axis = 3;
      return;
    }
    const unsigned n = (unsigned)prims.size();
    if (n < params.min_prims || depth >= params.max_depth) {
      node->axis = 3;
      node->prims = std::move(prims);
      return;
    }
    int axis = 0;
    double pos = 0;
    if (!sah_split(prims, parent_axis, cell, axis, pos)) {
      node->axis = 3;
      node->prims = std::move(prims);
      return;
    }
    std::vector<uint32_t> left, right;
    for (uint32_t p : prims) {
      const Aabb &b = bounds[p];
      if (b.is_point()) continue;
      if (b.hi[axis] > pos) right.push_back(p);
      if (b.lo[axis] <= pos) left.push_back(p);
    }
    // kdTreeAccel.cpp:586-597: with SAH, stop only if neither side shrinks.
    if (left.size() == n && right.size() == n) {
      node->axis = 3;
      node->prims = std::move(prims);
      return;
    }
    prims.clear();
    prims.shrink_to_fit();
    node->axis = axis;
    node->split = (float)pos;
    node->kids.reset(new BNode[2]);
    Aabb lcell = cell, rcell = cell;
    lcell.hi[axis] = pos;
    rcell.lo[axis] = pos;
    if ((int)depth < spawn_depth) {
      BNode *rk = &node->kids[1];
      std::thread t([this, rk, &right, axis, depth, rcell, spawn_depth]() {
        build_helper(rk, std::move(right), axis, depth + 1, rcell, spawn_depth);
      });
      build_helper(&node->kids[0], std::move(left), axis, depth + 1, lcell, spawn_depth);
      t.join();
    } else {
      build_helper(&node->kids[0], std::move(left), axis, depth + 1, lcell, spawn_depth);
      build_helper(&node->kids[1], std::move(right), axis, depth + 1, rcell, spawn_depth);
    }
  }

  // kdTreeAccel.cpp:722-811 (_computeSplitPlaneSAH), 813-853
  // (_computeSplitPlanesSAH), 855-877 (_computeSplitCostSAH).
  bool sah_split(const std::vector<uint32_t> &prims, int parent_axis, const Aabb &cell,
                 int &best_axis, double &best_pos) const {
    const unsigned n_prims = (unsigned)prims.size();
    double best_cost = kInf;
    std::vector<SplitPlane> planes;
    planes.reserve(2 * prims.size());
    for (int axis = 0; axis < 3; ++axis) {
      if (axis == parent_axis || cell.lo[axis] >= cell.hi[axis]) continue;
      planes.clear();
      for (uint32_t i = n_prims; i--;) {
        const Aabb &b = bounds[prims[i]];
        if (b.is_point()) continue;
        const double mn = b.lo[axis], mx = b.hi[axis];
        if (mn == mx) {
          planes.push_back({mn, SplitPlane::PLANE});
        } else {
          planes.push_back({mn, SplitPlane::MIN});
          planes.push_back({mx, SplitPlane::MAX});
        }
      }
      const unsigned n = (unsigned)planes.size();
      if (n == 0) continue;
      std::sort(planes.begin(), planes.end());
      unsigned index = 0, n_left = 0, n_right = n_prims;
      while (index < n - 1) {
        const double cur = planes[index].pos;
        unsigned cnt[3] = {0, 0, 0};
        do {
          ++cnt[planes[index].type];
        } while (index < n - 1 && planes[++index].pos == cur);
        const unsigned n_plane = cnt[SplitPlane::PLANE];
        n_right -= cnt[SplitPlane::MAX] + n_plane;
        if (cur > cell.lo[axis] && cur < cell.hi[axis]) {
          double la, ra;
          cell.split_areas(axis, cur, la, ra);
          const unsigned nl = n_left + n_plane;
          double cost = la * nl + ra * n_right;
          if (nl + n_plane == 0 || n_right == 0) cost *= params.empty_bias;
          if (cost < best_cost) {
            best_cost = cost;
            best_axis = axis;
            best_pos = cur;
          }
        }
        n_left += cnt[SplitPlane::MIN] + n_plane;
      }
    }
    if (best_cost == kInf) return false;
    const double leaf_cost = (n_prims - params.cost_traversal) * cell.area();
    if (best_cost > leaf_cost) return false;
    return true;
  }

  void flatten(BNode &root) {
    nodes.assign(1, KdNode{3, 0, 0, 0, 0});
    std::vector<std::pair<BNode *, uint32_t>> todo{{&root, 0}};
    std::vector<std::pair<BNode *, uint32_t>> next;
    uint64_t depth = 0;
    while (!todo.empty()) {
      next.clear();
      for (auto &e : todo) {
        BNode *b = e.first;
        KdNode &k = nodes[e.second];
        if (b->axis == 3) {
          k.axis = 3;
          k.first = (uint32_t)lists.size();
          k.count = (uint32_t)b->prims.size();
          lists.insert(lists.end(), b->prims.begin(), b->prims.end());
          ++n_leaves;
          max_depth_seen = std::max(max_depth_seen, depth);
        } else {
          const uint32_t child = (uint32_t)nodes.size();
          k.axis = (uint32_t)b->axis;
          k.split = b->split;
          k.child = child;
          nodes.push_back(KdNode{3, 0, 0, 0, 0});
          nodes.push_back(KdNode{3, 0, 0, 0, 0});
          next.push_back({&b->kids[0], child});
          next.push_back({&b->kids[1], child + 1});
          ++n_internal;
        }
      }
      todo.swap(next);
      ++depth;
    }
  }
};

// ---------------------------------------------------------------------------
// Materials (materials/*.cpp) and scene.
// ---------------------------------------------------------------------------
enum { MAT_DIFFUSE = 0, MAT_PHONG = 1, MAT_DIELECTRIC = 2 };
struct Material {
  int type;
  V3 kd, ks;
  double exponent, ior, transparency;
  double ior_gb[2];   // per-channel IOR of green and blue (0 = ior)
  V3 Le;
  bool emitter() const { return Le.x > 0 || Le.y > 0 || Le.z > 0; }
  bool specular() const { return type == MAT_DIELECTRIC; }   // DielectricBSDF.h:64
};
Material load_material(const float *m) {
  Material r;
  r.type = (int)m[0];
  r.kd = V3(m[1], m[2], m[3]);
  r.ks = V3(m[4], m[5], m[6]);
  r.exponent = m[7];
  r.ior = m[8];
  r.transparency = m[9];
  r.Le = V3(m[10], m[11], m[12]);
  r.ior_gb[0] = m[13];
  r.ior_gb[1] = m[14];
  return r;
}

struct Quad {   // shapes/Plane.cpp -- unit plane y=0 mapped by (ex, n, ez) + c
  V3 c, ex, ez, n;
  V3 rx, rz;    // rows of the inverse linear map for x and z
  double area;
};
struct Sphere { V3 c; double r; };

}  // namespace

struct orc_scene {
  std::vector<V3> verts, vnormals;
  std::vector<uint32_t> idx, tri_mat, sph_mat, quad_mat;
  std::vector<uint8_t> one_sided;   // standalone Triangle (back faces culled)
  std::vector<Sphere> spheres;
  std::vector<Quad> quads;
  std::vector<Material> mats;
  // Area lights (Scene::getLights): emitting quads, spheres, then meshes by
  // first triangle.  A sphere samples through the linear part of its
  // transform (frame: images of the object axes); a mesh is its triangles.
  struct Light {
    int kind;           // 0 quad, 1 sphere, 2 mesh
    uint32_t index;     // quad or sphere index
    uint32_t mat;
    double area;        // Shape::getSurfaceArea
    V3 frame[3];
    std::vector<uint32_t> tris;
  };
  std::vector<Light> lights;
  std::vector<uint32_t> tri_light;    // per triangle: its emitting mesh's light index, or ORC_MISS
  std::vector<double> tri_area;       // per triangle: its emitting mesh's area, or 0
  struct Point {
    V3 p, Le;
  };
  std::vector<Point> points;      // point lights: position, P / (4 pi)
  V3 env;
  // Environment image (EnvironmentMap, materials/emitters/EnvironmentMap.cpp):
  // texels in double, row-major from the top row; the filter kernel as
  // KernelFilter::_computeKernel tabulates it (m x m, row-major).
  uint32_t env_w = 0, env_h = 0, env_m = 0;
  std::vector<V3> env_img;
  std::vector<double> env_kernel;
  double env_rep_u = 1, env_rep_v = 1;
  uint32_t nt = 0, ns = 0, nq = 0;

  // Scene::getBackgroundRadiance(d) = background->getLe(-d) (core/Scene.cpp:116-118).
  V3 background(const V3 &d) const {
    if (!env_w) return env;
    // EnvironmentMap::getLe (EnvironmentMap.cpp:44-61) with getSphericalCoords
    // (:63-73) and getUV (:79-85) of w = -d
    const V3 w = -d;
    const double rho = std::sqrt(w.x * w.x + w.z * w.z);
    const double mag = std::sqrt(w.x * w.x + w.y * w.y + w.z * w.z);
    const double theta = rho < std::fabs(w.y) ? std::acos(rho / mag) * (w.y >= 0 ? 1 : -1) : std::asin(w.y / mag);
    const double phi = (w.x == 0 && w.z == 0) ? 0 : std::atan2(w.x, w.z);
    const double u = phi / (-2 * M_PI) + 0.5, v = theta / -M_PI + 0.5;
    const double ss = u * env_rep_u, tt = v * env_rep_v;
    auto clampd = [](double x, double lo, double hi) { return x >= hi ? hi : (x <= lo ? lo : x); };
    const double x = clampd((ss - std::floor(ss)) * env_w, 0, env_w - 1.0);
    const double y = clampd((tt - std::floor(tt)) * env_h, 0, env_h - 1.0);
    // KernelFilter::apply (filters/KernelFilter.cpp:124-160)
    const int xi = (int)std::floor(x), yi = (int)std::floor(y), half = (int)(env_m >> 1);
    V3 rgb;
    double sum = 0;
    size_t k = 0;
    for (int i = -half; i <= half; ++i) {
      const int row = std::min(std::max(i + yi, 0), (int)env_h - 1);
      for (int j = -half; j <= half; ++j) {
        const int col = std::min(std::max(j + xi, 0), (int)env_w - 1);
        const V3 &t = env_img[(size_t)row * env_w + col];
        const double wgt = env_kernel[k++];
        sum += wgt;
        rgb = rgb + t * wgt;
      }
    }
    return rgb / sum;
  }
  KdTree mesh_tree;   // over triangles (Mesh::init, shapes/Mesh.cpp:211-251)
  KdTree top_tree;    // over shapes: [mesh?, spheres..., quads...] (ShapeSet.cpp:36-80)
  std::vector<uint32_t> top_kind;   // 0 mesh, 1 sphere, 2 quad
  std::vector<uint32_t> top_index;

  // --- primitive tests ---------------------------------------------------
  // shapes/MeshTriangle.cpp:92-169, non-culling branch (:142-167); a
  // standalone Triangle takes the culling branch of shapes/Triangle.cpp:245-270.
  double tri_t(uint32_t i, const Ray &ray) const {
    const V3 &v0 = verts[idx[3 * i]], &v1 = verts[idx[3 * i + 1]], &v2 = verts[idx[3 * i + 2]];
    const V3 e1 = v1 - v0, e2 = v2 - v0;
    const V3 p = cross(ray.d, e2);
    const double det = dot(e1, p);
    if (one_sided[i]) {
      if (det < kEps) return kInf;
      const V3 tv = ray.o - v0;
      const double u = dot(tv, p);
      if (u < 0.0 || u > det) return kInf;
      const V3 q = cross(tv, e1);
      const double v = dot(ray.d, q);
      if (v < 0.0 || u + v > det) return kInf;
      return dot(e2, q) * (1.0 / det);
    }
    // EQ(det, 0) (common/utilities.h:240) reduces to |det| <= EPSILON.
    if (std::fabs(det) < kEps || std::fabs(det - 0.0 / det) <= kEps) return kInf;
    const double inv = 1.0 / det;
    const V3 tv = ray.o - v0;
    const double u = dot(tv, p) * inv;
    if (u < 0.0 || u > 1.0) return kInf;
    const V3 q = cross(tv, e1);
    const double v = dot(ray.d, q) * inv;
    if (v < 0.0 || u + v > 1.0) return kInf;
    return dot(e2, q) * inv;
  }
  // shapes/Sphere.cpp:16-45, in the object space of a unit-diameter sphere.
  double sphere_t(uint32_t i, const Ray &ray) const {
    const Sphere &s = spheres[i];
    const double k = 1.0 / (2 * s.r);
    const V3 P = (ray.o - s.c) * k, D = ray.d * k;
    const double A = 2 * dot(D, D);
    const double B = -2 * dot(P, D);
    const double C = dot(P, P) - 0.25;
    double disc = B * B - 2 * A * C;
    if (disc < 0) return kInf;
    disc = std::sqrt(disc);
    const double F = 1 / A;
    double t = (B - disc) * F;
    if (t > kEps) return t;
    t = (B + disc) * F;
    return t > kEps ? t : kInf;
  }
  // shapes/Sphere.cpp:47-73
  bool sphere_any(uint32_t i, const Ray &ray, double tmax) const {
    const Sphere &s = spheres[i];
    const double k = 1.0 / (2 * s.r);
    const V3 P = (ray.o - s.c) * k, D = ray.d * k;
    const double A = 2 * dot(D, D);
    const double B = -2 * dot(P, D);
    const double C = dot(P, P) - 0.25;
    double disc = B * B - 2 * A * C;
    if (disc < 0) return false;
    disc = std::sqrt(disc);
    const double F = 1 / A;
    double t = (B + disc) * F;
    if (t < tmax && t > kEps) return true;
    t = (B - disc) * F;
    return t < tmax && t > kEps;
  }
  // shapes/Plane.cpp:25-49 (object-space y = 0 plane, |x|,|z| <= 0.5).
  bool quad_obj(uint32_t i, const Ray &ray, double &t) const {
    const Quad &q = quads[i];
    const V3 rel = ray.o - q.c;
    const double Py = dot(q.n, rel), Dy = dot(q.n, ray.d);
    const double Px = dot(q.rx, rel), Dx = dot(q.rx, ray.d);
    const double Pz = dot(q.rz, rel), Dz = dot(q.rz, ray.d);
    t = -Py / Dy;
    return std::fabs(Px + t * Dx) <= 0.5 && std::fabs(Pz + t * Dz) <= 0.5;
  }
  double quad_t(uint32_t i, const Ray &ray) const {
    double t;
    const bool in = quad_obj(i, ray, t);
    return (t > kEps && in) ? t : kInf;
  }
  bool quad_any(uint32_t i, const Ray &ray, double tmax) const {
    double t;
    const bool in = quad_obj(i, ray, t);
    return t < tmax && t > kEps && in;
  }

  // Scene::getIntersection (core/Scene.cpp:104-108) -> ShapeSet -> top kd ->
  // Mesh::getIntersection (shapes/Mesh.cpp:379-395) -> mesh kd.
  double closest(const Ray &ray, uint32_t &gid) const {
    gid = ORC_MISS;
    uint32_t hit = ORC_MISS;
    const double t = top_tree.closest(ray, hit, [&](uint32_t shape, const Ray &r, uint32_t &sub) -> double {
      const uint32_t kind = top_kind[shape], k = top_index[shape];
      if (kind == 0) {
        uint32_t tri = ORC_MISS;
        const double tt = mesh_tree.closest(r, tri, [&](uint32_t p, const Ray &rr, uint32_t &s2) -> double {
          s2 = p;
          return tri_t(p, rr);
        });
        if (tt > kEps && tt < kInf) sub = tri;
        return tt;
      }
      if (kind == 1) { sub = nt + k; return sphere_t(k, r); }
      sub = nt + ns + k;
      return quad_t(k, r);
    });
    if (t > kEps && t < kInf) gid = hit;
    return t;
  }

  // Scene::intersects (core/Scene.cpp:110-114) passes tMax - EPSILON.
  bool occluded(const Ray &ray, double tmax) const {
    const double clip = tmax - kEps;
    return top_tree.any(ray, clip, [&](uint32_t shape, const Ray &r, double c) -> bool {
      const uint32_t kind = top_kind[shape], k = top_index[shape];
      if (kind == 0) {
        // Mesh::intersects (shapes/Mesh.cpp:397-406) -> mesh kd ->
        // Intersectable::intersects (shapes/Intersectable.cpp:20-25).
        return mesh_tree.any(r, c, [&](uint32_t p, const Ray &rr, double cc) -> bool {
          const double t = tri_t(p, rr);
          return t < cc - kEps && t > kEps;
        });
      }
      if (kind == 1) return sphere_any(k, r, c);
      return quad_any(k, r, c);
    });
  }

  // --- surface points ----------------------------------------------------
  // Shape::initSurfacePoint (shapes/Shape.cpp:17-27) with the shape's
  // _getGeometricNormal; normalS = normalG (Material.cpp:183-186, no bump map).
  V3 normal_at(uint32_t gid, V3 p) const {
    if (gid < nt) {
      // Mesh::_getGeometricNormal (shapes/Mesh.cpp:419-433) with
      // MeshTriangle::getBaryocentricCoords (shapes/MeshTriangle.cpp:344-361).
      const uint32_t a = idx[3 * gid], b = idx[3 * gid + 1], c = idx[3 * gid + 2];
      const V3 &v1 = verts[a], &v2 = verts[b], &v3 = verts[c];
      const double A = 1.0 / len(cross(v2 - v1, v3 - v1));
      const double pA = len(cross(v2 - p, v3 - p)) * A;
      const double pB = len(cross(p - v1, v3 - v1)) * A;
      const double pC = 1 - pA - pB;
      return normalized(vnormals[a] * pA + vnormals[b] * pB + vnormals[c] * pC);
    }
    if (gid < nt + ns) return normalized(p - spheres[gid - nt].c);   // Sphere.cpp:86-92
    return quads[gid - nt - ns].n;                                    // Plane.cpp:59-63
  }
  const Material &material_of(uint32_t gid) const {
    if (gid < nt) return mats[tri_mat[gid]];
    if (gid < nt + ns) return mats[sph_mat[gid - nt]];
    return mats[quad_mat[gid - nt - ns]];
  }
  // Shape::getSurfaceArea: Plane.cpp:95-106; Sphere.cpp:138-141 (unit sphere).
  double shape_area(uint32_t gid) const {
    if (gid >= nt + ns) return quads[gid - nt - ns].area;
    if (gid >= nt) return M_PI;
    return tri_area[gid];
  }

  // The light list from the scene's emitters; tri_shape (optional) names
  // the shape of each triangle (else triangles sharing an emitting material
  // form one shape), sphere_frames (optional) the spheres' transforms (else
  // 2r times the identity).
  // EmitterSampler::init (stats/samplers/EmitterSampler.cpp:19-48): every
  // light's power averaged over the channels, over their total; area lights
  // then point lights.  An area light's power is Le pi A (Le = P / (pi A),
  // materials/Emitter.cpp:44, emitters/OrientedEmitter.cpp:17-23), a point
  // light's P = 4 pi Le (emitters/OmniEmitter.cpp:15-17).
  std::vector<double> light_share;
  double emit_area = 0;   // EmitterSampler::m_area: the lights' summed surface area
  void update_shares() {
    std::vector<double> power;
    emit_area = 0;
    for (const Light &l : lights) {
      const V3 &le = mats[l.mat].Le;
      power.push_back((le.x + le.y + le.z) / 3.0 * M_PI * l.area);
      emit_area += l.area;
    }
    for (const Point &pl : points) power.push_back((pl.Le.x + pl.Le.y + pl.Le.z) / 3.0 * 4.0 * M_PI);
    double total = 0;
    for (double w : power) total += w;
    light_share.clear();
    for (double w : power) light_share.push_back(total > 0 ? w / total : 0.0);
  }
  // Random::sampleNormalizedCDF (stats/Random.h:125-139): the first light
  // whose running share reaches x, else the last
  uint32_t select_light(double x) const {
    const uint32_t nl = (uint32_t)light_share.size();
    double sum = 0;
    for (uint32_t i = 0; i + 1 < nl; ++i) {
      sum += light_share[i];
      if (x <= sum) return i;
    }
    return nl - 1;
  }

  // Shape::getPoint (shapes/Shape.cpp:52-64) of area light L at (u, v), with
  // r1, r2 the further uniforms of a mesh triangle: position and normal.
  void light_point(const Light &L, double u, double v, double r1, double r2, V3 &lp, V3 &ln) const;

  void build_lights(const uint32_t *tri_shape, const float *sphere_frames) {
    lights.clear();
    tri_light.assign(nt, ORC_MISS);
    tri_area.assign(nt, 0.0);
    for (uint32_t q = 0; q < nq; ++q)
      if (mats[quad_mat[q]].emitter()) lights.push_back(Light{0, q, quad_mat[q], quads[q].area, {}, {}});
    for (uint32_t i = 0; i < ns; ++i) {
      if (!mats[sph_mat[i]].emitter()) continue;
      // Sphere::_getSurfaceArea ignores the transform (shapes/Sphere.cpp:138-141)
      Light l{1, i, sph_mat[i], M_PI, {}, {}};
      for (int k = 0; k < 3; ++k)
        l.frame[k] = sphere_frames ? V3(sphere_frames[9 * i + 3 * k], sphere_frames[9 * i + 3 * k + 1],
                                        sphere_frames[9 * i + 3 * k + 2])
                                   : V3(k == 0, k == 1, k == 2) * (2.0 * spheres[i].r);
      lights.push_back(l);
    }
    std::vector<uint64_t> keys;
    std::vector<size_t> slot;
    for (uint32_t t = 0; t < nt; ++t) {
      if (!mats[tri_mat[t]].emitter()) continue;
      const uint64_t key = tri_shape ? ((uint64_t)1 << 32) | tri_shape[t] : tri_mat[t];
      size_t g = keys.size();
      for (size_t k = 0; k < keys.size(); ++k)
        if (keys[k] == key) g = k;
      if (g == keys.size()) {
        keys.push_back(key);
        slot.push_back(lights.size());
        lights.push_back(Light{2, 0, tri_mat[t], 0.0, {}, {}});
      }
      lights[slot[g]].tris.push_back(t);
    }
    for (size_t g = 0; g < slot.size(); ++g) {
      Light &l = lights[slot[g]];
      // Mesh::_getSurfaceArea (shapes/Mesh.cpp:452-466), in world space
      for (uint32_t t : l.tris) {
        const V3 &a = verts[idx[3 * t]], &b = verts[idx[3 * t + 1]], &c = verts[idx[3 * t + 2]];
        l.area += 0.5 * len(cross(b - a, c - a));
      }
      for (uint32_t t : l.tris) {
        tri_light[t] = (uint32_t)slot[g];
        tri_area[t] = l.area;
      }
    }
    update_shares();
  }
};

void orc_scene::light_point(const Light &L, double u, double v, double r1, double r2, V3 &lp, V3 &ln) const {
  if (L.kind == 0) {
    const Quad &lq = quads[L.index];
    lp = lq.c + lq.ex * (u - 0.5) + lq.ez * (v - 0.5);   // Plane.cpp:91-93
    ln = lq.n;
  } else if (L.kind == 1) {
    // Sphere::getPosition (shapes/Sphere.cpp:117-124) through the transform;
    // the normal is the point's object-space direction carried to world
    // space (Sphere.cpp:86-92)
    const double x = std::min(0.5, std::max(-0.5, std::tan((1.0 - u) / (2 * M_PI))));
    const double y = std::min(0.5, std::max(-0.5, std::sin((0.5 - v) / M_PI) * 0.5));
    const double z = std::sqrt(1 - x * x - y * y);
    const V3 q = L.frame[0] * x + L.frame[1] * y + L.frame[2] * z;
    lp = spheres[L.index].c + q;
    ln = normalized(q);
  } else {
    // Mesh::getPosition (shapes/Mesh.cpp:442-450, MESH_GET_INDEX :26) with
    // MeshTriangle::getPosition (MeshTriangle.cpp:299-333)
    const uint32_t n_tri = (uint32_t)L.tris.size();
    const uint32_t t = L.tris[std::min((uint32_t)std::floor(u * n_tri), n_tri - 1)];
    const double sum = r1 + v + r2;
    lp = verts[idx[3 * t]] * (r1 / sum) + verts[idx[3 * t + 1]] * (v / sum) + verts[idx[3 * t + 2]] * (r2 / sum);
    ln = normal_at(t, lp);   // Mesh::_getGeometricNormal at that point
  }
}

namespace {

// ---------------------------------------------------------------------------
// BSDFs.
// ---------------------------------------------------------------------------
// Vector::getOrthonormalBasis (common/math/Vector.inl:527-538).
inline void onb(V3 n, V3 &U, V3 &V) {
  n = normalized(n);
  V3 up(0, 1, 0);
  if (std::fabs(dot(up, n)) > 0.8) up = V3(1, 0, 0);
  U = normalized(cross(up, n));
  V = normalized(cross(n, U));
}
// Vector::fromSpherical (common/math/Vector.h:113-117).
inline V3 from_spherical(double theta, double phi) {
  return normalized(V3(std::sin(theta) * std::cos(phi), std::cos(theta), std::sin(theta) * std::sin(phi)));
}
// Vector3::cosRandom (common/math/Vector.cpp:21-29) + convertHemisphere
// (common/math/sisd/Vector3.inl:18-22).
inline V3 cos_random(V3 n, double u1, double u2) {
  const double theta = std::acos(std::sqrt(u1));
  const double phi = 2.0 * M_PI * u2;
  const V3 w = from_spherical(theta, phi);
  V3 U, V;
  onb(n, U, V);
  return U * w.x + normalized(n) * w.y + V * w.z;
}
// Vector::reflectVector (common/math/Vector.inl:460-462).
inline V3 reflect(V3 w, V3 n) { return -2 * n * dot(w, n) + w; }
// Vector::refractVector (common/math/Vector.inl:468-494).
inline V3 refract(V3 v, V3 normal, double in, double out) {
  const V3 w = -v;
  V3 norm;
  double n;
  if (dot(v, normal) < 0) { norm = normal; n = in / out; }
  else { norm = -normal; n = out / in; }
  const double d = dot(w, norm);
  const double det = 1 - n * n * (1 - d * d);
  if (det < 0) return V3();
  return -n * w + norm * (n * d - std::sqrt(det));
}
// Fresnel::getFr (materials/bsdfs/Fresnel.h:45-90).
inline double fresnel_cos(double ci, double n1, double n2) {
  const double snell = (n1 / n2) * std::sin(std::acos(ci));
  if (std::fabs(snell) >= 1) return 1;
  const double ct = std::cos(std::asin(snell));
  const double Rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
  const double Rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
  return (Rs * Rs + Rp * Rp) / 2.0;
}
inline double fresnel(V3 wi, V3 N, double n1, double n2) {
  const double c = dot(N, wi);
  return c < 0 ? fresnel_cos(-c, n1, n2) : fresnel_cos(c, n2, n1);
}

struct BsdfSample {
  V3 wo;        // zero if absorbed
  V3 fs;        // f / pdf
  double pdf = 0;
};

inline double avg(V3 a) { return (a.x + a.y + a.z) / 3.0; }

// f(wi, wo) for each BSDF.
V3 bsdf_eval(const Material &m, V3 n, V3 wi, V3 wo) {
  if (m.type == MAT_DIFFUSE) {
    // DiffuseBSDF::evaluate (materials/bsdfs/DiffuseBSDF.cpp:40-48): one-sided.
    if (dot(n, wi) >= 0 || dot(n, wo) <= 0) return V3();
    return m.kd / M_PI;
  }
  if (m.type == MAT_PHONG) {
    // ModifiedPhongBSDF::evaluate (ModifiedPhongBSDF.cpp:164-181).
    if (dot(n, wo) <= 0) return V3();
    const V3 r = reflect(wi, n);
    const double a = std::min(1.0, std::max(0.0, dot(r, wo)));
    const double kda = avg(m.kd), ksa = avg(m.ks);
    const double an = std::pow(a, m.exponent);
    const double fs_s = ksa > 0 ? (m.exponent + 2) / (2 * M_PI * ksa) * an : 0;
    const double fs_d = kda > 0 ? 1.0 / (M_PI * kda) : 0;
    return m.kd * fs_d + m.ks * fs_s;
  }
  return V3();   // a dielectric is a delta distribution: zero off its events
}

// pdf of a direction (projected solid angle measure) for diffuse and Phong.
double bsdf_pdf_diffuse(V3 n, V3 wi, V3 wo) {
  // DiffuseBSDF::getPd (DiffuseBSDF.cpp:30-38)
  if (dot(n, wi) > 0) return 0;
  return dot(n, wo) > 0 ? 1.0 / M_PI : 0;
}

BsdfSample bsdf_sample(const Material &m, V3 n, V3 wi, const double u[4], uint32_t ior_index) {
  BsdfSample s;
  if (m.type == MAT_DIFFUSE) {
    // DiffuseBSDF::sample (DiffuseBSDF.cpp:21-28)
    if (dot(n, wi) > 0) return s;
    s.wo = cos_random(n, u[0], u[1]);
    s.pdf = bsdf_pdf_diffuse(n, wi, s.wo);
    if (s.pdf > 0) s.fs = bsdf_eval(m, n, wi, s.wo) / s.pdf;
    return s;
  }
  if (m.type == MAT_PHONG) {
    // ModifiedPhongBSDF::sample / getPd (ModifiedPhongBSDF.cpp:84-162)
    const double kda = avg(m.kd), ksa = avg(m.ks);
    const double p0 = u[3] - kEps;
    int event;
    if (p0 <= kda) {
      event = 1;
      s.wo = cos_random(n, u[0], u[1]);
    } else if (p0 <= kda + ksa) {
      event = 2;
      const double alpha = std::acos(std::pow(u[0], 1.0 / (m.exponent + 1)));
      const double phi = 2.0 * M_PI * u[1];
      const V3 R = reflect(wi, n);
      V3 r = R, uu, vv;
      if (std::fabs(dot(n, r)) < .95) {
        uu = normalized(cross(r, n));
        vv = cross(uu, r);
      } else {
        onb(r, vv, uu);
        r = normalized(r);
      }
      const V3 w = from_spherical(alpha, phi);
      const V3 wo = uu * w.x + r * w.y + vv * w.z;
      s.wo = dot(n, wo) > 0 ? wo : R;
    } else {
      return s;   // absorbed
    }
    const double cosA = dot(n, s.wo);
    if (cosA <= 0) { s.wo = V3(); return s; }
    double pdf;
    if (event == 2) {
      const V3 r = reflect(wi, n);
      const double ca = std::fabs(dot(r, s.wo));
      pdf = std::min(1.0, (m.exponent + 1) / (2 * M_PI) * std::pow(ca, m.exponent));
    } else {
      pdf = std::fabs(dot(n, s.wo)) / M_PI;
    }
    s.pdf = pdf / cosA;
    if (!(s.pdf > 0)) { s.wo = V3(); return s; }
    s.fs = bsdf_eval(m, n, wi, s.wo) / s.pdf;
    return s;
  }
  // DielectricBSDF::sample / getPd / evaluate (DielectricBSDF.cpp:55-158).
  const V3 N = dot(n, wi) > 0 ? -n : n;
  // m_ior[index] with index = the path's channel (DielectricBSDF.cpp:65-87)
  const double ior = ior_index > 0 && m.ior_gb[ior_index - 1] > 0 ? m.ior_gb[ior_index - 1] : m.ior;
  const V3 wt = refract(wi, n, kAirIor, ior);
  bool refracted = false;
  if (!is_black(wt)) {
    const double Fr = fresnel(wi, n, kAirIor, ior);
    const double Fs = m.transparency * (1 - Fr);
    if (u[3] < Fs) refracted = true;
  }
  s.pdf = 1;
  if (refracted) {
    s.wo = wt;
    V3 sel;
    sel[ior_index] = 1;   // only the path's IOR channel transmits (:139-150)
    s.fs = mul(m.ks, sel);
  } else {
    s.wo = reflect(wi, N);
    s.fs = m.ks;
  }
  return s;
}

// Camera ray: ThinLensCamera::getWorldRay with aperture 0 (PinholeCamera.cpp:
// 22; ThinLensCamera.cpp:131-135, 181-239, 241-256).
struct Camera {
  V3 eye, u, v, n;
  double tan_w, tan_h;
  double aperture, focal_length, focal_distance;   // ThinLensCamera (m)
  // sc: resolves autofocus (ThinLensCamera::init, cameras/ThinLensCamera.cpp:
  // 97-118) when given
  explicit Camera(const double *c, const orc_scene *sc = nullptr) {
    eye = V3(c[0], c[1], c[2]);
    const V3 look(c[3], c[4], c[5]), up(c[6], c[7], c[8]);
    n = -look;
    u = cross(up, n);
    n = normalized(n);
    u = normalized(u);
    v = cross(n, u);
    const double ha = c[9] * (M_PI / 180);
    tan_h = std::tan(ha / 2);
    tan_w = c[10] * tan_h;
    aperture = c[11] > 0 ? c[11] : 0;
    focal_length = c[12];
    focal_distance = c[13];
    if (aperture > 0 && c[16] != 0 && sc) {
      // filmToWorld of (x - 1/2, 1/2 - y, -1): half the extent of getWorldRay
      const double a = std::min(std::max(c[14], 0.0), 1.0) - 0.5, b = 0.5 - std::min(std::max(c[15], 0.0), 1.0);
      uint32_t gid;
      const double t = sc->closest(Ray(eye, normalized(u * (a * tan_w) + v * (b * tan_h) - n)), gid);
      if (t > kEps && t < kInf) focal_distance = t;
    }
  }
  V3 dir(double x, double y) const {
    const double nx = 2 * x - 1, ny = 1 - 2 * y;
    return normalized(u * (nx * tan_w) + v * (ny * tan_h) - n);
  }
  // ThinLensCamera::getWorldRay (cameras/ThinLensCamera.cpp:131-172): with
  // an aperture, the origin moves by r = cof sin(e1) at angle e2 within the
  // lens plane, e1 = pi/2 l0, e2 = 2 pi l1, cof = aperture (l - D) / l with
  // l = f t / (t - f) and t the distance to the focal plane along the ray,
  // and the ray is re-aimed at the point where the pinhole ray meets the
  // focal plane.
  Ray primary(double x, double y, double l0, double l1) const {
    Ray ray(eye, dir(x, y));
    if (!(aperture > 0)) return ray;
    const double t = -focal_distance / dot(ray.d, n);
    if (t == focal_length) return ray;
    const V3 focus = ray.o + ray.d * t;
    const double l_image = focal_length * t / (t - focal_length);
    const double cof = aperture * (l_image - focal_distance) / l_image;
    const double e1 = l0 * (M_PI / 2), e2 = l1 * (2 * M_PI);
    const double r = cof * std::sin(e1);
    const V3 o = ray.o + u * (r * std::cos(e2)) + v * (r * std::sin(e2));
    return Ray(o, normalized(focus - o));
  }
};

// Sample position of (pixel, sample): the stratum of SuperSG +
// JitteredSG (SuperSampleGenerator.cpp:48-91, JitteredSampleGenerator.cpp:
// 30-31, core/Viewport.h:29-36).
inline void film_position(const orc_job &job, uint32_t pixel, uint32_t sample, double u0, double u1,
                          double &x, double &y) {
  const uint32_t side = (uint32_t)std::sqrt((double)job.spp);
  const uint32_t k = sample % (side * side);
  const uint32_t sx = k % side, sy = k / side;
  const uint32_t px = pixel % job.width, py = pixel / job.width;
  x = (px + (sx + u0) / side) / job.width;
  y = (py + (sy + u1) / side) / job.height;
}

struct PathCounters { uint64_t camera = 0, extension = 0, shadow = 0; };

// PathTracer::_evaluate (renderers/renderers/PathTracer.cpp:28-116) as an
// iteration, with DirectIllumination::evaluate (renderers/DirectIllumination.
// cpp:46-136) at each non-specular vertex.
V3 trace_path(const orc_scene &sc, const orc_job &job, const Camera &cam, uint32_t pixel,
              uint32_t sample, PathCounters &ctr) {
  const Uniforms cu = draw(job.seed, job.job_id, pixel, sample, kCameraDepth, 0);
  double fx, fy;
  film_position(job, pixel, sample, cu.u[0], cu.u[1], fx, fy);
  // the lens draws its own stream
  const Uniforms lu = cam.aperture > 0 ? draw(job.seed, job.job_id, pixel, sample, kCameraDepth, 1) : cu;
  Ray ray = cam.primary(fx, fy, lu.u[0], lu.u[1]);
  // PathTracer.cpp:32: one IOR channel per camera ray.
  const uint32_t ior_index = (uint32_t)(((uint64_t)(cu.raw[2] >> 8) * 3) >> 24);
  ctr.camera++;
  V3 L, beta(1, 1, 1);
  bool any_specular = false, d0_continued = false, emit_after_specular = false;
  for (uint32_t depth = 0; depth < job.max_depth; ++depth) {
    uint32_t gid;
    ctr.extension++;
    const double t = sc.closest(ray, gid);
    if (!(t > kEps && t < kInf)) {
      L = L + mul(beta, sc.background(ray.d));   // Scene::getBackgroundRadiance (Scene.cpp:116-118)
      break;
    }
    const V3 p = ray.o + ray.d * t;
    const V3 n = sc.normal_at(gid, p);
    const Material &m = sc.material_of(gid);
    if (m.specular()) any_specular = true;
    // PathTracer.cpp:50-69 -- emission at depth 0, or after a specular
    // bounce that left the surface on the normal's side; at every vertex
    // when efficientDirect is false (:51).
    if (m.emitter() && dot(n, -ray.d) > 0) {
      if (depth == 0 || job.direct_off) L = L + mul(beta, m.Le);
      else if (emit_after_specular) L = L + mul(beta, m.Le) * (sc.shape_area(gid) / (t * t));
    }
    const Uniforms bu = draw(job.seed, job.job_id, pixel, sample, depth, 0);
    // PathTracer.cpp:72-73: next-event estimation only with efficientDirect
    if (!m.specular() && !job.direct_off) {
      // DirectIllumination.cpp:90-135: per light, ns samples jittered over
      // floor(sqrt(ns))^2 strata of the light, weighted area / ns.
      const uint32_t ns = std::max(1u, job.direct_samples);
      uint32_t side = 1;
      while ((side + 1) * (side + 1) <= ns) ++side;
      // the vertex lies on area light li (DirectIllumination.cpp:93 skips it)
      auto on_light = [&](uint32_t li) {
        const orc_scene::Light &L = sc.lights[li];
        return (L.kind == 0 && gid == sc.nt + sc.ns + L.index) || (L.kind == 1 && gid == sc.nt + L.index) ||
               (L.kind == 2 && gid < sc.nt && sc.tri_light[gid] == li);
      };
      // f Le cos cos / r^2 of sample k (stratum k) of area light li, drawn
      // from random stream `stream`; 0 when occluded
      auto area_term = [&](uint32_t li, uint32_t k, uint32_t stream) {
        const orc_scene::Light &L = sc.lights[li];
        const Uniforms lu = draw(job.seed, job.job_id, pixel, sample, depth, stream);
        const uint32_t cell = k % (side * side);
        const double u = ((double)(cell % side) + lu.u[0]) / side, v = ((double)(cell / side) + lu.u[1]) / side;
        // Shape::getPoint (shapes/Shape.cpp:52-64): position and normal
        V3 lp, ln;
        sc.light_point(L, u, v, lu.u[2], lu.u[3], lp, ln);
        V3 wo = lp - p;
        const double dist = len(wo);
        wo = wo / dist;
        const double cos_wo = std::fabs(dot(n, wo));
        const double cos_wi = std::fabs(dot(ln, wo));
        const V3 fr = bsdf_eval(m, n, ray.d, wo);
        if (cos_wo > 0 && cos_wi > 0 && !is_black(fr)) {
          ctr.shadow++;
          if (!sc.occluded(Ray(p, wo), dist - kEps)) {
            const Material &lm = sc.mats[L.mat];
            const V3 Le = dot(ln, -wo) > 0 ? lm.Le : V3();   // OrientedEmitter.cpp:17-23
            return mul(fr, Le) * (cos_wo * cos_wi / (dist * dist));
          }
        }
        return V3();
      };
      // Point lights (DirectIllumination.cpp:96-121, isPoint): one sample, no
      // area weight; a PointShape has no normal, so SurfacePoint's normal
      // dot is 1 (core/SurfacePoint.h:143) and so is cos_wi.
      auto point_term = [&](uint32_t li) {
        const orc_scene::Point &pl = sc.points[li];
        V3 wo = pl.p - p;
        const double dist = len(wo);
        if (!(dist > 0)) return V3();
        wo = wo / dist;
        const double cos_wo = std::fabs(dot(n, wo));
        const V3 fr = bsdf_eval(m, n, ray.d, wo);
        if (cos_wo > 0 && !is_black(fr)) {
          ctr.shadow++;
          if (!sc.occluded(Ray(p, wo), dist - kEps)) return mul(fr, pl.Le) * (cos_wo / (dist * dist));
        }
        return V3();
      };
      V3 direct;
      if (job.light_select) {
        // each sample draws one light as EmitterSampler::sample does
        // (stats/samplers/EmitterSampler.cpp:51-61 with
        // Random::sampleNormalizedCDF, stats/Random.h:125-139) and divides
        // by its probability
        for (uint32_t k = 0; k < ns; ++k) {
          const double x = draw(job.seed, job.job_id, pixel, sample, depth, 1 + ns + k).u[0];
          const uint32_t j = sc.select_light(x);
          const double pj = sc.light_share[j];
          if (j < sc.lights.size()) {
            if (!on_light(j)) direct = direct + area_term(j, k, 1 + k) * (sc.lights[j].area / (ns * pj));
          } else {
            direct = direct + point_term(j - (uint32_t)sc.lights.size()) / (ns * pj);
          }
        }
      } else {
        for (uint32_t li = (uint32_t)sc.lights.size(); li--;) {
          if (on_light(li)) continue;
          V3 dl;
          for (uint32_t k = 0; k < ns; ++k) dl = dl + area_term(li, k, 1 + li * ns + k);
          direct = direct + dl * (sc.lights[li].area / ns);
        }
        for (uint32_t li = (uint32_t)sc.points.size(); li--;) direct = direct + point_term(li);
      }
      L = L + mul(beta, direct);
    }
    const BsdfSample bs = bsdf_sample(m, n, ray.d, bu.u, ior_index);
    if (is_black(bs.wo)) break;
    // PathTracer.cpp:88-92
    const double p_cont = depth < 2 ? (is_black(bs.fs) ? 0.0 : 1.0) : std::min(.95, max_comp(bs.fs));
    if (!(bu.u[2] < p_cont)) break;
    if (depth + 1 >= job.max_depth) break;
    if (depth == 0) d0_continued = true;
    emit_after_specular = m.specular() && dot(n, bs.wo) > kEps;
    beta = mul(beta, bs.fs) / p_cont;
    ray = Ray(p, bs.wo);
  }
  // PathTracer.cpp:107-114 -- with any specular vertex, keep the path's IOR
  // channel only, scaled by the channel count.
  if (any_specular && d0_continued) {
    V3 r;
    r[ior_index] = 3 * L[ior_index];
    L = r;
  }
  return L;
}

// ---------------------------------------------------------------------------
// Bidirectional path tracing: BidirectionalPathTracer::sample
// (renderers/renderers/BidirectionalPathTracer.cpp:34-174) over Path
// (renderers/utils/Path.cpp) and PathVertex (renderers/utils/PathVertex.cpp).
//
// One sample builds an eye subpath (generateE, :233-249) and, unless it ended
// on an emitter, a light subpath (generateL, :222-231); Path::append(Path)
// (Path.cpp:339-515) joins them and completes every vertex's cumulative
// light-side and eye-side values; every split (s, t) of every prefix/suffix
// length k is then weighted by the power heuristic over its k + 1 densities
// (:111-165) and the sum is clamped to [0, 1] per channel (RADIANCE_HACK,
// :167-171).
//
// Random numbers (shared with the device): the camera stream as for the
// path tracer; eye vertex j samples its continuation from (depth j, stream
// 2) plus stream 3 for a dielectric's channel and draws the event of its
// construction (PathVertex::_init, PathVertex.cpp:58-66, whose Phong lobe a
// last vertex keeps) from stream 4; the light subpath uses depths
// kLightDepth + i the same way, its start (EmitterSampler::sample) streams 0
// and 1 of depth kLightDepth; the join of the two (stream 0 of kJoinDepth).
// The device supports pinhole cameras and quad, mesh and point lights here;
// the oracle returns NaN for a thin lens or an emitting sphere.
// ---------------------------------------------------------------------------
const uint32_t kLightDepth = 1024;
const uint32_t kJoinDepth = 2047;   // the join's forced roulette (Path.cpp:364-376)

// EQ (common/utilities.h:240) with its precedence as written; Vector::
// operator== (common/math/Vector.inl:73-80) compares component by component.
inline bool eq_ref(double a, double b) {
  return std::fabs(a - b) < kEps || std::fabs(a - b / ((std::fabs(b) > std::fabs(a)) ? b : a)) <= kEps;
}
inline bool veq(V3 a, V3 b) { return eq_ref(a.x, b.x) && eq_ref(a.y, b.y) && eq_ref(a.z, b.z); }
// fs == SpectralSampleSet::black() (SpectralSampleSet.inl:141-153): every
// channel EQ 0, i.e. |c| <= EPSILON
inline bool black_ref(V3 a) { return std::fabs(a.x) <= kEps && std::fabs(a.y) <= kEps && std::fabs(a.z) <= kEps; }
// SpectralSampleSet::getMaxSample (SpectralSampleSet.inl:388-398)
inline double max_sample(V3 a) {
  double m = a.x;
  if (a.y > m) m = a.y;
  if (a.z > m) m = a.z;
  return m;
}

enum { BV_SURFACE = 0, BV_EMITTER = 1, BV_SENSOR = 2 };
enum { LOBE_ABSORB = 0, LOBE_DIFFUSE = 1, LOBE_SPECULAR = 2 };

struct BVert {
  V3 p, n;
  bool has_normal = true;      // a point light has none (SurfacePointNormal::dot = 1)
  int kind = BV_SURFACE;       // the BSDF it scatters with: material, emitter or sensor
  uint32_t gid = ORC_MISS;
  const Material *m = nullptr; // material of the surface (emission of an emitter vertex)
  int point = -1;              // point light of a light subpath that starts on one
  int lobe = LOBE_DIFFUSE;     // ModifiedPhongBSDF event index of PathVertex::event
  double area = 0;             // Shape::getSurfaceArea of its shape
  double GL = 1, GE = 1, pdfL = 1, pdfE = 1, pL = 1, pE = 1;
  V3 alphaL{1, 1, 1}, alphaE{1, 1, 1}, fs;
};

inline double ndot(const BVert &v, V3 w) { return v.has_normal ? dot(v.n, w) : 1.0; }

// pt->emitter->isEmitter() of the vertex's surface point
inline bool on_emitter(const BVert &v) { return v.point >= 0 || (v.m && v.m->emitter()); }
// BSDF::isSpecular: dielectric and sensor (DielectricBSDF.h:64, Sensor.h:153)
inline bool bv_specular(const BVert &v) {
  return v.kind == BV_SENSOR || (v.kind == BV_SURFACE && v.m->specular());
}

inline double channel_ior(const Material &m, int c) {
  return c > 0 && m.ior_gb[c - 1] > 0 ? m.ior_gb[c - 1] : m.ior;
}

// Emitter::getLe of the vertex's emitter: OrientedEmitter (emitters/
// OrientedEmitter.cpp:17-23) on the normal's side, OmniEmitter (OmniEmitter.
// cpp:15-17) everywhere; NullEmitter (NullEmitter.cpp:20-22) gives black.
inline V3 bv_Le(const orc_scene &sc, const BVert &v, V3 wo) {
  if (v.point >= 0) return sc.points[v.point].Le;
  if (!v.m || !v.m->emitter()) return V3();
  return dot(v.n, wo) > 0 ? v.m->Le : V3();
}
// Emitter::getLe0 = radiant exitance, power / area (Emitter.cpp:30-49,
// 64-66): pi Le for an oriented emitter, the power 4 pi Le for a point.
inline V3 bv_Le0(const orc_scene &sc, const BVert &v) {
  if (v.point >= 0) return sc.points[v.point].Le * (4 * M_PI);
  return v.m->Le * M_PI;
}

// BSDF::getPd(Event(dir, event)) of the vertex with its incident direction
// set to wi.
double bv_pdf(const BVert &v, V3 wi, V3 dir) {
  if (v.kind == BV_SENSOR) {
    // Sensor::getPd (materials/Sensor.cpp:85-95)
    if (veq(wi, -v.n)) return 1;
    return veq(dir, v.n) ? 1 : 0;
  }
  if (v.kind == BV_EMITTER) {
    // Emitter::getPd -> its sampler (Emitter.cpp:160-165): UniformOnSphere
    // returns 1 (stats/samplers/UniformOnSphereSampler.cpp:32-34); an
    // oriented emitter's DiffuseBSDF sees wi = -normalS always
    // (OrientedEmitter.cpp:25-31), so only the side of dir counts.
    if (v.point >= 0) return 1;
    return dot(v.n, dir) > 0 ? 1.0 / M_PI : 0.0;
  }
  const Material &m = *v.m;
  if (m.type == MAT_DIFFUSE) return bsdf_pdf_diffuse(v.n, wi, dir);   // DiffuseBSDF.cpp:30-38
  if (m.type == MAT_PHONG) {
    // ModifiedPhongBSDF::getPd (ModifiedPhongBSDF.cpp:131-162) by event index
    const double kda = avg(m.kd), ksa = avg(m.ks);
    if (v.lobe == LOBE_ABSORB) return 1 - (kda + ksa);
    const double cosA = dot(v.n, dir);
    if (cosA <= 0) return 0;
    double pdf;
    if (v.lobe == LOBE_SPECULAR) {
      const double c = std::fabs(dot(reflect(wi, v.n), dir));
      pdf = std::min(1.0, (m.exponent + 1) / (2 * M_PI) * std::pow(c, m.exponent));
    } else {
      pdf = std::fabs(dot(v.n, dir)) / M_PI;
    }
    return pdf / cosA;
  }
  // DielectricBSDF::getPd (DielectricBSDF.cpp:106-127) with iorIndex 0: the
  // share of channels whose refraction is dir, or 1 for the mirror image
  const V3 N = dot(v.n, wi) > 0 ? -v.n : v.n;
  if (dot(dir, N) < 0) {
    int cnt = 0;
    for (int c = 0; c < 3; ++c) cnt += veq(dir, refract(wi, v.n, kAirIor, channel_ior(m, c))) ? 1 : 0;
    return cnt / 3.0;
  }
  return veq(dir, reflect(wi, N)) ? 1 : 0;
}

// BSDF::evaluate(wi, dir) of the vertex (light arriving along wi, leaving
// along dir).
V3 bv_f(const BVert &v, V3 wi, V3 dir) {
  if (v.kind == BV_SENSOR) {
    // Sensor::evaluate = getWe1(-wi) = (-wi == normalS) (Sensor.cpp:47-62, 98-107)
    return veq(-wi, v.n) ? V3(1, 1, 1) : V3();
  }
  if (v.kind == BV_EMITTER) {
    // Emitter::evaluate = getLe1(dir) = getLe(dir) / getLe0() (Emitter.cpp:
    // 68-70, 167-171): 1 / (4 pi) for a point, 1 / pi on an oriented
    // emitter's side (a channel with no emission gives 0; the reference
    // asserts against dividing by it)
    const V3 le0 = v.point >= 0 ? V3(4 * M_PI, 4 * M_PI, 4 * M_PI) : V3(M_PI, M_PI, M_PI);
    V3 le = v.point >= 0 ? V3(1, 1, 1) : (dot(v.n, dir) > 0 ? V3(1, 1, 1) : V3());
    const V3 ch = v.point >= 0 ? V3(1, 1, 1) : v.m->Le;   // channels that emit
    V3 r;
    for (int c = 0; c < 3; ++c) r[c] = ch[c] != 0 ? le[c] / le0[c] : 0.0;
    return r;
  }
  const Material &m = *v.m;
  if (m.type != MAT_DIELECTRIC) return bsdf_eval(m, v.n, wi, dir);
  // DielectricBSDF::evaluate (DielectricBSDF.cpp:129-158)
  const V3 N = dot(v.n, wi) > 0 ? -v.n : v.n;
  if (dot(dir, N) < 0) {
    V3 s;
    for (int c = 0; c < 3; ++c) s[c] = veq(dir, refract(wi, v.n, kAirIor, channel_ior(m, c))) ? 1 : 0;
    return mul(m.ks, s);
  }
  return veq(dir, reflect(wi, N)) ? m.ks : V3();
}

// BSDF::sample of the vertex: the event's direction (zero when absorbed)
// and Phong lobe.  u: (u0, u1) direction, u2 lobe / refraction choice; uch
// a dielectric's channel (Random::sampleInt(0, 3), stats/Random.h:87-89).
V3 bv_sample(const BVert &v, V3 wi, const double u[4], double uch, int &lobe) {
  lobe = LOBE_DIFFUSE;
  if (v.kind == BV_SENSOR) return v.n;   // Sensor::sample (Sensor.cpp:72-83)
  if (v.kind == BV_EMITTER) {
    if (v.point >= 0) {
      // UniformOnSphereSampler::sample (UniformOnSphereSampler.cpp:22-30)
      const double z = 1 - 2 * u[0], r = std::sqrt(std::max(0.0, 1 - z * z)), phi = 2 * M_PI * u[1];
      return V3(r * std::cos(phi), r * std::sin(phi), z);
    }
    return cos_random(v.n, u[0], u[1]);   // DiffuseBSDF::sample with wi = -normalS
  }
  const Material &m = *v.m;
  if (m.type == MAT_DIFFUSE) {
    if (dot(v.n, wi) > 0) return V3();   // DiffuseBSDF.cpp:21-28
    return cos_random(v.n, u[0], u[1]);
  }
  if (m.type == MAT_PHONG) {
    // ModifiedPhongBSDF::sample (ModifiedPhongBSDF.cpp:84-129): the event's
    // direction as drawn, below the surface or not
    const double kda = avg(m.kd), ksa = avg(m.ks);
    const double p0 = u[2] - kEps;
    if (p0 <= kda) return cos_random(v.n, u[0], u[1]);
    if (p0 <= kda + ksa) {
      lobe = LOBE_SPECULAR;
      const double alpha = std::acos(std::pow(u[0], 1.0 / (m.exponent + 1)));
      const double phi = 2.0 * M_PI * u[1];
      const V3 R = reflect(wi, v.n);
      V3 r = R, uu, vv;
      if (std::fabs(dot(v.n, r)) < .95) {
        uu = normalized(cross(r, v.n));
        vv = cross(uu, r);
      } else {
        onb(r, vv, uu);
        r = normalized(r);
      }
      const V3 w = from_spherical(alpha, phi);
      const V3 wo = uu * w.x + r * w.y + vv * w.z;
      return dot(v.n, wo) > 0 ? wo : R;
    }
    lobe = LOBE_ABSORB;
    return V3();
  }
  // DielectricBSDF::sample (DielectricBSDF.cpp:55-104), iorIndex 0: a random
  // channel's IOR
  const V3 N = dot(v.n, wi) > 0 ? -v.n : v.n;
  const int c = std::min(2, (int)std::floor(uch * 3));
  const double ior = channel_ior(m, c);
  const V3 wt = refract(wi, v.n, kAirIor, ior);
  if (!is_black(wt)) {
    const double Fr = fresnel(wi, v.n, kAirIor, ior);
    if (u[2] < m.transparency * (1 - Fr)) return wt;
  }
  return reflect(wi, N);
}

// The Phong lobe of a new vertex's construction event (PathVertex::_init).
inline int init_lobe(const BVert &v, double u) {
  if (v.kind != BV_SURFACE || v.m->type != MAT_PHONG) return LOBE_DIFFUSE;
  const double kda = avg(v.m->kd), ksa = avg(v.m->ks);
  const double p0 = u - kEps;
  return p0 <= kda ? LOBE_DIFFUSE : (p0 <= kda + ksa ? LOBE_SPECULAR : LOBE_ABSORB);
}

// Path::_initE (Path.cpp:647-693): eye vertex z scattering from wi (toward
// the camera side) to wo (a vertex at distance t with point pt); rr: the
// uniform of the roulette, taken when `roulette` or when fs is black.
bool init_E(BVert &z, V3 wi, V3 wo, double t, const BVert &pt, bool roulette, double rr, V3 &alphaE, double &pE) {
  if (t < kEps) {   // GEOMETRY_BOUND
    z.GE = 0;
    return false;
  }
  z.GE = std::fabs(ndot(z, wo) * ndot(pt, -wo)) / (t * t);
  z.pdfL = bv_pdf(z, -wo, -wi);
  z.fs = bv_f(z, -wo, -wi);
  z.pdfE = bv_pdf(z, wi, wo);
  if (z.GE == 0 || z.pdfE == 0) return false;
  if (roulette || black_ref(z.fs)) {
    const double q = std::min(0.95, max_sample(z.fs) / z.pdfE);
    if (q <= rr) {
      z.pdfE = 0;
      return false;
    }
    z.pdfE *= q;
  }
  pE = (z.pdfE * z.GE) * z.pE;
  alphaE = mul(z.fs / z.pdfE, z.alphaE);
  return true;
}

// Path::_initL (Path.cpp:600-645): the light-side twin.
bool init_L(BVert &y, V3 wi, V3 wo, double t, const BVert &pt, bool roulette, double rr, V3 &alphaL, double &pL) {
  if (t < kEps) {
    y.GL = 0;
    return false;
  }
  y.GL = std::fabs(ndot(y, wo) * ndot(pt, -wo)) / (t * t);
  y.pdfE = bv_pdf(y, -wo, -wi);
  y.fs = bv_f(y, wi, wo);
  y.pdfL = bv_pdf(y, wi, wo);
  if (y.GL == 0 || y.pdfL == 0) return false;
  if (roulette || black_ref(y.fs)) {
    const double q = std::min(0.95, max_sample(y.fs) / y.pdfL);
    if (q <= rr) {
      y.pdfL = 0;
      return false;
    }
    y.pdfL *= q;
  }
  pL = (y.pdfL * y.GL) * y.pL;
  alphaL = mul(y.fs / y.pdfL, y.alphaL);
  return true;
}

// A surface vertex at a closest hit (SurfacePoint::init, core/SurfacePoint.
// cpp:49-78).
BVert surface_vertex(const orc_scene &sc, uint32_t gid, V3 p) {
  BVert v;
  v.p = p;
  v.gid = gid;
  v.n = sc.normal_at(gid, p);
  v.m = &sc.material_of(gid);
  // Shape::getSurfaceArea: a triangle's mesh is taken to have an area of at
  // least EPSILON; quads and spheres (pi) have theirs
  v.area = gid < sc.nt ? 1.0 : sc.shape_area(gid);
  return v;
}

struct Visibility {
  const orc_scene &sc;
  PathCounters &ctr;
  // Scene::intersects(Ray(a, dir(a -> b)), |b - a|) (core/Scene.cpp:110-114)
  bool visible(V3 a, V3 b) {
    V3 w = b - a;
    const double d = len(w);
    ctr.shadow++;
    return !sc.occluded(Ray(a, w / d), d);
  }
};

V3 bdpt_path(const orc_scene &sc, const orc_job &job, const Camera &cam, uint32_t pixel, uint32_t sample,
             PathCounters &ctr) {
  const uint32_t cap = job.max_depth + 1;   // vertices per subpath (the reference has no cap)
  const double nan = std::numeric_limits<double>::quiet_NaN();
  if (cam.aperture > 0) return V3(nan, nan, nan);
  for (const orc_scene::Light &l : sc.lights)
    if (l.kind == 1) return V3(nan, nan, nan);
  if (sc.lights.empty() && sc.points.empty()) return V3(nan, nan, nan);   // EmitterSampler.cpp:51-53 asserts
  auto rnd = [&](uint32_t depth, uint32_t stream) { return draw(job.seed, job.job_id, pixel, sample, depth, stream); };
  auto trace = [&](V3 o, V3 d, uint32_t &gid) {
    ctr.extension++;
    return sc.closest(Ray(o, d), gid);
  };

  // --- eye subpath (generateE; Path::prepend, :277-306 and _samplePathVertex
  // :517-598 with adjoint = true) ------------------------------------------
  std::vector<BVert> eye;
  {
    const Uniforms cu = rnd(kCameraDepth, 0);
    double fx, fy;
    film_position(job, pixel, sample, cu.u[0], cu.u[1], fx, fy);
    // Camera::getPoint at the film position (BidirectionalPathTracer.cpp:
    // 38-45): the eye, normal = the camera ray's direction
    // (Camera::_getGeometricNormal, cameras/Camera.cpp:48-50); pE = 1
    BVert c;
    c.p = cam.eye;
    c.n = cam.dir(fx, fy);
    c.kind = BV_SENSOR;
    c.area = 0;   // Camera::_getSurfaceArea is Shape's 0
    eye.push_back(c);
    ctr.camera++;
  }
  for (;;) {
    const uint32_t T = (uint32_t)eye.size();
    if (T >= cap) break;
    const bool roulette = T >= 2;
    BVert &v = eye.back();
    V3 wi;
    if (T > 1) wi = normalized(v.p - eye[T - 2].p);
    const Uniforms su = rnd(T - 1, 2);
    const double uch = rnd(T - 1, 3).u[0];
    int lobe;
    const V3 wo = bv_sample(v, wi, su.u, uch, lobe);
    v.lobe = lobe;
    if (is_black(wo)) break;
    uint32_t gid;
    const double t = trace(v.p, wo, gid);
    if (!(t > kEps && t < kInf)) break;
    if (T == 1) wi = -wo;
    BVert nv = surface_vertex(sc, gid, v.p + wo * t);
    V3 alphaE;
    double pE;
    if (!init_E(v, wi, wo, t, nv, roulette, su.u[3], alphaE, pE)) break;
    nv.GL = v.GE;
    nv.GE = 1;
    nv.alphaL = V3(1, 1, 1);
    nv.alphaE = alphaE;
    nv.pL = 1;
    nv.pE = pE;
    nv.lobe = init_lobe(nv, rnd(T, 4).u[0]);
    eye.push_back(nv);
    if (on_emitter(eye.back())) break;
  }

  // --- light subpath (generateL; Path::append, :255-275 with
  // EmitterSampler::sample / getPd, stats/samplers/EmitterSampler.cpp:51-96)
  std::vector<BVert> light;
  if (!on_emitter(eye.back())) {
    const Uniforms a = rnd(kLightDepth, 0), b = rnd(kLightDepth, 1);
    const uint32_t j = sc.select_light(a.u[0]);
    BVert e;
    e.kind = BV_EMITTER;
    if (j >= sc.lights.size()) {
      e.point = (int)(j - sc.lights.size());
      e.p = sc.points[e.point].p;
      e.has_normal = false;
      e.area = 0;
    } else {
      const orc_scene::Light &L = sc.lights[j];
      V3 ln;
      sc.light_point(L, a.u[1], a.u[2], a.u[3], b.u[0], e.p, ln);
      e.n = ln;
      e.m = &sc.mats[L.mat];
      e.area = L.area;
      e.gid = L.kind == 0 ? sc.nt + sc.ns + L.index : ORC_MISS;
    }
    const double pA = sc.emit_area > 0 ? 1.0 / sc.emit_area : 1.0;
    // PathVertex(pt, pA, 1) (PathVertex.cpp:31-56): alphaL = Le0 / pL
    e.pL = pA;
    e.pE = 1;
    e.alphaL = bv_Le0(sc, e) / pA;
    e.alphaE = V3(1, 1, 1);
    light.push_back(e);
    for (;;) {
      const uint32_t S = (uint32_t)light.size();
      if (S >= cap) break;
      BVert &v = light.back();
      V3 wi;
      if (S > 1) wi = normalized(v.p - light[S - 2].p);
      const Uniforms su = rnd(kLightDepth + S - 1, 2);
      const double uch = rnd(kLightDepth + S - 1, 3).u[0];
      int lobe;
      const V3 wo = bv_sample(v, wi, su.u, uch, lobe);
      v.lobe = lobe;
      if (is_black(wo)) break;
      uint32_t gid;
      const double t = trace(v.p, wo, gid);
      if (!(t > kEps && t < kInf)) break;
      if (S == 1) wi = -wo;
      BVert nv = surface_vertex(sc, gid, v.p + wo * t);
      V3 alphaL;
      double pL;
      if (!init_L(v, wi, wo, t, nv, true, su.u[3], alphaL, pL)) break;
      nv.GL = 1;
      nv.GE = v.GL;
      nv.alphaL = alphaL;
      nv.alphaE = V3(1, 1, 1);
      nv.pL = pL;
      nv.pE = 1;
      nv.lobe = init_lobe(nv, rnd(kLightDepth + S, 4).u[0]);
      light.push_back(nv);
    }
  }

  // --- Path::append(const Path &) (Path.cpp:339-515) ----------------------
  const uint32_t S = (uint32_t)light.size(), T = (uint32_t)eye.size(), n = S + T;
  std::vector<BVert> m(light);
  for (uint32_t j = T; j--;) m.push_back(eye[j]);
  if (on_emitter(m[0])) m[0].kind = BV_EMITTER;
  Visibility vis{sc, ctr};
  bool valid = true;
  if (S > 0 && T > 0) {
    BVert &y = m[S - 1], &z = m[S];
    V3 wo = z.p - y.p;
    const double d = len(wo);
    wo = wo / d;
    V3 alpha;
    double pA;
    // the roulette a black fs forces at either end draws from its own stream
    const Uniforms ju = rnd(kJoinDepth, 0);
    V3 wi = S > 1 ? normalized(y.p - m[S - 2].p) : -wo;
    if (!init_L(y, wi, wo, d, z, false, ju.u[0], alpha, pA)) valid = false;
    wi = T > 1 ? normalized(z.p - m[S + 1].p) : wo;
    if (!init_E(z, wi, -wo, d, y, false, ju.u[1], alpha, pA)) valid = false;
    if (valid && !vis.visible(y.p, z.p)) {
      y.GL = z.GE = 0;
      valid = false;
    }
  }
  // the eye-side values of the light vertices ...
  for (uint32_t i = S; i--;) {
    BVert &y = m[i];
    const BVert &z = m[i + 1];
    y.alphaE = (z.pdfE > 0 && z.GE > 0) ? mul(z.alphaE, z.fs / z.pdfE) : V3();
    y.pE = z.pE * z.pdfE * z.GE;
  }
  // ... and the light-side values of the eye vertices
  for (uint32_t i = 0; i < T; ++i) {
    BVert &z = m[S + i];
    if (S + i == 0) {
      // an eye subpath that ended on an emitter: it starts the light side
      // with the emitter's EmitterSampler density
      if (on_emitter(z)) {
        const double pA = sc.emit_area > 0 ? 1.0 / sc.emit_area : 1.0;
        const V3 wo = n > 1 ? normalized(m[1].p - z.p) : V3();
        z.kind = BV_EMITTER;
        z.alphaL = bv_Le0(sc, z) / pA;
        z.pL = pA;
        z.fs = bv_f(z, V3(), wo);
        z.pdfL = bv_pdf(z, V3(), wo);
      } else {
        z.alphaL = V3();
        z.pL = 0;
        z.fs = V3();
        z.pdfL = 0;
      }
    } else {
      const BVert &y = m[S + i - 1];
      z.alphaL = (y.pdfL > 0 && y.GL > 0) ? mul(y.alphaL, y.fs / y.pdfL) : V3();
      z.pL = y.pL * y.pdfL * y.GL;
    }
  }

  // --- getPd / getContribution of every split (Path.cpp:48-197) ----------
  // visibility of the pair (u, v), traced at most once
  std::vector<int8_t> seen((size_t)n * n, -1);
  auto pair_visible = [&](uint32_t u, uint32_t v) {
    int8_t &c = seen[(size_t)u * n + v];
    if (c < 0) c = vis.visible(m[u].p, m[v].p) ? 1 : 0;
    return c == 1;
  };
  auto pd = [&](uint32_t s, uint32_t t) -> double {
    const uint32_t u = s - 1, v = n - t;
    if (s == 0) return m[v].area < kEps ? 0.0 : m[v].pE;
    if (t == 0) return m[u].area < kEps ? 0.0 : m[u].pL;
    const BVert &y = m[u], &z = m[v];
    if (bv_specular(y) || bv_specular(z)) return 0;
    if (s + t == n) return y.pL * z.pE;
    if (len(z.p - y.p) < kEps) return 0;
    if (!pair_visible(u, v)) return 0;
    return y.pL * z.pE;
  };
  auto contribution = [&](uint32_t s, uint32_t t) -> V3 {
    const uint32_t u = s - 1, v = n - t;
    if (s == 0) {
      if (m[v].area < kEps) return V3();
      const V3 wo = v < n - 1 ? normalized(m[v + 1].p - m[v].p) : V3();
      return mul(bv_Le(sc, m[v], wo), m[v].alphaE);
    }
    if (t == 0) return V3();   // NullSensor::getWe is black off the camera, which has no area
    const BVert &y = m[u], &z = m[v];
    if (bv_specular(y) || bv_specular(z)) return V3();
    if (s + t == n) return mul(mul(mul(y.alphaL, y.fs) * y.GL, z.fs), z.alphaE);
    V3 wo = z.p - y.p;
    const double d = len(wo);
    if (d < kEps) return V3();
    if (!pair_visible(u, v)) return V3();
    wo = wo / d;
    const V3 ywi = u > 0 ? normalized(y.p - m[u - 1].p) : V3();
    const V3 wo2 = v < n - 1 ? normalized(m[v + 1].p - z.p) : -wo;
    const double G = std::fabs(ndot(y, wo) * ndot(z, -wo)) / (d * d);
    return mul(mul(mul(y.alphaL, bv_f(y, ywi, wo)) * G, bv_f(z, wo, wo2)), z.alphaE);
  };
  const uint32_t length = n - (valid ? 0 : 1);
  V3 L;
  std::vector<double> pdf(n + 2);
  for (uint32_t k = 2; k <= length; ++k) {
    double sum = 0;
    for (uint32_t s = 0; s <= k; ++s) {
      const double p = pd(s, k - s);
      pdf[s] = p * p;   // power heuristic, beta = 2
      sum += pdf[s];
    }
    if (sum == 0) continue;
    for (uint32_t s = 0; s <= k; ++s)
      if (pdf[s] > 0) L = L + contribution(s, k - s) * (pdf[s] / sum);
  }
  // RADIANCE_HACK: CLAMP(value, 0, 1) per channel (common/utilities.h:195-200)
  for (int c = 0; c < 3; ++c) L[c] = L[c] >= 1 ? 1.0 : (L[c] <= 0 ? 0.0 : L[c]);
  return L;
}

V3 sample_path(const orc_scene &sc, const orc_job &job, const Camera &cam, uint32_t pixel, uint32_t sample,
               PathCounters &ctr) {
  if (job.renderer == 1) return bdpt_path(sc, job, cam, pixel, sample, ctr);
  return trace_path(sc, job, cam, pixel, sample, ctr);
}

template <class F>
void parallel_for(uint64_t n, int nthreads, uint64_t chunk, F &&f) {
  if (nthreads <= 1 || n <= chunk) {
    f(0, n, 0);
    return;
  }
  std::atomic<uint64_t> next(0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) {
    th.emplace_back([&, t]() {
      for (;;) {
        const uint64_t b = next.fetch_add(chunk);
        if (b >= n) break;
        f(b, std::min(n, b + chunk), t);
      }
    });
  }
  for (auto &x : th) x.join();
}

#include "oracle_mlt.inc"

}  // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  philox(ctr, key, out);
}

// Bumped with every change of an exported signature; pyoracle refuses a
// library built from older sources.
int orc_interface_version(void) { return 13; }

orc_scene *orc_scene_create(uint32_t nv, const float *vertices, uint32_t nt, const uint32_t *indices,
                            const uint32_t *tri_material, uint32_t ns, const float *spheres,
                            const uint32_t *sphere_material, uint32_t nq, const float *quads,
                            const uint32_t *quad_material, uint32_t nm, const float *materials,
                            const float env_radiance[3], const uint32_t *tri_flags, const float *vertex_normals,
                            const uint8_t *vertex_normal_mask, uint32_t kd_min_prims, uint32_t kd_max_depth,
                            int use_naive, int nthreads) {
  std::unique_ptr<orc_scene> s(new orc_scene());
  s->nt = nt;
  s->ns = ns;
  s->nq = nq;
  s->verts.resize(nv);
  for (uint32_t i = 0; i < nv; ++i) s->verts[i] = V3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]);
  s->idx.assign(indices, indices + 3ull * nt);
  s->tri_mat.assign(tri_material, tri_material + nt);
  s->one_sided.assign(nt, 0);
  if (tri_flags)
    for (uint32_t i = 0; i < nt; ++i) s->one_sided[i] = (tri_flags[i] & 1u) ? 1 : 0;
  for (uint32_t i = 0; i < 3ull * nt; ++i)
    if (s->idx[i] >= nv) return nullptr;
  for (uint32_t i = 0; i < nm; ++i) s->mats.push_back(load_material(materials + 16ull * i));
  for (uint32_t i = 0; i < nt; ++i)
    if (s->tri_mat[i] >= nm || (s->mats[s->tri_mat[i]].emitter() && s->one_sided[i])) return nullptr;
  // Mesh::computeNormals (shapes/Mesh.cpp:335-377): sum of unit face normals.
  s->vnormals.assign(nv, V3());
  for (uint32_t i = 0; i < nt; ++i) {
    const uint32_t a = s->idx[3 * i], b = s->idx[3 * i + 1], c = s->idx[3 * i + 2];
    const V3 fn = normalized(cross(s->verts[b] - s->verts[a], s->verts[c] - s->verts[a]));
    s->vnormals[a] = s->vnormals[a] + fn;
    s->vnormals[b] = s->vnormals[b] + fn;
    s->vnormals[c] = s->vnormals[c] + fn;
  }
  // a mesh file's own normals replace the computed ones
  if (vertex_normals && vertex_normal_mask)
    for (uint32_t i = 0; i < nv; ++i)
      if (vertex_normal_mask[i])
        s->vnormals[i] = V3(vertex_normals[3 * i], vertex_normals[3 * i + 1], vertex_normals[3 * i + 2]);
  for (auto &v : s->vnormals) v = normalized(v);
  for (uint32_t i = 0; i < ns; ++i) {
    s->spheres.push_back(Sphere{V3(spheres[4 * i], spheres[4 * i + 1], spheres[4 * i + 2]), spheres[4 * i + 3]});
    if (sphere_material[i] >= nm) return nullptr;
  }
  s->sph_mat.assign(sphere_material, sphere_material + ns);
  for (uint32_t i = 0; i < nq; ++i) {
    const float *q = quads + 9ull * i;
    Quad Q;
    Q.c = V3(q[0], q[1], q[2]);
    Q.ex = V3(q[3], q[4], q[5]);
    Q.ez = V3(q[6], q[7], q[8]);
    Q.n = normalized(cross(Q.ez, Q.ex));
    Q.rx = Q.ex / dot(Q.ex, Q.ex);
    Q.rz = Q.ez / dot(Q.ez, Q.ez);
    Q.area = len(cross(Q.ex, Q.ez));
    s->quads.push_back(Q);
    if (quad_material[i] >= nm) return nullptr;
  }
  s->quad_mat.assign(quad_material, quad_material + nq);
  s->build_lights(nullptr, nullptr);
  s->env = V3(env_radiance[0], env_radiance[1], env_radiance[2]);

  KdParams kp;
  if (kd_min_prims) kp.min_prims = kd_min_prims;
  if (kd_max_depth) kp.max_depth = std::min<unsigned>(kd_max_depth, 60);
  // MeshTriangle::getAABB (shapes/MeshTriangle.cpp:186-198)
  std::vector<Aabb> tb(nt);
  for (uint32_t i = 0; i < nt; ++i)
    for (int k = 0; k < 3; ++k) tb[i].add(s->verts[s->idx[3 * i + k]]);
  s->mesh_tree.build(std::move(tb), kp, use_naive != 0, nthreads);
  std::vector<Aabb> top;
  if (nt) {
    top.push_back(s->mesh_tree.box);
    s->top_kind.push_back(0);
    s->top_index.push_back(0);
  }
  for (uint32_t i = 0; i < ns; ++i) {
    Aabb b;
    const Sphere &sp = s->spheres[i];
    b.add(sp.c - V3(sp.r, sp.r, sp.r));
    b.add(sp.c + V3(sp.r, sp.r, sp.r));
    top.push_back(b);
    s->top_kind.push_back(1);
    s->top_index.push_back(i);
  }
  for (uint32_t i = 0; i < nq; ++i) {
    const Quad &Q = s->quads[i];
    Aabb b;
    for (int a = -1; a <= 1; a += 2)
      for (int c = -1; c <= 1; c += 2) b.add(Q.c + Q.ex * (0.5 * a) + Q.ez * (0.5 * c));
    top.push_back(b);
    s->top_kind.push_back(2);
    s->top_index.push_back(i);
  }
  s->top_tree.build(std::move(top), kp, use_naive != 0, 1);
  return s.release();
}

void orc_scene_destroy(orc_scene *s) { delete s; }

void orc_kd_stats(const orc_scene *s, int check, uint64_t out[6]) {
  const KdTree &k = s->mesh_tree;
  uint64_t maxleaf = 0;
  for (const KdNode &n : k.nodes)
    if (n.axis == 3) maxleaf = std::max<uint64_t>(maxleaf, n.count);
  out[0] = k.n_internal;
  out[1] = k.n_leaves;
  out[2] = k.max_depth_seen;
  out[3] = k.lists.size();
  out[4] = maxleaf;
  out[5] = check ? k.check_invariant() : 0;
}

void orc_trace_closest(const orc_scene *s, uint64_t n, const double *rays, double *t_out, uint32_t *id_out,
                       int nthreads) {
  parallel_for(n, nthreads, 4096, [&](uint64_t b, uint64_t e, int) {
    for (uint64_t i = b; i < e; ++i) {
      const double *r = rays + 6 * i;
      uint32_t gid;
      const double t = s->closest(Ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5])), gid);
      const bool hit = t > kEps && t < kInf;
      t_out[i] = hit ? t : kInf;
      id_out[i] = hit ? gid : ORC_MISS;
    }
  });
}

void orc_prim_t(const orc_scene *s, uint64_t n, const double *rays, const uint32_t *prim, double *t_out) {
  for (uint64_t i = 0; i < n; ++i) {
    const double *r = rays + 6 * i;
    const Ray ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5]));
    const uint32_t g = prim[i];
    double t = kInf;
    if (g < s->nt) t = s->tri_t(g, ray);
    else if (g < s->nt + s->ns) t = s->sphere_t(g - s->nt, ray);
    else if (g < s->nt + s->ns + s->nq) t = s->quad_t(g - s->nt - s->ns, ray);
    t_out[i] = (t > kEps && t < kInf) ? t : kInf;
  }
}

void orc_trace_any(const orc_scene *s, uint64_t n, const double *rays, const double *tmax, uint8_t *occ,
                   int nthreads) {
  parallel_for(n, nthreads, 4096, [&](uint64_t b, uint64_t e, int) {
    for (uint64_t i = b; i < e; ++i) {
      const double *r = rays + 6 * i;
      occ[i] = s->occluded(Ray(V3(r[0], r[1], r[2]), V3(r[3], r[4], r[5])), tmax[i]) ? 1 : 0;
    }
  });
}

void orc_render(const orc_scene *s, const orc_job *job, double *sum, double *sumsq, uint64_t *counters,
                int nthreads) {
  const Camera cam(job->camera, s);
  const uint64_t np = job->pix_end - job->pix_begin;
  const int nt = std::max(1, nthreads);
  std::vector<PathCounters> pc(nt);
  parallel_for(np, nt, 64, [&](uint64_t b, uint64_t e, int tid) {
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t pixel = job->pix_begin + (uint32_t)i;
      V3 acc, acc2;
      for (uint32_t smp = 0; smp < job->spp; ++smp) {
        const V3 L = sample_path(*s, *job, cam, pixel, smp, pc[tid]);
        acc = acc + L;
        acc2 = acc2 + mul(L, L);
      }
      for (int c = 0; c < 3; ++c) {
        sum[3 * i + c] = acc[c];
        if (sumsq) sumsq[3 * i + c] = acc2[c];
      }
    }
  });
  if (counters) {
    PathCounters tot;
    for (auto &c : pc) {
      tot.camera += c.camera;
      tot.extension += c.extension;
      tot.shadow += c.shadow;
    }
    counters[0] = tot.camera;
    counters[1] = tot.extension;
    counters[2] = tot.shadow;
    counters[3] = np * job->spp;
  }
}

void orc_render_paths(const orc_scene *s, const orc_job *job, uint64_t n, const uint32_t *pixel,
                      const uint32_t *sample, double *L_out, int nthreads) {
  const Camera cam(job->camera, s);
  parallel_for(n, std::max(1, nthreads), 1024, [&](uint64_t b, uint64_t e, int) {
    PathCounters pc;
    for (uint64_t i = b; i < e; ++i) {
      const V3 L = sample_path(*s, *job, cam, pixel[i], sample[i], pc);
      L_out[3 * i] = L.x;
      L_out[3 * i + 1] = L.y;
      L_out[3 * i + 2] = L.z;
    }
  });
}

namespace {
double orc_filter(int type, double support, const double *p, double x, double y);
}

void orc_render_film(const orc_scene *s, const orc_job *job, double *rgbw_out, int nthreads) {
  const uint32_t W = job->width, H = job->height, spp = job->spp;
  const uint64_t n = (uint64_t)W * H * spp;
  std::vector<double> L(3 * n);
  const Camera cam(job->camera, s);
  parallel_for(n, std::max(1, nthreads), 1024, [&](uint64_t b, uint64_t e, int) {
    PathCounters pc;
    for (uint64_t i = b; i < e; ++i) {
      const V3 v = sample_path(*s, *job, cam, (uint32_t)(i / spp), (uint32_t)(i % spp), pc);
      L[3 * i] = v.x;
      L[3 * i + 1] = v.y;
      L[3 * i + 2] = v.z;
    }
  });
  std::fill(rgbw_out, rgbw_out + 4ull * W * H, 0.0);
  const int type = (int)job->filter - 1;
  const double support = job->filter_support ? job->filter_support : 2.0;
  double p[2] = {job->filter_params[0], job->filter_params[1]};
  if (type == 2 && p[0] == 0) p[0] = 1.0;
  if (type == 3 && p[0] == 0 && p[1] == 0) p[0] = p[1] = 1.0 / 3.0;
  if (type == 4 && p[0] == 0) p[0] = 3.0;
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t pixel = (uint32_t)(i / spp), sample = (uint32_t)(i % spp);
    if (job->filter == 0) {   // RenderOutput::addSample: its own pixel, weight 1
      for (int c = 0; c < 3; ++c) rgbw_out[4ull * pixel + c] += L[3 * i + c];
      rgbw_out[4ull * pixel + 3] += 1;
      continue;
    }
    const Uniforms cu = draw(job->seed, job->job_id, pixel, sample, kCameraDepth, 0);
    double fx, fy;
    film_position(*job, pixel, sample, cu.u[0], cu.u[1], fx, fy);
    // Viewport::getBin (core/Viewport.h:78-81), then the 3 x 3 neighbourhood
    // (the output filter's width: KernelFilter() keeps support 2's width 3)
    const int col = (int)std::min(std::max(std::floor(fx * W), 0.0), W - 1.0);
    const int row = (int)std::min(std::max(std::floor(fy * H), 0.0), H - 1.0);
    const double cx = fx * W, cy = fy * H;
    for (int y = std::max(row - 1, 0); y <= std::min(row + 1, (int)H - 1); ++y)
      for (int x = std::max(col - 1, 0); x <= std::min(col + 1, (int)W - 1); ++x) {
        const double w = orc_filter(type, support, p, x - cx, y - cy);
        if (w > 0) {
          double *o = rgbw_out + 4ull * ((uint64_t)y * W + x);
          for (int c = 0; c < 3; ++c) o[c] += w * L[3 * i + c];
          o[3] += w;
        }
      }
  }
}

void orc_camera_ray(const orc_job *job, uint32_t pixel, uint32_t sample, double *ray_out) {
  const Camera cam(job->camera);
  const Uniforms cu = draw(job->seed, job->job_id, pixel, sample, kCameraDepth, 0);
  double fx, fy;
  film_position(*job, pixel, sample, cu.u[0], cu.u[1], fx, fy);
  const Uniforms lu = draw(job->seed, job->job_id, pixel, sample, kCameraDepth, 1);
  const Ray r = cam.primary(fx, fy, lu.u[0], lu.u[1]);
  ray_out[0] = r.o.x; ray_out[1] = r.o.y; ray_out[2] = r.o.z;
  ray_out[3] = r.d.x; ray_out[4] = r.d.y; ray_out[5] = r.d.z;
}

void orc_bsdf_sample(const float *material, const double *n, const double *wi, const double *u,
                     uint32_t ior_index, double *wo, double *fs, double *pdf, int *specular) {
  const Material m = load_material(material);
  const BsdfSample s = bsdf_sample(m, V3(n[0], n[1], n[2]), V3(wi[0], wi[1], wi[2]), u, ior_index);
  for (int i = 0; i < 3; ++i) { wo[i] = s.wo[i]; fs[i] = s.fs[i]; }
  *pdf = s.pdf;
  *specular = m.specular() ? 1 : 0;
}

void orc_bsdf_eval(const float *material, const double *n, const double *wi, const double *wo,
                   double *f, double *pdf) {
  const Material m = load_material(material);
  const V3 N(n[0], n[1], n[2]), WI(wi[0], wi[1], wi[2]), WO(wo[0], wo[1], wo[2]);
  const V3 r = bsdf_eval(m, N, WI, WO);
  for (int i = 0; i < 3; ++i) f[i] = r[i];
  if (m.type == MAT_DIFFUSE) {
    *pdf = bsdf_pdf_diffuse(N, WI, WO);
  } else if (m.type == MAT_PHONG) {
    // Mixture pdf of the two lobes over projected solid angle.
    const double kda = avg(m.kd), ksa = avg(m.ks);
    const double cosA = dot(N, WO);
    if (cosA <= 0) { *pdf = 0; return; }
    const double ca = std::fabs(dot(reflect(WI, N), WO));
    double spec = std::min(1.0, (m.exponent + 1) / (2 * M_PI) * std::pow(ca, m.exponent));
    *pdf = (kda * (cosA / M_PI) + ksa * spec) / cosA;
  } else {
    *pdf = 0;
  }
}

double orc_fresnel(const double *wi, const double *n, double n1, double n2) {
  return fresnel(V3(wi[0], wi[1], wi[2]), V3(n[0], n[1], n[2]), n1, n2);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tone maps: RenderOutput::getFinalizedOutput (renderers/RenderOutput.cpp:
// 129-231, the non-MLT branch: value = numerator / denominator, black where
// the denominator is 0, ProgressiveFilterValue.h:27-32) followed by the
// selected map.  Input: W*H (numerator RGB, denominator) float quadruples.
// ---------------------------------------------------------------------------

namespace {

// utils/ColorUtils.cpp:115-133: the sRGB (D65) primaries matrices.
const double kRgbToXyz[9] = {0.41245322014416, 0.35757958129350, 0.18042258997054,
                             0.21267112134122, 0.71515920531078, 0.07216877677613,
                             0.01933381646199, 0.11919354020663, 0.95022692228971};
const double kXyzToRgb[9] = {3.240479, -1.537150, -0.498535, -0.969256, 1.875992,
                             0.041556, 0.055648,  -0.204043, 1.057311};

void mat3(const double *m, double a, double b, double c, double &x, double &y, double &z) {
  x = m[0] * a + m[1] * b + m[2] * c;
  y = m[3] * a + m[4] * b + m[5] * c;
  z = m[6] * a + m[7] * b + m[8] * c;
}

// (unsigned char) CLAMP(v, 0, 255) (common/utilities.h:195-200): values at or
// above 255 give 255, at or below 0 give 0, the rest truncate.  A NaN passes
// the clamp; x86-64's truncating conversion of it yields byte 0.
uint8_t to_byte(double v) {
  if (v >= 255.0) return 255;
  if (v <= 0.0 || v != v) return 0;
  return (uint8_t)(int)v;
}

}  // namespace

extern "C" void orc_tonemap(const float *rgbw, uint64_t n, int mode, uint8_t *rgba) {
  std::vector<double> v(3 * n);
  for (uint64_t i = 0; i < n; ++i) {
    const double w = rgbw[4 * i + 3];
    for (int c = 0; c < 3; ++c) v[3 * i + c] = w == 0 ? 0.0 : (double)rgbw[4 * i + c] / w;
  }
  const double correction = 1.0 / 2.22;
  if (mode != 0) {
    // NullToneMap::map (renderers/tonemaps/NullToneMap.cpp:21-46)
    for (uint64_t i = 0; i < n; ++i) {
      for (int c = 0; c < 3; ++c) rgba[4 * i + c] = to_byte(255 * std::pow(v[3 * i + c], correction));
      rgba[4 * i + 3] = 255;
    }
    return;
  }
  // ReinhardToneMap::map (renderers/tonemaps/ReinhardToneMap.cpp:25-98):
  // to XYZ, chromaticity x = X/(X+Y+Z), y = Y/(X+Y+Z); key = exp(mean of
  // log(1e-7 + Y)) over the image (visited last pixel first); white = 1 /
  // max(Y)^2.
  std::vector<double> cx(n), cy(n), Y(n);
  double key = 0, white = 0;
  for (uint64_t i = n; i--;) {
    double X, Yv, Z;
    mat3(kRgbToXyz, v[3 * i], v[3 * i + 1], v[3 * i + 2], X, Yv, Z);
    cx[i] = X / (X + Yv + Z);
    cy[i] = Yv / (X + Yv + Z);
    Y[i] = Yv;
    key += std::log(1e-7 + Yv);
    white = std::max(white, Yv);
  }
  key = std::exp(key / (double)n);
  white = white <= 0 ? 0 : 1.0 / (white * white);
  // scaled luminance L = (0.18 / key) Y, compressed to L (1 + L white) / (1 + L),
  // back through xyY -> XYZ -> RGB, then gamma 2.22
  const double scale = key <= 0 ? 0 : 0.18 / key;
  for (uint64_t i = n; i--;) {
    const double lum = scale * Y[i];
    const double Yd = lum * (1.0 + lum * white) / (1.0 + lum);
    const double X = Yd * cx[i] / cy[i];
    const double Z = Yd * (1 - cx[i] - cy[i]) / cy[i];
    double r, g, b;
    mat3(kXyzToRgb, X, Yd, Z, r, g, b);
    rgba[4 * i] = to_byte(255 * std::pow(r, correction));
    rgba[4 * i + 1] = to_byte(255 * std::pow(g, correction));
    rgba[4 * i + 2] = to_byte(255 * std::pow(b, correction));
    rgba[4 * i + 3] = 255;
  }
}

// ---------------------------------------------------------------------------
// Environment images.
// ---------------------------------------------------------------------------
namespace {

// Filter2D::evaluate of the five KernelFilter types
// (filters/{Box,Triangle,Gaussian,Mitchell,LanczosSinc}Filter.cpp).
double orc_filter(int type, double support, const double *p, double x, double y) {
  if (type == 0) return 2.0 * std::sqrt(x * x + y * y) < support;   // BoxFilter.cpp:14-16
  if (type == 1) {                                                    // TriangleFilter.cpp:14-18
    const double radius = support / 2.0;
    return std::max(0.0, 1.0 - x / radius) * std::max(0.0, 1.0 - y / radius);
  }
  if (type == 2) {   // GaussianFilter.cpp:22-31
    const double coeff = (-0.5) / (p[0] * p[0]);
    return std::max(0.0, (1.0 / (std::sqrt(2.0 * M_PI) * p[0])) * std::exp(coeff * (x * x + y * y)));
  }
  if (type == 3) {   // MitchellFilter.cpp:27-43
    auto f = [&](double t) {
      t = std::fabs(2.0 * t / support);
      const double B = p[0], C = p[1];
      if (t > 1.0) return (((-B / 6.0 - C) * t + (B + 5.0 * C)) * t + (-2.0 * B - 8.0 * C)) * t + (4.0 / 3.0 * B + 4.0 * C);
      return t * t * ((2.0 - 1.5 * B - C) * t + (-3.0 + 2.0 * B + C)) + (1.0 - B / 3.0);
    };
    return f(x) * f(y);
  }
  auto f = [&](double t) {   // LanczosSincFilter.cpp:26-45
    t = std::fabs(t / support);
    if (t < kEps) return 1.0;
    if (t > 1) return 0.0;
    t *= M_PI;
    return std::sin(t * p[0]) / (t * p[0]) * (std::sin(t) / t);
  };
  return f(x) * f(y);
}

}  // namespace

// Point lights: 6 floats each, position and power; radiance P / (4 pi)
// (emitters/OmniEmitter.cpp:15-17; Emitter.cpp:44-46 for a zero area).
// Emitting meshes and spheres: the shape of each triangle and the spheres'
// transforms (both optional), as in mpt_scene_desc.
extern "C" void orc_scene_set_emitters(orc_scene *s, const uint32_t *tri_shape, const float *sphere_frames) {
  s->build_lights(tri_shape, sphere_frames);
}

// EmitterSampler::sample / ::getPd (stats/samplers/EmitterSampler.cpp:51-96):
// a light by power (Random::sampleNormalizedCDF), then its getRandomPoint.
// Sample i's uniforms: Philox (seed, i), stream 0 for the light and (u, v),
// stream 1 for a mesh triangle's further two.
extern "C" int orc_emitter_sample(const orc_scene *s, uint64_t n, uint32_t seed, double *out) {
  const uint32_t nl = (uint32_t)(s->lights.size() + s->points.size());
  if (nl == 0) return -1;
  const double pdf_area = s->emit_area > 0 ? 1.0 / s->emit_area : 1.0;   // EmitterSampler.cpp:86-93
  for (uint64_t i = 0; i < n; ++i) {
    const Uniforms r = draw(seed, 0, (uint32_t)i, (uint32_t)(i >> 32), 0, 0);
    const uint32_t j = s->select_light(r.u[0]);
    const double u = r.u[1], v = r.u[2];
    V3 lp, ln;
    if (j >= s->lights.size()) {
      lp = s->points[j - s->lights.size()].p;   // PointShape: its position, no normal
    } else {
      const orc_scene::Light &L = s->lights[j];
      if (L.kind == 1) {
        // Sphere::getRandomPoint (shapes/Sphere.cpp:107-115): a uniform point
        // on the unit sphere (z = 1 - 2u, azimuth 2 pi v), Sphere::_getUV
        // (:79-86), Sphere::getPosition (:117-124) with the reference's
        // CLAMP, through which a NaN passes
        const double sz = 1.0 - 2.0 * u, rho = std::sqrt(std::max(0.0, 1.0 - sz * sz)), phi = 2.0 * M_PI * v;
        const double sx = rho * std::cos(phi), sy = rho * std::sin(phi);
        const double su = 1.0 - std::atan2(sz, sx) * (1.0 / (2.0 * M_PI));
        const double sv = 0.5 - std::asin(sy * 2.0) * (1.0 / M_PI);
        auto clamp = [](double x, double lo, double hi) { return x >= hi ? hi : (x <= lo ? lo : x); };
        const double x = clamp(std::tan((1.0 - su) / (2.0 * M_PI)), -0.5, 0.5);
        const double y = clamp(std::sin((0.5 - sv) / M_PI) * 0.5, -0.5, 0.5);
        const double z = std::sqrt(1.0 - x * x - y * y);
        const V3 q = L.frame[0] * x + L.frame[1] * y + L.frame[2] * z;
        lp = s->spheres[L.index].c + q;
        ln = q / len(q);
      } else {
        // Shape::getRandomPoint (shapes/Shape.cpp:45-50): getPoint at two
        // uniforms
        const Uniforms r2 = draw(seed, 0, (uint32_t)i, (uint32_t)(i >> 32), 0, 1);
        s->light_point(L, u, v, r2.u[0], r2.u[1], lp, ln);
      }
    }
    double *o = out + 9 * i;
    o[0] = lp.x;
    o[1] = lp.y;
    o[2] = lp.z;
    o[3] = ln.x;
    o[4] = ln.y;
    o[5] = ln.z;
    o[6] = j;
    o[7] = pdf_area;
    o[8] = s->light_share[j];
  }
  return 0;
}

extern "C" void orc_scene_set_points(orc_scene *s, uint32_t n, const float *pts) {
  s->points.clear();
  for (uint32_t i = 0; i < n; ++i) {
    const float *q = pts + 6ull * i;
    orc_scene::Point pl;
    pl.p = V3(q[0], q[1], q[2]);
    pl.Le = V3(q[3], q[4], q[5]) / (4.0 * M_PI);
    s->points.push_back(pl);
  }
  s->update_shares();
}

extern "C" int orc_scene_set_env_map(orc_scene *s, uint32_t w, uint32_t h, const float *rgb, float repeat_u,
                                     float repeat_v, int filter, uint32_t support, const float *params) {
  if (!s || !w || !h || !rgb || filter < 0 || filter > 4) return -1;
  support = support ? support : 2;
  // KernelFilter(support): m_width = ceil(support) | 1 (filters/KernelFilter.h:26-29)
  const uint32_t m = ((uint32_t)std::ceil((double)support)) | 1u;
  double p[2] = {params ? params[0] : 0.0, params ? params[1] : 0.0};
  if (filter == 2 && p[0] == 0) p[0] = 1.0;
  if (filter == 3 && p[0] == 0 && p[1] == 0) p[0] = p[1] = 1.0 / 3.0;
  if (filter == 4 && p[0] == 0) p[0] = 3.0;
  // KernelFilter::_computeKernel (filters/KernelFilter.cpp:162-176)
  const double half = m / 2.0;
  s->env_kernel.assign((size_t)m * m, 0.0);
  for (uint32_t i = m; i--;)
    for (uint32_t j = m; j--;) s->env_kernel[(size_t)i * m + j] = orc_filter(filter, support, p, i - half, j - half);
  s->env_m = m;
  s->env_w = w;
  s->env_h = h;
  s->env_img.resize((size_t)w * h);
  for (size_t i = 0; i < (size_t)w * h; ++i) s->env_img[i] = V3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
  s->env_rep_u = repeat_u ? repeat_u : 1.0;
  s->env_rep_v = repeat_v ? repeat_v : 1.0;
  return 0;
}

extern "C" void orc_background(const orc_scene *s, const double *d, double *rgb) {
  const V3 L = s->background(V3(d[0], d[1], d[2]));
  rgb[0] = L.x;
  rgb[1] = L.y;
  rgb[2] = L.z;
}

// ---------------------------------------------------------------------------
// Metropolis light transport entry points.
// ---------------------------------------------------------------------------
namespace {

// The scenes and cameras the MLT restatement covers: a pinhole and quad,
// mesh and point lights (as the bidirectional renderer).
bool mlt_supported(const orc_scene &sc, const Camera &cam, const orc_job &job, const orc_mlt &p) {
  if (cam.aperture > 0) return false;
  for (const orc_scene::Light &l : sc.lights)
    if (l.kind == 1) return false;
  if (sc.lights.empty() && sc.points.empty()) return false;
  if (job.width * (uint64_t)job.height < 2) return false;
  return p.max_depth >= 2 && p.chains > 0;
}

MltParams mlt_params(const orc_mlt &p) {
  MltParams r;
  r.chains = p.chains;
  r.mutations = p.mutations;
  r.seed_paths = p.seed_paths;
  r.max_depth = p.max_depth;
  r.max_rejections = p.max_rejections;
  r.filter_proposed = p.filter_proposed;
  for (int i = 0; i < 3; ++i) r.probs[i] = p.probs[i];
  return r;
}

}  // namespace

extern "C" {

int orc_mlt_seeds(const orc_scene *s, const orc_job *job, const orc_mlt *p, double *b_out, uint64_t *n_out,
                  double *w_out, uint32_t *len_out, uint64_t cap, int nthreads) {
  const Camera cam(job->camera, s);
  if (!mlt_supported(*s, cam, *job, *p)) return -1;
  PathCounters ctr;
  MltCtx cx{*s, cam, job->width, job->height, 1.0 / job->width, 1.0 / job->height, ctr};
  SeedSet S;
  if (!init_seeds(cx, *job, mlt_params(*p), S, nthreads)) return -2;
  *b_out = S.b;
  *n_out = S.w.size();
  for (uint64_t i = 0; i < std::min<uint64_t>(cap, S.w.size()); ++i) {
    if (w_out) w_out[i] = S.w[i];
    if (len_out) len_out[i] = S.paths[i].len();
  }
  return 0;
}

int orc_mlt_render(const orc_scene *s, const orc_job *job, const orc_mlt *p, double *film_out, double *num_out,
                   uint64_t *proposed_out, uint64_t *stats, double *b_out, int nthreads) {
  const Camera cam(job->camera, s);
  if (!mlt_supported(*s, cam, *job, *p)) return -1;
  const MltParams mp = mlt_params(*p);
  PathCounters ctr;
  MltCtx cx{*s, cam, job->width, job->height, 1.0 / job->width, 1.0 / job->height, ctr};
  SeedSet S;
  if (!init_seeds(cx, *job, mp, S, nthreads)) return -2;
  const Mutations mu(mp);
  const uint32_t P = job->width * job->height;
  const int nt = std::max(1, nthreads);
  std::vector<MltFilm> films(nt);
  std::vector<PathCounters> ctrs(nt);
  std::vector<uint64_t> acc(nt, 0);
  for (MltFilm &f : films) {
    f.num.assign(3ull * P, 0.0);
    f.proposed.assign(P, 0);
  }
  parallel_for(mp.chains, nthreads, 1, [&](uint64_t b, uint64_t e, int tid) {
    MltCtx tc{*s, cam, job->width, job->height, 1.0 / job->width, 1.0 / job->height, ctrs[tid]};
    for (uint64_t c = b; c < e; ++c) {
      Chain ch;
      chain_start(tc, S, (uint32_t)c, *job, ch);
      for (uint32_t j = 0; j < mp.mutations; ++j) {
        StepRec r;
        chain_step(tc, mp, mu, S.b, ch, (uint32_t)c, j, *job, films[tid], &r);
        acc[tid] += r.accepted;
      }
    }
  });
  MltFilm f;
  f.num.assign(3ull * P, 0.0);
  f.proposed.assign(P, 0);
  uint64_t accepted = 0;
  for (int t = 0; t < nt; ++t) {
    for (uint64_t i = 0; i < 3ull * P; ++i) f.num[i] += films[t].num[i];
    for (uint32_t i = 0; i < P; ++i) f.proposed[i] += films[t].proposed[i];
    f.samples += films[t].samples;
    ctr.extension += ctrs[t].extension;
    ctr.shadow += ctrs[t].shadow;
    accepted += acc[t];
  }
  finalize_film(f, job->width, job->height, mp.filter_proposed != 0, film_out);
  if (num_out) std::copy(f.num.begin(), f.num.end(), num_out);
  if (proposed_out) std::copy(f.proposed.begin(), f.proposed.end(), proposed_out);
  if (stats) {
    stats[0] = f.samples;
    stats[1] = accepted;
    stats[2] = (uint64_t)mp.chains * mp.mutations;
    stats[3] = ctr.extension;
    stats[4] = ctr.shadow;
    stats[5] = S.w.size();
  }
  if (b_out) *b_out = S.b;
  return 0;
}

int orc_mlt_trace(const orc_scene *s, const orc_job *job, const orc_mlt *p, uint32_t c0, uint32_t c1, uint32_t steps,
                  double *rec, int nthreads) {
  const Camera cam(job->camera, s);
  if (!mlt_supported(*s, cam, *job, *p) || c1 < c0) return -1;
  const MltParams mp = mlt_params(*p);
  PathCounters ctr;
  MltCtx cx{*s, cam, job->width, job->height, 1.0 / job->width, 1.0 / job->height, ctr};
  SeedSet S;
  if (!init_seeds(cx, *job, mp, S, nthreads)) return -2;
  const Mutations mu(mp);
  const uint32_t P = job->width * job->height;
  const int nt = std::max(1, nthreads);
  std::vector<MltFilm> films(nt);
  std::vector<PathCounters> ctrs(nt);
  for (MltFilm &f : films) {
    f.num.assign(3ull * P, 0.0);
    f.proposed.assign(P, 0);
  }
  parallel_for(c1 - c0, nthreads, 1, [&](uint64_t b, uint64_t e, int tid) {
    MltCtx tc{*s, cam, job->width, job->height, 1.0 / job->width, 1.0 / job->height, ctrs[tid]};
    for (uint64_t i = b; i < e; ++i) {
      const uint32_t c = c0 + (uint32_t)i;
      Chain ch;
      chain_start(tc, S, c, *job, ch);
      for (uint32_t j = 0; j < steps; ++j) {
        StepRec r;
        chain_step(tc, mp, mu, S.b, ch, c, j, *job, films[tid], &r);
        double *o = rec + 6 * ((uint64_t)i * steps + j);
        o[0] = r.mutation;
        o[1] = r.alpha;
        o[2] = r.accepted;
        o[3] = r.y_len;
        o[4] = r.tent_bin;
        o[5] = r.x_len;
      }
    }
  });
  return 0;
}

}  // extern "C"
