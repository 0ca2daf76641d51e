// Device functions of the path-tracing core (gfx950).
//
// Traversal replaces kdTreeAccel::getIntersection / intersects
// (accel/kdTreeAccel.cpp:893-1094) over the flattened layout of kdbuild.h.
// Intersection replaces MeshTriangle::getIntersection (shapes/MeshTriangle.cpp:
// 92-169), Sphere (shapes/Sphere.cpp:16-73) and Plane (shapes/Plane.cpp:25-49).
// Shading replaces SurfacePoint::init + Material::initSurfacePoint
// (core/SurfacePoint.cpp:49-78, materials/Material.cpp:157-177) and the BSDFs
// (materials/bsdfs/DiffuseBSDF.cpp, ModifiedPhongBSDF.cpp, DielectricBSDF.cpp).
#ifndef MPT_DEVICE_H_
#define MPT_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/milton_pt.h"

namespace mpt_dev {

// Traversal and intersection compile for the host too, so the host-side
// trace entry point runs exactly the code of the kernels.
#define MPT_HD __host__ __device__ __forceinline__

MPT_HD float mpt_u2f(uint32_t u) {
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}
MPT_HD uint32_t mpt_f2u(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  return u;
}

constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr float kEps = 1e-7f;            // EPSILON, common/arch.h:137
constexpr float kAirIor = 1.0003f;       // IndexOfRefraction::AIR
constexpr int kBlock = 256;              // threads per block of every traversal kernel
constexpr int kStack = 16;               // default short-stack entries per lane (LDS ring)
constexpr float kPi = 3.14159265358979323846f;

enum { KIND_TRIANGLE = 0, KIND_TRIANGLE_ONE_SIDED = 1, KIND_SPHERE = 2, KIND_QUAD = 3 };   // triangles: kind < 2

// An area light of next-event estimation (DirectIllumination.cpp:90-135 over
// Scene::getLights).  kind: LIGHT_QUAD (index = quad), LIGHT_SPHERE (index =
// sphere; frame = the linear part of its object-to-world transform, columns),
// LIGHT_MESH (triangles light_tris[first, first + count)).  area: the
// shape's getSurfaceArea; mat: its emitting material.
enum { LIGHT_QUAD = 0, LIGHT_SPHERE = 1, LIGHT_MESH = 2 };
struct LightDev {
  uint32_t kind, index, first, count;
  float area;
  uint32_t mat, pad[2];
  float4 frame[3];
};

// Read-only scene as seen by the kernels.
struct Scene {
  const uint2 *nodes;        // kd nodes (kdbuild.h)
  const float4 *recs;        // 3 float4 per leaf reference
  float box_lo[3], box_hi[3];
  const float4 *tshade;      // 3 per triangle: vertex normals (n0, material bits), (n1, area of the emitting
                             // shape or 0), (n2, its light index + 1 or 0)
  const float4 *spheres;     // (centre, r)
  const uint32_t *sphere_mat;
  const float4 *quads;       // 4 per quad: (c, area), (ex, 0), (ez, 0), (n, material bits)
  const mpt_material *mats;
  uint32_t top_bytes;        // the tree's top: node bytes [0, top_bytes), which a kernel may stage in LDS
  const struct LightDev *lights;   // area lights: emitting quads, spheres, meshes
  uint32_t num_lights;
  const float4 *light_tris;  // 3 per triangle of an emitting mesh: (v0, triangle id) (v1, 0) (v2, 0)
  uint32_t nt, ns, nq;
  float env[3];
  // environment image (env_w 0: the constant env): texels (r, g, b, 0) from
  // the top row, the filter kernel (env_m x env_m weights, then their sum)
  const float4 *env_tex;
  const double *env_kernel;
  uint32_t env_w, env_h, env_m;
  float env_rep_u, env_rep_v;
  // point lights: (position, 0), (radiance P / (4 pi), 0) per light
  const float4 *points;
  uint32_t num_points;
  // emitted power of the area lights then the point lights, normalized:
  // (sum through this light, this light's share) per light
  const float2 *light_sel;
  float emit_area;   // the area lights' summed surface area
};

MPT_HD float3 f3(float x, float y, float z) { return make_float3(x, y, z); }
MPT_HD float3 add(float3 a, float3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
MPT_HD float3 sub(float3 a, float3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
MPT_HD float3 mul(float3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
MPT_HD float3 mul(float3 a, float3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
MPT_HD float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MPT_HD float3 cross(float3 a, float3 b) {
  return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float3 normalize(float3 a) { return mul(a, rsqrtf(dot(a, a))); }
MPT_HD float len(float3 a) { return sqrtf(dot(a, a)); }
MPT_HD float3 xyz(float4 a) { return f3(a.x, a.y, a.z); }
MPT_HD float comp(float3 a, uint32_t i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// a / b through the hardware reciprocal (within 2 ulp): for the FP32 hit
// distances of the traversal, which only rank candidates and are recomputed
// in double for the hit that wins.
MPT_HD float div_approx(float a, float b) {
#ifdef __HIP_DEVICE_COMPILE__
  return a * __builtin_amdgcn_rcpf(b);
#else
  return a * (1.0f / b);
#endif
}
MPT_HD bool is_black(float3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
MPT_HD float max3(float3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }

// Per-ray constants: origin, direction, its reciprocal for the kd planes,
// and the shear of the watertight triangle test (Woop, Benthin, Wald, JCGT
// 2013): the axis permutation (kx, ky, kz packed in k) and the factors sx,
// sy, sz.
struct RaySetup {
  float3 o, d, inv;
  float sx, sy, sz;
  uint32_t k;   // kx | ky << 2 | kz << 4
};

// 1/x with |x| clamped to 1e-20 so axis-parallel rays stay finite.  Both
// zeros map to -1e20: inv < 0 then means d <= 0, which is what the kd walk
// uses to order the children of a plane the ray starts on.
MPT_HD float safe_inv(float x) {
  return x > 0.0f ? 1.0f / fmaxf(x, 1e-20f) : -1.0f / fmaxf(-x, 1e-20f);
}

MPT_HD RaySetup setup_ray(float3 o, float3 d) {
  RaySetup r;
  r.o = o;
  r.d = d;
  r.inv = f3(safe_inv(d.x), safe_inv(d.y), safe_inv(d.z));
  const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
  uint32_t kz = (ax >= ay && ax >= az) ? 0 : (ay >= az ? 1 : 2);
  uint32_t kx = kz == 2 ? 0 : kz + 1;
  uint32_t ky = kx == 2 ? 0 : kx + 1;
  const float dz = comp(d, kz);
  if (dz < 0.0f) {
    const uint32_t t = kx;
    kx = ky;
    ky = t;
  }
  r.sx = comp(d, kx) / dz;
  r.sy = comp(d, ky) / dz;
  r.sz = 1.0f / dz;
  r.k = kx | (ky << 2) | (kz << 4);
  return r;
}

// Two-sided watertight triangle test.  Matches the reference's acceptance
// rules: edges and vertices count as hits (inclusive u, v, u + v <= 1), and a
// triangle whose Moller-Trumbore determinant |d . ((v1-v0) x (v2-v0))| is at
// most EPSILON is rejected (EQ(det, 0), MeshTriangle.cpp:144).  The sheared
// determinant U + V + W equals that product divided by d[kz] = 1 / sz.
//
// tri_td leaves the division to the caller: on a hit it returns the scaled
// distance T and determinant det > 0 with t = T / det, so a leaf loop can
// reject candidates beyond its current best without dividing.
// Edge functions U, V, W of the sheared triangle (U weights v0, V v1, W v2)
// and the sheared z of its vertices.
struct TriEdges {
  float U, V, W, Az, Bz, Cz;
};

MPT_HD TriEdges tri_edges(const RaySetup &r, float3 v0, float3 v1, float3 v2) {
  const uint32_t kx = r.k & 3u, ky = (r.k >> 2) & 3u, kz = r.k >> 4;
  const float3 A = sub(v0, r.o), B = sub(v1, r.o), C = sub(v2, r.o);
  TriEdges e;
  e.Az = comp(A, kz);
  e.Bz = comp(B, kz);
  e.Cz = comp(C, kz);
  const float Ax = comp(A, kx) - r.sx * e.Az, Ay = comp(A, ky) - r.sy * e.Az;
  const float Bx = comp(B, kx) - r.sx * e.Bz, By = comp(B, ky) - r.sy * e.Bz;
  const float Cx = comp(C, kx) - r.sx * e.Cz, Cy = comp(C, ky) - r.sy * e.Cz;
  e.U = Cx * By - Cy * Bx;
  e.V = Ax * Cy - Ay * Cx;
  e.W = Bx * Ay - By * Ax;
  return e;
}

// one_sided: the reference's standalone Triangle, which culls back faces: it
// rejects e1 . (d x e2) < EPSILON (Triangle.cpp:245-270), and that product
// is (U + V + W) * |d[kz]| = (U + V + W) / |sz|.
// Everything is computed unconditionally and the tests are combined without
// branches: a leaf loop over triangles then runs straight-line code.
MPT_HD bool tri_td(const RaySetup &r, float3 v0, float3 v1, float3 v2, float &T, float &det,
                   bool one_sided = false) {
  const TriEdges e = tri_edges(r, v0, v1, v2);
  const float U = e.U, V = e.V, W = e.W;
  const bool mixed = (fminf(U, fminf(V, W)) < 0.0f) & (fmaxf(U, fmaxf(V, W)) > 0.0f);
  const float d = U + V + W;
  const bool flat = one_sided ? d < kEps * fabsf(r.sz) : fabsf(d) <= kEps * fabsf(r.sz);
  const float t = (U * e.Az + V * e.Bz + W * e.Cz) * r.sz;
  // (-T) / (-det) == T / det exactly, so the sign moves onto T
  T = d < 0.0f ? -t : t;
  det = fabsf(d);
  return !(mixed | flat);
}

// Sphere (centre c, radius rad): nearest root above EPSILON; a ray leaving the
// sphere it starts on skips the root at its own origin.
MPT_HD float sphere_t(const RaySetup &r, float3 c, float rad, bool from_self) {
  const float3 f = sub(r.o, c);
  const float b = -dot(f, r.d);
  const float3 l = add(f, mul(r.d, b));
  const float disc = rad * rad - dot(l, l);
  if (disc < 0.0f) return INFINITY;
  const float cc = dot(f, f) - rad * rad;
  const float q = b + copysignf(sqrtf(disc), b);
  float t0 = q != 0.0f ? cc / q : 0.0f, t1 = q;
  if (t0 > t1) {
    const float t = t0;
    t0 = t1;
    t1 = t;
  }
  const float lo = from_self ? fmaxf(kEps, 1e-3f * rad) : kEps;
  if (t0 > lo) return t0;
  if (t1 > lo) return t1;
  return INFINITY;
}

// Quad: the reference unit plane mapped by (ex, ez) around c, two-sided.
MPT_HD float quad_t(const RaySetup &r, float3 c, float3 ex, float3 ez) {
  const float3 n = cross(ez, ex);
  const float dn = dot(n, r.d);
  const float t = dot(n, sub(c, r.o)) / dn;
  const float3 p = sub(add(r.o, mul(r.d, t)), c);
  const float u = dot(p, ex) / dot(ex, ex), v = dot(p, ez) / dot(ez, ez);
  return (fabsf(u) <= 0.5f && fabsf(v) <= 0.5f) ? t : INFINITY;
}

// Recomputes the closest hit's t in double precision (one plane or sphere
// solve per ray), so t matches the reference's FP64 intersect
// (MeshTriangle.cpp:92-169 computes t in double) even for hits very close to
// the ray origin, where the FP32 edge-function t loses relative accuracy.
// With uv, a triangle hit also gets its barycentric weights of v1 and v2 at
// the refined hit point (the weight of v0 is 1 minus both); for a point on
// the triangle they equal the reference's ratios of unsigned sub-triangle
// areas (MeshTriangle.cpp:344-361).  Other kinds report (0, 0).
MPT_HD float refine_hit(const float4 *rec, float3 o, float3 d, float t32, float2 *uv = nullptr) {
  const float4 a = rec[0], b = rec[1], c = rec[2];
  const uint32_t kind = mpt_f2u(b.w);
  const double ox = o.x, oy = o.y, oz = o.z, dx = d.x, dy = d.y, dz = d.z;
  if (uv) *uv = make_float2(0.0f, 0.0f);
  if (kind == KIND_SPHERE) {
    const double fx = ox - a.x, fy = oy - a.y, fz = oz - a.z, rad = b.x;
    const double bb = fx * dx + fy * dy + fz * dz, dd = dx * dx + dy * dy + dz * dz;
    const double cc = fx * fx + fy * fy + fz * fz - rad * rad;
    const double disc = bb * bb - dd * cc;
    if (disc < 0) return t32;
    const double sq = sqrt(disc);
    const double t0 = (-bb - sq) / dd, t1 = (-bb + sq) / dd;
    // the root the FP32 test picked
    return (float)(fabs(t0 - t32) <= fabs(t1 - t32) ? t0 : t1);
  }
  double e1x, e1y, e1z, e2x, e2y, e2z;
  const bool tri = kind < KIND_SPHERE;
  if (tri) {
    e1x = (double)b.x - a.x; e1y = (double)b.y - a.y; e1z = (double)b.z - a.z;
    e2x = (double)c.x - a.x; e2y = (double)c.y - a.y; e2z = (double)c.z - a.z;
  } else {   // quad: plane through the centre spanned by ex, ez
    e1x = b.x; e1y = b.y; e1z = b.z;
    e2x = c.x; e2y = c.y; e2z = c.z;
  }
  const double nx = e1y * e2z - e1z * e2y, ny = e1z * e2x - e1x * e2z, nz = e1x * e2y - e1y * e2x;
  const double dn = dx * nx + dy * ny + dz * nz;
  if (dn == 0.0) return t32;
  const double t = (((double)a.x - ox) * nx + ((double)a.y - oy) * ny + ((double)a.z - oz) * nz) / dn;
  if (uv && tri) {
    // p - v0 = b1 e1 + b2 e2  =>  b1 = ((p - v0) x e2).n / n.n, b2 = (e1 x (p - v0)).n / n.n
    const double wx = ox + t * dx - a.x, wy = oy + t * dy - a.y, wz = oz + t * dz - a.z;
    const double nn = nx * nx + ny * ny + nz * nz;
    const double b1 = ((wy * e2z - wz * e2y) * nx + (wz * e2x - wx * e2z) * ny + (wx * e2y - wy * e2x) * nz) / nn;
    const double b2 = ((e1y * wz - e1z * wy) * nx + (e1z * wx - e1x * wz) * ny + (e1x * wy - e1y * wx) * nz) / nn;
    *uv = make_float2((float)b1, (float)b2);
  }
  return (float)t;
}

// kd traversal with a ring stack in LDS and kd-restart on overflow (Foley &
// Sugerman 2005).  `stk` points at this lane's column of a [STACK][kBlock]
// uint2 array: a ring of STACK rows that holds at most STACK - 1 entries, so
// the row after the newest entry is always free and every descent step can
// store into it whether or not it pushes (BLOCK: the threads of the block,
// one column each).  skip_origin is the primitive the
// ray leaves (self-intersection guard in place of the reference's FP64
// t > EPSILON); skip_target the triangle or quad a shadow ray aims at (its
// only crossing of that primitive is the end point, which the reference's
// FP64 shadow window t < tMax - 2 EPSILON leaves out).  ANY=true finishes at
// the first hit.
// COUNT=true adds the nodes fetched and references tested to n_nodes/n_refs.
//
// The walk is resumable: start() clips the ray to the scene box, each step()
// descends to the next leaf, tests its references and pops the next cell, and
// returns true when the ray is finished.  Splitting the walk into leaf steps
// lets a persistent wave hand a finished lane a new ray while its other lanes
// are still walking.
//
// Node references and ring positions are byte offsets, so the descent does
// no index arithmetic: a child is (word0 & ~3) + 8 * side, a ring row is
// row * kRow.
template <bool ANY, int STACK, int BLOCK = kBlock>
struct Walker {
  static constexpr uint32_t kRow = BLOCK * sizeof(uint2);            // bytes per ring row
  static constexpr uint32_t kCap = (STACK - 1) * kRow;               // entries held, in bytes
  static constexpr bool kPow2 = (STACK & (STACK - 1)) == 0;
  float tmin, tmax, troot, best;
  uint32_t best_rec, node, skip_origin;   // best_rec == kMiss: no hit yet; node: byte offset
  uint32_t skip_target;
  uint32_t depth, top;   // entries held and next write row, both as row * kRow

  MPT_HD bool start(const Scene &sc, const RaySetup &r, float tlimit, uint32_t so, uint32_t st = kMiss) {
    // Root slab test (accel/AABB.inl:25-56).
    const float ix = r.inv.x, iy = r.inv.y, iz = r.inv.z;
    const float tx0 = (sc.box_lo[0] - r.o.x) * ix, tx1 = (sc.box_hi[0] - r.o.x) * ix;
    const float ty0 = (sc.box_lo[1] - r.o.y) * iy, ty1 = (sc.box_hi[1] - r.o.y) * iy;
    const float tz0 = (sc.box_lo[2] - r.o.z) * iz, tz1 = (sc.box_hi[2] - r.o.z) * iz;
    tmin = fmaxf(0.0f, fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1))));
    tmax = fminf(tlimit, fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1))));
    // widen by a few ulps so rays grazing the root box still enter it
    tmin = fmaxf(0.0f, tmin * (1.0f - 1e-6f) - 1e-7f);
    tmax = tmax * (1.0f + 1e-6f) + 1e-7f;
    troot = tmax;
    best = tlimit;
    best_rec = kMiss;
    node = 0;
    depth = 0;
    top = 0;
    skip_origin = so;
    skip_target = st;
    return tmin <= tmax;
  }

  // TOP: node bytes below sc.top_bytes are read from `top_lds`, the
  // kernel's LDS copy of the top of the tree.
  template <bool COUNT, bool TOP = false>
  MPT_HD bool step(const Scene &sc, const RaySetup &r, uint2 *stk, uint32_t *n_nodes,
                                       uint32_t *n_refs, const uint2 *top_lds = nullptr) {
    const char *nodes = reinterpret_cast<const char *>(sc.nodes);
    const char *tops = reinterpret_cast<const char *>(top_lds);
    char *ring = reinterpret_cast<char *>(stk);
    auto fetch = [&](uint32_t off) {
      // the LDS copy is read as one 64-bit word: a load of another type than
      // the global one keeps the compiler from merging the two into a
      // generic-address load
      if (TOP && off < sc.top_bytes) {
        const unsigned long long v = *reinterpret_cast<const unsigned long long *>(tops + off);
        return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
      }
      return *reinterpret_cast<const uint2 *>(nodes + off);
    };
    uint2 nd = fetch(node);
    if (COUNT) ++*n_nodes;
    while ((nd.x & 3u) != 3u) {
      const uint32_t axis = nd.x & 3u;
      const float split = mpt_u2f(nd.y);
      const float oa = comp(r.o, axis), ia = comp(r.inv, axis);
      const float tp = (split - oa) * ia;
      // Near child first: the left one when the ray runs up the axis (ia > 0),
      // the right one otherwise (a zero direction has ia = -1e20).  Whatever
      // side the origin is on, the cell's segment [tmin, tmax] then lies
      // wholly in the far child when tp < tmin (this covers a plane behind
      // the origin), wholly in the near child when tp > tmax, and crosses
      // from near to far at tp otherwise.
      const uint32_t first = (nd.x & ~3u) | ((mpt_f2u(ia) >> 31) << 3);
      const uint32_t second = first ^ 8u;
      const bool far_only = tp < tmin;
      const bool near_only = tp > tmax;
      const bool both = !far_only && !near_only;
      // Branch-free push: the far cell always goes into the free row; only
      // when the ray spans both children does the ring advance over it.
      *reinterpret_cast<uint2 *>(ring + top) = make_uint2(second, mpt_f2u(tmax));
      const uint32_t inc = both ? kRow : 0u;
      const uint32_t held = depth + inc;
      depth = held < kCap ? held : kCap;
      if (kPow2) {
        top = (top + inc) & (STACK * kRow - 1);
      } else {
        top = top + inc == STACK * kRow ? 0u : top + inc;
      }
      tmax = both ? tp : tmax;
      node = far_only ? second : first;
      nd = fetch(node);
      if (COUNT) ++*n_nodes;
    }
    const uint32_t count = nd.x >> 2;
    if (COUNT) *n_refs += count;
    const float4 *rec = sc.recs + 3ull * nd.y;
    for (uint32_t k = 0; k < count; ++k, rec += 3) {
      // the whole 48-byte record in one go (every kind has three float4s)
      const float4 a = rec[0], b = rec[1], c = rec[2];
      const uint32_t id = mpt_f2u(a.w);
      const uint32_t kind = mpt_f2u(b.w);
      float t = INFINITY;
      if (kind < KIND_SPHERE) {
        float T, det;
        // divide only for candidates that can beat the current best; the
        // 2^-20 margins keep the prefilter conservative, the exact test
        // below decides
        const bool hit = tri_td(r, xyz(a), xyz(b), xyz(c), T, det, kind == KIND_TRIANGLE_ONE_SIDED);
        if (hit & (id != skip_origin) & (id != skip_target) & (T <= best * det * (1.0f + 0x1p-20f)) &
            (T >= kEps * det * (1.0f - 0x1p-20f)))
          t = div_approx(T, det);
      } else if (kind == KIND_SPHERE) {
        t = sphere_t(r, xyz(a), b.x, id == skip_origin);
      } else if (id != skip_origin && id != skip_target) {
        t = quad_t(r, xyz(a), xyz(b), xyz(c));
      }
      if (t > kEps && t < best) {
        best = t;
        best_rec = nd.y + k;
        if (ANY) return true;
      }
    }
    if (best <= tmax) return true;
    if (depth == 0) {
      // An empty stack ends the walk at the root cell's exit: the oldest
      // entry holds troot itself unless the ring overwrote it, and with no
      // entry lost tmax is back at troot here.  Otherwise restart from the
      // root past the cells already visited.
      if (tmax >= troot) return true;
      node = 0;
      tmin = tmax;
      tmax = troot;
      return false;
    }
    depth -= kRow;
    if (kPow2) {
      top = (top - kRow) & (STACK * kRow - 1);
    } else {
      top = top == 0 ? (STACK - 1) * kRow : top - kRow;
    }
    const uint2 e = *reinterpret_cast<const uint2 *>(ring + top);
    node = e.x;
    tmin = tmax;
    tmax = mpt_u2f(e.y);
    return tmin > best;
  }

  // Result of a finished walk: the closest hit's t is recomputed in double.
  // The closest hit as found: the FP32 t and the leaf record, or kMiss.
  MPT_HD void raw(float &t32, uint32_t &rec) const {
    t32 = best_rec == kMiss ? INFINITY : best;
    rec = best_rec;
  }

  // Result of a finished walk: the closest hit's t recomputed in double and
  // the primitive id (refine_hit); with uv also its barycentrics.
  MPT_HD bool result(const Scene &sc, const RaySetup &r, float &t_out, uint32_t &id_out,
                     float2 *uv = nullptr) const {
    if (best_rec == kMiss) {
      t_out = INFINITY;
      id_out = kMiss;
      if (uv) *uv = make_float2(0.0f, 0.0f);
      return false;
    }
    const float4 *rec = sc.recs + 3ull * best_rec;
    t_out = ANY ? best : refine_hit(rec, r.o, r.d, best, uv);
    id_out = mpt_f2u(rec[0].w);
    return true;
  }
};

// kd traversal of one ray to completion (ANY: first hit below tlimit).
template <bool ANY, bool COUNT = false, int STACK = kStack, int BLOCK = kBlock>
MPT_HD bool trace(const Scene &sc, const RaySetup &r, float tlimit, uint32_t skip_origin, uint2 *stk,
                  float &t_out, uint32_t &id_out, uint32_t *n_nodes = nullptr, uint32_t *n_refs = nullptr,
                  float2 *uv = nullptr, uint32_t skip_target = kMiss) {
  Walker<ANY, STACK, BLOCK> w;
  if (!w.start(sc, r, tlimit, skip_origin, skip_target)) {
    t_out = INFINITY;
    id_out = kMiss;
    if (uv) *uv = make_float2(0.0f, 0.0f);
    return false;
  }
  while (!w.template step<COUNT>(sc, r, stk, n_nodes, n_refs)) {
  }
  return w.result(sc, r, t_out, id_out, uv);
}

// ---------------------------------------------------------------------------
// Shading.
// ---------------------------------------------------------------------------
// Vector::getOrthonormalBasis (common/math/Vector.inl:527-538).
__device__ __forceinline__ void onb(float3 n, float3 &U, float3 &V) {
  const float3 up = fabsf(n.y) > 0.8f ? f3(1, 0, 0) : f3(0, 1, 0);
  U = normalize(cross(up, n));
  V = normalize(cross(n, U));
}

// Vector3::cosRandom (common/math/Vector.cpp:21-29) through convertHemisphere
// (common/math/sisd/Vector3.inl:18-22): theta = acos(sqrt(u1)), phi = 2 pi u2.
// The hardware sine/cosine take revolutions, so phi enters as u2 directly.
__device__ __forceinline__ float3 cos_random(float3 n, float u1, float u2) {
  const float ct = sqrtf(u1), st = sqrtf(fmaxf(0.0f, 1.0f - u1));
  const float cp = __builtin_amdgcn_cosf(u2), sp = __builtin_amdgcn_sinf(u2);
  float3 U, V;
  onb(n, U, V);
  return add(add(mul(U, st * cp), mul(n, ct)), mul(V, st * sp));
}

__device__ __forceinline__ float3 reflect(float3 w, float3 n) { return add(mul(n, -2.0f * dot(w, n)), w); }

// Vector::refractVector (common/math/Vector.inl:468-494).
__device__ __forceinline__ bool refract(float3 v, float3 normal, float in, float out, float3 &wt) {
  float3 norm;
  float n;
  if (dot(v, normal) < 0.0f) {
    norm = normal;
    n = in / out;
  } else {
    norm = mul(normal, -1.0f);
    n = out / in;
  }
  const float3 w = mul(v, -1.0f);
  const float d = dot(w, norm);
  const float det = 1.0f - n * n * (1.0f - d * d);
  if (det < 0.0f) return false;
  wt = add(mul(w, -n), mul(norm, n * d - sqrtf(det)));
  return true;
}

// Fresnel::getFr (materials/bsdfs/Fresnel.h:45-90), unpolarised.
__device__ __forceinline__ float fresnel(float3 wi, float3 N, float n1, float n2) {
  float ci = dot(N, wi);
  if (ci < 0.0f) {
    ci = -ci;
  } else {
    const float t = n1;
    n1 = n2;
    n2 = t;
  }
  const float snell = (n1 / n2) * sqrtf(fmaxf(0.0f, 1.0f - ci * ci));
  if (fabsf(snell) >= 1.0f) return 1.0f;
  const float ct = sqrtf(fmaxf(0.0f, 1.0f - snell * snell));
  const float Rs = (n1 * ci - n2 * ct) / (n1 * ci + n2 * ct);
  const float Rp = (n1 * ct - n2 * ci) / (n1 * ct + n2 * ci);
  return (Rs * Rs + Rp * Rp) * 0.5f;
}

__device__ __forceinline__ float3 kd_of(const mpt_material &m) { return f3(m.kd[0], m.kd[1], m.kd[2]); }
__device__ __forceinline__ float3 ks_of(const mpt_material &m) { return f3(m.ks[0], m.ks[1], m.ks[2]); }
__device__ __forceinline__ float3 le_of(const mpt_material &m) { return f3(m.Le[0], m.Le[1], m.Le[2]); }
__device__ __forceinline__ float avg3(float3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }
__device__ __forceinline__ bool emissive(const mpt_material &m) { return m.Le[0] > 0 || m.Le[1] > 0 || m.Le[2] > 0; }

// f(wi, wo) of diffuse (DiffuseBSDF.cpp:40-48) and modified Phong
// (ModifiedPhongBSDF.cpp:164-181); zero for the dielectric's delta lobes.
__device__ __forceinline__ float3 bsdf_eval(const mpt_material &m, float3 n, float3 wi, float3 wo) {
  if (m.type == MPT_BSDF_DIFFUSE) {
    if (dot(n, wi) >= 0.0f || dot(n, wo) <= 0.0f) return f3(0, 0, 0);
    return mul(kd_of(m), 1.0f / kPi);
  }
  if (m.type == MPT_BSDF_PHONG) {
    if (dot(n, wo) <= 0.0f) return f3(0, 0, 0);
    const float a = fminf(1.0f, fmaxf(0.0f, dot(reflect(wi, n), wo)));
    const float kda = avg3(kd_of(m)), ksa = avg3(ks_of(m));
    const float an = a > 0.0f ? exp2f(m.exponent * log2f(a)) : 0.0f;
    const float fs_s = ksa > 0.0f ? (m.exponent + 2.0f) / (2.0f * kPi * ksa) * an : 0.0f;
    const float fs_d = kda > 0.0f ? 1.0f / (kPi * kda) : 0.0f;
    return add(mul(kd_of(m), fs_d), mul(ks_of(m), fs_s));
  }
  return f3(0, 0, 0);
}

struct Sample {
  float3 wo;   // zero if absorbed
  float3 fs;   // f / pdf
};

// BSDF sampling; u = (u0, u1) direction, u3 lobe choice (Phong) or
// reflect-vs-refract (dielectric).
__device__ __forceinline__ Sample bsdf_sample(const mpt_material &m, float3 n, float3 wi, float u0, float u1,
                                              float u3, uint32_t ior_index) {
  Sample s;
  s.wo = f3(0, 0, 0);
  s.fs = f3(0, 0, 0);
  if (m.type == MPT_BSDF_DIFFUSE) {
    // DiffuseBSDF::sample / getPd (DiffuseBSDF.cpp:21-38): f/pdf = kd.
    if (dot(n, wi) > 0.0f) return s;
    const float3 wo = cos_random(n, u0, u1);
    if (dot(n, wo) <= 0.0f) return s;
    s.wo = wo;
    s.fs = mul(bsdf_eval(m, n, wi, wo), kPi);
    return s;
  }
  if (m.type == MPT_BSDF_PHONG) {
    // ModifiedPhongBSDF::sample / getPd (ModifiedPhongBSDF.cpp:84-162).
    const float kda = avg3(kd_of(m)), ksa = avg3(ks_of(m));
    const float p0 = u3 - kEps;
    float3 wo;
    bool spec;
    if (p0 <= kda) {
      spec = false;
      wo = cos_random(n, u0, u1);
    } else if (p0 <= kda + ksa) {
      spec = true;
      const float ca = exp2f(log2f(u0) / (m.exponent + 1.0f));
      const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
      const float cp = __builtin_amdgcn_cosf(u1), sp = __builtin_amdgcn_sinf(u1);
      const float3 R = reflect(wi, n);
      float3 rr = R, uu, vv;
      if (fabsf(dot(n, rr)) < 0.95f) {
        uu = normalize(cross(rr, n));
        vv = cross(uu, rr);
      } else {
        rr = normalize(rr);
        onb(rr, vv, uu);
      }
      const float3 w = add(add(mul(uu, sa * cp), mul(rr, ca)), mul(vv, sa * sp));
      wo = dot(n, w) > 0.0f ? w : R;
    } else {
      return s;
    }
    const float cosA = dot(n, wo);
    if (cosA <= 0.0f) return s;
    float pdf;
    if (spec) {
      const float c = fabsf(dot(reflect(wi, n), wo));
      const float cn = c > 0.0f ? exp2f(m.exponent * log2f(c)) : 0.0f;
      pdf = fminf(1.0f, (m.exponent + 1.0f) / (2.0f * kPi) * cn);
    } else {
      pdf = fabsf(cosA) / kPi;
    }
    pdf = pdf / cosA;
    if (!(pdf > 0.0f)) return s;
    s.wo = wo;
    s.fs = mul(bsdf_eval(m, n, wi, wo), 1.0f / pdf);
    return s;
  }
  // DielectricBSDF::sample / getPd / evaluate (DielectricBSDF.cpp:55-158):
  // pdf 1, f = ks, and a refraction keeps only the path's IOR channel.
  const float3 N = dot(n, wi) > 0.0f ? mul(n, -1.0f) : n;
  // the path's channel of a dispersive IOR (DielectricBSDF.cpp:65-87)
  const float ior = ior_index == 1 && m.ior_g > 0.0f ? m.ior_g : (ior_index == 2 && m.ior_b > 0.0f ? m.ior_b : m.ior);
  float3 wt;
  bool refracted = false;
  if (refract(wi, n, kAirIor, ior, wt)) {
    const float Fr = fresnel(wi, n, kAirIor, ior);
    const float Fs = m.transparency * (1.0f - Fr);
    refracted = u3 < Fs;
  }
  if (refracted) {
    s.wo = wt;
    s.fs = f3(ior_index == 0 ? m.ks[0] : 0.0f, ior_index == 1 ? m.ks[1] : 0.0f, ior_index == 2 ? m.ks[2] : 0.0f);
  } else {
    s.wo = reflect(wi, N);
    s.fs = ks_of(m);
  }
  return s;
}

}  // namespace mpt_dev

#endif
