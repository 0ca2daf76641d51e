// MI355X path-tracing core: wavefront kernels and the C ABI (include/milton_pt.h).
//
// One render job is processed in batches of whole pixels (all of a pixel's
// samples are in one batch, so the film needs no atomics and is
// deterministic).  Per batch:
//   k_raygen    camera rays of every (pixel, sample) slot
//   per depth:  k_extend  closest hit of the live ray queue
//               k_shade   emission, NEE (shadow-ray queue), BSDF sample,
//                         Russian roulette; survivors are appended to the next
//                         ray queue with one atomic per wavefront
//               k_shadow  any-hit test of the shadow queue
//   k_finalize  the depth-0 spectral-channel rule of PathTracer.cpp:107-114
//   k_film      per-pixel sum over the pixel's samples
// Queue counts stay on the device: every kernel reads its count and loops
// grid-stride, so a batch needs no host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <atomic>
#include <chrono>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"
#include "device.h"
#include "engine.h"
#include "envmap.h"
#include "filters.h"
#include "kdbuild.h"
#include "rng.h"

using namespace mpt_dev;
using namespace mpt_engine;

namespace {

// Path flags.
constexpr uint32_t F_IOR_MASK = 3u;          // the path's IOR channel (PathTracer.cpp:32)
constexpr uint32_t F_ANY_SPECULAR = 4u;      // data["specular"] (PathTracer.cpp:46-47)
constexpr uint32_t F_D0_CONTINUED = 8u;      // the depth-0 vertex recursed
constexpr uint32_t F_EMIT_AFTER_SPEC = 16u;  // emittedContribution > EPSILON (PathTracer.cpp:97-98)

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Appends `want` lanes to a queue with one atomic per block: the waves'
// counts meet in LDS (sh holds WAVES + 1 words) and one lane adds the block
// total.  Queue counters are hot (every block of a launch appends to the
// same two or three), so this divides their traffic by the waves per block.
// Every thread of the block must call it.
template <int WAVES>
__device__ __forceinline__ uint32_t block_append(bool want, uint32_t *counter, uint32_t *sh) {
  const uint64_t mask = __ballot(want);
  const uint32_t wave = threadIdx.x / 64;
  if (lane_id() == 0) sh[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = sh[w];
      sh[w] = total;
      total += c;
    }
    sh[WAVES] = total ? atomicAdd(counter, total) : 0u;
  }
  __syncthreads();
  const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
  const uint32_t idx = sh[WAVES] + sh[wave] + below;
  __syncthreads();   // sh is reused by the next append
  return idx;
}

// Two appends with one exchange: lanes with wa get consecutive indices of
// counter ca and lanes with wb of counter cb, with one atomic per counter per
// block.  sh holds 2 * (WAVES + 1) words per call; callers alternate between
// two such regions (`parity`), so a call needs no barrier at its end: the
// next call that writes the same region comes after another call's first
// barrier, which every thread reaches only after reading this call's result.
// Every thread of the block must call it.
template <int WAVES>
__device__ __forceinline__ void block_append2(bool wa, bool wb, uint32_t *ca, uint32_t *cb, uint32_t *sh_pair,
                                              uint32_t parity, uint32_t &ia, uint32_t &ib) {
  uint32_t *sh = sh_pair + parity * 2 * (WAVES + 1);
  const uint64_t ma = __ballot(wa), mb = __ballot(wb);
  const uint32_t wave = threadIdx.x / 64;
  if (lane_id() == 0) {
    sh[wave] = (uint32_t)__popcll(ma);
    sh[WAVES + 1 + wave] = (uint32_t)__popcll(mb);
  }
  __syncthreads();
  if (threadIdx.x < 2) {   // thread 0 books counter a, thread 1 counter b
    uint32_t *row = sh + threadIdx.x * (WAVES + 1);
    uint32_t total = 0;
    for (int w = 0; w < WAVES; ++w) {
      const uint32_t c = row[w];
      row[w] = total;
      total += c;
    }
    row[WAVES] = total ? atomicAdd(threadIdx.x ? cb : ca, total) : 0u;
  }
  __syncthreads();
  const uint32_t below_a = __builtin_amdgcn_mbcnt_hi((uint32_t)(ma >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ma, 0u));
  const uint32_t below_b = __builtin_amdgcn_mbcnt_hi((uint32_t)(mb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mb, 0u));
  ia = sh[WAVES] + sh[wave] + below_a;
  ib = sh[2 * WAVES + 1] + sh[WAVES + 1 + wave] + below_b;
}

// Sum of v over the wave (for statistics counters).
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
__global__ void k_reset(uint32_t *counts, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) counts[i] = 0;
}

__device__ __forceinline__ uint32_t chunk_of_slot(const BatchDesc *bd, uint32_t slot) {
  uint32_t lo = 0, hi = bd->hdr.n_chunks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (bd->chunks[mid].slot0 <= slot) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ uint32_t chunk_of_pixel(const BatchDesc *bd, uint32_t pix) {
  uint32_t lo = 0, hi = bd->hdr.n_chunks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (bd->chunks[mid].pixoff <= pix) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Camera rays: ThinLensCamera::getWorldRay with aperture 0
// (cameras/ThinLensCamera.cpp:131-135) at the jittered stratum position of
// SuperSG + JitteredSG (renderers/generators/SuperSampleGenerator.cpp:48-91,
// JitteredSampleGenerator.cpp:30-31).  Slot -> (job, pixel, sample) through
// the batch's chunk table.
__global__ __launch_bounds__(256) void k_raygen(const BatchDesc *__restrict__ bd, float4 *__restrict__ q,
                                                uint2 *__restrict__ pjs, float2 *__restrict__ pos,
                                                uint32_t *__restrict__ count0) {
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t n = bd->hdr.n_paths;
  if (slot == 0) *count0 = n;
  if (slot >= n) return;
  const ChunkDev ch = bd->chunks[chunk_of_slot(bd, slot)];
  const JobDev &job = bd->jobs[ch.job];
  const uint32_t spp = job.spp, side = job.side, width = job.width;
  const uint32_t local = slot - ch.slot0;
  const uint32_t pixel = ch.pix0 + local / spp, sample = local % spp;
  const mpt_rng::U4 r = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, mpt_rng::kCameraDepth, 0);
  const uint32_t k = sample % (side * side);
  const uint32_t sx = k % side, sy = k / side;
  const uint32_t px = pixel % width, py = pixel / width;
  const float inv_side = 1.0f / (float)side;
  const float x = ((float)px + ((float)sx + mpt_rng::unitf(r.x)) * inv_side) / (float)width;
  const float y = ((float)py + ((float)sy + mpt_rng::unitf(r.y)) * inv_side) / (float)job.height;
  // a filtered film needs the sample's position (in pixels) when it splats
  if (job.filter && pos) pos[slot] = make_float2(x * (float)width, y * (float)job.height);
  const CameraDev &cam = job.cam;
  const float nx = (2.0f * x - 1.0f) * cam.tan_w, ny = (1.0f - 2.0f * y) * cam.tan_h;
  const float3 U = f3(cam.u[0], cam.u[1], cam.u[2]), V = f3(cam.v[0], cam.v[1], cam.v[2]);
  const float3 N = f3(cam.n[0], cam.n[1], cam.n[2]);
  float3 d = normalize(sub(add(mul(U, nx), mul(V, ny)), N));
  float3 o = f3(cam.eye[0], cam.eye[1], cam.eye[2]);
  if (cam.aperture > 0.0f) {
    // ThinLensCamera::getWorldRay (cameras/ThinLensCamera.cpp:131-172): the
    // ray meets the focal plane at t; the origin moves within the circle of
    // confusion of radius aperture (l - D) / l, l = f t / (t - f), at
    // r = radius sin(e1), e1 ~ U(0, pi/2), angle e2 ~ U(0, 2 pi), and the ray
    // aims at the focal-plane point again.
    const float t = -cam.focal_distance / dot(d, N);
    if (t != cam.focal_length) {
      const float3 focus = add(o, mul(d, t));
      const float l_image = cam.focal_length * t / (t - cam.focal_length);
      const float cof = cam.aperture * (l_image - cam.focal_distance) / l_image;
      const mpt_rng::U4 lr = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, mpt_rng::kCameraDepth, 1);
      const float e1 = mpt_rng::unitf(lr.x) * (0.5f * kPi), e2 = mpt_rng::unitf(lr.y) * (2.0f * kPi);
      const float r = cof * sinf(e1);
      o = add(o, add(mul(U, r * cosf(e2)), mul(V, r * sinf(e2))));
      d = normalize(sub(focus, o));
    }
  }
  // queue entry: (origin, slot) (direction, origin primitive) (throughput, flags)
  q[3 * slot] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
  q[3 * slot + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(kMiss));
  q[3 * slot + 2] = make_float4(1.0f, 1.0f, 1.0f, __uint_as_float(mpt_rng::third(r.z)));
  pjs[slot] = make_uint2(pixel, (ch.job << kJobBits) | sample);
}

// Closest hit of the live queue (kdTreeAccel::getIntersection).  With COUNT
// the kernel also accumulates the kd nodes and references it visited.
template <bool COUNT, int STACK>
__global__ __launch_bounds__(256) void k_extend(Scene sc, const float4 *__restrict__ q,
                                                const uint32_t *__restrict__ count, float2 *__restrict__ hits,
                                                unsigned long long *__restrict__ counters) {
  __shared__ uint2 stack[STACK * kBlock];
  uint2 *stk = stack + threadIdx.x;
  const uint32_t n = *count;
  uint32_t n_nodes = 0, n_refs = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float4 a = q[3 * i], b = q[3 * i + 1];
    const RaySetup r = setup_ray(xyz(a), xyz(b));
    Walker<false, STACK> w;
    float t = INFINITY;
    uint32_t rec = kMiss;
    if (w.start(sc, r, INFINITY, __float_as_uint(b.w))) {
      while (!w.template step<COUNT>(sc, r, stk, &n_nodes, &n_refs)) {
      }
      w.raw(t, rec);
    }
    hits[i] = make_float2(t, __uint_as_float(rec));
  }
  if (COUNT) {
    atomicAdd(&counters[0], (unsigned long long)n_nodes);
    atomicAdd(&counters[1], (unsigned long long)n_refs);
  }
}

// Persistent variant of k_extend: each wave keeps its lanes busy by handing a
// lane whose ray finished the next ray of the wave's claimed chunk (refill
// taken when at least kRefill lanes are idle, or all are); a wave claims
// kChunk rays of the queue at a time with one atomic, so the claims balance
// the load across waves without one atomic per refill.  24 and 128 were the
// fastest of 8-32 and 64-1024 on the 1M-triangle soup.
#ifndef MPT_REFILL
#define MPT_REFILL 24
#endif
#ifndef MPT_CHUNK
#define MPT_CHUNK 128
#endif
constexpr uint32_t kRefill = MPT_REFILL;   // idle lanes that trigger a refill
constexpr uint32_t kChunk = MPT_CHUNK;     // rays per claim

// TOP: the block first copies the top of the tree (up to kTopNodes nodes,
// the breadth-first prefix of the node array) into LDS and reads those
// nodes from there.
constexpr uint32_t kTopNodes = 512;

template <bool COUNT, int STACK, bool TOP, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(STACK <= 8 ? 8 : (STACK <= 16 ? 4 : 2)))) void k_extend_persistent(
    Scene sc, const float4 *__restrict__ q, const uint32_t *__restrict__ count, float2 *__restrict__ hits,
    unsigned long long *__restrict__ counters, uint32_t *__restrict__ fetch) {
  __shared__ uint2 stack[STACK * BLOCK];
  __shared__ uint2 top_lds[TOP ? kTopNodes : 1];
  uint2 *stk = stack + threadIdx.x;
  if (TOP) {
    for (uint32_t i = threadIdx.x; i < sc.top_bytes / 8; i += blockDim.x) top_lds[i] = sc.nodes[i];
    __syncthreads();
  }
  const uint32_t n = *count;
  uint32_t next = 0, end = 0;   // wave-uniform: the unassigned part of the claimed chunk
  bool exhausted = false;       // wave-uniform: the queue has no unclaimed rays left
  uint32_t n_nodes = 0, n_refs = 0;
  uint32_t ray = kMiss;
  RaySetup r;
  Walker<false, STACK, BLOCK> w;
  for (;;) {
    const bool need = ray == kMiss;
    const uint64_t idle = __ballot(need);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= kRefill || n_idle == 64) {
      if (next >= end && !exhausted) {
        uint32_t base = 0;
        if (lane_id() == 0) base = atomicAdd(fetch, kChunk);
        base = __builtin_amdgcn_readfirstlane(base);
        next = min(base, n);
        end = min(base + kChunk, n);
        exhausted = base + kChunk >= n;
      }
      if (next < end) {
        const uint32_t idx = next + __builtin_amdgcn_mbcnt_hi((uint32_t)(idle >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)idle, 0u));
        next = min(end, next + n_idle);
        if (need && idx < end) {
          const float4 a = q[3 * idx], b = q[3 * idx + 1];
          r = setup_ray(xyz(a), xyz(b));
          if (w.start(sc, r, INFINITY, __float_as_uint(b.w))) {
            ray = idx;
          } else {
            hits[idx] = make_float2(INFINITY, __uint_as_float(kMiss));
          }
        }
      }
    }
    if (__ballot(ray != kMiss) == 0) {
      if (exhausted && next >= end) break;
      continue;
    }
    if (ray != kMiss && w.template step<COUNT, TOP>(sc, r, stk, &n_nodes, &n_refs, top_lds)) {
      float t;
      uint32_t rec;
      w.raw(t, rec);
      hits[ray] = make_float2(t, __uint_as_float(rec));
      ray = kMiss;
    }
  }
  if (COUNT) {
    atomicAdd(&counters[0], (unsigned long long)n_nodes);
    atomicAdd(&counters[1], (unsigned long long)n_refs);
  }
}

// Any hit of the shadow rays (kdTreeAccel::intersects) over t < the ray's
// limit, skipping the primitive it leaves and the light primitive it aims
// at.  One lane per group:
// the light samples of one vertex are a run of the shadow buffer
// (sr[slot * spv ...], spv = shadow rays per vertex of the batch), traced in
// (light, sample) order.  Their unoccluded contributions are summed in that
// order and added to the path's radiance once, as DirectIllumination sums
// per light in a fixed loop (renderers/DirectIllumination.cpp:90-135), so a
// path's value does not depend on how the rays were scheduled.  Each slot
// has at most one group per launch, so its radiance sees one addition here.
template <int STACK>
__global__ __launch_bounds__(256) void k_shadow(Scene sc, const BatchDesc *__restrict__ bd,
                                                const float4 *__restrict__ sg, const float4 *__restrict__ sr,
                                                const uint32_t *__restrict__ count, float4 *__restrict__ Ls,
                                                unsigned long long *__restrict__ total_rays) {
  __shared__ uint2 stack[STACK * kBlock];
  uint2 *stk = stack + threadIdx.x;
  const uint32_t n = *count;
  const uint32_t spv = sc.num_lights * bd->hdr.max_direct + sc.num_points;
  uint32_t rays = 0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    // group: (origin, slot) (origin primitive, ray count, 0, 0)
    const float4 g0 = sg[2ull * i], g1 = sg[2ull * i + 1];
    const uint32_t slot = __float_as_uint(g0.w), prim = __float_as_uint(g1.x), cnt = __float_as_uint(g1.y);
    // ray: (direction, distance limit) (contribution, light primitive or kMiss)
    const float4 *run = sr + 2ull * ((uint64_t)slot * spv);
    float3 sum = f3(0, 0, 0);
    for (uint32_t j = 0; j < cnt; ++j) {
      const float4 a = run[2 * j], c = run[2 * j + 1];
      const RaySetup r = setup_ray(xyz(g0), xyz(a));
      float t;
      uint32_t id;
      if (!trace<true, false, STACK>(sc, r, a.w, prim, stk, t, id, nullptr, nullptr, nullptr, __float_as_uint(c.w)))
        sum = add(sum, xyz(c));
    }
    rays += cnt;
    if (!is_black(sum)) {
      float4 *L = Ls + slot;
      atomicAdd(&L->x, sum.x);
      atomicAdd(&L->y, sum.y);
      atomicAdd(&L->z, sum.z);
    }
  }
  rays = wave_sum(rays);
  if (lane_id() == 0 && rays) atomicAdd(total_rays, (unsigned long long)rays);
}

struct HitInfo {
  float3 n;
  uint32_t mat;
  float area;       // shape surface area (quads, emitting meshes; Sphere::_getSurfaceArea for spheres)
  uint32_t light;   // the emitting mesh the hit triangle belongs to (light index), else kMiss
};

// Shading normal and material of a hit: Mesh::_getGeometricNormal with
// MeshTriangle::getBaryocentricCoords (shapes/Mesh.cpp:419-433,
// shapes/MeshTriangle.cpp:344-361), Sphere / Plane normals (shapes/Sphere.cpp:
// 86-92, shapes/Plane.cpp:59-63).  The normal is not flipped toward the ray.
__device__ __forceinline__ HitInfo hit_info(const Scene &sc, uint32_t prim, float3 p, float2 uv) {
  HitInfo h;
  h.area = 0.0f;
  h.light = kMiss;
  if (prim < sc.nt) {
    // Mesh::_getGeometricNormal (shapes/Mesh.cpp:419-433): the vertex normals
    // weighted by the hit's barycentric coordinates
    const float4 *ts = sc.tshade + 3ull * prim;
    const float4 n0 = ts[0], n1 = ts[1], n2 = ts[2];
    const float w0 = 1.0f - uv.x - uv.y;
    h.n = normalize(add(add(mul(xyz(n0), w0), mul(xyz(n1), uv.x)), mul(xyz(n2), uv.y)));
    h.mat = __float_as_uint(n0.w);
    h.area = n1.w;                          // the emitting shape's area (0 otherwise)
    h.light = __float_as_uint(n2.w) - 1u;   // light index + 1, 0 for a triangle that does not emit
  } else if (prim < sc.nt + sc.ns) {
    const float4 s = sc.spheres[prim - sc.nt];
    h.n = normalize(sub(p, xyz(s)));
    h.mat = sc.sphere_mat[prim - sc.nt];
    h.area = kPi;
  } else {
    const float4 *q = sc.quads + 4ull * (prim - sc.nt - sc.ns);
    const float4 n = q[3];
    h.n = xyz(n);
    h.mat = __float_as_uint(n.w);
    h.area = q[0].w;
  }
  return h;
}

// One vertex of PathTracer::_evaluate (renderers/renderers/PathTracer.cpp:
// 28-116) with DirectIllumination::evaluate (renderers/DirectIllumination.cpp:
// 46-136).  The recursion becomes a throughput `beta`; the emission terms, NEE
// and roulette follow the reference rules, each cited below.
// Scene::getBackgroundRadiance(d) = EnvironmentMap::getLe(-d)
// (core/Scene.cpp:116-118; materials/emitters/EnvironmentMap.cpp:44-85) in
// double precision: spherical coordinates of w = -d, the lat-long UV, the
// repeat, the texel position clamped into the image, then the kernel
// filter over the clamped neighbourhood (filters/KernelFilter.cpp:124-160),
// normalised by the weights' sum.
// Out of line, with the scene's fields by value: the miss path is rare and
// its FP64 trigonometry would otherwise raise the shading kernel's register
// pressure everywhere.
__device__ __noinline__ float3 env_radiance_at(const float4 *__restrict__ tex, const double *__restrict__ kern,
                                               uint32_t env_w, uint32_t env_h, uint32_t env_m, float rep_u,
                                               float rep_v, float3 dir) {
  constexpr double kPiD = 3.14159265358979323846;
  const double wx = -(double)dir.x, wy = -(double)dir.y, wz = -(double)dir.z;
  const double rho = sqrt(wx * wx + wz * wz);
  const double mag = sqrt(wx * wx + wy * wy + wz * wz);
  const double theta = rho < fabs(wy) ? acos(rho / mag) * (wy >= 0 ? 1.0 : -1.0) : asin(wy / mag);
  const double phi = (wx == 0 && wz == 0) ? 0.0 : atan2(wx, wz);
  const double u = phi / (-2 * kPiD) + 0.5, v = theta / -kPiD + 0.5;
  const double s = u * rep_u, t = v * rep_v;
  const double W = env_w, H = env_h;
  double x = (s - floor(s)) * W, y = (t - floor(t)) * H;
  x = x >= W - 1 ? W - 1 : (x <= 0 ? 0.0 : x);
  y = y >= H - 1 ? H - 1 : (y <= 0 ? 0.0 : y);
  const int xi = (int)floor(x), yi = (int)floor(y), half = (int)(env_m >> 1);
  const int wmax = (int)env_w - 1, hmax = (int)env_h - 1;
  double r = 0, g = 0, b = 0;
  const double *k = kern;
  for (int i = -half; i <= half; ++i) {
    const int row = min(max(i + yi, 0), hmax);
    for (int j = -half; j <= half; ++j) {
      const int col = min(max(j + xi, 0), wmax);
      const float4 tx = tex[(size_t)row * env_w + col];
      const double wgt = *k++;
      r += tx.x * wgt;
      g += tx.y * wgt;
      b += tx.z * wgt;
    }
  }
  const double sum = *k;
  return f3((float)(r / sum), (float)(g / sum), (float)(b / sum));
}

__device__ __forceinline__ float3 env_radiance(const Scene &sc, float3 dir) {
  return env_radiance_at(sc.env_tex, sc.env_kernel, sc.env_w, sc.env_h, sc.env_m, sc.env_rep_u, sc.env_rep_v, dir);
}

// Shape::getPoint (shapes/Shape.cpp:52-64) of an area light at (u, v), with
// r1, r2 the two further uniforms a mesh triangle takes: the light point,
// its normal, and the primitive a shadow ray to it may skip (a planar
// light's only crossing is the end point; kMiss for a sphere).
__device__ __forceinline__ void light_point(const Scene &sc, const LightDev &ld, float u, float v, float r1, float r2,
                                            float3 &lp, float3 &ln, uint32_t &target) {
  target = kMiss;
  if (ld.kind == LIGHT_QUAD) {
    // Plane::getPosition (shapes/Plane.cpp:91-93)
    const float4 *lq = sc.quads + 4ull * ld.index;
    const float4 c0 = lq[0], ex = lq[1], ez = lq[2], qn = lq[3];
    lp = add(xyz(c0), add(mul(xyz(ex), u - 0.5f), mul(xyz(ez), v - 0.5f)));
    ln = xyz(qn);
    target = sc.nt + sc.ns + ld.index;
  } else if (ld.kind == LIGHT_SPHERE) {
    // Sphere::getPosition (shapes/Sphere.cpp:117-124): an object-space point
    // at distance 1 from the centre near +z, mapped by the sphere's
    // transform; the normal is that point's direction (Sphere.cpp:86-92, a
    // similarity transform)
    const float x = fminf(0.5f, fmaxf(-0.5f, tanf((1.0f - u) * (0.5f / kPi))));
    const float y = fminf(0.5f, fmaxf(-0.5f, 0.5f * sinf((0.5f - v) * (1.0f / kPi))));
    const float z = sqrtf(1.0f - x * x - y * y);
    const float3 q = add(add(mul(xyz(ld.frame[0]), x), mul(xyz(ld.frame[1]), y)), mul(xyz(ld.frame[2]), z));
    lp = add(xyz(sc.spheres[ld.index]), q);
    ln = normalize(q);
  } else {
    // Mesh::getPosition (shapes/Mesh.cpp:442-450): the triangle picked by u,
    // barycentrics (r1, v, r2) / (r1 + v + r2) (MeshTriangle.cpp:299-333);
    // the normal interpolates the vertex normals there
    // (Mesh::_getGeometricNormal, :419-433)
    const uint32_t ti = ld.first + min((uint32_t)floor((double)u * ld.count), ld.count - 1u);
    const float4 *lt = sc.light_tris + 3ull * ti;
    const float4 t0 = lt[0], t1 = lt[1], t2 = lt[2];
    const float ks = 1.0f / (r1 + v + r2);
    const float b0 = r1 * ks, b1 = v * ks, b2 = r2 * ks;
    lp = add(add(mul(xyz(t0), b0), mul(xyz(t1), b1)), mul(xyz(t2), b2));
    target = __float_as_uint(t0.w);
    const float4 *ts = sc.tshade + 3ull * target;
    ln = normalize(add(add(mul(xyz(ts[0]), b0), mul(xyz(ts[1]), b1)), mul(xyz(ts[2]), b2)));
  }
}

// Waves per SIMD the shading kernel is compiled for (its register budget):
// the most at which the single-sample form does not spill; the form with
// several light samples per light gets one fewer.
#ifndef MPT_SHADE_WAVES
#define MPT_SHADE_WAVES 5
#endif

// kMulti: some job of the batch draws several light samples per light, or
// draws its lights by power; the common single-sample form keeps the NEE
// loop's counters out of registers.
// kEnv: the scene's background is an environment image (a constant
// otherwise).
//
// The grid holds as many blocks as fit on the device at once (every block
// strides over the queue the same number of times, give or take one), and
// each pass appends its shadow groups and extension rays with one exchange.
template <bool kMulti, bool kEnv, int kShadeBlock>
__global__ __launch_bounds__(kShadeBlock) __attribute__((amdgpu_waves_per_eu(kMulti ? MPT_SHADE_WAVES - 1 : MPT_SHADE_WAVES))) void k_shade(
    Scene sc, const BatchDesc *__restrict__ bd, uint32_t depth, const float4 *__restrict__ qin,
    const float2 *__restrict__ hits, const uint32_t *__restrict__ count_in, float4 *__restrict__ qout,
    uint32_t *__restrict__ count_out, float4 *__restrict__ sg, float4 *__restrict__ sr,
    uint32_t *__restrict__ count_groups, float4 *__restrict__ Ls, uint32_t *__restrict__ flags,
    const uint2 *__restrict__ pjs) {
  __shared__ uint32_t append_sh[4 * (kShadeBlock / 64 + 1)];
  const uint32_t n = *count_in;
  const uint32_t stride = gridDim.x * blockDim.x;
  // shadow rays per vertex: the stride of a slot's run in the shadow buffer
  const uint32_t spv = sc.num_lights * bd->hdr.max_direct + sc.num_points;
  uint32_t parity = 0;
  // block-uniform loop: block_append2 synchronises the block
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride, parity ^= 1u) {
    const uint32_t i = base + threadIdx.x;
    const bool active = i < n;
    bool want_ext = false;
    float3 p = f3(0, 0, 0), wo = f3(0, 0, 0);
    uint32_t slot = 0, prim = kMiss;
    float3 beta = f3(0, 0, 0);
    uint32_t fl = 0;
    float4 Lv = make_float4(0, 0, 0, 0);
    HitInfo hi;
    hi.n = f3(0, 1, 0);
    hi.mat = 0;
    hi.area = 0;
    bool surface = false;
    float thit = 0.0f;
    float3 d = f3(0, 0, 1);
    uint32_t pixel = 0, sample = 0, seed = 0, job_id = 0, max_depth = 0;
    uint32_t ndir_job = 1;   // light samples per light
    bool direct_off = false; // efficientDirect = false
    bool select = false;     // light samples drawn by emitted power
    if (active) {
      // (origin, slot) (direction, origin primitive) (throughput, flags)
      const float4 a = qin[3ull * i], b = qin[3ull * i + 1], c = qin[3ull * i + 2];
      const float2 h = hits[i];
      slot = __float_as_uint(a.w);
      d = xyz(b);
      const uint2 pj = pjs[slot];
      pixel = pj.x;
      sample = pj.y & ((1u << kJobBits) - 1);
      const JobDev &job = bd->jobs[pj.y >> kJobBits];
      seed = job.seed;
      job_id = job.job_id;
      max_depth = job.max_depth;
      if (kMulti) ndir_job = job.direct;
      direct_off = job.direct_off != 0;
      if (kMulti) select = job.light_select != 0;
      beta = xyz(c);
      fl = __float_as_uint(c.w);
      const uint32_t rec = __float_as_uint(h.y);
      if (rec == kMiss) {
        // Scene::getBackgroundRadiance (core/Scene.cpp:116-118), every depth.
        const float3 le = kEnv ? env_radiance(sc, d) : f3(sc.env[0], sc.env[1], sc.env[2]);
        Lv.x += beta.x * le.x;
        Lv.y += beta.y * le.y;
        Lv.z += beta.z * le.z;
      } else {
        // the closest hit's record: primitive id, and t and barycentrics
        // recomputed in double precision
        const float4 *rp = sc.recs + 3ull * rec;
        prim = __float_as_uint(rp[0].w);
        float2 uv;
        surface = true;
        thit = refine_hit(rp, xyz(a), d, h.x, &uv);
        p = add(xyz(a), mul(d, thit));
        hi = hit_info(sc, prim, p, uv);
      }
    }
    mpt_material m;
    if (surface) m = sc.mats[hi.mat];
    const bool specular = surface && m.type == MPT_BSDF_DIELECTRIC;
    if (surface) {
      if (specular) fl |= F_ANY_SPECULAR;
      // PathTracer.cpp:50-69: emission at depth 0, or after a specular vertex
      // that sent the ray out on its normal's side (Le * area / t^2); with
      // efficientDirect = false at every vertex (:51).
      if (emissive(m) && dot(hi.n, d) < 0.0f) {
        const float3 le = le_of(m);
        float w = 0.0f;
        if (depth == 0 || direct_off) {
          w = 1.0f;
        } else if (fl & F_EMIT_AFTER_SPEC) {
          w = hi.area / (thit * thit);
        }
        Lv.x += beta.x * le.x * w;
        Lv.y += beta.y * le.y * w;
        Lv.z += beta.z * le.z * w;
      }
    }
    if (active) {
      // A slot's radiance starts with its camera vertex (a store: k_raygen
      // leaves it unset); later vertices add with atomics that return
      // nothing.  Shadow contributions of the same depth come after, in
      // k_shadow, so each slot sees one ordered sequence of additions.
      if (depth == 0) {
        Ls[slot] = Lv;
      } else if (Lv.x != 0.0f || Lv.y != 0.0f || Lv.z != 0.0f) {
        atomicAdd(&Ls[slot].x, Lv.x);
        atomicAdd(&Ls[slot].y, Lv.y);
        atomicAdd(&Ls[slot].z, Lv.z);
      }
    }
    // Next-event estimation (DirectIllumination.cpp:90-135): per light, n
    // shadow rays to points jittered over floor(sqrt(n))^2 strata of the
    // light, each weighted area / n; skipped at specular vertices (:57) and
    // on the light itself (:93).  With light selection by power, each of the
    // n rays goes to one light drawn from the emitted-power table instead,
    // weighted by 1 / its probability.  The rays go to the slot's own run of
    // the shadow buffer in sample order; one group entry per vertex points
    // k_shadow at them.
    uint32_t n_sh = 0;
    if (surface && !specular && spv && !direct_off) {
      float4 *run = sr + 2ull * ((uint64_t)slot * spv);
      uint32_t side = 1;   // floor(sqrt(n)) strata per side
      if (kMulti)
        while ((side + 1) * (side + 1) <= ndir_job) ++side;
      const float inv = 1.0f / (float)side;
      // sample k of area light li from random stream `stream`, weighted
      // area / (n prob)
      auto area_sample = [&](uint32_t li, uint32_t k, uint32_t stream, float prob) {
        const LightDev &ld = sc.lights[li];
        // not the light the vertex lies on (DirectIllumination.cpp:93)
        if ((ld.kind == LIGHT_QUAD && prim == sc.nt + sc.ns + ld.index) ||
            (ld.kind == LIGHT_SPHERE && prim == sc.nt + ld.index) || (ld.kind == LIGHT_MESH && hi.light == li))
          return;
        // OrientedEmitter::getLe (emitters/OrientedEmitter.cpp:17-23)
        const float3 lle = le_of(sc.mats[ld.mat]);
        const mpt_rng::U4 r = mpt_rng::path_draw(seed, job_id, pixel, sample, depth, stream);
        const uint32_t cell = kMulti ? k % (side * side) : 0u;
        const float u = ((float)(cell % side) + mpt_rng::unitf(r.x)) * inv;
        const float v = ((float)(cell / side) + mpt_rng::unitf(r.y)) * inv;
        float3 lp, ln;
        uint32_t target;
        light_point(sc, ld, u, v, mpt_rng::unitf(r.z), mpt_rng::unitf(r.w), lp, ln, target);
        float3 w = sub(lp, p);
        const float dist = len(w);
        w = mul(w, 1.0f / dist);
        const float cos_wo = fabsf(dot(hi.n, w));
        const float cos_wi = fabsf(dot(ln, w));
        const float3 fr = bsdf_eval(m, hi.n, d, w);
        const bool lit = dot(ln, w) < 0.0f;
        if (cos_wo > 0.0f && cos_wi > 0.0f && !is_black(fr) && lit) {
          const float g = cos_wo * cos_wi / (dist * dist) * (ld.area / ((float)ndir_job * prob));
          const float3 contrib = mul(mul(beta, mul(fr, lle)), g);
          if (!is_black(contrib)) {
            // Scene::intersects(ray, t - EPSILON) over the kd window
            // t < tMax - EPSILON (DirectIllumination.cpp:123, Scene.cpp:113)
            run[2 * n_sh] = make_float4(w.x, w.y, w.z, dist - 2.0f * kEps);
            run[2 * n_sh + 1] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(target));
            ++n_sh;
          }
        }
      };
      // Point lights (DirectIllumination.cpp:96-121 with isPoint): one
      // sample, no area weight, and a point has no normal, so its cosine is
      // 1 (SurfacePoint's normal dot, core/SurfacePoint.h:143); divided by q.
      auto point_sample = [&](uint32_t li, float q) {
        const float4 lp = sc.points[2 * li], le = sc.points[2 * li + 1];
        float3 w = sub(xyz(lp), p);
        const float dist = len(w);
        w = mul(w, 1.0f / dist);
        const float cos_wo = fabsf(dot(hi.n, w));
        const float3 fr = bsdf_eval(m, hi.n, d, w);
        if (dist > 0.0f && cos_wo > 0.0f && !is_black(fr)) {
          const float3 contrib = mul(mul(beta, mul(fr, xyz(le))), cos_wo / (dist * dist) / q);
          if (!is_black(contrib)) {
            run[2 * n_sh] = make_float4(w.x, w.y, w.z, dist - 2.0f * kEps);
            run[2 * n_sh + 1] = make_float4(contrib.x, contrib.y, contrib.z, __uint_as_float(kMiss));
            ++n_sh;
          }
        }
      };
      if (select) {
        // EmitterSampler::sample's draw (EmitterSampler.cpp:51-61,
        // Random::sampleNormalizedCDF): the first light whose running share
        // reaches the uniform, else the last
        const uint32_t nl = sc.num_lights + sc.num_points;
        for (uint32_t k = 0; k < ndir_job; ++k) {
          const float x = mpt_rng::unitf(mpt_rng::path_draw(seed, job_id, pixel, sample, depth, 1 + ndir_job + k).x);
          uint32_t j = 0;
          while (j + 1 < nl && !(x <= sc.light_sel[j].x)) ++j;
          const float pj = sc.light_sel[j].y;
          if (j < sc.num_lights) area_sample(j, k, 1 + k, pj);
          else point_sample(j - sc.num_lights, (float)ndir_job * pj);
        }
      } else {
        for (uint32_t li = 0; li < sc.num_lights; ++li)
          for (uint32_t k = 0; k < ndir_job; ++k) area_sample(li, k, 1 + li * ndir_job + k, 1.0f);
        for (uint32_t li = 0; li < sc.num_points; ++li) point_sample(li, 1.0f);
      }
    }
    if (surface) {
      const mpt_rng::U4 r = mpt_rng::path_draw(seed, job_id, pixel, sample, depth, 0);
      const Sample s = bsdf_sample(m, hi.n, d, mpt_rng::unitf(r.x), mpt_rng::unitf(r.y), mpt_rng::unitf(r.w),
                                   fl & F_IOR_MASK);
      if (!is_black(s.wo)) {
        // Russian roulette (PathTracer.cpp:88-92).
        const float p_cont = depth < 2 ? (is_black(s.fs) ? 0.0f : 1.0f) : fminf(0.95f, max3(s.fs));
        if (mpt_rng::unitf(r.z) < p_cont && depth + 1 < max_depth) {
          want_ext = true;
          wo = s.wo;
          if (depth == 0) fl |= F_D0_CONTINUED;
          fl = (specular && dot(hi.n, s.wo) > kEps) ? (fl | F_EMIT_AFTER_SPEC) : (fl & ~F_EMIT_AFTER_SPEC);
          beta = mul(beta, mul(s.fs, 1.0f / p_cont));
        }
      }
    }
    uint32_t gi, ei;
    block_append2<kShadeBlock / 64>(n_sh != 0, want_ext, count_groups, count_out, append_sh, parity, gi, ei);
    if (n_sh) {
      sg[2ull * gi] = make_float4(p.x, p.y, p.z, __uint_as_float(slot));
      sg[2ull * gi + 1] = make_float4(__uint_as_float(prim), __uint_as_float(n_sh), 0.0f, 0.0f);
    }
    if (want_ext) {
      qout[3ull * ei] = make_float4(p.x, p.y, p.z, __uint_as_float(slot));
      qout[3ull * ei + 1] = make_float4(wo.x, wo.y, wo.z, __uint_as_float(prim));
      qout[3ull * ei + 2] = make_float4(beta.x, beta.y, beta.z, __uint_as_float(fl));
    }
    // the path's flags as k_finalize needs them, once, when it ends
    if (active && !want_ext) flags[slot] = fl;
  }
}

// PathTracer.cpp:107-114: a path with a specular vertex whose depth-0 vertex
// recursed keeps only its IOR channel, times the channel count.
__global__ __launch_bounds__(256) void k_finalize(const BatchDesc *__restrict__ bd, float4 *__restrict__ Ls,
                                                  const uint32_t *__restrict__ flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bd->hdr.n_paths) return;
  const uint32_t f = flags[i];
  if ((f & F_ANY_SPECULAR) && (f & F_D0_CONTINUED)) {
    float4 L = Ls[i];
    const uint32_t c = f & F_IOR_MASK;
    L = make_float4(c == 0 ? 3.0f * L.x : 0.0f, c == 1 ? 3.0f * L.y : 0.0f, c == 2 ? 3.0f * L.z : 0.0f, 0.0f);
    Ls[i] = L;
  }
}

// Film: ProgressiveFilterValue numerator/denominator per pixel
// (filters/ProgressiveFilterValue.h:20-37) with the pixel-bin weight 1 of
// RenderOutput::addSample (renderers/RenderOutput.cpp:72-76).  All samples of
// a pixel are in one batch and are summed in sample order, so a pixel's value
// does not depend on what else shares the batch.
__global__ __launch_bounds__(256) void k_film(const BatchDesc *__restrict__ bd, const float4 *__restrict__ Ls) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bd->hdr.n_pixels) return;
  const ChunkDev ch = bd->chunks[chunk_of_pixel(bd, i)];
  const JobDev &job = bd->jobs[ch.job];
  float4 *film = job.film;
  if (!film || job.filter) return;
  const uint32_t spp = job.spp;
  const float4 *L = Ls + ch.slot0 + (uint64_t)(i - ch.pixoff) * spp;
  float r = 0, g = 0, b = 0;
  // Neighbouring lanes' runs are spp * 16 bytes apart, so one load
  // instruction touches a cache line per lane.  Eight loads back to back use
  // the whole 128-byte line while it is still cached; the sum stays in
  // sample order.
  uint32_t s = 0;
  for (; s + 8 <= spp; s += 8) {
    float4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = L[s + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      r += v[k].x;
      g += v[k].y;
      b += v[k].z;
    }
  }
  for (; s < spp; ++s) {
    const float4 v = L[s];
    r += v.x;
    g += v.y;
    b += v.z;
  }
  film[ch.pix0 + (i - ch.pixoff)] = make_float4(r, g, b, (float)spp);
}

// ---------------------------------------------------------------------------
// Tone maps on the device, in double precision like the reference
// (renderers/tonemaps/ReinhardToneMap.cpp:25-98, NullToneMap.cpp:21-46).
// Reinhard needs two image-wide figures, the log-average luminance (key) and
// the largest luminance (white): k_tm_reduce folds a fixed grid of partials,
// k_tm_fold reduces them in a fixed order (the same bits every run), and
// k_tm_map maps each pixel.  HBM-bound: 16 B read per pixel per pass, 4 B
// written.
// ---------------------------------------------------------------------------
constexpr int kTmBlock = 256, kTmGrid = 512;

// utils/ColorUtils.cpp:115-133: the sRGB (D65) primaries matrices.
__constant__ double kRgbToXyz[9] = {0.41245322014416, 0.35757958129350, 0.18042258997054,
                                    0.21267112134122, 0.71515920531078, 0.07216877677613,
                                    0.01933381646199, 0.11919354020663, 0.95022692228971};
__constant__ double kXyzToRgb[9] = {3.240479, -1.537150, -0.498535, -0.969256, 1.875992,
                                    0.041556, 0.055648,  -0.204043, 1.057311};

// numerator / denominator, black where nothing was accumulated
// (filters/ProgressiveFilterValue.h:27-32)
__device__ __forceinline__ void film_value(float4 f, double &r, double &g, double &b) {
  const double w = f.w;
  r = w == 0 ? 0.0 : f.x / w;
  g = w == 0 ? 0.0 : f.y / w;
  b = w == 0 ? 0.0 : f.z / w;
}

// (unsigned char) CLAMP(v, 0, 255): at or above 255 -> 255, at or below 0
// or NaN -> 0, else truncation.
__device__ __forceinline__ uint8_t tm_byte(double v) {
  if (v >= 255.0) return 255;
  if (!(v > 0.0)) return 0;
  return (uint8_t)(int)v;
}

__global__ __launch_bounds__(kTmBlock) void k_tm_reduce(const float4 *__restrict__ film, uint32_t n,
                                                       double2 *__restrict__ partials) {
  __shared__ double2 sh[kTmBlock];
  double sum = 0.0, top = 0.0;
  for (uint32_t i = blockIdx.x * kTmBlock + threadIdx.x; i < n; i += kTmGrid * kTmBlock) {
    double r, g, b;
    film_value(film[i], r, g, b);
    const double Y = kRgbToXyz[3] * r + kRgbToXyz[4] * g + kRgbToXyz[5] * b;
    sum += log(1e-7 + Y);
    top = fmax(top, Y);
  }
  sh[threadIdx.x] = make_double2(sum, top);
  __syncthreads();
  for (int w = kTmBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x].x += sh[threadIdx.x + w].x;
      sh[threadIdx.x].y = fmax(sh[threadIdx.x].y, sh[threadIdx.x + w].y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// One block: folds the kTmGrid partials and writes (scale, white) after them.
__global__ __launch_bounds__(kTmBlock) void k_tm_fold(double2 *__restrict__ partials, uint32_t n) {
  __shared__ double2 sh[kTmBlock];
  double sum = 0.0, top = 0.0;
  for (int i = threadIdx.x; i < kTmGrid; i += kTmBlock) {
    sum += partials[i].x;
    top = fmax(top, partials[i].y);
  }
  sh[threadIdx.x] = make_double2(sum, top);
  __syncthreads();
  for (int w = kTmBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[threadIdx.x].x += sh[threadIdx.x + w].x;
      sh[threadIdx.x].y = fmax(sh[threadIdx.x].y, sh[threadIdx.x + w].y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double key = exp(sh[0].x / (double)n);
    const double top_y = sh[0].y;
    partials[kTmGrid] = make_double2(key <= 0 ? 0.0 : 0.18 / key, top_y <= 0 ? 0.0 : 1.0 / (top_y * top_y));
  }
}

__global__ __launch_bounds__(kTmBlock) void k_tm_map(const float4 *__restrict__ film, uint32_t n, uint32_t mode,
                                                    const double2 *__restrict__ params, uchar4 *__restrict__ out) {
  const uint32_t i = blockIdx.x * kTmBlock + threadIdx.x;
  if (i >= n) return;
  double r, g, b;
  film_value(film[i], r, g, b);
  const double correction = 1.0 / 2.22;
  if (mode == MPT_TONEMAP_REINHARD) {
    const double scale = params[kTmGrid].x, white = params[kTmGrid].y;
    const double X = kRgbToXyz[0] * r + kRgbToXyz[1] * g + kRgbToXyz[2] * b;
    const double Y = kRgbToXyz[3] * r + kRgbToXyz[4] * g + kRgbToXyz[5] * b;
    const double Z = kRgbToXyz[6] * r + kRgbToXyz[7] * g + kRgbToXyz[8] * b;
    // chromaticity kept, luminance compressed: L (1 + L white) / (1 + L)
    const double cx = X / (X + Y + Z), cy = Y / (X + Y + Z);
    const double lum = scale * Y;
    const double Yd = lum * (1.0 + lum * white) / (1.0 + lum);
    const double Xd = Yd * cx / cy, Zd = Yd * (1 - cx - cy) / cy;
    r = kXyzToRgb[0] * Xd + kXyzToRgb[1] * Yd + kXyzToRgb[2] * Zd;
    g = kXyzToRgb[3] * Xd + kXyzToRgb[4] * Yd + kXyzToRgb[5] * Zd;
    b = kXyzToRgb[6] * Xd + kXyzToRgb[7] * Yd + kXyzToRgb[8] * Zd;
  }
  out[i] = make_uchar4(tm_byte(255 * pow(r, correction)), tm_byte(255 * pow(g, correction)),
                       tm_byte(255 * pow(b, correction)), 255);
}

// Scratch memory of mpt_film_tonemap_noalloc, one copy per device for the
// life of the process: the tone map's partials (kTmGrid + 1 double2) and the
// 8-bit staging image of films of up to kNoAllocMaxPixels pixels.  Host code
// never passes `scratch` itself to a kernel: it looks the device address up
// once per device (scratch_setup) and passes that.
constexpr uint32_t kNoAllocMaxPixels = 1u << 22;
constexpr size_t kScratchPartialBytes = ((kTmGrid + 1) * sizeof(double2) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kScratchPartialBytes + (size_t)kNoAllocMaxPixels * sizeof(uchar4);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

__device__ __forceinline__ uint32_t chunk_of_target(const BatchDesc *bd, uint32_t t) {
  uint32_t lo = 0, hi = bd->hdr.n_chunks - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (bd->chunks[mid].tgt0 <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Filtered films: ReconstructionRenderOutput::addSample
// (renderers/outputs/ReconstructionRenderOutput.cpp:30-83) as a gather.  A
// sample splats onto the 3 x 3 pixels around the pixel it falls in (the
// reference's output filters keep that width whatever their support) with
// weight filter(x - X, y - Y) for the integer corner (x, y) of the target
// and the sample position (X, Y), when that weight is positive.  One thread
// per target pixel of a chunk sums its neighbours' samples in a fixed order
// (rows, columns, samples) in double and adds numerator and denominator to
// the film; a pixel near a chunk's edge also receives the neighbouring
// chunk's part (two additions, the same sum in either order).
__global__ __launch_bounds__(256) void k_film_filtered(const BatchDesc *__restrict__ bd,
                                                       const float4 *__restrict__ Ls,
                                                       const float2 *__restrict__ pos) {
  const uint32_t n = bd->hdr.n_targets;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const ChunkDev ch = bd->chunks[chunk_of_target(bd, t)];
    const JobDev &job = bd->jobs[ch.job];
    if (!job.film || !job.filter || t - ch.tgt0 >= ch.ntgt) continue;
    const uint32_t W = job.width, H = job.height, spp = job.spp;
    const uint32_t P = ch.trow0 * W + (t - ch.tgt0);
    const int px = (int)(P % W), py = (int)(P / W);
    const mpt_filter::Params fp = mpt_filter::resolve(job.filter - 1, job.filter_support, job.filter_p0,
                                                      job.filter_p1);
    double r = 0, g = 0, b = 0, wsum = 0;
    for (int dy = -1; dy <= 1; ++dy) {
      const int qy = py + dy;
      if (qy < 0 || qy >= (int)H) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        const int qx = px + dx;
        if (qx < 0 || qx >= (int)W) continue;
        const uint32_t Q = (uint32_t)qy * W + (uint32_t)qx;
        if (Q < ch.pix0 || Q - ch.pix0 >= ch.npix) continue;
        const uint64_t s0 = ch.slot0 + (uint64_t)(Q - ch.pix0) * spp;
        for (uint32_t s = 0; s < spp; ++s) {
          const float2 c = pos[s0 + s];
          const double w = mpt_filter::weight(fp, (double)px - (double)c.x, (double)py - (double)c.y);
          if (w > 0) {
            const float4 L = Ls[s0 + s];
            r += w * L.x;
            g += w * L.y;
            b += w * L.z;
            wsum += w;
          }
        }
      }
    }
    if (wsum > 0) {
      float4 *f = job.film + P;
      atomicAdd(&f->x, (float)r);
      atomicAdd(&f->y, (float)g);
      atomicAdd(&f->z, (float)b);
      atomicAdd(&f->w, (float)wsum);
    }
  }
}

// Camera and extension rays of a batch (shadow rays are counted by k_shadow).
__global__ void k_totals(const uint32_t *__restrict__ counts, uint32_t depth,
                         unsigned long long *__restrict__ totals) {
  if (threadIdx.x != 0) return;
  unsigned long long ext = 0;
  for (uint32_t d = 0; d < depth; ++d) ext += counts[d];
  totals[0] += counts[0];
  totals[1] += ext;
}

// Host rays as closest-hit queue entries (origin, -) (direction, no origin
// primitive), so the render's own closest-hit kernels can trace them.
__global__ __launch_bounds__(256) void k_rays_to_queue(const mpt_ray *__restrict__ rays, uint32_t n,
                                                       float4 *__restrict__ q) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const mpt_ray r = rays[i];
  q[3ull * i] = make_float4(r.o[0], r.o[1], r.o[2], __uint_as_float(i));
  q[3ull * i + 1] = make_float4(r.d[0], r.d[1], r.d[2], __uint_as_float(kMiss));
}

// Raw closest hits (FP32 t, leaf record) as (t recomputed in double,
// primitive id), as the shading kernel resolves them.
__global__ __launch_bounds__(256) void k_resolve_hits(Scene sc, const mpt_ray *__restrict__ rays,
                                                      const float2 *__restrict__ raw, uint32_t n,
                                                      mpt_hit *__restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 h = raw[i];
  const uint32_t rec = __float_as_uint(h.y);
  mpt_hit o;
  if (rec == kMiss) {
    o.t = INFINITY;
    o.prim = kMiss;
  } else {
    const float4 *rp = sc.recs + 3ull * rec;
    const mpt_ray r = rays[i];
    o.t = refine_hit(rp, f3(r.o[0], r.o[1], r.o[2]), f3(r.d[0], r.d[1], r.d[2]), h.x);
    o.prim = __float_as_uint(rp[0].w);
  }
  out[i] = o;
}

// EmitterSampler::sample and ::getPd (stats/samplers/EmitterSampler.cpp:
// 51-96): a light by emitted power, then the shape's getRandomPoint.  Sample
// i draws from Philox keyed by (seed, i): stream 0 gives the light's uniform
// and (u, v), stream 1 the mesh triangle's two further uniforms.
__device__ __forceinline__ float clamp_ref(float x, float lo, float hi) {
  // the reference's CLAMP: a NaN passes through
  return x >= hi ? hi : (x <= lo ? lo : x);
}

__global__ __launch_bounds__(256) void k_emitter_sample(Scene sc, uint32_t seed, uint64_t n,
                                                        mpt_emitter_point *__restrict__ out) {
  const uint32_t nl = sc.num_lights + sc.num_points;
  const float pdf_area = sc.emit_area > 0.0f ? 1.0f / sc.emit_area : 1.0f;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const mpt_rng::U4 r = mpt_rng::path_draw(seed, 0, (uint32_t)i, (uint32_t)(i >> 32), 0, 0);
    const float x = mpt_rng::unitf(r.x), u = mpt_rng::unitf(r.y), v = mpt_rng::unitf(r.z);
    uint32_t j = 0;
    while (j + 1 < nl && !(x <= sc.light_sel[j].x)) ++j;
    float3 lp = f3(0, 0, 0), ln = f3(0, 0, 0);
    if (j >= sc.num_lights) {
      // PointShape: its position, no normal
      lp = xyz(sc.points[2 * (j - sc.num_lights)]);
    } else {
      const LightDev &ld = sc.lights[j];
      if (ld.kind == LIGHT_SPHERE) {
        // Sphere::getRandomPoint (shapes/Sphere.cpp:107-115): a uniform point
        // s on the unit sphere, its (u, v) by Sphere::_getUV (:79-86), then
        // Sphere::getPosition (:117-124); asin leaves its domain for
        // |s_y| > 1/2 and the point is NaN there, as in the reference
        const float sz = 1.0f - 2.0f * u, rho = sqrtf(fmaxf(0.0f, 1.0f - sz * sz));
        const float phi = 2.0f * kPi * v;
        const float sx = rho * cosf(phi), sy = rho * sinf(phi);
        const float su = 1.0f - atan2f(sz, sx) * (0.5f / kPi);
        const float sv = 0.5f - asinf(2.0f * sy) * (1.0f / kPi);
        const float px = clamp_ref(tanf((1.0f - su) * (0.5f / kPi)), -0.5f, 0.5f);
        const float py = clamp_ref(0.5f * sinf((0.5f - sv) * (1.0f / kPi)), -0.5f, 0.5f);
        const float pz = sqrtf(1.0f - px * px - py * py);
        const float3 q = add(add(mul(xyz(ld.frame[0]), px), mul(xyz(ld.frame[1]), py)), mul(xyz(ld.frame[2]), pz));
        lp = add(xyz(sc.spheres[ld.index]), q);
        ln = mul(q, 1.0f / len(q));
      } else {
        // Shape::getRandomPoint (shapes/Shape.cpp:45-50): getPoint at two
        // uniforms (Plane::getPosition, or Mesh::getPosition with two more)
        const mpt_rng::U4 r2 = mpt_rng::path_draw(seed, 0, (uint32_t)i, (uint32_t)(i >> 32), 0, 1);
        uint32_t target;
        light_point(sc, ld, u, v, mpt_rng::unitf(r2.x), mpt_rng::unitf(r2.y), lp, ln, target);
      }
    }
    mpt_emitter_point e;
    e.p[0] = lp.x;
    e.p[1] = lp.y;
    e.p[2] = lp.z;
    e.light = j;
    e.n[0] = ln.x;
    e.n[1] = ln.y;
    e.n[2] = ln.z;
    e.pdf_area = pdf_area;
    e.pdf_select = sc.light_sel[j].y;
    e.pad = 0;
    out[i] = e;
  }
}

template <bool ANY>
__global__ __launch_bounds__(256) void k_trace_rays(Scene sc, const mpt_ray *__restrict__ rays,
                                                    const float *__restrict__ tmax, uint64_t n,
                                                    mpt_hit *__restrict__ hits, uint8_t *__restrict__ occ) {
  __shared__ uint2 stack[kStack * kBlock];
  uint2 *stk = stack + threadIdx.x;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const mpt_ray ry = rays[i];
    const RaySetup r = setup_ray(f3(ry.o[0], ry.o[1], ry.o[2]), f3(ry.d[0], ry.d[1], ry.d[2]));
    float t;
    uint32_t id;
    if (ANY) {
      occ[i] = trace<true>(sc, r, tmax[i], kMiss, stk, t, id) ? 1 : 0;
    } else {
      trace<false>(sc, r, INFINITY, kMiss, stk, t, id);
      hits[i].t = t;
      hits[i].prim = id;
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side: the batch engine.
// ---------------------------------------------------------------------------
namespace {

bool emissive_host(const mpt_material &m) { return m.Le[0] > 0 || m.Le[1] > 0 || m.Le[2] > 0; }

#include "bdpt.h"
#include "mlt.h"

int hip_fail(mpt_context *ctx, hipError_t e, const char *what) {
  return fail(ctx, MPT_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define MPT_HIP(ctx, call)                                  \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, #call); \
  } while (0)

#define MPT_LAUNCHED(ctx, what)                                  \
  do {                                                           \
    const hipError_t e_ = hipGetLastError();                     \
    if (e_ != hipSuccess) return hip_fail((ctx), e_, (what));    \
  } while (0)

template <class T>
int upload(mpt_context *ctx, DBuf<T> &buf, const T *host, size_t count) {
  MPT_HIP(ctx, buf.reserve(count));
  if (count) MPT_HIP(ctx, hipMemcpyAsync(buf.p, host, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return MPT_OK;
}

void normalize3(double *v) {
  const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= l;
  v[1] /= l;
  v[2] /= l;
}

uint32_t grid_for(const mpt_context *ctx, uint64_t n, uint32_t per_cu) {
  const uint64_t blocks = (n + kBlock - 1) / kBlock;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)ctx->num_cus * per_cu));
}

uint64_t g_scene_uid = 0;

// Device address of `scratch` per device, looked up once per device with that
// device current (the array's name is only a host-side stand-in).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// Called by mpt_open with `device` current.  A failed lookup is kept and
// reported by the entry point that needs the scratch array; nothing else
// depends on it.
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_done[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? ptr : nullptr;
  g_scratch_done[device] = true;
}

// The device address of this device's scratch array, or null (err set).
void *scratch_ptr(int device, hipError_t *err) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices || !g_scratch_done[device]) {
    *err = hipErrorNotInitialized;
    return nullptr;
  }
  *err = g_scratch_err[device];
  return g_scratch_ptr[device];
}



template <int STACK>
void launch_extend_t(mpt_context *ctx, hipStream_t st, uint32_t cap, uint32_t g, const Scene &sc, const float4 *q,
                     const uint32_t *count, float2 *hits, unsigned long long *counters, uint32_t *fetch) {
  if (ctx->persistent == 1) {
    if (ctx->tree_top) {
      if (counters)
        k_extend_persistent<true, STACK, true, 256><<<g, 256, 0, st>>>(sc, q, count, hits, counters, fetch);
      else
        k_extend_persistent<false, STACK, true, 256><<<g, 256, 0, st>>>(sc, q, count, hits, nullptr, fetch);
    } else if (ctx->trace_block == 128) {
      // the same waves in twice as many blocks
      if (counters)
        k_extend_persistent<true, STACK, false, 128><<<2 * g, 128, 0, st>>>(sc, q, count, hits, counters, fetch);
      else
        k_extend_persistent<false, STACK, false, 128><<<2 * g, 128, 0, st>>>(sc, q, count, hits, nullptr, fetch);
    } else {
      if (counters)
        k_extend_persistent<true, STACK, false, 256><<<g, 256, 0, st>>>(sc, q, count, hits, counters, fetch);
      else
        k_extend_persistent<false, STACK, false, 256><<<g, 256, 0, st>>>(sc, q, count, hits, nullptr, fetch);
    }
    return;
  }
  // form 0: grid-stride over a fixed grid of 8 blocks per CU; form 2: one
  // ray per lane over a grid sized by the batch capacity
  const uint32_t grid = ctx->persistent == 2 ? (uint32_t)((cap + kBlock - 1) / kBlock) : g;
  if (counters)
    k_extend<true, STACK><<<grid, kBlock, 0, st>>>(sc, q, count, hits, counters);
  else
    k_extend<false, STACK><<<grid, kBlock, 0, st>>>(sc, q, count, hits, nullptr);
}

// The short-stack depth is a run-time choice among compiled variants (it
// trades LDS per wave, hence occupancy, against kd-restarts).
void launch_extend(mpt_context *ctx, hipStream_t st, uint32_t cap, uint32_t g, const Scene &sc, const float4 *q,
                   const uint32_t *count, float2 *hits, unsigned long long *counters, uint32_t *fetch) {
  switch (ctx->stack) {
    case 8: launch_extend_t<8>(ctx, st, cap, g, sc, q, count, hits, counters, fetch); break;
    case 12: launch_extend_t<12>(ctx, st, cap, g, sc, q, count, hits, counters, fetch); break;
    case 24: launch_extend_t<24>(ctx, st, cap, g, sc, q, count, hits, counters, fetch); break;
    default: launch_extend_t<16>(ctx, st, cap, g, sc, q, count, hits, counters, fetch); break;
  }
}

void launch_shadow(mpt_context *ctx, hipStream_t st, uint32_t cap, uint32_t g, const Scene &sc, const BatchDesc *bd,
                   const float4 *sg, const float4 *sr, const uint32_t *count, float4 *Ls,
                   unsigned long long *total_rays) {
  // form 2: one group per lane over the group queue's capacity (one group per
  // path slot at most), as for closest hits
  if (ctx->persistent == 2) g = (uint32_t)(((uint64_t)cap + kBlock - 1) / kBlock);
  switch (ctx->stack) {
    case 8: k_shadow<8><<<g, kBlock, 0, st>>>(sc, bd, sg, sr, count, Ls, total_rays); break;
    case 12: k_shadow<12><<<g, kBlock, 0, st>>>(sc, bd, sg, sr, count, Ls, total_rays); break;
    case 24: k_shadow<24><<<g, kBlock, 0, st>>>(sc, bd, sg, sr, count, Ls, total_rays); break;
    default: k_shadow<16><<<g, kBlock, 0, st>>>(sc, bd, sg, sr, count, Ls, total_rays); break;
  }
}

// Blocks of a kernel resident on the whole device at once (at least one per
// CU), cached per kernel.
template <class K>
uint32_t resident_grid(mpt_context *ctx, K kernel, int block) {
  const void *key = reinterpret_cast<const void *>(kernel);
  for (const auto &e : ctx->grid_cache)
    if (e.first == key) return e.second;
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess || per_cu < 1) per_cu = 1;
  const uint32_t g = (uint32_t)ctx->num_cus * (uint32_t)per_cu;
  ctx->grid_cache.push_back({key, g});
  return g;
}

// Enqueues the kernel sequence of one batch (the part a graph captures).
int enqueue_batch(mpt_context *ctx, Lane &ln, mpt_scene *s, uint32_t D, bool multi, bool filtered,
                  bool instrument) {
  Workspace &w = ln.ws;
  const hipStream_t st = ln.stream;
  const uint32_t cap = w.cap, L = s->dev.num_lights + s->dev.num_points;
  const uint32_t g = grid_for(ctx, cap, 8);
  const uint32_t gcap = (cap + 255) / 256;
  int rc;
  auto tic = [&](int kind) -> int {
    if (!instrument || !ctx->timing) return MPT_OK;
    size_t used = 0;
    for (const auto &t : ctx->timed) used = std::max(used, t.second + 2);
    while (used + 2 > ctx->events.size()) {
      hipEvent_t e;
      MPT_HIP(ctx, hipEventCreate(&e));
      ctx->events.push_back(e);
    }
    ctx->timed.push_back({kind, used});
    MPT_HIP(ctx, hipEventRecord(ctx->events[used], st));
    return MPT_OK;
  };
  auto toc = [&]() -> int {
    if (!instrument || !ctx->timing) return MPT_OK;
    MPT_HIP(ctx, hipEventRecord(ctx->events[ctx->timed.back().second + 1], st));
    return MPT_OK;
  };
  k_reset<<<1, 64, 0, st>>>(w.counts.p, 3 * D + 1);
  k_raygen<<<gcap, 256, 0, st>>>(w.desc.p, w.q[0].p, w.pjs.p, filtered ? w.pos.p : nullptr, w.counts.p);
  MPT_LAUNCHED(ctx, "k_raygen");
  for (uint32_t d = 0; d < D; ++d) {
    float4 *qin = w.q[d & 1].p, *qout = w.q[(d + 1) & 1].p;
    if ((rc = tic(0))) return rc;
    launch_extend(ctx, st, cap, g, s->dev, qin, w.counts.p + d, w.hits.p,
                  instrument && ctx->counters ? w.totals.p + 3 : nullptr, w.counts.p + 2 * D + 1 + d);
    if ((rc = toc())) return rc;
    if ((rc = tic(1))) return rc;
    auto shade = [&](auto kern, int block) {
      const uint32_t sgrid = resident_grid(ctx, kern, block);
      kern<<<sgrid, block, 0, st>>>(s->dev, w.desc.p, d, qin, w.hits.p, w.counts.p + d, qout,
                                    w.counts.p + d + 1, w.sg.p, w.sr.p, w.counts.p + D + 1 + d, w.Ls.p,
                                    w.flags.p, w.pjs.p);
    };
    const bool env = s->dev.env_w != 0;
    if (ctx->shade_block == 512) {
      if (multi)
        env ? shade(k_shade<true, true, 512>, 512) : shade(k_shade<true, false, 512>, 512);
      else
        env ? shade(k_shade<false, true, 512>, 512) : shade(k_shade<false, false, 512>, 512);
    } else {
      if (multi)
        env ? shade(k_shade<true, true, 256>, 256) : shade(k_shade<true, false, 256>, 256);
      else
        env ? shade(k_shade<false, true, 256>, 256) : shade(k_shade<false, false, 256>, 256);
    }
    if ((rc = toc())) return rc;
    if (L) {
      if ((rc = tic(2))) return rc;
      launch_shadow(ctx, st, cap, g, s->dev, w.desc.p, w.sg.p, w.sr.p, w.counts.p + D + 1 + d, w.Ls.p,
                    w.totals.p + 2);
      if ((rc = toc())) return rc;
    }
    MPT_LAUNCHED(ctx, "depth kernels");
  }
  k_finalize<<<gcap, 256, 0, st>>>(w.desc.p, w.Ls.p, w.flags.p);
  k_film<<<gcap, 256, 0, st>>>(w.desc.p, w.Ls.p);
  if (filtered) k_film_filtered<<<gcap, 256, 0, st>>>(w.desc.p, w.Ls.p, w.pos.p);
  k_totals<<<1, 64, 0, st>>>(w.counts.p, D, w.totals.p);
  MPT_LAUNCHED(ctx, "film kernels");
  return MPT_OK;
}

// The kernel sequence of one bidirectional batch (the part a graph
// captures).  counts: eye subpath queue sizes per bounce [0, D], light
// subpath queue sizes [D + 1, 2D + 1], the pair queue's size, then the
// closest-hit claim cursors of the eye and the light bounces.
int enqueue_batch_bidir(mpt_context *ctx, Lane &ln, mpt_scene *s, uint32_t D, bool filtered, bool instrument) {
  Workspace &w = ln.ws;
  const hipStream_t st = ln.stream;
  const uint32_t cap = w.cap, V = 2 * (D + 1);
  const uint32_t g = grid_for(ctx, cap, 8);
  const uint32_t gcap = (cap + 255) / 256;
  uint32_t *cnt_eye = w.counts.p, *cnt_light = w.counts.p + (D + 1), *cnt_pairs = w.counts.p + 2 * (D + 1);
  uint32_t *cur_eye = cnt_pairs + 1, *cur_light = cur_eye + D;
  unsigned long long *ctr = instrument && ctx->counters ? w.totals.p + 3 : nullptr;
  int rc;
  auto tic = [&](int kind) -> int {
    if (!instrument || !ctx->timing) return MPT_OK;
    size_t used = 0;
    for (const auto &t : ctx->timed) used = std::max(used, t.second + 2);
    while (used + 2 > ctx->events.size()) {
      hipEvent_t e;
      MPT_HIP(ctx, hipEventCreate(&e));
      ctx->events.push_back(e);
    }
    ctx->timed.push_back({kind, used});
    MPT_HIP(ctx, hipEventRecord(ctx->events[used], st));
    return MPT_OK;
  };
  auto toc = [&]() -> int {
    if (!instrument || !ctx->timing) return MPT_OK;
    MPT_HIP(ctx, hipEventRecord(ctx->events[ctx->timed.back().second + 1], st));
    return MPT_OK;
  };
  k_reset<<<1, 64, 0, st>>>(w.counts.p, 4 * D + 3 + bdpt::kMaxVerts + 1);
  bdpt::k_bd_camera<<<gcap, 256, 0, st>>>(w.desc.p, w.q[0].p, w.pjs.p, filtered ? w.pos.p : nullptr, cnt_eye,
                                          w.bverts.p, w.binfo.p, V);
  MPT_LAUNCHED(ctx, "k_bd_camera");
  // the eye subpaths, then the light subpaths, one bounce per launch pair
  for (int side = 0; side < 2; ++side) {
    uint32_t *cnt = side == 0 ? cnt_eye : cnt_light, *cur = side == 0 ? cur_eye : cur_light;
    if (side == 1) {
      bdpt::k_bd_light_start<<<g, 256, 0, st>>>(s->dev, w.desc.p, w.q[0].p, cnt_light, w.bverts.p, w.binfo.p, V,
                                                w.pjs.p);
      MPT_LAUNCHED(ctx, "k_bd_light_start");
    }
    for (uint32_t b = 0; b < D; ++b) {
      float4 *qin = w.q[b & 1].p, *qout = w.q[(b + 1) & 1].p;
      if ((rc = tic(0))) return rc;
      launch_extend(ctx, st, cap, g, s->dev, qin, cnt + b, w.hits.p, ctr, cur + b);
      if ((rc = toc())) return rc;
      if ((rc = tic(1))) return rc;
      if (side == 0)
        bdpt::k_bd_step<true><<<g, 256, 0, st>>>(s->dev, w.desc.p, b, qin, w.hits.p, cnt + b, qout, cnt + b + 1,
                                                 w.bverts.p, w.binfo.p, V, w.pjs.p);
      else
        bdpt::k_bd_step<false><<<g, 256, 0, st>>>(s->dev, w.desc.p, b, qin, w.hits.p, cnt + b, qout, cnt + b + 1,
                                                  w.bverts.p, w.binfo.p, V, w.pjs.p);
      if ((rc = toc())) return rc;
      MPT_LAUNCHED(ctx, "subpath kernels");
    }
  }
  bdpt::k_bd_join<<<g, 256, 0, st>>>(s->dev, w.desc.p, w.bverts.p, w.binfo.p, V, w.pjs.p, w.bpairs.p, cnt_pairs,
                                     w.bmasks.p);
  MPT_LAUNCHED(ctx, "k_bd_join");
  if ((rc = tic(2))) return rc;
  switch (ctx->stack) {
    case 8: bdpt::k_bd_vis<8><<<g, 256, 0, st>>>(s->dev, w.bpairs.p, cnt_pairs, w.bverts.p, w.binfo.p, V, w.bmasks.p); break;
    case 12: bdpt::k_bd_vis<12><<<g, 256, 0, st>>>(s->dev, w.bpairs.p, cnt_pairs, w.bverts.p, w.binfo.p, V, w.bmasks.p); break;
    case 24: bdpt::k_bd_vis<24><<<g, 256, 0, st>>>(s->dev, w.bpairs.p, cnt_pairs, w.bverts.p, w.binfo.p, V, w.bmasks.p); break;
    default: bdpt::k_bd_vis<16><<<g, 256, 0, st>>>(s->dev, w.bpairs.p, cnt_pairs, w.bverts.p, w.binfo.p, V, w.bmasks.p); break;
  }
  if ((rc = toc())) return rc;
  MPT_LAUNCHED(ctx, "k_bd_vis");
  // slot indices by path length: combine's lanes then walk similar numbers
  // of splits
  uint32_t *len_bins = cur_light + D;
  bdpt::k_bd_len_count<<<g, 256, 0, st>>>(w.desc.p, w.binfo.p, len_bins);
  bdpt::k_bd_len_scan<<<1, 64, 0, st>>>(len_bins);
  bdpt::k_bd_len_scatter<<<g, 256, 0, st>>>(w.desc.p, w.binfo.p, len_bins, w.border.p);
  MPT_LAUNCHED(ctx, "path-length order");
  // per-lane arrays sized by the batch's path length: subpaths of 3, 5, 7
  // or 12 bounces
  auto combine = [&](auto kern) {
    kern<<<gcap, 256, 0, st>>>(s->dev, w.desc.p, w.bverts.p, w.binfo.p, V, w.bmasks.p, w.border.p, w.Ls.p);
  };
  if (V <= 8) combine(bdpt::k_bd_combine<8>);
  else if (V <= 12) combine(bdpt::k_bd_combine<12>);
  else if (V <= 16) combine(bdpt::k_bd_combine<16>);
  else combine(bdpt::k_bd_combine<bdpt::kMaxVerts>);
  MPT_LAUNCHED(ctx, "k_bd_combine");
  k_film<<<gcap, 256, 0, st>>>(w.desc.p, w.Ls.p);
  if (filtered) k_film_filtered<<<gcap, 256, 0, st>>>(w.desc.p, w.Ls.p, w.pos.p);
  bdpt::k_bd_totals<<<1, 64, 0, st>>>(w.counts.p, D, w.totals.p);
  MPT_LAUNCHED(ctx, "film kernels");
  return MPT_OK;
}

// Copies a batch plan into the next pinned staging descriptor and enqueues
// its upload.
int stage_plan(mpt_context *ctx, Lane &ln, const BatchPlan &plan) {
  const uint32_t slot = ln.staging_next++ & 3u;
  MPT_HIP(ctx, hipEventSynchronize(ln.staging_ev[slot]));
  BatchDesc &bd = ln.staging[slot];
  bd.hdr.n_chunks = (uint32_t)plan.chunks.size();
  bd.hdr.n_paths = plan.n_paths;
  bd.hdr.n_pixels = plan.n_pixels;
  bd.hdr.max_depth = plan.max_depth;
  bd.hdr.max_direct = std::max(1u, plan.max_direct);
  std::copy(plan.jobs.begin(), plan.jobs.end(), bd.jobs);
  std::copy(plan.chunks.begin(), plan.chunks.end(), bd.chunks);
  // the film pixels each filtered chunk's samples reach: whole rows from one
  // above its first pixel to one below its last
  uint32_t targets = 0;
  for (size_t c = 0; c < plan.chunks.size(); ++c) {
    ChunkDev &ch = bd.chunks[c];
    const JobDev &j = plan.jobs[ch.job];
    ch.tgt0 = targets;
    ch.trow0 = 0;
    ch.ntgt = 0;
    if (!j.filter || !ch.npix) continue;
    const uint32_t r0 = ch.pix0 / j.width, r1 = (ch.pix0 + ch.npix - 1) / j.width;
    ch.trow0 = r0 > 0 ? r0 - 1 : 0;
    const uint32_t last = std::min(r1 + 1, j.height - 1);
    const uint64_t n = (uint64_t)(last - ch.trow0 + 1) * j.width;
    if ((uint64_t)targets + n > 0xFFFFFFFFull) return fail(ctx, MPT_E_INVALID, "filtered batch too large");
    ch.ntgt = (uint32_t)n;
    targets += ch.ntgt;
  }
  bd.hdr.n_targets = targets;
  const size_t head = offsetof(BatchDesc, jobs) + plan.jobs.size() * sizeof(JobDev);
  MPT_HIP(ctx, hipMemcpyAsync(ln.ws.desc.p, &bd, head, hipMemcpyHostToDevice, ln.stream));
  MPT_HIP(ctx, hipMemcpyAsync(reinterpret_cast<char *>(ln.ws.desc.p) + offsetof(BatchDesc, chunks), bd.chunks,
                              plan.chunks.size() * sizeof(ChunkDev), hipMemcpyHostToDevice, ln.stream));
  MPT_HIP(ctx, hipEventRecord(ln.staging_ev[slot], ln.stream));
  return MPT_OK;
}

}  // namespace

namespace mpt_engine {

void drop_graphs(mpt_context *ctx, uint64_t scene_uid) {
  auto &c = ctx->graphs_cache;
  if (c.empty()) return;
  // a graph may still run on any lane
  for (Lane &ln : ctx->lanes)
    if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  for (auto it = c.begin(); it != c.end();) {
    if (scene_uid == 0 || (uint64_t)(uintptr_t)it->scene == scene_uid) {
      (void)hipGraphExecDestroy(it->exec);
      it = c.erase(it);
    } else {
      ++it;
    }
  }
}

int fail(mpt_context *ctx, int code, const std::string &msg) {
  if (ctx) {
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->err = msg;
  }
  return code;
}

// ThinLensCamera::setOrientation + _computeMatrix (cameras/ThinLensCamera.cpp:
// 181-256): n = -look, u = up x n, v = n x u, film scale tan(heightAngle/2).
JobDev make_job(const mpt_job &job, float4 *film) {
  const mpt_camera &c = job.camera;
  double n[3] = {-c.look[0], -c.look[1], -c.look[2]};
  double u[3] = {c.up[1] * n[2] - c.up[2] * n[1], c.up[2] * n[0] - c.up[0] * n[2], c.up[0] * n[1] - c.up[1] * n[0]};
  normalize3(n);
  normalize3(u);
  const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  JobDev j;
  std::memset(&j, 0, sizeof(j));
  for (int k = 0; k < 3; ++k) {
    j.cam.eye[k] = c.eye[k];
    j.cam.u[k] = (float)u[k];
    j.cam.v[k] = (float)v[k];
    j.cam.n[k] = (float)n[k];
  }
  const double th = std::tan(c.height_angle * M_PI / 360.0);
  j.cam.tan_h = (float)th;
  j.cam.tan_w = (float)(c.aspect * th);
  j.cam.aperture = c.aperture > 0 ? c.aperture : 0.0f;
  j.cam.focal_length = c.focal_length;
  j.cam.focal_distance = c.focal_distance;
  j.width = job.width;
  j.height = job.height;
  j.spp = job.spp;
  j.max_depth = job.max_depth;
  j.seed = job.seed;
  j.job_id = job.job_id;
  j.direct = std::max(1u, job.direct_samples);
  uint32_t side = std::max<uint32_t>(1, (uint32_t)std::sqrt((double)job.spp));
  while ((uint64_t)(side + 1) * (side + 1) <= job.spp) ++side;
  while ((uint64_t)side * side > job.spp) --side;
  j.side = side;
  j.film = film;
  j.filter = job.filter;
  j.filter_support = job.filter_support;
  j.filter_p0 = job.filter_params[0];
  j.filter_p1 = job.filter_params[1];
  j.direct_off = job.direct_off ? 1u : 0u;
  j.light_select = job.light_select ? 1u : 0u;
  j.renderer = job.renderer;
  return j;
}

int check_job(mpt_context *ctx, const mpt_job *job) {
  if (!job) return fail(ctx, MPT_E_INVALID, "job is NULL");
  if (job->renderer == MPT_RENDERER_MLT) {
    if (job->width == 0 || job->height == 0) return fail(ctx, MPT_E_INVALID, "job needs width and height > 0");
  } else if (job->width == 0 || job->height == 0 || job->spp == 0 || job->max_depth == 0) {
    return fail(ctx, MPT_E_INVALID, "job needs width, height, spp and max_depth > 0");
  }
  if (job->max_depth > 64) return fail(ctx, MPT_E_INVALID, "max_depth above 64");
  if ((uint64_t)job->width * job->height > (1ull << 30)) return fail(ctx, MPT_E_INVALID, "film too large");
  if (job->spp >= (1u << kJobBits)) return fail(ctx, MPT_E_INVALID, "spp at or above 2^20");
  if (job->direct_samples > kMaxDirect) return fail(ctx, MPT_E_INVALID, "direct_samples above 64");
  if (job->direct_off > 1) return fail(ctx, MPT_E_INVALID, "direct_off must be 0 or 1");
  if (job->light_select > 1) return fail(ctx, MPT_E_INVALID, "light_select must be 0 or 1");
  if (job->renderer > MPT_RENDERER_MLT) return fail(ctx, MPT_E_INVALID, "unknown renderer");
  if (job->renderer == MPT_RENDERER_MLT) {
    if (job->camera.aperture > 0)
      return fail(ctx, MPT_E_INVALID, "Metropolis light transport needs a pinhole camera (aperture 0)");
    if (job->mlt_chains == 0) return fail(ctx, MPT_E_INVALID, "mlt_chains must be at least 1");
    if (job->mlt_mutations > 32767)
      return fail(ctx, MPT_E_INVALID, "mlt_mutations above 32767 (the lens stratification's 16-bit counts)");
    if (job->mlt_max_depth < 2 || job->mlt_max_depth > mlt::kMaxLen)
      return fail(ctx, MPT_E_INVALID, "mlt_max_depth must be 2..16 path vertices");
    if ((uint64_t)job->width * job->height < 2) return fail(ctx, MPT_E_INVALID, "the film needs two or more pixels");
    if (job->filter) return fail(ctx, MPT_E_INVALID, "Metropolis light transport bins samples (filter 0)");
    if (job->mlt_seed_paths > (1u << 22)) return fail(ctx, MPT_E_INVALID, "mlt_seed_paths above 2^22");
    // a bin's proposal count is 32-bit and its fixed-point sums 64-bit
    // (at most 2^33 samples of at most 14 b each in units of b / 2^24)
    if ((uint64_t)job->mlt_chains * job->mlt_mutations >= (1ull << 32))
      return fail(ctx, MPT_E_INVALID, "mlt_chains x mlt_mutations at or above 2^32");
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(job->mlt_probs[k])) return fail(ctx, MPT_E_INVALID, "mlt_probs must be finite");
    const uint64_t strat = (uint64_t)job->mlt_chains * job->width * job->height * sizeof(uint16_t);
    if (strat > (64ull << 30))
      return fail(ctx, MPT_E_INVALID,
                  "mlt_chains x pixels x 2 B (the lens mutation's per-chain film counts) above 64 GiB");
  }
  if (job->renderer == MPT_RENDERER_BIDIR) {
    if (job->max_depth > bdpt::kMaxDepth)
      return fail(ctx, MPT_E_INVALID, "the bidirectional renderer takes max_depth up to 12");
    if (job->camera.aperture > 0)
      return fail(ctx, MPT_E_INVALID, "the bidirectional renderer needs a pinhole camera (aperture 0)");
  }
  const float *l = job->camera.look, *u = job->camera.up;
  const double ll = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
  const double cx = u[1] * l[2] - u[2] * l[1], cy = u[2] * l[0] - u[0] * l[2], cz = u[0] * l[1] - u[1] * l[0];
  if (!(ll > 0) || !(cx * cx + cy * cy + cz * cz > 0))
    return fail(ctx, MPT_E_INVALID, "camera look must be non-zero and not collinear with up");
  if (!(job->camera.height_angle > 0 && job->camera.height_angle < 180) || !(job->camera.aspect > 0))
    return fail(ctx, MPT_E_INVALID, "camera height_angle must be in (0, 180) and aspect > 0");
  const mpt_camera &c = job->camera;
  if (!(c.aperture >= 0) || !std::isfinite(c.aperture)) return fail(ctx, MPT_E_INVALID, "camera aperture must be >= 0");
  if (c.aperture > 0 && (!(c.focal_length > 0) || !std::isfinite(c.focal_length) || !(c.focal_distance > 0) ||
                         !std::isfinite(c.focal_distance)))
    return fail(ctx, MPT_E_INVALID, "a thin-lens camera needs focal_length > 0 and focal_distance > 0");
  if (job->filter > 1 + MPT_FILTER_LANCZOS) return fail(ctx, MPT_E_INVALID, "unknown film filter");
  if (job->filter && (job->filter_support > 64 || !std::isfinite(job->filter_params[0]) ||
                      !std::isfinite(job->filter_params[1]) ||
                      (job->filter == 1 + MPT_FILTER_GAUSSIAN && job->filter_params[0] < 0)))
    return fail(ctx, MPT_E_INVALID, "bad film filter parameters");
  return MPT_OK;
}

int resolve_focus(mpt_scene *s, mpt_job &job) {
  mpt_camera &c = job.camera;
  if (!(c.aperture > 0) || !c.autofocus) return MPT_OK;
  // ThinLensCamera::init (cameras/ThinLensCamera.cpp:97-118): the film point
  // (x - 1/2, 1/2 - y, -1) through filmToWorld, i.e. half the extent
  // getWorldRay gives the same film position; the hit distance if the ray
  // hits, else the given focal distance.
  double n[3] = {-c.look[0], -c.look[1], -c.look[2]};
  double u[3] = {c.up[1] * n[2] - c.up[2] * n[1], c.up[2] * n[0] - c.up[0] * n[2], c.up[0] * n[1] - c.up[1] * n[0]};
  normalize3(n);
  normalize3(u);
  const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
  const double th = std::tan(c.height_angle * M_PI / 360.0), tw = c.aspect * th;
  const double a = std::min(std::max((double)c.focal_point[0], 0.0), 1.0) - 0.5;
  const double b = 0.5 - std::min(std::max((double)c.focal_point[1], 0.0), 1.0);
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = a * tw * u[k] + b * th * v[k] - n[k];
  normalize3(d);
  mpt_ray ray;
  for (int k = 0; k < 3; ++k) {
    ray.o[k] = c.eye[k];
    ray.d[k] = (float)d[k];
  }
  mpt_hit hit;
  const int rc = mpt_trace_closest(s, &ray, 1, &hit);
  if (rc) return rc;
  if (hit.prim != MPT_MISS && std::isfinite(hit.t) && hit.t > 0) c.focal_distance = hit.t;
  c.autofocus = 0;
  return MPT_OK;
}

uint32_t batch_capacity(const mpt_context *ctx, const mpt_scene *s, uint32_t max_direct) {
  return (uint32_t)mpt_batch_paths_for(ctx->batch_paths, s->dev.num_lights, s->dev.num_points, max_direct);
}

uint64_t workspace_bytes(uint64_t cap, uint32_t depth, uint64_t shadow) {
  // q[2], hits, sg, sr, Ls, flags, pjs, pos, counts, totals, desc
  return cap * (2 * 3 * sizeof(float4) + sizeof(float2) + 2 * sizeof(float4) + sizeof(float4) + sizeof(uint32_t) +
                sizeof(uint2) + sizeof(float2)) +
         shadow * 2 * sizeof(float4) + (4ull * depth + 3 + bdpt::kMaxVerts + 1) * sizeof(uint32_t) +
         8 * sizeof(unsigned long long) +
         sizeof(BatchDesc);
}

uint64_t bidir_slot_bytes(uint32_t depth) {
  const uint64_t V = 2ull * (depth + 1), P = V * (V - 1) / 2;
  return V * bdpt::kVtx * sizeof(float4) + sizeof(uint4) + P * sizeof(uint2) + bdpt::kMaskWords * sizeof(uint32_t) +
         sizeof(uint32_t);
}

uint32_t bdpt_max_depth() { return bdpt::kMaxDepth; }

uint32_t bidir_capacity(const mpt_context *ctx, uint32_t depth, uint64_t lane_bytes) {
  const uint64_t V = 2ull * (depth + 1), P = V * (V - 1) / 2;
  uint64_t cap = std::min<uint64_t>(ctx->batch_paths, lane_bytes / bidir_slot_bytes(depth));
  // the pair queue's count is 32-bit
  cap = std::min<uint64_t>(cap, 0xFFFFFFFFull / P);
  cap = cap / 256 * 256;
  return (uint32_t)std::max<uint64_t>(cap, kMinBatchPaths);
}

int check_scene_job(mpt_context *ctx, const mpt_scene *s, const mpt_job *job) {
  if (job->renderer == MPT_RENDERER_PATH) return MPT_OK;
  const char *who = job->renderer == MPT_RENDERER_BIDIR ? "the bidirectional renderer" : "Metropolis light transport";
  // EmitterSampler::sample asserts a light (stats/samplers/EmitterSampler.cpp:51-53)
  if (s->dev.num_lights + s->dev.num_points == 0)
    return fail(ctx, MPT_E_INVALID, std::string(who) + " needs a light (an emitting quad, mesh or point)");
  // Sphere::getRandomPoint's uv round trip yields NaN points for half its
  // draws (shapes/Sphere.cpp:79-124)
  if (s->sphere_lights) return fail(ctx, MPT_E_INVALID, std::string(who) + " does not sample emitting spheres");
  return MPT_OK;
}

int ensure_lane(mpt_context *ctx, uint32_t k) {
  Lane &ln = ctx->lanes[k];
  if (ln.stream) return MPT_OK;
  MPT_HIP(ctx, hipStreamCreateWithFlags(&ln.stream, hipStreamNonBlocking));
  MPT_HIP(ctx, hipEventCreateWithFlags(&ln.mark, hipEventDisableTiming));
  MPT_HIP(ctx, hipHostMalloc((void **)&ln.staging, 4 * sizeof(BatchDesc), hipHostMallocDefault));
  for (hipEvent_t &e : ln.staging_ev) MPT_HIP(ctx, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return MPT_OK;
}

void release_lane(Lane &ln) {
  if (ln.stream) (void)hipStreamSynchronize(ln.stream);
  for (hipEvent_t &e : ln.staging_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  if (ln.staging) (void)hipHostFree(ln.staging);
  ln.staging = nullptr;
  if (ln.mark) (void)hipEventDestroy(ln.mark);
  ln.mark = nullptr;
  if (ln.stream) (void)hipStreamDestroy(ln.stream);
  ln.stream = nullptr;
}

int ensure_workspace(mpt_context *ctx, uint32_t k, uint32_t cap, uint32_t depth, uint32_t spv) {
  int rc = ensure_lane(ctx, k);
  if (rc) return rc;
  Lane &ln = ctx->lanes[k];
  Workspace &w = ln.ws;
  // the shadow buffer holds spv rays for each of this batch's slots
  const uint64_t shadow = std::max<uint64_t>(w.shadow_entries, (uint64_t)cap * spv);
  cap = std::max(cap, w.cap);
  depth = std::max(depth, w.depth);
  if (cap == w.cap && depth == w.depth && shadow == w.shadow_entries && w.desc.p) return MPT_OK;
  MPT_HIP(ctx, hipStreamSynchronize(ln.stream));
  MPT_HIP(ctx, w.q[0].reserve(3ull * cap));
  MPT_HIP(ctx, w.q[1].reserve(3ull * cap));
  MPT_HIP(ctx, w.hits.reserve(cap));
  MPT_HIP(ctx, w.Ls.reserve(cap));
  MPT_HIP(ctx, w.flags.reserve(cap));
  MPT_HIP(ctx, w.pjs.reserve(cap));
  MPT_HIP(ctx, w.sg.reserve(2ull * cap));
  MPT_HIP(ctx, w.sr.reserve(2ull * std::max<uint64_t>(shadow, 1)));
  MPT_HIP(ctx, w.counts.reserve(4ull * depth + 3 + bdpt::kMaxVerts + 1));
  MPT_HIP(ctx, w.desc.reserve(1));
  if (!w.totals.p) {
    // totals persist across reallocation: a render's counts are zeroed once
    // before its first batch
    MPT_HIP(ctx, w.totals.reserve(8));
    MPT_HIP(ctx, hipMemsetAsync(w.totals.p, 0, 8 * sizeof(unsigned long long), ln.stream));
  }
  w.cap = cap;
  w.depth = depth;
  w.shadow_entries = shadow;
  ++w.generation;
  drop_graphs(ctx, 0);
  return MPT_OK;
}

// Makes sure lane k holds the subpath records, pairs and masks of a
// bidirectional batch of `cap` paths with subpaths of `depth` bounces.
int ensure_bidir(mpt_context *ctx, uint32_t k, uint32_t cap, uint32_t depth) {
  Lane &ln = ctx->lanes[k];
  Workspace &w = ln.ws;
  const uint64_t V = 2ull * (depth + 1), P = V * (V - 1) / 2;
  const uint64_t nv = (uint64_t)cap * V * bdpt::kVtx, np = (uint64_t)cap * P, nm = (uint64_t)cap * bdpt::kMaskWords;
  if (w.bverts.n >= nv && w.binfo.n >= cap && w.bpairs.n >= np && w.bmasks.n >= nm && w.border.n >= cap)
    return MPT_OK;
  MPT_HIP(ctx, hipStreamSynchronize(ln.stream));
  MPT_HIP(ctx, w.bverts.reserve(std::max<uint64_t>(nv, w.bverts.n)));
  MPT_HIP(ctx, w.binfo.reserve(std::max<uint64_t>(cap, w.binfo.n)));
  MPT_HIP(ctx, w.bpairs.reserve(std::max<uint64_t>(np, w.bpairs.n)));
  MPT_HIP(ctx, w.bmasks.reserve(std::max<uint64_t>(nm, w.bmasks.n)));
  MPT_HIP(ctx, w.border.reserve(std::max<uint64_t>(cap, w.border.n)));
  ++w.generation;
  drop_graphs(ctx, 0);
  return MPT_OK;
}

uint32_t pick_lane(mpt_context *ctx) {
  const uint32_t n = std::max(1u, std::min(ctx->num_lanes, kMaxLanes));
  return ctx->next_lane++ % n;
}

int fork_lanes(mpt_context *ctx) {
  MPT_HIP(ctx, hipEventRecord(ctx->mark, ctx->stream));
  for (uint32_t k = 0; k < kMaxLanes; ++k) {
    if (k < ctx->num_lanes) {
      const int rc = ensure_lane(ctx, k);
      if (rc) return rc;
    }
    if (ctx->lanes[k].stream) MPT_HIP(ctx, hipStreamWaitEvent(ctx->lanes[k].stream, ctx->mark, 0));
  }
  return MPT_OK;
}

int lane_follow(mpt_context *ctx, uint32_t k) {
  const int rc = ensure_lane(ctx, k);
  if (rc) return rc;
  MPT_HIP(ctx, hipEventRecord(ctx->mark, ctx->stream));
  MPT_HIP(ctx, hipStreamWaitEvent(ctx->lanes[k].stream, ctx->mark, 0));
  return MPT_OK;
}

int join_lanes(mpt_context *ctx) {
  for (Lane &ln : ctx->lanes) {
    if (!ln.stream) continue;
    MPT_HIP(ctx, hipEventRecord(ln.mark, ln.stream));
    MPT_HIP(ctx, hipStreamWaitEvent(ctx->stream, ln.mark, 0));
  }
  return MPT_OK;
}

int reset_totals(mpt_context *ctx) {
  ctx->timed.clear();
  for (uint32_t k = 0; k < kMaxLanes; ++k) {
    Lane &ln = ctx->lanes[k];
    if (!ln.stream && k >= ctx->num_lanes) continue;
    int rc = ensure_lane(ctx, k);
    if (rc) return rc;
    MPT_HIP(ctx, ln.ws.totals.reserve(8));
    MPT_HIP(ctx, hipMemsetAsync(ln.ws.totals.p, 0, 8 * sizeof(unsigned long long), ctx->stream));
  }
  return MPT_OK;
}

int read_totals(mpt_context *ctx, unsigned long long out[5]) {
  unsigned long long tot[kMaxLanes][8];
  uint32_t used = 0;
  for (Lane &ln : ctx->lanes) {
    if (!ln.ws.totals.p) continue;
    MPT_HIP(ctx, hipMemcpyAsync(tot[used++], ln.ws.totals.p, sizeof(tot[0]), hipMemcpyDeviceToHost, ctx->stream));
  }
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < 5; ++i) {
    out[i] = 0;
    for (uint32_t k = 0; k < used; ++k) out[i] += tot[k][i];
  }
  return MPT_OK;
}

int run_batch(mpt_context *ctx, mpt_scene *s, const BatchPlan &plan, float *L_host, uint32_t lane) {
  if (plan.chunks.empty() || plan.n_paths == 0) return MPT_OK;
  if (plan.jobs.size() > kMaxJobs || plan.chunks.size() > kMaxChunks)
    return fail(ctx, MPT_E_INVALID, "batch plan exceeds the job or chunk table");
  if (lane >= kMaxLanes) return fail(ctx, MPT_E_INVALID, "bad lane");
  int rc;
  if (plan.bidir) {
    if (plan.max_depth > bdpt::kMaxDepth) return fail(ctx, MPT_E_INVALID, "bidirectional batch deeper than 12");
    const uint32_t cap = plan.capacity ? plan.capacity : bidir_capacity(ctx, plan.max_depth, kBidirLaneBytes);
    if ((rc = ensure_workspace(ctx, lane, std::max(plan.n_paths, cap), plan.max_depth, 1))) return rc;
    if ((rc = ensure_bidir(ctx, lane, std::max(plan.n_paths, cap), plan.max_depth))) return rc;
  } else {
    // shadow rays per vertex: one per light sample; their random streams are
    // numbered 1 + light * samples + sample below 256
    const uint32_t shadow_per_vertex = s->dev.num_lights * std::max(1u, plan.max_direct) + s->dev.num_points;
    if (shadow_per_vertex > 254) return fail(ctx, MPT_E_INVALID, "lights x direct_samples above 254");
    const uint32_t cap = plan.capacity ? plan.capacity : batch_capacity(ctx, s, plan.max_direct);
    if ((rc = ensure_workspace(ctx, lane, std::max(plan.n_paths, cap), plan.max_depth, shadow_per_vertex)))
      return rc;
  }
  Lane &ln = ctx->lanes[lane];
  if ((rc = stage_plan(ctx, ln, plan))) return rc;
  // injected failure: count this batch down, and fail it when the count ends
  for (uint32_t n = ctx->fail_batch.load(); n != 0;)
    if (ctx->fail_batch.compare_exchange_weak(n, n - 1)) {
      if (n == 1) return fail(ctx, MPT_E_DEVICE, "injected batch failure (MPT_OPT_FAIL_BATCH)");
      break;
    }
  const uint32_t D = plan.max_depth;
  // the multi-sample form also carries light selection by power
  bool multi = plan.max_direct > 1;
  for (const JobDev &j : plan.jobs) multi = multi || j.light_select != 0;
  bool filtered = false;
  for (const JobDev &j : plan.jobs) filtered = filtered || j.filter != 0;
  if (filtered && ln.ws.pos.n < ln.ws.cap) {
    // the sample positions of filtered films; a new buffer invalidates the
    // graphs that captured the old one
    MPT_HIP(ctx, hipStreamSynchronize(ln.stream));
    MPT_HIP(ctx, ln.ws.pos.reserve(ln.ws.cap));
    ++ln.ws.generation;
    drop_graphs(ctx, 0);
  }
  const bool instrument = ctx->timing || ctx->counters;
  if (ctx->graphs && !instrument) {
    Workspace &w = ln.ws;
    hipGraphExec_t exec = nullptr;
    const void *key = reinterpret_cast<const void *>((uintptr_t)s->uid);
    for (const GraphEntry &e : ctx->graphs_cache)
      if (e.scene == key && e.lane == lane && e.depth == D && e.cap == w.cap && e.lights == s->dev.num_lights &&
          e.multi == multi && e.filtered == filtered && e.bidir == plan.bidir && e.generation == w.generation) {
        exec = e.exec;
        break;
      }
    if (!exec) {
      hipGraph_t graph = nullptr;
      MPT_HIP(ctx, hipStreamBeginCapture(ln.stream, hipStreamCaptureModeThreadLocal));
      rc = plan.bidir ? enqueue_batch_bidir(ctx, ln, s, D, filtered, false)
                      : enqueue_batch(ctx, ln, s, D, multi, filtered, false);
      hipGraph_t captured = nullptr;
      const hipError_t ec = hipStreamEndCapture(ln.stream, &captured);
      if (rc) {
        if (captured) (void)hipGraphDestroy(captured);
        return rc;
      }
      if (ec != hipSuccess) return hip_fail(ctx, ec, "hipStreamEndCapture");
      graph = captured;
      const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (ei != hipSuccess) return hip_fail(ctx, ei, "hipGraphInstantiate");
      GraphEntry e;
      e.scene = key;
      e.lane = lane;
      e.depth = D;
      e.cap = w.cap;
      e.lights = s->dev.num_lights;
      e.multi = multi;
      e.filtered = filtered;
      e.bidir = plan.bidir;
      e.generation = w.generation;
      e.exec = exec;
      ctx->graphs_cache.push_back(e);
    }
    MPT_HIP(ctx, hipGraphLaunch(exec, ln.stream));
  } else {
    rc = plan.bidir ? enqueue_batch_bidir(ctx, ln, s, D, filtered, true)
                    : enqueue_batch(ctx, ln, s, D, multi, filtered, true);
    if (rc) return rc;
  }
  if (L_host) {
    MPT_HIP(ctx, hipMemcpy2DAsync(L_host, 3 * sizeof(float), ln.ws.Ls.p, sizeof(float4), 3 * sizeof(float),
                                  plan.n_paths, hipMemcpyDeviceToHost, ln.stream));
    MPT_HIP(ctx, hipStreamSynchronize(ln.stream));
  }
  return MPT_OK;
}

// Renders pixels [pix_begin, pix_end) of one job as consecutive batches of
// whole pixels.
int render_mlt(mpt_scene *s, const mpt_job *job, float4 *film, mpt_render_stats *stats, double *rec,
               uint32_t rec_steps, double *b_out, uint64_t *entries_out) {
  mpt_context *ctx = s->ctx;
  int rc = check_job(ctx, job);
  if (rc) return rc;
  if ((rc = check_scene_job(ctx, s, job))) return rc;
  using mlt::MltDesc;
  MltWork &w = ctx->mlt;
  const uint32_t W = job->width, H = job->height, P = W * H, C = job->mlt_chains;
  const uint32_t N = std::max(1u, job->mlt_seed_paths);
  hipStream_t st = ctx->stream;
  MltDesc d{};
  d.cam = make_job(*job, nullptr).cam;
  d.W = W;
  d.H = H;
  d.seed = job->seed;
  d.job_id = job->job_id;
  d.chains = C;
  d.mutations = job->mlt_mutations;
  d.seed_paths = N;
  d.max_len = job->mlt_max_depth;
  d.max_rejections = job->mlt_max_rejections;
  d.filter_proposed = (job->mlt_flags & 1u) ? 0u : 1u;
  {
    // MLTAggregatePathMutation::init (MLTAggregatePathMutation.cpp:34-91):
    // |weights|, all three when their sum is at most EPSILON, normalised
    double a[3] = {std::fabs((double)job->mlt_probs[0]), std::fabs((double)job->mlt_probs[1]),
                   std::fabs((double)job->mlt_probs[2])};
    if (a[0] + a[1] + a[2] <= 1e-7) a[0] = a[1] = a[2] = 1;
    d.n_mut = 0;
    for (uint32_t i = 0; i < 3; ++i)
      if (a[i] > 0) {
        d.mut_id[d.n_mut] = i;
        d.mut_w[d.n_mut++] = a[i];
      }
    double sum = 0;
    for (uint32_t i = d.n_mut; i--;) sum += d.mut_w[i];
    for (uint32_t i = d.n_mut; i--;) d.mut_w[i] /= sum;
  }
  // path records in runs of 64 lanes (mlt::Store)
  const uint64_t seed_recs = mlt::wave_runs(N) * mlt::kRun;
  MPT_HIP(ctx, w.seed_v.reserve(seed_recs));
  MPT_HIP(ctx, w.seed_tmp.reserve(seed_recs));
  MPT_HIP(ctx, w.seed_tmp2.reserve(seed_recs));
  MPT_HIP(ctx, w.seed_len.reserve(N));
  MPT_HIP(ctx, w.seed_att.reserve(N));
  MPT_HIP(ctx, w.seed_cnt.reserve(N));
  MPT_HIP(ctx, w.ent_w.reserve((uint64_t)N * mlt::kMaxEntries));
  MPT_HIP(ctx, w.ent_pos.reserve((uint64_t)N * mlt::kMaxEntries));
  MPT_HIP(ctx, w.ent_ks.reserve((uint64_t)N * mlt::kMaxEntries));
  MPT_HIP(ctx, w.cdf.reserve((uint64_t)N * mlt::kMaxEntries));
  MPT_HIP(ctx, w.order.reserve((uint64_t)N * mlt::kMaxEntries));
  MPT_HIP(ctx, w.chain_v.reserve(3 * mlt::wave_runs(C) * mlt::kRun));
  MPT_HIP(ctx, w.states.reserve(sizeof(mlt::ChainState) * (uint64_t)C));
  MPT_HIP(ctx, w.strat.reserve((uint64_t)C * P));
  MPT_HIP(ctx, w.num.reserve(3ull * P));
  MPT_HIP(ctx, w.proposed.reserve(P));
  MPT_HIP(ctx, w.counters.reserve(4));
  MPT_HIP(ctx, w.desc.reserve(sizeof(MltDesc)));
  const uint32_t steps = rec ? std::min(rec_steps, job->mlt_mutations) : job->mlt_mutations;
  if (rec) MPT_HIP(ctx, w.rec.reserve(std::max<uint64_t>(1, 6ull * C * steps)));
  mlt::MltBufs B;
  B.seed_v = w.seed_v.p;
  B.seed_tmp = w.seed_tmp.p;
  B.seed_tmp2 = w.seed_tmp2.p;
  B.seed_len = w.seed_len.p;
  B.seed_att = w.seed_att.p;
  B.seed_cnt = w.seed_cnt.p;
  B.ent_w = w.ent_w.p;
  B.ent_pos = w.ent_pos.p;
  B.ent_ks = w.ent_ks.p;
  B.cdf = w.cdf.p;
  B.order = w.order.p;
  B.chain_v = w.chain_v.p;
  B.st = reinterpret_cast<mlt::ChainState *>(w.states.p);
  B.strat = w.strat.p;
  B.num = w.num.p;
  B.proposed = w.proposed.p;
  B.counters = w.counters.p;
  MltDesc *dd = reinterpret_cast<MltDesc *>(w.desc.p);
  MPT_HIP(ctx, hipEventRecord(ctx->ev0, st));
  MPT_HIP(ctx, hipMemsetAsync(w.counters.p, 0, 4 * sizeof(unsigned long long), st));
  MPT_HIP(ctx, hipMemsetAsync(w.num.p, 0, 3ull * P * sizeof(unsigned long long), st));
  MPT_HIP(ctx, hipMemsetAsync(w.proposed.p, 0, (uint64_t)P * sizeof(uint32_t), st));
  MPT_HIP(ctx, hipMemsetAsync(w.strat.p, 0, (uint64_t)C * P * sizeof(uint16_t), st));
  // seed paths, round after round while no seed carries light
  // (MLTRenderer.cpp:144-205)
  bool found = false;
  for (uint32_t round = 0; round < mlt::kRounds && !found; ++round) {
    d.round = round;
    MPT_HIP(ctx, hipMemcpyAsync(dd, &d, sizeof(d), hipMemcpyHostToDevice, st));
    mlt::k_mlt_seed<<<(N + mlt::kSeedBlock - 1) / mlt::kSeedBlock, mlt::kSeedBlock, 0, st>>>(s->dev, dd, B);
    MPT_LAUNCHED(ctx, "k_mlt_seed");
    mlt::k_mlt_weights<<<1, mlt::kWeightsBlock, 0, st>>>(dd, B);
    MPT_LAUNCHED(ctx, "k_mlt_weights");
    MPT_HIP(ctx, hipMemcpyAsync(&d, dd, sizeof(d), hipMemcpyDeviceToHost, st));
    MPT_HIP(ctx, hipStreamSynchronize(st));
    found = d.has_seeds != 0;
    if (!found) d.n_prev += d.n_entries;
  }
  if (!found) return fail(ctx, MPT_E_INVALID, "no seed path of the Metropolis job carries light");
  const uint32_t gc = (C + mlt::kBlock - 1) / mlt::kBlock;
  mlt::k_mlt_start<<<gc, mlt::kBlock, 0, st>>>(s->dev, dd, B);
  MPT_LAUNCHED(ctx, "k_mlt_start");
  // runs of mutations: about 2^22 chain mutations per launch
  const uint32_t run = std::max<uint32_t>(1, std::min<uint64_t>(steps, (4ull << 20) / C));
  for (uint32_t j0 = 0; j0 < steps; j0 += run) {
    const uint32_t j1 = std::min(steps, j0 + run);
    mlt::k_mlt_step<<<gc, mlt::kBlock, 0, st>>>(s->dev, dd, B, j0, j1, rec ? w.rec.p : nullptr, steps);
    MPT_LAUNCHED(ctx, "k_mlt_step");
  }
  if (film) {
    mlt::k_mlt_finish<<<(P + 255) / 256, 256, 0, st>>>(dd, B, film);
    MPT_LAUNCHED(ctx, "k_mlt_finish");
  }
  MPT_HIP(ctx, hipEventRecord(ctx->ev1, st));
  unsigned long long ctr[4];
  MPT_HIP(ctx, hipMemcpyAsync(ctr, w.counters.p, sizeof(ctr), hipMemcpyDeviceToHost, st));
  if (rec) MPT_HIP(ctx, hipMemcpyAsync(rec, w.rec.p, 6ull * C * steps * sizeof(double), hipMemcpyDeviceToHost, st));
  MPT_HIP(ctx, hipStreamSynchronize(st));
  if (b_out) *b_out = d.b;
  if (entries_out) *entries_out = (uint64_t)d.n_entries + d.n_prev;
  if (stats) {
    float ms = 0;
    MPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::memset(stats, 0, sizeof(*stats));
    stats->paths = (uint64_t)C * steps;
    stats->extension_rays = ctr[2];
    stats->shadow_rays = ctr[3];
    stats->render_ms = ms;
    stats->mlt_accepted = ctr[1];
    stats->mlt_samples = ctr[0];
    stats->mlt_b = d.b;
  }
  return MPT_OK;
}

int render_range(mpt_scene *s, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end, float4 *film,
                 float *L_host, mpt_render_stats *stats) {
  mpt_context *ctx = s->ctx;
  int rc = check_job(ctx, job);
  if (rc) return rc;
  if (job->renderer == MPT_RENDERER_MLT)
    return fail(ctx, MPT_E_INVALID,
                "Metropolis light transport renders whole films (mpt_render, mpt_render_film), not pixel ranges");
  mpt_job focused = *job;
  if ((rc = resolve_focus(s, focused))) return rc;
  job = &focused;
  if ((rc = check_scene_job(ctx, s, job))) return rc;
  const uint32_t spp = job->spp;
  const bool bidir = job->renderer == MPT_RENDERER_BIDIR;
  const uint32_t cap = bidir ? bidir_capacity(ctx, job->max_depth, kBidirLaneBytes)
                             : batch_capacity(ctx, s, std::max(1u, job->direct_samples));
  const uint32_t batch_pix = std::max<uint32_t>(1, cap / spp);
  if ((rc = reset_totals(ctx))) return rc;
  MPT_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  if ((rc = fork_lanes(ctx))) return rc;
  BatchPlan plan;
  plan.jobs.push_back(make_job(*job, film));
  plan.max_depth = job->max_depth;
  plan.max_direct = plan.jobs[0].direct;
  plan.bidir = bidir;
  uint64_t batches = 0;
  for (uint32_t p0 = pix_begin; p0 < pix_end; p0 += batch_pix) {
    const uint32_t np = std::min(batch_pix, pix_end - p0);
    plan.chunks.assign(1, ChunkDev{0, p0, np, 0, 0, 0, 0, 0});
    plan.n_paths = np * spp;
    plan.n_pixels = np;
    if ((rc = run_batch(ctx, s, plan, L_host ? L_host + 3ull * (uint64_t)(p0 - pix_begin) * spp : nullptr,
                        pick_lane(ctx))))
      return rc;
    ++batches;
  }
  if ((rc = join_lanes(ctx))) return rc;
  MPT_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  unsigned long long tot[5];
  if ((rc = read_totals(ctx, tot))) return rc;
  if (stats) {
    float ms = 0;
    MPT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::memset(stats, 0, sizeof(*stats));
    stats->paths = (uint64_t)(pix_end - pix_begin) * spp;
    stats->camera_rays = tot[0];
    stats->extension_rays = tot[1];
    stats->shadow_rays = tot[2];
    stats->nodes_visited = tot[3];
    stats->prims_tested = tot[4];
    stats->render_ms = ms;
    stats->extend_launches = batches * job->max_depth;
    std::vector<std::pair<float, float>> span[3];   // per kind: [start, end) after ev0
    for (const auto &t : ctx->timed) {
      float e = 0, a = 0, b = 0;
      MPT_HIP(ctx, hipEventElapsedTime(&e, ctx->events[t.second], ctx->events[t.second + 1]));
      MPT_HIP(ctx, hipEventElapsedTime(&a, ctx->ev0, ctx->events[t.second]));
      MPT_HIP(ctx, hipEventElapsedTime(&b, ctx->ev0, ctx->events[t.second + 1]));
      (t.first == 0 ? stats->extend_ms : t.first == 1 ? stats->shade_ms : stats->shadow_ms) += e;
      span[t.first].push_back({a, b});
    }
    double *wall[3] = {&stats->extend_wall_ms, &stats->shade_wall_ms, &stats->shadow_wall_ms};
    for (int k = 0; k < 3; ++k) {
      // length of the union of the launch intervals
      std::sort(span[k].begin(), span[k].end());
      double total = 0, lo = 0, hi = -1;
      for (const auto &iv : span[k]) {
        if (iv.first > hi) {
          if (hi > lo) total += hi - lo;
          lo = iv.first;
          hi = iv.second;
        } else {
          hi = std::max(hi, (double)iv.second);
        }
      }
      if (hi > lo) total += hi - lo;
      *wall[k] = total;
    }
  }
  return MPT_OK;
}

}  // namespace mpt_engine

namespace {

// Host side of scene creation: validation, the device layouts and the kd
// build (used by mpt_scene_create and by the host-only kd entry points).
struct HostScene {
  std::vector<float4> verts, vnorm, sph, quads;
  std::vector<float4> tshade;   // device shading records, 3 per triangle
  std::vector<uint4> tris;
  std::vector<uint32_t> sphere_mat;
  std::vector<LightDev> lights;
  std::vector<float4> light_tris;   // 3 per triangle of an emitting mesh
  std::vector<mpt_material> mats;
  mpt_kd::Output ko;
  uint32_t nt = 0, ns = 0, nq = 0;
  uint32_t leaf_size = 0;   // effective kd leaf size
  float env[3] = {0, 0, 0};
  // environment image, device layout; kernel weights then their sum
  std::vector<float4> env_tex;
  std::vector<double> env_kernel;
  uint32_t env_w = 0, env_h = 0, env_m = 0;
  float env_rep_u = 1, env_rep_v = 1;
  std::vector<float4> points;   // 2 per point light, the device layout
  std::vector<float2> light_sel;   // Scene::light_sel
  double emit_area = 0;
};

int prepare_scene(mpt_context *ctx, const mpt_scene_desc *d, const mpt_kd_params *kd, HostScene &hs) {
  if (!d) return fail(ctx, MPT_E_INVALID, "NULL argument");
  if (d->version != MPT_ABI_VERSION) return fail(ctx, MPT_E_INVALID, "scene desc version mismatch");
  const uint32_t nv = d->num_vertices, nt = d->num_triangles, ns = d->num_spheres, nq = d->num_quads;
  const uint32_t nm = d->num_materials;
  if ((uint64_t)nt + ns + nq >= 0x3FFFFFFFull) return fail(ctx, MPT_E_INVALID, "too many primitives");
  if ((nt && (!d->indices || !d->tri_material || !d->vertices)) || (ns && (!d->spheres || !d->sphere_material)) ||
      (nq && (!d->quads || !d->quad_material)) || (nm && !d->materials))
    return fail(ctx, MPT_E_INVALID, "missing scene array");
  for (uint64_t i = 0; i < 3ull * nt; ++i)
    if (d->indices[i] >= nv) return fail(ctx, MPT_E_INVALID, "triangle index out of range");
  if ((d->vertex_normals == nullptr) != (d->vertex_normal_mask == nullptr))
    return fail(ctx, MPT_E_INVALID, "vertex_normals and vertex_normal_mask go together");
  if (d->tri_flags)
    for (uint32_t i = 0; i < nt; ++i)
      if (d->tri_flags[i] & ~(uint32_t)MPT_TRI_ONE_SIDED) return fail(ctx, MPT_E_INVALID, "unknown triangle flag");
  auto bad_mat = [&](uint32_t m) { return m >= nm; };
  for (uint32_t i = 0; i < nt; ++i) {
    if (bad_mat(d->tri_material[i])) return fail(ctx, MPT_E_INVALID, "triangle material out of range");
    // a standalone Triangle's light samples depend on its texture
    // coordinates (Triangle::getPosition, shapes/Triangle.cpp:330-353)
    if (emissive_host(d->materials[d->tri_material[i]]) && d->tri_flags && (d->tri_flags[i] & MPT_TRI_ONE_SIDED))
      return fail(ctx, MPT_E_INVALID, "a one-sided (standalone) triangle may not emit");
  }
  for (uint32_t i = 0; i < ns; ++i)
    if (bad_mat(d->sphere_material[i])) return fail(ctx, MPT_E_INVALID, "sphere material out of range");
  if (d->sphere_frames)
    for (uint64_t i = 0; i < 9ull * ns; ++i)
      if (!std::isfinite(d->sphere_frames[i])) return fail(ctx, MPT_E_INVALID, "bad sphere frame");
  for (uint32_t i = 0; i < nq; ++i)
    if (bad_mat(d->quad_material[i])) return fail(ctx, MPT_E_INVALID, "quad material out of range");
  if (d->num_points && !d->points) return fail(ctx, MPT_E_INVALID, "missing scene array");
  for (uint32_t i = 0; i < d->num_points; ++i) {
    const float *pt = d->points + 6ull * i;
    for (int k = 0; k < 6; ++k)
      if (!std::isfinite(pt[k]) || (k >= 3 && pt[k] < 0)) return fail(ctx, MPT_E_INVALID, "bad point light");
    // OmniEmitter::getLe (emitters/OmniEmitter.cpp:15-17) with the radiant
    // exitance of a zero-area shape equal to its power (Emitter.cpp:44-46)
    const double k4pi = 1.0 / (4.0 * M_PI);
    hs.points.push_back(make_float4(pt[0], pt[1], pt[2], 0.0f));
    hs.points.push_back(make_float4((float)(pt[3] * k4pi), (float)(pt[4] * k4pi), (float)(pt[5] * k4pi), 0.0f));
  }
  for (uint32_t i = 0; i < nm; ++i) {
    const mpt_material &mt = d->materials[i];
    if (mt.type > MPT_BSDF_DIELECTRIC) return fail(ctx, MPT_E_INVALID, "unknown BSDF type");
    if (!(mt.ior_g >= 0) || !(mt.ior_b >= 0)) return fail(ctx, MPT_E_INVALID, "negative per-channel IOR");
  }
  // Host copies in the device layouts.
  std::vector<float4> &verts = hs.verts, &vnorm = hs.vnorm;
  verts.resize(nv);
  vnorm.resize(nv);
  for (uint32_t i = 0; i < nv; ++i) verts[i] = make_float4(d->vertices[3 * i], d->vertices[3 * i + 1], d->vertices[3 * i + 2], 0);
  {
    // Mesh::computeNormals (shapes/Mesh.cpp:335-377) in double.
    std::vector<double> acc(3ull * nv, 0.0);
    for (uint32_t t = 0; t < nt; ++t) {
      const uint32_t *ix = d->indices + 3ull * t;
      const float *a = d->vertices + 3ull * ix[0], *b = d->vertices + 3ull * ix[1], *c = d->vertices + 3ull * ix[2];
      const double e1[3] = {(double)b[0] - a[0], (double)b[1] - a[1], (double)b[2] - a[2]};
      const double e2[3] = {(double)c[0] - a[0], (double)c[1] - a[1], (double)c[2] - a[2]};
      double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
      const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      if (l > 0)
        for (int k = 0; k < 3; ++k) n[k] /= l;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) acc[3ull * ix[j] + k] += n[k];
    }
    for (uint32_t i = 0; i < nv; ++i) {
      double *n = &acc[3ull * i];
      // a normal the mesh file gave wins over the averaged face normals
      if (d->vertex_normals && d->vertex_normal_mask[i])
        for (int k = 0; k < 3; ++k) n[k] = d->vertex_normals[3ull * i + k];
      const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      const double k = l > 0 ? 1.0 / l : 0.0;
      vnorm[i] = make_float4((float)(n[0] * k), (float)(n[1] * k), (float)(n[2] * k), 0);
    }
  }
  std::vector<uint4> &tris = hs.tris;
  tris.resize(nt);
  for (uint32_t t = 0; t < nt; ++t)
    tris[t] = make_uint4(d->indices[3 * t], d->indices[3 * t + 1], d->indices[3 * t + 2], d->tri_material[t]);
  hs.tshade.resize(3ull * nt);
  for (uint32_t t = 0; t < nt; ++t) {
    const float4 n0 = vnorm[tris[t].x], n1 = vnorm[tris[t].y], n2 = vnorm[tris[t].z];
    uint32_t mbits = tris[t].w;
    float mf;
    std::memcpy(&mf, &mbits, 4);
    hs.tshade[3ull * t] = make_float4(n0.x, n0.y, n0.z, mf);
    hs.tshade[3ull * t + 1] = make_float4(n1.x, n1.y, n1.z, 0);
    hs.tshade[3ull * t + 2] = make_float4(n2.x, n2.y, n2.z, 0);
  }
  std::vector<float4> &sph = hs.sph;
  sph.resize(ns);
  for (uint32_t i = 0; i < ns; ++i)
    sph[i] = make_float4(d->spheres[4 * i], d->spheres[4 * i + 1], d->spheres[4 * i + 2], d->spheres[4 * i + 3]);
  std::vector<float4> &quads = hs.quads;
  quads.resize(4ull * nq);
  for (uint32_t i = 0; i < nq; ++i) {
    const float *q = d->quads + 9ull * i;
    const double ex[3] = {q[3], q[4], q[5]}, ez[3] = {q[6], q[7], q[8]};
    double n[3] = {ez[1] * ex[2] - ez[2] * ex[1], ez[2] * ex[0] - ez[0] * ex[2], ez[0] * ex[1] - ez[1] * ex[0]};
    const double area = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (!(area > 0)) return fail(ctx, MPT_E_INVALID, "degenerate quad");
    normalize3(n);
    quads[4 * i] = make_float4(q[0], q[1], q[2], (float)area);
    quads[4 * i + 1] = make_float4(q[3], q[4], q[5], 0);
    quads[4 * i + 2] = make_float4(q[6], q[7], q[8], 0);
    uint32_t mbits = d->quad_material[i];
    float mf;
    std::memcpy(&mf, &mbits, 4);
    quads[4 * i + 3] = make_float4((float)n[0], (float)n[1], (float)n[2], mf);
  }
  // Area lights (Scene::getLights): emitting quads, then spheres, then
  // emitting meshes in the order of their first triangle.
  std::vector<LightDev> &lights = hs.lights;
  auto light = [](uint32_t kind, uint32_t index, uint32_t mat, double area) {
    LightDev l;
    std::memset(&l, 0, sizeof(l));
    l.kind = kind;
    l.index = index;
    l.mat = mat;
    l.area = (float)area;
    return l;
  };
  for (uint32_t i = 0; i < nq; ++i)
    if (emissive_host(d->materials[d->quad_material[i]]))
      lights.push_back(light(LIGHT_QUAD, i, d->quad_material[i], quads[4 * i].w));
  for (uint32_t i = 0; i < ns; ++i) {
    if (!emissive_host(d->materials[d->sphere_material[i]])) continue;
    // Sphere::_getSurfaceArea: 4 pi 0.25 whatever the scale (shapes/Sphere.cpp:138-141)
    LightDev l = light(LIGHT_SPHERE, i, d->sphere_material[i], M_PI);
    const float r2 = 2.0f * sph[i].w;
    for (int k = 0; k < 3; ++k) {
      const float *f = d->sphere_frames ? d->sphere_frames + 9ull * i + 3 * k : nullptr;
      l.frame[k] = f ? make_float4(f[0], f[1], f[2], 0) : make_float4(k == 0 ? r2 : 0, k == 1 ? r2 : 0, k == 2 ? r2 : 0, 0);
    }
    lights.push_back(l);
  }
  {
    // emitting triangles grouped by shape (tri_shape, else material); a
    // shape's world area is the sum of its triangles' (Mesh::_getSurfaceArea)
    std::vector<uint32_t> group_of(nt, kMiss);
    std::vector<uint32_t> order;               // groups by first triangle
    std::vector<std::vector<uint32_t>> members;
    std::vector<uint64_t> keys;
    for (uint32_t t = 0; t < nt; ++t) {
      if (!emissive_host(d->materials[d->tri_material[t]])) continue;
      const uint64_t key = d->tri_shape ? ((uint64_t)1 << 32) | d->tri_shape[t] : d->tri_material[t];
      uint32_t g = kMiss;
      for (uint32_t k = 0; k < keys.size(); ++k)
        if (keys[k] == key) g = k;
      if (g == kMiss) {
        g = (uint32_t)keys.size();
        keys.push_back(key);
        members.emplace_back();
      }
      if (d->tri_material[t] != d->tri_material[members[g].empty() ? t : members[g][0]])
        return fail(ctx, MPT_E_INVALID, "triangles of one emitting shape have different materials");
      members[g].push_back(t);
      group_of[t] = g;
    }
    for (uint32_t g = 0; g < members.size(); ++g) {
      double area = 0;
      for (uint32_t t : members[g]) {
        const float4 a = verts[tris[t].x], b = verts[tris[t].y], c = verts[tris[t].z];
        const double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
        const double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
        const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                     cz = e1[0] * e2[1] - e1[1] * e2[0];
        area += 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
      }
      const uint32_t li = (uint32_t)lights.size();
      LightDev l = light(LIGHT_MESH, 0, d->tri_material[members[g][0]], area);
      l.first = (uint32_t)(hs.light_tris.size() / 3);
      l.count = (uint32_t)members[g].size();
      for (uint32_t t : members[g]) {
        uint32_t tb = t;
        float tf;
        std::memcpy(&tf, &tb, 4);
        const float4 a = verts[tris[t].x], b = verts[tris[t].y], c = verts[tris[t].z];
        hs.light_tris.push_back(make_float4(a.x, a.y, a.z, tf));
        hs.light_tris.push_back(make_float4(b.x, b.y, b.z, 0));
        hs.light_tris.push_back(make_float4(c.x, c.y, c.z, 0));
        // the shading record carries the shape's area and light index + 1
        uint32_t lb = li + 1;
        float lf;
        std::memcpy(&lf, &lb, 4);
        hs.tshade[3ull * t + 1].w = (float)area;
        hs.tshade[3ull * t + 2].w = lf;
      }
      lights.push_back(l);
    }
  }
  {
    // EmitterSampler::init (stats/samplers/EmitterSampler.cpp:19-48): each
    // light's power averaged over the channels, normalized by the total;
    // the draw walks the running sum in light order.  An area light's power
    // is Le pi A (Le = P / (pi A), Emitter.cpp:44; OrientedEmitter.cpp:17-23),
    // a point light's the given P.
    std::vector<double> power;
    for (const LightDev &l : lights) {
      const mpt_material &lm = d->materials[l.mat];
      power.push_back((lm.Le[0] + (double)lm.Le[1] + lm.Le[2]) / 3.0 * M_PI * (double)l.area);
      hs.emit_area += (double)l.area;
    }
    for (uint32_t i = 0; i < d->num_points; ++i) {
      const float *pt = d->points + 6ull * i;
      power.push_back(((double)pt[3] + pt[4] + pt[5]) / 3.0);
    }
    double total = 0, run = 0;
    for (double w : power) total += w;
    for (double w : power) {
      const double share = total > 0 ? w / total : 0.0;
      run += share;
      hs.light_sel.push_back(make_float2((float)run, (float)share));
    }
  }

  // kd-tree over all primitives.
  std::vector<mpt_kd::Box> boxes((size_t)nt + ns + nq);
  for (uint32_t t = 0; t < nt; ++t) {
    mpt_kd::Box &b = boxes[t];
    for (int k = 0; k < 3; ++k) {
      const float v0 = (&verts[tris[t].x].x)[k], v1 = (&verts[tris[t].y].x)[k], v2 = (&verts[tris[t].z].x)[k];
      b.lo[k] = std::min(v0, std::min(v1, v2));
      b.hi[k] = std::max(v0, std::max(v1, v2));
    }
  }
  for (uint32_t i = 0; i < ns; ++i) {
    mpt_kd::Box &b = boxes[(size_t)nt + i];
    const float c[3] = {sph[i].x, sph[i].y, sph[i].z};
    for (int k = 0; k < 3; ++k) {
      b.lo[k] = std::nextafter(c[k] - sph[i].w, -INFINITY);
      b.hi[k] = std::nextafter(c[k] + sph[i].w, INFINITY);
    }
  }
  for (uint32_t i = 0; i < nq; ++i) {
    mpt_kd::Box &b = boxes[(size_t)nt + ns + i];
    const float4 c = quads[4 * i], ex = quads[4 * i + 1], ez = quads[4 * i + 2];
    for (int k = 0; k < 3; ++k) {
      const float cc = (&c.x)[k], a = 0.5f * std::fabs((&ex.x)[k]), e = 0.5f * std::fabs((&ez.x)[k]);
      b.lo[k] = std::nextafter(cc - a - e, -INFINITY);
      b.hi[k] = std::nextafter(cc + a + e, INFINITY);
    }
  }
  mpt_kd::Params kp;
  if (kd) {
    if (kd->leaf_size) kp.leaf_size = kd->leaf_size;
    if (kd->max_depth) kp.max_depth = kd->max_depth;
    if (kd->cost_traversal > 0) kp.cost_traversal = kd->cost_traversal;
    if (kd->cost_intersect > 0) kp.cost_intersect = kd->cost_intersect;
    if (kd->empty_bias > 0) kp.empty_bias = kd->empty_bias;
    kp.num_threads = kd->num_threads;
    if (kd->bins) kp.bins = kd->bins;
    kp.clip = kd->no_clip == 0;
  }
  hs.leaf_size = kp.leaf_size;
  std::vector<float> tri_verts(9ull * nt);
  for (uint32_t t = 0; t < nt; ++t) {
    const float4 a = verts[tris[t].x], b = verts[tris[t].y], c = verts[tris[t].z];
    const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
    std::memcpy(&tri_verts[9ull * t], v, sizeof(v));
  }
  mpt_kd::Output &ko = hs.ko;
  auto write = [&](uint32_t prim, float *r) {
    uint32_t kind, id = prim;
    std::memset(r, 0, 12 * sizeof(float));
    if (prim < nt) {
      kind = d->tri_flags && (d->tri_flags[prim] & MPT_TRI_ONE_SIDED) ? mpt_kd::KIND_TRIANGLE_ONE_SIDED
                                                                      : mpt_kd::KIND_TRIANGLE;
      const float4 a = verts[tris[prim].x], b = verts[tris[prim].y], c = verts[tris[prim].z];
      r[0] = a.x; r[1] = a.y; r[2] = a.z;
      r[4] = b.x; r[5] = b.y; r[6] = b.z;
      r[8] = c.x; r[9] = c.y; r[10] = c.z;
    } else if (prim < nt + ns) {
      kind = mpt_kd::KIND_SPHERE;
      const float4 sp = sph[prim - nt];
      r[0] = sp.x; r[1] = sp.y; r[2] = sp.z;
      r[4] = sp.w;
    } else {
      kind = mpt_kd::KIND_QUAD;
      const float4 *q = &quads[4ull * (prim - nt - ns)];
      r[0] = q[0].x; r[1] = q[0].y; r[2] = q[0].z;
      r[4] = q[1].x; r[5] = q[1].y; r[6] = q[1].z;
      r[8] = q[2].x; r[9] = q[2].y; r[10] = q[2].z;
    }
    std::memcpy(&r[3], &id, 4);
    std::memcpy(&r[7], &kind, 4);
  };
  try {
    mpt_kd::build(boxes, kp, write, ko, tri_verts.data(), nt);
  } catch (const std::bad_alloc &) {
    return fail(ctx, MPT_E_NOMEM, "kd build: host allocation");
  }
  if (ko.too_large) return fail(ctx, MPT_E_INVALID, "kd tree above 2^29 nodes");
  hs.nt = nt;
  hs.ns = ns;
  hs.nq = nq;
  hs.sphere_mat.assign(d->sphere_material, d->sphere_material + ns);
  hs.mats.assign(d->materials, d->materials + nm);
  for (int k = 0; k < 3; ++k) hs.env[k] = d->env_radiance[k];
  return MPT_OK;
}

// The environment image of a scene description, in the device layout.
int prepare_env(mpt_context *ctx, const mpt_scene_desc *d, HostScene &hs) {
  if (!d->env_map) return MPT_OK;
  const mpt_env_map &e = *d->env_map;
  if (!e.width || !e.height || !e.rgb) return fail(ctx, MPT_E_INVALID, "environment map without texels");
  if ((uint64_t)e.width * e.height > (1ull << 28)) return fail(ctx, MPT_E_INVALID, "environment map too large");
  std::string err;
  if (!mpt_env::kernel(e, hs.env_kernel, hs.env_m, err)) return fail(ctx, MPT_E_INVALID, err);
  const size_t n = (size_t)e.width * e.height;
  hs.env_tex.resize(n);
  for (size_t i = 0; i < n; ++i) hs.env_tex[i] = make_float4(e.rgb[3 * i], e.rgb[3 * i + 1], e.rgb[3 * i + 2], 0);
  hs.env_w = e.width;
  hs.env_h = e.height;
  hs.env_rep_u = e.repeat_u != 0 ? e.repeat_u : 1.0f;
  hs.env_rep_v = e.repeat_v != 0 ? e.repeat_v : 1.0f;
  return MPT_OK;
}

// Points a device-layout Scene at host (or device) arrays.
void bind_scene(Scene &sc, const uint2 *nodes, const float4 *recs, const mpt_kd::Box &box, const float4 *tshade,
                const float4 *sph, const uint32_t *sph_mat,
                const float4 *quads, const mpt_material *mats, const LightDev *lights, uint32_t nl,
                const float4 *light_tris, uint32_t nt,
                uint32_t ns, uint32_t nq, const float *env) {
  sc.nodes = nodes;
  sc.recs = recs;
  for (int k = 0; k < 3; ++k) {
    sc.box_lo[k] = box.lo[k];
    sc.box_hi[k] = box.hi[k];
  }
  sc.tshade = tshade;
  sc.spheres = sph;
  sc.sphere_mat = sph_mat;
  sc.quads = quads;
  sc.mats = mats;
  sc.lights = lights;
  sc.num_lights = nl;
  sc.light_tris = light_tris;
  sc.nt = nt;
  sc.ns = ns;
  sc.nq = nq;
  for (int k = 0; k < 3; ++k) sc.env[k] = env[k];
}

}  // namespace

struct mpt_kd_tree {
  HostScene hs;
  Scene sc{};
};

extern "C" {

int mpt_abi_version(void) { return (int)MPT_ABI_VERSION; }

uint64_t mpt_batch_paths_for(uint64_t batch_paths, uint32_t num_lights, uint32_t num_points,
                             uint32_t direct_samples) {
  const uint64_t spv = (uint64_t)num_lights * std::max(1u, direct_samples) + num_points;
  if (spv <= 1) return batch_paths;
  const uint64_t fit = std::max<uint64_t>(kMinBatchPaths, kShadowRayBudget / spv);
  return std::min<uint64_t>(batch_paths, fit);
}

int mpt_open(int device, mpt_context **out) {
  if (!out) return MPT_E_INVALID;
  *out = nullptr;
  std::unique_ptr<mpt_context> ctx(new (std::nothrow) mpt_context());
  if (!ctx) return MPT_E_NOMEM;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return MPT_E_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return MPT_E_DEVICE;
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
    ctx->num_cus = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return MPT_E_DEVICE;
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) return MPT_E_DEVICE;
  if (hipEventCreateWithFlags(&ctx->mark, hipEventDisableTiming) != hipSuccess) return MPT_E_DEVICE;
  scratch_setup(device);
  for (uint32_t k = 0; k < ctx->num_lanes; ++k)
    if (ensure_lane(ctx.get(), k)) return MPT_E_DEVICE;
  *out = ctx.release();
  return MPT_OK;
}

int mpt_close(mpt_context *ctx) {
  if (!ctx) return MPT_E_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  while (!ctx->servers.empty()) mpt_server_destroy(ctx->servers.back());
  while (!ctx->scenes.empty()) mpt_scene_destroy(ctx->scenes.back());
  while (!ctx->films.empty()) mpt_film_destroy(ctx->films.back());
  drop_graphs(ctx, 0);
  for (Lane &ln : ctx->lanes) release_lane(ln);
  if (ctx->mark) (void)hipEventDestroy(ctx->mark);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MPT_OK;
}

// The message is copied per calling thread, so a server's worker thread
// recording a failure does not change a string the caller is reading.
const char *mpt_last_error(const mpt_context *ctx) {
  if (!ctx) return "no context";
  thread_local std::string copy;
  std::lock_guard<std::mutex> lk(const_cast<mpt_context *>(ctx)->err_mu);
  copy = ctx->err;
  return copy.c_str();
}

int mpt_set_option(mpt_context *ctx, uint32_t option, uint64_t value) {
  if (!ctx) return MPT_E_INVALID;
  switch (option) {
    case MPT_OPT_BATCH_PATHS:
      if (value < 1024 || value > (1ull << 28)) return fail(ctx, MPT_E_INVALID, "batch paths out of [1024, 2^28]");
      ctx->batch_paths = (uint32_t)value;
      return MPT_OK;
    case MPT_OPT_TIMING:
      ctx->timing = value != 0;
      return MPT_OK;
    case MPT_OPT_COUNTERS:
      ctx->counters = value != 0;
      return MPT_OK;
    case MPT_OPT_GRAPHS:
      ctx->graphs = value != 0;
      return MPT_OK;
    case MPT_OPT_PERSISTENT:
      if (value > 2) return fail(ctx, MPT_E_INVALID, "persistent must be 0, 1 or 2");
      ctx->persistent = (uint32_t)value;
      drop_graphs(ctx, 0);
      return MPT_OK;
    case MPT_OPT_LANES:
      if (value < 1 || value > kMaxLanes) return fail(ctx, MPT_E_INVALID, "lanes must be 1..4");
      ctx->num_lanes = (uint32_t)value;
      ctx->next_lane = 0;
      return MPT_OK;
    case MPT_OPT_TREE_TOP:
      if (value > 1) return fail(ctx, MPT_E_INVALID, "tree_top must be 0 or 1");
      ctx->tree_top = value != 0;
      drop_graphs(ctx, 0);
      return MPT_OK;
    case MPT_OPT_TRACE_BLOCK:
      if (value != 128 && value != 256) return fail(ctx, MPT_E_INVALID, "trace block must be 128 or 256");
      ctx->trace_block = (uint32_t)value;
      drop_graphs(ctx, 0);
      return MPT_OK;
    case MPT_OPT_SHADE_BLOCK:
      if (value != 256 && value != 512) return fail(ctx, MPT_E_INVALID, "shade block must be 256 or 512");
      ctx->shade_block = (uint32_t)value;
      drop_graphs(ctx, 0);
      return MPT_OK;
    case MPT_OPT_FAIL_BATCH:
      if (value > 0xffffffffull) return fail(ctx, MPT_E_INVALID, "fail_batch out of range");
      ctx->fail_batch.store((uint32_t)value);
      return MPT_OK;
    case MPT_OPT_STACK:
      if (value != 8 && value != 12 && value != 16 && value != 24)
        return fail(ctx, MPT_E_INVALID, "stack depth must be 8, 12, 16 or 24");
      ctx->stack = (uint32_t)value;
      drop_graphs(ctx, 0);
      return MPT_OK;
    default:
      return fail(ctx, MPT_E_INVALID, "unknown option");
  }
}

int mpt_get_option(mpt_context *ctx, uint32_t option, uint64_t *value) {
  if (!ctx || !value) return MPT_E_INVALID;
  switch (option) {
    case MPT_OPT_BATCH_PATHS: *value = ctx->batch_paths; return MPT_OK;
    case MPT_OPT_TIMING: *value = ctx->timing; return MPT_OK;
    case MPT_OPT_COUNTERS: *value = ctx->counters; return MPT_OK;
    case MPT_OPT_GRAPHS: *value = ctx->graphs; return MPT_OK;
    case MPT_OPT_STACK: *value = ctx->stack; return MPT_OK;
    case MPT_OPT_PERSISTENT: *value = ctx->persistent; return MPT_OK;
    case MPT_OPT_LANES: *value = ctx->num_lanes; return MPT_OK;
    case MPT_OPT_FAIL_BATCH: *value = ctx->fail_batch.load(); return MPT_OK;
    case MPT_OPT_TREE_TOP: *value = ctx->tree_top; return MPT_OK;
    case MPT_OPT_TRACE_BLOCK: *value = ctx->trace_block; return MPT_OK;
    case MPT_OPT_SHADE_BLOCK: *value = ctx->shade_block; return MPT_OK;
    default: return fail(ctx, MPT_E_INVALID, "unknown option");
  }
}

int mpt_sync(mpt_context *ctx) {
  if (!ctx) return MPT_E_INVALID;
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_scene_create(mpt_context *ctx, const mpt_scene_desc *d, const mpt_kd_params *kd, mpt_scene **out) {
  if (!ctx || !d || !out) return fail(ctx, MPT_E_INVALID, "NULL argument");
  *out = nullptr;
  HostScene hs;
  int rc = prepare_scene(ctx, d, kd, hs);
  if (!rc) rc = prepare_env(ctx, d, hs);
  if (rc) return rc;
  MPT_HIP(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<mpt_scene> s(new (std::nothrow) mpt_scene());
  if (!s) return fail(ctx, MPT_E_NOMEM, "host allocation");
  s->ctx = ctx;
  const mpt_kd::Output &ko = hs.ko;
  const uint32_t nt = hs.nt, ns = hs.ns, nq = hs.nq;
  if ((rc = upload(ctx, s->nodes, reinterpret_cast<const uint2 *>(ko.nodes.data()), ko.nodes.size() / 2))) return rc;
  if ((rc = upload(ctx, s->recs, reinterpret_cast<const float4 *>(ko.records.data()), ko.records.size() / 4))) return rc;
  if ((rc = upload(ctx, s->tshade, hs.tshade.data(), hs.tshade.size()))) return rc;
  if ((rc = upload(ctx, s->spheres, hs.sph.data(), hs.sph.size()))) return rc;
  if ((rc = upload(ctx, s->sphere_mat, hs.sphere_mat.data(), hs.sphere_mat.size()))) return rc;
  if ((rc = upload(ctx, s->quads, hs.quads.data(), hs.quads.size()))) return rc;
  if ((rc = upload(ctx, s->mats, hs.mats.data(), hs.mats.size()))) return rc;
  if ((rc = upload(ctx, s->lights, hs.lights.data(), hs.lights.size()))) return rc;
  s->sphere_lights = 0;
  for (const LightDev &l : hs.lights) s->sphere_lights += l.kind == LIGHT_SPHERE ? 1u : 0u;
  if ((rc = upload(ctx, s->light_tris, hs.light_tris.data(), hs.light_tris.size()))) return rc;
  if ((rc = upload(ctx, s->light_sel, hs.light_sel.data(), hs.light_sel.size()))) return rc;
  if ((rc = upload(ctx, s->env_tex, hs.env_tex.data(), hs.env_tex.size()))) return rc;
  if ((rc = upload(ctx, s->env_kernel, hs.env_kernel.data(), hs.env_kernel.size()))) return rc;
  if ((rc = upload(ctx, s->points, hs.points.data(), hs.points.size()))) return rc;
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  bind_scene(s->dev, s->nodes.p, s->recs.p, ko.box, s->tshade.p, s->spheres.p,
             s->sphere_mat.p, s->quads.p, s->mats.p, s->lights.p, (uint32_t)hs.lights.size(), s->light_tris.p, nt, ns,
             nq, hs.env);
  s->dev.env_tex = s->env_tex.p;
  s->dev.env_kernel = s->env_kernel.p;
  s->dev.env_w = hs.env_w;
  s->dev.env_h = hs.env_h;
  s->dev.env_m = hs.env_m;
  s->dev.env_rep_u = hs.env_rep_u;
  s->dev.env_rep_v = hs.env_rep_v;
  s->dev.points = s->points.p;
  s->dev.num_points = (uint32_t)(hs.points.size() / 2);
  s->dev.light_sel = s->light_sel.p;
  s->dev.top_bytes = (uint32_t)(std::min<uint64_t>(ko.top_slots, kTopNodes) * sizeof(uint2));
  s->dev.emit_area = (float)hs.emit_area;
  s->stats[0] = ko.n_nodes;
  s->stats[1] = ko.n_leaves;
  s->stats[2] = ko.n_refs;
  s->stats[3] = ko.max_depth;
  s->stats[4] = s->device_bytes();
  s->stats[5] = (uint64_t)ko.build_ms;
  s->stats[6] = ko.max_leaf;
  s->stats[7] = nt;
  s->leaf_size = hs.leaf_size;
  s->uid = ++g_scene_uid;
  ctx->scenes.push_back(s.get());
  *out = s.release();
  return MPT_OK;
}

int mpt_kd_build(const mpt_scene_desc *d, const mpt_kd_params *kd, mpt_kd_tree **out) {
  if (!d || !out) return MPT_E_INVALID;
  *out = nullptr;
  std::unique_ptr<mpt_kd_tree> t(new (std::nothrow) mpt_kd_tree());
  if (!t) return MPT_E_NOMEM;
  int rc;
  try {
    rc = prepare_scene(nullptr, d, kd, t->hs);
  } catch (const std::bad_alloc &) {
    return MPT_E_NOMEM;
  }
  if (rc) return rc;
  HostScene &hs = t->hs;
  bind_scene(t->sc, reinterpret_cast<const uint2 *>(hs.ko.nodes.data()),
             reinterpret_cast<const float4 *>(hs.ko.records.data()), hs.ko.box, hs.tshade.data(),
             hs.sph.data(), hs.sphere_mat.data(), hs.quads.data(), hs.mats.data(),
             hs.lights.data(), (uint32_t)hs.lights.size(), hs.light_tris.data(), hs.nt, hs.ns, hs.nq, hs.env);
  *out = t.release();
  return MPT_OK;
}

int mpt_kd_stats(const mpt_kd_tree *t, uint64_t out[6]) {
  if (!t || !out) return MPT_E_INVALID;
  const mpt_kd::Output &ko = t->hs.ko;
  out[0] = ko.n_nodes;
  out[1] = ko.n_leaves;
  out[2] = ko.n_refs;
  out[3] = ko.max_depth;
  out[4] = (uint64_t)ko.build_ms;
  out[5] = ko.max_leaf;
  return MPT_OK;
}

int mpt_kd_trace(const mpt_kd_tree *t, const mpt_ray *rays, uint64_t n, mpt_hit *hits, uint64_t counters[2],
                 uint32_t nthreads) {
  if (!t || (n && (!rays || !hits))) return MPT_E_INVALID;
  const unsigned nth = std::max(1u, std::min(nthreads ? nthreads : std::thread::hardware_concurrency(), 64u));
  std::vector<uint64_t> nodes(nth, 0), refs(nth, 0);
  std::atomic<uint64_t> next(0);
  auto work = [&](unsigned tid) {
    // The LDS ring of a device lane becomes a strided host buffer.
    std::vector<uint2> stk((size_t)8 * kBlock);
    for (;;) {
      const uint64_t b = next.fetch_add(4096);
      if (b >= n) break;
      const uint64_t e = std::min<uint64_t>(n, b + 4096);
      for (uint64_t i = b; i < e; ++i) {
        const mpt_ray &ry = rays[i];
        const RaySetup r = setup_ray(make_float3(ry.o[0], ry.o[1], ry.o[2]), make_float3(ry.d[0], ry.d[1], ry.d[2]));
        float tt;
        uint32_t id, nn = 0, nr = 0;
        trace<false, true, 8>(t->sc, r, INFINITY, kMiss, stk.data(), tt, id, &nn, &nr);
        hits[i].t = tt;
        hits[i].prim = id;
        nodes[tid] += nn;
        refs[tid] += nr;
      }
    }
  };
  std::vector<std::thread> th;
  for (unsigned k = 1; k < nth; ++k) th.emplace_back(work, k);
  work(0);
  for (auto &x : th) x.join();
  if (counters) {
    counters[0] = counters[1] = 0;
    for (unsigned k = 0; k < nth; ++k) {
      counters[0] += nodes[k];
      counters[1] += refs[k];
    }
  }
  return MPT_OK;
}

int mpt_kd_destroy(mpt_kd_tree *t) {
  delete t;
  return MPT_OK;
}

int mpt_scene_destroy(mpt_scene *s) {
  if (!s) return MPT_E_INVALID;
  mpt_context *ctx = s->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->scenes.erase(std::remove(ctx->scenes.begin(), ctx->scenes.end(), s), ctx->scenes.end());
  drop_graphs(ctx, s->uid);
  delete s;
  return MPT_OK;
}

int mpt_scene_stats(const mpt_scene *s, uint64_t out[8]) {
  if (!s || !out) return MPT_E_INVALID;
  std::memcpy(out, s->stats, sizeof(s->stats));
  return MPT_OK;
}

int mpt_trace_closest(mpt_scene *s, const mpt_ray *rays, uint64_t n, mpt_hit *hits) {
  if (!s || (n && (!rays || !hits))) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  mpt_context *ctx = s->ctx;
  if (!n) return MPT_OK;
  DBuf<mpt_ray> dr;
  DBuf<mpt_hit> dh;
  MPT_HIP(ctx, dr.reserve(n));
  MPT_HIP(ctx, dh.reserve(n));
  MPT_HIP(ctx, hipMemcpyAsync(dr.p, rays, n * sizeof(mpt_ray), hipMemcpyHostToDevice, ctx->stream));
  k_trace_rays<false><<<grid_for(ctx, n, 8), kBlock, 0, ctx->stream>>>(s->dev, dr.p, nullptr, n, dh.p, nullptr);
  MPT_LAUNCHED(ctx, "k_trace_rays");
  MPT_HIP(ctx, hipMemcpyAsync(hits, dh.p, n * sizeof(mpt_hit), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_trace_closest_stats(mpt_scene *s, const mpt_ray *rays, uint64_t n, mpt_hit *hits, uint64_t counters[2]) {
  if (!s || !counters || (n && (!rays || !hits))) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  mpt_context *ctx = s->ctx;
  counters[0] = counters[1] = 0;
  if (!n) return MPT_OK;
  if (n > (1ull << 31)) return fail(ctx, MPT_E_INVALID, "at most 2^31 rays per call");
  const uint32_t n32 = (uint32_t)n;
  DBuf<mpt_ray> dr;
  DBuf<float4> q;
  DBuf<float2> raw;
  DBuf<mpt_hit> dh;
  DBuf<uint32_t> cnt;   // [queue size, claim cursor]
  DBuf<unsigned long long> ctr;
  MPT_HIP(ctx, dr.reserve(n));
  MPT_HIP(ctx, q.reserve(3ull * n));
  MPT_HIP(ctx, raw.reserve(n));
  MPT_HIP(ctx, dh.reserve(n));
  MPT_HIP(ctx, cnt.reserve(2));
  MPT_HIP(ctx, ctr.reserve(2));
  const uint32_t init[2] = {n32, 0u};
  MPT_HIP(ctx, hipMemcpyAsync(dr.p, rays, n * sizeof(mpt_ray), hipMemcpyHostToDevice, ctx->stream));
  MPT_HIP(ctx, hipMemcpyAsync(cnt.p, init, sizeof(init), hipMemcpyHostToDevice, ctx->stream));
  MPT_HIP(ctx, hipMemsetAsync(ctr.p, 0, 2 * sizeof(unsigned long long), ctx->stream));
  const uint32_t blocks = (n32 + 255) / 256;
  k_rays_to_queue<<<blocks, 256, 0, ctx->stream>>>(dr.p, n32, q.p);
  launch_extend(ctx, ctx->stream, n32, grid_for(ctx, n, 8), s->dev, q.p, cnt.p, raw.p, ctr.p, cnt.p + 1);
  k_resolve_hits<<<blocks, 256, 0, ctx->stream>>>(s->dev, dr.p, raw.p, n32, dh.p);
  MPT_LAUNCHED(ctx, "k_trace closest with counters");
  unsigned long long c[2];
  MPT_HIP(ctx, hipMemcpyAsync(hits, dh.p, n * sizeof(mpt_hit), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipMemcpyAsync(c, ctr.p, sizeof(c), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  counters[0] = c[0];
  counters[1] = c[1];
  return MPT_OK;
}

int mpt_trace_any(mpt_scene *s, const mpt_ray *rays, const float *tmax, uint64_t n, uint8_t *occluded) {
  if (!s || (n && (!rays || !tmax || !occluded))) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  mpt_context *ctx = s->ctx;
  if (!n) return MPT_OK;
  DBuf<mpt_ray> dr;
  DBuf<float> dt;
  DBuf<uint8_t> dout;
  MPT_HIP(ctx, dr.reserve(n));
  MPT_HIP(ctx, dt.reserve(n));
  MPT_HIP(ctx, dout.reserve(n));
  MPT_HIP(ctx, hipMemcpyAsync(dr.p, rays, n * sizeof(mpt_ray), hipMemcpyHostToDevice, ctx->stream));
  MPT_HIP(ctx, hipMemcpyAsync(dt.p, tmax, n * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  k_trace_rays<true><<<grid_for(ctx, n, 8), kBlock, 0, ctx->stream>>>(s->dev, dr.p, dt.p, n, nullptr, dout.p);
  MPT_LAUNCHED(ctx, "k_trace_rays");
  MPT_HIP(ctx, hipMemcpyAsync(occluded, dout.p, n, hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_emitter_sample(mpt_scene *s, uint64_t n, uint32_t seed, mpt_emitter_point *out) {
  if (!s || (n && !out)) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  mpt_context *ctx = s->ctx;
  if (s->dev.num_lights + s->dev.num_points == 0) return fail(ctx, MPT_E_INVALID, "the scene has no lights");
  if (!n) return MPT_OK;
  DBuf<mpt_emitter_point> dout;
  MPT_HIP(ctx, dout.reserve(n));
  k_emitter_sample<<<grid_for(ctx, n, 8), 256, 0, ctx->stream>>>(s->dev, seed, n, dout.p);
  MPT_LAUNCHED(ctx, "k_emitter_sample");
  MPT_HIP(ctx, hipMemcpyAsync(out, dout.p, n * sizeof(mpt_emitter_point), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_film_create(mpt_context *ctx, uint32_t width, uint32_t height, mpt_film **out) {
  if (!ctx || !out || !width || !height) return fail(ctx, MPT_E_INVALID, "bad film arguments");
  *out = nullptr;
  std::unique_ptr<mpt_film> f(new (std::nothrow) mpt_film());
  if (!f) return fail(ctx, MPT_E_NOMEM, "host allocation");
  f->ctx = ctx;
  f->width = width;
  f->height = height;
  MPT_HIP(ctx, f->data.reserve((size_t)width * height));
  MPT_HIP(ctx, hipMemsetAsync(f->data.p, 0, f->data.bytes(), ctx->stream));
  ctx->films.push_back(f.get());
  *out = f.release();
  return MPT_OK;
}

int mpt_film_destroy(mpt_film *f) {
  if (!f) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  (void)hipStreamSynchronize(ctx->stream);
  ctx->films.erase(std::remove(ctx->films.begin(), ctx->films.end(), f), ctx->films.end());
  delete f;
  return MPT_OK;
}

int mpt_render_film(mpt_scene *s, const mpt_job *job, mpt_film *film, mpt_render_stats *stats) {
  if (!s || !film) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  if (!job || job->width != film->width || job->height != film->height)
    return fail(s->ctx, MPT_E_INVALID, "film size does not match the job");
  if (job->renderer == MPT_RENDERER_MLT)
    return render_mlt(s, job, film->data.p, stats, nullptr, 0, nullptr, nullptr);
  // a filtered render adds into the film
  if (job->filter) MPT_HIP(s->ctx, hipMemsetAsync(film->data.p, 0, film->data.bytes(), s->ctx->stream));
  return render_range(s, job, 0, job->width * job->height, film->data.p, nullptr, stats);
}

int mpt_mlt_trace(mpt_scene *s, const mpt_job *job, uint32_t steps, double *rec, double *b, uint64_t *entries) {
  if (!s || !job || !rec) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  if (job->renderer != MPT_RENDERER_MLT) return fail(s->ctx, MPT_E_INVALID, "not a Metropolis job");
  if (steps == 0 || steps > job->mlt_mutations)
    return fail(s->ctx, MPT_E_INVALID, "steps must be 1..mlt_mutations");
  return render_mlt(s, job, nullptr, nullptr, rec, steps, b, entries);
}

int mpt_render_film_range(mpt_scene *s, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end, mpt_film *film,
                          mpt_render_stats *stats) {
  if (!s || !film) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  if (!job || job->width != film->width || job->height != film->height)
    return fail(s->ctx, MPT_E_INVALID, "film size does not match the job");
  if (pix_begin >= pix_end || (uint64_t)pix_end > (uint64_t)job->width * job->height)
    return fail(s->ctx, MPT_E_INVALID, "bad pixel range");
  return render_range(s, job, pix_begin, pix_end, film->data.p, nullptr, stats);
}

int mpt_film_download(mpt_film *f, float *rgb) {
  if (!f || !rgb) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  std::vector<float4> host;
  try {
    host.resize((size_t)f->width * f->height);
  } catch (const std::bad_alloc &) {
    return fail(ctx, MPT_E_NOMEM, "host allocation");
  }
  MPT_HIP(ctx, hipMemcpyAsync(host.data(), f->data.p, f->data.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i < host.size(); ++i) {
    const float w = host[i].w;
    const float k = w > 0 ? 1.0f / w : 0.0f;
    rgb[3 * i] = host[i].x * k;
    rgb[3 * i + 1] = host[i].y * k;
    rgb[3 * i + 2] = host[i].z * k;
  }
  return MPT_OK;
}

int mpt_film_download_raw(mpt_film *f, float *rgbw) {
  if (!f || !rgbw) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  MPT_HIP(ctx, hipMemcpyAsync(rgbw, f->data.p, f->data.bytes(), hipMemcpyDeviceToHost, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_film_upload_raw(mpt_film *f, const float *rgbw) {
  if (!f || !rgbw) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  MPT_HIP(ctx, hipMemcpyAsync(f->data.p, rgbw, f->data.bytes(), hipMemcpyHostToDevice, ctx->stream));
  MPT_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MPT_OK;
}

int mpt_film_tonemap(mpt_film *f, uint32_t tonemap, uint8_t *rgba) {
  if (!f || !rgba) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  if (tonemap != MPT_TONEMAP_REINHARD && tonemap != MPT_TONEMAP_GAMMA)
    return fail(ctx, MPT_E_INVALID, "unknown tone map");
  const uint32_t n = f->width * f->height;
  MPT_HIP(ctx, f->ldr.reserve(n));
  MPT_HIP(ctx, f->tm.reserve(kTmGrid + 1));
  hipStream_t st = ctx->stream;
  if (tonemap == MPT_TONEMAP_REINHARD) {
    k_tm_reduce<<<kTmGrid, kTmBlock, 0, st>>>(f->data.p, n, f->tm.p);
    k_tm_fold<<<1, kTmBlock, 0, st>>>(f->tm.p, n);
  }
  k_tm_map<<<(n + kTmBlock - 1) / kTmBlock, kTmBlock, 0, st>>>(f->data.p, n, tonemap, f->tm.p, f->ldr.p);
  MPT_LAUNCHED(ctx, "tone map");
  MPT_HIP(ctx, hipMemcpyAsync(rgba, f->ldr.p, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  MPT_HIP(ctx, hipStreamSynchronize(st));
  return MPT_OK;
}

int mpt_film_tonemap_noalloc(mpt_film *f, uint32_t tonemap, uint8_t *rgba) {
  if (!f || !rgba) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  if (tonemap != MPT_TONEMAP_REINHARD && tonemap != MPT_TONEMAP_GAMMA)
    return fail(ctx, MPT_E_INVALID, "unknown tone map");
  const uint64_t n64 = (uint64_t)f->width * f->height;
  if (n64 > kNoAllocMaxPixels)
    return fail(ctx, MPT_E_INVALID,
                "film of " + std::to_string(n64) + " pixels exceeds the allocation-free tone map's scratch limit of " +
                    std::to_string(kNoAllocMaxPixels) + " pixels");
  hipError_t e = hipSuccess;
  char *base = static_cast<char *>(scratch_ptr(ctx->device, &e));
  if (!base) return hip_fail(ctx, e, "scratch array address (hipGetSymbolAddress)");
  double2 *partials = reinterpret_cast<double2 *>(base);
  uchar4 *ldr = reinterpret_cast<uchar4 *>(base + kScratchPartialBytes);
  const uint32_t n = (uint32_t)n64;
  hipStream_t st = ctx->stream;
  // every partial is written by k_tm_reduce before k_tm_fold reads it, so the
  // array needs no clearing between calls
  if (tonemap == MPT_TONEMAP_REINHARD) {
    k_tm_reduce<<<kTmGrid, kTmBlock, 0, st>>>(f->data.p, n, partials);
    k_tm_fold<<<1, kTmBlock, 0, st>>>(partials, n);
  }
  k_tm_map<<<(n + kTmBlock - 1) / kTmBlock, kTmBlock, 0, st>>>(f->data.p, n, tonemap, partials, ldr);
  MPT_LAUNCHED(ctx, "tone map");
  MPT_HIP(ctx, hipMemcpyAsync(rgba, ldr, (size_t)n * 4, hipMemcpyDeviceToHost, st));
  MPT_HIP(ctx, hipStreamSynchronize(st));
  return MPT_OK;
}

int mpt_film_save(mpt_film *f, const char *path, uint32_t tonemap, uint32_t flags) {
  if (!f || !path) return MPT_E_INVALID;
  mpt_context *ctx = f->ctx;
  std::string ext(path);
  const size_t dot = ext.rfind('.');
  ext = dot == std::string::npos ? std::string() : ext.substr(dot + 1);
  for (char &c : ext) c = (char)std::tolower((unsigned char)c);
  const bool hdr = ext == "hdr" || ext == "rgbe" || ext == "pfm";
  if (!hdr && ext != "png" && ext != "ppm")
    return fail(ctx, MPT_E_IO, "unsupported image format '" + ext + "' (png, ppm, hdr, rgbe, pfm)");
  const size_t n = (size_t)f->width * f->height;
  int rc;
  try {
    if (hdr && (flags & MPT_SAVE_RADIANCE)) {
      std::vector<float> rgb(3 * n);
      if ((rc = mpt_film_download(f, rgb.data()))) return rc;
      rc = mpt_image_write_hdr(path, f->width, f->height, rgb.data());
    } else {
      std::vector<uint8_t> rgba(4 * n);
      if ((rc = mpt_film_tonemap(f, tonemap, rgba.data()))) return rc;
      rc = mpt_image_write_ldr(path, f->width, f->height, rgba.data());
    }
  } catch (const std::bad_alloc &) {
    return fail(ctx, MPT_E_NOMEM, "host allocation");
  }
  if (rc) return fail(ctx, rc, std::string("cannot write ") + path);
  return MPT_OK;
}

int mpt_render(mpt_scene *s, const mpt_job *job, float *film_rgb, mpt_render_stats *stats) {
  if (!s || !film_rgb) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  int rc = check_job(s->ctx, job);
  if (rc) return rc;
  mpt_film *f = nullptr;
  rc = mpt_film_create(s->ctx, job->width, job->height, &f);
  if (rc) return rc;
  rc = mpt_render_film(s, job, f, stats);
  if (!rc) rc = mpt_film_download(f, film_rgb);
  mpt_film_destroy(f);
  return rc;
}

int mpt_render_paths(mpt_scene *s, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end, float *L_out) {
  if (!s || !L_out) return fail(s ? s->ctx : nullptr, MPT_E_INVALID, "NULL argument");
  int rc = check_job(s->ctx, job);
  if (rc) return rc;
  if (pix_begin >= pix_end || pix_end > job->width * job->height)
    return fail(s->ctx, MPT_E_INVALID, "bad pixel range");
  return render_range(s, job, pix_begin, pix_end, nullptr, L_out, nullptr);
}

}  // extern "C"
