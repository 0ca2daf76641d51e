// Bidirectional path tracing on the device: BidirectionalPathTracer::sample
// (renderers/renderers/BidirectionalPathTracer.cpp:34-174) over Path
// (renderers/utils/Path.cpp) and PathVertex (renderers/utils/PathVertex.cpp),
// in wavefront form.
//
// Per batch, every path slot holds its eye subpath (camera vertex first) and
// its light subpath (emitter vertex first) as vertex records in HBM:
//   k_bd_camera        the camera vertex and its ray (Camera::getPoint at the
//                      film position, BidirectionalPathTracer.cpp:38-45)
//   per bounce b:      closest hit of the subpath rays (the path tracer's
//                      k_extend), then k_bd_step: Path::_initE / _initL on
//                      vertex b (geometry term, both directions' pdfs, the
//                      roulette), the new vertex b + 1 with its cumulative
//                      contribution and density, and its next event
//                      (Path::_samplePathVertex, Path.cpp:517-598)
//   k_bd_light_start   EmitterSampler::sample / getPd for the slots whose eye
//                      subpath did not end on an emitter (generateL)
//   k_bd_join          Path::append(const Path &) (Path.cpp:339-515): the
//                      join's _initL / _initE, and every vertex pair whose
//                      visibility a split (s, t) needs, appended to a pair
//                      queue with one atomic per block
//   k_bd_vis           any hit of every queued pair (Scene::intersects,
//                      core/Scene.cpp:110-114); an occluded pair sets its bit
//   k_bd_combine       per slot in double precision: the join's visibility,
//                      the light-side values of the eye vertices and the
//                      eye-side values of the light vertices, getPd and
//                      getContribution of every split, the power heuristic,
//                      and the [0, 1] clamp (RADIANCE_HACK)
// The film kernels of the path tracer then sum each pixel's samples.
//
// Included by mpt.hip after the shading helpers it uses (hit_info,
// light_point, block appends).
#ifndef MPT_BDPT_H_
#define MPT_BDPT_H_

namespace bdpt {

// Random streams (shared with the oracle): eye vertex j continues from
// (depth j, stream 2) with stream 3 for a dielectric's channel and draws its
// construction event from stream 4; light vertex i uses depth kLightDepth + i
// the same way, the light's selection and point streams 0 and 1 of depth
// kLightDepth; the join's forced roulette stream 0 of kJoinDepth.
constexpr uint32_t kLightDepth = 1024;
constexpr uint32_t kJoinDepth = 2047;
// Deepest subpath (bounces) a bidirectional job may ask for: a path of
// 2 (12 + 1) vertices, whose pairs fit the per-slot bit masks below.
constexpr uint32_t kMaxDepth = 12;
constexpr uint32_t kMaxVerts = 2 * (kMaxDepth + 1);

enum { K_SURFACE = 0, K_EMITTER = 1, K_SENSOR = 2 };
enum { LOBE_ABSORB = 0, LOBE_DIFFUSE = 1, LOBE_SPECULAR = 2 };
// info word of a vertex: kind | lobe << 2 | flags | material (or point light) << 8
constexpr uint32_t I_NO_NORMAL = 16u;   // a point light: SurfacePointNormal::dot is 1
constexpr uint32_t I_AREA = 32u;        // its shape's area is at least EPSILON
constexpr uint32_t I_POINT = 64u;       // the emitter is point light (info >> 8)
constexpr uint32_t I_SPEC = 128u;       // k_bd_combine's copy: the vertex scatters specularly
// slot flags
constexpr uint32_t S_EYE_ON_EMITTER = 1u;   // the eye subpath ended on an emitter: no light subpath
constexpr uint32_t S_JOIN_OK = 2u;          // the join's _initL and _initE succeeded; its pair is queued

constexpr uint32_t kVtx = 5;   // float4 per vertex record

// A subpath vertex.  alpha and pcum are the vertex's own side's cumulative
// contribution and density (alphaE, pE on the eye subpath; alphaL, pL on the
// light subpath); gprev / gnext the geometry terms of the edges to its
// predecessor / successor in its own subpath; pdf_fwd the projected-solid-
// angle pdf of scattering onward along its subpath (pdfE for eye vertices,
// pdfL for light vertices), pdf_rev the other direction's; fs the BSDF of
// light flowing from the light side to the eye side through it.
struct Vtx {
  float3 p;
  uint32_t prim;   // primitive the vertex lies on (kMiss: camera, point light)
  float3 n;
  uint32_t info;
  float3 alpha;
  float gprev;
  float3 fs;
  float gnext;
  float pdf_fwd, pdf_rev;
  double pcum;
};

__device__ __forceinline__ Vtx load_vtx(const float4 *r) {
  const float4 a = r[0], b = r[1], c = r[2], d = r[3], e = r[4];
  Vtx v;
  v.p = xyz(a);
  v.prim = __float_as_uint(a.w);
  v.n = xyz(b);
  v.info = __float_as_uint(b.w);
  v.alpha = xyz(c);
  v.gprev = c.w;
  v.fs = xyz(d);
  v.gnext = d.w;
  v.pdf_fwd = e.x;
  v.pdf_rev = e.y;
  v.pcum = __hiloint2double(__float_as_int(e.w), __float_as_int(e.z));
  return v;
}

__device__ __forceinline__ void store_vtx(float4 *r, const Vtx &v) {
  r[0] = make_float4(v.p.x, v.p.y, v.p.z, __uint_as_float(v.prim));
  r[1] = make_float4(v.n.x, v.n.y, v.n.z, __uint_as_float(v.info));
  r[2] = make_float4(v.alpha.x, v.alpha.y, v.alpha.z, v.gprev);
  r[3] = make_float4(v.fs.x, v.fs.y, v.fs.z, v.gnext);
  r[4] = make_float4(v.pdf_fwd, v.pdf_rev, __int_as_float(__double2loint(v.pcum)),
                     __int_as_float(__double2hiint(v.pcum)));
}

__device__ __forceinline__ uint32_t kind_of(uint32_t info) { return info & 3u; }
__device__ __forceinline__ uint32_t lobe_of(uint32_t info) { return (info >> 2) & 3u; }
__device__ __forceinline__ uint32_t with_lobe(uint32_t info, uint32_t lobe) { return (info & ~12u) | (lobe << 2); }
__device__ __forceinline__ uint32_t mat_of(uint32_t info) { return info >> 8; }

// SurfacePointNormal::dot (core/SurfacePoint.h:141-144)
__device__ __forceinline__ float ndot(const Vtx &v, float3 w) { return (v.info & I_NO_NORMAL) ? 1.0f : dot(v.n, w); }

// pt->emitter->isEmitter() of the vertex's surface point
__device__ __forceinline__ bool on_emitter(const Scene &sc, uint32_t info) {
  if (info & I_POINT) return true;
  if (kind_of(info) == K_SENSOR) return false;
  return emissive(sc.mats[mat_of(info)]);
}

// BSDF::isSpecular: dielectrics and the sensor (DielectricBSDF.h:64, Sensor.h:153)
__device__ __forceinline__ bool specular_of(const Scene &sc, uint32_t info) {
  const uint32_t k = kind_of(info);
  if (k == K_SENSOR) return true;
  if (k == K_EMITTER) return false;
  return sc.mats[mat_of(info)].type == MPT_BSDF_DIELECTRIC;
}

// Vector::operator== (common/math/Vector.inl:73-80) of two unit directions:
// the reference's EQ is 1e-7 in double; these directions carry FP32
// rounding, so they count as equal within 1e-5 per component (distinct
// refraction channels differ by far more).
__device__ __forceinline__ bool same_dir(float3 a, float3 b) {
  return fabsf(a.x - b.x) < 1e-5f && fabsf(a.y - b.y) < 1e-5f && fabsf(a.z - b.z) < 1e-5f;
}

__device__ __forceinline__ float channel_ior(const mpt_material &m, uint32_t c) {
  return c == 1 && m.ior_g > 0.0f ? m.ior_g : (c == 2 && m.ior_b > 0.0f ? m.ior_b : m.ior);
}

// BSDF::getPd(Event(dir, event)) of vertex v with its incident direction set
// to wi (projected solid angle).
__device__ float pdf_v(const Scene &sc, const Vtx &v, float3 wi, float3 dir) {
  const uint32_t k = kind_of(v.info);
  if (k == K_SENSOR) {
    // Sensor::getPd (materials/Sensor.cpp:85-95)
    if (same_dir(wi, mul(v.n, -1.0f))) return 1.0f;
    return same_dir(dir, v.n) ? 1.0f : 0.0f;
  }
  if (k == K_EMITTER) {
    // Emitter::getPd -> its sampler (Emitter.cpp:160-165): UniformOnSphere
    // returns 1 (UniformOnSphereSampler.cpp:32-34); an oriented emitter's
    // DiffuseBSDF sees wi = -normalS (OrientedEmitter.cpp:25-31)
    if (v.info & I_POINT) return 1.0f;
    return dot(v.n, dir) > 0.0f ? 1.0f / kPi : 0.0f;
  }
  const mpt_material &m = sc.mats[mat_of(v.info)];
  if (m.type == MPT_BSDF_DIFFUSE) {
    // DiffuseBSDF::getPd (DiffuseBSDF.cpp:30-38)
    if (dot(v.n, wi) > 0.0f) return 0.0f;
    return dot(v.n, dir) > 0.0f ? 1.0f / kPi : 0.0f;
  }
  if (m.type == MPT_BSDF_PHONG) {
    // ModifiedPhongBSDF::getPd (ModifiedPhongBSDF.cpp:131-162) by event index
    const float kda = avg3(kd_of(m)), ksa = avg3(ks_of(m));
    const uint32_t lobe = lobe_of(v.info);
    if (lobe == LOBE_ABSORB) return 1.0f - (kda + ksa);
    const float cosA = dot(v.n, dir);
    if (cosA <= 0.0f) return 0.0f;
    float pdf;
    if (lobe == LOBE_SPECULAR) {
      const float c = fabsf(dot(reflect(wi, v.n), dir));
      const float cn = c > 0.0f ? exp2f(m.exponent * log2f(c)) : 0.0f;
      pdf = fminf(1.0f, (m.exponent + 1.0f) / (2.0f * kPi) * cn);
    } else {
      pdf = fabsf(dot(v.n, dir)) / kPi;
    }
    return pdf / cosA;
  }
  // DielectricBSDF::getPd (DielectricBSDF.cpp:106-127), iorIndex 0: the share
  // of channels whose refraction is dir, or 1 for the mirror image
  const float3 N = dot(v.n, wi) > 0.0f ? mul(v.n, -1.0f) : v.n;
  if (dot(dir, N) < 0.0f) {
    uint32_t cnt = 0;
    for (uint32_t c = 0; c < 3; ++c) {
      float3 wt;
      if (refract(wi, v.n, kAirIor, channel_ior(m, c), wt) && same_dir(dir, wt)) ++cnt;
    }
    return (float)cnt * (1.0f / 3.0f);
  }
  return same_dir(dir, reflect(wi, N)) ? 1.0f : 0.0f;
}

// BSDF::evaluate(wi, dir) of vertex v: light arriving along wi, leaving
// along dir.
__device__ float3 f_v(const Scene &sc, const Vtx &v, float3 wi, float3 dir) {
  const uint32_t k = kind_of(v.info);
  if (k == K_SENSOR) {
    // Sensor::evaluate = getWe1(-wi) (Sensor.cpp:47-62, 98-107)
    return same_dir(mul(wi, -1.0f), v.n) ? f3(1, 1, 1) : f3(0, 0, 0);
  }
  if (k == K_EMITTER) {
    // Emitter::evaluate = getLe(dir) / getLe0() (Emitter.cpp:68-70, 167-171):
    // 1 / (4 pi) for a point, 1 / pi on an oriented emitter's side; a
    // channel without emission gives 0 (the reference asserts against it)
    if (v.info & I_POINT) return f3(0.25f / kPi, 0.25f / kPi, 0.25f / kPi);
    if (!(dot(v.n, dir) > 0.0f)) return f3(0, 0, 0);
    const float3 le = le_of(sc.mats[mat_of(v.info)]);
    const float s = 1.0f / kPi;
    return f3(le.x != 0.0f ? s : 0.0f, le.y != 0.0f ? s : 0.0f, le.z != 0.0f ? s : 0.0f);
  }
  const mpt_material &m = sc.mats[mat_of(v.info)];
  if (m.type != MPT_BSDF_DIELECTRIC) return bsdf_eval(m, v.n, wi, dir);
  // DielectricBSDF::evaluate (DielectricBSDF.cpp:129-158)
  const float3 N = dot(v.n, wi) > 0.0f ? mul(v.n, -1.0f) : v.n;
  if (dot(dir, N) < 0.0f) {
    float s[3];
    for (uint32_t c = 0; c < 3; ++c) {
      float3 wt;
      s[c] = refract(wi, v.n, kAirIor, channel_ior(m, c), wt) && same_dir(dir, wt) ? 1.0f : 0.0f;
    }
    return f3(m.ks[0] * s[0], m.ks[1] * s[1], m.ks[2] * s[2]);
  }
  return same_dir(dir, reflect(wi, N)) ? ks_of(m) : f3(0, 0, 0);
}

// BSDF::sample of vertex v: the event's direction (zero when absorbed) and
// Phong lobe.  u0, u1 direction, u2 lobe / refraction choice, uch a
// dielectric's channel (Random::sampleInt(0, 3), stats/Random.h:87-89).
__device__ float3 sample_v(const Scene &sc, const Vtx &v, float3 wi, float u0, float u1, float u2, float uch,
                           uint32_t &lobe) {
  lobe = LOBE_DIFFUSE;
  const uint32_t k = kind_of(v.info);
  if (k == K_SENSOR) return v.n;   // Sensor::sample (Sensor.cpp:72-83)
  if (k == K_EMITTER) {
    if (v.info & I_POINT) {
      // UniformOnSphereSampler::sample (UniformOnSphereSampler.cpp:22-30)
      const float z = 1.0f - 2.0f * u0, r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
      return f3(r * __builtin_amdgcn_cosf(u1), r * __builtin_amdgcn_sinf(u1), z);
    }
    return cos_random(v.n, u0, u1);   // DiffuseBSDF::sample with wi = -normalS
  }
  const mpt_material &m = sc.mats[mat_of(v.info)];
  if (m.type == MPT_BSDF_DIFFUSE) {
    if (dot(v.n, wi) > 0.0f) return f3(0, 0, 0);   // DiffuseBSDF.cpp:21-28
    return cos_random(v.n, u0, u1);
  }
  if (m.type == MPT_BSDF_PHONG) {
    // ModifiedPhongBSDF::sample (ModifiedPhongBSDF.cpp:84-129): the event's
    // direction as drawn, below the surface or not
    const float kda = avg3(kd_of(m)), ksa = avg3(ks_of(m));
    const float p0 = u2 - kEps;
    if (p0 <= kda) return cos_random(v.n, u0, u1);
    if (p0 <= kda + ksa) {
      lobe = LOBE_SPECULAR;
      const float ca = exp2f(log2f(u0) / (m.exponent + 1.0f));
      const float sa = sqrtf(fmaxf(0.0f, 1.0f - ca * ca));
      const float cp = __builtin_amdgcn_cosf(u1), sp = __builtin_amdgcn_sinf(u1);
      const float3 R = reflect(wi, v.n);
      float3 rr = R, uu, vv;
      if (fabsf(dot(v.n, rr)) < 0.95f) {
        uu = normalize(cross(rr, v.n));
        vv = cross(uu, rr);
      } else {
        rr = normalize(rr);
        onb(rr, vv, uu);
      }
      const float3 w = add(add(mul(uu, sa * cp), mul(rr, ca)), mul(vv, sa * sp));
      return dot(v.n, w) > 0.0f ? w : R;
    }
    lobe = LOBE_ABSORB;
    return f3(0, 0, 0);
  }
  // DielectricBSDF::sample (DielectricBSDF.cpp:55-104), iorIndex 0
  const float3 N = dot(v.n, wi) > 0.0f ? mul(v.n, -1.0f) : v.n;
  const uint32_t c = min(2u, (uint32_t)(uch * 3.0f));
  const float ior = channel_ior(m, c);
  float3 wt;
  if (refract(wi, v.n, kAirIor, ior, wt)) {
    const float Fr = fresnel(wi, v.n, kAirIor, ior);
    if (u2 < m.transparency * (1.0f - Fr)) return wt;
  }
  return reflect(wi, N);
}

// The Phong lobe of a new vertex's construction event (PathVertex::_init).
__device__ __forceinline__ uint32_t init_lobe(const Scene &sc, uint32_t info, float u) {
  if (kind_of(info) != K_SURFACE) return LOBE_DIFFUSE;
  const mpt_material &m = sc.mats[mat_of(info)];
  if (m.type != MPT_BSDF_PHONG) return LOBE_DIFFUSE;
  const float kda = avg3(kd_of(m)), ksa = avg3(ks_of(m));
  const float p0 = u - kEps;
  return p0 <= kda ? LOBE_DIFFUSE : (p0 <= kda + ksa ? LOBE_SPECULAR : LOBE_ABSORB);
}

// fs == black (SpectralSampleSet.inl:141-153): every channel EQ 0
__device__ __forceinline__ bool black_ref(float3 a) {
  return fabsf(a.x) <= kEps && fabsf(a.y) <= kEps && fabsf(a.z) <= kEps;
}
// SpectralSampleSet::getMaxSample (SpectralSampleSet.inl:388-398)
__device__ __forceinline__ float max_sample(float3 a) {
  float m = a.x;
  if (a.y > m) m = a.y;
  if (a.z > m) m = a.z;
  return m;
}

// Path::_initE (EYE, Path.cpp:647-693) / Path::_initL (Path.cpp:600-645):
// vertex v scattering from wi to wo, toward the point `to` at distance t.
// Updates v's geometry term toward `to` (gnext), both pdfs, fs; with the
// roulette's uniform rr (taken when `roulette` or fs is black) and on
// success the onward cumulative contribution and density.
template <bool EYE>
__device__ bool init_side(const Scene &sc, Vtx &v, float3 wi, float3 wo, float t, const Vtx &to, bool roulette,
                          float rr, float3 &alpha_out, double &pcum_out) {
  if (t < kEps) {   // GEOMETRY_BOUND
    v.gnext = 0.0f;
    return false;
  }
  const float3 mwo = mul(wo, -1.0f), mwi = mul(wi, -1.0f);
  v.gnext = fabsf(ndot(v, wo) * ndot(to, mwo)) / (t * t);
  v.pdf_rev = pdf_v(sc, v, mwo, mwi);
  v.fs = EYE ? f_v(sc, v, mwo, mwi) : f_v(sc, v, wi, wo);
  v.pdf_fwd = pdf_v(sc, v, wi, wo);
  if (v.gnext == 0.0f || v.pdf_fwd == 0.0f) return false;
  if (roulette || black_ref(v.fs)) {
    const float q = fminf(0.95f, max_sample(v.fs) / v.pdf_fwd);
    if (q <= rr) {
      v.pdf_fwd = 0.0f;
      return false;
    }
    v.pdf_fwd *= q;
  }
  pcum_out = ((double)v.pdf_fwd * (double)v.gnext) * v.pcum;
  alpha_out = mul(mul(v.fs, 1.0f / v.pdf_fwd), v.alpha);
  return true;
}

// A surface vertex at a closest hit (SurfacePoint::init, core/SurfacePoint.
// cpp:49-78).  A triangle's mesh counts as having an area of at least
// EPSILON; quads and spheres (pi) have theirs.
__device__ __forceinline__ Vtx surface_vtx(const Scene &sc, uint32_t prim, float3 p, float2 uv) {
  const HitInfo hi = hit_info(sc, prim, p, uv);
  Vtx v;
  v.p = p;
  v.prim = prim;
  v.n = hi.n;
  const bool area = prim < sc.nt || prim < sc.nt + sc.ns || hi.area >= kEps;
  v.info = K_SURFACE | (area ? I_AREA : 0u) | (hi.mat << 8);
  v.alpha = f3(1, 1, 1);
  v.gprev = 1.0f;
  v.fs = f3(0, 0, 0);
  v.gnext = 1.0f;
  v.pdf_fwd = 1.0f;
  v.pdf_rev = 1.0f;
  v.pcum = 1.0;
  return v;
}

struct SlotView {
  float4 *base;        // the slot's vertex records
  uint32_t E;          // eye records before the light records
};

// Camera vertices and rays: Camera::getPoint at the jittered film position
// of (pixel, sample), normal = the pinhole ray's direction
// (Camera::_getGeometricNormal, cameras/Camera.cpp:48-50), pE = 1.
__global__ __launch_bounds__(256) void k_bd_camera(const BatchDesc *__restrict__ bd, float4 *__restrict__ q,
                                                   uint2 *__restrict__ pjs, float2 *__restrict__ pos,
                                                   uint32_t *__restrict__ count0, float4 *__restrict__ verts,
                                                   uint4 *__restrict__ info, uint32_t V) {
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
  if (job.filter && pos) pos[slot] = make_float2(x * (float)width, y * (float)job.height);
  const CameraDev &cam = job.cam;
  const float nx = (2.0f * x - 1.0f) * cam.tan_w, ny = (1.0f - 2.0f * y) * cam.tan_h;
  const float3 U = f3(cam.u[0], cam.u[1], cam.u[2]), Vv = f3(cam.v[0], cam.v[1], cam.v[2]);
  const float3 N = f3(cam.n[0], cam.n[1], cam.n[2]);
  const float3 d = normalize(sub(add(mul(U, nx), mul(Vv, ny)), N));
  const float3 o = f3(cam.eye[0], cam.eye[1], cam.eye[2]);
  Vtx c;
  c.p = o;
  c.prim = kMiss;
  c.n = d;
  c.info = K_SENSOR | (LOBE_DIFFUSE << 2);   // Camera::_getSurfaceArea is 0
  c.alpha = f3(1, 1, 1);
  c.gprev = 1.0f;
  c.fs = f3(0, 0, 0);
  c.gnext = 1.0f;
  c.pdf_fwd = 1.0f;
  c.pdf_rev = 1.0f;
  c.pcum = 1.0;
  store_vtx(verts + (uint64_t)slot * V * kVtx, c);
  info[slot] = make_uint4(1u, 0u, 0u, 0u);
  q[3ull * slot] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
  q[3ull * slot + 1] = make_float4(d.x, d.y, d.z, __uint_as_float(kMiss));
  pjs[slot] = make_uint2(pixel, (ch.job << kJobBits) | sample);
}

// One bounce of the eye (EYE) or light subpaths: vertex b of each queued
// ray's slot scattered into the ray; the hit becomes vertex b + 1
// (Path::_samplePathVertex with _initE / _initL, Path.cpp:517-598), which
// draws its next event unless the subpath ends there (generateE stops on an
// emitter, BidirectionalPathTracer.cpp:233-249; the job's max_depth caps
// each subpath's bounces).
template <bool EYE>
__global__ __launch_bounds__(256) void k_bd_step(Scene sc, const BatchDesc *__restrict__ bd, uint32_t b,
                                                 const float4 *__restrict__ qin, const float2 *__restrict__ hits,
                                                 const uint32_t *__restrict__ count_in, float4 *__restrict__ qout,
                                                 uint32_t *__restrict__ count_out, float4 *__restrict__ verts,
                                                 uint4 *__restrict__ info, uint32_t V,
                                                 const uint2 *__restrict__ pjs) {
  __shared__ uint32_t append_sh[256 / 64 + 1];
  const uint32_t n = *count_in;
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t E = V / 2;
  const uint32_t dbase = EYE ? 0u : kLightDepth;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    bool want = false;
    float3 np = f3(0, 0, 0), ndir = f3(0, 0, 0);
    uint32_t slot = 0, nprim = kMiss;
    if (i < n) {
      const float4 a = qin[3ull * i], dq = qin[3ull * i + 1];
      const float2 h = hits[i];
      slot = __float_as_uint(a.w);
      const float3 o = xyz(a), d = xyz(dq);
      const uint32_t rec = __float_as_uint(h.y);
      const uint2 pj = pjs[slot];
      const JobDev &job = bd->jobs[pj.y >> kJobBits];
      const uint32_t pixel = pj.x, sample = pj.y & ((1u << kJobBits) - 1);
      float4 *vs = verts + (uint64_t)slot * V * kVtx + (EYE ? 0u : E * kVtx);
      if (rec != kMiss) {
        const float4 *rp = sc.recs + 3ull * rec;
        const uint32_t prim = __float_as_uint(rp[0].w);
        float2 uv;
        const float t = refine_hit(rp, o, d, h.x, &uv);
        if (t > kEps) {
          Vtx v = load_vtx(vs + b * kVtx);
          const float3 wi = b == 0 ? mul(d, -1.0f) : normalize(sub(v.p, load_vtx(vs + (b - 1) * kVtx).p));
          Vtx nv = surface_vtx(sc, prim, add(o, mul(d, t)), uv);
          const mpt_rng::U4 su = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, dbase + b, 2);
          const bool roulette = EYE ? b >= 1 : true;
          float3 alpha;
          double pcum;
          if (init_side<EYE>(sc, v, wi, d, t, nv, roulette, mpt_rng::unitf(su.w), alpha, pcum)) {
            store_vtx(vs + b * kVtx, v);
            nv.gprev = v.gnext;
            nv.alpha = alpha;
            nv.pcum = pcum;
            const mpt_rng::U4 iu = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, dbase + b + 1, 4);
            nv.info = with_lobe(nv.info, init_lobe(sc, nv.info, mpt_rng::unitf(iu.x)));
            uint32_t *cnt = EYE ? &info[slot].x : &info[slot].y;
            *cnt = b + 2;
            bool end = b + 2 >= job.max_depth + 1;
            if (EYE && on_emitter(sc, nv.info)) {
              end = true;
              info[slot].z |= S_EYE_ON_EMITTER;
            }
            if (!end) {
              // the next event, from the incident direction Path::
              // _samplePathVertex recomputes from the positions
              const float3 wi2 = normalize(sub(nv.p, v.p));
              const mpt_rng::U4 nu = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, dbase + b + 1, 2);
              const float uch = mpt_rng::unitf(mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, dbase + b + 1, 3).x);
              uint32_t lobe;
              const float3 dir = sample_v(sc, nv, wi2, mpt_rng::unitf(nu.x), mpt_rng::unitf(nu.y),
                                          mpt_rng::unitf(nu.z), uch, lobe);
              nv.info = with_lobe(nv.info, lobe);
              if (!is_black(dir)) {
                want = true;
                np = nv.p;
                ndir = dir;
                nprim = prim;
              }
            }
            store_vtx(vs + (b + 1) * kVtx, nv);
          }
        }
      }
    }
    const uint32_t ei = block_append<256 / 64>(want, count_out, append_sh);
    if (want) {
      qout[3ull * ei] = make_float4(np.x, np.y, np.z, __uint_as_float(slot));
      qout[3ull * ei + 1] = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(nprim));
    }
  }
}

// The light subpath's first vertex and ray (generateL's first Path::append,
// Path.cpp:255-275): a light by power and a point on it (EmitterSampler::
// sample, stats/samplers/EmitterSampler.cpp:51-61, getRandomPoint of its
// shape), pA = 1 / the area lights' summed area (getPd, :63-96), the vertex
// PathVertex(pt, pA, 1) with alphaL = Le0 / pA (PathVertex.cpp:31-56), and
// its emitted direction.  Slots whose eye subpath ended on an emitter get
// none.
__global__ __launch_bounds__(256) void k_bd_light_start(Scene sc, const BatchDesc *__restrict__ bd,
                                                        float4 *__restrict__ q, uint32_t *__restrict__ count0,
                                                        float4 *__restrict__ verts, uint4 *__restrict__ info,
                                                        uint32_t V, const uint2 *__restrict__ pjs) {
  __shared__ uint32_t append_sh[256 / 64 + 1];
  const uint32_t n = bd->hdr.n_paths;
  const uint32_t E = V / 2;
  const uint32_t nl = sc.num_lights + sc.num_points;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t slot = base + threadIdx.x;
    bool want = false;
    float3 lp = f3(0, 0, 0), dir = f3(0, 0, 0);
    uint32_t target = kMiss;
    if (slot < n && !(info[slot].z & S_EYE_ON_EMITTER)) {
      const uint2 pj = pjs[slot];
      const JobDev &job = bd->jobs[pj.y >> kJobBits];
      const uint32_t pixel = pj.x, sample = pj.y & ((1u << kJobBits) - 1);
      const mpt_rng::U4 a = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, kLightDepth, 0);
      const mpt_rng::U4 b2 = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, kLightDepth, 1);
      // Random::sampleNormalizedCDF over the lights' power shares
      const float x = mpt_rng::unitf(a.x);
      uint32_t j = 0;
      while (j + 1 < nl && !(x <= sc.light_sel[j].x)) ++j;
      Vtx e;
      float3 le0;
      if (j < sc.num_lights) {
        const LightDev &ld = sc.lights[j];
        float3 ln;
        light_point(sc, ld, mpt_rng::unitf(a.y), mpt_rng::unitf(a.z), mpt_rng::unitf(a.w), mpt_rng::unitf(b2.x), lp,
                    ln, target);
        e.n = ln;
        e.info = K_EMITTER | (LOBE_DIFFUSE << 2) | (ld.area >= kEps ? I_AREA : 0u) | (ld.mat << 8);
        le0 = mul(le_of(sc.mats[ld.mat]), kPi);   // radiant exitance = power / area
      } else {
        const uint32_t pi = j - sc.num_lights;
        lp = xyz(sc.points[2 * pi]);
        e.n = f3(0, 1, 0);
        e.info = K_EMITTER | (LOBE_DIFFUSE << 2) | I_NO_NORMAL | I_POINT | (pi << 8);
        le0 = mul(xyz(sc.points[2 * pi + 1]), 4.0f * kPi);   // the point's power
      }
      const float pA = sc.emit_area > 0.0f ? 1.0f / sc.emit_area : 1.0f;
      e.p = lp;
      e.prim = target;
      e.alpha = mul(le0, 1.0f / pA);
      e.gprev = 1.0f;
      e.fs = f3(0, 0, 0);
      e.gnext = 1.0f;
      e.pdf_fwd = 1.0f;
      e.pdf_rev = 1.0f;
      e.pcum = (double)pA;
      info[slot].y = 1u;
      if (job.max_depth >= 1) {
        const mpt_rng::U4 u = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, kLightDepth, 2);
        uint32_t lobe;
        dir = sample_v(sc, e, f3(0, 0, 0), mpt_rng::unitf(u.x), mpt_rng::unitf(u.y), mpt_rng::unitf(u.z), 0.0f, lobe);
        want = !is_black(dir);
      }
      store_vtx(verts + ((uint64_t)slot * V + E) * kVtx, e);
    }
    const uint32_t ei = block_append<256 / 64>(want, count0, append_sh);
    if (want) {
      q[3ull * ei] = make_float4(lp.x, lp.y, lp.z, __uint_as_float(slot));
      q[3ull * ei + 1] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(target));
    }
  }
}

// Appends cnt entries per thread to a queue with one atomic per block (a
// wave scan, the waves' totals in LDS).  Every thread of the block must
// call it; returns the thread's first index.
template <int WAVES>
__device__ __forceinline__ uint32_t block_append_n(uint32_t cnt, uint32_t *counter, uint32_t *sh) {
  uint32_t x = cnt;
  const uint32_t lane = lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  const uint32_t wave = threadIdx.x / 64;
  if (lane == 63) sh[wave] = x;
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
  const uint32_t idx = sh[WAVES] + sh[wave] + x - cnt;
  __syncthreads();
  return idx;
}

// Merged path index -> vertex record: light vertices 0..S-1 first, then the
// eye subpath from its last vertex back to the camera (Path::append(Path),
// Path.cpp:345-348).
__device__ __forceinline__ uint32_t rec_of(uint32_t i, uint32_t S, uint32_t T, uint32_t E) {
  return i < S ? E + i : T - 1 - (i - S);
}

// Occlusion bit of the merged pair (u, v) in a slot's mask.
// V: the batch's vertex records per slot.
__device__ __forceinline__ uint32_t pair_bit(uint32_t u, uint32_t v, uint32_t V) { return u * V + v; }
constexpr uint32_t kMaskWords = (kMaxVerts * kMaxVerts + 31) / 32;

// Whether split (u, v) of a path needs a visibility test (both ends
// scatter diffusely and are at least EPSILON apart: Path::getPd, :142-190).
__device__ __forceinline__ bool pair_needed(const Scene &sc, const Vtx &a, const Vtx &b) {
  if (specular_of(sc, a.info) || specular_of(sc, b.info)) return false;
  return len(sub(b.p, a.p)) >= kEps;
}

// Path::append(const Path &) (Path.cpp:339-515) up to the visibility tests:
// the emitter an eye subpath ended on scatters as that emitter; the join of
// the light subpath's last vertex y and the eye subpath's last vertex z
// (_initL on y, _initE on z, stored back); then every vertex pair two or
// more apart that a split's getPd / getContribution tests for visibility,
// and the join itself when both ends succeeded, go to the pair queue as
// (slot, u | v << 8).  The slot's occlusion mask is cleared.
__global__ __launch_bounds__(256) void k_bd_join(Scene sc, const BatchDesc *__restrict__ bd,
                                                 float4 *__restrict__ verts, uint4 *__restrict__ info, uint32_t V,
                                                 const uint2 *__restrict__ pjs, uint2 *__restrict__ pairs,
                                                 uint32_t *__restrict__ pair_count, uint32_t *__restrict__ masks) {
  __shared__ uint32_t append_sh[256 / 64 + 1];
  const uint32_t n_slots = bd->hdr.n_paths;
  const uint32_t E = V / 2;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n_slots; base += gridDim.x * blockDim.x) {
    const uint32_t slot = base + threadIdx.x;
    uint32_t cnt = 0, S = 0, T = 0, flags = 0;
    float4 *vs = verts + (uint64_t)slot * V * kVtx;
    if (slot < n_slots) {
      const uint4 in = info[slot];
      T = in.x;
      S = in.z & S_EYE_ON_EMITTER ? 0u : in.y;
      flags = in.z & S_EYE_ON_EMITTER;
      uint32_t *mask = masks + (uint64_t)slot * kMaskWords;
      for (uint32_t k = 0; k < kMaskWords; ++k) mask[k] = 0u;
      if (S > 0) {
        const uint2 pj = pjs[slot];
        const JobDev &job = bd->jobs[pj.y >> kJobBits];
        const uint32_t pixel = pj.x, sample = pj.y & ((1u << kJobBits) - 1);
        const mpt_rng::U4 ju = mpt_rng::path_draw(job.seed, job.job_id, pixel, sample, kJoinDepth, 0);
        Vtx y = load_vtx(vs + rec_of(S - 1, S, T, E) * kVtx), z = load_vtx(vs + rec_of(S, S, T, E) * kVtx);
        float3 wo = sub(z.p, y.p);
        const float d = len(wo);
        wo = mul(wo, 1.0f / d);
        const float3 wiy = S > 1 ? normalize(sub(y.p, load_vtx(vs + rec_of(S - 2, S, T, E) * kVtx).p)) : mul(wo, -1.0f);
        float3 alpha;
        double pc;
        const bool ok_y = init_side<false>(sc, y, wiy, wo, d, z, false, mpt_rng::unitf(ju.x), alpha, pc);
        const float3 wiz = T > 1 ? normalize(sub(z.p, load_vtx(vs + rec_of(S + 1, S, T, E) * kVtx).p)) : wo;
        const bool ok_z = init_side<true>(sc, z, wiz, mul(wo, -1.0f), d, y, false, mpt_rng::unitf(ju.y), alpha, pc);
        store_vtx(vs + rec_of(S - 1, S, T, E) * kVtx, y);
        store_vtx(vs + rec_of(S, S, T, E) * kVtx, z);
        if (ok_y && ok_z) {
          flags |= S_JOIN_OK;
          ++cnt;
        }
      }
      const uint32_t nv = S + T;
      for (uint32_t u = 0; u + 2 < nv; ++u) {
        const Vtx a = load_vtx(vs + rec_of(u, S, T, E) * kVtx);
        const uint32_t ka = S == 0 && u == 0 ? K_EMITTER : kind_of(a.info);
        Vtx a2 = a;
        a2.info = (a.info & ~3u) | ka;
        for (uint32_t v = u + 2; v < nv; ++v)
          if (pair_needed(sc, a2, load_vtx(vs + rec_of(v, S, T, E) * kVtx))) ++cnt;
      }
      info[slot].z = flags;
      info[slot].y = S;
    }
    uint32_t at = block_append_n<256 / 64>(cnt, pair_count, append_sh);
    if (cnt) {
      const uint32_t nv = S + T;
      if (flags & S_JOIN_OK) pairs[at++] = make_uint2(slot, (S - 1) | (S << 8));
      for (uint32_t u = 0; u + 2 < nv; ++u) {
        Vtx a = load_vtx(vs + rec_of(u, S, T, E) * kVtx);
        if (S == 0 && u == 0) a.info = (a.info & ~3u) | K_EMITTER;
        for (uint32_t v = u + 2; v < nv; ++v)
          if (pair_needed(sc, a, load_vtx(vs + rec_of(v, S, T, E) * kVtx))) pairs[at++] = make_uint2(slot, u | (v << 8));
      }
    }
  }
}

// Any hit between the two vertices of every queued pair (Scene::intersects
// over (EPSILON, t - EPSILON), core/Scene.cpp:110-114): the ray leaves the
// first vertex's primitive and ends 2 EPSILON before the second; a planar
// target primitive is skipped (its only crossing is the end point), a sphere
// target ends the window 1e-4 of the distance early instead.
template <int STACK>
__global__ __launch_bounds__(256) void k_bd_vis(Scene sc, const uint2 *__restrict__ pairs,
                                                const uint32_t *__restrict__ pair_count,
                                                const float4 *__restrict__ verts, const uint4 *__restrict__ info,
                                                uint32_t V, uint32_t *__restrict__ masks) {
  __shared__ uint2 stack[STACK * kBlock];
  uint2 *stk = stack + threadIdx.x;
  const uint32_t n = *pair_count;
  const uint32_t E = V / 2;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint2 e = pairs[i];
    const uint32_t slot = e.x, u = e.y & 255u, v = e.y >> 8;
    const uint4 in = info[slot];
    const uint32_t S = in.y, T = in.x;
    const float4 *vs = verts + (uint64_t)slot * V * kVtx;
    const float4 a = vs[rec_of(u, S, T, E) * kVtx], b = vs[rec_of(v, S, T, E) * kVtx];
    float3 w = sub(xyz(b), xyz(a));
    const float d = len(w);
    w = mul(w, 1.0f / d);
    const uint32_t pa = __float_as_uint(a.w), pb = __float_as_uint(b.w);
    const bool sphere_target = pb != kMiss && pb >= sc.nt && pb < sc.nt + sc.ns;
    const float tlimit = sphere_target ? d * (1.0f - 1e-4f) : d - 2.0f * kEps;
    const RaySetup r = setup_ray(xyz(a), w);
    float t;
    uint32_t id;
    if (trace<true, false, STACK>(sc, r, tlimit, pa, stk, t, id, nullptr, nullptr, nullptr,
                                  sphere_target ? kMiss : pb)) {
      const uint32_t bit = pair_bit(u, v, V);
      atomicOr(masks + (uint64_t)slot * kMaskWords + bit / 32, 1u << (bit % 32));
    }
  }
}

// The combine kernel's work grows with the square of a path's vertex count,
// so its lanes take the slots ordered by that count (a counting sort of the
// slot indices): a wave then holds paths of one length instead of waiting
// for its longest.  The order within one length is free (results are per
// slot).
__global__ __launch_bounds__(256) void k_bd_len_count(const BatchDesc *__restrict__ bd,
                                                      const uint4 *__restrict__ info, uint32_t *__restrict__ bins) {
  __shared__ uint32_t sh[kMaxVerts + 1];
  for (uint32_t k = threadIdx.x; k <= kMaxVerts; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  const uint32_t n = bd->hdr.n_paths;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint4 in = info[i];
    atomicAdd(&sh[min(in.x + in.y, kMaxVerts)], 1u);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k <= kMaxVerts; k += blockDim.x)
    if (sh[k]) atomicAdd(&bins[k], sh[k]);
}

// bins -> exclusive offsets
__global__ void k_bd_len_scan(uint32_t *__restrict__ bins) {
  if (threadIdx.x != 0) return;
  uint32_t sum = 0;
  for (uint32_t k = 0; k <= kMaxVerts; ++k) {
    const uint32_t c = bins[k];
    bins[k] = sum;
    sum += c;
  }
}

__global__ __launch_bounds__(256) void k_bd_len_scatter(const BatchDesc *__restrict__ bd,
                                                        const uint4 *__restrict__ info,
                                                        uint32_t *__restrict__ cursors, uint32_t *__restrict__ order) {
  __shared__ uint32_t cnt[kMaxVerts + 1], base[kMaxVerts + 1];
  const uint32_t n = bd->hdr.n_paths;
  for (uint32_t b0 = blockIdx.x * blockDim.x; b0 < n; b0 += gridDim.x * blockDim.x) {
    for (uint32_t k = threadIdx.x; k <= kMaxVerts; k += blockDim.x) cnt[k] = 0;
    __syncthreads();
    const uint32_t i = b0 + threadIdx.x;
    uint32_t len = 0, rank = 0;
    if (i < n) {
      const uint4 in = info[i];
      len = min(in.x + in.y, kMaxVerts);
      rank = atomicAdd(&cnt[len], 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k <= kMaxVerts; k += blockDim.x)
      base[k] = cnt[k] ? atomicAdd(&cursors[k], cnt[k]) : 0u;
    __syncthreads();
    if (i < n) order[base[len] + rank] = i;
    __syncthreads();
  }
}

// The rest of BidirectionalPathTracer::sample, one slot per lane: Path::
// append's cumulative values across the join (Path.cpp:397-512), getPd and
// getContribution of every split (:48-197), the power heuristic per path
// length k (BidirectionalPathTracer.cpp:111-165), and RADIANCE_HACK's clamp
// to [0, 1] (:167-171).  Densities and the weights are double precision.
// VM: the most vertices a path of the batch can have (the per-lane arrays
// are sized by it).
template <int VM>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_bd_combine(Scene sc, const BatchDesc *__restrict__ bd,
                                                    const float4 *__restrict__ verts, const uint4 *__restrict__ info,
                                                    uint32_t V, const uint32_t *__restrict__ masks,
                                                    const uint32_t *__restrict__ order, float4 *__restrict__ Ls) {
  const uint32_t lane_slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (lane_slot >= bd->hdr.n_paths) return;
  const uint32_t slot = order[lane_slot];
  const uint32_t E = V / 2;
  const uint4 in = info[slot];
  const uint32_t T = in.x, S = in.y, flags = in.z, n = S + T;
  const float4 *vs = verts + (uint64_t)slot * V * kVtx;
  // the slot's occlusion mask in registers: its words are picked by a select
  // chain over a compile-time count, so no lookup goes back to memory
  constexpr uint32_t kW = (VM * VM + 31) / 32;
  uint32_t mw[kW];
  const uint32_t *mask = masks + (uint64_t)slot * kMaskWords;
#pragma unroll
  for (uint32_t j = 0; j < kW; ++j) mw[j] = mask[j];
  auto occluded = [&](uint32_t u, uint32_t v) {
    const uint32_t bit = pair_bit(u, v, V), w = bit >> 5;
    uint32_t word = mw[0];
#pragma unroll
    for (uint32_t j = 1; j < kW; ++j) word = w == j ? mw[j] : word;
    return (word >> (bit & 31u)) & 1u;
  };
  // the merged path's values in the reference's terms: GL[i] is the
  // geometry term of edge (i, i + 1), GE[i] of edge (i - 1, i)
  float3 P[VM], Nn[VM], fs[VM], aL[VM], aE[VM];
  double pL[VM], pE[VM];
  float pdfL[VM], pdfE[VM], GL[VM], GE[VM];
  uint32_t inf[VM];
  uint32_t emits = 0;   // bit i: vertex i's surface emits
  for (uint32_t i = 0; i < n; ++i) {
    const Vtx v = load_vtx(vs + rec_of(i, S, T, E) * kVtx);
    P[i] = v.p;
    Nn[i] = v.n;
    fs[i] = v.fs;
    inf[i] = v.info;
    const bool light = i < S;
    (light ? aL[i] : aE[i]) = v.alpha;
    (light ? pL[i] : pE[i]) = v.pcum;
    (light ? pdfL[i] : pdfE[i]) = v.pdf_fwd;
    (light ? pdfE[i] : pdfL[i]) = v.pdf_rev;
    (light ? GE[i] : GL[i]) = v.gprev;
    (light ? GL[i] : GE[i]) = v.gnext;
  }
  // the emitter an eye subpath ended on scatters as that emitter
  if (S == 0) inf[0] = (inf[0] & ~3u) | K_EMITTER;
  // specular and emitting flags, looked up once per vertex
  for (uint32_t i = 0; i < n; ++i) {
    if (specular_of(sc, inf[i])) inf[i] |= I_SPEC;
    if (on_emitter(sc, inf[i])) emits |= 1u << i;
  }
  bool valid = true;
  if (S > 0) {
    if (!(flags & S_JOIN_OK)) {
      valid = false;
    } else if (occluded(S - 1, S)) {
      GL[S - 1] = GE[S] = 0.0f;
      valid = false;
    }
  }
  // the eye-side values of the light vertices ...
  for (uint32_t i = S; i--;) {
    const bool ok = pdfE[i + 1] > 0.0f && GE[i + 1] > 0.0f;
    aE[i] = ok ? mul(mul(aE[i + 1], fs[i + 1]), 1.0f / pdfE[i + 1]) : f3(0, 0, 0);
    pE[i] = pE[i + 1] * (double)pdfE[i + 1] * (double)GE[i + 1];
  }
  const float pA = sc.emit_area > 0.0f ? 1.0f / sc.emit_area : 1.0f;
  // ... and the light-side values of the eye vertices
  for (uint32_t i = S; i < n; ++i) {
    if (i == 0) {
      // an eye subpath that ended on an emitter starts the light side with
      // the emitter's EmitterSampler density (Path.cpp:466-497)
      Vtx z;
      z.p = P[0];
      z.n = Nn[0];
      z.info = inf[0];
      const float3 wo = normalize(sub(P[1], P[0]));
      aL[0] = mul(le_of(sc.mats[mat_of(inf[0])]), kPi / pA);
      pL[0] = (double)pA;
      fs[0] = f_v(sc, z, f3(0, 0, 0), wo);
      pdfL[0] = pdf_v(sc, z, f3(0, 0, 0), wo);
    } else {
      const bool ok = pdfL[i - 1] > 0.0f && GL[i - 1] > 0.0f;
      aL[i] = ok ? mul(mul(aL[i - 1], fs[i - 1]), 1.0f / pdfL[i - 1]) : f3(0, 0, 0);
      pL[i] = pL[i - 1] * (double)pdfL[i - 1] * (double)GL[i - 1];
    }
  }
  auto has_area = [&](uint32_t i) { return (inf[i] & I_AREA) != 0u; };
  // Path::getPd (Path.cpp:142-197)
  auto pd = [&](uint32_t s, uint32_t t) -> double {
    const uint32_t u = s - 1, v = n - t;
    if (s == 0) return has_area(v) ? pE[v] : 0.0;
    if (t == 0) return has_area(u) ? pL[u] : 0.0;
    if ((inf[u] | inf[v]) & I_SPEC) return 0.0;
    if (s + t == n) return pL[u] * pE[v];
    if (len(sub(P[v], P[u])) < kEps) return 0.0;
    if (occluded(u, v)) return 0.0;
    return pL[u] * pE[v];
  };
  // Path::getContribution (Path.cpp:48-140)
  auto contribution = [&](uint32_t s, uint32_t t) -> float3 {
    const uint32_t u = s - 1, v = n - t;
    if (s == 0) {
      if (!has_area(v)) return f3(0, 0, 0);
      const float3 wo = v < n - 1 ? normalize(sub(P[v + 1], P[v])) : f3(0, 0, 0);
      // Emitter::getLe of the vertex's surface (OrientedEmitter.cpp:17-23;
      // NullEmitter is black)
      if (!((emits >> v) & 1u) || !(dot(Nn[v], wo) > 0.0f)) return f3(0, 0, 0);
      return mul(le_of(sc.mats[mat_of(inf[v])]), aE[v]);
    }
    if (t == 0) return f3(0, 0, 0);   // NullSensor::getWe is black off the camera, which has no area
    if ((inf[u] | inf[v]) & I_SPEC) return f3(0, 0, 0);
    if (s + t == n) return mul(mul(mul(aL[u], fs[u]), GL[u]), mul(fs[v], aE[v]));
    float3 wo = sub(P[v], P[u]);
    const float d = len(wo);
    if (d < kEps) return f3(0, 0, 0);
    if (occluded(u, v)) return f3(0, 0, 0);
    wo = mul(wo, 1.0f / d);
    Vtx y, z;
    y.p = P[u];
    y.n = Nn[u];
    y.info = inf[u];
    z.p = P[v];
    z.n = Nn[v];
    z.info = inf[v];
    const float3 ywi = u > 0 ? normalize(sub(P[u], P[u - 1])) : f3(0, 0, 0);
    const float3 wo2 = v < n - 1 ? normalize(sub(P[v + 1], P[v])) : mul(wo, -1.0f);
    const float G = fabsf(ndot(y, wo) * ndot(z, mul(wo, -1.0f))) / (d * d);
    return mul(mul(mul(aL[u], f_v(sc, y, ywi, wo)), G), mul(f_v(sc, z, wo, wo2), aE[v]));
  };
  const uint32_t length = n - (valid ? 0u : 1u);
  double L[3] = {0.0, 0.0, 0.0};
  double pdf[VM + 1];
  for (uint32_t k = 2; k <= length; ++k) {
    double sum = 0.0;
    for (uint32_t s = 0; s <= k; ++s) {
      const double p = pd(s, k - s);
      pdf[s] = p * p;   // power heuristic, beta = 2
      sum += pdf[s];
    }
    if (sum == 0.0) continue;
    for (uint32_t s = 0; s <= k; ++s) {
      if (!(pdf[s] > 0.0)) continue;
      const float3 c = contribution(s, k - s);
      const double w = pdf[s] / sum;
      L[0] += w * c.x;
      L[1] += w * c.y;
      L[2] += w * c.z;
    }
  }
  // RADIANCE_HACK: CLAMP(value, 0, 1) (common/utilities.h:195-200)
  float o[3];
  for (int c = 0; c < 3; ++c) o[c] = (float)(L[c] >= 1.0 ? 1.0 : (L[c] <= 0.0 ? 0.0 : L[c]));
  Ls[slot] = make_float4(o[0], o[1], o[2], 0.0f);
}

// Camera, extension (eye and light subpath) and visibility rays of a batch.
__global__ void k_bd_totals(const uint32_t *__restrict__ counts, uint32_t D,
                            unsigned long long *__restrict__ totals) {
  if (threadIdx.x != 0) return;
  unsigned long long ext = 0;
  for (uint32_t b = 0; b < 2 * (D + 1); ++b) ext += counts[b];
  totals[0] += counts[0];
  totals[1] += ext;
  totals[2] += counts[2 * (D + 1)];
}

}  // namespace bdpt

#endif
