// Metropolis light transport on the device: MLTRenderer (renderers/mlt/
// MLTRenderer.cpp:70-235), MLTMarkovProcess (MLTMarkovProcess.cpp:38-176),
// the aggregate of MLTBidirPathMutation, MLTLensSubpathMutation and
// MLTPerturbationPathMutation, over Path's general operations
// (renderers/utils/Path.cpp:208-720) and RenderOutput's MLT film
// (renderers/RenderOutput.cpp:72-216).
//
// One lane is one Markov chain (the reference runs one per render thread).
// Kernels:
//   k_mlt_seed      one seed path per lane (BidirectionalPathTracer::
//                   generate, retried until it has at most mltMaxDepth
//                   vertices) and every split (s, t) of it joined, weighted
//                   by its contribution's luminance; the random-stream
//                   position of each split's join is kept so a chain can
//                   rebuild the split it starts from
//   k_mlt_weights   one block: the entries in the reference's order, their
//                   sum, b = sum / entries, the normalised running sum that
//                   Random::sampleNormalizedCDF walks
//   k_mlt_start     per chain: its seed entry, the split rebuilt, the first
//                   sample
//   k_mlt_step      per chain, a run of mutations: the aggregate's choice,
//                   the mutation, acceptance, and the two samples each adds
//                   to the film (fixed-point sums, so films are the same on
//                   every run), the proposal counts and the sample count
//   k_mlt_finish    per pixel: the 3 x 3 median of the proposal counts, the
//                   scale, 2 / samples
// Rays are traced inline by the lane (the kd walk of the path tracer).
// Random numbers: sequential streams shared with the oracle, keyed (seed
// path, attempt), (chain) and (chain, mutation); a BSDF, emitter or sensor
// sample takes 4 uniforms, the emitter sampler 5, a random camera point 4, a
// roulette 1.
//
// Included by mpt.hip after bdpt.h, whose vertex helpers it uses.
#ifndef MPT_MLT_H_
#define MPT_MLT_H_

namespace mlt {

using namespace bdpt;

constexpr uint32_t kSeedKind = 0xF0, kInitKind = 0xF1, kStepKind = 0xF2;
constexpr uint32_t kAttempts = 64;   // generate() retries of one seed path (the reference retries forever)
constexpr uint32_t kRounds = 16;     // rounds of seed paths while every weight is 0
constexpr uint32_t kMaxLen = 16;     // most vertices of a chain's path (mltMaxDepth)
constexpr uint32_t kCap = kMaxLen + 1;   // vertex records per path buffer: a subpath may grow one past the cap
// splits (k, s) of one seed path of kMaxLen vertices: sum over k = 2..16 of k + 1
constexpr uint32_t kMaxEntries = (kMaxLen + 1) * (kMaxLen + 2) / 2 - 3;
constexpr uint32_t kSlots = 12;      // float4 per vertex record
constexpr uint32_t kWave = 64;       // lanes that share a run of path records
constexpr uint64_t kRun = (uint64_t)kCap * kSlots * kWave;   // float4 of one wave's path buffer
// runs of path records for n lanes
__host__ __device__ constexpr uint64_t wave_runs(uint64_t n) { return (n + kWave - 1) / kWave; }
constexpr uint32_t kBlock = 128;     // threads per block of the chain kernels
// threads per block of the seed kernel: its few waves (10,000 seed paths
// by default) spread over more CUs; the kernels' LDS rings are laid out for
// kBlock lanes, so any smaller block stays within them
constexpr uint32_t kSeedBlock = 64;
static_assert(kSeedBlock <= kBlock, "the seed kernel's ring rows hold kBlock lanes");
constexpr uint32_t kStackRows = 8;   // traversal ring rows per lane
constexpr double kFix = 16777216.0;  // film sums: units of b / 2^24

// event bits of a vertex: the event carries DielectricMetaData (reflected or
// refracted, and its channel); the vertex is the camera's point
constexpr uint32_t E_META = 1u, E_REFL = 2u, E_CAM = 16u;
__device__ __forceinline__ uint32_t ev_ch(uint32_t e) { return (e >> 2) & 3u; }

// Spectra and the values built from them in double precision: a Phong
// vertex's pdf is that of its event's lobe alone, and the contributions of
// seed splits and mutated paths reach 1e44 and more, past FP32's range.
struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 d3(float3 a) { return D3{a.x, a.y, a.z}; }
__device__ __forceinline__ D3 dmul(D3 a, D3 b) { return D3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ D3 dscale(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ bool dzero(D3 a) { return a.x == 0 && a.y == 0 && a.z == 0; }

// PathVertex (renderers/utils/PathVertex.h:38-120): position and primitive,
// normal and info word (the BSDF kind it scatters with, Phong lobe of its
// event, flags, material), wi, the event's direction and mode, the film
// point of a camera vertex; in double both sides' geometry terms, pdfs,
// cumulative densities and contributions, and fs.
struct MVtx {
  float3 p;
  uint32_t prim;
  float3 n;
  uint32_t info;
  float3 wi;
  uint32_t evb;
  float3 ev;
  float fu, fv;
  double GL, GE, pdfL, pdfE, pL, pE;
  D3 alphaL, alphaE, fs;
};

// A chain's (or seed lane's) path buffer: slot k of vertex i at
// base[(i * kSlots + k) * stride].  Buffers are laid out by wave: the 64
// lanes of a wave interleave (stride 64), so a wave's load of one slot is
// 1 KiB contiguous, and each wave's records are one contiguous run
// (kCap * kSlots * 1 KiB), which keeps a wave's accesses within a few pages.
struct Store {
  float4 *base;
  uint32_t stride;
  __device__ __forceinline__ float4 &at(uint32_t i, uint32_t k) const {
    return base[((uint64_t)i * kSlots + k) * stride];
  }
  __device__ __forceinline__ void get2(uint32_t i, uint32_t k, double &a, double &b) const {
    const float4 q = at(i, k);
    a = __hiloint2double(__float_as_int(q.y), __float_as_int(q.x));
    b = __hiloint2double(__float_as_int(q.w), __float_as_int(q.z));
  }
  __device__ __forceinline__ void put2(uint32_t i, uint32_t k, double a, double b) const {
    at(i, k) = make_float4(__int_as_float(__double2loint(a)), __int_as_float(__double2hiint(a)),
                           __int_as_float(__double2loint(b)), __int_as_float(__double2hiint(b)));
  }
  __device__ MVtx get(uint32_t i) const {
    const float4 a = at(i, 0), b = at(i, 1), c = at(i, 2), d = at(i, 3);
    MVtx v;
    v.p = xyz(a);
    v.prim = __float_as_uint(a.w);
    v.n = xyz(b);
    v.info = __float_as_uint(b.w);
    v.wi = xyz(c);
    v.evb = __float_as_uint(c.w);
    v.ev = xyz(d);
    v.fu = d.w;
    double fv;
    get2(i, 4, v.GL, v.GE);
    get2(i, 5, v.pdfL, v.pdfE);
    get2(i, 6, v.pL, v.pE);
    get2(i, 7, v.alphaL.x, v.alphaL.y);
    get2(i, 8, v.alphaL.z, v.alphaE.x);
    get2(i, 9, v.alphaE.y, v.alphaE.z);
    get2(i, 10, v.fs.x, v.fs.y);
    get2(i, 11, v.fs.z, fv);
    v.fv = (float)fv;
    return v;
  }
  __device__ void put(uint32_t i, const MVtx &v) const {
    at(i, 0) = make_float4(v.p.x, v.p.y, v.p.z, __uint_as_float(v.prim));
    at(i, 1) = make_float4(v.n.x, v.n.y, v.n.z, __uint_as_float(v.info));
    at(i, 2) = make_float4(v.wi.x, v.wi.y, v.wi.z, __uint_as_float(v.evb));
    at(i, 3) = make_float4(v.ev.x, v.ev.y, v.ev.z, v.fu);
    put2(i, 4, v.GL, v.GE);
    put2(i, 5, v.pdfL, v.pdfE);
    put2(i, 6, v.pL, v.pE);
    put2(i, 7, v.alphaL.x, v.alphaL.y);
    put2(i, 8, v.alphaL.z, v.alphaE.x);
    put2(i, 9, v.alphaE.y, v.alphaE.z);
    put2(i, 10, v.fs.x, v.fs.y);
    put2(i, 11, v.fs.z, (double)v.fv);
  }
  __device__ void copy(uint32_t dst, const Store &src, uint32_t i) const {
#pragma unroll
    for (uint32_t k = 0; k < kSlots; ++k) at(dst, k) = src.at(i, k);
  }
};

// Lane `lane`'s records in a buffer of runs of 64 lanes.
__device__ __forceinline__ Store lane_store(float4 *buf, uint32_t lane) {
  return Store{buf + (uint64_t)(lane / kWave) * kRun + lane % kWave, kWave};
}

// Sequential uniforms of one stream (the oracle's Seq).
struct Seq {
  uint32_t seed, job, x, y, kind, n;
  mpt_rng::U4 blk;
  __device__ Seq(uint32_t seed_, uint32_t job_, uint32_t x_, uint32_t y_, uint32_t kind_, uint32_t n_ = 0)
      : seed(seed_), job(job_), x(x_), y(y_), kind(kind_), n(n_) {
    if (n & 3u) load(n >> 2);
  }
  __device__ __forceinline__ void load(uint32_t block) {
    blk = mpt_rng::path_draw(seed, job, x, y, (kind << 16) | (block >> 8), block & 255u);
  }
  __device__ float next() {
    if ((n & 3u) == 0) load(n >> 2);
    const uint32_t k = n++ & 3u;
    return mpt_rng::unitf(k == 0 ? blk.x : (k == 1 ? blk.y : (k == 2 ? blk.z : blk.w)));
  }
};

// What one chain kernel needs besides the scene.
struct Ctx {
  CameraDev cam;
  uint32_t W, H;
  float invW, invH;
  double pA_film;   // invW * invH, the film bin's area
  uint2 *stk;       // the lane's traversal ring
  uint32_t n_ext, n_vis;
};

__device__ __forceinline__ Vtx view(const MVtx &v) {
  Vtx t;
  t.n = v.n;
  t.info = v.info;
  return t;
}
__device__ __forceinline__ bool spec(const Scene &sc, const MVtx &v) { return specular_of(sc, v.info); }
__device__ __forceinline__ bool is_cam(const MVtx &v) { return (v.evb & E_CAM) != 0u; }
__device__ __forceinline__ bool on_em(const Scene &sc, const MVtx &v) { return !is_cam(v) && on_emitter(sc, v.info); }
__device__ __forceinline__ bool has_area(const MVtx &v) { return (v.info & I_AREA) != 0u; }
// Shape::getMaterial of the vertex's point (the camera's is its own)
__device__ __forceinline__ uint32_t material_key(const MVtx &v) {
  if (is_cam(v)) return 0xFFFFFFFEu;
  if (v.info & I_POINT) return 0x80000000u | mat_of(v.info);
  return mat_of(v.info);
}
__device__ __forceinline__ double ddot(float3 a, float3 b) {
  return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z;
}
__device__ __forceinline__ double ndot_m(const MVtx &v, float3 w) { return (v.info & I_NO_NORMAL) ? 1.0 : ddot(v.n, w); }
__device__ __forceinline__ double lum(D3 c) { return 0.2126 * c.x + 0.7152 * c.y + 0.0722 * c.z; }
__device__ __forceinline__ double lum_ratio(D3 y, D3 x) {
  return 0.2126 * (y.x / x.x) + 0.7152 * (y.y / x.y) + 0.0722 * (y.z / x.z);
}
// MIN (common/utilities.h:161-165)
__device__ __forceinline__ double ref_min(double x, double y) { return x < y ? x : y; }

// Emitter::getLe0: pi Le for an area light, 4 pi Le for a point
__device__ D3 le0_of(const Scene &sc, const MVtx &v) {
  if (v.info & I_POINT) return dscale(d3(xyz(sc.points[2 * mat_of(v.info) + 1])), 4.0 * M_PI);
  const mpt_material &m = sc.mats[mat_of(v.info)];
  return emissive(m) ? dscale(d3(le_of(m)), M_PI) : d3(0, 0, 0);
}
// Emitter::getLe of the vertex's emitter toward wo
__device__ D3 le_toward(const Scene &sc, const MVtx &v, float3 wo) {
  if (v.info & I_POINT) return d3(xyz(sc.points[2 * mat_of(v.info) + 1]));
  if (is_cam(v)) return d3(0, 0, 0);
  const mpt_material &m = sc.mats[mat_of(v.info)];
  return dot(v.n, wo) > 0.0f ? d3(le_of(m)) : d3(0, 0, 0);
}

// BSDF::getPd(Event(dir, event)) of vertex v with incident direction wi, in
// double: the bidirectional kernels' pdf_v with every product widened.
__device__ double pdf_d(const Scene &sc, const MVtx &v, float3 wi, float3 dir) {
  const uint32_t k = kind_of(v.info);
  if (k == K_SENSOR) {
    if (same_dir(wi, mul(v.n, -1.0f))) return 1.0;
    return same_dir(dir, v.n) ? 1.0 : 0.0;
  }
  if (k == K_EMITTER) {
    if (v.info & I_POINT) return 1.0;
    return dot(v.n, dir) > 0.0f ? 1.0 / M_PI : 0.0;
  }
  const mpt_material &m = sc.mats[mat_of(v.info)];
  if (m.type == MPT_BSDF_DIFFUSE) {
    if (dot(v.n, wi) > 0.0f) return 0.0;
    return dot(v.n, dir) > 0.0f ? 1.0 / M_PI : 0.0;
  }
  if (m.type == MPT_BSDF_PHONG) {
    const double kda = ((double)m.kd[0] + m.kd[1] + m.kd[2]) / 3.0, ksa = ((double)m.ks[0] + m.ks[1] + m.ks[2]) / 3.0;
    const uint32_t lobe = lobe_of(v.info);
    if (lobe == LOBE_ABSORB) return 1.0 - (kda + ksa);
    const double cosA = ddot(v.n, dir);
    if (cosA <= 0) return 0.0;
    double pdf;
    if (lobe == LOBE_SPECULAR) {
      const double c = fabs(ddot(reflect(wi, v.n), dir));
      pdf = fmin(1.0, ((double)m.exponent + 1.0) / (2.0 * M_PI) * pow(c, (double)m.exponent));
    } else {
      pdf = fabs(ddot(v.n, dir)) / M_PI;
    }
    return pdf / cosA;
  }
  return (double)pdf_v(sc, view(v), wi, dir);   // a dielectric: 0, 1/3, 2/3 or 1
}

// BSDF::evaluate(wi, dir) of vertex v in double.
__device__ D3 f_d(const Scene &sc, const MVtx &v, float3 wi, float3 dir) {
  const uint32_t k = kind_of(v.info);
  if (k == K_SENSOR) return same_dir(mul(wi, -1.0f), v.n) ? d3(1, 1, 1) : d3(0, 0, 0);
  if (k == K_EMITTER) {
    if (v.info & I_POINT) {
      const double f = 0.25 / M_PI;
      return d3(f, f, f);
    }
    if (!(dot(v.n, dir) > 0.0f)) return d3(0, 0, 0);
    const float3 le = le_of(sc.mats[mat_of(v.info)]);
    const double f = 1.0 / M_PI;
    return d3(le.x != 0.0f ? f : 0.0, le.y != 0.0f ? f : 0.0, le.z != 0.0f ? f : 0.0);
  }
  const mpt_material &m = sc.mats[mat_of(v.info)];
  if (m.type == MPT_BSDF_DIFFUSE) {
    if (dot(v.n, wi) >= 0.0f || dot(v.n, dir) <= 0.0f) return d3(0, 0, 0);
    return dscale(d3(kd_of(m)), 1.0 / M_PI);
  }
  if (m.type == MPT_BSDF_PHONG) {
    if (dot(v.n, dir) <= 0.0f) return d3(0, 0, 0);
    const double a = fmin(1.0, fmax(0.0, ddot(reflect(wi, v.n), dir)));
    const double kda = ((double)m.kd[0] + m.kd[1] + m.kd[2]) / 3.0, ksa = ((double)m.ks[0] + m.ks[1] + m.ks[2]) / 3.0;
    const double an = pow(a, (double)m.exponent);
    const double fs_s = ksa > 0 ? ((double)m.exponent + 2.0) / (2.0 * M_PI * ksa) * an : 0.0;
    const double fs_d = kda > 0 ? 1.0 / (M_PI * kda) : 0.0;
    return d3(m.kd[0] * fs_d + m.ks[0] * fs_s, m.kd[1] * fs_d + m.ks[1] * fs_s, m.kd[2] * fs_d + m.ks[2] * fs_s);
  }
  return d3(f_v(sc, view(v), wi, dir));   // a dielectric: ks on its events
}
__device__ __forceinline__ bool black_d(D3 a) { return fabs(a.x) <= kEps && fabs(a.y) <= kEps && fabs(a.z) <= kEps; }
__device__ __forceinline__ double max_d(D3 a) {
  double m = a.x;
  if (a.y > m) m = a.y;
  if (a.z > m) m = a.z;
  return m;
}

// BSDF::sample(e) of v with incident direction wi; keep: the event whose
// scattering mode a dielectric preserves (DielectricBSDF.cpp:55-104), null
// for sample().  Sets v's event.
__device__ void mv_sample(const Scene &sc, MVtx &v, float3 wi, Seq &rng, const MVtx *keep) {
  const float u0 = rng.next(), u1 = rng.next(), u2 = rng.next(), u3 = rng.next();
  if (kind_of(v.info) == K_SURFACE && sc.mats[mat_of(v.info)].type == MPT_BSDF_DIELECTRIC) {
    const mpt_material &m = sc.mats[mat_of(v.info)];
    const float3 N = dot(v.n, wi) > 0.0f ? mul(v.n, -1.0f) : v.n;
    const uint32_t c = min(2u, (uint32_t)((double)u3 * 3.0));   // Random::sampleInt(0, 3)
    if (keep && (keep->evb & E_META)) {
      if (keep->evb & E_REFL) {
        v.evb = (v.evb & E_CAM) | E_META | E_REFL | (c << 2);
        v.ev = reflect(wi, N);
      } else {
        const uint32_t kc = ev_ch(keep->evb);
        v.evb = (v.evb & E_CAM) | E_META | (kc << 2);
        float3 wt;
        v.ev = refract(wi, v.n, kAirIor, channel_ior(m, kc), wt) ? wt : f3(0, 0, 0);
      }
      return;
    }
    const float ior = channel_ior(m, c);
    float3 wt;
    if (refract(wi, v.n, kAirIor, ior, wt)) {
      const float Fr = fresnel(wi, v.n, kAirIor, ior);
      if (u2 < m.transparency * (1.0f - Fr)) {
        v.evb = (v.evb & E_CAM) | E_META | (c << 2);
        v.ev = wt;
        return;
      }
    }
    v.evb = (v.evb & E_CAM) | E_META | E_REFL | (c << 2);
    v.ev = reflect(wi, N);
    return;
  }
  v.evb &= E_CAM;
  uint32_t lobe;
  v.ev = sample_v(sc, view(v), wi, u0, u1, u2, u3, lobe);
  v.info = with_lobe(v.info, lobe);
}

// Path::_initL (Path.cpp:615-667) with the roulette's uniform drawn when it
// is taken.
__device__ bool init_L(const Scene &sc, MVtx &y, float3 wi, float3 wo, double t, const MVtx &pt, bool roulette,
                       Seq &rng, D3 &alphaL, double &pL) {
  if (t < kEps) {
    y.GL = 0;
    return false;
  }
  const float3 mwo = mul(wo, -1.0f), mwi = mul(wi, -1.0f);
  y.GL = fabs(ndot_m(y, wo) * ndot_m(pt, mwo)) / (t * t);
  y.pdfE = pdf_d(sc, y, mwo, mwi);
  y.fs = f_d(sc, y, wi, wo);
  y.pdfL = pdf_d(sc, y, wi, wo);
  if (y.GL == 0 || y.pdfL == 0) {
    pL = 0;
    alphaL = d3(0, 0, 0);
    return false;
  }
  if (roulette || black_d(y.fs)) {
    const double q = fmin(0.95, max_d(y.fs) / y.pdfL);
    if (q <= (double)rng.next()) {
      y.pdfL = 0;
      pL = 0;
      alphaL = d3(0, 0, 0);
      return false;
    }
    y.pdfL *= q;
  }
  pL = (y.pdfL * y.GL) * y.pL;
  alphaL = dmul(dscale(y.fs, 1.0 / y.pdfL), y.alphaL);
  return true;
}

// Path::_initE (Path.cpp:669-720).
__device__ bool init_E(const Scene &sc, MVtx &z, float3 wi, float3 wo, double t, const MVtx &pt, bool roulette,
                       Seq &rng, D3 &alphaE, double &pE) {
  if (t < kEps) {
    z.GE = 0;
    return false;
  }
  const float3 mwo = mul(wo, -1.0f), mwi = mul(wi, -1.0f);
  z.GE = fabs(ndot_m(z, wo) * ndot_m(pt, mwo)) / (t * t);
  z.pdfL = pdf_d(sc, z, mwo, mwi);
  z.fs = f_d(sc, z, mwo, mwi);
  z.pdfE = pdf_d(sc, z, wi, wo);
  if (z.GE == 0 || z.pdfE == 0) {
    pE = 0;
    alphaE = d3(0, 0, 0);
    return false;
  }
  if (roulette || black_d(z.fs)) {
    const double q = fmin(0.95, max_d(z.fs) / z.pdfE);
    if (q <= (double)rng.next()) {
      z.pdfE = 0;
      pE = 0;
      alphaE = d3(0, 0, 0);
      return false;
    }
    z.pdfE *= q;
  }
  pE = (z.pdfE * z.GE) * z.pE;
  alphaE = dmul(dscale(z.fs, 1.0 / z.pdfE), z.alphaE);
  return true;
}

// The camera's ray direction through film point (x, y).
__device__ __forceinline__ float3 camera_dir(const CameraDev &cam, float x, float y) {
  const float nx = (2.0f * x - 1.0f) * cam.tan_w, ny = (1.0f - 2.0f * y) * cam.tan_h;
  const float3 U = f3(cam.u[0], cam.u[1], cam.u[2]), Vv = f3(cam.v[0], cam.v[1], cam.v[2]);
  const float3 N = f3(cam.n[0], cam.n[1], cam.n[2]);
  return normalize(sub(add(mul(U, nx), mul(Vv, ny)), N));
}

// Camera::getPoint at film point (fx, fy) as PathVertex(pt, 1, pA)
// (PathVertex.cpp:31-66), with its event.
__device__ MVtx camera_vertex(const Scene &sc, const Ctx &cx, float fx, float fy, double pA, Seq &rng) {
  MVtx c;
  c.p = f3(cx.cam.eye[0], cx.cam.eye[1], cx.cam.eye[2]);
  c.prim = kMiss;
  c.n = camera_dir(cx.cam, fx, fy);
  c.info = K_SENSOR | (LOBE_DIFFUSE << 2);   // Camera::_getSurfaceArea is 0
  c.evb = E_CAM;
  c.wi = mul(c.n, -1.0f);
  c.ev = f3(0, 0, 0);
  c.alphaL = d3(1, 1, 1);
  c.alphaE = d3(1.0 / pA, 1.0 / pA, 1.0 / pA);
  c.GL = c.GE = 1.0;
  c.fs = d3(0, 0, 0);
  c.pdfL = c.pdfE = 1.0;
  c.fu = fx;
  c.fv = fy;
  c.pL = 1.0;
  c.pE = pA;
  mv_sample(sc, c, c.wi, rng, nullptr);
  return c;
}

__device__ MVtx hit_vertex(const Scene &sc, uint32_t prim, float3 p, float2 uv) {
  const Vtx s = surface_vtx(sc, prim, p, uv);
  MVtx v;
  v.p = p;
  v.prim = prim;
  v.n = s.n;
  v.info = s.info;
  v.evb = 0;
  v.ev = f3(0, 0, 0);
  v.alphaL = v.alphaE = d3(1, 1, 1);
  v.GL = v.GE = 1.0;
  v.fs = d3(0, 0, 0);
  v.pdfL = v.pdfE = 1.0;
  v.fu = v.fv = 0.0f;
  v.pL = v.pE = 1.0;
  return v;
}

// A subpath being grown at its open end: A[len - 1] is the end (the back of
// a light subpath; the front of an eye subpath, stored camera first).
struct Sub {
  Store s;
  uint32_t len;
};

// Path::_samplePathVertex (Path.cpp:517-610): samples (or, without
// sampleBSDF, takes) the end's event, traces it and adds the vertex hit.
__device__ bool sample_vertex(const Scene &sc, Ctx &cx, Sub &A, bool roulette, bool adjoint, bool sampleBSDF,
                              Seq &rng) {
  const uint32_t n = A.len;
  MVtx v = A.s.get(n - 1);
  float3 wi = v.wi;   // the BSDF's last incident direction (a one-vertex path is an emitter or the camera)
  if (n > 1) wi = normalize(sub(v.p, A.s.get(n - 2).p));
  if (sampleBSDF) mv_sample(sc, v, wi, rng, nullptr);
  const float3 wo = v.ev;
  if (is_black(wo)) {
    A.s.put(n - 1, v);
    return false;
  }
  float t;
  uint32_t id;
  float2 uv;
  ++cx.n_ext;
  trace<false, false, kStackRows, kBlock>(sc, setup_ray(v.p, wo), INFINITY, v.prim, cx.stk, t, id, nullptr, nullptr, &uv);
  if (!(t > kEps && t < INFINITY)) {
    A.s.put(n - 1, v);
    return false;
  }
  if (n == 1) wi = mul(wo, -1.0f);
  v.wi = wi;
  MVtx nv = hit_vertex(sc, id, add(v.p, mul(wo, t)), uv);
  if (adjoint) {
    D3 alphaE;
    double pE;
    const bool ok = init_E(sc, v, wi, wo, t, nv, roulette, rng, alphaE, pE);
    A.s.put(n - 1, v);
    if (!ok) return false;
    nv.wi = wo;
    nv.GL = v.GE;
    nv.GE = 1.0;
    nv.alphaL = d3(1, 1, 1);
    nv.alphaE = alphaE;
    nv.pL = 1.0;
    nv.pE = pE;
  } else {
    D3 alphaL;
    double pL;
    const bool ok = init_L(sc, v, wi, wo, t, nv, roulette, rng, alphaL, pL);
    A.s.put(n - 1, v);
    if (!ok) return false;
    nv.wi = wo;
    nv.GL = 1.0;
    nv.GE = v.GL;
    nv.alphaL = alphaL;
    nv.alphaE = d3(1, 1, 1);
    nv.pL = pL;
    nv.pE = 1.0;
  }
  mv_sample(sc, nv, nv.wi, rng, nullptr);   // PathVertex::_init
  A.s.put(n, nv);
  A.len = n + 1;
  return true;
}

// Path::append(bool) on a light subpath (Path.cpp:234-249): an empty one
// starts on a light by EmitterSampler (EmitterSampler.cpp:51-96).
__device__ bool append_light(const Scene &sc, Ctx &cx, Sub &A, bool roulette, bool sampleBSDF, Seq &rng) {
  if (A.len > 0) return sample_vertex(sc, cx, A, roulette, false, sampleBSDF, rng);
  const float u0 = rng.next(), u1 = rng.next(), u2 = rng.next(), u3 = rng.next(), u4 = rng.next();
  const uint32_t nl = sc.num_lights + sc.num_points;
  uint32_t j = 0;
  while (j + 1 < nl && !(u0 <= sc.light_sel[j].x)) ++j;
  MVtx e;
  e.evb = 0;
  e.ev = f3(0, 0, 0);
  if (j < sc.num_lights) {
    const LightDev &ld = sc.lights[j];
    float3 lp, ln;
    uint32_t target;
    light_point(sc, ld, u1, u2, u3, u4, lp, ln, target);
    e.p = lp;
    e.n = ln;
    e.prim = target;
    e.info = K_EMITTER | (LOBE_DIFFUSE << 2) | (ld.area >= kEps ? I_AREA : 0u) | (ld.mat << 8);
    e.wi = mul(ln, -1.0f);
  } else {
    const uint32_t pi = j - sc.num_lights;
    e.p = xyz(sc.points[2 * pi]);
    e.n = f3(0, 1, 0);
    e.prim = kMiss;
    e.info = K_EMITTER | (LOBE_DIFFUSE << 2) | I_NO_NORMAL | I_POINT | (pi << 8);
    e.wi = f3(0, 0, 0);
  }
  const double pA = sc.emit_area > 0.0f ? 1.0 / (double)sc.emit_area : 1.0;
  e.pL = pA;
  e.pE = 1.0;
  e.alphaL = dscale(le0_of(sc, e), 1.0 / pA);
  e.alphaE = d3(1, 1, 1);
  e.GL = e.GE = 1.0;
  e.fs = d3(0, 0, 0);
  e.pdfL = e.pdfE = 1.0;
  e.fu = e.fv = 0.0f;
  mv_sample(sc, e, e.wi, rng, nullptr);
  A.s.put(0, e);
  A.len = 1;
  return true;
}

// Path::prepend(bool) on an eye subpath (Path.cpp:251-275): an empty one
// starts at a random film point (Shape::getRandomPoint), pA = the film
// bin's area.
__device__ bool prepend_eye(const Scene &sc, Ctx &cx, Sub &A, bool roulette, bool sampleBSDF, Seq &rng) {
  if (A.len > 0) return sample_vertex(sc, cx, A, roulette, true, sampleBSDF, rng);
  rng.next();
  rng.next();
  const float fx = rng.next(), fy = rng.next();
  A.s.put(0, camera_vertex(sc, cx, fx, fy, cx.pA_film, rng));
  A.len = 1;
  return true;
}

// Any hit between two vertices (Scene::intersects, core/Scene.cpp:110-114),
// with the window the bidirectional visibility kernel uses.
__device__ bool occluded(const Scene &sc, Ctx &cx, const MVtx &a, const MVtx &b) {
  float3 w = sub(b.p, a.p);
  const float d = len(w);
  w = mul(w, 1.0f / d);
  const bool sphere_target = b.prim != kMiss && b.prim >= sc.nt && b.prim < sc.nt + sc.ns;
  const float tlimit = sphere_target ? d * (1.0f - 1e-4f) : d - 2.0f * kEps;
  float t;
  uint32_t id;
  ++cx.n_vis;
  return trace<true, false, kStackRows, kBlock>(sc, setup_ray(a.p, w), tlimit, a.prim, cx.stk, t, id, nullptr, nullptr,
                                        nullptr, sphere_target ? kMiss : b.prim);
}

// Path::_computeRadiance (Path.cpp:208-222).
__device__ D3 path_radiance(const Scene &sc, const Store &P, uint32_t n) {
  if (n < 2) return d3(0, 0, 0);
  const MVtx f = P.get(0);
  D3 r = (!is_cam(f) && on_emitter(sc, f.info)) ? le0_of(sc, f) : d3(0, 0, 0);
  for (uint32_t i = 0; i < n; ++i) {
    const MVtx v = i == 0 ? f : P.get(i);
    r = dscale(dmul(r, v.fs), v.GL);
  }
  return r;
}

// Path::append(const Path &) (Path.cpp:348-515): light subpath Y (len s)
// and eye subpath T (len t, camera first) joined into Y, both sides'
// cumulative values completed, the radiance.
__device__ bool join(const Scene &sc, Ctx &cx, Sub &Y, const Sub &T, Seq &rng, D3 &rad) {
  const uint32_t s = Y.len, t = T.len, k = s + t;
  for (uint32_t i = 0; i < t; ++i) Y.s.copy(s + i, T.s, t - 1 - i);
  Y.len = k;
  bool invalid = false;
  if (k > 0) {
    MVtx m0 = Y.s.get(0);
    if (on_em(sc, m0)) {
      m0.info = (m0.info & ~3u) | K_EMITTER;
      Y.s.put(0, m0);
    }
  }
  if (s > 0 && t > 0) {
    MVtx y = Y.s.get(s - 1), z = Y.s.get(s);
    float3 wo = sub(z.p, y.p);
    const float d = len(wo);
    wo = mul(wo, 1.0f / d);
    D3 alpha;
    double pA;
    float3 wi = s > 1 ? normalize(sub(y.p, Y.s.get(s - 2).p)) : mul(wo, -1.0f);
    if (!init_L(sc, y, wi, wo, d, z, false, rng, alpha, pA)) invalid = true;
    wi = t > 1 ? normalize(sub(z.p, Y.s.get(s + 1).p)) : wo;
    if (!init_E(sc, z, wi, mul(wo, -1.0f), d, y, false, rng, alpha, pA)) invalid = true;
    if (!invalid && occluded(sc, cx, y, z)) {
      y.GL = z.GE = 0;
      invalid = true;
    }
    Y.s.put(s - 1, y);
    Y.s.put(s, z);
  }
  for (uint32_t i = s; i--;) {
    MVtx y = Y.s.get(i);
    if (i >= k - 1) {
      if (is_cam(y)) {
        const double pA = cx.pA_film;   // the camera's area is 0
        if (kind_of(y.info) != K_SENSOR) {
          y.info = (y.info & ~3u) | K_SENSOR;
          mv_sample(sc, y, y.wi, rng, nullptr);
        }
        const float3 wi = i > 0 ? normalize(sub(y.p, Y.s.get(i - 1).p)) : y.wi;
        y.alphaE = d3(1.0 / pA, 1.0 / pA, 1.0 / pA);
        y.pE = pA;
        y.fs = f_d(sc, y, wi, mul(wi, -1.0f));
        y.pdfE = pdf_d(sc, y, wi, mul(wi, -1.0f));
      } else {
        y.alphaE = d3(0, 0, 0);
        y.pE = 0;
        y.fs = d3(0, 0, 0);
        y.pdfE = 0;
      }
    } else {
      const MVtx z = Y.s.get(i + 1);
      y.alphaE = (z.pdfE > 0 && z.GE > 0) ? dmul(z.alphaE, dscale(z.fs, 1.0 / z.pdfE)) : d3(0, 0, 0);
      y.pE = z.pE * z.pdfE * z.GE;
    }
    Y.s.put(i, y);
  }
  for (uint32_t i = 0; i < t; ++i) {
    MVtx z = Y.s.get(s + i);
    if (s + i == 0) {
      if (on_em(sc, z)) {
        const double pA = sc.emit_area > 0.0f ? 1.0 / (double)sc.emit_area : 1.0;
        const float3 wo = s + i < k - 1 ? normalize(sub(Y.s.get(s + i + 1).p, z.p)) : mul(z.wi, -1.0f);
        z.alphaL = dscale(le0_of(sc, z), 1.0 / pA);
        z.pL = pA;
        z.fs = f_d(sc, z, f3(0, 0, 0), wo);
        z.pdfL = pdf_d(sc, z, f3(0, 0, 0), wo);
      } else {
        z.alphaL = d3(0, 0, 0);
        z.pL = 0;
        z.fs = d3(0, 0, 0);
        z.pdfL = 0;
      }
    } else {
      const MVtx y = Y.s.get(s + i - 1);
      z.alphaL = (y.pdfL > 0 && y.GL > 0) ? dmul(y.alphaL, dscale(y.fs, 1.0 / y.pdfL)) : d3(0, 0, 0);
      z.pL = y.pL * y.pdfL * y.GL;
    }
    Y.s.put(s + i, z);
  }
  rad = path_radiance(sc, Y.s, k);
  return !invalid;
}

// Path::getContribution (Path.cpp:47-150) of split (s, t) of path P (len n).
__device__ D3 contribution_st(const Scene &sc, Ctx &cx, const Store &P, uint32_t n, uint32_t s, uint32_t t) {
  const uint32_t u = s - 1, v = n - t;
  if (s == 0) {
    const MVtx z = P.get(v);
    if (!has_area(z) || is_cam(z)) return d3(0, 0, 0);
    const float3 wo = v < n - 1 ? normalize(sub(P.get(v + 1).p, z.p)) : mul(z.wi, -1.0f);
    return dmul(le_toward(sc, z, wo), z.alphaE);
  }
  if (t == 0) return d3(0, 0, 0);   // the camera has no area; NullSensor::getWe is black elsewhere
  const MVtx y = P.get(u), z = P.get(v);
  if (spec(sc, y) || spec(sc, z)) return d3(0, 0, 0);
  if (s + t == n) return dmul(dmul(dscale(dmul(y.alphaL, y.fs), y.GL), z.fs), z.alphaE);
  float3 wo = sub(z.p, y.p);
  const float d = len(wo);
  if (d < kEps) return d3(0, 0, 0);
  if (occluded(sc, cx, y, z)) return d3(0, 0, 0);
  wo = mul(wo, 1.0f / d);
  const float3 ywi = u > 0 ? normalize(sub(y.p, P.get(u - 1).p)) : y.wi;
  const float3 wo2 = v < n - 1 ? normalize(sub(P.get(v + 1).p, z.p)) : mul(wo, -1.0f);
  const double G = fabs(ndot_m(y, wo) * ndot_m(z, mul(wo, -1.0f))) / ((double)d * d);
  return dmul(dmul(dscale(dmul(y.alphaL, f_d(sc, y, ywi, wo)), G), f_d(sc, z, wo, wo2)), z.alphaE);
}

// getPd of a full-length split (no visibility test), as the mutations use it.
__device__ double pd_full(const Scene &sc, const Store &P, uint32_t n, uint32_t s) {
  const uint32_t t = n - s;
  if (s == 0) {
    const MVtx z = P.get(0);
    return has_area(z) ? z.pE : 0.0;
  }
  if (t == 0) {
    const MVtx y = P.get(n - 1);
    return has_area(y) ? y.pL : 0.0;
  }
  const MVtx y = P.get(s - 1), z = P.get(s);
  if (spec(sc, y) || spec(sc, z)) return 0.0;
  return y.pL * z.pE;
}

// Camera::getProjection (cameras/ThinLensCamera.cpp:265-270) of a pinhole
// and Viewport::getBin (core/Viewport.h:78-81), in double.
__device__ void project(const CameraDev &cam, float3 p, double &x, double &y) {
  const double dx = (double)p.x - cam.eye[0], dy = (double)p.y - cam.eye[1], dz = (double)p.z - cam.eye[2];
  const double c = -(dx * cam.n[0] + dy * cam.n[1] + dz * cam.n[2]);
  x = ((dx * cam.u[0] + dy * cam.u[1] + dz * cam.u[2]) / (c * cam.tan_w) + 1) * 0.5;
  y = (1 - (dx * cam.v[0] + dy * cam.v[1] + dz * cam.v[2]) / (c * cam.tan_h)) * 0.5;
}
__device__ uint32_t film_bin(const Ctx &cx, double x, double y) {
  const double fc = floor(x * cx.W), fr = floor(y * cx.H);
  const double hc = cx.W - 1.0, hr = cx.H - 1.0;
  const uint32_t col = (uint32_t)(fc >= hc ? hc : (fc <= 0 ? 0.0 : fc));
  const uint32_t row = (uint32_t)(fr >= hr ? hr : (fr <= 0 ? 0.0 : fr));
  return row * cx.W + col;
}
__device__ __forceinline__ float clamp01(float v) { return v >= 1.0f ? 1.0f : (v <= 0.0f ? 0.0f : v); }

// BidirectionalPathTracer::generate (BidirectionalPathTracer.cpp:212-249)
// into L (the light subpath, then the whole path) with T as the eye
// subpath; false once the path would exceed `cap` vertices.
__device__ bool generate(const Scene &sc, Ctx &cx, Sub &L, Sub &T, uint32_t cap, Seq &rng, D3 &rad) {
  T.len = 0;
  L.len = 0;
  for (;;) {
    const bool roulette = T.len >= 2;
    if (!prepend_eye(sc, cx, T, roulette, true, rng)) break;
    if (T.len > cap) return false;
    if (on_em(sc, T.s.get(T.len - 1))) break;
  }
  if (!on_em(sc, T.s.get(T.len - 1))) {
    for (;;) {
      const bool roulette = L.len > 0;
      if (!append_light(sc, cx, L, roulette, true, rng)) break;
      if (L.len + T.len > cap) return false;
    }
  }
  (void)join(sc, cx, L, T, rng, rad);
  return true;
}

// MLTMarkovProcess::_initSample (MLTMarkovProcess.cpp:114-152): a black
// value keeps the sample's previous film bin.
__device__ void init_sample(const Ctx &cx, const Store &P, uint32_t n, D3 rad, double b, uint32_t &bin, D3 &val) {
  if (n < 2 || dzero(rad)) {
    val = d3(0, 0, 0);
    return;
  }
  val = dscale(rad, b / lum(rad));
  double x, y;
  project(cx.cam, P.get(n - 2).p, x, y);
  bin = film_bin(cx, x, y);
}

__device__ double pd1_of(uint32_t n, uint32_t kd) {   // MLTBidirPathMutation::_getPd1 (:440-463)
  if (kd == 0) return 0.25;
  if (kd == 1) return 0.5;
  if (kd < n) return 1.0 / (double)(1ull << ((kd + 1) & 63));
  double p = 0;
  for (uint32_t i = 2; i < n; ++i) p += 1.0 / (double)(1ull << ((i + 1) & 63));
  p = 1.0 - (p + 0.25 + 0.5);
  if (p < 0) return 1.0 / (double)(1ull << ((kd + 1) & 63));
  return p;
}
__device__ double pa1_of(uint32_t kd, uint32_t ka) {   // _getPa1 (:478-490)
  const uint32_t diff = kd > ka ? kd - ka : ka - kd;
  if (diff == 0) return 0.5;
  if (diff == 1) return 0.2;
  return 0.2 * (1.0 / (double)(1ull << (diff & 63)));
}
// _getMutationR (:205-365): T(y|x) over every split of the added subpath.
__device__ double mutation_R(const Scene &sc, const Store &P, uint32_t n, D3 rad, uint32_t ka, double pd,
                             double pa1, int l) {
  if (lum(rad) <= 0) return 0;
  double T = 0;
  for (uint32_t i = 0; i <= ka; ++i) T += (1.0 / (double)(ka + 1)) * pd_full(sc, P, n, (uint32_t)(l + (int)i + 1));
  T *= pa1 * pd;
  return T <= 0 ? 0 : T;
}

// The chain's paths: X (the current one), Y (the mutation) and T (an eye
// subpath being grown).
struct ChainPaths {
  Store X, Y, T;
  uint32_t xlen;
  D3 xrad;
};

// MLTBidirPathMutation::mutate (:29-202).
__device__ double mutate_bidir(const Scene &sc, Ctx &cx, const ChainPaths &C, Sub &Y, D3 &yrad, uint32_t max_len,
                               Seq &rng) {
  const uint32_t n = C.xlen;
  uint32_t kd = 0, ka = 0;
  double pd1 = 1, pa1 = 1;
  {
    const double e1 = rng.next();
    if (e1 < 0.25) {
      kd = 0;
      pd1 = 0.25;
    } else if (e1 < 0.75) {
      kd = 1;
      pd1 = 0.5;
    } else {
      double p = 0.75;
      kd = 1;
      do {
        ++kd;
        pd1 = 1.0 / (double)(1ull << ((kd + 1) & 63));
        p += pd1;
      } while (e1 - p > 0 && kd < n);
      if (kd == n && p < 1) pd1 += 1 - p;
    }
  }
  const int l = (int)floor((double)rng.next() * (double)((int)(n - kd) + 1) - 1.0);
  const int m = l + (int)kd + 1;
  const double pd2 = 1.0 / (double)(n - kd + 1);
  {
    double e2 = rng.next();
    if (e2 < 0.5) {
      ka = kd;
      pa1 = 0.5;
    } else {
      e2 -= 0.5;
      if (e2 < 0.2) {
        ka = kd + 1;
        pa1 = 0.2;
      } else {
        e2 -= 0.2;
        bool found = false;
        if (kd >= 1) {
          if (e2 < 0.2) {
            ka = kd - 1;
            pa1 = 0.2;
            found = true;
          } else {
            e2 -= 0.2;
          }
        }
        if (!found) {
          uint32_t k = 1;
          do {
            ++k;
            ka = kd + k;
            pa1 = 0.2 * (1.0 / (double)(1ull << (k & 63)));
            if (kd >= k) {
              if (e2 < pa1) ka = kd - k;
              e2 -= pa1;
            }
            e2 -= pa1;
          } while (e2 > 0);
        }
      }
    }
  }
  const uint32_t lP = (uint32_t)floor((double)rng.next() * (double)(ka + 1));
  const uint32_t mP = ka - lP;
  if (kd >= n && ka == 0) return 0;
  if (n - kd + ka > max_len) return 0;   // rejected for its length whatever happens below
  for (uint32_t i = 0; i < (uint32_t)(l + 1); ++i) Y.s.copy(i, C.X, i);
  Y.len = (uint32_t)(l + 1);
  Sub T{C.T, n - (uint32_t)m};
  for (uint32_t i = 0; i < T.len; ++i) T.s.copy(i, C.X, n - 1 - i);
  for (uint32_t i = 0; i < lP; ++i)
    if (!append_light(sc, cx, Y, false, true, rng)) return 0;
  for (uint32_t i = 0; i < mP; ++i)
    if (!prepend_eye(sc, cx, T, false, true, rng)) return 0;
  if (!join(sc, cx, Y, T, rng, yrad)) return 0;
  if (kind_of(Y.s.get(0).info) != K_EMITTER || kind_of(Y.s.get(Y.len - 1).info) != K_SENSOR) return 0;
  const double Rxy = mutation_R(sc, Y.s, Y.len, yrad, ka, pd1 * pd2, pa1, l);
  const uint32_t rkd = ka, rka = kd;
  const double rpd = pd1_of(Y.len, rkd) * (1.0 / (double)(Y.len - rkd + 1));
  const double Ryx = mutation_R(sc, C.X, n, C.xrad, rka, rpd, pa1_of(rkd, rka), l);
  const double Fx = lum(C.xrad);
  if (Fx == 0 || Rxy == 0) return 0;
  return ref_min(1.0, lum_ratio(yrad, C.xrad) * (Ryx / Rxy));
}

// One lens vertex behind T's front on X's material at that position,
// sampling with X's event there (MLTLensSubpathMutation.cpp:115-135).
__device__ bool lens_step(const Scene &sc, Ctx &cx, const ChainPaths &C, Sub &T, Seq &rng) {
  const uint32_t n = C.xlen;
  if (!prepend_eye(sc, cx, T, false, false, rng)) return false;
  if (T.len > n) return false;
  const MVtx xv = C.X.get(n - T.len);
  MVtx f = T.s.get(T.len - 1);
  if (material_key(f) != material_key(xv)) return false;
  mv_sample(sc, f, f.wi, rng, &xv);
  T.s.put(T.len - 1, f);
  return true;
}

// The lens subpath mutation's film stratification (MLTLensSubpathMutation.
// cpp:57-103) and the dissolve rover (DissolveSampleGenerator.cpp:111-140).
struct Strat {
  uint16_t *cnt;   // the chain's per-bin counts
  uint32_t max_samples, cur_samples, seq;
};

__device__ void rover_next(Strat &st, uint32_t W, uint32_t H, uint32_t &col, uint32_t &row) {
  const uint32_t size = W * H;
  uint32_t bits = 0;
  for (uint32_t v = size - 1; v; v >>= 1) ++bits;
  const uint32_t masks[32] = {
      0x00000001u, 0x00000003u, 0x00000006u, 0x0000000Cu, 0x00000014u, 0x00000030u, 0x00000060u, 0x000000B8u,
      0x00000110u, 0x00000240u, 0x00000500u, 0x00000CA0u, 0x00001B00u, 0x00003500u, 0x00006000u, 0x0000B400u,
      0x00012000u, 0x00020400u, 0x00072000u, 0x00090000u, 0x00140000u, 0x00300000u, 0x00400000u, 0x00D80000u,
      0x01200000u, 0x03880000u, 0x07200000u, 0x09000000u, 0x14000000u, 0x32800000u, 0x48000000u, 0xA3000000u};
  const uint32_t mask = masks[bits - 1];
  double x, y;
  if (st.seq == 1) {
    x = 0.5 / W;
    y = 0.5 / H;
  } else {
    y = (0.5 + (double)(st.seq / W)) / H;
    x = (0.5 + (double)(st.seq % W)) / W;
  }
  do {
    st.seq = (st.seq >> 1) ^ ((st.seq & 1u) * mask);
  } while (st.seq >= size);
  col = min((uint32_t)(x * W), W - 1);
  row = min((uint32_t)(y * H), H - 1);
}

// MLTLensSubpathMutation::mutate (:35-160).
__device__ double mutate_lens(const Scene &sc, Ctx &cx, const ChainPaths &C, Sub &Y, D3 &yrad, Strat &st,
                              Seq &rng) {
  const uint32_t n = C.xlen;
  // delete the lens subpath: the camera, zero or more specular vertices,
  // one more, repeated while the end is specular
  uint32_t s = n;
  do {
    do {
      --s;
      if (s == 0) return 0;   // X starts on an emitter, so this never triggers
    } while (spec(sc, C.X.get(s - 1)));
    --s;
  } while (s > 0 && spec(sc, C.X.get(s - 1)));
  const uint32_t W = cx.W, H = cx.H, size = W * H;
  uint32_t col = (uint32_t)floor((double)rng.next() * W), row = (uint32_t)floor((double)rng.next() * H);
  for (;;) {
    uint16_t &c = st.cnt[row * W + col];
    if (c < st.max_samples) {
      if ((uint32_t)(++c) == st.max_samples) ++st.cur_samples;
      break;
    } else if (++st.cur_samples >= size) {
      ++c;
      st.max_samples = (uint32_t)c + 1u;
      st.cur_samples = 0;
      break;
    }
    rover_next(st, W, H, col, row);
  }
  const float fx = (float)(((double)col + rng.next()) * (double)cx.invW);
  const float fy = (float)(((double)row + rng.next()) * (double)cx.invH);
  Sub T{C.T, 1};
  T.s.put(0, camera_vertex(sc, cx, fx, fy, cx.pA_film, rng));
  do {
    do {
      if (!lens_step(sc, cx, C, T, rng)) return 0;
    } while (spec(sc, T.s.get(T.len - 1)) && s + T.len < n);
  } while (s + T.len < n);
  for (uint32_t i = 0; i < s; ++i) Y.s.copy(i, C.X, i);
  Y.len = s;
  if (!join(sc, cx, Y, T, rng, yrad)) return 0;
  if (kind_of(Y.s.get(0).info) != K_EMITTER || kind_of(Y.s.get(Y.len - 1).info) != K_SENSOR) return 0;
  const double Fx = lum(C.xrad), Txy = Y.s.get(s).pE;
  if (Txy * Fx == 0) return 0;
  const double Tyx = C.X.get(s).pE;
  return ref_min(1.0, lum_ratio(yrad, C.xrad) * (Tyx / Txy));
}

// _perturbVector (MLTPerturbationPathMutation.cpp:344-355), t1 0.05, t2 0.3.
__device__ float3 perturb(float3 N, Seq &rng) {
  float3 U, V;
  N = normalize(N);
  onb(N, U, V);
  const double phi = (double)rng.next() * (2.0 * M_PI);
  const double r = 0.3 * exp(-log(0.3 / 0.05) * (double)rng.next());
  const float rc = (float)(r * cos(phi)), rs = (float)(r * sin(phi));
  return normalize(add(add(N, mul(U, rc)), mul(V, rs)));
}

// MLTPerturbationPathMutation::mutate (:60-271).
__device__ double mutate_perturb(const Scene &sc, Ctx &cx, const ChainPaths &C, Sub &Y, D3 &yrad, Seq &rng) {
  enum { LENS, MULTI, CAUSTIC };
  const uint32_t n = C.xlen;
  int type;
  uint32_t s = n - 1;   // Y = X without its camera vertex
  if (!spec(sc, C.X.get(s - 1)) && s >= 2 && spec(sc, C.X.get(s - 2))) {
    type = CAUSTIC;
    --s;
    while (spec(sc, C.X.get(s - 1))) {
      --s;
      if (s == 0) return 0;
    }
  } else {
    type = LENS;
    while (spec(sc, C.X.get(s - 1))) {
      --s;
      if (s == 0) return 0;
    }
    --s;
    while (s >= 2 && spec(sc, C.X.get(s - 1))) {
      type = MULTI;
      do {
        --s;
        if (s == 0) return 0;
      } while (spec(sc, C.X.get(s - 1)));
      --s;
    }
    if (type == LENS && s > 0 && spec(sc, C.X.get(s - 1))) return 0;
  }
  for (uint32_t i = 0; i < s; ++i) Y.s.copy(i, C.X, i);
  Y.len = s;
  const double pA = cx.pA_film;
  Sub T{C.T, 0};
  if (type != CAUSTIC) {
    const MVtx xc = C.X.get(n - 1);
    float fx = (float)((double)xc.fu + ((double)rng.next() * 10.0 - 5.0) * (double)cx.invW);
    float fy = (float)((double)xc.fv + ((double)rng.next() * 10.0 - 5.0) * (double)cx.invH);
    fx = clamp01(fx);
    fy = clamp01(fy);
    T.s.put(0, camera_vertex(sc, cx, fx, fy, pA, rng));
    T.len = 1;
    do {
      if (!lens_step(sc, cx, C, T, rng)) return 0;
    } while (spec(sc, T.s.get(T.len - 1)));
    if (type == MULTI) {
      do {
        const uint32_t t = n - T.len;
        if (t == 0 || Y.len + T.len >= n) break;
        const float3 wo = perturb(normalize(sub(C.X.get(t - 1).p, C.X.get(t).p)), rng);
        MVtx f = T.s.get(T.len - 1);
        f.ev = wo;
        T.s.put(T.len - 1, f);
        do {
          if (!lens_step(sc, cx, C, T, rng)) return 0;
        } while (spec(sc, T.s.get(T.len - 1)));
      } while (T.len + Y.len < n);
    }
  } else {
    const float3 wo = perturb(normalize(sub(C.X.get(s).p, C.X.get(s - 1).p)), rng);
    MVtx b = Y.s.get(s - 1);
    b.ev = wo;
    Y.s.put(s - 1, b);
    for (;;) {
      if (!append_light(sc, cx, Y, false, false, rng)) return 0;
      const MVtx xb = C.X.get(Y.len - 1);
      MVtx yb = Y.s.get(Y.len - 1);
      if (material_key(yb) != material_key(xb)) return 0;
      if (Y.len >= n - 1) break;
      mv_sample(sc, yb, yb.wi, rng, &xb);
      Y.s.put(Y.len - 1, yb);
    }
    double x, y;
    project(cx.cam, Y.s.get(Y.len - 1).p, x, y);
    const float fx = clamp01((float)x), fy = clamp01((float)y);
    T.s.put(0, camera_vertex(sc, cx, fx, fy, pA, rng));
    T.len = 1;
  }
  if (!join(sc, cx, Y, T, rng, yrad)) return 0;
  if (kind_of(Y.s.get(0).info) != K_EMITTER || kind_of(Y.s.get(Y.len - 1).info) != K_SENSOR) return 0;
  const double Txy = type == CAUSTIC ? Y.s.get(Y.len - 1).pL : Y.s.get(s).pE;
  const double Tyx = type == CAUSTIC ? C.X.get(n - 1).pL : C.X.get(s).pE;
  const double Fx = lum(C.xrad);
  if (Txy * Fx == 0) return 0;
  return ref_min(1.0, lum_ratio(yrad, C.xrad) * (Tyx / Txy));
}

// ---------------------------------------------------------------------------
// Job-level state and kernels.
// ---------------------------------------------------------------------------

// Per-job constants of the MLT kernels.
struct MltDesc {
  CameraDev cam;
  uint32_t W, H, seed, job_id;
  uint32_t chains, mutations, seed_paths, max_len, max_rejections, filter_proposed;
  uint32_t n_mut;            // enabled mutations
  uint32_t mut_id[3];
  double mut_w[3];           // normalised weights of the enabled mutations
  uint32_t round;            // seed round being generated
  uint32_t has_seeds;        // a round found light
  uint32_t n_entries;        // seed entries of this round
  uint32_t n_prev;           // entries of earlier rounds (all of weight 0)
  uint32_t first_pos, last_pos;   // first and last entry of positive weight
  double b, sum;
};

// Per chain (MLTMarkovProcess's state and the lens mutation's).
struct ChainState {
  uint32_t xbuf, xlen, cur_bin, tent_bin;
  uint32_t visits, smax, scur, seq;
  D3 cur_val, xrad;
};

struct MltBufs {
  // seed paths: vertex records (runs of 64 paths), length and attempt per
  // path; per entry (path-major, kMaxEntries per path): weight, stream
  // position of its join, (k, s)
  float4 *seed_v;
  float4 *seed_tmp;       // the seed lanes' eye subpath
  float4 *seed_tmp2;      // the seed lanes' split being joined
  uint32_t *seed_len, *seed_att, *seed_cnt;
  double *ent_w;
  uint32_t *ent_pos;
  uint32_t *ent_ks;
  // the normalised running sum in the reference's entry order, and each
  // order position's (path, entry)
  double *cdf;
  uint32_t *order;
  // chains: three path buffers (runs of 64 chains each), states, stratification
  float4 *chain_v;
  ChainState *st;
  uint16_t *strat;
  // film: fixed-point sums per bin and channel, proposal counts, the count
  // of samples added, counters (accepted, extension rays, visibility rays)
  unsigned long long *num;
  uint32_t *proposed;
  unsigned long long *counters;   // [0] samples, [1] accepted, [2] extension, [3] visibility
};

__device__ __forceinline__ void flush_counters(unsigned long long *ctr, uint32_t a, uint32_t b, uint32_t c,
                                               uint32_t d) {
  if (a) atomicAdd(ctr + 0, (unsigned long long)a);
  if (b) atomicAdd(ctr + 1, (unsigned long long)b);
  if (c) atomicAdd(ctr + 2, (unsigned long long)c);
  if (d) atomicAdd(ctr + 3, (unsigned long long)d);
}

// One seed path per lane (MLTRenderer::_initSeedPaths, MLTRenderer.cpp:
// 128-190): generate until the path has at most max_len vertices, then every
// split (k, s) joined in the reference's order, weighted by the luminance
// of the original path's contribution of that split.
__global__ __launch_bounds__(kBlock) void k_mlt_seed(Scene sc, const MltDesc *__restrict__ dd, MltBufs B) {
  __shared__ uint2 stack[kStackRows * kBlock];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const MltDesc &d = *dd;
  const uint32_t N = d.seed_paths;
  if (i >= N) return;
  Ctx cx;
  cx.cam = d.cam;
  cx.W = d.W;
  cx.H = d.H;
  cx.invW = 1.0f / d.W;
  cx.invH = 1.0f / d.H;
  cx.pA_film = (1.0 / d.W) * (1.0 / d.H);
  cx.stk = stack + threadIdx.x;
  cx.n_ext = cx.n_vis = 0;
  Sub G{lane_store(B.seed_v, i), 0};
  Sub T{lane_store(B.seed_tmp, i), 0};
  const uint32_t idx = d.round * N + i;
  uint32_t cnt = 0;
  for (uint32_t a = 0; a < kAttempts; ++a) {
    Seq rng(d.seed, d.job_id, idx, a, kSeedKind);
    D3 rad;
    if (!generate(sc, cx, G, T, d.max_len, rng, rad)) continue;
    if (G.len > d.max_len) continue;
    B.seed_att[i] = a;
    const uint32_t L = G.len;
    for (uint32_t k = 2; k <= L; ++k) {
      for (uint32_t s = 0; s <= k; ++s) {
        const uint32_t t = k - s;
        const uint32_t e = i * kMaxEntries + cnt;
        B.ent_pos[e] = rng.n;
        B.ent_ks[e] = (k << 8) | s;
        Sub P{lane_store(B.seed_tmp2, i), s};
        for (uint32_t j = 0; j < s; ++j) P.s.copy(j, G.s, j);
        Sub R{T.s, t};
        for (uint32_t j = 0; j < t; ++j) R.s.copy(j, G.s, L - 1 - j);
        D3 prad;
        const bool v = join(sc, cx, P, R, rng, prad);
        B.ent_w[e] = v ? lum(contribution_st(sc, cx, G.s, L, s, t)) : 0.0;
        ++cnt;
      }
    }
    break;
  }
  B.seed_len[i] = G.len;
  B.seed_cnt[i] = cnt;
  flush_counters(B.counters, 0, 0, cx.n_ext, cx.n_vis);
}

// The seed entries in the reference's order (the entries of path N - 1
// first, MLTRenderer.cpp:150) compacted into cdf / order, their sum in that
// order taken backwards as the reference does (MLTRenderer.cpp:202-205),
// b = sum / entries, and the normalised running sum Random::
// sampleNormalizedCDF walks.  One block: the sums are sequential (so they
// round as the oracle's), one lane adding what the block stages in LDS.
constexpr uint32_t kWeightsBlock = 256, kWeightsChunk = 2048;
__global__ __launch_bounds__(kWeightsBlock) void k_mlt_weights(MltDesc *dd, MltBufs B) {
  __shared__ uint32_t scan[kWeightsBlock];
  __shared__ double buf[kWeightsChunk];
  __shared__ double sum_sh;
  MltDesc &d = *dd;
  const uint32_t N = d.seed_paths, t = threadIdx.x;
  uint32_t base = 0;
  for (uint32_t top = 0; top < N; top += kWeightsBlock) {
    const uint32_t k = top + t;   // the k-th path in the reference's order
    const uint32_t i = k < N ? N - 1 - k : 0;
    const uint32_t c = k < N ? B.seed_cnt[i] : 0u;
    scan[t] = c;
    __syncthreads();
    for (uint32_t o = 1; o < kWeightsBlock; o <<= 1) {
      const uint32_t v = t >= o ? scan[t - o] : 0u;
      __syncthreads();
      scan[t] += v;
      __syncthreads();
    }
    const uint32_t off = base + scan[t] - c;
    for (uint32_t j = 0; j < c; ++j) {
      B.cdf[off + j] = B.ent_w[i * kMaxEntries + j];
      B.order[off + j] = i * kMaxEntries + j;
    }
    base += scan[kWeightsBlock - 1];
    __syncthreads();
  }
  const uint32_t n = base;
  double sum = 0;
  for (uint32_t hi = n; hi > 0;) {
    const uint32_t lo = hi > kWeightsChunk ? hi - kWeightsChunk : 0u;
    for (uint32_t g = lo + t; g < hi; g += kWeightsBlock) buf[g - lo] = B.cdf[g];
    __syncthreads();
    if (t == 0) {
      // the same adds in the same order, the LDS reads of eight issued
      // before them
      uint32_t g = hi - lo;
      for (; g >= 8; g -= 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = buf[g - 1 - q];
#pragma unroll
        for (int q = 0; q < 8; ++q) sum += v[q];
      }
      while (g > 0) sum += buf[--g];
    }
    __syncthreads();
    hi = lo;
  }
  if (t == 0) sum_sh = sum;
  __syncthreads();
  sum = sum_sh;
  const bool found = n > 0 && sum > 0;
  if (t == 0) {
    d.n_entries = n;
    d.sum = sum;
    d.has_seeds = found ? 1u : 0u;
    if (found) d.b = sum / (double)(n + d.n_prev);
  }
  if (!found) return;
  // the first and last entries of positive weight, found by every thread
  // (scan and first_last share no storage with the sums)
  uint32_t *first_last = scan;
  if (t == 0) {
    first_last[0] = ~0u;
    first_last[1] = 0;
  }
  __syncthreads();
  double acc = 0;
  for (uint32_t lo = 0; lo < n; lo += kWeightsChunk) {
    const uint32_t hi = min(n, lo + kWeightsChunk);
    for (uint32_t g = lo + t; g < hi; g += kWeightsBlock) {
      const double w = B.cdf[g] / sum;
      buf[g - lo] = w;
      if (w > 0) {
        atomicMin(first_last, g);
        atomicMax(first_last + 1, g);
      }
    }
    __syncthreads();
    if (t == 0) {
      uint32_t g = 0;
      for (; g + 8 <= hi - lo; g += 8) {
        double w[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) w[q] = buf[g + q];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          acc += w[q];
          buf[g + q] = acc;
        }
      }
      for (; g < hi - lo; ++g) {
        acc += buf[g];
        buf[g] = acc;
      }
    }
    __syncthreads();
    for (uint32_t g = lo + t; g < hi; g += kWeightsBlock) B.cdf[g] = buf[g - lo];
    __syncthreads();
  }
  if (t == 0) {
    d.first_pos = first_last[0];
    d.last_pos = first_last[1];
  }
}

__device__ __forceinline__ Ctx chain_ctx(const MltDesc &d, uint2 *stk) {
  Ctx cx;
  cx.cam = d.cam;
  cx.W = d.W;
  cx.H = d.H;
  cx.invW = 1.0f / d.W;
  cx.invH = 1.0f / d.H;
  cx.pA_film = (1.0 / d.W) * (1.0 / d.H);
  cx.stk = stk;
  cx.n_ext = cx.n_vis = 0;
  return cx;
}

__device__ __forceinline__ Store chain_store(const MltBufs &B, const MltDesc &d, uint32_t which, uint32_t c) {
  return Store{B.chain_v + ((uint64_t)which * wave_runs(d.chains) + c / kWave) * kRun + c % kWave, kWave};
}

// A chain's start (MLTRenderer.cpp:100-112): its seed entry by the weights
// (an entry of weight 0 is never taken), the split rebuilt from its seed
// path with the join's own random numbers, the first sample.
__global__ __launch_bounds__(kBlock) void k_mlt_start(Scene sc, const MltDesc *__restrict__ dd, MltBufs B) {
  __shared__ uint2 stack[kStackRows * kBlock];
  const MltDesc &d = *dd;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.chains) return;
  Ctx cx = chain_ctx(d, stack + threadIdx.x);
  Seq pick(d.seed, d.job_id, c, 0, kInitKind);
  const double x = pick.next();
  uint32_t lo = 0, hi = d.n_entries;   // first g with cdf[g] >= x
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (B.cdf[mid] >= x) hi = mid;
    else lo = mid + 1;
  }
  uint32_t g = lo;
  if (g >= d.n_entries) g = d.last_pos;
  else if (B.ent_w[B.order[g]] <= 0) g = d.first_pos;
  const uint32_t e = B.order[g];
  const uint32_t i = e / kMaxEntries;
  const uint32_t ks = B.ent_ks[e], k = ks >> 8, s = ks & 255u, t = k - s;
  const uint32_t N = d.seed_paths;
  const Store G = lane_store(B.seed_v, i);
  const uint32_t L = B.seed_len[i];
  Sub X{chain_store(B, d, 0, c), s};
  for (uint32_t j = 0; j < s; ++j) X.s.copy(j, G, j);
  Sub R{chain_store(B, d, 2, c), t};
  for (uint32_t j = 0; j < t; ++j) R.s.copy(j, G, L - 1 - j);
  Seq rng(d.seed, d.job_id, d.round * N + i, B.seed_att[i], kSeedKind, B.ent_pos[e]);
  D3 rad;
  (void)join(sc, cx, X, R, rng, rad);
  ChainState st;
  st.xbuf = 0;
  st.xlen = X.len;
  st.cur_bin = st.tent_bin = 0;
  st.visits = 0;
  st.smax = 1;
  st.scur = 0;
  st.seq = 1;
  D3 val;
  uint32_t bin = 0;
  init_sample(cx, X.s, X.len, rad, d.b, bin, val);
  st.cur_bin = bin;
  st.cur_val = val;
  st.xrad = rad;
  B.st[c] = st;
  flush_counters(B.counters, 0, 0, cx.n_ext, cx.n_vis);
}

__device__ __forceinline__ void add_sample(const MltBufs &B, double b, uint32_t bin, D3 val, double w) {
  // value / b * w in units of 2^-24 (values are at most 1 / 0.0722 times b)
  const double s = w / b * kFix;
  unsigned long long *o = B.num + 3ull * bin;
  const double q[3] = {val.x * s, val.y * s, val.z * s};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (q[k] > 0) atomicAdd(o + k, (unsigned long long)llrint(q[k]));
}

// Mutations [j0, j1) of each chain (MLTMarkovProcess::run, :38-112).  rec
// (optional): per chain and mutation, the parity record (mutation, alpha,
// accepted, Y length, tentative bin, X length after).  The step waits on
// dependent path-record loads and kd walks most of its cycles; held to 256
// registers (about 60 spilled), two waves share each SIMD to hide that, which
// pays once a job has two waves of chains per SIMD (131,072 on 256 CUs).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_mlt_step(Scene sc, const MltDesc *__restrict__ dd, MltBufs B, uint32_t j0, uint32_t j1, double *rec,
                uint32_t rec_steps) {
  __shared__ uint2 stack[kStackRows * kBlock];
  const MltDesc &d = *dd;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.chains) return;
  Ctx cx = chain_ctx(d, stack + threadIdx.x);
  ChainState st = B.st[c];
  Strat sr{B.strat + (uint64_t)c * d.W * d.H, st.smax, st.scur, st.seq};
  uint32_t n_samples = 0, n_acc = 0;
  for (uint32_t j = j0; j < j1; ++j) {
    ChainPaths C;
    C.X = chain_store(B, d, st.xbuf, c);
    C.Y = chain_store(B, d, 1 - st.xbuf, c);
    C.T = chain_store(B, d, 2, c);
    C.xlen = st.xlen;
    C.xrad = st.xrad;
    Seq rng(d.seed, d.job_id, c, j, kStepKind);
    const double x = rng.next();
    uint32_t which = d.mut_id[d.n_mut - 1];
    {
      double sum = 0;
      for (uint32_t i = 0; i + 1 < d.n_mut; ++i) {
        sum += d.mut_w[i];
        if (x <= sum) {
          which = d.mut_id[i];
          break;
        }
      }
    }
    Sub Y{C.Y, 0};
    D3 yrad = d3(0, 0, 0);
    double alpha;
    if (which == 0) alpha = mutate_bidir(sc, cx, C, Y, yrad, d.max_len, rng);
    else if (which == 1) alpha = mutate_lens(sc, cx, C, Y, yrad, sr, rng);
    else alpha = mutate_perturb(sc, cx, C, Y, yrad, rng);
    alpha = ref_min(1.0, alpha);
    if (alpha == 0 || Y.len < 2 || Y.len > d.max_len) {
      Y.len = 0;
      yrad = d3(0, 0, 0);
      alpha = 0;
    }
    D3 tval;
    init_sample(cx, Y.s, Y.len, yrad, d.b, st.tent_bin, tval);
    ++st.visits;
    if (d.max_rejections == 0 || st.visits < d.max_rejections) {
      if (1 - alpha > 0) {
        add_sample(B, d.b, st.cur_bin, st.cur_val, 1 - alpha);
        ++n_samples;
      }
      if (alpha > 0) {
        atomicAdd(B.proposed + st.tent_bin, 1u);
        add_sample(B, d.b, st.tent_bin, tval, alpha);
        ++n_samples;
      }
    }
    const bool acc = (double)rng.next() < alpha;
    if (acc) {
      st.visits = 0;
      st.xbuf = 1 - st.xbuf;
      st.xlen = Y.len;
      st.xrad = yrad;
      st.cur_bin = st.tent_bin;
      st.cur_val = tval;
      ++n_acc;
    }
    if (rec && j < rec_steps) {
      double *o = rec + 6 * ((uint64_t)c * rec_steps + j);
      o[0] = which;
      o[1] = alpha;
      o[2] = acc ? 1.0 : 0.0;
      o[3] = Y.len;
      o[4] = st.tent_bin;
      o[5] = st.xlen;
    }
  }
  st.smax = sr.max_samples;
  st.scur = sr.cur_samples;
  st.seq = sr.seq;
  B.st[c] = st;
  flush_counters(B.counters, n_samples, n_acc, cx.n_ext, cx.n_vis);
}

// RenderOutput::getFinalizedOutput's MLT branch (RenderOutput.cpp:153-216):
// per bin, the numerator times the proposal scale (the 3 x 3 median of the
// proposal counts over this bin's count, clamped at the film's edge) times
// 2 / (samples added).
__global__ void k_mlt_finish(const MltDesc *__restrict__ dd, MltBufs B, float4 *__restrict__ film) {
  const MltDesc &d = *dd;
  const uint32_t o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= d.W * d.H) return;
  const uint32_t i = o / d.W, j = o % d.W;
  const unsigned long long ns = B.counters[0];
  const double inv = ns ? 2.0 / (double)ns : 0.0;
  double scale = 1;
  if (d.filter_proposed) {
    uint32_t t[9];
    int n = 0;
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        const int r = min(max((int)i + y, 0), (int)d.H - 1);
        const int cc = min(max((int)j + x, 0), (int)d.W - 1);
        t[n++] = B.proposed[(uint32_t)r * d.W + (uint32_t)cc];
      }
    for (int a = 1; a < 9; ++a) {   // insertion sort of nine
      const uint32_t v = t[a];
      int b = a - 1;
      while (b >= 0 && t[b] > v) {
        t[b + 1] = t[b];
        --b;
      }
      t[b + 1] = v;
    }
    const double mid = (double)t[4];
    const uint32_t prop = B.proposed[o];
    scale = mid > 0 && prop > 0 ? mid / (double)prop : 1.0;
  }
  const double f = d.b / kFix * scale * inv;
  const unsigned long long *q = B.num + 3ull * o;
  film[o] = make_float4((float)((double)q[0] * f), (float)((double)q[1] * f), (float)((double)q[2] * f), 1.0f);
}

}  // namespace mlt

#endif  // MPT_MLT_H_
