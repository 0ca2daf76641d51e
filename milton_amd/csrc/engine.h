// Internal interface of the wavefront engine shared by the C ABI (mpt.hip)
// and the batching / serving layer (serve.cpp).
//
// A batch is a set of path slots filled from one scene and any number of
// jobs.  Its contents live in a device-side descriptor (job table + chunk
// table); every kernel of a batch reads its sizes from that descriptor and
// runs on a grid sized by the batch capacity, so the kernel sequence of a
// batch is identical for every batch of a (scene, depth) pair and is captured
// once into a hipGraph and replayed.
#ifndef MPT_ENGINE_H_
#define MPT_ENGINE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/milton_pt.h"
#include "device.h"

namespace mpt_engine {

constexpr uint32_t kMaxJobs = 256;      // jobs per batch
constexpr uint32_t kMaxDirect = 64;     // light samples per light per vertex
constexpr uint32_t kMaxChunks = 1024;   // pixel ranges per batch
constexpr uint32_t kJobBits = 20;       // sample index bits in a path's (job, sample) word
// Shadow rays a batch may hold (32 B each, 8 GiB per lane): a batch of a job
// with many light samples per vertex gets fewer paths.
constexpr uint64_t kShadowRayBudget = 1ull << 28;
constexpr uint32_t kMinBatchPaths = 1024;

struct CameraDev {
  float eye[3], u[3], v[3], n[3];
  float tan_w, tan_h;
  float aperture, focal_length, focal_distance;   // thin lens; aperture 0 = pinhole
};

struct JobDev {
  CameraDev cam;
  uint32_t width, height, spp, max_depth;
  uint32_t seed, job_id, side, direct;   // direct: light samples per light (>= 1)
  float4 *film;   // W*H (sum r, g, b, samples) or, filtered, (numerator, denominator)
  uint32_t filter;                       // 0: pixel bin; else 1 + MPT_FILTER_*
  uint32_t filter_support;
  float filter_p0, filter_p1;
  uint32_t direct_off;   // efficientDirect = false: emission at every vertex, no NEE
  uint32_t light_select; // 1: each light sample draws its light by emitted power
  uint32_t renderer;     // MPT_RENDERER_PATH or MPT_RENDERER_BIDIR
  uint32_t pad;
};

struct ChunkDev {
  uint32_t job;     // index into the batch job table
  uint32_t pix0;    // first pixel of the chunk in the job's film
  uint32_t npix;
  uint32_t slot0;   // first path slot of the chunk in the batch
  uint32_t pixoff;  // first batch pixel of the chunk
  // filtered jobs: the film pixels the chunk's samples reach (whole rows
  // from trow0) are targets [tgt0, tgt0 + ntgt) of the batch
  uint32_t tgt0, trow0, ntgt;
};

struct BatchHdr {
  uint32_t n_chunks, n_paths, n_pixels, max_depth;
  uint32_t max_direct;   // most light samples per light of the batch's jobs
  uint32_t n_targets;    // film pixels of filtered chunks (ChunkDev::tgt0)
  uint32_t pad[2];
};

struct BatchDesc {
  BatchHdr hdr;
  JobDev jobs[kMaxJobs];
  ChunkDev chunks[kMaxChunks];
};

template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t reserve(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    const hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    else p = nullptr;
    return e;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Metropolis light transport's device buffers (the kernels' view is
// mlt::MltBufs), kept across renders and grown as jobs need.
struct MltWork {
  DBuf<float4> seed_v, seed_tmp, seed_tmp2, chain_v;
  DBuf<uint32_t> seed_len, seed_att, seed_cnt, ent_pos, ent_ks, order, proposed;
  DBuf<double> ent_w, cdf, rec;
  DBuf<unsigned char> desc, states;
  DBuf<uint16_t> strat;
  DBuf<unsigned long long> num, counters;
  uint64_t bytes() const {
    return seed_v.bytes() + seed_tmp.bytes() + seed_tmp2.bytes() + chain_v.bytes() + seed_len.bytes() +
           seed_att.bytes() + seed_cnt.bytes() + ent_pos.bytes() + ent_ks.bytes() + order.bytes() +
           proposed.bytes() + ent_w.bytes() + cdf.bytes() + rec.bytes() + desc.bytes() + states.bytes() +
           strat.bytes() + num.bytes() + counters.bytes();
  }
};

// Path and queue buffers, sized by the batch capacity.
struct Workspace {
  uint32_t cap = 0, depth = 0;
  uint64_t shadow_entries = 0;   // shadow rays the shadow buffer holds
  uint64_t generation = 0;   // bumped on every reallocation (invalidates graphs)
  DBuf<float4> q[2];
  DBuf<float2> hits;         // (FP32 t, leaf record of the closest hit or MPT_MISS)
  // NEE: one group per shading vertex with light samples, (origin, slot)
  // (origin primitive, ray count, 0, 0); its rays sit at sr[slot * spv ...],
  // (direction, distance limit) (contribution, 0) each, spv = shadow rays per
  // vertex of the batch
  DBuf<float4> sg;
  DBuf<float4> sr;
  DBuf<float4> Ls;
  DBuf<uint32_t> flags;
  DBuf<uint2> pjs;           // (pixel, job << 20 | sample) per slot
  DBuf<float2> pos;          // film position of each camera sample, in pixels (filtered jobs)
  DBuf<uint32_t> counts;     // [0, D] queue sizes, [D+1, 2D] shadow group counts, [2D+1, 3D] claim cursors
  DBuf<unsigned long long> totals;
  DBuf<BatchDesc> desc;
  // Bidirectional batches: per slot its subpaths' vertex records (V per slot,
  // 5 float4 each), (eye vertices, light vertices, flags), the visibility
  // pairs it queues and its occlusion mask.
  DBuf<float4> bverts;
  DBuf<uint4> binfo;
  DBuf<uint2> bpairs;
  DBuf<uint32_t> bmasks;
  DBuf<uint32_t> border;     // slot indices ordered by path length for the combine kernel
  size_t bytes() const {
    return q[0].bytes() + q[1].bytes() + hits.bytes() + sg.bytes() + sr.bytes() + Ls.bytes() + flags.bytes() +
           pjs.bytes() + pos.bytes() + counts.bytes() + totals.bytes() + desc.bytes() + bverts.bytes() +
           binfo.bytes() + bpairs.bytes() + bmasks.bytes() + border.bytes();
  }
};

// An engine lane: a stream with its own workspace and descriptor staging
// ring.  Batches on different lanes are independent (disjoint film pixels,
// separate path state), so a context alternates its batches over its lanes:
// the tail of one batch's launches (the last waves of a kernel) overlaps the
// bulk of another's.
struct Lane {
  hipStream_t stream = nullptr;
  Workspace ws;
  BatchDesc *staging = nullptr;   // pinned ring of descriptors
  hipEvent_t staging_ev[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t staging_next = 0;
  hipEvent_t mark = nullptr;      // fork / join point
};

constexpr uint32_t kMaxLanes = 4;

// One captured batch sequence.
struct GraphEntry {
  const void *scene = nullptr;
  uint32_t lane = 0;
  uint32_t depth = 0, cap = 0, lights = 0;
  bool multi = false;      // several light samples per light
  bool filtered = false;   // the batch splats filtered samples
  bool bidir = false;      // bidirectional path tracing
  uint64_t generation = 0;
  hipGraphExec_t exec = nullptr;
};

// A batch as planned on the host.
struct BatchPlan {
  std::vector<JobDev> jobs;
  std::vector<ChunkDev> chunks;
  uint32_t n_paths = 0, n_pixels = 0, max_depth = 0;
  uint32_t max_direct = 1;   // most light samples per light of its jobs
  bool filtered = false;     // some job reconstructs its film through a filter
  bool bidir = false;        // every job renders with MPT_RENDERER_BIDIR
  uint32_t capacity = 0;     // paths the workspace is sized for (0: the engine's batch capacity)
};

}  // namespace mpt_engine

struct mpt_server;

struct mpt_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int num_cus = 256;
  std::mutex err_mu;                 // err is written by a server's worker thread too
  std::string err;
  mpt_engine::Lane lanes[mpt_engine::kMaxLanes];
  uint32_t num_lanes = 3;            // lanes batches alternate over
  uint32_t next_lane = 0;
  hipEvent_t mark = nullptr;         // fork point on the context stream
  uint32_t batch_paths = 1u << 25;   // paths per batch (queue depth)
  bool timing = false;
  bool counters = false;
  bool graphs = true;
  uint32_t stack = 8;                // short-stack entries of the traversal kernels
  bool tree_top = false;             // persistent closest hit reads the tree's top from LDS
  uint32_t trace_block = 128;        // threads per block of the persistent closest-hit kernel (128 or 256)
  uint32_t shade_block = 512;        // threads per block of the shading kernel (256 or 512)
  uint32_t persistent = 1;           // closest-hit kernel form: 0 grid-stride, 1 persistent waves with
                                     // lane refill, 2 one ray per lane over a capacity-sized grid
  std::atomic<uint32_t> fail_batch{0};   // MPT_OPT_FAIL_BATCH countdown (0: off); a server's worker reads it
  std::vector<hipEvent_t> events;    // timing pool, pairs per launch
  std::vector<std::pair<int, size_t>> timed;   // instrumented launches: (kind, first event) since the render began
  std::vector<mpt_scene *> scenes;   // live scenes, films and servers, released by mpt_close
  std::vector<mpt_film *> films;
  std::vector<mpt_server *> servers;
  std::vector<mpt_engine::GraphEntry> graphs_cache;
  std::vector<std::pair<const void *, uint32_t>> grid_cache;   // resident blocks per kernel
  mpt_engine::MltWork mlt;           // Metropolis light transport buffers
};

struct mpt_scene {
  mpt_context *ctx = nullptr;
  uint64_t uid = 0;   // unique per process, keys the graph cache
  mpt_engine::DBuf<uint2> nodes;
  mpt_engine::DBuf<float4> recs, tshade, spheres, quads;
  mpt_engine::DBuf<uint32_t> sphere_mat;
  mpt_engine::DBuf<mpt_dev::LightDev> lights;
  mpt_engine::DBuf<float4> light_tris;
  mpt_engine::DBuf<float2> light_sel;
  mpt_engine::DBuf<mpt_material> mats;
  mpt_engine::DBuf<float4> env_tex, points;
  mpt_engine::DBuf<double> env_kernel;
  mpt_dev::Scene dev{};
  uint64_t stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t leaf_size = 0;   // kd leaf size the tree was built with
  uint32_t sphere_lights = 0;   // emitting spheres among the area lights
  uint64_t device_bytes() const {
    return nodes.bytes() + recs.bytes() + tshade.bytes() + spheres.bytes() + quads.bytes() + sphere_mat.bytes() +
           lights.bytes() + light_tris.bytes() + light_sel.bytes() + mats.bytes() + env_tex.bytes() + env_kernel.bytes();
  }
};

struct mpt_film {
  mpt_context *ctx = nullptr;
  uint32_t width = 0, height = 0;
  mpt_engine::DBuf<float4> data;
  mpt_engine::DBuf<uchar4> ldr;     // tone-mapped image
  mpt_engine::DBuf<double2> tm;     // tone-map partials, then (scale, white)
};

namespace mpt_engine {

int fail(mpt_context *ctx, int code, const std::string &msg);
int check_job(mpt_context *ctx, const mpt_job *job);
JobDev make_job(const mpt_job &job, float4 *film);
// A thin-lens job with autofocus gets its focal distance from the scene: the
// hit distance of the ray through its focal point.
int resolve_focus(mpt_scene *s, mpt_job &job);

// Creates lane k's stream, events and staging ring if it has none yet.
int ensure_lane(mpt_context *ctx, uint32_t k);
// Makes sure lane k's workspace holds a batch of `cap` paths, `depth`
// bounces and `spv` shadow rays per vertex.
int ensure_workspace(mpt_context *ctx, uint32_t k, uint32_t cap, uint32_t depth, uint32_t spv);
// Paths per batch for a scene and the largest light-sample count of the
// batch's jobs: the context's batch size, lowered so that the shadow rays
// of one batch (paths x shadow rays per vertex) fit kShadowRayBudget.
uint32_t batch_capacity(const mpt_context *ctx, const mpt_scene *s, uint32_t max_direct);
// Device bytes of one lane's workspace for batches of `cap` paths, `depth`
// bounces and `shadow` shadow rays (what ensure_workspace allocates).
uint64_t workspace_bytes(uint64_t cap, uint32_t depth, uint64_t shadow);
// Bidirectional batches: the device bytes per path slot of their subpath
// vertices, visibility pairs and masks for subpaths of `depth` bounces, and
// the paths per batch that fit `lane_bytes` of them (at most the context's
// batch size, at least kMinBatchPaths).
constexpr uint64_t kBidirLaneBytes = 4ull << 30;
uint64_t bidir_slot_bytes(uint32_t depth);
uint32_t bidir_capacity(const mpt_context *ctx, uint32_t depth, uint64_t lane_bytes);
// Deepest subpaths (bounces) of a bidirectional job.
uint32_t bdpt_max_depth();
// Scene-dependent conditions of a job: a bidirectional job needs a light
// (quad, mesh or point) and no emitting sphere.
int check_scene_job(mpt_context *ctx, const mpt_scene *s, const mpt_job *job);

// Metropolis light transport of a whole job into film (W*H, finalised
// values with weight 1), or, with rec, its first rec_steps mutations of
// every chain recorded (mpt_mlt_trace); b_out / entries_out optional.
int render_mlt(mpt_scene *s, const mpt_job *job, float4 *film, mpt_render_stats *stats, double *rec,
               uint32_t rec_steps, double *b_out, uint64_t *entries_out);

// Enqueues one batch on lane `lane` (asynchronous).  With L_host the
// per-path radiance of the batch is copied to L_host (slot order) and the
// call synchronises.  Uses a captured graph when the context allows it and
// no per-launch instrumentation is on.
int run_batch(mpt_context *ctx, mpt_scene *scene, const BatchPlan &plan, float *L_host, uint32_t lane);
// The lane the next batch goes to (round robin over num_lanes).
uint32_t pick_lane(mpt_context *ctx);
// Orders every lane after the work queued so far on the context stream, and
// the context stream after the work queued so far on every lane.
int fork_lanes(mpt_context *ctx);
int join_lanes(mpt_context *ctx);
// Orders lane k after the work queued so far on the context stream.
int lane_follow(mpt_context *ctx, uint32_t k);

// Zeroes the device ray totals of every lane (on the context stream, before
// a fork); read_totals syncs and returns their sums:
// [camera, extension, shadow, nodes, refs].
int reset_totals(mpt_context *ctx);
// Renders pixels [pix_begin, pix_end) of one job into film (accumulating) or
// per-path radiance into L_host; stats optional.
int render_range(mpt_scene *s, const mpt_job *job, uint32_t pix_begin, uint32_t pix_end, float4 *film,
                 float *L_host, mpt_render_stats *stats);
// Drops captured graphs of one scene uid (0: all).
void drop_graphs(mpt_context *ctx, uint64_t scene_uid);
int read_totals(mpt_context *ctx, unsigned long long out[5]);

}  // namespace mpt_engine

#endif
