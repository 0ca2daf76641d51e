// Batching and serving over the wavefront engine (declared in
// include/milton_pt.h, "Batching and serving").
//
// The reference renders one job at a time: PointSampleRenderer::render runs a
// producer thread and N consumer threads over one film and returns when it is
// done (renderers/PointSampleRenderer.cpp:24-81).  Here a server keeps
// scenes resident and runs continuous batching:
//   admission   the HBM budget covers the lanes' ray queues and the films:
//               the server's batch size is set so that every lane's
//               workspace fits half the budget, and queued jobs are
//               admitted in arrival order while their films fit the rest
//               (and the active-job limit);
//   packing     a round picks the scene of the oldest admitted job with work
//               left and fills one batch with whole-pixel chunks of every
//               admitted job on that scene, oldest first;
//   overlap     up to max_inflight batches are enqueued before the oldest is
//               waited for, so planning and result copies overlap the device;
//   completion  when a job's last batch finishes its film is copied to host
//               memory and its device film is freed (admitting the next job).
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "engine.h"

using namespace mpt_engine;

namespace {

struct Job {
  uint64_t handle = 0;
  mpt_scene *scene = nullptr;
  mpt_job job{};
  uint32_t state = MPT_JOB_QUEUED;
  uint32_t error = 0;
  uint64_t npix = 0, next_pixel = 0, pixels_done = 0;
  uint32_t inflight = 0;        // batches in flight that reference the film
  DBuf<float4> film;
  std::vector<float> host;      // RGB result
  double submit_ms = 0, start_ms = 0, done_ms = 0;
  uint64_t film_bytes() const { return npix * sizeof(float4); }
};

struct Inflight {
  hipEvent_t done = nullptr;
  std::vector<std::pair<Job *, uint32_t>> parts;   // job, pixels in the batch
  uint64_t paths = 0;
};

}  // namespace

struct mpt_server {
  mpt_context *ctx = nullptr;
  mpt_server_config cfg{};
  std::mutex mu;
  std::condition_variable cv;
  std::map<uint64_t, std::unique_ptr<Job>> jobs;
  std::deque<Job *> queued;     // arrival order
  std::vector<Job *> active;    // admission order
  std::deque<Inflight> inflight;
  std::vector<hipEvent_t> free_events;
  uint64_t next_handle = 1;
  uint64_t film_bytes = 0, peak_film_bytes = 0;
  uint32_t cap = 0;           // paths per batch (the lanes' workspaces fit queue_bytes)
  uint32_t bidir_cap = 0;     // paths per bidirectional batch (their subpath records fit bidir_lane_bytes)
  uint64_t bidir_lane_bytes = 0;
  uint64_t queue_bytes = 0;   // device bytes reserved for the lanes' ray queues and subpath records
  uint64_t batches = 0, paths = 0, jobs_done = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::thread worker;
  std::atomic<bool> running{false};   // read without the mutex by step / drain
  bool stop = false;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

// Frees what a finished, cancelled or failed job holds on the device once no
// batch in flight still writes its film.
void release_film(mpt_server *sv, Job *j) {
  if (j->inflight == 0 && j->film.p) {
    j->film.release();
    sv->film_bytes -= j->film_bytes();
  }
}

void admit(mpt_server *sv) {
  while (!sv->queued.empty() && sv->active.size() < sv->cfg.max_active) {
    Job *j = sv->queued.front();
    if (j->state == MPT_JOB_CANCELLED) {
      sv->queued.pop_front();
      continue;
    }
    if (sv->queue_bytes + sv->film_bytes + j->film_bytes() > sv->cfg.hbm_budget) break;
    if (j->film.reserve(j->npix) != hipSuccess) break;
    // a filtered film is accumulated into, so it starts at zero; the clear
    // goes on the context stream, which every lane waits for before a batch
    if (j->job.filter && hipMemsetAsync(j->film.p, 0, j->film.bytes(), sv->ctx->stream) != hipSuccess) break;
    if (resolve_focus(j->scene, j->job) != MPT_OK) {
      // the job's autofocus ray could not be traced: it fails on its own
      sv->queued.pop_front();
      j->film.release();
      j->state = MPT_JOB_FAILED;
      j->error = (uint32_t)-MPT_E_DEVICE;
      j->done_ms = sv->now_ms();
      continue;
    }
    sv->film_bytes += j->film_bytes();
    sv->peak_film_bytes = std::max(sv->peak_film_bytes, sv->film_bytes);
    sv->queued.pop_front();
    j->state = MPT_JOB_RUNNING;
    j->start_ms = sv->now_ms();
    sv->active.push_back(j);
  }
}

// Waits for the oldest batch in flight and books its pixels.
int retire_one(mpt_server *sv, std::unique_lock<std::mutex> &lk) {
  Inflight b = std::move(sv->inflight.front());
  sv->inflight.pop_front();
  lk.unlock();
  const hipError_t e = hipEventSynchronize(b.done);
  lk.lock();
  sv->free_events.push_back(b.done);
  for (auto &part : b.parts) {
    Job *j = part.first;
    --j->inflight;
    if (e != hipSuccess && j->state == MPT_JOB_RUNNING) {
      j->state = MPT_JOB_FAILED;
      j->error = (uint32_t)(-MPT_E_DEVICE);
      j->done_ms = sv->now_ms();
    }
    if (j->state != MPT_JOB_RUNNING) {
      release_film(sv, j);
      continue;
    }
    j->pixels_done += part.second;
    if (j->pixels_done == j->npix) {
      // Film (numerator, denominator) -> RGB on the host.
      std::vector<float4> raw(j->npix);
      const hipError_t ce = hipMemcpy(raw.data(), j->film.p, j->film_bytes(), hipMemcpyDeviceToHost);
      if (ce != hipSuccess) {
        j->state = MPT_JOB_FAILED;
        j->error = (uint32_t)(-MPT_E_DEVICE);
      } else {
        j->host.resize(3 * j->npix);
        for (uint64_t i = 0; i < j->npix; ++i) {
          const float w = raw[i].w, k = w > 0 ? 1.0f / w : 0.0f;
          j->host[3 * i] = raw[i].x * k;
          j->host[3 * i + 1] = raw[i].y * k;
          j->host[3 * i + 2] = raw[i].z * k;
        }
        j->state = MPT_JOB_DONE;
        ++sv->jobs_done;
      }
      j->done_ms = sv->now_ms();
      release_film(sv, j);
    }
  }
  sv->cv.notify_all();
  return e == hipSuccess ? MPT_OK : MPT_E_DEVICE;
}

// One scheduling round: admit, plan one batch, enqueue it.  Returns the
// number of batches enqueued (0 when idle).
int round(mpt_server *sv, std::unique_lock<std::mutex> &lk, uint32_t *ran) {
  *ran = 0;
  admit(sv);
  // Drop jobs that left the running state from the active list.
  sv->active.erase(std::remove_if(sv->active.begin(), sv->active.end(),
                                  [](Job *j) { return j->state != MPT_JOB_RUNNING || j->next_pixel == j->npix; }),
                   sv->active.end());
  if (sv->active.empty()) {
    while (!sv->inflight.empty()) retire_one(sv, lk);
    admit(sv);
    sv->active.erase(std::remove_if(sv->active.begin(), sv->active.end(),
                                    [](Job *j) { return j->state != MPT_JOB_RUNNING || j->next_pixel == j->npix; }),
                     sv->active.end());
    if (sv->active.empty()) return MPT_OK;
  }
  mpt_context *ctx = sv->ctx;
  // a batch holds jobs of one scene and one renderer
  mpt_scene *scene = sv->active.front()->scene;
  const uint32_t renderer = sv->active.front()->job.renderer;
  const bool bidir = renderer == MPT_RENDERER_BIDIR;
  BatchPlan plan;
  plan.bidir = bidir;
  std::vector<std::pair<Job *, uint32_t>> parts;
  for (Job *j : sv->active) {
    if (j->scene != scene || j->job.renderer != renderer) continue;
    if (plan.jobs.size() == kMaxJobs || plan.chunks.size() == kMaxChunks) break;
    const uint32_t spp = j->job.spp;
    // a job with more light samples per vertex lowers the batch's capacity:
    // its shadow rays (paths x light samples) stay within the lane's queue
    const uint32_t md = std::max(plan.max_direct, std::max(1u, j->job.direct_samples));
    const uint32_t spv = std::max(1u, scene->dev.num_lights * md + scene->dev.num_points);
    const uint32_t cap = bidir ? sv->bidir_cap
                               : std::max(1u, std::min(std::min(sv->cap, batch_capacity(ctx, scene, md)), sv->cap / spv));
    const uint64_t room = plan.n_paths >= cap ? 0 : (cap - plan.n_paths) / spp;
    if (room == 0) {
      if (plan.n_paths == 0) {
        // a single pixel wider than the capacity: run it alone
      } else {
        break;
      }
    }
    const uint32_t np = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(room, 1), j->npix - j->next_pixel);
    const uint32_t jidx = (uint32_t)plan.jobs.size();
    plan.jobs.push_back(make_job(j->job, j->film.p));
    ChunkDev ch{};
    ch.job = jidx;
    ch.pix0 = (uint32_t)j->next_pixel;
    ch.npix = np;
    ch.slot0 = plan.n_paths;
    ch.pixoff = plan.n_pixels;
    plan.chunks.push_back(ch);
    plan.n_paths += np * spp;
    plan.n_pixels += np;
    plan.max_depth = std::max(plan.max_depth, j->job.max_depth);
    plan.max_direct = std::max(plan.max_direct, std::max(1u, j->job.direct_samples));
    j->next_pixel += np;
    ++j->inflight;
    parts.push_back({j, np});
    if (plan.n_paths >= cap) break;
  }
  if (plan.chunks.empty()) return MPT_OK;
  plan.capacity = bidir ? sv->bidir_cap
                        : std::max(1u, std::min(sv->cap, sv->cap / std::max(1u, scene->dev.num_lights * plan.max_direct +
                                                                                    scene->dev.num_points)));
  while (sv->inflight.size() >= sv->cfg.max_inflight) retire_one(sv, lk);
  // Batches alternate over the context's lanes; a lane first waits for the
  // context stream, where new films are cleared.
  const uint32_t lane = pick_lane(ctx);
  int rc = lane_follow(ctx, lane);
  if (!rc) rc = run_batch(ctx, scene, plan, nullptr, lane);
  hipEvent_t ev = nullptr;
  if (!rc) {
    if (!sv->free_events.empty()) {
      ev = sv->free_events.back();
      sv->free_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      rc = MPT_E_DEVICE;
    }
    if (!rc && hipEventRecord(ev, ctx->lanes[lane].stream) != hipSuccess) rc = MPT_E_DEVICE;
  }
  if (rc) {
    // whatever the failed batch did enqueue finishes before its films are
    // released (their memory may be handed to the next admitted job)
    (void)hipStreamSynchronize(ctx->lanes[lane].stream);
    for (auto &part : parts) {
      Job *j = part.first;
      --j->inflight;
      j->state = MPT_JOB_FAILED;
      j->error = (uint32_t)(-rc);
      j->done_ms = sv->now_ms();
      release_film(sv, j);
    }
    if (ev) sv->free_events.push_back(ev);
    sv->cv.notify_all();
    return rc;
  }
  Inflight b;
  b.done = ev;
  b.parts = std::move(parts);
  b.paths = plan.n_paths;
  sv->inflight.push_back(std::move(b));
  ++sv->batches;
  sv->paths += plan.n_paths;
  *ran = 1;
  return MPT_OK;
}

bool idle(const mpt_server *sv) {
  if (!sv->inflight.empty()) return false;
  for (const Job *j : sv->active)
    if (j->state == MPT_JOB_RUNNING) return false;
  for (const Job *j : sv->queued)
    if (j->state == MPT_JOB_QUEUED) return false;
  return true;
}

void worker_loop(mpt_server *sv) {
  (void)hipSetDevice(sv->ctx->device);
  std::unique_lock<std::mutex> lk(sv->mu);
  while (!sv->stop) {
    uint32_t ran = 0;
    round(sv, lk, &ran);
    if (!ran) {
      while (!sv->inflight.empty()) retire_one(sv, lk);
      if (idle(sv)) sv->cv.wait_for(lk, std::chrono::milliseconds(2));
    }
  }
  while (!sv->inflight.empty()) retire_one(sv, lk);
}

}  // namespace

extern "C" {

int mpt_server_create(mpt_context *ctx, const mpt_server_config *cfg, mpt_server **out) {
  if (!ctx || !out) return MPT_E_INVALID;
  *out = nullptr;
  std::unique_ptr<mpt_server> sv(new (std::nothrow) mpt_server());
  if (!sv) return MPT_E_NOMEM;
  sv->ctx = ctx;
  if (cfg) sv->cfg = *cfg;
  if (!sv->cfg.hbm_budget) sv->cfg.hbm_budget = 64ull << 30;
  if (!sv->cfg.max_active) sv->cfg.max_active = 256;
  // two batches in flight per lane by default; at most three per lane, since
  // each lane's descriptor staging ring holds four
  const uint32_t lanes = std::max(1u, std::min(ctx->num_lanes, kMaxLanes));
  if (!sv->cfg.max_inflight) sv->cfg.max_inflight = 2 * lanes;
  sv->cfg.max_inflight = std::min<uint32_t>(sv->cfg.max_inflight, 3 * lanes);
  // The lanes' ray queues take at most half the budget: halve the batch
  // until every lane's workspace (batches up to depth 64, one shadow ray per
  // path) and the subpath records of a bidirectional batch (half as many
  // bytes, at most kBidirLaneBytes) fit there.
  auto lane_bytes = [](uint64_t c) {
    const uint64_t ws = workspace_bytes(c, 64, c);
    return ws + std::min(kBidirLaneBytes, ws / 2);
  };
  uint64_t cap = ctx->batch_paths;
  while (cap > kMinBatchPaths && lanes * lane_bytes(cap) > sv->cfg.hbm_budget / 2) cap /= 2;
  if (lanes * lane_bytes(cap) > sv->cfg.hbm_budget / 2)
    return fail(ctx, MPT_E_BUDGET, "HBM budget too small for the ray queues of the server's lanes");
  sv->cap = (uint32_t)cap;
  sv->bidir_lane_bytes = lane_bytes(cap) - workspace_bytes(cap, 64, cap);
  // sized for the deepest subpaths, so every bidirectional batch fits
  sv->bidir_cap = std::min<uint32_t>(sv->cap, bidir_capacity(ctx, bdpt_max_depth(), sv->bidir_lane_bytes));
  sv->queue_bytes = lanes * lane_bytes(cap);
  if (reset_totals(ctx)) return MPT_E_DEVICE;
  ctx->servers.push_back(sv.get());
  *out = sv.release();
  return MPT_OK;
}

int mpt_server_destroy(mpt_server *sv) {
  if (!sv) return MPT_E_INVALID;
  mpt_server_stop(sv);
  {
    std::unique_lock<std::mutex> lk(sv->mu);
    while (!sv->inflight.empty()) retire_one(sv, lk);
  }
  for (hipEvent_t e : sv->free_events) (void)hipEventDestroy(e);
  auto &v = sv->ctx->servers;
  v.erase(std::remove(v.begin(), v.end(), sv), v.end());
  delete sv;
  return MPT_OK;
}

int mpt_submit(mpt_server *sv, mpt_scene *scene, const mpt_job *job, uint64_t *handle) {
  if (!sv || !scene || !job || !handle) return MPT_E_INVALID;
  if (scene->ctx != sv->ctx) return fail(sv->ctx, MPT_E_INVALID, "scene belongs to another context");
  int rc = check_job(sv->ctx, job);
  if (!rc) rc = check_scene_job(sv->ctx, scene, job);
  if (rc) return rc;
  // a Metropolis job's chains splat over the whole film, so it cannot share
  // batches of pixel ranges with other jobs
  if (job->renderer == MPT_RENDERER_MLT)
    return fail(sv->ctx, MPT_E_INVALID,
                "Metropolis light transport jobs render through mpt_render / mpt_render_film, not the server");
  std::unique_ptr<Job> j(new (std::nothrow) Job());
  if (!j) return MPT_E_NOMEM;
  j->scene = scene;
  j->job = *job;
  j->npix = (uint64_t)job->width * job->height;
  if (sv->queue_bytes + j->film_bytes() > sv->cfg.hbm_budget)
    return fail(sv->ctx, MPT_E_BUDGET, "film exceeds the HBM budget left beside the ray queues");
  std::lock_guard<std::mutex> lk(sv->mu);
  j->handle = sv->next_handle++;
  j->submit_ms = sv->now_ms();
  *handle = j->handle;
  sv->queued.push_back(j.get());
  sv->jobs[j->handle] = std::move(j);
  sv->cv.notify_all();
  return MPT_OK;
}

int mpt_status(mpt_server *sv, uint64_t handle, mpt_job_status *st) {
  if (!sv || !st) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(handle);
  if (it == sv->jobs.end()) return MPT_E_NOTFOUND;
  const Job &j = *it->second;
  st->state = j.state;
  st->error = j.error;
  st->pixels_done = j.pixels_done;
  st->pixels_total = j.npix;
  st->submit_ms = j.submit_ms;
  st->start_ms = j.start_ms;
  st->done_ms = j.done_ms;
  return MPT_OK;
}

int mpt_cancel(mpt_server *sv, uint64_t handle) {
  if (!sv) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(handle);
  if (it == sv->jobs.end()) return MPT_E_NOTFOUND;
  Job *j = it->second.get();
  if (j->state == MPT_JOB_QUEUED || j->state == MPT_JOB_RUNNING) {
    j->state = MPT_JOB_CANCELLED;
    j->done_ms = sv->now_ms();
    release_film(sv, j);
    sv->cv.notify_all();
  }
  return MPT_OK;
}

int mpt_result(mpt_server *sv, uint64_t handle, float *rgb) {
  if (!sv || !rgb) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  auto it = sv->jobs.find(handle);
  if (it == sv->jobs.end()) return MPT_E_NOTFOUND;
  Job *j = it->second.get();
  if (j->state == MPT_JOB_CANCELLED) return MPT_E_CANCELLED;
  if (j->state == MPT_JOB_FAILED) return -(int)j->error;
  if (j->state != MPT_JOB_DONE) return fail(sv->ctx, MPT_E_INVALID, "job not finished");
  std::memcpy(rgb, j->host.data(), j->host.size() * sizeof(float));
  sv->jobs.erase(it);
  return MPT_OK;
}

int mpt_wait(mpt_server *sv, uint64_t handle, uint32_t timeout_ms) {
  if (!sv) return MPT_E_INVALID;
  std::unique_lock<std::mutex> lk(sv->mu);
  const auto deadline = std::chrono::steady_clock::now() + std::chrono::milliseconds(timeout_ms);
  for (;;) {
    auto it = sv->jobs.find(handle);
    if (it == sv->jobs.end()) return MPT_E_NOTFOUND;
    const uint32_t s = it->second->state;
    if (s != MPT_JOB_QUEUED && s != MPT_JOB_RUNNING) return 0;
    if (sv->cv.wait_until(lk, deadline) == std::cv_status::timeout) return 1;
  }
}

int mpt_server_step(mpt_server *sv, uint32_t *batches_run) {
  if (!sv) return MPT_E_INVALID;
  if (sv->running) return fail(sv->ctx, MPT_E_INVALID, "the worker thread drives this server");
  std::unique_lock<std::mutex> lk(sv->mu);
  uint32_t ran = 0;
  const int rc = round(sv, lk, &ran);
  if (!ran)
    while (!sv->inflight.empty()) retire_one(sv, lk);
  if (batches_run) *batches_run = ran;
  return rc;
}

int mpt_server_drain(mpt_server *sv) {
  if (!sv) return MPT_E_INVALID;
  if (sv->running) {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv.wait(lk, [&] { return idle(sv); });
    return MPT_OK;
  }
  int first_error = MPT_OK;
  for (;;) {
    uint32_t ran = 0;
    const int rc = mpt_server_step(sv, &ran);
    if (rc && !first_error) first_error = rc;
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!ran && idle(sv)) break;
  }
  return first_error;
}

int mpt_server_start(mpt_server *sv) {
  if (!sv) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  if (sv->running) return MPT_OK;
  sv->stop = false;
  sv->running = true;
  sv->worker = std::thread(worker_loop, sv);
  return MPT_OK;
}

int mpt_server_stop(mpt_server *sv) {
  if (!sv) return MPT_E_INVALID;
  {
    std::lock_guard<std::mutex> lk(sv->mu);
    if (!sv->running) return MPT_OK;
    sv->stop = true;
    sv->cv.notify_all();
  }
  sv->worker.join();
  sv->running = false;
  return MPT_OK;
}

int mpt_server_memory(mpt_server *sv, uint64_t out[6]) {
  if (!sv || !out) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  out[0] = sv->cfg.hbm_budget;
  out[1] = sv->queue_bytes;
  out[2] = sv->film_bytes;
  out[3] = sv->peak_film_bytes;
  out[4] = sv->cap;
  uint64_t held = 0;
  for (const Lane &ln : sv->ctx->lanes) held += ln.ws.bytes();
  out[5] = held;
  return MPT_OK;
}

int mpt_server_stats(mpt_server *sv, uint64_t out[6]) {
  if (!sv || !out) return MPT_E_INVALID;
  std::lock_guard<std::mutex> lk(sv->mu);
  unsigned long long tot[5] = {0, 0, 0, 0, 0};
  if (read_totals(sv->ctx, tot)) return MPT_E_DEVICE;
  out[0] = sv->batches;
  out[1] = sv->paths;
  out[2] = tot[1];
  out[3] = tot[2];
  out[4] = sv->jobs_done;
  out[5] = sv->peak_film_bytes;
  return MPT_OK;
}

}  // extern "C"
