// Autotuner of the wavefront engine (declared in include/milton_pt.h,
// "Autotuning").
//
// The reference has no tunables on this path beyond kdTreeAccel::BuildParams
// (accel/kdTreeAccel.h:68-116), which it fixes per scene file.  Here the
// settings that move throughput without changing results are chosen by
// measurement on a sample of the job itself:
//   persistent  closest-hit kernel form (one ray per lane, or persistent
//               waves with lane refill),
//   stack       short-stack entries of the kd walk (8, 12, 16),
//   blocks      threads per block of the persistent closest-hit kernel (128,
//               256) and of the shading kernel (256, 512),
//   batch paths queue depth of a wavefront batch (2^20 .. 2^26; a sample
//               smaller than a candidate runs as one batch),
//   leaf size   kd leaf size (1 .. 16, mpt_tune_scene only; rebuilds the tree).
// The sample is the job's own camera, depth and seed at a reduced number of
// samples per pixel over a centred window of pixels, about sample_paths paths.
// Each candidate is rendered once untimed (graph capture, workspace growth)
// and then timed twice with device events; the faster time counts.  The search
// is one coordinate pass in the order above, starting from the context's
// current settings, so the result is never slower than the start on the
// sample.
#include <algorithm>
#include <cstring>
#include <vector>

#include "engine.h"

using namespace mpt_engine;

namespace {

struct Sample {
  mpt_job job{};
  uint32_t pix_begin = 0, pix_end = 0;
  uint64_t paths = 0;
};

Sample make_sample(const mpt_job &job, uint64_t sample_paths) {
  Sample s;
  s.job = job;
  const uint64_t npix = (uint64_t)job.width * job.height;
  const uint64_t spp = std::max<uint64_t>(1, std::min<uint64_t>(job.spp, sample_paths / std::max<uint64_t>(npix, 1)));
  s.job.spp = (uint32_t)spp;
  const uint64_t win = std::min<uint64_t>(npix, std::max<uint64_t>(1, sample_paths / spp));
  s.pix_begin = (uint32_t)((npix - win) / 2);
  s.pix_end = (uint32_t)(s.pix_begin + win);
  s.paths = win * spp;
  return s;
}

// Best of two timed renders after one untimed one; paths per second.
int time_sample(mpt_scene *scene, const Sample &smp, mpt_film *film, double &rate) {
  mpt_render_stats st;
  int rc = render_range(scene, &smp.job, smp.pix_begin, smp.pix_end, film->data.p, nullptr, &st);
  if (rc) return rc;
  double best_ms = 1e30;
  for (int rep = 0; rep < 2; ++rep) {
    if ((rc = render_range(scene, &smp.job, smp.pix_begin, smp.pix_end, film->data.p, nullptr, &st))) return rc;
    best_ms = std::min(best_ms, (double)st.render_ms);
  }
  rate = best_ms > 0 ? (double)smp.paths / (best_ms * 1e-3) : 0.0;
  return MPT_OK;
}

struct Settings {
  uint32_t persistent, stack, batch_paths, trace_block, shade_block;
  bool operator==(const Settings &o) const {
    return persistent == o.persistent && stack == o.stack && batch_paths == o.batch_paths &&
           trace_block == o.trace_block && shade_block == o.shade_block;
  }
};

Settings current(const mpt_context *ctx) {
  return {ctx->persistent, ctx->stack, ctx->batch_paths, ctx->trace_block, ctx->shade_block};
}

void apply(mpt_context *ctx, const Settings &s) {
  if (ctx->persistent != s.persistent || ctx->stack != s.stack || ctx->trace_block != s.trace_block ||
      ctx->shade_block != s.shade_block)
    drop_graphs(ctx, 0);
  ctx->persistent = s.persistent;
  ctx->stack = s.stack;
  ctx->batch_paths = s.batch_paths;
  ctx->trace_block = s.trace_block;
  ctx->shade_block = s.shade_block;
}

// One coordinate pass over the engine settings on `scene`.
int tune_engine(mpt_scene *scene, const Sample &smp, mpt_film *film, mpt_tune_result &res) {
  mpt_context *ctx = scene->ctx;
  Settings best = current(ctx);
  double best_rate = 0;
  int rc = time_sample(scene, smp, film, best_rate);
  if (rc) return rc;
  ++res.trials;
  res.baseline_paths_per_s = best_rate;
  auto trial = [&](Settings s) -> int {
    if (s == best) return MPT_OK;
    apply(ctx, s);
    double r = 0;
    int e = time_sample(scene, smp, film, r);
    if (e) return e;
    ++res.trials;
    if (r > best_rate) {
      best_rate = r;
      best = s;
    }
    apply(ctx, best);
    return MPT_OK;
  };
  for (uint32_t p : {0u, 1u, 2u}) {
    Settings s = best;
    s.persistent = p;
    if ((rc = trial(s))) return rc;
  }
  for (uint32_t k : {8u, 12u, 16u}) {
    Settings s = best;
    s.stack = k;
    if ((rc = trial(s))) return rc;
  }
  if (best.persistent == 1) {
    for (uint32_t t : {128u, 256u}) {
      Settings s = best;
      s.trace_block = t;
      if ((rc = trial(s))) return rc;
    }
  }
  for (uint32_t t : {256u, 512u}) {
    Settings s = best;
    s.shade_block = t;
    if ((rc = trial(s))) return rc;
  }
  for (uint32_t b = 20; b <= 26; ++b) {
    Settings s = best;
    s.batch_paths = 1u << b;
    if ((rc = trial(s))) return rc;
  }
  apply(ctx, best);
  res.persistent = best.persistent;
  res.stack = best.stack;
  res.batch_paths = best.batch_paths;
  res.trace_block = best.trace_block;
  res.shade_block = best.shade_block;
  res.paths_per_s = best_rate;
  return MPT_OK;
}

int check_tune_args(mpt_context *ctx, const mpt_job *job, uint64_t sample_paths, mpt_tune_result *res) {
  if (!job || !res) return fail(ctx, MPT_E_INVALID, "NULL argument");
  if (sample_paths < 1024) return fail(ctx, MPT_E_INVALID, "sample_paths below 1024");
  return check_job(ctx, job);
}

}  // namespace

extern "C" {

int mpt_tune(mpt_scene *scene, const mpt_job *job, uint64_t sample_paths, mpt_tune_result *result) {
  if (!scene) return MPT_E_INVALID;
  mpt_context *ctx = scene->ctx;
  int rc = check_tune_args(ctx, job, sample_paths, result);
  if (rc) return rc;
  std::memset(result, 0, sizeof(*result));
  result->leaf_size = scene->leaf_size;
  const Sample smp = make_sample(*job, sample_paths);
  mpt_film *film = nullptr;
  if ((rc = mpt_film_create(ctx, job->width, job->height, &film))) return rc;
  rc = tune_engine(scene, smp, film, *result);
  mpt_film_destroy(film);
  return rc;
}

int mpt_tune_scene(mpt_context *ctx, const mpt_scene_desc *desc, const mpt_kd_params *kd, const mpt_job *job,
                   uint64_t sample_paths, mpt_scene **scene, mpt_tune_result *result) {
  if (!ctx || !desc || !scene) return fail(ctx, MPT_E_INVALID, "NULL argument");
  *scene = nullptr;
  int rc = check_tune_args(ctx, job, sample_paths, result);
  if (rc) return rc;
  std::memset(result, 0, sizeof(*result));
  const Sample smp = make_sample(*job, sample_paths);
  mpt_film *film = nullptr;
  if ((rc = mpt_film_create(ctx, job->width, job->height, &film))) return rc;
  mpt_kd_params kp{};
  if (kd) kp = *kd;
  const uint32_t start = kp.leaf_size ? std::min(16u, std::max(1u, kp.leaf_size)) : 4u;

  // Hill climb over leaf sizes 1, 2, 4, 8, 16 (and the start value) at the
  // context's engine settings; each candidate is a full tree build.
  mpt_scene *best = nullptr;
  double best_rate = 0;
  uint32_t best_leaf = 0;
  auto build_and_time = [&](uint32_t leaf, bool &better) -> int {
    better = false;
    kp.leaf_size = leaf;
    mpt_scene *s = nullptr;
    int e = mpt_scene_create(ctx, desc, &kp, &s);
    if (e) return e;
    double r = 0;
    if ((e = time_sample(s, smp, film, r))) {
      mpt_scene_destroy(s);
      return e;
    }
    ++result->trials;
    if (!best || r > best_rate) {
      if (best) mpt_scene_destroy(best);
      best = s;
      best_rate = r;
      best_leaf = leaf;
      better = true;
    } else {
      mpt_scene_destroy(s);
    }
    return MPT_OK;
  };
  bool better = false;
  if ((rc = build_and_time(start, better))) goto done;
  result->baseline_paths_per_s = best_rate;
  for (int dir : {+1, -1}) {
    uint32_t leaf = start;
    for (;;) {
      const uint32_t next = dir > 0 ? leaf * 2 : leaf / 2;
      if (next < 1 || next > 16 || next == leaf) break;
      if ((rc = build_and_time(next, better))) goto done;
      if (!better) break;
      leaf = next;
    }
    if (best_leaf != start) break;   // improved upward: no need to look down
  }
  {
    const double leaf_baseline = result->baseline_paths_per_s;
    rc = tune_engine(best, smp, film, *result);
    result->baseline_paths_per_s = leaf_baseline;
    result->leaf_size = best_leaf;
  }
done:
  mpt_film_destroy(film);
  if (rc) {
    if (best) mpt_scene_destroy(best);
    return rc;
  }
  *scene = best;
  return MPT_OK;
}

}  // extern "C"
