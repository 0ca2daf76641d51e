// Test driver, not part of the library: exercises the threaded host code,
// under ThreadSanitizer where the process can run with it (built by
// tests/tsan/Makefile, run by tests/test_tsan.py and
// tests/test_gpu_server_threads.py).  Each mode also checks that the
// threaded results equal the single-threaded ones bit for bit.
//   kd      mpt_kd_build (subtrees on 8 threads) twice and on 1 thread over a
//           200k-triangle soup, then mpt_kd_trace of 200k rays on 8 threads
//           and on 1 thread: the same tree, hits and counters;
//   oracle  the oracle's threaded kd build, closest hits and render (8
//           threads) against 1 thread;
//   server  (needs a GPU; built without the sanitizer, which keeps the HIP
//           runtime from finding a device) a server with its worker
//           thread, fed by four submitting threads that poll status, wait
//           and cancel while it runs; every finished film equals the job
//           rendered alone;
//   selftest a planted race on a plain counter, which the sanitizer must
//           report (the instrumentation is live).
// Exit status 0 when every check passes; ThreadSanitizer's own exit code
// (TSAN_OPTIONS exitcode) on a reported race.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "milton_pt.h"
#include "milton_pt_scenes.h"
#include "../../oracle/oracle_pt.h"

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

namespace {

int kd_mode() {
  mpt_gen *g = nullptr;
  CHECK(mpt_gen_create(MPT_SCENE_SOUP, 200000, 1, &g) == 0);
  mpt_scene_desc d;
  mpt_camera cam;
  CHECK(mpt_gen_describe(g, &d, &cam) == 0);
  mpt_kd_tree *t[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < 3; ++k) {
    mpt_kd_params p{};
    p.num_threads = k < 2 ? 8 : 1;
    CHECK(mpt_kd_build(&d, &p, &t[k]) == 0);
  }
  uint64_t st[3][6];
  for (int k = 0; k < 3; ++k) CHECK(mpt_kd_stats(t[k], st[k]) == 0);
  for (int k = 1; k < 3; ++k)
    for (int i : {0, 1, 2, 3, 5}) CHECK(st[k][i] == st[0][i]);   // [4] is the build time
  const uint64_t n = 200000;
  std::vector<mpt_ray> rays(n);
  CHECK(mpt_gen_rays(g, n, 7, 1920, 1080, rays.data()) == 0);
  std::vector<mpt_hit> h[3];
  uint64_t cnt[3][2];
  for (int k = 0; k < 3; ++k) {
    h[k].resize(n);
    CHECK(mpt_kd_trace(t[k], rays.data(), n, h[k].data(), cnt[k], k < 2 ? 8 : 1) == 0);
  }
  for (int k = 1; k < 3; ++k) {
    CHECK(std::memcmp(h[k].data(), h[0].data(), n * sizeof(mpt_hit)) == 0);
    CHECK(cnt[k][0] == cnt[0][0] && cnt[k][1] == cnt[0][1]);
  }
  std::printf("kd: %llu nodes, %llu refs, %llu nodes visited by %llu rays on 8 and 1 threads: identical\n",
              (unsigned long long)st[0][0], (unsigned long long)st[0][2], (unsigned long long)cnt[0][0],
              (unsigned long long)n);
  for (auto *x : t) mpt_kd_destroy(x);
  mpt_gen_destroy(g);
  return 0;
}

std::vector<float> oracle_materials(const mpt_scene_desc &d) {
  std::vector<float> m(16ull * d.num_materials, 0.0f);
  for (uint32_t i = 0; i < d.num_materials; ++i) {
    const mpt_material &s = d.materials[i];
    float *o = &m[16ull * i];
    o[0] = (float)s.type;
    for (int c = 0; c < 3; ++c) {
      o[1 + c] = s.kd[c];
      o[4 + c] = s.ks[c];
      o[10 + c] = s.Le[c];
    }
    o[7] = s.exponent;
    o[8] = s.ior;
    o[9] = s.transparency;
    o[13] = s.ior_g;
    o[14] = s.ior_b;
  }
  return m;
}

orc_scene *oracle_scene(const mpt_scene_desc &d, const std::vector<float> &mats, int threads) {
  return orc_scene_create(d.num_vertices, d.vertices, d.num_triangles, d.indices, d.tri_material, d.num_spheres,
                          d.spheres, d.sphere_material, d.num_quads, d.quads, d.quad_material, d.num_materials,
                          mats.data(), d.env_radiance, d.tri_flags, d.vertex_normals, d.vertex_normal_mask, 0, 0,
                          0, threads);
}

int oracle_mode() {
  // closest hits on a 50k soup, kd built on 8 threads and on 1
  mpt_gen *g = nullptr;
  CHECK(mpt_gen_create(MPT_SCENE_SOUP, 50000, 2, &g) == 0);
  mpt_scene_desc d;
  mpt_camera cam;
  CHECK(mpt_gen_describe(g, &d, &cam) == 0);
  const std::vector<float> mats = oracle_materials(d);
  orc_scene *s8 = oracle_scene(d, mats, 8), *s1 = oracle_scene(d, mats, 1);
  CHECK(s8 && s1);
  uint64_t k8[6], k1[6];
  orc_kd_stats(s8, 0, k8);
  orc_kd_stats(s1, 0, k1);
  CHECK(std::memcmp(k8, k1, sizeof(k8)) == 0);
  const uint64_t n = 50000;
  std::vector<mpt_ray> rays(n);
  CHECK(mpt_gen_rays(g, n, 9, 640, 480, rays.data()) == 0);
  std::vector<double> r(6 * n), t8(n), t1(n);
  for (uint64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      r[6 * i + c] = rays[i].o[c];
      r[6 * i + 3 + c] = rays[i].d[c];
    }
  std::vector<uint32_t> i8(n), i1(n);
  orc_trace_closest(s8, n, r.data(), t8.data(), i8.data(), 8);
  orc_trace_closest(s1, n, r.data(), t1.data(), i1.data(), 1);
  CHECK(std::memcmp(t8.data(), t1.data(), n * sizeof(double)) == 0 && i8 == i1);
  orc_scene_destroy(s8);
  orc_scene_destroy(s1);
  mpt_gen_destroy(g);
  // the room rendered on 8 threads and on 1
  CHECK(mpt_gen_create(MPT_SCENE_ROOM, 0, 1, &g) == 0);
  CHECK(mpt_gen_describe(g, &d, &cam) == 0);
  const std::vector<float> rm = oracle_materials(d);
  orc_scene *room = oracle_scene(d, rm, 8);
  CHECK(room);
  orc_job job;
  std::memset(&job, 0, sizeof(job));
  const double camera[11] = {cam.eye[0], cam.eye[1], cam.eye[2], cam.look[0], cam.look[1], cam.look[2],
                             cam.up[0],  cam.up[1],  cam.up[2],  cam.height_angle, cam.aspect};
  std::memcpy(job.camera, camera, sizeof(camera));
  job.width = job.height = 32;
  job.spp = 16;
  job.max_depth = 5;
  job.seed = 1;
  job.pix_begin = 0;
  job.pix_end = 32 * 32;
  job.direct_samples = 1;
  const uint64_t np = 32 * 32;
  std::vector<double> a(3 * np), a2(3 * np), b(3 * np), b2(3 * np);
  uint64_t ca[4], cb[4];
  orc_render(room, &job, a.data(), a2.data(), ca, 8);
  orc_render(room, &job, b.data(), b2.data(), cb, 1);
  CHECK(a == b && a2 == b2 && std::memcmp(ca, cb, sizeof(ca)) == 0);
  std::printf("oracle: %llu closest hits and a %llu-path render on 8 and 1 threads: identical\n",
              (unsigned long long)n, (unsigned long long)ca[3]);
  orc_scene_destroy(room);
  mpt_gen_destroy(g);
  return 0;
}

// From the HIP runtime the library links: names why a device did not open.
extern "C" int hipInit(unsigned flags);
extern "C" const char *hipGetErrorString(int err);

int server_mode() {
  mpt_context *ctx = nullptr;
  const int hip = hipInit(0);
  if (hip != 0) std::fprintf(stderr, "hipInit: %d %s\n", hip, hipGetErrorString(hip));
  CHECK(mpt_open(0, &ctx) == 0);
  mpt_gen *g = nullptr;
  CHECK(mpt_gen_create(MPT_SCENE_ROOM, 0, 1, &g) == 0);
  mpt_scene_desc d;
  mpt_camera cam;
  CHECK(mpt_gen_describe(g, &d, &cam) == 0);
  mpt_scene *sc = nullptr;
  CHECK(mpt_scene_create(ctx, &d, nullptr, &sc) == 0);
  CHECK(mpt_set_option(ctx, MPT_OPT_BATCH_PATHS, 1 << 14) == 0);   // many small batches
  const int per_thread = 6, threads = 4, n = per_thread * threads;
  std::vector<mpt_job> jobs(n);
  for (int i = 0; i < n; ++i) {
    mpt_job &j = jobs[i];
    std::memset(&j, 0, sizeof(j));
    j.camera = cam;
    j.width = 24 + 8 * (i % 5);
    j.height = 16 + 8 * (i % 3);
    j.spp = 1 + 3 * (i % 3);
    j.max_depth = 3 + i % 4;
    j.seed = 100 + i;
    j.job_id = i;
    j.direct_samples = 1 + (i % 2);
  }
  // films of the jobs rendered alone, before the server exists
  std::vector<std::vector<float>> alone(n);
  for (int i = 0; i < n; ++i) {
    alone[i].resize(3ull * jobs[i].width * jobs[i].height);
    mpt_render_stats rs;
    CHECK(mpt_render(sc, &jobs[i], alone[i].data(), &rs) == 0);
  }
  mpt_server *sv = nullptr;
  CHECK(mpt_server_create(ctx, nullptr, &sv) == 0);
  CHECK(mpt_server_start(sv) == 0);
  std::atomic<int> bad{0};
  std::vector<uint64_t> handles(n, 0);
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t)
    th.emplace_back([&, t]() {
      for (int k = 0; k < per_thread; ++k) {
        const int i = t * per_thread + k;
        if (mpt_submit(sv, sc, &jobs[i], &handles[i]) != 0) {
          ++bad;
          continue;
        }
        mpt_job_status st;
        if (mpt_status(sv, handles[i], &st) != 0) ++bad;
        if (i % 7 == 3 && mpt_cancel(sv, handles[i]) != 0) ++bad;   // some cancelled while queued or running
        std::this_thread::sleep_for(std::chrono::microseconds(300));
      }
      for (int k = 0; k < per_thread; ++k) {
        const int i = t * per_thread + k;
        if (handles[i] && mpt_wait(sv, handles[i], 120000) != 0) ++bad;
      }
    });
  for (auto &x : th) x.join();
  CHECK(bad.load() == 0);
  int done = 0, cancelled = 0;
  for (int i = 0; i < n; ++i) {
    mpt_job_status st;
    CHECK(mpt_status(sv, handles[i], &st) == 0);
    if (st.state == MPT_JOB_CANCELLED) {
      CHECK(i % 7 == 3);
      ++cancelled;
      continue;
    }
    CHECK(st.state == MPT_JOB_DONE);
    std::vector<float> film(alone[i].size());
    CHECK(mpt_result(sv, handles[i], film.data()) == 0);
    CHECK(film == alone[i]);
    ++done;
  }
  CHECK(mpt_server_stop(sv) == 0);
  CHECK(mpt_server_destroy(sv) == 0);
  std::printf("server: %d jobs from %d threads, %d identical to alone, %d cancelled\n", n, threads, done,
              cancelled);
  mpt_scene_destroy(sc);
  mpt_gen_destroy(g);
  mpt_close(ctx);
  return 0;
}

// Two threads add to one plain counter: ThreadSanitizer must report it
// (the check that the instrumentation is live in this build).
int selftest_mode() {
  static volatile uint64_t counter = 0;
  auto add = []() {
    for (int i = 0; i < 100000; ++i) counter = counter + 1;
  };
  std::thread a(add), b(add);
  a.join();
  b.join();
  std::printf("selftest: counter %llu\n", (unsigned long long)counter);
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  const char *mode = argc > 1 ? argv[1] : "";
  int rc = 2;
  if (!std::strcmp(mode, "kd")) rc = kd_mode();
  else if (!std::strcmp(mode, "oracle")) rc = oracle_mode();
  else if (!std::strcmp(mode, "server")) rc = server_mode();
  else if (!std::strcmp(mode, "selftest")) rc = selftest_mode();
  else std::fprintf(stderr, "usage: %s kd|oracle|server|selftest\n", argv[0]);
  std::fflush(stdout);
  return rc;
}
