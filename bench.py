"""Benchmark of the MI355X path-tracing core on the kd-tree stress scene.

Workload (one step = one whole render job): 1,000,000 random triangles in the
unit cube (lognormal edge lengths, mu = ln 0.005, sigma = 0.5), diffuse albedo
0.5, constant environment radiance 1, rendered at 1920x1080 with 64 samples per
pixel and maximum path depth 4, FP32.  The scene and its kd-tree are resident
in HBM before the timed region; the film stays on the device.

Prints one JSON line: paths/s (whole job), Mrays/s, the roofline of the
closest-hit kernel (algorithmic bytes per launch over its event-timed average
duration when launches do not overlap, against 8 TB/s HBM), a CPU baseline
(the double-precision oracle restatement of the reference PathTracer on a
bounded random sample of the same job's pixels at 65,536 spp, on all of the
process's host cores, which also gives the RMSE against an FP64 reference),
and the systems workloads: `batched` (64 jobs over the room, the soup and the
2048^2 terrain through one server, against the same jobs one at a time) and
`serving` (Poisson arrivals, p50 / p99 latency), and `bidir`: the
bidirectional renderer on the room at 256^2 and 1024^2 beside its oracle.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTH, HEIGHT, SPP, DEPTH = 1920, 1080, 64, 4
TRIANGLES = 1_000_000
HBM_PEAK_GBS = 8000.0
NODE_BYTES, RECORD_BYTES, RAY_BYTES, HIT_BYTES = 8, 48, 32, 8
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_soup1M_1920x1080_64spp.md")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(arrays, job, film_rgb, seconds, threads, nproc, max_pixels=256, ref_spp=65536):
    """The oracle (the FP64 restatement of the reference PathTracer) on host
    threads over a bounded sample of the same job: random pixels of its film
    at ref_spp samples each, i.e. a 65,536-spp double-precision reference of
    those pixels.  Pixels are added in groups of 16 until max_pixels or
    `seconds` of render time.  Returns the timing (paths/s; paths/s does not
    depend on spp) and the RMSE of the device image (film_rgb, the job's own
    spp) against that reference at the sampled pixels."""
    from oracle import pyoracle as po
    t0 = time.perf_counter()
    osc = po.OracleScene(arrays, threads=threads)
    build_s = time.perf_counter() - t0
    rng = np.random.default_rng(12345)
    npix = job.width * job.height
    order = rng.choice(npix, max_pixels, replace=False).astype(np.uint32)
    ref_job = job.__class__.from_buffer_copy(job)
    ref_job.spp, ref_job.seed = ref_spp, job.seed + 7919
    samples = np.arange(ref_spp, dtype=np.uint32)
    means, done, render_s = [], 0, 0.0
    while done < max_pixels and (done < 64 or render_s < seconds):
        group = order[done:done + 16]
        px = np.repeat(group, ref_spp)
        sm = np.tile(samples, len(group))
        t = time.perf_counter()
        L = osc.render_paths(ref_job, px, sm, threads=threads)
        render_s += time.perf_counter() - t
        means.append(L.reshape(len(group), ref_spp, 3).mean(axis=1))
        done += len(group)
    osc.close()
    ref = np.concatenate(means)
    pix = order[:done]
    img = film_rgb.reshape(-1, 3)[pix].astype(np.float64)
    rmse = float(np.sqrt(np.mean((img - ref) ** 2)))
    paths = done * ref_spp
    log(f"cpu baseline: {paths} paths in {render_s:.1f}s on {threads} threads; fp64 rmse {rmse:.5f}")
    base = {"value": paths / render_s, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"{done} random pixels x {ref_spp} spp of the same camera, scene and depth "
                      f"({paths} paths, {render_s:.1f} s on {threads} threads; oracle kd build {build_s:.1f} s "
                      f"excluded)",
            "nproc": nproc, "threads": threads, "cpu": cpu_model()}
    rm = {"value": rmse, "reference": "fp64 oracle", "reference_spp": ref_spp, "pixels": int(done),
          "mean_reference": float(ref.mean())}
    return base, rm


def cpu_model():
    """The host CPU's model name (the first "model name" of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_calibration(nbytes=256 << 20):
    """A plain device-to-device copy of known size, the reference dispatch
    that scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


RMSE_REF_SPP, RMSE_ROWS = 65536, 8


def rmse_vs_reference(m, ctx, scene, job):
    """Per-channel linear RMSE of one job-sized render (job.spp) against a
    65,536-spp reference of the same camera with an independent seed, over
    RMSE_ROWS evenly spaced full rows (the reference costs ~1e9 paths)."""
    rows = [int((k + 0.5) * job.height / RMSE_ROWS) for k in range(RMSE_ROWS)]
    img = m.Film(ctx, job.width, job.height)
    img.render(scene, job)
    ref_job = m.Job.from_buffer_copy(job)
    ref_job.spp, ref_job.seed = RMSE_REF_SPP, job.seed + 7919
    ref = m.Film(ctx, job.width, job.height)
    t0 = time.perf_counter()
    for r in rows:
        ref.render_range(scene, ref_job, r * job.width, (r + 1) * job.width)
    ctx.sync()
    ref_s = time.perf_counter() - t0
    a, b = img.download()[rows], ref.download()[rows]
    img.close()
    ref.close()
    rmse = float(np.sqrt(np.mean((a.astype(np.float64) - b) ** 2)))
    log(f"rmse {rmse:.5f} vs {RMSE_REF_SPP}-spp reference over {len(rows)} rows ({ref_s:.1f} s)")
    return {"value": rmse, "reference_spp": RMSE_REF_SPP, "pixels": len(rows) * job.width,
            "rows": rows, "mean_reference": float(b.mean())}


def measured_traffic(engine, kd_leaf):
    """HBM bytes per k_extend launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), if it was taken with the same
    batch size, stack, kernel form, block size and kd leaf size; else None."""
    import re
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch k_extend batch_paths=(\d+) stack=(\d+) persistent=(\d+) "
                  r"trace_block=(\d+) kd_leaf=(\d+): (\d+)", text)
    if not m:
        return None
    bp, st, pe, tb, leaf, val = (int(x) for x in m.groups())
    if (bp, st, pe, tb, leaf) != (engine["batch_paths"], engine["stack"], engine["persistent"], engine["trace_block"],
                                  kd_leaf):
        return None
    return val


def engine_settings(ctx, m):
    return {"batch_paths": ctx.get_option(m.MPT_OPT_BATCH_PATHS), "stack": ctx.get_option(m.MPT_OPT_STACK),
            "persistent": ctx.get_option(m.MPT_OPT_PERSISTENT), "lanes": ctx.get_option(m.MPT_OPT_LANES),
            "tree_top": ctx.get_option(m.MPT_OPT_TREE_TOP), "trace_block": ctx.get_option(m.MPT_OPT_TRACE_BLOCK),
            "shade_block": ctx.get_option(m.MPT_OPT_SHADE_BLOCK)}


def system_scenes(m, ctx, reuse=None):
    """The three resident scenes of the batched and serving workloads;
    `reuse` maps a scene kind to an (arrays, Scene) pair already built."""
    from milton_amd import workloads
    scenes, build = dict(reuse or {}), {}
    for kind in workloads.SCENES:
        if kind in scenes:
            continue
        arrays = m.generate(kind, workloads.SCENE_COUNTS[kind], 1)
        t0 = time.perf_counter()
        scenes[kind] = (arrays, m.Scene(ctx, arrays))
        build[kind] = round(time.perf_counter() - t0, 2)
    log(f"scenes built: {build}")
    return scenes, build


def bench_batched(m, ctx, scenes):
    """64 mixed jobs three ways: one server packing all of them into shared
    batches (graph-replayed, over the context's lanes); each job alone
    through the same engine (per-job); and each job alone on one stream with
    kernels launched one by one (a classic wavefront renderer).  Every film
    must come out bit-identical, so the RMSE is equal by construction."""
    from milton_amd import workloads
    cams = {k: v[0].camera for k, v in scenes.items()}
    jobs = workloads.batched_jobs(cams, 64, seed=11)
    total = sum(j.width * j.height * j.spp for _, j in jobs)
    for kind, (_, sc) in scenes.items():   # warm up graphs and caches per scene
        sc.render(m.make_job(cams[kind], 64, 64, 4, 8, seed=1))
    films = [m.Film(ctx, j.width, j.height) for _, j in jobs]
    job_s = np.zeros(len(jobs))

    def per_job():
        ctx.sync()
        t0 = time.perf_counter()
        for k, ((kind, j), f) in enumerate(zip(jobs, films)):
            t1 = time.perf_counter()
            f.render(scenes[kind][1], j)
            job_s[k] = time.perf_counter() - t1
        ctx.sync()
        return time.perf_counter() - t0

    saved = {o: ctx.get_option(o) for o in (m.MPT_OPT_LANES, m.MPT_OPT_GRAPHS)}
    ctx.set_option(m.MPT_OPT_LANES, 1)
    ctx.set_option(m.MPT_OPT_GRAPHS, 0)
    per_job()                                # warm-up of the classic form
    t_classic = per_job()
    classic = [f.download() for f in films]
    for o, v in saved.items():
        ctx.set_option(o, v)
    per_job()
    t_plain = per_job()
    srv = m.Server(ctx)
    ctx.sync()
    t0 = time.perf_counter()
    handles = [srv.submit(scenes[kind][1], j) for kind, j in jobs]
    srv.drain()
    t_batched = time.perf_counter() - t0
    st = srv.stats()
    same = same_classic = 0
    for h, (kind, j), f, c in zip(handles, jobs, films, classic):
        got = srv.result(h, j.width, j.height)
        same += int(np.array_equal(got, f.download()))
        same_classic += int(np.array_equal(got, c))
        f.close()
    srv.close()
    # where the per-job time goes: jobs under 4 M paths against the rest
    paths = np.array([j.width * j.height * j.spp for _, j in jobs], np.float64)
    small = paths < 4e6
    sat = saturated_bound(m, ctx, scenes, jobs, cams)
    log(f"batched: {len(jobs)} jobs, {total} paths, classic per-job {t_classic:.3f}s, per-job {t_plain:.3f}s, "
        f"batched {t_batched:.3f}s, {st['batches']} batches, identical films {same}/{len(jobs)} "
        f"(vs classic {same_classic}/{len(jobs)}); small jobs {int(small.sum())} = "
        f"{job_s[small].sum() / job_s.sum():.3f} of per-job time")
    return {"jobs": len(jobs), "paths": total, "seconds": t_batched, "paths_per_s": total / t_batched,
            "per_job_seconds": t_plain, "per_job_paths_per_s": total / t_plain,
            "classic_seconds": t_classic, "classic_paths_per_s": total / t_classic,
            "speedup_vs_per_job": t_plain / t_batched, "speedup_vs_classic": t_classic / t_batched,
            "identical_films": same, "identical_films_classic": same_classic, "batches": st["batches"],
            "extension_rays": st["extension_rays"], "shadow_rays": st["shadow_rays"],
            "small_jobs": int(small.sum()), "small_job_share_of_per_job_time": float(job_s[small].sum() / job_s.sum()),
            "saturated_seconds": sat["seconds"], "saturated_rates": sat["rates"],
            "batched_of_saturated": sat["seconds"] / t_batched, "per_job_of_saturated": sat["seconds"] / t_plain,
            "max_speedup_vs_per_job": t_plain / sat["seconds"]}


def saturated_bound(m, ctx, scenes, jobs, cams, side=1024, spp=100):
    """The mix's time if every job ran at its scene's saturated throughput:
    for each (scene, max depth) in the mix, one large job of that scene's
    camera (1024 x 1024 x 100 = 105 M paths, several full batches over the
    lanes) gives paths/s with the GPU full; the bound is the sum over jobs of
    paths / that rate.  Work per path does not depend on the resolution or on
    the sample count, so no schedule of the mix's jobs can finish much
    sooner than this."""
    rates = {}
    film = m.Film(ctx, side, side)
    for kind, d in sorted({(kind, j.max_depth) for kind, j in jobs}):
        cam = dict(cams[kind])
        cam["aspect"] = 1.0
        big = m.make_job(cam, side, side, spp, d, seed=5, job_id=0)
        film.render(scenes[kind][1], big)      # warm-up (graphs, workspace)
        ctx.sync()
        t0 = time.perf_counter()
        film.render(scenes[kind][1], big)
        ctx.sync()
        rates[f"{kind}/{d}"] = side * side * spp / (time.perf_counter() - t0)
    film.close()
    seconds = sum(j.width * j.height * j.spp / rates[f"{kind}/{j.max_depth}"] for kind, j in jobs)
    log(f"saturated bound: {seconds:.3f}s; rates (scene/depth: G paths/s) "
        f"{ {k: round(v / 1e9, 3) for k, v in rates.items()} }")
    return {"seconds": seconds, "rates": rates}


def bench_serving(m, ctx, scenes, rate, seconds):
    """Poisson arrivals of jobs drawn as in the batched mix at `rate` jobs/s
    for `seconds`; latency is submit -> film ready (host clock of the
    server)."""
    from milton_amd import workloads
    cams = {k: v[0].camera for k, v in scenes.items()}
    arrivals = workloads.poisson_arrivals(rate, seconds, seed=13)
    rng = np.random.default_rng(13)
    jobs = [workloads.draw_job(rng, cams, 1000 + i) for i in range(len(arrivals))]
    srv = m.Server(ctx)
    srv.start()
    t0 = time.perf_counter()
    handles = []
    for t_arr, (kind, j) in zip(arrivals, jobs):
        delay = t_arr - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        handles.append(srv.submit(scenes[kind][1], j))
    for h in handles:
        srv.wait(h, 600000)
    t_end = time.perf_counter() - t0
    lat, paths = [], 0
    for h, (_, j) in zip(handles, jobs):
        s = srv.status(h)
        lat.append(s["done_ms"] - s["submit_ms"])
        paths += j.width * j.height * j.spp
    srv.stop()
    srv.close()
    lat = np.array(lat)
    out = {"rate_jobs_s": rate, "jobs": len(jobs), "seconds": t_end, "paths": paths, "paths_per_s": paths / t_end,
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "max_ms": float(lat.max())}
    log(f"serving: {out}")
    return out


def systems(m, ctx, args, reuse=None):
    """The batched and serving workloads on the three resident scenes."""
    scenes, build = system_scenes(m, ctx, reuse)
    out = {"scene_build_s": build}
    tuned = None
    if not args.no_systems_tune:
        # engine settings picked by the autotuner on a sample of the mix's
        # largest job, 2^27 paths so that every candidate batch size runs
        # several batches over the lanes (a one-batch sample would favour
        # small batches, which overlap); the per-job, classic and batched runs
        # all use them
        from milton_amd import workloads
        cams = {k: v[0].camera for k, v in scenes.items()}
        kind, big = max(workloads.batched_jobs(cams, 64, seed=11),
                        key=lambda kj: kj[1].width * kj[1].height * kj[1].spp)
        tuned = scenes[kind][1].tune(big, 1 << 27)
        tuned["scene"] = int(kind)
        log(f"systems tuned on the largest job: {tuned}")
    out["batched"] = bench_batched(m, ctx, scenes)
    out["batched"]["tuned"] = tuned
    out["batched"]["engine"] = engine_settings(ctx, m)
    if args.serve_seconds > 0:
        # Poisson arrivals at 60% of the job rate the batched run sustained
        out["serving"] = bench_serving(m, ctx, scenes, 0.6 * out["batched"]["jobs"] / out["batched"]["seconds"],
                                       args.serve_seconds)
    for kind, (_, sc) in scenes.items():
        if not reuse or kind not in reuse:
            sc.close()
    return out


def bench_bidir(m, ctx, args, threads):
    """The bidirectional renderer (MPT_RENDERER_BIDIR) on the room: the
    CPU-runnable room job (256^2, 16 spp, depth 5) and the same camera at
    1024^2 (16.8 M paths), paths/s and rays/s (camera, subpath extension and
    connection visibility rays), beside the oracle's bidirectional
    restatement on host threads over a bounded sample of the small job."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    scene = m.Scene(ctx, arrays)
    out = {"renderer": "bidirectional (BidirectionalPathTracer.cpp:34-174)", "scene": "room"}
    for key, (w, h) in (("room_256", (256, 256)), ("room_1024", (1024, 1024))):
        job = m.make_job(arrays.camera, w, h, 16, 5, seed=1, renderer=m.MPT_RENDERER_BIDIR)
        film = m.Film(ctx, w, h)
        for _ in range(2):
            film.render(scene, job)
        ctx.sync()
        steps = 5
        t0 = time.perf_counter()
        runs = [film.render(scene, job) for _ in range(steps)]
        ctx.sync()
        el = time.perf_counter() - t0
        r = runs[0]
        rays = r["extension_rays"] + r["shadow_rays"]
        out[key] = {"width": w, "height": h, "spp": 16, "max_depth": 5, "ms_per_job": 1000 * el / steps,
                    "paths_per_s": w * h * 16 * steps / el, "mrays_per_s": rays * steps / el / 1e6,
                    "extension_rays": r["extension_rays"], "visibility_rays": r["shadow_rays"]}
        film.close()
        log(f"bidir {key}: {out[key]}")
    if not args.no_cpu:
        # the oracle on host threads over whole rows of the 1024^2 job, in a
        # fixed random order, until about cpu_seconds (at most 10 s) of work
        from oracle import pyoracle as po
        job = m.make_job(arrays.camera, 1024, 1024, 16, 5, seed=1, renderer=m.MPT_RENDERER_BIDIR)
        osc = po.OracleScene(arrays, threads=threads)
        rows = np.random.default_rng(4321).permutation(1024)
        done, t_cpu = 0, 0.0
        while done < 1024 and t_cpu < min(args.cpu_seconds, 10.0):
            px = np.repeat(np.arange(rows[done] * 1024, (rows[done] + 1) * 1024, dtype=np.uint32), 16)
            sm = np.tile(np.arange(16, dtype=np.uint32), 1024)
            t = time.perf_counter()
            osc.render_paths(job, px, sm, threads=threads)
            t_cpu += time.perf_counter() - t
            done += 1
        osc.close()
        paths = done * 1024 * 16
        out["cpu_baseline"] = {"value": paths / t_cpu, "unit": "paths/s", "cores": threads, "kind": "port",
                               "sample": f"{done} of the 1024 rows (in a fixed random order) of the 1024^2 x 16 "
                                         f"spp room job ({paths} paths, {t_cpu:.1f} s on {threads} threads)"}
        out["speedup_vs_cpu"] = out["room_1024"]["paths_per_s"] / out["cpu_baseline"]["value"]
        log(f"bidir cpu baseline: {out['cpu_baseline']}")
    scene.close()
    return out


def apply_engine_args(m, ctx, args):
    if args.batch_paths:
        ctx.set_option(m.MPT_OPT_BATCH_PATHS, args.batch_paths)
    if args.stack:
        ctx.set_option(m.MPT_OPT_STACK, args.stack)
    if args.persistent >= 0:
        ctx.set_option(m.MPT_OPT_PERSISTENT, args.persistent)
    if args.lanes:
        ctx.set_option(m.MPT_OPT_LANES, args.lanes)
    if args.tree_top >= 0:
        ctx.set_option(m.MPT_OPT_TREE_TOP, args.tree_top)
    if args.trace_block:
        ctx.set_option(m.MPT_OPT_TRACE_BLOCK, args.trace_block)
    if args.shade_block:
        ctx.set_option(m.MPT_OPT_SHADE_BLOCK, args.shade_block)


def main_systems(args):
    """--workload batched / serving: only the systems workloads."""
    import milton_amd as m
    ctx = m.Context(0)
    apply_engine_args(m, ctx, args)
    if args.workload == "batched":
        args.serve_seconds = 0
    sy = systems(m, ctx, args)
    b = sy["batched"]
    out = {"metric": "paths/s", "value": b["paths_per_s"], "unit": "paths/s", "n_gpus": 1,
           "steps": 1, "warmup": 1, "ms_per_step": 1000 * b["seconds"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batched: 64 jobs over room / 1M soup / 2048^2 terrain, sides 128-1024, "
                                  "spp 4-256, depth 4-8", "scene_build_s": sy["scene_build_s"],
                      "engine": engine_settings(ctx, m)}}
    out.update({k: v for k, v in sy.items() if k != "scene_build_s"})
    ctx.close()
    print(json.dumps(out), flush=True)


# two waves of chains per SIMD of the 256 CUs, the step kernel's occupancy
MLT_W, MLT_H, MLT_CHAINS, MLT_MUTATIONS = 256, 256, 131072, 128


def bench_mlt(m, ctx, args, threads):
    """Metropolis light transport (MLTRenderer, renderers/mlt/) of the room:
    a 256^2 film, 131,072 chains of 128 mutations (16.8 M mutations), the
    reference's defaults otherwise (10,000 seed paths, mltMaxDepth 10, all
    three mutations, the proposal filter).  The oracle's rate on the host
    threads comes from the same job, less its seed phase."""
    arrays = m.generate(m.MPT_SCENE_ROOM)
    scene = m.Scene(ctx, arrays)
    film = m.Film(ctx, MLT_W, MLT_H)
    nch = args.mlt_chains or MLT_CHAINS
    nmut = args.mlt_mutations or MLT_MUTATIONS
    job = m.make_job(arrays.camera, MLT_W, MLT_H, 1, 1, seed=1, renderer=m.MPT_RENDERER_MLT,
                     mlt=dict(chains=nch, mutations=nmut))
    # the seed phase alone: seeds, chain starts and one mutation
    one = m.make_job(arrays.camera, MLT_W, MLT_H, 1, 1, seed=1, renderer=m.MPT_RENDERER_MLT,
                     mlt=dict(chains=nch, mutations=1))
    film.render(scene, job)   # warm-up
    t = time.perf_counter()
    st1 = film.render(scene, one)
    seed_s = time.perf_counter() - t
    steps = 3
    ctx.sync()
    t = time.perf_counter()
    for _ in range(steps):
        st = film.render(scene, job)
    ctx.sync()
    el = (time.perf_counter() - t) / steps
    muts = nch * nmut
    out = {"renderer": "Metropolis light transport (MLTRenderer.cpp:70-235)", "scene": "room",
           "width": MLT_W, "height": MLT_H, "chains": nch, "mutations_per_chain": nmut,
           "ms_per_job": 1000 * el, "mutations_per_s": muts / el,
           "seed_phase_ms": 1000 * seed_s,
           "mutations_per_s_after_seeds": muts / max(1e-9, el - seed_s),
           "accepted": st["mlt_accepted"], "samples": st["mlt_samples"], "b": st["mlt_b"],
           "extension_rays": st["extension_rays"], "visibility_rays": st["shadow_rays"],
           "device_ms": st["render_ms"], "device_ms_seed_phase": st1["render_ms"]}
    log(f"mlt room: {out}")
    if not args.no_cpu:
        from oracle import pyoracle as po
        chains = nch   # the whole job: about 2 s on 16 threads
        cjob = m.make_job(arrays.camera, MLT_W, MLT_H, 1, 1, seed=1, renderer=m.MPT_RENDERER_MLT,
                          mlt=dict(chains=chains, mutations=nmut))
        osc = po.OracleScene(arrays, threads=threads)
        osc.mlt_seeds(cjob, threads=threads)   # warm-up
        t = time.perf_counter()
        osc.mlt_seeds(cjob, threads=threads)
        cseed = time.perf_counter() - t
        t = time.perf_counter()
        osc.mlt_render(cjob, threads=threads)
        cel = time.perf_counter() - t
        osc.close()
        cm = chains * nmut
        out["cpu_baseline"] = {"value": cm / max(1e-9, cel - cseed), "unit": "mutations/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{chains} chains x {nmut} mutations of the same job "
                                         f"({cel:.1f} s on {threads} threads, less its {cseed:.2f} s seed phase)"}
        out["speedup_vs_cpu"] = out["mutations_per_s_after_seeds"] / out["cpu_baseline"]["value"]
        log(f"mlt cpu baseline: {out['cpu_baseline']}")
    film.close()
    scene.close()
    return out


def main_mlt(args):
    """--workload mlt: only the Metropolis light transport room job."""
    import milton_amd as m
    ctx = m.Context(0)
    apply_engine_args(m, ctx, args)
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    r = bench_mlt(m, ctx, args, threads)
    out = {"metric": "mutations/s", "value": r["mutations_per_s"], "unit": "mutations/s", "n_gpus": 1,
           "steps": 3, "warmup": 1, "ms_per_step": r["ms_per_job"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"Metropolis light transport of the room, {MLT_W}x{MLT_H}, {r['chains']} chains x "
                                  f"{r['mutations_per_chain']} mutations", "engine": engine_settings(ctx, m)},
           "cpu_baseline": r.get("cpu_baseline"), "mlt": r}
    ctx.close()
    print(json.dumps(out), flush=True)


def main_bidir(args):
    """--workload bidir: only the bidirectional renderer's room jobs."""
    import milton_amd as m
    ctx = m.Context(0)
    apply_engine_args(m, ctx, args)
    nproc = os.cpu_count() or 1
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    b = bench_bidir(m, ctx, args, threads)
    r = b["room_1024"]
    out = {"metric": "paths/s", "value": r["paths_per_s"], "unit": "paths/s", "n_gpus": 1, "steps": 5,
           "warmup": 2, "ms_per_step": r["ms_per_job"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "bidirectional path tracing of the room, 1024x1024, 16 spp, depth 5",
                      "engine": engine_settings(ctx, m)},
           "cpu_baseline": b.get("cpu_baseline"), "bidir": b}
    ctx.close()
    print(json.dumps(out), flush=True)


def timed_render(m, ctx, film, scene, job, lanes):
    """One render with per-launch HIP event timing on `lanes` lanes."""
    saved = ctx.get_option(m.MPT_OPT_LANES)
    ctx.set_option(m.MPT_OPT_LANES, lanes)
    ctx.set_option(m.MPT_OPT_TIMING, 1)
    try:
        return film.render(scene, job)
    finally:
        ctx.set_option(m.MPT_OPT_TIMING, 0)
        ctx.set_option(m.MPT_OPT_LANES, saved)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (and the FP64 RMSE it yields)")
    ap.add_argument("--no-rmse", action="store_true", help="skip the RMSE against a 65,536-spp device reference")
    ap.add_argument("--no-systems", action="store_true", help="skip the batched and serving workloads")
    ap.add_argument("--no-bidir", action="store_true", help="skip the bidirectional renderer's room jobs")
    ap.add_argument("--no-mlt", action="store_true", help="skip the Metropolis light transport room job")
    ap.add_argument("--mlt-chains", type=int, default=0, help="Markov chains of the Metropolis job (0: the default)")
    ap.add_argument("--mlt-mutations", type=int, default=0,
                    help="mutations per chain of the Metropolis job (0: the default)")
    ap.add_argument("--batch-paths", type=int, default=0)
    ap.add_argument("--stack", type=int, default=0, help="traversal short-stack rows (8, 12, 16, 24)")
    ap.add_argument("--persistent", type=int, default=-1,
                    help="closest-hit kernel form: 0 grid-stride, 1 persistent waves, 2 one ray per lane")
    ap.add_argument("--lanes", type=int, default=0, help="engine lanes (streams) batches alternate over")
    ap.add_argument("--tree-top", type=int, default=-1,
                    help="1: the persistent closest-hit kernel reads the kd-tree's top from LDS (-1: library default)")
    ap.add_argument("--trace-block", type=int, default=0, help="threads per persistent closest-hit block (128, 256)")
    ap.add_argument("--shade-block", type=int, default=0, help="threads per shading block (256, 512)")
    ap.add_argument("--kd-leaf", type=int, default=0, help="kd leaf size (0: library default)")
    ap.add_argument("--kd-ci", type=float, default=0.0, help="kd SAH intersection cost (0: default)")
    ap.add_argument("--kd-eb", type=float, default=0.0, help="kd SAH empty-side factor (0: default)")
    ap.add_argument("--no-clip", action="store_true", help="kd build without triangle clipping")
    ap.add_argument("--tune", action="store_true",
                    help="autotune (kd leaf size and engine settings) on a sample of the job before timing")
    ap.add_argument("--workload", default="soup", choices=["soup", "batched", "serving", "bidir", "mlt"],
                    help="soup: the headline kd-tree stress scene (plus the systems workloads unless "
                         "--no-systems); batched / serving: only the systems workloads; bidir: only the "
                         "bidirectional renderer's room jobs")
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--no-systems-tune", action="store_true",
                    help="batched / serving at the context's settings instead of the autotuner's")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a 256 MiB device-to-device copy (reference dispatch for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X; --gpus {args.gpus} is not "
              f"supported (paths shard by pixel range; see the design notes)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "bidir":
        main_bidir(args)
        return
    if args.workload == "mlt":
        main_mlt(args)
        return
    if args.workload != "soup":
        main_systems(args)
        return

    import milton_amd as m
    ctx = m.Context(0)
    apply_engine_args(m, ctx, args)
    if args.copy_calibration:
        copy_calibration()
    t0 = time.perf_counter()
    arrays = m.generate(m.MPT_SCENE_SOUP, TRIANGLES, 1)
    gen_s = time.perf_counter() - t0
    job = m.make_job(arrays.camera, WIDTH, HEIGHT, SPP, DEPTH, seed=1)
    kp = m.KdParams()
    kp.leaf_size, kp.cost_intersect, kp.empty_bias, kp.no_clip = args.kd_leaf, args.kd_ci, args.kd_eb, args.no_clip
    t0 = time.perf_counter()
    tuned = None
    if args.tune:
        scene, tuned = m.Scene.tuned(ctx, arrays, job, kd=kp)
        log(f"tuned: {tuned}")
    else:
        scene = m.Scene(ctx, arrays, kp)
    build_s = time.perf_counter() - t0
    st = scene.stats()
    log(f"scene: gen {gen_s:.2f}s build+upload{' +tune' if tuned else ''} {build_s:.2f}s kd {st}")
    film = m.Film(ctx, WIDTH, HEIGHT)

    # Counter pass (untimed): kd nodes and references visited by every
    # closest-hit ray of the job; the job is deterministic, so the timed runs
    # trace exactly these rays.
    ctx.set_option(m.MPT_OPT_COUNTERS, 1)
    cst = film.render(scene, job)
    ctx.set_option(m.MPT_OPT_COUNTERS, 0)
    log(f"counters: {cst}")
    ext_rays = cst["extension_rays"]
    alg_bytes = (NODE_BYTES * cst["nodes_visited"] + RECORD_BYTES * cst["prims_tested"]
                 + (RAY_BYTES + HIT_BYTES) * ext_rays)

    # Timed steps: the production configuration (graph-replayed batches over
    # the context's lanes, no per-launch instrumentation).
    for _ in range(args.warmup):
        film.render(scene, job)
    ctx.sync()
    t0 = time.perf_counter()
    runs = [film.render(scene, job) for _ in range(args.steps)]
    ctx.sync()
    elapsed = time.perf_counter() - t0
    for r in runs[:3]:
        log(f"step: {r}")
    paths = WIDTH * HEIGHT * SPP
    rays = runs[0]["extension_rays"] + runs[0]["shadow_rays"]
    assert runs[0]["extension_rays"] == ext_rays
    ms_step = 1000.0 * elapsed / args.steps
    film_rgb = film.download()

    # Roofline of the closest-hit kernel: algorithmic bytes of one launch over
    # its duration measured alone (one lane, so no other batch's kernels
    # share the device), HIP events on the lane's stream.  Beside it, all
    # launches' bytes over the time during which at least one ran in the
    # production configuration.
    excl = [timed_render(m, ctx, film, scene, job, 1) for _ in range(2)][-1]
    busy = timed_render(m, ctx, film, scene, job, ctx.get_option(m.MPT_OPT_LANES))
    log(f"timed one lane: {excl}")
    log(f"timed all lanes: {busy}")
    launches = excl["extend_launches"]
    avg_launch_ms = excl["extend_ms"] / launches
    achieved = (alg_bytes / launches) / (avg_launch_ms * 1e-3) / 1e9
    achieved_busy = alg_bytes / (busy["extend_wall_ms"] * 1e-3) / 1e9
    engine = engine_settings(ctx, m)
    default_tree = not (args.tune or args.kd_ci or args.kd_eb or args.no_clip)
    traffic = measured_traffic(engine, args.kd_leaf or 4) if default_tree else None
    out = {
        "metric": "paths/s",
        "value": paths * args.steps / elapsed,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "mrays_per_s": rays * args.steps / elapsed / 1e6,
        "config": {"workload": "kd-tree stress scene: 1M random triangles, 1920x1080, 64 spp, max depth 4",
                   "triangles": TRIANGLES, "width": WIDTH, "height": HEIGHT, "spp": SPP, "max_depth": DEPTH,
                   "rays_per_step": rays, "kd_nodes": st["nodes"], "kd_refs": st["refs"],
                   "scene_build_s": round(build_s, 3), "tuned": tuned, "engine": engine},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                     "kernel": "k_extend", "alg_bytes_per_launch": alg_bytes / launches,
                     "avg_launch_ms": avg_launch_ms, "launch_timing": "one lane (launches do not overlap)",
                     "launches_per_step": launches, "busy_ms_per_step": busy["extend_wall_ms"],
                     "achieved_busy": achieved_busy, "frac_busy": achieved_busy / HBM_PEAK_GBS,
                     "nodes_per_ray": cst["nodes_visited"] / ext_rays,
                     "refs_per_ray": cst["prims_tested"] / ext_rays,
                     "shade_ms_per_step_one_lane": excl["shade_ms"]},
    }
    if not args.no_rmse:
        out["rmse"] = rmse_vs_reference(m, ctx, scene, job)
    if not args.no_cpu:
        nproc = os.cpu_count() or 1
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
        out["cpu_baseline"], out["rmse_fp64"] = cpu_baseline(arrays, job, film_rgb, args.cpu_seconds, threads,
                                                             nproc)
    else:
        out["cpu_baseline"] = None
    film.close()
    nproc = os.cpu_count() or 1
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    if not args.no_bidir:
        out["bidir"] = bench_bidir(m, ctx, args, threads)
    if not args.no_mlt:
        out["mlt"] = bench_mlt(m, ctx, args, threads)
    if not args.no_systems:
        out.update(systems(m, ctx, args, reuse={m.MPT_SCENE_SOUP: (arrays, scene)}))
    scene.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
