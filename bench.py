#!/usr/bin/env python3
"""Benchmark of the gfx950 Tracer on the material-spheres grid workload.

Workload: 7x7 UV spheres (64x32 segments) on a 100-tri ground plane under a 2048x1024
float32 procedural sky rotated 30 deg, BVH leaves of at most 4 triangles, rendered at
1920x1080 with 256 samples per pixel and the reference's estimator (bounces 5, energy
cut-off 0.01, per-sample ACES).  One step = one full frame (530,841,600 camera paths).

Printed: one JSON line with Msamples/s (camera paths per second over the timed steps, scene
build and upload excluded as in main.cpp:125-131), the roofline of the dominant kernel
(extend = closest-hit traversal) and the CPU baseline (the oracle built with the
reference's release flags, mt19937 per tile, all worker threads of this host's share).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--spp 256] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p, plain vs cached at equal RMSE; fwd+adjoint ms/iter"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one closest-hit ray: world AABB 24 B per instance tested, inverse
# matrix 64 B per instance entered, 32 B per BVH node popped, 36 B positions + 4 B index per
# triangle test
B_AABB, B_ENTER, B_NODE, B_TRI = 24, 64, 32, 40
# algorithmic bytes of one shaded path step: the scene data the reference reads per hit
# (vertex normals 36 B, texture coordinates 24 B, material record 64 B; the grid has no
# textures) or per miss (four float RGB taps of the environment, 48 B)
B_HIT, B_MISS = 36 + 24 + 64, 4 * 12


def load_profile_json(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--paths-in-flight", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate wall time of the CPU-baseline sample")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the large-scene leg (plain traversals, CPU oracle, cache)")
    ap.add_argument("--cache-reference", action="store_true",
                    help="run the radiance-cache comparison with the reference traversal")
    ap.add_argument("--no-cache", action="store_true",
                    help="skip the plain vs radiance-cached comparison on the large scene")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-requests workload")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the fwd+adjoint timing of the inverse-rendering workload")
    ap.add_argument("--no-rmse", action="store_true",
                    help="skip the RMSE of the plain renders against 4096-spp references")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run a device-to-device copy of 1 GiB (counter calibration)")
    return ap.parse_args()


def host_cpu_info():
    """What the host offers: os.cpu_count(), the affinity set of this process, the CPU model
    (lscpu / /proc/cpuinfo) and the thread share the job is given (OMP_NUM_THREADS, which the
    GPU hosts set to the CPU share of one GPU)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    share = os.environ.get("OMP_NUM_THREADS", "")
    return {"os_cpu_count": os.cpu_count(), "affinity": affinity, "model": model,
            "omp_num_threads": int(share) if share.isdigit() else None}


def cpu_threads():
    """Worker threads of the CPU baseline: every CPU of the affinity set, capped by the job's
    thread share when the host states one (OMP_NUM_THREADS): a GPU host shares its cores
    between the jobs of its GPUs, and one GPU's job is given that many."""
    info = host_cpu_info()
    n = info["affinity"]
    if info["omp_num_threads"]:
        n = min(n, info["omp_num_threads"])
    return max(1, n)


def cpu_baseline(scene, settings_fn, seconds):
    """The oracle (reference release flags, mt19937 per 32x32 tile, faithful per-ray inverse)
    on full 1080p frames at a reduced spp; per-sample cost does not depend on spp."""
    from oracle import pyoracle as po
    o = po.OracleScene(scene, fast=True)
    threads = cpu_threads()
    total_s, total_n, total_r, spp_done = 0.0, 0, 0, 0
    spp = 1
    while True:
        st = settings_fn(samples=spp, seed=1000 + spp_done)
        _, _, w, secs = o.render(st, threads=threads, rng=po.ORC_RNG_MT19937, faithful_inverse=True)
        total_s += secs
        total_n += w["samples"]
        total_r += w["rays"]
        spp_done += spp
        if total_s >= seconds * 0.5 or spp_done >= 64:
            break
        spp = max(1, min(64 - spp_done, int((seconds - total_s) / max(secs / spp, 1e-3))))
    o.close()
    return {"value": total_n / total_s / 1e6, "unit": "Msamples/s", "cores": threads,
            "kind": "port", "host": host_cpu_info(),
            "sample": f"1920x1080 at {spp_done} spp ({total_n} camera paths, {total_r} rays) in "
                      f"{total_s:.1f} s; oracle with -Ofast -flto -march=x86-64-v3, mt19937 per "
                      f"32x32 tile, {threads} threads",
            "mrays_per_s": total_r / total_s / 1e6}


def converged(g, settings_fn, spp, ref_frames, seed0):
    """RMSE of one `spp`-sample render against the mean of `ref_frames` renders of `spp` samples
    with other seeds (a reference of spp * ref_frames samples per pixel; the reference averages
    tone-mapped samples, as every render does, so both estimate the same per-pixel mean), and
    the share of pixel channels within 3 sigma of it (sigma of an `spp`-sample mean, from the
    spread of the reference's frames)"""
    import numpy as np
    img, _, _ = g.render(settings_fn(samples=spp, seed=seed0))
    acc = np.zeros(img.shape, np.float64)
    acc2 = np.zeros(img.shape, np.float64)
    for i in range(ref_frames):
        f, _, _ = g.render(settings_fn(samples=spp, seed=seed0 + 1 + i))
        acc += f
        acc2 += f.astype(np.float64) ** 2
    ref = acc / ref_frames
    var = np.maximum(acc2 / ref_frames - ref ** 2, 0.0) * ref_frames / (ref_frames - 1)
    sigma = np.sqrt(var * (1.0 + 1.0 / ref_frames))
    d = img - ref
    return {"spp": spp, "reference_spp": spp * ref_frames, "rmse": float(np.sqrt(np.mean(d ** 2))),
            "mean": float(ref.mean()), "rmse_over_mean": float(np.sqrt(np.mean(d ** 2)) / ref.mean()),
            "within_3_sigma": float(np.mean(np.abs(d) <= 3.0 * sigma + 1e-7))}


def copy_calibration(nbytes=1 << 30):
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def adjoint_timing(reps=5):
    """fwd+adjoint ms per iteration of the inverse-rendering workload: the sphere grid with the
    shared 512x512 float texture, one 512x512 pose at 32 spp (8,388,608 paths), every
    parameter differentiated.  Device buffers come from the HIP runtime the library uses."""
    import ctypes as C
    import numpy as np
    from mitrace_amd import abi, scenegen
    from mitrace_amd._lib import check
    from mitrace_amd.tracer import GpuScene
    hip = C.CDLL("libamdhip64.so.7")
    sc, p = scenegen.inverse_scene()
    st = scenegen.inverse_settings()
    g = GpuScene(sc)
    L = g.L
    bufs = []

    def dev(a):
        a = np.ascontiguousarray(a, np.float32)
        ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(ptr), C.c_size_t(max(a.nbytes, 4))) == 0
        assert hip.hipMemcpy(ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0
        bufs.append(ptr)
        return ptr.value

    prm = abi.mt_diff_params(dev(p["tex"]), 512, 512, 0, 49, dev(p["rough"]), dev(p["metal"]),
                             dev(p["emis"]), len(sc.materials))
    npix = st.width * st.height
    img = dev(np.zeros(npix * 3))
    gimg = dev(np.random.default_rng(0).random(npix * 3))
    grads = abi.mt_diff_grads(dev(np.zeros(p["tex"].size)), dev(np.zeros(len(sc.materials))),
                              dev(np.zeros(len(sc.materials))), dev(np.zeros(len(sc.materials) * 3)))
    h = C.c_void_p()
    check(L.mt_diff_create(g.h, C.byref(st.to_desc()), C.byref(h)))
    cam = scenegen.inverse_poses(1)[0].to_desc()
    fw, bw = [], []
    for i in range(reps + 1):
        t0 = time.perf_counter()
        check(L.mt_diff_forward(h, C.byref(cam), C.byref(prm), None, 100 + i, img, 0))
        t1 = time.perf_counter()
        check(L.mt_diff_backward(h, C.byref(cam), C.byref(prm), 100 + i, gimg, C.byref(grads), 0))
        t2 = time.perf_counter()
        if i:  # the first pair is a warm-up
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    L.mt_diff_destroy(h)
    for b in bufs:
        hip.hipFree(b)
    g.close()
    f, b = 1e3 * float(np.median(fw)), 1e3 * float(np.median(bw))
    return {"workload": "inverse-rendering sphere grid (512x512 float texture, 4 point lights), "
                        "512x512 at 32 spp, one pose", "fwd_ms": f, "adjoint_ms": b,
            "fwd_plus_adjoint_ms": f + b, "ratio_to_fwd": (f + b) / f}


def batch_leg(rate_hz=16.0):
    """The batched-requests workload: 64 seeded jobs over 8 scene instances (3 Cornell boxes,
    3 sphere grids, 2 terrain scenes at 512x512 terrain vertices and 1000 icospheres),
    128^2-1024^2 pixels, 16-256 spp, depth 8-16, Poisson arrivals at rate_hz, served by one
    running batch whose options come from the autotuner (calibrated on one 512x512 64-spp
    job per scene: 16.8 M paths, about the workload's mean job of 18 M samples).  The main
    figures are the persistent mode's (one kernel launch serves the workload); the graph
    mode's, with the same options otherwise, are reported under "graphs"."""
    from mitrace_amd import batch as B
    from mitrace_amd.tracer import GpuScene
    t0 = time.perf_counter()
    scenes = B.workload_scenes()
    gscenes = [GpuScene(sc) for sc, _ in scenes]
    fns = [fn for _, fn in scenes]
    build_s = time.perf_counter() - t0
    calib = []
    for i, g in enumerate(gscenes):
        job = {"width": 512, "height": 512, "spp": 64, "bounces": 8, "seed": 7 + i,
               "pose_seed": 11 + i}
        st, cam = B.job_settings(job, fns[i], g.scene)
        calib.append((g, st, cam))
    t1 = time.perf_counter()
    best, trials = B.autotune(calib)
    tune_s = time.perf_counter() - t1
    print(f"# batch: scenes built in {build_s:.1f} s, autotuned in {tune_s:.1f} s: {best}",
          file=sys.stderr, flush=True)
    jobs = B.workload_jobs(rate_hz=rate_hz)
    # the workload served both ways with the tuned shape: one persistent launch (the main
    # figures) and a captured graph per iteration (beside it)
    # ("streams" is the graph mode's concurrent streams and the persistent mode's lane groups:
    # a count tuned for the other mode is not carried over, 0 = that mode's default)
    modes = {}
    for name, mode in (("graphs", B.GRAPHS), ("persistent", B.PERSISTENT)):
        opts = dict(best, mode=mode)
        if best["mode"] != mode:
            opts["streams"] = 0
        r, _ = B.run_stream(gscenes, fns, jobs, options=opts)
        modes[name] = r
        print(f"# batch ({name}): {r['msamples_per_s']:.1f} Msamples/s, p50 "
              f"{r['completion_ms_p50']:.0f} ms, p99 {r['completion_ms_p99']:.0f} ms, "
              f"launches {r['stats']['launches']}, iterations {r['stats']['iterations']}",
              file=sys.stderr, flush=True)
    rep = dict(modes["persistent"])
    rep["graphs"] = {k: modes["graphs"][k] for k in ("msamples_per_s", "seconds", "completion_ms_p50",
                                                      "completion_ms_p99", "stats")}
    # the same jobs one after another through mt_render (no arrivals: pure device time)
    t2 = time.perf_counter()
    for j in jobs:
        st, cam = B.job_settings(j, fns[j["scene"]], gscenes[j["scene"]].scene)
        gscenes[j["scene"]].render(st, camera=cam)
    seq_s = time.perf_counter() - t2
    rep["sequential_mt_render"] = {"seconds": seq_s, "msamples_per_s": rep["samples"] / seq_s / 1e6}
    for g in gscenes:
        g.close()
    rep.update({"workload": "64 jobs over 8 scene instances, 128^2-1024^2 px, 16-256 spp, "
                            "depth 8-16", "arrival_rate_hz": rate_hz,
                "autotuned": best, "autotune_trials": trials, "autotune_s": tune_s})
    return rep


def inverse_loop_timing(iters=8):
    """ms per iteration of the inverse-rendering loop as it runs (8 poses at 512x512, each a
    32-spp forward for the residual plus a 32-spp forward and full adjoint through the torch
    autograd wrapper, then the Adam step), median over `iters` iterations.  It runs in a child
    process (the loop uses torch; this script does not), with 64-spp targets and no image
    evaluation, which are outside the timed iterations."""
    import subprocess
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "inverse.json")
        r = subprocess.run([sys.executable, "-m", "mitrace_amd.inverse", "optimize", "--iters", str(iters),
                            "--tex-lr", "1e-3", "--target-spp", "64", "--eval-spp", "0", "--json", out],
                           cwd=os.path.dirname(os.path.abspath(__file__)), capture_output=True, text=True,
                           timeout=900)
        if r.returncode != 0:
            raise RuntimeError("inverse-rendering loop failed: " + r.stderr[-2000:])
        with open(out) as f:
            rep = json.load(f)
    return {"workload": "inverse rendering: 8 poses at 512x512, 32 spp, residual forward + forward "
                        "+ full adjoint per pose, Adam", "iters_timed": iters,
            "ms_per_iter": rep["ms_per_iter"], "poses": rep["poses"]}


def large_cpu(sc, settings_fn, seconds, threads, tlas):
    """The oracle (release flags, mt19937 per 32x32 tile, faithful per-ray inverse) on a band of
    rows of the 1080p frame at 1 spp, grown until about `seconds` of CPU time; with `tlas` the
    instances are culled through a tree over their world boxes (same results)."""
    from oracle import pyoracle as po
    o = po.OracleScene(sc, fast=True)
    rows, r0, total_s, total_n, total_r = 16, 532, 0.0, 0, 0
    while True:
        st = settings_fn(samples=1, seed=2000 + r0)
        _, _, w, secs = o.render(st, threads=threads, rng=po.ORC_RNG_MT19937,
                                 faithful_inverse=True, rows=(r0, r0 + rows), tlas=tlas)
        total_s += secs
        total_n += w["samples"]
        total_r += w["rays"]
        r0 += rows
        if total_s >= seconds * 0.5 or r0 + rows > 1080:
            break
        rows = max(1, min(1080 - r0, int(rows * (seconds - total_s) / max(secs, 1e-3))))
    o.close()
    return {"value": total_n / total_s / 1e6, "unit": "Msamples/s", "cores": threads,
            "kind": "port", "tlas": bool(tlas), "mrays_per_s": total_r / total_s / 1e6,
            "sample": f"rows 532-{r0} of the 1920x1080 frame at 1 spp ({total_n} camera paths, "
                      f"{total_r} rays) in {total_s:.1f} s; oracle with -Ofast -flto "
                      f"-march=x86-64-v3, mt19937 per 32x32 tile, {threads} threads"
                      + ("; instances culled by a tree over their world boxes" if tlas else
                         "; the reference's linear loop over all 5,001 instances")}


def large_leg(args):
    """The large terrain scene (2048x2048-vertex fBm terrain, 5000 icospheres, 16 point
    lights, four 2048^2 textures, sky) at 1920x1080: the plain tracer with the reference
    traversal and with the fast traversal, the CPU oracle beside them (linear instance loop and
    instance tree), and the radiance-cache comparison (64 spp; RMSE ladders against a 4096-spp
    plain render at 320x180; the speed-up at the plain render's RMSE)."""
    from mitrace_amd import abi, cache_eval, scenegen
    from mitrace_amd.tracer import GpuScene
    t0 = time.perf_counter()
    sc = scenegen.large_scene()
    sc.bvh.fast = True
    g = GpuScene(sc)
    build_s = time.perf_counter() - t0
    print(f"# large: scene generated and built (both BVHs) in {build_s:.1f} s", file=sys.stderr, flush=True)
    out = {"workload": "large terrain scene (8.4M-triangle terrain, 5000 icospheres, 16 lights), "
                       "1920x1080", "scene_build_s": build_s}
    for name, trav, spp in (("gpu_reference_traversal", abi.MT_TRAVERSAL_REFERENCE, 8),
                            ("gpu_fast_traversal", abi.MT_TRAVERSAL_FAST, 64)):
        g.render(scenegen.large_scene_settings(samples=1, seed=4, traversal=trav))
        _, _, cnt = g.render(scenegen.large_scene_settings(samples=1, seed=5, traversal=trav,
                                                           count_work=True))
        t = time.perf_counter()
        _, _, s = g.render(scenegen.large_scene_settings(samples=spp, seed=6, traversal=trav))
        dt = time.perf_counter() - t
        out[name] = {"spp": spp, "seconds": dt, "msamples_per_s": s["samples"] / dt / 1e6,
                     "mrays_per_s": s["rays"] / dt / 1e6,
                     "per_ray": {k: cnt[k] / cnt["rays"] for k in
                                 ("inst_aabb_tests", "inst_entered", "node_pops", "tri_tests")}}
        print(f"# large: {name} {out[name]['msamples_per_s']:.1f} Msamples/s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        threads = cpu_threads()
        for tlas in (False, True):
            key = "cpu_oracle_instance_tree" if tlas else "cpu_oracle_linear_loop"
            out[key] = large_cpu(sc, scenegen.large_scene_settings, args.cpu_seconds, threads, tlas)
            print(f"# large: {key} {out[key]['value']:.3f} Msamples/s", file=sys.stderr, flush=True)
        ref = out["gpu_reference_traversal"]["msamples_per_s"]
        fast = out["gpu_fast_traversal"]["msamples_per_s"]
        out["gpu_over_cpu"] = {
            "reference_traversal_vs_linear_loop": ref / out["cpu_oracle_linear_loop"]["value"],
            "reference_traversal_vs_instance_tree": ref / out["cpu_oracle_instance_tree"]["value"],
            "fast_traversal_vs_instance_tree": fast / out["cpu_oracle_instance_tree"]["value"]}
    cache = None
    if not args.no_cache:
        # against the fastest plain form (fast traversal) unless asked otherwise; two cache
        # configurations: after-light queries at the first bounce's hit and at the camera hit
        trav = abi.MT_TRAVERSAL_REFERENCE if args.cache_reference else abi.MT_TRAVERSAL_FAST

        def fn(**kw):
            kw.setdefault("traversal", trav)
            return scenegen.large_scene_settings(**kw)
        ref = cache_eval.reference(g, fn, 4096)
        cache = {"workload": out["workload"] + ", 64 spp, linear radiance",
                 "traversal": "reference" if args.cache_reference else "fast",
                 "reference": "plain render at 4096 spp, 1920x1080 (%.1f s)" % ref[1], "configs": []}
        # the camera ray's hit queries the cache after its light sample; the training paths
        # estimate the rest of the path with environment samples (same expectation, far less
        # variance) and the network learns it with the mean-keeping L2 loss; beside it the
        # same query trained on the reference estimator's own paths with the relative loss
        for opts in (cache_eval.LARGE_SCENE_CACHE, {"query_depth": 0, "query_after_light": 1}):
            cache["configs"].append(cache_eval.compare(
                g, fn, cache_opts=opts, log=lambda m: print(m, file=sys.stderr, flush=True), ref=ref))
    g.close()
    return out, cache


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "a render would shard by pixel tiles, one GPU per tile set, and this measures "
              "one GPU's share", file=sys.stderr)
        return 2
    from mitrace_amd import scenegen
    from mitrace_amd.tracer import GpuScene

    scene = scenegen.sphere_grid(seed=1)
    scene.bvh.fast = True  # also the fast traversal's BVHs (the timed frames use the reference ones)

    def settings(**kw):
        kw.setdefault("samples", args.spp)
        kw.setdefault("paths_in_flight", args.paths_in_flight)
        return scenegen.sphere_grid_settings(**kw)

    g = GpuScene(scene)
    W, H = 1920, 1080
    # work counts of one frame (same seed => the timed frames do exactly this work)
    _, _, cnt = g.render(settings(count_work=True))
    print(f"# counted frame: {cnt['samples']} samples {cnt['rays']} rays", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        g.render(settings())
        print(f"# warmup {i} done", file=sys.stderr, flush=True)
    times, stats = [], []
    for i in range(args.steps):
        t0 = time.perf_counter()
        _, _, s = g.render(settings())  # returns after the device work and the readback
        times.append(time.perf_counter() - t0)
        stats.append(s)
        print(f"# step {i}: {times[-1]:.3f} s", file=sys.stderr, flush=True)
    if args.copy_calibration:
        copy_calibration()
    # one more frame on a single stream: each launch then runs alone, so its HIP-event time is
    # the kernel's exclusive duration (the timed frames overlap three streams)
    os.environ["MT_STREAMS"] = "1"
    _, _, ex = g.render(settings())
    del os.environ["MT_STREAMS"]
    print(f"# exclusive frame (one stream): extend {ex['kernel_ms_extend'] / ex['extend_launches']:.2f} "
          f"ms per launch", file=sys.stderr, flush=True)
    total_t = sum(times)
    samples = sum(s["samples"] for s in stats)
    rays = sum(s["rays"] for s in stats)
    ext_ms = sum(s["kernel_ms_extend"] for s in stats)
    ext_launches = sum(s["extend_launches"] for s in stats)
    bytes_frame = (B_AABB * cnt["inst_aabb_tests"] + B_ENTER * cnt["inst_entered"] +
                   B_NODE * cnt["node_pops"] + B_TRI * cnt["tri_tests"])
    launches_frame = cnt["extend_launches"]
    bytes_per_launch = bytes_frame / launches_frame
    avg_launch_s = ext_ms / 1e3 / ext_launches
    achieved = bytes_per_launch / avg_launch_s / 1e9
    ext_traffic = load_profile_json("extend_traffic_sphere_grid_1080p.json")
    traffic = ext_traffic.get("hbm_bytes_per_launch") if ext_traffic else None
    ex_ext_s = ex["kernel_ms_extend"] / 1e3 / ex["extend_launches"]
    ex_shade_s = ex["kernel_ms_shade"] / 1e3 / ex["extend_launches"]
    # shading: one launch per bounce and chunk, as extend
    shade_bytes_launch = (B_HIT * cnt["hits_shaded"] + B_MISS * cnt["misses"]) / launches_frame
    shade_traffic = load_profile_json("shade_traffic_sphere_grid_1080p.json")
    shade_hbm = shade_traffic.get("hbm_bytes_per_launch") if shade_traffic else None
    line = {
        "metric": METRIC,
        "value": samples / total_t / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total_t / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "material-spheres grid, plain GPU tracer", "width": W, "height": H,
                   "spp": args.spp, "bounces": 5, "triangles": scene.triangle_count,
                   "instances": len(scene.instances), "bvh_max_triangles": 4,
                   "env": "2048x1024 float32 equirect, rotated 30 deg",
                   "samples_per_step": samples // args.steps},
        "mrays_per_s": rays / total_t / 1e6,
        "kernel_ms_per_step": sum(s["kernel_ms_total"] for s in stats) / args.steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_trace<false,0,12u,0> (extend: closest hit, reference traversal)",
                     "level": "algorithmic node/triangle bytes, most served by L2 and MALL "
                              "(hbm: the measured HBM traffic)",
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     # chunks run on concurrent streams, so launches overlap: the whole frame's
                     # extend bytes over the frame time is a lower bound of the aggregate rate
                     "achieved_aggregate_lower_bound": bytes_frame / (total_t / args.steps) / 1e9,
                     "frac_aggregate_lower_bound": bytes_frame / (total_t / args.steps) / 1e9 / HBM_PEAK_GBS,
                     # the same kernel alone on the GPU (single-stream frame)
                     "exclusive": {"avg_launch_ms": ex_ext_s * 1e3,
                                   "achieved": bytes_per_launch / ex_ext_s / 1e9,
                                   "frac": bytes_per_launch / ex_ext_s / 1e9 / HBM_PEAK_GBS},
                     "hbm": None if traffic is None else {
                         "bytes_per_launch": traffic, "achieved": traffic / ex_ext_s / 1e9,
                         "frac": traffic / ex_ext_s / 1e9 / HBM_PEAK_GBS,
                         "l2_hit_rate": ext_traffic.get("l2_hit_rate")},
                     "per_ray": {k: cnt[k] / cnt["rays"] for k in
                                 ("inst_aabb_tests", "inst_entered", "node_pops", "tri_tests")}},
        "cpu_baseline": None,
        # the shading kernel (scenes without point lights), alone on the GPU
        "shade_roofline": {
            "kernel": "k_shade_unlit<false> (shade)", "bound": "hbm", "unit": "GB/s",
            "peak": HBM_PEAK_GBS, "avg_launch_ms": ex_shade_s * 1e3,
            "avg_launch_ms_overlapped": sum(s["kernel_ms_shade"] for s in stats) / ext_launches,
            "algorithmic_bytes_per_launch": shade_bytes_launch,
            "achieved": shade_bytes_launch / ex_shade_s / 1e9,
            "frac": shade_bytes_launch / ex_shade_s / 1e9 / HBM_PEAK_GBS,
            "hbm": None if shade_hbm is None else {
                "bytes_per_launch": shade_hbm, "achieved": shade_hbm / ex_shade_s / 1e9,
                "frac": shade_hbm / ex_shade_s / 1e9 / HBM_PEAK_GBS,
                "l2_hit_rate": shade_traffic.get("l2_hit_rate")},
            "per_launch": {"hits": cnt["hits_shaded"] / launches_frame,
                           "misses": cnt["misses"] / launches_frame}},
    }
    # the same frame with the fast traversal (true closest hits; not the headline value)
    from mitrace_amd import abi
    g.render(settings(traversal=abi.MT_TRAVERSAL_FAST))
    _, _, fc = g.render(settings(traversal=abi.MT_TRAVERSAL_FAST, count_work=True))
    ft = []
    for i in range(max(1, args.steps)):
        t0 = time.perf_counter()
        _, _, fs = g.render(settings(traversal=abi.MT_TRAVERSAL_FAST))
        ft.append(time.perf_counter() - t0)
    line["fast_traversal"] = {
        "msamples_per_s": fs["samples"] * len(ft) / sum(ft) / 1e6,
        "mrays_per_s": fs["rays"] * len(ft) / sum(ft) / 1e6,
        "per_ray": {k: fc[k] / fc["rays"] for k in
                    ("inst_aabb_tests", "inst_entered", "node_pops", "tri_tests")}}
    print(f"# fast traversal: {line['fast_traversal']['msamples_per_s']:.1f} Msamples/s",
          file=sys.stderr, flush=True)
    if not args.no_rmse:
        # the plain tracer's image error at the configurations' sample counts, against
        # 4096-spp references of the same estimator (the oracle's samples are these bit for bit,
        # so its RMSE at equal spp is the same)
        from mitrace_amd import abi as A
        line["rmse_vs_4096spp"] = {"sphere_grid_1080p": converged(g, settings, args.spp, 4096 // args.spp, 7000)}
        cg = GpuScene(scenegen.cornell())
        line["rmse_vs_4096spp"]["cornell_256_rr"] = converged(
            cg, lambda **kw: scenegen.cornell_settings(mode=A.MT_MODE_FIXED, rr_depth=3, bounces=8,
                                                       ev_exposure=2.0, **kw),
            64, 64, 7100)
        cg.close()
        print(f"# rmse: {line['rmse_vs_4096spp']}", file=sys.stderr, flush=True)
    if not args.no_large:
        line["large"], cache = large_leg(args)
        if cache is not None:
            line["cache"] = cache
    if not args.no_batch:
        line["batch"] = batch_leg()
        print(f"# batch: {line['batch']['msamples_per_s']:.1f} Msamples/s, p50 "
              f"{line['batch']['completion_ms_p50']:.0f} ms, p99 "
              f"{line['batch']['completion_ms_p99']:.0f} ms", file=sys.stderr, flush=True)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_timing()
        line["adjoint"]["inverse_loop"] = inverse_loop_timing()
        print(f"# adjoint: {line['adjoint']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(scene, settings, args.cpu_seconds)
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = line["value"] / cb["value"]
    g.close()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
