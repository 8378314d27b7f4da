"""Batched rendering over the C-ABI (mt_batch_*): many render jobs of uneven size, sample count,
depth and scene served either by one persistent kernel launch fed from a pinned queue
(options["mode"] = PERSISTENT, the default) or by one device work queue per scene and a captured
iteration graph per scene (GRAPHS), plus the autotuner of the iteration's shape.

    b = Batch(options)            # options: dict or None (defaults)
    j = b.submit(gpu_scene, settings, camera)
    b.start(); ...; b.finish()    # submit() may also be called while the batch runs
    img = b.image(j); info = b.info(j)
"""
import ctypes as C
import math
import time

import numpy as np

from . import abi, scenegen
from ._lib import check, load

OPTION_NAMES = ("paths_per_iteration", "tile_pixels", "iteration_tiles", "trace_blocks_per_cu",
                "streams", "mode")
PERSISTENT, GRAPHS = 1, 2  # MT_BATCH_PERSISTENT, MT_BATCH_GRAPHS


def _options(d):
    o = abi.mt_batch_options()
    for k, v in (d or {}).items():
        setattr(o, k, int(v))
    return o


class Batch:
    def __init__(self, options=None):
        self.L = load()
        h = C.c_void_p()
        opt = _options(options)
        check(self.L.mt_batch_create(C.byref(opt), C.byref(h)))
        self.h = h
        self._images = {}
        self._keep = []

    def close(self):
        if self.h:
            self.L.mt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, gscene, settings, camera=None, want_image=True):
        cam = (camera or gscene.scene.camera).to_desc()
        st = settings.to_desc()
        img = np.zeros((settings.height, settings.width, 3), np.float32) if want_image else None
        jid = C.c_uint32()
        check(self.L.mt_batch_submit(self.h, gscene.h, C.byref(cam), C.byref(st),
                                     abi.fptr(img) if want_image else None, C.byref(jid)))
        self._images[jid.value] = img
        self._keep.append(gscene)
        return jid.value

    def start(self):
        check(self.L.mt_batch_start(self.h))

    def finish(self):
        check(self.L.mt_batch_finish(self.h))

    def image(self, job):
        return self._images[job]

    def info(self, job):
        i = abi.mt_batch_job_info()
        check(self.L.mt_batch_get_job_info(self.h, job, C.byref(i)))
        return {"done": bool(i.done), "status": i.status, "tiles": i.tiles, "samples": i.samples,
                "submit_ms": i.submit_ms, "complete_ms": i.complete_ms}

    def stats(self):
        s = abi.mt_batch_stats()
        check(self.L.mt_batch_get_stats(self.h, C.byref(s)))
        return {"iterations": s.iterations, "graphs_captured": s.graphs_captured,
                "launches": s.launches, "samples": s.samples, "jobs": s.jobs, "scenes": s.scenes,
                "options": s.options.as_dict()}


def autotune(jobs, max_trials=32):
    """jobs: [(gpu_scene, settings, camera or None)] calibration set.  Each trial serves the
    jobs once at 1 spp (untimed: graph captures, first touches), then times them as given,
    all submitted at once.  Returns (best options dict, trials [{options, seconds,
    msamples_per_s}])."""
    L = load()
    n = len(jobs)
    scenes = (C.c_void_p * n)(*[j[0].h.value for j in jobs])
    cams = (abi.mt_camera_desc * n)(*[(j[2] or j[0].scene.camera).to_desc() for j in jobs])
    sets = (abi.mt_settings * n)(*[j[1].to_desc() for j in jobs])
    best = abi.mt_batch_options()
    trials = (abi.mt_autotune_trial * max_trials)()
    nt = C.c_uint32()
    check(L.mt_batch_autotune(scenes, cams, sets, n, C.byref(best), trials, max_trials,
                              C.byref(nt)))
    return best.as_dict(), [{"options": t.options.as_dict(), "seconds": t.seconds,
                             "msamples_per_s": t.msamples_per_s} for t in trials[:nt.value]]


# ------------------------------------------------------------------ the batched-requests workload
def _scene_center(scene):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for inst in scene.instances:
        p = scene.meshes[inst.mesh].positions
        w = (inst.transform[:3, :3] @ p.T).T + inst.transform[:3, 3]
        lo = np.minimum(lo, w.min(0))
        hi = np.maximum(hi, w.max(0))
    return (lo + hi) / 2


def _orbit(scene, rng, yaw_deg=20.0, pitch_deg=8.0):
    """the scene's camera turned by a seeded yaw/pitch about the point of its view axis nearest
    to the scene's centre"""
    from .scene import Camera, look_at
    camera = scene.camera
    M = np.asarray(camera.camera_to_world, np.float64).reshape(4, 4)
    eye = M[:3, 3]
    fwd = -M[:3, 2]
    target = eye + max(0.5, float(np.dot(_scene_center(scene) - eye, fwd))) * fwd
    v = eye - target
    yaw = np.radians(rng.uniform(-yaw_deg, yaw_deg))
    pitch = np.radians(rng.uniform(-pitch_deg, pitch_deg))
    cy, sy = np.cos(yaw), np.sin(yaw)
    v = np.array([cy * v[0] + sy * v[2], v[1], -sy * v[0] + cy * v[2]])
    r = np.linalg.norm(v)
    el = np.arcsin(np.clip(v[1] / r, -0.99, 0.99)) + pitch
    az = np.arctan2(v[2], v[0])
    v = r * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    return Camera(type=camera.type, camera_to_world=look_at(target + v, target), yfov=camera.yfov,
                  aspect=camera.aspect, ortho_size=camera.ortho_size)


def workload_scenes(seed=3, large_terrain=512, large_spheres=1000, large_tex=512):
    """8 scene instances over the three generators: 3 Cornell boxes, 3 sphere grids (seeded base
    colours), 2 terrain scenes (reduced: terrain_res x terrain_res vertices, large_spheres
    icospheres, large_tex textures).  Returns [(scene, settings_fn)]."""
    out = []
    for i in range(3):
        out.append((scenegen.cornell(), scenegen.cornell_settings))
    for s in (1, 5, 6):
        out.append((scenegen.sphere_grid(seed=s), scenegen.sphere_grid_settings))
    for s in (2, 7):
        out.append((scenegen.large_scene(seed=s, terrain_res=large_terrain, n_spheres=large_spheres,
                                         tex_res=large_tex), scenegen.large_scene_settings))
    return out


def workload_jobs(n_scenes=8, n_jobs=64, seed=3, rate_hz=8.0):
    """64 seeded jobs: scene uniform over the instances, pixels log-uniform in [128^2, 1024^2]
    with aspect 1 or 16:9, spp log-uniform in [16, 256], bounces 8..16, a seeded camera pose
    per job, and Poisson arrival times (rate_hz).  Returns [dict]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(n_jobs):
        pix = math.exp(rng.uniform(math.log(128 * 128), math.log(1024 * 1024)))
        aspect = 1.0 if rng.random() < 0.5 else 16.0 / 9.0
        w = max(16, int(round(math.sqrt(pix * aspect))))
        h = max(16, int(round(pix / w)))
        spp = int(round(math.exp(rng.uniform(math.log(16), math.log(256)))))
        jobs.append({"scene": int(rng.integers(n_scenes)), "width": w, "height": h, "spp": spp,
                     "bounces": int(rng.integers(8, 17)), "seed": int(rng.integers(1 << 31)),
                     "pose_seed": int(rng.integers(1 << 31)), "arrival_s": t})
        t += rng.exponential(1.0 / rate_hz)
    return jobs


def job_settings(job, settings_fn, scene):
    st = settings_fn(width=job["width"], height=job["height"], samples=job["spp"],
                     bounces=job["bounces"], seed=job["seed"])
    cam = _orbit(scene, np.random.default_rng(job["pose_seed"]))
    cam.aspect = job["width"] / job["height"]
    return st, cam


def run_stream(gscenes, settings_fns, jobs, options=None, want_images=False):
    """submits the jobs at their arrival times to a running batch; returns the report"""
    b = Batch(options)
    ids = []
    prepared = []
    for j in jobs:
        g = gscenes[j["scene"]]
        st, cam = job_settings(j, settings_fns[j["scene"]], g.scene)
        prepared.append((g, st, cam))
    b.start()
    t0 = time.perf_counter()
    for j, (g, st, cam) in zip(jobs, prepared):
        wait = j["arrival_s"] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ids.append(b.submit(g, st, cam, want_image=want_images))
    b.finish()
    infos = [b.info(i) for i in ids]
    first = min(x["submit_ms"] for x in infos)
    last = max(x["complete_ms"] for x in infos)
    lat = np.array([x["complete_ms"] - x["submit_ms"] for x in infos])
    samples = sum(x["samples"] for x in infos)
    rep = {"jobs": len(jobs), "samples": int(samples),
           "seconds": (last - first) / 1e3,
           "msamples_per_s": samples / ((last - first) / 1e3) / 1e6,
           "completion_ms_p50": float(np.percentile(lat, 50)),
           "completion_ms_p99": float(np.percentile(lat, 99)),
           "stats": b.stats()}
    images = [b.image(i) for i in ids] if want_images else None
    b.close()
    return rep, images
