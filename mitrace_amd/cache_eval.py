"""Plain vs radiance-cached rendering of one scene at equal RMSE.

RMSE is measured against a high-spp plain render of the same view at a reduced resolution
(the per-pixel error depends on the samples per pixel, not on the image size); throughput is
measured at the full resolution, wall clock around each render call (the cached one includes
its warm-up passes and all training).  The samples the cached form needs to reach the plain
form's RMSE come from log-log interpolation of its RMSE ladder; the speed-up at equal RMSE is
the plain render's time at `spp` over the cached render's time at that many samples (rounded
up), both at full resolution.
"""
import time

import numpy as np

from . import abi
from .cache import RadianceCache

DEFAULT_CACHE = {"train_every": 16, "query_depth": 2, "warmup_passes": 4, "steps_per_pass": 8,
                 "train_batch": 16384, "learning_rate": 5e-3, "spp_per_pass": 4, "seed": 1}


def rmse(a, b):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - b) ** 2)))


def _spp_for(rmse_target, ladder):
    """samples needed to reach rmse_target by log-log interpolation of [(spp, rmse)]"""
    s = np.log([x[0] for x in ladder])
    r = np.log([x[1] for x in ladder])
    t = np.log(rmse_target)
    if t >= r[0]:  # already below the target at the smallest count measured: at most that
        return float(ladder[0][0])
    for i in range(len(s) - 1):
        if r[i + 1] <= t <= r[i]:
            f = (t - r[i]) / (r[i + 1] - r[i])
            return float(np.exp(s[i] + f * (s[i + 1] - s[i])))
    return None  # never reached: the cache's bias is above the target


def compare(gscene, settings_fn, spp=64, small=(320, 180), ref_spp=4096, ladder=(4, 8, 16, 32, 64),
            cache_opts=None, log=print):
    opts = dict(DEFAULT_CACHE, **(cache_opts or {}))
    out = {"cache_options": opts, "spp": spp, "output": "linear"}
    # linear radiance on both sides: the reference tone-maps every sample before averaging
    # (Tracer.cpp:212-227), so any estimator with less variance would change the expected tone-mapped image
    user_fn = settings_fn

    def settings_fn(**kw):
        kw.setdefault("output", abi.MT_OUTPUT_LINEAR)
        return user_fn(**kw)
    # throughput at full resolution
    st = settings_fn(samples=spp, seed=5)
    gscene.render(settings_fn(samples=1, seed=4))  # first-touch warm-up
    t0 = time.perf_counter()
    _, _, sp = gscene.render(st)
    t_plain = time.perf_counter() - t0
    cache = RadianceCache(gscene, opts)
    t0 = time.perf_counter()
    _, sc = cache.render(st)
    t_cached = time.perf_counter() - t0
    cache.close()
    n = st.width * st.height * spp
    out["plain"] = {"seconds": t_plain, "msamples_per_s": n / t_plain / 1e6,
                    "mrays_per_s": sp["rays"] / t_plain / 1e6}
    out["cached"] = {"seconds": t_cached, "msamples_per_s": n / t_cached / 1e6,
                     "mrays_per_s": sc["rays"] / t_cached / 1e6, "trained_samples": int(sc["diag"][8])}
    log(f"# cache: 1080p plain {out['plain']['msamples_per_s']:.1f} Msamples/s, cached "
        f"{out['cached']['msamples_per_s']:.1f} Msamples/s")
    # RMSE ladders at reduced resolution
    W, H = small
    ref, _, _ = gscene.render(settings_fn(width=W, height=H, samples=ref_spp, seed=999))
    lad_p, lad_c = [], []
    for s in ladder:
        img, _, _ = gscene.render(settings_fn(width=W, height=H, samples=s, seed=17))
        lad_p.append((s, rmse(img, ref)))
        cache = RadianceCache(gscene, opts)
        img_c, _ = cache.render(settings_fn(width=W, height=H, samples=s, seed=17))
        cache.close()
        lad_c.append((s, rmse(img_c, ref)))
        if s == ladder[-1]:
            out["cached_bias"] = float(np.mean(img_c.astype(np.float64) - ref))
            out["plain_bias"] = float(np.mean(img.astype(np.float64) - ref))
    out["rmse_plain"] = lad_p
    out["rmse_cached"] = lad_c
    ref_mean = float(np.mean(ref.astype(np.float64)))
    out["reference"] = {"width": W, "height": H, "spp": ref_spp, "mean": ref_mean}
    out["cached_bias_relative"] = out["cached_bias"] / ref_mean if ref_mean else None
    target = lad_p[-1][1]
    need = _spp_for(target, lad_c)
    out["equal_rmse"] = {"rmse": target, "plain_spp": ladder[-1], "cached_spp": need,
                         "cached_spp_is_upper_bound": bool(need is not None and need <= ladder[0]
                                                           and lad_c[0][1] <= target)}
    if need:
        # the cached render at full resolution with the samples it needs for that RMSE
        # (warm-up and training included), against the plain render at `spp`
        per = opts["spp_per_pass"]
        k = max(per, int(np.ceil(need / per)) * per)
        cache = RadianceCache(gscene, opts)
        t0 = time.perf_counter()
        cache.render(settings_fn(samples=k, seed=5))
        t_c = time.perf_counter() - t0
        cache.close()
        out["equal_rmse"].update({"cached_spp_rendered": k, "cached_seconds": t_c,
                                  "plain_seconds": t_plain, "speedup": t_plain / t_c})
    else:
        out["equal_rmse"]["speedup"] = None
    log(f"# cache: RMSE plain {lad_p} cached {lad_c} -> {out['equal_rmse']}")
    return out


def mlp_bench(gscene, n_infer=1 << 22, n_train=1 << 16, reps=10, log=print):
    """the cache's kernels alone: reps inferences of n_infer queries and reps training steps of
    n_train samples on random inputs (wall clock includes the host copies of the test hooks;
    the kernels' own durations come from a kernel trace of this call)"""
    rng = np.random.default_rng(0)
    cache = RadianceCache(gscene, {"seed": 1, "train_batch": n_train})
    pos = rng.random((n_infer, 3)).astype(np.float32) * 10
    d = rng.normal(size=(n_infer, 3)).astype(np.float32)
    tgt = rng.random((n_train, 3)).astype(np.float32)
    cache.infer(pos[:1024], d[:1024])
    t0 = time.perf_counter()
    for _ in range(reps):
        cache.infer(pos, d)
    t_inf = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        cache.train(pos[:n_train], d[:n_train], tgt)
    t_tr = (time.perf_counter() - t0) / reps
    cache.close()
    flop_q = 2 * (32 * 64 + 3 * 64 * 64 + 64 * 3)
    out = {"infer_queries": n_infer, "infer_s": t_inf, "train_samples": n_train, "train_s": t_tr,
           "flop_per_query": flop_q}
    log(f"# mlp: infer {n_infer} in {t_inf * 1e3:.1f} ms (with host copies), train step "
        f"{n_train} in {t_tr * 1e3:.1f} ms")
    return out


if __name__ == "__main__":
    import json
    import sys
    from . import scenegen
    from .tracer import GpuScene
    g = GpuScene(scenegen.cornell())
    print(json.dumps(mlp_bench(g, reps=int(sys.argv[1]) if len(sys.argv) > 1 else 10)))
