"""Plain vs radiance-cached rendering of one scene at equal RMSE.

RMSE and bias are measured against a high-spp plain render of the same view at the same
resolution (the cache's training depends on the number of paths per pass, so the ladder is not
taken at a reduced size); time is the wall clock around each render call (the cached one
includes its warm-up passes and all training).  The samples the cached form needs to reach the
plain form's RMSE come from log-log interpolation of its RMSE ladder; the speed-up at equal
RMSE is the plain render's time at `spp` over the cached render's time at that many samples
(whole passes).
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


def reference(gscene, settings_fn, ref_spp=4096):
    """the plain render that judges the cached form: (linear image as float64, seconds)"""
    t0 = time.perf_counter()
    ref, _, _ = gscene.render(settings_fn(samples=ref_spp, seed=999, output=abi.MT_OUTPUT_LINEAR))
    return ref.astype(np.float64), time.perf_counter() - t0


def compare(gscene, settings_fn, spp=64, ref_spp=4096, ladder=(4, 8, 16, 32, 64), cache_opts=None,
            log=print, ref=None):
    """Plain and cached renders at the full resolution of settings_fn for every sample count of
    the ladder (each cached render with a fresh cache: warm-up and training included in its
    time), their RMSE and relative bias against a ref_spp plain render of the same view, and
    the speed-up at the plain render's RMSE at `spp`."""
    opts = dict(DEFAULT_CACHE, **(cache_opts or {}))
    out = {"cache_options": opts, "spp": spp, "output": "linear"}
    # linear radiance on both sides: the reference tone-maps every sample before averaging
    # (Tracer.cpp:212-227), so any estimator with less variance would change the expected tone-mapped image
    user_fn = settings_fn

    def settings_fn(**kw):
        kw.setdefault("output", abi.MT_OUTPUT_LINEAR)
        return user_fn(**kw)
    gscene.render(settings_fn(samples=1, seed=4))  # first-touch warm-up
    ref, t_ref = ref if ref is not None else reference(gscene, user_fn, ref_spp)
    ref_mean = float(np.mean(ref))
    st0 = settings_fn(samples=1, seed=0)
    out["reference"] = {"width": st0.width, "height": st0.height, "spp": ref_spp, "mean": ref_mean,
                        "seconds": t_ref}
    lad_p, lad_c, bias_c, t_p, t_c = [], [], [], {}, {}
    for s in ladder:
        t0 = time.perf_counter()
        img, _, sp = gscene.render(settings_fn(samples=s, seed=17))
        t_p[s] = time.perf_counter() - t0
        lad_p.append((s, rmse(img, ref)))
        cache = RadianceCache(gscene, opts)
        t0 = time.perf_counter()
        img_c, sc = cache.render(settings_fn(samples=s, seed=17))
        t_c[s] = time.perf_counter() - t0
        cache.close()
        lad_c.append((s, rmse(img_c, ref)))
        bias_c.append((s, float(np.mean(img_c, dtype=np.float64)) / ref_mean - 1.0 if ref_mean else None))
        if s == spp:
            n = st0.width * st0.height * spp
            out["plain"] = {"seconds": t_p[s], "msamples_per_s": n / t_p[s] / 1e6,
                            "mrays_per_s": sp["rays"] / t_p[s] / 1e6}
            out["cached"] = {"seconds": t_c[s], "msamples_per_s": n / t_c[s] / 1e6,
                             "mrays_per_s": sc["rays"] / t_c[s] / 1e6,
                             "trained_samples": int(sc["diag"][8])}
            out["plain_bias_relative"] = float(np.mean(img, dtype=np.float64)) / ref_mean - 1.0
            out["cached_bias_relative"] = bias_c[-1][1]
        log(f"# cache: {s} spp: plain {t_p[s]:.3f} s RMSE {lad_p[-1][1]:.1f}, cached {t_c[s]:.3f} s "
            f"RMSE {lad_c[-1][1]:.1f} bias {bias_c[-1][1]:+.4f}")
    out["rmse_plain"] = lad_p
    out["rmse_cached"] = lad_c
    out["cached_bias_relative_by_spp"] = bias_c
    out["seconds_plain_by_spp"] = [(s, t_p[s]) for s in ladder]
    out["seconds_cached_by_spp"] = [(s, t_c[s]) for s in ladder]
    target = dict(lad_p)[spp]
    need = _spp_for(target, lad_c)
    out["equal_rmse"] = {"rmse": target, "plain_spp": spp, "cached_spp": need,
                         "cached_spp_is_upper_bound": bool(need is not None and need <= ladder[0]
                                                           and lad_c[0][1] <= target)}
    if need:
        # the cached render with the samples it needs for that RMSE (whole passes), against the
        # plain render at `spp`
        per = opts["spp_per_pass"]
        k = max(per, int(np.ceil(need / per - 1e-9)) * per)
        if k in t_c:
            tk = t_c[k]
            bk = dict(bias_c)[k]
        else:
            cache = RadianceCache(gscene, opts)
            t0 = time.perf_counter()
            img_k, _ = cache.render(settings_fn(samples=k, seed=17))
            tk = time.perf_counter() - t0
            cache.close()
            bk = float(np.mean(img_k, dtype=np.float64)) / ref_mean - 1.0
        out["equal_rmse"].update({"cached_spp_rendered": k, "cached_seconds": tk,
                                  "cached_bias_relative": bk, "plain_seconds": t_p[spp],
                                  "speedup": t_p[spp] / tk})
    else:
        out["equal_rmse"]["speedup"] = None
    log(f"# cache: -> {out['equal_rmse']}")
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


# the bench's configuration for the large scene: the camera ray's hit queries the cache after
# its light sample, training paths with environment samples, L2 loss
LARGE_SCENE_CACHE = {"query_depth": 0, "query_after_light": 1, "train_nee": 1, "loss": 1,
                     "learning_rate": 2e-3, "warmup_passes": 4, "steps_per_pass": 16}


def render_large(spp=16, log=print):
    """one plain and one cached render of the large scene at 1080p (fast traversal, linear
    radiance, the bench's cache configuration): the workload of the cache's kernel trace"""
    from . import scenegen
    from .tracer import GpuScene
    sc = scenegen.large_scene()
    sc.bvh.fast = True
    g = GpuScene(sc)

    def st(**kw):
        return scenegen.large_scene_settings(traversal=abi.MT_TRAVERSAL_FAST, output=abi.MT_OUTPUT_LINEAR, **kw)
    g.render(st(samples=1, seed=4))
    t0 = time.perf_counter()
    _, _, sp = g.render(st(samples=spp, seed=5))
    t_plain = time.perf_counter() - t0
    cache = RadianceCache(g, dict(DEFAULT_CACHE, **LARGE_SCENE_CACHE))
    t0 = time.perf_counter()
    _, sc_ = cache.render(st(samples=spp, seed=5))
    t_cached = time.perf_counter() - t0
    cache.close()
    g.close()
    log(f"# large scene 1080p {spp} spp: plain {t_plain:.3f} s ({sp['rays']} rays), cached "
        f"{t_cached:.3f} s ({sc_['rays']} rays, {int(sc_['diag'][8])} trained samples)")


if __name__ == "__main__":
    import json
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "large":
        render_large(int(sys.argv[2]) if len(sys.argv) > 2 else 16)
    else:
        from . import scenegen
        from .tracer import GpuScene
        g = GpuScene(scenegen.cornell())
        print(json.dumps(mlp_bench(g, reps=int(sys.argv[1]) if len(sys.argv) > 1 else 10)))
