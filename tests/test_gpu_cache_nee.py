"""Radiance cache with environment-sampled training paths (mt_cache_options.train_nee).

The training paths' estimate of the rest of their path must have the reference estimator's
expectation (Tracer.cpp:134-348), quirks included: the sky term's extra base colour, the
throughput that moves only where a light was chosen, the energy cut-off, and the true density
of the bounce direction (the cosine directions that replace specular draws reflected below the
surface).  With every path training and the test hook MT_CACHE_TRAINING_IMAGE=1, each image
sample is that full estimate; its image must agree with the plain render's within the Monte
Carlo error of both, per block of pixels.  Then the trained cache on a sun-lit terrain: lower
RMSE than the plain render at equal samples, with a small relative bias.
"""
import numpy as np
import pytest

from helpers import load_scene, small_settings
from mitrace_amd import abi

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes(gpu):
    """GpuScene by name; the scenes are built with the fast traversal's BVHs as well"""
    from mitrace_amd.tracer import GpuScene
    out = {}

    def get(name):
        if name not in out:
            if name == "sunlit_terrain":
                # the large scene's generator at a small size (96^2-vertex terrain, 300
                # spheres, 64^2 textures, 16 point lights) under the full 2048x1024 sky with
                # its 0.5 deg sun of radiance 5e4
                from mitrace_amd import scenegen
                sc = scenegen.large_scene(terrain_res=96, n_spheres=300, tex_res=64, env_width=2048,
                                          env_height=1024)
            else:
                sc = load_scene(name)
            sc.bvh.fast = True
            out[name] = GpuScene(sc)
        return out[name]
    yield get
    for g in out.values():
        g.close()


def _blocks(img, n=4):
    h, w, _ = img.shape
    return np.array([img[i * h // n:(i + 1) * h // n, j * w // n:(j + 1) * w // n].mean()
                     for i in range(n) for j in range(n)])


def _replicates(fn, reps):
    """block means and image means of `reps` independent renders"""
    b = np.array([_blocks(fn(r)) for r in range(reps)])
    return b.mean(0), b.std(0, ddof=1) / np.sqrt(reps)


ROT = dict(hdri_rotation=30.0, hdri_rotation_apply=True)
CASES = {
    # no point lights: every training vertex draws an environment sample, and the reference
    # throughput never moves (no light is ever chosen)
    "sphere_grid": ("sphere_grid", dict(bounces=5)),
    # point lights and sky: the shadow ray goes to either, each with weight 2
    "large_small": ("large_small", dict(bounces=5)),
    # Russian roulette instead of the energy cut-off
    "sphere_grid_rr": ("sphere_grid", dict(bounces=8, mode=abi.MT_MODE_FIXED, rr_depth=2)),
    # the rotated sky, and the fast traversal (closest hits, any-hit shadow rays whose
    # environment samples have no light distance)
    "sphere_grid_fast_rotated": ("sphere_grid", dict(bounces=5, traversal=abi.MT_TRAVERSAL_FAST, **ROT)),
    "large_small_fast_rotated": ("large_small", dict(bounces=5, traversal=abi.MT_TRAVERSAL_FAST, **ROT)),
    # the sun: sharp in the environment (its rotation moves it), sampled by the training paths
    "sunlit_terrain_fast_rotated": ("sunlit_terrain", dict(bounces=5, traversal=abi.MT_TRAVERSAL_FAST,
                                                           width=128, height=96, plain_spp=1536, **ROT)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_training_estimate_has_the_reference_expectation(scenes, case, monkeypatch):
    from mitrace_amd.cache import RadianceCache
    name, extra = CASES[case]
    g = scenes(name)
    kw = dict(width=64, height=48, output=abi.MT_OUTPUT_LINEAR)
    kw.update(extra)
    plain_spp = kw.pop("plain_spp", 256)
    reps = 24

    def plain(r):
        img, _, _ = g.render(small_settings(samples=plain_spp, seed=1000 + r, **kw))
        return img

    monkeypatch.setenv("MT_CACHE_TRAINING_IMAGE", "1")

    def nee(r):
        c = RadianceCache(g, {"train_every": 1, "query_depth": 0, "query_after_light": 1,
                              "train_nee": 1, "spp_per_pass": 64, "steps_per_pass": 1,
                              "train_batch": 1024, "learning_rate": 1e-30})
        img, _ = c.render(small_settings(samples=64, seed=2000 + r, **kw))
        c.close()
        return img

    mp, sp = _replicates(plain, reps)
    mn, sn = _replicates(nee, reps)
    z = (mn - mp) / np.sqrt(sp ** 2 + sn ** 2 + 1e-30)
    # the image mean (blocks of equal size): its difference over its standard error
    se = np.sqrt((sp ** 2).sum() + (sn ** 2).sum()) / len(mp)
    zi = (mn.mean() - mp.mean()) / se
    print(f"{case}: image mean plain {mp.mean():.5g} training estimate {mn.mean():.5g} "
          f"(z {zi:+.2f}, se {se:.3g}); block z max {np.abs(z).max():.2f}; "
          f"standard error plain/estimate {np.sqrt((sp ** 2).sum() / (sn ** 2).sum()):.2f}")
    assert np.all(np.abs(z) < 4.5), z
    assert abs(zi) < 4.0
    assert se < 0.01 * mp.mean()  # precise enough to see a 4 % error


def test_trained_nee_cache_on_the_sunlit_terrain(scenes):
    """The sun (radiance 5e4) reaches the scene only through bounce rays, so the plain render's
    error is its fireflies; the camera-hit cache trained on environment-sampled paths (L2 loss)
    removes them with a small bias."""
    from mitrace_amd import scenegen
    from mitrace_amd.cache import RadianceCache
    g = scenes("sunlit_terrain")
    kw = dict(width=480, height=270, traversal=abi.MT_TRAVERSAL_FAST, output=abi.MT_OUTPUT_LINEAR)
    ref, _, _ = g.render(scenegen.large_scene_settings(samples=4096, seed=300, **kw))
    plain, _, _ = g.render(scenegen.large_scene_settings(samples=16, seed=301, **kw))
    c = RadianceCache(g, {"train_every": 8, "query_depth": 0, "query_after_light": 1,
                          "train_nee": 1, "loss": 1, "warmup_passes": 8, "spp_per_pass": 4,
                          "steps_per_pass": 16, "train_batch": 16384, "learning_rate": 2e-3})
    img, st = c.render(scenegen.large_scene_settings(samples=16, seed=301, **kw))
    c.close()
    # trained samples: at most the Adam steps' batches of the 8 warm-up and 4 counted passes
    assert 0 < st["diag"][8] <= 12 * 16 * 16384, st["diag"][8]
    e_plain = np.sqrt(np.mean((plain - ref) ** 2))
    e_cache = np.sqrt(np.mean((img - ref) ** 2))
    bias = (img.mean() - ref.mean()) / ref.mean()
    print(f"rmse plain {e_plain:.4g} cached {e_cache:.4g} relative bias {bias:+.4f}")
    # measured: RMSE 272 against 3058, bias -1.0 % to -1.4 % over three seeds
    assert e_cache < 0.2 * e_plain
    assert abs(bias) < 0.04


def test_train_nee_needs_after_light_queries(scenes):
    from mitrace_amd._lib import MiTraceError
    from mitrace_amd.cache import RadianceCache
    g = scenes("cornell_gen")
    c = RadianceCache(g, {"train_nee": 1, "query_depth": 1})
    with pytest.raises(MiTraceError):
        c.render(small_settings(width=8, height=8, samples=4))
    c.close()


@pytest.mark.parametrize("mode", [abi.MT_MODE_PARITY, abi.MT_MODE_FIXED])
def test_training_estimate_without_environment_is_the_plain_estimate(scenes, mode, monkeypatch):
    """No environment to sample: the training paths run the reference estimator itself, so with
    every path training their full estimates are the plain render's samples bit for bit."""
    from mitrace_amd.cache import RadianceCache
    g = scenes("cornell_gen")
    kw = dict(width=24, height=20, samples=8, bounces=8, seed=31, mode=mode, rr_depth=3,
              output=abi.MT_OUTPUT_LINEAR)
    ref, _, _ = g.render(small_settings(**kw))
    monkeypatch.setenv("MT_CACHE_TRAINING_IMAGE", "1")
    c = RadianceCache(g, {"train_every": 1, "query_depth": 0, "query_after_light": 1, "train_nee": 1,
                          "loss": 1, "spp_per_pass": 8, "steps_per_pass": 1, "train_batch": 1024})
    img, _ = c.render(small_settings(**kw))
    c.close()
    np.testing.assert_array_equal(img, ref)
