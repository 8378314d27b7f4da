"""Fast traversal (MT_TRAVERSAL_FAST; a GPU speed mode with no reference counterpart): per-mesh
SAH BVHs with tight boxes (built on the host, or on the GPU as linear BVHs), nodes and
instances culled by entry distance, shadow rays ended at their first occluder.

Bars: every closest hit is the exact closest hit of the oracle's float64 brute-force search
(same instance and triangle, or a hit at the same distance within 1e-4 relative: a shared
edge or coplanar tie), hit/miss disagreements at most 1 in 2000 rays (float32 edge cases);
occlusion answers equal the brute-force answer for bounds away from the hit distance, and in
the reference traversal they equal its own closest hit bit for bit; rendered images equal the
reference traversal's within 3 sigma of the per-pixel Monte Carlo error; batched jobs with the
fast traversal equal their single-job renders bit for bit.
"""
import numpy as np
import pytest

from helpers import load_scene, seeded_rays, small_settings
from mitrace_amd import abi
from mitrace_amd._lib import MiTraceError
from mitrace_amd.tracer import GpuScene
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FAST = abi.MT_TRAVERSAL_FAST
REF = abi.MT_TRAVERSAL_REFERENCE
# ":lbvh": the fast trees of every mesh above 4 triangles built on the GPU as linear BVHs
# (mt_bvh_options.fast_device_build_min_tris) instead of the host's binned SAH
SCENES = ["cornell_gen", "material_spheres_bvh4", "sphere_grid", "large_small", "sphere_grid:lbvh",
          "large_small:lbvh", "material_spheres_bvh4:lbvh"]


@pytest.fixture(scope="module")
def scenes(gpu):
    cache = {}

    def get(name):
        if name not in cache:
            base, _, variant = name.partition(":")
            sc = load_scene(base)
            sc.bvh.fast = True
            if variant == "lbvh":
                sc.bvh.fast_device_build_min_tris = 1
            cache[name] = (sc, GpuScene(sc), po.OracleScene(sc))
        return cache[name]
    yield get
    for sc, g, o in cache.values():
        g.close()
        o.close()


@pytest.mark.parametrize("name", SCENES)
def test_fast_closest_hit_is_the_exact_closest_hit(scenes, name):
    sc, g, o = scenes(name)
    rays = seeded_rays(sc, 6000, seed=21)
    fh = g.trace(rays, FAST)
    bh, _ = o.trace(rays, brute=True)
    rh = g.trace(rays, REF)
    print(f"{name}: case {len(rays)} seeded rays, seed 21", flush=True)
    assert (bh["instance"] >= 0).sum() > 100
    hitmiss = (fh["instance"] >= 0) != (bh["instance"] >= 0)
    assert hitmiss.sum() <= len(rays) // 2000, np.where(hitmiss)[0]
    both = (fh["instance"] >= 0) & (bh["instance"] >= 0)
    same = both & (fh["instance"] == bh["instance"]) & (fh["triangle"] == bh["triangle"])
    other = both & ~same
    # a different triangle only at the same distance (shared edge, coplanar tie)
    np.testing.assert_allclose(fh["distance"][other], bh["distance"][other], rtol=1e-4, atol=1e-6)
    # float32 vs float64 on the same triangle; grazing hits are ill-conditioned, hence 1e-3
    np.testing.assert_allclose(fh["distance"][same], bh["distance"][same], rtol=1e-3, atol=1e-5)
    # never farther than the reference traversal's hit (which may be non-closest)
    rb = both & (rh["instance"] >= 0)
    assert np.all(fh["distance"][rb] <= rh["distance"][rb] * (1 + 1e-4) + 1e-6)
    # the world hit point lies on the ray at the reported distance
    o_, d_ = rays[:, :3], rays[:, 3:]
    p = o_[both] + d_[both] * fh["distance"][both][:, None]
    np.testing.assert_allclose(fh["world_pos"][both], p, rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("name", SCENES)
def test_occlusion_queries(scenes, name):
    sc, g, o = scenes(name)
    rays = seeded_rays(sc, 4000, seed=22)
    bh, _ = o.trace(rays, brute=True)
    rh = g.trace(rays, REF)
    rng = np.random.default_rng(5)
    d = np.where(bh["instance"] >= 0, bh["distance"], 50.0).astype(np.float64)
    bound = d * rng.choice([0.5, 0.9, 1.1, 2.0], size=len(rays))
    r7 = np.concatenate([rays, bound[:, None]], 1).astype(np.float32)
    print(f"{name}: case {len(rays)} occlusion rays, seed 22", flush=True)
    occ_f = g.occluded(r7, FAST)
    occ_r = g.occluded(r7, REF)
    # the reference query is its closest hit compared with the bound (Tracer.cpp:322)
    expect_r = (rh["instance"] >= 0) & ~(rh["distance"] > r7[:, 6])
    assert np.array_equal(occ_r.astype(bool), expect_r)
    exact = (bh["instance"] >= 0) & (bh["distance"] <= bound)
    bad = np.where(occ_f.astype(bool) != exact)[0]
    assert len(bad) <= len(rays) // 2000, bad
    assert exact.sum() > 100 and (~exact).sum() > 100


@pytest.mark.parametrize("name", ["sphere_grid", "sphere_grid:lbvh"])
def test_fast_render_matches_reference_traversal(scenes, name):
    sc, g, _ = scenes(name)
    for extra in ({}, {"mode": abi.MT_MODE_FIXED, "bounces": 8}):
        st = small_settings(width=48, height=32, samples=64, output=abi.MT_OUTPUT_LINEAR,
                            hdri_rotation=30.0, hdri_rotation_apply=True, **extra)
        ref, rs, rstat = g.render(st, want_samples=True)
        st.traversal = FAST
        fast, fs, fstat = g.render(st, want_samples=True)
        assert fstat["samples"] == rstat["samples"]
        sigma = np.sqrt(rs.var(axis=1) / st.samples + fs.var(axis=1) / st.samples).reshape(ref.shape)
        assert np.all(np.abs(fast - ref) <= 3 * sigma + 1e-4 * np.abs(ref) + 1e-6)
        # most samples take the same path with the same hits
        same = np.all(np.isclose(rs, fs, rtol=1e-4, atol=1e-6), axis=2).mean()
        assert same > 0.95, same


def test_fast_requires_fast_bvhs(gpu):
    sc = load_scene("cornell_gen")
    g = GpuScene(sc)
    st = small_settings(traversal=FAST)
    with pytest.raises(MiTraceError):
        g.render(st)
    with pytest.raises(MiTraceError):
        g.trace(np.zeros((1, 6), np.float32), FAST)
    with pytest.raises(MiTraceError):
        g.render(small_settings(traversal=7))
    g.close()


def test_batched_fast_jobs_equal_single_renders(scenes):
    from mitrace_amd.batch import Batch
    jobs = [("sphere_grid", small_settings(width=40, height=24, samples=16, seed=3, traversal=FAST)),
            ("large_small", small_settings(width=32, height=32, samples=8, seed=4, traversal=FAST)),
            ("sphere_grid", small_settings(width=24, height=24, samples=8, seed=5))]
    b = Batch({"paths_per_iteration": 1 << 14, "tile_pixels": 64, "iteration_tiles": 64})
    ids = [b.submit(scenes(n)[1], st) for n, st in jobs]
    b.start()
    b.finish()
    for (n, st), j in zip(jobs, ids):
        ref, _, _ = scenes(n)[1].render(st)
        np.testing.assert_array_equal(b.image(j), ref, err_msg=f"{n} {st}")
    b.close()
