"""Edge cases of the render path, the GPU against the oracle sample for sample through the
C-ABI (the reference's own tests hold no numbers for these; the oracle restates its rules):

* empty scenes -- no instances, with a sky (every camera ray misses into it) and without one
  (black) -- and a scene without lights or sky;
* zero and one bounce, and the maximum of 64 bounces in a closed box;
* a 1x1 image at 1 spp; ragged images whose chunks do not divide the render
  (paths_in_flight smaller than the render, the last chunk partial);
* a single triangle, and degenerate triangles (zero area: repeated and collinear vertices)
  beside it, for closest hits and renders;
* invalid settings rejected with the C-ABI's error codes, not run.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import load_scene, small_settings
from mitrace_amd import abi, scenegen
from mitrace_amd.scene import Instance, Light, Material, Mesh, Scene
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def same_render(sc, st):
    from mitrace_amd.tracer import GpuScene
    g, o = GpuScene(sc), po.OracleScene(sc)
    try:
        grgb, gs, gk = g.render(st, want_samples=True)
        orgb, os_, ow, _ = o.render(st, want_samples=True)
        diff = gs.view(np.uint32) != os_.view(np.uint32)
        assert not diff.any(), f"{diff.any(axis=2).mean():.4%} of samples differ"
        assert np.array_equal(grgb, orgb)
        assert gk["rays"] == ow["rays"] and gk["samples"] == ow["samples"]
        return grgb, gk
    finally:
        g.close()
        o.close()


def test_empty_scene_with_and_without_sky(gpu):
    base = scenegen.sphere_grid(env_width=64, env_height=32)
    st = small_settings(width=16, height=9, samples=3, bounces=5)
    sky = Scene(meshes=[], materials=[], instances=[], lights=[], camera=base.camera, env=base.env)
    rgb, k = same_render(sky, st)
    assert rgb.mean() > 0.0 and k["rays"] == 16 * 9 * 3  # camera rays only: every one misses
    dark = Scene(meshes=[], materials=[], instances=[], lights=[], camera=base.camera)
    rgb, k = same_render(dark, st)
    assert not rgb.any() and k["rays"] == 16 * 9 * 3


def test_scene_without_lights_or_sky(gpu):
    sc = load_scene("sphere_grid_coarse")
    sc.env = None
    sc.lights = []
    rgb, _ = same_render(sc, small_settings(width=20, height=12, samples=4, bounces=4))
    assert not rgb.any()


@pytest.mark.parametrize("bounces", [0, 1, 64])
def test_bounce_limits(gpu, bounces):
    sc = scenegen.cornell() if bounces == 64 else load_scene("sphere_grid_coarse")
    rgb, k = same_render(sc, small_settings(width=12, height=10, samples=4, bounces=bounces))
    if bounces == 0:
        assert not rgb.any()  # no path step: nothing is collected


@pytest.mark.parametrize("w,h,spp,chunk", [(1, 1, 1, 0), (7, 3, 5, 20), (13, 11, 3, 7), (5, 4, 64, 64)])
def test_tiny_and_ragged_renders(gpu, w, h, spp, chunk):
    """chunk = paths_in_flight (0: automatic), rounded down to whole pixels: 20 paths = 4
    pixels of 5 samples, so the 21 pixels take 6 chunks with a partial last one; 7 paths at
    3 spp = 2 pixels per chunk; 64 paths at 64 spp = one pixel per chunk"""
    sc = load_scene("cornell_gen")
    same_render(sc, small_settings(width=w, height=h, samples=spp, bounces=6, paths_in_flight=chunk))


def degenerate_scene():
    """one ordinary triangle facing the camera plus zero-area triangles beside it (repeated
    vertices, collinear vertices), under a point light"""
    tris = np.array([
        [[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [0.0, 1.0, 0.0]],   # ordinary
        [[0.5, 0.5, 0.2], [0.5, 0.5, 0.2], [0.5, 0.5, 0.2]],      # a point
        [[-1.0, 0.0, 0.1], [0.0, 0.0, 0.1], [1.0, 0.0, 0.1]],     # collinear
        [[0.2, -0.5, 0.3], [0.2, -0.5, 0.3], [0.8, 0.1, 0.3]],    # two equal corners
    ], np.float32)
    pos = tris.reshape(-1, 3)
    nrm = np.tile(np.array([0.0, 0.0, 1.0], np.float32), (len(pos), 1))
    base = load_scene("cornell_gen")
    cam = base.camera
    m = Mesh("tri", pos, nrm)
    sc = Scene(meshes=[m], materials=[Material(base_color_factor=(0.8, 0.7, 0.6, 1.0))],
               instances=[Instance(0, 0, np.eye(4, dtype=np.float32))],
               lights=[Light(abi.MT_LIGHT_POINT, (0.0, 0.0, 3.0), (20.0, 20.0, 20.0), 0.0)],
               camera=cam, env=None)
    return sc


def test_single_and_degenerate_triangles(gpu):
    from mitrace_amd.tracer import GpuScene
    sc = degenerate_scene()
    g, o = GpuScene(sc), po.OracleScene(sc)
    rng = np.random.default_rng(2)
    o_ = np.concatenate([rng.uniform(-1.2, 1.2, (4000, 2)), np.full((4000, 1), 2.0)], 1)
    d = np.concatenate([rng.normal(scale=0.1, size=(4000, 2)), -np.ones((4000, 1))], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o_, d], 1).astype(np.float32)
    gh = g.trace(rays)
    oh, _ = o.trace(rays)
    assert (oh["instance"] >= 0).sum() > 500
    for f in ("instance", "triangle", "bv_index", "bvh_tests", "triangle_tests"):
        assert np.array_equal(gh[f], oh[f]), f
    hit = oh["instance"] >= 0
    for f in ("u", "v", "distance"):
        np.testing.assert_allclose(gh[f][hit], oh[f][hit], rtol=1e-6, atol=1e-7, err_msg=f)
    assert set(np.unique(oh["triangle"][hit])) == {0}  # zero-area triangles are never hit
    g.close()
    o.close()
    same_render(sc, small_settings(width=16, height=16, samples=4, bounces=3))


def test_invalid_settings_are_rejected(gpu):
    from mitrace_amd.tracer import GpuScene
    g = GpuScene(load_scene("cornell_gen"))
    cam = g.scene.camera.to_desc()
    rgb = np.zeros((16, 16, 3), np.float32)
    stats = abi.mt_stats()
    for kw, code in [({"width": 0}, abi.MT_E_INVALID), ({"samples": 0}, abi.MT_E_INVALID),
                     ({"bounces": 65}, abi.MT_E_UNSUPPORTED), ({"traversal": 7}, abi.MT_E_INVALID),
                     ({"traversal": abi.MT_TRAVERSAL_FAST}, abi.MT_E_INVALID)]:
        st = small_settings(width=16, height=16, **kw).to_desc() if "width" not in kw else \
            small_settings(height=16, **kw).to_desc()
        rc = g.L.mt_render(g.h, C.byref(cam), C.byref(st), abi.fptr(rgb), None, C.byref(stats))
        assert rc == code, (kw, rc)
        assert g.L.mt_last_error()
    g.close()
