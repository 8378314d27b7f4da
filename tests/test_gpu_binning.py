"""Ray binning (MT_BIN_RAYS=1): before the extend of every secondary bounce the queue's rays are
put in order of direction octant and origin cell, and extend takes them in that order while
writing each hit at the ray's own queue slot.  The order in which rays are traced must not
change anything: samples, ray totals and work counters equal the unbinned render's bit for bit
(and the oracle's), over several chunks on several streams, with point lights (shadow rays and
lit shading), the instance tree, the fast traversal, Russian roulette and a debug view."""
import numpy as np
import pytest

from helpers import load_scene, small_settings
from mitrace_amd import abi
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

WORK = ("rays", "samples", "inst_aabb_tests", "inst_entered", "node_pops", "tri_tests",
        "hits_shaded", "misses")


def render_both(monkeypatch, g, st):
    out = []
    for flag in ("0", "1"):
        monkeypatch.setenv("MT_BIN_RAYS", flag)
        out.append(g.render(st, want_samples=True))
    return out


def assert_same(a, b):
    (ra, sa, ka), (rb, sb, kb) = a, b
    diff = sa.view(np.uint32) != sb.view(np.uint32)
    assert not diff.any(), f"{diff.any(axis=2).mean():.4%} of samples differ"
    assert np.array_equal(ra.view(np.uint32), rb.view(np.uint32))
    for f in WORK:
        assert ka[f] == kb[f], f


@pytest.mark.parametrize("name,kw", [
    ("sphere_grid_coarse", {}),
    ("cornell_gen", {"mode": abi.MT_MODE_FIXED, "rr_depth": 3, "bounces": 12}),
    ("light_creep", {}),
    ("large_small", {}),
])
def test_binned_render_equals_unbinned(gpu, monkeypatch, name, kw):
    from mitrace_amd.tracer import GpuScene
    g = GpuScene(load_scene(name))
    try:
        # 1536 paths per chunk: 16 chunks over the default three streams
        st = small_settings(width=64, height=48, samples=8, bounces=kw.pop("bounces", 6),
                            paths_in_flight=1536, count_work=True, **kw)
        a, b = render_both(monkeypatch, g, st)
        assert a[2]["rays"] > a[2]["samples"]  # secondary bounces ran, so binning did
        assert_same(a, b)
    finally:
        g.close()


def test_binned_render_equals_oracle(gpu, monkeypatch):
    from mitrace_amd.tracer import GpuScene
    sc = load_scene("sphere_grid_coarse")
    g, o = GpuScene(sc), po.OracleScene(sc)
    try:
        st = small_settings(width=40, height=30, samples=6, bounces=5, paths_in_flight=1200)
        monkeypatch.setenv("MT_BIN_RAYS", "1")
        grgb, gs, gk = g.render(st, want_samples=True)
        orgb, os_, ow, _ = o.render(st, want_samples=True)
        assert not (gs.view(np.uint32) != os_.view(np.uint32)).any()
        assert np.array_equal(grgb, orgb) and gk["rays"] == ow["rays"]
    finally:
        g.close()
        o.close()


def test_binned_fast_traversal_and_debug_view(gpu, monkeypatch):
    from mitrace_amd.tracer import GpuScene
    sc = load_scene("large_small")
    sc.bvh.fast = True
    g = GpuScene(sc)
    try:
        st = small_settings(width=48, height=32, samples=4, bounces=5, paths_in_flight=1024,
                            traversal=abi.MT_TRAVERSAL_FAST)
        assert_same(*render_both(monkeypatch, g, st))
        st = small_settings(width=48, height=32, samples=2, bounces=4, paths_in_flight=1024,
                            debug_mode=abi.MT_DEBUG_BOUNCES)
        assert_same(*render_both(monkeypatch, g, st))
    finally:
        g.close()
