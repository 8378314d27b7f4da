"""Debug views (Config::DebugMode, ProcessDebugRay Tracer.cpp:366-488, the heat-map post-process
Tracer.cpp:642-716).  CPU: the oracle's views on the Cornell box against closed forms (the wall
albedos 0.96*b + 0.04, flat-wall normals, camera distances, RandomColorFromU16 colours, the
heat-map ramp).  GPU: every view through mt_render against the oracle on the same inputs."""
import numpy as np
import pytest

from helpers import load_scene, small_settings
from mitrace_amd import abi, scenegen
from oracle import pyoracle as po

ALL_MODES = list(range(1, 23))
# views computed from first-hit records and integer counters: equal up to libm last bits
GEOMETRIC = [1, 2, 3, 4, 5, 6, 7, 8, 12, 13, 14, 15, 16, 17, 18, 19, 20]


def _render(orc, st):
    img, _, _, _ = orc.render(st, faithful_inverse=False)
    return img


def color_from_u16(v):
    x = v & 0xFFFF
    x ^= x >> 7
    x = (x * 0x2C1B3C6D) & 0xFFFFFFFF
    x ^= x >> 12
    x = (x * 0x297A2D39) & 0xFFFFFFFF
    x ^= x >> 15
    v = x & 0xFFFF
    return np.array([((v & 0x1C0) >> 6) / 7.0, ((v & 0x38) >> 3) / 7.0, (v & 7) / 7.0], np.float32)


@pytest.fixture(scope="module")
def cornell():
    sc = scenegen.cornell()
    return sc, po.OracleScene(sc)


def test_albedo_view_holds_the_wall_colours(cornell):
    sc, orc = cornell
    st = small_settings(width=48, height=48, samples=1, bounces=3, seed=1, debug_mode=abi.MT_DEBUG_ALBEDO)
    img = _render(orc, st).reshape(-1, 3)
    expected = [0.96 * np.array(m.base_color_factor[:3], np.float32) + 0.04 for m in sc.materials]
    close = np.zeros(len(img), bool)
    for e in expected:
        close |= np.abs(img - e).max(axis=1) < 1e-6
    assert close.mean() > 0.95, close.mean()


def test_normal_and_depth_views(cornell):
    sc, orc = cornell
    st = small_settings(width=32, height=32, samples=1, bounces=2, seed=2,
                        debug_mode=abi.MT_DEBUG_GEOMETRIC_NORMAL)
    n = _render(orc, st).reshape(-1, 3) * 2 - 1
    # every first hit is on an axis-aligned face or a box rotated about y: unit normals
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    st.debug_mode = abi.MT_DEBUG_DEPTH
    d = _render(orc, st)  # grey heat map: depth / max depth
    assert d.max() == pytest.approx(1.0) and d.min() > 0.3


def test_color_per_mesh_is_random_color_from_u16(cornell):
    sc, orc = cornell
    st = small_settings(width=32, height=32, samples=1, bounces=2, seed=3,
                        debug_mode=abi.MT_DEBUG_COLOR_PER_MESH)
    img = _render(orc, st).reshape(-1, 3)
    palette = np.array([color_from_u16(i) for i in range(len(sc.instances))])
    dist = np.abs(img[:, None, :] - palette[None]).max(axis=2).min(axis=1)
    assert (dist < 1e-6).mean() > 0.95


def test_heat_map_ramp_and_bounce_view(cornell):
    sc, orc = cornell
    st = small_settings(width=24, height=24, samples=2, bounces=6, seed=4,
                        debug_mode=abi.MT_DEBUG_BOUNCES)
    img = _render(orc, st).reshape(-1, 3)
    # colourful ramp blue -> cyan -> green -> yellow -> red: the maximum pixel is red
    assert np.allclose(img[np.argmax(img[:, 0] - img[:, 2])], [1, 0, 0], atol=1e-6)
    assert np.all(img >= 0) and np.all(img <= 1)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ALL_MODES)
def test_gpu_debug_view_matches_oracle(gpu, mode):
    from mitrace_amd.tracer import GpuScene
    sc = load_scene("cornell_box")
    g = GpuScene(sc)
    orc = po.OracleScene(sc)
    st = small_settings(width=40, height=30, samples=4, bounces=5, seed=13, debug_mode=mode,
                        firefly_enable=mode in (21, 22))
    ref = _render(orc, st)
    img, _, _ = g.render(st)
    scale = max(float(np.abs(ref).max()), 1e-6)
    close = np.isclose(img, ref, rtol=1e-4, atol=1e-5 * scale).all(axis=-1)
    # every pixel of every view: the paths are the oracle's sample for sample (same counter RNG
    # and elementary functions); sums over samples may differ in their last bits
    assert close.all(), (mode, close.mean())
