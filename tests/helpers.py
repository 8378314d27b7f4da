"""Shared scene and ray builders for the test-suite."""
import os

import numpy as np

from mitrace_amd import abi, gltf, scenegen
from mitrace_amd.scene import BvhOptions, Settings

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fixtures")


def fixture_path(name):
    return {
        "cornell_box": os.path.join(FIX, "cornell_box_test", "cornell_box.gltf"),
        "light_creep": os.path.join(FIX, "light_creep_test", "light_creep_test.gltf"),
        "material_spheres": os.path.join(FIX, "material_spheres_test", "material_spheres.gltf"),
    }[name]


def render_params(name):
    """the --width/--height/--samples/--ev-exposure of the fixture's render_params.txt"""
    d = os.path.dirname(fixture_path(name))
    toks = open(os.path.join(d, "render_params.txt")).read().split()
    out = {}
    for i in range(0, len(toks) - 1, 2):
        out[toks[i].lstrip("-")] = float(toks[i + 1])
    return out


def load_scene(name, bvh=None):
    """Named scenes: the reference's test fixtures and the generated workloads (small forms)."""
    if name in ("cornell_box", "light_creep", "material_spheres"):
        return gltf.load(fixture_path(name), bvh=bvh)
    if name == "material_spheres_bvh4":
        return gltf.load(fixture_path("material_spheres"), bvh=BvhOptions(False, 4, 32))
    if name == "cornell_gen":
        return scenegen.cornell()
    if name == "sphere_grid":
        return scenegen.sphere_grid(env_width=256, env_height=128)
    if name == "large_small":
        return scenegen.large_scene(terrain_res=96, n_spheres=300, tex_res=64, env_width=64,
                                    env_height=32)
    if name == "sphere_grid_coarse":
        return scenegen.sphere_grid(env_width=128, env_height=64, segments=16, rings=8)
    raise KeyError(name)


def scene_bounds(sc):
    lo = np.full(3, np.inf)
    hi = np.full(3, -np.inf)
    for inst in sc.instances:
        p = sc.meshes[inst.mesh].positions
        w = (inst.transform[:3, :3] @ p.T).T + inst.transform[:3, 3]
        lo = np.minimum(lo, w.min(0))
        hi = np.maximum(hi, w.max(0))
    return lo, hi


def seeded_rays(sc, n, seed=0, settings=None):
    """Camera rays of the scene plus random rays through its bounds, plus axis-parallel rays
    (direction components exactly 0, the inf/NaN cases of the slab test)."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    n_cam = n // 2
    aspect = sc.camera.aspect
    uv = rng.random((n_cam, 2)).astype(np.float32)
    cam = po.camera_rays(sc.camera, uv, aspect)
    lo, hi = scene_bounds(sc)
    c = (lo + hi) / 2
    ext = (hi - lo)
    n_rnd = n - n_cam - 24
    o = c + (rng.random((n_rnd, 3)) - 0.5) * ext * 1.2
    d = rng.normal(size=(n_rnd, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rnd = np.concatenate([o, d], 1).astype(np.float32)
    ax = []
    for k in range(24):
        o = c + (rng.random(3) - 0.5) * ext * 0.8
        dd = np.zeros(3)
        dd[k % 3] = 1.0 if (k // 3) % 2 == 0 else -1.0
        ax.append(np.concatenate([o, dd]))
    return np.concatenate([cam, rnd, np.array(ax, np.float32)]).astype(np.float32)


def small_settings(**kw):
    s = Settings(width=32, height=24, samples=8, bounces=5, seed=7)
    for k, v in kw.items():
        setattr(s, k, v)
    return s
