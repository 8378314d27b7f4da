"""The reference-rule BVH built on the GPU (mt_bvh_options.device_build_min_tris) equals the
host build bit for bit -- node words, boxes and the triangle index list (BVH.cpp:117-272) -- on
meshes of every shape the builder meets: UV spheres, terrain grids up to the large scene's
2048x2048-vertex terrain, random soups, coincident triangles (an unsplittable root), depth limits
that force leaves, and leaf sizes from 1 to 32."""
import time

import numpy as np
import pytest

from mitrace_amd import scenegen
from mitrace_amd.scene import BvhOptions, Camera, Instance, Material, Mesh, Scene

pytestmark = pytest.mark.gpu


def _scene(mesh, bvh):
    return Scene(meshes=[mesh], materials=[Material()], instances=[Instance(0, 0, np.eye(4, dtype=np.float32))],
                 camera=Camera(type=0, camera_to_world=np.eye(4, dtype=np.float32), yfov=0.5), bvh=bvh)


def _soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-5, 5, (n, 1, 3))
    p = (c + rng.normal(scale=0.3, size=(n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return Mesh("soup", p, np.tile(np.array([0, 1, 0], np.float32), (3 * n, 1)))


def _coincident(n):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    p = np.tile(tri, (n, 1))
    return Mesh("stack", p, np.tile(np.array([0, 0, 1], np.float32), (3 * n, 1)))


def _terrain(res, seed=2):
    return scenegen.terrain_mesh(res, 400.0, 30.0, np.random.default_rng(seed))[0]


def _rays(mesh, n, seed):
    """rays toward random points of the mesh's box from outside it"""
    rng = np.random.default_rng(seed)
    p = mesh.positions
    lo, hi = p.min(0), p.max(0)
    target = lo + rng.random((n, 3)) * (hi - lo)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = target - d * (np.linalg.norm(hi - lo) + 1.0)
    return np.concatenate([o, d], 1).astype(np.float32)


def _both(mesh, max_tri=4, max_depth=32):
    """the mesh's exported BVH from a host build and a device build, the scene-creation times,
    and closest hits through both scenes (the device build's pair and prim records are made on
    the device)"""
    from mitrace_amd.tracer import GpuScene
    out, times, hits = [], [], []
    rays = _rays(mesh, 4000, 3)
    for dev in (0, 1):
        t0 = time.perf_counter()
        g = GpuScene(_scene(mesh, BvhOptions(False, max_tri, max_depth, device_build_min_tris=dev)))
        times.append(time.perf_counter() - t0)
        out.append(g.mesh_bvh(0))
        hits.append(g.trace(rays))
        g.close()
    assert (hits[0]["instance"] >= 0).any()
    for f in hits[0].dtype.names:
        assert np.array_equal(hits[0][f], hits[1][f]), f
    return out, times


@pytest.mark.parametrize("name,mesh,max_tri,max_depth", [
    ("uv_sphere", lambda: scenegen.uv_sphere(), 4, 32),
    ("uv_sphere_leaf32", lambda: scenegen.uv_sphere(), 32, 32),
    ("terrain_96", lambda: _terrain(96), 4, 32),
    ("terrain_96_depth6", lambda: _terrain(96), 4, 6),
    ("terrain_96_depth1", lambda: _terrain(96), 2, 1),
    ("soup_20000", lambda: _soup(20000, 1), 4, 32),
    ("soup_leaf1", lambda: _soup(3000, 2), 1, 40),
    ("coincident", lambda: _coincident(300), 4, 32),
    ("icosphere", lambda: scenegen.icosphere(3), 4, 32),
])
def test_device_build_equals_host_build(gpu, name, mesh, max_tri, max_depth):
    (host, dev), _ = _both(mesh(), max_tri, max_depth)
    assert host[0].shape == dev[0].shape, (name, host[0].shape, dev[0].shape)
    assert np.array_equal(host[0], dev[0]), name
    assert np.array_equal(host[1], dev[1]), name


def test_device_build_large_terrain(gpu):
    """the large scene's 2048x2048-vertex terrain (8,380,418 triangles)"""
    import sys
    mesh = _terrain(2048)
    (host, dev), (th, td) = _both(mesh)
    print(f"terrain 2048: {host[0].shape[0]} nodes, {host[1].shape[0]} leaf entries; scene create "
          f"with host build {th:.2f} s, with device build {td:.2f} s", file=sys.stderr, flush=True)
    assert host[0].shape == dev[0].shape
    assert np.array_equal(host[0], dev[0]) and np.array_equal(host[1], dev[1])
