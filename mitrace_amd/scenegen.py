"""Seeded procedural scenes for the benchmark workloads.

* :func:`cornell` -- the CPU-runnable Cornell box: a 1x1x1 box open at +z (5 quads, 10 tris),
  a tall 0.3x0.6x0.3 box rotated 15 deg and a short 0.3^3 box rotated -18 deg (12 tris each),
  one point light at (0.5, 0.98, 0.5) of radius 0.05 and intensity 15, a 40 deg camera
  looking along -z.  34 triangles, every mesh at most 32, so the reference defaults scan
  them linearly (Mesh.cpp:37-40).
* :func:`sphere_grid` -- 7x7 UV spheres (64 segments x 32 rings, 3968 tris each) on a
  100-tri ground plane, roughness linspace(0.05, 1) by column, metallic 0 for rows 0-3 and 1
  for rows 4-6, seeded base colours in [0.2, 0.9]^3, lit by a procedural 2048x1024 float32
  equirect sky (vertical gradient plus a 0.5 deg sun disk of radiance 5e4), BVH leaves of at
  most 4 triangles.
"""
import numpy as np

from . import abi
from .scene import (BvhOptions, Camera, Environment, Instance, Light, Material, Mesh, Scene,
                    Settings, look_at, trs)


def _quad(p0, p1, p2, p3, n):
    """two triangles (p0,p1,p2), (p0,p2,p3) with a flat normal."""
    pos = np.array([p0, p1, p2, p0, p2, p3], np.float32)
    nrm = np.tile(np.array(n, np.float32), (6, 1))
    return pos, nrm


def _mesh_from_quads(name, quads):
    P, N = zip(*[_quad(*q) for q in quads])
    return Mesh(name=name, positions=np.concatenate(P), normals=np.concatenate(N))


def unit_cube(name="cube"):
    """a cube spanning [-0.5, 0.5]^3 with outward flat normals (12 tris)."""
    h = 0.5
    q = [
        ((-h, -h, h), (h, -h, h), (h, h, h), (-h, h, h), (0, 0, 1)),
        ((h, -h, -h), (-h, -h, -h), (-h, h, -h), (h, h, -h), (0, 0, -1)),
        ((h, -h, h), (h, -h, -h), (h, h, -h), (h, h, h), (1, 0, 0)),
        ((-h, -h, -h), (-h, -h, h), (-h, h, h), (-h, h, -h), (-1, 0, 0)),
        ((-h, h, h), (h, h, h), (h, h, -h), (-h, h, -h), (0, 1, 0)),
        ((-h, -h, -h), (h, -h, -h), (h, -h, h), (-h, -h, h), (0, -1, 0)),
    ]
    return _mesh_from_quads(name, q)


def cornell(seed=0):
    """The Cornell-box workload (256x256 at 64 spp in the benchmark table)."""
    del seed  # fully determined; the seed only keys the RNG of a render
    white = _mesh_from_quads("white_walls", [
        ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1), (0, 1, 0)),      # floor
        ((0, 1, 1), (1, 1, 1), (1, 1, 0), (0, 1, 0), (0, -1, 0)),     # ceiling
        ((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 1)),      # back
    ])
    left = _mesh_from_quads("left_wall", [((0, 0, 1), (0, 1, 1), (0, 1, 0), (0, 0, 0), (1, 0, 0))])
    right = _mesh_from_quads("right_wall", [((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1), (-1, 0, 0))])
    cube = unit_cube()
    mats = [
        Material("white", (0.73, 0.73, 0.73, 1.0), metallic_factor=0.0, roughness_factor=1.0),
        Material("red", (0.63, 0.065, 0.05, 1.0), metallic_factor=0.0, roughness_factor=1.0),
        Material("green", (0.14, 0.45, 0.09, 1.0), metallic_factor=0.0, roughness_factor=1.0),
    ]
    eye = np.eye(4, dtype=np.float32)
    inst = [
        Instance(0, 0, eye), Instance(1, 1, eye), Instance(2, 2, eye),
        Instance(3, 0, trs((0.33, 0.3, 0.37), 15.0, (0.3, 0.6, 0.3))),
        Instance(3, 0, trs((0.66, 0.15, 0.66), -18.0, (0.3, 0.3, 0.3))),
    ]
    light = Light(type=abi.MT_LIGHT_POINT, position=(0.5, 0.98, 0.5), color=(15.0, 15.0, 15.0),
                  point_size=0.05)
    cam = Camera(type=abi.MT_CAMERA_PERSPECTIVE,
                 camera_to_world=trs((0.5, 0.5, 0.5 + 0.5 / np.tan(np.radians(20.0)) + 0.02)),
                 yfov=float(np.radians(40.0)), aspect=1.0)
    return Scene(meshes=[white, left, right, cube], materials=mats, instances=inst,
                 lights=[light], camera=cam)


def cornell_settings(**kw):
    s = Settings(width=256, height=256, samples=64, bounces=8)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def uv_sphere(segments=64, rings=32, radius=1.0, name="uv_sphere"):
    """UV sphere with smooth normals: 2*segments cap triangles + 2*segments*(rings-2) quads
    split in two (3968 triangles for 64x32)."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, segments + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    V = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    tris = []
    for r in range(rings):
        for s in range(segments):
            a, b = V[r, s], V[r, s + 1]
            c, d = V[r + 1, s], V[r + 1, s + 1]
            if r == 0:
                tris.append((a, d, c))
            elif r == rings - 1:
                tris.append((a, b, c))
            else:
                tris.append((a, b, d))
                tris.append((a, d, c))
    n = np.array(tris, np.float64).reshape(-1, 3)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return Mesh(name=name, positions=(n * radius).astype(np.float32), normals=n.astype(np.float32))


def ground_plane(nx=10, nz=5, x0=-3.0, x1=9.0, z0=-3.0, z1=9.0, y=0.0):
    xs = np.linspace(x0, x1, nx + 1)
    zs = np.linspace(z0, z1, nz + 1)
    quads = []
    for i in range(nx):
        for j in range(nz):
            quads.append(((xs[i], y, zs[j + 1]), (xs[i + 1], y, zs[j + 1]),
                          (xs[i + 1], y, zs[j]), (xs[i], y, zs[j]), (0, 1, 0)))
    return _mesh_from_quads("ground", quads)


def procedural_sky(width=2048, height=1024, sun_elevation_deg=35.0, sun_azimuth_deg=60.0,
                   sun_radius_deg=0.5, sun_radiance=5e4):
    """float32 RGB equirect in the reference's mapping (Image.cpp:63-75): u = (atan2(z,x)+pi)/2pi,
    v = acos(y)/pi.  Vertical gradient from horizon to zenith, dark ground below the horizon,
    and a sun disk."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = u * 2.0 * np.pi - np.pi
    phi = v * np.pi
    y = np.cos(phi)[:, None]
    sx = (np.sin(phi)[:, None] * np.cos(theta)[None, :])
    sz = (np.sin(phi)[:, None] * np.sin(theta)[None, :])
    zen = np.array([0.25, 0.45, 0.95])
    hor = np.array([0.85, 0.9, 1.0])
    gnd = np.array([0.25, 0.22, 0.2])
    t = np.clip(y, 0.0, 1.0)[..., None] ** 0.5
    sky = hor * (1 - t) + zen * t
    img = np.where((y >= 0)[..., None], sky, np.broadcast_to(gnd, sky.shape))
    img = np.broadcast_to(img, (height, width, 3)).copy()
    el, az = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sd = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    cosang = sx * sd[0] + y * sd[1] + sz * sd[2]
    img[cosang >= np.cos(np.radians(sun_radius_deg))] = sun_radiance
    return img.astype(np.float32)


def sphere_grid(seed=1, env_width=2048, env_height=1024, segments=64, rings=32):
    """The material-spheres grid workload (1920x1080 at 256 spp in the benchmark table)."""
    rng = np.random.default_rng(seed)
    sphere = uv_sphere(segments, rings, 1.0)
    ground = ground_plane()
    meshes = [sphere, ground]
    mats, inst = [], []
    rough = np.linspace(0.05, 1.0, 7)
    radius, spacing = 0.4, 1.0
    for row in range(7):
        for col in range(7):
            base = rng.uniform(0.2, 0.9, 3)
            mats.append(Material(f"sphere_{row}_{col}", (float(base[0]), float(base[1]), float(base[2]), 1.0),
                                 metallic_factor=0.0 if row < 4 else 1.0,
                                 roughness_factor=float(rough[col])))
            inst.append(Instance(0, len(mats) - 1,
                                 trs((col * spacing, radius, row * spacing), 0.0, (radius,) * 3)))
    mats.append(Material("ground", (0.5, 0.5, 0.5, 1.0), metallic_factor=0.0, roughness_factor=0.8))
    inst.append(Instance(1, len(mats) - 1, np.eye(4, dtype=np.float32)))
    cam = Camera(type=abi.MT_CAMERA_PERSPECTIVE,
                 camera_to_world=look_at((3.0, 4.2, 11.0), (3.0, 0.2, 3.0)),
                 yfov=float(np.radians(40.0)), aspect=16.0 / 9.0)
    env = Environment(texels=procedural_sky(env_width, env_height), filter=abi.MT_FILTER_LINEAR)
    return Scene(meshes=meshes, materials=mats, instances=inst, lights=[], camera=cam, env=env,
                 bvh=BvhOptions(disable=False, max_triangles=4, max_depth=32))


def sphere_grid_settings(**kw):
    s = Settings(width=1920, height=1080, samples=256, bounces=5, hdri_rotation=30.0,
                 hdri_rotation_apply=True)
    for k, v in kw.items():
        setattr(s, k, v)
    return s
