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
* :func:`inverse_scene` -- the sphere grid with one shared float texture as every sphere's
  base colour, scalar roughness/metallic per sphere and four point lights; 8 camera poses from
  :func:`inverse_poses`.
"""
import numpy as np

from . import abi
from .scene import (BvhOptions, Camera, Environment, Instance, Light, Material, Mesh, Scene,
                    Settings, Texture, look_at, trs)


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


def uv_sphere(segments=64, rings=32, radius=1.0, name="uv_sphere", with_uv=False):
    """UV sphere with smooth normals: 2*segments cap triangles + 2*segments*(rings-2) quads
    split in two (3968 triangles for 64x32).  with_uv adds texcoords u = phi/2pi, v = theta/pi."""
    th = np.linspace(0.0, np.pi, rings + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, segments + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    V = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], axis=-1)
    UV = np.stack([P / (2.0 * np.pi), T / np.pi], axis=-1)
    idx = []
    for r in range(rings):
        for s in range(segments):
            a, b, c, d = (r, s), (r, s + 1), (r + 1, s), (r + 1, s + 1)
            if r == 0:
                idx.append((a, d, c))
            elif r == rings - 1:
                idx.append((a, b, c))
            else:
                idx.append((a, b, d))
                idx.append((a, d, c))
    ij = np.array(idx, np.int64).reshape(-1, 2)
    n = V[ij[:, 0], ij[:, 1]].astype(np.float64)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    uv = UV[ij[:, 0], ij[:, 1]].astype(np.float32) if with_uv else None
    return Mesh(name=name, positions=(n * radius).astype(np.float32), normals=n.astype(np.float32),
                texcoord0=uv)


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


def smooth_noise_texture(res, seed):
    """seeded smooth RGB noise in [0.1, 0.9]: three octaves of value noise per channel"""
    rng = np.random.default_rng(seed)
    ch = [0.1 + 0.8 * fbm(res, 3, rng, base_cells=4) for _ in range(3)]
    return np.ascontiguousarray(np.stack(ch, -1), np.float32)


def inverse_scene(seed=4, tex_res=512, env_width=2048, env_height=1024, segments=64, rings=32):
    """The sphere grid for inverse rendering: every sphere takes its base colour from one shared
    float RGB texture (seeded smooth noise, spherical UVs) and has scalar roughness and
    metallic.  Metallic is 0.1 / 0.9 (rows 0-3 / 4-6) rather than 0 / 1 so that both sides of
    every scalar are inside the BRDF's clamps and central differences exist.  Four point lights
    above the grid: without a light the reference's path throughput is never updated and the
    estimate would not depend on roughness or metallic except through the sampled directions.

    Returns (scene, params) with params = {tex, tex_first, tex_count, rough, metal, emis}
    (the ground-truth parameters; the spheres are materials 0..48, the ground is 49)."""
    rng = np.random.default_rng(seed)
    sc = sphere_grid(seed=1, env_width=env_width, env_height=env_height, segments=segments,
                     rings=rings)
    sc.meshes[0] = uv_sphere(segments, rings, 1.0, with_uv=True)
    for i in range(49):
        m = sc.materials[i]
        m.metallic_factor = 0.1 if i // 7 < 4 else 0.9
        m.base_color_factor = (1.0, 1.0, 1.0, 1.0)
    for x, z in [(0.5, 0.5), (5.5, 0.5), (0.5, 5.5), (5.5, 5.5)]:
        sc.lights.append(Light(type=abi.MT_LIGHT_POINT, position=(x, 3.0, z),
                               color=(40.0, 38.0, 36.0), point_size=0.1))
    # a wider, dimmer sun (3 deg, radiance 100): the 0.5 deg / 5e4 disk of the throughput
    # workload makes 32-sample estimates dominated by rare sun hits
    sc.env = Environment(texels=procedural_sky(env_width, env_height, sun_radius_deg=3.0,
                                               sun_radiance=100.0), filter=abi.MT_FILTER_LINEAR)
    tex = smooth_noise_texture(tex_res, int(rng.integers(1 << 31)))
    params = {
        "tex": tex, "tex_first": 0, "tex_count": 49,
        "rough": np.array([m.roughness_factor for m in sc.materials], np.float32),
        "metal": np.array([m.metallic_factor for m in sc.materials], np.float32),
        "emis": np.array([m.emissive_factor for m in sc.materials], np.float32).reshape(-1),
    }
    return sc, params


def inverse_poses(n=8, seed=4):
    """n seeded cameras on a ring around the grid centre, looking at it (square aspect)"""
    rng = np.random.default_rng(seed + 1000)
    cams = []
    for i in range(n):
        az = 2.0 * np.pi * (i + rng.uniform(-0.3, 0.3)) / n
        el = rng.uniform(np.radians(25.0), np.radians(55.0))
        r = rng.uniform(8.0, 10.0)
        c = np.array([3.0, 0.2, 3.0])
        eye = c + r * np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
        cams.append(Camera(type=abi.MT_CAMERA_PERSPECTIVE, camera_to_world=look_at(eye, c),
                           yfov=float(np.radians(40.0)), aspect=1.0))
    return cams


def inverse_settings(**kw):
    """512x512, linear radiance, Russian roulette from depth 3, 5 bounces"""
    s = Settings(width=512, height=512, samples=32, bounces=5, hdri_rotation=30.0,
                 hdri_rotation_apply=True, hdri_primary_intensity=1.0,
                 hdri_secondary_intensity=1.0, mode=abi.MT_MODE_FIXED, rr_depth=3,
                 output=abi.MT_OUTPUT_LINEAR)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def sphere_grid_settings(**kw):
    s = Settings(width=1920, height=1080, samples=256, bounces=5, hdri_rotation=30.0,
                 hdri_rotation_apply=True)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


# ------------------------------------------------------------------ large terrain scene
def _value_noise(res, cells, rng):
    """bilinearly interpolated random lattice of cells x cells values, sampled at res x res."""
    g = rng.random((cells + 1, cells + 1))
    t = np.linspace(0, cells, res, endpoint=False)
    i = t.astype(np.int64)
    f = t - i
    f = f * f * (3 - 2 * f)
    a = g[np.ix_(i, i)]
    b = g[np.ix_(i, i + 1)]
    c = g[np.ix_(i + 1, i)]
    d = g[np.ix_(i + 1, i + 1)]
    fx = f[None, :]
    fy = f[:, None]
    return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy


def fbm(res, octaves, rng, base_cells=4):
    h = np.zeros((res, res))
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = min(base_cells * (2 ** o), res - 1)
        h += amp * _value_noise(res, cells, rng)
        total += amp
        amp *= 0.5
    return h / total


def terrain_mesh(res, extent, height, rng, uv_repeat=8.0):
    """heightfield of res x res vertices over [-extent, extent]^2: 2 (res-1)^2 triangles."""
    h = (fbm(res, 6, rng) - 0.5) * height
    xs = np.linspace(-extent, extent, res)
    X, Z = np.meshgrid(xs, xs, indexing="xy")
    P = np.stack([X, h, Z], -1)
    dx = (xs[1] - xs[0])
    gx = np.gradient(h, dx, axis=1)
    gz = np.gradient(h, dx, axis=0)
    N = np.stack([-gx, np.ones_like(h), -gz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    U = np.stack([(X + extent) / (2 * extent) * uv_repeat, (Z + extent) / (2 * extent) * uv_repeat], -1)
    i0 = np.arange(res - 1)
    I, J = np.meshgrid(i0, i0, indexing="ij")
    a = (I * res + J).ravel()
    b = a + 1
    c = a + res
    d = c + 1
    idx = np.stack([np.stack([a, c, b], 1), np.stack([b, c, d], 1)], 1).reshape(-1)
    Pf, Nf, Uf = P.reshape(-1, 3), N.reshape(-1, 3), U.reshape(-1, 2)
    return Mesh(name="terrain", positions=Pf[idx].astype(np.float32), normals=Nf[idx].astype(np.float32),
                texcoord0=Uf[idx].astype(np.float32)), h, xs


def icosphere(subdiv=3, name="icosphere"):
    """subdivided icosahedron on the unit sphere: 20 * 4^subdiv triangles, spherical UVs."""
    t = (1 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
         (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    tris = np.array([[V[i] for i in f] for f in F], np.float64)
    tris /= np.linalg.norm(tris, axis=2, keepdims=True)
    for _ in range(subdiv):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tris = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                         np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    P = tris.reshape(-1, 3)
    uv = np.stack([np.arctan2(P[:, 2], P[:, 0]) / (2 * np.pi) + 0.5, np.arccos(np.clip(P[:, 1], -1, 1)) / np.pi], 1)
    return Mesh(name=name, positions=P.astype(np.float32), normals=P.astype(np.float32),
                texcoord0=uv.astype(np.float32))


def noise_texture(res, rng, tint):
    n = fbm(res, 5, rng, base_cells=8)[..., None]
    img = np.clip(n * np.asarray(tint) * 1.4, 0, 1)
    return (img * 255).astype(np.uint8)


def large_scene(seed=2, terrain_res=2048, n_spheres=5000, tex_res=2048, env_width=2048,
                env_height=1024, n_lights=16, extent=100.0):
    """The large-scene workload: fBm terrain (2 (res-1)^2 triangles), icospheres of 1280
    triangles at seeded positions with log-uniform radii in [0.2, 3], per-object roughness
    U[0.05, 1] and metallic Bernoulli(0.3), four value-noise base-colour textures, a
    metallic-roughness map and a tangent-space normal map, 16 point lights and the procedural
    sky; BVH leaves of at most 4 triangles, depth 32."""
    rng = np.random.default_rng(seed)
    terrain, h, xs = terrain_mesh(terrain_res, extent, height=20.0, rng=rng)
    sphere = icosphere(3)
    tints = [(0.6, 0.5, 0.35), (0.8, 0.3, 0.2), (0.3, 0.5, 0.8), (0.5, 0.8, 0.4)]
    images = [noise_texture(tex_res, rng, t) for t in tints]
    mr = np.zeros((tex_res, tex_res, 3), np.uint8)
    mr[..., 1] = (fbm(tex_res, 4, rng, 8) * 255).astype(np.uint8)   # roughness in G
    mr[..., 2] = (fbm(tex_res, 4, rng, 8) > 0.6).astype(np.uint8) * 255  # metallic in B
    images.append(mr)
    hn = fbm(tex_res, 5, rng, 16)
    gy, gx = np.gradient(hn)
    nrm = np.stack([-gx * 40, -gy * 40, np.ones_like(hn)], -1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    images.append(((nrm * 0.5 + 0.5) * 255).astype(np.uint8))
    materials = [Material("terrain", (1.0, 1.0, 1.0, 1.0), metallic_factor=1.0, roughness_factor=1.0,
                          base_color_tex=Texture(0), metallic_roughness_tex=Texture(4), normal_tex=Texture(5))]
    instances = [Instance(0, 0, np.eye(4, dtype=np.float32))]
    pos = rng.uniform(-extent * 0.9, extent * 0.9, (n_spheres, 2))
    radii = np.exp(rng.uniform(np.log(0.2), np.log(3.0), n_spheres))
    rough = rng.uniform(0.05, 1.0, n_spheres)
    metal = (rng.random(n_spheres) < 0.3).astype(np.float64)
    cols = rng.uniform(0.3, 1.0, (n_spheres, 3))
    step = xs[1] - xs[0]
    for i in range(n_spheres):
        gi = int(np.clip((pos[i, 1] + extent) / step, 0, terrain_res - 1))
        gj = int(np.clip((pos[i, 0] + extent) / step, 0, terrain_res - 1))
        y = h[gi, gj] + radii[i] * 0.6
        materials.append(Material(f"sphere_{i}", (float(cols[i, 0]), float(cols[i, 1]), float(cols[i, 2]), 1.0),
                                  metallic_factor=float(metal[i]), roughness_factor=float(rough[i]),
                                  base_color_tex=Texture(1 + i % 3)))
        instances.append(Instance(1, len(materials) - 1,
                                  trs((pos[i, 0], y, pos[i, 1]), float(rng.uniform(0, 360)), (radii[i],) * 3)))
    lights = []
    for i in range(n_lights):
        lx, lz = rng.uniform(-extent * 0.7, extent * 0.7, 2)
        col = rng.uniform(0.6, 1.0, 3) * 5000.0
        lights.append(Light(type=abi.MT_LIGHT_POINT, position=(float(lx), float(rng.uniform(15, 30)), float(lz)),
                            color=tuple(float(c) for c in col), point_size=0.5))
    cam = Camera(type=abi.MT_CAMERA_PERSPECTIVE,
                 camera_to_world=look_at((0.0, 35.0, extent * 1.05), (0.0, 0.0, 0.0)),
                 yfov=float(np.radians(50.0)), aspect=16.0 / 9.0)
    env = Environment(texels=procedural_sky(env_width, env_height), filter=abi.MT_FILTER_LINEAR)
    return Scene(meshes=[terrain, sphere], images=images, materials=materials, instances=instances,
                 lights=lights, camera=cam, env=env, bvh=BvhOptions(False, 4, 32))


def large_scene_settings(**kw):
    s = Settings(width=1920, height=1080, samples=64, bounces=5, hdri_rotation=30.0,
                 hdri_rotation_apply=True)
    for k, v in kw.items():
        setattr(s, k, v)
    return s
