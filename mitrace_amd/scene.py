"""Host-side scene model mirroring the reference's ``Scene`` layer.

``Scene`` holds what ``Scene::Scene`` (src/Scene/Scene.cpp:53-59) builds from a loaded glTF:
de-indexed triangle-soup meshes (src/Scene/Mesh.cpp:24-41), per-instance materials and
model matrices (src/Scene/Mesh.hpp:44-72), lights (src/Scene/Light.hpp:36-53), the camera
(src/Scene/Camera.cpp:5-27), decoded 8-bit RGB images (src/Scene/Image.cpp:77-104) and an
optional environment map.  ``Scene.to_desc()`` packs it into the C-ABI ``mt_scene_desc``.

Matrices are kept as numpy (4, 4) float32 arrays in math convention (``M @ p``); the ABI
takes them column-major, which is ``M.T.ravel()``.
"""
from dataclasses import dataclass, field
from typing import List, Optional
import ctypes as C

import numpy as np

from . import abi


@dataclass
class Mesh:
    name: str
    positions: np.ndarray            # (3T, 3) float32 soup
    normals: np.ndarray              # (3T, 3) float32
    texcoord0: Optional[np.ndarray] = None  # (3T, 2) float32

    @property
    def tri_count(self):
        return self.positions.shape[0] // 3


@dataclass
class Texture:
    image: int = -1
    filter: int = abi.MT_FILTER_LINEAR
    offset: tuple = (0.0, 0.0)
    scale: tuple = (1.0, 1.0)


@dataclass
class Material:
    # Scene::Material defaults for a glTF material that is present but lacks factors
    # (GLTF.cpp:449-469); Loader::Material{} (no material) uses 0.5/0.5 (Types.hpp:65-82).
    name: str = "material"
    base_color_factor: tuple = (1.0, 1.0, 1.0, 1.0)
    emissive_factor: tuple = (0.0, 0.0, 0.0)
    normal_scale: float = 1.0
    metallic_factor: float = 1.0
    roughness_factor: float = 1.0
    occlusion_strength: float = 1.0
    base_color_tex: Texture = field(default_factory=Texture)
    metallic_roughness_tex: Texture = field(default_factory=Texture)
    normal_tex: Texture = field(default_factory=Texture)
    occlusion_tex: Texture = field(default_factory=Texture)
    emissive_tex: Texture = field(default_factory=Texture)


@dataclass
class Instance:
    mesh: int
    material: int
    transform: np.ndarray  # (4, 4) float32, model-to-world


@dataclass
class Light:
    type: int
    position: tuple
    color: tuple          # colour * intensity (Light.hpp:39)
    point_size: float = 0.1
    spot_inner_cos: float = 1.0
    spot_outer_cos: float = 0.70710678
    direction: tuple = (0.0, 0.0, -1.0)


@dataclass
class Camera:
    type: int
    camera_to_world: np.ndarray
    yfov: float = 0.0
    ortho_size: tuple = (0.0, 0.0)
    aspect: float = 1.0

    def to_desc(self):
        d = abi.mt_camera_desc()
        d.type = self.type
        d.camera_to_world[:] = [float(x) for x in np.asarray(self.camera_to_world, np.float32).T.ravel()]
        d.yfov = self.yfov
        d.ortho_size[:] = list(self.ortho_size)
        d.aspect = self.aspect
        return d


@dataclass
class Environment:
    texels: np.ndarray  # (H, W, 3) uint8 or float32
    filter: int = abi.MT_FILTER_LINEAR

    @property
    def format(self):
        return abi.MT_ENV_RGB8 if self.texels.dtype == np.uint8 else abi.MT_ENV_RGB32F


@dataclass
class BvhOptions:
    disable: bool = False
    max_triangles: int = 32
    max_depth: int = 32
    fast: bool = False  # also build the SAH BVHs of MT_TRAVERSAL_FAST
    # meshes with at least this many triangles get their reference-rule BVH built on the GPU
    # (bit-identical to the host build, about 3.5x faster scene creation on the large
    # terrain); 0 = every mesh on the host
    device_build_min_tris: int = 65536
    # with fast: meshes with at least this many triangles get the fast traversal's tree built on
    # the GPU as a linear BVH instead of the host's binned SAH; 0 = every mesh on the host
    fast_device_build_min_tris: int = 0


@dataclass
class Scene:
    meshes: List[Mesh] = field(default_factory=list)
    images: List[np.ndarray] = field(default_factory=list)  # (H, W, 3) uint8
    materials: List[Material] = field(default_factory=list)
    instances: List[Instance] = field(default_factory=list)
    lights: List[Light] = field(default_factory=list)
    camera: Optional[Camera] = None
    env: Optional[Environment] = None
    bvh: BvhOptions = field(default_factory=BvhOptions)

    @property
    def triangle_count(self):
        return sum(self.meshes[i.mesh].tri_count for i in self.instances)

    def to_desc(self):
        """Pack into an ``mt_scene_desc``; returns (desc, keepalive) -- keep both alive."""
        keep = []

        def arr(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a

        meshes = (abi.mt_mesh_desc * max(1, len(self.meshes)))()
        for i, m in enumerate(self.meshes):
            p = arr(m.positions, np.float32)
            n = arr(m.normals, np.float32)
            if p.shape != n.shape or p.shape[0] % 3:
                raise ValueError(f"mesh {m.name}: positions/normals must be (3T,3)")
            meshes[i].tri_count = p.shape[0] // 3
            meshes[i].positions = abi.fptr(p)
            meshes[i].normals = abi.fptr(n)
            if m.texcoord0 is not None:
                meshes[i].texcoord0 = abi.fptr(arr(m.texcoord0, np.float32))
        images = (abi.mt_image_desc * max(1, len(self.images)))()
        for i, im in enumerate(self.images):
            a = arr(im, np.uint8)
            if a.ndim != 3 or a.shape[2] != 3:
                raise ValueError("images must be (H, W, 3) uint8")
            images[i].width = a.shape[1]
            images[i].height = a.shape[0]
            images[i].rgb = a.ctypes.data_as(C.POINTER(C.c_uint8))

        def tex(t, out):
            out.image = t.image
            out.filter = t.filter
            out.offset[:] = list(t.offset)
            out.scale[:] = list(t.scale)

        mats = (abi.mt_material_desc * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            d = mats[i]
            d.base_color_factor[:] = list(m.base_color_factor)
            d.emissive_factor[:] = list(m.emissive_factor)
            d.normal_scale = m.normal_scale
            d.metallic_factor = m.metallic_factor
            d.roughness_factor = m.roughness_factor
            d.occlusion_strength = m.occlusion_strength
            tex(m.base_color_tex, d.base_color_tex)
            tex(m.metallic_roughness_tex, d.metallic_roughness_tex)
            tex(m.normal_tex, d.normal_tex)
            tex(m.occlusion_tex, d.occlusion_tex)
            tex(m.emissive_tex, d.emissive_tex)
        insts = (abi.mt_instance_desc * max(1, len(self.instances)))()
        for i, inst in enumerate(self.instances):
            insts[i].mesh = inst.mesh
            insts[i].material = inst.material
            insts[i].transform[:] = [float(x) for x in np.asarray(inst.transform, np.float32).T.ravel()]
        lights = (abi.mt_light_desc * max(1, len(self.lights)))()
        for i, l in enumerate(self.lights):
            d = lights[i]
            d.type = l.type
            d.position[:] = list(l.position)
            d.color[:] = list(l.color)
            d.point_size = l.point_size
            d.spot_inner_cos = l.spot_inner_cos
            d.spot_outer_cos = l.spot_outer_cos
            d.direction[:] = list(l.direction)
        desc = abi.mt_scene_desc()
        desc.meshes, desc.mesh_count = meshes, len(self.meshes)
        desc.images, desc.image_count = images, len(self.images)
        desc.materials, desc.material_count = mats, len(self.materials)
        desc.instances, desc.instance_count = insts, len(self.instances)
        desc.lights, desc.light_count = lights, len(self.lights)
        if self.env is not None:
            t = arr(self.env.texels, self.env.texels.dtype)
            desc.env.format = self.env.format
            desc.env.width = t.shape[1]
            desc.env.height = t.shape[0]
            desc.env.texels = t.ctypes.data
            desc.env.filter = self.env.filter
        desc.bvh.disable = int(self.bvh.disable)
        desc.bvh.max_triangles = self.bvh.max_triangles
        desc.bvh.max_depth = self.bvh.max_depth
        desc.bvh.fast = int(self.bvh.fast)
        desc.bvh.device_build_min_tris = int(self.bvh.device_build_min_tris)
        desc.bvh.fast_device_build_min_tris = int(self.bvh.fast_device_build_min_tris)
        keep.extend([meshes, images, mats, insts, lights])
        return desc, keep


@dataclass
class Settings:
    """The ``ConfigOptions`` fields the Tracer reads (src/CLI/Config.hpp:43-80) with the
    reference's defaults, plus the GPU-side additions of ``mt_settings``."""
    width: int = 1280
    height: int = 720
    samples: int = 64
    bounces: int = 5
    ev_exposure: float = 8.0
    hdri_rotation: float = 0.0
    hdri_secondary_intensity: float = 350.0
    hdri_primary_intensity: float = 350.0
    emission_base_intensity: float = 1.0
    # the CLI flag is inverted (Config.cpp:166-170, 239): the effective default is OFF
    firefly_enable: bool = False
    firefly_threshold: float = 6.0
    terminate_energy: float = 0.01
    image_block_size: int = 32
    seed: int = 0
    mode: int = abi.MT_MODE_PARITY
    rr_depth: int = 3
    output: int = abi.MT_OUTPUT_TONEMAPPED
    hdri_rotation_apply: bool = False
    paths_in_flight: int = 0
    count_work: bool = False
    debug_mode: int = 0  # MT_DEBUG_* (Config::DebugMode)
    traversal: int = abi.MT_TRAVERSAL_REFERENCE

    def to_desc(self):
        s = abi.mt_settings()
        for name, _ in abi.mt_settings._fields_:
            v = getattr(self, name)
            setattr(s, name, int(v) if isinstance(v, bool) else v)
        return s


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """camera-to-world matrix of a camera at ``eye`` looking at ``target`` along its -z."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    M = np.eye(4)
    M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3] = r, u, -f, eye
    return M.astype(np.float32)


def trs(translation=(0, 0, 0), rotation_y_deg=0.0, scale=(1, 1, 1)):
    """T * R_y * S in float32 (the composition order of GLTF.cpp:275-306)."""
    t = np.eye(4, dtype=np.float32)
    t[:3, 3] = translation
    a = np.radians(rotation_y_deg)
    r = np.eye(4, dtype=np.float32)
    r[0, 0], r[0, 2], r[2, 0], r[2, 2] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)
    s = np.diag([scale[0], scale[1], scale[2], 1.0]).astype(np.float32)
    return (t @ r @ s).astype(np.float32)
