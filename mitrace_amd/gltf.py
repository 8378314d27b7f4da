"""glTF 2.0 reader with the reference loader's semantics (src/Loader/GLTF.cpp).

It reads the subset the reference accepts -- external ``.bin`` buffers, triangle primitives
(mode 4) with indices, tightly packed FLOAT/UINT32/UINT16 accessors (byteStride ignored),
pbrMetallicRoughness materials with the reference's defaults (Types.hpp:65-82), KHR_texture_transform
offset/scale, KHR_lights_punctual and EXT_light_attributes area lights, perspective and
orthographic cameras -- and produces a :class:`mitrace_amd.scene.Scene`, i.e. what
``Scene::Scene(loader.LoadScene(i))`` plus ``SetCamera`` build in main.cpp:100-107.

The EXT_sky environment (doc/gltfs/EXT_sky.md) is parsed by the reference loader
(GLTF.cpp:902-912) but dropped by ``Scene::Scene`` (Scene.cpp:53-59), so by default it is not
read here and an environment comes only from ``hdri_override``; ``load(..., ext_sky=True)``
keeps it.  The reference loads 8-bit images only (Image.cpp:77-104 rejects ``.hdr``/``.exr``);
``load_env_image`` also reads Radiance ``.hdr`` files as float RGB, which the float
environment path of the GPU library takes.
"""
import json
import os

import numpy as np

from . import abi
from .scene import (BvhOptions, Camera, Environment, Instance, Light, Material, Mesh, Scene,
                    Texture)

_COMP = {5126: np.float32, 5123: np.uint16, 5125: np.uint32}
_NCOMP = {"SCALAR": 1, "VEC2": 2, "VEC3": 3, "VEC4": 4}


def _quat_to_mat3(q):
    """glm::mat4_cast of a unit quaternion (x, y, z, w), float32."""
    x, y, z, w = [np.float32(v) for v in q]
    two = np.float32(2)
    one = np.float32(1)
    return np.array([
        [one - two * (y * y + z * z), two * (x * y - w * z), two * (x * z + w * y)],
        [two * (x * y + w * z), one - two * (x * x + z * z), two * (y * z - w * x)],
        [two * (x * z - w * y), two * (y * z + w * x), one - two * (x * x + y * y)],
    ], dtype=np.float32)


def load_image_rgb8(path):
    """stbi_load(path, ..., 3) equivalent: 8-bit RGB, alpha dropped (Image.cpp:77-104)."""
    if path.endswith(".hdr") or path.endswith(".exr"):
        raise RuntimeError("Cannot load float image")  # Image.cpp:83-88
    from PIL import Image as PILImage
    with PILImage.open(path) as im:
        return np.asarray(im.convert("RGB"), dtype=np.uint8).copy()


def load_hdr(path):
    """Radiance RGBE (.hdr) as float32 (H, W, 3): '-Y H +X W' orientation, flat or new-style
    run-length encoded scanlines; a texel (m, e) decodes to m * 2^(e - 136), 0 when e = 0 (the
    decode stb_image uses)."""
    with open(path, "rb") as f:
        data = f.read()
    if not (data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")):
        raise RuntimeError("not a Radiance HDR file")
    pos = 0
    fmt_ok = False
    while True:  # header lines up to the blank line
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line == b"":
            break
        if line.startswith(b"FORMAT="):
            fmt_ok = line == b"FORMAT=32-bit_rle_rgbe"
    if not fmt_ok:
        raise RuntimeError("unsupported HDR format (only 32-bit_rle_rgbe)")
    end = data.index(b"\n", pos)
    tok = data[pos:end].split()
    pos = end + 1
    if len(tok) != 4 or tok[0] != b"-Y" or tok[2] != b"+X":
        raise RuntimeError("unsupported HDR orientation (only -Y H +X W)")
    h, w = int(tok[1]), int(tok[3])
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and \
                (data[pos + 2] << 8 | data[pos + 3]) == w:
            pos += 4
            for ch in range(4):  # each channel as runs: >128 = repeat, else literal bytes
                x = 0
                while x < w:
                    c = data[pos]
                    pos += 1
                    if c > 128:
                        n = c - 128
                        if n == 0 or x + n > w:
                            raise RuntimeError("bad HDR run length")
                        out[y, x:x + n, ch] = data[pos]
                        pos += 1
                    else:
                        if c == 0 or x + c > w:
                            raise RuntimeError("bad HDR literal length")
                        out[y, x:x + c, ch] = np.frombuffer(data, np.uint8, c, pos)
                        pos += c
                        n = c
                    x += n
        else:
            row = np.frombuffer(data, np.uint8, w * 4, pos)
            if w >= 1 and row[0] == 1 and row[1] == 1 and row[2] == 1:
                raise RuntimeError("old-style HDR run-length encoding is not supported")
            out[y] = row.reshape(w, 4)
            pos += w * 4
    e = out[..., 3].astype(np.int32)
    scale = np.where(e == 0, 0.0, np.ldexp(1.0, e - 136)).astype(np.float32)
    return (out[..., :3].astype(np.float32) * scale[..., None]).astype(np.float32)


def load_env_image(path):
    """An environment image: float RGB from a Radiance .hdr, else 8-bit RGB as the reference."""
    if path.lower().endswith(".hdr"):
        return load_hdr(path)
    return load_image_rgb8(path)


class GLTF:
    def __init__(self, path):
        self.path = path
        self.base = os.path.dirname(os.path.abspath(path))
        with open(path, "r") as f:
            self.j = json.load(f)
        self._buffers = {}
        self._meshes = {}
        self._images = {}     # path -> image index (Texture::imageCache, Texture.cpp:17-25)
        self.images = []

    # ---- buffers / accessors (GLTF.cpp:41-214)
    def _buffer(self, i):
        if i not in self._buffers:
            b = self.j["buffers"][i]
            if b["uri"].startswith("data:"):
                raise RuntimeError("Data URI buffers are not supported")
            with open(os.path.join(self.base, b["uri"]), "rb") as f:
                data = f.read(b["byteLength"])
            if len(data) < b["byteLength"]:
                raise RuntimeError("Failed to read buffer file")
            self._buffers[i] = data
        return self._buffers[i]

    def _accessor(self, i, expect_type, expect_comp):
        a = self.j["accessors"][i]
        if a["type"] not in _NCOMP or a["componentType"] not in _COMP:
            raise RuntimeError("Unsupported accessor type")
        if a["type"] != expect_type or a["componentType"] != expect_comp:
            raise RuntimeError("Attribute type does not match expected type")
        bv = self.j["bufferViews"][a["bufferView"]]
        data = self._buffer(bv["buffer"])
        n = _NCOMP[a["type"]]
        dt = _COMP[a["componentType"]]
        off = a.get("byteOffset", 0) + bv.get("byteOffset", 0)
        count = a["count"]
        arr = np.frombuffer(data, dtype=dt, count=count * n, offset=off)  # tightly packed
        return arr.reshape(count, n) if n > 1 else arr

    def _indices(self, i):
        a = self.j["accessors"][i]
        if a["type"] == "SCALAR" and a["componentType"] == 5125:
            return self._accessor(i, "SCALAR", 5125).astype(np.uint32)
        if a["type"] == "SCALAR" and a["componentType"] == 5123:
            return self._accessor(i, "SCALAR", 5123).astype(np.uint32)
        raise RuntimeError("Indices accessor has unsupported component type")

    # ---- meshes (GLTF.cpp:308-409; Mesh.cpp:9-35)
    def mesh(self, mi, pi):
        key = (mi, pi)
        if key in self._meshes:
            return self._meshes[key]
        m = self.j["meshes"][mi]
        prim = m["primitives"][pi]
        if "attributes" not in prim or "indices" not in prim:
            raise RuntimeError("Invalid GLTF primitive data: missing attributes or indices")
        if prim.get("mode", 4) != 4:
            raise RuntimeError("Only triangle primitives (mode 4) are supported")
        at = prim["attributes"]
        if "POSITION" not in at or "NORMAL" not in at:
            raise RuntimeError("POSITION and NORMAL attributes are required")
        idx = self._indices(prim["indices"])
        pos = self._accessor(at["POSITION"], "VEC3", 5126)
        nrm = self._accessor(at["NORMAL"], "VEC3", 5126)
        uv = self._accessor(at["TEXCOORD_0"], "VEC2", 5126) if "TEXCOORD_0" in at else None
        mesh = Mesh(name=m.get("name", "Unnamed_Mesh"),
                    positions=np.ascontiguousarray(pos[idx], np.float32),
                    normals=np.ascontiguousarray(nrm[idx], np.float32),
                    texcoord0=None if uv is None else np.ascontiguousarray(uv[idx], np.float32))
        self._meshes[key] = mesh
        return mesh

    # ---- textures and materials (GLTF.cpp:216-273, 411-506)
    def _image(self, ii):
        im = self.j["images"][ii]
        if "uri" not in im:
            raise RuntimeError("Only URI-based images are supported")
        path = os.path.normpath(os.path.join(self.base, im["uri"]))
        if path not in self._images:
            self._images[path] = len(self.images)
            self.images.append(load_image_rgb8(path))
        return self._images[path]

    def _texture(self, info):
        t = self.j["textures"][info["index"]]
        filt = abi.MT_FILTER_LINEAR
        if "sampler" in t:
            mf = self.j["samplers"][t["sampler"]].get("minFilter", 9729)
            table = {9728: 0, 9984: 0, 9985: 0, 9729: 1, 9986: 1, 9987: 1}
            if mf not in table:
                raise RuntimeError("Unsupported filter mode in texture sampler")
            filt = table[mf]
        if info.get("texCoord", 0) != 0:
            raise RuntimeError("Only TEXCOORD_0 is supported for texture info")
        tex = Texture(image=self._image(t["source"]), filter=filt)
        ext = info.get("extensions", {}).get("KHR_texture_transform")
        if ext is not None:
            if "rotation" in ext:
                raise RuntimeError("KHR_texture_transform rotation is not supported")
            if "texCoord" in ext:
                raise RuntimeError("KHR_texture_transform texCoord override is not supported")
            if "offset" in ext:
                tex.offset = tuple(float(v) for v in ext["offset"])
            if "scale" in ext:
                tex.scale = tuple(float(v) for v in ext["scale"])
        return tex

    def material(self, mi, pi):
        prim = self.j["meshes"][mi]["primitives"][pi]
        if "material" not in prim:
            return Material(name="empty", metallic_factor=0.5, roughness_factor=0.5,
                            occlusion_strength=0.0)  # Loader::Material{} (Types.hpp:65-82)
        md = self.j["materials"][prim["material"]]
        if "pbrMetallicRoughness" not in md:
            return Material(name="empty", metallic_factor=0.5, roughness_factor=0.5,
                            occlusion_strength=0.0)
        pbr = md["pbrMetallicRoughness"]
        mat = Material(name=md.get("name", "Unnamed_Material"))
        if "emissiveTexture" in md:
            mat.emissive_tex = self._texture(md["emissiveTexture"])
        em = np.array(md.get("emissiveFactor", [0.0, 0.0, 0.0]), np.float32)
        strength = md.get("extensions", {}).get("KHR_materials_emissive_strength")
        if strength is not None:
            em = em * np.float32(strength.get("emissiveStrength", 1.0))
        mat.emissive_factor = tuple(float(v) for v in em)
        mat.base_color_factor = tuple(float(v) for v in pbr.get("baseColorFactor", [1, 1, 1, 1]))
        mat.metallic_factor = float(pbr.get("metallicFactor", 1.0))
        mat.roughness_factor = float(pbr.get("roughnessFactor", 1.0))
        if "baseColorTexture" in pbr:
            mat.base_color_tex = self._texture(pbr["baseColorTexture"])
        if "metallicRoughnessTexture" in pbr:
            mat.metallic_roughness_tex = self._texture(pbr["metallicRoughnessTexture"])
        if "normalTexture" in md:
            mat.normal_tex = self._texture(md["normalTexture"])
        mat.normal_scale = float(md.get("normalScale", 1.0))  # material-level key (GLTF.cpp:480)
        if "occlusionTexture" in md:
            mat.occlusion_tex = self._texture(md["occlusionTexture"])
        mat.occlusion_strength = float(md.get("occlusionStrength", 1.0))
        if md.get("alphaMode", "OPAQUE") not in ("OPAQUE", "MASK", "BLEND"):
            raise RuntimeError("Unsupported alpha mode in material")
        return mat

    # ---- nodes (GLTF.cpp:275-306, 681-892)
    def node_transform(self, ni):
        nd = self.j["nodes"][ni]
        if "matrix" in nd:
            return np.array(nd["matrix"], np.float32).reshape(4, 4).T.copy()
        m = np.eye(4, dtype=np.float32)
        if "translation" in nd:
            m[:3, 3] = np.array(nd["translation"], np.float32)
        if "rotation" in nd:
            r = np.eye(4, dtype=np.float32)
            r[:3, :3] = _quat_to_mat3(nd["rotation"])
            m = (m @ r).astype(np.float32)
        if "scale" in nd:
            s = np.array(nd["scale"], np.float32)
            m[:, 0] *= s[0]
            m[:, 1] *= s[1]
            m[:, 2] *= s[2]
        return m

    def _walk(self, ni, parent, visit):
        world = (parent @ self.node_transform(ni)).astype(np.float32)
        visit(ni, world)
        for c in self.j["nodes"][ni].get("children", []):
            self._walk(c, world, visit)

    def _scene_nodes(self, si):
        return self.j["scenes"][si]["nodes"]

    def light(self, li, world):
        ld = self.j["extensions"]["KHR_lights_punctual"]["lights"][li]
        kinds = {"point": abi.MT_LIGHT_POINT, "directional": abi.MT_LIGHT_DIRECTIONAL,
                 "spot": abi.MT_LIGHT_SPOT, "area": abi.MT_LIGHT_AREA}
        if ld.get("type") not in kinds:
            raise RuntimeError("Unsupported light type in GLTF")
        t = kinds[ld["type"]]
        col = np.array(ld.get("color", [1.0, 1.0, 1.0]), np.float32) * np.float32(ld.get("intensity", 1.0))
        size = 0.1
        inner, outer = 0.0, 0.0
        if t == abi.MT_LIGHT_POINT:
            size = float(ld.get("size", 0.1))
        elif t == abi.MT_LIGHT_SPOT:
            size = float(ld.get("size", 0.0))
            inner = float(ld.get("spot", {}).get("innerConeAngle", 0.0))
            outer = float(ld.get("spot", {}).get("outerConeAngle", np.pi / 4))
        return self._make_light(t, world, col, size, inner, outer)

    @staticmethod
    def _make_light(t, world, col, size, inner, outer):
        d = -world[:3, 2].astype(np.float64)
        n = np.linalg.norm(d)
        d = d / n if n > 0 else np.array([0.0, 0.0, -1.0])
        return Light(type=t, position=tuple(float(v) for v in world[:3, 3]),
                     color=tuple(float(v) for v in col), point_size=size,
                     spot_inner_cos=float(np.cos(np.float32(inner))),
                     spot_outer_cos=float(np.cos(np.float32(outer))),
                     direction=tuple(float(v) for v in d))

    def area_light(self, ai, world):
        ld = self.j["extensions"]["EXT_light_attributes"]["areaLights"][ai]
        if "shape" not in ld:
            raise RuntimeError("Area light missing 'shape' property")
        if ld["shape"] not in ("rectangle", "disc"):
            raise RuntimeError("Unsupported area light shape")
        col = np.array(ld.get("color", [1.0, 1.0, 1.0]), np.float32) * np.float32(ld.get("intensity", 1.0))
        return self._make_light(abi.MT_LIGHT_AREA, world, col, 0.1, 0.0, 0.0)

    def camera(self, ci, world):
        cd = self.j["cameras"][ci]
        if cd["type"] == "perspective":
            p = cd["perspective"]
            return Camera(type=abi.MT_CAMERA_PERSPECTIVE, camera_to_world=world,
                          yfov=float(np.float32(p["yfov"])), aspect=float(p.get("aspectRatio", 1.0)))
        if cd["type"] == "orthographic":
            o = cd["orthographic"]
            return Camera(type=abi.MT_CAMERA_ORTHOGONAL, camera_to_world=world,
                          ortho_size=(float(o["xmag"]), float(o["ymag"])),
                          aspect=float(np.float32(o["xmag"]) / np.float32(o["ymag"])))
        raise RuntimeError("Unsupported camera type in GLTF")

    def load_scene(self, scene_index=0, camera_index=0):
        meshes, mesh_ids, instances, materials, lights, cameras = [], {}, [], [], [], []

        def visit(ni, world):
            nd = self.j["nodes"][ni]
            if "mesh" in nd:
                mi = nd["mesh"]
                for pi in range(len(self.j["meshes"][mi]["primitives"])):
                    m = self.mesh(mi, pi)
                    if id(m) not in mesh_ids:  # dedupe by mesh (Scene.cpp:14-22)
                        mesh_ids[id(m)] = len(meshes)
                        meshes.append(m)
                    materials.append(self.material(mi, pi))
                    instances.append(Instance(mesh=mesh_ids[id(m)], material=len(materials) - 1,
                                              transform=world))
            ext = nd.get("extensions", {})
            if "KHR_lights_punctual" in ext:
                lights.append(self.light(ext["KHR_lights_punctual"]["light"], world))
            if "areaLight" in ext.get("EXT_light_attributes", {}):
                lights.append(self.area_light(ext["EXT_light_attributes"]["areaLight"], world))
            if "camera" in nd:
                cameras.append(self.camera(nd["camera"], world))

        # the reference walks meshes, then lights, then cameras, each in DFS node order; the
        # per-kind order is the same in a single walk
        for ni in self._scene_nodes(scene_index):
            self._walk(ni, np.eye(4, dtype=np.float32), visit)
        if not cameras:
            raise RuntimeError("No camera found in the scene")
        cam = cameras[camera_index] if camera_index < len(cameras) else cameras[0]
        return Scene(meshes=meshes, images=self.images, materials=materials,
                     instances=instances, lights=lights, camera=cam)


def load(path, scene_index=0, camera_index=0, hdri_override=None, bvh=None, ext_sky=False):
    """Loader::GLTF(path).LoadScene + Scene::Scene + optional --hdri-override.  ext_sky=True
    keeps the EXT_sky texture (LoadSceneEnvironmentTexture, GLTF.cpp:902-912) as the
    environment, which Scene::Scene drops; hdri_override still takes precedence, and the
    equirect is always sampled bilinearly (Texture.cpp:35-50)."""
    g = GLTF(path)
    sc = g.load_scene(scene_index, camera_index)
    if ext_sky and hdri_override is None:
        sky = g.j.get("extensions", {}).get("EXT_sky", {})
        if "sky_texture" in sky:
            tex = g.j["textures"][sky["sky_texture"]]
            im = g.j["images"][tex["source"]]
            if "uri" not in im:
                raise RuntimeError("Only URI-based images are supported")
            sc.env = Environment(texels=load_env_image(os.path.join(g.base, im["uri"])),
                                 filter=abi.MT_FILTER_LINEAR)
    if hdri_override is not None:
        sc.env = Environment(texels=load_env_image(hdri_override), filter=abi.MT_FILTER_LINEAR)
    if bvh is not None:
        sc.bvh = bvh
    return sc
