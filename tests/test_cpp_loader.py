"""The C++ scene loader (include/mitrace_gpu_loader.hpp) through its test driver
tests/cpp/loader_shim.cpp: no GPU needed.

* The reference's glTF fixtures load to the same mt_scene_desc as the Python loader
  (mitrace_amd/gltf.py, which follows src/Loader/GLTF.cpp and src/Scene/Scene.cpp): meshes,
  images, materials, instances, lights and camera field by field.
* Its PNG decoder returns what stbi_load(..., 3) returns (Image.cpp:77-104) for PNGs written by
  an independent encoder in this file -- every colour type and bit depth, all five row filters,
  stored / fixed-Huffman / dynamic-Huffman deflate blocks, Adam7 interlacing -- and agrees with
  PIL where PIL decodes the same file.
* Its Radiance .hdr reader agrees with gltf.load_hdr on flat and run-length scanlines.
* Errors carry the reference's messages (data-URI buffers, float textures, no camera), and
  corrupt files are rejected.
A GPU test renders a fixture with the command-line renderer (mitrace_amd/bin/mitrace_gpu_cli)
and checks the PNG it writes against mt_render of the Python-loaded scene, byte for byte.
"""
import ctypes as C
import json
import os
import shutil
import struct
import subprocess
import zlib

import numpy as np
import pytest

from mitrace_amd import abi, gltf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
SHIM = os.path.join(CPP, "build", "libloader_shim.so")
FIX = os.path.join(ROOT, "tests", "golden", "fixtures")
FIXTURES = ["cornell_box_test/cornell_box.gltf", "material_spheres_test/material_spheres.gltf",
            "light_creep_test/light_creep_test.gltf"]

_shim = None


def shim():
    global _shim
    if _shim is None:
        subprocess.run(["make", "-s", "-C", CPP, "build/libloader_shim.so"], check=True)
        L = C.CDLL(SHIM)
        L.ldr_open.restype = C.c_void_p
        L.ldr_open.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_int]
        L.ldr_close.argtypes = [C.c_void_p]
        L.ldr_desc.argtypes = [C.c_void_p, C.POINTER(abi.mt_scene_desc), C.POINTER(abi.mt_camera_desc)]
        L.ldr_image.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.c_void_p, C.c_char_p, C.c_int]
        L.ldr_default_settings.argtypes = [C.POINTER(abi.mt_settings)]
        _shim = L
    return _shim


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    ct = {np.float32: C.c_float, np.uint8: C.c_uint8}[dtype]
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).copy()


def _tex(t):
    return (t.image, t.filter, tuple(t.offset), tuple(t.scale))


def read_desc(d, cam):
    """plain Python values of everything an mt_scene_desc and camera hold"""
    out = {"meshes": [], "images": [], "materials": [], "instances": [], "lights": []}
    for i in range(d.mesh_count):
        m = d.meshes[i]
        t = m.tri_count
        out["meshes"].append((_arr(m.positions, 9 * t, np.float32), _arr(m.normals, 9 * t, np.float32),
                              _arr(m.texcoord0, 6 * t, np.float32) if m.texcoord0 else None))
    for i in range(d.image_count):
        im = d.images[i]
        out["images"].append((im.width, im.height, _arr(im.rgb, im.width * im.height * 3, np.uint8)))
    for i in range(d.material_count):
        m = d.materials[i]
        out["materials"].append((tuple(m.base_color_factor), tuple(m.emissive_factor), m.normal_scale,
                                 m.metallic_factor, m.roughness_factor, m.occlusion_strength,
                                 _tex(m.base_color_tex), _tex(m.metallic_roughness_tex), _tex(m.normal_tex),
                                 _tex(m.occlusion_tex), _tex(m.emissive_tex)))
    for i in range(d.instance_count):
        x = d.instances[i]
        out["instances"].append((x.mesh, x.material, np.array(x.transform[:], np.float32)))
    for i in range(d.light_count):
        l = d.lights[i]
        out["lights"].append({"type": l.type, "position": tuple(l.position), "color": tuple(l.color),
                              "point_size": l.point_size, "inner": l.spot_inner_cos,
                              "outer": l.spot_outer_cos, "direction": np.array(l.direction[:])})
    e = d.env
    out["env"] = None
    if e.texels:
        n = e.width * e.height * 3
        dt = np.uint8 if e.format == abi.MT_ENV_RGB8 else np.float32
        out["env"] = (e.format, e.width, e.height, _arr(e.texels, n, dt))
    out["camera"] = {"type": cam.type, "m": np.array(cam.camera_to_world[:], np.float32), "yfov": cam.yfov,
                     "ortho": tuple(cam.ortho_size), "aspect": cam.aspect}
    return out


def load_cpp(path, scene=0, camera=0, ext_sky=False, env=""):
    L = shim()
    err = C.create_string_buffer(512)
    h = L.ldr_open(path.encode(), scene, camera, int(ext_sky), env.encode(), err, 512)
    if not h:
        raise RuntimeError(err.value.decode())
    try:
        d, cam = abi.mt_scene_desc(), abi.mt_camera_desc()
        L.ldr_desc(h, C.byref(d), C.byref(cam))
        return read_desc(d, cam)
    finally:
        L.ldr_close(h)


def load_py(path, **kw):
    sc = gltf.load(path, **kw)
    d, keep = sc.to_desc()
    out = read_desc(d, sc.camera.to_desc())
    del keep
    return out


def assert_same_scene(a, b, where):
    assert len(a["meshes"]) == len(b["meshes"]), where
    for k, (ma, mb) in enumerate(zip(a["meshes"], b["meshes"])):
        for x, y in zip(ma, mb):
            assert (x is None) == (y is None), (where, k)
            if x is not None:
                assert np.array_equal(x, y), (where, k)
    assert len(a["images"]) == len(b["images"]), where
    for ia, ib in zip(a["images"], b["images"]):
        assert ia[:2] == ib[:2] and np.array_equal(ia[2], ib[2]), where
    assert a["materials"] == b["materials"], where
    assert len(a["instances"]) == len(b["instances"]), where
    for ia, ib in zip(a["instances"], b["instances"]):
        assert ia[:2] == ib[:2], where
        # node products: glm's summation order here, numpy's in the Python loader
        np.testing.assert_allclose(ia[2], ib[2], rtol=1e-6, atol=1e-6, err_msg=where)
    assert len(a["lights"]) == len(b["lights"]), where
    for la, lb in zip(a["lights"], b["lights"]):
        for k in ("type", "position", "color", "point_size"):
            assert la[k] == lb[k], (where, k)
        for k in ("inner", "outer"):  # cosf here, numpy's float32 cos there
            assert abs(la[k] - lb[k]) <= 1e-6, (where, k)
        np.testing.assert_allclose(la["direction"], lb["direction"], atol=1e-6, err_msg=where)
    ca, cb = a["camera"], b["camera"]
    assert (ca["type"], ca["yfov"], ca["ortho"], ca["aspect"]) == (cb["type"], cb["yfov"], cb["ortho"], cb["aspect"])
    np.testing.assert_allclose(ca["m"], cb["m"], rtol=1e-6, atol=1e-6, err_msg=where)
    assert (a["env"] is None) == (b["env"] is None), where
    if a["env"] is not None:
        assert a["env"][:3] == b["env"][:3] and np.array_equal(a["env"][3], b["env"][3]), where


@pytest.mark.parametrize("fixture", FIXTURES)
def test_fixture_loads_like_the_python_loader(fixture):
    path = os.path.join(FIX, fixture)
    a, b = load_cpp(path), load_py(path)
    assert_same_scene(a, b, fixture)
    assert a["meshes"] and a["instances"] and a["lights"]
    # the fixtures' node transforms are reproduced exactly, not only within rounding
    for ia, ib in zip(a["instances"], b["instances"]):
        assert np.array_equal(ia[2], ib[2]), fixture


def test_material_spheres_texture_is_the_checkerboard_png():
    from PIL import Image
    a = load_cpp(os.path.join(FIX, FIXTURES[1]))
    assert len(a["images"]) == 1
    w, h, rgb = a["images"][0]
    with Image.open(os.path.join(FIX, "textures", "checkerboard.png")) as im:
        want = np.asarray(im.convert("RGB"))
    assert (h, w) == want.shape[:2] and np.array_equal(rgb.reshape(h, w, 3), want)


# ---------------------------------------------------------------------------------- PNG
def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c))


def _filter_rows(raw, bpp, kinds):
    """PNG row filters applied by an encoder (ISO/IEC 15948 section 9): raw is (rows, stride)
    uint8, kinds one filter type per row"""
    out = []
    prev = np.zeros(raw.shape[1], np.int32)
    for row, k in zip(raw.astype(np.int32), kinds):
        left = np.concatenate([np.zeros(bpp, np.int32), row[:-bpp]]) if bpp < len(row) else np.zeros_like(row)
        upleft = np.concatenate([np.zeros(bpp, np.int32), prev[:-bpp]]) if bpp < len(row) else np.zeros_like(row)
        pred = [np.zeros_like(row), left, prev, (left + prev) >> 1, _paeth(left, prev, upleft)][k]
        out.append(np.concatenate([[k], (row - pred) & 0xFF]).astype(np.uint8))
        prev = row
    return np.concatenate(out).tobytes() if out else b""


def _pack(samples, depth):
    """rows of samples (rows, n) to packed big-endian bytes per row"""
    if depth == 8:
        return samples.astype(np.uint8)
    if depth == 16:
        return samples.astype(">u2").view(np.uint8).reshape(samples.shape[0], -1)
    per = 8 // depth
    rows, n = samples.shape
    pad = (-n) % per
    s = np.concatenate([samples, np.zeros((rows, pad), samples.dtype)], 1).reshape(rows, -1, per)
    shifts = np.array([8 - depth * (i + 1) for i in range(per)])
    return (s.astype(np.uint32) << shifts).sum(-1).astype(np.uint8)


def write_png(path, samples, ctype, depth, palette=None, interlace=False, level=6, fixed=False, seed=0):
    """samples: (h, w, channels) integers; filters chosen at random per row"""
    rng = np.random.default_rng(seed)
    h, w, ch = samples.shape
    bpp = max(1, ch * depth // 8)

    def encode(img):
        if img.shape[0] == 0 or img.shape[1] == 0:
            return b""
        packed = _pack(img.reshape(img.shape[0], -1), depth)
        return _filter_rows(packed, bpp, rng.integers(0, 5, img.shape[0]))

    if not interlace:
        data = encode(samples)
    else:
        data = b""
        for x0, y0, dx, dy in [(0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4),
                               (1, 0, 2, 2), (0, 1, 1, 2)]:
            data += encode(samples[y0::dy, x0::dx])
    if fixed:
        co = zlib.compressobj(level, zlib.DEFLATED, 15, 9, zlib.Z_FIXED)
        z = co.compress(data) + co.flush()
    else:
        z = zlib.compress(data, level)

    def chunk(t, b):
        return struct.pack(">I", len(b)) + t + b + struct.pack(">I", zlib.crc32(t + b) & 0xffffffff)

    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, int(interlace)))
    if palette is not None:
        png += chunk(b"PLTE", palette.astype(np.uint8).tobytes())
    png += chunk(b"IDAT", z) + chunk(b"IEND", b"")
    with open(path, "wb") as f:
        f.write(png)


def expected_rgb8(samples, ctype, depth, palette=None):
    """what stbi_load(..., 3) returns"""
    s = samples.astype(np.int64)
    if depth == 16:
        s = s >> 8
    if ctype == 3:
        return palette[s[..., 0]].astype(np.uint8)
    if ctype in (0, 4):
        g = s[..., 0] * (255 // ((1 << depth) - 1)) if depth < 8 else s[..., 0]
        return np.repeat(g[..., None], 3, -1).astype(np.uint8)
    return s[..., :3].astype(np.uint8)


def load_png_cpp(path):
    L = shim()
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    if L.ldr_image(path.encode(), 0, C.byref(w), C.byref(h), None, err, 512):
        raise RuntimeError(err.value.decode())
    out = np.zeros((h.value, w.value, 3), np.uint8)
    assert L.ldr_image(path.encode(), 0, C.byref(w), C.byref(h), out.ctypes.data, err, 512) == 0
    return out


CASES = [  # (colour type, bit depth, channels)
    (0, 1, 1), (0, 2, 1), (0, 4, 1), (0, 8, 1), (0, 16, 1), (2, 8, 3), (2, 16, 3), (3, 1, 1),
    (3, 2, 1), (3, 4, 1), (3, 8, 1), (4, 8, 2), (4, 16, 2), (6, 8, 4), (6, 16, 4)]


@pytest.mark.parametrize("ctype,depth,ch", CASES)
def test_png_decoder_matches_stb_semantics(tmp_path, ctype, depth, ch):
    from PIL import Image
    rng = np.random.default_rng(ctype * 100 + depth)
    for (w, h), interlace, level, fixed in [((1, 1), False, 6, False), ((13, 7), False, 0, False),
                                            ((37, 29), False, 9, False), ((37, 29), True, 9, False),
                                            ((64, 3), False, 6, True), ((5, 11), True, 1, True)]:
        maxv = (1 << depth) - 1
        palette = None
        if ctype == 3:
            n = 1 << depth
            palette = rng.integers(0, 256, (n, 3))
        samples = rng.integers(0, maxv + 1, (h, w, ch))
        path = str(tmp_path / f"t_{ctype}_{depth}_{w}x{h}_{int(interlace)}_{level}_{int(fixed)}.png")
        write_png(path, samples, ctype, depth, palette, interlace, level, fixed, seed=w * h)
        got = load_png_cpp(path)
        want = expected_rgb8(samples, ctype, depth, palette)
        assert np.array_equal(got, want), path
        if depth == 8 or (depth < 8 and ctype == 3):  # where PIL's RGB conversion is stb's
            with Image.open(path) as im:
                assert np.array_equal(got, np.asarray(im.convert("RGB"))), path


def test_png_rejects_corrupt_files(tmp_path):
    path = str(tmp_path / "ok.png")
    write_png(path, np.zeros((4, 4, 3), np.uint8), 2, 8)
    data = bytearray(open(path, "rb").read())
    bad = bytearray(data)
    bad[40] ^= 0xFF  # inside IDAT: its CRC no longer matches
    (tmp_path / "crc.png").write_bytes(bytes(bad))
    with pytest.raises(RuntimeError, match="CRC"):
        load_png_cpp(str(tmp_path / "crc.png"))
    (tmp_path / "short.png").write_bytes(bytes(data[:30]))
    with pytest.raises(RuntimeError):
        load_png_cpp(str(tmp_path / "short.png"))
    (tmp_path / "x.hdr").write_bytes(b"#?RADIANCE\n")
    with pytest.raises(RuntimeError, match="Cannot load float image"):
        load_png_cpp(str(tmp_path / "x.hdr"))


# ---------------------------------------------------------------------------------- HDR
def _rgbe(px):
    m = px.max(-1)
    e = np.zeros(m.shape, np.int32)
    mant, ex = np.frexp(m)
    ok = m >= 1e-32
    scale = np.where(ok, mant * 256.0 / np.where(ok, m, 1.0), 0.0)
    out = np.zeros(px.shape[:2] + (4,), np.uint8)
    out[..., :3] = np.where(ok[..., None], (px * scale[..., None]).astype(np.int64), 0)
    out[..., 3] = np.where(ok, ex + 128, 0)
    return out


def _rle_scanline(row):
    """new-style run-length scanline: each component coded in turn, runs of >= 3 as runs"""
    w = row.shape[0]
    out = bytearray([2, 2, w >> 8, w & 0xFF])
    for c in range(4):
        v = row[:, c]
        i = 0
        while i < w:
            j = i
            while j < w and j - i < 127 and v[j] == v[i]:
                j += 1
            if j - i >= 3:
                out += bytes([128 + j - i, v[i]])
                i = j
            else:
                k = i
                while k < w and k - i < 128 and not (k + 2 < w and v[k] == v[k + 1] == v[k + 2]):
                    k += 1
                out += bytes([k - i]) + bytes(v[i:k])
                i = k
    return bytes(out)


@pytest.mark.parametrize("rle", [False, True])
def test_hdr_reader_matches_the_python_reader(tmp_path, rle):
    rng = np.random.default_rng(3)
    px = (rng.random((9, 40, 3)) ** 4 * 30.0).astype(np.float32)
    px[2, 5:20] = 1.5  # runs
    px[4] = 0.0
    e = _rgbe(px)
    body = b"".join(_rle_scanline(r) for r in e) if rle else e.tobytes()
    path = tmp_path / "t.hdr"
    path.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 9 +X 40\n" + body)
    L = shim()
    w, h = C.c_uint32(), C.c_uint32()
    err = C.create_string_buffer(512)
    out = np.zeros((9, 40, 3), np.float32)
    assert L.ldr_image(str(path).encode(), 1, C.byref(w), C.byref(h), out.ctypes.data, err, 512) == 0, err.value
    assert (w.value, h.value) == (40, 9)
    assert np.array_equal(out, gltf.load_hdr(str(path)))


# ---------------------------------------------------------------------------------- glTF
def _tiny_gltf(tmp_path, **over):
    """a one-triangle scene with a camera, written next to its buffer"""
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 1]] * 3, np.float32)
    idx = np.array([0, 1, 2], np.uint16)
    blob = pos.tobytes() + nrm.tobytes() + idx.tobytes() + b"\0\0"
    (tmp_path / "t.bin").write_bytes(blob)
    j = {"asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 1]}],
         "nodes": [{"mesh": 0, "translation": [1.0, 2.0, 3.0], "rotation": [0.0, 0.38268343, 0.0, 0.9238795],
                    "scale": [2.0, 1.0, 0.5]},
                   {"camera": 0, "children": [2], "matrix": [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 5, 1]},
                   {"extensions": {"KHR_lights_punctual": {"light": 0}}, "translation": [0, 3, 0],
                    "rotation": [0.70710677, 0.0, 0.0, 0.70710677]}],
         "meshes": [{"name": "tri", "primitives": [{"attributes": {"POSITION": 0, "NORMAL": 1}, "indices": 2,
                                                    "material": 0}]}],
         "materials": [{"pbrMetallicRoughness": {"baseColorFactor": [0.8, 0.2, 0.1, 1.0], "metallicFactor": 0.3},
                        "emissiveFactor": [1.0, 0.5, 0.25],
                        "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 3.0}}}],
         "accessors": [{"bufferView": 0, "componentType": 5126, "count": 3, "type": "VEC3"},
                       {"bufferView": 1, "componentType": 5126, "count": 3, "type": "VEC3"},
                       {"bufferView": 2, "componentType": 5123, "count": 3, "type": "SCALAR"}],
         "bufferViews": [{"buffer": 0, "byteOffset": 0, "byteLength": 36},
                         {"buffer": 0, "byteOffset": 36, "byteLength": 36},
                         {"buffer": 0, "byteOffset": 72, "byteLength": 6}],
         "buffers": [{"uri": "t.bin", "byteLength": len(blob)}],
         "cameras": [{"type": "perspective", "perspective": {"yfov": 0.7, "aspectRatio": 1.5, "znear": 0.1}}],
         "extensions": {"KHR_lights_punctual": {"lights": [{"type": "spot", "color": [1.0, 0.9, 0.8], "intensity": 20.0,
                                                            "spot": {"innerConeAngle": 0.3, "outerConeAngle": 0.6}}]}}}
    for k, v in over.items():
        j[k] = v
    path = tmp_path / "t.gltf"
    path.write_text(json.dumps(j))
    return str(path), j


def test_node_tree_lights_and_materials_like_the_python_loader(tmp_path):
    path, _ = _tiny_gltf(tmp_path)
    a, b = load_cpp(path), load_py(path)
    assert_same_scene(a, b, "tiny")
    assert a["lights"][0]["type"] == abi.MT_LIGHT_SPOT
    # KHR_materials_emissive_strength multiplies the factor
    assert a["materials"][0][1] == (np.float32(3.0), np.float32(1.5), np.float32(0.75))


def test_environment_override_and_ext_sky(tmp_path):
    from PIL import Image
    sky = (np.arange(8 * 4 * 3) % 251).astype(np.uint8).reshape(4, 8, 3)
    Image.fromarray(sky).save(tmp_path / "sky.png")
    path, _ = _tiny_gltf(tmp_path, images=[{"uri": "sky.png"}], textures=[{"source": 0}])
    j = json.loads(open(path).read())
    j["extensions"]["EXT_sky"] = {"sky_texture": 0}
    open(path, "w").write(json.dumps(j))
    assert load_cpp(path)["env"] is None  # Scene::Scene drops EXT_sky
    a = load_cpp(path, ext_sky=True)
    assert a["env"][:3] == (abi.MT_ENV_RGB8, 8, 4) and np.array_equal(a["env"][3], sky.ravel())
    assert_same_scene(a, load_py(path, ext_sky=True), "ext_sky")
    # a float .hdr override (the reference rejects float images; the GPU path takes them)
    px = np.full((2, 4, 3), 0.5, np.float32)
    (tmp_path / "o.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 2 +X 4\n" + _rgbe(px).tobytes())
    a = load_cpp(path, env=str(tmp_path / "o.hdr"))
    assert a["env"][:3] == (abi.MT_ENV_RGB32F, 4, 2) and np.array_equal(a["env"][3], px.ravel())


def test_errors_carry_the_reference_messages(tmp_path):
    path, j = _tiny_gltf(tmp_path)
    j2 = dict(j, buffers=[{"uri": "data:application/octet-stream;base64,AAAA", "byteLength": 3}])
    (tmp_path / "d.gltf").write_text(json.dumps(j2))
    with pytest.raises(RuntimeError, match="Data URI buffers are not supported"):
        load_cpp(str(tmp_path / "d.gltf"))
    j3 = dict(j, scenes=[{"nodes": [0]}])
    (tmp_path / "c.gltf").write_text(json.dumps(j3))
    with pytest.raises(RuntimeError, match="No camera found in the scene"):
        load_cpp(str(tmp_path / "c.gltf"))
    (tmp_path / "f.hdr").write_bytes(b"")
    j4 = dict(j, images=[{"uri": "f.hdr"}], textures=[{"source": 0}])
    j4["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": 0}}}]
    (tmp_path / "t4.gltf").write_text(json.dumps(j4))
    with pytest.raises(RuntimeError, match="Cannot load float image"):
        load_cpp(str(tmp_path / "t4.gltf"))
    (tmp_path / "bad.gltf").write_text("{\"asset\": ")
    with pytest.raises(RuntimeError, match="invalid JSON"):
        load_cpp(str(tmp_path / "bad.gltf"))


def test_cli_options_default_to_the_reference_config():
    s = abi.mt_settings()
    shim().ldr_default_settings(C.byref(s))
    assert (s.width, s.height, s.samples, s.bounces) == (1280, 720, 64, 5)
    assert (s.ev_exposure, s.hdri_primary_intensity, s.hdri_secondary_intensity) == (8.0, 350.0, 350.0)
    assert (s.firefly_enable, s.image_block_size, s.mode, s.output) == (0, 32, abi.MT_MODE_PARITY,
                                                                       abi.MT_OUTPUT_TONEMAPPED)
    assert abs(s.terminate_energy - 0.01) < 1e-9 and s.firefly_threshold == 6.0


@pytest.mark.gpu
def test_cli_renders_a_fixture_like_mt_render(gpu, tmp_path):
    """mitrace_gpu_cli on the Cornell box fixture writes the PNG of mt_render's image of the
    Python-loaded scene (same settings and seed), byte for byte"""
    from PIL import Image
    from mitrace_amd.build import CLI
    from mitrace_amd.scene import Settings
    from mitrace_amd.tracer import GpuScene
    assert os.path.exists(CLI), "the command-line renderer is not built"
    path = os.path.join(FIX, FIXTURES[0])
    out = str(tmp_path / "cli.png")
    r = subprocess.run([CLI, path, "-o", out, "-w", "96", "-h", "64", "-s", "16", "-b", "6", "-e", "6",
                        "--seed", "7", "-q"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = GpuScene(gltf.load(path))
    img, _, _ = g.render(Settings(width=96, height=64, samples=16, bounces=6, ev_exposure=6.0, seed=7))
    g.close()
    want = np.clip(img * np.float32(255.0), 0.0, 255.0).astype(np.uint8)
    with Image.open(out) as im:
        assert np.array_equal(np.asarray(im), want)
    # an unknown output type is refused before anything is loaded (main.cpp:90-95)
    r = subprocess.run([CLI, path, "-o", str(tmp_path / "x.bmp")], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "Invalid output file type" in r.stderr
