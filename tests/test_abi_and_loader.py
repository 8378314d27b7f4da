"""CPU checks of the boundary: the built library exports every function include/mitrace_gpu.h
declares with matching struct layouts, and the glTF reader reproduces the reference loader's
view of the reference's own test scenes (no GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from helpers import fixture_path, render_params
from mitrace_amd import abi, gltf, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mitrace_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(mt_[a-z0-9_]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)
    assert L.mt_abi_version() == 3


def test_struct_layouts_match_the_header():
    structs = ["mt_mesh_desc", "mt_image_desc", "mt_texture_desc", "mt_material_desc",
               "mt_instance_desc", "mt_light_desc", "mt_camera_desc", "mt_env_desc",
               "mt_bvh_options", "mt_scene_desc", "mt_settings", "mt_stats", "mt_hit",
               "mt_diff_params", "mt_diff_grads", "mt_batch_options", "mt_batch_job_info",
               "mt_batch_stats", "mt_autotune_trial", "mt_cache_options"]
    src = '#include <stdio.h>\n#include <stddef.h>\n#include "%s"\nint main(void){\n' % HEADER
    for s in structs:
        src += f'printf("{s} %zu\\n", sizeof({s}));\n'
    src += 'printf("mt_stats.diag %zu\\n", offsetof(mt_stats, diag));\nreturn 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sizes.c")
        exe = os.path.join(d, "sizes")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c99", c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")
    sizes = dict(line.split() for line in out if line)
    for s in structs:
        assert int(sizes[s]) == C.sizeof(getattr(abi, s)), s
    assert int(sizes["mt_stats.diag"]) == abi.mt_stats.diag.offset


def test_errors_without_device_are_reported_not_raised():
    L = _lib.load()
    n = C.c_int(-1)
    rc = L.mt_device_count(C.byref(n))
    assert rc in (0, abi.MT_E_HIP)
    if n.value == 0 or rc != 0:
        h = C.c_void_p()
        assert L.mt_scene_create(None, C.byref(h)) != 0
        assert L.mt_last_error()


def test_cornell_fixture_loads_like_the_reference():
    sc = gltf.load(fixture_path("cornell_box"))
    # 3 glTF meshes; the walls mesh has 4 primitives -> 6 instances
    assert len(sc.instances) == 6 and len(sc.meshes) == 6
    assert sum(m.tri_count for m in sc.meshes) == 42
    # EXT_light_attributes area light -> a point light of zero jitter at the node (GLTF.cpp:558-571)
    (l,) = sc.lights
    assert l.type == abi.MT_LIGHT_AREA
    np.testing.assert_allclose(l.position, (0, 1.99, 1.0550744533538818), rtol=1e-6)
    assert l.color == (543.0, 543.0, 543.0)
    cam = sc.camera
    assert cam.type == abi.MT_CAMERA_PERSPECTIVE and cam.yfov == pytest.approx(0.6911112070083618)
    np.testing.assert_allclose(cam.camera_to_world[:3, 3], (0, 1, -3))
    # material defaults (Types.hpp:65-82): present without textures
    red = [m for m in sc.materials if m.name == "WallRed"][0]
    assert red.metallic_factor == 0 and red.roughness_factor == 0.5
    assert render_params("cornell_box") == {"width": 512, "height": 512, "samples": 512, "ev-exposure": 6}


def test_material_spheres_fixture():
    sc = gltf.load(fixture_path("material_spheres"))
    assert len(sc.instances) == 17 and len(sc.lights) == 4 and len(sc.images) == 1
    assert sc.images[0].shape == (256, 256, 3)
    assert sum(sc.meshes[i.mesh].tri_count for i in sc.instances) == 7214
    m = sc.materials[3]  # metallicFactor absent -> 1.0 (GLTF.cpp:468)
    assert m.metallic_factor == 1.0 and m.roughness_factor == 0.0
    tex = sc.materials[0].base_color_tex
    assert tex.image == 0 and tex.filter == abi.MT_FILTER_LINEAR  # minFilter 9987
    assert sc.lights[0].color == pytest.approx((2717.5706532941126, 2717.5706532941126 * 0.85, 2717.5706532941126 * 0.7), rel=1e-6)
    assert sc.lights[0].point_size == 0.1  # KHR_lights_punctual "size" default (GLTF.cpp:558-571)


def test_light_creep_orthographic_camera():
    sc = gltf.load(fixture_path("light_creep"))
    assert sc.camera.type == abi.MT_CAMERA_ORTHOGONAL
    assert sc.camera.ortho_size == pytest.approx((0.85, 0.478125), rel=1e-6)
    assert sc.camera.aspect == pytest.approx(0.85 / 0.478125, rel=1e-6)


def _write_hdr(path, w, rows):
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n-Y %d +X %d\n" % (len(rows), w))
        for r in rows:
            f.write(bytes(r))


def test_hdr_reader_flat_and_run_length_scanlines(tmp_path):
    """Radiance RGBE: a texel (m, e) is m * 2^(e - 136) (stb_image's decode), e = 0 is black.
    Row 0 is a new-style run-length scanline (2, 2, width; per channel: a byte above 128 repeats
    the next byte, a byte up to 128 counts literal bytes), row 1 a flat one."""
    w = 10
    rle = [2, 2, 0, w,
           128 + 10, 200,                          # R: one run of 200
           10, *range(10),                         # G: ten literals 0..9
           128 + 5, 7, 5, 1, 2, 3, 4, 5,           # B: run of 7, then 5 literals
           128 + 9, 129, 1, 0]                     # E: run of 129, then a literal 0 (black)
    flat = []
    for x in range(w):
        flat += [x * 10, 255 - x, 3, 140 if x % 2 else 0]
    p = tmp_path / "t.hdr"
    _write_hdr(p, w, [rle, flat])
    img = gltf.load_hdr(str(p))
    assert img.shape == (2, w, 3) and img.dtype == np.float32
    s = 2.0 ** (129 - 136)
    want0 = np.stack([np.full(w, 200.0), np.arange(10.0), np.array([7] * 5 + [1, 2, 3, 4, 5], float)], 1) * s
    want0[9] = 0.0  # exponent byte 0
    assert np.array_equal(img[0], want0.astype(np.float32))
    for x in range(w):
        e = 140 if x % 2 else 0
        sc = 2.0 ** (e - 136) if e else 0.0
        assert np.array_equal(img[1, x], np.float32([x * 10 * sc, (255 - x) * sc, 3 * sc]))
    q = tmp_path / "bad.hdr"
    q.write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\0\0\0\0")
    with pytest.raises(RuntimeError):
        gltf.load_hdr(str(q))


def test_ext_sky_is_dropped_by_default_and_kept_on_request(tmp_path):
    """EXT_sky (doc/gltfs/EXT_sky.md): the reference loader reads sky_texture
    (GLTF.cpp:902-912) and Scene::Scene drops it (Scene.cpp:53-59).  load() does the same by
    default; ext_sky=True keeps it as the environment (8-bit PNG or float Radiance .hdr)."""
    import json
    from PIL import Image
    src = fixture_path("cornell_box")
    j = json.load(open(src))
    base = os.path.dirname(os.path.abspath(src))
    for b in j["buffers"]:
        b["uri"] = os.path.join(base, b["uri"])
    for im in j.get("images", []):
        im["uri"] = os.path.join(base, im["uri"])
    sky = (np.arange(4 * 8 * 3) % 251).astype(np.uint8).reshape(4, 8, 3)
    Image.fromarray(sky).save(tmp_path / "sky.png")
    _write_hdr(tmp_path / "sky.hdr", 2, [[128, 64, 32, 129, 0, 0, 0, 0]])
    for name, want in (("sky.png", sky), ("sky.hdr", np.float32([[[1.0, 0.5, 0.25], [0, 0, 0]]]))):
        jj = json.loads(json.dumps(j))
        jj.setdefault("images", []).append({"uri": str(tmp_path / name)})
        jj.setdefault("textures", []).append({"source": len(jj["images"]) - 1})
        jj.setdefault("extensions", {})["EXT_sky"] = {"sky_texture": len(jj["textures"]) - 1}
        path = tmp_path / "scene.gltf"
        path.write_text(json.dumps(jj))
        assert gltf.load(str(path)).env is None
        env = gltf.load(str(path), ext_sky=True).env
        assert env.texels.dtype == want.dtype and np.array_equal(env.texels, want)
        assert env.filter == abi.MT_FILTER_LINEAR
