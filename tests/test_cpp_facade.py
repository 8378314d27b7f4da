"""The C++20 host side (include/mitrace_gpu_tracer.hpp) through its test driver
tests/cpp/facade_shim.cpp.

CPU: the driver builds with g++ -Werror against the library and loads; DefaultSettings holds
the reference's Config defaults; RenderBuffer's RGB8 conversion and DrawSubBuffer follow
RenderBuffer.cpp:3-58 (checked against values fixed by hand and against the Python mirror);
a failing C call surfaces as mitrace_gpu::Error with the C-ABI's code and message;
SaveToFilePNG / SaveToFileHDR write files that decode (PIL, zlib, the Radiance reader) to the
buffer's pixels, with known RGBE bytes pinned by hand; SaveToFileJPG writes baseline JFIF that
PIL decodes, with libjpeg's quantisation tables for the reference's effective quality.
GPU: the facade's Start/Wait render equals the synchronous mt_render of the same seed bit for
bit and the oracle within 3 sigma; StatThread-style polling sees monotone progress; a kill
ends the render early without an error and a second StartRender completes.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from helpers import small_settings
from mitrace_amd import abi, scenegen
from mitrace_amd._lib import load
from mitrace_amd.scene import Settings
from mitrace_amd.tracer import GpuScene, RenderBuffer
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
SHIM = os.path.join(CPP, "build", "libfacade_shim.so")

_shim = None


def shim():
    global _shim
    if _shim is None:
        load()  # the product library first (fails loudly when it is not built)
        subprocess.run(["make", "-s", "-C", CPP], check=True)
        L = C.CDLL(SHIM)
        f = C.POINTER(C.c_float)
        L.fac_default_settings.argtypes = [C.POINTER(abi.mt_settings)]
        L.fac_render.argtypes = [C.POINTER(abi.mt_scene_desc), C.POINTER(abi.mt_camera_desc),
                                 C.POINTER(abi.mt_settings), f, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_int), C.c_char_p, C.c_int]
        L.fac_kill.argtypes = [C.POINTER(abi.mt_scene_desc), C.POINTER(abi.mt_camera_desc),
                               C.POINTER(abi.mt_settings), C.POINTER(C.c_uint64), C.c_char_p,
                               C.c_int]
        L.fac_scene_error.argtypes = [C.POINTER(abi.mt_scene_desc), C.c_char_p, C.c_int]
        L.fac_rgb8.argtypes = [f, C.c_uint, C.c_uint, C.POINTER(C.c_uint8)]
        L.fac_draw_sub.argtypes = [f, C.c_uint, C.c_uint, f, C.c_uint, C.c_uint, C.c_uint,
                                   C.c_uint, f]
        _shim = L
    return _shim


def fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_default_settings_are_the_reference_config_defaults():
    s = abi.mt_settings()
    shim().fac_default_settings(C.byref(s))
    # src/CLI/Config.hpp:43-80: 1280x720, 64 spp, 5 bounces, ev 8, HDRI intensities 350,
    # emission 1, firefly threshold 6 (flag effectively off), energy cut-off 0.01, 32-px tiles
    assert (s.width, s.height, s.samples, s.bounces, s.image_block_size) == (1280, 720, 64, 5, 32)
    assert (s.ev_exposure, s.hdri_primary_intensity, s.hdri_secondary_intensity) == (8.0, 350.0, 350.0)
    assert s.emission_base_intensity == 1.0 and s.firefly_enable == 0
    assert s.firefly_threshold == 6.0 and abs(s.terminate_energy - 0.01) < 1e-9
    assert (s.mode, s.output, s.traversal, s.debug_mode) == (
        abi.MT_MODE_PARITY, abi.MT_OUTPUT_TONEMAPPED, abi.MT_TRAVERSAL_REFERENCE, abi.MT_DEBUG_NONE)
    py = Settings().to_desc()  # the Python mirror holds the same defaults
    for name, _ in abi.mt_settings._fields_:
        assert getattr(s, name) == pytest.approx(getattr(py, name)), name


def test_render_buffer_rgb8_truncates_and_clamps():
    vals = np.array([-1.0, 0.0, 0.5, 1.0 / 255.0, 0.99999, 1.0, 2.0, 254.9 / 255.0, 0.1],
                    np.float32)
    px = np.tile(vals, 2).reshape(2, 3, 3)  # 3 x 2 pixels
    out = np.zeros(px.size, np.uint8)
    shim().fac_rgb8(fp(px), 3, 2, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    # trunc(clamp(v*255)), no rounding and no sRGB curve
    want = np.tile(np.array([0, 0, 127, 1, 254, 255, 255, 254, 25], np.uint8), 2)
    assert np.array_equal(out, want)
    rb = RenderBuffer(3, 2)
    rb.pixels[:] = px
    assert np.array_equal(out.reshape(2, 3, 3), rb.get_pixels_rgb8())


def test_render_buffer_draw_sub_buffer_clips_at_the_edges():
    rng = np.random.default_rng(3)
    dst = rng.random((5, 7, 3), dtype=np.float32)
    src = rng.random((3, 4, 3), dtype=np.float32)
    for ox, oy in [(0, 0), (2, 1), (5, 3), (7, 5), (6, 4)]:
        out = np.zeros_like(dst)
        shim().fac_draw_sub(fp(dst), 7, 5, fp(src), 4, 3, ox, oy, fp(out))
        want = dst.copy()
        h = max(0, min(3, 5 - oy))
        w = max(0, min(4, 7 - ox))
        want[oy:oy + h, ox:ox + w] = src[:h, :w]
        assert np.array_equal(out, want), (ox, oy)


def test_failing_call_throws_error_with_code_and_message():
    """A bad description (here: no meshes array for one mesh) fails in mt_scene_create; with no
    GPU the call fails at the device query instead.  Either way the facade throws Error with the
    C-ABI's code and mt_last_error's text."""
    d = abi.mt_scene_desc()
    d.mesh_count = 1  # meshes stays NULL
    err = C.create_string_buffer(512)
    code = shim().fac_scene_error(C.byref(d), err, 512)
    msg = err.value.decode()
    assert code in (abi.MT_E_INVALID, abi.MT_E_HIP), (code, msg)
    assert msg.startswith("mt_scene_create failed (%d)" % code)
    want = "meshes is NULL" if code == abi.MT_E_INVALID else "no HIP device"
    assert want in msg, msg


def _cornell(w=48, h=40, spp=16, **kw):
    sc = scenegen.cornell()
    st = small_settings(width=w, height=h, samples=spp, bounces=8, ev_exposure=2.0, seed=11)
    for k, v in kw.items():
        setattr(st, k, v)
    return sc, st


@pytest.mark.gpu
def test_facade_render_equals_mt_render_and_oracle(gpu):
    sc, st = _cornell()
    desc, keep = sc.to_desc()
    cam = sc.camera.to_desc()
    sd = st.to_desc()
    rgb = np.zeros((st.height, st.width, 3), np.float32)
    stats = (C.c_uint64 * 2)()
    polls = (C.c_int * 3)()
    err = C.create_string_buffer(512)
    rc = shim().fac_render(C.byref(desc), C.byref(cam), C.byref(sd), fp(rgb), stats, polls, err, 512)
    assert rc == 0, err.value.decode()
    del keep
    n = st.width * st.height * st.samples
    assert stats[1] == n and stats[0] >= n
    assert polls[0] >= 1 and polls[1] == polls[0], list(polls)  # progress in [0,1], monotone
    assert polls[2] == 1  # IsDone false before StartRender, true after WaitForRender

    g = GpuScene(sc)
    ref, _, s2 = g.render(st)
    g.close()
    assert s2["rays"] == stats[0]
    assert np.array_equal(rgb, ref.reshape(rgb.shape))  # same library, same seed: bit for bit

    o = po.OracleScene(sc)
    orgb, osamples, _, _ = o.render(st, want_samples=True)
    o.close()
    sigma = osamples.std(axis=1) / np.sqrt(st.samples)
    ok = (np.abs(rgb.reshape(-1, 3) - orgb.reshape(-1, 3)) <= 3 * sigma + 1e-4).mean()
    assert ok > 0.99, ok


@pytest.mark.gpu
def test_facade_kill_ends_early_and_restarts(gpu):
    # 64 chunks of 16 Ki paths, so a kill after the first chunk leaves most of the frame undone
    sc, st = _cornell(w=128, h=128, spp=64, paths_in_flight=16384)
    desc, keep = sc.to_desc()
    cam = sc.camera.to_desc()
    sd = st.to_desc()
    out = (C.c_uint64 * 3)()
    err = C.create_string_buffer(512)
    rc = shim().fac_kill(C.byref(desc), C.byref(cam), C.byref(sd), out, err, 512)
    assert rc == 0, err.value.decode()
    del keep
    full = st.width * st.height * st.samples
    assert out[0] < full, (out[0], full)
    assert out[1] == 1  # IsDone after a killed render's WaitForRender
    assert out[2] == 1  # the next StartRender rendered every sample


def _save(px, path, which):
    h, w, _ = px.shape
    px = np.ascontiguousarray(px, np.float32)
    err = C.create_string_buffer(256)
    L = shim()
    L.fac_save.argtypes = [C.POINTER(C.c_float), C.c_uint, C.c_uint, C.c_char_p, C.c_int,
                           C.c_char_p, C.c_int]
    return L.fac_save(fp(px), w, h, str(path).encode(), which, err, 256), err.value.decode()


def test_save_png_decodes_to_the_rgb8_pixels(tmp_path):
    """SaveToFilePNG (RenderBuffer.cpp:60-66): a standard PNG whose pixels are GetPixelsRGB8's.
    Decoded twice: by PIL and by hand (chunk CRCs, zlib stream, filter bytes)."""
    import struct
    import zlib
    from PIL import Image
    rng = np.random.default_rng(5)
    for w, h in [(1, 1), (7, 3), (300, 250)]:  # 300x250: the scanlines span several stored blocks
        px = (rng.random((h, w, 3)) * 1.4 - 0.2).astype(np.float32)
        path = tmp_path / f"img_{w}x{h}.png"
        rc, msg = _save(px, path, 0)
        assert rc == 0, msg
        want = np.clip(px * np.float32(255.0), 0.0, 255.0).astype(np.uint8)
        with Image.open(path) as im:
            assert im.mode == "RGB" and im.size == (w, h)
            assert np.array_equal(np.asarray(im), want)
        data = path.read_bytes()
        assert data[:8] == b"\x89PNG\r\n\x1a\n"
        pos, idat = 8, b""
        while pos < len(data):
            n, typ = struct.unpack(">I4s", data[pos:pos + 8])
            body = data[pos + 8:pos + 8 + n]
            crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
            assert crc == zlib.crc32(typ + body) & 0xffffffff, typ
            if typ == b"IDAT":
                idat += body
            pos += 12 + n
        raw = zlib.decompress(idat)  # checks the Adler-32 as well
        rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 3 * w)
        assert (rows[:, 0] == 0).all()
        assert np.array_equal(rows[:, 1:].reshape(h, w, 3), want)


def test_save_hdr_rgbe_round_trip(tmp_path):
    """SaveToFileHDR (RenderBuffer.cpp:77-84): Radiance RGBE.  Known texels: (1, 0.5, 0.25)
    shares exponent 1 (frexp(1) = 0.5 * 2^1), mantissas 128/64/32, stored exponent 129; black
    and negative maxima encode as zeros.  Random texels decode within the format's precision
    (a component loses less than 2^-8 of the texel's largest component)."""
    from mitrace_amd.gltf import load_hdr
    px = np.array([[[1.0, 0.5, 0.25], [0.0, 0.0, 0.0], [-1.0, -2.0, -3.0], [3.0, 0.0, 1e-3]]],
                  np.float32)
    path = tmp_path / "k.hdr"
    rc, msg = _save(px, path, 1)
    assert rc == 0, msg
    data = path.read_bytes()
    head = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 4\n"
    assert data.startswith(head)
    body = np.frombuffer(data[len(head):], np.uint8).reshape(4, 4)
    assert body[0].tolist() == [128, 64, 32, 129]
    assert body[1].tolist() == [0, 0, 0, 0] and body[2].tolist() == [0, 0, 0, 0]
    assert body[3].tolist() == [192, 0, 0, 130]  # 3 = 0.75 * 2^2
    back = load_hdr(str(path))
    assert np.array_equal(back[0, 0], [1.0, 0.5, 0.25]) and not back[0, 1:3].any()
    rng = np.random.default_rng(9)
    px = (rng.random((37, 53, 3)) ** 3 * np.float32(50.0)).astype(np.float32)
    rc, msg = _save(px, tmp_path / "r.hdr", 1)
    assert rc == 0, msg
    back = load_hdr(str(tmp_path / "r.hdr"))
    mx = px.max(axis=2, keepdims=True)
    assert back.shape == px.shape
    assert (back <= px).all() and (px - back <= mx / 128.0).all()


def test_save_jpg_decodes_close_to_the_rgb8_pixels(tmp_path):
    """SaveToFileJPG (RenderBuffer.cpp:69-75): a baseline JFIF file that PIL (libjpeg) decodes.
    The reference hands stb the row stride (width * 3) as the quality, so the quality is
    min(100, 3 * width): from 34 pixels wide every quantisation step is 1 and the decoded
    pixels stay within a few levels of GetPixelsRGB8's; narrower images carry libjpeg's own
    tables for that quality (the IJG rule over the T.81 Annex K.1 tables) and lose no more than
    libjpeg's own 4:4:4 encoding at that quality.  Edge blocks (sizes not multiples of 8)
    decode to the image's size."""
    import io
    from PIL import Image
    rng = np.random.default_rng(11)
    yy, xx = np.mgrid[0:250, 0:300]
    smooth = np.stack([xx / 300.0, yy / 250.0, 0.5 + 0.5 * np.sin(xx / 17.0 + yy / 23.0)], -1)
    cases = [(smooth + rng.normal(scale=0.05, size=smooth.shape)).astype(np.float32),
             (rng.random((37, 50, 3)) * 1.4 - 0.2).astype(np.float32),
             rng.random((3, 7, 3)).astype(np.float32), rng.random((1, 1, 3)).astype(np.float32),
             smooth[:40, :10].astype(np.float32), smooth[:9, :20].astype(np.float32)]
    for px in cases:
        h, w, _ = px.shape
        path = tmp_path / f"img_{w}x{h}.jpg"
        rc, msg = _save(px, path, 2)
        assert rc == 0, msg
        want = np.clip(px * np.float32(255.0), 0.0, 255.0).astype(np.uint8)
        quality = min(100, 3 * w)
        ref = io.BytesIO()
        Image.fromarray(want).save(ref, "JPEG", quality=quality, subsampling=0)
        with Image.open(path) as im, Image.open(ref) as lib:
            assert im.format == "JPEG" and im.mode == "RGB" and im.size == (w, h)
            assert im.quantization == lib.quantization, (w, h)
            got = np.asarray(im).astype(np.int32)
            libpx = np.asarray(lib).astype(np.int32)
        err = np.abs(got - want.astype(np.int32))
        if quality == 100:
            assert all(v == 1 for t in im.quantization.values() for v in t)
            assert err.max() <= 4 and err.mean() < 1.0, (w, h, err.max(), err.mean())
        # the same quantisation as libjpeg's encoder (integer DCT): the same distortion, up to
        # coefficients near a half step that the two DCTs round apart
        rms = lambda d: float(np.sqrt(np.mean(d.astype(np.float64) ** 2)))
        assert rms(got - want) <= 1.1 * rms(libpx - want) + 0.5, (w, h, rms(got - want),
                                                                   rms(libpx - want))


def test_save_to_unwritable_path_throws_reference_message(tmp_path):
    px = np.zeros((2, 2, 3), np.float32)
    for which in (0, 1, 2):
        rc, msg = _save(px, tmp_path / "missing_dir" / "x.img", which)
        assert rc == 1 and msg == "Failed to save image to file"
