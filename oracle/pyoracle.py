"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Imported by tests/ as the checker and by bench.py for the CPU-baseline leg; the product
package never imports it.
"""
import ctypes as C
import os

import numpy as np

from mitrace_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
ORC_RNG_COUNTER, ORC_RNG_MT19937 = 0, 1


class orc_render_opts(C.Structure):
    _fields_ = [("threads", C.c_int32), ("rng", C.c_int32), ("faithful_inverse", C.c_int32),
                ("row_begin", C.c_uint32), ("row_end", C.c_uint32), ("tlas", C.c_int32)]


class orc_work(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "samples", "inst_aabb_tests", "inst_entered",
                                         "node_pops", "tri_tests", "hits_shaded", "misses")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class orc_diff_params(C.Structure):
    _fields_ = [("tex", C.c_void_p), ("tex_w", C.c_uint32), ("tex_h", C.c_uint32),
                ("tex_first", C.c_uint32), ("tex_count", C.c_uint32), ("rough", C.c_void_p),
                ("metal", C.c_void_p), ("emis", C.c_void_p)]


def _diff_struct(d, keep):
    """dict {tex (H,W,3), tex_first, tex_count, rough, metal, emis} -> orc_diff_params"""
    out = orc_diff_params()
    if d is None:
        return out
    for key in ("tex", "rough", "metal", "emis"):
        a = d.get(key)
        if a is not None:
            a = np.ascontiguousarray(a, np.float32)
            keep.append(a)
            setattr(out, key, a.ctypes.data)
            if key == "tex":
                out.tex_h, out.tex_w = a.shape[0], a.shape[1]
    out.tex_first = d.get("tex_first", 0)
    out.tex_count = d.get("tex_count", 0)
    return out


def lib_path(fast=False, mutant=0):
    """fast: the release-flag build; mutant k > 0: the strict build with planted mistake k
    (oracle/Makefile target `mutants`, for the tests that show each pin rejects a mistake)"""
    if mutant:
        return os.path.join(_HERE, "build", f"liboracle_m{mutant}.so")
    return os.path.join(_HERE, "build", "liboracle_fast.so" if fast else "liboracle.so")


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_libs = {}


def load(fast=False, mutant=0):
    key = (bool(fast), int(mutant))
    if key in _libs:
        return _libs[key]
    p = lib_path(fast, mutant)
    if not os.path.exists(p):
        raise FileNotFoundError(f"oracle library {p} is not built (make -C oracle)")
    L = C.CDLL(p)
    L.orc_scene_create.restype = C.c_void_p
    L.orc_scene_create.argtypes = [C.POINTER(abi.mt_scene_desc)]
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_last_error.restype = C.c_char_p
    L.orc_scene_mesh_bvh.argtypes = [C.c_void_p, C.c_uint32] + [C.POINTER(C.c_uint32)] * 4
    L.orc_scene_world_aabbs.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    L.orc_trace_closest.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                    C.POINTER(abi.mt_hit), C.POINTER(orc_work)]
    L.orc_trace_bruteforce.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float),
                                       C.POINTER(abi.mt_hit)]
    L.orc_slab.restype = C.c_float
    L.orc_slab.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_triangle_test.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_invert4.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_brdf_eval.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_sample_bounce.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_env_sample.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_int32,
                                 C.c_float, C.POINTER(C.c_float)]
    L.orc_camera_rays.argtypes = [C.POINTER(abi.mt_camera_desc), C.c_uint32, C.POINTER(C.c_float),
                                  C.c_float, C.POINTER(C.c_float)]
    L.orc_aces.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.c_float, C.POINTER(C.c_float)]
    L.orc_cleanup_fireflies.restype = C.c_uint32
    L.orc_cleanup_fireflies.argtypes = [C.POINTER(C.c_float), C.c_uint32, C.c_float]
    u32p = C.POINTER(C.c_uint32)
    L.orc_rng_draws.argtypes = [C.c_uint64, C.c_uint32, u32p, u32p, u32p, C.POINTER(C.c_float)]
    L.orc_render.argtypes = [C.c_void_p, C.POINTER(abi.mt_camera_desc), C.POINTER(abi.mt_settings),
                             C.POINTER(orc_render_opts), C.POINTER(C.c_float), C.POINTER(C.c_float),
                             C.POINTER(orc_work), C.POINTER(C.c_double)]
    L.orc_brdf_tangent.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float)]
    L.orc_render_diff.argtypes = [C.c_void_p, C.POINTER(abi.mt_camera_desc),
                                  C.POINTER(abi.mt_settings), C.POINTER(orc_diff_params),
                                  C.POINTER(orc_diff_params), C.c_int32, C.c_int32,
                                  C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_bounce_terms.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_math.argtypes = [C.c_int32, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.orc_sample_image.argtypes = [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                   C.POINTER(C.c_float), C.c_int32, C.POINTER(C.c_float)]
    L.orc_compute_normal.argtypes = [C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]
    L.orc_light_sample.argtypes = [C.c_void_p, C.c_uint32, C.POINTER(C.c_float), C.c_uint64,
                                   C.c_uint32, C.POINTER(C.c_int32), C.POINTER(C.c_float),
                                   C.POINTER(C.c_float), C.POINTER(C.c_float)]
    _libs[key] = L
    return L


HIT_DTYPE = np.dtype([("instance", np.int32), ("triangle", np.uint32), ("bv_index", np.uint32),
                      ("u", np.float32), ("v", np.float32), ("distance", np.float32),
                      ("world_pos", np.float32, 3), ("bvh_tests", np.uint32),
                      ("triangle_tests", np.uint32)])
assert HIT_DTYPE.itemsize == C.sizeof(abi.mt_hit)


class OracleScene:
    """The oracle's Scene::Scene over a :class:`mitrace_amd.scene.Scene`."""

    def __init__(self, scene, fast=False, mutant=0):
        self.L = load(fast, mutant)
        self.scene = scene
        desc, keep = scene.to_desc()
        self.h = self.L.orc_scene_create(C.byref(desc))
        if not self.h:
            raise RuntimeError(self.L.orc_last_error().decode())

    def close(self):
        if self.h:
            self.L.orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def mesh_bvh(self, mesh):
        nn, ni = C.c_uint32(), C.c_uint32()
        self.L.orc_scene_mesh_bvh(self.h, mesh, None, C.byref(nn), None, C.byref(ni))
        nodes = np.zeros((nn.value, 8), np.uint32)
        idx = np.zeros(ni.value, np.uint32)
        self.L.orc_scene_mesh_bvh(self.h, mesh, nodes.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  C.byref(nn), idx.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(ni))
        return nodes, idx

    def world_aabbs(self):
        out = np.zeros((len(self.scene.instances), 6), np.float32)
        self.L.orc_scene_world_aabbs(self.h, abi.fptr(out))
        return out

    def trace(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        hits = np.zeros(len(rays), HIT_DTYPE)
        hp = hits.ctypes.data_as(C.POINTER(abi.mt_hit))
        w = orc_work()
        if brute:
            self.L.orc_trace_bruteforce(self.h, len(rays), abi.fptr(rays), hp)
        else:
            self.L.orc_trace_closest(self.h, len(rays), abi.fptr(rays), hp, C.byref(w))
        return hits, w.as_dict()

    def env_sample(self, dirs, apply_rotation=False, rotation_deg=0.0):
        dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros_like(dirs)
        self.L.orc_env_sample(self.h, len(dirs), abi.fptr(dirs), int(apply_rotation),
                              rotation_deg, abi.fptr(out))
        return out

    def light_sample(self, pos, seed=0, base_dim=2):
        """One light choice and light point per hit position (pixel i, sample 0 of `seed`):
        (chosen light or -1, selection probability, light point, light colour)"""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        n = len(pos)
        chosen = np.zeros(n, np.int32)
        pdf = np.zeros(n, np.float32)
        lpos = np.zeros((n, 3), np.float32)
        lcol = np.zeros((n, 3), np.float32)
        self.L.orc_light_sample(self.h, n, abi.fptr(pos), seed, base_dim,
                                chosen.ctypes.data_as(C.POINTER(C.c_int32)), abi.fptr(pdf),
                                abi.fptr(lpos), abi.fptr(lcol))
        return chosen, pdf, lpos, lcol

    def render(self, settings, camera=None, threads=0, rng=ORC_RNG_COUNTER, faithful_inverse=True,
               rows=None, want_samples=False, tlas=False):
        cam = (camera or self.scene.camera).to_desc()
        st = settings.to_desc()
        W, H, S = settings.width, settings.height, settings.samples
        rgb = np.zeros((H, W, 3), np.float32)
        samples = np.zeros((H * W, S, 3), np.float32) if want_samples else None
        o = orc_render_opts(threads=threads, rng=rng, faithful_inverse=int(faithful_inverse),
                            row_begin=rows[0] if rows else 0, row_end=rows[1] if rows else 0,
                            tlas=int(tlas))
        w = orc_work()
        secs = C.c_double()
        rc = self.L.orc_render(self.h, C.byref(cam), C.byref(st), C.byref(o), abi.fptr(rgb),
                               abi.fptr(samples) if want_samples else None, C.byref(w), C.byref(secs))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return rgb, samples, w.as_dict(), secs.value


    def render_diff(self, settings, params, tangent=None, camera=None, threads=0,
                    sampling_terms=False):
        """The differentiable estimator: (mean linear radiance, directional derivative along
        `tangent`) per pixel.  params/tangent: dicts {tex, tex_first, tex_count, rough, metal,
        emis} of numpy arrays (missing = the scene's value / zero tangent).  sampling_terms adds
        the sampling decisions' dependence on the parameters (full derivative)."""
        cam = (camera or self.scene.camera).to_desc()
        st = settings.to_desc()
        W, H = settings.width, settings.height
        img = np.zeros((H, W, 3), np.float32)
        dimg = np.zeros((H, W, 3), np.float32)
        keep = []
        p = _diff_struct(params, keep)
        t = _diff_struct(tangent, keep)
        rc = self.L.orc_render_diff(self.h, C.byref(cam), C.byref(st), C.byref(p), C.byref(t),
                                    int(sampling_terms), threads, abi.fptr(img), abi.fptr(dimg))
        if rc != 0:
            raise RuntimeError(self.L.orc_last_error().decode())
        return img, dimg


def brdf_eval(inp):
    L = load()
    inp = np.ascontiguousarray(inp, np.float32).reshape(-1, 14)
    out = np.zeros((len(inp), 5), np.float32)
    L.orc_brdf_eval(len(inp), abi.fptr(inp), abi.fptr(out))
    return out


def brdf_tangent(inp, tin):
    """(f, df) of EvaluateBRDF for inputs (n, 14) along tangents (n, 5) {dbase, drough, dmetal}"""
    L = load()
    inp = np.ascontiguousarray(inp, np.float32).reshape(-1, 14)
    tin = np.ascontiguousarray(tin, np.float32).reshape(-1, 5)
    out = np.zeros((len(inp), 6), np.float32)
    L.orc_brdf_tangent(len(inp), abi.fptr(inp), abi.fptr(tin), abi.fptr(out))
    return out[:, :3], out[:, 3:]


def sample_bounce(inp):
    L = load()
    inp = np.ascontiguousarray(inp, np.float32).reshape(-1, 16)
    out = np.zeros((len(inp), 8), np.float32)
    L.orc_sample_bounce(len(inp), abi.fptr(inp), abi.fptr(out))
    return out


def camera_rays(camera, uv, aspect):
    L = load()
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 6), np.float32)
    cd = camera.to_desc()
    L.orc_camera_rays(C.byref(cd), len(uv), abi.fptr(uv), aspect, abi.fptr(out))
    return out


def aces(rgb, ev):
    L = load()
    rgb = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    out = np.zeros_like(rgb)
    L.orc_aces(len(rgb), abi.fptr(rgb), ev, abi.fptr(out))
    return out


def cleanup_fireflies(samples, threshold):
    L = load()
    s = np.ascontiguousarray(samples, np.float32).reshape(-1, 3).copy()
    n = L.orc_cleanup_fireflies(abi.fptr(s), len(s), threshold)
    return s, n


def rng_draws(seed, pixel, sample, dim):
    L = load()
    pixel = np.ascontiguousarray(pixel, np.uint32)
    sample = np.ascontiguousarray(sample, np.uint32)
    dim = np.ascontiguousarray(dim, np.uint32)
    out = np.zeros(len(pixel), np.float32)
    p = lambda a: a.ctypes.data_as(C.POINTER(C.c_uint32))
    L.orc_rng_draws(seed, len(pixel), p(pixel), p(sample), p(dim), abi.fptr(out))
    return out


def slab(ray, box):
    ray = np.ascontiguousarray(ray, np.float32)
    box = np.ascontiguousarray(box, np.float32)
    return float(load().orc_slab(abi.fptr(ray), abi.fptr(box)))


def triangle_test(ray, tri):
    ray = np.ascontiguousarray(ray, np.float32)
    tri = np.ascontiguousarray(tri, np.float32).reshape(9)
    uvt = np.zeros(3, np.float32)
    hit = load().orc_triangle_test(abi.fptr(ray), abi.fptr(tri), abi.fptr(uvt))
    return bool(hit), uvt


def invert4(m_colmajor):
    m = np.ascontiguousarray(m_colmajor, np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    load().orc_invert4(abi.fptr(m), abi.fptr(out))
    return out


def sample_image(rgb8, uv, filter_mode, mutant=0):
    """Image::Sample of an (H, W, 3) uint8 image at (n, 2) uv pairs -> (n, 4) RGBA floats"""
    img = np.ascontiguousarray(rgb8, np.uint8)
    h, w = img.shape[:2]
    uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
    out = np.zeros((len(uv), 4), np.float32)
    rc = load(mutant=mutant).orc_sample_image(w, h, img.ctypes.data, len(uv), abi.fptr(uv), filter_mode,
                                 abi.fptr(out))
    assert rc == 0
    return out


def compute_normal(surf, tex, mutant=0):
    surf = np.ascontiguousarray(surf, np.float32).reshape(-1, 3)
    tex = np.ascontiguousarray(np.broadcast_to(tex, surf.shape), np.float32)
    out = np.zeros_like(surf)
    load(mutant=mutant).orc_compute_normal(len(surf), abi.fptr(surf), abi.fptr(tex), abi.fptr(out))
    return out


def math_eval(op, x, y=None):
    """the oracle's elementary function `op` (abi.MT_MATH_*) at x (and y for ATAN2: atan2(x, y))"""
    x = np.asarray(x, np.float32).ravel()
    y = np.zeros_like(x) if y is None else np.asarray(y, np.float32).ravel()
    inp = np.ascontiguousarray(np.stack([x, y], 1), np.float32)
    out = np.zeros(len(x), np.float32)
    load().orc_math(op, len(x), abi.fptr(inp), abi.fptr(out))
    return out


def bounce_terms(inp):
    """orc_bounce_terms over (n, 22) inputs -> (n, 13)"""
    inp = np.ascontiguousarray(inp, np.float32).reshape(-1, 22)
    out = np.zeros((len(inp), 13), np.float32)
    load().orc_bounce_terms(len(inp), abi.fptr(inp), abi.fptr(out))
    return out
