"""Python host side over the C-ABI, mirroring the reference's ``Tracer`` surface.

``Tracer(buffer, scene)`` with ``StartRender / WaitForRender / KillRender / GetStats /
IsDone`` follows src/Tracer/Tracer.hpp:121-128; ``render()`` is the synchronous form used by
main.cpp:121-142.  ``GpuScene`` is ``Scene::Scene`` uploaded to HBM (mt_scene_create).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import abi
from ._lib import check, load

HIT_DTYPE = np.dtype([("instance", np.int32), ("triangle", np.uint32), ("bv_index", np.uint32),
                      ("u", np.float32), ("v", np.float32), ("distance", np.float32),
                      ("world_pos", np.float32, 3), ("bvh_tests", np.uint32),
                      ("triangle_tests", np.uint32)])


class GpuScene:
    def __init__(self, scene):
        self.L = load()
        self.scene = scene
        desc, keep = scene.to_desc()
        h = C.c_void_p()
        check(self.L.mt_scene_create(C.byref(desc), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.mt_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def mesh_bvh(self, mesh):
        nn, ni = C.c_uint32(), C.c_uint32()
        check(self.L.mt_scene_mesh_bvh(self.h, mesh, None, C.byref(nn), None, C.byref(ni)))
        nodes = np.zeros((nn.value, 8), np.uint32)
        idx = np.zeros(ni.value, np.uint32)
        u32 = C.POINTER(C.c_uint32)
        check(self.L.mt_scene_mesh_bvh(self.h, mesh, nodes.ctypes.data_as(u32), C.byref(nn),
                                       idx.ctypes.data_as(u32), C.byref(ni)))
        return nodes, idx

    def trace(self, rays, traversal=None):
        """IntersectScene on the GPU for rays (N, 6) {o, d}; returns a HIT_DTYPE array.
        ``traversal`` (MT_TRAVERSAL_*) selects mt_trace_closest_ex; None = mt_trace_closest."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        hits = np.zeros(len(rays), HIT_DTYPE)
        hp = hits.ctypes.data_as(C.POINTER(abi.mt_hit))
        if traversal is None:
            check(self.L.mt_trace_closest(self.h, len(rays), abi.fptr(rays), hp))
        else:
            check(self.L.mt_trace_closest_ex(self.h, len(rays), abi.fptr(rays), hp, traversal))
        return hits

    def occluded(self, rays, traversal=abi.MT_TRAVERSAL_REFERENCE):
        """Shadow-ray query for rays (N, 7) {o, d, distance}: 1 where the closest hit is at
        most that far (Tracer.cpp:317-323)."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 7)
        occ = np.zeros(len(rays), np.uint32)
        check(self.L.mt_trace_occluded(self.h, len(rays), abi.fptr(rays),
                                       occ.ctypes.data_as(C.POINTER(C.c_uint32)), traversal))
        return occ

    def render(self, settings, camera=None, want_samples=False):
        """Synchronous render: returns (rgb (H, W, 3), samples or None, stats dict)."""
        cam = (camera or self.scene.camera).to_desc()
        st = settings.to_desc()
        W, H, S = settings.width, settings.height, settings.samples
        rgb = np.zeros((H, W, 3), np.float32)
        samples = np.zeros((H * W, S, 3), np.float32) if want_samples else None
        stats = abi.mt_stats()
        check(self.L.mt_render(self.h, C.byref(cam), C.byref(st), abi.fptr(rgb),
                               abi.fptr(samples) if want_samples else None, C.byref(stats)))
        return rgb, samples, stats_dict(stats)


def stats_dict(s):
    d = {n: getattr(s, n) for n, _ in abi.mt_stats._fields_}
    d["diag"] = list(s.diag)
    return d


def brdf_eval(inp):
    L = load()
    inp = np.ascontiguousarray(inp, np.float32).reshape(-1, 14)
    out = np.zeros((len(inp), 5), np.float32)
    check(L.mt_brdf_eval(len(inp), abi.fptr(inp), abi.fptr(out)))
    return out


class RenderBuffer:
    """W x H float RGB, row-major (src/Tracer/RenderBuffer.hpp:11-35)."""

    def __init__(self, width, height):
        self.width, self.height = width, height
        self.pixels = np.zeros((height, width, 3), np.float32)

    def get_pixels_rgb8(self):
        # trunc(clamp(v * 255)) with no sRGB (RenderBuffer.cpp:37-58)
        return np.clip(self.pixels * 255.0, 0.0, 255.0).astype(np.uint8)


@dataclass
class Stats:
    progress: float
    timeElapsed: float
    estimatedTimeRemaining: float
    rays: int
    samples: int


class Tracer:
    """The reference Tracer's method set over mt_render_start/wait/kill/stats."""

    def __init__(self, buffer, scene, settings):
        self.L = load()
        self.buffer = buffer
        self.scene = scene if isinstance(scene, GpuScene) else GpuScene(scene)
        self.settings = settings
        if (settings.width, settings.height) != (buffer.width, buffer.height):
            raise ValueError("settings and buffer sizes differ")
        self.job = None

    def StartRender(self):
        cam = self.scene.scene.camera.to_desc()
        st = self.settings.to_desc()
        h = C.c_void_p()
        check(self.L.mt_render_start(self.scene.h, C.byref(cam), C.byref(st),
                                     abi.fptr(self.buffer.pixels), C.byref(h)))
        self.job = h

    def WaitForRender(self):
        rc = self.L.mt_render_wait(self.job)
        if rc not in (0, abi.MT_E_KILLED):
            check(rc)
        return rc

    def KillRender(self):
        if self.job:
            self.L.mt_render_kill(self.job)

    def GetStats(self):
        s = abi.mt_stats()
        check(self.L.mt_render_stats(self.job, C.byref(s)))
        return Stats(s.progress, s.time_elapsed, s.estimated_time_remaining, s.rays, s.samples)

    def IsDone(self):
        return bool(self.job) and bool(self.L.mt_render_is_done(self.job))

    def close(self):
        if self.job:
            self.L.mt_job_destroy(self.job)
            self.job = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
