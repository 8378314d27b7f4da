"""Radiance cache over the C-ABI (mt_cache_*): an MLP trained online on the tracer's own
training paths and queried by the other paths at a fixed bounce.

    cache = RadianceCache(gpu_scene, {"train_every": 16, "query_depth": 1})
    rgb, stats = cache.render(settings)           # mt_render_cached
    out = cache.infer(pos, dir)                   # test hooks over host arrays
    loss, grad = cache.train(pos, dir, target)
"""
import ctypes as C

import numpy as np

from . import abi
from ._lib import check, load
from .tracer import stats_dict

N_PARAMS = 14787
LAYERS = [(32, 64), (64, 64), (64, 64), (64, 64), (64, 3)]  # (in, out)


def unpack(weights):
    """flat parameters -> [(W[out][in], b[out])] per layer"""
    out, o = [], 0
    for i, n in LAYERS:
        W = weights[o:o + n * i].reshape(n, i)
        o += n * i
        b = weights[o:o + n]
        o += n
        out.append((W, b))
    return out


class RadianceCache:
    def __init__(self, gpu_scene, options=None):
        self.L = load()
        self.gscene = gpu_scene
        o = abi.mt_cache_options()
        for k, v in (options or {}).items():
            setattr(o, k, v)
        h = C.c_void_p()
        check(self.L.mt_cache_create(gpu_scene.h, C.byref(o), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.mt_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def render(self, settings, camera=None):
        cam = (camera or self.gscene.scene.camera).to_desc()
        st = settings.to_desc()
        rgb = np.zeros((settings.height, settings.width, 3), np.float32)
        stats = abi.mt_stats()
        check(self.L.mt_render_cached(self.gscene.h, C.byref(cam), C.byref(st), self.h,
                                      abi.fptr(rgb), C.byref(stats)))
        return rgb, stats_dict(stats)

    def weights(self):
        w = np.zeros(N_PARAMS, np.float32)
        check(self.L.mt_cache_get_weights(self.h, abi.fptr(w)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == N_PARAMS
        check(self.L.mt_cache_set_weights(self.h, abi.fptr(w)))

    def infer(self, pos, d):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        out = np.zeros_like(pos)
        check(self.L.mt_cache_infer(self.h, len(pos), abi.fptr(pos), abi.fptr(d), abi.fptr(out)))
        return out

    def train(self, pos, d, target, want_grad=False):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        target = np.ascontiguousarray(target, np.float32).reshape(-1, 3)
        loss = C.c_float()
        grad = np.zeros(N_PARAMS, np.float32) if want_grad else None
        check(self.L.mt_cache_train(self.h, len(pos), abi.fptr(pos), abi.fptr(d),
                                    abi.fptr(target), C.byref(loss),
                                    abi.fptr(grad) if want_grad else None))
        return loss.value, grad
