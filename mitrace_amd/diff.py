"""Differentiable rendering over the path-replay adjoint of the C-ABI (mt_diff_*).

The estimator is the Tracer's own (same counter RNG, same decisions) with linear radiance
output and no firefly rejection.  The parameters that can be differentiated are a float RGB
texture that replaces the base colour of a range of materials, and per-material roughness,
metallic and RGB emission.  Gradients are those of the expected image: the values the path
carries are differentiated, and so is the sampling -- the lobe choice's probability, the GGX
half-vector density and the mixture pdf the throughput divides by -- in score-function form
(the radiance collected after a bounce times the derivative of the log density of its
decisions).  With detach_sampling the sampling decisions and pdfs are held fixed instead (the
derivative of the frozen-sampling estimator of forward(sampling=...)).

``torch`` is imported before the library is loaded so that both share torch's HIP runtime.
"""
import ctypes as C

import torch

from . import abi
from ._lib import check, load

PARAM_NAMES = ("tex", "rough", "metal", "emis")


def _ptr(t):
    return t.data_ptr() if t is not None else None


class DiffRenderer:
    """mt_diff_create over a :class:`mitrace_amd.tracer.GpuScene` for one image size."""

    def __init__(self, gpu_scene, settings, tex_first=0, tex_count=0):
        self.L = load()
        self.gscene = gpu_scene
        self.settings = settings
        self.tex_first = tex_first
        self.tex_count = tex_count
        self.material_count = len(gpu_scene.scene.materials)
        h = C.c_void_p()
        st = settings.to_desc()
        check(self.L.mt_diff_create(gpu_scene.h, C.byref(st), C.byref(h)))
        self.h = h
        # the per-path store holds the data of the latest forward call only
        self.generation = 0

    def close(self):
        if self.h:
            self.L.mt_diff_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _params(self, tex, rough, metal, emis):
        p = abi.mt_diff_params()
        for t, n in ((tex, "tex"), (rough, "rough"), (metal, "metal"), (emis, "emis")):
            if t is None:
                continue
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
                raise ValueError(f"{n} must be a contiguous float32 tensor on the GPU")
        if tex is not None:
            if tex.dim() != 3 or tex.shape[2] != 3:
                raise ValueError("tex must be (height, width, 3)")
            p.base_texture_dev = _ptr(tex)
            p.tex_height, p.tex_width = tex.shape[0], tex.shape[1]
            p.tex_material_first = self.tex_first
            p.tex_material_count = self.tex_count
        for t, n, k in ((rough, "roughness_dev", 1), (metal, "metallic_dev", 1),
                        (emis, "emission_dev", 3)):
            if t is not None:
                if t.numel() != self.material_count * k:
                    raise ValueError(f"{n} needs {self.material_count * k} values")
                setattr(p, n, _ptr(t))
        p.material_count = self.material_count
        return p

    def forward(self, camera, seed, tex=None, rough=None, metal=None, emis=None, sampling=None):
        """linear image (H, W, 3) on the GPU; records the per-path data for backward().
        sampling: optional dict {tex, rough, metal, emis} that takes every sampling decision
        (frozen sampling) while the given parameters only set the values."""
        s = self.settings
        img = torch.empty((s.height, s.width, 3), dtype=torch.float32, device="cuda")
        p = self._params(tex, rough, metal, emis)
        ps = None
        if sampling is not None:
            ps = C.byref(self._params(*(sampling.get(n) for n in PARAM_NAMES)))
        cam = camera.to_desc()
        stream = torch.cuda.current_stream().cuda_stream
        check(self.L.mt_diff_forward(self.h, C.byref(cam), C.byref(p), ps, seed, img.data_ptr(),
                                     stream))
        self.generation += 1
        return img

    def backward(self, camera, seed, grad_image, tex=None, rough=None, metal=None, emis=None,
                 want=("tex", "rough", "metal", "emis"), detach_sampling=False):
        """gradients of sum(grad_image * image) for the parameters given (and in `want`)"""
        p = self._params(tex, rough, metal, emis)
        g = abi.mt_diff_grads()
        g.detach_sampling = int(bool(detach_sampling))
        out = {}
        for t, n, field in ((tex, "tex", "base_texture_dev"), (rough, "rough", "roughness_dev"),
                            (metal, "metal", "metallic_dev"), (emis, "emis", "emission_dev")):
            if t is not None and n in want:
                out[n] = torch.zeros_like(t)
                setattr(g, field, out[n].data_ptr())
        gi = grad_image.to(dtype=torch.float32).contiguous()
        if gi.shape != (self.settings.height, self.settings.width, 3):
            raise ValueError("grad_image must be (height, width, 3)")
        cam = camera.to_desc()
        stream = torch.cuda.current_stream().cuda_stream
        check(self.L.mt_diff_backward(self.h, C.byref(cam), C.byref(p), seed, gi.data_ptr(),
                                      C.byref(g), stream))
        return out

    def render(self, camera, seed, tex=None, rough=None, metal=None, emis=None):
        """autograd-aware render: image = f(tex, rough, metal, emis)"""
        return _RenderFn.apply(self, camera, seed, tex, rough, metal, emis)


class _RenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, renderer, camera, seed, tex, rough, metal, emis):
        tensors = [None if t is None else t.detach().contiguous() for t in (tex, rough, metal, emis)]
        ctx.renderer, ctx.camera, ctx.seed = renderer, camera, seed
        ctx.save_for_backward(*[t if t is not None else torch.empty(0) for t in tensors])
        ctx.present = [t is not None for t in tensors]
        img = renderer.forward(camera, seed, *tensors)
        ctx.generation = renderer.generation
        return img

    @staticmethod
    def backward(ctx, grad_image):
        saved = ctx.saved_tensors
        tensors = [saved[i] if ctx.present[i] else None for i in range(4)]
        want = tuple(n for n, need in zip(PARAM_NAMES, ctx.needs_input_grad[3:]) if need)
        if ctx.renderer.generation != ctx.generation:
            # another forward call has overwritten the per-path store: record this one again
            ctx.renderer.forward(ctx.camera, ctx.seed, *tensors)
        g = ctx.renderer.backward(ctx.camera, ctx.seed, grad_image, *tensors, want=want)
        return (None, None, None) + tuple(g.get(n) for n in PARAM_NAMES)
