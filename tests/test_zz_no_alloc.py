"""mt_trace_closest_noalloc: closest hits of device-resident rays with the traversal's scratch in
a static device array.  Its hits equal mt_trace_closest_ex's bit for bit, on two calls in a row
(the second with fewer rays, so stale counters and spill contents from the first are present),
for both traversals, on the caller's stream."""
import ctypes as C

import numpy as np
import pytest

from helpers import load_scene, seeded_rays
from mitrace_amd import abi

pytestmark = pytest.mark.gpu

HIT_FIELDS = ("instance", "triangle", "bv_index", "u", "v", "distance", "world_pos", "bvh_tests",
              "triangle_tests")


def _noalloc(g, rays, traversal):
    import torch
    from mitrace_amd._lib import check
    from oracle.pyoracle import HIT_DTYPE
    n = len(rays)
    d_rays = torch.from_numpy(np.ascontiguousarray(rays, np.float32)).cuda()
    d_hits = torch.full((max(n, 1) * HIT_DTYPE.itemsize,), 0xAB, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    check(g.L.mt_trace_closest_noalloc(g.h, n, C.c_void_p(d_rays.data_ptr()),
                                       C.c_void_p(d_hits.data_ptr()), traversal, stream))
    return d_hits.cpu().numpy()[: n * HIT_DTYPE.itemsize].view(HIT_DTYPE)


@pytest.mark.parametrize("name,traversal", [
    ("sphere_grid", abi.MT_TRAVERSAL_REFERENCE),
    ("large_small", abi.MT_TRAVERSAL_REFERENCE),
    ("large_small", abi.MT_TRAVERSAL_FAST),
])
def test_noalloc_equals_trace_closest(gpu, name, traversal):
    from mitrace_amd.tracer import GpuScene
    sc = load_scene(name)
    sc.bvh.fast = True
    g = GpuScene(sc)
    for n, seed in ((20000, 31), (3001, 32)):
        rays = seeded_rays(sc, n, seed=seed)
        want = g.trace(rays, traversal=traversal)
        got = _noalloc(g, rays, traversal)
        assert (want["instance"] >= 0).sum() > n // 10
        for f in HIT_FIELDS:
            assert np.array_equal(got[f].view(np.uint32), want[f].view(np.uint32)), (name, n, f)
    g.close()


def test_noalloc_arguments(gpu):
    from mitrace_amd.tracer import GpuScene
    g = GpuScene(load_scene("cornell_gen"))
    assert g.L.mt_trace_closest_noalloc(g.h, 0, None, None, abi.MT_TRAVERSAL_REFERENCE, 0) == 0
    assert g.L.mt_trace_closest_noalloc(g.h, 4, None, None, abi.MT_TRAVERSAL_REFERENCE, 0) == abi.MT_E_INVALID
    rays = seeded_rays(g.scene, 100, seed=1)
    import torch
    d = torch.from_numpy(rays).cuda()
    h = torch.zeros(100 * 48, dtype=torch.uint8, device="cuda")
    assert g.L.mt_trace_closest_noalloc(g.h, 100, C.c_void_p(d.data_ptr()), C.c_void_p(h.data_ptr()),
                                        7, 0) == abi.MT_E_INVALID
    # a scene built without fast BVHs rejects the fast traversal
    assert g.L.mt_trace_closest_noalloc(g.h, 100, C.c_void_p(d.data_ptr()), C.c_void_p(h.data_ptr()),
                                        abi.MT_TRAVERSAL_FAST, 0) == abi.MT_E_INVALID
    g.close()
