"""Path-replay adjoint on the GPU (mt_diff_*) against the oracle's forward-mode tangents and
against central finite differences with common random numbers.

Tolerances: the adjoint and the oracle differentiate the same estimator on the same counter-RNG
paths (every sample is bit-identical), so their directional derivatives agree to float32
accumulation error: 1e-3 on a sum over every pixel, for the full derivative (with the sampling
decisions' dependence on the parameters) and for the detached one.  Central differences of the
estimator with its sampling frozen (decisions from the unperturbed parameters) differentiate
the same function as the detached adjoint: 5 % per parameter at h = 1e-3.  Central differences
with common random numbers see the changed sampling decisions too, which the full adjoint
includes: 5 % on the mean over seeds.
"""
import numpy as np
import pytest

from helpers import load_scene
from mitrace_amd import abi, scenegen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GROUPS = ("tex", "rough", "metal", "emis")


@pytest.fixture(scope="module")
def torch_mod(gpu):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("torch sees no GPU")
    return torch


def _scene(segments=16, rings=8, tex_res=64):
    return scenegen.inverse_scene(tex_res=tex_res, env_width=256, env_height=128,
                                  segments=segments, rings=rings)


def _dev(torch, params):
    out = {}
    for k in GROUPS:
        out[k] = torch.tensor(np.ascontiguousarray(params[k], np.float32), device="cuda")
    return out


def _renderer(sc, st, params):
    from mitrace_amd.diff import DiffRenderer
    from mitrace_amd.tracer import GpuScene
    gs = GpuScene(sc)
    return gs, DiffRenderer(gs, st, params["tex_first"], params["tex_count"])


def test_diff_forward_equals_plain_linear_render(torch_mod):
    torch = torch_mod
    sc, p = _scene()
    st = scenegen.inverse_settings(width=40, height=32, samples=8, seed=3)
    gs, dr = _renderer(sc, st, p)
    img = dr.forward(sc.camera, 3).cpu().numpy()
    ref, _, _ = gs.render(st)
    # the scene's own values (the texture is not a scene texture, so only without overrides)
    np.testing.assert_array_equal(img, ref)
    # with the parameters as overrides the image changes only through the texture
    d = _dev(torch, p)
    img2 = dr.forward(sc.camera, 3, **d).cpu().numpy()
    assert np.isfinite(img2).all() and not np.array_equal(img2, img)


def test_diff_forward_matches_oracle_estimator(torch_mod):
    torch = torch_mod
    sc, p = _scene()
    st = scenegen.inverse_settings(width=32, height=32, samples=16, seed=5)
    gs, dr = _renderer(sc, st, p)
    img = dr.forward(sc.camera, 5, **_dev(torch, p)).cpu().numpy()
    ref, _ = po.OracleScene(sc).render_diff(st, p)
    # the same paths sample for sample: every pixel equal up to the order of its sum
    close = np.isclose(img, ref, rtol=1e-5, atol=1e-6).all(axis=-1)
    assert close.all(), close.mean()


@pytest.mark.parametrize("detach", [False, True])
@pytest.mark.parametrize("group", GROUPS)
def test_adjoint_matches_oracle_tangent(torch_mod, group, detach):
    """adjoint . v = the oracle's forward-mode tangent along v, for the full derivative (the
    sampling decisions' dependence included) and for the detached one"""
    torch = torch_mod
    sc, p = _scene()
    p = dict(p)
    p["emis"] = np.full_like(p["emis"], 0.25)  # so that emission has an effect
    st = scenegen.inverse_settings(width=32, height=32, samples=16, seed=7)
    gs, dr = _renderer(sc, st, p)
    rng = np.random.default_rng(17)
    gimg = rng.normal(size=(32, 32, 3)).astype(np.float32)
    v = rng.normal(size=np.shape(p[group])).astype(np.float32)
    d = _dev(torch, p)
    dr.forward(sc.camera, 7, **d)
    g = dr.backward(sc.camera, 7, torch.tensor(gimg, device="cuda"), detach_sampling=detach, **d)
    adj = float((g[group].cpu().numpy().astype(np.float64) * v).sum())
    _, dimg = po.OracleScene(sc).render_diff(st, p, {group: v}, sampling_terms=not detach)
    ref = float((dimg.astype(np.float64) * gimg).sum())
    print(f"{group}: adjoint {adj:.6g} oracle tangent {ref:.6g}")
    assert abs(ref) > 1e-3
    assert abs(adj - ref) <= 1e-3 * abs(ref), (adj, ref)


def test_autograd_function_and_store_regeneration(torch_mod):
    torch = torch_mod
    sc, p = _scene()
    st = scenegen.inverse_settings(width=24, height=24, samples=8, seed=2)
    gs, dr = _renderer(sc, st, p)
    d = {k: t.clone().requires_grad_(True) for k, t in _dev(torch, p).items()}
    cams = scenegen.inverse_poses(2)
    target = torch.rand(24, 24, 3, device="cuda")
    # two renders before one backward: the first one's per-path store must be recorded again
    img_a = dr.render(cams[0], 11, **d)
    img_b = dr.render(cams[1], 12, **d)
    loss = ((img_a - target) ** 2).sum() + ((img_b - target) ** 2).sum()
    loss.backward()
    auto = {k: d[k].grad.clone() for k in GROUPS}
    # the same gradients from direct forward/backward pairs
    ref = {k: torch.zeros_like(d[k]) for k in GROUPS}
    plain = {k: t.detach() for k, t in d.items()}
    for cam, seed in ((cams[0], 11), (cams[1], 12)):
        img = dr.forward(cam, seed, **plain)
        g = dr.backward(cam, seed, 2.0 * (img - target), **plain)
        for k in GROUPS:
            ref[k] += g[k]
    for k in GROUPS:
        a, b = auto[k].cpu().numpy(), ref[k].cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-6 * (np.abs(b).max() + 1e-12))
        assert np.abs(b).max() > 0 or k == "emis"


def test_emission_gradient_is_exact(torch_mod):
    """the image is linear in emission and no decision depends on it: the adjoint equals the
    central difference of two renders with the same seed to float32 rounding"""
    torch = torch_mod
    sc, p = _scene()
    p = dict(p)
    p["emis"] = np.full_like(p["emis"], 0.5)
    st = scenegen.inverse_settings(width=32, height=32, samples=16, seed=4)
    gs, dr = _renderer(sc, st, p)
    rng = np.random.default_rng(3)
    w = torch.tensor(rng.normal(size=(32, 32, 3)).astype(np.float32), device="cuda")
    d = _dev(torch, p)
    dr.forward(sc.camera, 4, **d)
    g = dr.backward(sc.camera, 4, w, **d)["emis"].cpu().numpy()
    # the image is exactly linear in emission, so a large step loses nothing; the loss is
    # summed in float64 so that its rounding stays far below the difference
    h = 0.25
    w64 = w.double()
    for mi in (0, 24, 48, 49):
        for c in range(3):
            dp, dm = dict(d), dict(d)
            e = d["emis"].clone()
            e[mi * 3 + c] += h
            dp["emis"] = e
            e2 = d["emis"].clone()
            e2[mi * 3 + c] -= h
            dm["emis"] = e2
            lp = float((dr.forward(sc.camera, 4, **dp).double() * w64).sum())
            lm = float((dr.forward(sc.camera, 4, **dm).double() * w64).sum())
            fd = (lp - lm) / (2 * h)
            assert abs(g[mi * 3 + c] - fd) <= 2e-3 * max(abs(fd), 1e-2), (mi, c, g[mi * 3 + c], fd)


def test_adjoint_vs_frozen_sampling_central_differences(torch_mod):
    """per parameter: 24 random texels, every sphere's roughness and metallic, 4 emission
    entries; central differences (h = 1e-3) of the estimator with its sampling frozen at the
    unperturbed parameters -- the estimator the adjoint differentiates"""
    torch = torch_mod
    sc, p = _scene(segments=32, rings=16, tex_res=32)
    p = dict(p)
    p["emis"] = np.full_like(p["emis"], 0.25)
    st = scenegen.inverse_settings(width=64, height=64, samples=64, seed=31)
    gs, dr = _renderer(sc, st, p)
    rng = np.random.default_rng(29)
    w = torch.tensor(rng.uniform(0.0, 1.0, size=(64, 64, 3)).astype(np.float32), device="cuda")
    d = _dev(torch, p)
    dr.forward(sc.camera, 31, **d)
    g = {k: v.cpu().numpy() for k, v in dr.backward(sc.camera, 31, w, detach_sampling=True, **d).items()}
    h = 1e-3
    cases = [("tex", int(i)) for i in rng.choice(p["tex"].size, 24, replace=False)]
    cases += [("rough", i) for i in range(49) if p["rough"][i] > 0.05 + 6 * h]
    cases += [("metal", i) for i in range(49)]
    cases += [("emis", i) for i in (0, 7, 100, 149)]

    def central(group, i, step):
        flat = d[group].reshape(-1)
        dp_, dm_ = dict(d), dict(d)
        e = flat.clone()
        e[i] += step
        dp_[group] = e.reshape(d[group].shape)
        e = flat.clone()
        e[i] -= step
        dm_[group] = e.reshape(d[group].shape)
        ip = dr.forward(sc.camera, 31, sampling=d, **dp_).double()
        im = dr.forward(sc.camera, 31, sampling=d, **dm_).double()
        return float(((ip - im) * w).sum()) / (2 * step)

    within, worst = 0, 0.0
    for group, i in cases:
        fd = central(group, i, h)
        # the float32 image rounds each sample: the difference of two step sizes bounds that
        # noise for samples whose values are large next to the derivative
        noise = abs(fd - central(group, i, 3 * h))
        adj = float(g[group].reshape(-1)[i])
        floor = 1e-3 * np.abs(g[group]).max()
        err = abs(adj - fd)
        within += err <= 0.05 * abs(fd) + floor
        worst = max(worst, err / max(abs(fd), floor))
        assert err <= 0.05 * abs(fd) + 2.0 * noise + floor, (group, i, adj, fd, noise)
    print(f"{len(cases)} parameters, {within} within 5 %, worst relative difference {worst:.2e}")
    assert within >= 0.95 * len(cases)


def test_adjoint_vs_crn_central_differences_texture(torch_mod):
    """sampling not frozen: central differences with common random numbers also see the
    changed sampling decisions, which are zero-mean; mean over 4 seeds, 64x64 at 2048 spp"""
    torch = torch_mod
    sc, p = _scene(segments=64, rings=32, tex_res=16)
    st = scenegen.inverse_settings(width=64, height=64, samples=2048, seed=21,
                                   paths_in_flight=1 << 23)
    gs, dr = _renderer(sc, st, p)
    rng = np.random.default_rng(23)
    w = torch.tensor(rng.uniform(0.0, 1.0, size=(64, 64, 3)).astype(np.float32), device="cuda")
    d = _dev(torch, p)
    vt = torch.tensor(rng.normal(size=p["tex"].shape).astype(np.float32), device="cuda")
    h = 1e-2
    adj, fd = [], []
    for seed in (21, 22, 23, 24):
        dr.forward(sc.camera, seed, **d)
        adj.append(float((dr.backward(sc.camera, seed, w, **d)["tex"] * vt).sum()))
        dp_, dm_ = dict(d), dict(d)
        dp_["tex"] = (d["tex"] + h * vt).contiguous()
        dm_["tex"] = (d["tex"] - h * vt).contiguous()
        lp = float((dr.forward(sc.camera, seed, **dp_).double() * w).sum())
        lm = float((dr.forward(sc.camera, seed, **dm_).double() * w).sum())
        fd.append((lp - lm) / (2 * h))
    a, f = np.mean(adj), np.mean(fd)
    print(f"adjoint {adj} central differences {fd}")
    assert abs(a - f) <= 0.05 * abs(f), (a, f)


@pytest.mark.parametrize("group", ["rough", "metal"])
def test_full_adjoint_vs_crn_central_differences(torch_mod, group):
    """The full derivative against central differences of the unfrozen estimator with common
    random numbers (h = 1e-2, 4096 spp, 128x128, the mean over 8 seeds): the sampled directions
    and lobe choices move with the parameters there, which the adjoint takes in through the
    score of the sampling decisions and the derivative of the mixture pdf.  Roughness along a
    random direction over the spheres of roughness >= 0.3 (the smoother spheres' highlights make
    both sides too noisy at this size); metallic over every sphere.  The detached adjoint misses
    the sampling's share (measured: 7.5 % off for roughness, 3.7 % for metallic)."""
    torch = torch_mod
    sc, p = scenegen.inverse_scene(env_width=512, env_height=256)
    st = scenegen.inverse_settings(width=128, height=128, samples=4096, seed=1, paths_in_flight=1 << 24)
    gs, dr = _renderer(sc, st, p)
    cam = scenegen.inverse_poses(1)[0]
    d = _dev(torch, p)
    rng = np.random.default_rng(3)
    w = torch.tensor(rng.uniform(0.0, 1.0, size=(128, 128, 3)).astype(np.float32), device="cuda")
    vv = rng.normal(size=np.shape(p[group])).astype(np.float32)
    if group == "rough":
        vv[p["rough"] < 0.3] = 0.0
    v = torch.tensor(vv, device="cuda")
    h = 1e-2
    full, det, fd = [], [], []
    for seed in range(21, 29):
        dr.forward(cam, seed, **d)
        full.append(float((dr.backward(cam, seed, w, **d)[group].double() * v).sum()))
        det.append(float((dr.backward(cam, seed, w, detach_sampling=True, **d)[group].double() * v).sum()))
        dp_, dm_ = dict(d), dict(d)
        dp_[group] = (d[group] + h * v).contiguous()
        dm_[group] = (d[group] - h * v).contiguous()
        lp = float((dr.forward(cam, seed, **dp_).double() * w).sum())
        lm = float((dr.forward(cam, seed, **dm_).double() * w).sum())
        fd.append((lp - lm) / (2 * h))
    a, b, f = np.mean(full), np.mean(det), np.mean(fd)
    print(f"{group}: full adjoint {a:.2f}, detached {b:.2f}, central differences {f:.2f}")
    assert abs(a - f) <= 0.05 * abs(f), (a, f)
    assert abs(a - f) < abs(b - f), (a, b, f)


def test_diff_rejects_bad_arguments(gpu):
    from mitrace_amd.tracer import GpuScene
    import ctypes as C
    sc = load_scene("cornell_box")
    gs = GpuScene(sc)
    st = scenegen.inverse_settings(width=8, height=8, samples=1, bounces=40).to_desc()
    h = C.c_void_p()
    assert gpu.mt_diff_create(gs.h, C.byref(st), C.byref(h)) == abi.MT_E_UNSUPPORTED
    assert not h.value
