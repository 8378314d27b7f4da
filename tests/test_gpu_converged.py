"""Converged images: the plain tracer's renders at the configurations' sample counts against
4096-spp references of the same estimator.  Each pixel channel lies within 3 sigma of the
reference (sigma from the render's own samples), and the image RMSE equals the Monte Carlo
error the samples predict (no bias); the Cornell render is the oracle's sample for sample, so
the CPU tracer has the same RMSE at the same sample count."""
import numpy as np
import pytest

from mitrace_amd import abi, scenegen
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _reference(g, settings_fn, spp, frames, seed0):
    acc = None
    for i in range(frames):
        f, _, _ = g.render(settings_fn(samples=spp, seed=seed0 + i))
        acc = f.astype(np.float64) if acc is None else acc + f
    return acc / frames


def _check(img, samples, ref, spp, ref_spp):
    sigma = samples.std(axis=1).reshape(img.shape) / np.sqrt(spp)
    sig = np.sqrt(sigma ** 2 * (1.0 + spp / ref_spp))  # the reference's own error too
    d = img - ref
    within = (np.abs(d) <= 3.0 * sig + 1e-7).mean()
    rmse = float(np.sqrt(np.mean(d ** 2)))
    predicted = float(np.sqrt(np.mean(sig ** 2)))
    print(f"rmse {rmse:.4g} predicted {predicted:.4g} within 3 sigma {within:.4f}")
    assert within > 0.98
    assert 0.8 * predicted <= rmse <= 1.25 * predicted
    return rmse


def test_cornell_configuration_converges(gpu):
    """the Cornell configuration as stated: 256x256, 64 spp, Russian roulette after depth 3,
    at most 8 bounces"""
    from mitrace_amd.tracer import GpuScene
    sc = scenegen.cornell()
    g = GpuScene(sc)
    fn = lambda **kw: scenegen.cornell_settings(mode=abi.MT_MODE_FIXED, rr_depth=3, bounces=8,
                                                ev_exposure=2.0, **kw)
    st = fn(samples=64, seed=11)
    img, samples, _ = g.render(st, want_samples=True)
    orgb, osamples, _, _ = po.OracleScene(sc).render(st, want_samples=True)
    assert np.array_equal(samples, osamples) and np.array_equal(img, orgb)
    ref = _reference(g, fn, 64, 64, 100)
    _check(img, samples, ref, 64, 4096)
    g.close()


def test_sphere_grid_converges(gpu):
    """the material-spheres grid at 256 spp (240x135, the configuration's sample count)"""
    from mitrace_amd.tracer import GpuScene
    g = GpuScene(scenegen.sphere_grid(env_width=512, env_height=256))
    fn = lambda **kw: scenegen.sphere_grid_settings(width=240, height=135, **kw)
    img, samples, _ = g.render(fn(samples=256, seed=12), want_samples=True)
    ref = _reference(g, fn, 256, 16, 200)
    _check(img, samples, ref, 256, 4096)
    g.close()
