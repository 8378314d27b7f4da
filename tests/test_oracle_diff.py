"""The oracle's differentiable estimator (forward-mode tangents, detached sampling) pinned
against finite differences of the oracle's own functions.

* EvaluateBRDF tangents vs central differences of EvaluateBRDF (float32, h = 1e-3).
* The primal image of orc_render_diff equals orc_render's linear output.
* With metallic 0 the lobe probability does not depend on the base colour, and with Russian
  roulette off no decision depends on the texture or the emission: the estimator is then a
  smooth function of those parameters, and its central difference with common random numbers
  must equal the tangent image (up to O(h^2) and float32 rounding).
"""
import numpy as np
import pytest

from helpers import load_scene
from mitrace_amd import abi, scenegen
from oracle import pyoracle as po


def _brdf_cases(n, seed):
    rng = np.random.default_rng(seed)
    N = np.tile([0.0, 0.0, 1.0], (n, 1))
    def hemi():
        v = rng.normal(size=(n, 3))
        v[:, 2] = np.abs(v[:, 2]) + 0.2
        return v / np.linalg.norm(v, axis=1, keepdims=True)
    L, V = hemi(), hemi()
    base = rng.uniform(0.05, 0.95, (n, 3))
    rough = rng.uniform(0.08, 1.0, n)
    metal = rng.uniform(0.02, 0.98, n)
    return np.concatenate([L, V, N, base, rough[:, None], metal[:, None]], 1).astype(np.float32)


@pytest.mark.parametrize("which", ["base_r", "base_g", "base_b", "rough", "metal"])
def test_brdf_tangent_matches_central_difference(which):
    inp = _brdf_cases(512, 3)
    col = {"base_r": 9, "base_g": 10, "base_b": 11, "rough": 12, "metal": 13}[which]
    tin = np.zeros((len(inp), 5), np.float32)
    tin[:, col - 9] = 1.0
    f, df = po.brdf_tangent(inp, tin)
    np.testing.assert_allclose(f, po.brdf_eval(inp)[:, :3], rtol=1e-6, atol=1e-7)
    h = 1e-3
    ip, im = inp.copy(), inp.copy()
    ip[:, col] += h
    im[:, col] -= h
    fd = (po.brdf_eval(ip)[:, :3].astype(np.float64) - po.brdf_eval(im)[:, :3]) / (2 * h)
    scale = np.abs(fd).max(axis=1, keepdims=True) + 1e-3
    err = np.abs(df - fd) / scale
    assert np.percentile(err, 99) < 5e-3 and err.max() < 5e-2, (which, err.max())
    # the tangent is not trivially zero
    assert np.abs(df).max() > 1e-2


def test_brdf_tangent_at_clamps():
    """roughness below 0.05 and metallic outside [0, 1] are clamped: zero derivative there"""
    inp = _brdf_cases(64, 4)
    inp[:, 12] = 0.02
    tin = np.zeros((len(inp), 5), np.float32)
    tin[:, 3] = 1.0
    _, df = po.brdf_tangent(inp, tin)
    assert np.all(df == 0.0)


def _diff_scene():
    sc, params = scenegen.inverse_scene(tex_res=64, env_width=128, env_height=64, segments=16,
                                        rings=8)
    return sc, params


def _settings(**kw):
    s = scenegen.inverse_settings(width=24, height=24, samples=4, seed=11)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_diff_primal_equals_linear_render():
    sc = load_scene("cornell_box")
    orc = po.OracleScene(sc)
    st = _settings(width=20, height=20, samples=4)
    st.mode = abi.MT_MODE_PARITY
    img, dimg = orc.render_diff(st, None)
    ref, _, _, _ = orc.render(st, faithful_inverse=False)
    np.testing.assert_allclose(img, ref, rtol=2e-6, atol=1e-7)
    assert np.all(dimg == 0.0)


@pytest.mark.parametrize("group", ["tex", "emis"])
def test_diff_tangent_matches_crn_difference_when_decisions_are_fixed(group):
    sc, p = _diff_scene()
    orc = po.OracleScene(sc)
    st = _settings(rr_depth=99)  # no Russian roulette
    params = dict(p)
    params["metal"] = np.zeros_like(p["metal"])  # lobe probability independent of the base
    rng = np.random.default_rng(5)
    if group == "tex":
        tan = {"tex": rng.normal(size=p["tex"].shape).astype(np.float32)}
        h = 1e-2
    else:
        params["emis"] = np.full_like(p["emis"], 0.5)
        tan = {"emis": rng.normal(size=p["emis"].shape).astype(np.float32)}
        h = 1e-2
    img, dimg = orc.render_diff(st, params, tan)
    pp, pm = dict(params), dict(params)
    pp[group] = params[group] + h * tan[group]
    pm[group] = params[group] - h * tan[group]
    ip, _ = orc.render_diff(st, pp)
    im, _ = orc.render_diff(st, pm)
    fd = (ip.astype(np.float64) - im) / (2 * h)
    assert np.abs(fd).max() > 1e-3
    err = np.abs(dimg - fd)
    tol = 2e-3 * np.abs(fd).max() + 1e-4 * (np.abs(ip).max())
    assert err.max() < tol, (err.max(), tol)


def test_diff_tangent_is_linear_in_the_direction():
    """tangent(a*u + b*v) = a*tangent(u) + b*tangent(v) over all parameter groups at once"""
    sc, p = _diff_scene()
    orc = po.OracleScene(sc)
    st = _settings(width=16, height=16)
    rng = np.random.default_rng(9)
    def rnd():
        return {k: rng.normal(size=np.shape(p[k])).astype(np.float32)
                for k in ("tex", "rough", "metal", "emis")}
    u, v = rnd(), rnd()
    w = {k: (2.0 * u[k] - 0.5 * v[k]).astype(np.float32) for k in u}
    _, du = orc.render_diff(st, p, u)
    _, dv_ = orc.render_diff(st, p, v)
    _, dw = orc.render_diff(st, p, w)
    np.testing.assert_allclose(dw, 2.0 * du - 0.5 * dv_, rtol=1e-3,
                               atol=1e-4 * (np.abs(dw).max() + 1e-6))


# ---------------------------------------------------------------- the sampling's dependence
def _bounce_expectation(view, n, base, rough, metal, tan, R, G=192, G2=8):
    """E over one bounce's uniforms of w(x) R(x) (w = f cos / pdf, Tracer.cpp:177-200) by
    quadrature: the lobe choice weighted by its probability, the direction's two uniforms on a
    G x G midpoint grid, and for the specular samples whose reflection falls below the surface
    the cosine fallback's two uniforms on a G2 x G2 grid.  Also the expectations of the two
    derivative integrands: detached (dw R) and full (dw R + w R (score - dlog pdf))."""
    def run(u):
        inp = np.zeros((len(u), 22), np.float32)
        inp[:, 0:3], inp[:, 3:6], inp[:, 6:9] = view, n, base
        inp[:, 9], inp[:, 10], inp[:, 11:16], inp[:, 16:21] = rough, metal, tan, u
        return po.bounce_terms(inp)

    def parts(o):
        Rx = R(o[:, 0:3])[:, None]
        val = o[:, 3:6] * Rx
        return val, o[:, 6:9] * Rx + val * (o[:, 9:10] - o[:, 10:11]), o[:, 6:9] * Rx

    g = (np.arange(G) + 0.5) / G
    U1, U2 = np.meshgrid(g, g, indexing="ij")
    res = {}
    for lobe, u0 in (("spec", 0.0), ("diff", 0.99999)):
        u = np.zeros((G * G, 5), np.float32)
        u[:, 0], u[:, 1], u[:, 2], u[:, 3], u[:, 4] = u0, U1.ravel(), U2.ravel(), 0.5, 0.5
        o = run(u)
        val, full, det = parts(o)
        fb = np.nonzero(o[:, 12] > 0.5)[0]
        if len(fb):
            g2 = (np.arange(G2) + 0.5) / G2
            F3, F4 = np.meshgrid(g2, g2, indexing="ij")
            uu = np.repeat(u[fb], G2 * G2, axis=0)
            uu[:, 3], uu[:, 4] = np.tile(F3.ravel(), len(fb)), np.tile(F4.ravel(), len(fb))
            v2, f2, d2 = (x.reshape(len(fb), -1, 3).mean(1) for x in parts(run(uu)))
            val[fb], full[fb], det[fb] = v2, f2, d2
        res[lobe] = (o[0, 11], val.mean(0), full.mean(0), det.mean(0), len(fb) / len(o))
    ps = res["spec"][0]
    mix = lambda k: ps * res["spec"][k] + (1 - ps) * res["diff"][k]
    return mix(1), mix(2), mix(3), res["spec"][4]


@pytest.mark.parametrize("view,rough,metal", [
    ((0.2, 0.5, 0.84), 0.7, 0.5),   # a third of the GGX samples fall back to cosine sampling
    ((0.9, 0.2, 0.1), 0.8, 0.3),    # grazing view: almost half fall back
    ((0.0, 1.0, 0.0), 0.4, 0.9),    # normal view, shiny metal: few fall back
])
def test_full_derivative_of_one_bounce_equals_finite_differences(view, rough, metal):
    """The sampling's dependence on the parameters (Tracer.cpp:49-82, 177-200): the lobe is
    chosen with a probability that depends on base colour and metallic, the GGX half vector's
    density D(h) n.h depends on roughness, and where the reflection falls below the surface a
    cosine direction replaces it while the estimator still divides by the mixture pdf, so the
    direction's actual density is not that pdf.  By quadrature over the bounce's uniforms, the
    expectation of the full derivative integrand (score-function form) equals central
    differences of the expected bounce contribution for roughness, metallic and base colour,
    while the detached integrand misses the fallback's share."""
    n = np.array([0.0, 1.0, 0.0], np.float32)
    v = np.array(view, np.float32)
    v /= np.linalg.norm(v)
    base = np.array([0.7, 0.4, 0.2], np.float32)
    a = np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5])
    R = lambda x: 1.0 + 3.0 * np.maximum(0.0, x @ a) ** 3   # a smooth radiance field
    h = 1e-2
    for name, tan in (("rough", [0, 0, 0, 1, 0]), ("metal", [0, 0, 0, 0, 1]), ("base", [1, 0.5, -0.3, 0, 0])):
        tan = np.array(tan, np.float32)
        _, full, det, fb = _bounce_expectation(v, n, base, rough, metal, tan, R)
        Ep, *_ = _bounce_expectation(v, n, base + h * tan[:3], rough + h * tan[3], metal + h * tan[4], tan, R)
        Em, *_ = _bounce_expectation(v, n, base - h * tan[:3], rough - h * tan[3], metal - h * tan[4], tan, R)
        fd = (Ep - Em) / (2 * h)
        scale = np.abs(fd).max()
        # the quadrature resolves the fallback boundary to about 1/G
        assert np.abs(full - fd).max() <= 0.02 * scale, (name, fd, full)
        if fb > 0.3 and name != "base":  # there the detached derivative is far off
            assert np.abs(det - fd).max() > 0.1 * scale, (name, fd, det)
