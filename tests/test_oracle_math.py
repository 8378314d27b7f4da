"""The oracle's building blocks pinned against closed forms, textbook identities, brute force
and sampling distributions (no GPU).  Each test names the reference function it pins."""
import math

import numpy as np
import pytest

from oracle import pyoracle as po

f32 = np.float32


# ---------------------------------------------------------------- Moller-Trumbore (Intersect.cpp:34)
def test_triangle_closed_form():
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    hit, (u, v, t) = po.triangle_test([0.25, 0.25, 1, 0, 0, -1], tri)
    assert hit and (u, v, t) == (0.25, 0.25, 1.0)
    # two-sided: the reversed winding is hit too, with (u, v) swapped roles
    hit, (u, v, t) = po.triangle_test([0.25, 0.25, 1, 0, 0, -1], [0, 0, 0, 0, 1, 0, 1, 0, 0])
    assert hit and t == 1.0 and (u, v) == (0.25, 0.25)
    # from behind the plane the distance is negative: no t-range test inside the routine
    hit, (u, v, t) = po.triangle_test([0.25, 0.25, -2, 0, 0, -1], tri)
    assert hit and t == -2.0


def test_triangle_edges_and_parallel():
    tri = [0, 0, 0, 1, 0, 0, 0, 1, 0]
    assert po.triangle_test([0.0, 0.5, 1, 0, 0, -1], tri)[0]          # on edge u = 0
    assert po.triangle_test([0.5, 0.5, 1, 0, 0, -1], tri)[0]          # on edge u + v = 1
    assert po.triangle_test([0.0, 0.0, 1, 0, 0, -1], tri)[0]          # vertex
    assert not po.triangle_test([0.51, 0.5, 1, 0, 0, -1], tri)[0]
    assert not po.triangle_test([-0.01, 0.5, 1, 0, 0, -1], tri)[0]
    assert not po.triangle_test([0.2, 0.2, 1, 1, 0, 0], tri)[0]       # parallel: |det| = 0
    # |det| just above / below FLT_EPSILON
    tiny = [0, 0, 0, 1e-4, 0, 0, 0, 1e-4, 0]
    assert not po.triangle_test([2e-5, 2e-5, 1, 0, 0, -1], tiny)[0]   # det = 1e-8 < eps
    small = [0, 0, 0, 1e-3, 0, 0, 0, 1e-3, 0]
    assert po.triangle_test([2e-4, 2e-4, 1, 0, 0, -1], small)[0]      # det = 1e-6 > eps


def test_triangle_against_float64():
    rng = np.random.default_rng(5)
    checked = 0
    for _ in range(3000):
        tri = rng.uniform(-1, 1, 9)
        o = rng.uniform(-2, 2, 3)
        target = tri[:3] * 0.3 + tri[3:6] * 0.3 + tri[6:] * 0.4 + rng.normal(0, 0.3, 3)
        d = target - o
        d /= np.linalg.norm(d)
        ray = np.concatenate([o, d]).astype(np.float32)
        hit, (u, v, t) = po.triangle_test(ray, tri.astype(np.float32))
        # float64 reference of the same ray against the float32-rounded triangle
        v0, v1, v2 = (tri.astype(np.float32).astype(np.float64).reshape(3, 3))
        e1, e2 = v1 - v0, v2 - v0
        dd, oo = ray[3:].astype(np.float64), ray[:3].astype(np.float64)
        p = np.cross(dd, e2)
        det = e1 @ p
        if abs(det) < 1e-4:
            continue
        s = oo - v0
        uu = (s @ p) / det
        q = np.cross(s, e1)
        vv = (dd @ q) / det
        tt = (e2 @ q) / det
        inside = uu >= 0 and vv >= 0 and uu + vv <= 1
        margin = min(uu, vv, 1 - uu - vv)
        if abs(margin) < 1e-5:
            continue  # too close to an edge for a float32 decision to be pinned
        assert hit == inside
        if hit:
            assert abs(u - uu) < 1e-4 and abs(v - vv) < 1e-4 and abs(t - tt) < 1e-4 * max(1, abs(tt))
            checked += 1
    assert checked > 300


# ---------------------------------------------------------------- slab test (Intersect.cpp:9-18)
def test_slab_returns_exit_distance():
    box = [0, 0, 0, 1, 1, 1]
    assert po.slab([0.5, 0.5, -1, 0, 0, 1], box) == 2.0      # exit, not entry (1.0)
    assert po.slab([0.5, 0.5, 0.5, 0, 0, 1], box) == 0.5     # origin inside
    assert po.slab([0.5, 0.5, 2, 0, 0, 1], box) == -1.0      # box behind: exit < 0 -> returned as is? no: -1
    assert po.slab([2, 2, -1, 0, 0, 1], box) == -1.0         # miss
    d = np.array([1, 1, 1], np.float32) / np.float32(math.sqrt(3))
    t = po.slab([-1, -1, -1, *d], box)
    assert abs(t - 2 * math.sqrt(3)) < 1e-5


def test_slab_behind_box_reports_negative_exit():
    # when the whole box is behind the origin, compMin(tmax) is negative and the reference
    # returns that negative exit distance (culled by the caller's < 0 test)
    t = po.slab([0.5, 0.5, 3, 0, 0, 1], [0, 0, 0, 1, 1, 1])
    assert t < 0


# ---------------------------------------------------------------- 4x4 inverse (glm::inverse stand-in)
def test_inverse_matches_float64():
    rng = np.random.default_rng(2)
    for _ in range(50):
        M = np.eye(4)
        M[:3, :3] = rng.normal(size=(3, 3))
        M[:3, 3] = rng.normal(size=3)
        inv = po.invert4(M.astype(np.float32).T.ravel()).reshape(4, 4).T
        ref = np.linalg.inv(M.astype(np.float32).astype(np.float64))
        np.testing.assert_allclose(inv, ref, rtol=2e-4, atol=2e-5)


# ---------------------------------------------------------------- BRDF (BRDF.hpp)
def brdf(L, V, N, base, r, m):
    inp = np.concatenate([L, V, N, base, [r, m]]).astype(np.float32)[None]
    return po.brdf_eval(inp)[0]


@pytest.mark.parametrize("r,m", [(0.5, 0.0), (0.2, 1.0), (1.0, 0.3), (0.01, 0.0)])
def test_brdf_at_normal_incidence_closed_form(r, m):
    N = np.array([0, 0, 1.0])
    base = np.array([0.8, 0.4, 0.2])
    out = brdf(N, N, N, base, r, m)
    rr = max(r, 0.05)
    a2 = (rr * rr) ** 2
    D = 1.0 / (math.pi * a2)                       # NH = 1: D = a2 / (pi * a2^2)
    F = 0.04 * (1 - m) + base * m                  # F(1) = F0
    spec = D * 1.0 * F / 4.0                       # G(1, 1) = 1, 4 NV NL = 4
    diffuse = (1 - F) * (1 - m) * base / math.pi
    # float32 evaluates the GGX denominator NH^2 (a^2 - 1) + 1 with a cancellation at NH = 1
    # (relative error ~ 6e-8 / a^2), so the tolerance scales with 1 / a^2
    tol = 2e-5 + 1.2e-7 / a2
    np.testing.assert_allclose(out[:3], diffuse + spec, rtol=tol)
    assert out[3] == pytest.approx(1 / math.pi, rel=1e-6)        # cos / pi at cos = 1
    assert out[4] == pytest.approx(D / 4, rel=tol)               # D NH / (4 VH)


def test_fresnel_grazing_limit_and_below_horizon():
    N = np.array([0, 0, 1.0])
    below = brdf(np.array([0, 0.6, -0.8]), N, N, [0.5] * 3, 0.5, 0.0)
    assert np.all(below == 0)  # NL <= 0 -> 0 for value and both pdfs
    # specular probability saturates at F = 1 at grazing view (F(0) = 1), clamped to 0.95
    V = np.array([1.0, 0, 1e-4]) / np.linalg.norm([1.0, 0, 1e-4])
    inp = np.concatenate([V, N, [0.5] * 3, [0.5, 0.0], [0.99, 0.5, 0.5, 0.5, 0.5]])[None]
    out = po.sample_bounce(inp)
    assert out[0, 7] == 0.0  # u0 = 0.99 > 0.95: diffuse lobe chosen


def _sphere_dirs(n, rng):
    v = rng.normal(size=(n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def test_pdfs_integrate_over_the_sphere():
    """Monte Carlo over uniform sphere directions: the diffuse pdf integrates to 1 and the
    GGX reflection pdf D*NH/(4 VH) at normal view integrates to the visible GGX mass."""
    rng = np.random.default_rng(11)
    n = 400000
    L = _sphere_dirs(n, rng)
    N = np.tile([0, 0, 1.0], (n, 1))
    inp = np.concatenate([L, N, N, np.full((n, 3), 0.5), np.full((n, 1), 0.3), np.zeros((n, 1))], 1)
    out = po.brdf_eval(inp.astype(np.float32))
    area = 4 * math.pi
    diffuse_int = out[:, 3].mean() * area
    spec_int = out[:, 4].mean() * area
    assert diffuse_int == pytest.approx(1.0, abs=0.01)
    assert 0.95 < spec_int < 1.01


# ---------------------------------------------------------------- BRDF sampling (Tracer.cpp:34-82)
def test_cosine_sampling_distribution():
    """diffuse lobe: P(cos theta < c) = c^2 (Kolmogorov-Smirnov)."""
    rng = np.random.default_rng(3)
    n = 50000
    u = rng.random((n, 5))
    u[:, 0] = 0.999  # always the diffuse lobe
    inp = np.concatenate([np.tile([0, 0, 1.0], (n, 1)), np.tile([0, 0, 1.0], (n, 1)),
                          np.full((n, 3), 0.5), np.full((n, 1), 0.5), np.zeros((n, 1)), u], 1)
    out = po.sample_bounce(inp.astype(np.float32))
    c = np.sort(out[:, 2])
    emp = np.arange(1, n + 1) / n
    assert np.max(np.abs(emp - c ** 2)) < 1.63 / math.sqrt(n) * 1.5
    assert np.all(out[:, 7] == 0)


@pytest.mark.parametrize("r", [0.3, 0.6])
def test_ggx_half_vector_distribution(r):
    """specular lobe at normal view: H has tan^2(theta) = a^2 u/(1-u), i.e.
    P(cos theta_h >= c) = t/(a^2 + t) with t = tan^2 -> check via the reflected direction."""
    rng = np.random.default_rng(4)
    n = 40000
    u = rng.random((n, 5))
    u[:, 0] = 0.0  # always the specular lobe
    N = np.tile([0, 0, 1.0], (n, 1))
    inp = np.concatenate([N, N, np.full((n, 3), 0.5), np.full((n, 1), r), np.zeros((n, 1)), u], 1)
    out = po.sample_bounce(inp.astype(np.float32))
    L = out[:, :3].astype(np.float64)
    H = L + np.array([0, 0, 1.0])
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    ct = H[:, 2]
    tan2 = (1 - ct ** 2) / ct ** 2
    a = max(r * r, 0.0025)
    a2 = a * a
    cdf = tan2 / (a2 + tan2)  # the u that generated it
    # at normal view the reflection stays above the surface while theta_h < 45 deg, i.e.
    # u1 < 1 / (1 + a^2); beyond that the lobe falls back to cosine sampling (Tracer.cpp:70)
    kept = u[:, 1] < 1 / (1 + a2) - 1e-4
    fell = u[:, 1] > 1 / (1 + a2) + 1e-4
    assert kept.mean() > 0.8
    np.testing.assert_allclose(cdf[kept], u[kept, 1], atol=2e-3)
    if fell.any():  # fallback directions are cosine samples driven by u3, u4: z = sqrt(1 - u3)
        np.testing.assert_allclose(out[fell, 2], np.sqrt(1 - u[fell, 3]), atol=1e-5)


def test_sampling_energy_is_f_cos_over_pdf():
    rng = np.random.default_rng(8)
    n = 2000
    V = _sphere_dirs(n, rng)
    V[:, 2] = np.abs(V[:, 2]) + 0.05
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    N = np.tile([0, 0, 1.0], (n, 1))
    base = rng.uniform(0.2, 0.9, (n, 3))
    rough = rng.uniform(0, 1, (n, 1))
    metal = rng.uniform(0, 1, (n, 1))
    u = rng.random((n, 5))
    out = po.sample_bounce(np.concatenate([V, N, base, rough, metal, u], 1).astype(np.float32))
    f = po.brdf_eval(np.concatenate([out[:, :3], V, N, base, rough, metal], 1).astype(np.float32))
    nl = np.maximum(out[:, 2], 0)
    ok = out[:, 3] > 0
    np.testing.assert_allclose(out[ok, 4:7], f[ok, :3] * (nl[ok] / out[ok, 3])[:, None], rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------------- camera (Tracer/Camera.cpp:17-50)
def test_camera_perspective_corners():
    from mitrace_amd.scene import Camera
    cam = Camera(type=0, camera_to_world=np.eye(4, dtype=np.float32), yfov=math.radians(40), aspect=1.0)
    r = po.camera_rays(cam, [[0.5, 0.5], [0.0, 0.0], [1.0, 1.0]], 16 / 9)
    np.testing.assert_allclose(r[0], [0, 0, 0, 0, 0, -1], atol=1e-7)
    t = math.tan(math.radians(20))
    for row, sx, sy in ((1, -1, 1), (2, 1, -1)):  # v grows downward: row 0 is the top
        d = np.array([sx * t * 16 / 9, sy * t, -1.0])
        np.testing.assert_allclose(r[row, 3:], d / np.linalg.norm(d), rtol=1e-6)


def test_camera_orthographic():
    from mitrace_amd.scene import Camera
    M = np.eye(4, dtype=np.float32)
    M[:3, 3] = [1, 2, 3]
    cam = Camera(type=1, camera_to_world=M, ortho_size=(2.0, 1.0), aspect=2.0)
    r = po.camera_rays(cam, [[1.0, 0.0]], 2.0)
    np.testing.assert_allclose(r[0], [1 + 2, 2 + 1, 3, 0, 0, -1], atol=1e-6)


# ---------------------------------------------------------------- tone map, fireflies, RNG
def test_aces_closed_form():
    x = np.array([[0, 0, 0], [1, 2, 4], [1e6, 1e6, 1e6]], np.float32)
    out = po.aces(x, 0.0)
    c = x.astype(np.float64)
    ref = (c * (2.51 * c + 0.03)) / (c * (2.43 * c + 0.59) + 0.14)
    np.testing.assert_allclose(out, ref, rtol=1e-6)
    assert out[2, 0] == pytest.approx(2.51 / 2.43, rel=1e-5)
    np.testing.assert_allclose(po.aces(x * 256, 8.0), out, rtol=1e-6)  # L * 2^-ev


def test_firefly_rejection_hand_computed():
    s = np.full((100, 3), 0.5, np.float32)
    s[17] = 100.0
    out, n = po.cleanup_fireflies(s, 6.0)
    assert n == 99
    assert np.all(out[17] == 0) and out.sum(0)[0] / n == pytest.approx(0.5)
    # all samples equal: sigma = 0, nothing is strictly beyond 0, divisor = count
    out, n = po.cleanup_fireflies(np.ones((8, 3), np.float32), 6.0)
    assert n == 8


def test_counter_rng_uniform_and_decorrelated():
    n = 200000
    rng = np.random.default_rng(0)
    pix = rng.integers(0, 1 << 21, n).astype(np.uint32)
    smp = rng.integers(0, 256, n).astype(np.uint32)
    a = po.rng_draws(7, pix, smp, np.zeros(n, np.uint32))
    b = po.rng_draws(7, pix, smp, np.ones(n, np.uint32))
    c = po.rng_draws(8, pix, smp, np.zeros(n, np.uint32))
    for x in (a, b, c):
        assert x.min() >= 0 and x.max() < 1
        xs = np.sort(x)
        assert np.max(np.abs(xs - np.arange(1, n + 1) / n)) < 2.0 / math.sqrt(n)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.01 and abs(np.corrcoef(a, c)[0, 1]) < 0.01
    # replayable: the same key gives the same draw
    assert np.array_equal(a, po.rng_draws(7, pix, smp, np.zeros(n, np.uint32)))


# ---------------------------------------------------------------- elementary functions
# The path's sin, cos, atan2, acos and x^5 are this build's own (the oracle and the kernels
# compute them with the same operations); pinned against float64 numpy within a few units in
# the last place, with the special values and symmetries the C library guarantees.
from mitrace_amd import abi


def _ulp(got, ref):
    return np.abs(got.astype(np.float64) - ref) / np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)


def test_sin_cos_within_two_ulp():
    """the angles of the path: 2 pi u (cosine and GGX sampling, light-point longitude) and
    theta in [-pi/2, pi/2) (light-point latitude), plus a margin; results within 2 ulp of the
    float64 values (1e-8 absolute where the value crosses zero)"""
    x = np.linspace(-4.0, 8.0, 1_200_001).astype(np.float32)
    for op, fn in ((abi.MT_MATH_SIN, np.sin), (abi.MT_MATH_COS, np.cos)):
        g = po.math_eval(op, x)
        r = fn(x.astype(np.float64))
        big = np.abs(r) > 1e-3
        assert _ulp(g[big], r[big]).max() <= 2.0
        assert np.abs(g - r).max() < 1e-7
    for v, s, c in ((0.0, 0.0, 1.0), (-0.0, -0.0, 1.0)):
        assert po.math_eval(abi.MT_MATH_SIN, [v])[0] == s and po.math_eval(abi.MT_MATH_COS, [v])[0] == c
    # sin(-x) = -sin(x), cos(-x) = cos(x) inside the first reduction interval (at its end the
    # rounding of x 2/pi + 1/2 can pick neighbouring quadrants for x and -x)
    x = np.linspace(0.0, 0.78, 10001).astype(np.float32)
    assert np.array_equal(po.math_eval(abi.MT_MATH_SIN, -x), -po.math_eval(abi.MT_MATH_SIN, x))
    assert np.array_equal(po.math_eval(abi.MT_MATH_COS, -x), po.math_eval(abi.MT_MATH_COS, x))
    # sin^2 + cos^2 = 1 within float rounding
    x = np.random.default_rng(1).uniform(0, 2 * np.pi, 100000).astype(np.float32)
    s, c = po.math_eval(abi.MT_MATH_SIN, x).astype(np.float64), po.math_eval(abi.MT_MATH_COS, x).astype(np.float64)
    assert np.abs(s * s + c * c - 1).max() < 4e-7


def test_atan2_within_three_ulp_and_signed_zeros():
    """atan2 over every octant (random and axis-aligned arguments) within 3 ulp; the signed-zero
    cases of C's atan2: (+0, +0) = +0, (+0, -0) = pi, (-0, -0) = -pi, (-0, +0) = -0"""
    rng = np.random.default_rng(2)
    y = rng.normal(size=1_000_000).astype(np.float32)
    x = rng.normal(size=1_000_000).astype(np.float32)
    x[:1000] = 0.0
    y[1000:2000] = 0.0
    x[2000:3000] = y[2000:3000]            # the octant boundary t = 1
    x[3000:4000] = -y[3000:4000]
    g = po.math_eval(abi.MT_MATH_ATAN2, y, x)
    r = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    nz = np.abs(r) > 1e-3
    assert _ulp(g[nz], r[nz]).max() <= 3.0
    assert np.abs(g - r).max() < 3e-7
    for (a, b), want in (((0.0, 0.0), 0.0), ((0.0, -0.0), math.pi), ((-0.0, -0.0), -math.pi),
                         ((-0.0, 0.0), -0.0), ((2.0, 0.0), math.pi / 2), ((0.0, -3.0), math.pi)):
        got = po.math_eval(abi.MT_MATH_ATAN2, [a], [b])[0]
        assert got == np.float32(want) and math.copysign(1, got) == math.copysign(1, want)
    # the longitude of SampleEquirectangular: (atan2(z, x) + pi) / 2 pi stays in [0, 1]
    u = (g + np.float32(math.pi)) / np.float32(2 * math.pi)
    assert u.min() >= 0.0 and u.max() <= 1.0


def test_acos_within_four_ulp():
    """acos on [-1, 1] (the polar angle y/|d| of SampleEquirectangular) within 4 ulp, exact at
    the ends and the middle; NaN outside [-1, 1] like the C library"""
    x = np.linspace(-1.0, 1.0, 2_000_001).astype(np.float32)
    g = po.math_eval(abi.MT_MATH_ACOS, x)
    r = np.arccos(x.astype(np.float64))
    nz = r > 1e-3
    assert _ulp(g[nz], r[nz]).max() <= 4.0
    assert np.abs(g - r).max() < 4e-7
    ends = po.math_eval(abi.MT_MATH_ACOS, [1.0, -1.0, 0.0])
    assert ends[0] == 0.0 and ends[1] == np.float32(math.pi) and ends[2] == np.float32(math.pi / 2)
    assert np.isnan(po.math_eval(abi.MT_MATH_ACOS, [1.0000001, -1.5])).all()


def test_pow5_exact_products():
    """x^5 of Schlick's Fresnel (BRDF.hpp:9-13) as (x^2 x^2) x: within 3 ulp of float64 (three
    roundings), exact
    where the products are exact (0, 1, 1/2, 2^-k)"""
    x = np.linspace(0.0, 1.0, 1_000_001).astype(np.float32)
    g = po.math_eval(abi.MT_MATH_POW5, x)
    r = x.astype(np.float64) ** 5
    nz = r > 1e-30
    assert _ulp(g[nz], r[nz]).max() <= 3.0
    assert np.array_equal(po.math_eval(abi.MT_MATH_POW5, [0.0, 1.0, 0.5, 0.25]),
                          np.float32([0.0, 1.0, 1 / 32, 1 / 1024]))
