// mt_device.hpp -- device-side arithmetic of the gfx950 Tracer: vector math with a fixed
// operation order, the counter RNG, texture/environment lookups, the BRDF, BRDF sampling and
// the closest-hit traversal.  Included by mt_render.hip only.
//
// Operation-order contract (the CPU oracle follows the same one, written separately):
//   float32, no FMA contraction (built with -ffp-contract=off), IEEE-rounded division/sqrt;
//   dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z; normalize(v) = v * (1/sqrt(dot(v,v)));
//   min(a,b) = (b < a) ? b : a; max(a,b) = (a < b) ? b : a (GLM's scalar forms);
//   mat4*vec4 row r = (m[0][r]*x + m[1][r]*y) + (m[2][r]*z + m[3][r]*w).
#pragma once

#include <hip/hip_runtime.h>

#include "mt_internal.hpp"

namespace mt {
namespace dev {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kFltMax = 3.402823466e+38f;
constexpr float kFltEps = 1.192092896e-07f;
constexpr float kMinDelta = 0.1f;       // traversal slack, Intersect.cpp:58
constexpr float kPullout = 0.0001f;     // Tracer.cpp:24
constexpr uint32_t kLinearTag = 0xFFFFFFFFu;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 mk(float a) { return {a, a, a}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 mul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 divv(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
__device__ __forceinline__ V3 scale(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ V3 vmin(V3 a, V3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
__device__ __forceinline__ V3 vmax(V3 a, V3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__device__ __forceinline__ float len(V3 v) { return __builtin_sqrtf(dot(v, v)); }
__device__ __forceinline__ V3 normalize(V3 v) { return scale(v, 1.0f / __builtin_sqrtf(dot(v, v))); }
__device__ __forceinline__ float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ V3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// ------------------------------------------------------------------ elementary functions
// sin, cos, atan2, acos and x^5 of the path, operation by operation as the CPU oracle
// computes them, so that renders agree with it bit for bit (the device library's sinf/cosf/
// atan2f/acosf/powf and the C library's differ in the last bits).  sin/cos: Cody & Waite
// (1980) reduction x - k pi/2 with pi/2 split into three floats (the first two with short
// mantissas, so k times them is exact for |k| < 2^12), Taylor polynomials of sin (through r^9)
// and cos (through r^10) on [-pi/4, pi/4].  atan2: octant symmetry to t in [0, 1], atan(t) =
// pi/4 + atan((t-1)/(t+1)) to |t| <= tan(pi/8), Taylor series through t^19.
// acos(x) = atan2(sqrt((1-x)(1+x)), x); x^5 = (x^2 x^2) x.
namespace em {
constexpr float kTwoOverPi = 0x1.45f306p-1f;
constexpr float kPio2a = 0x1.92p+0f, kPio2b = 0x1.fb6p-12f, kPio2c = -0x1.777a5cp-25f;
constexpr float kPiF = 0x1.921fb6p+1f, kPio2 = 0x1.921fb6p+0f, kPio4 = 0x1.921fb6p-1f;
constexpr float kTanPi8 = 0x1.a8279ap-2f;
__device__ __forceinline__ float sin_poly(float r) {
    const float r2 = r * r;
    float p = 0x1.71de3ap-19f;
    p = p * r2 + -0x1.a01a02p-13f;
    p = p * r2 + 0x1.111112p-7f;
    p = p * r2 + -0x1.555556p-3f;
    return r + (r * r2) * p;
}
__device__ __forceinline__ float cos_poly(float r) {
    const float r2 = r * r;
    float p = -0x1.27e4fcp-22f;
    p = p * r2 + 0x1.a01a02p-16f;
    p = p * r2 + -0x1.6c16c2p-10f;
    p = p * r2 + 0x1.555556p-5f;
    p = p * r2 + -0.5f;
    return 1.0f + r2 * p;
}
__device__ __forceinline__ void sincos(float x, float& s, float& c) {
    const float k = floorf(x * kTwoOverPi + 0.5f);
    const float r = ((x - k * kPio2a) - k * kPio2b) - k * kPio2c;
    const float ps = sin_poly(r), pc = cos_poly(r);
    const int q = static_cast<int>(k) & 3;
    s = q == 0 ? ps : (q == 1 ? pc : (q == 2 ? -ps : -pc));
    c = q == 0 ? pc : (q == 1 ? -ps : (q == 2 ? -pc : ps));
}
__device__ __forceinline__ float atan_poly(float t) {
    const float t2 = t * t;
    float p = -0x1.af286cp-5f;
    p = p * t2 + 0x1.e1e1e2p-5f;
    p = p * t2 + -0x1.111112p-4f;
    p = p * t2 + 0x1.3b13b2p-4f;
    p = p * t2 + -0x1.745d18p-4f;
    p = p * t2 + 0x1.c71c72p-4f;
    p = p * t2 + -0x1.24924ap-3f;
    p = p * t2 + 0x1.99999ap-3f;
    p = p * t2 + -0x1.555556p-2f;
    return t + (t * t2) * p;
}
__device__ __forceinline__ float atan2(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ay < ax ? ax : ay, mn = ay < ax ? ay : ax;
    float a;
    if (mx == 0.0f) {
        a = 0.0f;
    } else {
        float t = mn / mx;
        float off = 0.0f;
        if (t > kTanPi8) {
            t = (t - 1.0f) / (t + 1.0f);
            off = kPio4;
        }
        a = off + atan_poly(t);
    }
    if (ax < ay) a = kPio2 - a;
    if (signbit(x)) a = kPiF - a;
    return signbit(y) ? -a : a;
}
__device__ __forceinline__ float acos(float x) {
    return atan2(__builtin_sqrtf((1.0f - x) * (1.0f + x)), x);
}
__device__ __forceinline__ float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}
}  // namespace em

// column-major 4x4 times (p, w)
__device__ __forceinline__ V3 xform(const float* m, V3 p, float w) {
    V3 o;
    o.x = (m[0] * p.x + m[4] * p.y) + (m[8] * p.z + m[12] * w);
    o.y = (m[1] * p.x + m[5] * p.y) + (m[9] * p.z + m[13] * w);
    o.z = (m[2] * p.x + m[6] * p.y) + (m[10] * p.z + m[14] * w);
    return o;
}

// ------------------------------------------------------------------ counter RNG
__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
__device__ __forceinline__ uint32_t path_key(uint32_t seed_key, uint32_t pixel, uint32_t sample) {
    return mix32(mix32(seed_key ^ pixel) + sample * 0x9E3779B9u);
}
__device__ __forceinline__ float draw(uint32_t key, uint32_t dim) {
    const uint32_t x = mix32(key ^ mix32(dim + 0x7F4A7C15u));
    return static_cast<float>(x >> 8) * (1.0f / 16777216.0f);
}
// per-bounce dimension offsets; bounce k starts at 2 + k * (8 + light_count)
enum : uint32_t { kDimLobe = 0, kDimDir = 1, kDimFallback = 3, kDimLightPt = 5, kDimRR = 7,
                  kDimLightSel = 8 };

// ------------------------------------------------------------------ textures (Image.cpp)
struct Tex8 {
    const uint32_t* texels;
    int w, h;
    __device__ __forceinline__ float4 operator()(int x, int y) const {
        const uint32_t t = texels[y * w + x];
        return make_float4(static_cast<float>(t & 0xFFu) / 255.0f,
                           static_cast<float>((t >> 8) & 0xFFu) / 255.0f,
                           static_cast<float>((t >> 16) & 0xFFu) / 255.0f, 1.0f);
    }
};
struct TexF {
    const float4* texels;
    int w, h;
    __device__ __forceinline__ float4 operator()(int x, int y) const { return texels[y * w + x]; }
};
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}
template <typename F>
__device__ float4 sample_image(const F& f, float uvx, float uvy, int filter) {
    const int w = f.w, h = f.h;
    const float x = uvx - floorf(uvx);
    const float y = uvy - floorf(uvy);
    if (filter == MT_FILTER_NEAREST) {
        int ix = static_cast<int>(x * static_cast<float>(w)) % w;
        int iy = static_cast<int>(y * static_cast<float>(h)) % h;
        if (ix < 0) ix += w;
        if (iy < 0) iy += h;
        ix = ix > 0 ? ix : 0;
        iy = iy > 0 ? iy : 0;
        return f(ix, iy);
    }
    // weights in uv units (Image.cpp:26-44); first texel clamped where x*W rounds up to W
    int x0 = static_cast<int>(floorf(x * static_cast<float>(w)));
    int y0 = static_cast<int>(floorf(y * static_cast<float>(h)));
    x0 = x0 < w - 1 ? x0 : w - 1;
    y0 = y0 < h - 1 ? y0 : h - 1;
    x0 = x0 > 0 ? x0 : 0;
    y0 = y0 > 0 ? y0 : 0;
    const int x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1;
    const int y1 = y0 + 1 < h - 1 ? y0 + 1 : h - 1;
    const float sx = x - static_cast<float>(x0) / static_cast<float>(w);
    const float sy = y - static_cast<float>(y0) / static_cast<float>(h);
    const float4 c0 = mix4(f(x0, y0), f(x1, y0), sx);
    const float4 c1 = mix4(f(x0, y1), f(x1, y1), sx);
    return mix4(c0, c1, sy);
}
__device__ __forceinline__ float4 sample_tex(const DevScene& S, const DevTex& t, float u, float v) {
    const DevImage im = S.images[t.image];
    Tex8 f{S.texels + im.offset, static_cast<int>(im.w), static_cast<int>(im.h)};
    return sample_image(f, u * t.sx + t.ox, v * t.sy + t.oy, t.filter);
}
// Scene::SampleSkybox without the intensity (Scene.cpp:61-73 -> Image.cpp:63-75)
__device__ V3 sample_env(const DevScene& S, V3 d, int apply_rot, float rot_offset) {
    if (!S.has_env) return mk(0.0f);
    const float theta = em::atan2(d.z, d.x);
    const float phi = em::acos(d.y / len(d));
    float u = (theta + kPi) / (2.0f * kPi);
    const float v = phi / kPi;
    if (apply_rot) u = u * 1.0f + rot_offset;
    TexF f{S.env, static_cast<int>(S.env_w), static_cast<int>(S.env_h)};
    const float4 c = sample_image(f, u, v, S.env_filter);
    return {c.x, c.y, c.z};
}

// ------------------------------------------------------------------ BRDF (BRDF.hpp)
__device__ __forceinline__ V3 fresnel(float c, V3 f0) {
    c = clampf(c, 0.0f, 1.0f);
    const float p = em::pow5(1.0f - c);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
__device__ __forceinline__ V3 base_refl(V3 base, float m) {
    m = clampf(m, 0.0f, 1.0f);
    return {mixf(0.04f, base.x, m), mixf(0.04f, base.y, m), mixf(0.04f, base.z, m)};
}
__device__ __forceinline__ float ggx_d(float nh, float r) {
    nh = clampf(nh, 0.0f, 1.0f);
    const float a = r * r;
    const float a2 = a * a;
    float den = (nh * nh) * (a2 - 1.0f) + 1.0f;
    den = kPi * den * den;
    return a2 / den;
}
__device__ __forceinline__ float smith_g1(float nv, float r) {
    nv = clampf(nv, 0.0f, 1.0f);
    const float rr = r + 1.0f;
    const float k = (rr * rr) / 8.0f;
    return nv / (nv * (1.0f - k) + k);
}
__device__ V3 eval_brdf(V3 L, V3 V, V3 N, V3 base, float rough, float metal) {
    rough = gmax(rough, 0.05f);
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return mk(0.0f);
    const V3 h = normalize(add(V, L));
    const float nh = gmax(dot(N, h), 0.0f);
    const float hv = gmax(dot(h, V), 0.0f);
    const V3 f = fresnel(hv, base_refl(base, metal));
    const float d = ggx_d(nh, rough);
    const float g = smith_g1(nv, rough) * smith_g1(nl, rough);
    const V3 kd = scale(sub(mk(1.0f), f), 1.0f - metal);
    const V3 diffuse = divs(mul(kd, base), kPi);
    const V3 spec = divs(scale(f, d * g), gmax(4.0f * nv * nl, 0.001f));
    return vmax(add(diffuse, spec), mk(0.0f));
}
// EvaluateBRDF with its partial derivatives (for the adjoint): df_dbase[c] = d f_c / d base_c
// (channels are independent), df_drough, df_dmetal per channel.  A clamp contributes the
// derivative of the branch it selects: zero for roughness < 0.05, metallic outside [0, 1],
// a negative sum in the final max(., 0), and NL/NV <= 0.
struct BrdfGrad {
    V3 f, df_dbase, df_drough, df_dmetal;
};
__device__ BrdfGrad eval_brdf_grad(V3 L, V3 V, V3 N, V3 base, float rough_in, float metal) {
    BrdfGrad g;
    g.f = g.df_dbase = g.df_drough = g.df_dmetal = mk(0.0f);
    const float rough = gmax(rough_in, 0.05f);
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return g;
    const V3 h = normalize(add(V, L));
    const float nh = clampf(gmax(dot(N, h), 0.0f), 0.0f, 1.0f);
    const float hv = clampf(gmax(dot(h, V), 0.0f), 0.0f, 1.0f);
    const float mc = clampf(metal, 0.0f, 1.0f);
    const float p5 = em::pow5(1.0f - hv);
    const float a = rough * rough;
    const float a2 = a * a;
    const float den = (nh * nh) * (a2 - 1.0f) + 1.0f;
    const float D = a2 / (kPi * den * den);
    const float k = ((rough + 1.0f) * (rough + 1.0f)) / 8.0f;
    const float q1 = nv * (1.0f - k) + k, q2 = nl * (1.0f - k) + k;
    const float g1 = nv / q1, g2 = nl / q2;
    const float G = g1 * g2;
    const float den4 = gmax(4.0f * nv * nl, 0.001f);
    // roughness derivatives (zero below the 0.05 clamp)
    float dD = 0.0f, dG = 0.0f;
    if (!(rough_in < 0.05f)) {  // the max(rough, 0.05) passes rough through
        const float dD_da2 = (den - 2.0f * a2 * nh * nh) / (kPi * den * den * den);
        dD = dD_da2 * 4.0f * rough * rough * rough;
        const float dk = (rough + 1.0f) / 4.0f;
        const float dg1 = -nv * (1.0f - nv) / (q1 * q1) * dk;
        const float dg2 = -nl * (1.0f - nl) / (q2 * q2) * dk;
        dG = dg1 * g2 + g1 * dg2;
    }
    const float bc[3] = {base.x, base.y, base.z};
    float fo[3], db[3], dr[3], dm[3];
    for (int c = 0; c < 3; ++c) {
        const float f0 = 0.04f * (1.0f - mc) + bc[c] * mc;
        const float F = f0 + (1.0f - f0) * p5;
        const float kd = (1.0f - F) * (1.0f - metal);
        const float diffuse = kd * bc[c] / kPi;
        const float spec = D * G * F / den4;
        fo[c] = gmax(diffuse + spec, 0.0f);
        const bool live = !((diffuse + spec) < 0.0f);  // max(x, 0) passes x through
        const float dF_db = mc * (1.0f - p5);
        const float dkd_db = -dF_db * (1.0f - metal);
        db[c] = live ? (dkd_db * bc[c] + kd) / kPi + D * G * dF_db / den4 : 0.0f;
        dr[c] = live ? (dD * G + D * dG) * F / den4 : 0.0f;
        const float dF_dm = (metal >= 0.0f && metal <= 1.0f) ? (bc[c] - 0.04f) * (1.0f - p5) : 0.0f;
        const float dkd_dm = -dF_dm * (1.0f - metal) - (1.0f - F);
        dm[c] = live ? dkd_dm * bc[c] / kPi + D * G * dF_dm / den4 : 0.0f;
    }
    g.f = {fo[0], fo[1], fo[2]};
    g.df_dbase = {db[0], db[1], db[2]};
    g.df_drough = {dr[0], dr[1], dr[2]};
    g.df_dmetal = {dm[0], dm[1], dm[2]};
    return g;
}

__device__ __forceinline__ float diffuse_pdf(V3 L, V3 N) { return gmax(dot(N, L), 0.0f) / kPi; }
__device__ float specular_pdf(V3 L, V3 V, V3 N, float rough) {
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
    const V3 h = normalize(add(V, L));
    const float nh = gmax(dot(N, h), 0.0f);
    const float vh = gmax(dot(V, h), 0.0f);
    if (vh <= 0.0f) return 0.0f;
    return ggx_d(nh, gmax(rough, 0.05f)) * nh / (4.0f * vh);
}

// ------------------------------------------------------------------ sampling (Tracer.cpp:26-121)
__device__ __forceinline__ void tbn(V3 n, V3& t, V3& b) {
    const V3 up = (fabsf(n.z) < 0.999f) ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    t = normalize(cross(up, n));
    b = cross(n, t);
}
__device__ V3 cosine_dir(float u1, float u2, V3 n) {
    const float r = __builtin_sqrtf(u1);
    const float phi = 2.0f * kPi * u2;
    V3 t, b;
    tbn(n, t, b);
    float s, c;
    em::sincos(phi, s, c);
    const float lz = __builtin_sqrtf(gmax(0.0f, 1.0f - u1));
    return normalize(add(add(scale(t, r * c), scale(b, r * s)), scale(n, lz)));
}
__device__ bool ggx_dir(float u1, float u2, V3 incident, V3 n, float rough, V3& out, float& hcos) {
    const float alpha = gmax(rough * rough, 0.0025f);
    const float alpha2 = alpha * alpha;
    const float phi = 2.0f * kPi * u2;
    const float tan2 = alpha2 * u1 / gmax(1.0f - u1, 0.0001f);
    const float ct = 1.0f / __builtin_sqrtf(1.0f + tan2);
    hcos = ct;  // n.h of the sampled half vector, before any flip
    const float st = __builtin_sqrtf(gmax(0.0f, 1.0f - ct * ct));
    V3 t, b;
    tbn(n, t, b);
    float s, c;
    em::sincos(phi, s, c);
    V3 h = normalize(add(add(scale(t, st * c), scale(b, st * s)), scale(n, ct)));
    if (dot(h, neg(incident)) < 0.0f) h = neg(h);
    const V3 refl = sub(incident, scale(scale(h, dot(h, incident)), 2.0f));  // glm::reflect
    if (dot(refl, n) > 0.0f) {
        out = normalize(refl);
        return true;
    }
    return false;
}
// How one bounce's sampling depends on base colour (base' = 0.96 base + 0.04), roughness and
// metallic, for the full derivative of the expected image (score-function form): the tangent
// factor d log p(decisions) - [throughput updated] d log pdf(direction) per parameter.  The
// decisions: the lobe, chosen with p = clamp(mean_c F_c, 0.05, 0.95), F_c = f0_c + (1 - f0_c)
// (1 - n.v)^5, f0_c = mix(0.04, base'_c, clamp(metal, 0, 1)) (Tracer.cpp:74-82); on the
// specular lobe the half vector, drawn with density D(h) n.h at roughness max(r, 0.05)
// (Tracer.cpp:49-72; drawn also when its reflection falls below the surface and a cosine
// direction replaces it).  The estimator divides by the mixture pdf (Tracer.cpp:190-196).
struct SampleScore {
    float base, rough, metal;  // base: the same for every colour channel
};
__device__ __forceinline__ float dlog_ggx_d(float nh, float r) {  // d log D(h) / d roughness
    if (r < 0.05f) return 0.0f;
    nh = clampf(nh, 0.0f, 1.0f);
    const float a = r * r;
    const float A = a * a;
    const float nh2 = nh * nh;
    const float den = nh2 * (A - 1.0f) + 1.0f;
    return 4.0f * r * r * r * (1.0f / A - 2.0f * nh2 / den);
}
__device__ SampleScore sample_score(V3 view, V3 N, V3 basec, float rough, float metal, float ps,
                                    bool spec_lobe, float hcos, V3 dir, float pdf, bool pdf_term) {
    const float c = clampf(gmax(dot(N, view), 0.0f), 0.0f, 1.0f);
    const float p5 = em::pow5(1.0f - c);
    const float mc = clampf(metal, 0.0f, 1.0f);
    const float bc[3] = {basec.x, basec.y, basec.z};
    float fsum = 0.0f, dm = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float f0 = 0.04f * (1.0f - mc) + bc[k] * mc;
        fsum += f0 + (1.0f - f0) * p5;
        dm += (bc[k] - 0.04f) * (1.0f - p5);
    }
    const float pbar = fsum / 3.0f;
    const bool live = !(pbar < 0.05f) && !(0.95f < pbar);  // the clamp passes p through
    const float dps_m = (live && metal >= 0.0f && metal <= 1.0f) ? dm / 3.0f : 0.0f;
    const float dps_b = live ? mc * (1.0f - p5) / 3.0f : 0.0f;
    SampleScore s;
    if (spec_lobe) {
        s.metal = dps_m / ps;
        s.base = dps_b / ps;
        s.rough = dlog_ggx_d(hcos, rough);
    } else {
        s.metal = -dps_m / (1.0f - ps);
        s.base = -dps_b / (1.0f - ps);
        s.rough = 0.0f;
    }
    if (pdf_term && pdf > 0.0f) {
        const float dp = diffuse_pdf(dir, N), sp = specular_pdf(dir, view, N, rough);
        float dsp_r = 0.0f;
        if (sp > 0.0f) {
            const V3 h = normalize(add(view, dir));
            dsp_r = sp * dlog_ggx_d(gmax(dot(N, h), 0.0f), rough);
        }
        s.metal -= dps_m * (sp - dp) / pdf;
        s.base -= dps_b * (sp - dp) / pdf;
        s.rough -= ps * dsp_r / pdf;
    }
    return s;
}

__device__ __forceinline__ float spec_probability(V3 view, V3 n, V3 base, float metal) {
    const V3 f = fresnel(gmax(dot(n, view), 0.0f), base_refl(base, metal));
    return clampf((f.x + f.y + f.z) / 3.0f, 0.05f, 0.95f);
}
__device__ V3 shading_normal(V3 surf, V3 texn) {  // Tracer.cpp:95-105
    const V3 n = normalize(surf);
    const V3 up = (fabsf(n.y) < 0.9f) ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 t = normalize(cross(n, up));
    const V3 b = normalize(cross(t, n));
    return normalize(add(add(scale(t, texn.x), scale(b, texn.y)), scale(n, texn.z)));
}
__device__ __forceinline__ V3 aces(V3 c, float expo) {  // Tracer.cpp:212-222
    const V3 x = scale(c, expo);
    const V3 num = mul(x, add(scale(x, 2.51f), mk(0.03f)));
    const V3 den = add(mul(x, add(scale(x, 2.43f), mk(0.59f))), mk(0.14f));
    return divv(num, den);
}

// ------------------------------------------------------------------ traversal (Intersect.cpp)
__device__ __forceinline__ float slab(V3 o, V3 inv, V3 mn, V3 mx) {
    const V3 t1 = mul(sub(mn, o), inv);
    const V3 t2 = mul(sub(mx, o), inv);
    const V3 lo = vmin(t1, t2), hi = vmax(t1, t2);
    const float tmin = gmax(gmax(lo.x, lo.y), lo.z);
    const float tmax = gmin(gmin(hi.x, hi.y), hi.z);
    return tmin <= tmax ? tmax : -1.0f;
}

// Slab test with hardware min/max.  For finite t values it makes the same decisions as
// slab(): v_min/v_max and the GLM ternaries differ only in the sign of a zero result, and
// every consumer of the exit distance compares it (>= 0, <= best + slack, <, > best + slack),
// where -0 and +0 are equal.  NaN t values occur only for an infinite inverse direction
// component (a direction component of exactly 0), and callers use slab() for those rays.
__device__ __forceinline__ float slab_fast(V3 o, V3 inv, V3 mn, V3 mx) {
    const V3 t1 = mul(sub(mn, o), inv);
    const V3 t2 = mul(sub(mx, o), inv);
    const float tmin = fmaxf(fmaxf(fminf(t1.x, t2.x), fminf(t1.y, t2.y)), fminf(t1.z, t2.z));
    const float tmax = fminf(fminf(fmaxf(t1.x, t2.x), fmaxf(t1.y, t2.y)), fmaxf(t1.z, t2.z));
    return tmin <= tmax ? tmax : -1.0f;
}
__device__ __forceinline__ bool finite3(V3 v) {
    return __builtin_isfinite(v.x) && __builtin_isfinite(v.y) && __builtin_isfinite(v.z);
}

// two-sided Moller-Trumbore with the bounds tested against det before scaling
__device__ __forceinline__ bool tri_test(V3 o, V3 d, V3 v0, V3 e1, V3 e2, float& u, float& v, float& t) {
    const V3 p = cross(d, e2);
    const float det = dot(e1, p);
    V3 q;
    if (det > kFltEps) {
        const V3 s = sub(o, v0);
        u = dot(s, p);
        if (u < 0.0f || u > det) return false;
        q = cross(s, e1);
        v = dot(d, q);
        if (v < 0.0f || u + v > det) return false;
    } else if (det < -kFltEps) {
        const V3 s = sub(o, v0);
        u = dot(s, p);
        if (u > 0.0f || u < det) return false;
        q = cross(s, e1);
        v = dot(d, q);
        if (v > 0.0f || u + v < det) return false;
    } else {
        return false;
    }
    const float id = 1.0f / det;
    t = dot(e2, q) * id;
    u = u * id;
    v = v * id;
    return true;
}

struct Hit {
    int inst;
    uint32_t tri, bv;
    float u, v, dist;
    V3 wpos;
    uint32_t bvh_tests, tri_tests;  // of the winning instance
};

struct Work {
    uint32_t aabb, enter, pops, tris;
    // traversal diagnostics (counting kernels only): lane steps by kind, wave-level iterations
    // and iterations in which any lane took each kind of step
    uint32_t st_setup, st_cull, st_inner, st_leaf, it_wave, it_setup, it_inner, it_leaf, it_mask;
};

// Traversal stack: entries [0, lds_depth) in LDS (lds[i * lds_stride]), deeper entries in a
// global spill area (spill[(i - lds_depth) * spill_stride]).  cap is the allocated depth; an
// overflow sets *err and drops the entry (never writes out of bounds).
struct Stack {
    uint2* lds;
    uint32_t lds_stride, lds_depth;
    uint2* spill;
    uint32_t spill_stride;
    uint32_t cap;
    uint32_t* err;
    uint32_t sp;
    __device__ __forceinline__ void push(uint32_t node, float d) {
        const uint2 e = make_uint2(node, __float_as_uint(d));
        if (sp < lds_depth) {
            lds[sp * lds_stride] = e;
        } else if (sp < cap) {
            spill[(sp - lds_depth) * spill_stride] = e;
        } else {
            atomicOr(err, 1u);
            return;
        }
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        return sp < lds_depth ? lds[sp * lds_stride] : spill[(sp - lds_depth) * spill_stride];
    }
};

// Per-lane state of the closest-hit engine: IntersectScene (Intersect.cpp:201-219) over
// IntersectMeshInstance (143-199), IntersectTrianglesBVH (47-141) and
// IntersectTrianglesLinear (20-45).  Instances are culled by their world AABB in groups of
// 64 (instance_mask); a lane then walks its entered instances in index order.  Work is split
// into small steps -- instance set-up/finish, one node pop, one triangle test -- so the
// kernel can batch lanes that need the same kind of step and retire/refill lanes
// independently.  Each lane performs exactly the reference's sequence of node pops and
// triangle tests; the per-instance results and the strict '<' across instances in index
// order are those of the reference.
constexpr uint32_t kNoLeaf = 0xFFFFFFFFu;
enum : uint32_t { kStepSetup = 0, kStepCull = 1, kStepInner = 2, kStepLeaf = 3 };

struct Trav {
    V3 o, d, invd;
    Hit R;
    uint64_t mask;
    uint32_t cb;   // first instance of the current group of 64
    int cur;       // instance (within the group) being traversed, -1 = none
    uint32_t leaf; // next prim slot of the leaf being tested, kNoLeaf = none
    uint32_t nxt_word;  // the node to pop next, kept in registers (has_next)
    float nxt_dist;
    bool has_next;
    bool fast;     // local ray has a finite inverse direction: slab_fast is exact
    bool tl;       // instances come from the top-level BVH (else groups of 64 in index order)
    uint32_t base; // stack depth below which the entries belong to the top-level BVH
    V3 lo, ld, linv;
    float ldist, bu, bvv;
    uint32_t btri, bbv, btests, ttests, pair_base, prim_base;
};

__device__ __forceinline__ void trav_begin(Trav& t, V3 o, V3 d) {
    t.o = o;
    t.d = d;
    t.invd = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
    t.R.inst = -1;
    t.R.dist = kFltMax;
    t.R.tri = t.R.bv = 0;
    t.R.u = t.R.v = 0.0f;
    t.R.wpos = mk(0.0f);
    t.R.bvh_tests = t.R.tri_tests = 0;
    t.mask = 0;
    t.cb = 0;
    t.cur = -1;
    t.leaf = kNoLeaf;
    t.has_next = false;
    t.fast = false;
    t.tl = false;
    t.base = 0;
}

__device__ __forceinline__ uint32_t trav_instance(const DevScene& S, const Trav& t, int j) {
    return t.tl ? S.tlas_inst[t.cb + j] : t.cb + static_cast<uint32_t>(j);
}

// Start the top-level traversal of a ray whose inverse direction is finite: push the root.
__device__ __forceinline__ void trav_tlas_begin(const DevScene& S, Trav& t, Stack& st) {
    t.tl = true;
    t.mask = 0;
    t.cur = -1;
    t.leaf = kNoLeaf;
    t.has_next = false;
    t.base = 0;
    st.sp = 0;
    const V3 mn = {S.tlas_root_min[0], S.tlas_root_min[1], S.tlas_root_min[2]};
    const V3 mx = {S.tlas_root_max[0], S.tlas_root_max[1], S.tlas_root_max[2]};
    if (slab_fast(t.o, t.invd, mn, mx) >= 0.0f) st.push(S.tlas_root_word, 0.0f);
    t.base = st.sp;  // the lane starts with a set-up step, which walks the top level
}

// world-AABB test of instances [cb, cb + 64) (Intersect.cpp:147); cb must be wave-uniform so
// the instance records are read with scalar loads
template <bool kCount, bool kFast>
__device__ __forceinline__ uint64_t instance_mask(const DevScene& S, uint32_t cb, V3 o, V3 invd,
                                                  Work& w) {
    const uint32_t nch = (S.instance_count - cb) < 64u ? (S.instance_count - cb) : 64u;
    uint64_t mask = 0;
    for (uint32_t j = 0; j < nch; ++j) {
        const DevInstance* I = S.instances + cb + j;
        const V3 mn = {I->wmin[0], I->wmin[1], I->wmin[2]};
        const V3 mx = {I->wmax[0], I->wmax[1], I->wmax[2]};
        const float te = kFast ? slab_fast(o, invd, mn, mx) : slab(o, invd, mn, mx);
        if (te >= 0.0f) mask |= (1ull << j);
    }
    if (kCount) w.aabb += nch;
    return mask;
}

// The lane needs a set-up step: no node is pending and no leaf is open.
__device__ __forceinline__ bool trav_needs_setup(const Trav& t, const Stack& st) {
    return st.sp == t.base && !t.has_next && t.leaf == kNoLeaf;
}

// Set-up step: finish the current instance (combine its hit), then start the next entered
// instance.  Returns true when the ray is finished; sets need_mask when the lane moved on to
// the next group of 64 instances.
template <bool kCount, bool kDiag>
__device__ __forceinline__ bool trav_setup(const DevScene& S, Trav& t, Stack& st, Work& w, bool& need_mask) {
    if (t.cur >= 0) {
        if (kCount) {
            w.pops += t.btests;
            w.tris += t.ttests;
        }
        if (t.ldist < kFltMax) {
            const uint32_t ii = trav_instance(S, t, t.cur);
            const float* M = S.instances[ii].M;
            const V3 lp = add(t.lo, scale(t.ld, t.ldist));
            const V3 wp = xform(M, lp, 1.0f);
            const float wd = len(sub(wp, t.o));
            // strict '<' in instance order (Intersect.cpp:210); the top level visits
            // instances out of order, so an equal distance goes to the lower index
            if (wd < t.R.dist || (wd == t.R.dist && static_cast<int>(ii) < t.R.inst)) {
                t.R.dist = wd;
                t.R.inst = static_cast<int>(ii);
                t.R.tri = t.btri;
                t.R.u = t.bu;
                t.R.v = t.bvv;
                t.R.wpos = wp;
                if (kDiag) {
                    t.R.bv = t.bbv;
                    t.R.bvh_tests = t.btests;
                    t.R.tri_tests = t.ttests;
                }
            }
        }
        t.cur = -1;
    }
    if (t.mask == 0) {
        if (!t.tl) {
            if (t.cb + 64u < S.instance_count) {
                t.cb += 64u;
                need_mask = true;
                return false;
            }
            return true;
        }
        // top level: pop nodes until a leaf yields instances whose world box the ray crosses
        // (the reference's own test, Intersect.cpp:147)
        while (true) {
            if (st.sp == 0) return true;
            const uint32_t word = st.pop().x;
            if (word & kLeafBit) {
                const uint32_t first = word & 0x00FFFFFFu, cnt = (word >> 24) & 0x7Fu;
                uint64_t m = 0;
                for (uint32_t j = 0; j < cnt; ++j) {
                    const DevInstance* I = S.instances + S.tlas_inst[first + j];
                    const V3 mn = {I->wmin[0], I->wmin[1], I->wmin[2]};
                    const V3 mx = {I->wmax[0], I->wmax[1], I->wmax[2]};
                    if (slab_fast(t.o, t.invd, mn, mx) >= 0.0f) m |= (1ull << j);
                }
                if (kCount) w.aabb += cnt;
                if (m) {
                    t.cb = first;
                    t.mask = m;
                    break;
                }
            } else {
                const float4* q = S.tlas_pairs + 4ull * word;
                const float4 a0 = q[0], a1 = q[1], b0 = q[2], b1 = q[3];
                if (slab_fast(t.o, t.invd, xyz(a0), xyz(a1)) >= 0.0f) st.push(__float_as_uint(a0.w), 0.0f);
                if (slab_fast(t.o, t.invd, xyz(b0), xyz(b1)) >= 0.0f) st.push(__float_as_uint(a1.w), 0.0f);
            }
        }
    }
    t.cur = __builtin_ctzll(t.mask);
    t.mask &= t.mask - 1;
    t.base = st.sp;
    const DevInstance* I = S.instances + trav_instance(S, t, t.cur);
    float inv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) inv[k] = I->inv[k];
    t.lo = xform(inv, t.o, 1.0f);
    t.ld = normalize(xform(inv, t.d, 0.0f));
    t.linv = {1.0f / t.ld.x, 1.0f / t.ld.y, 1.0f / t.ld.z};
    t.ldist = kFltMax;
    t.bu = t.bvv = 0.0f;
    t.btri = t.bbv = t.btests = t.ttests = 0;
    const DevMesh& m = S.meshes[I->mesh];
    t.pair_base = m.pair_base;
    t.prim_base = m.prim_base;
    if (kCount) w.enter++;
    t.fast = finite3(t.linv) && finite3(t.lo);
    if (m.flags & kMeshHasBvh) {
        const V3 rmn = {m.root_min[0], m.root_min[1], m.root_min[2]};
        const V3 rmx = {m.root_max[0], m.root_max[1], m.root_max[2]};
        t.nxt_word = m.root_word;
        t.nxt_dist = t.fast ? slab_fast(t.lo, t.linv, rmn, rmx) : slab(t.lo, t.linv, rmn, rmx);
        t.has_next = true;
    } else {
        t.leaf = 0;  // linear scan: the whole mesh is one leaf, no node pops counted
    }
    return false;
}

// One node pop (Intersect.cpp:67-78, 109-139).  The nearer accepted child is kept in
// registers as the next node instead of being pushed and popped again, which is the same
// visiting order.  Returns false when the pop was culled.
template <bool kFast, bool kDiag>
__device__ __forceinline__ bool trav_node(const DevScene& S, Trav& t, Stack& st) {
    uint32_t word;
    float nd;
    if (t.has_next) {
        word = t.nxt_word;
        nd = t.nxt_dist;
        t.has_next = false;
    } else {
        const uint2 e = st.pop();
        word = e.x;
        nd = __uint_as_float(e.y);
    }
    if (kDiag) t.btests++;
    if (nd < 0.0f) return false;
    if (nd > t.ldist + kMinDelta) return false;
    if (word & kLeafBit) {
        t.leaf = word & ~kLeafBit;
        return true;
    }
    // the children's boxes and words in one 64-byte record
    const float4* q = S.pairs + 4ull * (t.pair_base + word);
    const float4 a0 = q[0], a1 = q[1], b0 = q[2], b1 = q[3];
    float da, db;
    if (kFast) {
        da = slab_fast(t.lo, t.linv, xyz(a0), xyz(a1));
        db = slab_fast(t.lo, t.linv, xyz(b0), xyz(b1));
    } else {
        da = slab(t.lo, t.linv, xyz(a0), xyz(a1));
        db = slab(t.lo, t.linv, xyz(b0), xyz(b1));
    }
    const uint32_t wa = __float_as_uint(a0.w), wb = __float_as_uint(a1.w);
    const bool pa = (da >= 0.0f && da <= t.ldist + kMinDelta);
    const bool pb = (db >= 0.0f && db <= t.ldist + kMinDelta);
    if (pa && pb) {
        // the reference pushes the farther child first and pops the nearer one next
        const bool a_first = da < db;
        st.push(a_first ? wb : wa, a_first ? db : da);
        t.nxt_word = a_first ? wa : wb;
        t.nxt_dist = a_first ? da : db;
        t.has_next = true;
    } else if (pa || pb) {
        t.nxt_word = pa ? wa : wb;
        t.nxt_dist = pa ? da : db;
        t.has_next = true;
    }
    return true;
}

// One triangle test of the open leaf (Intersect.cpp:84-105 / 24-44).  kDiag keeps the per-hit
// diagnostics (tests of the winning instance, leaf index) of the counting and test builds.
template <bool kDiag>
__device__ __forceinline__ void trav_tri(const DevScene& S, Trav& t) {
    const float4* P = S.prims + 3ull * (t.prim_base + t.leaf);
    const float4 p0 = P[0], p1 = P[1], p2 = P[2];
    float u, v, tt;
    if (kDiag) t.ttests++;
    if (tri_test(t.lo, t.ld, xyz(p0), xyz(p1), xyz(p2), u, v, tt)) {
        if (tt > 0.0f && tt < t.ldist) {
            t.ldist = tt;
            t.bu = u;
            t.bvv = v;
            if (kDiag) t.bbv = __float_as_uint(p1.w);
            t.btri = __float_as_uint(p0.w);
        }
    }
    t.leaf = (__float_as_uint(p2.w) & kPrimLast) ? kNoLeaf : t.leaf + 1;
}

// ---- fast traversal (MT_TRAVERSAL_FAST) ----------------------------------------------------
// Per-mesh SAH BVHs with tight boxes (DevScene::fast_meshes).  The local ray keeps the length
// of the world direction (no re-normalisation), so one parameter t measures the hit in every
// instance and in world space; nodes, top-level entries and instances are culled when their
// entry distance is not below the closest hit so far (t.R.dist while entering, t.ldist inside
// an instance).  Shadow rays (kAny) start with the light distance as the bound and stop at the
// first occluder.

// entry distance of the ray into the box, clamped at 0; +inf when the box is missed or lies
// beyond tmax
__device__ __forceinline__ float slab_enter(V3 o, V3 inv, V3 mn, V3 mx, float tmax) {
    const V3 t1 = mul(sub(mn, o), inv);
    const V3 t2 = mul(sub(mx, o), inv);
    const float tn = fmaxf(fmaxf(fminf(t1.x, t2.x), fminf(t1.y, t2.y)), fmaxf(fminf(t1.z, t2.z), 0.0f));
    const float tf = fminf(fminf(fmaxf(t1.x, t2.x), fmaxf(t1.y, t2.y)), fminf(fmaxf(t1.z, t2.z), tmax));
    return tn <= tf ? tn : __builtin_huge_valf();
}

__device__ __forceinline__ void ftrav_tlas_begin(const DevScene& S, Trav& t, Stack& st) {
    t.tl = true;
    t.mask = 0;
    t.cur = -1;
    t.leaf = kNoLeaf;
    t.has_next = false;
    st.sp = 0;
    const V3 mn = {S.tlas_root_min[0], S.tlas_root_min[1], S.tlas_root_min[2]};
    const V3 mx = {S.tlas_root_max[0], S.tlas_root_max[1], S.tlas_root_max[2]};
    const float te = slab_enter(t.o, t.invd, mn, mx, t.R.dist);
    if (te < t.R.dist) st.push(S.tlas_root_word, te);
    t.base = st.sp;
}

// Set-up step of the fast traversal: finish the current instance, then enter the next one
// whose world box the ray enters before the closest hit so far.  Returns true when the ray is
// finished (for kAny also as soon as an occluder is found).
template <bool kCount, bool kAny, bool kDiag>
__device__ __forceinline__ bool ftrav_setup(const DevScene& S, Trav& t, Stack& st, Work& w, bool& need_mask) {
    if (t.cur >= 0) {
        if (kCount) {
            w.pops += t.btests;
            w.tris += t.ttests;
        }
        if (t.ldist < t.R.dist) {
            const uint32_t ii = trav_instance(S, t, t.cur);
            const V3 lp = add(t.lo, scale(t.ld, t.ldist));
            t.R.dist = t.ldist;
            t.R.inst = static_cast<int>(ii);
            t.R.tri = t.btri;
            t.R.u = t.bu;
            t.R.v = t.bvv;
            t.R.wpos = xform(S.instances[ii].M, lp, 1.0f);
            if (kDiag) {
                t.R.bv = t.bbv;
                t.R.bvh_tests = t.btests;
                t.R.tri_tests = t.ttests;
            }
            if (kAny) return true;
        }
        t.cur = -1;
    }
    const DevInstance* I;
    while (true) {
        if (t.mask == 0) {
            if (!t.tl) {
                if (t.cb + 64u < S.instance_count) {
                    t.cb += 64u;
                    need_mask = true;
                    return false;
                }
                return true;
            }
            // top level, nearest entry first
            while (true) {
                if (st.sp == 0) return true;
                const uint2 e = st.pop();
                if (__uint_as_float(e.y) >= t.R.dist) continue;
                const uint32_t word = e.x;
                if (word & kLeafBit) {
                    const uint32_t first = word & 0x00FFFFFFu, cnt = (word >> 24) & 0x7Fu;
                    uint64_t m = 0;
                    for (uint32_t j = 0; j < cnt; ++j) {
                        const DevInstance* J = S.instances + S.tlas_inst[first + j];
                        const V3 mn = {J->wmin[0], J->wmin[1], J->wmin[2]};
                        const V3 mx = {J->wmax[0], J->wmax[1], J->wmax[2]};
                        if (slab_enter(t.o, t.invd, mn, mx, t.R.dist) < t.R.dist) m |= (1ull << j);
                    }
                    if (kCount) w.aabb += cnt;
                    if (m) {
                        t.cb = first;
                        t.mask = m;
                        break;
                    }
                } else {
                    const float4* q = S.tlas_pairs + 4ull * word;
                    const float4 a0 = q[0], a1 = q[1], b0 = q[2], b1 = q[3];
                    const float da = slab_enter(t.o, t.invd, xyz(a0), xyz(a1), t.R.dist);
                    const float db = slab_enter(t.o, t.invd, xyz(b0), xyz(b1), t.R.dist);
                    const uint32_t wa = __float_as_uint(a0.w), wb = __float_as_uint(a1.w);
                    const bool pa = da < t.R.dist, pb = db < t.R.dist;
                    if (pa && pb) {  // the nearer child is popped first
                        const bool a_first = da <= db;
                        st.push(a_first ? wb : wa, a_first ? db : da);
                        st.push(a_first ? wa : wb, a_first ? da : db);
                    } else if (pa) {
                        st.push(wa, da);
                    } else if (pb) {
                        st.push(wb, db);
                    }
                }
            }
        }
        t.cur = __builtin_ctzll(t.mask);
        t.mask &= t.mask - 1;
        I = S.instances + trav_instance(S, t, t.cur);
        // the closest hit may have moved nearer since the box was accepted
        const V3 mn = {I->wmin[0], I->wmin[1], I->wmin[2]};
        const V3 mx = {I->wmax[0], I->wmax[1], I->wmax[2]};
        if (slab_enter(t.o, t.invd, mn, mx, t.R.dist) < t.R.dist) break;
        t.cur = -1;
    }
    t.base = st.sp;
    float inv[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) inv[k] = I->inv[k];
    t.lo = xform(inv, t.o, 1.0f);
    t.ld = xform(inv, t.d, 0.0f);
    t.linv = {1.0f / t.ld.x, 1.0f / t.ld.y, 1.0f / t.ld.z};
    t.ldist = t.R.dist;
    t.bu = t.bvv = 0.0f;
    t.btri = t.bbv = t.btests = t.ttests = 0;
    const DevMesh& m = S.fast_meshes[I->mesh];
    t.pair_base = m.pair_base;
    t.prim_base = m.prim_base;
    if (kCount) w.enter++;
    if (m.flags & kMeshHasBvh) {
        const V3 rmn = {m.root_min[0], m.root_min[1], m.root_min[2]};
        const V3 rmx = {m.root_max[0], m.root_max[1], m.root_max[2]};
        const float te = slab_enter(t.lo, t.linv, rmn, rmx, t.ldist);
        if (te < t.ldist) {
            t.nxt_word = m.root_word;
            t.nxt_dist = te;
            t.has_next = true;
        }
    } else if (m.prim_count) {
        t.leaf = 0;
    }
    return false;
}

// One node pop of the fast traversal; returns false when the pop was culled.
template <bool kDiag>
__device__ __forceinline__ bool ftrav_node(const DevScene& S, Trav& t, Stack& st) {
    uint32_t word;
    float nd;
    if (t.has_next) {
        word = t.nxt_word;
        nd = t.nxt_dist;
        t.has_next = false;
    } else {
        const uint2 e = st.pop();
        word = e.x;
        nd = __uint_as_float(e.y);
    }
    if (kDiag) t.btests++;
    if (nd >= t.ldist) return false;
    if (word & kLeafBit) {
        t.leaf = word & ~kLeafBit;
        return true;
    }
    const float4* q = S.pairs + 4ull * (t.pair_base + word);
    const float4 a0 = q[0], a1 = q[1], b0 = q[2], b1 = q[3];
    const float da = slab_enter(t.lo, t.linv, xyz(a0), xyz(a1), t.ldist);
    const float db = slab_enter(t.lo, t.linv, xyz(b0), xyz(b1), t.ldist);
    const uint32_t wa = __float_as_uint(a0.w), wb = __float_as_uint(a1.w);
    const bool pa = da < t.ldist, pb = db < t.ldist;
    if (pa && pb) {
        const bool a_first = da <= db;
        st.push(a_first ? wb : wa, a_first ? db : da);
        t.nxt_word = a_first ? wa : wb;
        t.nxt_dist = a_first ? da : db;
        t.has_next = true;
    } else if (pa || pb) {
        t.nxt_word = pa ? wa : wb;
        t.nxt_dist = pa ? da : db;
        t.has_next = true;
    }
    return true;
}

// One triangle test of the fast traversal; with kAny the first hit closes the instance.
template <bool kAny, bool kDiag>
__device__ __forceinline__ void ftrav_tri(const DevScene& S, Trav& t, Stack& st) {
    const float4* P = S.prims + 3ull * (t.prim_base + t.leaf);
    const float4 p0 = P[0], p1 = P[1], p2 = P[2];
    float u, v, tt;
    if (kDiag) t.ttests++;
    t.leaf = (__float_as_uint(p2.w) & kPrimLast) ? kNoLeaf : t.leaf + 1;
    if (tri_test(t.lo, t.ld, xyz(p0), xyz(p1), xyz(p2), u, v, tt)) {
        if (tt > 0.0f && tt < t.ldist) {
            t.ldist = tt;
            t.bu = u;
            t.bvv = v;
            if (kDiag) t.bbv = __float_as_uint(p1.w);
            t.btri = __float_as_uint(p0.w);
            if (kAny) {
                t.leaf = kNoLeaf;
                t.has_next = false;
                st.sp = t.base;
            }
        }
    }
}

}  // namespace dev
}  // namespace mt
