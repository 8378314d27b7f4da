// oracle.cpp -- CPU oracle for the MiTrace Tracer path.  TEST INFRASTRUCTURE ONLY.
//
// A plain C++ restatement of the reference's CPU path, used by tests/ as the checker of the
// GPU library and by bench.py as the CPU baseline.  It is never linked into the product.
//
// Arithmetic contract (shared with the GPU kernels, written independently there):
//   * float32 everywhere, no FMA contraction (-ffp-contract=off), IEEE division and sqrt;
//   * dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z;  normalize(v) = v * (1/sqrt(dot(v,v)));
//   * min(a,b) = (b < a) ? b : a,  max(a,b) = (a < b) ? b : a  (GLM's scalar forms);
//   * mat4*vec4 row r = (m[0][r]*x + m[1][r]*y) + (m[2][r]*z + m[3][r]*w);
//   * 4x4 inverse by the 2x2-minor (Laplace expansion) scheme of invert4() below.
// GLM 1.0.3 (CMakeLists.txt:57-63) is not available here, so these orders are this build's
// definition; the reference's exact bits under -Ofast are not reproducible (parity unpinned
// against real GLM, see the test-suite notes).

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

constexpr float kPi = 3.14159265358979323846f;

// ---------------------------------------------------------------- vector math (GLM contract)
struct V3 {
    float x, y, z;
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
struct V4 {
    float x, y, z, w;
};
inline V3 v3(float a) { return {a, a, a}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float gmin(float a, float b) { return (b < a) ? b : a; }
inline float gmax(float a, float b) { return (a < b) ? b : a; }
inline V3 vmin(V3 a, V3 b) { return {gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)}; }
inline float compMin(V3 v) { return gmin(gmin(v.x, v.y), v.z); }
inline float compMax(V3 v) { return gmax(gmax(v.x, v.y), v.z); }
inline float clampf(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
inline float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
inline float length(V3 v) { return std::sqrt(dot(v, v)); }
inline V3 normalize(V3 v) { return v * (1.0f / std::sqrt(dot(v, v))); }
inline float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
inline float get(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

struct M4 {
    float m[16];  // column-major: m[c*4 + r]
    float at(int c, int r) const { return m[c * 4 + r]; }
};
inline V4 mul(const M4& M, V4 v) {
    V4 o;
    float* po = &o.x;
    for (int r = 0; r < 4; ++r)
        po[r] = (M.at(0, r) * v.x + M.at(1, r) * v.y) + (M.at(2, r) * v.z + M.at(3, r) * v.w);
    return o;
}
inline V3 xfPoint(const M4& M, V3 p) {
    V4 o = mul(M, {p.x, p.y, p.z, 1.0f});
    return {o.x, o.y, o.z};
}
inline V3 xfDir(const M4& M, V3 d) {
    V4 o = mul(M, {d.x, d.y, d.z, 0.0f});
    return {o.x, o.y, o.z};
}

// 4x4 inverse (stands in for glm::inverse, Intersect.cpp:152).  a[r][c] is row r, column c.
M4 invert4(const M4& M) {
    float a[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) a[r][c] = M.at(c, r);
    const float s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    const float s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
    const float s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3];
    const float s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    const float s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3];
    const float s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
    const float c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3];
    const float c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
    const float c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2];
    const float c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
    const float c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2];
    const float c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const float id = 1.0f / det;
    float b[4][4];
    b[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id;
    b[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id;
    b[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id;
    b[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id;
    b[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id;
    b[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id;
    b[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id;
    b[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id;
    b[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id;
    b[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id;
    b[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id;
    b[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id;
    b[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * id;
    b[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * id;
    b[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * id;
    b[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * id;
    M4 out;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out.m[c * 4 + r] = b[r][c];
    return out;
}

// ---------------------------------------------------------------- elementary functions
// The path's sin, cos, atan2, acos and x^5, written out operation by operation so that the
// oracle and the GPU kernels (mt_device.hpp, the same operations in the same order) return the
// same bits; the C library's and the device library's versions differ in the last bits, which
// made a few per cent of the samples differ.  Rounding is IEEE float32 throughout (no
// contraction).  sin/cos: the argument reduction of Cody & Waite (1980), x - k pi/2 with pi/2
// split into three floats whose first two have short mantissas (k (pi/2)_1 and k (pi/2)_2 are
// exact for |k| < 2^12), then the Taylor polynomials of sin (through r^9) and cos (through
// r^10) on [-pi/4, pi/4], where the first omitted terms are below 2e-9.  atan2: octant
// symmetry folds the argument into t in [0, 1], the identity atan(t) = pi/4 + atan((t-1)/(t+1))
// into |t| <= tan(pi/8), then the Taylor series of atan through t^19 (next term < 5e-10).
// acos(x) = atan2(sqrt((1-x)(1+x)), x).  x^5 = (x^2 x^2) x.
namespace em {
constexpr float kTwoOverPi = 0x1.45f306p-1f;
constexpr float kPio2a = 0x1.92p+0f, kPio2b = 0x1.fb6p-12f, kPio2c = -0x1.777a5cp-25f;
constexpr float kPi = 0x1.921fb6p+1f, kPio2 = 0x1.921fb6p+0f, kPio4 = 0x1.921fb6p-1f;
constexpr float kTanPi8 = 0x1.a8279ap-2f;
inline float sinPoly(float r) {
    const float r2 = r * r;
    float p = 0x1.71de3ap-19f;           //  1/9!
    p = p * r2 + -0x1.a01a02p-13f;        // -1/7!
    p = p * r2 + 0x1.111112p-7f;          //  1/5!
    p = p * r2 + -0x1.555556p-3f;         // -1/3!
    return r + (r * r2) * p;
}
inline float cosPoly(float r) {
    const float r2 = r * r;
    float p = -0x1.27e4fcp-22f;           // -1/10!
    p = p * r2 + 0x1.a01a02p-16f;         //  1/8!
    p = p * r2 + -0x1.6c16c2p-10f;        // -1/6!
    p = p * r2 + 0x1.555556p-5f;          //  1/4!
    p = p * r2 + -0.5f;                   // -1/2!
    return 1.0f + r2 * p;
}
inline void sincos(float x, float& s, float& c) {
    const float k = std::floor(x * kTwoOverPi + 0.5f);
    const float r = ((x - k * kPio2a) - k * kPio2b) - k * kPio2c;
    const float ps = sinPoly(r), pc = cosPoly(r);
    const int q = static_cast<int>(k) & 3;
    s = q == 0 ? ps : (q == 1 ? pc : (q == 2 ? -ps : -pc));
    c = q == 0 ? pc : (q == 1 ? -ps : (q == 2 ? -pc : ps));
}
inline float sin(float x) {
    float s, c;
    sincos(x, s, c);
    return s;
}
inline float cos(float x) {
    float s, c;
    sincos(x, s, c);
    return c;
}
inline float atanPoly(float t) {  // |t| <= tan(pi/8)
    const float t2 = t * t;
    float p = -0x1.af286cp-5f;            // -1/19
    p = p * t2 + 0x1.e1e1e2p-5f;          //  1/17
    p = p * t2 + -0x1.111112p-4f;         // -1/15
    p = p * t2 + 0x1.3b13b2p-4f;          //  1/13
    p = p * t2 + -0x1.745d18p-4f;         // -1/11
    p = p * t2 + 0x1.c71c72p-4f;          //  1/9
    p = p * t2 + -0x1.24924ap-3f;         // -1/7
    p = p * t2 + 0x1.99999ap-3f;          //  1/5
    p = p * t2 + -0x1.555556p-2f;         // -1/3
    return t + (t * t2) * p;
}
inline float atan2(float y, float x) {
    const float ax = std::fabs(x), ay = std::fabs(y);
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
        a = off + atanPoly(t);
    }
    if (ax < ay) a = kPio2 - a;
    if (std::signbit(x)) a = kPi - a;
    return std::signbit(y) ? -a : a;
}
inline float acos(float x) { return atan2(std::sqrt((1.0f - x) * (1.0f + x)), x); }
inline float pow5(float x) {
    const float x2 = x * x;
    return (x2 * x2) * x;
}
}  // namespace em

// ---------------------------------------------------------------- planted mistakes
// A build with -DORC_MUTANT=k (oracle/Makefile, target `mutants`) changes one formula, so that
// the test-suite can show that each of its pins fails on a plausible mistake.  Every normal
// build has ORC_MUTANT 0, where every MUT(k) is a compile-time false.
#ifndef ORC_MUTANT
#define ORC_MUTANT 0
#endif
#define MUT(k) (ORC_MUTANT == (k))

// ---------------------------------------------------------------- counter RNG
// Keyed integer hash; the GPU kernels implement the same function.
inline uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}
inline uint32_t seedKey(uint64_t seed) {
    const uint32_t k0 = mix32(static_cast<uint32_t>(seed) ^ 0xA511E9B3u);
    return mix32(k0 ^ static_cast<uint32_t>(seed >> 32) ^ 0x63D83595u);
}
inline uint32_t pathKey(uint32_t sk, uint32_t pixel, uint32_t sample) {
    return mix32(mix32(sk ^ pixel) + sample * 0x9E3779B9u);
}
inline float drawU(uint32_t key, uint32_t dim) {
    const uint32_t x = mix32(key ^ mix32(dim + 0x7F4A7C15u));
    return static_cast<float>(x >> 8) * (1.0f / 16777216.0f);
}
// dimension layout: 0,1 camera jitter; bounce k uses base = 2 + k*(8 + lightCount)
enum : uint32_t { kDimLobe = 0, kDimDir = 1, kDimFallback = 3, kDimLightPt = 5, kDimRR = 7,
                  kDimLightSel = 8 };

struct Rng {
    // counter mode
    uint32_t key = 0;
    uint32_t base = 0;
    // mt mode
    std::mt19937* mt = nullptr;
    std::uniform_real_distribution<float> unit{0.0f, 1.0f};
    // given uniforms (test hooks): dimension d of the current base reads fixed[d]
    const float* fixed = nullptr;
    float u(uint32_t dimOffset) {
        if (fixed) return fixed[dimOffset];
        if (mt) return unit(*mt);
        return drawU(key, base + dimOffset);
    }
    float uRange(uint32_t dimOffset, float a, float b) {
        if (mt) {
            std::uniform_real_distribution<float> d(a, b);
            return d(*mt);
        }
        return drawU(key, base + dimOffset) * (b - a) + a;
    }
};

// ---------------------------------------------------------------- scene
struct Node {  // BVHNode, BVH.hpp:74-119 (32 bytes)
    V3 mn, mx;
    uint32_t a, b;
};
constexpr uint32_t kLeafBit = 0x80000000u;

struct Mesh {
    std::vector<V3> pos, nrm;
    std::vector<float> uv;  // 2 per corner, may be empty
    bool hasBvh = false;
    std::vector<Node> nodes;
    std::vector<uint32_t> indices;
    V3 aabbMin, aabbMax;
};
struct Tex {
    int image = -1;
    int filter = 0;
    float ox = 0, oy = 0, sx = 1, sy = 1;
};
struct Material {
    float bcf[4];
    V3 emissive;
    float normalScale, metallic, roughness;
    Tex baseTex, mrTex, nrmTex, emTex;
};
struct Image {
    int w = 0, h = 0;
    std::vector<float> rgb;  // decoded texels (u8/255 or float), 3 per texel
};
struct Instance {
    uint32_t mesh, material;
    M4 M, inv;
    V3 wmin, wmax;
};
struct Light {
    int type;
    V3 pos, color, dir;
    float size, inner, outer;
    float intensity;  // PrepareLights simplified intensity
};
// Median-split tree over the instances' world boxes (leaves of at most 4).  It only skips
// instances whose box the ray misses: the entered set, its order and every result are those of
// the reference's linear loop (Intersect.cpp:201-219).  Used when orc_render_opts.tlas is set,
// to time the CPU tracer without its 5,000-instance loop.
struct TlasNode {
    V3 mn, mx;
    int a = -1, b = -1;
    uint32_t first = 0, count = 0;
};
struct Scene {
    std::vector<Mesh> meshes;
    std::vector<Image> images;
    std::vector<Material> materials;
    std::vector<Instance> instances;
    std::vector<Light> lights;
    Image env;
    bool hasEnv = false;
    int envFilter = 1;
    std::vector<TlasNode> tlas;
    std::vector<uint32_t> tlasIds;
};

// ---------------------------------------------------------------- BVH build (BVH.cpp)
struct TreeNode {
    std::vector<uint32_t> idx;
    V3 mn, mx;
    bool leaf = true;
    std::unique_ptr<TreeNode> ca, cb;
};

void computeAABB(const std::vector<V3>& P, TreeNode& n) {  // BVH.cpp:29-50
    V3 mn = v3(FLT_MAX), mx = v3(-FLT_MAX);
    for (uint32_t i : n.idx) {
        const V3 &a = P[i * 3], &b = P[i * 3 + 1], &c = P[i * 3 + 2];
        mn = vmin(mn, vmin(a, vmin(b, c)));
        mx = vmax(mx, vmax(a, vmax(b, c)));
    }
    n.mn = mn;
    n.mx = mx;
}

void subdivide(const std::vector<V3>& P, TreeNode& n, unsigned maxTri, unsigned maxDepth) {
    // BVH.cpp:117-195
    if (n.idx.size() <= static_cast<size_t>(maxTri)) return;
    V3 sum = v3(0.0f);  // FindCenterOfMass, BVH.cpp:52-63
    for (uint32_t i : n.idx) sum = sum + ((P[i * 3] + P[i * 3 + 1]) + P[i * 3 + 2]);
    const V3 com = sum / static_cast<float>(n.idx.size() * 3);
    std::vector<uint32_t> L[3], R[3];
    float score[3];
    for (int ax = 0; ax < 3; ++ax) {  // SplitTriangles, BVH.cpp:65-102
        const float sp = get(com, ax);
        for (uint32_t i : n.idx) {
            const V3 &a = P[i * 3], &b = P[i * 3 + 1], &c = P[i * 3 + 2];
            const V3 lo = vmin(a, vmin(b, c)), hi = vmax(a, vmax(b, c));
            if (get(lo, ax) < sp) L[ax].push_back(i);
            if (get(hi, ax) >= sp) R[ax].push_back(i);
        }
        // ComputeSplitScore, BVH.cpp:104-115
        if (L[ax].size() == n.idx.size() || R[ax].size() == n.idx.size())
            score[ax] = FLT_MAX;
        else
            score[ax] = static_cast<float>(L[ax].size() + R[ax].size()) /
                        static_cast<float>(n.idx.size());
    }
    if (score[0] == FLT_MAX && score[1] == FLT_MAX && score[2] == FLT_MAX) return;
    int ax;
    if (score[0] < score[1] && score[0] < score[2])
        ax = 0;
    else if (score[1] < score[2])
        ax = 1;
    else
        ax = 2;
    n.ca = std::make_unique<TreeNode>();
    n.cb = std::make_unique<TreeNode>();
    n.ca->idx = std::move(L[ax]);
    n.cb->idx = std::move(R[ax]);
    n.ca->mn = n.mn;
    n.ca->mx = n.mx;
    n.cb->mn = n.mn;
    n.cb->mx = n.mx;
    const float sp = get(com, ax);
    (ax == 0 ? n.ca->mx.x : ax == 1 ? n.ca->mx.y : n.ca->mx.z) = sp;
    (ax == 0 ? n.cb->mn.x : ax == 1 ? n.cb->mn.y : n.cb->mn.z) = sp;
    n.leaf = false;
    if (maxDepth > 0) {
        subdivide(P, *n.ca, maxTri, maxDepth - 1);
        subdivide(P, *n.cb, maxTri, maxDepth - 1);
    }
}

void appendLeaf(Mesh& m, const TreeNode& t, Node& out) {  // BVH.cpp:214-221
    const uint32_t index = static_cast<uint32_t>(m.indices.size());
    m.indices.insert(m.indices.end(), t.idx.begin(), t.idx.end());
    out.a = index | kLeafBit;
    out.b = static_cast<uint32_t>(t.idx.size());
}

std::pair<uint32_t, uint32_t> flattenPair(Mesh& m, const TreeNode& A, const TreeNode& B) {
    // BVH.cpp:223-259
    const uint32_t ia = static_cast<uint32_t>(m.nodes.size()), ib = ia + 1;
    m.nodes.push_back(Node{});
    m.nodes.push_back(Node{});
    m.nodes[ia].mn = A.mn;
    m.nodes[ia].mx = A.mx;
    m.nodes[ib].mn = B.mn;
    m.nodes[ib].mx = B.mx;
    if (A.leaf) {
        appendLeaf(m, A, m.nodes[ia]);
    } else {
        auto c = flattenPair(m, *A.ca, *A.cb);
        m.nodes[ia].a = c.first;
        m.nodes[ia].b = c.second;
    }
    if (B.leaf) {
        appendLeaf(m, B, m.nodes[ib]);
    } else {
        auto c = flattenPair(m, *B.ca, *B.cb);
        m.nodes[ib].a = c.first;
        m.nodes[ib].b = c.second;
    }
    return {ia, ib};
}

void buildBvh(Mesh& m, unsigned maxTri, unsigned maxDepth) {
    // BVH::BVH (BVH.cpp:274-282) -> BVHTree (204-212) -> Flatten (261-272)
    TreeNode root;
    const uint32_t T = static_cast<uint32_t>(m.pos.size() / 3);
    root.idx.resize(T);
    for (uint32_t i = 0; i < T; ++i) root.idx[i] = i;
    computeAABB(m.pos, root);
    subdivide(m.pos, root, maxTri, maxDepth - 1);
    m.nodes.clear();
    m.indices.clear();
    m.nodes.push_back(Node{});
    m.nodes.push_back(Node{});  // slot 1 stays zero (value-initialised BVHNode)
    m.nodes[0].mn = root.mn;
    m.nodes[0].mx = root.mx;
    if (root.leaf) {
        // the reference dereferences null children here; this build keeps one leaf root
        appendLeaf(m, root, m.nodes[0]);
    } else {
        auto c = flattenPair(m, *root.ca, *root.cb);
        m.nodes[0].a = c.first;
        m.nodes[0].b = c.second;
    }
    m.hasBvh = true;
}

// ---------------------------------------------------------------- intersection (Intersect.cpp)
struct Ray {
    V3 o, d;
};

inline float slab(const Ray& r, V3 mn, V3 mx) {  // Intersect.cpp:9-18
    const V3 inv = V3{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    const V3 t1 = (mn - r.o) * inv;
    const V3 t2 = (mx - r.o) * inv;
    const V3 tmin3 = vmin(t1, t2), tmax3 = vmax(t1, t2);
    return compMax(tmin3) <= compMin(tmax3) ? compMin(tmax3) : -1.0f;
}

// Moller & Trumbore (1997), two-sided, |det| > FLT_EPSILON, bounds tested against det before
// scaling by 1/det, no t-range test (the caller applies it) -- glm::intersectRayTriangle's
// contract as used at Intersect.cpp:34 and :94.
inline bool triangleTest(V3 o, V3 d, V3 v0, V3 v1, V3 v2, float& u, float& v, float& t) {
    const V3 e1 = v1 - v0;
    const V3 e2 = v2 - v0;
    const V3 p = cross(d, e2);
    const float det = dot(e1, p);
    V3 q;
    if (det > FLT_EPSILON) {
        const V3 s = o - v0;
        u = dot(s, p);
        if (u < 0.0f || u > det) return false;
        q = cross(s, e1);
        v = dot(d, q);
        if (v < 0.0f || u + v > det) return false;
    } else if (det < -FLT_EPSILON) {
        const V3 s = o - v0;
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

struct Counters {
    uint64_t aabb = 0, enter = 0, pops = 0, tris = 0;
};

void traverseLinear(const Mesh& m, const Ray& lr, float& dist, float& bu, float& bv, uint32_t& tri,
                    uint32_t& triTests) {  // Intersect.cpp:20-45
    const uint32_t T = static_cast<uint32_t>(m.pos.size() / 3);
    for (uint32_t i = 0; i < T; ++i) {
        float u, v, t;
        triTests++;
        if (triangleTest(lr.o, lr.d, m.pos[i * 3], m.pos[i * 3 + 1], m.pos[i * 3 + 2], u, v, t)) {
            if (t > 0.0f && t < dist) {
                dist = t;
                bu = u;
                bv = v;
                tri = i;
            }
        }
    }
}

void traverseBvh(const Mesh& m, const Ray& lr, float& dist, float& bu, float& bv, uint32_t& bvIdx,
                 uint32_t& tri, uint32_t& bvhTests, uint32_t& triTests) {  // Intersect.cpp:47-141
    struct Entry {
        uint32_t node;
        float dist;
    };
    thread_local std::vector<Entry> stack;
    const float minDelta = 0.1f;
    stack.clear();
    stack.push_back({0, slab(lr, m.nodes[0].mn, m.nodes[0].mx)});
    while (!stack.empty()) {
        const Entry e = stack.back();
        stack.pop_back();
        bvhTests++;
        const Node& n = m.nodes[e.node];
        if (e.dist < 0.0f) continue;
        if (e.dist > dist + minDelta) continue;
        if (n.a & kLeafBit) {
            const uint32_t first = n.a & ~kLeafBit, count = n.b & ~kLeafBit;
            for (uint32_t i = first; i < first + count; ++i) {
                const uint32_t t3 = m.indices[i];
                float u, v, t;
                triTests++;
                if (triangleTest(lr.o, lr.d, m.pos[t3 * 3], m.pos[t3 * 3 + 1], m.pos[t3 * 3 + 2], u,
                                 v, t)) {
                    if (t > 0.0f && t < dist) {
                        dist = t;
                        bu = u;
                        bv = v;
                        bvIdx = e.node;
                        tri = t3;
                    }
                }
            }
        } else {
            const Node& A = m.nodes[n.a];
            const Node& B = m.nodes[n.b];
            const float da = slab(lr, A.mn, A.mx);
            const float db = slab(lr, B.mn, B.mx);
            const bool pa = (da >= 0.0f && da <= dist + minDelta);
            const bool pb = (db >= 0.0f && db <= dist + minDelta);
            if (pa && pb) {
                if (da < db) {
                    stack.push_back({n.b, db});
                    stack.push_back({n.a, da});
                } else {
                    stack.push_back({n.a, da});
                    stack.push_back({n.b, db});
                }
            } else if (pa) {
                stack.push_back({n.a, da});
            } else if (pb) {
                stack.push_back({n.b, db});
            }
        }
    }
}

struct Hit {
    bool valid = false;
    int inst = -1;
    uint32_t tri = 0, bv = 0, bvhTests = 0, triTests = 0;
    float u = 0, v = 0, dist = 0;
    V3 wpos{0, 0, 0};
    V3 o{0, 0, 0}, d{0, 0, 0};
};

bool intersectInstance(const Scene& S, int ii, const Ray& r, bool faithfulInverse, Hit& h,
                       Counters* c) {  // Intersect.cpp:143-199
    const Instance& I = S.instances[ii];
    if (c) c->aabb++;
    if (slab(r, I.wmin, I.wmax) < 0.0f) return false;
    if (c) c->enter++;
    const M4 inv = faithfulInverse ? invert4(I.M) : I.inv;
    Ray lr;
    lr.o = xfPoint(inv, r.o);
    lr.d = normalize(xfDir(inv, r.d));
    float dist = FLT_MAX, bu = 0, bv = 0;
    uint32_t bvIdx = 0, tri = 0, bvhTests = 0, triTests = 0;
    const Mesh& m = S.meshes[I.mesh];
    if (m.hasBvh)
        traverseBvh(m, lr, dist, bu, bv, bvIdx, tri, bvhTests, triTests);
    else
        traverseLinear(m, lr, dist, bu, bv, tri, triTests);
    if (c) {
        c->pops += bvhTests;
        c->tris += triTests;
    }
    if (dist < FLT_MAX) {
        const V3 lp = lr.o + dist * lr.d;
        h.wpos = xfPoint(I.M, lp);
        h.dist = length(h.wpos - r.o);
        h.tri = tri;
        h.bv = bvIdx;
        h.u = bu;
        h.v = bv;
        h.bvhTests = bvhTests;
        h.triTests = triTests;
        h.o = r.o;
        h.d = r.d;
        h.inst = ii;
        h.valid = true;
        return true;
    }
    return false;
}

int buildTlasNode(Scene& S, uint32_t lo, uint32_t hi) {
    const int id = static_cast<int>(S.tlas.size());
    S.tlas.emplace_back();
    V3 mn = v3(FLT_MAX), mx = v3(-FLT_MAX), cmn = v3(FLT_MAX), cmx = v3(-FLT_MAX);
    for (uint32_t k = lo; k < hi; ++k) {
        const Instance& I = S.instances[S.tlasIds[k]];
        const V3 c = 0.5f * (I.wmin + I.wmax);
        for (int a = 0; a < 3; ++a) {
            mn[a] = std::min(mn[a], I.wmin[a]);
            mx[a] = std::max(mx[a], I.wmax[a]);
            cmn[a] = std::min(cmn[a], c[a]);
            cmx[a] = std::max(cmx[a], c[a]);
        }
    }
    S.tlas[id].mn = mn;
    S.tlas[id].mx = mx;
    if (hi - lo <= 4) {
        S.tlas[id].first = lo;
        S.tlas[id].count = hi - lo;
        return id;
    }
    int ax = 0;
    for (int a = 1; a < 3; ++a)
        if (cmx[a] - cmn[a] > cmx[ax] - cmn[ax]) ax = a;
    const uint32_t mid = lo + (hi - lo) / 2;
    std::nth_element(S.tlasIds.begin() + lo, S.tlasIds.begin() + mid, S.tlasIds.begin() + hi,
                     [&](uint32_t x, uint32_t y) {
                         const float cx = S.instances[x].wmin[ax] + S.instances[x].wmax[ax];
                         const float cy = S.instances[y].wmin[ax] + S.instances[y].wmax[ax];
                         return cx < cy;
                     });
    const int a = buildTlasNode(S, lo, mid);
    const int b = buildTlasNode(S, mid, hi);
    S.tlas[id].a = a;
    S.tlas[id].b = b;
    return id;
}

void buildTlas(Scene& S) {
    S.tlas.clear();
    S.tlasIds.resize(S.instances.size());
    for (uint32_t i = 0; i < S.tlasIds.size(); ++i) S.tlasIds[i] = i;
    if (!S.instances.empty()) buildTlasNode(S, 0, static_cast<uint32_t>(S.instances.size()));
}

Hit intersectScene(const Scene& S, const Ray& r, bool faithfulInverse, Counters* c,
                   bool useTlas = false) {
    // Intersect.cpp:201-219
    float lowest = FLT_MAX;
    Hit best;
    auto visit = [&](int i) {
        Hit h;
        if (intersectInstance(S, i, r, faithfulInverse, h, c)) {
            if (h.dist < lowest) {
                lowest = h.dist;
                best = h;
            }
        }
    };
    if (!useTlas || S.tlas.empty()) {
        for (int i = 0; i < static_cast<int>(S.instances.size()); ++i) visit(i);
        return best;
    }
    // the instances whose box the tree does not rule out, visited in index order
    thread_local std::vector<uint32_t> cand, stack;
    cand.clear();
    stack.assign(1, 0u);
    while (!stack.empty()) {
        const TlasNode& n = S.tlas[stack.back()];
        stack.pop_back();
        if (slab(r, n.mn, n.mx) < 0.0f) continue;
        if (n.a < 0) {
            for (uint32_t k = 0; k < n.count; ++k) cand.push_back(S.tlasIds[n.first + k]);
        } else {
            stack.push_back(static_cast<uint32_t>(n.a));
            stack.push_back(static_cast<uint32_t>(n.b));
        }
    }
    std::sort(cand.begin(), cand.end());
    for (uint32_t i : cand) visit(static_cast<int>(i));
    return best;
}

// ---------------------------------------------------------------- textures (Image.cpp)
struct RGBA {
    float r, g, b, a;
};
inline RGBA texel(const Image& im, int x, int y) {  // Image.cpp:11-15
    const size_t i = (static_cast<size_t>(y) * im.w + x) * 3;
    return {im.rgb[i], im.rgb[i + 1], im.rgb[i + 2], 1.0f};
}
inline RGBA mix4(RGBA a, RGBA b, float t) {
    return {mixf(a.r, b.r, t), mixf(a.g, b.g, t), mixf(a.b, b.b, t), mixf(a.a, b.a, t)};
}
RGBA sampleImage(const Image& im, float uvx, float uvy, int filter) {  // Image.cpp:46-61
    const float x = uvx - std::floor(uvx);
    const float y = uvy - std::floor(uvy);
    if (filter == MT_FILTER_NEAREST) {  // Image.cpp:17-24
        int ix = static_cast<int>(x * static_cast<float>(im.w) + (MUT(1) ? 0.5f : 0.0f)) % im.w;
        int iy = static_cast<int>(y * static_cast<float>(im.h) + (MUT(1) ? 0.5f : 0.0f)) % im.h;
        if (ix < 0) ix += im.w;
        if (iy < 0) iy += im.h;
        ix = std::max(ix, 0);  // a NaN coordinate converts to INT_MIN on x86
        iy = std::max(iy, 0);
        return texel(im, ix, iy);
    }
    // "linear" with weights in uv units (Image.cpp:26-44); x0/y0 are clamped to the last
    // texel where the reference would read past the row (x*W rounding up to W).
    int x0 = static_cast<int>(std::floor(x * static_cast<float>(im.w)));
    int y0 = static_cast<int>(std::floor(y * static_cast<float>(im.h)));
    if (MUT(4)) {
        x0 %= im.w;
        y0 %= im.h;
    }
    x0 = std::max(std::min(x0, im.w - 1), 0);
    y0 = std::max(std::min(y0, im.h - 1), 0);
    const int x1 = MUT(3) ? (x0 + 1) % im.w : std::min(x0 + 1, im.w - 1);
    const int y1 = MUT(3) ? (y0 + 1) % im.h : std::min(y0 + 1, im.h - 1);
    const float sx = MUT(2) ? x * static_cast<float>(im.w) - static_cast<float>(x0)
                            : x - static_cast<float>(x0) / static_cast<float>(im.w);
    const float sy = MUT(2) ? y * static_cast<float>(im.h) - static_cast<float>(y0)
                            : y - static_cast<float>(y0) / static_cast<float>(im.h);
    const RGBA c0 = mix4(texel(im, x0, y0), texel(im, x1, y0), sx);
    const RGBA c1 = mix4(texel(im, x0, y1), texel(im, x1, y1), sx);
    return mix4(c0, c1, sy);
}
RGBA sampleTex(const Scene& S, const Tex& t, float u, float v) {  // Texture.cpp:12-15,52-56
    const Image& im = S.images[t.image];
    return sampleImage(im, u * t.sx + t.ox, v * t.sy + t.oy, t.filter);
}
V3 sampleEnv(const Scene& S, V3 d, bool applyRot, float rotDeg) {  // Image.cpp:63-75
    if (!S.hasEnv) return v3(0.0f);
    const float theta = MUT(5) ? em::atan2(d.x, d.z) : em::atan2(d.z, d.x);
    const float phi = em::acos(MUT(6) ? d.y : d.y / length(d));
    float u = (theta + kPi) / (2.0f * kPi);
    const float v = phi / kPi;
    if (applyRot) u = u * 1.0f + (MUT(7) ? -rotDeg : rotDeg) / 360.0f;  // TransformUv offset that SampleEquirectangular bypasses
    const RGBA c = sampleImage(S.env, u, v, S.envFilter);
    return {c.r, c.g, c.b};
}

// ---------------------------------------------------------------- BRDF (BRDF.hpp)
inline V3 fresnel(float cosT, V3 f0) {  // BRDF.hpp:9-13
    cosT = clampf(cosT, 0.0f, 1.0f);
    const float p = em::pow5(1.0f - cosT);
    return {f0.x + (1.0f - f0.x) * p, f0.y + (1.0f - f0.y) * p, f0.z + (1.0f - f0.z) * p};
}
inline V3 baseRefl(V3 base, float metallic) {  // BRDF.hpp:16-20
    metallic = clampf(metallic, 0.0f, 1.0f);
    return {mixf(0.04f, base.x, metallic), mixf(0.04f, base.y, metallic),
            mixf(0.04f, base.z, metallic)};
}
inline float distGGX(float nh, float r) {  // BRDF.hpp:22-33
    nh = clampf(nh, 0.0f, 1.0f);
    const float a = r * r;
    const float a2 = a * a;
    const float nh2 = nh * nh;
    float den = nh2 * (a2 - 1.0f) + 1.0f;
    den = kPi * den * den;
    return a2 / den;
}
inline float geomSchlick(float nv, float r) {  // BRDF.hpp:35-43
    nv = clampf(nv, 0.0f, 1.0f);
    const float rr = r + 1.0f;
    const float k = (rr * rr) / 8.0f;
    const float den = nv * (1.0f - k) + k;
    return nv / den;
}
V3 evalBrdf(V3 L, V3 V, V3 N, V3 base, float rough, float metal) {  // BRDF.hpp:52-77
    rough = gmax(rough, 0.05f);
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return v3(0.0f);
    const V3 h = normalize(V + L);
    const float nh = gmax(dot(N, h), 0.0f);
    const float hv = gmax(dot(h, V), 0.0f);
    const V3 f0 = baseRefl(base, metal);
    const V3 f = fresnel(hv, f0);
    const float d = distGGX(nh, rough);
    const float g = geomSchlick(nv, rough) * geomSchlick(nl, rough);
    const V3 kd = (v3(1.0f) - f) * (1.0f - metal);
    const V3 diffuse = kd * base / kPi;
    const V3 spec = (d * g * f) / gmax(4.0f * nv * nl, 0.001f);
    return vmax(diffuse + spec, v3(0.0f));
}
inline float diffusePdf(V3 L, V3 N) { return gmax(dot(N, L), 0.0f) / kPi; }  // BRDF.hpp:79-82
float specularPdf(V3 L, V3 V, V3 N, float rough) {  // BRDF.hpp:84-97
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return 0.0f;
    const V3 h = normalize(V + L);
    const float nh = gmax(dot(N, h), 0.0f);
    const float vh = gmax(dot(V, h), 0.0f);
    if (vh <= 0.0f) return 0.0f;
    return distGGX(nh, gmax(rough, 0.05f)) * nh / (4.0f * vh);
}

// ---------------------------------------------------------------- sampling (Tracer.cpp:26-121)
inline void buildTBN(V3 n, V3& t, V3& b) {  // Tracer.cpp:26-32
    const V3 up = (std::fabs(n.z) < 0.999f) ? V3{0.0f, 0.0f, 1.0f} : V3{1.0f, 0.0f, 0.0f};
    t = normalize(cross(up, n));
    b = cross(n, t);
}
V3 cosineHemisphere(float u1, float u2, V3 n) {  // Tracer.cpp:34-47
    const float r = std::sqrt(u1);
    const float phi = 2.0f * kPi * u2;
    V3 t, b;
    buildTBN(n, t, b);
    float sp, cp;
    em::sincos(phi, sp, cp);
    const float lx = r * cp, ly = r * sp;
    const float lz = std::sqrt(gmax(0.0f, 1.0f - u1));
    return normalize(t * lx + b * ly + n * lz);
}
// SampleGGXDirection (Tracer.cpp:49-72).  Returns false when the reflection falls below the
// surface and the caller must fall back to cosine sampling with two more uniforms.
bool ggxDirection(float u1, float u2, V3 incident, V3 n, float rough, V3& out, float* hcos = nullptr) {
    const float alpha = gmax(rough * rough, 0.0025f);
    const float alpha2 = alpha * alpha;
    const float phi = 2.0f * kPi * u2;
    const float tan2 = alpha2 * u1 / gmax(1.0f - u1, 0.0001f);
    const float cosT = 1.0f / std::sqrt(1.0f + tan2);
    if (hcos) *hcos = cosT;  // n.h of the sampled half vector, before any flip
    const float sinT = std::sqrt(gmax(0.0f, 1.0f - cosT * cosT));
    V3 t, b;
    buildTBN(n, t, b);
    float sp, cp;
    em::sincos(phi, sp, cp);
    V3 h = normalize(t * (sinT * cp) + b * (sinT * sp) + n * cosT);
    if (dot(h, -incident) < 0.0f) h = -h;
    const V3 refl = incident - h * dot(h, incident) * 2.0f;  // glm::reflect
    if (dot(refl, n) > 0.0f) {
        out = normalize(refl);
        return true;
    }
    return false;
}
inline float specProbability(V3 view, V3 n, V3 base, float metallic) {  // Tracer.cpp:74-82
    const V3 f0 = baseRefl(base, metallic);
    const V3 f = fresnel(gmax(dot(n, view), 0.0f), f0);
    const float p = (f.x + f.y + f.z) / 3.0f;
    return clampf(p, 0.05f, 0.95f);
}
V3 computeNormal(V3 surf, V3 tex) {  // Tracer.cpp:95-105
    const V3 n = normalize(surf);
    const V3 up = ((MUT(9) ? n.y : std::fabs(n.y)) < 0.9f) ? V3{0.0f, 1.0f, 0.0f} : V3{1.0f, 0.0f, 0.0f};
    const V3 t = MUT(8) ? normalize(cross(up, n)) : normalize(cross(n, up));
    const V3 b = normalize(cross(t, n));
    return normalize(t * tex.x + b * tex.y + n * tex.z);
}
V3 aces(V3 c, float expo) {  // Tracer.cpp:212-222
    const V3 col = c * expo;
    const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
    return (col * (a * col + v3(b))) / (col * (cc * col + v3(d)) + v3(e));
}

// ---------------------------------------------------------------- path (Tracer.cpp:134-348)
struct Settings {
    mt_settings s;
    float expo;
};

struct PathStep {  // Tracer.hpp:34-56 (the fields the estimator and the debug views read)
    Ray ray;
    V3 hitPos, base, geomNormal, normal, energy, emission;
    float metallic, roughness;
    float brdfIntensity = 0.0f, pdf = 0.0f;
    uint32_t bvhTests = 0, triangleTests = 0;
    uint32_t triangleIndex = 0, bvIndex = 0, meshIndex = 0;  // & 0xffff
    bool didHit;
};

struct LightOut {  // Tracer.hpp LightOutput: the four sums before their total
    V3 emitted, direct, indirect, sky;
};

struct ShadeOut {
    V3 dir, energy;
    float pdf;
    bool spec;
    float brdfIntensity;
    float hcos = 0.0f;  // specular lobe: n.h of the sampled half vector (also on the fallback)
    bool fallback = false;  // specular lobe whose reflection fell below the surface
};

// the BRDF-sampling half of one GeneratePath iteration (Tracer.cpp:177-200)
ShadeOut sampleBounce(V3 view, V3 n, V3 base, float rough, float metal, Rng& rng) {
    ShadeOut o;
    const float p = specProbability(view, n, base, metal);
    const V3 incident = -view;
    if (rng.u(kDimLobe) < p) {
        o.spec = true;
        const float u1 = rng.u(kDimDir), u2 = rng.u(kDimDir + 1);
        if (!ggxDirection(u1, u2, incident, n, rough, o.dir, &o.hcos)) {
            o.fallback = true;
            const float f1 = rng.u(kDimFallback), f2 = rng.u(kDimFallback + 1);
            o.dir = cosineHemisphere(f1, f2, n);
        }
    } else {
        o.spec = false;
        const float u1 = rng.u(kDimDir), u2 = rng.u(kDimDir + 1);
        o.dir = cosineHemisphere(u1, u2, n);
    }
    const float dp = diffusePdf(o.dir, n);
    const float sp = specularPdf(o.dir, view, n, rough);
    o.pdf = (1.0f - p) * dp + p * sp;
    const V3 f = evalBrdf(o.dir, view, n, base, rough, metal);
    const float nl = gmax(dot(n, o.dir), 0.0f);
    o.energy = (o.pdf > 0.0f) ? f * (nl / o.pdf) : v3(0.0f);
    o.brdfIntensity = (f.x + f.y) + f.z;  // brdf.r + brdf.g + brdf.b (Tracer.cpp:203)
    return o;
}

// Light selection of one path step (Tracer.cpp:257-279): a one-pass weighted reservoir with one
// uniform per light, score = PrepareLights' mean intensity / (d^2 + 1).  Returns the chosen
// light (-1: none) and its probability score / sum of scores.
int chooseLight(const Scene& S, V3 hitPos, Rng& rng, float& pdf) {
    int chosen = -1;
    float chosenScore = 0.0f, totalScore = 0.0f;
    for (size_t i = 0; i < S.lights.size(); ++i) {  // Tracer.cpp:263-275
        const V3 lp = S.lights[i].pos - hitPos;
        const float d2 = dot(lp, lp);
        const float score = S.lights[i].intensity / (MUT(10) ? d2 : d2 + 1.0f);
        totalScore += score;
        const float r = rng.u(kDimLightSel + static_cast<uint32_t>(i)) * totalScore;
        if (MUT(11) ? r > score : r < score) {
            chosen = static_cast<int>(i);
            chosenScore = score;
        }
    }
    pdf = chosenScore / totalScore;
    return chosen;
}

// The sampled point of light L seen from hitPos and its colour there (Tracer.cpp:282-315):
// point and spot lights are jittered by GenerateRandomDirection (Tracer.cpp:84-93; phi and
// theta uniform, so not uniform on the sphere) times the point size, other types sit at their
// position; spot lights scale their colour by GetSpotIntensity (Light.hpp:58-63).
V3 lightPoint(const Light& L, V3 hitPos, Rng& rng, V3& lcol) {
    V3 lpos = L.pos;
    if (MUT(15) || L.type == MT_LIGHT_POINT || L.type == MT_LIGHT_SPOT) {
        const float phi = rng.uRange(kDimLightPt, 0.0f, 2.0f * kPi);
        const float th = MUT(12) ? std::asin(2.0f * rng.u(kDimLightPt + 1) - 1.0f)
                                 : rng.uRange(kDimLightPt + 1, -kPi * 0.5f, kPi * 0.5f);
        float sp, cp, st, ct;
        em::sincos(phi, sp, cp);
        em::sincos(th, st, ct);
        const V3 rd = normalize(V3{cp * ct, st, sp * ct});
        lpos = L.pos + rd * L.size;
    }
    lcol = L.color;
    if (L.type == MT_LIGHT_SPOT) {
        const V3 ldir = normalize(lpos - hitPos);
        const float dd = MUT(13) ? dot(ldir, L.dir) : dot(-ldir, L.dir);
        const float f = MUT(14) ? (dd - L.inner) / (L.outer - L.inner) : (dd - L.outer) / (L.inner - L.outer);
        lcol = L.color * clampf(f, 0.0f, 1.0f);
    }
    return lpos;
}

struct Tracer {
    const Scene& S;
    Settings cfg;
    uint32_t lightStride;
    bool faithfulInverse;
    bool useTlas = false;

    V3 materialBase(const Material& m, float u, float v, float& metal, float& rough, V3& nrmTex,
                    V3& emission) const {  // Material.hpp:82-135
        V3 base{m.bcf[0], m.bcf[1], m.bcf[2]};
        if (m.baseTex.image >= 0) {
            const RGBA c = sampleTex(S, m.baseTex, u, v);
            base = V3{c.r * m.bcf[0], c.g * m.bcf[1], c.b * m.bcf[2]};
        }
        if (m.mrTex.image >= 0) {
            const RGBA c = sampleTex(S, m.mrTex, u, v);
            metal = c.b * m.metallic;
            rough = c.g * m.roughness;
        } else {
            metal = m.metallic;
            rough = m.roughness;
        }
        if (m.nrmTex.image >= 0) {
            const RGBA c = sampleTex(S, m.nrmTex, u, v);
            nrmTex = normalize(V3{c.r * m.normalScale * 2.0f - 1.0f, c.g * m.normalScale * 2.0f - 1.0f,
                                  c.b * m.normalScale * 2.0f - 1.0f});
        } else {
            nrmTex = V3{0.0f, 0.0f, 1.0f};
        }
        if (m.emTex.image >= 0) {
            const RGBA c = sampleTex(S, m.emTex, u, v);
            emission = V3{c.r, c.g, c.b} * m.emissive;
        } else {
            emission = m.emissive;
        }
        return base;
    }

    // GeneratePath (Tracer.cpp:134-210) + ProcessRayForward (229-348) for one camera ray.
    V3 forward(const Ray& camRay, Rng& rng, std::vector<PathStep>& path, Counters* c,
               uint64_t& rays, uint64_t& hits, uint64_t& misses, LightOut* parts = nullptr) const {
        const mt_settings& st = cfg.s;
        path.clear();
        Ray cur = camRay;
        V3 total = v3(1.0f);
        const uint32_t dimBase = 2;
        while (path.size() < st.bounces) {
            const uint32_t k = static_cast<uint32_t>(path.size());
            if (!rng.mt) rng.base = dimBase + k * lightStride;
            PathStep step;
            const Hit h = intersectScene(S, cur, faithfulInverse, c, useTlas);
            rays++;
            if (!h.valid) {
                step.ray = cur;
                step.didHit = false;
                path.push_back(step);
                misses++;
                break;
            }
            hits++;
            // RayHitGeometryInfo (Ray.cpp:7-48)
            const Instance& I = S.instances[h.inst];
            const Mesh& m = S.meshes[I.mesh];
            const float w = 1.0f - h.u - h.v;
            const uint32_t t3 = h.tri * 3;
            const V3 ni = normalize(w * m.nrm[t3] + h.u * m.nrm[t3 + 1] + h.v * m.nrm[t3 + 2]);
            const V3 gN = normalize(xfDir(I.M, ni));
            float tu = 0.0f, tv = 0.0f;
            if (!m.uv.empty()) {
                tu = w * m.uv[t3 * 2] + h.u * m.uv[t3 * 2 + 2] + h.v * m.uv[t3 * 2 + 4];
                tv = w * m.uv[t3 * 2 + 1] + h.u * m.uv[t3 * 2 + 3] + h.v * m.uv[t3 * 2 + 5];
            }
            const Material& mat = S.materials[I.material];
            float metal, rough;
            V3 ntex, emission;
            V3 base = materialBase(mat, tu, tv, metal, rough, ntex, emission);
            base = base * 0.96f + v3(0.04f);  // Tracer.cpp:159
            step.ray = cur;
            step.hitPos = h.wpos;
            step.base = base;
            step.roughness = rough;
            step.metallic = metal;
            step.geomNormal = gN;
            step.normal = computeNormal(gN, ntex);
            step.emission = emission;
            step.didHit = true;
            step.bvhTests = h.bvhTests;  // of the winning instance (Tracer.cpp:170-175)
            step.triangleTests = h.triTests;
            step.bvIndex = h.bv & 0xffffu;
            step.triangleIndex = h.tri & 0xffffu;
            step.meshIndex = static_cast<uint32_t>(h.inst) & 0xffffu;  // instance, not a pointer
            const V3 view = -cur.d;
            const ShadeOut so = sampleBounce(view, step.normal, base, rough, metal, rng);
            step.brdfIntensity = so.brdfIntensity;
            step.pdf = so.pdf;
            Ray nr;
            nr.o = h.o + h.d * h.dist + 0.0001f * step.normal;  // ReflectSpecular/Diffuse
            nr.d = so.dir;
            step.energy = so.energy;
            path.push_back(step);
            cur = nr;
            total = total * step.energy;
            if (st.mode == MT_MODE_FIXED) {
                if (k + 1 >= st.rr_depth) {
                    const float q = clampf(gmax(gmax(total.x, total.y), total.z), 0.05f, 1.0f);
                    if (rng.u(kDimRR) >= q) break;
                    path.back().energy = path.back().energy / q;
                    total = total / q;
                }
            } else if (length(total) < st.terminate_energy) {
                break;
            }
        }

        // pass 2 (Tracer.cpp:235-340)
        V3 cE = v3(1.0f);
        V3 emitted = v3(0.0f), direct = v3(0.0f), indirect = v3(0.0f), sky = v3(0.0f);
        for (size_t si = 0; si < path.size(); ++si) {
            const PathStep& step = path[si];
            if (!rng.mt) rng.base = dimBase + static_cast<uint32_t>(si) * lightStride;
            if (!step.didHit) {
                const V3 col = (si > 0) ? path[si - 1].base : v3(1.0f);
                const bool prim = si == 0;
                const float mult = prim ? st.hdri_primary_intensity : st.hdri_secondary_intensity;
                const V3 sk = sampleEnv(S, step.ray.d, st.hdri_rotation_apply != 0,
                                        st.hdri_rotation) * mult;
                sky = sky + col * cE * sk;
                break;
            }
            emitted = emitted + cE * st.emission_base_intensity * step.emission;
            float chosenPdf = 0.0f;
            const int chosen = chooseLight(S, step.hitPos, rng, chosenPdf);
            if (chosen == -1) continue;  // Tracer.cpp:278-279: skips the throughput update
            V3 lcol;
            const V3 lpos = lightPoint(S.lights[chosen], step.hitPos, rng, lcol);
            const V3 lvec = lpos - step.hitPos;
            const V3 ldir = normalize(lvec);
            Ray sr;
            sr.o = step.hitPos + lvec * 0.0001f;
            sr.d = ldir;
            const Hit sh = intersectScene(S, sr, faithfulInverse, c, useTlas);
            rays++;
            const float llen = length(lvec);
            if (!sh.valid || sh.dist > llen) {
                const V3 f = evalBrdf(ldir, -step.ray.d, step.normal, step.base, step.roughness,
                                      step.metallic);
                const float nl = gmax(dot(step.normal, ldir), 0.0f);
                const float fall = 1.0f / (llen * llen + 1.0f);
                const V3 le = f * nl * fall;
                const V3 contrib = cE * le * lcol / chosenPdf;
                if (si == 0)
                    direct = direct + contrib;
                else
                    indirect = indirect + contrib;
            }
            cE = cE * step.energy;
        }
        if (parts) *parts = LightOut{emitted, direct, indirect, sky};
        return ((emitted + direct) + indirect) + sky;
    }
};

// RandomColorFromU16 (Tracer.cpp:350-364)
V3 colorFromU16(uint32_t value) {
    uint32_t x = value & 0xffffu;
    x ^= x >> 7;
    x *= 0x2C1B3C6Du;
    x ^= x >> 12;
    x *= 0x297A2D39u;
    x ^= x >> 15;
    const uint32_t v = x & 0xffffu;
    return {static_cast<float>((v & 0x01C0u) >> 6) / 7.0f, static_cast<float>((v & 0x0038u) >> 3) / 7.0f,
            static_cast<float>(v & 0x0007u) / 7.0f};
}

// ProcessDebugRay (Tracer.cpp:366-488) for a path already traced; `total` is the sample's
// regular colour (used by the modes the function does not handle)
V3 debugColor(int mode, const std::vector<PathStep>& path, const LightOut& L, const Ray& ray,
              V3 total, float expo) {
    const bool hit = !path.empty() && path[0].didHit;
    const V3 black = v3(0.0f);
    switch (mode) {
        case MT_DEBUG_PRIMARY_BVH_TESTS: return hit ? v3(static_cast<float>(path[0].bvhTests)) : black;
        case MT_DEBUG_TOTAL_BVH_TESTS: {
            float sum = 0.0f;
            for (const PathStep& st : path) sum += static_cast<float>(st.didHit ? st.bvhTests : 0u);
            return hit ? v3(sum) : black;
        }
        case MT_DEBUG_PRIMARY_TRIANGLE_TESTS: return hit ? v3(static_cast<float>(path[0].triangleTests)) : black;
        case MT_DEBUG_TOTAL_TRIANGLE_TESTS: {
            float sum = 0.0f;
            for (const PathStep& st : path) sum += static_cast<float>(st.didHit ? st.triangleTests : 0u);
            return hit ? v3(sum) : black;
        }
        case MT_DEBUG_BOUNCES: return hit ? v3(static_cast<float>(path.size())) : black;
        case MT_DEBUG_ALBEDO: return hit ? path[0].base : black;
        case MT_DEBUG_METALLIC_ROUGHNESS: return hit ? V3{0.0f, path[0].metallic, path[0].roughness} : black;
        case MT_DEBUG_GEOMETRIC_NORMAL: return hit ? path[0].geomNormal * 0.5f + v3(0.5f) : black;
        case MT_DEBUG_SHADING_NORMAL: return hit ? path[0].normal * 0.5f + v3(0.5f) : black;
        case MT_DEBUG_DIRECT_ONLY: return aces(L.direct, expo);
        case MT_DEBUG_INDIRECT_ONLY: return aces(L.indirect + L.sky, expo);
        case MT_DEBUG_EMISSION: return aces(L.emitted, expo);
        case MT_DEBUG_COLOR_PER_MESH: return hit ? colorFromU16(path[0].meshIndex) : black;
        case MT_DEBUG_COLOR_PER_TRIANGLE: return hit ? colorFromU16(path[0].triangleIndex) : black;
        case MT_DEBUG_COLOR_PER_BOUNDING_VOLUME: return hit ? colorFromU16(path[0].bvIndex) : black;
        case MT_DEBUG_DEPTH: return hit ? v3(length(path[0].hitPos - ray.o)) : black;
        case MT_DEBUG_FIRST_HIT_FRESNEL: {
            if (!hit) return black;
            const V3 f0 = baseRefl(path[0].base, path[0].metallic);
            return fresnel(gmax(dot(path[0].normal, -ray.d), 0.0f), f0);
        }
        case MT_DEBUG_FIRST_HIT_BRDF: return hit ? v3(path[0].brdfIntensity) : black;
        case MT_DEBUG_FIRST_HIT_PDF: return hit ? v3(path[0].pdf) : black;
        case MT_DEBUG_REFLECTED_DIRECTION: return path.size() > 1 ? path[1].ray.d * 0.5f + v3(0.5f) : black;
        default: return total;
    }
}

// ApplyDebugPostProcessing (Tracer.cpp:642-716): heat map of (r+g+b)/3 over the image maximum
V3 heatmap(float t) {
    t = clampf(t, 0.0f, 1.0f);
    static const V3 colors[5] = {{0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}};
    const float x = t * 4.0f;
    const int i = std::min(static_cast<int>(x), 3);
    const float f = x - static_cast<float>(i);
    return {mixf(colors[i].x, colors[i + 1].x, f), mixf(colors[i].y, colors[i + 1].y, f),
            mixf(colors[i].z, colors[i + 1].z, f)};
}
void debugPost(int mode, float* rgb, size_t npix) {
    bool colorful;
    switch (mode) {
        case MT_DEBUG_PRIMARY_BVH_TESTS: case MT_DEBUG_TOTAL_BVH_TESTS:
        case MT_DEBUG_PRIMARY_TRIANGLE_TESTS: case MT_DEBUG_TOTAL_TRIANGLE_TESTS:
        case MT_DEBUG_BOUNCES: case MT_DEBUG_FIREFLY_ELIMINATION:
            colorful = true;
            break;
        case MT_DEBUG_DEPTH: case MT_DEBUG_FIRST_HIT_BRDF: case MT_DEBUG_PIXEL_STANDARD_DEVIATION:
            colorful = false;
            break;
        default:
            return;
    }
    float mx = 0.0f;
    for (size_t i = 0; i < npix; ++i) mx = std::max(mx, (rgb[i * 3] + rgb[i * 3 + 1] + rgb[i * 3 + 2]) / 3.0f);
    for (size_t i = 0; i < npix; ++i) {
        const float in = (rgb[i * 3] + rgb[i * 3 + 1] + rgb[i * 3 + 2]) / 3.0f;
        const float t = mx > 0.0f ? in / mx : 0.0f;
        const V3 c = colorful ? heatmap(t) : v3(t);
        rgb[i * 3] = c.x;
        rgb[i * 3 + 1] = c.y;
        rgb[i * 3 + 2] = c.z;
    }
}

// ---------------------------------------------------------------- differentiable estimator
// Forward-mode (tangent) derivative of the per-sample radiance along one direction in parameter
// space, with the sampling detached: every decision (lobe, direction, light, termination) is
// taken from the primal values and the derivative flows only through the BRDF values, the
// base colour, the emission and the throughput products.  This is an automatic-differentiation
// restatement, independent of the hand-derived reverse-mode (adjoint) kernels it checks.
struct Dv {
    float v, d;
};
inline Dv dv(float v) { return {v, 0.0f}; }
inline Dv operator+(Dv a, Dv b) { return {a.v + b.v, a.d + b.d}; }
inline Dv operator-(Dv a, Dv b) { return {a.v - b.v, a.d - b.d}; }
inline Dv operator*(Dv a, Dv b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dv operator/(Dv a, Dv b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dv operator+(Dv a, float b) { return a + dv(b); }
inline Dv operator-(float a, Dv b) { return dv(a) - b; }
inline Dv operator*(Dv a, float b) { return a * dv(b); }
inline Dv operator*(float a, Dv b) { return dv(a) * b; }
inline Dv operator/(Dv a, float b) { return a / dv(b); }
inline Dv operator-(Dv a, float b) { return a - dv(b); }
inline Dv dmax(Dv a, float b) { return (a.v < b) ? dv(b) : a; }  // same selection as gmax
inline Dv dmin(Dv a, float b) { return (b < a.v) ? dv(b) : a; }  // same selection as gmin
struct D3 {
    Dv c[3];
};
inline D3 d3(V3 v) { return {{dv(v.x), dv(v.y), dv(v.z)}}; }
inline D3 d3(V3 v, V3 t) { return {{{v.x, t.x}, {v.y, t.y}, {v.z, t.z}}}; }
inline D3 operator+(const D3& a, const D3& b) { return {{a.c[0] + b.c[0], a.c[1] + b.c[1], a.c[2] + b.c[2]}}; }
inline D3 operator*(const D3& a, const D3& b) { return {{a.c[0] * b.c[0], a.c[1] * b.c[1], a.c[2] * b.c[2]}}; }
inline D3 operator*(const D3& a, float s) { return {{a.c[0] * s, a.c[1] * s, a.c[2] * s}}; }
inline V3 tangent(const D3& a) { return {a.c[0].d, a.c[1].d, a.c[2].d}; }

// EvaluateBRDF (BRDF.hpp:52-77) over tangent numbers in base, roughness and metallic
D3 evalBrdfDv(V3 L, V3 V, V3 N, const D3& base, Dv rough, Dv metal) {
    D3 out = d3(v3(0.0f));
    rough = dmax(rough, 0.05f);
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return out;
    const V3 h = normalize(V + L);
    const float nh = clampf(gmax(dot(N, h), 0.0f), 0.0f, 1.0f);
    const float hv = clampf(gmax(dot(h, V), 0.0f), 0.0f, 1.0f);
    const Dv mc = dmin(dmax(metal, 0.0f), 1.0f);
    const float p5 = em::pow5(1.0f - hv);
    const Dv a = rough * rough;
    const Dv a2 = a * a;
    Dv den = (a2 - dv(1.0f)) * (nh * nh) + 1.0f;
    den = (den * den) * kPi;
    const Dv D = a2 / den;
    const Dv rr = rough + 1.0f;
    const Dv k = rr * rr / 8.0f;
    const Dv g1 = dv(nv) / ((1.0f - k) * nv + k);
    const Dv g2 = dv(nl) / ((1.0f - k) * nl + k);
    const float den4 = gmax(4.0f * nv * nl, 0.001f);
    for (int c = 0; c < 3; ++c) {
        const Dv f0 = (1.0f - mc) * 0.04f + base.c[c] * mc;
        const Dv F = f0 + (1.0f - f0) * p5;
        const Dv kd = (1.0f - F) * (1.0f - metal);
        const Dv diffuse = kd * base.c[c] / kPi;
        const Dv spec = D * (g1 * g2) * F / den4;
        out.c[c] = dmax(diffuse + spec, 0.0f);
    }
    return out;
}

// The sampling's dependence on the parameters, for the full derivative of the expected image
// (score-function form): the specular-lobe probability (Tracer.cpp:74-82), the GGX density of
// the sampled half vector (Tracer.cpp:49-72: tan^2 = alpha^2 u / (1 - u) samples D(h) n.h) and
// the mixture pdf the estimator divides by (Tracer.cpp:190-196), over tangent numbers.
Dv distGGXDv(float nh, Dv r) {  // D(h) of BRDF.hpp:22-33 at roughness max(r, 0.05)
    nh = clampf(nh, 0.0f, 1.0f);
    const Dv rc = dmax(r, 0.05f);
    const Dv a = rc * rc;
    const Dv a2 = a * a;
    const float nh2 = nh * nh;
    const Dv den = nh2 * (a2 - 1.0f) + 1.0f;
    return a2 / (kPi * den * den);
}
Dv specProbabilityDv(V3 view, V3 n, const D3& base, Dv metal) {
    const float c = clampf(gmax(dot(n, view), 0.0f), 0.0f, 1.0f);
    const float p5 = em::pow5(1.0f - c);
    const Dv m = dmin(dmax(metal, 0.0f), 1.0f);
    Dv sum = dv(0.0f);
    for (int k = 0; k < 3; ++k) {
        const Dv f0 = dv(0.04f) * (1.0f - m) + base.c[k] * m;
        sum = sum + (f0 + (1.0f - f0) * p5);
    }
    const Dv p = sum / 3.0f;
    return (p.v < 0.05f) ? dv(0.05f) : ((0.95f < p.v) ? dv(0.95f) : p);
}
Dv specularPdfDv(V3 L, V3 V, V3 N, Dv rough) {
    const float nl = gmax(dot(N, L), 0.0f);
    const float nv = gmax(dot(N, V), 0.0f);
    if (nl <= 0.0f || nv <= 0.0f) return dv(0.0f);
    const V3 h = normalize(V + L);
    const float nh = gmax(dot(N, h), 0.0f);
    const float vh = gmax(dot(V, h), 0.0f);
    if (vh <= 0.0f) return dv(0.0f);
    return distGGXDv(nh, rough) * (nh / (4.0f * vh));
}

struct DiffSet {  // host copies of the differentiable parameters (value or tangent)
    bool hasTex = false;
    Image tex;
    uint32_t texFirst = 0, texCount = 0;
    std::vector<float> rough, metal, emis;
};

struct DiffTracer {
    const Scene& S;
    mt_settings st;
    uint32_t lightStride;
    const DiffSet& P;  // values
    const DiffSet& T;  // tangents (same shapes, empty = zero)
    bool samplingTerms = false;  // add the sampling decisions' dependence (full derivative)

    struct MatD {
        D3 base, emission;
        Dv rough, metal;
        V3 ntex;
    };

    // Material::SampleMaterial (Material.hpp:82-135) with the parameter overrides
    MatD material(uint32_t mi, float u, float v) const {
        const Material& m = S.materials[mi];
        MatD o;
        if (P.hasTex && mi >= P.texFirst && mi - P.texFirst < P.texCount) {
            const RGBA c = sampleImage(P.tex, u, v, MT_FILTER_LINEAR);
            V3 t = v3(0.0f);
            if (T.hasTex) {
                const RGBA ct = sampleImage(T.tex, u, v, MT_FILTER_LINEAR);
                t = V3{ct.r, ct.g, ct.b};
            }
            o.base = d3(V3{c.r, c.g, c.b}, t);
        } else if (m.baseTex.image >= 0) {
            const RGBA c = sampleTex(S, m.baseTex, u, v);
            o.base = d3(V3{c.r * m.bcf[0], c.g * m.bcf[1], c.b * m.bcf[2]});
        } else {
            o.base = d3(V3{m.bcf[0], m.bcf[1], m.bcf[2]});
        }
        const Dv rf = P.rough.empty() ? dv(m.roughness) : Dv{P.rough[mi], T.rough.empty() ? 0.0f : T.rough[mi]};
        const Dv mf = P.metal.empty() ? dv(m.metallic) : Dv{P.metal[mi], T.metal.empty() ? 0.0f : T.metal[mi]};
        if (m.mrTex.image >= 0) {
            const RGBA c = sampleTex(S, m.mrTex, u, v);
            o.metal = mf * c.b;
            o.rough = rf * c.g;
        } else {
            o.metal = mf;
            o.rough = rf;
        }
        if (m.nrmTex.image >= 0) {
            const RGBA c = sampleTex(S, m.nrmTex, u, v);
            o.ntex = normalize(V3{c.r * m.normalScale * 2.0f - 1.0f, c.g * m.normalScale * 2.0f - 1.0f,
                                  c.b * m.normalScale * 2.0f - 1.0f});
        } else {
            o.ntex = V3{0.0f, 0.0f, 1.0f};
        }
        D3 em = d3(m.emissive);
        if (!P.emis.empty()) {
            for (int c = 0; c < 3; ++c)
                em.c[c] = Dv{P.emis[mi * 3 + c], T.emis.empty() ? 0.0f : T.emis[mi * 3 + c]};
        }
        if (m.emTex.image >= 0) {
            const RGBA c = sampleTex(S, m.emTex, u, v);
            o.emission = em * d3(V3{c.r, c.g, c.b});
        } else {
            o.emission = em;
        }
        return o;
    }

    struct StepD {
        Ray ray;
        V3 hitPos, normal;
        D3 base, energy, emission;
        Dv rough, metal;
        bool didHit;
        // full derivative: tangents of log p(decisions of this bounce) and of log pdf(direction)
        float score = 0.0f, dlogpdf = 0.0f;
    };

    // GeneratePath + ProcessRayForward (Tracer.cpp:134-348) for one camera ray, tangent numbers
    D3 forward(const Ray& camRay, Rng& rng, std::vector<StepD>& path) const {
        path.clear();
        Ray cur = camRay;
        V3 total = v3(1.0f);
        while (path.size() < st.bounces) {
            const uint32_t k = static_cast<uint32_t>(path.size());
            rng.base = 2 + k * lightStride;
            StepD step;
            const Hit h = intersectScene(S, cur, false, nullptr);
            if (!h.valid) {
                step.ray = cur;
                step.didHit = false;
                path.push_back(step);
                break;
            }
            const Instance& I = S.instances[h.inst];
            const Mesh& m = S.meshes[I.mesh];
            const float w = 1.0f - h.u - h.v;
            const uint32_t t3 = h.tri * 3;
            const V3 ni = normalize(w * m.nrm[t3] + h.u * m.nrm[t3 + 1] + h.v * m.nrm[t3 + 2]);
            const V3 gN = normalize(xfDir(I.M, ni));
            float tu = 0.0f, tv = 0.0f;
            if (!m.uv.empty()) {
                tu = w * m.uv[t3 * 2] + h.u * m.uv[t3 * 2 + 2] + h.v * m.uv[t3 * 2 + 4];
                tv = w * m.uv[t3 * 2 + 1] + h.u * m.uv[t3 * 2 + 3] + h.v * m.uv[t3 * 2 + 5];
            }
            const MatD md = material(I.material, tu, tv);
            step.base = md.base * 0.96f;
            for (int c = 0; c < 3; ++c) step.base.c[c] = step.base.c[c] + 0.04f;  // Tracer.cpp:159
            step.ray = cur;
            step.hitPos = h.wpos;
            step.rough = md.rough;
            step.metal = md.metal;
            step.normal = computeNormal(gN, md.ntex);
            step.emission = md.emission;
            step.didHit = true;
            const V3 view = -cur.d;
            const V3 basev{step.base.c[0].v, step.base.c[1].v, step.base.c[2].v};
            const ShadeOut so = sampleBounce(view, step.normal, basev, md.rough.v, md.metal.v, rng);
            // d(energy) = d(f) * nl / pdf with the sampled direction and its pdf held fixed
            const D3 fd = evalBrdfDv(so.dir, view, step.normal, step.base, md.rough, md.metal);
            const float nl = gmax(dot(step.normal, so.dir), 0.0f);
            const V3 ed = (so.pdf > 0.0f) ? tangent(fd) * (nl / so.pdf) : v3(0.0f);
            step.energy = d3(so.energy, ed);
            if (samplingTerms) {
                // the bounce's decisions: the lobe (probability p_s or 1 - p_s) and, on the
                // specular lobe, the half vector (density D(h) n.h; drawn also when the
                // reflection falls below the surface and a cosine direction replaces it)
                const Dv ps = specProbabilityDv(view, step.normal, step.base, md.metal);
                if (so.spec) {
                    const Dv dh = distGGXDv(so.hcos, md.rough);
                    step.score = ps.d / ps.v + (dh.v > 0.0f ? dh.d / dh.v : 0.0f);
                } else {
                    step.score = -ps.d / (1.0f - ps.v);
                }
                const Dv pdf = (1.0f - ps) * diffusePdf(so.dir, step.normal) +
                               ps * specularPdfDv(so.dir, view, step.normal, md.rough);
                step.dlogpdf = pdf.v > 0.0f ? pdf.d / pdf.v : 0.0f;
            }
            Ray nr;
            nr.o = h.o + h.d * h.dist + 0.0001f * step.normal;
            nr.d = so.dir;
            path.push_back(step);
            cur = nr;
            total = total * so.energy;
            if (st.mode == MT_MODE_FIXED) {
                if (k + 1 >= st.rr_depth) {
                    const float q = clampf(gmax(gmax(total.x, total.y), total.z), 0.05f, 1.0f);
                    if (rng.u(kDimRR) >= q) break;
                    path.back().energy = path.back().energy * (1.0f / q);
                    for (int c = 0; c < 3; ++c) path.back().energy.c[c].v = get(so.energy, c) / q;
                    total = total / q;
                }
            } else if (length(total) < st.terminate_energy) {
                break;
            }
        }
        D3 cE = d3(v3(1.0f));
        D3 L = d3(v3(0.0f));
        // full derivative: (radiance so far, tangent factor) after each bounce decision; what
        // is collected after the decision times the factor is added at the end
        struct Mark {
            V3 sofar;
            float factor;
        };
        std::vector<Mark> marks;
        auto mark = [&](const StepD& step, bool chosen) {
            if (!samplingTerms) return;
            // the pdf divides the throughput only when the step updates it (a light was chosen,
            // Tracer.cpp:278-279); the decisions steer the rest of the path either way
            marks.push_back({V3{L.c[0].v, L.c[1].v, L.c[2].v},
                             step.score - (chosen ? step.dlogpdf : 0.0f)});
        };
        for (size_t si = 0; si < path.size(); ++si) {
            const StepD& step = path[si];
            rng.base = 2 + static_cast<uint32_t>(si) * lightStride;
            if (!step.didHit) {
                const D3 col = (si > 0) ? path[si - 1].base : d3(v3(1.0f));
                const float mult = si == 0 ? st.hdri_primary_intensity : st.hdri_secondary_intensity;
                const V3 sk = sampleEnv(S, step.ray.d, st.hdri_rotation_apply != 0, st.hdri_rotation) * mult;
                L = L + col * cE * d3(sk);
                break;
            }
            L = L + cE * step.emission * st.emission_base_intensity;
            float chosenPdf = 0.0f;
            const int chosen = chooseLight(S, step.hitPos, rng, chosenPdf);
            if (chosen == -1) {
                mark(step, false);
                continue;
            }
            V3 lcol;
            const V3 lpos = lightPoint(S.lights[chosen], step.hitPos, rng, lcol);
            const V3 lvec = lpos - step.hitPos;
            const V3 ldir = normalize(lvec);
            Ray sr;
            sr.o = step.hitPos + lvec * 0.0001f;
            sr.d = ldir;
            const Hit sh = intersectScene(S, sr, false, nullptr);
            const float llen = length(lvec);
            if (!sh.valid || sh.dist > llen) {
                const D3 f = evalBrdfDv(ldir, -step.ray.d, step.normal, step.base, step.rough, step.metal);
                const float nl = gmax(dot(step.normal, ldir), 0.0f);
                const float fall = 1.0f / (llen * llen + 1.0f);
                L = L + cE * f * d3(lcol * (nl * fall / chosenPdf));
            }
            mark(step, true);
            cE = cE * step.energy;
        }
        for (const Mark& mk : marks)
            for (int c = 0; c < 3; ++c) L.c[c].d = L.c[c].d + (L.c[c].v - get(mk.sofar, c)) * mk.factor;
        return L;
    }
};

DiffSet makeDiffSet(const orc_diff_params* p, uint32_t nmat) {
    DiffSet d;
    if (!p) return d;
    if (p->tex) {
        d.hasTex = true;
        d.tex.w = static_cast<int>(p->tex_w);
        d.tex.h = static_cast<int>(p->tex_h);
        d.tex.rgb.assign(p->tex, p->tex + static_cast<size_t>(p->tex_w) * p->tex_h * 3);
        d.texFirst = p->tex_first;
        d.texCount = p->tex_count;
    }
    if (p->rough) d.rough.assign(p->rough, p->rough + nmat);
    if (p->metal) d.metal.assign(p->metal, p->metal + nmat);
    if (p->emis) d.emis.assign(p->emis, p->emis + nmat * 3);
    return d;
}

// the camera (Tracer/Camera.cpp:17-50)
Ray cameraRay(const mt_camera_desc& cam, float u, float v, float aspect) {
    M4 M;
    std::memcpy(M.m, cam.camera_to_world, sizeof(M.m));
    if (cam.type == MT_CAMERA_PERSPECTIVE) {
        const float fs = std::tan(cam.yfov * 0.5f);
        const float px = (2.0f * u - 1.0f) * fs * aspect;
        const float py = (2.0f * v - 1.0f) * fs;
        const V3 dc = normalize(V3{px, -py, -1.0f});
        Ray r;
        r.o = xfPoint(M, v3(0.0f));
        r.d = normalize(xfDir(M, dc));
        return r;
    }
    const float px = (2.0f * u - 1.0f) * cam.ortho_size[0] * aspect / cam.aspect;
    const float py = (2.0f * v - 1.0f) * cam.ortho_size[1];
    const V3 dc = normalize(V3{0.0f, 0.0f, -1.0f});
    Ray r;
    r.o = xfPoint(M, V3{px, -py, 0.0f});
    r.d = normalize(xfDir(M, dc));
    return r;
}

uint32_t cleanupFireflies(std::vector<V3>& buf, float thr, int debugMode = 0) {  // Tracer.cpp:490-548
    auto inten = [](V3 c) { return (c.x + c.y + c.z) / 3.0f; };
    size_t n = buf.size();
    float sum = 0.0f;
    for (const V3& c : buf) sum = sum + inten(c);
    const float avg = sum / static_cast<float>(n);
    float var = 0.0f;
    for (const V3& c : buf) {
        const float d = inten(c) - avg;
        var = var + d * d;
    }
    var = var / static_cast<float>(n);
    const float sd = std::sqrt(var);
    if (debugMode == MT_DEBUG_PIXEL_STANDARD_DEVIATION) {  // Tracer.cpp:515-523
        for (V3& c : buf) c = v3(sd);
        return static_cast<uint32_t>(buf.size());
    }
    const float t = thr * sd;
    for (V3& c : buf) {
        if (std::fabs(inten(c) - avg) > t) {
            c = v3(0.0f);
            n--;
        }
    }
    if (debugMode == MT_DEBUG_FIREFLY_ELIMINATION) {  // Tracer.cpp:536-544
        for (V3& c : buf) c = v3(static_cast<float>(buf.size() - n));
        return static_cast<uint32_t>(buf.size());
    }
    return n == 0 ? 1u : static_cast<uint32_t>(n);
}

bool loadScene(const mt_scene_desc* d, Scene& S) {
    if (!d) {
        g_err = "null scene description";
        return false;
    }
    const unsigned maxTri = d->bvh.max_triangles, maxDepth = d->bvh.max_depth;
    S.meshes.resize(d->mesh_count);
    for (uint32_t i = 0; i < d->mesh_count; ++i) {
        const mt_mesh_desc& md = d->meshes[i];
        Mesh& m = S.meshes[i];
        const size_t nc = static_cast<size_t>(md.tri_count) * 3;
        m.pos.resize(nc);
        m.nrm.resize(nc);
        for (size_t k = 0; k < nc; ++k) {
            m.pos[k] = V3{md.positions[k * 3], md.positions[k * 3 + 1], md.positions[k * 3 + 2]};
            m.nrm[k] = V3{md.normals[k * 3], md.normals[k * 3 + 1], md.normals[k * 3 + 2]};
        }
        if (md.texcoord0) m.uv.assign(md.texcoord0, md.texcoord0 + nc * 2);
        // Mesh::CalculateAABB (Mesh.cpp:43-54)
        V3 mn = v3(FLT_MAX), mx = v3(-FLT_MAX);
        for (const V3& p : m.pos) {
            mn = vmin(mn, p);
            mx = vmax(mx, p);
        }
        m.aabbMin = mn;
        m.aabbMax = mx;
        // Mesh::Mesh BVH trigger (Mesh.cpp:37-40)
        if (!d->bvh.disable && md.tri_count > maxTri) buildBvh(m, maxTri, maxDepth - 1);
    }
    S.images.resize(d->image_count);
    for (uint32_t i = 0; i < d->image_count; ++i) {
        const mt_image_desc& id = d->images[i];
        Image& im = S.images[i];
        im.w = static_cast<int>(id.width);
        im.h = static_cast<int>(id.height);
        im.rgb.resize(static_cast<size_t>(im.w) * im.h * 3);
        for (size_t k = 0; k < im.rgb.size(); ++k) im.rgb[k] = static_cast<float>(id.rgb[k]) / 255.0f;
    }
    auto tex = [&](const mt_texture_desc& t) {
        Tex o;
        o.image = (t.image >= 0 && static_cast<uint32_t>(t.image) < d->image_count) ? t.image : -1;
        o.filter = t.filter;
        o.ox = t.offset[0];
        o.oy = t.offset[1];
        o.sx = t.scale[0];
        o.sy = t.scale[1];
        return o;
    };
    S.materials.resize(d->material_count);
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const mt_material_desc& md = d->materials[i];
        Material& m = S.materials[i];
        std::memcpy(m.bcf, md.base_color_factor, sizeof(m.bcf));
        m.emissive = V3{md.emissive_factor[0], md.emissive_factor[1], md.emissive_factor[2]};
        m.normalScale = md.normal_scale;
        m.metallic = md.metallic_factor;
        m.roughness = md.roughness_factor;
        m.baseTex = tex(md.base_color_tex);
        m.mrTex = tex(md.metallic_roughness_tex);
        m.nrmTex = tex(md.normal_tex);
        m.emTex = tex(md.emissive_tex);
    }
    S.instances.resize(d->instance_count);
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const mt_instance_desc& id = d->instances[i];
        if (id.mesh >= d->mesh_count || id.material >= d->material_count) {
            g_err = "instance references a missing mesh or material";
            return false;
        }
        Instance& I = S.instances[i];
        I.mesh = id.mesh;
        I.material = id.material;
        std::memcpy(I.M.m, id.transform, sizeof(I.M.m));
        I.inv = invert4(I.M);
        // MeshInstance::CalculateWorldAABB (Mesh.cpp:56-80)
        const Mesh& m = S.meshes[id.mesh];
        const V3 a = m.aabbMin, b = m.aabbMax;
        const V3 corners[8] = {{a.x, a.y, a.z}, {a.x, a.y, b.z}, {a.x, b.y, a.z}, {a.x, b.y, b.z},
                               {b.x, a.y, a.z}, {b.x, a.y, b.z}, {b.x, b.y, a.z}, {b.x, b.y, b.z}};
        I.wmin = v3(FLT_MAX);
        I.wmax = v3(-FLT_MAX);
        for (const V3& c : corners) {
            const V3 t = xfPoint(I.M, c);
            I.wmin = vmin(I.wmin, t);
            I.wmax = vmax(I.wmax, t);
        }
    }
    S.lights.resize(d->light_count);
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const mt_light_desc& ld = d->lights[i];
        Light& L = S.lights[i];
        L.type = ld.type;
        L.pos = V3{ld.position[0], ld.position[1], ld.position[2]};
        L.color = V3{ld.color[0], ld.color[1], ld.color[2]};
        L.dir = V3{ld.direction[0], ld.direction[1], ld.direction[2]};
        L.size = ld.point_size;
        L.inner = ld.spot_inner_cos;
        L.outer = ld.spot_outer_cos;
        L.intensity = (L.color.x + L.color.y + L.color.z) / 3.0f;  // PrepareLights, Tracer.cpp:123-132
    }
    S.hasEnv = false;
    if (d->env.format != MT_ENV_NONE && d->env.texels && d->env.width && d->env.height) {
        S.hasEnv = true;
        S.envFilter = d->env.filter;
        S.env.w = static_cast<int>(d->env.width);
        S.env.h = static_cast<int>(d->env.height);
        const size_t n = static_cast<size_t>(S.env.w) * S.env.h * 3;
        S.env.rgb.resize(n);
        if (d->env.format == MT_ENV_RGB8) {
            const uint8_t* p = static_cast<const uint8_t*>(d->env.texels);
            for (size_t k = 0; k < n; ++k) S.env.rgb[k] = static_cast<float>(p[k]) / 255.0f;
        } else {
            const float* p = static_cast<const float*>(d->env.texels);
            std::copy(p, p + n, S.env.rgb.begin());
        }
    }
    return true;
}

void toHit(const Hit& h, mt_hit& o) {
    o.instance = h.valid ? h.inst : -1;
    o.triangle = h.tri;
    o.bv_index = h.bv;
    o.u = h.u;
    o.v = h.v;
    o.distance = h.valid ? h.dist : 0.0f;
    o.world_pos[0] = h.wpos.x;
    o.world_pos[1] = h.wpos.y;
    o.world_pos[2] = h.wpos.z;
    o.bvh_tests = h.bvhTests;
    o.triangle_tests = h.triTests;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

void* orc_scene_create(const mt_scene_desc* desc) {
    try {
        auto* S = new Scene();
        if (!loadScene(desc, *S)) {
            delete S;
            return nullptr;
        }
        buildTlas(*S);
        return S;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_scene_destroy(void* s) { delete static_cast<Scene*>(s); }

int orc_scene_mesh_bvh(const void* sp, uint32_t mesh, uint32_t* nodes_out, uint32_t* node_count,
                       uint32_t* indices_out, uint32_t* index_count) {
    const Scene& S = *static_cast<const Scene*>(sp);
    if (mesh >= S.meshes.size()) return MT_E_INVALID;
    const Mesh& m = S.meshes[mesh];
    if (node_count) *node_count = static_cast<uint32_t>(m.nodes.size());
    if (index_count) *index_count = static_cast<uint32_t>(m.indices.size());
    if (nodes_out)
        for (size_t i = 0; i < m.nodes.size(); ++i) std::memcpy(nodes_out + i * 8, &m.nodes[i], 32);
    if (indices_out) std::copy(m.indices.begin(), m.indices.end(), indices_out);
    return MT_OK;
}

int orc_scene_world_aabbs(const void* sp, float* out) {
    const Scene& S = *static_cast<const Scene*>(sp);
    for (size_t i = 0; i < S.instances.size(); ++i) {
        const Instance& I = S.instances[i];
        const float v[6] = {I.wmin.x, I.wmin.y, I.wmin.z, I.wmax.x, I.wmax.y, I.wmax.z};
        std::memcpy(out + i * 6, v, sizeof(v));
    }
    return MT_OK;
}

int orc_trace_closest(const void* sp, uint32_t n, const float* rays, mt_hit* hits, orc_work* work) {
    const Scene& S = *static_cast<const Scene*>(sp);
    Counters c;
    for (uint32_t i = 0; i < n; ++i) {
        Ray r;
        r.o = V3{rays[i * 6], rays[i * 6 + 1], rays[i * 6 + 2]};
        r.d = V3{rays[i * 6 + 3], rays[i * 6 + 4], rays[i * 6 + 5]};
        toHit(intersectScene(S, r, true, &c), hits[i]);
    }
    if (work) {
        std::memset(work, 0, sizeof(*work));
        work->rays = n;
        work->inst_aabb_tests = c.aabb;
        work->inst_entered = c.enter;
        work->node_pops = c.pops;
        work->tri_tests = c.tris;
    }
    return MT_OK;
}

float orc_slab(const float* ray, const float* box) {
    Ray r;
    r.o = V3{ray[0], ray[1], ray[2]};
    r.d = V3{ray[3], ray[4], ray[5]};
    return slab(r, V3{box[0], box[1], box[2]}, V3{box[3], box[4], box[5]});
}

int orc_triangle_test(const float* ray, const float* tri, float* uvt) {
    float u = 0, v = 0, t = 0;
    const bool hit = triangleTest(V3{ray[0], ray[1], ray[2]}, V3{ray[3], ray[4], ray[5]},
                                  V3{tri[0], tri[1], tri[2]}, V3{tri[3], tri[4], tri[5]},
                                  V3{tri[6], tri[7], tri[8]}, u, v, t);
    uvt[0] = u;
    uvt[1] = v;
    uvt[2] = t;
    return hit ? 1 : 0;
}

void orc_invert4(const float* m, float* out) {
    M4 M;
    std::memcpy(M.m, m, sizeof(M.m));
    const M4 r = invert4(M);
    std::memcpy(out, r.m, sizeof(r.m));
}

int orc_trace_bruteforce(const void* sp, uint32_t n, const float* rays, mt_hit* hits) {
    const Scene& S = *static_cast<const Scene*>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const double o[3] = {rays[i * 6], rays[i * 6 + 1], rays[i * 6 + 2]};
        const double d[3] = {rays[i * 6 + 3], rays[i * 6 + 4], rays[i * 6 + 5]};
        double best = 1e300;
        mt_hit h{};
        h.instance = -1;
        for (size_t ii = 0; ii < S.instances.size(); ++ii) {
            const Instance& I = S.instances[ii];
            const Mesh& m = S.meshes[I.mesh];
            const size_t T = m.pos.size() / 3;
            for (size_t t = 0; t < T; ++t) {
                double P[3][3];
                for (int k = 0; k < 3; ++k) {
                    const V3 p = m.pos[t * 3 + k];
                    for (int r = 0; r < 3; ++r)
                        P[k][r] = static_cast<double>(I.M.at(0, r)) * p.x +
                                  static_cast<double>(I.M.at(1, r)) * p.y +
                                  static_cast<double>(I.M.at(2, r)) * p.z + I.M.at(3, r);
                }
                double e1[3], e2[3], pv[3], s[3], q[3];
                for (int r = 0; r < 3; ++r) {
                    e1[r] = P[1][r] - P[0][r];
                    e2[r] = P[2][r] - P[0][r];
                    s[r] = o[r] - P[0][r];
                }
                pv[0] = d[1] * e2[2] - d[2] * e2[1];
                pv[1] = d[2] * e2[0] - d[0] * e2[2];
                pv[2] = d[0] * e2[1] - d[1] * e2[0];
                const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
                if (std::fabs(det) < 1e-30) continue;
                const double u = (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]) / det;
                if (u < 0 || u > 1) continue;
                q[0] = s[1] * e1[2] - s[2] * e1[1];
                q[1] = s[2] * e1[0] - s[0] * e1[2];
                q[2] = s[0] * e1[1] - s[1] * e1[0];
                const double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
                if (v < 0 || u + v > 1) continue;
                const double tt = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
                const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                const double wd = tt * dl;
                if (tt > 0 && wd < best) {
                    best = wd;
                    h.instance = static_cast<int32_t>(ii);
                    h.triangle = static_cast<uint32_t>(t);
                    h.u = static_cast<float>(u);
                    h.v = static_cast<float>(v);
                    h.distance = static_cast<float>(wd);
                }
            }
        }
        hits[i] = h;
    }
    return MT_OK;
}

void orc_brdf_eval(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = in + i * 14;
        const V3 L{p[0], p[1], p[2]}, V{p[3], p[4], p[5]}, N{p[6], p[7], p[8]};
        const V3 base{p[9], p[10], p[11]};
        const V3 f = evalBrdf(L, V, N, base, p[12], p[13]);
        float* o = out + i * 5;
        o[0] = f.x;
        o[1] = f.y;
        o[2] = f.z;
        o[3] = diffusePdf(L, N);
        o[4] = specularPdf(L, V, N, p[12]);
    }
}

void orc_sample_bounce(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* p = in + i * 16;
        const V3 V{p[0], p[1], p[2]}, N{p[3], p[4], p[5]}, base{p[6], p[7], p[8]};
        const float rough = p[9], metal = p[10];
        const float* u = p + 11;
        const float pr = specProbability(V, N, base, metal);
        ShadeOut so;
        if (u[0] < pr) {
            so.spec = true;
            if (!ggxDirection(u[1], u[2], -V, N, rough, so.dir)) so.dir = cosineHemisphere(u[3], u[4], N);
        } else {
            so.spec = false;
            so.dir = cosineHemisphere(u[1], u[2], N);
        }
        so.pdf = (1.0f - pr) * diffusePdf(so.dir, N) + pr * specularPdf(so.dir, V, N, rough);
        const V3 f = evalBrdf(so.dir, V, N, base, rough, metal);
        const float nl = gmax(dot(N, so.dir), 0.0f);
        so.energy = (so.pdf > 0.0f) ? f * (nl / so.pdf) : v3(0.0f);
        float* o = out + i * 8;
        o[0] = so.dir.x;
        o[1] = so.dir.y;
        o[2] = so.dir.z;
        o[3] = so.pdf;
        o[4] = so.energy.x;
        o[5] = so.energy.y;
        o[6] = so.energy.z;
        o[7] = so.spec ? 1.0f : 0.0f;
    }
}

int orc_env_sample(const void* sp, uint32_t n, const float* dirs, int32_t apply_rotation,
                   float rotation_deg, float* rgb) {
    const Scene& S = *static_cast<const Scene*>(sp);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 c = sampleEnv(S, V3{dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]},
                               apply_rotation != 0, rotation_deg);
        rgb[i * 3] = c.x;
        rgb[i * 3 + 1] = c.y;
        rgb[i * 3 + 2] = c.z;
    }
    return MT_OK;
}

void orc_math(int32_t op, uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float x = in[2 * i], y = in[2 * i + 1];
        switch (op) {
            case MT_MATH_SIN: out[i] = em::sin(x); break;
            case MT_MATH_COS: out[i] = em::cos(x); break;
            case MT_MATH_ATAN2: out[i] = em::atan2(x, y); break;
            case MT_MATH_ACOS: out[i] = em::acos(x); break;
            case MT_MATH_POW5: out[i] = em::pow5(x); break;
            default: out[i] = 0.0f;
        }
    }
}

void orc_bounce_terms(uint32_t n, const float* in, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + 22ull * i;
        const V3 view{q[0], q[1], q[2]}, nrm{q[3], q[4], q[5]};
        const V3 base{q[6], q[7], q[8]};
        const float rough = q[9], metal = q[10];
        // the bounce's uniforms in the sampler's dimensions (lobe 0, direction 1-2, fallback 3-4)
        Rng rng;
        rng.fixed = q + 16;
        const ShadeOut so = sampleBounce(view, nrm, base, rough, metal, rng);
        const D3 bd = d3(base, V3{q[11], q[12], q[13]});
        const Dv rd{rough, q[14]}, md{metal, q[15]};
        const D3 f = evalBrdfDv(so.dir, view, nrm, bd, rd, md);
        const float nl = gmax(dot(nrm, so.dir), 0.0f);
        const Dv ps = specProbabilityDv(view, nrm, bd, md);
        float score;
        if (so.spec) {
            const Dv dh = distGGXDv(so.hcos, rd);
            score = ps.d / ps.v + (dh.v > 0.0f ? dh.d / dh.v : 0.0f);
        } else {
            score = -ps.d / (1.0f - ps.v);
        }
        const Dv pdf = (1.0f - ps) * diffusePdf(so.dir, nrm) + ps * specularPdfDv(so.dir, view, nrm, rd);
        float* o = out + 13ull * i;
        o[0] = so.dir.x;
        o[1] = so.dir.y;
        o[2] = so.dir.z;
        for (int c = 0; c < 3; ++c) {
            o[3 + c] = get(so.energy, c);                                   // f cos / pdf
            o[6 + c] = so.pdf > 0.0f ? f.c[c].d * (nl / so.pdf) : 0.0f;    // detached tangent
        }
        o[9] = score;
        o[10] = pdf.v > 0.0f ? pdf.d / pdf.v : 0.0f;
        o[11] = ps.v;
        o[12] = so.fallback ? 1.0f : 0.0f;
    }
}

int orc_sample_image(uint32_t w, uint32_t h, const uint8_t* rgb8, uint32_t n, const float* uv,
                     int32_t filter, float* out) {
    if (w == 0 || h == 0 || !rgb8) return MT_E_INVALID;
    Image im;
    im.w = static_cast<int>(w);
    im.h = static_cast<int>(h);
    im.rgb.resize(3ull * w * h);
    for (size_t i = 0; i < im.rgb.size(); ++i) im.rgb[i] = static_cast<float>(rgb8[i]) / 255.0f;
    for (uint32_t i = 0; i < n; ++i) {
        const RGBA c = sampleImage(im, uv[2 * i], uv[2 * i + 1], filter);
        out[4 * i] = c.r;
        out[4 * i + 1] = c.g;
        out[4 * i + 2] = c.b;
        out[4 * i + 3] = c.a;
    }
    return MT_OK;
}

void orc_compute_normal(uint32_t n, const float* surf, const float* tex, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const V3 r = computeNormal(V3{surf[3 * i], surf[3 * i + 1], surf[3 * i + 2]},
                                   V3{tex[3 * i], tex[3 * i + 1], tex[3 * i + 2]});
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

int orc_light_sample(const void* sp, uint32_t n, const float* pos, uint64_t seed, uint32_t base_dim,
                     int32_t* chosen, float* pdf, float* lpos, float* lcol) {
    const Scene& S = *static_cast<const Scene*>(sp);
    const uint32_t sk = seedKey(seed);
    for (uint32_t i = 0; i < n; ++i) {
        Rng rng;
        rng.key = pathKey(sk, i, 0);
        rng.base = base_dim;
        const V3 p{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        float pr = 0.0f;
        const int c = chooseLight(S, p, rng, pr);
        chosen[i] = c;
        if (pdf) pdf[i] = pr;
        V3 lp{0.0f, 0.0f, 0.0f}, lc{0.0f, 0.0f, 0.0f};
        if (c >= 0) lp = lightPoint(S.lights[c], p, rng, lc);
        if (lpos) {
            lpos[3 * i] = lp.x;
            lpos[3 * i + 1] = lp.y;
            lpos[3 * i + 2] = lp.z;
        }
        if (lcol) {
            lcol[3 * i] = lc.x;
            lcol[3 * i + 1] = lc.y;
            lcol[3 * i + 2] = lc.z;
        }
    }
    return MT_OK;
}

void orc_camera_rays(const mt_camera_desc* cam, uint32_t n, const float* uv, float aspect,
                     float* rays) {
    for (uint32_t i = 0; i < n; ++i) {
        const Ray r = cameraRay(*cam, uv[i * 2], uv[i * 2 + 1], aspect);
        const float v[6] = {r.o.x, r.o.y, r.o.z, r.d.x, r.d.y, r.d.z};
        std::memcpy(rays + i * 6, v, sizeof(v));
    }
}

void orc_aces(uint32_t n, const float* rgb, float ev, float* out) {
    const float expo = std::pow(2.0f, -ev);
    for (uint32_t i = 0; i < n; ++i) {
        const V3 c = aces(V3{rgb[i * 3], rgb[i * 3 + 1], rgb[i * 3 + 2]}, expo);
        out[i * 3] = c.x;
        out[i * 3 + 1] = c.y;
        out[i * 3 + 2] = c.z;
    }
}

uint32_t orc_cleanup_fireflies(float* s, uint32_t count, float threshold) {
    std::vector<V3> buf(count);
    for (uint32_t i = 0; i < count; ++i) buf[i] = V3{s[i * 3], s[i * 3 + 1], s[i * 3 + 2]};
    const uint32_t r = cleanupFireflies(buf, threshold);
    for (uint32_t i = 0; i < count; ++i) {
        s[i * 3] = buf[i].x;
        s[i * 3 + 1] = buf[i].y;
        s[i * 3 + 2] = buf[i].z;
    }
    return r;
}

void orc_rng_draws(uint64_t seed, uint32_t n, const uint32_t* pixel, const uint32_t* sample,
                   const uint32_t* dim, float* out) {
    const uint32_t sk = seedKey(seed);
    for (uint32_t i = 0; i < n; ++i) out[i] = drawU(pathKey(sk, pixel[i], sample[i]), dim[i]);
}

int orc_render(const void* sp, const mt_camera_desc* cam, const mt_settings* st,
               const orc_render_opts* opts, float* rgb_out, float* sample_rgb_out, orc_work* work,
               double* seconds) {
    const Scene& S = *static_cast<const Scene*>(sp);
    if (!cam || !st || !rgb_out || st->width == 0 || st->height == 0 || st->samples == 0) {
        g_err = "invalid render arguments";
        return MT_E_INVALID;
    }
    orc_render_opts o{};
    if (opts) o = *opts;
    Tracer tr{S, {*st, std::pow(2.0f, -st->ev_exposure)},
              8u + static_cast<uint32_t>(S.lights.size()), o.faithful_inverse != 0, o.tlas != 0};
    const uint32_t W = st->width, H = st->height, spp = st->samples;
    const uint32_t rb = (o.row_end > o.row_begin) ? o.row_begin : 0;
    const uint32_t re = (o.row_end > o.row_begin) ? std::min(o.row_end, H) : H;
    const uint32_t bs = st->image_block_size ? st->image_block_size : 32;
    struct Block {
        uint32_t x, y, w, h, id;
    };
    std::queue<Block> blocks;  // Tracer::StartRender tile queue (Tracer.cpp:727-737)
    uint32_t id = 0;
    for (uint32_t y = rb; y < re; y += bs)
        for (uint32_t x = 0; x < W; x += bs) blocks.push({x, y, std::min(bs, W - x), std::min(bs, re - y), id++});
    std::mutex mtx;
    std::atomic<uint64_t> rays{0}, samples{0}, hitsN{0}, missN{0};
    std::atomic<uint64_t> cA{0}, cE{0}, cP{0}, cT{0};
    const float aspect = static_cast<float>(W) / static_cast<float>(H);
    const float pxW = 1.0f / static_cast<float>(W), pxH = 1.0f / static_cast<float>(H);
    const uint32_t sk = seedKey(st->seed);
    const bool wantCounters = st->count_work != 0;
    auto worker = [&]() {
        std::vector<PathStep> path;
        std::vector<V3> sbuf;
        Counters c;
        uint64_t lr = 0, ls = 0, lh = 0, lm = 0;
        while (true) {
            Block b;
            {
                std::scoped_lock lock(mtx);
                if (blocks.empty()) break;
                b = blocks.front();
                blocks.pop();
            }
            std::mt19937 mt(static_cast<uint32_t>(st->seed) * 2654435761u + b.id);
            Rng rng;
            if (o.rng == ORC_RNG_MT19937) rng.mt = &mt;
            for (uint32_t y = b.y; y < b.y + b.h; ++y) {
                for (uint32_t x = b.x; x < b.x + b.w; ++x) {
                    const uint32_t pix = y * W + x;
                    sbuf.clear();
                    const float bu = static_cast<float>(x) * pxW;
                    const float bv = static_cast<float>(y) * pxH;
                    for (uint32_t s = 0; s < spp; ++s) {
                        rng.key = pathKey(sk, pix, s);
                        rng.base = 0;
                        const float ju = bu + rng.uRange(0, 0.0f, pxW);
                        const float jv = bv + rng.uRange(1, 0.0f, pxH);
                        const Ray r = cameraRay(*cam, ju, jv, aspect);
                        LightOut parts;
                        V3 L = tr.forward(r, rng, path, wantCounters ? &c : nullptr, lr, lh, lm, &parts);
                        ls++;
                        V3 col = (st->output == MT_OUTPUT_LINEAR) ? L : aces(L, tr.cfg.expo);
                        if (st->debug_mode) col = debugColor(st->debug_mode, path, parts, r, col, tr.cfg.expo);
                        sbuf.push_back(col);
                        if (sample_rgb_out) {
                            float* q = sample_rgb_out + (static_cast<size_t>(pix) * spp + s) * 3;
                            q[0] = col.x;
                            q[1] = col.y;
                            q[2] = col.z;
                        }
                    }
                    uint32_t cnt = spp;
                    if (st->firefly_enable) cnt = cleanupFireflies(sbuf, st->firefly_threshold, st->debug_mode);
                    V3 acc = v3(0.0f);
                    for (const V3& c2 : sbuf) acc = acc + c2;
                    const V3 col = acc / static_cast<float>(cnt);
                    rgb_out[pix * 3] = col.x;
                    rgb_out[pix * 3 + 1] = col.y;
                    rgb_out[pix * 3 + 2] = col.z;
                }
            }
        }
        rays += lr;
        samples += ls;
        hitsN += lh;
        missN += lm;
        cA += c.aabb;
        cE += c.enter;
        cP += c.pops;
        cT += c.tris;
    };
    int nt = o.threads > 0 ? o.threads : static_cast<int>(std::thread::hardware_concurrency());
    if (nt < 1) nt = 1;
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
    const auto t1 = std::chrono::steady_clock::now();
    if (st->debug_mode) debugPost(st->debug_mode, rgb_out, static_cast<size_t>(W) * H);
    if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
    if (work) {
        work->rays = rays;
        work->samples = samples;
        work->inst_aabb_tests = cA;
        work->inst_entered = cE;
        work->node_pops = cP;
        work->tri_tests = cT;
        work->hits_shaded = hitsN;
        work->misses = missN;
    }
    return MT_OK;
}

void orc_brdf_tangent(uint32_t n, const float* in, const float* tin, float* out) {
    for (uint32_t i = 0; i < n; ++i) {
        const float* q = in + i * 14;
        const float* t = tin + i * 5;
        const D3 base = d3(V3{q[9], q[10], q[11]}, V3{t[0], t[1], t[2]});
        const D3 f = evalBrdfDv(V3{q[0], q[1], q[2]}, V3{q[3], q[4], q[5]}, V3{q[6], q[7], q[8]},
                                base, Dv{q[12], t[3]}, Dv{q[13], t[4]});
        float* o = out + i * 6;
        for (int c = 0; c < 3; ++c) {
            o[c] = f.c[c].v;
            o[3 + c] = f.c[c].d;
        }
    }
}

int orc_render_diff(const void* sp, const mt_camera_desc* cam, const mt_settings* st,
                    const orc_diff_params* params, const orc_diff_params* tangent_dir,
                    int32_t sampling_terms, int32_t threads, float* image, float* dimage) {
    const Scene& S = *static_cast<const Scene*>(sp);
    if (!cam || !st || !image || st->width == 0 || st->height == 0 || st->samples == 0) {
        g_err = "invalid render arguments";
        return MT_E_INVALID;
    }
    const uint32_t nm = static_cast<uint32_t>(S.materials.size());
    const DiffSet P = makeDiffSet(params, nm);
    DiffSet T = makeDiffSet(tangent_dir, nm);
    T.texFirst = P.texFirst;
    T.texCount = P.texCount;
    if (T.hasTex && (!P.hasTex || T.tex.w != P.tex.w || T.tex.h != P.tex.h)) {
        g_err = "tangent texture does not match the parameter texture";
        return MT_E_INVALID;
    }
    const DiffTracer tr{S, *st, 8u + static_cast<uint32_t>(S.lights.size()), P, T, sampling_terms != 0};
    const uint32_t W = st->width, H = st->height, spp = st->samples;
    const float aspect = static_cast<float>(W) / static_cast<float>(H);
    const float pxW = 1.0f / static_cast<float>(W), pxH = 1.0f / static_cast<float>(H);
    const uint32_t sk = seedKey(st->seed);
    std::atomic<uint32_t> nextRow{0};
    auto worker = [&]() {
        std::vector<DiffTracer::StepD> path;
        Rng rng;
        while (true) {
            const uint32_t y = nextRow++;
            if (y >= H) break;
            for (uint32_t x = 0; x < W; ++x) {
                const uint32_t pix = y * W + x;
                V3 acc = v3(0.0f), dacc = v3(0.0f);
                for (uint32_t s = 0; s < spp; ++s) {
                    rng.key = pathKey(sk, pix, s);
                    rng.base = 0;
                    const float ju = static_cast<float>(x) * pxW + rng.uRange(0, 0.0f, pxW);
                    const float jv = static_cast<float>(y) * pxH + rng.uRange(1, 0.0f, pxH);
                    const D3 L = tr.forward(cameraRay(*cam, ju, jv, aspect), rng, path);
                    acc = acc + V3{L.c[0].v, L.c[1].v, L.c[2].v};
                    dacc = dacc + tangent(L);
                }
                const float inv = 1.0f / static_cast<float>(spp);
                image[pix * 3] = acc.x * inv;
                image[pix * 3 + 1] = acc.y * inv;
                image[pix * 3 + 2] = acc.z * inv;
                if (dimage) {
                    dimage[pix * 3] = dacc.x * inv;
                    dimage[pix * 3 + 1] = dacc.y * inv;
                    dimage[pix * 3 + 2] = dacc.z * inv;
                }
            }
        }
    };
    int nt = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    if (nt < 1) nt = 1;
    std::vector<std::thread> th;
    for (int i = 0; i < nt; ++i) th.emplace_back(worker);
    for (auto& t : th) t.join();
    return MT_OK;
}

}  // extern "C"
