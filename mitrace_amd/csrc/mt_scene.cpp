// mt_scene.cpp -- host side of mt_scene_create: the reference's per-mesh BVH build and the
// flattening of the Scene into the HBM layout described in mt_internal.hpp.
//
// The BVH rules follow src/Scene/BVH.cpp:29-272: split plane at the vertex-mean centroid,
// all three axes tried, score = (|L|+|R|)/|N| (FLT_MAX when one side keeps everything),
// straddling triangles go to both children, child boxes are the parent box clipped at the
// plane, children of a node are an adjacent pair in DFS order, node 1 is an unused slot.
// The tree is built with per-triangle bounds and centroid sums computed once; the float
// operations and their order are those of BVH.cpp, so the nodes are bit-identical.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>

#include "mt_internal.hpp"

namespace mt {

namespace {

inline float fmin_g(float a, float b) { return (b < a) ? b : a; }  // glm::min scalar form
inline float fmax_g(float a, float b) { return (a < b) ? b : a; }  // glm::max scalar form

struct TriInfo {
    float lo[3], hi[3];  // min(v0, min(v1, v2)) and max(...) (BVH.cpp:75-76)
    float sum[3];        // (v0 + v1) + v2 (BVH.cpp:59)
};

struct TmpNode {
    float mn[3], mx[3];
    int32_t ca = -1, cb = -1;      // children in the temporary array (-1 = leaf)
    std::vector<uint32_t> tris;    // leaf triangles, split order preserved
};

class Builder {
   public:
    Builder(const std::vector<TriInfo>& info, uint32_t max_tri) : info_(info), max_tri_(max_tri) {}

    // BVHTreeNode::Subdivide (BVH.cpp:117-195); returns the index of the node in nodes_
    int32_t subdivide(std::vector<uint32_t>&& tris, const float* mn, const float* mx,
                      uint32_t depth_left) {
        const int32_t id = static_cast<int32_t>(nodes_.size());
        nodes_.emplace_back();
        std::memcpy(nodes_[id].mn, mn, sizeof(float) * 3);
        std::memcpy(nodes_[id].mx, mx, sizeof(float) * 3);
        if (tris.size() <= static_cast<size_t>(max_tri_)) {
            nodes_[id].tris = std::move(tris);
            return id;
        }
        float sum[3] = {0.0f, 0.0f, 0.0f};
        for (uint32_t t : tris)
            for (int k = 0; k < 3; ++k) sum[k] = sum[k] + info_[t].sum[k];
        const float denom = static_cast<float>(tris.size() * 3);
        const float com[3] = {sum[0] / denom, sum[1] / denom, sum[2] / denom};
        // count both sides of every axis first; only the chosen axis is materialised
        size_t nl[3] = {0, 0, 0}, nr[3] = {0, 0, 0};
        for (uint32_t t : tris) {
            const TriInfo& ti = info_[t];
            for (int k = 0; k < 3; ++k) {
                nl[k] += ti.lo[k] < com[k];
                nr[k] += ti.hi[k] >= com[k];
            }
        }
        float score[3];
        for (int k = 0; k < 3; ++k) {
            if (nl[k] == tris.size() || nr[k] == tris.size())
                score[k] = FLT_MAX;
            else
                score[k] = static_cast<float>(nl[k] + nr[k]) / static_cast<float>(tris.size());
        }
        if (score[0] == FLT_MAX && score[1] == FLT_MAX && score[2] == FLT_MAX) {
            nodes_[id].tris = std::move(tris);
            return id;
        }
        const int ax = (score[0] < score[1] && score[0] < score[2]) ? 0 : (score[1] < score[2] ? 1 : 2);
        std::vector<uint32_t> left, right;
        left.reserve(nl[ax]);
        right.reserve(nr[ax]);
        for (uint32_t t : tris) {
            if (info_[t].lo[ax] < com[ax]) left.push_back(t);
            if (info_[t].hi[ax] >= com[ax]) right.push_back(t);
        }
        std::vector<uint32_t>().swap(tris);
        float amx[3] = {mx[0], mx[1], mx[2]};
        float bmn[3] = {mn[0], mn[1], mn[2]};
        amx[ax] = com[ax];
        bmn[ax] = com[ax];
        int32_t a, b;
        if (depth_left > 0) {
            a = subdivide(std::move(left), mn, amx, depth_left - 1);
            b = subdivide(std::move(right), bmn, mx, depth_left - 1);
        } else {
            a = leaf(std::move(left), mn, amx);
            b = leaf(std::move(right), bmn, mx);
        }
        nodes_[id].ca = a;
        nodes_[id].cb = b;
        return id;
    }

    int32_t leaf(std::vector<uint32_t>&& tris, const float* mn, const float* mx) {
        const int32_t id = static_cast<int32_t>(nodes_.size());
        nodes_.emplace_back();
        std::memcpy(nodes_[id].mn, mn, sizeof(float) * 3);
        std::memcpy(nodes_[id].mx, mx, sizeof(float) * 3);
        nodes_[id].tris = std::move(tris);
        return id;
    }

    // BVH::Flatten / FlattenRecursive (BVH.cpp:223-272)
    void flatten(int32_t root, HostBvh& out) {
        out.nodes.clear();
        out.indices.clear();
        out.nodes.resize(2);
        std::memset(out.nodes.data(), 0, sizeof(HostNode) * 2);
        set_box(out.nodes[0], nodes_[root]);
        uint32_t depth = 1;
        if (nodes_[root].ca < 0) {
            // a root that could not be split (the reference would dereference null children)
            set_leaf(out, 0, nodes_[root]);
        } else {
            auto c = pair(out, nodes_[root].ca, nodes_[root].cb, 2, depth);
            out.nodes[0].a = c.first;
            out.nodes[0].b = c.second;
        }
        out.depth = depth;
    }

   private:
    void set_box(HostNode& h, const TmpNode& t) {
        std::memcpy(h.mn, t.mn, sizeof(h.mn));
        std::memcpy(h.mx, t.mx, sizeof(h.mx));
    }
    void set_leaf(HostBvh& out, uint32_t slot, TmpNode& t) {
        const uint32_t first = static_cast<uint32_t>(out.indices.size());
        out.indices.insert(out.indices.end(), t.tris.begin(), t.tris.end());
        out.nodes[slot].a = first | kLeafBit;
        out.nodes[slot].b = static_cast<uint32_t>(t.tris.size());
        std::vector<uint32_t>().swap(t.tris);
    }
    std::pair<uint32_t, uint32_t> pair(HostBvh& out, int32_t a, int32_t b, uint32_t level,
                                       uint32_t& depth) {
        depth = std::max(depth, level);
        const uint32_t ia = static_cast<uint32_t>(out.nodes.size()), ib = ia + 1;
        out.nodes.resize(out.nodes.size() + 2);
        set_box(out.nodes[ia], nodes_[a]);
        set_box(out.nodes[ib], nodes_[b]);
        for (int side = 0; side < 2; ++side) {
            const int32_t n = side ? b : a;
            const uint32_t slot = side ? ib : ia;
            if (nodes_[n].ca < 0) {
                set_leaf(out, slot, nodes_[n]);
            } else {
                auto c = pair(out, nodes_[n].ca, nodes_[n].cb, level + 1, depth);
                out.nodes[slot].a = c.first;
                out.nodes[slot].b = c.second;
            }
        }
        return {ia, ib};
    }

    const std::vector<TriInfo>& info_;
    uint32_t max_tri_;
    std::vector<TmpNode> nodes_;
};

}  // namespace

void build_bvh(const float* P, uint32_t tri_count, uint32_t max_triangles, uint32_t max_depth,
               HostBvh& out) {
    std::vector<TriInfo> info(tri_count);
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t t = 0; t < tri_count; ++t) {
        const float* v0 = P + t * 9;
        const float* v1 = v0 + 3;
        const float* v2 = v0 + 6;
        TriInfo& ti = info[t];
        for (int k = 0; k < 3; ++k) {
            ti.lo[k] = fmin_g(v0[k], fmin_g(v1[k], v2[k]));
            ti.hi[k] = fmax_g(v0[k], fmax_g(v1[k], v2[k]));
            ti.sum[k] = (v0[k] + v1[k]) + v2[k];
            // BVHTreeNode::ComputeAABB of the root (BVH.cpp:29-50)
            mn[k] = fmin_g(mn[k], ti.lo[k]);
            mx[k] = fmax_g(mx[k], ti.hi[k]);
        }
    }
    std::vector<uint32_t> all(tri_count);
    for (uint32_t t = 0; t < tri_count; ++t) all[t] = t;
    Builder b(info, max_triangles);
    // BVHTree passes maxDepth - 1 to the root's Subdivide (BVH.cpp:204-212), unsigned
    const int32_t root = b.subdivide(std::move(all), mn, mx, max_depth - 1u);
    b.flatten(root, out);
}

void mul4(const float* m, const float* v, float* o) {
    for (int r = 0; r < 4; ++r)
        o[r] = (m[0 * 4 + r] * v[0] + m[1 * 4 + r] * v[1]) + (m[2 * 4 + r] * v[2] + m[3 * 4 + r] * v[3]);
}

// 4x4 inverse by 2x2 minors of the top and bottom row pairs (Laplace expansion).
void invert4(const float* m, float* out) {
    auto A = [&](int r, int c) { return m[c * 4 + r]; };
    const float s0 = A(0, 0) * A(1, 1) - A(1, 0) * A(0, 1);
    const float s1 = A(0, 0) * A(1, 2) - A(1, 0) * A(0, 2);
    const float s2 = A(0, 0) * A(1, 3) - A(1, 0) * A(0, 3);
    const float s3 = A(0, 1) * A(1, 2) - A(1, 1) * A(0, 2);
    const float s4 = A(0, 1) * A(1, 3) - A(1, 1) * A(0, 3);
    const float s5 = A(0, 2) * A(1, 3) - A(1, 2) * A(0, 3);
    const float c5 = A(2, 2) * A(3, 3) - A(3, 2) * A(2, 3);
    const float c4 = A(2, 1) * A(3, 3) - A(3, 1) * A(2, 3);
    const float c3 = A(2, 1) * A(3, 2) - A(3, 1) * A(2, 2);
    const float c2 = A(2, 0) * A(3, 3) - A(3, 0) * A(2, 3);
    const float c1 = A(2, 0) * A(3, 2) - A(3, 0) * A(2, 2);
    const float c0 = A(2, 0) * A(3, 1) - A(3, 0) * A(2, 1);
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    const float id = 1.0f / det;
    float B[4][4];
    B[0][0] = (A(1, 1) * c5 - A(1, 2) * c4 + A(1, 3) * c3) * id;
    B[0][1] = (-A(0, 1) * c5 + A(0, 2) * c4 - A(0, 3) * c3) * id;
    B[0][2] = (A(3, 1) * s5 - A(3, 2) * s4 + A(3, 3) * s3) * id;
    B[0][3] = (-A(2, 1) * s5 + A(2, 2) * s4 - A(2, 3) * s3) * id;
    B[1][0] = (-A(1, 0) * c5 + A(1, 2) * c2 - A(1, 3) * c1) * id;
    B[1][1] = (A(0, 0) * c5 - A(0, 2) * c2 + A(0, 3) * c1) * id;
    B[1][2] = (-A(3, 0) * s5 + A(3, 2) * s2 - A(3, 3) * s1) * id;
    B[1][3] = (A(2, 0) * s5 - A(2, 2) * s2 + A(2, 3) * s1) * id;
    B[2][0] = (A(1, 0) * c4 - A(1, 1) * c2 + A(1, 3) * c0) * id;
    B[2][1] = (-A(0, 0) * c4 + A(0, 1) * c2 - A(0, 3) * c0) * id;
    B[2][2] = (A(3, 0) * s4 - A(3, 1) * s2 + A(3, 3) * s0) * id;
    B[2][3] = (-A(2, 0) * s4 + A(2, 1) * s2 - A(2, 3) * s0) * id;
    B[3][0] = (-A(1, 0) * c3 + A(1, 1) * c1 - A(1, 2) * c0) * id;
    B[3][1] = (A(0, 0) * c3 - A(0, 1) * c1 + A(0, 2) * c0) * id;
    B[3][2] = (-A(3, 0) * s3 + A(3, 1) * s1 - A(3, 2) * s0) * id;
    B[3][3] = (A(2, 0) * s3 - A(2, 1) * s1 + A(2, 2) * s0) * id;
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) out[c * 4 + r] = B[r][c];
}


namespace {

// Binned SAH tree over axis-aligned boxes (6 floats each: min xyz, max xyz).  All three axes
// are binned in one pass over the node's items (16 bins on the box centres); a node with at
// most leaf_max items is a leaf.  The first levels hand one child to a new thread, so large
// meshes build on several cores; every thread owns a disjoint range of ids.
struct SahTree {
    struct Node {
        float mn[3], mx[3];
        int32_t a = -1, b = -1;  // children, -1 = leaf
        uint32_t first = 0, count = 0;
    };
    static constexpr int kBins = 16;
    const float* boxes = nullptr;
    std::vector<float> cent;      // box centres
    std::vector<uint32_t> ids;    // item order: a leaf holds ids[first, first + count)
    std::vector<Node> nodes;
    uint32_t leaf_max = 4;
    uint32_t depth = 0;           // levels of the tree

    static float area(const float* mn, const float* mx) {
        const float dx = mx[0] - mn[0], dy = mx[1] - mn[1], dz = mx[2] - mn[2];
        return dx * dy + dy * dz + dz * dx;
    }

    // best split of ids[lo, hi): returns the split position (lo < mid < hi)
    uint32_t split(uint32_t lo, uint32_t hi, float* mn, float* mx) {
        float cmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, cmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int k = 0; k < 3; ++k) {
            mn[k] = FLT_MAX;
            mx[k] = -FLT_MAX;
        }
        for (uint32_t i = lo; i < hi; ++i) {
            const float* bx = boxes + 6ull * ids[i];
            const float* c = &cent[3ull * ids[i]];
            for (int k = 0; k < 3; ++k) {
                mn[k] = std::min(mn[k], bx[k]);
                mx[k] = std::max(mx[k], bx[3 + k]);
                cmn[k] = std::min(cmn[k], c[k]);
                cmx[k] = std::max(cmx[k], c[k]);
            }
        }
        const uint32_t n = hi - lo;
        if (n <= leaf_max) return lo;
        float scale[3];
        for (int k = 0; k < 3; ++k) {
            const float ext = cmx[k] - cmn[k];
            scale[k] = ext > 0.0f ? kBins / ext : 0.0f;
        }
        float bmn[3][kBins][3], bmx[3][kBins][3];
        uint32_t cnt[3][kBins] = {};
        for (int ax = 0; ax < 3; ++ax)
            for (int b = 0; b < kBins; ++b)
                for (int k = 0; k < 3; ++k) {
                    bmn[ax][b][k] = FLT_MAX;
                    bmx[ax][b][k] = -FLT_MAX;
                }
        for (uint32_t i = lo; i < hi; ++i) {
            const float* bx = boxes + 6ull * ids[i];
            const float* c = &cent[3ull * ids[i]];
            for (int ax = 0; ax < 3; ++ax) {
                const int b = std::min(kBins - 1, static_cast<int>((c[ax] - cmn[ax]) * scale[ax]));
                cnt[ax][b]++;
                for (int k = 0; k < 3; ++k) {
                    bmn[ax][b][k] = std::min(bmn[ax][b][k], bx[k]);
                    bmx[ax][b][k] = std::max(bmx[ax][b][k], bx[3 + k]);
                }
            }
        }
        float best = FLT_MAX;
        int best_ax = -1, best_bin = 0;
        for (int ax = 0; ax < 3; ++ax) {
            if (scale[ax] == 0.0f) continue;
            // suffix areas and counts, then a prefix sweep
            float ra[kBins];
            uint32_t rc[kBins];
            float smn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, smx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t sc = 0;
            for (int b = kBins - 1; b > 0; --b) {
                for (int k = 0; k < 3; ++k) {
                    smn[k] = std::min(smn[k], bmn[ax][b][k]);
                    smx[k] = std::max(smx[k], bmx[ax][b][k]);
                }
                sc += cnt[ax][b];
                ra[b] = sc ? area(smn, smx) : 0.0f;
                rc[b] = sc;
            }
            float pmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, pmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
            uint32_t pc = 0;
            for (int b = 1; b < kBins; ++b) {
                for (int k = 0; k < 3; ++k) {
                    pmn[k] = std::min(pmn[k], bmn[ax][b - 1][k]);
                    pmx[k] = std::max(pmx[k], bmx[ax][b - 1][k]);
                }
                pc += cnt[ax][b - 1];
                if (!pc || !rc[b]) continue;
                const float cost = pc * area(pmn, pmx) + rc[b] * ra[b];
                if (cost < best) {
                    best = cost;
                    best_ax = ax;
                    best_bin = b;
                }
            }
        }
        uint32_t mid = lo + n / 2;  // coincident centres: split the list
        if (best_ax >= 0) {
            const int ax = best_ax;
            auto it = std::partition(ids.begin() + lo, ids.begin() + hi, [&](uint32_t id) {
                const int b = std::min(kBins - 1, static_cast<int>((cent[3ull * id + ax] - cmn[ax]) * scale[ax]));
                return b < best_bin;
            });
            const uint32_t m = static_cast<uint32_t>(it - ids.begin());
            if (m > lo && m < hi) mid = m;
        }
        return mid;
    }

    // builds the subtree of ids[lo, hi) into out; returns its root index in out
    int32_t grow(std::vector<Node>& out, uint32_t lo, uint32_t hi, uint32_t level, uint32_t& dep,
                 int par_levels) {
        dep = std::max(dep, level);
        const int32_t id = static_cast<int32_t>(out.size());
        out.emplace_back();
        float mn[3], mx[3];
        const uint32_t mid = split(lo, hi, mn, mx);
        std::memcpy(out[id].mn, mn, sizeof(mn));
        std::memcpy(out[id].mx, mx, sizeof(mx));
        if (mid == lo) {
            out[id].first = lo;
            out[id].count = hi - lo;
            return id;
        }
        int32_t a, b;
        if (par_levels > 0 && hi - lo > 65536) {
            std::vector<Node> sub;
            uint32_t dep_a = 0;
            int32_t ra = 0;
            std::thread th([&] { ra = grow(sub, lo, mid, level + 1, dep_a, par_levels - 1); });
            b = grow(out, mid, hi, level + 1, dep, par_levels - 1);
            th.join();
            dep = std::max(dep, dep_a);
            const int32_t off = static_cast<int32_t>(out.size());
            for (Node& nd : sub) {
                if (nd.a >= 0) {
                    nd.a += off;
                    nd.b += off;
                }
                out.push_back(nd);
            }
            a = ra + off;
        } else {
            a = grow(out, lo, mid, level + 1, dep, 0);
            b = grow(out, mid, hi, level + 1, dep, 0);
        }
        out[id].a = a;
        out[id].b = b;
        return id;
    }

    int32_t build(const float* bx, uint32_t n, uint32_t lmax) {
        boxes = bx;
        leaf_max = lmax;
        cent.resize(3ull * n);
        ids.resize(n);
        for (uint32_t i = 0; i < n; ++i) {
            ids[i] = i;
            for (int k = 0; k < 3; ++k) cent[3ull * i + k] = 0.5f * (bx[6ull * i + k] + bx[6ull * i + 3 + k]);
        }
        nodes.clear();
        nodes.reserve(n ? 2ull * n / std::max(1u, lmax) + 2 : 1);
        depth = 1;
        const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        int par = 0;
        while ((1u << par) < hw) ++par;
        return grow(nodes, 0, n, 1, depth, n > 65536 ? par : 0);
    }
};

}  // namespace

uint32_t build_tlas(const std::vector<float>& boxes, std::vector<float4>& pairs,
                    std::vector<uint32_t>& indices, float root_min[3], float root_max[3],
                    uint32_t& root_word) {
    SahTree tb;
    const uint32_t n = static_cast<uint32_t>(boxes.size() / 6);
    const int32_t root = tb.build(boxes.data(), n, 4);
    indices = tb.ids;
    // flatten: every inner node's children become one pair record, pair index = word
    pairs.clear();
    auto word_of = [&](int32_t id, auto& self) -> uint32_t {
        const SahTree::Node& nd = tb.nodes[id];
        if (nd.a < 0) return kLeafBit | (nd.count << 24) | nd.first;
        const uint32_t q = static_cast<uint32_t>(pairs.size() / 4);
        pairs.resize(pairs.size() + 4);
        const uint32_t wa = self(nd.a, self), wb = self(nd.b, self);
        const SahTree::Node& A = tb.nodes[nd.a];
        const SahTree::Node& B = tb.nodes[nd.b];
        float4* r = &pairs[4 * q];
        r[0] = make_float4(A.mn[0], A.mn[1], A.mn[2], 0.0f);
        r[1] = make_float4(A.mx[0], A.mx[1], A.mx[2], 0.0f);
        r[2] = make_float4(B.mn[0], B.mn[1], B.mn[2], 0.0f);
        r[3] = make_float4(B.mx[0], B.mx[1], B.mx[2], 0.0f);
        std::memcpy(&r[0].w, &wa, 4);
        std::memcpy(&r[1].w, &wb, 4);
        return q;
    };
    root_word = word_of(root, word_of);
    std::memcpy(root_min, tb.nodes[root].mn, sizeof(float) * 3);
    std::memcpy(root_max, tb.nodes[root].mx, sizeof(float) * 3);
    return tb.depth;
}

uint32_t build_fast_bvh(const float* P, uint32_t tri_count, std::vector<float4>& pairs,
                        std::vector<float4>& prims, DevMesh& dm) {
    dm.pair_base = static_cast<uint32_t>(pairs.size() / 4);
    dm.prim_base = static_cast<uint32_t>(prims.size() / 3);
    dm.prim_count = tri_count;
    auto emit = [&](uint32_t t, uint32_t leaf, bool last) {
        const float* v = P + static_cast<uint64_t>(t) * 9;
        const uint32_t lf = last ? kPrimLast : 0u;
        float4 a = make_float4(v[0], v[1], v[2], 0.0f), b = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], 0.0f),
               c = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], 0.0f);
        std::memcpy(&a.w, &t, 4);
        std::memcpy(&b.w, &leaf, 4);
        std::memcpy(&c.w, &lf, 4);
        prims.push_back(a);
        prims.push_back(b);
        prims.push_back(c);
    };
    if (tri_count <= 4) {  // one leaf: a linear scan
        dm.root_word = 0;
        for (uint32_t t = 0; t < tri_count; ++t) emit(t, 0, t + 1 == tri_count);
        return 1;
    }
    std::vector<float> boxes(6ull * tri_count);
    for (uint32_t t = 0; t < tri_count; ++t) {
        const float* v = P + static_cast<uint64_t>(t) * 9;
        for (int k = 0; k < 3; ++k) {
            boxes[6ull * t + k] = std::min(v[k], std::min(v[3 + k], v[6 + k]));
            boxes[6ull * t + 3 + k] = std::max(v[k], std::max(v[3 + k], v[6 + k]));
        }
    }
    SahTree tb;
    const int32_t root = tb.build(boxes.data(), tri_count, 4);
    // DFS flatten: inner node -> its children's pair record (mesh-local index), leaf -> first
    // prim slot (mesh-local) | kLeafBit, its triangles contiguous with the last one flagged
    uint32_t slot = 0;
    auto word_of = [&](int32_t id, auto& self) -> uint32_t {
        const SahTree::Node& nd = tb.nodes[id];
        if (nd.a < 0) {
            const uint32_t w = kLeafBit | slot;
            for (uint32_t k = 0; k < nd.count; ++k)
                emit(tb.ids[nd.first + k], static_cast<uint32_t>(id), k + 1 == nd.count);
            slot += nd.count;
            return w;
        }
        const size_t q = pairs.size() / 4;
        pairs.resize(pairs.size() + 4);
        const uint32_t wa = self(nd.a, self), wb = self(nd.b, self);
        const SahTree::Node& A = tb.nodes[nd.a];
        const SahTree::Node& B = tb.nodes[nd.b];
        float4* r = &pairs[4 * q];
        r[0] = make_float4(A.mn[0], A.mn[1], A.mn[2], 0.0f);
        r[1] = make_float4(A.mx[0], A.mx[1], A.mx[2], 0.0f);
        r[2] = make_float4(B.mn[0], B.mn[1], B.mn[2], 0.0f);
        r[3] = make_float4(B.mx[0], B.mx[1], B.mx[2], 0.0f);
        std::memcpy(&r[0].w, &wa, 4);
        std::memcpy(&r[1].w, &wb, 4);
        return static_cast<uint32_t>(q - dm.pair_base);
    };
    dm.root_word = word_of(root, word_of);
    dm.flags |= kMeshHasBvh;
    for (int k = 0; k < 3; ++k) {
        dm.root_min[k] = tb.nodes[root].mn[k];
        dm.root_max[k] = tb.nodes[root].mx[k];
    }
    return tb.depth;
}

namespace {

template <typename T>
int dev_upload(mt_scene* s, const std::vector<T>& host, const T** out) {
    *out = nullptr;
    if (host.empty()) return MT_OK;
    void* p = nullptr;
    const size_t bytes = host.size() * sizeof(T);
    if (device_malloc(&p, bytes)) return fail(MT_E_NOMEM, "device allocation failed in scene upload");
    s->allocations.push_back(p);
    if (hipMemcpy(p, host.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return fail(MT_E_HIP, "hipMemcpy failed in scene upload");
    *out = static_cast<const T*>(p);
    return MT_OK;
}

}  // namespace

int upload_scene(const mt_scene_desc* d, mt_scene* s) {
    if (!d) return fail(MT_E_INVALID, "null scene description");
    if (d->mesh_count && !d->meshes) return fail(MT_E_INVALID, "meshes is NULL");
    if (d->instance_count && !d->instances) return fail(MT_E_INVALID, "instances is NULL");
    if (d->material_count && !d->materials) return fail(MT_E_INVALID, "materials is NULL");
    if (d->light_count && !d->lights) return fail(MT_E_INVALID, "lights is NULL");
    if (d->instance_count >= (1u << 24)) return fail(MT_E_UNSUPPORTED, "more than 2^24 instances");
    if (d->image_count && !d->images) return fail(MT_E_INVALID, "images is NULL");
    s->mesh_count = d->mesh_count;
    s->instance_count = d->instance_count;
    s->light_count = d->light_count;

    std::vector<DevMesh> meshes(d->mesh_count);
    std::vector<DevMesh> fast_meshes(d->bvh.fast ? d->mesh_count : 0u);
    uint64_t fast_pairs = 0, fast_slots = 0;
    std::vector<float4> pairs, prims;
    std::vector<uint8_t> last_flags;
    // meshes whose BVH was built on the device: their pair records and leaf slots stay there and
    // go after the host-built records in the scene's arrays
    struct DevPart {
        uint32_t mesh;
        DeviceBvhRecords r;
        bool fast;  // a record set of the fast traversal's trees (fast_meshes[mesh])
    };
    std::vector<DevPart> dev_parts;
    struct PartsGuard {
        std::vector<DevPart>& v;
        ~PartsGuard() {
            for (DevPart& p : v) {
                std::vector<void*> b{p.r.pairs, p.r.prims};
                free_blocks(b);
            }
        }
    } parts_guard{dev_parts};
    std::vector<float> normals, uvs;
    std::vector<std::pair<std::array<float, 3>, std::array<float, 3>>> local_aabb(d->mesh_count);
    s->bvhs.resize(d->mesh_count);
    s->mesh_has_bvh.assign(d->mesh_count, 0);
    uint64_t tri_total = 0;
    for (uint32_t i = 0; i < d->mesh_count; ++i) {
        const mt_mesh_desc& md = d->meshes[i];
        if (md.tri_count && (!md.positions || !md.normals))
            return fail(MT_E_INVALID, "mesh without positions or normals");
        const float* P = md.positions;
        DevMesh& dm = meshes[i];
        std::memset(&dm, 0, sizeof(dm));
        dm.attr_base = static_cast<uint32_t>(tri_total);
        dm.tri_count = md.tri_count;
        dm.flags = 0;
        tri_total += md.tri_count;
        // Mesh::CalculateAABB (Mesh.cpp:43-54)
        std::array<float, 3> mn{FLT_MAX, FLT_MAX, FLT_MAX}, mx{-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (uint64_t c = 0; c < static_cast<uint64_t>(md.tri_count) * 3; ++c)
            for (int k = 0; k < 3; ++k) {
                mn[k] = fmin_g(mn[k], P[c * 3 + k]);
                mx[k] = fmax_g(mx[k], P[c * 3 + k]);
            }
        local_aabb[i] = {mn, mx};
        // BVH trigger (Mesh.cpp:37-40): max depth passed as bvhMaxDepth - 1 (unsigned)
        const bool bvh = !d->bvh.disable && md.tri_count > d->bvh.max_triangles;
        std::vector<uint32_t> order;
        std::vector<uint32_t> leaf_of_slot;  // reference node index of every leaf slot
        const bool on_device = bvh && d->bvh.device_build_min_tris && md.tri_count >= d->bvh.device_build_min_tris;
        if (bvh) {
            if (on_device) {
                DeviceBvhRecords recs{};
                const int rc = build_bvh_device(P, md.tri_count, d->bvh.max_triangles, d->bvh.max_depth - 1u,
                                                s->bvhs[i], &recs);
                if (rc) return rc;
                dev_parts.push_back({i, recs, false});
            } else {
                build_bvh(P, md.tri_count, d->bvh.max_triangles, d->bvh.max_depth - 1u, s->bvhs[i]);
            }
            s->mesh_has_bvh[i] = 1;
            s->max_bvh_depth = std::max(s->max_bvh_depth, s->bvhs[i].depth);
            const HostBvh& hb = s->bvhs[i];
            dm.flags |= kMeshHasBvh;
            // a node's word: inner -> index of its children's pair record, leaf -> first slot
            auto word = [&](const HostNode& h) {
                return (h.a & kLeafBit) ? h.a : (h.a >> 1);
            };
            for (int k = 0; k < 3; ++k) {
                dm.root_min[k] = hb.nodes[0].mn[k];
                dm.root_max[k] = hb.nodes[0].mx[k];
            }
            dm.root_word = word(hb.nodes[0]);
            dm.pair_base = static_cast<uint32_t>(pairs.size() / 4);
            for (size_t q = 0; !on_device && q < hb.nodes.size() / 2; ++q) {
                const HostNode& A = hb.nodes[2 * q];
                const HostNode& B = hb.nodes[2 * q + 1];
                const uint32_t wa = word(A), wb = word(B), zero = 0;
                float4 r[4];
                r[0].x = A.mn[0]; r[0].y = A.mn[1]; r[0].z = A.mn[2]; std::memcpy(&r[0].w, &wa, 4);
                r[1].x = A.mx[0]; r[1].y = A.mx[1]; r[1].z = A.mx[2]; std::memcpy(&r[1].w, &wb, 4);
                r[2].x = B.mn[0]; r[2].y = B.mn[1]; r[2].z = B.mn[2]; std::memcpy(&r[2].w, &zero, 4);
                r[3].x = B.mx[0]; r[3].y = B.mx[1]; r[3].z = B.mx[2]; std::memcpy(&r[3].w, &zero, 4);
                pairs.insert(pairs.end(), r, r + 4);
            }
            if (!on_device) order = hb.indices;
            leaf_of_slot.assign(order.size(), 0);
            std::vector<uint8_t> last(order.size(), 0);
            for (size_t n = 0; !on_device && n < hb.nodes.size(); ++n) {
                const HostNode& h = hb.nodes[n];
                if (!(h.a & kLeafBit) || (n == 1)) continue;
                const uint32_t first = h.a & ~kLeafBit, count = h.b & ~kLeafBit;
                for (uint32_t k = first; k < first + count; ++k) leaf_of_slot[k] = static_cast<uint32_t>(n);
                if (count) last[first + count - 1] = 1;
            }
            last_flags.swap(last);
        } else {
            dm.pair_base = 0;
            dm.root_word = 0;
            order.resize(md.tri_count);
            for (uint32_t t = 0; t < md.tri_count; ++t) order[t] = t;
            leaf_of_slot.assign(order.size(), 0);
            last_flags.assign(order.size(), 0);
            if (!order.empty()) last_flags.back() = 1;
        }
        dm.prim_base = static_cast<uint32_t>(prims.size() / 3);
        dm.prim_count = static_cast<uint32_t>(on_device ? dev_parts.back().r.nslots : order.size());
        for (size_t k = 0; k < order.size(); ++k) {
            const uint32_t t = order[k];
            const float* v = P + static_cast<uint64_t>(t) * 9;
            const uint32_t lf = last_flags[k] ? kPrimLast : 0u;
            float4 a, b, c;
            a.x = v[0];
            a.y = v[1];
            a.z = v[2];
            std::memcpy(&a.w, &t, 4);
            b.x = v[3] - v[0];
            b.y = v[4] - v[1];
            b.z = v[5] - v[2];
            std::memcpy(&b.w, &leaf_of_slot[k], 4);
            c.x = v[6] - v[0];
            c.y = v[7] - v[1];
            c.z = v[8] - v[2];
            std::memcpy(&c.w, &lf, 4);
            prims.push_back(a);
            prims.push_back(b);
            prims.push_back(c);
        }
        normals.insert(normals.end(), md.normals, md.normals + static_cast<uint64_t>(md.tri_count) * 9);
        dm.uv_base = static_cast<uint32_t>(uvs.size() / 6);
        if (md.texcoord0) {
            dm.flags |= kMeshHasUv;
            uvs.insert(uvs.end(), md.texcoord0, md.texcoord0 + static_cast<uint64_t>(md.tri_count) * 6);
        }
        if (d->bvh.fast) {
            DevMesh& fm = fast_meshes[i];
            fm = dm;
            fm.flags &= ~kMeshHasBvh;
            if (d->bvh.fast_device_build_min_tris && md.tri_count >= d->bvh.fast_device_build_min_tris &&
                md.tri_count > 4) {
                DeviceBvhRecords recs{};
                uint32_t levels = 0;
                const int rc = build_fast_bvh_device(P, md.tri_count, fm, recs, levels);
                if (rc) return rc;
                dev_parts.push_back({i, recs, true});
                s->max_fast_depth = std::max(s->max_fast_depth, levels);
                fast_pairs += recs.npairs;
                fast_slots += recs.nslots;
            } else {
                const size_t p0 = pairs.size(), s0 = prims.size();
                s->max_fast_depth = std::max(s->max_fast_depth, build_fast_bvh(P, md.tri_count, pairs, prims, fm));
                fast_pairs += (pairs.size() - p0) / 4;
                fast_slots += (prims.size() - s0) / 3;
            }
        }
    }
    // the device-built parts follow the host-built records
    uint64_t tot_pairs = pairs.size() / 4, tot_prims = prims.size() / 3;
    for (DevPart& p : dev_parts) {
        DevMesh& m = p.fast ? fast_meshes[p.mesh] : meshes[p.mesh];
        m.pair_base = static_cast<uint32_t>(tot_pairs);
        m.prim_base = static_cast<uint32_t>(tot_prims);
        tot_pairs += p.r.npairs;
        tot_prims += p.r.nslots;
    }
    if (tot_pairs >= kLeafBit || tot_prims >= kLeafBit)
        return fail(MT_E_UNSUPPORTED, "more than 2^31 BVH node pairs or leaf slots");
    s->prim_slots = tot_prims - fast_slots;
    s->node_count = (tot_pairs - fast_pairs) * 2;
    s->fast_record_bytes = fast_pairs * 64 + fast_slots * 48;
    s->triangle_count = tri_total;

    std::vector<DevInstance> inst(d->instance_count);
    for (uint32_t i = 0; i < d->instance_count; ++i) {
        const mt_instance_desc& id = d->instances[i];
        if (id.mesh >= d->mesh_count || id.material >= d->material_count)
            return fail(MT_E_INVALID, "instance references a missing mesh or material");
        DevInstance& di = inst[i];
        std::memset(&di, 0, sizeof(di));
        std::memcpy(di.M, id.transform, sizeof(di.M));
        invert4(di.M, di.inv);
        di.mesh = id.mesh;
        di.material = id.material;
        di.attr_base = meshes[id.mesh].attr_base;
        di.uv_base = meshes[id.mesh].uv_base;
        di.mesh_flags = meshes[id.mesh].flags;
        // MeshInstance::CalculateWorldAABB (Mesh.cpp:56-80)
        const auto& bb = local_aabb[id.mesh];
        const float* a = bb.first.data();
        const float* b = bb.second.data();
        const float corners[8][3] = {{a[0], a[1], a[2]}, {a[0], a[1], b[2]}, {a[0], b[1], a[2]},
                                     {a[0], b[1], b[2]}, {b[0], a[1], a[2]}, {b[0], a[1], b[2]},
                                     {b[0], b[1], a[2]}, {b[0], b[1], b[2]}};
        for (int k = 0; k < 3; ++k) {
            di.wmin[k] = FLT_MAX;
            di.wmax[k] = -FLT_MAX;
        }
        for (const auto& c : corners) {
            const float v[4] = {c[0], c[1], c[2], 1.0f};
            float t[4];
            mul4(di.M, v, t);
            for (int k = 0; k < 3; ++k) {
                di.wmin[k] = fmin_g(di.wmin[k], t[k]);
                di.wmax[k] = fmax_g(di.wmax[k], t[k]);
            }
        }
    }

    // bounds of the whole scene (the radiance cache's position encoding)
    for (int k = 0; k < 3; ++k) {
        s->world_lo[k] = FLT_MAX;
        s->world_hi[k] = -FLT_MAX;
    }
    for (uint32_t i = 0; i < d->instance_count; ++i)
        for (int k = 0; k < 3; ++k) {
            s->world_lo[k] = fmin_g(s->world_lo[k], inst[i].wmin[k]);
            s->world_hi[k] = fmax_g(s->world_hi[k], inst[i].wmax[k]);
        }
    // top-level BVH for scenes with more instances than one wave-wide mask covers
    // (MT_TLAS_MIN: above how many instances, default 64); the fast traversal also uses one
    // above MT_TLAS_MIN_FAST instances (default 8: measured faster on the 50-instance sphere
    // grid, where the reference traversal's wave-wide instance masks stay faster)
    std::vector<float4> tlas_pairs;
    std::vector<uint32_t> tlas_inst;
    uint32_t tlas_min = 64, tlas_min_fast = 8;
    if (const char* v = std::getenv("MT_TLAS_MIN")) tlas_min = static_cast<uint32_t>(std::strtoul(v, nullptr, 10));
    if (const char* v = std::getenv("MT_TLAS_MIN_FAST"))
        tlas_min_fast = static_cast<uint32_t>(std::strtoul(v, nullptr, 10));
    s->dev.use_tlas = (d->instance_count > tlas_min ? kTlasReference | kTlasFast : 0u) |
                      (d->bvh.fast && d->instance_count > tlas_min_fast ? kTlasFast : 0u);
    if (s->dev.use_tlas) {
        std::vector<float> boxes(static_cast<size_t>(d->instance_count) * 6);
        for (uint32_t i = 0; i < d->instance_count; ++i) {
            std::memcpy(&boxes[6 * i], inst[i].wmin, sizeof(float) * 3);
            std::memcpy(&boxes[6 * i + 3], inst[i].wmax, sizeof(float) * 3);
        }
        s->tlas_depth = build_tlas(boxes, tlas_pairs, tlas_inst, s->dev.tlas_root_min,
                                   s->dev.tlas_root_max, s->dev.tlas_root_word);
    }

    std::vector<DevImage> images(d->image_count);
    std::vector<uint32_t> texels;
    for (uint32_t i = 0; i < d->image_count; ++i) {
        const mt_image_desc& im = d->images[i];
        if (!im.rgb || !im.width || !im.height) return fail(MT_E_INVALID, "empty image");
        images[i] = DevImage{static_cast<uint32_t>(texels.size()), im.width, im.height, 0};
        const uint64_t n = static_cast<uint64_t>(im.width) * im.height;
        for (uint64_t k = 0; k < n; ++k)
            texels.push_back(static_cast<uint32_t>(im.rgb[k * 3]) | (static_cast<uint32_t>(im.rgb[k * 3 + 1]) << 8) |
                             (static_cast<uint32_t>(im.rgb[k * 3 + 2]) << 16) | 0xFF000000u);
    }
    auto tex = [&](const mt_texture_desc& t) {
        DevTex o{};
        o.image = (t.image >= 0 && static_cast<uint32_t>(t.image) < d->image_count) ? t.image : -1;
        o.filter = t.filter;
        o.ox = t.offset[0];
        o.oy = t.offset[1];
        o.sx = t.scale[0];
        o.sy = t.scale[1];
        return o;
    };
    s->material_count = d->material_count;
    std::vector<DevMaterial> mats(d->material_count);
    for (uint32_t i = 0; i < d->material_count; ++i) {
        const mt_material_desc& m = d->materials[i];
        DevMaterial& o = mats[i];
        std::memset(&o, 0, sizeof(o));
        std::memcpy(o.base, m.base_color_factor, sizeof(o.base));
        std::memcpy(o.emissive, m.emissive_factor, sizeof(o.emissive));
        o.normal_scale = m.normal_scale;
        o.metallic = m.metallic_factor;
        o.roughness = m.roughness_factor;
        o.tex[0] = tex(m.base_color_tex);
        o.tex[1] = tex(m.metallic_roughness_tex);
        o.tex[2] = tex(m.normal_tex);
        o.tex[3] = tex(m.emissive_tex);
    }
    std::vector<DevLight> lights(d->light_count);
    for (uint32_t i = 0; i < d->light_count; ++i) {
        const mt_light_desc& l = d->lights[i];
        DevLight& o = lights[i];
        std::memset(&o, 0, sizeof(o));
        std::memcpy(o.pos, l.position, sizeof(o.pos));
        o.type = l.type;
        std::memcpy(o.color, l.color, sizeof(o.color));
        o.size = l.point_size;
        std::memcpy(o.dir, l.direction, sizeof(o.dir));
        o.inner = l.spot_inner_cos;
        o.outer = l.spot_outer_cos;
        o.intensity = (l.color[0] + l.color[1] + l.color[2]) / 3.0f;
    }
    std::vector<float4> env;
    s->dev.has_env = 0;
    if (d->env.format != MT_ENV_NONE && d->env.texels && d->env.width && d->env.height) {
        const uint64_t n = static_cast<uint64_t>(d->env.width) * d->env.height;
        env.resize(n);
        for (uint64_t k = 0; k < n; ++k) {
            float c[3];
            if (d->env.format == MT_ENV_RGB8) {
                const uint8_t* p = static_cast<const uint8_t*>(d->env.texels) + k * 3;
                for (int j = 0; j < 3; ++j) c[j] = static_cast<float>(p[j]) / 255.0f;  // Image.cpp:11-15
            } else if (d->env.format == MT_ENV_RGB32F) {
                const float* p = static_cast<const float*>(d->env.texels) + k * 3;
                for (int j = 0; j < 3; ++j) c[j] = p[j];
            } else {
                return fail(MT_E_INVALID, "unknown environment format");
            }
            env[k].x = c[0];
            env[k].y = c[1];
            env[k].z = c[2];
            env[k].w = 1.0f;
        }
        s->dev.has_env = 1;
        s->dev.env_w = d->env.width;
        s->dev.env_h = d->env.height;
        s->dev.env_filter = d->env.filter;
    }

    int rc;
    if ((rc = dev_upload(s, inst, &s->dev.instances))) return rc;
    if ((rc = dev_upload(s, meshes, &s->dev.meshes))) return rc;
    {
        // host-built records first, then each device-built part
        void *dp = nullptr, *dq = nullptr;
        if (device_malloc(&dp, sizeof(float4) * 4 * std::max<uint64_t>(tot_pairs, 1)))
            return fail(MT_E_NOMEM, "device allocation failed in scene upload");
        s->allocations.push_back(dp);
        if (device_malloc(&dq, sizeof(float4) * 3 * std::max<uint64_t>(tot_prims, 1)))
            return fail(MT_E_NOMEM, "device allocation failed in scene upload");
        s->allocations.push_back(dq);
        float4* fp = static_cast<float4*>(dp);
        float4* fq = static_cast<float4*>(dq);
        if (!pairs.empty() && hipMemcpy(fp, pairs.data(), sizeof(float4) * pairs.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(MT_E_HIP, "hipMemcpy failed in scene upload");
        if (!prims.empty() && hipMemcpy(fq, prims.data(), sizeof(float4) * prims.size(), hipMemcpyHostToDevice) != hipSuccess)
            return fail(MT_E_HIP, "hipMemcpy failed in scene upload");
        for (const DevPart& p : dev_parts) {
            const DevMesh& m = p.fast ? fast_meshes[p.mesh] : meshes[p.mesh];
            if (hipMemcpy(fp + 4ull * m.pair_base, p.r.pairs, sizeof(float4) * 4 * p.r.npairs,
                          hipMemcpyDeviceToDevice) != hipSuccess ||
                hipMemcpy(fq + 3ull * m.prim_base, p.r.prims, sizeof(float4) * 3 * p.r.nslots,
                          hipMemcpyDeviceToDevice) != hipSuccess)
                return fail(MT_E_HIP, "hipMemcpy of device-built BVH records failed");
        }
        s->dev.pairs = fp;
        s->dev.prims = fq;
    }
    if ((rc = dev_upload(s, normals, &s->dev.normals))) return rc;
    if ((rc = dev_upload(s, uvs, &s->dev.uvs))) return rc;
    if ((rc = dev_upload(s, mats, &s->dev.materials))) return rc;
    if ((rc = dev_upload(s, texels, &s->dev.texels))) return rc;
    if ((rc = dev_upload(s, images, &s->dev.images))) return rc;
    if ((rc = dev_upload(s, env, &s->dev.env))) return rc;
    if ((rc = dev_upload(s, lights, &s->dev.lights))) return rc;
    if ((rc = dev_upload(s, fast_meshes, &s->dev.fast_meshes))) return rc;
    if ((rc = dev_upload(s, tlas_pairs, &s->dev.tlas_pairs))) return rc;
    if ((rc = dev_upload(s, tlas_inst, &s->dev.tlas_inst))) return rc;
    s->dev.instance_count = d->instance_count;
    s->dev.light_count = d->light_count;
    return MT_OK;
}

}  // namespace mt

mt_scene::~mt_scene() {
    for (void* p : allocations) (void)hipFree(p);
}
