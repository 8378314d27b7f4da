// mt_bvh_build.hip -- the reference-rule BVH of one mesh built on the GPU, level by level.
//
// The result is the host builder's (mt_scene.cpp build_bvh, BVH.cpp:29-272) bit for bit:
//   * per triangle: lo = min(v0, min(v1, v2)), hi = max(...), sum = (v0 + v1) + v2, and the
//     root box as the first minimum / first maximum over the triangles in order (glm::min/max
//     keep their first argument on ties, so the reduction combines partial results left to
//     right, which keeps the sequential result, signed zeros included);
//   * per node: the vertex-mean centroid is the float sum of the triangles' sums IN THE NODE'S
//     TRIANGLE ORDER, divided by 3 x count -- a sequential sum, so one wave walks each large
//     node's list and adds the lanes' values in order (readlane), small nodes take a thread;
//   * side counts per axis (lo < c, hi >= c) are integer sums, aggregated per wave;
//   * score (|L| + |R|) / |N|, FLT_MAX when a side keeps everything, the axis rule of
//     BVH.cpp:85-101, and the stable partition (straddlers in both lists, order kept) through
//     exclusive scans of the side flags;
//   * flattening: children of a node are an adjacent pair in DFS order (node 1 unused), leaf
//     triangle lists appended in the same order; every node's slot and first index follow from
//     its subtree's slot and leaf-entry counts (bottom-up), assigned top-down.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "mt_internal.hpp"

namespace mt {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kSmallNode = 512;  // nodes up to this many triangles sum in one thread

__device__ __forceinline__ float dmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float dmax(float a, float b) { return (a < b) ? b : a; }

__global__ void k_tri_info(const float* P, uint32_t n, float4* lo, float4* hi, float4* sum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* v = P + 9ull * i;
    float l[3], h[3], s[3];
    for (int k = 0; k < 3; ++k) {
        l[k] = dmin(v[k], dmin(v[3 + k], v[6 + k]));
        h[k] = dmax(v[k], dmax(v[3 + k], v[6 + k]));
        s[k] = (v[k] + v[3 + k]) + v[6 + k];
    }
    lo[i] = make_float4(l[0], l[1], l[2], 0.0f);
    hi[i] = make_float4(h[0], h[1], h[2], 0.0f);
    sum[i] = make_float4(s[0], s[1], s[2], 0.0f);
}

// root box, pass 1: each thread folds a contiguous range in order (first min / first max)
__global__ void k_box_partial(const float4* lo, const float4* hi, uint32_t n, uint32_t per,
                              uint32_t parts, float4* pmn, float4* pmx) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= parts) return;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    const uint64_t b = static_cast<uint64_t>(t) * per;
    const uint64_t e = std::min<uint64_t>(b + per, n);
    for (uint64_t i = b; i < e; ++i) {
        const float4 l = lo[i], h = hi[i];
        mn[0] = dmin(mn[0], l.x);
        mn[1] = dmin(mn[1], l.y);
        mn[2] = dmin(mn[2], l.z);
        mx[0] = dmax(mx[0], h.x);
        mx[1] = dmax(mx[1], h.y);
        mx[2] = dmax(mx[2], h.z);
    }
    pmn[t] = make_float4(mn[0], mn[1], mn[2], 0.0f);
    pmx[t] = make_float4(mx[0], mx[1], mx[2], 0.0f);
}

// root box, pass 2: the partial results left to right in one thread; writes level 0's node
__global__ void k_box_final(const float4* pmn, const float4* pmx, uint32_t parts, float4* mn_out,
                            float4* mx_out) {
    if (blockIdx.x || threadIdx.x) return;
    float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t i = 0; i < parts; ++i) {
        const float4 a = pmn[i], b = pmx[i];
        mn[0] = dmin(mn[0], a.x);
        mn[1] = dmin(mn[1], a.y);
        mn[2] = dmin(mn[2], a.z);
        mx[0] = dmax(mx[0], b.x);
        mx[1] = dmax(mx[1], b.y);
        mx[2] = dmax(mx[2], b.z);
    }
    *mn_out = make_float4(mn[0], mn[1], mn[2], 0.0f);
    *mx_out = make_float4(mx[0], mx[1], mx[2], 0.0f);
}

__global__ void k_iota(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}
__global__ void k_fill(uint32_t* p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// the nodes of one level (structure of arrays)
struct Level {
    uint32_t *start, *cnt, *depth, *forced, *gid;
    float4 *mn, *mx;
    float4* com;        // centroid (w unused)
    uint32_t* state;    // 0 leaf, 1 candidate (needs its centroid and counts), 2 split
    uint32_t* counts;   // [node * 6]: |L| x, y, z, |R| x, y, z
    uint32_t* ax;
    uint32_t *ce, *cn, *cl;     // child elements, child nodes, leaf elements
    uint32_t *be, *bn, *bl;     // their exclusive scans
};

// which nodes are leaves; the centroid sums of the candidates with few triangles; the others
// go to the list of large nodes
__global__ void k_prep(Level L, uint32_t m, uint32_t max_tri, const uint32_t* idx,
                       const float4* tsum, uint32_t* big, uint32_t* nbig) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    for (int k = 0; k < 6; ++k) L.counts[i * 6 + k] = 0;
    const uint32_t c = L.cnt[i];
    if (c <= max_tri || L.forced[i]) {
        L.state[i] = 0;
        return;
    }
    L.state[i] = 1;
    if (c > kSmallNode) {
        big[atomicAdd(nbig, 1u)] = i;
        return;
    }
    float s[3] = {0.0f, 0.0f, 0.0f};
    const uint32_t b = L.start[i];
    for (uint32_t e = b; e < b + c; ++e) {
        const float4 v = tsum[idx[e]];
        s[0] = s[0] + v.x;
        s[1] = s[1] + v.y;
        s[2] = s[2] + v.z;
    }
    const float denom = static_cast<float>(static_cast<uint64_t>(c) * 3ull);
    L.com[i] = make_float4(s[0] / denom, s[1] / denom, s[2] / denom, 0.0f);
}

__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// one wave per large node: the lanes load 64 consecutive triangles' sums, then every lane adds
// them in order (the same values in every lane), so the sum is the sequential one
__global__ __launch_bounds__(64) void k_sum_big(Level L, const uint32_t* big, const uint32_t* nbig,
                                                const uint32_t* idx, const float4* tsum) {
    if (blockIdx.x >= *nbig) return;
    const uint32_t node = big[blockIdx.x];
    const uint32_t b = L.start[node], c = L.cnt[node];
    const uint32_t lane = threadIdx.x;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f;
    for (uint32_t base = 0; base < c; base += 64) {
        const uint32_t e = base + lane;
        float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (e < c) v = tsum[idx[b + e]];
        const int m = static_cast<int>(std::min<uint32_t>(64u, c - base));
        for (int j = 0; j < m; ++j) {
            sx = sx + lane_value(v.x, j);
            sy = sy + lane_value(v.y, j);
            sz = sz + lane_value(v.z, j);
        }
    }
    if (lane == 0) {
        const float denom = static_cast<float>(static_cast<uint64_t>(c) * 3ull);
        L.com[node] = make_float4(sx / denom, sy / denom, sz / denom, 0.0f);
    }
}

// side counts of the candidates: lo < c (left) and hi >= c (right) per axis, summed per wave
// for each node present in the wave
__global__ void k_count(Level L, const uint32_t* idx, const uint32_t* own, uint32_t E,
                        const float4* tlo, const float4* thi) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t node = 0xFFFFFFFFu;
    bool cand = false;
    uint32_t f = 0;
    if (e < E) {
        node = own[e];
        cand = L.state[node] == 1;
        if (cand) {
            const uint32_t t = idx[e];
            const float4 c = L.com[node], l = tlo[t], h = thi[t];
            f = (l.x < c.x ? 1u : 0u) | (l.y < c.y ? 2u : 0u) | (l.z < c.z ? 4u : 0u) |
                (h.x >= c.x ? 8u : 0u) | (h.y >= c.y ? 16u : 0u) | (h.z >= c.z ? 32u : 0u);
        }
    }
    uint64_t pending = __ballot(cand);
    while (pending) {
        const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
        const uint32_t o = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(node), leader));
        const bool mine = cand && node == o;
        const uint64_t same = __ballot(mine);
        for (int k = 0; k < 6; ++k) {
            const uint32_t n = static_cast<uint32_t>(__popcll(__ballot(mine && (f >> k & 1u))));
            if (static_cast<int>(threadIdx.x % 64) == leader && n) atomicAdd(&L.counts[o * 6 + k], n);
        }
        pending &= ~same;
    }
}

// BVHTreeNode::Subdivide's choice (BVH.cpp:85-101): the scores, the axis, leaf or split
__global__ void k_decide(Level L, uint32_t m) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t ce = 0, cn = 0, cl = 0;
    const uint32_t c = L.cnt[i];
    if (L.state[i] == 1) {
        float score[3];
        for (int k = 0; k < 3; ++k) {
            const uint32_t nl = L.counts[i * 6 + k], nr = L.counts[i * 6 + 3 + k];
            if (nl == c || nr == c)
                score[k] = FLT_MAX;
            else
                score[k] = static_cast<float>(static_cast<uint64_t>(nl) + nr) / static_cast<float>(static_cast<uint64_t>(c));
        }
        if (score[0] == FLT_MAX && score[1] == FLT_MAX && score[2] == FLT_MAX) {
            L.state[i] = 0;
        } else {
            const uint32_t ax = (score[0] < score[1] && score[0] < score[2]) ? 0 : (score[1] < score[2] ? 1 : 2);
            L.ax[i] = ax;
            L.state[i] = 2;
            ce = L.counts[i * 6 + ax] + L.counts[i * 6 + 3 + ax];
            cn = 2;
        }
    }
    if (L.state[i] == 0) cl = c;
    L.ce[i] = ce;
    L.cn[i] = cn;
    L.cl[i] = cl;
}

// side flags of the chosen axis: left in the low word, right in the high word
__global__ void k_flags(Level L, const uint32_t* idx, const uint32_t* own, uint32_t E,
                        const float4* tlo, const float4* thi, unsigned long long* flags) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t node = own[e];
    unsigned long long f = 0;
    if (L.state[node] == 2) {
        const uint32_t t = idx[e], ax = L.ax[node];
        const float4 c = L.com[node], l = tlo[t], h = thi[t];
        const float cc = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
        const float ll = ax == 0 ? l.x : (ax == 1 ? l.y : l.z);
        const float hh = ax == 0 ? h.x : (ax == 1 ? h.y : h.z);
        f = (ll < cc ? 1ull : 0ull) | ((hh >= cc ? 1ull : 0ull) << 32);
    }
    flags[e] = f;
}

// the children's triangle lists in the next level's arrays (stable: ranks from the exclusive
// scan of the side flags), and the leaves' lists into the leaf store
__global__ void k_partition(Level L, const uint32_t* idx, const uint32_t* own, uint32_t E,
                            const unsigned long long* flags, const unsigned long long* scan,
                            uint32_t* idx_next, uint32_t* own_next, uint32_t* leaf_store,
                            uint32_t leaf_base) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const uint32_t node = own[e], t = idx[e];
    if (L.state[node] == 2) {
        const unsigned long long s0 = scan[L.start[node]], sv = scan[e], f = flags[e];
        const uint32_t rl = static_cast<uint32_t>(sv) - static_cast<uint32_t>(s0);
        const uint32_t rr = static_cast<uint32_t>(sv >> 32) - static_cast<uint32_t>(s0 >> 32);
        const uint32_t nl = L.counts[node * 6 + L.ax[node]];
        const uint32_t base = L.be[node], child = L.bn[node];
        if (f & 1ull) {
            idx_next[base + rl] = t;
            own_next[base + rl] = child;
        }
        if (f >> 32) {
            idx_next[base + nl + rr] = t;
            own_next[base + nl + rr] = child + 1;
        }
    } else {
        leaf_store[leaf_base + L.bl[node] + (e - L.start[node])] = t;
    }
}

// the tree's records of this level's nodes, and the next level's nodes
struct Tree {
    float4 *mn, *mx;
    int32_t* child;     // global id of the first child (children are consecutive), -1 = leaf
    uint32_t *cnt, *leaf_off;
};

__global__ void k_children(Level L, uint32_t m, Level N, Tree T, uint32_t next_gid_base,
                           uint32_t leaf_base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = L.gid[i];
    const float4 mn = L.mn[i], mx = L.mx[i];
    T.mn[g] = mn;
    T.mx[g] = mx;
    T.cnt[g] = L.cnt[i];
    if (L.state[i] != 2) {
        T.child[g] = -1;
        T.leaf_off[g] = leaf_base + L.bl[i];
        return;
    }
    const uint32_t ax = L.ax[i];
    const uint32_t nl = L.counts[i * 6 + ax], nr = L.counts[i * 6 + 3 + ax];
    const uint32_t a = L.bn[i];
    T.child[g] = static_cast<int32_t>(next_gid_base + a);
    T.leaf_off[g] = 0;
    const float4 c = L.com[i];
    const float cc = ax == 0 ? c.x : (ax == 1 ? c.y : c.z);
    float4 amx = mx, bmn = mn;
    if (ax == 0) { amx.x = cc; bmn.x = cc; }
    if (ax == 1) { amx.y = cc; bmn.y = cc; }
    if (ax == 2) { amx.z = cc; bmn.z = cc; }
    const uint32_t d = L.depth[i];
    const uint32_t forced = d > 0 ? 0u : 1u;
    const uint32_t cd = d > 0 ? d - 1 : 0u;
    N.start[a] = L.be[i];
    N.cnt[a] = nl;
    N.mn[a] = mn;
    N.mx[a] = amx;
    N.start[a + 1] = L.be[i] + nl;
    N.cnt[a + 1] = nr;
    N.mn[a + 1] = bmn;
    N.mx[a + 1] = mx;
    for (int s = 0; s < 2; ++s) {
        N.depth[a + s] = cd;
        N.forced[a + s] = forced;
        N.gid[a + s] = next_gid_base + a + s;
    }
}

// flattening, bottom-up: flattened slots below a node and leaf entries of its subtree
__global__ void k_sizes(Tree T, uint32_t g0, uint32_t m, uint32_t* size, uint32_t* lcount) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = g0 + i;
    const int32_t a = T.child[g];
    if (a < 0) {
        size[g] = 0;
        lcount[g] = T.cnt[g];
    } else {
        size[g] = 2 + size[a] + size[a + 1];
        lcount[g] = lcount[a] + lcount[a + 1];
    }
}

// flattening, top-down: every node's record at its slot; its children's slots, pair slots and
// first leaf entries (slot, pair, ibase of the root are set by the host: 0, 2, 0)
__global__ void k_place(Tree T, uint32_t g0, uint32_t m, const uint32_t* size, const uint32_t* lcount,
                        uint32_t* slot, uint32_t* pair, uint32_t* ibase, const uint32_t* leaf_store,
                        HostNode* out, uint32_t* out_idx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t g = g0 + i;
    HostNode h;
    const float4 mn = T.mn[g], mx = T.mx[g];
    h.mn[0] = mn.x; h.mn[1] = mn.y; h.mn[2] = mn.z;
    h.mx[0] = mx.x; h.mx[1] = mx.y; h.mx[2] = mx.z;
    const int32_t a = T.child[g];
    if (a < 0) {
        const uint32_t c = T.cnt[g], b = ibase[g], src = T.leaf_off[g];
        h.a = b | kLeafBit;
        h.b = c;
        for (uint32_t k = 0; k < c; ++k) out_idx[b + k] = leaf_store[src + k];
    } else {
        const uint32_t p = pair[g];
        h.a = p;
        h.b = p + 1;
        slot[a] = p;
        slot[a + 1] = p + 1;
        pair[a] = p + 2;
        pair[a + 1] = p + 2 + size[a];
        ibase[a] = ibase[g];
        ibase[a + 1] = ibase[g] + lcount[a];
    }
    out[slot[g]] = h;
}

// the scene's records of a built tree: one 64-byte record per sibling pair {A.min, A.word}
// {A.max, B.word} {B.min, 0} {B.max, 0}, word = child pair index or first slot | leaf bit
__global__ void k_pair_records(const HostNode* nodes, uint32_t npairs, float4* pairs) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npairs) return;
    const HostNode A = nodes[2 * q], B = nodes[2 * q + 1];
    const uint32_t wa = (A.a & kLeafBit) ? A.a : (A.a >> 1);
    const uint32_t wb = (B.a & kLeafBit) ? B.a : (B.a >> 1);
    pairs[4 * q + 0] = make_float4(A.mn[0], A.mn[1], A.mn[2], __uint_as_float(wa));
    pairs[4 * q + 1] = make_float4(A.mx[0], A.mx[1], A.mx[2], __uint_as_float(wb));
    pairs[4 * q + 2] = make_float4(B.mn[0], B.mn[1], B.mn[2], 0.0f);
    pairs[4 * q + 3] = make_float4(B.mx[0], B.mx[1], B.mx[2], 0.0f);
}

// every leaf marks its slots with its node index and flags its last slot (node 1 is unused)
__global__ void k_leaf_marks(const HostNode* nodes, uint32_t nslots, uint32_t* leaf_of, uint32_t* last) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nslots || n == 1) return;
    const HostNode h = nodes[n];
    if (!(h.a & kLeafBit)) return;
    const uint32_t first = h.a & ~kLeafBit, count = h.b & ~kLeafBit;
    for (uint32_t k = first; k < first + count; ++k) {
        leaf_of[k] = n;
        last[k] = k + 1 == first + count ? kPrimLast : 0u;
    }
}

// one 48-byte slot per leaf entry: {v0, triangle} {v1 - v0, leaf node} {v2 - v0, last flag}
__global__ void k_prim_records(const float* P, const uint32_t* order, uint32_t n, const uint32_t* leaf_of,
                               const uint32_t* last, float4* prims) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = order[k];
    const float* v = P + 9ull * t;
    prims[3ull * k + 0] = make_float4(v[0], v[1], v[2], __uint_as_float(t));
    prims[3ull * k + 1] = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], __uint_as_float(leaf_of[k]));
    prims[3ull * k + 2] = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], __uint_as_float(last[k]));
}

// device buffers of one build, freed together
struct Bufs {
    std::vector<void*> v;
    ~Bufs() { free_blocks(v); }
    template <typename T>
    int get(T** p, size_t n) {
        void* q = nullptr;
        if (device_malloc(&q, std::max<size_t>(n, 1) * sizeof(T))) return MT_E_NOMEM;
        v.push_back(q);
        *p = static_cast<T*>(q);
        return MT_OK;
    }
    template <typename T>
    void drop(T* p) {  // free one block early
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i] == p) {
                std::vector<void*> one{p};
                free_blocks(one);
                v.erase(v.begin() + static_cast<long>(i));
                return;
            }
    }
};

int alloc_level(Bufs& B, Level& L, uint32_t m) {
    int rc;
    if ((rc = B.get(&L.start, m)) || (rc = B.get(&L.cnt, m)) || (rc = B.get(&L.depth, m)) ||
        (rc = B.get(&L.forced, m)) || (rc = B.get(&L.gid, m)) || (rc = B.get(&L.mn, m)) ||
        (rc = B.get(&L.mx, m)) || (rc = B.get(&L.com, m)) || (rc = B.get(&L.state, m)) ||
        (rc = B.get(&L.counts, static_cast<size_t>(m) * 6)) || (rc = B.get(&L.ax, m)) ||
        (rc = B.get(&L.ce, m)) || (rc = B.get(&L.cn, m)) || (rc = B.get(&L.cl, m)) ||
        (rc = B.get(&L.be, m)) || (rc = B.get(&L.bn, m)) || (rc = B.get(&L.bl, m)))
        return rc;
    return MT_OK;
}
void free_level(Bufs& B, Level& L) {
    for (void* p : {static_cast<void*>(L.start), static_cast<void*>(L.cnt), static_cast<void*>(L.depth),
                    static_cast<void*>(L.forced), static_cast<void*>(L.gid), static_cast<void*>(L.mn),
                    static_cast<void*>(L.mx), static_cast<void*>(L.com), static_cast<void*>(L.state),
                    static_cast<void*>(L.counts), static_cast<void*>(L.ax), static_cast<void*>(L.ce),
                    static_cast<void*>(L.cn), static_cast<void*>(L.cl), static_cast<void*>(L.be),
                    static_cast<void*>(L.bn), static_cast<void*>(L.bl)})
        B.drop(p);
    L = Level{};
}

#define BB_TRY(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            return fail(MT_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

inline uint32_t grid(uint64_t n) { return static_cast<uint32_t>(std::max<uint64_t>(1, (n + kThreads - 1) / kThreads)); }

// exclusive sums of n values (hipCUB), temporary storage grown on demand
template <typename T>
int scan(Bufs& B, void*& tmp, size_t& tmp_bytes, const T* in, T* out, uint32_t n, hipStream_t s) {
    size_t need = 0;
    BB_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, static_cast<int>(n), s));
    if (need > tmp_bytes) {
        if (tmp) B.drop(static_cast<unsigned char*>(tmp));
        unsigned char* t = nullptr;
        if (B.get(&t, need)) return MT_E_NOMEM;
        tmp = t;
        tmp_bytes = need;
    }
    BB_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, static_cast<int>(n), s));
    return MT_OK;
}

template <typename T>
int read_last(const T* scan, const T* vals, uint32_t n, T& total) {
    T a = 0, b = 0;
    BB_TRY(hipMemcpy(&a, scan + n - 1, sizeof(T), hipMemcpyDeviceToHost));
    BB_TRY(hipMemcpy(&b, vals + n - 1, sizeof(T), hipMemcpyDeviceToHost));
    total = a + b;
    return MT_OK;
}

// ------------------------------------------------------------------ the fast traversal's tree
// A linear BVH (Karras 2012, "Maximizing Parallelism in the Construction of BVHs, Octrees, and
// k-d Trees"): triangles sorted by the 30-bit Morton code of their box centre on the grid of
// the centres' bounds (ties by index), the binary radix tree over the sorted keys, every
// subtree of at most kFastLeaf triangles collapsed into one leaf, boxes fitted from the leaves
// upward a level at a time, records in the layout of the host's SAH build.
constexpr uint32_t kFastLeaf = 4;
constexpr uint32_t kLeafChild = 0x80000000u;  // child word of the radix tree: a sorted position

// 10-bit value -> its bits at every third position
__device__ __forceinline__ uint32_t spread3(uint32_t v) {
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_lbvh_boxes(const float* P, uint32_t n, float4* lo, float4* hi, float4* cen) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* v = P + 9ull * i;
    float l[3], h[3];
    for (int k = 0; k < 3; ++k) {
        l[k] = dmin(v[k], dmin(v[3 + k], v[6 + k]));
        h[k] = dmax(v[k], dmax(v[3 + k], v[6 + k]));
    }
    lo[i] = make_float4(l[0], l[1], l[2], 0.0f);
    hi[i] = make_float4(h[0], h[1], h[2], 0.0f);
    cen[i] = make_float4(0.5f * (l[0] + h[0]), 0.5f * (l[1] + h[1]), 0.5f * (l[2] + h[2]), 0.0f);
}

__global__ void k_lbvh_keys(const float4* cen, uint32_t n, const float4* cmn, const float4* cmx, uint32_t* key,
                            uint32_t* id) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = cen[i], a = *cmn, b = *cmx;
    const float cc[3] = {c.x, c.y, c.z}, lo[3] = {a.x, a.y, a.z}, hi[3] = {b.x, b.y, b.z};
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = hi[k] - lo[k];
        const float u = ext > 0.0f ? (cc[k] - lo[k]) / ext : 0.0f;
        q[k] = static_cast<uint32_t>(fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f));
    }
    key[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    id[i] = i;
}

// common prefix length of sorted keys i and j (ties broken by the positions); -1 outside
__device__ __forceinline__ int lbvh_delta(const uint32_t* key, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = key[i], b = key[j];
    if (a != b) return __clz(a ^ b);
    return 32 + __clz(static_cast<uint32_t>(i) ^ static_cast<uint32_t>(j));
}

// internal node i of the radix tree: its range of sorted positions [first, last], its split
// and its children (kLeafChild | position for a single triangle); parents of both kinds
__global__ void k_lbvh_nodes(const uint32_t* key, int n, uint32_t* child, uint32_t* first, uint32_t* last,
                             uint32_t* par_node, uint32_t* par_leaf) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (lbvh_delta(key, n, i, i + 1) - lbvh_delta(key, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin_ = lbvh_delta(key, n, i, i - d);
    int lmax = 2;
    while (lbvh_delta(key, n, i, i + lmax * d) > dmin_) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lbvh_delta(key, n, i, i + (l + t) * d) > dmin_) l += t;
    const int j = i + l * d;
    const int dnode = lbvh_delta(key, n, i, j);
    int sp = 0;
    for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
        if (lbvh_delta(key, n, i, i + (sp + t) * d) > dnode) sp += t;
        if (t == 1) break;
    }
    const int g = i + sp * d + (d < 0 ? -1 : 0);
    const int lo = d > 0 ? i : j, hi = d > 0 ? j : i;
    const uint32_t a = (lo == g) ? (kLeafChild | static_cast<uint32_t>(g)) : static_cast<uint32_t>(g);
    const uint32_t b = (hi == g + 1) ? (kLeafChild | static_cast<uint32_t>(g + 1)) : static_cast<uint32_t>(g + 1);
    child[2 * i] = a;
    child[2 * i + 1] = b;
    first[i] = static_cast<uint32_t>(lo);
    last[i] = static_cast<uint32_t>(hi);
    if (a & kLeafChild) par_leaf[a & ~kLeafChild] = static_cast<uint32_t>(i); else par_node[a] = static_cast<uint32_t>(i);
    if (b & kLeafChild) par_leaf[b & ~kLeafChild] = static_cast<uint32_t>(i); else par_node[b] = static_cast<uint32_t>(i);
}

// depth of every internal node (the root's is 0), by walking to the root
__global__ void k_lbvh_depth(const uint32_t* par_node, uint32_t m, uint32_t* depth, uint32_t* max_depth) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t d = 0, p = i;
    while (p != 0 && d < m) {
        p = par_node[p];
        ++d;
    }
    depth[i] = d;
    atomicMax(max_depth, d);
}

// boxes of the internal nodes at one depth from their children's (deeper levels are done)
__global__ void k_lbvh_fit(const uint32_t* child, const uint32_t* depth, uint32_t m, uint32_t level,
                           const uint32_t* id, const float4* tlo, const float4* thi, float4* nmn, float4* nmx) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || depth[i] != level) return;
    float4 a0, a1, b0, b1;
    const uint32_t ca = child[2 * i], cb = child[2 * i + 1];
    if (ca & kLeafChild) { a0 = tlo[id[ca & ~kLeafChild]]; a1 = thi[id[ca & ~kLeafChild]]; } else { a0 = nmn[ca]; a1 = nmx[ca]; }
    if (cb & kLeafChild) { b0 = tlo[id[cb & ~kLeafChild]]; b1 = thi[id[cb & ~kLeafChild]]; } else { b0 = nmn[cb]; b1 = nmx[cb]; }
    nmn[i] = make_float4(dmin(a0.x, b0.x), dmin(a0.y, b0.y), dmin(a0.z, b0.z), 0.0f);
    nmx[i] = make_float4(dmax(a1.x, b1.x), dmax(a1.y, b1.y), dmax(a1.z, b1.z), 0.0f);
}

// internal nodes that stay nodes: more than kFastLeaf triangles below them
__global__ void k_lbvh_keep(const uint32_t* first, const uint32_t* last, uint32_t m, uint32_t* keep) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) keep[i] = (last[i] - first[i] + 1u > kFastLeaf) ? 1u : 0u;
}

// one pair record per kept node (index: the scan of keep); children of at most kFastLeaf
// triangles become leaves over their sorted positions, marked for the prim records
__global__ void k_lbvh_pairs(const uint32_t* child, const uint32_t* first, const uint32_t* last,
                             const uint32_t* keep, const uint32_t* cidx, const uint32_t* depth, uint32_t m,
                             const uint32_t* id, const float4* tlo, const float4* thi, const float4* nmn,
                             const float4* nmx, float4* pairs, uint32_t* leaf_of, uint32_t* last_of,
                             uint32_t* max_level) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m || !keep[i]) return;
    uint32_t w[2];
    float4 mn[2], mx[2];
    bool leaf_child = false;
    for (int side = 0; side < 2; ++side) {
        const uint32_t c = child[2 * i + side];
        uint32_t f, l;
        if (c & kLeafChild) {
            f = l = c & ~kLeafChild;
            mn[side] = tlo[id[f]];
            mx[side] = thi[id[f]];
        } else {
            f = first[c];
            l = last[c];
            mn[side] = nmn[c];
            mx[side] = nmx[c];
        }
        if (!(c & kLeafChild) && keep[c]) {
            w[side] = cidx[c];
        } else {
            leaf_child = true;
            w[side] = kLeafBit | f;
            const uint32_t lid = 2u * i + static_cast<uint32_t>(side);
            for (uint32_t k = f; k <= l; ++k) {
                leaf_of[k] = lid;
                last_of[k] = (k == l) ? kPrimLast : 0u;
            }
        }
    }
    // levels as the host build counts them: the root is level 1, a leaf one below its parent
    atomicMax(max_level, depth[i] + (leaf_child ? 2u : 1u));
    float4* r = pairs + 4ull * cidx[i];
    r[0] = make_float4(mn[0].x, mn[0].y, mn[0].z, __uint_as_float(w[0]));
    r[1] = make_float4(mx[0].x, mx[0].y, mx[0].z, __uint_as_float(w[1]));
    r[2] = make_float4(mn[1].x, mn[1].y, mn[1].z, 0.0f);
    r[3] = make_float4(mx[1].x, mx[1].y, mx[1].z, 0.0f);
}

// prim slot k: the triangle at sorted position k {v0, triangle} {v1 - v0, leaf} {v2 - v0, last}
__global__ void k_lbvh_prims(const float* P, const uint32_t* id, uint32_t n, const uint32_t* leaf_of,
                             const uint32_t* last_of, float4* prims) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t t = id[k];
    const float* v = P + 9ull * t;
    float4* r = prims + 3ull * k;
    r[0] = make_float4(v[0], v[1], v[2], __uint_as_float(t));
    r[1] = make_float4(v[3] - v[0], v[4] - v[1], v[5] - v[2], __uint_as_float(leaf_of[k]));
    r[2] = make_float4(v[6] - v[0], v[7] - v[1], v[8] - v[2], __uint_as_float(last_of[k]));
}

}  // namespace

int build_bvh_device(const float* positions, uint32_t tri_count, uint32_t max_triangles,
                     uint32_t max_depth, HostBvh& out, DeviceBvhRecords* recs) {
    out = HostBvh{};
    if (tri_count == 0) return fail(MT_E_INVALID, "device BVH build of an empty mesh");
    hipStream_t s = nullptr;
    Bufs B;
    int rc;
    const uint32_t n = tri_count;
    float* P;
    float4 *tlo, *thi, *tsum;
    if ((rc = B.get(&P, 9ull * n)) || (rc = B.get(&tlo, n)) || (rc = B.get(&thi, n)) || (rc = B.get(&tsum, n)))
        return fail(rc, "device allocation failed in the BVH build");
    BB_TRY(hipMemcpy(P, positions, sizeof(float) * 9ull * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_tri_info, dim3(grid(n)), dim3(kThreads), 0, s, P, n, tlo, thi, tsum);
    BB_TRY(hipGetLastError());
    if (!recs) B.drop(P);
    // level 0: the root over every triangle, in order
    Level cur{};
    if ((rc = alloc_level(B, cur, 1))) return fail(rc, "device allocation failed in the BVH build");
    {
        const uint32_t parts = std::min<uint32_t>(65536u, std::max<uint32_t>(1u, n / 64u));
        const uint32_t per = (n + parts - 1) / parts;
        float4 *pmn, *pmx;
        if ((rc = B.get(&pmn, parts)) || (rc = B.get(&pmx, parts))) return fail(rc, "device allocation failed");
        hipLaunchKernelGGL(k_box_partial, dim3((parts + kThreads - 1) / kThreads), dim3(kThreads), 0, s,
                           tlo, thi, n, per, parts, pmn, pmx);
        hipLaunchKernelGGL(k_box_final, dim3(1), dim3(1), 0, s, pmn, pmx, parts, cur.mn, cur.mx);
        BB_TRY(hipGetLastError());
    }
    const uint32_t zero = 0;
    BB_TRY(hipMemcpy(cur.start, &zero, 4, hipMemcpyHostToDevice));
    BB_TRY(hipMemcpy(cur.cnt, &n, 4, hipMemcpyHostToDevice));
    // BVHTree passes maxDepth - 1 to the root's Subdivide (BVH.cpp:204-212), unsigned, as the
    // host builder does with the value it is given
    const uint32_t root_depth = max_depth - 1u;
    BB_TRY(hipMemcpy(cur.depth, &root_depth, 4, hipMemcpyHostToDevice));
    BB_TRY(hipMemcpy(cur.forced, &zero, 4, hipMemcpyHostToDevice));
    BB_TRY(hipMemcpy(cur.gid, &zero, 4, hipMemcpyHostToDevice));
    uint32_t *idx, *own;
    if ((rc = B.get(&idx, n)) || (rc = B.get(&own, n))) return fail(rc, "device allocation failed");
    hipLaunchKernelGGL(k_iota, dim3(grid(n)), dim3(kThreads), 0, s, idx, n);
    hipLaunchKernelGGL(k_fill, dim3(grid(n)), dim3(kThreads), 0, s, own, n, 0u);
    BB_TRY(hipGetLastError());

    // the tree grows by levels; its arrays are reallocated when a level does not fit
    Tree T{};
    uint64_t tcap = 0, ntree = 0;
    uint32_t* leaf_store = nullptr;
    uint64_t lcap = 0, nleaf = 0;
    auto grow_tree = [&](uint64_t need) -> int {
        if (need <= tcap) return MT_OK;
        const uint64_t cap = std::max<uint64_t>(need, tcap * 2);
        if (cap >= (1ull << 31)) return fail(MT_E_UNSUPPORTED, "more than 2^31 BVH nodes");
        Tree U{};
        int r;
        if ((r = B.get(&U.mn, cap)) || (r = B.get(&U.mx, cap)) || (r = B.get(&U.child, cap)) ||
            (r = B.get(&U.cnt, cap)) || (r = B.get(&U.leaf_off, cap)))
            return fail(r, "device allocation failed in the BVH build");
        if (ntree) {
            BB_TRY(hipMemcpy(U.mn, T.mn, sizeof(float4) * ntree, hipMemcpyDeviceToDevice));
            BB_TRY(hipMemcpy(U.mx, T.mx, sizeof(float4) * ntree, hipMemcpyDeviceToDevice));
            BB_TRY(hipMemcpy(U.child, T.child, sizeof(int32_t) * ntree, hipMemcpyDeviceToDevice));
            BB_TRY(hipMemcpy(U.cnt, T.cnt, sizeof(uint32_t) * ntree, hipMemcpyDeviceToDevice));
            BB_TRY(hipMemcpy(U.leaf_off, T.leaf_off, sizeof(uint32_t) * ntree, hipMemcpyDeviceToDevice));
        }
        if (tcap) {
            B.drop(T.mn);
            B.drop(T.mx);
            B.drop(T.child);
            B.drop(T.cnt);
            B.drop(T.leaf_off);
        }
        T = U;
        tcap = cap;
        return MT_OK;
    };
    auto grow_leaves = [&](uint64_t need) -> int {
        if (need <= lcap) return MT_OK;
        const uint64_t cap = std::max<uint64_t>(need, lcap * 2);
        if (cap >= (1ull << 31)) return fail(MT_E_UNSUPPORTED, "more than 2^31 BVH leaf entries");
        uint32_t* u;
        if (B.get(&u, cap)) return fail(MT_E_NOMEM, "device allocation failed in the BVH build");
        if (nleaf) BB_TRY(hipMemcpy(u, leaf_store, sizeof(uint32_t) * nleaf, hipMemcpyDeviceToDevice));
        if (lcap) B.drop(leaf_store);
        leaf_store = u;
        lcap = cap;
        return MT_OK;
    };

    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    uint32_t *big, *nbig;
    if ((rc = B.get(&nbig, 1))) return fail(rc, "device allocation failed");
    big = nullptr;
    uint64_t big_cap = 0;
    std::vector<std::pair<uint32_t, uint32_t>> levels;  // (first global id, node count)
    uint32_t m = 1, E = n;
    while (m) {
        if ((rc = grow_tree(ntree + m))) return rc;
        levels.push_back({static_cast<uint32_t>(ntree), m});
        if (big_cap < m) {
            if (big) B.drop(big);
            big_cap = std::max<uint64_t>(m, 1024);
            if ((rc = B.get(&big, big_cap))) return fail(rc, "device allocation failed");
        }
        BB_TRY(hipMemset(nbig, 0, 4));
        hipLaunchKernelGGL(k_prep, dim3(grid(m)), dim3(kThreads), 0, s, cur, m, max_triangles, idx, tsum, big, nbig);
        // one wave per large node (at most m blocks; blocks beyond the list's length return)
        uint32_t hb = 0;
        BB_TRY(hipMemcpy(&hb, nbig, 4, hipMemcpyDeviceToHost));
        if (hb) hipLaunchKernelGGL(k_sum_big, dim3(hb), dim3(64), 0, s, cur, big, nbig, idx, tsum);
        hipLaunchKernelGGL(k_count, dim3(grid(E)), dim3(kThreads), 0, s, cur, idx, own, E, tlo, thi);
        hipLaunchKernelGGL(k_decide, dim3(grid(m)), dim3(kThreads), 0, s, cur, m);
        BB_TRY(hipGetLastError());
        if ((rc = scan(B, tmp, tmp_bytes, cur.ce, cur.be, m, s)) || (rc = scan(B, tmp, tmp_bytes, cur.cn, cur.bn, m, s)) ||
            (rc = scan(B, tmp, tmp_bytes, cur.cl, cur.bl, m, s)))
            return rc;
        uint32_t e_next = 0, m_next = 0, l_add = 0;
        if ((rc = read_last(cur.be, cur.ce, m, e_next)) || (rc = read_last(cur.bn, cur.cn, m, m_next)) ||
            (rc = read_last(cur.bl, cur.cl, m, l_add)))
            return rc;
        if ((rc = grow_leaves(nleaf + l_add))) return rc;
        unsigned long long *flags, *fscan;
        uint32_t *idx_next = nullptr, *own_next = nullptr;
        if ((rc = B.get(&flags, E)) || (rc = B.get(&fscan, E)) || (rc = B.get(&idx_next, e_next)) ||
            (rc = B.get(&own_next, e_next)))
            return fail(rc, "device allocation failed in the BVH build");
        hipLaunchKernelGGL(k_flags, dim3(grid(E)), dim3(kThreads), 0, s, cur, idx, own, E, tlo, thi, flags);
        BB_TRY(hipGetLastError());
        if ((rc = scan(B, tmp, tmp_bytes, flags, fscan, E, s))) return rc;
        hipLaunchKernelGGL(k_partition, dim3(grid(E)), dim3(kThreads), 0, s, cur, idx, own, E, flags, fscan,
                           idx_next, own_next, leaf_store, static_cast<uint32_t>(nleaf));
        BB_TRY(hipGetLastError());
        Level nxt{};
        if (m_next && (rc = alloc_level(B, nxt, m_next))) return fail(rc, "device allocation failed");
        hipLaunchKernelGGL(k_children, dim3(grid(m)), dim3(kThreads), 0, s, cur, m, nxt, T,
                           static_cast<uint32_t>(ntree + m), static_cast<uint32_t>(nleaf));
        BB_TRY(hipGetLastError());
        BB_TRY(hipDeviceSynchronize());
        B.drop(flags);
        B.drop(fscan);
        B.drop(idx);
        B.drop(own);
        free_level(B, cur);
        cur = nxt;
        idx = idx_next;
        own = own_next;
        ntree += m;
        nleaf += l_add;
        m = m_next;
        E = e_next;
    }
    // flatten
    uint32_t *size, *lcount, *slot, *pair, *ibase;
    if ((rc = B.get(&size, ntree)) || (rc = B.get(&lcount, ntree)) || (rc = B.get(&slot, ntree)) ||
        (rc = B.get(&pair, ntree)) || (rc = B.get(&ibase, ntree)))
        return fail(rc, "device allocation failed in the BVH build");
    for (size_t l = levels.size(); l-- > 0;)
        hipLaunchKernelGGL(k_sizes, dim3(grid(levels[l].second)), dim3(kThreads), 0, s, T, levels[l].first,
                           levels[l].second, size, lcount);
    BB_TRY(hipGetLastError());
    uint32_t root_size = 0, root_leaf = 0;
    BB_TRY(hipMemcpy(&root_size, size, 4, hipMemcpyDeviceToHost));
    BB_TRY(hipMemcpy(&root_leaf, lcount, 4, hipMemcpyDeviceToHost));
    const uint64_t nslots = 2ull + root_size;
    if (nslots >= (1ull << 31)) return fail(MT_E_UNSUPPORTED, "more than 2^31 BVH nodes");
    HostNode* dnodes;
    uint32_t* didx;
    if ((rc = B.get(&dnodes, nslots)) || (rc = B.get(&didx, root_leaf))) return fail(rc, "device allocation failed");
    BB_TRY(hipMemset(dnodes, 0, sizeof(HostNode) * nslots));
    const uint32_t two = 2;
    BB_TRY(hipMemcpy(slot, &zero, 4, hipMemcpyHostToDevice));
    BB_TRY(hipMemcpy(pair, &two, 4, hipMemcpyHostToDevice));
    BB_TRY(hipMemcpy(ibase, &zero, 4, hipMemcpyHostToDevice));
    for (const auto& lv : levels)
        hipLaunchKernelGGL(k_place, dim3(grid(lv.second)), dim3(kThreads), 0, s, T, lv.first, lv.second, size,
                           lcount, slot, pair, ibase, leaf_store, dnodes, didx);
    BB_TRY(hipGetLastError());
    out.nodes.resize(nslots);
    out.indices.resize(root_leaf);
    BB_TRY(hipMemcpy(out.nodes.data(), dnodes, sizeof(HostNode) * nslots, hipMemcpyDeviceToHost));
    if (root_leaf) BB_TRY(hipMemcpy(out.indices.data(), didx, sizeof(uint32_t) * root_leaf, hipMemcpyDeviceToHost));
    out.depth = static_cast<uint32_t>(levels.size());
    if (recs) {
        // the scene's pair and prim records, built here; the caller owns them
        *recs = DeviceBvhRecords{};
        uint32_t *leaf_of, *last;
        if ((rc = B.get(&leaf_of, root_leaf)) || (rc = B.get(&last, root_leaf)))
            return fail(rc, "device allocation failed in the BVH build");
        void *dp = nullptr, *dq = nullptr;
        const uint64_t npairs = nslots / 2;
        if (device_malloc(&dp, sizeof(float4) * 4 * npairs)) return MT_E_NOMEM;
        if (device_malloc(&dq, sizeof(float4) * 3 * std::max<uint64_t>(root_leaf, 1))) {
            std::vector<void*> one{dp};
            free_blocks(one);
            return MT_E_NOMEM;
        }
        recs->pairs = static_cast<float4*>(dp);
        recs->prims = static_cast<float4*>(dq);
        recs->npairs = npairs;
        recs->nslots = root_leaf;
        hipLaunchKernelGGL(k_pair_records, dim3(grid(npairs)), dim3(kThreads), 0, s, dnodes,
                           static_cast<uint32_t>(npairs), recs->pairs);
        hipLaunchKernelGGL(k_leaf_marks, dim3(grid(nslots)), dim3(kThreads), 0, s, dnodes,
                           static_cast<uint32_t>(nslots), leaf_of, last);
        hipLaunchKernelGGL(k_prim_records, dim3(grid(root_leaf)), dim3(kThreads), 0, s, P, didx, root_leaf,
                           leaf_of, last, recs->prims);
        BB_TRY(hipGetLastError());
        BB_TRY(hipDeviceSynchronize());
    }
    return MT_OK;
}


int build_fast_bvh_device(const float* positions, uint32_t tri_count, DevMesh& dm, DeviceBvhRecords& recs,
                          uint32_t& levels) {
    recs = DeviceBvhRecords{};
    levels = 0;
    const uint32_t n = tri_count;
    if (n <= kFastLeaf) return fail(MT_E_INVALID, "device build of a mesh that is one leaf");
    if (n >= (1u << 31)) return fail(MT_E_UNSUPPORTED, "more than 2^31 triangles in one mesh");
    hipStream_t s = nullptr;
    Bufs B;
    int rc;
    const uint32_t m = n - 1;  // internal nodes of the radix tree
    float* P;
    float4 *tlo, *thi, *cen, *cmn, *cmx, *nmn, *nmx;
    uint32_t *key, *key2, *id, *id2, *child, *first, *last, *par_node, *par_leaf, *depth, *keep, *cidx, *leaf_of,
        *last_of, *counters;
    if ((rc = B.get(&P, 9ull * n)) || (rc = B.get(&tlo, n)) || (rc = B.get(&thi, n)) || (rc = B.get(&cen, n)) ||
        (rc = B.get(&cmn, 1)) || (rc = B.get(&cmx, 1)) || (rc = B.get(&key, n)) || (rc = B.get(&key2, n)) ||
        (rc = B.get(&id, n)) || (rc = B.get(&id2, n)) || (rc = B.get(&child, 2ull * m)) ||
        (rc = B.get(&first, m)) || (rc = B.get(&last, m)) || (rc = B.get(&par_node, m)) ||
        (rc = B.get(&par_leaf, n)) || (rc = B.get(&depth, m)) || (rc = B.get(&keep, m)) || (rc = B.get(&cidx, m)) ||
        (rc = B.get(&nmn, m)) || (rc = B.get(&nmx, m)) || (rc = B.get(&leaf_of, n)) || (rc = B.get(&last_of, n)) ||
        (rc = B.get(&counters, 2)))
        return fail(rc, "device allocation failed in the fast BVH build");
    BB_TRY(hipMemcpy(P, positions, sizeof(float) * 9ull * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_lbvh_boxes, dim3(grid(n)), dim3(kThreads), 0, s, P, n, tlo, thi, cen);
    {  // bounds of the box centres
        const uint32_t parts = std::min<uint32_t>(65536u, std::max<uint32_t>(1u, n / 64u));
        const uint32_t per = (n + parts - 1) / parts;
        float4 *pmn, *pmx;
        if ((rc = B.get(&pmn, parts)) || (rc = B.get(&pmx, parts))) return fail(rc, "device allocation failed");
        hipLaunchKernelGGL(k_box_partial, dim3((parts + kThreads - 1) / kThreads), dim3(kThreads), 0, s, cen, cen,
                           n, per, parts, pmn, pmx);
        hipLaunchKernelGGL(k_box_final, dim3(1), dim3(1), 0, s, pmn, pmx, parts, cmn, cmx);
    }
    hipLaunchKernelGGL(k_lbvh_keys, dim3(grid(n)), dim3(kThreads), 0, s, cen, n, cmn, cmx, key, id);
    BB_TRY(hipGetLastError());
    {  // stable sort by key: equal keys keep the triangle order
        size_t bytes = 0;
        BB_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key, key2, id, id2, static_cast<int>(n), 0, 30, s));
        unsigned char* tmp = nullptr;
        if ((rc = B.get(&tmp, bytes))) return fail(rc, "device allocation failed in the fast BVH build");
        BB_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, key, key2, id, id2, static_cast<int>(n), 0, 30, s));
    }
    const uint32_t zero2[2] = {0, 0};
    BB_TRY(hipMemcpy(counters, zero2, sizeof(zero2), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_fill, dim3(grid(m)), dim3(kThreads), 0, s, par_node, m, 0u);
    hipLaunchKernelGGL(k_lbvh_nodes, dim3(grid(m)), dim3(kThreads), 0, s, key2, static_cast<int>(n), child, first,
                       last, par_node, par_leaf);
    hipLaunchKernelGGL(k_lbvh_depth, dim3(grid(m)), dim3(kThreads), 0, s, par_node, m, depth, counters);
    BB_TRY(hipGetLastError());
    uint32_t max_depth = 0;
    BB_TRY(hipMemcpy(&max_depth, counters, 4, hipMemcpyDeviceToHost));
    for (uint32_t lv = max_depth + 1; lv-- > 0;)
        hipLaunchKernelGGL(k_lbvh_fit, dim3(grid(m)), dim3(kThreads), 0, s, child, depth, m, lv, id2, tlo, thi, nmn,
                           nmx);
    hipLaunchKernelGGL(k_lbvh_keep, dim3(grid(m)), dim3(kThreads), 0, s, first, last, m, keep);
    BB_TRY(hipGetLastError());
    void* tmp = nullptr;
    size_t tmp_bytes = 0;
    if ((rc = scan(B, tmp, tmp_bytes, keep, cidx, m, s))) return rc;
    uint32_t npairs = 0;
    if ((rc = read_last(cidx, keep, m, npairs))) return rc;
    void *dp = nullptr, *dq = nullptr;
    if (device_malloc(&dp, sizeof(float4) * 4 * std::max<uint32_t>(npairs, 1))) return MT_E_NOMEM;
    if (device_malloc(&dq, sizeof(float4) * 3 * n)) {
        std::vector<void*> one{dp};
        free_blocks(one);
        return MT_E_NOMEM;
    }
    recs.pairs = static_cast<float4*>(dp);
    recs.prims = static_cast<float4*>(dq);
    recs.npairs = npairs;
    recs.nslots = n;
    hipLaunchKernelGGL(k_lbvh_pairs, dim3(grid(m)), dim3(kThreads), 0, s, child, first, last, keep, cidx, depth, m,
                       id2, tlo, thi, nmn, nmx, recs.pairs, leaf_of, last_of, counters + 1);
    hipLaunchKernelGGL(k_lbvh_prims, dim3(grid(n)), dim3(kThreads), 0, s, P, id2, n, leaf_of, last_of, recs.prims);
    BB_TRY(hipGetLastError());
    float4 rmn, rmx;
    BB_TRY(hipMemcpy(&rmn, nmn, sizeof(float4), hipMemcpyDeviceToHost));
    BB_TRY(hipMemcpy(&rmx, nmx, sizeof(float4), hipMemcpyDeviceToHost));
    BB_TRY(hipMemcpy(&levels, counters + 1, 4, hipMemcpyDeviceToHost));
    BB_TRY(hipDeviceSynchronize());
    dm.prim_count = n;
    dm.root_word = 0;  // the root is kept (n > kFastLeaf) and numbered first
    dm.flags |= kMeshHasBvh;
    dm.root_min[0] = rmn.x;
    dm.root_min[1] = rmn.y;
    dm.root_min[2] = rmn.z;
    dm.root_max[0] = rmx.x;
    dm.root_max[1] = rmx.y;
    dm.root_max[2] = rmx.z;
    return MT_OK;
}

}  // namespace mt
