// mt_render.hip -- wavefront path tracing kernels for gfx950 and the render driver.
//
// One render = a sequence of chunks.  A chunk is a set of pixels with all their samples
// (npix * spp paths, path p = local pixel * spp + sample, so a wave holds the samples of one
// pixel -- or of 64/spp neighbouring pixels -- and its rays are as coherent as possible).
// Per chunk:
//   generate                camera rays, path state init             (Tracer.cpp:556-586)
//   for bounce k < bounces:
//     extend   closest hit of every active path                      (Intersect.cpp:201-219)
//     shade    miss -> sky; hit -> material, BRDF sampling, emission,
//              light choice, shadow-ray record, termination          (Tracer.cpp:134-340)
//     shadow   closest hit of every shadow ray, visible -> add light  (Tracer.cpp:317-336)
//   finalize  per-sample total and ACES                               (Tracer.cpp:344-346)
//   resolve   per-pixel mean (with firefly rejection when enabled)    (Tracer.cpp:490-601)
// Active paths and shadow rays live in SoA queues compacted with wave ballots and one atomic
// per wave.  extend/shadow are persistent kernels: each wave takes 64 queue entries at a time
// from an atomic work counter until the queue is drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "mt_device.hpp"
#include "mt_mlp.hpp"
#include "mt_render.hpp"

namespace mt {

using namespace dev;

namespace {

constexpr int kBlock = 256;
// idle lanes of a wave that trigger a refill of the traversal (measured best on the sphere grid:
// the reference traversal's instance culling runs per refill, so it prefers mid-size refills;
// the fast traversal's rays are short, so refilling less often wins)
constexpr uint32_t kRefillMin = 16;
constexpr uint32_t kRefillMinFast = 48;
constexpr uint32_t kSetupMin = 12;   // lanes needing an instance set-up that trigger one
constexpr uint32_t kGrab = 256;      // queue positions a traversal wave takes per atomic
constexpr uint32_t kGrabFast = 1024;
// the same settings, and the set-up thresholds of both traversals, for scenes whose records
// exceed the last-level cache
constexpr uint32_t kRefillMinLarge = 24;
constexpr uint32_t kGrabLarge = 64;
constexpr uint32_t kSetupMinLarge = 4, kSetupMinLargeFast = 48;
constexpr uint32_t kDiffChunks = 2, kDiffSlots = 2;  // differentiable renders: chunks and slots

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// wave-aggregated queue slot: one atomic per wave, ballot + prefix for the lane offsets
__device__ __forceinline__ uint32_t wave_push(uint32_t* counter, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return 0;
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    uint32_t base = 0;
    if (static_cast<int>(lane_id()) == leader) base = atomicAdd(counter, static_cast<uint32_t>(__popcll(m)));
    base = __shfl(base, leader);
    const uint64_t below = m & ((1ull << lane_id()) - 1ull);
    return base + static_cast<uint32_t>(__popcll(below));
}

// block-aggregated queue slot: one atomic per block and call.  Device-scope atomics on one
// address are serialised outside the XCDs' L2s (measured: one more atomic per wave-level push
// in the shading kernel, 4.4 -> 8.3 ms per launch), so the shading kernel pushes per block.
// Every thread of the block calls it at a block-uniform point.  Consecutive calls need no
// third barrier: a wave writes only its own s_off entry, after reading its previous results,
// and thread 0 rewrites the entries and s_base only after the first barrier of the next call,
// which every wave reaches after its reads.
__device__ __forceinline__ uint32_t block_push(uint32_t* counter, bool pred) {
    __shared__ uint32_t s_off[kBlock / 64];
    __shared__ uint32_t s_base;
    const uint64_t m = __ballot(pred);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) s_off[w] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t tot = 0;
        for (uint32_t i = 0; i < (blockDim.x >> 6); ++i) {
            const uint32_t n = s_off[i];
            s_off[i] = tot;
            tot += n;
        }
        s_base = tot ? atomicAdd(counter, tot) : 0u;
    }
    __syncthreads();
    const uint64_t below = m & ((1ull << lane_id()) - 1ull);
    return s_base + s_off[w] + static_cast<uint32_t>(__popcll(below));
}

__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, uint32_t v) {
    // wave reduction then one atomic
    unsigned long long x = v;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane_id() == 0 && x) atomicAdd(dst, x);
}

__device__ __forceinline__ void path_coords(const ChunkParams& c, uint32_t p, uint32_t& pix,
                                            uint32_t& s) {
    // a training-only pass of a cached render runs every path_stride-th path of the chunk
    const uint32_t q = c.path_stride > 1u ? p * c.path_stride + c.path_first : p;
    const uint32_t lp = q / c.spp;  // pixel-major: a wave holds the samples of one pixel
    s = q - lp * c.spp + c.sample_base;
    pix = c.pix_begin + lp;
}

// camera ray and path state of path p (pixel pix, sample s) of a render with parameters J
__device__ __forceinline__ void init_path(const ChunkParams& J, const PathBuffers& b, uint32_t p,
                                          uint32_t pix, uint32_t s) {
    const uint32_t x = pix % J.width, y = pix / J.width;
    const uint32_t key = path_key(J.seed_key, pix, s);
    // jitter: u = x/W + U[0, 1/W) (Tracer.cpp:567-586)
    const float u = static_cast<float>(x) * J.px_w + (draw(key, 0) * (J.px_w - 0.0f) + 0.0f);
    const float v = static_cast<float>(y) * J.px_h + (draw(key, 1) * (J.px_h - 0.0f) + 0.0f);
    V3 o, d;
    if (J.cam_type == MT_CAMERA_PERSPECTIVE) {  // Tracer/Camera.cpp:17-29
        const float px = (2.0f * u - 1.0f) * J.fov_scale * J.aspect;
        const float py = (2.0f * v - 1.0f) * J.fov_scale;
        const V3 dc = normalize(V3{px, -py, -1.0f});
        o = xform(J.cam, mk(0.0f), 1.0f);
        d = normalize(xform(J.cam, dc, 0.0f));
    } else {  // Camera.cpp:31-44
        const float px = (2.0f * u - 1.0f) * J.ortho_x * J.aspect / J.cam_aspect;
        const float py = (2.0f * v - 1.0f) * J.ortho_y;
        const V3 dc = normalize(V3{0.0f, 0.0f, -1.0f});
        o = xform(J.cam, V3{px, -py, 0.0f}, 1.0f);
        d = normalize(xform(J.cam, dc, 0.0f));
    }
    // bounce 0's queue is the identity (slot p holds path p) and every path starts with
    // throughputs and previous base colour 1: the kernels of bounce 0 take those from the slot
    // index and constants, so only the camera ray and the accumulators are written here
    b.ray_o[0][p] = f4(o, 0.0f);
    b.ray_d[0][p] = f4(d, 0.0f);
    const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    b.acc_e[p] = z;
    b.acc_d[p] = z;
    b.acc_i[p] = z;
    b.acc_s[p] = z;
}

__global__ __launch_bounds__(kBlock) void k_generate(ChunkParams c, PathBuffers b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        // a cached render reads the overflow flag once at the end: keep it across chunks
        for (uint32_t k = 0; k < kMaxCounters; ++k)
            if (k != kCntError || !c.cache) b.counters[k] = 0;
        b.counters[kCntExtend] = c.paths;
    }
    if (p >= c.paths) return;
    uint32_t pix, s;
    path_coords(c, p, pix, s);
    init_path(c, b, p, pix, s);
    if (c.debug) {
        b.dbg_i[p] = make_uint4(0, 0, 0, 0);
        b.dbg_j[p] = make_uint4(0, 0, 0, 0);
    }
    if (c.cache) {  // one path in train_every is a training path (the set moves every pass)
        const uint32_t q = p + c.train_offset;
        uint32_t slot = kNoSlot;
        if (c.train_every != kNoSlot && q % c.train_every == 0) slot = q / c.train_every;
        b.train_slot[p] = slot;
        if (c.cache_nee) b.acc_n[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (slot < c.train_cap) {
            b.tr_path[slot] = p;
            for (uint32_t j = 0; j < c.train_nrec; ++j)
                b.tr_w[static_cast<size_t>(slot) * c.train_nrec + j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    if (c.diff) {
        b.vis[p] = 0u;
        b.adj_s[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        b.adj_prev[p] = make_float4(0.0f, 0.0f, __uint_as_float(0xFFFFFFFFu), 0.0f);
    }
}

// a queue size that atomics (wave_push) update: read it with an agent-scope atomic load, which
// sees every earlier atomic on it even where only L1 is invalidated between the writer and the
// reader (the persistent batch kernel's barriers)
__device__ __forceinline__ uint32_t counter(const PathBuffers& b, uint32_t i) {
    return __hip_atomic_load(b.counters + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ ray binning
// Before the extend of a secondary bounce the queue's rays are binned by direction octant and
// origin cell, so that the rays a wave takes up together start close to each other and head
// the same way.  Only the order in which extend takes the queue's slots changes: it reads the
// ray of slot perm[i] and writes that slot's hit record, so hits and shading are unchanged.
// Counting sort over at most kBinMax bins: per-block histograms in LDS over contiguous slot
// ranges, one global histogram, its exclusive scan, then every block reserves its run of each
// bin with one atomic and places its slots (in any order inside the run).
constexpr uint32_t kBinMax = 4096;
constexpr uint32_t kBinKeyBits = 12;  // direction bits (3 per axis step) + origin bits

__device__ __forceinline__ uint32_t ray_bin(const ChunkParams& c, V3 o, V3 d) {
    uint32_t key = 0;
    const float dv[3] = {d.x, d.y, d.z}, ov[3] = {o.x, o.y, o.z};
    const uint32_t dtop = (1u << c.bin_dir_bits) - 1u;
    const float dscale = 0.5f * static_cast<float>(1u << c.bin_dir_bits);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float f = (dv[a] + 1.0f) * dscale;
        key = (key << c.bin_dir_bits) |
              (f > 0.0f ? (f < static_cast<float>(dtop) ? static_cast<uint32_t>(f) : dtop) : 0u);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const uint32_t top = (1u << c.bin_bits[a]) - 1u;
        const float f = (ov[a] - c.bin_lo[a]) * c.bin_scale[a];
        // cells outside the box (and NaN) clamp to the edge cells
        const uint32_t q = f > 0.0f ? (f < static_cast<float>(top) ? static_cast<uint32_t>(f) : top) : 0u;
        key = (key << c.bin_bits[a]) | q;
    }
    return key;
}

// the contiguous slot range [lo, hi) of this block out of n
__device__ __forceinline__ void bin_range(uint32_t n, uint32_t& lo, uint32_t& hi) {
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    lo = min(blockIdx.x * per, n);
    hi = min(lo + per, n);
}

__device__ __forceinline__ void bin_block_histogram(const PathBuffers& b, uint32_t lo, uint32_t hi,
                                                    uint32_t* h) {
    for (uint32_t j = threadIdx.x; j < kBinMax; j += kBlock) h[j] = 0;
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) atomicAdd(&h[b.bin_key[i] & (kBinMax - 1)], 1u);
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_bin_count(PathBuffers b, uint32_t k) {
    __shared__ uint32_t h[kBinMax];
    uint32_t lo, hi;
    bin_range(counter(b, kCntExtend + k), lo, hi);
    bin_block_histogram(b, lo, hi, h);
    for (uint32_t j = threadIdx.x; j < kBinMax; j += kBlock)
        if (h[j]) atomicAdd(b.bin_hist + j, h[j]);
}

// one block: bin_cur = exclusive scan of bin_hist, and bin_hist back to zero for the next bounce
__global__ __launch_bounds__(kBlock) void k_bin_scan(PathBuffers b) {
    constexpr uint32_t kPer = kBinMax / kBlock;
    __shared__ uint32_t part[kBlock];
    uint32_t v[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        v[j] = b.bin_hist[threadIdx.x * kPer + j];
        sum += v[j];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t i = 0; i < kBlock; ++i) {
            const uint32_t t = part[i];
            part[i] = run;
            run += t;
        }
    }
    __syncthreads();
    uint32_t run = part[threadIdx.x];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        b.bin_cur[threadIdx.x * kPer + j] = run;
        b.bin_hist[threadIdx.x * kPer + j] = 0;
        run += v[j];
    }
}

__global__ __launch_bounds__(kBlock) void k_bin_scatter(PathBuffers b, uint32_t k) {
    __shared__ uint32_t h[kBinMax];
    const uint32_t n = counter(b, kCntExtend + k);
    uint32_t lo, hi;
    bin_range(n, lo, hi);
    bin_block_histogram(b, lo, hi, h);
    for (uint32_t j = threadIdx.x; j < kBinMax; j += kBlock)
        if (h[j]) h[j] = atomicAdd(b.bin_cur + j, h[j]);  // this block's first position in bin j
    __syncthreads();
    for (uint32_t i = lo + threadIdx.x; i < hi; i += kBlock) {
        const uint32_t pos = atomicAdd(&h[b.bin_key[i] & (kBinMax - 1)], 1u);
        if (pos < n) b.perm[pos] = i;
    }
}

// ------------------------------------------------------------------ batched rendering
// One iteration of a batch: the scheduler takes tiles from the scene's device queue in order,
// from its head up to the end the host wrote for this iteration (q_state[1], the same greedy
// rule: while their paths fit the iteration's capacity, always at least one), lays the paths
// out tile after tile (pixel-major inside a tile) and resets the counters.
__global__ void k_batch_schedule(PathBuffers b) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t head = b.q_state[0];
    const uint32_t tail = b.q_state[1];
    uint32_t paths = 0, npix = 0, n = 0;
    while (head < tail && n < b.it_tile_cap) {
        const uint4 t = b.q_tiles[head];
        if (n > 0 && paths + t.w > b.it_capacity) break;
        b.it_tiles[n] = make_uint4(t.x, t.y, t.z, paths);
        b.it_pix_off[n] = npix;
        paths += t.w;
        npix += t.z;
        ++n;
        ++head;
    }
    b.it_pix_off[n] = npix;
    b.q_state[0] = head;
    for (uint32_t k = 0; k < kMaxCounters; ++k) b.counters[k] = 0;
    b.counters[kCntExtend] = paths;
    b.counters[kCntPaths] = paths;
    b.counters[kCntPixels] = npix;
    b.counters[kCntTiles] = n;
}

// last tile whose offset (in paths: .w, in pixels: it_pix_off) is <= x
__device__ __forceinline__ uint32_t tile_of_path(const PathBuffers& b, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b.it_tiles[mid].w <= x) lo = mid; else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ uint32_t tile_of_pixel(const PathBuffers& b, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (b.it_pix_off[mid] <= x) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool kBatch>
__device__ __forceinline__ const ChunkParams& job_params(const ChunkParams& c, const PathBuffers& b,
                                                         uint32_t p) {
    if constexpr (kBatch) return b.jobs[b.path_job[p]].p;
    else return c;
}
template <bool kBatch>
__device__ __forceinline__ void coords(const ChunkParams& c, const PathBuffers& b, uint32_t p,
                                       uint32_t& pix, uint32_t& s) {
    if constexpr (kBatch) {
        pix = b.path_pix[p];
        s = b.path_smp[p];
    } else {
        path_coords(c, p, pix, s);
    }
}

// path p (< the iteration's paths) of a batched iteration: its tile, job, pixel and sample
__device__ __forceinline__ void batch_generate_path(const PathBuffers& b, uint32_t p) {
    const uint4 t = b.it_tiles[tile_of_path(b, b.counters[kCntTiles], p)];
    const ChunkParams& J = b.jobs[t.x].p;
    const uint32_t local = p - t.w;
    const uint32_t lp = local / J.spp;
    const uint32_t s = local - lp * J.spp;
    const uint32_t pix = t.y + lp;
    b.path_job[p] = t.x;
    b.path_pix[p] = pix;
    b.path_smp[p] = s;
    init_path(J, b, p, pix, s);
}

__global__ __launch_bounds__(kBlock) void k_batch_generate(PathBuffers b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= b.counters[kCntPaths]) return;
    batch_generate_path(b, p);
}

// rays of the test hooks (mode 2: closest hits, mode 3: occlusion up to a distance)
struct TestIO {
    const float* rays;  // mode 2: 6 floats per ray, mode 3: 7 (origin, direction, distance)
    mt_hit* out;
    uint32_t n;
    uint32_t* fetch;
    uint32_t* occ;
};

// Persistent closest-hit kernel.  mode 0: extend (queue of path ids), 1: shadow records,
// 2/3: test rays.  kAccel 0 runs the reference traversal, 1 the fast one (for shadow rays an
// any-hit query bounded by the light distance).  A wave keeps 64 lanes busy: whenever at least refill_min lanes are idle it
// hands them that many new work items (ballot + prefix for the lane slots) from its own batch
// of queue positions, taken `grab` at a time with one atomic, runs the instance culling for
// them together, and every lane then advances its own ray by one unit of work per loop
// iteration.  Waves leave the loop when the queue is drained.
#ifndef MT_TRACE_WAVES_PER_SIMD
#define MT_TRACE_WAVES_PER_SIMD 6
#endif
// The unit steps of one wave step of the traversal, in order: 'N' a node pop (for lanes not in
// a leaf), 'T' a triangle test (for lanes in a leaf); the reference and the fast traversal
#ifndef MT_STEP_SCHEDULE
#define MT_STEP_SCHEDULE "NTTNTNT"
#endif
#ifndef MT_STEP_SCHEDULE_FAST
#define MT_STEP_SCHEDULE_FAST "NNTNNT"
#endif
constexpr char kStepSchedule[] = MT_STEP_SCHEDULE;
constexpr char kStepScheduleFast[] = MT_STEP_SCHEDULE_FAST;
constexpr int kStepUnits = static_cast<int>(sizeof(kStepSchedule)) - 1;
constexpr int kStepUnitsFast = static_cast<int>(sizeof(kStepScheduleFast)) - 1;
static_assert(kStepSchedule[0] == 'N' && kStepScheduleFast[0] == 'N', "a wave step starts with a node pop");
// The body runs in k_trace (gtid/nthreads over the grid) and in the batch's persistent kernel
// (over one lane group's blocks); lds_stack holds kDepth * kBlock entries of the block.
template <bool kCount, int kMode, uint32_t kDepth, int kAccel>
__device__ __forceinline__ void trace_body(const DevScene& S, const ChunkParams& c, const PathBuffers& b,
                                           uint32_t bounce, const TestIO& io, uint2* lds_stack,
                                           uint32_t gtid, uint32_t nthreads) {
    Stack st{lds_stack + threadIdx.x, kBlock, kDepth, b.spill + gtid,
             nthreads, c.stack_cap, b.counters + kCntError, 0};
    uint32_t count;
    uint32_t* fetch;
    if (kMode == 0) {
        count = counter(b, kCntExtend + bounce);
        fetch = b.counters + kCntFetchExt + bounce;
    } else if (kMode == 1) {
        count = counter(b, kCntShadow + bounce);
        fetch = b.counters + kCntFetchSh + bounce;
    } else {
        count = io.n;
        fetch = io.fetch;
    }
    constexpr bool kAny = kAccel == 1 && (kMode == 1 || kMode == 3);
    // per-hit diagnostics (the winning instance's tests, the leaf index): counting builds (work
    // statistics and debug views) and the test hooks only
    constexpr bool kDiag = kCount || kMode >= 2;
    const uint32_t* q = b.queue[bounce & 1];
    const bool binned = kMode == 0 && c.bin && bounce > 0;
    const uint32_t refill_min = c.refill_min;
    // a wave takes work items from the queue in batches (one atomic per batch) and hands them
    // to its idle lanes from that batch: one atomic per refill on one counter for every wave of
    // the chip serialises at the L2 when rays are short.  A batch is at most c.grab positions
    // and at most an eighth of a wave's share of the queue, so short queues still spread over
    // every wave; never below 64 (one refill of a whole wave).
    const uint32_t grab = max(64u, min(c.grab, count / (max(nthreads / 64u, 1u) * 8u)));
    const uint64_t lanes_below = (1ull << lane_id()) - 1ull;
    Work w{};
    Trav t;
    bool active = false, need = false, exhausted = false, gdone = false;
    uint32_t item = 0, lb = 0, le = 0;  // the wave's batch [lb, le) of queue positions
    while (true) {
        const uint64_t idle = __ballot(!active);
        const uint32_t n_idle = static_cast<uint32_t>(__popcll(idle));
        if (!exhausted && n_idle >= refill_min) {
            const uint32_t rem = le - lb;
            uint32_t nb = 0;
            const bool take = rem < n_idle && !gdone;
            if (take) {
                // every lane of the wave runs this loop, so lane 0 is the first active lane;
                // readfirstlane keeps the batch bounds wave-uniform (scalar registers)
                if (lane_id() == 0) nb = atomicAdd(fetch, grab);
                nb = __builtin_amdgcn_readfirstlane(nb);
                if (nb >= count || count - nb <= grab) gdone = true;
            }
            // idle lane of rank r: position lb + r from the batch, past its end the new batch's
            const uint32_t r = static_cast<uint32_t>(__popcll(idle & lanes_below));
            const uint32_t my = r < rem ? lb + r : (take ? nb + (r - rem) : count);
            if (rem >= n_idle) {
                lb += n_idle;
            } else if (take) {
                lb = nb + (n_idle - rem);
                le = nb < count ? (count - nb < grab ? count : nb + grab) : nb;
                if (lb > le) lb = le;
            } else {
                lb = le;
            }
            if (gdone && lb >= le) exhausted = true;
            if (!active) {
                if (my < count) {
                    // binned bounces: the my-th slot in bin order (clamped: a slot index
                    // never leaves the queue)
                    if (kMode == 0 && binned) item = min(b.perm[my], count - 1u);
                    else item = my;
                    V3 o, d;
                    float bound = kFltMax;
                    if (kMode == 0) {
                        o = xyz(b.ray_o[bounce & 1][item]);
                        d = xyz(b.ray_d[bounce & 1][item]);
                    } else if (kMode == 1) {
                        const float4 so = b.sh_o[my];
                        o = xyz(so);
                        d = xyz(b.sh_d[my]);
                        bound = so.w;
                    } else if (kMode == 2) {
                        o = V3{io.rays[my * 6 + 0], io.rays[my * 6 + 1], io.rays[my * 6 + 2]};
                        d = V3{io.rays[my * 6 + 3], io.rays[my * 6 + 4], io.rays[my * 6 + 5]};
                    } else {
                        o = V3{io.rays[my * 7 + 0], io.rays[my * 7 + 1], io.rays[my * 7 + 2]};
                        d = V3{io.rays[my * 7 + 3], io.rays[my * 7 + 4], io.rays[my * 7 + 5]};
                        bound = io.rays[my * 7 + 6];
                    }
                    trav_begin(t, o, d);
                    // an occluder counts up to and including the light distance (Tracer.cpp:322)
                    if (kAny) t.R.dist = bound >= 0.0f ? __uint_as_float(__float_as_uint(bound) + 1u) : bound;
                    active = true;
                    if ((S.use_tlas & (kAccel ? kTlasFast : kTlasReference)) && finite3(t.invd) && finite3(o)) {
                        if (kAccel) {
                            ftrav_tlas_begin(S, t, st);
                        } else {
                            trav_tlas_begin(S, t, st);
                        }
                        need = false;
                    } else if (kMode <= 1 && !kAccel && !kCount && b.inst_mask) {
                        // the mask computed before the traversal (k_instance_masks), and the
                        // start of the instance walk as the culling below would begin it
                        t.mask = b.inst_mask[item];
                        t.cur = -1;
                        t.leaf = kNoLeaf;
                        t.has_next = false;
                        st.sp = 0;
                        need = false;
                    } else {
                        need = true;
                    }
                }
            }
        }
        // instance culling for lanes that start a (group of 64) instances, one group at a time
        while (true) {
            const uint64_t nm = __ballot(need);
            if (nm == 0) break;
            if (kCount) w.it_mask++;
            const int l = __ffsll(static_cast<unsigned long long>(nm)) - 1;
            const uint32_t ucb = __builtin_amdgcn_readfirstlane(__shfl(t.cb, l));
            const bool mine = need && t.cb == ucb;
            const bool fast = finite3(t.invd) && finite3(t.o);
            if (__ballot(mine && !fast) == 0) {
                if (mine) t.mask = instance_mask<kCount, true>(S, ucb, t.o, t.invd, w);
            } else {
                if (mine) t.mask = instance_mask<kCount, false>(S, ucb, t.o, t.invd, w);
            }
            if (mine) {
                t.cur = -1;
                t.leaf = kNoLeaf;
                t.has_next = false;
                st.sp = 0;
                need = false;
            }
        }
        if (__ballot(active) == 0) {
            if (exhausted) break;
            continue;
        }
        // one step per lane; set-up steps (instance finish/start, the costliest branch) run only
        // when at least setup_min lanes need one or no lane can do anything else
        const bool wants_setup = active && trav_needs_setup(t, st);
        const uint32_t n_setup = static_cast<uint32_t>(__popcll(__ballot(wants_setup)));
        const uint32_t n_active = static_cast<uint32_t>(__popcll(__ballot(active)));
        const bool run_setup = n_setup >= c.setup_min || n_setup == n_active;
        uint32_t kind = 0xFFu;
        bool finished = false;
        bool nm = false;
        if (active) {
            if (wants_setup) {
                if (run_setup) {
                    kind = kStepSetup;
                    if (kAccel) {
                        finished = ftrav_setup<kCount, kAny, kDiag>(S, t, st, w, nm);
                    } else {
                        finished = trav_setup<kCount, kDiag>(S, t, st, w, nm);
                    }
                }
            } else {
                // A step is a fixed schedule of unit steps, each taken by the lanes it applies
                // to (reference traversal: node pop, triangle, triangle, node pop, triangle,
                // node pop, triangle; fast traversal: node, node, triangle, node, node,
                // triangle).  A lane's tests keep the traversal's order; only more of them
                // share one wave iteration, whose refill and set-up bookkeeping they amortise
                // (sphere grid: 527 Msamples/s with one unit per iteration, 632 with the
                // schedule; longer ones measured no faster, a third triangle in a row slower;
                // the fast traversal's schedule: 1396 -> 1459 on the sphere grid).
#pragma unroll
                for (int u = 0; u < (kAccel ? kStepUnitsFast : kStepUnits); ++u) {
                    if ((kAccel ? kStepScheduleFast[u] : kStepSchedule[u]) == 'N') {
                        if (t.leaf == kNoLeaf && (u == 0 || !trav_needs_setup(t, st))) {
                            bool inner;
                            if (kAccel) {
                                inner = ftrav_node<kDiag>(S, t, st);
                            } else if (t.fast) {
                                inner = trav_node<true, kDiag>(S, t, st);
                            } else {
                                inner = trav_node<false, kDiag>(S, t, st);
                            }
                            if (u == 0) kind = inner ? kStepInner : kStepCull;
                        }
                    } else if (t.leaf != kNoLeaf) {
                        if (kind == 0xFFu) kind = kStepLeaf;
                        if (kAccel) {
                            ftrav_tri<kAny, kDiag>(S, t, st);
                        } else {
                            trav_tri<kDiag>(S, t);
                        }
                    }
                }
            }
        }
        if (kCount) {
            w.it_wave++;
            w.st_setup += kind == kStepSetup;
            w.st_cull += kind == kStepCull;
            w.st_inner += kind == kStepInner;
            w.st_leaf += kind == kStepLeaf;
            w.it_setup += __ballot(kind == kStepSetup) != 0;
            w.it_inner += __ballot(kind == kStepInner || kind == kStepCull) != 0;
            w.it_leaf += __ballot(kind == kStepLeaf) != 0;
        }
        if (active) {
            if (finished) {
                active = false;
                const Hit& h = t.R;
                if (kMode == 0) {
                    // one 16-byte record in queue order, so that shading gets the instance,
                    // triangle and barycentrics with a single coalesced load
                    b.hit_rec[item] = make_uint4(static_cast<uint32_t>(h.inst), h.tri, __float_as_uint(h.u),
                                                 __float_as_uint(h.v));
                    b.hit_pos[item] = f4(h.wpos, h.dist);
                    if (kDiag && c.debug && h.inst >= 0) {  // the winning instance's tests (Tracer.cpp:170-175)
                        const uint32_t p = bounce == 0 ? item : q[item];
                        uint4 di = b.dbg_i[p];
                        if (bounce == 0) {
                            di.x = h.bvh_tests;
                            di.y = h.tri_tests;
                            uint4 dj = b.dbg_j[p];
                            dj.z = h.bv & 0xffffu;
                            b.dbg_j[p] = dj;
                        }
                        di.z += h.bvh_tests;
                        di.w += h.tri_tests;
                        b.dbg_i[p] = di;
                    }
                } else if (kMode == 1) {
                    const float4 so = b.sh_o[item];
                    if (h.inst < 0 || h.dist > so.w) {  // Tracer.cpp:322
                        const uint32_t p = __float_as_uint(b.sh_d[item].w);
                        const float4 cc = b.sh_c[item];
                        const uint32_t kind = __float_as_uint(cc.w);  // direct, indirect, environment sample
                        float4* acc = kind == 0u ? b.acc_d : (kind == 1u ? b.acc_i : b.acc_n);
                        const float4 a = acc[p];
                        acc[p] = make_float4(a.x + cc.x, a.y + cc.y, a.z + cc.z, 0.0f);
                        // one shadow ray per path and bounce: no other lane writes this word
                        if (c.diff) b.vis[p] |= 1u << bounce;
                    }
                } else if (kMode == 3) {
                    io.occ[item] = (h.inst >= 0 && !(h.dist > io.rays[item * 7 + 6])) ? 1u : 0u;
                } else {
                    mt_hit r;
                    r.instance = h.inst;
                    r.triangle = h.tri;
                    r.bv_index = h.bv;
                    r.u = h.u;
                    r.v = h.v;
                    r.distance = h.inst >= 0 ? h.dist : 0.0f;
                    r.world_pos[0] = h.wpos.x;
                    r.world_pos[1] = h.wpos.y;
                    r.world_pos[2] = h.wpos.z;
                    r.bvh_tests = h.bvh_tests;
                    r.triangle_tests = h.tri_tests;
                    io.out[item] = r;
                }
            }
            need = nm;
        }
    }
    if (kCount) {
        wave_add_u64(b.work + 0, w.aabb);
        wave_add_u64(b.work + 1, w.enter);
        wave_add_u64(b.work + 2, w.pops);
        wave_add_u64(b.work + 3, w.tris);
        wave_add_u64(b.work + 8, w.st_setup);
        wave_add_u64(b.work + 9, w.st_cull);
        wave_add_u64(b.work + 10, w.st_inner);
        wave_add_u64(b.work + 11, w.st_leaf);
        if (lane_id() == 0) {  // wave-level counts are identical on every lane
            atomicAdd(b.work + 12, static_cast<unsigned long long>(w.it_wave));
            atomicAdd(b.work + 13, static_cast<unsigned long long>(w.it_setup));
            atomicAdd(b.work + 14, static_cast<unsigned long long>(w.it_inner));
            atomicAdd(b.work + 15, static_cast<unsigned long long>(w.it_leaf));
            atomicAdd(b.work + 7, static_cast<unsigned long long>(w.it_mask));
        }
    }
}

template <bool kCount, int kMode, uint32_t kDepth, int kAccel>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(MT_TRACE_WAVES_PER_SIMD, 8))) void k_trace(DevScene S, ChunkParams c, PathBuffers b,
                                                  uint32_t bounce, TestIO io) {
    __shared__ uint2 lds_stack[kDepth * kBlock];
    trace_body<kCount, kMode, kDepth, kAccel>(S, c, b, bounce, io, lds_stack,
                                              blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// The instance masks of a queue's rays, computed before the reference traversal of a scene of
// at most 64 instances: bit j of slot i's mask is set when ray i enters instance j's world box
// (Intersect.cpp:206-216 tests every instance's world box).  The traversal kernel would
// otherwise test the 64 boxes in its refill steps with only the refilled lanes active.
template <int kMode>
__global__ __launch_bounds__(kBlock) void k_instance_masks(DevScene S, PathBuffers b, uint32_t bounce) {
    const uint32_t count = counter(b, (kMode == 0 ? kCntExtend : kCntShadow) + bounce);
    Work w{};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        V3 o, d;
        if (kMode == 0) {
            o = xyz(b.ray_o[bounce & 1][i]);
            d = xyz(b.ray_d[bounce & 1][i]);
        } else {
            o = xyz(b.sh_o[i]);
            d = xyz(b.sh_d[i]);
        }
        const V3 invd = {1.0f / d.x, 1.0f / d.y, 1.0f / d.z};  // as trav_begin
        b.inst_mask[i] = (finite3(invd) && finite3(o)) ? instance_mask<false, true>(S, 0, o, invd, w)
                                                       : instance_mask<false, false>(S, 0, o, invd, w);
    }
}

template <int kMode, int kAccel>
void launch_trace_as(uint32_t depth, bool count, uint32_t blocks, hipStream_t s, const DevScene& S,
                     const ChunkParams& c, const PathBuffers& b, uint32_t k, TestIO io) {
    if constexpr (kMode >= 2) {  // test hooks: no counting build, one stack depth
        hipLaunchKernelGGL((k_trace<false, kMode, 16, kAccel>), dim3(blocks), dim3(kBlock), 0, s, S, c, b, k, io);
        return;
    } else {
#define MT_TRACE_CASE(D)                                                                          \
    case D:                                                                                       \
        if (count)                                                                                \
            hipLaunchKernelGGL((k_trace<true, kMode, D, kAccel>), dim3(blocks), dim3(kBlock), 0, s, S, c, b, k, io); \
        else                                                                                      \
            hipLaunchKernelGGL((k_trace<false, kMode, D, kAccel>), dim3(blocks), dim3(kBlock), 0, s, S, c, b, k, io); \
        break;
        if (kMode <= 1 && !kAccel && !count && b.inst_mask)
            hipLaunchKernelGGL(k_instance_masks<kMode>, dim3(blocks), dim3(kBlock), 0, s, S, b, k);
        switch (depth) {
            MT_TRACE_CASE(4)
            MT_TRACE_CASE(8)
            MT_TRACE_CASE(12)
            default:
            MT_TRACE_CASE(16)
        }
#undef MT_TRACE_CASE
    }
}

// the traversal is chosen per chunk (ChunkParams::accel)
template <int kMode>
void launch_trace(uint32_t depth, bool count, uint32_t blocks, hipStream_t s, const DevScene& S,
                  const ChunkParams& c, const PathBuffers& b, uint32_t k, TestIO io = {}) {
    count = count || c.debug;  // debug views read the per-hit diagnostics of the counting build
    if (c.accel)
        launch_trace_as<kMode, 1>(depth, count, blocks, s, S, c, b, k, io);
    else
        launch_trace_as<kMode, 0>(depth, count, blocks, s, S, c, b, k, io);
}

// Material::SampleMaterial (Material.hpp:82-135) plus the differentiable overrides.
struct MatSample {
    V3 base, ntex, emission;
    float metal, rough;
    // adjoint bookkeeping: scale of each parameter in the sampled value and the texel taps of
    // the base-colour texture override
    float rough_scale, metal_scale;
    V3 em_scale;
    int tex;            // 1 when the base colour comes from the differentiable texture
    uint32_t tap[4];    // texel indices
    float tw[4];        // bilinear weights (Image.cpp:26-44 weights in uv units)
};

__device__ __forceinline__ void sample_param_texture(const DiffParams& dp, float uvx, float uvy,
                                                     MatSample& ms) {
    const int w = static_cast<int>(dp.tex_w), h = static_cast<int>(dp.tex_h);
    const float x = uvx - floorf(uvx);
    const float y = uvy - floorf(uvy);
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
    ms.tap[0] = y0 * w + x0;
    ms.tap[1] = y0 * w + x1;
    ms.tap[2] = y1 * w + x0;
    ms.tap[3] = y1 * w + x1;
    ms.tw[0] = (1.0f - sx) * (1.0f - sy);
    ms.tw[1] = sx * (1.0f - sy);
    ms.tw[2] = (1.0f - sx) * sy;
    ms.tw[3] = sx * sy;
    float c[3];
    for (int ch = 0; ch < 3; ++ch) {
        const float c00 = dp.tex[ms.tap[0] * 3 + ch], c10 = dp.tex[ms.tap[1] * 3 + ch];
        const float c01 = dp.tex[ms.tap[2] * 3 + ch], c11 = dp.tex[ms.tap[3] * 3 + ch];
        const float c0 = mixf(c00, c10, sx), c1 = mixf(c01, c11, sx);
        c[ch] = mixf(c0, c1, sy);
    }
    ms.base = {c[0], c[1], c[2]};
    ms.tex = 1;
}

template <int kDiff>
__device__ void eval_material(const DevScene& S, const DiffParams& dp, uint32_t mi, float u,
                              float v, MatSample& ms) {
    // the adjoint bookkeeping (tex, *_scale, taps) is written only in the differentiable
    // builds: in the plain ones its stores stayed in private memory (20 B of scratch per lane)
    const DevMaterial& m = S.materials[mi];
    if constexpr (kDiff != 0) ms.tex = 0;
    ms.base = V3{m.base[0], m.base[1], m.base[2]};
    if (kDiff && dp.tex && mi >= dp.tex_first && mi - dp.tex_first < dp.tex_count) {
        sample_param_texture(dp, u, v, ms);
    } else if (m.tex[0].image >= 0) {
        const float4 t = sample_tex(S, m.tex[0], u, v);
        ms.base = V3{t.x * m.base[0], t.y * m.base[1], t.z * m.base[2]};
    }
    const float rf = (kDiff && dp.rough) ? dp.rough[mi] : m.roughness;
    const float mf = (kDiff && dp.metal) ? dp.metal[mi] : m.metallic;
    if (m.tex[1].image >= 0) {
        const float4 t = sample_tex(S, m.tex[1], u, v);
        ms.metal = t.z * mf;
        ms.rough = t.y * rf;
        if constexpr (kDiff != 0) {
            ms.metal_scale = t.z;
            ms.rough_scale = t.y;
        }
    } else {
        ms.metal = mf;
        ms.rough = rf;
        if constexpr (kDiff != 0) ms.metal_scale = ms.rough_scale = 1.0f;
    }
    if (m.tex[2].image >= 0) {
        const float4 t = sample_tex(S, m.tex[2], u, v);
        const float ns = m.normal_scale;
        ms.ntex = normalize(V3{t.x * ns * 2.0f - 1.0f, t.y * ns * 2.0f - 1.0f, t.z * ns * 2.0f - 1.0f});
    } else {
        ms.ntex = V3{0.0f, 0.0f, 1.0f};
    }
    const V3 em = (kDiff && dp.emis) ? V3{dp.emis[mi * 3], dp.emis[mi * 3 + 1], dp.emis[mi * 3 + 2]}
                                     : V3{m.emissive[0], m.emissive[1], m.emissive[2]};
    if (m.tex[3].image >= 0) {
        const float4 t = sample_tex(S, m.tex[3], u, v);
        if constexpr (kDiff != 0) ms.em_scale = V3{t.x, t.y, t.z};
        ms.emission = mul(V3{t.x, t.y, t.z}, em);
    } else {
        if constexpr (kDiff != 0) ms.em_scale = mk(1.0f);
        ms.emission = em;
    }
}

// per-wave reduction of a float addend per key (a material id), then one atomic per key
__device__ __forceinline__ void wave_atomic_by_key(float* base, uint32_t key, float v, bool pred) {
    uint64_t todo = __ballot(pred);
    while (todo) {
        const int leader = __ffsll(static_cast<unsigned long long>(todo)) - 1;
        const uint32_t lk = __shfl(key, leader);
        const bool mine = pred && key == lk;
        float x = mine ? v : 0.0f;
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
        if (static_cast<int>(lane_id()) == leader && x != 0.0f) atomicAdd(base + lk, x);
        todo &= ~__ballot(mine);
    }
}

// Per-material gradient sums of one block of the adjoint shading kernel, in LDS: roughness,
// metallic and the three emission channels of material m at [5m .. 5m + 4].  Each hit adds
// its five addends with LDS atomics; the block adds its nonzero sums to the gradient arrays
// once, after its last path (one wave-level reduction and global atomic per material present
// in a wave, five times per hit, had cost the adjoint 9 of its 37 ms on the inverse-rendering
// sphere grid).  Scenes with more materials keep the wave-level reduction.
constexpr uint32_t kGradLdsMats = 256;
__device__ __forceinline__ float* grad_lds() {
    __shared__ float s[kGradLdsMats * 5];
    return s;
}
__device__ __forceinline__ bool grad_in_lds(const DiffGradPtrs& gr) {
    return gr.materials > 0 && gr.materials <= kGradLdsMats && (gr.rough || gr.metal || gr.emis);
}

// Sum of three addends per key over runs of consecutive lanes with the same key, then one
// atomic per run and channel from the run's last lane.  Lanes that add nothing take a key of
// their own, so they end runs; a key that recurs apart adds more than once, which is still
// exact.  The samples of one pixel are neighbouring lanes and mostly hit the same texels, so
// this removes most of the adjoint's per-lane texel atomics (which wait on each other at one
// address).  Every lane of the wave calls it.
__device__ __forceinline__ void wave_run_add3(float* base, uint32_t key, V3 v, bool pred) {
    const uint32_t lane = lane_id();
    const uint32_t kk = pred ? key : (0x80000000u | lane);  // texel indices stay below 2^31
    const uint32_t kprev = __shfl_up(kk, 1);
    const uint64_t starts = __ballot(lane == 0 || kprev != kk);
    const uint32_t run = static_cast<uint32_t>(__popcll(starts & ((2ull << lane) - 1ull)));
    float x = pred ? v.x : 0.0f, y = pred ? v.y : 0.0f, z = pred ? v.z : 0.0f;
    for (uint32_t off = 1; off < 64; off <<= 1) {  // segmented inclusive scan over the runs
        const uint32_t ro = __shfl_up(run, off);
        const float xo = __shfl_up(x, off), yo = __shfl_up(y, off), zo = __shfl_up(z, off);
        if (lane >= off && ro == run) {
            x += xo;
            y += yo;
            z += zo;
        }
    }
    const uint32_t rnext = __shfl_down(run, 1);
    if (pred && (lane == 63 || rnext != run)) {
        if (x != 0.0f) atomicAdd(base + key * 3 + 0, x);
        if (y != 0.0f) atomicAdd(base + key * 3 + 1, y);
        if (z != 0.0f) atomicAdd(base + key * 3 + 2, z);
    }
}

// a lane's texel-gradient addend of this bounce: the taps and weights of the vertex's base
// colour texture and d(loss)/d(base colour), added to the texels at the loop's end (the base
// colour is b' = 0.96 * texel + 0.04, so d/d(texel) = 0.96 * tap weight)
struct TexelGrad {
    uint32_t tap[4];
    float tw[4];
    V3 g;
    bool on;
};
__device__ __forceinline__ void set_texel_grad(TexelGrad& tg, const DiffGradPtrs& gr, const MatSample& ms, V3 g) {
    if (!gr.tex || !ms.tex) return;
    for (int t = 0; t < 4; ++t) {
        tg.tap[t] = ms.tap[t];
        tg.tw[t] = ms.tw[t];
    }
    tg.g = g;
    tg.on = true;
}

// ------------------------------------------------------------------ environment samples
// (training paths of a cached render with cache_nee)
constexpr int kPbCos = 129, kPbRough = 65;
constexpr uint32_t kDimNee = 0x9E000000u;  // dimensions of the training paths' extra draws

// Chance that the specular lobe's half-vector draw (Tracer.cpp:49-72) reflects the view below
// the surface, so that a cosine direction replaces it: entry (ri, ci) for roughness
// ri / (kPbRough - 1) and cos(theta_o) = ci / (kPbCos - 1).  The half vector's polar angle t
// follows from u1 (tan^2 t = a^2 u1 / max(1 - u1, 1e-4), a = max(r^2, 0.0025)) and its azimuth
// is uniform; flipping h toward the view does not move the reflection, which lies below the
// surface where 2 (h.v)(h.n) <= cos(theta_o), i.e. where
// sin(2t) sin(theta_o) cos(phi) <= -cos(theta_o) cos(2t): a fraction 1 - acos(c) / pi of the
// azimuths, c = the right side over sin(2t) sin(theta_o).  Midpoint rule over u1, in double.
// One wave per entry (launched with 64 threads per block, one block per entry): each lane
// takes every 64th midpoint and the wave adds the lanes' sums.
__global__ __launch_bounds__(64) void k_pbelow_table(float* table) {
    const int e = static_cast<int>(blockIdx.x);
    if (e >= kPbCos * kPbRough) return;
    const int ri = e / kPbCos, ci = e % kPbCos;
    const double co = static_cast<double>(ci) / (kPbCos - 1), r = static_cast<double>(ri) / (kPbRough - 1);
    const double so = sqrt(fmax(0.0, 1.0 - co * co));
    const double a = fmax(r * r, 0.0025), a2 = a * a;
    constexpr int n = 8192;
    double below = 0.0;
    for (int i = static_cast<int>(threadIdx.x); i < n; i += 64) {
        const double u = (i + 0.5) / n;
        const double t2 = a2 * u / fmax(1.0 - u, 1e-4);
        const double ct = 1.0 / sqrt(1.0 + t2), st = sqrt(fmax(0.0, 1.0 - ct * ct));
        const double num = -co * (ct * ct - st * st), den = 2.0 * st * ct * so;
        if (den > 0.0) below += 1.0 - acos(fmin(1.0, fmax(-1.0, num / den))) / 3.14159265358979323846;
        else below += (num >= 0.0) ? 1.0 : 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) below += __shfl_xor(below, off);
    if (threadIdx.x == 0) table[ri * kPbCos + ci] = static_cast<float>(below / n);
}

// The environment's sampling tables (the luminance-weighted 2-D CDF of env_sample/env_pdf):
// block j builds row j's conditional CDF, cond[j * (W + 1) + i] for i = 0..W, from the weights
// (lum(i, j) + blend * (lum(i + 1, j) + lum(i, j + 1) + lum(i + 1, j + 1))) * sin(theta_j), with
// a contiguous run of texels per thread and a scan of the threads' sums, in double; rows[j] =
// the row's total.  k_env_marginal then scans the rows into marg[0..H] and total[0].
__device__ __forceinline__ double env_lum(const float4* env, uint32_t W, uint32_t H, uint32_t i, uint32_t j) {
    const float4 t = env[static_cast<size_t>(min(j, H - 1)) * W + min(i, W - 1)];
    const double m = (static_cast<double>(t.x) + t.y + t.z) / 3.0;
    return (m > 0.0 && isfinite(m)) ? m : 0.0;
}
__global__ __launch_bounds__(256) void k_env_rows(const float4* env, uint32_t W, uint32_t H, double blend,
                                                  float* cond, double* rows) {
    __shared__ double part[256];
    const uint32_t j = blockIdx.x;
    const uint32_t per = (W + 255) / 256;
    const uint32_t i0 = min(threadIdx.x * per, W), i1 = min(i0 + per, W);
    const double sth = sin(3.14159265358979323846 * (j + 0.5) / H);
    double sum = 0.0;
    for (uint32_t i = i0; i < i1; ++i)
        sum += (env_lum(env, W, H, i, j) + blend * (env_lum(env, W, H, i + 1, j) + env_lum(env, W, H, i, j + 1) +
                                                    env_lum(env, W, H, i + 1, j + 1))) * sth;
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 256; off <<= 1) {  // inclusive scan of the threads' sums
        const double v = threadIdx.x >= off ? part[threadIdx.x - off] : 0.0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    const double total = part[255];
    float* row = cond + static_cast<size_t>(j) * (W + 1);
    double cum = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0.0;
    for (uint32_t i = i0; i < i1; ++i) {
        cum += (env_lum(env, W, H, i, j) + blend * (env_lum(env, W, H, i + 1, j) + env_lum(env, W, H, i, j + 1) +
                                                    env_lum(env, W, H, i + 1, j + 1))) * sth;
        row[i + 1] = total > 0.0 ? static_cast<float>(cum / total) : static_cast<float>(i + 1) / W;
    }
    if (threadIdx.x == 0) {
        row[0] = 0.0f;
        rows[j] = total;
    }
    __syncthreads();
    if (threadIdx.x == 0) row[W] = 1.0f;
}
__global__ void k_env_marginal(const double* rows, uint32_t H, float* marg, double* total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t j = 0; j < H; ++j) t += rows[j];
    double acc = 0.0;
    marg[0] = 0.0f;
    for (uint32_t j = 0; j < H; ++j) {
        acc += rows[j];
        marg[j + 1] = t > 0.0 ? static_cast<float>(acc / t) : 0.0f;
    }
    marg[H] = 1.0f;
    total[0] = t;
}

__device__ __forceinline__ float pbelow(const float* tab, float co, float r) {
    const float x = clampf(co, 0.0f, 1.0f) * static_cast<float>(kPbCos - 1);
    const float y = clampf(r, 0.0f, 1.0f) * static_cast<float>(kPbRough - 1);
    const int x0 = min(static_cast<int>(x), kPbCos - 2), y0 = min(static_cast<int>(y), kPbRough - 2);
    const float fx = x - static_cast<float>(x0), fy = y - static_cast<float>(y0);
    const float* t = tab + y0 * kPbCos + x0;
    return mixf(mixf(t[0], t[1], fx), mixf(t[kPbCos], t[kPbCos + 1], fx), fy);
}

// The true density of the reference's bounce direction at l (Tracer.cpp:177-200): the mixture
// pdf it divides by, plus the cosine directions that replace specular draws reflected below.
__device__ __forceinline__ float bounce_density(const ChunkParams& c, V3 l, V3 view, V3 N, float ps, float rough,
                                                float pdf_ref) {
    const float nl = dot(N, l);
    if (!(nl > 0.0f)) return 0.0f;
    return pdf_ref + ps * pbelow(c.pbelow, dot(N, view), rough) * nl * (1.0f / kPi);
}

// The environment's sampling distribution: cell (i, j) of its W x H texel grid with chance
// (marg[j+1] - marg[j]) (cond_j[i+1] - cond_j[i]), uniform in (u, v) inside the cell, where
// u = (atan2(d.z, d.x) + pi) / 2pi (+ the rotation offset) and v = acos(d.y) / pi as in
// sample_env; solid-angle density = cell chance * W * H / (2 pi^2 sin(pi v)).
__device__ __forceinline__ float env_cell_chance(const ChunkParams& c, uint32_t W, int i, int j) {
    const float* row = c.env_cond + static_cast<size_t>(j) * (W + 1);
    return (c.env_marg[j + 1] - c.env_marg[j]) * (row[i + 1] - row[i]);
}

__device__ float env_pdf(const ChunkParams& c, const DevScene& S, const ChunkParams& J, V3 d) {
    const int W = static_cast<int>(S.env_w), H = static_cast<int>(S.env_h);
    const float l = len(d);
    const float cy = d.y / l;
    float u = (em::atan2(d.z, d.x) + kPi) / (2.0f * kPi);
    if (J.hdri_rot_apply) u = u * 1.0f + J.hdri_rot_offset;
    const float v = em::acos(cy) / kPi;
    const float x = u - floorf(u), y = v - floorf(v);
    const int i = min(max(static_cast<int>(x * static_cast<float>(W)), 0), W - 1);
    const int j = min(max(static_cast<int>(y * static_cast<float>(H)), 0), H - 1);
    const float sphi = __builtin_sqrtf(gmax(0.0f, 1.0f - cy * cy));
    if (!(sphi > 0.0f)) return 0.0f;
    return env_cell_chance(c, S.env_w, i, j) * static_cast<float>(W) * static_cast<float>(H) /
           (2.0f * kPi * kPi * sphi);
}

__device__ bool env_sample(const ChunkParams& c, const DevScene& S, const ChunkParams& J, float u1, float u2,
                           float u3, float u4, V3& dir, float& pdf) {
    const int W = static_cast<int>(S.env_w), H = static_cast<int>(S.env_h);
    int lo = 0, hi = H;  // the last row j with marg[j] <= u1 (marg[0] = 0, marg[H] = 1)
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (c.env_marg[mid] <= u1) lo = mid; else hi = mid;
    }
    const int j = lo;
    const float* row = c.env_cond + static_cast<size_t>(j) * (W + 1);
    lo = 0;
    hi = W;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row[mid] <= u2) lo = mid; else hi = mid;
    }
    const int i = lo;
    const float chance = (c.env_marg[j + 1] - c.env_marg[j]) * (row[i + 1] - row[i]);
    if (!(chance > 0.0f)) return false;
    float u = (static_cast<float>(i) + u3) / static_cast<float>(W);
    const float v = (static_cast<float>(j) + u4) / static_cast<float>(H);
    if (J.hdri_rot_apply) u -= J.hdri_rot_offset;
    u -= floorf(u);
    float sp, cp, sth, cth;
    em::sincos(v * kPi, sp, cp);
    em::sincos(u * (2.0f * kPi) - kPi, sth, cth);
    if (!(sp > 0.0f)) return false;
    dir = V3{sp * cth, cp, sp * sth};
    pdf = chance * static_cast<float>(W) * static_cast<float>(H) / (2.0f * kPi * kPi * sp);
    return true;
}

// The shading of one bounce for every path in the extension queue (Tracer.cpp:134-340).
// kDiff 0: plain render; 1: differentiable forward (parameter overrides, shadow visibility
// bits); 2: adjoint replay -- the same decisions, no shadow rays (visibility comes from the
// forward pass), and gradients of the per-path radiance accumulated with atomics:
//   emission term  T*ebi*e          -> d/d(emission)
//   light term     T*f_l*NL*fall*C/p -> d f_l / d(base, roughness, metallic)
//   throughput w = f*cos/pdf        -> (L - S) * (dw/dtheta) / w with the sampling detached
//   sky term       base'_{k-1}*T*sky -> d/d(base of the previous hit)
// where T is the radiance throughput of the reference and S the radiance collected so far.
// kExt compiles in the debug-view records and the radiance-cache bookkeeping (plain renders
// that use neither run the leaner kExt = false build: fewer registers, more waves in flight);
// kLights = false is the build for scenes without point lights (no light choice, no shadow ray).
#ifndef MT_SHADE_WAVES_PER_SIMD
#define MT_SHADE_WAVES_PER_SIMD 4
#endif
// blk / nblk: this block and the blocks sharing the work (the grid, or one lane group of the
// batch's persistent kernel)
template <int kDiff, bool kBatch, bool kExt, bool kLights>
__device__ __forceinline__ void shade_body(const DevScene& S, const ChunkParams& c, const PathBuffers& b,
                                           uint32_t k, uint32_t blk, uint32_t nblk) {
    const uint32_t count = counter(b, kCntExtend + k);
    const uint32_t* qin = b.queue[k & 1];
    uint32_t* qout = b.queue[(k + 1) & 1];
    const uint32_t stride = nblk * blockDim.x;
    uint32_t hits = 0, misses = 0;
    const DiffParams& dp = c.dp;
    const DiffGradPtrs& gr = c.gr;
    if constexpr (kDiff == 2) {
        if (grad_in_lds(gr)) {
            float* sg = grad_lds();
            for (uint32_t i = threadIdx.x; i < gr.materials * 5u; i += blockDim.x) sg[i] = 0.0f;
            __syncthreads();
        }
    }
    for (uint32_t base = blk * blockDim.x; base < count; base += stride) {
        const uint32_t idx = base + threadIdx.x;
        const bool valid = idx < count;
        bool cont = false, shadow = false;
        V3 next_o = mk(0.0f), next_d = mk(0.0f);  // the continuing path's next ray
        V3 next_term = mk(0.0f), next_base = mk(0.0f), next_trad = mk(0.0f);
        float next_q = 0.0f;  // training paths (cache_nee): true density of the bounce direction
        float4 sho{}, shd{}, shc{};
        uint32_t p = 0;
        // cache query of a path that stops here
        bool query = false;
        float4 qpos{}, qdir{}, qthr{};  // and the throughput at the query vertex
        // adjoint addends for the per-material wave reductions
        uint32_t gmat = 0;
        float g_r = 0.0f, g_m = 0.0f;
        V3 g_e = mk(0.0f);
        bool gvalid = false;
        TexelGrad tgr{};
        if (valid) {
            p = k == 0 ? idx : qin[idx];
            // the render this path belongs to (batched: its job's settings)
            const ChunkParams& J = job_params<kBatch>(c, b, p);
            const bool last = k + 1 >= J.bounces;
            // adjoint replay: dLoss/dL of this path and what the forward pass recorded for it
            V3 dl = mk(0.0f), ltot = mk(0.0f);
            uint32_t visbits = 0;
            if (kDiff == 2) {
                uint32_t px, sm;
                path_coords(c, p, px, sm);
                const float4 rec = c.path_store[static_cast<size_t>(px) * c.spp + sm];
                ltot = xyz(rec);
                visbits = __float_as_uint(rec.w);
                const float* gi = c.grad_img + static_cast<size_t>(px) * 3;
                dl = scale(V3{gi[0], gi[1], gi[2]}, c.inv_spp);
            }
            const V3 d = xyz(b.ray_d[k & 1][idx]);
            const uint4 hrec = b.hit_rec[idx];
            const int inst = static_cast<int>(hrec.x);
            const V3 trad = k == 0 ? mk(1.0f) : xyz(b.t_rad[k & 1][idx]);
            if (kDiff == 0 && !kBatch && kExt && c.cache && c.cache_after && k > c.cache_depth) {
                // the training record of the previous vertex: everything collected up to and
                // including its light sample (the rest is what the bounce ray brought)
                const uint32_t slot = b.train_slot[p];
                const uint32_t j = k - 1 - c.cache_depth;
                if (slot < c.train_cap && j < c.train_nrec) {
                    const size_t r = static_cast<size_t>(slot) * c.train_nrec + j;
                    b.tr_s[r] = f4(add(add(add(xyz(b.acc_e[p]), xyz(b.acc_d[p])), xyz(b.acc_i[p])),
                                       xyz(b.acc_s[p])), 0.0f);
                    const float4 tw = b.tr_w[r];
                    b.tr_w[r] = make_float4(tw.x, tw.y, tw.z, 1.0f);
                }
            }
            if (kDiff == 0 && !kBatch && kExt && c.debug) {  // one more path step; the second one's ray
                uint4 dj = b.dbg_j[p];
                dj.w += 4u;
                if (k == 1) {
                    dj.w |= 2u;
                    b.dbg_e[p] = f4(d, b.dbg_e[p].w);
                }
                b.dbg_j[p] = dj;
            }
            if (inst < 0) {  // miss: sky term (Tracer.cpp:244-252)
                misses++;
                const V3 col = (k > 0) ? xyz(b.prev_base[k & 1][idx]) : mk(1.0f);
                const float mult = (k == 0) ? J.hdri_primary : J.hdri_secondary;
                V3 sk = scale(sample_env(S, d, J.hdri_rot_apply, J.hdri_rot_offset), mult);
                if (kDiff == 0 && !kBatch && kExt && c.env_pc > 0.0f && k > c.cache_depth &&
                    b.train_slot[p] < c.train_cap) {
                    // a training path's bounce ray from a vertex that also drew environment
                    // samples: balance heuristic against them (the true density of this
                    // direction was stored with the previous base colour)
                    const float q = b.prev_base[k & 1][idx].w;
                    const float pe = c.env_pc * env_pdf(c, S, J, d);
                    if (pe > 0.0f) sk = scale(sk, q / (q + pe));
                }
                b.acc_s[p] = f4(mul(mul(col, trad), sk), 0.0f);
                if (kDiff == 2 && k > 0) {
                    const float4 pv = b.adj_prev[p];
                    const uint32_t pm = __float_as_uint(pv.z);
                    if (pm != 0xFFFFFFFFu) {
                        MatSample pms;
                        eval_material<2>(S, dp, pm, pv.x, pv.y, pms);
                        set_texel_grad(tgr, gr, pms, mul(dl, mul(trad, sk)));
                    }
                }
            } else if (kDiff == 0 && !kBatch && kExt && c.cache && !c.cache_after && k == c.cache_depth &&
                       b.train_slot[p] == kNoSlot) {
                // the cache supplies the radiance leaving this vertex toward the previous one
                hits++;
                query = true;
                qpos = b.hit_pos[idx];
                qdir = f4(neg(d), 0.0f);
                qthr = f4(trad, 0.0f);
            } else {
                hits++;
                if (kDiff == 0 && !kBatch && kExt && c.cache && k >= c.cache_depth) {
                    const uint32_t slot = b.train_slot[p];
                    if (slot < c.train_cap && k - c.cache_depth < c.train_nrec) {
                        // training vertex: radiance collected so far, throughput at the vertex
                        const size_t r = static_cast<size_t>(slot) * c.train_nrec + (k - c.cache_depth);
                        const V3 sofar = add(add(add(xyz(b.acc_e[p]), xyz(b.acc_d[p])), xyz(b.acc_i[p])),
                                             xyz(b.acc_s[p]));
                        b.tr_pos[r] = b.hit_pos[idx];
                        b.tr_dir[r] = f4(neg(d), 0.0f);
                        // after-light records become valid when the path goes on (next bounce)
                        b.tr_w[r] = f4(trad, c.cache_after ? 0.0f : 1.0f);
                        b.tr_s[r] = f4(sofar, 0.0f);
                    }
                }
                uint32_t pix, s;
                coords<kBatch>(c, b, p, pix, s);
                const uint32_t key = path_key(J.seed_key, pix, s);
                const uint32_t dim0 = 2u + k * J.light_stride;
                const V3 o = xyz(b.ray_o[k & 1][idx]);
                const float4 hp = b.hit_pos[idx];
                const V3 wpos = xyz(hp);
                const float dist = hp.w;
                const float2 huv = make_float2(__uint_as_float(hrec.z), __uint_as_float(hrec.w));
                const uint32_t tri = hrec.y;
                const DevInstance* I = S.instances + inst;
                // RayHitGeometryInfo (Ray.cpp:7-48)
                const float w = 1.0f - huv.x - huv.y;
                const float* nn = S.normals + 9ull * (I->attr_base + tri);
                const V3 n0 = {nn[0], nn[1], nn[2]}, n1 = {nn[3], nn[4], nn[5]}, n2 = {nn[6], nn[7], nn[8]};
                const V3 ni = normalize(add(add(scale(n0, w), scale(n1, huv.x)), scale(n2, huv.y)));
                float M[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) M[j] = I->M[j];
                const V3 gN = normalize(xform(M, ni, 0.0f));
                float tu = 0.0f, tv = 0.0f;
                if (I->mesh_flags & kMeshHasUv) {
                    const float* uv = S.uvs + 6ull * (I->uv_base + tri);
                    tu = (w * uv[0] + huv.x * uv[2]) + huv.y * uv[4];
                    tv = (w * uv[1] + huv.x * uv[3]) + huv.y * uv[5];
                }
                const uint32_t mi = I->material;
                MatSample ms;
                eval_material<kDiff>(S, dp, mi, tu, tv, ms);
                const V3 basec = add(scale(ms.base, 0.96f), mk(0.04f));  // Tracer.cpp:159
                const float metal = ms.metal, rough = ms.rough;
                const V3 N = shading_normal(gN, ms.ntex);
                const V3 view = neg(d);
                // frozen sampling (differentiable forward only): every decision is taken with
                // the sampling parameter set, the values are evaluated with the other one
                V3 basec_s = basec;
                float metal_s = metal, rough_s = rough;
                const bool frozen = kDiff == 1 && c.frozen;
                if (frozen) {
                    MatSample mss;
                    eval_material<1>(S, c.dps, mi, tu, tv, mss);
                    basec_s = add(scale(mss.base, 0.96f), mk(0.04f));
                    metal_s = mss.metal;
                    rough_s = mss.rough;
                }
                // BRDF sampling (Tracer.cpp:177-200)
                const float ps = spec_probability(view, N, basec_s, metal_s);
                V3 nd;
                bool spec_lobe = false;
                float hcos = 0.0f;
                if (draw(key, dim0 + kDimLobe) < ps) {
                    spec_lobe = true;
                    const float u1 = draw(key, dim0 + kDimDir), u2 = draw(key, dim0 + kDimDir + 1);
                    if (!ggx_dir(u1, u2, d, N, rough_s, nd, hcos))
                        nd = cosine_dir(draw(key, dim0 + kDimFallback), draw(key, dim0 + kDimFallback + 1), N);
                } else {
                    nd = cosine_dir(draw(key, dim0 + kDimDir), draw(key, dim0 + kDimDir + 1), N);
                }
                const float pdf = (1.0f - ps) * diffuse_pdf(nd, N) + ps * specular_pdf(nd, view, N, rough_s);
                const V3 fb = eval_brdf(nd, view, N, basec, rough, metal);
                const float nl = gmax(dot(N, nd), 0.0f);
                V3 energy = (pdf > 0.0f) ? scale(fb, nl / pdf) : mk(0.0f);
                if (kDiff == 0 && !kBatch && kExt && c.debug && k == 0) {  // first-hit debug record
                    const V3 fr = fresnel(gmax(dot(N, view), 0.0f), base_refl(basec, metal));
                    b.dbg_a[p] = f4(basec, metal);
                    b.dbg_b[p] = make_float4(rough, gN.x, gN.y, gN.z);
                    b.dbg_c[p] = f4(N, len(sub(wpos, o)));
                    b.dbg_d[p] = f4(fr, (fb.x + fb.y) + fb.z);
                    b.dbg_e[p] = make_float4(0.0f, 0.0f, 0.0f, pdf);
                    uint4 dj = b.dbg_j[p];
                    dj.x = static_cast<uint32_t>(inst) & 0xffffu;
                    dj.y = tri & 0xffffu;
                    dj.w |= 1u;
                    b.dbg_j[p] = dj;
                }
                // termination bookkeeping of GeneratePath (Tracer.cpp:206-208)
                V3 energy_t = energy;
                if (frozen)
                    energy_t = (pdf > 0.0f) ? scale(eval_brdf(nd, view, N, basec_s, rough_s, metal_s), nl / pdf)
                                            : mk(0.0f);
                V3 tterm = mul(k == 0 ? mk(1.0f) : xyz(b.t_term[k & 1][idx]), energy_t);
                bool stop;
                float rr_q = 1.0f;
                if (J.mode == MT_MODE_FIXED) {
                    stop = false;
                    if (k + 1 >= J.rr_depth) {
                        const float q = clampf(gmax(gmax(tterm.x, tterm.y), tterm.z), 0.05f, 1.0f);
                        if (draw(key, dim0 + kDimRR) >= q) {
                            stop = true;
                        } else {
                            energy = divs(energy, q);
                            tterm = divs(tterm, q);
                            rr_q = q;
                        }
                    }
                } else {
                    stop = len(tterm) < J.terminate_energy;
                }
                // ProcessRayForward for this step (Tracer.cpp:254-339)
                const V3 e_term = mul(scale(trad, J.emission_base), ms.emission);
                // the sum starts at +0 and e_term is never -0 (throughput and emission are not
                // negative), so adding an all-zero e_term leaves it bit-identical: skip the
                // read-modify-write then (NaN components compare unequal and are added)
                if (e_term.x != 0.0f || e_term.y != 0.0f || e_term.z != 0.0f) {
                    const float4 ae = b.acc_e[p];
                    b.acc_e[p] = make_float4(ae.x + e_term.x, ae.y + e_term.y, ae.z + e_term.z, 0.0f);
                }
                V3 ssum = mk(0.0f);
                if (kDiff == 2) {
                    ssum = add(xyz(b.adj_s[p]), e_term);
                    gmat = mi;
                    gvalid = true;
                    g_e = mul(mul(dl, scale(trad, J.emission_base)), ms.em_scale);
                }
                // a training path's vertex from its query vertex on (cache_nee): its shadow ray
                // goes to an environment sample with probability env_pc, else to the
                // reference's light choice with weight 1 / (1 - env_pc)
                const bool nee_v = kDiff == 0 && !kBatch && kExt && c.env_pc > 0.0f && k >= c.cache_depth &&
                                   b.train_slot[p] < c.train_cap;
                const bool use_env = nee_v && (c.env_pc >= 1.0f || draw(key, kDimNee + k * 8u) < c.env_pc);
                if (nee_v) next_q = bounce_density(c, nd, view, N, ps, rough_s, pdf);
                int chosen = -1;
                float chosen_pdf = 0.0f, total = 0.0f;
                const uint32_t n_lights = kLights ? S.light_count : 0u;
                for (uint32_t li = 0; li < n_lights; ++li) {
                    const DevLight* L = S.lights + li;
                    const V3 lp = sub(V3{L->pos[0], L->pos[1], L->pos[2]}, wpos);
                    const float score = L->intensity / (dot(lp, lp) + 1.0f);
                    total += score;
                    if (draw(key, dim0 + kDimLightSel + li) * total < score) {
                        chosen = static_cast<int>(li);
                        chosen_pdf = score;
                    }
                }
                chosen_pdf /= total;
                V3 trad_next = trad;
                V3 g_base = mk(0.0f);
                if (chosen >= 0) {
                    const DevLight L = S.lights[chosen];
                    V3 lpos = {L.pos[0], L.pos[1], L.pos[2]};
                    if (L.type == MT_LIGHT_POINT || L.type == MT_LIGHT_SPOT) {  // Tracer.cpp:84-93
                        const float ph = draw(key, dim0 + kDimLightPt) * (2.0f * kPi - 0.0f) + 0.0f;
                        const float th = draw(key, dim0 + kDimLightPt + 1) * (kPi * 0.5f - (-kPi * 0.5f)) + (-kPi * 0.5f);
                        float sp, cp, st, ct;
                        em::sincos(ph, sp, cp);
                        em::sincos(th, st, ct);
                        const V3 rd = normalize(V3{cp * ct, st, sp * ct});
                        lpos = add(lpos, scale(rd, L.size));
                    }
                    const V3 lvec = sub(lpos, wpos);
                    const V3 ldir = normalize(lvec);
                    V3 lcol = {L.color[0], L.color[1], L.color[2]};
                    if (L.type == MT_LIGHT_SPOT) {  // Light.hpp:58-63
                        const float dd = dot(neg(ldir), V3{L.dir[0], L.dir[1], L.dir[2]});
                        lcol = scale(lcol, clampf((dd - L.outer) / (L.inner - L.outer), 0.0f, 1.0f));
                    }
                    const float llen = len(lvec);
                    const V3 fl = eval_brdf(ldir, view, N, basec, rough, metal);
                    const float nll = gmax(dot(N, ldir), 0.0f);
                    const float fall = 1.0f / (llen * llen + 1.0f);
                    const V3 le = scale(scale(fl, nll), fall);
                    const V3 contrib = divs(mul(mul(trad, le), lcol), chosen_pdf);
                    if (kDiff == 2) {
                        if ((visbits >> k) & 1u) {  // the forward pass found the light visible
                            ssum = add(ssum, contrib);
                            const BrdfGrad bg = eval_brdf_grad(ldir, view, N, basec, rough, metal);
                            const V3 wgt = divs(mul(mul(dl, trad), lcol), chosen_pdf);
                            const float sc = nll * fall;
                            g_base = add(g_base, scale(mul(wgt, bg.df_dbase), sc));
                            g_r += dot(scale(wgt, sc), bg.df_drough) * ms.rough_scale;
                            g_m += dot(scale(wgt, sc), bg.df_dmetal) * ms.metal_scale;
                        }
                        // throughput factor w = f*cos/pdf of this step scales everything after it
                        const V3 rem = sub(ltot, ssum);
                        const BrdfGrad bb = eval_brdf_grad(nd, view, N, basec, rough, metal);
                        if (pdf > 0.0f) {
                            const float sc = nl / pdf / rr_q;
                            const float cw[3] = {energy.x, energy.y, energy.z};
                            const float rm[3] = {rem.x, rem.y, rem.z}, dd3[3] = {dl.x, dl.y, dl.z};
                            const float db[3] = {bb.df_dbase.x, bb.df_dbase.y, bb.df_dbase.z};
                            const float dr[3] = {bb.df_drough.x, bb.df_drough.y, bb.df_drough.z};
                            const float dm[3] = {bb.df_dmetal.x, bb.df_dmetal.y, bb.df_dmetal.z};
                            float gb[3] = {0, 0, 0};
                            for (int ch = 0; ch < 3; ++ch) {
                                if (!(cw[ch] > 0.0f)) continue;
                                const float a = dd3[ch] * rm[ch] / cw[ch] * sc;
                                gb[ch] = a * db[ch];
                                g_r += a * dr[ch] * ms.rough_scale;
                                g_m += a * dm[ch] * ms.metal_scale;
                            }
                            g_base = add(g_base, V3{gb[0], gb[1], gb[2]});
                        }
                    } else if (!use_env) {
                        shadow = true;
                        const float wl = nee_v ? 1.0f / (1.0f - c.env_pc) : 1.0f;
                        sho = f4(add(wpos, scale(lvec, kPullout)), llen);
                        shd = make_float4(ldir.x, ldir.y, ldir.z, __uint_as_float(p));
                        shc = make_float4(contrib.x * wl, contrib.y * wl, contrib.z * wl,
                                          __uint_as_float(k == 0 ? 0u : 1u));
                    }
                    trad_next = mul(trad, energy);
                }
                if (use_env && !last) {
                    // environment sample: the sky term the next bounce would add
                    // (Tracer.cpp:244-252: base' x throughput x sky x secondary intensity, the
                    // throughput updated only where a light was chosen, Tracer.cpp:273, 336),
                    // where the path would go on past this vertex (Tracer.cpp:206-208), by the
                    // balance heuristic with the bounce's own draw of the same term
                    V3 ed;
                    float pe;
                    const uint32_t dn = kDimNee + k * 8u;
                    if (env_sample(c, S, J, draw(key, dn + 1), draw(key, dn + 2), draw(key, dn + 3),
                                   draw(key, dn + 4), ed, pe)) {
                        const float nle = dot(N, ed);
                        const float pref = (1.0f - ps) * diffuse_pdf(ed, N) + ps * specular_pdf(ed, view, N, rough_s);
                        if (nle > 0.0f && pref > 0.0f) {
                            const float q = bounce_density(c, ed, view, N, ps, rough_s, pref);
                            const V3 ee = scale(eval_brdf(ed, view, N, basec, rough, metal), nle / pref);
                            // chance that the path goes on past this vertex in direction ed:
                            // the energy cut-off, or the roulette on the termination throughput
                            const V3 tt = mul(k == 0 ? mk(1.0f) : xyz(b.t_term[k & 1][idx]), ee);
                            float keep = 1.0f;
                            if (J.mode == MT_MODE_FIXED) {
                                if (k + 1 >= J.rr_depth) keep = clampf(gmax(gmax(tt.x, tt.y), tt.z), 0.05f, 1.0f);
                            } else if (len(tt) < J.terminate_energy) {
                                keep = 0.0f;
                            }
                            if (keep > 0.0f && q > 0.0f) {
                                // where a light was chosen the throughput takes the energy over
                                // the roulette's chance (which cancels it), elsewhere it stays
                                // and only the chance of going on weighs the term
                                const V3 thr = (chosen >= 0) ? mul(trad, ee) : scale(trad, keep);
                                const V3 sky = scale(sample_env(S, ed, J.hdri_rot_apply, J.hdri_rot_offset),
                                                     J.hdri_secondary);
                                const V3 val = scale(mul(mul(basec, thr), sky), q / (q + c.env_pc * pe));
                                if (val.x > 0.0f || val.y > 0.0f || val.z > 0.0f) {
                                    shadow = true;
                                    // no light distance: any occluder counts (kFltMax, whose
                                    // next float, the any-hit bound, is +inf)
                                    sho = f4(add(add(o, scale(d, dist)), scale(N, kPullout)), kFltMax);
                                    shd = make_float4(ed.x, ed.y, ed.z, __uint_as_float(p));
                                    shc = make_float4(val.x, val.y, val.z, __uint_as_float(2u));
                                }
                            }
                        }
                    }
                }
                if (kDiff == 2) {
                    if (gr.sampling) {
                        // the bounce's decisions steer everything collected after this vertex:
                        // (L - S) times the tangent factor of the decisions' density, and of
                        // the pdf where it divides the throughput (a light was chosen)
                        const V3 rem = sub(ltot, ssum);
                        const float wsum = (dl.x * rem.x + dl.y * rem.y) + dl.z * rem.z;
                        if (wsum != 0.0f) {
                            const SampleScore sc = sample_score(view, N, basec, rough, metal, ps, spec_lobe, hcos,
                                                                nd, pdf, chosen >= 0);
                            g_r += wsum * sc.rough * ms.rough_scale;
                            g_m += wsum * sc.metal * ms.metal_scale;
                            g_base = add(g_base, mk(wsum * sc.base));
                        }
                    }
                    set_texel_grad(tgr, gr, ms, g_base);
                    b.adj_s[p] = f4(ssum, 0.0f);
                    b.adj_prev[p] = make_float4(tu, tv, __uint_as_float(ms.tex ? mi : 0xFFFFFFFFu), 0.0f);
                }
                const bool after_q = (!stop || c.cache_nee) && !last && kDiff == 0 && !kBatch && kExt &&
                                     c.cache && c.cache_after && k == c.cache_depth;
                if (after_q && stop && b.train_slot[p] < c.train_cap) {
                    // cache_nee: a training path that stops here has its record completed when
                    // the path ends (all it collects after this point is environment samples)
                    const size_t r = static_cast<size_t>(b.train_slot[p]) * c.train_nrec;
                    const float4 tw = b.tr_w[r];
                    b.tr_w[r] = make_float4(tw.x, tw.y, tw.z, -1.0f);
                }
                if (after_q) {
                    // a path that would go on asks the cache for what the bounce ray would
                    // bring (scaled by the throughput at this vertex); a training path asks too
                    // (its image takes the same estimate) and goes on to record its target
                    query = true;
                    qpos = hp;
                    qdir = f4(view, 0.0f);
                    qthr = f4(trad, 0.0f);
                }
                if (after_q && b.train_slot[p] == kNoSlot) {
                    // stops here
                } else if (!stop && !last) {
                    cont = true;
                    next_o = add(add(o, scale(d, dist)), scale(N, kPullout));  // stored below
                    next_d = nd;
                    next_trad = trad_next;
                    next_term = tterm;
                    next_base = basec;
                }
            }
        }
        if (kDiff == 2) {
            if (gr.tex) {  // every lane of the wave is here
                const V3 gc = scale(tgr.g, 0.96f);
                for (int t = 0; t < 4; ++t)
                    wave_run_add3(gr.tex, tgr.tap[t], scale(gc, tgr.tw[t]), tgr.on && tgr.tw[t] != 0.0f);
            }
            if (grad_in_lds(gr)) {
                if (gvalid) {
                    float* sg = grad_lds() + gmat * 5u;
                    if (gr.rough && g_r != 0.0f) atomicAdd(sg + 0, g_r);
                    if (gr.metal && g_m != 0.0f) atomicAdd(sg + 1, g_m);
                    if (gr.emis) {
                        if (g_e.x != 0.0f) atomicAdd(sg + 2, g_e.x);
                        if (g_e.y != 0.0f) atomicAdd(sg + 3, g_e.y);
                        if (g_e.z != 0.0f) atomicAdd(sg + 4, g_e.z);
                    }
                }
            } else {
                if (gr.rough) wave_atomic_by_key(gr.rough, gmat, g_r, gvalid && g_r != 0.0f);
                if (gr.metal) wave_atomic_by_key(gr.metal, gmat, g_m, gvalid && g_m != 0.0f);
                if (gr.emis) {
                    wave_atomic_by_key(gr.emis, gmat * 3 + 0, g_e.x, gvalid && g_e.x != 0.0f);
                    wave_atomic_by_key(gr.emis, gmat * 3 + 1, g_e.y, gvalid && g_e.y != 0.0f);
                    wave_atomic_by_key(gr.emis, gmat * 3 + 2, g_e.z, gvalid && g_e.z != 0.0f);
                }
            }
        } else {
            const uint32_t ss = block_push(b.counters + kCntShadow + k, shadow);
            if (shadow) {
                b.sh_o[ss] = sho;
                b.sh_d[ss] = shd;
                b.sh_c[ss] = shc;
            }
            if (kDiff == 0 && !kBatch && kExt && c.cache) {
                const uint32_t qs = block_push(b.counters + kCntQuery, query);
                if (query) {
                    b.cq_path[qs] = p;
                    b.cq_pos[qs] = qpos;
                    b.cq_dir[qs] = qdir;
                    b.cq_t[qs] = qthr;
                }
            }
        }
        const uint32_t cs = block_push(b.counters + kCntExtend + k + 1, cont);
        if (cont) {
            qout[cs] = p;
            b.ray_o[(k + 1) & 1][cs] = f4(next_o, 0.0f);
            b.ray_d[(k + 1) & 1][cs] = f4(next_d, 0.0f);
            if (!kBatch && c.bin) b.bin_key[cs] = ray_bin(c, next_o, next_d);
            b.t_term[(k + 1) & 1][cs] = f4(next_term, 0.0f);
            b.t_rad[(k + 1) & 1][cs] = f4(next_trad, 0.0f);
            b.prev_base[(k + 1) & 1][cs] = f4(next_base, next_q);
        }
    }
    if constexpr (kDiff == 2) {
        if (grad_in_lds(gr)) {  // the block's sums, once
            __syncthreads();
            const float* sg = grad_lds();
            for (uint32_t i = threadIdx.x; i < gr.materials * 5u; i += blockDim.x) {
                const float v = sg[i];
                if (v == 0.0f) continue;
                const uint32_t m = i / 5u, j = i - m * 5u;
                if (j == 0) atomicAdd(gr.rough + m, v);
                else if (j == 1) atomicAdd(gr.metal + m, v);
                else atomicAdd(gr.emis + m * 3u + (j - 2u), v);
            }
        }
    }
    if (c.count_work) {
        wave_add_u64(b.work + 4, hits);
        wave_add_u64(b.work + 5, misses);
    }
}

template <int kDiff, bool kBatch, bool kExt, bool kLights = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MT_SHADE_WAVES_PER_SIMD, 8)))
void k_shade(DevScene S, ChunkParams c, PathBuffers b, uint32_t k) {
    shade_body<kDiff, kBatch, kExt, kLights>(S, c, b, k, blockIdx.x, gridDim.x);
}

// plain shading of a scene without point lights: small enough to run five waves per SIMD
#ifndef MT_SHADE_UNLIT_WAVES_PER_SIMD
#define MT_SHADE_UNLIT_WAVES_PER_SIMD 5
#endif
template <bool kBatch>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(MT_SHADE_UNLIT_WAVES_PER_SIMD, 8)))
void k_shade_unlit(DevScene S, ChunkParams c, PathBuffers b, uint32_t k) {
    shade_body<0, kBatch, false, false>(S, c, b, k, blockIdx.x, gridDim.x);
}

// RandomColorFromU16 (Tracer.cpp:350-364)
__device__ __forceinline__ V3 color_from_u16(uint32_t value) {
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

// ProcessDebugRay (Tracer.cpp:366-488) from the path's debug records; `total` is the sample's
// regular colour, kept by the modes the function does not handle
__device__ V3 debug_color(const ChunkParams& c, const PathBuffers& b, uint32_t p, V3 e, V3 d, V3 i,
                          V3 s, V3 total) {
    const uint4 di = b.dbg_i[p], dj = b.dbg_j[p];
    const bool hit = dj.w & 1u;
    const V3 black = mk(0.0f);
    switch (c.debug) {
        case MT_DEBUG_PRIMARY_BVH_TESTS: return hit ? mk(static_cast<float>(di.x)) : black;
        case MT_DEBUG_TOTAL_BVH_TESTS: return hit ? mk(static_cast<float>(di.z)) : black;
        case MT_DEBUG_PRIMARY_TRIANGLE_TESTS: return hit ? mk(static_cast<float>(di.y)) : black;
        case MT_DEBUG_TOTAL_TRIANGLE_TESTS: return hit ? mk(static_cast<float>(di.w)) : black;
        case MT_DEBUG_BOUNCES: return hit ? mk(static_cast<float>(dj.w >> 2)) : black;
        case MT_DEBUG_ALBEDO: return hit ? xyz(b.dbg_a[p]) : black;
        case MT_DEBUG_METALLIC_ROUGHNESS: return hit ? V3{0.0f, b.dbg_a[p].w, b.dbg_b[p].x} : black;
        case MT_DEBUG_GEOMETRIC_NORMAL: {
            const float4 q = b.dbg_b[p];
            return hit ? add(scale(V3{q.y, q.z, q.w}, 0.5f), mk(0.5f)) : black;
        }
        case MT_DEBUG_SHADING_NORMAL: return hit ? add(scale(xyz(b.dbg_c[p]), 0.5f), mk(0.5f)) : black;
        case MT_DEBUG_DIRECT_ONLY: return aces(d, c.exposure);
        case MT_DEBUG_INDIRECT_ONLY: return aces(add(i, s), c.exposure);
        case MT_DEBUG_EMISSION: return aces(e, c.exposure);
        case MT_DEBUG_COLOR_PER_MESH: return hit ? color_from_u16(dj.x) : black;
        case MT_DEBUG_COLOR_PER_TRIANGLE: return hit ? color_from_u16(dj.y) : black;
        case MT_DEBUG_COLOR_PER_BOUNDING_VOLUME: return hit ? color_from_u16(dj.z) : black;
        case MT_DEBUG_DEPTH: return hit ? mk(b.dbg_c[p].w) : black;
        case MT_DEBUG_FIRST_HIT_FRESNEL: return hit ? xyz(b.dbg_d[p]) : black;
        case MT_DEBUG_FIRST_HIT_BRDF: return hit ? mk(b.dbg_d[p].w) : black;
        case MT_DEBUG_FIRST_HIT_PDF: return hit ? mk(b.dbg_e[p].w) : black;
        case MT_DEBUG_REFLECTED_DIRECTION:
            return (dj.w & 2u) ? add(scale(xyz(b.dbg_e[p]), 0.5f), mk(0.5f)) : black;
        default: return total;
    }
}

template <bool kBatch>
__device__ __forceinline__ void finalize_path(const ChunkParams& c, const PathBuffers& b, uint32_t p) {
    const ChunkParams& J = job_params<kBatch>(c, b, p);
    const V3 e = xyz(b.acc_e[p]), d = xyz(b.acc_d[p]), i = xyz(b.acc_i[p]), s = xyz(b.acc_s[p]);
    const V3 L = add(add(add(e, d), i), s);  // Tracer.cpp:344-345
    V3 Limg = L;
    if (!kBatch && c.cache && c.cache_after) {
        // a training path that went on past its query vertex: its image takes the cached
        // estimate (radiance up to and including the vertex's light sample + its query)
        const uint32_t slot = b.train_slot[p];
        if (slot < c.train_cap) {
            const size_t r = static_cast<size_t>(slot) * c.train_nrec;
            const float4 tw = b.tr_w[r];
            if (tw.w < 0.0f) {
                // cache_nee: the path stopped after its query vertex, so everything it collected
                // (environment samples aside) came up to and including that vertex's light sample
                b.tr_s[r] = f4(L, 0.0f);
                b.tr_w[r] = make_float4(tw.x, tw.y, tw.z, 1.0f);
                Limg = add(L, xyz(b.tr_img[p]));
            } else if (tw.w > 0.0f) {
                Limg = add(xyz(b.tr_s[r]), xyz(b.tr_img[p]));
            }
        }
    }
    if (!kBatch && c.cache_full_image && b.train_slot[p] < c.train_cap)
        Limg = c.cache_nee ? add(L, xyz(b.acc_n[p])) : L;
    V3 col = J.output_linear ? Limg : aces(Limg, J.exposure);
    if (!kBatch && c.debug) col = debug_color(c, b, p, e, d, i, s, col);
    b.sample_rgb[p] = f4(col, 0.0f);
    if (!kBatch && c.diff) {
        uint32_t pix, s;
        path_coords(c, p, pix, s);
        c.path_store[static_cast<size_t>(pix) * c.spp + s] = make_float4(L.x, L.y, L.z, __uint_as_float(b.vis[p]));
    }
}

template <bool kBatch>
__global__ __launch_bounds__(kBlock) void k_finalize(ChunkParams c, PathBuffers b) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (kBatch ? b.counters[kCntPaths] : c.paths)) return;
    finalize_path<kBatch>(c, b, p);
}

// per-pixel mean of the samples sp[0..n) in sample order, with CleanupFireflies
// (Tracer.cpp:490-548)
__device__ __forceinline__ void resolve_pixel(const ChunkParams& J, const float4* sp, float* out) {
    const uint32_t n = J.spp;
    uint32_t cnt = n;
    float avg = 0.0f, thr = 0.0f;
    if (J.firefly) {
        float sum = 0.0f;
        for (uint32_t s = 0; s < n; ++s) {
            const float4 v = sp[s];
            sum = sum + (v.x + v.y + v.z) / 3.0f;
        }
        avg = sum / static_cast<float>(n);
        float var = 0.0f;
        for (uint32_t s = 0; s < n; ++s) {
            const float4 v = sp[s];
            const float dd = (v.x + v.y + v.z) / 3.0f - avg;
            var = var + dd * dd;
        }
        var = var / static_cast<float>(n);
        const float sd = __builtin_sqrtf(var);
        if (J.debug == MT_DEBUG_PIXEL_STANDARD_DEVIATION) {  // Tracer.cpp:515-523
            out[0] = out[1] = out[2] = sd;
            return;
        }
        thr = J.firefly_threshold * sd;
    }
    V3 acc = mk(0.0f);
    for (uint32_t s = 0; s < n; ++s) {
        const float4 v = sp[s];
        V3 col = xyz(v);
        if (J.firefly && fabsf((v.x + v.y + v.z) / 3.0f - avg) > thr) {
            col = mk(0.0f);
            cnt--;
        }
        acc = add(acc, col);
    }
    if (J.firefly && J.debug == MT_DEBUG_FIREFLY_ELIMINATION) {  // Tracer.cpp:536-544
        out[0] = out[1] = out[2] = static_cast<float>(n - cnt);
        return;
    }
    if (cnt == 0) cnt = 1;
    const V3 r = divs(acc, static_cast<float>(cnt));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

__global__ __launch_bounds__(kBlock) void k_resolve(ChunkParams c, PathBuffers b, float* img) {
    const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
    if (lp >= c.npix) return;
    resolve_pixel(c, b.sample_rgb + static_cast<size_t>(lp) * c.spp, img + (c.pix_begin + lp) * 3ull);
}

// batched: pixel q of the iteration, written into its job's image
__device__ __forceinline__ void batch_resolve_pixel(const PathBuffers& b, uint32_t q) {
    const uint32_t ti = tile_of_pixel(b, b.counters[kCntTiles], q);
    const uint4 t = b.it_tiles[ti];
    const BatchJob& job = b.jobs[t.x];
    const uint32_t lp = q - b.it_pix_off[ti];
    resolve_pixel(job.p, b.sample_rgb + t.w + static_cast<size_t>(lp) * job.p.spp,
                  job.img + (t.y + lp) * 3ull);
}

__global__ __launch_bounds__(kBlock) void k_batch_resolve(PathBuffers b) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.counters[kCntPixels]) return;
    batch_resolve_pixel(b, q);
}

// ------------------------------------------------------------------ one persistent launch
// The grid is split into lane groups of group_blocks consecutive blocks, each with its own
// path state; every group loops: its first block takes tiles from the shared queue (scene of
// the first tile, in queue order while their paths fit), then the group runs generate, the
// bounces (extend, shade, shadow), finalize and resolve of that iteration with a group
// barrier between phases, and marks the jobs whose last tile it resolved.  The launch is
// cooperative, so all blocks are resident and the barriers can be waited on.  Every wait
// ends: a group exits when the host has closed the queue and it is drained, or after
// idle_ticks without tiles, and a barrier that waits longer than stall_ticks sets the abort
// word, on which every waiting block returns.
constexpr uint32_t kSelNone = 0xFFFFFFFFu;  // no tiles yet
constexpr uint32_t kSelExit = 0xFFFFFFFEu;  // the group leaves the kernel

// polls are relaxed loads (an acquire load invalidates the cache hierarchy on every poll,
// evicting the BVH the other lane groups are traversing); one acquire fence follows the wait
__device__ __forceinline__ uint32_t load_sys(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the XCD (accelerator complex die) this wave runs on
__device__ __forceinline__ uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & 0xfu;
}

// barrier of one lane group's blocks; false when the launch is aborting (the caller returns).
// Release: every wave of the block makes its own stores visible at agent scope (L2 write-back)
// before the block barrier (__syncthreads waits only for LDS and scalar operations), then one
// thread arrives.  Acquire: that thread invalidates L1 and L2 and waits for the invalidation
// before the block's other waves go on.  The wait polls with relaxed loads: an acquire load
// would invalidate the caches on every poll, evicting what the other groups are working on.
__device__ bool group_sync(const PersistCtl& ctl, uint32_t* bar, uint32_t* flag) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        const uint32_t gen = load_agent(bar + 1);
        if (atomicAdd(bar, 1u) == ctl.group_blocks - 1u) {
            __hip_atomic_store(bar, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(bar + 1, gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const uint64_t t0 = wall_clock64();
            while (load_agent(bar + 1) == gen) {
                if (load_agent(ctl.state + kPsAbort)) {
                    ok = 0;
                    break;
                }
                if (wall_clock64() - t0 > ctl.stall_ticks) {
                    atomicExch(ctl.state + kPsAbort, 1u);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(16);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        *flag = ok;
    }
    __syncthreads();
    return *flag != 0;
}

// 16-byte words [first, end) of the published host records into the device copies
__device__ __forceinline__ void copy_records(uint4* dst, const uint4* src, uint64_t first, uint64_t end) {
    for (uint64_t i = first + threadIdx.x; i < end; i += kBlock) dst[i] = src[i];
}

// the group's first block: copy what the host published since the last look, then take the
// next iteration's tiles; returns the scene slot, or kSelExit
__device__ uint32_t persist_take(const PersistCtl& ctl, const PathBuffers& b, uint32_t* sh) {
    const uint64_t t0 = wall_clock64();
    constexpr uint32_t kJobWords = sizeof(BatchJob) / 16, kSlotWords = sizeof(PersistSlot) / 16;
    static_assert(sizeof(BatchJob) % 16 == 0 && sizeof(PersistSlot) % 16 == 0, "record layout");
    while (true) {
        if (threadIdx.x == 0) {
            uint32_t ok = 1;
            while (atomicCAS(ctl.state + kPsLock, 0u, 1u) != 0u) {
                if (load_agent(ctl.state + kPsAbort)) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(4);
            }
            __threadfence();
            sh[0] = ok;
            if (ok) {
                // the tail first: the host publishes it after the job and slot counts, so
                // every tile read below has its job's record and scene slot
                sh[5] = ctl.state[kPsTiles];
                sh[6] = load_sys(&ctl.host->tail);
                sh[3] = ctl.state[kPsJobs];
                sh[4] = load_sys(&ctl.host->jobs);
                sh[1] = ctl.state[kPsSlots];
                sh[2] = load_sys(&ctl.host->slots);
                __atomic_thread_fence(__ATOMIC_ACQUIRE);  // system scope: the records behind the counts
            }
        }
        __syncthreads();
        if (!sh[0]) return kSelExit;
        copy_records(reinterpret_cast<uint4*>(ctl.slots), reinterpret_cast<const uint4*>(ctl.host_slots),
                     uint64_t(sh[1]) * kSlotWords, uint64_t(sh[2]) * kSlotWords);
        copy_records(reinterpret_cast<uint4*>(ctl.jobs), reinterpret_cast<const uint4*>(ctl.host_jobs),
                     uint64_t(sh[3]) * kJobWords, uint64_t(sh[4]) * kJobWords);
        for (uint32_t j = sh[3] + threadIdx.x; j < sh[4]; j += kBlock) {
            ctl.job_slot[j] = ctl.host_job_slot[j];
            ctl.tiles_left[j] = ctl.host_job_tiles[j];
        }
        copy_records(ctl.tiles, ctl.host_tiles, sh[5], sh[6]);
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            ctl.state[kPsSlots] = sh[2];
            ctl.state[kPsJobs] = sh[4];
            ctl.state[kPsTiles] = sh[6];
            uint32_t head = ctl.state[kPsHead];
            uint32_t n = 0, paths = 0, npix = 0, slot = kSelNone;
            while (head < sh[6] && n < b.it_tile_cap) {
                const uint4 t = ctl.tiles[head];
                const uint32_t s = ctl.job_slot[t.x];
                if (n > 0 && (s != slot || paths + t.w > b.it_capacity)) break;
                slot = s;
                b.it_tiles[n] = make_uint4(t.x, t.y, t.z, paths);
                b.it_pix_off[n] = npix;
                paths += t.w;
                npix += t.z;
                ++n;
                ++head;
            }
            b.it_pix_off[n] = npix;
            ctl.state[kPsHead] = head;
            __threadfence();
            __hip_atomic_store(ctl.state + kPsLock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t sel = kSelNone;
            if (n > 0) {
                for (uint32_t k = 0; k < kMaxCounters; ++k) b.counters[k] = 0;
                b.counters[kCntExtend] = paths;
                b.counters[kCntPaths] = paths;
                b.counters[kCntPixels] = npix;
                b.counters[kCntTiles] = n;
                sel = slot;
            } else if (load_sys(&ctl.host->closed) && load_sys(&ctl.host->tail) == sh[6]) {
                sel = kSelExit;  // closed after its last tiles, and every tile taken
            } else if (wall_clock64() - t0 > ctl.idle_ticks) {
                sel = kSelExit;
            }
            sh[7] = sel;
        }
        __syncthreads();
        if (sh[7] != kSelNone) return sh[7];
        if (threadIdx.x == 0)
            for (int i = 0; i < 16; ++i) __builtin_amdgcn_s_sleep(127);
        __syncthreads();
    }
}

template <uint32_t kDepth>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
void k_batch_persistent(PersistCtl ctl) {
    __shared__ uint2 lds_stack[kDepth * kBlock];
    __shared__ uint32_t sh[8];
    uint32_t g, blk;
    if (ctl.xcds) {
        // groups within XCDs: the hardware deals blocks round-robin over the XCDs, so block b
        // is the (b / xcds)-th block of XCD b % xcds; each XCD's blocks are dealt over its
        // groups.  A group's barriers then write back and invalidate only the L2 its own
        // blocks use, not the other groups'.  (The placement affects speed only: the barriers
        // are agent-scope either way.)
        const uint32_t x = blockIdx.x % ctl.xcds, w = blockIdx.x / ctl.xcds;
        const uint32_t gpx = ctl.groups / ctl.xcds;
        g = x * gpx + w % gpx;
        blk = w / gpx;
    } else {
        g = blockIdx.x / ctl.group_blocks;
        blk = blockIdx.x - g * ctl.group_blocks;
    }
    if (g >= ctl.groups || blk >= ctl.group_blocks) return;  // the grid is groups * group_blocks blocks
    const PathBuffers& b = ctl.lanes[g];
    uint32_t* bar = ctl.bar + g * kPsBarWords;
    const uint32_t nblk = ctl.group_blocks;
    const uint32_t nthreads = nblk * kBlock;
    const uint32_t gtid = blk * kBlock + threadIdx.x;
    const TestIO io{};
    while (true) {
        if (blk == 0) {
            const uint32_t sel = persist_take(ctl, b, sh);
            if (threadIdx.x == 0) bar[2] = sel;
        }
        if (!group_sync(ctl, bar, sh)) return;
        const uint32_t slot = __builtin_amdgcn_readfirstlane(bar[2]);
        if (slot == kSelExit) return;
        const PersistSlot& P = ctl.slots[slot];
        const bool lights = __builtin_amdgcn_readfirstlane(P.lights) != 0;
        const bool fast = __builtin_amdgcn_readfirstlane(P.accel) != 0;
        const uint32_t paths = __builtin_amdgcn_readfirstlane(b.counters[kCntPaths]);
        for (uint32_t p = gtid; p < paths; p += nthreads) batch_generate_path(b, p);
        if (!group_sync(ctl, bar, sh)) return;
        for (uint32_t k = 0; k < kMaxBounces; ++k) {
            if (__builtin_amdgcn_readfirstlane(counter(b, kCntExtend + k)) == 0) break;
            if (fast) trace_body<false, 0, kDepth, 1>(P.S, P.c, b, k, io, lds_stack, gtid, nthreads);
            else trace_body<false, 0, kDepth, 0>(P.S, P.c, b, k, io, lds_stack, gtid, nthreads);
            if (!group_sync(ctl, bar, sh)) return;
            if (lights) shade_body<0, true, false, true>(P.S, P.c, b, k, blk, nblk);
            else shade_body<0, true, false, false>(P.S, P.c, b, k, blk, nblk);
            if (!group_sync(ctl, bar, sh)) return;
            if (lights) {
                if (fast) trace_body<false, 1, kDepth, 1>(P.S, P.c, b, k, io, lds_stack, gtid, nthreads);
                else trace_body<false, 1, kDepth, 0>(P.S, P.c, b, k, io, lds_stack, gtid, nthreads);
                if (!group_sync(ctl, bar, sh)) return;
            }
        }
        for (uint32_t p = gtid; p < paths; p += nthreads) finalize_path<true>(P.c, b, p);
        if (!group_sync(ctl, bar, sh)) return;
        const uint32_t npix = __builtin_amdgcn_readfirstlane(b.counters[kCntPixels]);
        for (uint32_t q = gtid; q < npix; q += nthreads) batch_resolve_pixel(b, q);
        if (!group_sync(ctl, bar, sh)) return;
        if (blk == 0 && threadIdx.x == 0) {
            // the jobs whose last tile this iteration resolved are final
            const uint32_t n = b.counters[kCntTiles];
            for (uint32_t i = 0; i < n; ++i) {
                const uint32_t job = b.it_tiles[i].x;
                if (atomicSub(ctl.tiles_left + job, 1u) == 1u)
                    __hip_atomic_store(ctl.job_done + job, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
            // only this thread writes the group's count: it survives relaunches
            const unsigned long long it = __hip_atomic_load(ctl.host_iters + g, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(ctl.host_iters + g, it + 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// ------------------------------------------------------------------ radiance cache passes
// adds throughput x prediction to every path that stopped at a cache query
__global__ __launch_bounds__(kBlock) void k_cache_apply(ChunkParams c, PathBuffers b) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= b.counters[kCntQuery]) return;
    const uint32_t p = b.cq_path[q];
    const float4 o = b.cq_out[q];
    const float sc = c.cache_scale[0];
    const V3 pred = {gmax(o.x, 0.0f) * sc, gmax(o.y, 0.0f) * sc, gmax(o.z, 0.0f) * sc};
    const V3 add_ = mul(xyz(b.cq_t[q]), pred);
    if (c.cache_after && b.train_slot[p] < c.train_cap) {
        // a training path keeps its full radiance for its target; its image value is formed
        // in k_finalize from the radiance up to the query vertex and this estimate
        b.tr_img[p] = f4(add_, 0.0f);
        return;
    }
    const float4 a = b.acc_i[p];
    b.acc_i[p] = make_float4(a.x + add_.x, a.y + add_.y, a.z + add_.z, 0.0f);
}

// per chunk pixel: adds its samples of this pass, in sample order, to the pixel's sum
__global__ __launch_bounds__(kBlock) void k_cache_accum(ChunkParams c, PathBuffers b) {
    const uint32_t lp = blockIdx.x * blockDim.x + threadIdx.x;
    if (lp >= c.npix) return;
    float* dst = b.img_sum + (c.pix_begin + lp) * 3ull;
    V3 acc = {dst[0], dst[1], dst[2]};
    const float4* sp = b.sample_rgb + static_cast<size_t>(lp) * c.spp;
    for (uint32_t s = 0; s < c.spp; ++s) acc = add(acc, xyz(sp[s]));
    dst[0] = acc.x;
    dst[1] = acc.y;
    dst[2] = acc.z;
}

// training samples from the recorded vertices of the training paths: target = radiance the
// path collected from the vertex on, divided by the throughput at the vertex
__global__ __launch_bounds__(kBlock) void k_cache_targets(ChunkParams c, PathBuffers b, uint32_t slot0,
                                                          uint32_t slots) {
    const uint32_t slot = slot0 + blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok_slot = slot < slot0 + slots && slot < c.train_cap;
    V3 L = mk(0.0f);
    uint32_t p = 0;
    if (ok_slot) {
        p = b.tr_path[slot];
        L = add(add(add(xyz(b.acc_e[p]), xyz(b.acc_d[p])), xyz(b.acc_i[p])), xyz(b.acc_s[p]));
        if (c.cache_nee) L = add(L, xyz(b.acc_n[p]));
    }
    for (uint32_t j = 0; j < c.train_nrec; ++j) {
        bool ok = false;
        float4 pos{}, dir{}, tgt{};
        if (ok_slot) {
            const size_t r = static_cast<size_t>(slot) * c.train_nrec + j;
            const float4 w = b.tr_w[r];
            if (w.w > 0.0f && w.x > 0.0f && w.y > 0.0f && w.z > 0.0f) {
                const V3 s = xyz(b.tr_s[r]);
                const V3 t = {(L.x - s.x) / w.x, (L.y - s.y) / w.y, (L.z - s.z) / w.z};
                if (finite3(t)) {
                    ok = true;
                    pos = b.tr_pos[r];
                    dir = b.tr_dir[r];
                    tgt = make_float4(gmax(t.x, 0.0f), gmax(t.y, 0.0f), gmax(t.z, 0.0f), 0.0f);
                }
            }
        }
        const uint32_t ts = wave_push(b.counters + kCntTrainRec, ok);
        if (ok) {
            b.tp_pos[ts] = pos;
            b.tp_dir[ts] = dir;
            b.tp_tgt[ts] = tgt;
        }
    }
}

// The scale of the cache's targets, set on the first training chunk that the cache sees: with
// the L2 loss the mean target (over channels and samples; 1 if that is not positive and
// finite), with the relative loss 1.  One block.
__global__ __launch_bounds__(1024) void k_cache_scale_init(ChunkParams c, PathBuffers b) {
    __shared__ double part[1024];
    if (c.cache_scale[1] != 0.0f) return;  // set already (uniform over the block)
    const uint32_t n = b.counters[kCntTrainRec];
    double sum = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const float4 t = b.tp_tgt[i];
        sum += (static_cast<double>(t.x) + t.y) + t.z;
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = n ? part[0] / (3.0 * n) : 0.0;
        // the relative loss does not depend on the units (but for its 0.01): no scale there
        c.cache_scale[0] = (c.cache_l2 && m > 0.0 && m < 3.0e38) ? static_cast<float>(m) : 1.0f;
        c.cache_scale[1] = 1.0f;
    }
}

// The chunk's training samples in the order the Adam steps take them, targets in the cache's
// units.  The records come in path order (pixel rows), and a pass's steps may use only part of
// them: sample i is record (i * (2^31 - 1)) mod n, a permutation for every n < 2^31 (the stride
// is prime) that spreads every step's consecutive samples over the whole chunk.
__global__ __launch_bounds__(kBlock) void k_cache_shuffle(ChunkParams c, PathBuffers b) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = b.counters[kCntTrainRec];
    if (i >= n) return;
    const uint32_t j = static_cast<uint32_t>((static_cast<uint64_t>(i) * 2147483647ull) % n);
    const float inv = 1.0f / c.cache_scale[0];
    const float4 t = b.tp_tgt[j];
    b.ts_pos[i] = b.tp_pos[j];
    b.ts_dir[i] = b.tp_dir[j];
    b.ts_tgt[i] = make_float4(t.x * inv, t.y * inv, t.z * inv, t.w);
}

__global__ __launch_bounds__(kBlock) void k_cache_resolve(const float* sum, uint32_t n, float spp,
                                                          float* img) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) img[i] = sum[i] / spp;
}

// adds the chunk's rays (extension + shadow, every bounce) to a running total
__global__ void k_count_rays(PathBuffers b, uint32_t bounces, unsigned long long* total) {
    if (threadIdx.x != 0) return;
    unsigned long long r = 0;
    for (uint32_t k = 0; k < bounces; ++k) r += b.counters[kCntExtend + k] + b.counters[kCntShadow + k];
    atomicAdd(total, r);
}

__global__ void k_fill(float* p, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_brdf(uint32_t n, const float* in, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* q = in + i * 14;
    const V3 L = {q[0], q[1], q[2]}, V = {q[3], q[4], q[5]}, N = {q[6], q[7], q[8]};
    const V3 base = {q[9], q[10], q[11]};
    const V3 f = eval_brdf(L, V, N, base, q[12], q[13]);
    float* o = out + i * 5;
    o[0] = f.x;
    o[1] = f.y;
    o[2] = f.z;
    o[3] = diffuse_pdf(L, N);
    o[4] = specular_pdf(L, V, N, q[12]);
}

__global__ void k_math(int32_t op, uint32_t n, const float* in, float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = in[2 * i], y = in[2 * i + 1];
    float r = 0.0f, s, c;
    switch (op) {
        case MT_MATH_SIN: em::sincos(x, s, c); r = s; break;
        case MT_MATH_COS: em::sincos(x, s, c); r = c; break;
        case MT_MATH_ATAN2: r = em::atan2(x, y); break;
        case MT_MATH_ACOS: r = em::acos(x); break;
        case MT_MATH_POW5: r = em::pow5(x); break;
        default: break;
    }
    out[i] = r;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(MT_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

uint32_t env_u32(const char* name, uint32_t def) {
    const char* v = std::getenv(name);
    if (!v || !*v) return def;
    const long x = std::strtol(v, nullptr, 10);
    return x > 0 ? static_cast<uint32_t>(x) : def;
}

int device_cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            cus = prop.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return cus;
}

// Device blocks of finished renderers are kept for the next renderer that asks for a block of
// the same size on the same device: consecutive renders of one configuration ask for identical
// path buffers (up to 48 GB at 1080p), and reusing them spares each render mapping and
// unmapping that memory (measured: a 1080p 256-spp frame in 64 Mi-path chunks 1.41 s instead
// of 0.45 s of kernel time without the pool on the fast traversal, 517 against 495
// Msamples/s with it on the reference one).  At most MT_POOL_GB gibibytes (default 64, 0 =
// off) stay pooled, oldest released first; a failing hipMalloc empties the pool and retries;
// mt_pool_trim empties it on request.  Only Renderer::release returns blocks to the pool,
// after the device has finished every launch that used them.
struct PoolBlock {
    int device;
    size_t bytes;
    void* ptr;
};
std::mutex g_pool_mu;
std::deque<PoolBlock> g_pool;                           // oldest first
size_t g_pool_bytes = 0;
std::unordered_map<void*, std::pair<int, size_t>> g_block;  // blocks made by alloc(): device, bytes

size_t pool_limit() {
    static const size_t lim = [] {
        const char* v = std::getenv("MT_POOL_GB");
        const long gb = (v && *v) ? std::strtol(v, nullptr, 10) : 64;
        return gb > 0 ? static_cast<size_t>(gb) << 30 : size_t(0);
    }();
    return lim;
}

void pool_flush_locked() {
    for (const PoolBlock& b : g_pool) {
        g_block.erase(b.ptr);
        (void)hipFree(b.ptr);
    }
    g_pool.clear();
    g_pool_bytes = 0;
}

// return alloc()'s blocks: to the pool (recycle; their users have finished) or to hipFree
void release_blocks(std::vector<void*>& owned, bool recycle) {
    std::scoped_lock lk(g_pool_mu);
    for (void* p : owned) {
        auto it = g_block.find(p);
        if (recycle && it != g_block.end() && pool_limit() > 0) {
            g_pool.push_back({it->second.first, it->second.second, p});
            g_pool_bytes += it->second.second;
        } else {
            if (it != g_block.end()) g_block.erase(it);
            (void)hipFree(p);
        }
    }
    owned.clear();
    while (g_pool_bytes > pool_limit() && !g_pool.empty()) {
        const PoolBlock b = g_pool.front();
        g_pool.pop_front();
        g_pool_bytes -= b.bytes;
        g_block.erase(b.ptr);
        (void)hipFree(b.ptr);
    }
}

}  // namespace

void pool_trim() {
    std::scoped_lock lk(g_pool_mu);
    pool_flush_locked();
}

int device_malloc(void** p, size_t bytes) {
    *p = nullptr;
    if (hipMalloc(p, std::max<size_t>(bytes, 1)) == hipSuccess) return MT_OK;
    (void)hipGetLastError();
    {
        std::scoped_lock lk(g_pool_mu);
        pool_flush_locked();
    }
    if (hipMalloc(p, std::max<size_t>(bytes, 1)) == hipSuccess) return MT_OK;
    (void)hipGetLastError();
    *p = nullptr;
    return fail(MT_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
}

void free_blocks(std::vector<void*>& blocks) { release_blocks(blocks, false); }

namespace {

template <typename T>
int alloc(std::vector<void*>& owned, T** p, size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    int dev = 0;
    (void)hipGetDevice(&dev);
    void* q = nullptr;
    {
        std::scoped_lock lk(g_pool_mu);
        for (auto it = g_pool.begin(); it != g_pool.end(); ++it)
            if (it->device == dev && it->bytes == bytes) {
                q = it->ptr;
                g_pool_bytes -= bytes;
                g_pool.erase(it);
                break;
            }
        if (!q) {
            if (hipMalloc(&q, bytes) != hipSuccess) {
                (void)hipGetLastError();
                pool_flush_locked();
                if (hipMalloc(&q, bytes) != hipSuccess) return fail(MT_E_NOMEM, "hipMalloc of path buffers failed");
            }
            g_block[q] = {dev, bytes};
        }
    }
    owned.push_back(q);
    *p = static_cast<T*>(q);
    return MT_OK;
}

}  // namespace

int cu_count() { return device_cu_count(); }

uint32_t stack_capacity(const mt_scene* s) {
    const uint32_t blas = std::max<uint32_t>(std::max(s->max_bvh_depth, s->max_fast_depth) + 1u, 2u);
    return s->dev.use_tlas ? blas + s->tlas_depth + 1u : blas;
}

// Measured on the material-spheres grid at 6 blocks per CU: a 12-entry reference stack (24 KB
// per block, all 6 blocks resident) 447 -> 472 Msamples/s over 16 entries (5 resident); the
// fast traversal's deeper SAH trees lose 2 % at 12.
uint32_t lds_stack_depth(bool fast, uint32_t blocks_per_cu) {
    if (fast) return 16;
    constexpr uint32_t kLdsBytes = 160u * 1024u, kEntryBytes = 8u * kBlock;
    for (uint32_t d : {16u, 12u, 8u})
        if (blocks_per_cu * d * kEntryBytes <= kLdsBytes) return d;
    return 4;
}

int trace_rays(mt_scene* s, uint32_t n, const float* rays, mt_hit* hits, int32_t traversal,
               uint32_t* occluded) {
    if (traversal != MT_TRAVERSAL_REFERENCE && traversal != MT_TRAVERSAL_FAST)
        return fail(MT_E_INVALID, "unknown traversal");
    if (traversal == MT_TRAVERSAL_FAST && !s->dev.fast_meshes && s->mesh_count)
        return fail(MT_E_INVALID, "the scene was created without fast BVHs (mt_bvh_options.fast)");
    if (n == 0) return MT_OK;
    std::vector<void*> owned;
    struct Guard {
        std::vector<void*>& v;
        ~Guard() { release_blocks(v, false); }
    } guard{owned};
    const uint32_t stride = occluded ? 7u : 6u;
    float* drays;
    mt_hit* dhits = nullptr;
    uint32_t* docc = nullptr;
    PathBuffers b{};
    const uint32_t cap = stack_capacity(s);
    const uint32_t spill_depth = cap > 16u ? cap - 16u : 1;  // the test kernels keep 16 in LDS
    const uint32_t blocks = std::max<uint32_t>(
        1u, std::min<uint32_t>((n + kBlock - 1) / kBlock, static_cast<uint32_t>(device_cu_count() * 8)));
    int rc;
    if ((rc = alloc(owned, &drays, static_cast<size_t>(n) * stride))) return rc;
    if (occluded) {
        if ((rc = alloc(owned, &docc, n))) return rc;
    } else {
        if ((rc = alloc(owned, &dhits, n))) return rc;
    }
    if ((rc = alloc(owned, &b.spill, static_cast<size_t>(spill_depth) * blocks * kBlock))) return rc;
    if ((rc = alloc(owned, &b.counters, kMaxCounters + 1))) return rc;
    HIP_TRY(hipMemcpy(drays, rays, sizeof(float) * stride * n, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(b.counters, 0, sizeof(uint32_t) * (kMaxCounters + 1)));
    ChunkParams c{};
    c.stack_cap = cap;
    c.refill_min = kRefillMin;
    c.setup_min = kSetupMin;
    c.grab = traversal == MT_TRAVERSAL_FAST ? kGrabFast : kGrab;
    c.accel = traversal == MT_TRAVERSAL_FAST ? 1 : 0;
    TestIO io{drays, dhits, n, b.counters + kMaxCounters, docc};
    if (occluded)
        launch_trace<3>(16, false, blocks, nullptr, s->dev, c, b, 0, io);
    else
        launch_trace<2>(16, false, blocks, nullptr, s->dev, c, b, 0, io);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    uint32_t herr = 0;
    HIP_TRY(hipMemcpy(&herr, b.counters + kCntError, sizeof(herr), hipMemcpyDeviceToHost));
    if (herr) return fail(MT_E_INVALID, "traversal stack overflow");
    if (occluded)
        HIP_TRY(hipMemcpy(occluded, docc, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    else
        HIP_TRY(hipMemcpy(hits, dhits, sizeof(mt_hit) * n, hipMemcpyDeviceToHost));
    return MT_OK;
}

// ------------------------------------------------------------------ closest hits without allocation
// mt_trace_closest_noalloc keeps its scratch -- the traversal kernel's work counters and error
// flag, then the per-thread stack spill area -- in this file-scope device array instead of
// allocating it.  Every GPU that runs the library holds its own copy (kNoAllocScratchBytes)
// until the process ends; its address on each device comes from hipGetSymbolAddress, taken
// once per device with that device current.  All calls on one GPU share that GPU's array:
// calls that overlap on one GPU (other threads or streams) race.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
constexpr size_t kNoAllocCounterBytes = 4096;  // >= (kMaxCounters + 1) words, keeps the spill 4 KiB-aligned
static_assert((kMaxCounters + 1) * sizeof(uint32_t) <= kNoAllocCounterBytes, "counter block");
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kNoAllocMaxDevices] = {};

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// the device address of `scratch` on the current device
int scratch_address(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return fail(MT_E_UNSUPPORTED, "device index above 63");
    std::scoped_lock lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MT_E_HIP, std::string("hipGetSymbolAddress of the scratch array failed: ") +
                                      hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[dev];
    return MT_OK;
}
}  // namespace

void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
}

int trace_rays_noalloc(mt_scene* s, uint32_t n, const float* rays_dev, mt_hit* hits_dev,
                       int32_t traversal, hipStream_t stream) {
    if (traversal != MT_TRAVERSAL_REFERENCE && traversal != MT_TRAVERSAL_FAST)
        return fail(MT_E_INVALID, "unknown traversal");
    if (traversal == MT_TRAVERSAL_FAST && !s->dev.fast_meshes && s->mesh_count)
        return fail(MT_E_INVALID, "the scene was created without fast BVHs (mt_bvh_options.fast)");
    if (n == 0) return MT_OK;
    HIP_TRY(hipSetDevice(s->device));
    unsigned char* base = nullptr;
    int rc = scratch_address(s->device, &base);
    if (rc) return rc;
    // the test kernel keeps 16 stack entries in LDS and spills the rest, (cap - 16) entries of
    // 8 bytes per thread of the grid: the grid is as large as the spill area allows
    const uint32_t cap = stack_capacity(s);
    const uint64_t spill_depth = cap > 16u ? cap - 16u : 1u;
    const uint64_t per_block = spill_depth * kBlock * sizeof(uint2);
    const uint64_t fit = (kNoAllocScratchBytes - kNoAllocCounterBytes) / per_block;
    if (fit == 0)
        return fail(MT_E_UNSUPPORTED, "the scene's BVH needs a traversal stack deeper than the " +
                                          std::to_string(kNoAllocScratchBytes >> 20) +
                                          " MiB scratch of mt_trace_closest_noalloc holds");
    const uint32_t blocks = static_cast<uint32_t>(std::max<uint64_t>(
        1u, std::min<uint64_t>({(static_cast<uint64_t>(n) + kBlock - 1) / kBlock,
                                static_cast<uint64_t>(device_cu_count()) * 8u, fit})));
    PathBuffers b{};
    b.counters = reinterpret_cast<uint32_t*>(base);
    b.spill = reinterpret_cast<uint2*>(base + kNoAllocCounterBytes);
    // the counters start at zero (the last call's values are still there): a fill kernel,
    // no memset
    constexpr uint32_t kWords = kMaxCounters + 1;
    hipLaunchKernelGGL(k_zero_u32, dim3((kWords + 255) / 256), dim3(256), 0, stream, b.counters, kWords);
    ChunkParams c{};
    c.stack_cap = cap;
    c.refill_min = kRefillMin;
    c.setup_min = kSetupMin;
    c.grab = traversal == MT_TRAVERSAL_FAST ? kGrabFast : kGrab;
    c.accel = traversal == MT_TRAVERSAL_FAST ? 1 : 0;
    TestIO io{rays_dev, hits_dev, n, b.counters + kMaxCounters, nullptr};
    launch_trace<2>(16, false, blocks, stream, s->dev, c, b, 0, io);
    HIP_TRY(hipGetLastError());
    uint32_t herr = 0;
    HIP_TRY(hipMemcpyAsync(&herr, b.counters + kCntError, sizeof(herr), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (herr) return fail(MT_E_INVALID, "traversal stack overflow");
    return MT_OK;
}

int brdf_eval(uint32_t n, const float* in, float* out) {
    if (n == 0) return MT_OK;
    std::vector<void*> owned;
    struct Guard {
        std::vector<void*>& v;
        ~Guard() { release_blocks(v, false); }
    } guard{owned};
    float *din, *dout;
    int rc;
    if ((rc = alloc(owned, &din, static_cast<size_t>(n) * 14))) return rc;
    if ((rc = alloc(owned, &dout, static_cast<size_t>(n) * 5))) return rc;
    HIP_TRY(hipMemcpy(din, in, sizeof(float) * 14 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_brdf, dim3((n + 255) / 256), dim3(256), 0, 0, n, din, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout, sizeof(float) * 5 * n, hipMemcpyDeviceToHost));
    return MT_OK;
}

int math_eval(int32_t op, uint32_t n, const float* in, float* out) {
    if (op < MT_MATH_SIN || op > MT_MATH_POW5) return fail(MT_E_INVALID, "unknown elementary function");
    if (n == 0) return MT_OK;
    std::vector<void*> owned;
    struct Guard {
        std::vector<void*>& v;
        ~Guard() { release_blocks(v, false); }
    } guard{owned};
    float *din, *dout;
    int rc;
    if ((rc = alloc(owned, &din, static_cast<size_t>(n) * 2))) return rc;
    if ((rc = alloc(owned, &dout, static_cast<size_t>(n)))) return rc;
    HIP_TRY(hipMemcpy(din, in, sizeof(float) * 2 * n, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_math, dim3((n + 255) / 256), dim3(256), 0, 0, op, n, din, dout);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(out, dout, sizeof(float) * n, hipMemcpyDeviceToHost));
    return MT_OK;
}

// ------------------------------------------------------------------ render driver

Renderer::~Renderer() { release(); }

void Renderer::release() {
    if (tail_stream_) (void)hipStreamSynchronize(tail_stream_);
    for (size_t i = 0; i < slots_.size(); ++i) {
        StreamSlot& S = slots_[i];
        if (S.ev_bounce) (void)hipEventDestroy(S.ev_bounce);
        if (S.ev_tail) (void)hipEventDestroy(S.ev_tail);
        if (i == 0) continue;  // slot 0's stream, events and counters are the renderer's own
        if (S.stream) (void)hipStreamSynchronize(S.stream);
        for (hipEvent_t e : S.events) (void)hipEventDestroy(e);
        if (S.host_counters) (void)hipHostFree(S.host_counters);
        if (S.stream) (void)hipStreamDestroy(S.stream);
    }
    slots_.clear();
    if (tail_stream_) (void)hipStreamDestroy(tail_stream_);
    tail_stream_ = nullptr;
    if (stream_) (void)hipStreamSynchronize(stream_);
    // the blocks go back to the pool only when nothing can still use them: the differentiable
    // renderer launches on its caller's stream, so wait for the device as hipFree would
    (void)hipDeviceSynchronize();
    release_blocks(owned_, true);
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    events_.clear();
    diff_ev_ = nullptr;  // destroyed with events_
    if (stream_) (void)hipStreamDestroy(stream_);
    stream_ = nullptr;
    if (host_counters_) (void)hipHostFree(host_counters_);
    host_counters_ = nullptr;
}

void Renderer::set_camera(const mt_camera_desc& cam) {
    cam_ = cam;
    ChunkParams& c = params_;
    c.cam_type = cam.type;
    std::memcpy(c.cam, cam.camera_to_world, sizeof(c.cam));
    c.fov_scale = std::tan(cam.yfov * 0.5f);
    c.ortho_x = cam.ortho_size[0];
    c.ortho_y = cam.ortho_size[1];
    c.cam_aspect = cam.aspect;
}

uint32_t seed_key(uint64_t seed) {
    const uint32_t k0 = seed_mix(static_cast<uint32_t>(seed) ^ 0xA511E9B3u);
    return seed_mix(k0 ^ static_cast<uint32_t>(seed >> 32) ^ 0x63D83595u);
}

void Renderer::set_seed(uint64_t seed) {
    st_.seed = seed;
    params_.seed_key = seed_key(seed);
}

void fill_params(ChunkParams& c, const mt_scene* scene, const mt_camera_desc& cam,
                 const mt_settings& st) {
    std::memset(&c, 0, sizeof(c));
    c.width = st.width;
    c.spp = st.samples;
    c.bounces = st.bounces;
    c.px_w = 1.0f / static_cast<float>(st.width);
    c.px_h = 1.0f / static_cast<float>(st.height);
    c.aspect = static_cast<float>(st.width) / static_cast<float>(st.height);
    c.cam_type = cam.type;
    std::memcpy(c.cam, cam.camera_to_world, sizeof(c.cam));
    c.fov_scale = std::tan(cam.yfov * 0.5f);
    c.ortho_x = cam.ortho_size[0];
    c.ortho_y = cam.ortho_size[1];
    c.cam_aspect = cam.aspect;
    c.seed_key = seed_key(st.seed);
    c.light_stride = 8u + scene->light_count;
    c.mode = st.mode;
    c.rr_depth = st.rr_depth;
    c.terminate_energy = st.terminate_energy;
    c.emission_base = st.emission_base_intensity;
    c.hdri_primary = st.hdri_primary_intensity;
    c.hdri_secondary = st.hdri_secondary_intensity;
    c.hdri_rot_apply = st.hdri_rotation_apply;
    c.hdri_rot_offset = st.hdri_rotation / 360.0f;
    c.exposure = std::pow(2.0f, -st.ev_exposure);
    c.output_linear = st.output == MT_OUTPUT_LINEAR;
    c.firefly = st.firefly_enable;
    c.firefly_threshold = st.firefly_threshold;
    c.count_work = st.count_work;
    c.debug = st.debug_mode;
    c.accel = st.traversal == MT_TRAVERSAL_FAST ? 1 : 0;
    c.stack_cap = stack_capacity(scene);
    c.refill_min = std::min<uint32_t>(64u, std::max<uint32_t>(1u, c.accel ? env_u32("MT_REFILL_MIN_FAST", kRefillMinFast)
                                                                          : env_u32("MT_REFILL_MIN", kRefillMin)));
    c.setup_min = std::min<uint32_t>(64u, std::max<uint32_t>(1u, env_u32("MT_SETUP_MIN", kSetupMin)));
    c.grab = std::min<uint32_t>(1u << 20, std::max<uint32_t>(64u, c.accel ? env_u32("MT_GRAB_FAST", kGrabFast)
                                                                         : env_u32("MT_GRAB", kGrab)));
    c.inv_spp = 1.0f / static_cast<float>(st.samples);
}

void launch_batch_iteration(const mt_scene* scene, const ChunkParams& c, const PathBuffers& b,
                            const BatchLaunch& L, hipStream_t s) {
    const uint32_t grid = (b.it_capacity + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_batch_schedule, dim3(1), dim3(64), 0, s, b);
    hipLaunchKernelGGL(k_batch_generate, dim3(grid), dim3(kBlock), 0, s, b);
    for (uint32_t k = 0; k < L.bounces; ++k) {
        launch_trace<0>(L.lds_depth, false, L.trace_blocks, s, scene->dev, c, b, k);
        if (scene->light_count)
            hipLaunchKernelGGL((k_shade<0, true, false>), dim3(L.shade_blocks), dim3(kBlock), 0, s, scene->dev, c, b, k);
        else
            hipLaunchKernelGGL((k_shade_unlit<true>), dim3(L.shade_blocks), dim3(kBlock), 0, s, scene->dev, c, b, k);
        if (scene->light_count) launch_trace<1>(L.lds_depth, false, L.trace_blocks, s, scene->dev, c, b, k);
    }
    hipLaunchKernelGGL(k_finalize<true>, dim3(grid), dim3(kBlock), 0, s, c, b);
    hipLaunchKernelGGL(k_batch_resolve, dim3(grid), dim3(kBlock), 0, s, b);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, s, p, v);
}

int persistent_blocks_per_cu(uint32_t* blocks) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_batch_persistent<kPersistDepth>, kBlock, 0) !=
            hipSuccess || n <= 0)
        return fail(MT_E_HIP, "occupancy query of the persistent batch kernel failed");
    *blocks = static_cast<uint32_t>(n);
    return MT_OK;
}

int launch_batch_persistent(const PersistCtl& ctl, hipStream_t s) {
    void* args[] = {const_cast<PersistCtl*>(&ctl)};
    const hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&k_batch_persistent<kPersistDepth>),
                                                    dim3(ctl.groups * ctl.group_blocks), dim3(kBlock), args, 0, s);
    if (e != hipSuccess)
        return fail(MT_E_HIP, std::string("cooperative launch of the persistent batch kernel failed: ") +
                                  hipGetErrorString(e));
    return MT_OK;
}

__global__ void k_xcc_ids(uint32_t* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

int xcd_count(uint32_t* xcds) {
    const uint32_t n = static_cast<uint32_t>(device_cu_count()) * 4u;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, sizeof(uint32_t) * n) != hipSuccess) return fail(MT_E_NOMEM, "XCD probe allocation failed");
    std::vector<uint32_t> h(n);
    hipLaunchKernelGGL(k_xcc_ids, dim3(n), dim3(64), 0, nullptr, d);
    const hipError_t e = hipMemcpy(h.data(), d, sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(MT_E_HIP, "XCD probe failed");
    uint32_t mx = 0;
    for (uint32_t v : h) mx = std::max(mx, v);
    *xcds = mx + 1u;
    return MT_OK;
}

int wall_ticks_per_ms(uint64_t* ticks) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        return fail(MT_E_HIP, "wall clock rate query failed");
    *ticks = static_cast<uint64_t>(khz);
    return MT_OK;
}

int alloc_path_buffers(std::vector<void*>& owned, PathBuffers& b, uint64_t P, uint64_t spill_entries,
                       bool diff, bool batch, bool debug) {
    int rc;
    if ((rc = alloc(owned, &b.ray_o[0], P)) || (rc = alloc(owned, &b.ray_d[0], P)) ||
        (rc = alloc(owned, &b.ray_o[1], P)) || (rc = alloc(owned, &b.ray_d[1], P)) ||
        (rc = alloc(owned, &b.t_rad[0], P)) || (rc = alloc(owned, &b.t_rad[1], P)) ||
        (rc = alloc(owned, &b.t_term[0], P)) ||
        (rc = alloc(owned, &b.t_term[1], P)) || (rc = alloc(owned, &b.prev_base[0], P)) ||
        (rc = alloc(owned, &b.prev_base[1], P)) || (rc = alloc(owned, &b.acc_e, P)) ||
        (rc = alloc(owned, &b.acc_d, P)) || (rc = alloc(owned, &b.acc_i, P)) ||
        (rc = alloc(owned, &b.acc_s, P)) || (rc = alloc(owned, &b.hit_rec, P)) ||
        (rc = alloc(owned, &b.hit_pos, P)) || (rc = alloc(owned, &b.queue[0], P)) ||
        (rc = alloc(owned, &b.queue[1], P)) || (rc = alloc(owned, &b.sh_o, P)) ||
        (rc = alloc(owned, &b.sh_d, P)) || (rc = alloc(owned, &b.sh_c, P)) ||
        (rc = alloc(owned, &b.sample_rgb, P)) || (rc = alloc(owned, &b.spill, spill_entries)) ||
        (rc = alloc(owned, &b.counters, kMaxCounters)) || (rc = alloc(owned, &b.work, 16)))
        return rc;
    if (diff && ((rc = alloc(owned, &b.vis, P)) || (rc = alloc(owned, &b.adj_s, P)) ||
                 (rc = alloc(owned, &b.adj_prev, P))))
        return rc;
    if (batch && ((rc = alloc(owned, &b.path_job, P)) || (rc = alloc(owned, &b.path_pix, P)) ||
                  (rc = alloc(owned, &b.path_smp, P))))
        return rc;
    if (debug && ((rc = alloc(owned, &b.dbg_a, P)) || (rc = alloc(owned, &b.dbg_b, P)) ||
                  (rc = alloc(owned, &b.dbg_c, P)) || (rc = alloc(owned, &b.dbg_d, P)) ||
                  (rc = alloc(owned, &b.dbg_e, P)) || (rc = alloc(owned, &b.dbg_i, P)) ||
                  (rc = alloc(owned, &b.dbg_j, P))))
        return rc;
    return MT_OK;
}

// the ray binning's buffers of one slot; the histogram starts at zero (the scan keeps it there)
int alloc_bins(std::vector<void*>& owned, PathBuffers& b, uint64_t P, hipStream_t s) {
    int rc;
    if ((rc = alloc(owned, &b.bin_key, P)) || (rc = alloc(owned, &b.perm, P)) ||
        (rc = alloc(owned, &b.bin_hist, kBinMax)) || (rc = alloc(owned, &b.bin_cur, kBinMax)))
        return rc;
    HIP_TRY(hipMemsetAsync(b.bin_hist, 0, sizeof(uint32_t) * kBinMax, s));
    return MT_OK;
}

// bins: dir_bits per direction component, the remaining key bits for the origin's cell in the
// scene's box, each bit to the axis with the longest cells (no binning for an empty box)
bool bin_grid(const mt_scene* scene, uint32_t dir_bits, uint32_t key_bits, ChunkParams& c) {
    c.bin_dir_bits = dir_bits;
    float ext[3];
    for (int a = 0; a < 3; ++a) {
        ext[a] = scene->world_hi[a] - scene->world_lo[a];
        if (!(ext[a] >= 0.0f) || !std::isfinite(ext[a])) return false;
        c.bin_bits[a] = 0;
        c.bin_lo[a] = scene->world_lo[a];
    }
    for (uint32_t i = 3 * dir_bits; i < key_bits; ++i) {
        int best = -1;
        float longest = 0.0f;
        for (int a = 0; a < 3; ++a) {
            const float cell = ext[a] / static_cast<float>(1u << c.bin_bits[a]);
            if (cell > longest) {
                longest = cell;
                best = a;
            }
        }
        if (best < 0) break;
        ++c.bin_bits[best];
    }
    for (int a = 0; a < 3; ++a)
        c.bin_scale[a] = c.bin_bits[a] ? static_cast<float>(1u << c.bin_bits[a]) / ext[a] : 0.0f;
    return true;
}

int Renderer::setup(mt_scene* scene, const mt_camera_desc& cam, const mt_settings& st, bool diff,
                    uint32_t max_slots) {
    scene_ = scene;
    cam_ = cam;
    st_ = st;
    if (st.width == 0 || st.height == 0 || st.samples == 0)
        return fail(MT_E_INVALID, "width, height and samples must be positive");
    if (st.bounces > kMaxBounces) return fail(MT_E_UNSUPPORTED, "bounces above 64 are not supported");
    if (st.traversal != MT_TRAVERSAL_REFERENCE && st.traversal != MT_TRAVERSAL_FAST)
        return fail(MT_E_INVALID, "unknown traversal");
    if (st.traversal == MT_TRAVERSAL_FAST && !scene->dev.fast_meshes && scene->mesh_count)
        return fail(MT_E_INVALID, "the scene was created without fast BVHs (mt_bvh_options.fast)");
    const uint64_t npix_total = static_cast<uint64_t>(st.width) * st.height;
    if (npix_total * 3 > 0xFFFFFFFFull) return fail(MT_E_UNSUPPORTED, "image too large");
    // Paths per chunk by default: an eighth of the render, at least 16 Mi and at most 64 Mi
    // (about 256 B of path state per path: 4 to 16 GB per stream slot).  Measured on the
    // 1920x1080 256-spp sphere grid (530.8 M paths), reference traversal: 4 Mi 375, 8 Mi 416,
    // 16 Mi 473, 24 Mi 481, 32 Mi 492, 48 Mi 494, 64 Mi 502 Msamples/s -- larger chunks spend
    // less of the frame in their sparse last bounces -- and 517 at 64 Mi once path buffers
    // are pooled across renders (without the pool, setting up and releasing 48 GB per render
    // cost the fast traversal's frame two thirds of its rate).  Eight chunks keep the three
    // stream slots busy on smaller renders (1920x1080 at 64 spp stays at 16.6 Mi).
    const uint64_t all_paths = npix_total * st.samples;
    // differentiable renders (a 512x512 pose at 32 spp is 8 Mi paths) split into two chunks
    // on two stream slots, so that one chunk's last, sparse bounces overlap the other's full
    // ones (alone, a traversal launch of the last bounces takes 1-1.6 ms for a few hundred
    // thousand rays).  Measured on that pose: forward 39.3 -> 36.9 ms, adjoint 28.4 -> 26.7;
    // six chunks cost more in launches than the overlap gains (47.8 and 33.1 ms on one to
    // three slots).
    uint64_t target = st.paths_in_flight ? st.paths_in_flight
                      : diff            ? std::clamp<uint64_t>(all_paths / kDiffChunks, 1ull << 20, 1ull << 26)
                                        : std::clamp<uint64_t>(all_paths / 8, 1ull << 24, 1ull << 26);
    target = std::max<uint64_t>(target, st.samples);
    target = std::min<uint64_t>(target, 1ull << 26);
    npix_chunk_ = static_cast<uint32_t>(std::max<uint64_t>(1, target / st.samples));
    npix_chunk_ = static_cast<uint32_t>(std::min<uint64_t>(npix_chunk_, npix_total));
    const uint64_t P = static_cast<uint64_t>(npix_chunk_) * st.samples;
    if (P > 0xFFFFFFFFull / 2) return fail(MT_E_UNSUPPORTED, "too many paths in flight");
    capacity_ = static_cast<uint32_t>(P);
    chunks_ = static_cast<uint32_t>((npix_total + npix_chunk_ - 1) / npix_chunk_);

    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    const int cus = device_cu_count();
    const uint32_t blocks_per_cu = env_u32("MT_TRACE_BLOCKS_PER_CU", 6);
    lds_depth_ = env_u32("MT_LDS_DEPTH", lds_stack_depth(st.traversal == MT_TRAVERSAL_FAST, blocks_per_cu));
    if (lds_depth_ != 4 && lds_depth_ != 8 && lds_depth_ != 12) lds_depth_ = 16;
    trace_blocks_ = static_cast<uint32_t>(cus * blocks_per_cu);
    shade_blocks_ = static_cast<uint32_t>(
        std::min<uint64_t>((P + kBlock - 1) / kBlock, cus * static_cast<uint64_t>(env_u32("MT_SHADE_BLOCKS_PER_CU", 16))));
    const uint32_t cap = stack_capacity(scene);
    const uint32_t spill_depth = cap > lds_depth_ ? cap - lds_depth_ : 1;
    const size_t trace_threads = static_cast<size_t>(trace_blocks_) * kBlock;
    PathBuffers& b = buf_;
    int rc;
    // instance masks computed before the reference traversal: scenes of 1 to 64 instances that
    // do not start at an instance tree
    premask_ = st.traversal == MT_TRAVERSAL_REFERENCE && scene->instance_count >= 1 &&
               scene->instance_count <= 64 && !(scene->dev.use_tlas & kTlasReference);
    if ((rc = alloc_path_buffers(owned_, b, P, spill_depth * trace_threads, diff, false, st.debug_mode != 0)) ||
        (premask_ && (rc = alloc(owned_, &b.inst_mask, P))) || (rc = alloc(owned_, &img_, npix_total * 3)))
        return rc;
    if (diff && (rc = alloc(owned_, &store_, npix_total * st.samples))) return rc;
    HIP_TRY(hipMemsetAsync(b.work, 0, sizeof(unsigned long long) * 16, stream_));
    HIP_TRY(hipMemsetAsync(img_, 0, sizeof(float) * npix_total * 3, stream_));
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&host_counters_), sizeof(uint32_t) * kMaxCounters));

    ChunkParams& c = params_;
    fill_params(c, scene, cam, st);
    // ray binning for the plain chunk loop (enqueue_chunk sets c.bin in its own copy): by
    // default for scenes whose traversal records exceed the 256 MB last-level cache (measured:
    // the large terrain scene +2 % reference / +5 % fast traversal; the sphere grid's 12 MB of
    // records, 98 % served by L2, gains nothing), MT_BIN_RAYS=0/1 to force it
    const uint64_t records = st.traversal == MT_TRAVERSAL_FAST
                                 ? scene->fast_record_bytes
                                 : scene->node_count * 32 + scene->prim_slots * 48;
    const bool beyond_llc = records > (256ull << 20);
    bool bin = beyond_llc;
    if (const char* v = std::getenv("MT_BIN_RAYS"); v && *v) bin = std::strtol(v, nullptr, 10) != 0;
    bin_blocks_ = 0;
    if (!diff && bin &&
        bin_grid(scene, std::min<uint32_t>(4u, std::max<uint32_t>(1u, env_u32("MT_BIN_DIR_BITS", 1))),
                 std::min<uint32_t>(kBinKeyBits, env_u32("MT_BIN_KEY_BITS", kBinKeyBits)), c)) {
        bin_blocks_ = static_cast<uint32_t>(cus) * 2u;
        if ((rc = alloc_bins(owned_, b, P, stream_))) return rc;
    }
    if (beyond_llc && !c.accel) {
        // records beyond the last-level cache: the reference traversal keeps its rays in flight
        // few and close together in queue order (measured on the large terrain scene: refill at
        // 24 idle lanes and batches of 64 positions, 17.0 Msamples/s, against 16.4 with the
        // settings of scenes that fit, which the sphere grid prefers: 700 against 665; the fast
        // traversal's short rays keep their settings there: 334 against 324 with these), and
        // its set-up steps (the instance tree's walk and instance entries, 4.8 per ray) run as
        // soon as 4 lanes want one (set-up threshold 2, 4, 8, 12, 20: 19.0, 18.9-19.1, 17.8,
        // 16.8, 15.0 Msamples/s)
        c.grab = std::max<uint32_t>(64u, std::min<uint32_t>(1u << 20, env_u32("MT_GRAB", kGrabLarge)));
        c.refill_min = std::min<uint32_t>(64u, std::max<uint32_t>(1u, env_u32("MT_REFILL_MIN", kRefillMinLarge)));
        c.setup_min = std::min<uint32_t>(64u, std::max<uint32_t>(1u, env_u32("MT_SETUP_MIN", kSetupMinLarge)));
    } else if (beyond_llc) {
        // and the fast traversal there, whose set-up steps are cheap next to its short node
        // walks, gathers more lanes for each (threshold 12, 20, 32, 48, 64: 345, 367, 382, 386,
        // 364 Msamples/s)
        c.setup_min = std::min<uint32_t>(64u, std::max<uint32_t>(1u, env_u32("MT_SETUP_MIN", kSetupMinLargeFast)));
    }
    c.stack_cap = cap;
    c.inv_spp = 1.0f / static_cast<float>(st.samples);
    if (diff) {  // the differentiable image is the linear radiance, no firefly rejection
        c.output_linear = 1;
        c.firefly = 0;
        c.path_store = store_;
    }
    for (uint32_t i = 0; i < 2 + 4 * st.bounces; ++i) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        events_.push_back(e);
    }
    if (diff) {  // the slots' streams start after the caller's stream (run_diff)
        HIP_TRY(hipEventCreateWithFlags(&diff_ev_, hipEventDisableTiming));
        events_.push_back(diff_ev_);
    }
    slots_.clear();
    slots_.push_back(StreamSlot{buf_, stream_, host_counters_, events_, -1, {}});
    // more slots for plain renders with several chunks (MT_STREAMS, default 3)
    // an explicit max_slots is taken as is (cached renders pipeline across passes too)
    const uint32_t want = diff ? std::min<uint32_t>(std::min<uint32_t>(env_u32("MT_STREAMS", kDiffSlots), 8), chunks_)
                          : max_slots ? std::min<uint32_t>(max_slots, 8)
                                      : std::min<uint32_t>(std::min<uint32_t>(env_u32("MT_STREAMS", 3), 8), chunks_);
    for (uint32_t i = 1; i < want; ++i) {
        StreamSlot L;
        HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        slots_.push_back(L);  // registered first, so that release() frees what follows
        StreamSlot& S = slots_.back();
        if ((rc = alloc_path_buffers(owned_, S.b, P, spill_depth * trace_threads, diff, false,
                                     st.debug_mode != 0)) ||
            (premask_ && (rc = alloc(owned_, &S.b.inst_mask, P))))
            return rc;
        if (bin_blocks_ && (rc = alloc_bins(owned_, S.b, P, S.stream))) return rc;
        HIP_TRY(hipMemsetAsync(S.b.work, 0, sizeof(unsigned long long) * 16, S.stream));
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&S.host_counters), sizeof(uint32_t) * kMaxCounters));
        for (uint32_t k = 0; k < 2 + 4 * st.bounces; ++k) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            S.events.push_back(e);
        }
    }
    return MT_OK;
}

uint32_t seed_mix(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

int Renderer::run_chunk(uint32_t chunk, float* sample_rgb_out) {
    const int rc = enqueue_chunk(0, chunk);
    return rc ? rc : complete_chunk(0, sample_rgb_out);
}

int Renderer::enqueue_chunk(uint32_t slot, uint32_t chunk) {
    StreamSlot& L = slots_[slot];
    const uint64_t npix_total = static_cast<uint64_t>(st_.width) * st_.height;
    ChunkParams c = params_;
    c.pix_begin = chunk * npix_chunk_;
    c.npix = static_cast<uint32_t>(std::min<uint64_t>(npix_chunk_, npix_total - c.pix_begin));
    c.paths = c.npix * st_.samples;
    c.bin = bin_blocks_ != 0;
    const uint32_t gen_blocks = (c.paths + kBlock - 1) / kBlock;
    hipStream_t s = L.stream;
    const PathBuffers& b = L.b;
    std::vector<hipEvent_t>& events = L.events;
    // events: [0] chunk start, [1] chunk end, then per bounce 4: before extend, after extend,
    // after shade, after shadow (all on this slot's stream, read after the chunk's sync)
    HIP_TRY(hipEventRecord(events[0], s));
    hipLaunchKernelGGL(k_generate, dim3(gen_blocks), dim3(kBlock), 0, s, c, b);
    for (uint32_t k = 0; k < st_.bounces; ++k) {
        hipEvent_t* ev = &events[2 + 4 * k];
        if (c.bin && k > 0) {  // the bins of the rays shading queued for this bounce
            hipLaunchKernelGGL(k_bin_count, dim3(bin_blocks_), dim3(kBlock), 0, s, b, k);
            hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(kBlock), 0, s, b);
            hipLaunchKernelGGL(k_bin_scatter, dim3(bin_blocks_), dim3(kBlock), 0, s, b, k);
        }
        HIP_TRY(hipEventRecord(ev[0], s));
        launch_trace<0>(lds_depth_, c.count_work != 0, trace_blocks_, s, scene_->dev, c, b, k);
        HIP_TRY(hipEventRecord(ev[1], s));
        if (c.debug || c.cache)
            hipLaunchKernelGGL((k_shade<0, false, true>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
        else if (scene_->light_count)
            hipLaunchKernelGGL((k_shade<0, false, false>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
        else
            hipLaunchKernelGGL((k_shade_unlit<false>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
        HIP_TRY(hipEventRecord(ev[2], s));
        if (scene_->light_count)
            launch_trace<1>(lds_depth_, c.count_work != 0, trace_blocks_, s, scene_->dev, c, b, k);
        HIP_TRY(hipEventRecord(ev[3], s));
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_finalize<false>, dim3(gen_blocks), dim3(kBlock), 0, s, c, b);
    hipLaunchKernelGGL(k_resolve, dim3((c.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c, b, img_);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(events[1], s));
    HIP_TRY(hipMemcpyAsync(L.host_counters, b.counters, sizeof(uint32_t) * kMaxCounters,
                           hipMemcpyDeviceToHost, s));
    L.chunk = chunk;
    L.c = c;
    return MT_OK;
}

int Renderer::complete_chunk(uint32_t slot, float* sample_rgb_out) {
    StreamSlot& L = slots_[slot];
    if (L.chunk < 0) return MT_OK;
    const ChunkParams& c = L.c;
    L.chunk = -1;
    HIP_TRY(hipStreamSynchronize(L.stream));
    const std::vector<hipEvent_t>& events = L.events;
    const uint32_t* hc = L.host_counters;
    float total = 0, ms_ext = 0, ms_shade = 0, ms_sh = 0;
    HIP_TRY(hipEventElapsedTime(&total, events[0], events[1]));
    for (uint32_t k = 0; k < st_.bounces; ++k) {
        const hipEvent_t* ev = &events[2 + 4 * k];
        float a = 0, b2 = 0, d = 0;
        HIP_TRY(hipEventElapsedTime(&a, ev[0], ev[1]));
        HIP_TRY(hipEventElapsedTime(&b2, ev[1], ev[2]));
        HIP_TRY(hipEventElapsedTime(&d, ev[2], ev[3]));
        ms_ext += a;
        ms_shade += b2;
        ms_sh += d;
    }
    stats_.extend_launches += st_.bounces;
    if (scene_->light_count) stats_.shadow_launches += st_.bounces;
    if (hc[kCntError]) return fail(MT_E_INVALID, "traversal stack overflow");
    uint64_t rays = 0;
    for (uint32_t k = 0; k < st_.bounces; ++k)
        rays += static_cast<uint64_t>(hc[kCntExtend + k]) + hc[kCntShadow + k];
    stats_.rays += rays;
    stats_.samples += c.paths;
    stats_.kernel_ms_total += total;
    stats_.kernel_ms_extend += ms_ext;
    stats_.kernel_ms_shade += ms_shade;
    stats_.kernel_ms_shadow += ms_sh;
    if (sample_rgb_out) {
        std::vector<float4> tmp(c.paths);
        HIP_TRY(hipMemcpy(tmp.data(), L.b.sample_rgb, sizeof(float4) * c.paths, hipMemcpyDeviceToHost));
        for (uint32_t p = 0; p < c.paths; ++p) {
            const uint32_t lp = p / c.spp, smp = p % c.spp;
            float* q = sample_rgb_out + (static_cast<size_t>(c.pix_begin + lp) * st_.samples + smp) * 3;
            q[0] = tmp[p].x;
            q[1] = tmp[p].y;
            q[2] = tmp[p].z;
        }
    }
    return MT_OK;
}

// one chunk of the differentiable forward (mode 1) or of the adjoint replay (mode 2), enqueued
// on slot L's stream with L's path state
int Renderer::run_diff_chunk(uint32_t chunk, int mode, float* img, StreamSlot& L) {
    const uint64_t npix_total = static_cast<uint64_t>(st_.width) * st_.height;
    ChunkParams c = params_;
    c.diff = mode;
    c.pix_begin = chunk * npix_chunk_;
    c.npix = static_cast<uint32_t>(std::min<uint64_t>(npix_chunk_, npix_total - c.pix_begin));
    c.paths = c.npix * st_.samples;
    const PathBuffers& b = L.b;
    const hipStream_t s = L.stream;
    const uint32_t gen_blocks = (c.paths + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_generate, dim3(gen_blocks), dim3(kBlock), 0, s, c, b);
    for (uint32_t k = 0; k < st_.bounces; ++k) {
        launch_trace<0>(lds_depth_, false, trace_blocks_, s, scene_->dev, c, b, k);
        if (mode == 1) {
            hipLaunchKernelGGL((k_shade<1, false, false>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
            if (scene_->light_count)
                launch_trace<1>(lds_depth_, false, trace_blocks_, s, scene_->dev, c, b, k);
        } else {
            hipLaunchKernelGGL((k_shade<2, false, false>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
        }
        HIP_TRY(hipGetLastError());
    }
    if (mode == 1) {
        hipLaunchKernelGGL(k_finalize<false>, dim3(gen_blocks), dim3(kBlock), 0, s, c, b);
        hipLaunchKernelGGL(k_resolve, dim3((c.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, s, c, b, img);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(L.host_counters, b.counters, sizeof(uint32_t) * kMaxCounters,
                           hipMemcpyDeviceToHost, s));
    L.chunk = chunk;
    L.c = c;
    return MT_OK;
}

int Renderer::finish_diff_chunk(StreamSlot& L, int mode) {
    if (L.chunk < 0) return MT_OK;
    L.chunk = -1;
    HIP_TRY(hipStreamSynchronize(L.stream));
    const uint32_t* hc = L.host_counters;
    if (hc[kCntError]) return fail(MT_E_INVALID, "traversal stack overflow");
    uint64_t rays = 0;
    for (uint32_t k = 0; k < st_.bounces; ++k)
        rays += static_cast<uint64_t>(hc[kCntExtend + k]) + (mode == 1 ? hc[kCntShadow + k] : 0u);
    stats_.rays += rays;
    stats_.samples += L.c.paths;
    return MT_OK;
}

// every chunk of one differentiable pass, round-robin over the slots; the slots' streams start
// after the caller's stream s, and the call returns when every chunk has finished
int Renderer::run_diff(int mode, float* img, hipStream_t s) {
    HIP_TRY(hipEventRecord(diff_ev_, s ? s : stream_));
    for (StreamSlot& L : slots_) HIP_TRY(hipStreamWaitEvent(L.stream, diff_ev_, 0));
    int rc = MT_OK;
    const uint32_t n = static_cast<uint32_t>(slots_.size());
    for (uint32_t ch = 0; ch < chunks_ && rc == MT_OK; ++ch) {
        StreamSlot& L = slots_[ch % n];
        rc = finish_diff_chunk(L, mode);
        if (rc == MT_OK) rc = run_diff_chunk(ch, mode, img, L);
    }
    for (StreamSlot& L : slots_) {  // drain every slot, also after an error
        const int r = finish_diff_chunk(L, mode);
        if (rc == MT_OK) rc = r;
    }
    return rc;
}

int Renderer::diff_forward(const DiffParams& dp, const DiffParams* sampling, float* img_dev,
                           hipStream_t s) {
    if (!store_) return fail(MT_E_INVALID, "renderer was not set up for differentiable rendering");
    params_.dp = dp;
    params_.frozen = sampling != nullptr;
    params_.dps = sampling ? *sampling : dp;
    params_.gr = DiffGradPtrs{};
    params_.grad_img = nullptr;
    return run_diff(1, img_dev, s);
}

int Renderer::diff_adjoint(const DiffParams& dp, const DiffGradPtrs& gr, const float* grad_img,
                           hipStream_t s) {
    if (!store_) return fail(MT_E_INVALID, "renderer was not set up for differentiable rendering");
    params_.dp = dp;
    params_.frozen = 0;
    params_.gr = gr;
    params_.grad_img = grad_img;
    return run_diff(2, nullptr, s);
}

int Renderer::setup_cached(mt_scene* scene, const mt_camera_desc& cam, const mt_settings& st,
                           const mt_cache_options& opt) {
    copt_ = opt;
    if (copt_.spp_per_pass == 0) copt_.spp_per_pass = 1;
    if (copt_.train_every == 0) copt_.train_every = 16;
    if (copt_.query_depth == 0 && !copt_.query_after_light) copt_.query_depth = 1;
    if (copt_.train_batch == 0) copt_.train_batch = 16384;
    if (copt_.steps_per_pass == 0) copt_.steps_per_pass = 8;
    if (st.samples % copt_.spp_per_pass)
        return fail(MT_E_INVALID, "samples must be a multiple of spp_per_pass");
    if (copt_.query_depth >= st.bounces)
        return fail(MT_E_INVALID, "query_depth must be below the bounce count");
    if (st.debug_mode) return fail(MT_E_INVALID, "debug views are rendered by mt_render");
    if (copt_.train_nee && !copt_.query_after_light)
        return fail(MT_E_INVALID, "train_nee needs query_after_light");
    total_spp_ = st.samples;
    mt_settings pass = st;
    pass.samples = copt_.spp_per_pass;
    pass.firefly_enable = 0;
    const uint64_t npix_total = static_cast<uint64_t>(st.width) * st.height;
    int rc = setup(scene, cam, pass, false, std::min<uint32_t>(env_u32("MT_STREAMS", 3), 8));
    if (rc) return rc;
    const uint64_t P = capacity_;
    // after-light: the querying paths ask only at query_depth, so the cache learns only there
    // (deeper vertices have fewer bounces left and would pull its estimate down)
    const uint32_t nrec = copt_.query_after_light ? 1u : st.bounces - copt_.query_depth;
    const uint64_t tslots = P + 1;  // warm-up passes may train on every path
    float* img_sum = nullptr;
    if ((rc = alloc(owned_, &img_sum, npix_total * 3)) ||
        (rc = alloc(owned_, &rays_dev_, 1)))
        return rc;
    for (StreamSlot& S : slots_) {
        PathBuffers& b = S.b;
        if ((rc = alloc(owned_, &b.train_slot, P)) || (rc = alloc(owned_, &b.tr_path, tslots)) ||
            (rc = alloc(owned_, &b.tr_pos, tslots * nrec)) || (rc = alloc(owned_, &b.tr_dir, tslots * nrec)) ||
            (rc = alloc(owned_, &b.tr_w, tslots * nrec)) || (rc = alloc(owned_, &b.tr_s, tslots * nrec)) ||
            (rc = alloc(owned_, &b.cq_path, P)) || (rc = alloc(owned_, &b.cq_pos, P)) ||
            (rc = alloc(owned_, &b.cq_dir, P)) || (rc = alloc(owned_, &b.cq_out, P)) ||
            (rc = alloc(owned_, &b.cq_t, P)) ||
            (rc = alloc(owned_, &b.tr_img, P)) ||
            (rc = alloc(owned_, &b.tp_pos, tslots * nrec)) || (rc = alloc(owned_, &b.tp_dir, tslots * nrec)) ||
            (rc = alloc(owned_, &b.tp_tgt, tslots * nrec)) ||
            (rc = alloc(owned_, &b.ts_pos, tslots * nrec)) || (rc = alloc(owned_, &b.ts_dir, tslots * nrec)) ||
            (rc = alloc(owned_, &b.ts_tgt, tslots * nrec)))
            return rc;
        b.img_sum = img_sum;
        HIP_TRY(hipEventCreateWithFlags(&S.ev_bounce, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&S.ev_tail, hipEventDisableTiming));
    }
    buf_ = slots_[0].b;
    HIP_TRY(hipStreamCreateWithFlags(&tail_stream_, hipStreamNonBlocking));
    params_.cache = 1;
    params_.cache_depth = copt_.query_depth;
    params_.cache_after = copt_.query_after_light ? 1 : 0;
    params_.cache_l2 = copt_.loss ? 1 : 0;
    params_.train_cap = static_cast<uint32_t>(tslots);
    params_.train_nrec = nrec;
    // test hook: every training path's image sample is the full estimate its target comes from
    params_.cache_full_image = env_u32("MT_CACHE_TRAINING_IMAGE", 0) ? 1 : 0;
    if (copt_.train_nee) {
        params_.cache_nee = 1;
        for (StreamSlot& S : slots_) {
            if ((rc = alloc(owned_, &S.b.acc_n, P))) return rc;
        }
        buf_ = slots_[0].b;
        if ((rc = setup_env_sampling())) return rc;
    }
    return MT_OK;
}

// The training paths' environment sampling (cache_nee): the P_below table of the bounce
// density and, for a scene with a non-black environment, the two-level CDF over its texels.
// A texel's weight is its mean channel value, plus (linear filter) the three neighbours the
// filter blends in with weights up to one texel in uv units, times sin(theta) of its row.
int Renderer::setup_env_sampling() {
    int rc;
    float* tab = nullptr;
    if ((rc = alloc(owned_, &tab, static_cast<size_t>(kPbCos) * kPbRough))) return rc;
    hipLaunchKernelGGL(k_pbelow_table, dim3(kPbCos * kPbRough), dim3(64), 0, stream_, tab);
    HIP_TRY(hipGetLastError());
    params_.pbelow = tab;
    params_.env_pc = 0.0f;
    const DevScene& D = scene_->dev;
    if (!D.has_env || !D.env_w || !D.env_h) return MT_OK;
    const uint32_t W = D.env_w, H = D.env_h;
    const double blend = D.env_filter == MT_FILTER_NEAREST ? 0.0 : 1.0 / std::min(W, H);
    float *dm = nullptr, *dc = nullptr;
    double* drows = nullptr;
    if ((rc = alloc(owned_, &dm, static_cast<size_t>(H) + 1)) ||
        (rc = alloc(owned_, &dc, static_cast<size_t>(H) * (W + 1))) ||
        (rc = alloc(owned_, &drows, static_cast<size_t>(H) + 1)))
        return rc;
    hipLaunchKernelGGL(k_env_rows, dim3(H), dim3(256), 0, stream_, D.env, W, H, blend, dc, drows);
    hipLaunchKernelGGL(k_env_marginal, dim3(1), dim3(64), 0, stream_, drows, H, dm, drows + H);
    HIP_TRY(hipGetLastError());
    double total = 0.0;
    HIP_TRY(hipMemcpyAsync(&total, drows + H, sizeof(double), hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    if (!(total > 0.0)) return MT_OK;  // a black environment: nothing to sample
    params_.env_marg = dm;
    params_.env_cond = dc;
    params_.env_pc = scene_->light_count ? 0.5f : 1.0f;
    return MT_OK;
}

// One chunk of a cached pass.  The bounce loop runs on the slot's stream; everything that
// touches the cache (inference, its use, finalize/accumulate, training samples, Adam steps)
// runs in chunk order on one tail stream, so the cache is trained on chunk c before chunk c+1
// queries it and no two kernels use its buffers at once.  Nothing waits on the host.
int Renderer::run_cache_chunk(Mlp& mlp, uint32_t slot, uint32_t chunk, uint32_t sample_base,
                              uint32_t train_every, uint32_t train_offset, bool accumulate) {
    StreamSlot& S = slots_[slot];
    const uint64_t npix_total = static_cast<uint64_t>(st_.width) * st_.height;
    ChunkParams c = params_;
    c.pix_begin = chunk * npix_chunk_;
    c.npix = static_cast<uint32_t>(std::min<uint64_t>(npix_chunk_, npix_total - c.pix_begin));
    c.paths = c.npix * st_.samples;
    c.sample_base = sample_base;
    c.train_every = train_every;
    c.train_offset = train_offset % train_every;
    if (!accumulate) {
        // a training-only (warm-up) pass: only its training paths, the chunk's paths p with
        // (p + offset) % train_every == 0, run, as paths 0, 1, ... of the pass, each its own
        // training slot (the other paths' image samples would be thrown away)
        if (train_every == kNoSlot) return MT_OK;
        const uint32_t first = (train_every - c.train_offset) % train_every;
        if (c.paths <= first) return MT_OK;
        c.path_stride = train_every;
        c.path_first = first;
        c.paths = (c.paths - 1 - first) / train_every + 1;
        c.train_every = 1;
        c.train_offset = 0;
    }
    const uint32_t gen_blocks = (c.paths + kBlock - 1) / kBlock;
    hipStream_t s = S.stream, m = tail_stream_;
    const PathBuffers& b = S.b;
    HIP_TRY(hipStreamWaitEvent(s, S.ev_tail, 0));  // the slot's previous chunk is finished
    hipLaunchKernelGGL(k_generate, dim3(gen_blocks), dim3(kBlock), 0, s, c, b);
    for (uint32_t k = 0; k < st_.bounces; ++k) {
        launch_trace<0>(lds_depth_, false, trace_blocks_, s, scene_->dev, c, b, k);
        hipLaunchKernelGGL((k_shade<0, false, true>), dim3(shade_blocks_), dim3(kBlock), 0, s, scene_->dev, c, b, k);
        if (scene_->light_count || c.env_pc > 0.0f)
            launch_trace<1>(lds_depth_, false, trace_blocks_, s, scene_->dev, c, b, k);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(S.ev_bounce, s));
    HIP_TRY(hipStreamWaitEvent(m, S.ev_bounce, 0));
    hipLaunchKernelGGL(k_count_rays, dim3(1), dim3(64), 0, m, b, st_.bounces, rays_dev_);
    int rc = mlp.infer(c.paths, b.cq_pos, b.cq_dir, b.cq_out, m, b.counters + kCntQuery);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cache_apply, dim3(gen_blocks), dim3(kBlock), 0, m, c, b);
    hipLaunchKernelGGL(k_finalize<false>, dim3(gen_blocks), dim3(kBlock), 0, m, c, b);
    if (accumulate)
        hipLaunchKernelGGL(k_cache_accum, dim3((c.npix + kBlock - 1) / kBlock), dim3(kBlock), 0, m, c, b);
    // training slots of this chunk: paths p with (p + offset) % train_every == 0, numbered
    // (p + offset) / train_every: [slot0, slot0 + nslots)
    const uint32_t te = c.train_every, off = c.train_offset;
    const uint32_t first = (te - off) % te;
    const uint32_t nslots = (te != kNoSlot && c.paths > first) ? (c.paths - 1 - first) / te + 1 : 0;
    const uint32_t slot0 = (first + off) / te;
    if (nslots) {
        hipLaunchKernelGGL(k_cache_targets, dim3((nslots + kBlock - 1) / kBlock), dim3(kBlock), 0, m, c, b,
                           slot0, nslots);
        hipLaunchKernelGGL(k_cache_scale_init, dim3(1), dim3(1024), 0, m, c, b);
        hipLaunchKernelGGL(k_cache_shuffle, dim3((nslots * c.train_nrec + kBlock - 1) / kBlock), dim3(kBlock),
                           0, m, c, b);
        const uint32_t batch = copt_.train_batch;
        for (uint32_t st = 0; st < copt_.steps_per_pass; ++st)
            if ((rc = mlp.train(batch, b.ts_pos, b.ts_dir, b.ts_tgt, m, nullptr, nullptr,
                                b.counters + kCntTrainRec, st * batch)))
                return rc;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(S.ev_tail, m));
    if (accumulate) stats_.samples += c.paths;
    return MT_OK;
}

int Renderer::render_cached(Mlp& mlp, float* rgb_out) {
    const uint64_t npix_total = static_cast<uint64_t>(st_.width) * st_.height;
    hipStream_t m = tail_stream_;
    mlp.set_bounds(scene_->world_lo, scene_->world_hi);
    params_.cache_scale = mlp.scale();
    const uint64_t trained0 = mlp.trained_samples();
    HIP_TRY(hipDeviceSynchronize());
    hipLaunchKernelGGL(k_fill, dim3(static_cast<uint32_t>((npix_total * 3 + kBlock - 1) / kBlock)), dim3(kBlock), 0, m,
                       buf_.img_sum, npix_total * 3);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, m, reinterpret_cast<float*>(rays_dev_), 2ull);
    HIP_TRY(hipEventRecord(events_[0], m));
    for (StreamSlot& S : slots_) {  // counters start at zero (the overflow flag is sticky here)
        hipLaunchKernelGGL(k_fill, dim3((kMaxCounters + kBlock - 1) / kBlock), dim3(kBlock), 0, m,
                           reinterpret_cast<float*>(S.b.counters), static_cast<uint64_t>(kMaxCounters));
        HIP_TRY(hipEventRecord(S.ev_tail, m));  // slots start after the fills
    }
    const uint32_t te = copt_.train_every;
    const uint32_t nslot = static_cast<uint32_t>(slots_.size());
    uint32_t seq = 0;
    // warm-up: training-only passes with more training paths, on samples the image never uses
    for (uint32_t w = 0; w < copt_.warmup_passes; ++w)
        for (uint32_t ch = 0; ch < chunks_; ++ch, ++seq) {
            const int rc = run_cache_chunk(mlp, seq % nslot, ch, 0x40000000u + w * copt_.spp_per_pass,
                                           te == kNoSlot ? te : std::max<uint32_t>(1u, te / 4), w, false);
            if (rc) return rc;
        }
    const uint32_t passes = total_spp_ / copt_.spp_per_pass;
    for (uint32_t pass = 0; pass < passes; ++pass)
        for (uint32_t ch = 0; ch < chunks_; ++ch, ++seq) {
            const int rc = run_cache_chunk(mlp, seq % nslot, ch, pass * copt_.spp_per_pass, te, pass * 7u, true);
            if (rc) return rc;
        }
    hipLaunchKernelGGL(k_cache_resolve, dim3(static_cast<uint32_t>((npix_total * 3 + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, m, buf_.img_sum, static_cast<uint32_t>(npix_total * 3), static_cast<float>(total_spp_), img_);
    HIP_TRY(hipEventRecord(events_[1], m));
    HIP_TRY(hipDeviceSynchronize());
    float total = 0.0f;
    HIP_TRY(hipEventElapsedTime(&total, events_[0], events_[1]));
    stats_.kernel_ms_total += total;
    unsigned long long rays = 0;
    HIP_TRY(hipMemcpy(&rays, rays_dev_, sizeof(rays), hipMemcpyDeviceToHost));
    stats_.rays += rays;
    stats_.diag[8] = mlp.trained_samples() - trained0;
    for (const StreamSlot& S : slots_) {
        uint32_t err = 0;
        HIP_TRY(hipMemcpy(&err, S.b.counters + kCntError, sizeof(err), hipMemcpyDeviceToHost));
        if (err) return fail(MT_E_INVALID, "traversal stack overflow");
    }
    if (rgb_out) HIP_TRY(hipMemcpy(rgb_out, img_, sizeof(float) * npix_total * 3, hipMemcpyDeviceToHost));
    return MT_OK;
}

int Renderer::finish(float* rgb_out) {
    const uint64_t npix_total = static_cast<uint64_t>(st_.width) * st_.height;
    if (st_.count_work) {
        unsigned long long w[16] = {};
        for (const StreamSlot& S : slots_) {
            unsigned long long ws[16];
            HIP_TRY(hipMemcpy(ws, S.b.work, sizeof(ws), hipMemcpyDeviceToHost));
            for (int i = 0; i < 16; ++i) w[i] += ws[i];
        }
        for (int i = 0; i < 9; ++i) stats_.diag[i] = w[7 + i];
        stats_.inst_aabb_tests = w[0];
        stats_.inst_entered = w[1];
        stats_.node_pops = w[2];
        stats_.tri_tests = w[3];
        stats_.hits_shaded = w[4];
        stats_.misses = w[5];
    }
    if (rgb_out) {
        HIP_TRY(hipMemcpy(rgb_out, img_, sizeof(float) * npix_total * 3, hipMemcpyDeviceToHost));
        if (st_.debug_mode) debug_post(st_.debug_mode, rgb_out, npix_total);
    }
    return MT_OK;
}

// ApplyDebugPostProcessing (Tracer.cpp:642-716): a heat map (colourful or grey) of
// (r+g+b)/3 over the image's maximum, for the counting and measuring views
void debug_post(int mode, float* rgb, uint64_t npix) {
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
    static const float ramp[5][3] = {{0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}};
    float mx = 0.0f;
    for (uint64_t i = 0; i < npix; ++i) mx = std::max(mx, (rgb[i * 3] + rgb[i * 3 + 1] + rgb[i * 3 + 2]) / 3.0f);
    for (uint64_t i = 0; i < npix; ++i) {
        const float in = (rgb[i * 3] + rgb[i * 3 + 1] + rgb[i * 3 + 2]) / 3.0f;
        float t = mx > 0.0f ? in / mx : 0.0f;
        if (!colorful) {
            rgb[i * 3] = rgb[i * 3 + 1] = rgb[i * 3 + 2] = t;
            continue;
        }
        t = std::min(std::max(t, 0.0f), 1.0f);
        const float x = t * 4.0f;
        const int k = std::min(static_cast<int>(x), 3);
        const float f = x - static_cast<float>(k);
        for (int ch = 0; ch < 3; ++ch) rgb[i * 3 + ch] = ramp[k][ch] * (1.0f - f) + ramp[k + 1][ch] * f;
    }
}

}  // namespace mt
