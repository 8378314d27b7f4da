// mt_render.hpp -- render driver of the gfx950 Tracer (internal).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "mt_internal.hpp"

namespace mt {

constexpr uint32_t kMaxBounces = 64;
// layout of the per-chunk device counter array
constexpr uint32_t kCntExtend = 0;                              // [kMaxBounces + 1] queue sizes
constexpr uint32_t kCntShadow = kCntExtend + kMaxBounces + 1;   // [kMaxBounces] shadow queue sizes
constexpr uint32_t kCntFetchExt = kCntShadow + kMaxBounces;     // [kMaxBounces] work fetch
constexpr uint32_t kCntFetchSh = kCntFetchExt + kMaxBounces;    // [kMaxBounces] work fetch
constexpr uint32_t kCntError = kCntFetchSh + kMaxBounces;       // traversal stack overflow flag
constexpr uint32_t kCntPaths = kCntError + 1;                   // batched: paths this iteration
constexpr uint32_t kCntPixels = kCntPaths + 1;                  // batched: pixels this iteration
constexpr uint32_t kCntTiles = kCntPixels + 1;                  // batched: tiles this iteration
constexpr uint32_t kCntQuery = kCntTiles + 1;                   // cached: cache queries
constexpr uint32_t kCntTrainRec = kCntQuery + 1;                // cached: training samples
constexpr uint32_t kMaxCounters = kCntTrainRec + 1;
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// differentiable parameters (device pointers, null = the scene's value): a float RGB texture that
// replaces the base colour of materials [tex_first, tex_first + tex_count), and per-material
// roughness, metallic and RGB emission
struct DiffParams {
    const float* tex;
    uint32_t tex_w, tex_h, tex_first, tex_count;
    const float* rough;
    const float* metal;
    const float* emis;
};
// gradient accumulators of the same layout (null = not wanted); sampling = 1: include the
// sampling decisions' dependence on the parameters (score-function terms)
struct DiffGradPtrs {
    float* tex;
    float* rough;
    float* metal;
    float* emis;
    int32_t sampling;
    uint32_t materials;  // the scene's material count (per-material gradients are indexed by it)
};

// per-chunk constants, passed to kernels by value
struct ChunkParams {
    uint32_t width, spp, bounces;
    uint32_t pix_begin, npix, paths;
    float px_w, px_h, aspect;
    int32_t cam_type;
    float cam[16];
    float fov_scale, ortho_x, ortho_y, cam_aspect;
    uint32_t seed_key, light_stride;
    int32_t mode;
    uint32_t rr_depth;
    float terminate_energy, emission_base, hdri_primary, hdri_secondary;
    int32_t hdri_rot_apply;
    float hdri_rot_offset;
    float exposure;
    int32_t output_linear, firefly;
    float firefly_threshold;
    int32_t count_work;
    uint32_t stack_cap;
    uint32_t refill_min;  // idle lanes of a wave that trigger a refill of the traversal
    uint32_t setup_min;   // lanes needing an instance set-up that trigger a set-up step
    uint32_t grab;        // queue positions a wave of the traversal takes per atomic (>= 64)
    // differentiable rendering: 0 plain, 1 forward with per-path store, 2 adjoint replay
    int32_t diff;
    int32_t frozen;        // differentiable forward: decisions from dps, values from dp
    DiffParams dp;
    DiffParams dps;
    DiffGradPtrs gr;
    float4* path_store;    // [npix_total * spp] {linear radiance, shadow visibility bits}
    const float* grad_img; // [npix_total * 3] dLoss/dImage (adjoint replay)
    float inv_spp;
    uint32_t sample_base;  // sample index of the chunk's first sample of every pixel
    // radiance cache: paths that are not training paths stop at their hit of bounce
    // cache_depth and take the rest of their radiance from the cache
    int32_t cache;
    uint32_t cache_depth, train_every, train_offset, train_cap, train_nrec;
    int32_t cache_after;  // the query follows the vertex's emission and light sample
    // cache_nee: the query is made whenever a bounce is left, a stopped training path's record
    // is completed at the end of the path, and from its query vertex on a training path draws
    // environment samples with probability env_pc per vertex (0: no environment sampling)
    int32_t cache_nee;
    float env_pc;
    const float* env_marg;  // [env_h + 1] CDF of the environment's sampling weight over rows
    const float* env_cond;  // [env_h][env_w + 1] CDF over the cells of every row
    const float* pbelow;    // [kPbRough][kPbCos] chance that a specular draw reflects below
    int32_t cache_full_image;  // test hook: a training path's sample is its full estimate
    // {scale, set}: the cache learns targets divided by the scale (the first training chunk's
    // mean target) and its predictions are multiplied by it
    float* cache_scale;
    int32_t cache_l2;  // the cache trains with the L2 loss (scaled targets)
    // a training-only pass runs the chunk's paths path_first + k * path_stride (stride 0 or 1:
    // every path)
    uint32_t path_stride, path_first;
    int32_t debug;  // MT_DEBUG_* view (plain renders)
    int32_t accel;  // 1 = fast traversal (MT_TRAVERSAL_FAST)
    // ray binning before the extend of bounces >= 1 (plain renders): shading writes every
    // continuing ray's bin (direction octant, then the origin's cell on a grid of
    // 2^bin_bits[a] cells per axis over the scene's box), extend takes rays bin by bin
    int32_t bin;
    uint32_t bin_dir_bits;  // cells per direction component: 2^bin_dir_bits over [-1, 1]
    uint32_t bin_bits[3];
    float bin_lo[3], bin_scale[3];
};

// SoA path state, queues and scratch of one render (capacity = paths per chunk)
struct PathBuffers {
    // rays of bounce k in queue order (slot i = the path at queue[k & 1][i]), double-buffered by
    // bounce parity, so that extend and shading read them coalesced, not through the queue
    float4 *ray_o[2], *ray_d[2];
    // radiance throughput, termination throughput and previous base colour in queue order, by
    // bounce parity (shading reads bounce k's at its queue slot and writes the survivors' next
    // values at their slots in queue k + 1)
    float4 *t_rad[2], *t_term[2], *prev_base[2];
    float4 *acc_e, *acc_d, *acc_i, *acc_s;
    uint4* hit_rec;   // closest hit in queue order: {instance (as int, -1 = miss), triangle, u, v bits}
    float4* hit_pos;  // world position, world distance (queue order)
    uint32_t* queue[2];
    // ray binning: bin of every queue slot of the next bounce, the order in which extend takes
    // the slots, and the per-bin counts and cursors
    uint32_t *bin_key, *perm, *bin_hist, *bin_cur;
    float4 *sh_o, *sh_d, *sh_c;  // shadow rays: {o, max dist} {d, path} {contribution, slot}
    // reference traversal of scenes of at most 64 instances: every queue slot's mask of the
    // instances whose world box its ray enters, computed before the traversal (or null)
    uint64_t* inst_mask;
    float4* sample_rgb;
    uint2* spill;
    uint32_t* counters;
    unsigned long long* work;  // aabb, entered, pops, tris, hits, misses
    // differentiable rendering
    uint32_t* vis;      // bit k: the shadow ray of bounce k reached its light
    float4* adj_s;      // radiance collected so far during the replay
    float4* adj_prev;   // {u, v, material or ~0u} of the previous hit (sky term adjoint)
    // batched rendering
    uint32_t *path_job, *path_pix, *path_smp;  // owner job, pixel and sample of every path
    uint4* it_tiles;        // this iteration's tiles {job, first pixel, pixels, first path}
    uint32_t* it_pix_off;   // [tiles + 1] first iteration pixel of every tile
    const uint4* q_tiles;   // the scene's tile queue {job, first pixel, pixels, paths}
    uint32_t* q_state;      // {head, end of the tiles the next iteration takes} of that queue
    const struct BatchJob* jobs;
    uint32_t it_capacity;   // paths per iteration
    uint32_t it_tile_cap;   // tiles per iteration
    // radiance cache
    uint32_t* train_slot;   // per path: training slot or kNoSlot
    uint32_t* tr_path;      // per slot: its path
    float4 *tr_pos, *tr_dir, *tr_w, *tr_s;  // per slot and recorded bounce: vertex, direction
                                            // to the previous vertex, throughput (w = valid),
                                            // radiance collected before the vertex
    uint32_t* cq_path;      // cache queries: path, position, direction, prediction and the
    float4 *cq_pos, *cq_dir, *cq_out, *cq_t;  // throughput at the query vertex
    float4* tr_img;         // per path (after-light mode): throughput x prediction of a
                            // training path's own query, the cached estimate its image uses
    float4 *tp_pos, *tp_dir, *tp_tgt;       // compact training samples
    float4 *ts_pos, *ts_dir, *ts_tgt;       // the same in the order the Adam steps take them
    float4* acc_n;          // per path (cache_nee): radiance of the environment samples
    float* img_sum;         // per pixel: sum of the accumulated passes' samples
    // debug views (first-hit records of ProcessDebugRay)
    float4 *dbg_a, *dbg_b, *dbg_c, *dbg_d, *dbg_e;  // {base', metal} {rough, geom normal}
                                                     // {normal, depth} {Fresnel, brdf sum}
                                                     // {second ray direction, pdf}
    uint4 *dbg_i, *dbg_j;   // {primary bvh, primary tri, total bvh, total tri tests}
                            // {instance, triangle, leaf (all & 0xffff), bounces<<2 | dir<<1 | hit}
};

// one job of a batch, resident in device memory
struct alignas(16) BatchJob {
    ChunkParams p;  // the job's settings and camera (pix_begin/npix/paths unused)
    float* img;     // width * height * 3, device
};

// one stream with its own path state: chunks on different slots run concurrently, so that
// one chunk's last, sparsely populated bounces overlap another chunk's full ones
struct StreamSlot {
    PathBuffers b{};
    hipStream_t stream = nullptr;
    uint32_t* host_counters = nullptr;
    std::vector<hipEvent_t> events;
    int64_t chunk = -1;  // chunk in flight, -1 = idle
    ChunkParams c{};
    hipEvent_t ev_bounce = nullptr, ev_tail = nullptr;  // cached renders
};

class Renderer {
   public:
    ~Renderer();
    // max_slots: concurrent chunk streams (0 = MT_STREAMS or 3, at most the chunk count)
    int setup(mt_scene* scene, const mt_camera_desc& cam, const mt_settings& st, bool diff = false,
              uint32_t max_slots = 0);
    void set_camera(const mt_camera_desc& cam);
    void set_seed(uint64_t seed);
    int run_chunk(uint32_t chunk, float* sample_rgb_out);
    // pipelined chunks: enqueue on an idle slot, complete (wait + statistics) in FIFO order
    uint32_t slots() const { return static_cast<uint32_t>(slots_.size()); }
    bool slot_busy(uint32_t i) const { return slots_[i].chunk >= 0; }
    int enqueue_chunk(uint32_t slot, uint32_t chunk);
    int complete_chunk(uint32_t slot, float* sample_rgb_out);
    // differentiable rendering over the whole image on the caller's stream (0 = own stream):
    // forward writes the linear image (mean radiance per pixel) to img_dev and the per-path
    // store; adjoint replays the same paths and accumulates into the gradient buffers
    int diff_forward(const DiffParams& dp, const DiffParams* sampling, float* img_dev,
                     hipStream_t s);
    int diff_adjoint(const DiffParams& dp, const DiffGradPtrs& gr, const float* grad_img,
                     hipStream_t s);
    // render with the radiance cache: passes of spp_pass samples per pixel; rgb_out (host)
    int setup_cached(mt_scene* scene, const mt_camera_desc& cam, const mt_settings& st,
                     const mt_cache_options& opt);
    int render_cached(class Mlp& mlp, float* rgb_out);
    int setup_env_sampling();
    int finish(float* rgb_out);
    uint32_t chunks() const { return chunks_; }
    const mt_stats& stats() const { return stats_; }
    mt_stats& stats() { return stats_; }

   private:
    void release();
    // differentiable chunks: enqueue one on a slot's stream; wait for a slot's chunk and take
    // its counters; all chunks of one pass over the slots, after the caller's stream
    int run_diff_chunk(uint32_t chunk, int mode, float* img, StreamSlot& L);
    int finish_diff_chunk(StreamSlot& L, int mode);
    int run_diff(int mode, float* img, hipStream_t s);
    hipEvent_t diff_ev_ = nullptr;
    int run_cache_chunk(class Mlp& mlp, uint32_t slot, uint32_t chunk, uint32_t sample_base,
                        uint32_t train_every, uint32_t train_offset, bool accumulate);
    hipStream_t tail_stream_ = nullptr;
    unsigned long long* rays_dev_ = nullptr;
    mt_cache_options copt_{};
    uint32_t total_spp_ = 0;
    mt_scene* scene_ = nullptr;
    mt_camera_desc cam_{};
    mt_settings st_{};
    ChunkParams params_{};
    PathBuffers buf_{};
    std::vector<void*> owned_;
    std::vector<hipEvent_t> events_;
    hipStream_t stream_ = nullptr;
    bool premask_ = false;  // the path buffers carry instance masks (PathBuffers::inst_mask)
    uint32_t* host_counters_ = nullptr;
    float* img_ = nullptr;
    float4* store_ = nullptr;
    uint32_t capacity_ = 0, npix_chunk_ = 0, chunks_ = 0;
    uint32_t trace_blocks_ = 0, shade_blocks_ = 0, lds_depth_ = 16;
    uint32_t bin_blocks_ = 0;  // > 0: plain chunks bin the rays of bounces >= 1 (MT_BIN_RAYS)
    mt_stats stats_{};
    std::vector<StreamSlot> slots_;  // slots_[0] uses buf_, stream_, host_counters_, events_
};

// ---- one persistent launch for a whole batch (mt_batch_options.mode = MT_BATCH_PERSISTENT)
// The host publishes scene slots, job records and tiles in coherent pinned memory and bumps
// the published counts last; the kernel's lane groups copy what is new into device memory
// under the queue lock, take tiles in queue order and run whole wavefront iterations between
// group barriers.
struct alignas(16) PersistSlot {  // a scene and traversal: what an iteration's kernels need
    DevScene S;
    ChunkParams c;           // scene-level launch constants (per-job ones: the job table)
    uint32_t lights;         // point lights: lit shading and shadow rays
    uint32_t accel;          // 1 = fast traversal
    uint32_t pad[2];
};
struct PersistHost {         // pinned, written by the host (published counts, then closed)
    uint32_t tail;           // tiles published
    uint32_t jobs;           // job records published
    uint32_t slots;          // scene slots published
    uint32_t closed;         // 1: nothing more will be published; groups exit once drained
};
// device state words (PersistCtl::state)
constexpr uint32_t kPsHead = 0;    // next tile to take
constexpr uint32_t kPsLock = 1;    // queue lock (group leaders)
constexpr uint32_t kPsAbort = 2;   // set when a barrier waited past stall_ticks
constexpr uint32_t kPsTiles = 3;   // tiles copied to the device queue
constexpr uint32_t kPsJobs = 4;    // job records copied
constexpr uint32_t kPsSlots = 5;   // scene slots copied
constexpr uint32_t kPsWords = 16;
constexpr uint32_t kPsBarWords = 64;  // per lane group: {arrived, generation, selection}, own 256 B
struct PersistCtl {
    const PersistHost* host;
    const uint4* host_tiles;         // {job, first pixel, pixels, paths}
    const BatchJob* host_jobs;
    const uint32_t* host_job_slot;   // per job: its scene slot
    const uint32_t* host_job_tiles;  // per job: its tile count
    const PersistSlot* host_slots;
    uint32_t* job_done;              // pinned, written by the device: 1 once the job's image is final
    unsigned long long* host_iters;  // pinned, written by the device: iterations per lane group
    uint4* tiles;                    // device copies of the published records
    BatchJob* jobs;
    uint32_t* job_slot;
    uint32_t* tiles_left;            // per job: tiles not yet resolved
    PersistSlot* slots;
    uint32_t* state;                 // kPs* words
    uint32_t* bar;                   // kPsBarWords per lane group
    const PathBuffers* lanes;        // per lane group
    uint32_t groups, group_blocks;
    // xcds > 0: lane groups within XCDs (groups a multiple of xcds; the hardware deals blocks
    // round-robin over the XCDs, block b to XCD b % xcds), so that a group's barrier fences
    // touch only its own XCD's L2.  0: groups of consecutive blocks (spanning XCDs).
    uint32_t xcds;
    uint64_t idle_ticks;             // a group exits after this long without tiles (wall clock)
    uint64_t stall_ticks;            // a barrier wait this long aborts the launch
};
constexpr uint32_t kPersistDepth = 16;  // traversal stack entries in LDS in the persistent kernel
// blocks of the persistent kernel one CU holds (occupancy of its registers and LDS)
int persistent_blocks_per_cu(uint32_t* blocks);
// cooperative launch of groups * group_blocks blocks; fails instead of launching a grid whose
// blocks could not all be resident (the group barriers need every block running)
int launch_batch_persistent(const PersistCtl& ctl, hipStream_t s);
// wall-clock ticks per millisecond of the device's constant-rate counter
int wall_ticks_per_ms(uint64_t* ticks);
// XCDs (accelerator complex dies) of the current device, from the XCC_ID of a grid of blocks
int xcd_count(uint32_t* xcds);
void launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t s);

// launch shape of one batched iteration
struct BatchLaunch {
    uint32_t bounces;       // extend/shade/shadow rounds (the largest of the scene's jobs)
    uint32_t lds_depth;     // traversal stack entries kept in LDS
    uint32_t trace_blocks;  // persistent traversal blocks
    uint32_t shade_blocks;
};

uint32_t seed_mix(uint32_t x);
uint32_t seed_key(uint64_t seed);
uint32_t stack_capacity(const mt_scene* s);
// traversal stack entries kept in LDS for `blocks_per_cu` persistent traversal blocks of 256
// threads: the fast traversal keeps 16; the reference traversal the deepest of 16, 12, 8 and 4
// whose stacks (8 B per entry and thread) let that many blocks share a CU's 160 KB of LDS,
// since a shallower stack only adds spills once LDS no longer limits the blocks per CU
uint32_t lds_stack_depth(bool fast, uint32_t blocks_per_cu);
int cu_count();
// the per-render parameters of a render of `scene` with this camera and these settings
void fill_params(ChunkParams& c, const mt_scene* scene, const mt_camera_desc& cam,
                 const mt_settings& st);
// hands every pooled device block back to the device (mt_pool_trim)
void pool_trim();
int alloc_path_buffers(std::vector<void*>& owned, PathBuffers& b, uint64_t paths,
                       uint64_t spill_entries, bool diff, bool batch, bool debug = false);
// enqueue one batched iteration (schedule, generate, bounces, finalize, resolve): no host
// synchronisation and no memset, so that it can be captured into a graph
void launch_batch_iteration(const mt_scene* scene, const ChunkParams& c, const PathBuffers& b,
                            const BatchLaunch& launch, hipStream_t s);
void launch_set_u32(uint32_t* p, uint32_t v, hipStream_t s);
// the heat-map post-process of the counting and measuring debug views (host, in place)
void debug_post(int mode, float* rgb, uint64_t npix);
// test hooks: closest hits (occluded == null, 6 floats per ray) or occlusion up to a distance
// (7 floats per ray)
int trace_rays(mt_scene* s, uint32_t n, const float* rays, mt_hit* hits, int32_t traversal,
               uint32_t* occluded);
// closest hits of n device-resident rays into device-resident hits, with the traversal's
// scratch in a static device array (no allocation): mt_trace_closest_noalloc
int trace_rays_noalloc(mt_scene* s, uint32_t n, const float* rays_dev, mt_hit* hits_dev,
                       int32_t traversal, hipStream_t stream);
int brdf_eval(uint32_t n, const float* in, float* out);
int math_eval(int32_t op, uint32_t n, const float* in, float* out);

}  // namespace mt
