// mt_batch.cpp -- batched rendering, two ways: one persistent cooperative launch fed from
// pinned memory (MT_BATCH_PERSISTENT), or per-scene device tile queues with one captured graph
// per scene for a whole wavefront iteration (MT_BATCH_GRAPHS); a host scheduler thread; and
// the autotuner.
//
// Persistent: the scheduler thread appends job records and tiles to pinned coherent memory
// and publishes their counts; the kernel's lane groups copy them to the device and take tiles
// (k_batch_persistent).  The thread watches the per-job completion flags the kernel writes,
// starts a new launch when one ended (after 1 s with an empty queue) while work is pending,
// and closes the queue at mt_batch_finish, upon which the launch drains it and ends.
//
// Graphs: each scene has one queue and one captured graph per lane (a stream with its own path state);
// a job's tiles are dealt round-robin over the lanes, so iterations of one scene run
// concurrently and one lane's sparse last bounces overlap another's full ones.
// The scheduler keeps an exact host mirror of the device queues: the device scheduler kernel
// and the host apply the same greedy rule (tiles in queue order while their paths fit the
// iteration, at least one), so the host knows which iteration finishes each job without
// reading anything back, and only waits on events to bound the number of queued iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mt_internal.hpp"
#include "mt_render.hpp"

namespace {

constexpr uint32_t kMaxJobs = 1u << 14;
constexpr uint32_t kQueueTiles = 1u << 20;  // per scene (graphs), per batch (persistent)
constexpr uint32_t kMaxSlots = 256;         // scene slots of a persistent batch
constexpr uint64_t kStallMs = 60000;        // persistent: a group barrier waiting this long aborts
                                            // (MT_BATCH_STALL_MS overrides, at least 1)
constexpr uint64_t kIdleMs = 1000;          // persistent: a lane group exits after this long idle
constexpr uint32_t kMaxStack = 64;          // traversal stack entries (LDS + spill)
constexpr uint32_t kBlockSize = 256;
constexpr uint32_t kInflight = 3;           // iterations queued ahead of the device

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

constexpr uint32_t kMaxLanes = 8;         // graphs: streams
constexpr uint32_t kMaxGroups = 32;       // persistent: lane groups
constexpr uint64_t kMaxPersistentPaths = 1ull << 28;  // persistent: paths in flight over all groups

struct HostJob {
    mt_scene* scene = nullptr;
    uint32_t slot = 0;
    mt_camera_desc cam{};
    mt_settings st{};
    float* rgb_out = nullptr;
    float* img_dev = nullptr;
    uint32_t tiles = 0, tiles_left = 0;
    uint32_t lane_left[kMaxLanes] = {};  // unscheduled tiles per lane
    uint64_t samples = 0;
    double submit_ms = 0.0;
    std::vector<hipEvent_t> done_evs;         // one per lane that had tiles; complete before `recorded`
    std::atomic<bool> recorded{false};
    // persistent mode: the completion flag seen by the scheduler thread, and when
    std::atomic<bool> finished{false};
    double complete_ms = -1.0;  // written before `finished`
};

// coherent pinned host memory that the device reads (and partly writes) while it runs
template <typename T>
int pinned_alloc(T** p, size_t n) {
    void* q = nullptr;
    if (hipHostMalloc(&q, sizeof(T) * n, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        return mt::fail(MT_E_NOMEM, "pinned host allocation of the batch queue failed");
    std::memset(q, 0, sizeof(T) * n);
    *p = static_cast<T*>(q);
    return MT_OK;
}
template <typename T>
T* device_view(T* host) {
    void* d = nullptr;
    if (hipHostGetDevicePointer(&d, host, 0) != hipSuccess) return nullptr;
    return static_cast<T*>(d);
}

// one stream with its own path state; every scene has a queue and a graph per lane
struct Lane {
    hipStream_t stream = nullptr;
    mt::PathBuffers buf{};
    std::deque<hipEvent_t> inflight;
};

struct SceneQueue {
    uint4* q_dev = nullptr;        // device queue
    uint32_t* q_state = nullptr;   // device {head, tail}
    uint4* q_host = nullptr;       // pinned, append-only copy of the queue
    uint32_t head = 0, tail = 0;   // host mirror
    uint32_t graph_bounces = 0;    // bounce count of the captured graph
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    mt::PathBuffers b{};
};

struct SceneSlot {
    mt_scene* scene = nullptr;
    uint32_t bounces = 0;          // largest bounce count of the scene's jobs
    mt::ChunkParams c{};
    SceneQueue q[kMaxLanes];
};

}  // namespace

struct mt_batch {
    int device = 0;
    mt_batch_options opt{};
    hipStream_t stream = nullptr, capture = nullptr;  // uploads; graph capture
    // Device blocks are freed when the batch is destroyed.  The scheduler thread adds tile
    // queues while submitters add job images, so the vector is guarded by mu (init runs
    // before the thread exists).
    std::vector<void*> owned;
    std::vector<Lane> lanes;
    uint32_t next_lane = 0;
    mt::BatchJob* jobs_dev = nullptr;
    mt::BatchJob* jobs_host = nullptr;  // pinned staging, one record per job
    uint32_t trace_blocks = 0, shade_blocks = 0;
    uint32_t lds_ref = 16, lds_fast = 16;  // traversal stack entries in LDS per traversal
    std::chrono::steady_clock::time_point created;
    // shared with submitters and readers: jobs, slots, owned and error under mu; status,
    // iterations and captures are atomics (written by the scheduler thread, read anywhere)
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<HostJob>> jobs;
    std::vector<std::unique_ptr<SceneSlot>> slots;
    uint32_t admitted = 0;  // jobs [0, admitted) are on the device queues
    std::vector<HostJob*> view;  // scheduler's copy of the job pointers
    std::atomic<bool> finishing{false};
    std::atomic<bool> started{false};
    std::thread worker;
    std::atomic<int> status{MT_OK};
    std::string error;
    // scheduler state
    hipEvent_t e0 = nullptr;
    double t0_ms = 0.0;
    std::atomic<uint64_t> iterations{0}, captures{0};
    uint32_t first_open = 0;  // oldest job with tiles not yet scheduled
    std::vector<hipEvent_t> spare;
    // persistent mode
    bool persistent = false;
    mt::PersistCtl ctl{};
    mt::PersistHost* ph = nullptr;        // pinned records and counts (host addresses)
    uint4* p_tiles = nullptr;
    mt::BatchJob* p_jobs = nullptr;
    uint32_t *p_job_slot = nullptr, *p_job_tiles = nullptr, *p_job_done = nullptr;
    mt::PersistSlot* p_slots = nullptr;
    unsigned long long* p_iters = nullptr;
    std::vector<std::pair<mt_scene*, int32_t>> p_slot_keys;  // (scene, traversal) per slot
    uint32_t p_tail = 0;
    uint32_t p_first_undone = 0;
    hipEvent_t p_end = nullptr;           // recorded after the running launch
    bool p_running = false;
    std::atomic<uint64_t> launches{0};
    uint64_t stall_ms = kStallMs;         // MT_BATCH_STALL_MS

    ~mt_batch();
    int init(const mt_batch_options* o);
    int init_persistent();
    int admit(uint32_t id);
    int admit_persistent(uint32_t id);
    int ensure_graph(SceneSlot& s, uint32_t lane);
    int run();
    int run_persistent();
    // records the first error of the batch (the scheduler thread stops at it)
    int fail(int code, const std::string& m) {
        {
            std::scoped_lock lk(mu);
            if (status.load() == MT_OK) {
                error = m;
                status.store(code);
            }
        }
        mt::set_error(m);
        return code;
    }
    // the batch's error for a caller thread: code, and its message as that thread's last error
    int report() {
        std::scoped_lock lk(mu);
        const int code = status.load();
        if (code) mt::set_error(error);
        return code;
    }
    // a device block this batch frees at destruction
    void own(void* p) {
        std::scoped_lock lk(mu);
        owned.push_back(p);
    }
};

#define BT_TRY(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(MT_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
    } while (0)

mt_batch::~mt_batch() {
    if (worker.joinable()) {
        finishing.store(true);
        cv.notify_all();
        worker.join();
    }
    // a launch still running (the scheduler stopped at an error) drains what was published
    // and ends once the queue is closed
    if (ph) __atomic_store_n(&ph->closed, 1u, __ATOMIC_RELEASE);
    if (stream) (void)hipStreamSynchronize(stream);
    for (Lane& L : lanes)
        if (L.stream) (void)hipStreamSynchronize(L.stream);
    for (auto& s : slots)
        for (SceneQueue& q : s->q) {
            if (q.exec) (void)hipGraphExecDestroy(q.exec);
            if (q.graph) (void)hipGraphDestroy(q.graph);
            if (q.q_host) (void)hipHostFree(q.q_host);
        }
    for (auto& j : jobs)
        for (hipEvent_t e : j->done_evs) (void)hipEventDestroy(e);
    for (Lane& L : lanes) {
        for (hipEvent_t e : L.inflight) (void)hipEventDestroy(e);
        if (L.stream) (void)hipStreamDestroy(L.stream);
    }
    for (hipEvent_t e : spare) (void)hipEventDestroy(e);
    if (e0) (void)hipEventDestroy(e0);
    // the lanes' path buffers come from the renderer's block allocator, which keeps track of
    // them: free_blocks hands every block back and clears that bookkeeping
    mt::free_blocks(owned);
    if (jobs_host) (void)hipHostFree(jobs_host);
    for (void* h : {static_cast<void*>(ph), static_cast<void*>(p_tiles), static_cast<void*>(p_jobs),
                    static_cast<void*>(p_job_slot), static_cast<void*>(p_job_tiles),
                    static_cast<void*>(p_job_done), static_cast<void*>(p_slots), static_cast<void*>(p_iters)})
        if (h) (void)hipHostFree(h);
    if (p_end) (void)hipEventDestroy(p_end);
    if (stream) (void)hipStreamDestroy(stream);
    if (capture) (void)hipStreamDestroy(capture);
}

int mt_batch::init(const mt_batch_options* o) {
    created = std::chrono::steady_clock::now();
    if (o) opt = *o;
    if (!opt.paths_per_iteration) opt.paths_per_iteration = 1u << 22;
    if (!opt.tile_pixels) opt.tile_pixels = 1024;
    if (!opt.iteration_tiles) opt.iteration_tiles = 4096;
    if (!opt.trace_blocks_per_cu) opt.trace_blocks_per_cu = 8;
    if (!opt.mode) opt.mode = MT_BATCH_PERSISTENT;
    persistent = opt.mode == MT_BATCH_PERSISTENT;
    BT_TRY(hipGetDevice(&device));
    uint32_t xcds = 1;
    if (persistent && mt::xcd_count(&xcds)) return fail(MT_E_HIP, mt_last_error());
    if (!opt.streams) opt.streams = persistent ? xcds : 3;  // persistent: one lane group per XCD
    if (opt.paths_per_iteration > (1u << 26) || opt.iteration_tiles > (1u << 20) ||
        opt.trace_blocks_per_cu > 32 || opt.streams > (persistent ? kMaxGroups : kMaxLanes) ||
        (persistent && uint64_t(opt.streams) * opt.paths_per_iteration > kMaxPersistentPaths) ||
        (opt.mode != MT_BATCH_PERSISTENT && opt.mode != MT_BATCH_GRAPHS))
        return fail(MT_E_UNSUPPORTED, "batch options out of range");
    BT_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    BT_TRY(hipStreamCreateWithFlags(&capture, hipStreamNonBlocking));
    const int cus = mt::cu_count();
    trace_blocks = static_cast<uint32_t>(cus) * opt.trace_blocks_per_cu;
    const uint64_t P = opt.paths_per_iteration;
    shade_blocks = static_cast<uint32_t>(std::min<uint64_t>((P + kBlockSize - 1) / kBlockSize, cus * 16ull));
    lds_ref = mt::lds_stack_depth(false, opt.trace_blocks_per_cu);
    lds_fast = mt::lds_stack_depth(true, opt.trace_blocks_per_cu);
    uint64_t spill = static_cast<uint64_t>(kMaxStack - std::min(lds_ref, lds_fast)) * trace_blocks * kBlockSize;
    if (persistent) {
        // the launch holds as many blocks as the CUs do, split evenly into the lane groups
        uint32_t per_cu = 0;
        if (mt::persistent_blocks_per_cu(&per_cu)) return fail(MT_E_HIP, mt_last_error());
        const uint32_t blocks = static_cast<uint32_t>(cus) * per_cu;
        ctl.groups = opt.streams;
        // lane groups inside XCDs when the counts divide evenly, else groups of consecutive blocks
        if (xcds > 1 && opt.streams % xcds == 0 && blocks % xcds == 0 &&
            (blocks / xcds) % (opt.streams / xcds) == 0) {
            ctl.xcds = xcds;
            ctl.group_blocks = blocks / opt.streams;
        } else {
            ctl.xcds = 0;
            ctl.group_blocks = blocks / opt.streams;
        }
        if (ctl.group_blocks == 0) return fail(MT_E_UNSUPPORTED, "more lanes than resident blocks");
        spill = static_cast<uint64_t>(kMaxStack - mt::kPersistDepth) * ctl.group_blocks * kBlockSize;
    }
    auto dev_alloc = [&](void** p, size_t bytes) -> int {
        if (mt::device_malloc(p, bytes)) return fail(MT_E_NOMEM, "device allocation of batch buffers failed");
        own(*p);
        return MT_OK;
    };
    int rc;
    if ((rc = dev_alloc(reinterpret_cast<void**>(&jobs_dev), sizeof(mt::BatchJob) * kMaxJobs))) return rc;
    lanes.resize(opt.streams);
    for (Lane& L : lanes) {
        if (!persistent) BT_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
        mt::PathBuffers& b = L.buf;
        if ((rc = mt::alloc_path_buffers(owned, b, P, spill, false, true))) return fail(rc, mt_last_error());
        if ((rc = dev_alloc(reinterpret_cast<void**>(&b.it_tiles), sizeof(uint4) * opt.iteration_tiles)) ||
            (rc = dev_alloc(reinterpret_cast<void**>(&b.it_pix_off), sizeof(uint32_t) * (opt.iteration_tiles + 1))))
            return rc;
        b.jobs = jobs_dev;
        b.it_capacity = opt.paths_per_iteration;
        b.it_tile_cap = opt.iteration_tiles;
    }
    BT_TRY(hipHostMalloc(reinterpret_cast<void**>(&jobs_host), sizeof(mt::BatchJob) * kMaxJobs));
    BT_TRY(hipEventCreate(&e0));
    return persistent ? init_persistent() : MT_OK;
}

int mt_batch::init_persistent() {
    int rc;
    auto dev_alloc = [&](void** p, size_t bytes) -> int {
        if (mt::device_malloc(p, bytes)) return fail(MT_E_NOMEM, "device allocation of batch buffers failed");
        own(*p);
        return MT_OK;
    };
    if ((rc = pinned_alloc(&ph, 1)) || (rc = pinned_alloc(&p_tiles, kQueueTiles)) ||
        (rc = pinned_alloc(&p_jobs, kMaxJobs)) || (rc = pinned_alloc(&p_job_slot, kMaxJobs)) ||
        (rc = pinned_alloc(&p_job_tiles, kMaxJobs)) || (rc = pinned_alloc(&p_job_done, kMaxJobs)) ||
        (rc = pinned_alloc(&p_slots, kMaxSlots)) || (rc = pinned_alloc(&p_iters, kMaxGroups)))
        return fail(rc, mt_last_error());
    ctl.host = device_view(ph);
    ctl.host_tiles = device_view(p_tiles);
    ctl.host_jobs = device_view(p_jobs);
    ctl.host_job_slot = device_view(p_job_slot);
    ctl.host_job_tiles = device_view(p_job_tiles);
    ctl.host_slots = device_view(p_slots);
    ctl.job_done = device_view(p_job_done);
    ctl.host_iters = device_view(p_iters);
    if (!ctl.host || !ctl.host_tiles || !ctl.host_jobs || !ctl.host_job_slot || !ctl.host_job_tiles ||
        !ctl.host_slots || !ctl.job_done || !ctl.host_iters)
        return fail(MT_E_HIP, "device addresses of the pinned batch queue failed");
    void* p = nullptr;
    if ((rc = dev_alloc(&p, sizeof(uint4) * kQueueTiles))) return rc;
    ctl.tiles = static_cast<uint4*>(p);
    ctl.jobs = jobs_dev;
    if ((rc = dev_alloc(&p, sizeof(uint32_t) * kMaxJobs))) return rc;
    ctl.job_slot = static_cast<uint32_t*>(p);
    if ((rc = dev_alloc(&p, sizeof(uint32_t) * kMaxJobs))) return rc;
    ctl.tiles_left = static_cast<uint32_t*>(p);
    if ((rc = dev_alloc(&p, sizeof(mt::PersistSlot) * kMaxSlots))) return rc;
    ctl.slots = static_cast<mt::PersistSlot*>(p);
    if ((rc = dev_alloc(&p, sizeof(uint32_t) * mt::kPsWords))) return rc;
    ctl.state = static_cast<uint32_t*>(p);
    if ((rc = dev_alloc(&p, sizeof(uint32_t) * mt::kPsBarWords * kMaxGroups))) return rc;
    ctl.bar = static_cast<uint32_t*>(p);
    if ((rc = dev_alloc(&p, sizeof(mt::PathBuffers) * lanes.size()))) return rc;
    ctl.lanes = static_cast<const mt::PathBuffers*>(p);
    std::vector<mt::PathBuffers> lb;
    for (const Lane& L : lanes) lb.push_back(L.buf);
    BT_TRY(hipMemcpyAsync(p, lb.data(), sizeof(mt::PathBuffers) * lb.size(), hipMemcpyHostToDevice, stream));
    mt::launch_zero_u32(ctl.state, mt::kPsWords, stream);
    mt::launch_zero_u32(ctl.bar, mt::kPsBarWords * kMaxGroups, stream);
    BT_TRY(hipGetLastError());
    BT_TRY(hipStreamSynchronize(stream));
    uint64_t per_ms = 0;
    if (mt::wall_ticks_per_ms(&per_ms)) return fail(MT_E_HIP, mt_last_error());
    ctl.idle_ticks = kIdleMs * per_ms;
    if (const char* v = std::getenv("MT_BATCH_STALL_MS"))
        stall_ms = std::max<uint64_t>(1, std::strtoull(v, nullptr, 10));
    ctl.stall_ticks = stall_ms * per_ms;
    BT_TRY(hipEventCreateWithFlags(&p_end, hipEventDisableTiming));
    return MT_OK;
}

// writes job `id`'s scene slot (new scenes only), record and tiles into the pinned queue; the
// counts are published by run_persistent after each round of admissions
int mt_batch::admit_persistent(uint32_t id) {
    HostJob& j = *view[id];
    const int32_t accel = j.st.traversal == MT_TRAVERSAL_FAST ? 1 : 0;
    uint32_t slot = static_cast<uint32_t>(p_slot_keys.size());
    for (uint32_t i = 0; i < p_slot_keys.size(); ++i)
        if (p_slot_keys[i].first == j.scene && p_slot_keys[i].second == accel) slot = i;
    if (slot == p_slot_keys.size()) {
        if (slot >= kMaxSlots) return fail(MT_E_UNSUPPORTED, "too many scenes in one batch");
        if (mt::stack_capacity(j.scene) > kMaxStack)
            return fail(MT_E_UNSUPPORTED, "scene BVH too deep for batched rendering");
        mt::PersistSlot& ps = p_slots[slot];
        ps.S = j.scene->dev;
        mt::fill_params(ps.c, j.scene, j.cam, j.st);
        ps.c.count_work = 0;
        ps.lights = j.scene->light_count ? 1u : 0u;
        ps.accel = static_cast<uint32_t>(accel);
        std::scoped_lock lk(mu);
        p_slot_keys.emplace_back(j.scene, accel);
    }
    j.slot = slot;
    mt::BatchJob& rec = p_jobs[id];
    mt::fill_params(rec.p, j.scene, j.cam, j.st);
    rec.p.count_work = 0;
    rec.img = j.img_dev;
    p_job_slot[id] = slot;
    const uint32_t npix = j.st.width * j.st.height;
    const uint32_t tp = std::max<uint32_t>(1u, std::min<uint32_t>(opt.tile_pixels, opt.paths_per_iteration / j.st.samples));
    uint32_t count = 0;
    for (uint32_t p0 = 0; p0 < npix; p0 += tp, ++count) {
        if (p_tail >= kQueueTiles) return fail(MT_E_UNSUPPORTED, "tile queue of the batch is full");
        const uint32_t n = std::min(tp, npix - p0);
        p_tiles[p_tail++] = make_uint4(id, p0, n, n * j.st.samples);
    }
    p_job_tiles[id] = count;
    j.tiles_left = count;
    std::scoped_lock lk(mu);
    j.tiles = count;
    return MT_OK;
}

int mt_batch::run_persistent() {
    if (hipSetDevice(device) != hipSuccess) return fail(MT_E_HIP, "hipSetDevice failed");
    bool closed = false;
    while (true) {
        {
            std::scoped_lock lk(mu);
            for (size_t i = view.size(); i < jobs.size(); ++i) view.push_back(jobs[i].get());
        }
        const uint32_t total = static_cast<uint32_t>(view.size());
        if (admitted < total) {
            while (admitted < total) {
                const int rc = admit_persistent(admitted);
                if (rc) return rc;
                ++admitted;
            }
            // records before counts, and the tail last: a group that sees a tile has its
            // job's record and scene slot
            __atomic_store_n(&ph->slots, static_cast<uint32_t>(p_slot_keys.size()), __ATOMIC_RELEASE);
            __atomic_store_n(&ph->jobs, admitted, __ATOMIC_RELEASE);
            __atomic_store_n(&ph->tail, p_tail, __ATOMIC_RELEASE);
        }
        // whether the launch has ended first, then the completion flags: a launch that ended
        // wrote every flag of the tiles it took before it ended
        if (p_running) {
            const hipError_t q = hipEventQuery(p_end);
            if (q == hipSuccess) {
                p_running = false;
                uint32_t abort_flag = 0;
                BT_TRY(hipMemcpy(&abort_flag, ctl.state + mt::kPsAbort, sizeof(uint32_t), hipMemcpyDeviceToHost));
                if (abort_flag)
                    return fail(MT_E_HIP, "persistent batch kernel aborted: a lane group waited " +
                                              std::to_string(stall_ms) + " ms at a barrier");
            } else if (q != hipErrorNotReady) {
                return fail(MT_E_HIP, std::string("persistent batch kernel failed: ") + hipGetErrorString(q));
            }
        }
        for (uint32_t i = p_first_undone; i < admitted; ++i) {
            HostJob& j = *view[i];
            if (!j.finished.load() && __atomic_load_n(p_job_done + i, __ATOMIC_ACQUIRE)) {
                j.complete_ms = now_ms(created);
                j.finished.store(true, std::memory_order_release);
            }
        }
        while (p_first_undone < admitted && view[p_first_undone]->finished.load()) ++p_first_undone;
        const bool pending = p_first_undone < admitted;
        if (!p_running && pending) {
            // a new launch: the queue state (head, copied counts) carries over; the lock, the
            // abort word and the group barriers start at zero
            mt::launch_zero_u32(ctl.state + mt::kPsLock, 2, stream);
            mt::launch_zero_u32(ctl.bar, mt::kPsBarWords * kMaxGroups, stream);
            BT_TRY(hipGetLastError());
            if (const int rc = mt::launch_batch_persistent(ctl, stream)) return fail(rc, mt_last_error());
            BT_TRY(hipEventRecord(p_end, stream));
            p_running = true;
            ++launches;
        }
        std::unique_lock lk(mu);
        if (finishing.load() && jobs.size() == admitted) {
            if (!closed) {
                __atomic_store_n(&ph->closed, 1u, __ATOMIC_RELEASE);
                closed = true;
            }
            if (!pending && !p_running) break;
        }
        cv.wait_for(lk, std::chrono::microseconds(100));
    }
    uint64_t it = 0;
    for (uint32_t g = 0; g < ctl.groups; ++g) it += __atomic_load_n(p_iters + g, __ATOMIC_ACQUIRE);
    iterations.store(it);
    BT_TRY(hipStreamSynchronize(stream));
    return MT_OK;
}

// puts job `id` on its scene's device queues, its tiles dealt round-robin over the lanes
// (scheduler thread; uploads on the upload stream, which every lane waits for)
int mt_batch::admit(uint32_t id) {
    HostJob& j = *view[id];
    const uint32_t nl = static_cast<uint32_t>(lanes.size());
    SceneSlot* s = nullptr;
    // one slot per scene and traversal: the slot's launch constants pick the traversal kernel
    const int32_t accel = j.st.traversal == MT_TRAVERSAL_FAST ? 1 : 0;
    for (auto& x : slots)  // only this thread changes slots: reading without the lock is safe
        if (x->scene == j.scene && x->c.accel == accel) s = x.get();
    if (!s) {
        auto ns = std::make_unique<SceneSlot>();
        ns->scene = j.scene;
        if (mt::stack_capacity(j.scene) > kMaxStack)
            return fail(MT_E_UNSUPPORTED, "scene BVH too deep for batched rendering");
        // scene-level launch constants (the per-job ones come from the job table)
        mt::fill_params(ns->c, j.scene, j.cam, j.st);
        ns->c.count_work = 0;
        for (uint32_t l = 0; l < nl; ++l) {
            SceneQueue& q = ns->q[l];
            void* qd = nullptr;
            void* qs = nullptr;
            if (mt::device_malloc(&qd, sizeof(uint4) * kQueueTiles)) return fail(MT_E_NOMEM, "device allocation of a tile queue failed");
            own(qd);
            if (mt::device_malloc(&qs, sizeof(uint32_t) * 2)) return fail(MT_E_NOMEM, "device allocation of a tile queue failed");
            own(qs);
            q.q_dev = static_cast<uint4*>(qd);
            q.q_state = static_cast<uint32_t*>(qs);
            BT_TRY(hipHostMalloc(reinterpret_cast<void**>(&q.q_host), sizeof(uint4) * kQueueTiles));
            mt::launch_set_u32(q.q_state, 0u, stream);
            mt::launch_set_u32(q.q_state + 1, 0u, stream);
            q.b = lanes[l].buf;
            q.b.q_tiles = q.q_dev;
            q.b.q_state = q.q_state;
        }
        {
            std::scoped_lock lk(mu);
            slots.push_back(std::move(ns));
            s = slots.back().get();
        }
    }
    j.slot = 0;
    for (uint32_t i = 0; i < slots.size(); ++i)
        if (slots[i].get() == s) j.slot = i;
    // job record
    mt::BatchJob& rec = jobs_host[id];
    mt::fill_params(rec.p, j.scene, j.cam, j.st);
    rec.p.count_work = 0;
    rec.img = j.img_dev;
    BT_TRY(hipMemcpyAsync(jobs_dev + id, &rec, sizeof(rec), hipMemcpyHostToDevice, stream));
    // tiles: whole pixels, at most paths_per_iteration paths each
    const uint32_t npix = j.st.width * j.st.height;
    const uint32_t tp = std::max<uint32_t>(1u, std::min<uint32_t>(opt.tile_pixels, opt.paths_per_iteration / j.st.samples));
    uint32_t first[kMaxLanes];
    for (uint32_t l = 0; l < nl; ++l) first[l] = s->q[l].tail;
    uint32_t count = 0;
    for (uint32_t p0 = 0; p0 < npix; p0 += tp, ++count) {
        const uint32_t l = (count + id) % nl;
        SceneQueue& q = s->q[l];
        if (q.tail >= kQueueTiles) return fail(MT_E_UNSUPPORTED, "tile queue of a scene is full");
        const uint32_t n = std::min(tp, npix - p0);
        q.q_host[q.tail++] = make_uint4(id, p0, n, n * j.st.samples);
        ++j.lane_left[l];
    }
    j.tiles_left = count;
    {
        std::scoped_lock lk(mu);
        j.tiles = count;
    }
    // only the tiles go to the device here: the end of the tiles an iteration may take is
    // written on the lane's own stream right before that iteration (run()), so an iteration
    // launched before this admission never takes this job's tiles, even when it runs on the
    // device after the upload has landed
    for (uint32_t l = 0; l < nl; ++l) {
        SceneQueue& q = s->q[l];
        if (q.tail == first[l]) continue;
        BT_TRY(hipMemcpyAsync(q.q_dev + first[l], q.q_host + first[l], sizeof(uint4) * (q.tail - first[l]),
                              hipMemcpyHostToDevice, stream));
    }
    BT_TRY(hipGetLastError());
    // every lane sees the uploads before its next iteration
    hipEvent_t up;
    if (spare.empty()) {
        BT_TRY(hipEventCreateWithFlags(&up, hipEventDisableTiming));
    } else {
        up = spare.back();
        spare.pop_back();
    }
    BT_TRY(hipEventRecord(up, stream));
    for (Lane& L : lanes) BT_TRY(hipStreamWaitEvent(L.stream, up, 0));
    spare.push_back(up);  // reusable: a wait refers to the record at the time of the call
    s->bounces = std::max(s->bounces, j.st.bounces);
    return MT_OK;
}

int mt_batch::ensure_graph(SceneSlot& s, uint32_t lane) {
    SceneQueue& q = s.q[lane];
    if (q.exec && q.graph_bounces >= s.bounces) return MT_OK;
    if (q.exec) {
        BT_TRY(hipStreamSynchronize(lanes[lane].stream));
        BT_TRY(hipGraphExecDestroy(q.exec));
        BT_TRY(hipGraphDestroy(q.graph));
        q.exec = nullptr;
        q.graph = nullptr;
    }
    mt::BatchLaunch L{s.bounces, s.c.accel ? lds_fast : lds_ref, trace_blocks, shade_blocks};
    BT_TRY(hipStreamBeginCapture(capture, hipStreamCaptureModeRelaxed));
    mt::launch_batch_iteration(s.scene, s.c, q.b, L, capture);
    const hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(capture, &g);
    if (le != hipSuccess || ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(MT_E_HIP, std::string("graph capture failed: ") +
                                  hipGetErrorString(le != hipSuccess ? le : ce));
    }
    q.graph = g;
    BT_TRY(hipGraphInstantiate(&q.exec, q.graph, nullptr, nullptr, 0));
    q.graph_bounces = s.bounces;
    ++captures;
    return MT_OK;
}

int mt_batch::run() {
    if (hipSetDevice(device) != hipSuccess) return fail(MT_E_HIP, "hipSetDevice failed");
    const uint32_t nl = static_cast<uint32_t>(lanes.size());
    while (true) {
        // admit everything submitted so far
        {
            std::scoped_lock lk(mu);
            for (size_t i = view.size(); i < jobs.size(); ++i) view.push_back(jobs[i].get());
        }
        const uint32_t total = static_cast<uint32_t>(view.size());
        while (admitted < total) {
            const int rc = admit(admitted);
            if (rc) return rc;
            ++admitted;
        }
        // the oldest job that still has unscheduled tiles decides the scene; its lanes with
        // pending tiles take turns
        while (first_open < admitted && view[first_open]->tiles_left == 0) ++first_open;
        if (first_open == admitted) {
            std::unique_lock lk(mu);
            if (finishing.load() && jobs.size() == admitted) break;
            cv.wait_for(lk, std::chrono::milliseconds(1));
            continue;
        }
        SceneSlot& s = *slots[view[first_open]->slot];
        uint32_t lane = nl;
        for (uint32_t k = 0; k < nl; ++k) {
            const uint32_t l = (next_lane + k) % nl;
            if (s.q[l].head < s.q[l].tail) {
                lane = l;
                break;
            }
        }
        next_lane = (lane + 1) % nl;
        SceneQueue& q = s.q[lane];
        Lane& L = lanes[lane];
        int rc = ensure_graph(s, lane);
        if (rc) return rc;
        // the tiles this iteration takes: k_batch_schedule's greedy rule (tiles in queue order
        // while their paths fit, at least one) over the queue as admitted so far; their end
        // goes to the device on the lane's stream ahead of the iteration, so the device takes
        // exactly these tiles
        uint32_t n = 0, end = q.head;
        uint64_t paths = 0;
        while (end < q.tail && n < opt.iteration_tiles) {
            const uint4 t = q.q_host[end];
            if (n > 0 && paths + t.w > opt.paths_per_iteration) break;
            paths += t.w;
            ++n;
            ++end;
        }
        mt::launch_set_u32(q.q_state + 1, end, L.stream);
        BT_TRY(hipGetLastError());
        BT_TRY(hipGraphLaunch(q.exec, L.stream));
        ++iterations;
        for (; q.head < end; ++q.head) {
            const uint4 t = q.q_host[q.head];
            HostJob& j = *view[t.x];
            --j.tiles_left;
            if (--j.lane_left[lane] == 0) {
                hipEvent_t ev;
                BT_TRY(hipEventCreate(&ev));
                BT_TRY(hipEventRecord(ev, L.stream));
                j.done_evs.push_back(ev);
                if (j.tiles_left == 0) j.recorded.store(true, std::memory_order_release);
            }
        }
        // keep a few iterations queued per lane, so that new jobs are admitted promptly
        hipEvent_t ev;
        if (spare.empty()) {
            BT_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        } else {
            ev = spare.back();
            spare.pop_back();
        }
        BT_TRY(hipEventRecord(ev, L.stream));
        L.inflight.push_back(ev);
        while (L.inflight.size() > kInflight) {
            BT_TRY(hipEventSynchronize(L.inflight.front()));
            spare.push_back(L.inflight.front());
            L.inflight.pop_front();
        }
    }
    for (Lane& L : lanes) BT_TRY(hipStreamSynchronize(L.stream));
    BT_TRY(hipStreamSynchronize(stream));
    return MT_OK;
}

namespace {

bool valid_job(mt_scene* scene, const mt_camera_desc* cam, const mt_settings* st, uint32_t cap) {
    if (!scene || !cam || !st) {
        mt::set_error("null scene, camera or settings");
        return false;
    }
    if (st->width == 0 || st->height == 0 || st->samples == 0) {
        mt::set_error("width, height and samples must be positive");
        return false;
    }
    if (st->debug_mode) {
        mt::set_error("debug views are rendered by mt_render / mt_render_start");
        return false;
    }
    if (st->traversal != MT_TRAVERSAL_REFERENCE && st->traversal != MT_TRAVERSAL_FAST) {
        mt::set_error("unknown traversal");
        return false;
    }
    if (st->traversal == MT_TRAVERSAL_FAST && !scene->dev.fast_meshes && scene->mesh_count) {
        mt::set_error("the scene was created without fast BVHs (mt_bvh_options.fast)");
        return false;
    }
    if (st->bounces > mt::kMaxBounces || st->samples > cap ||
        static_cast<uint64_t>(st->width) * st->height * 3 > 0xFFFFFFFFull) {
        mt::set_error("job exceeds the batch limits (bounces <= 64, samples <= paths per iteration)");
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

int mt_batch_create(const mt_batch_options* options, mt_batch** out) {
    if (!out) return mt::fail(MT_E_INVALID, "null output handle");
    *out = nullptr;
    try {
        auto b = std::make_unique<mt_batch>();
        const int rc = b->init(options);
        if (rc) return rc;
        *out = b.release();
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_NOMEM, e.what());
    }
}

int mt_batch_submit(mt_batch* batch, mt_scene* scene, const mt_camera_desc* camera,
                    const mt_settings* settings, float* rgb_out, uint32_t* job_id) {
    if (!batch) return mt::fail(MT_E_INVALID, "null batch");
    if (!valid_job(scene, camera, settings, batch->opt.paths_per_iteration)) return MT_E_INVALID;
    if (batch->finishing.load()) return mt::fail(MT_E_INVALID, "batch is finishing");
    if (const int rc = batch->report()) return rc;
    try {
        auto j = std::make_unique<HostJob>();
        j->scene = scene;
        j->cam = *camera;
        j->st = *settings;
        j->rgb_out = rgb_out;
        j->samples = static_cast<uint64_t>(settings->width) * settings->height * settings->samples;
        j->submit_ms = now_ms(batch->created);
        if (hipSetDevice(batch->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        const size_t bytes = sizeof(float) * 3ull * settings->width * settings->height;
        void* img = nullptr;
        if (mt::device_malloc(&img, bytes)) return MT_E_NOMEM;
        j->img_dev = static_cast<float*>(img);
        std::scoped_lock lk(batch->mu);
        // checked again under the lock: the scheduler exits once it sees finishing with every
        // job admitted, so a job pushed after that would never run
        if (batch->jobs.size() >= kMaxJobs || batch->finishing.load()) {
            (void)hipFree(j->img_dev);
            return batch->finishing.load() ? mt::fail(MT_E_INVALID, "batch is finishing")
                                           : mt::fail(MT_E_UNSUPPORTED, "too many jobs in one batch");
        }
        batch->owned.push_back(j->img_dev);
        if (job_id) *job_id = static_cast<uint32_t>(batch->jobs.size());
        batch->jobs.push_back(std::move(j));
    } catch (const std::exception& e) {
        return mt::fail(MT_E_NOMEM, e.what());
    }
    batch->cv.notify_all();
    return MT_OK;
}

int mt_batch_start(mt_batch* batch) {
    if (!batch) return mt::fail(MT_E_INVALID, "null batch");
    if (batch->started.exchange(true)) return mt::fail(MT_E_INVALID, "batch already started");
    if (hipSetDevice(batch->device) != hipSuccess ||
        hipEventRecord(batch->e0, batch->stream) != hipSuccess ||
        hipEventSynchronize(batch->e0) != hipSuccess)
        return mt::fail(MT_E_HIP, "batch start event failed");
    batch->t0_ms = now_ms(batch->created);
    // the scheduler thread: an error (or a C++ exception, e.g. std::bad_alloc from a vector
    // it grows) stops it and becomes the batch's status, reported to the caller by
    // mt_batch_finish; nothing may escape the thread, or the process would terminate
    batch->worker = std::thread([batch]() {
        int rc;
        try {
            rc = batch->persistent ? batch->run_persistent() : batch->run();
            if (rc) batch->fail(rc, mt_last_error());
        } catch (const std::bad_alloc&) {
            batch->fail(MT_E_NOMEM, "host allocation failed in the batch scheduler");
        } catch (const std::exception& e) {
            batch->fail(MT_E_INVALID, e.what());
        }
    });
    return MT_OK;
}

int mt_batch_finish(mt_batch* batch) {
    if (!batch) return mt::fail(MT_E_INVALID, "null batch");
    if (!batch->started.load()) {
        const int rc = mt_batch_start(batch);
        if (rc) return rc;
    }
    batch->finishing.store(true);
    batch->cv.notify_all();
    if (batch->worker.joinable()) batch->worker.join();
    if (const int rc = batch->report()) return rc;
    if (hipSetDevice(batch->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
    for (auto& j : batch->jobs) {
        if (!j->rgb_out) continue;
        const size_t bytes = sizeof(float) * 3ull * j->st.width * j->st.height;
        if (hipMemcpy(j->rgb_out, j->img_dev, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            return mt::fail(MT_E_HIP, "copy of a job image failed");
    }
    return MT_OK;
}

int mt_batch_get_job_info(const mt_batch* batch, uint32_t job_id, mt_batch_job_info* info) {
    if (!batch || !info) return mt::fail(MT_E_INVALID, "null batch or info");
    auto* b = const_cast<mt_batch*>(batch);
    std::scoped_lock lk(b->mu);
    if (job_id >= b->jobs.size()) return mt::fail(MT_E_INVALID, "unknown job id");
    const HostJob& j = *b->jobs[job_id];
    std::memset(info, 0, sizeof(*info));
    info->status = b->status.load();
    info->tiles = j.tiles;
    info->samples = j.samples;
    info->submit_ms = j.submit_ms;
    info->complete_ms = -1.0;
    if (b->persistent) {
        if (j.finished.load(std::memory_order_acquire)) {
            info->done = 1;
            info->complete_ms = j.complete_ms;
        }
    } else if (j.recorded.load(std::memory_order_acquire)) {
        bool all = true;
        float latest = 0.0f;
        for (hipEvent_t e : j.done_evs) {
            float ms = 0.0f;
            if (hipEventQuery(e) != hipSuccess || hipEventElapsedTime(&ms, b->e0, e) != hipSuccess) {
                all = false;
                break;
            }
            latest = std::max(latest, ms);
        }
        if (all) {
            info->done = 1;
            info->complete_ms = b->t0_ms + latest;
        }
    }
    return MT_OK;
}

int mt_batch_get_stats(const mt_batch* batch, mt_batch_stats* stats) {
    if (!batch || !stats) return mt::fail(MT_E_INVALID, "null batch or stats");
    auto* b = const_cast<mt_batch*>(batch);
    std::scoped_lock lk(b->mu);
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = b->iterations.load();
    if (b->persistent && b->p_iters) {
        uint64_t it = 0;  // the device's running counts
        for (uint32_t g = 0; g < b->ctl.groups; ++g) it += __atomic_load_n(b->p_iters + g, __ATOMIC_ACQUIRE);
        stats->iterations = it;
    }
    stats->graphs_captured = b->captures.load();
    stats->launches = b->launches.load();
    stats->jobs = static_cast<uint32_t>(b->jobs.size());
    stats->scenes = static_cast<uint32_t>(b->persistent ? b->p_slot_keys.size() : b->slots.size());
    for (auto& j : b->jobs) stats->samples += j->samples;
    stats->options = b->opt;
    return MT_OK;
}

int mt_batch_destroy(mt_batch* batch) {
    delete batch;
    return MT_OK;
}

int mt_batch_autotune(mt_scene* const* scenes, const mt_camera_desc* cameras,
                      const mt_settings* settings, uint32_t n, mt_batch_options* best,
                      mt_autotune_trial* trials, uint32_t max_trials, uint32_t* n_trials) {
    if (!scenes || !cameras || !settings || !best || n == 0)
        return mt::fail(MT_E_INVALID, "autotune needs calibration jobs and an output");
    if (n_trials) *n_trials = 0;
    uint32_t done = 0;
    auto measure = [&](const mt_batch_options& o, double& msps) -> int {
        mt_batch* b = nullptr;
        int rc = mt_batch_create(&o, &b);
        if (rc) return rc;
        std::unique_ptr<mt_batch> guard(b);
        // untimed round: every calibration job at 1 sample per pixel, so that each scene's
        // graphs are captured and the buffers touched before the timed round
        for (uint32_t i = 0; i < n; ++i) {
            if (settings[i].samples > o.paths_per_iteration) return mt::fail(MT_E_INVALID, "calibration job too large");
            mt_settings st = settings[i];
            st.samples = 1;
            if ((rc = mt_batch_submit(b, scenes[i], &cameras[i], &st, nullptr, nullptr))) return rc;
        }
        if ((rc = mt_batch_start(b))) return rc;
        const auto tw = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < n; ++i) {
            mt_batch_job_info info{};
            for (;;) {
                if ((rc = mt_batch_get_job_info(b, i, &info))) return rc;
                if (info.status) return mt_batch_finish(b);  // reports the batch's error
                if (info.done) break;
                if (now_ms(tw) > 600e3) return mt::fail(MT_E_HIP, "autotune warm-up round did not complete");
                std::this_thread::sleep_for(std::chrono::microseconds(200));
            }
        }
        // timed round: the calibration jobs as given, submitted at once to the running batch
        uint64_t samples = 0;
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < n; ++i) {
            if ((rc = mt_batch_submit(b, scenes[i], &cameras[i], &settings[i], nullptr, nullptr))) return rc;
            samples += static_cast<uint64_t>(settings[i].width) * settings[i].height * settings[i].samples;
        }
        if ((rc = mt_batch_finish(b))) return rc;
        const double s = now_ms(t0) / 1e3;
        msps = samples / s / 1e6;
        if (trials && done < max_trials) trials[done] = {o, s, msps};
        ++done;
        if (n_trials) *n_trials = std::min(done, max_trials);
        return MT_OK;
    };
    // streams 0: the mode's default (graphs 3 streams, persistent one lane group per XCD)
    mt_batch_options cur{1u << 22, 1024, 4096, 8, 0, MT_BATCH_PERSISTENT};
    uint32_t xcds = 1;
    if (mt::xcd_count(&xcds)) return MT_E_HIP;
    double cur_v = 0.0;
    int rc = measure(cur, cur_v);
    if (rc) return rc;
    const uint32_t paths[] = {1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24};
    const uint32_t tiles[] = {256, 1024, 4096, 16384};
    const uint32_t blocks[] = {4, 8, 12};
    const uint32_t itiles[] = {1024, 4096, 16384};
    const uint32_t nstreams[] = {1, 2, 3, 4};
    const uint32_t ngroups[] = {xcds, 2 * xcds, 4 * xcds};  // persistent: groups within XCDs
    const uint32_t modes[] = {MT_BATCH_PERSISTENT, MT_BATCH_GRAPHS};
    auto sweep = [&](auto field, const uint32_t* vals, size_t nv) -> int {
        mt_batch_options b0 = cur;
        for (size_t i = 0; i < nv; ++i) {
            mt_batch_options o = b0;
            o.*field = vals[i];
            if (o.*field == b0.*field) continue;
            // lanes of the mode (0 = its default), and only shapes whose path state fits
            const uint32_t lanes = o.streams ? o.streams : (o.mode == MT_BATCH_GRAPHS ? 3u : xcds);
            if (o.mode == MT_BATCH_GRAPHS && lanes > kMaxLanes) continue;
            if (o.mode != MT_BATCH_GRAPHS && (lanes > kMaxGroups ||
                                              uint64_t(lanes) * o.paths_per_iteration > kMaxPersistentPaths))
                continue;
            double v = 0.0;
            const int r = measure(o, v);
            if (r) return r;
            if (v > cur_v) {
                cur = o;
                cur_v = v;
            }
        }
        return MT_OK;
    };
    // the launch mode first (each trial of the others then refines the better one); the
    // persistent launch holds as many blocks as the CUs do, so traversal blocks per CU only
    // matter to the graphs
    if ((rc = sweep(&mt_batch_options::mode, modes, 2)) ||
        (rc = sweep(&mt_batch_options::paths_per_iteration, paths, 5)) ||
        (rc = sweep(&mt_batch_options::tile_pixels, tiles, 4)) ||
        (cur.mode == MT_BATCH_GRAPHS && (rc = sweep(&mt_batch_options::trace_blocks_per_cu, blocks, 3))) ||
        (rc = sweep(&mt_batch_options::iteration_tiles, itiles, 3)) ||
        (rc = cur.mode == MT_BATCH_GRAPHS ? sweep(&mt_batch_options::streams, nstreams, 4)
                                          : sweep(&mt_batch_options::streams, ngroups, 3)))
        return rc;
    *best = cur;
    return MT_OK;
}

}  // extern "C"
