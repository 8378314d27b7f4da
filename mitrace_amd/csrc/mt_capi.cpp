// mt_capi.cpp -- the extern "C" entry points of include/mitrace_gpu.h.
//
// Errors never cross the ABI as exceptions: every entry point catches, records a message in a
// thread-local slot (mt_last_error) and returns a negative MT_E* code.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "mt_internal.hpp"
#include "mt_mlp.hpp"
#include "mt_render.hpp"

namespace mt {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
int fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}
}  // namespace mt

// Tracer state behind mt_job*: one host thread drives the chunk loop on its own HIP stream;
// kill is checked between chunks (the reference checks it between tiles, Tracer.cpp:620).
struct mt_job {
    mt_scene* scene = nullptr;
    mt_camera_desc cam{};
    mt_settings st{};
    float* rgb_out = nullptr;
    mt::Renderer renderer;
    std::thread worker;
    std::atomic<bool> killed{false};
    std::atomic<bool> done{false};
    std::atomic<uint32_t> chunks_done{0};
    uint32_t chunks_total = 0;
    int status = MT_OK;
    std::string error;
    mutable std::mutex mu;
    mt_stats snapshot{};
    std::chrono::steady_clock::time_point start;
};

// Radiance cache behind mt_cache*: the MLP and its options; it belongs to one scene.
struct mt_cache {
    mt_scene* scene = nullptr;
    mt_cache_options opt{};
    mt::Mlp mlp;
};

// Differentiable renderer behind mt_diff*: a Renderer set up once with the per-path store.
struct mt_diff {
    mt_scene* scene = nullptr;
    mt_settings st{};
    mt::Renderer renderer;
    bool ready = false;
};

namespace {

int run_job(mt_job* j, float* sample_rgb_out) {
    if (hipSetDevice(j->scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
    int rc = j->renderer.setup(j->scene, j->cam, j->st);
    if (rc) return rc;
    {
        std::scoped_lock lk(j->mu);
        j->chunks_total = j->renderer.chunks();
    }
    // chunks go round-robin to the renderer's stream slots; a slot's previous chunk is
    // completed (waited for, counted) before it takes the next, so completion is in order
    mt::Renderer& R = j->renderer;
    const uint32_t n = R.chunks(), slots = R.slots();
    uint32_t done = 0;
    auto complete = [&](uint32_t slot) -> int {
        if (!R.slot_busy(slot)) return MT_OK;
        const int r = R.complete_chunk(slot, sample_rgb_out);
        if (r) return r;
        j->chunks_done.store(++done);
        std::scoped_lock lk(j->mu);
        j->snapshot = R.stats();
        return MT_OK;
    };
    auto drain = [&]() {  // error path: wait for whatever is in flight
        for (uint32_t k = 0; k < slots; ++k) (void)complete(k);
    };
    for (uint32_t c = 0; c < n; ++c) {
        const uint32_t slot = c % slots;
        if ((rc = complete(slot))) {
            drain();
            return rc;
        }
        if (j->killed.load()) {
            drain();
            return mt::fail(MT_E_KILLED, "render killed");
        }
        if ((rc = R.enqueue_chunk(slot, c))) {
            drain();
            return rc;
        }
    }
    const uint32_t oldest = done;  // the remaining chunks, oldest first
    for (uint32_t k = 0; k < slots; ++k)
        if ((rc = complete((oldest + k) % slots))) {
            drain();
            return rc;
        }
    rc = R.finish(j->rgb_out);
    std::scoped_lock lk(j->mu);
    j->snapshot = j->renderer.stats();
    return rc;
}

void fill_stats(const mt_job* j, mt_stats* out) {
    std::scoped_lock lk(j->mu);
    *out = j->snapshot;
    const float total = j->chunks_total ? static_cast<float>(j->chunks_total) : 1.0f;
    out->progress = static_cast<float>(j->chunks_done.load()) / total;
    out->time_elapsed =
        std::chrono::duration<float>(std::chrono::steady_clock::now() - j->start).count();
    out->estimated_time_remaining =
        out->progress > 0.0f ? (out->time_elapsed / out->progress) * (1.0f - out->progress) : 0.0f;
}

bool valid_args(mt_scene* scene, const mt_camera_desc* cam, const mt_settings* st) {
    if (!scene || !cam || !st) {
        mt::set_error("null scene, camera or settings");
        return false;
    }
    return true;
}

int diff_params(const mt_diff* d, const mt_diff_params* p, mt::DiffParams& out) {
    out = mt::DiffParams{};
    if (!p) return MT_OK;
    const uint32_t nm = d->scene->material_count;
    if ((p->roughness_dev || p->metallic_dev || p->emission_dev) && p->material_count != nm)
        return mt::fail(MT_E_INVALID, "mt_diff_params.material_count does not match the scene");
    if (p->base_texture_dev) {
        if (p->tex_width == 0 || p->tex_height == 0)
            return mt::fail(MT_E_INVALID, "base texture needs a positive width and height");
        if (p->tex_material_first > nm || p->tex_material_count > nm - p->tex_material_first)
            return mt::fail(MT_E_INVALID, "textured material range exceeds the scene's materials");
        if (static_cast<uint64_t>(p->tex_width) * p->tex_height * 3 > 0xFFFFFFFFull)
            return mt::fail(MT_E_INVALID, "base texture too large");
    }
    out.tex = p->base_texture_dev;
    out.tex_w = p->tex_width;
    out.tex_h = p->tex_height;
    out.tex_first = p->tex_material_first;
    out.tex_count = p->base_texture_dev ? p->tex_material_count : 0;
    out.rough = p->roughness_dev;
    out.metal = p->metallic_dev;
    out.emis = p->emission_dev;
    return MT_OK;
}

}  // namespace

extern "C" {

int mt_abi_version(void) { return MT_ABI_VERSION; }

const char* mt_last_error(void) { return mt::g_error.c_str(); }

int mt_pool_trim(void) {
    mt::pool_trim();
    return MT_OK;
}

int mt_device_count(int* count) {
    if (!count) return mt::fail(MT_E_INVALID, "null count");
    if (hipGetDeviceCount(count) != hipSuccess) {
        *count = 0;
        return mt::fail(MT_E_HIP, "hipGetDeviceCount failed");
    }
    return MT_OK;
}

int mt_scene_create(const mt_scene_desc* desc, mt_scene** out) {
    if (!out) return mt::fail(MT_E_INVALID, "null output handle");
    *out = nullptr;
    try {
        auto s = std::make_unique<mt_scene>();
        if (hipGetDevice(&s->device) != hipSuccess) return mt::fail(MT_E_HIP, "no HIP device");
        const int rc = mt::upload_scene(desc, s.get());
        if (rc) return rc;
        *out = s.release();
        return MT_OK;
    } catch (const std::bad_alloc&) {
        return mt::fail(MT_E_NOMEM, "host allocation failed in mt_scene_create");
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_scene_destroy(mt_scene* scene) {
    delete scene;
    return MT_OK;
}

int mt_scene_mesh_bvh(const mt_scene* scene, uint32_t mesh, uint32_t* nodes_out, uint32_t* node_count,
                      uint32_t* indices_out, uint32_t* index_count) {
    if (!scene || mesh >= scene->mesh_count) return mt::fail(MT_E_INVALID, "bad scene or mesh index");
    const mt::HostBvh& b = scene->bvhs[mesh];
    if (node_count) *node_count = static_cast<uint32_t>(b.nodes.size());
    if (index_count) *index_count = static_cast<uint32_t>(b.indices.size());
    if (nodes_out) std::memcpy(nodes_out, b.nodes.data(), b.nodes.size() * sizeof(mt::HostNode));
    if (indices_out) std::memcpy(indices_out, b.indices.data(), b.indices.size() * sizeof(uint32_t));
    return MT_OK;
}

int mt_fast_bvh_build(const float* positions, uint32_t tri_count, float* pairs_out,
                      uint32_t* pair_count, uint32_t* slot_tris_out, uint32_t* slot_last_out,
                      uint32_t* slot_count, uint32_t* root_word, float* root_box) {
    if (tri_count && !positions) return mt::fail(MT_E_INVALID, "positions is NULL");
    try {
        std::vector<float4> pairs, prims;
        mt::DevMesh dm{};
        mt::build_fast_bvh(positions, tri_count, pairs, prims, dm);
        const uint32_t np = static_cast<uint32_t>(pairs.size() / 4);
        const uint32_t ns = static_cast<uint32_t>(prims.size() / 3);
        if (pair_count) *pair_count = np;
        if (slot_count) *slot_count = ns;
        if (root_word) *root_word = (dm.flags & mt::kMeshHasBvh) ? dm.root_word : (mt::kLeafBit | 0u);
        if (root_box) {
            if (dm.flags & mt::kMeshHasBvh) {
                for (int k = 0; k < 3; ++k) {
                    root_box[k] = dm.root_min[k];
                    root_box[3 + k] = dm.root_max[k];
                }
            } else {  // one leaf: the bounds of its triangles
                for (int k = 0; k < 3; ++k) {
                    root_box[k] = FLT_MAX;
                    root_box[3 + k] = -FLT_MAX;
                }
                for (uint64_t c = 0; c < 3ull * tri_count; ++c)
                    for (int k = 0; k < 3; ++k) {
                        root_box[k] = std::min(root_box[k], positions[3 * c + k]);
                        root_box[3 + k] = std::max(root_box[3 + k], positions[3 * c + k]);
                    }
            }
        }
        if (pairs_out) std::memcpy(pairs_out, pairs.data(), sizeof(float4) * pairs.size());
        for (uint32_t i = 0; i < ns; ++i) {
            uint32_t t, last;
            std::memcpy(&t, &prims[3 * i].w, 4);
            std::memcpy(&last, &prims[3 * i + 2].w, 4);
            if (slot_tris_out) slot_tris_out[i] = t;
            if (slot_last_out) slot_last_out[i] = last & mt::kPrimLast;
        }
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_NOMEM, e.what());
    }
}

int mt_render(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
              float* rgb_out, float* sample_rgb_out, mt_stats* stats) {
    if (!valid_args(scene, camera, settings)) return MT_E_INVALID;
    try {
        mt_job j;
        j.scene = scene;
        j.cam = *camera;
        j.st = *settings;
        j.rgb_out = rgb_out;
        j.start = std::chrono::steady_clock::now();
        const int rc = run_job(&j, sample_rgb_out);
        if (stats) fill_stats(&j, stats);
        return rc;
    } catch (const std::bad_alloc&) {
        return mt::fail(MT_E_NOMEM, "host allocation failed in mt_render");
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_render_start(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
                    float* rgb_out, mt_job** job) {
    if (!job) return mt::fail(MT_E_INVALID, "null job handle");
    *job = nullptr;
    if (!valid_args(scene, camera, settings)) return MT_E_INVALID;
    try {
        auto j = std::make_unique<mt_job>();
        j->scene = scene;
        j->cam = *camera;
        j->st = *settings;
        j->rgb_out = rgb_out;
        j->start = std::chrono::steady_clock::now();
        mt_job* raw = j.get();
        // nothing may escape the render thread (an exception there would terminate the
        // process): host allocation failures become MT_E_NOMEM, other exceptions MT_E_INVALID
        raw->worker = std::thread([raw]() {
            int rc;
            try {
                rc = run_job(raw, nullptr);
            } catch (const std::bad_alloc&) {
                rc = mt::fail(MT_E_NOMEM, "host allocation failed in the render thread");
            } catch (const std::exception& e) {
                rc = mt::fail(MT_E_INVALID, e.what());
            }
            raw->status = rc;
            if (rc) raw->error = mt_last_error();
            raw->done.store(true);
        });
        *job = j.release();
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_render_wait(mt_job* job) {
    if (!job) return mt::fail(MT_E_INVALID, "null job");
    if (job->worker.joinable()) job->worker.join();
    if (job->status) mt::set_error(job->error);
    return job->status;
}

int mt_render_kill(mt_job* job) {
    if (!job) return mt::fail(MT_E_INVALID, "null job");
    job->killed.store(true);
    return MT_OK;
}

int mt_render_stats(const mt_job* job, mt_stats* stats) {
    if (!job || !stats) return mt::fail(MT_E_INVALID, "null job or stats");
    fill_stats(job, stats);
    return MT_OK;
}

int mt_render_is_done(const mt_job* job) { return job && job->done.load() ? 1 : 0; }

int mt_job_destroy(mt_job* job) {
    if (!job) return MT_OK;
    if (job->worker.joinable()) {
        job->killed.store(true);
        job->worker.join();
    }
    delete job;
    return MT_OK;
}

int mt_trace_closest(mt_scene* scene, uint32_t n, const float* rays, mt_hit* hits) {
    return mt_trace_closest_ex(scene, n, rays, hits, MT_TRAVERSAL_REFERENCE);
}

int mt_trace_closest_ex(mt_scene* scene, uint32_t n, const float* rays, mt_hit* hits,
                        int32_t traversal) {
    if (!scene || (n && (!rays || !hits))) return mt::fail(MT_E_INVALID, "bad arguments");
    try {
        return mt::trace_rays(scene, n, rays, hits, traversal, nullptr);
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_trace_closest_noalloc(mt_scene* scene, uint32_t n, const float* rays_dev, mt_hit* hits_dev,
                             int32_t traversal, uint64_t stream) {
    if (!scene || (n && (!rays_dev || !hits_dev))) return mt::fail(MT_E_INVALID, "bad arguments");
    return mt::trace_rays_noalloc(scene, n, rays_dev, hits_dev, traversal,
                                  reinterpret_cast<hipStream_t>(stream));
}

int mt_trace_occluded(mt_scene* scene, uint32_t n, const float* rays, uint32_t* occluded,
                      int32_t traversal) {
    if (!scene || (n && (!rays || !occluded))) return mt::fail(MT_E_INVALID, "bad arguments");
    try {
        return mt::trace_rays(scene, n, rays, nullptr, traversal, occluded);
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_brdf_eval(uint32_t n, const float* in, float* out) {
    if (n && (!in || !out)) return mt::fail(MT_E_INVALID, "bad arguments");
    try {
        return mt::brdf_eval(n, in, out);
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_math_eval(int32_t op, uint32_t n, const float* in, float* out) {
    if (n && (!in || !out)) return mt::fail(MT_E_INVALID, "bad arguments");
    try {
        return mt::math_eval(op, n, in, out);
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_diff_create(mt_scene* scene, const mt_settings* settings, mt_diff** out) {
    if (!out) return mt::fail(MT_E_INVALID, "null output handle");
    *out = nullptr;
    if (!scene || !settings) return mt::fail(MT_E_INVALID, "null scene or settings");
    if (settings->bounces > 32)
        return mt::fail(MT_E_UNSUPPORTED, "differentiable rendering supports at most 32 bounces");
    if (settings->debug_mode)
        return mt::fail(MT_E_INVALID, "debug views are rendered by mt_render / mt_render_start");
    try {
        auto d = std::make_unique<mt_diff>();
        d->scene = scene;
        d->st = *settings;
        if (hipSetDevice(scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        mt_camera_desc cam{};
        cam.type = MT_CAMERA_PERSPECTIVE;
        cam.yfov = 1.0f;
        cam.aspect = 1.0f;
        for (int i = 0; i < 4; ++i) cam.camera_to_world[i * 5] = 1.0f;
        const int rc = d->renderer.setup(scene, cam, d->st, true);
        if (rc) return rc;
        d->ready = true;
        *out = d.release();
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_NOMEM, e.what());
    }
}

int mt_diff_forward(mt_diff* diff, const mt_camera_desc* camera, const mt_diff_params* params,
                    const mt_diff_params* sampling_params, uint64_t seed, float* image_dev,
                    uint64_t stream) {
    if (!diff || !diff->ready || !camera || !image_dev)
        return mt::fail(MT_E_INVALID, "null diff handle, camera or image");
    try {
        mt::DiffParams dp, dps;
        int rc = diff_params(diff, params, dp);
        if (rc) return rc;
        if (sampling_params && (rc = diff_params(diff, sampling_params, dps))) return rc;
        if (hipSetDevice(diff->scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        if (!stream && hipDeviceSynchronize() != hipSuccess)
            return mt::fail(MT_E_HIP, "hipDeviceSynchronize failed");
        diff->renderer.set_camera(*camera);
        diff->renderer.set_seed(seed);
        rc = diff->renderer.diff_forward(dp, sampling_params ? &dps : nullptr, image_dev,
                                         reinterpret_cast<hipStream_t>(stream));
        return rc;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_diff_backward(mt_diff* diff, const mt_camera_desc* camera, const mt_diff_params* params,
                     uint64_t seed, const float* grad_image_dev, const mt_diff_grads* grads,
                     uint64_t stream) {
    if (!diff || !diff->ready || !camera || !grad_image_dev || !grads)
        return mt::fail(MT_E_INVALID, "null diff handle, camera, gradient image or gradients");
    try {
        mt::DiffParams dp;
        int rc = diff_params(diff, params, dp);
        if (rc) return rc;
        mt::DiffGradPtrs gr{};
        gr.tex = dp.tex ? grads->base_texture_dev : nullptr;
        gr.rough = grads->roughness_dev;
        gr.metal = grads->metallic_dev;
        gr.emis = grads->emission_dev;
        gr.sampling = grads->detach_sampling ? 0 : 1;
        gr.materials = diff->scene->material_count;
        if ((gr.rough || gr.metal || gr.emis) && (!params || params->material_count != diff->scene->material_count))
            return mt::fail(MT_E_INVALID, "per-material gradients need material_count in the params");
        if (hipSetDevice(diff->scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        if (!stream && hipDeviceSynchronize() != hipSuccess)
            return mt::fail(MT_E_HIP, "hipDeviceSynchronize failed");
        diff->renderer.set_camera(*camera);
        diff->renderer.set_seed(seed);
        rc = diff->renderer.diff_adjoint(dp, gr, grad_image_dev, reinterpret_cast<hipStream_t>(stream));
        return rc;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_diff_destroy(mt_diff* diff) {
    delete diff;
    return MT_OK;
}

int mt_cache_create(mt_scene* scene, const mt_cache_options* options, mt_cache** out) {
    if (!out) return mt::fail(MT_E_INVALID, "null output handle");
    *out = nullptr;
    if (!scene) return mt::fail(MT_E_INVALID, "null scene");
    try {
        auto c = std::make_unique<mt_cache>();
        c->scene = scene;
        if (options) c->opt = *options;
        if (!c->opt.seed) c->opt.seed = 1;
        if (!c->opt.train_batch) c->opt.train_batch = 16384;
        if (!(c->opt.learning_rate > 0.0f)) c->opt.learning_rate = 5e-3f;
        if (!c->opt.train_every) c->opt.train_every = 16;
        if (!c->opt.query_depth && !c->opt.query_after_light) c->opt.query_depth = 1;
        if (!c->opt.steps_per_pass) c->opt.steps_per_pass = 8;
        if (!c->opt.spp_per_pass) c->opt.spp_per_pass = 1;
        if (c->opt.train_batch > (1u << 22)) return mt::fail(MT_E_UNSUPPORTED, "train_batch above 4M");
        if (hipSetDevice(scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        const int rc = c->mlp.init(c->opt.seed, std::max<uint32_t>(c->opt.train_batch, 1u << 16),
                                   c->opt.learning_rate);
        if (rc) return rc;
        c->mlp.set_loss(c->opt.loss);
        c->mlp.set_bounds(scene->world_lo, scene->world_hi);
        *out = c.release();
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_NOMEM, e.what());
    }
}

int mt_cache_destroy(mt_cache* cache) {
    delete cache;
    return MT_OK;
}

int mt_render_cached(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
                     mt_cache* cache, float* rgb_out, mt_stats* stats) {
    if (!valid_args(scene, camera, settings)) return MT_E_INVALID;
    if (!cache || cache->scene != scene) return mt::fail(MT_E_INVALID, "the cache belongs to another scene");
    try {
        if (hipSetDevice(scene->device) != hipSuccess) return mt::fail(MT_E_HIP, "hipSetDevice failed");
        const auto t0 = std::chrono::steady_clock::now();
        mt::Renderer r;
        int rc = r.setup_cached(scene, *camera, *settings, cache->opt);
        if (rc) return rc;
        rc = r.render_cached(cache->mlp, rgb_out);
        if (rc) return rc;
        if (stats) {
            *stats = r.stats();
            stats->progress = 1.0f;
            stats->time_elapsed =
                std::chrono::duration<float>(std::chrono::steady_clock::now() - t0).count();
            stats->estimated_time_remaining = 0.0f;
        }
        return MT_OK;
    } catch (const std::exception& e) {
        return mt::fail(MT_E_INVALID, e.what());
    }
}

int mt_cache_get_weights(mt_cache* cache, float* weights) {
    if (!cache || !weights) return mt::fail(MT_E_INVALID, "null cache or weights");
    return cache->mlp.get_master(weights);
}

int mt_cache_set_weights(mt_cache* cache, const float* weights) {
    if (!cache || !weights) return mt::fail(MT_E_INVALID, "null cache or weights");
    return cache->mlp.set_master(weights);
}

namespace {

// host float3 arrays -> device float4 arrays
int upload3(std::vector<void*>& owned, uint32_t n, const float* src, float4** dst) {
    std::vector<float4> tmp(n);
    for (uint32_t i = 0; i < n; ++i) tmp[i] = make_float4(src[i * 3], src[i * 3 + 1], src[i * 3 + 2], 0.0f);
    void* p = nullptr;
    if (mt::device_malloc(&p, sizeof(float4) * std::max<uint32_t>(n, 1))) return MT_E_NOMEM;
    owned.push_back(p);
    *dst = static_cast<float4*>(p);
    if (n && hipMemcpy(p, tmp.data(), sizeof(float4) * n, hipMemcpyHostToDevice) != hipSuccess)
        return mt::fail(MT_E_HIP, "upload failed");
    return MT_OK;
}

struct FreeAll {
    std::vector<void*>& v;
    ~FreeAll() {
        for (void* p : v) (void)hipFree(p);
    }
};

}  // namespace

int mt_cache_infer(mt_cache* cache, uint32_t n, const float* pos, const float* dir, float* out) {
    if (!cache || (n && (!pos || !dir || !out))) return mt::fail(MT_E_INVALID, "bad arguments");
    if (n == 0) return MT_OK;
    std::vector<void*> owned;
    FreeAll guard{owned};
    float4 *dp, *dd, *dout;
    int rc;
    if ((rc = upload3(owned, n, pos, &dp)) || (rc = upload3(owned, n, dir, &dd)) ||
        (rc = upload3(owned, n, pos, &dout)))
        return rc;
    if ((rc = cache->mlp.infer(n, dp, dd, dout, nullptr))) return rc;
    std::vector<float4> tmp(n);
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(tmp.data(), dout, sizeof(float4) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return mt::fail(MT_E_HIP, "inference failed");
    for (uint32_t i = 0; i < n; ++i) {
        out[i * 3] = tmp[i].x;
        out[i * 3 + 1] = tmp[i].y;
        out[i * 3 + 2] = tmp[i].z;
    }
    return MT_OK;
}

int mt_cache_train(mt_cache* cache, uint32_t n, const float* pos, const float* dir,
                   const float* target, float* loss, float* grad) {
    if (!cache || (n && (!pos || !dir || !target))) return mt::fail(MT_E_INVALID, "bad arguments");
    if (n == 0) return MT_OK;
    std::vector<void*> owned;
    FreeAll guard{owned};
    float4 *dp, *dd, *dt;
    int rc;
    if ((rc = upload3(owned, n, pos, &dp)) || (rc = upload3(owned, n, dir, &dd)) ||
        (rc = upload3(owned, n, target, &dt)))
        return rc;
    float l = 0.0f;
    if ((rc = cache->mlp.train(n, dp, dd, dt, nullptr, &l, grad))) return rc;
    if (loss) *loss = l;
    return MT_OK;
}

}  // extern "C"
