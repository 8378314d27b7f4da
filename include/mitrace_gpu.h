/*
 * mitrace_gpu.h -- C-ABI of the MI355X (gfx950) implementation of MiTrace's Tracer stage.
 *
 * The reference has no FFI: its Tracer is a concrete C++ class that is linked statically
 * (src/Tracer/Tracer.hpp:19-129).  This header is the drop-in boundary that replaces it:
 *
 *   reference                                         this ABI
 *   -----------------------------------------------   -----------------------------------
 *   Scene::Scene(const Loader::Scene&)                mt_scene_create / mt_scene_destroy
 *     (src/Scene/Scene.cpp:53-59; BVH build            (host BVH build with the same rules,
 *      src/Scene/BVH.cpp:117-282, Mesh.cpp:24-41)       flattened into HBM)
 *   Tracer::Tracer(buf, scene) + StartRender()        mt_render_start
 *     (src/Tracer/Tracer.cpp:633-640, 718-779)
 *   Tracer::WaitForRender()  (Tracer.cpp:781-799)     mt_render_wait
 *   Tracer::KillRender()     (Tracer.cpp:801-804)     mt_render_kill
 *   Tracer::GetStats()       (Tracer.cpp:806-821)     mt_render_stats
 *   Tracer::IsDone()         (Tracer.cpp:823-828)     mt_render_is_done
 *   main.cpp:121-142 (start + wait, synchronous)      mt_render
 *   IntersectScene (src/Tracer/Intersect.cpp:201-219) mt_trace_closest   (test hook)
 *   BRDF::EvaluateBRDF / DiffusePdf / SpecularPdf     mt_brdf_eval       (test hook)
 *   sinf/cosf/atan2/acos/pow of the path              mt_math_eval       (test hook)
 *     (src/Tracer/BRDF.hpp:52-97)
 *
 * Conventions: plain C types only; every pointer argument is a HOST pointer unless its name
 * ends in _dev; every function returns 0 on success or a negative MT_E* code, and
 * mt_last_error() gives a message for the calling thread.  No C++ exception crosses the ABI.
 * Matrices are 4x4 float, column-major (m[col*4 + row]), as glTF and GLM store them.
 */
#ifndef MITRACE_GPU_H
#define MITRACE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MT_ABI_VERSION 3

/* ---- status codes ---------------------------------------------------------------------- */
#define MT_OK 0
#define MT_E_INVALID (-1)   /* bad argument or scene description                        */
#define MT_E_HIP (-2)       /* a HIP runtime call failed                                 */
#define MT_E_NOMEM (-3)     /* device or host allocation failed                          */
#define MT_E_UNSUPPORTED (-4)
#define MT_E_KILLED (-5)    /* render was stopped by mt_render_kill before it finished   */

/* ---- enums mirroring the reference ----------------------------------------------------- */
/* Scene::Light::LightType (src/Scene/Light.hpp:14-20) */
#define MT_LIGHT_POINT 0
#define MT_LIGHT_DIRECTIONAL 1
#define MT_LIGHT_SPOT 2
#define MT_LIGHT_AREA 3
/* Scene::Camera::Type (src/Scene/Camera.hpp:13-17) */
#define MT_CAMERA_PERSPECTIVE 0
#define MT_CAMERA_ORTHOGONAL 1
/* Scene::Image::FilterMode (src/Scene/Image.hpp:15-19) */
#define MT_FILTER_NEAREST 0
#define MT_FILTER_LINEAR 1
/* environment texel formats: the reference loads 8-bit RGB only (Image.cpp:77-104);
 * RGB32F is the float equirect the large configurations need. */
#define MT_ENV_NONE 0
#define MT_ENV_RGB8 1
#define MT_ENV_RGB32F 2
/* estimator modes */
#define MT_MODE_PARITY 0 /* the reference estimator, quirks included (Tracer.cpp:134-348) */
#define MT_MODE_FIXED 1  /* Russian roulette instead of the energy cut-off                  */
/* per-sample output */
/* Traversal.  REFERENCE visits exactly the reference's nodes and triangles (Intersect.cpp:9-219: the
 * reference's BVHs, exit-distance culling with its 0.1 slack, every instance whose world box
 * the ray line crosses, closest-hit shadow rays), so hits are the reference's, including its
 * occasional non-closest ones.  FAST uses per-mesh SAH BVHs with tight boxes (built when the
 * scene sets mt_bvh_options.fast), culls nodes and instances by entry distance against the
 * closest hit so far, and ends a shadow ray at its first occluder: hits are the true closest
 * hits, and images differ from REFERENCE ones only where the reference returns a non-closest
 * hit.  No reference counterpart (a GPU speed mode). */
#define MT_TRAVERSAL_REFERENCE 0
#define MT_TRAVERSAL_FAST 1

#define MT_OUTPUT_TONEMAPPED 0 /* ACES per sample before averaging (Tracer.cpp:212-227,346) */
#define MT_OUTPUT_LINEAR 1     /* mean of linear radiance                                   */
/* debug views: Config::DebugMode (src/CLI/Config.hpp:16-41) in the same order, rendered as
 * ProcessDebugRay (src/Tracer/Tracer.cpp:366-488) and post-processed as
 * ApplyDebugPostProcessing (Tracer.cpp:642-716) */
#define MT_DEBUG_NONE 0
#define MT_DEBUG_PRIMARY_BVH_TESTS 1
#define MT_DEBUG_TOTAL_BVH_TESTS 2
#define MT_DEBUG_PRIMARY_TRIANGLE_TESTS 3
#define MT_DEBUG_TOTAL_TRIANGLE_TESTS 4
#define MT_DEBUG_ALBEDO 5
#define MT_DEBUG_METALLIC_ROUGHNESS 6
#define MT_DEBUG_GEOMETRIC_NORMAL 7
#define MT_DEBUG_SHADING_NORMAL 8
#define MT_DEBUG_DIRECT_ONLY 9
#define MT_DEBUG_INDIRECT_ONLY 10
#define MT_DEBUG_EMISSION 11
#define MT_DEBUG_COLOR_PER_MESH 12
#define MT_DEBUG_COLOR_PER_TRIANGLE 13
#define MT_DEBUG_COLOR_PER_BOUNDING_VOLUME 14
#define MT_DEBUG_BOUNCES 15
#define MT_DEBUG_DEPTH 16
#define MT_DEBUG_FIRST_HIT_FRESNEL 17
#define MT_DEBUG_FIRST_HIT_BRDF 18
#define MT_DEBUG_FIRST_HIT_PDF 19
#define MT_DEBUG_REFLECTED_DIRECTION 20
#define MT_DEBUG_PIXEL_STANDARD_DEVIATION 21
#define MT_DEBUG_FIREFLY_ELIMINATION 22

/* ---- scene description (mirrors Scene::Mesh / MeshInstance / Material / Light / Camera) - */

/* A de-indexed triangle soup (Scene::Mesh, src/Scene/Mesh.hpp:13-42): corner k of
 * triangle t is element 3*t+k of every per-corner array. */
typedef struct mt_mesh_desc {
    uint32_t tri_count;
    const float* positions; /* tri_count*9 floats                      */
    const float* normals;   /* tri_count*9 floats                      */
    const float* texcoord0; /* tri_count*6 floats, or NULL             */
} mt_mesh_desc;

/* 8-bit RGB image (Scene::Image, src/Scene/Image.hpp:12-51): row-major, 3 bytes per texel. */
typedef struct mt_image_desc {
    uint32_t width;
    uint32_t height;
    const uint8_t* rgb;
} mt_image_desc;

/* Scene::Texture (src/Scene/Texture.hpp:14-40): image index (-1 = no texture), filter and the
 * KHR_texture_transform offset/scale. */
typedef struct mt_texture_desc {
    int32_t image;
    int32_t filter;
    float offset[2];
    float scale[2];
} mt_texture_desc;

/* Scene::Material (src/Scene/Material.hpp:32-44) */
typedef struct mt_material_desc {
    float base_color_factor[4];
    float emissive_factor[3];
    float normal_scale;
    float metallic_factor;
    float roughness_factor;
    float occlusion_strength;
    mt_texture_desc base_color_tex;
    mt_texture_desc metallic_roughness_tex;
    mt_texture_desc normal_tex;
    mt_texture_desc occlusion_tex;
    mt_texture_desc emissive_tex;
} mt_material_desc;

/* Scene::MeshInstance (src/Scene/Mesh.hpp:44-72) */
typedef struct mt_instance_desc {
    uint32_t mesh;
    uint32_t material;
    float transform[16]; /* model-to-world, column-major */
} mt_instance_desc;

/* Scene::Light (src/Scene/Light.hpp:14-66).  color is already color*intensity. */
typedef struct mt_light_desc {
    int32_t type;
    float position[3];
    float color[3];
    float point_size;
    float spot_inner_cos;
    float spot_outer_cos;
    float direction[3]; /* quat * (0,0,-1) for spot/directional lights */
} mt_light_desc;

/* Scene::Camera (src/Scene/Camera.hpp:10-36) */
typedef struct mt_camera_desc {
    int32_t type;
    float camera_to_world[16];
    float yfov;
    float ortho_size[2];
    float aspect;
} mt_camera_desc;

/* Environment map, sampled as an equirect (Image::SampleEquirectangular, Image.cpp:63-75). */
typedef struct mt_env_desc {
    int32_t format; /* MT_ENV_* */
    uint32_t width;
    uint32_t height;
    const void* texels; /* RGB8: w*h*3 bytes; RGB32F: w*h*3 floats */
    int32_t filter;     /* the reference always uses MT_FILTER_LINEAR (Texture.cpp:35-50) */
} mt_env_desc;

/* BVH build options read from the Config singleton by Mesh::Mesh (src/Scene/Mesh.cpp:37-40). */
typedef struct mt_bvh_options {
    int32_t disable;        /* --disable-bvh        (default 0)  */
    uint32_t max_triangles; /* --bvh-max-triangles  (default 32) */
    uint32_t max_depth;     /* --bvh-max-depth      (default 32) */
    int32_t fast;           /* 1 = also build the SAH BVHs of MT_TRAVERSAL_FAST (no reference
                               counterpart; the reference builds only the trees above)       */
    uint32_t device_build_min_tris; /* build the reference-rule BVH of every mesh with at least
                               this many triangles on the GPU, bit-identical to the host
                               build (BVH.cpp:117-272 rules); 0 = all on the host            */
    uint32_t fast_device_build_min_tris; /* with fast: build the fast traversal's tree of every
                               mesh with at least this many triangles on the GPU as a linear
                               BVH (triangles in Morton order of their box centres, subtrees of
                               at most 4 triangles as leaves) instead of the host's binned SAH;
                               0 = all on the host (no reference counterpart)                */
} mt_bvh_options;

typedef struct mt_scene_desc {
    const mt_mesh_desc* meshes;
    uint32_t mesh_count;
    const mt_image_desc* images;
    uint32_t image_count;
    const mt_material_desc* materials;
    uint32_t material_count;
    const mt_instance_desc* instances;
    uint32_t instance_count;
    const mt_light_desc* lights;
    uint32_t light_count;
    mt_env_desc env;
    mt_bvh_options bvh;
} mt_scene_desc;

/* ---- render settings (the fields of Config::ConfigOptions the path reads, plus GPU ones) - */
typedef struct mt_settings {
    uint32_t width;               /* imageWidth             (Config.hpp:50)           */
    uint32_t height;              /* imageHeight            (Config.hpp:51)           */
    uint32_t samples;             /* samples                (Config.hpp:52)           */
    uint32_t bounces;             /* bounces                (Config.hpp:53)           */
    float ev_exposure;            /* evExposure             (Config.hpp:56)           */
    float hdri_rotation;          /* degrees; used only when hdri_rotation_apply != 0 */
    float hdri_secondary_intensity;
    float hdri_primary_intensity;
    float emission_base_intensity;
    int32_t firefly_enable;       /* useStatisticFireflyElimination (CLI default off, Config.cpp:166-170) */
    float firefly_threshold;
    float terminate_energy;
    uint32_t image_block_size;    /* tile edge for progress/kill granularity          */
    /* additions */
    uint64_t seed;                /* counter-based RNG key                            */
    int32_t mode;                 /* MT_MODE_*                                        */
    uint32_t rr_depth;            /* MT_MODE_FIXED: roulette starts after this depth  */
    int32_t output;               /* MT_OUTPUT_*                                      */
    int32_t hdri_rotation_apply;  /* 0 = ignore rotation like the reference (Texture.cpp:35-61) */
    uint32_t paths_in_flight;     /* wavefront size; 0 = automatic                    */
    int32_t count_work;           /* 1 = gather traversal work counters (slower)      */
    int32_t debug_mode;           /* debugMode (Config.hpp:74), MT_DEBUG_*            */
    int32_t traversal;            /* MT_TRAVERSAL_*                                   */
} mt_settings;

/* Tracer::Stats (src/Tracer/Tracer.hpp:78-85) plus work counters. */
typedef struct mt_stats {
    float progress;
    float time_elapsed;
    float estimated_time_remaining;
    uint64_t rays;    /* camera + extension + shadow rays (Tracer.cpp:146, 319) */
    uint64_t samples; /* camera paths                           */
    /* traversal work, filled when mt_settings.count_work != 0 */
    uint64_t inst_aabb_tests;
    uint64_t inst_entered;
    uint64_t node_pops;
    uint64_t tri_tests;
    uint64_t hits_shaded;
    uint64_t misses;
    float kernel_ms_extend;  /* summed HIP-event time of the extend launches   */
    float kernel_ms_shadow;  /* summed HIP-event time of the shadow launches   */
    float kernel_ms_shade;   /* summed HIP-event time of the shade launches    */
    float kernel_ms_total;   /* summed HIP-event time of all render launches   */
    uint32_t extend_launches;
    uint32_t shadow_launches;
    /* traversal diagnostics, filled when count_work != 0: [0] instance-culling rounds,
     * lane steps [1] set-up/finish [2] culled pop [3] inner node [4] leaf, wave iterations
     * [5] all [6] with a set-up lane [7] with an inner-node lane [8] with a leaf lane */
    uint64_t diag[9];
} mt_stats;

/* One closest-hit record (Tracer::RayHit, src/Tracer/Ray.hpp:20-39). */
typedef struct mt_hit {
    int32_t instance;    /* -1 = miss */
    uint32_t triangle;   /* triangle index inside the mesh (soup order) */
    uint32_t bv_index;   /* BVH leaf node that produced the hit (0 for linear meshes) */
    float u, v;          /* barycentrics of corners 1 and 2 */
    float distance;      /* world-space |P - O| */
    float world_pos[3];
    uint32_t bvh_tests;      /* node pops in the winning instance (Intersect.cpp:71)  */
    uint32_t triangle_tests; /* triangle tests in the winning instance (Intersect.cpp:93) */
} mt_hit;

/* Differentiable-rendering parameters (no reference counterpart: the inverse-rendering form).
 * All pointers are DEVICE pointers (e.g. torch tensors' data_ptr()); NULL keeps the scene's
 * value.  Materials [tex_material_first, tex_material_first + tex_material_count) take their
 * base colour from base_texture (RGB float, sampled like the reference's "linear" filter). */
typedef struct mt_diff_params {
    const float* base_texture_dev; /* tex_height*tex_width*3 */
    uint32_t tex_width;
    uint32_t tex_height;
    uint32_t tex_material_first;
    uint32_t tex_material_count;
    const float* roughness_dev; /* material_count */
    const float* metallic_dev;  /* material_count */
    const float* emission_dev;  /* material_count*3 */
    uint32_t material_count;
} mt_diff_params;

/* Gradient accumulators (DEVICE pointers, same layouts as mt_diff_params; NULL = skip).
 * mt_diff_backward ADDS to them.  detach_sampling = 0 (default): the derivative of the expected
 * image, including how the sampling decisions depend on the parameters (the lobe choice's
 * probability, the GGX half-vector density and the mixture pdf, Tracer.cpp:49-82, 177-200, in
 * score-function form); 1: those decisions and pdfs held fixed (the derivative of the
 * frozen-sampling estimator of mt_diff_forward's sampling_params). */
typedef struct mt_diff_grads {
    float* base_texture_dev;
    float* roughness_dev;
    float* metallic_dev;
    float* emission_dev;
    int32_t detach_sampling;
} mt_diff_grads;

/* Batched rendering (no reference counterpart): knobs of the iteration loop.
 * Zero fields take the defaults in brackets. */
#define MT_BATCH_PERSISTENT 1 /* one cooperative launch serves the batch from a device queue */
#define MT_BATCH_GRAPHS 2     /* a host thread replays a captured graph per iteration        */
typedef struct mt_batch_options {
    uint32_t paths_per_iteration; /* wavefront width: camera paths in flight [4194304]   */
    uint32_t tile_pixels;         /* pixels per work item of the device queue [1024]      */
    uint32_t iteration_tiles;     /* queue capacity: tiles one iteration may take [4096] */
    uint32_t trace_blocks_per_cu; /* graphs: persistent traversal blocks per CU [8]
                                     (persistent: as many blocks as the CUs hold)        */
    uint32_t streams;             /* concurrent lanes, each with its own path state: streams
                                     (graphs) or groups of the launch's blocks (persistent) [3] */
    uint32_t mode;                /* MT_BATCH_PERSISTENT or MT_BATCH_GRAPHS [persistent]  */
} mt_batch_options;

typedef struct mt_batch_job_info {
    int32_t done;       /* 1 once the iteration holding the job's last tile has completed */
    int32_t status;     /* MT_OK, or the error that stopped the batch                      */
    uint32_t tiles;
    uint32_t pad;
    uint64_t samples;   /* camera paths = width * height * samples                         */
    double submit_ms;   /* host clock, ms since mt_batch_create                            */
    double complete_ms; /* device completion of the job's last iteration, same clock       */
} mt_batch_job_info;

typedef struct mt_batch_stats {
    uint64_t iterations;       /* wavefront iterations (graphs: graph launches)        */
    uint64_t graphs_captured;  /* captures (one per scene and bounce count; graphs)    */
    uint64_t samples;          /* camera paths of all jobs                             */
    uint32_t jobs;
    uint32_t scenes;
    mt_batch_options options;  /* the values in effect                                 */
    uint32_t pad;
    uint64_t launches;         /* kernel launches of the persistent mode: 1 unless the
                                  queue stayed empty for 1 s (the launch then ends and
                                  the next submission starts another)                   */
} mt_batch_stats;

/* one trial of mt_batch_autotune */
typedef struct mt_autotune_trial {
    mt_batch_options options;
    double seconds;
    double msamples_per_s;
} mt_autotune_trial;

/* Radiance cache (no reference counterpart): an MLP (32 frequency-encoded inputs from the
 * position and the direction, 4 hidden layers of 64, 3 outputs; BF16 on MFMA, FP32
 * accumulation) trained online on the radiance of the tracer's own training paths.
 * Zero fields take the defaults in brackets. */
typedef struct mt_cache_options {
    uint32_t seed;            /* weight initialisation [1]                                 */
    uint32_t train_batch;     /* samples per Adam step [16384]                             */
    float learning_rate;      /* Adam step size [0.005]                                    */
    uint32_t train_every;     /* one full-length training path per this many paths [16];
                                 0xFFFFFFFF: no training paths                              */
    uint32_t query_depth;     /* bounce at which the other paths stop and query [1]        */
    uint32_t warmup_passes;   /* training-only passes before the first counted pass (0: none) */
    uint32_t steps_per_pass;  /* at most this many Adam steps after every chunk [8]        */
    uint32_t spp_per_pass;    /* samples per pixel of one pass [1]                          */
    uint32_t query_after_light; /* 1: a querying path takes the emission and the light sample
                                   of its query vertex and stops instead of tracing the next
                                   bounce; the cache then stands for the light arriving
                                   through that bounce; training paths record only their
                                   query_depth vertex, and query_depth 0 (the camera ray's
                                   hit) is allowed [0]                                      */
    uint32_t train_nee;       /* 1 (needs query_after_light): the query is made whenever the
                                 path has a bounce left (not only when it would go on), so
                                 the cache stands for the expected rest of the path, and the
                                 training paths estimate that rest with the same expectation
                                 but less variance: from their query vertex on, each vertex's
                                 shadow ray goes to the environment (drawn in proportion to
                                 its radiance) or, with probability 1/2 when the scene has
                                 point lights, to the reference's light choice (weight 2);
                                 environment samples and bounce rays that reach the sky are
                                 combined by the balance heuristic with the reference's true
                                 sampling density [0]                                       */
    uint32_t loss;            /* training loss per output channel: 0 = relative L2,
                                 ((y - t) / sqrt(y^2 + 0.01))^2 with the denominator held,
                                 Huber-clipped at 10; 1 = L2, (y - t)^2, which keeps the mean:
                                 at a stationary point the output biases make the mean residual
                                 over the training samples zero; the render passes then train
                                 on targets divided by the first training chunk's mean target
                                 and multiply the predictions by it [0]                        */
} mt_cache_options;

typedef struct mt_scene mt_scene;
typedef struct mt_job mt_job;
typedef struct mt_cache mt_cache;
typedef struct mt_diff mt_diff;
typedef struct mt_batch mt_batch;

/* ---- functions -------------------------------------------------------------------------- */
int mt_abi_version(void);
const char* mt_last_error(void);
int mt_device_count(int* count);
/* Renders keep their finished path buffers in a per-process device pool for the next render of
 * the same size (at most MT_POOL_GB GiB, default 64; 0 = off).  This hands every pooled block
 * back to the device, e.g. before a caller allocates much device memory elsewhere.  No reference
 * counterpart (the reference tracer holds no device memory). */
int mt_pool_trim(void);

/* Scene::Scene(loaderScene): builds each mesh's BVH with the reference rules and uploads the
 * flattened scene to HBM of the current device.  The description is copied; the caller may
 * free it afterwards. */
int mt_scene_create(const mt_scene_desc* desc, mt_scene** out);
int mt_scene_destroy(mt_scene* scene);
/* Export of mesh `mesh`'s flattened BVH (BVH::Flatten, BVH.cpp:261-272): 8 words per node
 * {min.xyz, max.xyz, a, b} as raw bits, and the triangle index list.  Pass NULL buffers to
 * query the sizes. */
int mt_scene_mesh_bvh(const mt_scene* scene, uint32_t mesh, uint32_t* nodes_out,
                      uint32_t* node_count, uint32_t* indices_out, uint32_t* index_count);

/* The SAH BVH that mt_bvh_options.fast builds for one mesh, computed on the host only (no
 * device is needed): positions = tri_count*9 floats (the mesh soup).  pairs_out receives 16
 * floats per sibling pair {A.min, A.word} {A.max, B.word} {B.min, 0} {B.max, 0} (words as raw
 * bits: pair index of an inner node, or first slot | 0x80000000 of a leaf), slot_tris_out the
 * triangle of every leaf slot and slot_last_out 1 on a leaf's last slot.  Pass NULL buffers to
 * query the counts.  No reference counterpart: the host-side test hook of the fast traversal's
 * builder. */
int mt_fast_bvh_build(const float* positions, uint32_t tri_count, float* pairs_out,
                      uint32_t* pair_count, uint32_t* slot_tris_out, uint32_t* slot_last_out,
                      uint32_t* slot_count, uint32_t* root_word, float* root_box);

/* Synchronous render (main.cpp:121-142): writes width*height*3 floats to rgb_out.
 * sample_rgb_out, when not NULL, receives every per-sample colour ([pixel][sample][3]). */
int mt_render(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
              float* rgb_out, float* sample_rgb_out, mt_stats* stats);

/* Asynchronous render with the Tracer's start/wait/kill/stats surface. */
int mt_render_start(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
                    float* rgb_out, mt_job** job);
int mt_render_wait(mt_job* job);
int mt_render_kill(mt_job* job);
int mt_render_stats(const mt_job* job, mt_stats* stats);
int mt_render_is_done(const mt_job* job);
int mt_job_destroy(mt_job* job);

/* IntersectScene over n rays; rays holds 6 floats per ray {origin.xyz, direction.xyz}. */
int mt_trace_closest(mt_scene* scene, uint32_t n, const float* rays, mt_hit* hits);

/* mt_trace_closest with a choice of traversal (MT_TRAVERSAL_*).  No reference counterpart:
 * the test hook of the fast traversal. */
int mt_trace_closest_ex(mt_scene* scene, uint32_t n, const float* rays, mt_hit* hits,
                        int32_t traversal);

/* mt_trace_closest_ex without any device allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  rays_dev (6 floats per ray) and hits_dev (n
 * mt_hit) are DEVICE pointers the caller owns; the traversal's scratch (work counters, error
 * flag, stack spill area) lives in a 16 MiB static device array per GPU, held until the process
 * ends.  A scene whose BVH needs a deeper stack than that array holds is rejected
 * (MT_E_UNSUPPORTED).  Runs on `stream` (a hipStream_t, 0 = the null stream) and returns after
 * the work has completed.  Calls that overlap on one GPU share its scratch and race. */
int mt_trace_closest_noalloc(mt_scene* scene, uint32_t n, const float* rays_dev, mt_hit* hits_dev,
                             int32_t traversal, uint64_t stream);

/* Shadow-ray test hook (Tracer.cpp:317-323): rays = n x {ox, oy, oz, dx, dy, dz, dist} host
 * floats; occluded[i] = 1 when the closest hit is at distance <= dist (REFERENCE: the full
 * closest-hit query, FAST: the first occluder found), else 0. */
int mt_trace_occluded(mt_scene* scene, uint32_t n, const float* rays, uint32_t* occluded,
                      int32_t traversal);

/* BRDF::EvaluateBRDF, DiffusePdf, SpecularPdf on the GPU.  in: 14 floats per case
 * {L.xyz, V.xyz, N.xyz, base.rgb, roughness, metallic}; out: 5 floats per case
 * {f.rgb, diffuse_pdf, specular_pdf}. */
int mt_brdf_eval(uint32_t n, const float* in, float* out);

/* The path's elementary functions as the kernels compute them (the reference calls the C
 * library's sinf/cosf (Tracer.cpp:34-93), atan2/acos (Image.cpp:63-75) and glm::pow (BRDF.hpp:9-13)
 * on the CPU; this build defines them once, identically on the host oracle and on the GPU, so
 * that renders agree bit for bit).  Test hook: in = n pairs {x, y}, out = n results;
 * MT_MATH_ATAN2 computes atan2(x, y), the others use x only. */
#define MT_MATH_SIN 0
#define MT_MATH_COS 1
#define MT_MATH_ATAN2 2
#define MT_MATH_ACOS 3
#define MT_MATH_POW5 4
int mt_math_eval(int32_t op, uint32_t n, const float* in, float* out);

/* Path-replay adjoint (no reference counterpart; for inverse rendering).  mt_diff_create
 * sizes per-path storage (16 bytes per camera path: linear radiance and shadow visibility
 * bits) for settings->width*height*samples paths; the estimator is the one of *settings
 * with MT_OUTPUT_LINEAR and no firefly rejection.  mt_diff_forward writes the linear image
 * (width*height*3 floats, device) and stores the per-path data; mt_diff_backward replays every
 * path from its counter-RNG key and accumulates d(loss)/d(params) given grad_image_dev =
 * d(loss)/d(image).  sampling_params (NULL = params) freezes the sampling: every decision
 * (lobe, direction, pdf, termination) is taken with sampling_params and only the values are
 * evaluated with params -- the estimator whose derivative the adjoint computes, so central
 * differences of it check the adjoint exactly.  seed replaces settings->seed (a backward call must use the seed, camera
 * and params of the forward call it follows).  stream: a hipStream_t to run on (0 = the
 * library's own stream, after a device synchronisation); both calls return after their work
 * on that stream has completed. */
int mt_diff_create(mt_scene* scene, const mt_settings* settings, mt_diff** out);
int mt_diff_forward(mt_diff* diff, const mt_camera_desc* camera, const mt_diff_params* params,
                    const mt_diff_params* sampling_params, uint64_t seed, float* image_dev,
                    uint64_t stream);
int mt_diff_backward(mt_diff* diff, const mt_camera_desc* camera, const mt_diff_params* params,
                     uint64_t seed, const float* grad_image_dev, const mt_diff_grads* grads,
                     uint64_t stream);
int mt_diff_destroy(mt_diff* diff);

/* Radiance cache.  mt_cache_create sizes the training buffers and initialises the weights;
 * the position encoding spans the scene's bounds.  mt_render_cached renders like mt_render
 * (same settings, camera, output and stats; no firefly rejection) in passes of spp_per_pass
 * samples per pixel: 1 path in train_every runs the plain estimator to its end and records
 * training samples (position, direction, radiance of the rest of the path) at its hits from
 * query_depth on; every other path stops at its hit of bounce query_depth and adds
 * throughput x cache(position, direction).  The cache is trained after every chunk, and keeps
 * its weights across calls.  The remaining calls are test hooks over host arrays (3 floats per
 * position, direction, target or output; weights: 14787 floats, W then b per layer, W[out][in]). */
int mt_cache_create(mt_scene* scene, const mt_cache_options* options, mt_cache** out);
int mt_cache_destroy(mt_cache* cache);
int mt_render_cached(mt_scene* scene, const mt_camera_desc* camera, const mt_settings* settings,
                     mt_cache* cache, float* rgb_out, mt_stats* stats);
int mt_cache_get_weights(mt_cache* cache, float* weights);
int mt_cache_set_weights(mt_cache* cache, const float* weights);
int mt_cache_infer(mt_cache* cache, uint32_t n, const float* pos, const float* dir, float* out);
/* one Adam step on n samples; loss (optional) = mean relative L2 of the step's forward pass,
 * grad (optional, 14787 floats) = the gradient it applied */
int mt_cache_train(mt_cache* cache, uint32_t n, const float* pos, const float* dir,
                   const float* target, float* loss, float* grad);

/* Batched rendering (no reference counterpart; for serving many render requests).  Jobs of
 * any scene, resolution, sample count and depth are split into tiles of tile_pixels pixels.
 * MT_BATCH_PERSISTENT: one cooperative kernel launch serves the whole batch; the host
 * appends job records and tiles to a queue in pinned memory, and each lane (a group of the
 * launch's blocks with its own path state) takes the next tiles of one scene, runs the whole
 * wavefront iteration (generate, extend/shade/shadow per bounce, finalize, resolve into each
 * job's image) between group barriers and marks the jobs it completed in pinned memory.
 * MT_BATCH_GRAPHS: per-scene device queues (one per lane); one host thread replays, per
 * iteration, a captured graph of the same steps for the scene of the oldest unfinished job.
 * Every job's image equals mt_render's with the same scene, camera and settings.
 * mt_batch_submit may be called from any thread before or while the batch runs; rgb_out
 * (host, width*height*3 floats, may be NULL) is written by mt_batch_finish. */
int mt_batch_create(const mt_batch_options* options, mt_batch** out);
int mt_batch_submit(mt_batch* batch, mt_scene* scene, const mt_camera_desc* camera,
                    const mt_settings* settings, float* rgb_out, uint32_t* job_id);
int mt_batch_start(mt_batch* batch);
/* no more submissions: waits for every job and copies the images out */
int mt_batch_finish(mt_batch* batch);
int mt_batch_get_job_info(const mt_batch* batch, uint32_t job_id, mt_batch_job_info* info);
int mt_batch_get_stats(const mt_batch* batch, mt_batch_stats* stats);
int mt_batch_destroy(mt_batch* batch);
/* Picks mode, paths_per_iteration, tile_pixels, trace_blocks_per_cu (graphs), iteration_tiles
 * and streams by coordinate descent over measured runs of the n calibration jobs.  Each trial
 * creates a batch, serves the jobs once at 1 sample per pixel (graph captures, first touches;
 * untimed), then submits them as given, all at once, and times that round to its finish.
 * trials (may be NULL) receives up to max_trials measurements. */
int mt_batch_autotune(mt_scene* const* scenes, const mt_camera_desc* cameras,
                      const mt_settings* settings, uint32_t n, mt_batch_options* best,
                      mt_autotune_trial* trials, uint32_t max_trials, uint32_t* n_trials);

#ifdef __cplusplus
}
#endif
#endif /* MITRACE_GPU_H */
