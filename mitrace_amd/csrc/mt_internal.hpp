// mt_internal.hpp -- internal declarations of the gfx950 Tracer library (not part of the ABI).
//
// HBM layout of a scene (all arrays device-resident, built once by mt_scene_create):
//   instances  DevInstance[I]  192 B: inverse and model matrices, world AABB, mesh/material ids,
//                                     the mesh's shading-attribute offsets
//   meshes     DevMesh[M]       64 B: root box and word, offsets of the mesh's pair records,
//                                     leaf-ordered triangles and shading attributes, flags
//   pairs      float4[4*Q]      64 B per sibling pair of BVH nodes (the reference keeps siblings
//                                     adjacent, BVH.cpp:223-259): {A.min, A.word} {A.max, B.word}
//                                     {B.min, 0} {B.max, 0}; word = child pair index of an inner
//                                     node, or first leaf slot | kLeafBit; popping a node needs
//                                     only the record of its children
//   prims      float4[3*P]      48 B per leaf slot: {v0.xyz, triangle id} {e1.xyz, leaf node}
//                                     {e2.xyz, last-in-leaf flag}, in BVH::indices_ order, so a
//                                     leaf's triangles are contiguous (straddling duplicates kept)
//   normals    float[9*T]       per-corner normals of every triangle (soup order)
//   uvs        float[6*T]       per-corner TEXCOORD_0 (meshes without UVs have none)
//   materials  DevMaterial[]    factors and texture slots
//   texels     uint32[]         RGBA8 texels of all images; images DevImage[] {offset, w, h}
//   env        float4[W*H]      equirect texels decoded to float (u8/255 or float input)
//   lights     DevLight[]
//   fast_meshes DevMesh[M]      (scenes built with mt_bvh_options.fast) the same records for
//                                     per-mesh SAH BVHs, whose pair records and prim slots are
//                                     appended to pairs/prims: tight boxes, no duplicates
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "../../include/mitrace_gpu.h"

namespace mt {

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMeshHasBvh = 1u;
constexpr uint32_t kMeshHasUv = 2u;
constexpr uint32_t kTlasReference = 1u, kTlasFast = 2u;  // DevScene::use_tlas bits
constexpr uint32_t kPrimLast = 1u;  // prims[3i+2].w: last slot of its leaf

struct alignas(16) DevInstance {
    float inv[16];  // column-major inverse of the model matrix (glm::inverse, Intersect.cpp:152)
    float M[16];    // column-major model matrix
    float wmin[3];
    uint32_t mesh;
    float wmax[3];
    uint32_t material;
    // the mesh's shading-attribute offsets and flags (DevMesh::attr_base, uv_base, flags),
    // repeated here so that shading reaches a hit's normals in one dependent load fewer
    uint32_t attr_base;
    uint32_t uv_base;
    uint32_t mesh_flags;
    uint32_t pad[5];
};
static_assert(sizeof(DevInstance) == 192, "DevInstance layout");

struct alignas(16) DevMesh {
    float root_min[3];
    uint32_t root_word;  // the root node's word (pair index or leaf slot | kLeafBit)
    float root_max[3];
    uint32_t flags;      // kMeshHas*
    uint32_t pair_base;  // first sibling-pair record of this mesh
    uint32_t prim_base;  // first leaf slot in the global prim array
    uint32_t prim_count; // leaf slots (= triangles for a linear mesh)
    uint32_t attr_base;  // first triangle of this mesh in the normals array
    uint32_t tri_count;
    uint32_t uv_base;    // first triangle of this mesh in the uvs array
    uint32_t pad[2];
};
static_assert(sizeof(DevMesh) == 64, "DevMesh layout");

struct DevTex {
    int32_t image;  // -1 = none
    int32_t filter;
    float ox, oy, sx, sy;
    float pad[2];
};

struct alignas(16) DevMaterial {
    float base[4];
    float emissive[3];
    float normal_scale;
    float metallic, roughness, pad0, pad1;
    DevTex tex[4];  // base colour, metallic-roughness, normal, emissive
};

struct DevImage {
    uint32_t offset, w, h, pad;
};

struct alignas(16) DevLight {
    float pos[3];
    int32_t type;
    float color[3];
    float size;
    float dir[3];
    float inner;
    float outer;
    float intensity;  // PrepareLights (Tracer.cpp:123-132)
    float pad[2];
};

// device pointers and sizes of an uploaded scene, passed to kernels by value
struct DevScene {
    const DevInstance* instances;
    const DevMesh* meshes;
    const float4* pairs;
    const float4* prims;
    const float* normals;
    const float* uvs;
    const DevMaterial* materials;
    const uint32_t* texels;
    const DevImage* images;
    const float4* env;
    const DevLight* lights;
    uint32_t instance_count;
    uint32_t light_count;
    // top-level BVH over instance world boxes (used when instance_count > 64): pair records
    // like the meshes', leaf word = kLeafBit | count << 24 | first into tlas_inst
    const DevMesh* fast_meshes;  // per mesh: the SAH BVH of the fast traversal (or null)
    const float4* tlas_pairs;
    const uint32_t* tlas_inst;
    float tlas_root_min[3], tlas_root_max[3];
    uint32_t tlas_root_word;
    uint32_t use_tlas;  // kTlasReference | kTlasFast: the traversals that start at the top level
    uint32_t env_w, env_h;
    int32_t env_filter;
    int32_t has_env;
};

// host-side product of the BVH build for one mesh (kept for mt_scene_mesh_bvh)
struct HostNode {
    float mn[3], mx[3];
    uint32_t a, b;
};
struct HostBvh {
    std::vector<HostNode> nodes;
    std::vector<uint32_t> indices;
    uint32_t depth = 0;  // levels of the flattened tree
};

// BVH build with the reference's rules (BVH.cpp:13-282); positions is the mesh soup (9 floats
// per triangle).  max_depth is the value Mesh::Mesh passes (bvhMaxDepth - 1).
void build_bvh(const float* positions, uint32_t tri_count, uint32_t max_triangles,
               uint32_t max_depth, HostBvh& out);
// the same tree built on the current GPU (mt_bvh_build.hip); bit-identical to build_bvh.
// With recs, the mesh's pair records and leaf slots in the scene's device format (see the
// layout above; words relative to the mesh's pair/prim base) are built on the device too and
// handed to the caller (free with free_blocks).
struct DeviceBvhRecords {
    float4* pairs;
    float4* prims;
    uint64_t npairs, nslots;
};
int build_bvh_device(const float* positions, uint32_t tri_count, uint32_t max_triangles,
                     uint32_t max_depth, HostBvh& out, DeviceBvhRecords* recs = nullptr);
// the fast traversal's tree of one mesh (more than 4 triangles) built on the device as a linear
// BVH: records in recs (owned by the caller, mesh-local words), root box/word and flags in dm,
// levels = the tree's depth as build_fast_bvh counts it
int build_fast_bvh_device(const float* positions, uint32_t tri_count, DevMesh& dm, DeviceBvhRecords& recs,
                          uint32_t& levels);
// glm::inverse stand-in; column-major in and out
void invert4(const float* m, float* out);
// M * (x, y, z, w) with the library's fixed operation order
void mul4(const float* m, const float* v, float* out);

}  // namespace mt

// the opaque handle behind mt_scene*
struct mt_scene {
    int device = 0;
    mt::DevScene dev{};
    std::vector<void*> allocations;
    std::vector<mt::HostBvh> bvhs;  // per mesh; empty nodes = linear scan
    std::vector<uint32_t> mesh_has_bvh;
    uint32_t max_bvh_depth = 0;
    uint32_t tlas_depth = 0;
    uint32_t max_fast_depth = 0;  // levels of the deepest fast-traversal BVH (0 = none built)
    uint32_t mesh_count = 0;
    uint32_t instance_count = 0;
    uint32_t light_count = 0;
    uint32_t material_count = 0;
    float world_lo[3] = {0, 0, 0}, world_hi[3] = {0, 0, 0};  // union of the instance boxes
    uint64_t prim_slots = 0;        // leaf slots (48 bytes) of the reference-rule trees
    uint64_t node_count = 0;        // nodes of the reference-rule trees (64 bytes per pair)
    uint64_t fast_record_bytes = 0; // pair and slot records of the fast traversal's trees
    uint64_t triangle_count = 0;
    ~mt_scene();
};

namespace mt {
// error reporting (thread-local last error)
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
// Every device allocation of the library goes through device_malloc: hipMalloc, and when it
// fails, the pooled path buffers of finished renders are handed back to the device and the
// allocation is tried once more.  free_blocks frees such blocks (pooled path buffers included,
// whose bookkeeping it clears) and empties the vector.
int device_malloc(void** p, size_t bytes);
void free_blocks(std::vector<void*>& blocks);
int upload_scene(const mt_scene_desc* desc, mt_scene* scene);
// top-level BVH over boxes (6 floats each): SAH over 12 bins per axis, up to 4 boxes a leaf.
// pairs/indices use the device format of DevScene::tlas_*; returns the number of levels.
// SAH BVH of one mesh for the fast traversal: tight boxes, every triangle in exactly one leaf
// of at most 4; appends its pair records and prim slots (formats of DevScene::pairs/prims) and
// fills dm's root, pair_base, prim_base, prim_count and flags.  Returns the number of levels.
uint32_t build_fast_bvh(const float* positions, uint32_t tri_count, std::vector<float4>& pairs,
                        std::vector<float4>& prims, DevMesh& dm);
uint32_t build_tlas(const std::vector<float>& boxes, std::vector<float4>& pairs,
                    std::vector<uint32_t>& indices, float root_min[3], float root_max[3],
                    uint32_t& root_word);
}  // namespace mt
