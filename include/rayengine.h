/*
 * rayengine.h - C ABI of the MI355X (gfx950) ray-query library.
 *
 * Plain C: POD structs, raw pointers and sizes; no HIP or torch types cross
 * this boundary.  Every call returns an re_status and never exits the
 * process; re_last_error() describes the last failure.  A context is not
 * thread-safe (one context per host thread, or external locking).  Each
 * entry point is stream-ordered on the context's stream and returns after
 * the work is complete unless RE_ASYNC is passed in flags.
 *
 * Each declaration names the reference interface it replaces
 * (hadalhw17/RayEngine, paths relative to the repository root).
 */
#ifndef RAYENGINE_H
#define RAYENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RE_API __attribute__((visibility("default")))

typedef enum re_status {
  RE_OK = 0,
  RE_ERR_INVALID = 1,     /* bad argument, shape or id */
  RE_ERR_HIP = 2,         /* HIP runtime failure */
  RE_ERR_OOM = 3,         /* device or host allocation failed */
  RE_ERR_NOT_FOUND = 4,   /* unknown scene / volume / network id */
  RE_ERR_UNSUPPORTED = 5, /* no gfx950 device, or a feature not built */
  RE_ERR_IO = 6           /* file missing, short or malformed */
} re_status;

/* flags */
#define RE_DEVICE_PTRS 0x1u /* ray and result pointers are device memory */
#define RE_ASYNC 0x2u       /* return without synchronising the stream */

typedef struct re_ctx re_ctx;
typedef struct re_kdtree re_kdtree;

/* ---------------------------------------------------------------- context */

/* Replaces the reference's TU-global device state
 * (RayEngine/src/cuda-src/CUDARayTracing.cuh:15-85). */
RE_API re_status re_create(int device, re_ctx** out);
/* Replaces free_memory (RayEngine/src/cuda-src/sphere_tracing.cuh:757-767). */
RE_API void re_destroy(re_ctx* ctx);
/* Replaces gpuErrchk/gpuAssert (RayEngine/src/RayEngine/RayEngine.h:56-63):
 * errors are reported, not fatal.  ctx may be NULL (last error of a call that
 * had no context). */
RE_API const char* re_last_error(const re_ctx* ctx);
/* Library version string. */
RE_API const char* re_version(void);

/* --------------------------------------------- host scene preparation */

/* Subdivided icosphere, outward counter-clockwise triangles, as a soup of
 * 9 floats per triangle (v0, v1, v2).  n_tris = 20 * 4^subdiv.  Called with
 * out == NULL it only reports the count. */
RE_API re_status re_gen_icosphere(int subdiv, float radius, float* out, int64_t* n_tris);

/* n random triangles: centroid ~ U[lo,hi]^3, two edges of length
 * ~ U[len_lo,len_hi] along random unit directions (PCG32, seed). */
RE_API re_status re_gen_random_tris(uint64_t seed, int64_t n, float lo, float hi, float len_lo,
                                    float len_hi, float* out);

/* n triangles in n_clusters Gaussian clusters: cluster centre ~ U[lo,hi]^3,
 * triangle centroid ~ centre + N(0, sigma^2 I), edge lengths
 * ~ LogNormal(ln_mu, ln_sigma) along random unit directions (PCG32, seed). */
RE_API re_status re_gen_cluster_soup(uint64_t seed, int64_t n, int n_clusters, float lo, float hi,
                                     float sigma, float ln_mu, float ln_sigma, float* out);

/* Pinhole camera, reference convention.  RCamera
 * (RayEngine/src/Primitives/Camera.h:7-16) and generate_ray
 * (RayEngine/src/cuda-src/ray_functions.cuh:13-55). */
typedef struct re_camera {
  float pos[3];
  float view[3];    /* viewing direction (normalised by the ray generator) */
  float camdown[3]; /* reference default (0,-1,0) */
  float fov[2];     /* full field of view in degrees */
  float focal_distance;
} re_camera;

/* width*height primary rays, pixel (x, y) at index y*width + x; origins and
 * directions as 3 floats per ray. */
RE_API re_status re_gen_camera_rays(const re_camera* cam, int width, int height, float* ray_o,
                                    float* ray_d);

/* One cosine-weighted diffuse bounce ray per hit (prim_id >= 0): origin =
 * hit point + 1e-4 * n, n the geometric normal facing the incoming ray;
 * direction sampled about n (PCG32 stream (seed, ray index)).  tri_v is the
 * triangle soup the prim ids index.  Writes at most n rays and returns the
 * count in n_out. */
RE_API re_status re_gen_bounce_rays(const float* tri_v, int64_t n_tris, const float* ray_o,
                                    const float* ray_d, const int32_t* prim_id, const float* t,
                                    int64_t n, uint64_t seed, float* out_o, float* out_d,
                                    int64_t* n_out);

/* ------------------------------------------------------ kd-tree build */

/* Replaces RKDTreeCPU (RayEngine/src/Primitives/KDTree.h:39-99,
 * KDTree.cpp:117-834) and its flattening RKDThreeGPU
 * (RayEngine/src/Primitives/KDThreeGPU.cpp:155-194): binned SAH build with
 * ropes (Havran/Popov), split boxes on leaves, six ropes per leaf. */
typedef struct re_kd_build_params {
  int32_t max_leaf_tris;  /* leaf when a node holds <= this many (default 8) */
  int32_t max_depth;      /* leaf at this depth (root depth 1; default 24) */
  float cost_traversal;   /* SAH traversal cost (default 1.0) */
  float cost_intersect;   /* SAH intersection cost (default 1.5) */
  float empty_bonus;      /* SAH factor for a split with an empty side (default 0.8) */
  int32_t n_bins;         /* bins per axis for large nodes (default 32) */
  int32_t clip;           /* 1: a triangle straddling a split goes only to the children its
                             clipped polygon reaches, with the clipped box ("perfect
                             splits"); 0: by its bounding box (reference rule,
                             KDTree.cpp:581-600).  Default 1. */
  int32_t split_rule;     /* 0: SAH (default); 1: the reference's spatial median, the
                             midpoint lo + (hi - lo) / 2 of the node box's longest side
                             (KDTree.cpp:192-199,368-389), with exact split boxes for the
                             children instead of the reference's boxes grown by +-1e-4.
                             Host build only (re_kd_build_gpu refuses 1). */
} re_kd_build_params;

typedef struct re_kd_info {
  int64_t n_tris;    /* input triangles */
  int64_t n_nodes;   /* 8-byte node entries */
  int64_t n_leaves;  /* 64-byte leaf records */
  int64_t n_trirecs; /* 40-byte triangle records (with straddler copies) */
  int32_t depth;     /* deepest leaf, root = 1 */
  int32_t max_leaf;  /* largest leaf population */
  float root_min[3];
  float root_max[3];
} re_kd_info;

/* Fills p with the defaults above. */
RE_API void re_kd_default_params(re_kd_build_params* p);
/* Builds a tree over the soup tri_v (9 floats per triangle; prim ids are
 * 0..n_tris-1).  params may be NULL. */
RE_API re_status re_kd_build(const float* tri_v, int64_t n_tris, const re_kd_build_params* params,
                             re_kdtree** out);
RE_API void re_kd_free(re_kdtree* tree);
RE_API re_status re_kd_get_info(const re_kdtree* tree, re_kd_info* info);
/* Copies the flattened arrays out (sizes from re_kd_get_info): nodes
 * n_nodes*8 bytes, leaves n_leaves*64 bytes, tris n_trirecs*40 bytes.  Any
 * pointer may be NULL to skip it. */
RE_API re_status re_kd_export(const re_kdtree* tree, void* nodes, void* leaves, void* tris);
/* Builds the same tree as re_kd_build on the context's GPU: breadth first,
 * one wave per node choosing its split plane, references partitioned (and
 * straddlers clipped) in parallel, then preorder numbering, ropes and the
 * flat arrays on the device.  Replaces the host build the reference runs
 * before every upload (RKDTreeCPU::build / buildRopeStructure /
 * optimizeRopes, RayEngine/src/Primitives/KDTree.cpp:335-834, timed at
 * RTScene.cpp:186-198) and the flattening (RKDThreeGPU::buildTree,
 * KDThreeGPU.cpp:155-194).  The result is a tree object identical to
 * re_kd_build's for the same input and parameters (n_bins at most 64);
 * tri_v is host memory.  device_ms (nullable) gets the device time of the
 * build, upload of the soup and download of the arrays excluded. */
RE_API re_status re_kd_build_gpu(re_ctx* ctx, const float* tri_v, int64_t n_tris, const re_kd_build_params* params,
                                 re_kdtree** out, float* device_ms);

/* ------------------------------------------------------ scene upload */

/* One traced object: a tree plus its transform.  GPUSceneObject
 * (RayEngine/src/cuda-src/gpu_structs.h:69-90); rotation enters as the
 * sine and cosine of rotation.y, evaluated by the caller. */
typedef struct re_object {
  const re_kdtree* tree;
  float location[3];
  float sin_rot_y;
  float cos_rot_y;
  int32_t is_character;
  int32_t prim_offset; /* added to the tree's prim ids in the results */
} re_object;

/* Replaces copy_memory (RayEngine/src/cuda-src/cuda_memory_functions.cuh:58-211)
 * and update_objects (:50-56): concatenates the objects' trees and uploads
 * them.  Host data is copied; the context owns the device copy. */
RE_API re_status re_scene_upload(re_ctx* ctx, const re_object* objs, int32_t n_obj,
                                 int32_t* scene_id);
RE_API re_status re_scene_release(re_ctx* ctx, int32_t scene_id);
/* Rewrites the objects' transforms (location, rotation, is_character) of a
 * resident scene; trees and prim offsets stay.  Replaces update_objects
 * (RayEngine/src/cuda-src/cuda_memory_functions.cuh:50-56), called per tick
 * by RScene::Tick (RayEngine/src/World/Scene.cpp:79-91). */
RE_API re_status re_scene_update_objects(re_ctx* ctx, int32_t scene_id, const re_object* objs,
                                         int32_t n_obj);
/* Device footprint of a scene in bytes. */
RE_API re_status re_scene_bytes(re_ctx* ctx, int32_t scene_id, int64_t* bytes);
/* sizes[0..3] = node entries, leaf records, triangle records, objects. */
RE_API re_status re_scene_info(re_ctx* ctx, int32_t scene_id, int64_t sizes[4]);
/* Copies the resident scene back to host memory (sizes from re_scene_info;
 * objects are 52-byte records {location[3], sin_y, cos_y, is_character,
 * root entry, root_min[3], root_max[3]}).  Any pointer may be NULL.  The
 * reference's tree dump, RKDThreeGPU::PrintCPUAndGPUTrees
 * (RayEngine/src/Primitives/KDThreeGPU.cpp:137-153), read the same data. */
RE_API re_status re_scene_export(re_ctx* ctx, int32_t scene_id, void* nodes, void* leaves,
                                 void* tris, void* objs);

/* ------------------------------------------------------ closest hit */

typedef struct re_trace_params {
  int32_t max_leaf_visits; /* reference cap 50 (ray_functions.cuh:284); <= 0 = unlimited */
  int32_t one_sided;       /* 1: reference one-sided test (ray_functions.cuh:182) */
  int32_t secondary;       /* is_secondary of stackless_kdtree_traversal */
  int32_t block;           /* threads per block, 0 = default (64; 256 for variant 1) */
  int32_t variant;         /* 0: one ray per thread, blocks in ray order;
                              1: persistent waves with per-lane ray refill;
                              2: one ray per thread, last block of rays first;
                              3: as 2 with a 6-wave register budget;
                              4: as 3, each leaf visit done by the whole wave:
                                 the triangles of all lanes' leaves fetched as
                                 one coalesced list and tested one triangle
                                 per lane, or, when every visiting lane is at
                                 the same leaf, staged once in LDS (default);
                              5: as 4 in persistent waves: once half a wave's
                                 lanes have finished their rays, they take the
                                 next rays at the next leaf round (rays taken
                                 from the end of the array); faster than 4 on
                                 incoherent rays.
                              Results are identical for every variant. */
} re_trace_params;

RE_API void re_trace_default_params(re_trace_params* p);

/* Closest hit for n rays (origins/directions 3 floats per ray).  Replaces
 * trace_primary_rays (RayEngine/src/cuda-src/renderer.cu:389-410) over
 * trace_scene / stackless_kdtree_traversal
 * (RayEngine/src/cuda-src/ray_functions.cuh:230-396).  Misses give
 * prim_id = -1 and t = 1e20; obj_id may be NULL. */
RE_API re_status re_trace_closest(re_ctx* ctx, int32_t scene_id, const float* ray_o,
                                  const float* ray_d, int64_t n, const re_trace_params* params,
                                  int32_t* prim_id, float* t, int32_t* obj_id, uint32_t flags);
/* Closest hit with walk-start hints and hit leaves, for device-resident
 * pipelines that trace secondary rays from the hits of an earlier launch.
 * No counterpart in the reference, whose walks always start at the root
 * (RayEngine/src/cuda-src/ray_functions.cuh:273-291).  As re_trace_closest,
 * and:
 *   start_leaf (may be NULL): per ray, a leaf id of the scene (as hit_leaf
 *     returns them) or -1.  A walk of an object starts at that leaf instead
 *     of descending from the root when the leaf belongs to the object and
 *     holds the walk's first point on the same side of every split as the
 *     descent would take it (bmin <= p < bmax on each axis); otherwise, and
 *     for ids out of range, at the root.  Results are those of
 *     re_trace_closest for any hints; only the work differs.
 *   hit_leaf (may be NULL): per ray, the last leaf visited by the walk
 *     that found the closest hit, -1 for a miss.  The walk goes on past a
 *     taken hit until its entry point passes it, so that leaf's cell holds
 *     the hit point (unless the walk ended at the visit cap or a missing
 *     rope): a secondary ray leaving the hit point passes it as its
 *     start_leaf.
 *   counts (may be NULL): host array of 4, the summed work counters as
 *     re_trace_counts returns them.
 * Variants 4 and 5 only (RE_ERR_INVALID otherwise); all arrays are device
 * pointers (RE_DEVICE_PTRS is required). */
RE_API re_status re_trace_closest_hinted(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d,
                                         int64_t n, const re_trace_params* params, const int32_t* start_leaf,
                                         int32_t* prim_id, float* t, int32_t* obj_id, int32_t* hit_leaf,
                                         uint64_t counts[4], uint32_t flags);
/* Bounds-checked debug build (librayengine_debug.so, built with
 * RE_DEBUG_BOUNDS): the closest-hit, any-hit and status kernels check every
 * node entry, leaf record and triangle run index against the scene's arrays
 * before reading it; an index outside sets a bit (1 node entry, 2 leaf
 * record, 4 triangle run) and the read goes to index 0 instead.  This call
 * returns the bits set since the last call (after the context's stream has
 * finished) and clears them; RE_ERR_UNSUPPORTED in the normal build.  No
 * counterpart in the reference. */
RE_API re_status re_debug_bounds_errors(re_ctx* ctx, uint32_t* mask);
/* Closest hit with a status byte per ray, a debugging aid with no
 * counterpart in the reference (whose walk ends silently,
 * ray_functions.cuh:284-347): prim_id and t as re_trace_closest, and in
 * status how the walk of the last object ended - 0 t_entry >= t_exit (the
 * normal end), 1 the leaf-visit cap, 2 a leaf slab test missed, 3 no exit
 * face within the 1e-4 tolerance, 4 the exit rope is -1 (left the root box),
 * 5 the root box was missed, 6 the walk's state repeated (stopped, its
 * result final) - plus 0x80 when the ray's origin or direction holds a NaN or
 * an infinity or the direction is zero.  One ray per thread, no tuning. */
RE_API re_status re_trace_status(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d,
                                 int64_t n, const re_trace_params* params, int32_t* prim_id, float* t,
                                 uint8_t* status, uint32_t flags);

/* Per-triangle hit attributes of a scene, the per-vertex data of
 * copy_memory's normals and uvs (RayEngine/src/cuda-src/cuda_memory_functions.cuh:58-211;
 * packed per triangle as RTScene::init_triangles does, RayEngine/src/World/RTScene.cpp:243-257):
 * normals holds n0, n1, n2 (9 floats) per prim id, uvs uv0, uv1, uv2 (6
 * floats) per prim id or NULL; n_prims covers every prim id of the scene.
 * Host arrays, copied. */
RE_API re_status re_scene_set_attributes(re_ctx* ctx, int32_t scene_id, const float* normals, const float* uvs,
                                         int64_t n_prims);
/* Hit attributes of traced rays (prim_id and obj_id as re_trace_closest
 * returned them; obj_id may be NULL for a one-object scene; params: the
 * trace's one_sided and secondary): normal (3 floats per ray) = the vertex
 * normals interpolated at the hit and normalised (gpu_get_tri_normal,
 * RayEngine/src/cuda-src/ray_functions.cuh:146-155); uv (2 per ray) = the
 * vertex uvs interpolated (get_uvs, :157-166), or the barycentrics when the
 * scene has no uvs (:206-213); bary (2 per ray) = the barycentrics (u, v) of
 * the hit; hit_point (3 per ray) = ro + t rd in object space (:204).  Every
 * output may be NULL.  Misses get zeros. */
RE_API re_status re_hit_attributes(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                                   const re_trace_params* params, const int32_t* prim_id, const int32_t* obj_id,
                                   float* normal, float* uv, float* bary, float* hit_point, uint32_t flags);

/* re_trace_closest that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * work counters (counts, may be NULL), the persistent variant's ray counter
 * and the staging of host-pointer rays and results live in a fixed device
 * area of the library (about 9 MB per GPU, held until the process ends),
 * cleared by a kernel on every call; rays and results are the caller's.
 * Host pointers: at most re_trace_closest_noalloc_max_rays() rays
 * (RE_ERR_INVALID beyond); device pointers (RE_DEVICE_PTRS): any n.
 * Results equal re_trace_closest's (and counts re_trace_counts').  Every call
 * on one GPU shares that GPU's area: calls that overlap on one GPU (two
 * contexts or two host threads) race. */
RE_API re_status re_trace_closest_noalloc(re_ctx* ctx, int32_t scene_id, const float* ray_o,
                                          const float* ray_d, int64_t n, const re_trace_params* params,
                                          int32_t* prim_id, float* t, int32_t* obj_id, uint64_t counts[4],
                                          uint32_t flags);
RE_API int64_t re_trace_closest_noalloc_max_rays(void);

/* Same traversal, also summing its work counters: counts[0] node entries
 * read, counts[1] leaf records read, counts[2] triangle records tested,
 * counts[3] rays with a hit.  These give the algorithmic bytes of a launch:
 * 24 + 8 per ray, 8 per entry, 56 per leaf, 40 per triangle. */
RE_API re_status re_trace_counts(re_ctx* ctx, int32_t scene_id, const float* ray_o,
                                 const float* ray_d, int64_t n, const re_trace_params* params,
                                 int32_t* prim_id, float* t, uint64_t counts[4], uint32_t flags);

/* Any-hit (shadow) query.  Replaces trace_shadow
 * (RayEngine/src/cuda-src/ray_functions.cuh:365-374), whose callers
 * (phong_light / shade, RayEngine/src/cuda-src/material_functions.cuh:68-119)
 * use only its `hits` flag: occluded[i] = 1 when the closest-hit walk of some
 * object detects an intersection (same walk, visit cap and acceptance rule as
 * re_trace_closest; a walk stops at its first accepted hit).  params NULL
 * means the defaults with secondary = 1, as trace_shadow traverses. */
RE_API re_status re_trace_occluded(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d,
                                   int64_t n, const re_trace_params* params, uint8_t* occluded, uint32_t flags);

/* ------------------------------------------------------ terrain volume */

/* Doubled permutation table (512 entries) of RPerlinNoise(seed)
 * (RayEngine/src/World/PerlinNoise.h:16-52; reference default seed 2016). */
RE_API re_status re_perlin_table(uint32_t seed, uint32_t* perm);

/* Terrain field f(p) = (p.y - y_offset) - sum_{o<octaves} gain^o *
 * noise(p * base_freq * lacunarity^o) sampled at voxel (x,y,z) -> p =
 * (x,y,z) * spacing.  noise is eval (RayEngine/src/cuda-src/sdf_functions.cuh:83-135)
 * with hash P[P[P[x]+y]+z]. */
typedef struct re_noise_params {
  int32_t dim[3];     /* default 256^3 */
  float spacing[3];   /* default 1/256 (extent 1) */
  uint32_t seed;      /* permutation seed, default 2016 */
  int32_t octaves;    /* default 5 */
  float lacunarity;   /* default 2.0 */
  float gain;         /* default 0.5 */
  float base_freq;    /* default 4.0 (= 4 / extent) */
  float y_offset;     /* default 0.4 */
  float origin[3];    /* field position of voxel (0,0,0), default 0: voxel (x,y,z) samples the
                         field at origin + (x,y,z) * spacing.  cuda_update_chunk_gen passes the
                         chunk's location here (renderer.cu:982-1000) */
} re_noise_params;

RE_API void re_noise_default_params(re_noise_params* p);
/* Generates a volume on the device (bricked 8^3, Morton order inside a
 * brick).  Replaces generate_noise / cuda_update_chunk_gen
 * (RayEngine/src/cuda-src/renderer.cu:105-126,812-836,982-1005). */
RE_API re_status re_noise_generate(re_ctx* ctx, const re_noise_params* params, int32_t* vol_id);
/* Uploads a linear x-fastest grid (index x + dim.x*(y + dim.y*z)).
 * Replaces initialize_volume_render's grid upload
 * (RayEngine/src/cuda-src/sphere_tracing.cuh:617-755). */
RE_API re_status re_volume_upload(re_ctx* ctx, const int32_t* dim, const float* spacing,
                                  const float* linear, int32_t* vol_id);
/* Uploads a chunk's grid as linear x-fastest (distance, material) float
 * pairs, 2 * dim.x*dim.y*dim.z floats, keeping the material channel.
 * Replaces cuda_update_chunk's upload of the chunk's Grid voxels
 * (RayEngine/src/cuda-src/sphere_tracing.cuh:565-615).  Box [0, dim * spacing]. */
RE_API re_status re_volume_upload_pairs(re_ctx* ctx, const int32_t* dim, const float* spacing,
                                        const float* pairs, int32_t* vol_id);
/* Reads a volume back as a linear x-fastest grid (save_map's D2H copy,
 * RayEngine/src/cuda-src/renderer.cu:838-870). */
RE_API re_status re_volume_read(re_ctx* ctx, int32_t vol_id, float* linear);
RE_API re_status re_volume_info(re_ctx* ctx, int32_t vol_id, int32_t* dim, float* spacing,
                                int64_t* bytes);
RE_API re_status re_volume_release(re_ctx* ctx, int32_t vol_id);

/* ---- frame output -------------------------------------------------------- */

typedef struct re_frame_params {
  float sun_dir[3];         /* Atmosphere::sunDirection, towards the sun; default (0, 1, 0), the
                               direction Craze sets at angle 0 (renderer.cu:290-300) */
  float earth_radius;       /* Atmosphere defaults (Atmosphere.cuh): 6360e3 */
  float atmosphere_radius;  /* 6420e3 */
  float hr;                 /* Rayleigh scale height, 79945 as in the reference */
  float hm;                 /* Mie scale height, 1200 */
  float sky_origin[3];      /* viewpoint of the sky, (0, earth_radius + 10000, 300000) (renderer.cu:262) */
  float base_color[3];      /* colour of hit pixels before simple_shade, default 0 */
  int32_t gamma;            /* RenderingSettings.gamma, default 0 */
  int32_t vignetting;       /* RenderingSettings.vignetting, default 0 */
  float vignetting_k;       /* default 0.25 */
} re_frame_params;

RE_API void re_frame_default_params(re_frame_params* p);
/* Colours a width x height frame (pixel index y * width + x) from its rays
 * and hits, the per-pixel tail of render_sphere_trace
 * (RayEngine/src/cuda-src/renderer.cu:257-283): a missed pixel gets the
 * single-scattering sky (compute_incident_light, material_functions.cuh:
 * 295-344; 32 view and 8 light samples, the reference's expression kept)
 * and its tone curve; a hit pixel gets the normal-visualisation material
 * (simple_shade, cuda_helper_functions.h:137-140) on base_color; then
 * optional gamma and vignetting, clipping to [0, 1] and packing to RGBA8
 * (rgbaFloatToInt, cuda_helper_functions.h:60-68: R in the low byte, alpha
 * 255).  ray_d: 3 floats per pixel; hit: 0/1 per pixel; normal: 3 floats
 * per pixel (read for hits only).  Host pointers, or device pointers with
 * RE_DEVICE_PTRS. */
RE_API re_status re_frame_shade(re_ctx* ctx, const float* ray_d, const uint8_t* hit, const float* normal,
                                int32_t width, int32_t height, const re_frame_params* params, uint32_t* rgba,
                                uint32_t flags);
/* Writes RGBA8 pixels (as re_frame_shade packs them) as a binary PPM (P6,
 * RGB, rows top to bottom), the headless counterpart of the reference's
 * window output. */
RE_API re_status re_write_ppm(const char* path, const uint32_t* rgba, int32_t width, int32_t height);

/* ---- .rsdf files and SDF edits ---------------------------------------- */

/* Writes a volume as a .rsdf file: int32 dim[3], float spacing[3], then
 * dim.x*dim.y*dim.z pairs (distance, material) of float32, x fastest
 * (save_map, RayEngine/src/cuda-src/renderer.cu:838-871; the reference
 * indexes its array x + dim.y*(y + dim.x*z), which equals x + dim.x*(y +
 * dim.y*z) for the cubic grids it writes).  Material is 0 for volumes that
 * never had one (the noise volumes). */
RE_API re_status re_volume_save_rsdf(re_ctx* ctx, int32_t vol_id, const char* path);
/* Reads a .rsdf file into a new volume with its material channel (Grid::
 * load_volume_floam_file, RayEngine/src/World/Grid.cpp:23-64, and load_map,
 * renderer.cu:873-947).  As in load_map the volume box is [0, (dim - 1) *
 * spacing], the point of the last voxel.  RE_ERR_IO if the file is missing,
 * short or longer than its header says. */
RE_API re_status re_volume_load_rsdf(re_ctx* ctx, const char* path, int32_t* vol_id);
/* Checks a .rsdf file without a context or a device: RE_OK with its dims and
 * spacing when the header is valid (dims in [2,4096], spacing > 0) and the
 * file holds exactly the 24-byte header and dim.x * dim.y * dim.z
 * (distance, material) pairs; RE_ERR_IO otherwise.  The header is read as
 * load_map reads it (renderer.cu:873-890); nothing of the size it claims is
 * allocated before the file's length confirms it. */
RE_API re_status re_rsdf_info(const char* path, int32_t dim[3], float spacing[3]);
/* Reads the material channel back as a linear x-fastest grid (zeros if the
 * volume has none). */
RE_API re_status re_volume_read_material(re_ctx* ctx, int32_t vol_id, float* linear);

typedef enum re_brush_type {
  RE_BRUSH_SPHERE_ADD = 0, /* TerrainBrushType, RayEngine/src/RayEngine/RayEngine.h:77-83 */
  RE_BRUSH_SPHERE_SUBTRACT = 1,
  RE_BRUSH_CUBE_ADD = 2,
  RE_BRUSH_CUBE_SUBTRACT = 3
} re_brush_type;

typedef struct re_brush {
  int32_t type;         /* re_brush_type */
  float center[3];      /* brush position in volume coordinates (the brush ray's hit point) */
  float extent[3];      /* sphere radius = extent[0]; cube half extents (the reference's int3, as floats) */
  float material;       /* material written where the brush wins (material_index) */
  int32_t snap_to_grid; /* snap x and z of the voxel-to-brush offset to multiples of extent[0] */
  float smooth_k;       /* smooth-minimum width of SPHERE_ADD; the reference's 0.1 */
} re_brush;

RE_API void re_brush_default(re_brush* b);
/* Applies a brush to every voxel of a volume, distance and material
 * (insert_sphere_to_texture, renderer.cu:30-92, with sdf_smin / sdf_fmaxf /
 * fminf, sdf_functions.cuh:270-289, and sphere_distance / aabb_distance,
 * :422-437).  The reference traces the brush ray in every thread to find the
 * hit point; here the caller passes that point (re_sdf_trace gives it).
 * Adds a material channel (zeros) to a volume that has none. */
RE_API re_status re_volume_edit(re_ctx* ctx, int32_t vol_id, const re_brush* brush);

typedef struct re_sdf_params {
  int32_t max_steps; /* default 256 */
  float eps;         /* hit when d <= eps * t; default 1e-3 (renderer.cu:210-211) */
  float step_scale;  /* t += d * step_scale; default 0.8 */
  int32_t step_rule; /* 0: step_scale; 1: reference clamp(0.6 t, 0.45, 0.8) */
  int32_t block;     /* threads per block, 0 = default */
} re_sdf_params;

RE_API void re_sdf_default_params(re_sdf_params* p);
/* Sphere traces n rays through the volume (clipped to its box [0, dim *
 * spacing]), trilinear samples.  t = post-step distance of the hit (1e20 on
 * a miss), hit = 0/1; steps (nullable) = samples taken; normal (nullable, 3
 * floats per ray) = forward-difference normal at the hit.  Replaces the
 * trace part of render_sphere_trace / single_ray_sphere_trace
 * (RayEngine/src/cuda-src/renderer.cu:142-235, sphere_tracing.cuh:182-216). */
RE_API re_status re_sdf_trace(re_ctx* ctx, int32_t vol_id, const re_sdf_params* params,
                              const float* ray_o, const float* ray_d, int64_t n, float* t,
                              uint8_t* hit, int32_t* steps, float* normal, uint32_t flags);

/* ------------------------------------------------------ SDF surrogate
 *
 * A learned stand-in for the terrain field with no counterpart in the
 * reference: 16-frequency Fourier features of p, 4 hidden SiLU layers of 64,
 * scalar output, BF16 MFMA with FP32 accumulation.  Parameters are one flat
 * float vector of re_mlp_param_count() values:
 *   W0[64][99] b0[64] W1[64][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[64] b4
 * It is trained on the signed Euclidean distance to the field's zero set. */
RE_API int32_t re_mlp_param_count(void);
/* New network with seeded uniform initialisation (variance 1/fan_in). */
RE_API re_status re_mlp_create(re_ctx* ctx, uint64_t seed, int32_t* mlp_id);
RE_API re_status re_mlp_release(re_ctx* ctx, int32_t mlp_id);
/* Replaces the parameters (and resets the Adam state). */
RE_API re_status re_mlp_set_params(re_ctx* ctx, int32_t mlp_id, const float* params);
/* Reads the parameters; adam_m, adam_v, step may be NULL. */
RE_API re_status re_mlp_get_params(re_ctx* ctx, int32_t mlp_id, float* params, float* adam_m,
                                   float* adam_v, int32_t* step);
/* Network output at n points (3 floats each). */
RE_API re_status re_mlp_forward(re_ctx* ctx, int32_t mlp_id, const float* pts, int64_t n,
                                float* out, uint32_t flags);
/* Loss mean_i w_i (out_i - target_i)^2 (w_i = over_weight when out_i >
 * target_i, else 1) over n samples (x, y, z, target), n a multiple of 64:
 * grad (nullable) gets its gradient, loss (nullable) its value; with adam
 * != 0 one Adam step (beta 0.9/0.999, eps 1e-8) at rate lr is applied. */
RE_API re_status re_mlp_batch_step(re_ctx* ctx, int32_t mlp_id, const float* samples, int64_t n,
                                   float over_weight, int32_t adam, float lr, float* grad,
                                   float* loss);
/* Signed Euclidean distance (world units, positive where the field is
 * positive) from each voxel to the field's zero set: exact separable
 * transform of sign-change seeds; a new volume. */
RE_API re_status re_sdf_distance_volume(re_ctx* ctx, int32_t vol_id, int32_t* dist_vol_id);
/* n samples (x, y, z, signed distance) from a distance volume: even indices
 * uniform in the volume, odd indices within near_band of the surface. */
RE_API re_status re_mlp_sample(re_ctx* ctx, int32_t dist_vol_id, int64_t n, uint64_t seed,
                               float near_band, float* out, uint32_t flags);

typedef struct re_mlp_train_params {
  int64_t n_samples;      /* default 2^22 */
  int32_t batch;          /* default 65536 */
  int32_t steps;          /* default 20000 */
  float lr;               /* default 1e-3 */
  float beta1, beta2;     /* default 0.9, 0.999 */
  float eps;              /* default 1e-8 */
  float over_weight;      /* loss weight of over-estimates, default 4 */
  float near_band_voxels; /* near-surface band, default 2 voxels */
  uint64_t seed;
} re_mlp_train_params;

RE_API void re_mlp_default_train_params(re_mlp_train_params* p);
/* Adam training on the GPU on samples drawn from the distance volume;
 * loss_history (nullable) receives one loss per step.  The sample pool is
 * drawn from (volume, seed, n_samples, near band); step s reads batch s of
 * it, wrapping.  A call whose seed, n_samples, batch and near band equal the
 * network's last call continues at the batch where that call stopped (the
 * caller passes the same volume), so N steps and then M steps train exactly
 * as N + M steps in one call. */
RE_API re_status re_mlp_train(re_ctx* ctx, int32_t mlp_id, int32_t dist_vol_id,
                              const re_mlp_train_params* params, float* loss_history);
/* Training checkpoint: parameters, Adam moments, Adam step and the sample
 * pool's key and batch cursor, so that training resumed from it reproduces
 * an uninterrupted run bit for bit.  The file is a 64-byte header ("RESURR"
 * 0 1, param count, step, pool seed, pool size, cursor, batch, near band)
 * and three float32 vectors of re_mlp_param_count() values.  The reference
 * persists nothing but .rsdf volumes (renderer.cu:838-870); the network
 * itself has no counterpart there. */
RE_API re_status re_mlp_save(re_ctx* ctx, int32_t mlp_id, const char* path);
/* A new network from a checkpoint; RE_ERR_IO when the file is missing, is
 * not a checkpoint of this network or has the wrong length. */
RE_API re_status re_mlp_load(re_ctx* ctx, const char* path, int32_t* mlp_id);

/* Hybrid terrain tracer.  Surrogate phase: while the network's distance s
 * exceeds tau, t += s - margin.  Exact phase: march the trilinear field with
 * step max(f * k_exact, h_min); at the first sample with f <= 0, bisect the
 * crossing n_bisect times; after rejoin of travel without a crossing the
 * surrogate resumes.  With mlp_id < 0 the exact phase runs from the volume
 * entry: the exact-field tracer at the same accuracy.  Distances are in
 * voxels of the volume. */
typedef struct re_hybrid_params {
  float tau_voxels;    /* default 2; the surrogate steps while its distance exceeds tau */
  float margin_voxels; /* default 1; subtracted from each surrogate step, must be < tau */
  float k_exact;       /* default 0.05 */
  float h_min_voxels;  /* default 0.5 */
  float rejoin_voxels; /* default 4 */
  int32_t n_bisect;    /* default 8 */
  int32_t max_iters;   /* default 4096 */
  int32_t block;       /* reserved, 0 */
} re_hybrid_params;

RE_API void re_hybrid_default_params(re_hybrid_params* p);
/* t = hit distance (1e20 on a miss), hit = 0/1; evals (nullable) = network
 * evaluations per ray; samples (nullable) = exact-field samples per ray. */
RE_API re_status re_hybrid_trace(re_ctx* ctx, int32_t vol_id, int32_t mlp_id,
                                 const re_hybrid_params* params, const float* ray_o,
                                 const float* ray_d, int64_t n, float* t, uint8_t* hit,
                                 int32_t* evals, int32_t* samples, uint32_t flags);

/* ------------------------------------------------------ batched workload
 *
 * Many triangle scenes and terrain tiles, each with its own slice of one ray
 * array, traced by one persistent work-queue kernel: the rays are cut into
 * items of item_rays, waves take items from a device counter, and the wave
 * finishing a target's last item stamps its completion time.  The launch
 * sequence is captured once into a hipGraph and replayed by re_batch_run.
 * Results equal the single-workload entry points ray for ray.  No
 * counterpart in the reference (one synchronous launch per frame,
 * RayEngine/src/cuda-src/renderer.cu:701-738). */
typedef struct re_batch_scene {
  const re_kdtree* tree;
  int64_t ray_begin;
  int64_t ray_count;
} re_batch_scene;

typedef struct re_batch_tile {
  int32_t vol_id;
  int32_t mode; /* 0 exact-field march, 1 reference sphere trace, 2 surrogate hybrid */
  int64_t ray_begin;
  int64_t ray_count;
  int32_t mlp_id; /* network of a surrogate tile; -1: re_batch_config.mlp_id */
  int32_t reserved;
} re_batch_tile;

typedef struct re_batch_config {
  int32_t item_rays; /* rays per work item, default 256 */
  int32_t block;     /* threads per block (64, 128, 256), default 256 */
  int32_t mlp_id;    /* network of surrogate tiles, -1 none */
  re_trace_params trace;
  re_sdf_params sdf;     /* reference sphere-trace tiles (mode 1) */
  re_hybrid_params hybrid; /* surrogate tiles (mode 2) */
  re_hybrid_params march;  /* exact-march tiles (mode 0): k_exact, h_min_voxels, n_bisect, max_iters */
  float side_share;        /* with surrogate tiles: the fraction of the device's work-group slots the
                              persistent kernel leaves to the surrogate tiles' launches, which run
                              beside it on their own graph branch; in [0, 0.9], default 0 */
} re_batch_config;

RE_API void re_batch_default_config(re_batch_config* c);
/* Uploads the scenes, tiles and rays (host arrays) and captures the graph:
 * a reset kernel, one persistent work-queue kernel over every scene and
 * every exact / reference-march tile (items of item_rays rays, taken by
 * whole waves from a device counter), then one hybrid-trace launch per
 * surrogate tile.  Results of a triangle ray are (prim id, t); of a terrain
 * ray (0 hit / -1 miss, t); each equals the single-workload entry point's
 * result for that ray bit for bit.  The per-ray loops are bounded by the
 * config's trace, sdf and hybrid parameters, validated as the single-workload
 * entry points validate them. */
RE_API re_status re_batch_create(re_ctx* ctx, const re_batch_scene* scenes, int32_t n_scenes,
                                 const re_batch_tile* tiles, int32_t n_tiles, const float* ray_o,
                                 const float* ray_d, int64_t n_rays, const re_batch_config* config,
                                 int32_t* batch_id);
/* The work items re_batch_create cuts for these targets, with no context or
 * device: targets (slots: scenes 0..n_scenes-1, then tiles) in ascending ray
 * count, ties in slot order, each target's rays cut into items of item_rays
 * rays.  items receives 24-byte records {int32 kind (0 triangles, 1 exact
 * march, 2 reference sphere trace), int32 target (scene or tile index),
 * int64 first ray, int32 count, int32 slot}; surrogate tiles (mode 2) get no
 * items (their launches run beside the persistent kernel) and count one
 * completion.  remaining (nullable, n_scenes + n_tiles entries) receives each
 * slot's item count: the counter the wave finishing a slot's last item takes
 * to zero.  items may be NULL to query *n_items.  The tree pointers of scenes
 * are not read.  No counterpart in the reference. */
RE_API re_status re_batch_plan(const re_batch_scene* scenes, int32_t n_scenes, const re_batch_tile* tiles,
                               int32_t n_tiles, int64_t n_rays, int32_t item_rays, void* items, int64_t max_items,
                               int64_t* n_items, int32_t* remaining);
/* Replays the captured graph (RE_ASYNC: do not wait). */
RE_API re_status re_batch_run(re_ctx* ctx, int32_t batch_id, uint32_t flags);
RE_API re_status re_batch_info(re_ctx* ctx, int32_t batch_id, int64_t* n_rays, int64_t* n_items,
                               int32_t* n_slots);
RE_API re_status re_batch_results(re_ctx* ctx, int32_t batch_id, int32_t* prim, float* t);
/* Completion time of every target of the last run (scenes first, then
 * tiles), milliseconds after the run started, from device timestamps. */
RE_API re_status re_batch_completion_ms(re_ctx* ctx, int32_t batch_id, float* ms);
RE_API re_status re_batch_release(re_ctx* ctx, int32_t batch_id);

/* ------------------------------------------------------ autotuner */

/* Candidates of re_autotune_trace; a NULL list (or count 0) takes the default
 * in brackets. */
typedef struct re_tune_space {
  const int32_t* leaf_sizes; /* kd-tree max_leaf_tris, the tree rebuilt on the device for each [4, 8, 12, 16] */
  int32_t n_leaf_sizes;
  const int32_t* blocks;     /* work-group sizes, 64, 128 or 256 [64, 128, 256] */
  int32_t n_blocks;
  const int32_t* variants;   /* re_trace_params.variant values [3, 4, 5] */
  int32_t n_variants;
  int32_t reps;              /* timed launches per candidate, at most 64 [10] */
} re_tune_space;

typedef struct re_tune_result {
  int32_t max_leaf_tris;
  int32_t block;
  int32_t variant;
  float ms; /* mean kernel time of the winning candidate (HIP events) */
} re_tune_result;

/* Picks the leaf size, work-group size and kernel variant of the closest-hit
 * query for one scene (the soup tri_v, 9 floats per triangle, host memory)
 * and a ray set, by timing on the context's device: for every leaf size the
 * tree is built with re_kd_build_gpu (other fields from build, NULL =
 * defaults) and uploaded; every (block, variant) pair is launched once
 * untimed and then reps times, and the lowest mean kernel time wins (the
 * first such candidate on a tie).  base supplies the other trace parameters
 * (cap, one_sided, secondary; NULL = defaults).  Rays are host memory, or
 * device memory with RE_DEVICE_PTRS in flags.  ms_all (nullable) receives
 * every candidate's mean, leaf sizes outermost, variants innermost.  The
 * tuner allocates its own device buffers for the results (and for host rays)
 * and frees them, with the trees, before it returns.  No counterpart in the
 * reference, whose launch shape was fixed (renderer.cu:389-410); the batched
 * launch has its own tuner in the Python host mirror. */
RE_API re_status re_autotune_trace(re_ctx* ctx, const float* tri_v, int64_t n_tris, const re_kd_build_params* build,
                                   const float* ray_o, const float* ray_d, int64_t n, const re_trace_params* base,
                                   const re_tune_space* space, uint32_t flags, re_tune_result* best, float* ms_all);

/* Elapsed milliseconds of the last kernel launch of the context, measured
 * with HIP events recorded around the launch on its stream (waits for it).
 * Replaces the reference's FPS counter (RayEngine/src/Engine/Window.cpp:46-56). */
RE_API re_status re_last_kernel_ms(re_ctx* ctx, float* ms);
/* Durations of the last min(max, 256) kernel launches, oldest first (waits
 * for them); *count gets the number written. */
RE_API re_status re_kernel_times(re_ctx* ctx, float* ms, int32_t max, int32_t* count);
/* For the last k = min(max, 256) launches, oldest first: the time from the
 * start of the first of them to the end of each (waits for them), so idle
 * gaps and host issue delays between launches on the stream are included;
 * *count gets k. */
RE_API re_status re_launch_end_times(re_ctx* ctx, float* ms, int32_t max, int32_t* count);

#ifdef __cplusplus
}
#endif

#endif /* RAYENGINE_H */
