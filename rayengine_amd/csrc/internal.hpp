// Internal declarations shared between the C-ABI layer (capi.cpp) and the
// kernel launchers (kernels/*.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "../../include/rayengine.h"
#include "host/batch_plan.hpp"
#include "mlp_layout.hpp"
#include "re_math.hpp"
#include "terrain_math.hpp"

struct re_kdtree;

namespace re {

const std::vector<NodeEntry>& kd_nodes(const re_kdtree* t);
const std::vector<LeafRec>& kd_leaves(const re_kdtree* t);
const std::vector<TriRec>& kd_tris(const re_kdtree* t);

// A scene resident in HBM: the concatenated trees of its objects (leaf
// records in the device order of DevLeaf; the host keeps the flat records for
// re_scene_export).
struct DevScene {
  NodeEntry* nodes = nullptr;
  DevLeaf* leaves = nullptr;
  TriRec* tris = nullptr;
  ObjXform* objs = nullptr;
  int32_t n_obj = 0;
  int64_t n_nodes = 0, n_leaves = 0, n_trirecs = 0;
  int64_t bytes = 0;
  std::vector<LeafRec> h_leaves;
  // hit attributes (re_scene_set_attributes), indexed by prim id
  float* attr_geom = nullptr;  // v0, e1, e2 (9 floats) per prim
  float* attr_nrm = nullptr;   // vertex normals n0, n1, n2 (9 floats) per prim
  float* attr_uv = nullptr;    // vertex uvs uv0, uv1, uv2 (6 floats) per prim, or null
  int64_t n_attr = 0;
};

struct AttrLaunch {
  const DevScene* scene;
  const float* ray_o;
  const float* ray_d;
  const int32_t* prim;
  const int32_t* obj;  // null: object 0
  int64_t n;
  int32_t one_sided, secondary;
  float* normal;     // 3 per ray or null
  float* uv;         // 2 per ray or null
  float* bary;       // 2 per ray or null
  float* hit_point;  // 3 per ray or null
};
hipError_t launch_hit_attributes(const AttrLaunch& L, hipStream_t s);

// Closest hit with each ray's status (a debugging aid): how the walk of the
// last object ended, as the oracle numbers it (0 t_entry >= t_exit, 1 visit
// cap, 2 leaf slab miss, 3 no exit face within the tolerance, 4 rope -1,
// 5 root box missed, 6 walk state repeated), plus 0x80 when the ray's origin
// or direction holds a NaN or an infinity, or the direction is zero.
struct StatusLaunch {
  const DevScene* scene;
  const float* ray_o;
  const float* ray_d;
  int64_t n;
  int32_t cap, one_sided, secondary;
  int32_t* prim;
  float* t;
  uint8_t* status;
};
hipError_t launch_trace_status(const StatusLaunch& L, hipStream_t s);
// The bounds-check error mask of the debug build since the last call (read
// and cleared); hipErrorNotSupported in a build without RE_DEBUG_BOUNDS.
hipError_t debug_bounds_take(uint32_t* mask);

// Per-ray work counters of a traversal: node entries read, leaf records read,
// triangle records tested.
struct TraceCounts {
  unsigned long long entries, leaves, tris, rays_hit;
};

struct TraceLaunch {
  const DevScene* scene;
  const float* ray_o;
  const float* ray_d;
  int64_t n;
  int32_t cap;
  int32_t one_sided;
  int32_t secondary;
  int32_t block;
  int32_t* prim;
  float* t;
  int32_t* obj;
  TraceCounts* counts;  // device pointer or null
  uint8_t* occluded;    // any-hit query: 1 when some object detects a hit
  int32_t variant;      // re_trace_params.variant
  int64_t grid_blocks;  // persistent grid size (blocks)
  unsigned long long* queue;  // persistent ray counter, zeroed before launch
  // Walk-start hints and hit leaves (variants 4 and 5 only; device pointers
  // or null): start_leaf[i] a leaf the walk of ray i may start at,
  // hit_leaf[i] receives the last leaf visited by the walk that found ray
  // i's closest hit (-1 for a miss).
  const int32_t* start_leaf = nullptr;
  int32_t* hit_leaf = nullptr;
};

// Blocks of `block` threads that fill the device at the persistent kernel's
// occupancy.
int64_t trace_persistent_grid(int block);
// The same for the cooperative refill kernel (variant 5), whose blocks also
// hold one CoopWave of LDS per wave.
int64_t trace_coop_refill_grid(int block);

// Threads per block of a closest-hit launch: the requested size, or the
// default (256 for the persistent per-lane variant 1, one wave otherwise).
inline int trace_launch_block(int block, int variant) { return block > 0 ? block : (variant == 1 ? 256 : 64); }

hipError_t launch_trace_closest(const TraceLaunch& L, hipStream_t s);
// Zeroes n 32-bit words with a kernel (no memset node when a caller captures).
hipError_t launch_zero_words(uint32_t* p, int32_t n, hipStream_t s);

// Ray counters of the persistent kernels, from byte 64 of the counter
// header: kQueueSlots of them (one per XCD for the refill variant; the
// persistent per-lane variant uses the first), kQueueStride 64-bit words apart
// so that each has a 128-byte line of its own.
constexpr int kQueueSlots = 8;
constexpr int kQueueStride = 16;
constexpr int32_t kQueueWords = kQueueSlots * kQueueStride * 2;  // 32-bit words to clear
static_assert((kQueueSlots & (kQueueSlots - 1)) == 0, "a power of two: the XCD index is masked with it");
// The counter header (ctx->counts_buf, and the start of the fixed area):
// work counters at byte 0, the ray counters at byte 64.
constexpr size_t kCountsBytes = 2048;
static_assert(64 + 4 * (size_t)kQueueWords <= kCountsBytes, "ray counters inside the header");
// A fixed device area standing in for the allocated buffers of a call: the
// counter header, then host-pointer staging from kScratchHeader on.
constexpr size_t kScratchHeader = kCountsBytes;
struct ScratchArea {
  void* base;
  size_t bytes;
};

hipError_t launch_trace_occluded(const TraceLaunch& L, hipStream_t s);

// A volume resident in HBM: 8x8x8 Morton-ordered bricks.
struct DevVolume {
  float* data = nullptr;
  float* mat = nullptr;  // material channel, same layout as data; null until loaded or edited
  int32_t dim[3] = {0, 0, 0};
  int32_t bdim[3] = {0, 0, 0};
  float spacing[3] = {0, 0, 0};
  float box_min[3] = {0, 0, 0};
  float box_max[3] = {0, 0, 0};
  int64_t bytes = 0;
  VolView view() const {
    VolView v;
    v.data = data;
    for (int a = 0; a < 3; ++a) {
      v.dim[a] = dim[a];
      v.bdim[a] = bdim[a];
      v.spacing[a] = spacing[a];
    }
    return v;
  }
};

struct NoiseLaunch {
  float* out;
  const uint32_t* perm;
  int32_t dim[3];
  int32_t bdim[3];
  float spacing[3];
  float origin[3];
  FbmParams fbm;
};

struct SdfLaunch {
  VolView vol;
  float box_min[3], box_max[3];
  const float* ray_o;
  const float* ray_d;
  int64_t n;
  int32_t max_steps;
  float eps, step_scale;
  int32_t step_rule;
  int32_t block;
  float* t;
  uint8_t* hit;
  int32_t* steps;
  float* nrm;
};

hipError_t launch_noise_fill(const NoiseLaunch& L, hipStream_t s);

// Frame colouring (sky for misses, simple_shade for hits, tone, pack).
struct FrameLaunch {
  const float* ray_d;
  const uint8_t* hit;
  const float* normal;
  int32_t width, height;
  float sun[3], er, ar, hr, hm, origin[3], base[3];
  int32_t gamma, vignetting;
  float vignetting_k;
  uint32_t* rgba;
};
hipError_t launch_frame(const FrameLaunch& L, hipStream_t s);

// One brush edit over a whole volume (distance and material channels).
struct EditLaunch {
  float* data;
  float* mat;
  int32_t dim[3], bdim[3];
  float spacing[3];
  int32_t type;
  float center[3], extent[3];
  float material;
  int32_t snap;
  float k;
};
hipError_t launch_volume_edit(const EditLaunch& L, hipStream_t s);
hipError_t launch_sdf_trace(const SdfLaunch& L, hipStream_t s);

// Doubled Perlin permutation table of RPerlinNoise(seed).
void perlin_table(uint32_t seed, uint32_t* perm512);

// ---- SDF surrogate (kernels/surrogate.hip) ----
struct DevMlp {
  float* params = nullptr;   // FP32 master parameters (flat layout, mlp_layout.hpp)
  float* m1 = nullptr;       // Adam first moment
  float* m2 = nullptr;       // Adam second moment
  void* pack = nullptr;      // bf16 MFMA operands: training (64 KB), then inference (40 KB)
  float* small = nullptr;    // FP32 b0..b3, W4, b4: training, then inference
  float* partial = nullptr;  // per-block gradient partials (+ loss)
  float* grad = nullptr;     // summed gradient of the last batch
  float* loss = nullptr;     // per-step loss history
  int64_t loss_cap = 0;
  int32_t partial_blocks = 0;
  int32_t step = 0;          // Adam steps taken
  // Training sample pool: the key it was drawn with and the batches taken
  // from it, so that a later re_mlp_train call with the same key continues
  // where the last one stopped.
  uint64_t pool_seed = 0;
  int64_t pool_n = 0;
  int32_t pool_batch = 0;
  uint32_t pool_band_bits = 0;  // near_band_voxels as float bits
  int64_t cursor = -1;          // batches taken; -1: no pool yet
};

size_t mlp_train_lds_bytes();

// The inference operands (mlp_layout.hpp: pre-activations carried as
// -log2(e) z) stored after the training ones in the same buffers.
inline const void* infer_pack(const DevMlp& m) { return (const uint16_t*)m.pack + mlp::oPackInfer; }
inline const float* infer_small(const DevMlp& m) { return m.small + mlp::oSmallInfer; }

struct HybridParams {
  float tau;      // surrogate distance below which the exact field takes over
  float margin;   // subtracted from every surrogate step
  float k_exact;  // exact-phase step = max(f * k_exact, h_min)
  float h_min;
  float rejoin;   // exact-phase travel without a crossing before the surrogate resumes
  int32_t n_bisect;
  int32_t max_iters;
};

hipError_t launch_hybrid_trace(const DevVolume& field, const void* pack, const float* small, const float* ray_o,
                               const float* ray_d, int64_t n, const HybridParams& hp, bool use_mlp, float* t,
                               uint8_t* hit, int32_t* evals, int32_t* samples, hipStream_t s);

// ---- batched work queue (kernels/batch.hip); items: host/batch_plan.hpp ----

// A terrain tile of a batch: its volume, box and exact-march parameters in
// world units (scaled by that tile's own voxel size, as re_hybrid_trace
// scales them).
struct BatchTile {
  VolView vol;
  float box_min[3], box_max[3];
  HybridParams hp;
};

struct BatchSdf {
  int32_t max_steps;
  float eps, step_scale;
  int32_t step_rule;
};

struct BatchLaunch {
  const NodeEntry* nodes;
  const DevLeaf* leaves;
  const TriRec* tris;
  const ObjXform* objs;
  const BatchTile* tiles;
  const BatchItem* items;
  int64_t n_items;
  const float* ray_o;
  const float* ray_d;
  int32_t* prim;
  float* t;
  unsigned long long* ctl;           // [0] item counter (32-bit, low word), [1] start timestamp
  int32_t* remaining;                // per slot: items left
  const int32_t* remaining_init;
  unsigned long long* done_ts;       // per slot completion timestamp
  int32_t n_slots;
  int32_t cap, one_sided, secondary;
  BatchSdf sdf;
  int32_t block;
  int64_t grid_blocks;
};

// A batch resident on the device with its launch sequence captured once.
struct SurrogateTile {
  int32_t tile, slot;
  int64_t begin, count;
  int32_t mlp_id;  // the tile's network
};

struct DevBatch {
  DevScene arena;  // every scene's tree, one object per scene
  BatchTile* tiles = nullptr;
  std::vector<int32_t> tile_vol;  // volume id per tile
  std::vector<int32_t> tile_mode;
  float* ray_o = nullptr;
  float* ray_d = nullptr;
  int64_t n_rays = 0;
  int32_t* prim = nullptr;
  float* t = nullptr;
  uint8_t* hit_scratch = nullptr;
  BatchItem* items = nullptr;
  int64_t n_items = 0;
  unsigned long long* ctl = nullptr;
  int32_t* remaining = nullptr;
  int32_t* remaining_init = nullptr;
  unsigned long long* done_ts = nullptr;
  int32_t n_slots = 0, n_scenes = 0, n_tiles = 0;
  std::vector<SurrogateTile> surrogate;
  int32_t mlp_id = -1;
  BatchLaunch launch;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // surrogate tiles run on a branch of the graph beside the persistent kernel
  // (captured from this stream between a fork and a join event)
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double wall_clock_khz = 100000.0;
};

// Frees everything a batch holds (graph, branch stream and events, device
// buffers) and resets it.
void release_batch(DevBatch& b);

hipError_t launch_batch_reset(const BatchLaunch& L, hipStream_t s);
hipError_t launch_hit_to_prim(const uint8_t* hit, int32_t* prim, int64_t n, hipStream_t s);
hipError_t launch_batch(const BatchLaunch& L, hipStream_t s);
hipError_t launch_batch_stamp(unsigned long long* done_ts, int32_t slot, int32_t* remaining, hipStream_t s);
int64_t batch_grid(int block);
hipError_t launch_mlp_infer(const void* pack, const float* small, const float* pts, int64_t n, float* out,
                            hipStream_t s);
hipError_t launch_mlp_train(const void* pack, const float* small, const float4* samples, int64_t n_samples,
                            int64_t batch_start, int32_t batch, float over_weight, float* partial, int32_t blocks,
                            hipStream_t s);
hipError_t launch_mlp_update(float* params, float* m1, float* m2, float* partial, int32_t n_blocks,
                             int32_t step, float lr, float b1, float b2, float eps, int32_t do_adam, float* grad_out,
                             void* pack, float* small, float* loss_out, hipStream_t s);
hipError_t launch_edt(const DevVolume& field, float* scratch_a, float* scratch_b, float* out_bricked, hipStream_t s);
hipError_t launch_sample(const DevVolume& edt, int64_t n, uint64_t seed, float near_band, int32_t max_tries,
                         float4* out, hipStream_t s);

}  // namespace re

struct re_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // A pair of HIP events is recorded around every kernel launch into a ring,
  // so the durations of asynchronous launches can be read back later.
  static constexpr int kRing = 256;
  hipEvent_t ring[2 * kRing] = {};
  int64_t n_launch = 0;
  int32_t next_id = 1;
  std::map<int32_t, re::DevScene> scenes;
  std::map<int32_t, re::DevVolume> volumes;
  std::map<int32_t, re::DevMlp> mlps;
  std::map<int32_t, re::DevBatch> batches;
  // Staging for host-pointer calls, grown on demand.
  void* counts_buf = nullptr;  // kCountsBytes: work counters at 0, ray counters at 64
  void* noalloc_area = nullptr;  // this device's fixed area of re_trace_closest_noalloc (capi_noalloc.cpp)
  int64_t persistent_grid[3] = {0, 0, 0};  // per block size 64, 128, 256
  int64_t coop_refill_grid[3] = {0, 0, 0};  // the same for variant 5
  void* stage = nullptr;
  size_t stage_bytes = 0;
  std::string err;

  hipError_t launch_begin() { return hipEventRecord(ring[2 * (n_launch % kRing)], stream); }
  hipError_t launch_end() {
    hipError_t e = hipEventRecord(ring[2 * (n_launch % kRing) + 1], stream);
    ++n_launch;
    return e;
  }
};

namespace re {
// Resolves the device address of the fixed area of re_trace_closest_noalloc
// on the context's device (current), once per device, into ctx->noalloc_area.
re_status noalloc_setup(re_ctx* ctx);

// The closest-hit entry points' shared body (capi.cpp).
re_status trace_common(re_ctx* ctx, int32_t scene_id, const float* ray_o, const float* ray_d, int64_t n,
                       const re_trace_params* params, int32_t* prim_id, float* t, int32_t* obj_id,
                       uint32_t flags, TraceCounts* counts, const ScratchArea* sa,
                       const int32_t* start_leaf = nullptr, int32_t* hit_leaf = nullptr);
}  // namespace re
