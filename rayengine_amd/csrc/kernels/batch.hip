// Batched mixed workload in one persistent launch: many triangle scenes and
// terrain tiles, each with its own ray batch, cut into work items of a few
// dozen to a few hundred rays.  Every wave takes the next item from a device
// counter and runs its rays with the same per-ray paths as the single-
// workload kernels (so batched results equal unbatched ones bit for bit);
// the wave that finishes a target's last item stamps its completion time.
// The launch sequence (reset, persistent kernel) is captured once in a
// hipGraph and replayed.  No counterpart in the reference, which launched
// one kernel per frame and synchronised after each (renderer.cu:701-738).
#include "../device_paths.hpp"

namespace re {
namespace {

__global__ void batch_reset_kernel(unsigned long long* __restrict__ ctl, int32_t* __restrict__ remaining,
                                   const int32_t* __restrict__ remaining_init,
                                   unsigned long long* __restrict__ done_ts, int32_t n_slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    ctl[0] = 0ull;  // item counter (low 32 bits)
    ctl[1] = wall_clock64();
  }
  if (i < n_slots) {
    remaining[i] = remaining_init[i];
    done_ts[i] = 0ull;
  }
}

__global__ void batch_stamp_kernel(unsigned long long* __restrict__ done_ts, int32_t slot,
                                   int32_t* __restrict__ remaining) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    remaining[slot] = 0;
    done_ts[slot] = wall_clock64();
  }
}

__global__ void hit_to_prim_kernel(const uint8_t* __restrict__ hit, int32_t* __restrict__ prim, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) prim[i] = hit[i] ? 0 : -1;
}

// Wave-uniform fetch-and-add: every lane executes the atomic, lane 0 with
// `add` and the others with 0, so the wave adds `add` once; the value before
// the add is read back from lane 0 (the first active lane: the caller's wave
// is converged) with readfirstlane, which puts it in a scalar register.
// Control flow that depends on the result is therefore a scalar branch taken
// by the whole wave.
//
// Why not "if (lane == 0) v = atomicAdd(...); v = __shfl(v, 0);": in a
// loop whose exit depends on v, that form leaves the compiler free to treat
// the exit as divergent.  The gfx950 code it gave for this kernel re-entered
// the block holding the shuffle with lanes 1-63 only (lane 0 had left to
// issue the next atomic), so those lanes read a stale zero, took item 0 again
// and never left the loop: the launch did not finish.
__device__ __forceinline__ uint32_t wave_fetch_add(uint32_t* p, uint32_t add, int lane) {
  const uint32_t v = atomicAdd(p, lane == 0 ? add : 0u);
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ __forceinline__ int32_t wave_fetch_sub(int32_t* p, int32_t sub, int lane) {
  const int32_t v = atomicSub(p, lane == 0 ? sub : 0);
  return __builtin_amdgcn_readfirstlane(v);
}

template <bool ONE_SIDED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) batch_kernel(BatchLaunch L) {
  // one CoopWave per wave of the block (dynamic LDS sized by the launch)
  extern __shared__ CoopWave waves[];
  CoopWave& wv = waves[threadIdx.x >> 6];
  const int lane = threadIdx.x & 63;
  uint32_t* const queue = reinterpret_cast<uint32_t*>(L.ctl);  // ctl[0], low word
  for (;;) {
    const uint32_t it = wave_fetch_add(queue, 1u, lane);
    if ((int64_t)it >= L.n_items) break;  // scalar: `it` is wave-uniform
    const BatchItem item = L.items[it];
    if (item.kind == kItemTri) {
      // a scene's rays, 64 at a time, walked by the whole wave together
      // (the cooperative traversal of the single-scene kernel's default)
      for (int r0 = 0; r0 < item.count; r0 += 64) {
        const int r = r0 + lane;
        const bool has = r < item.count;
        const int64_t i = item.begin + (has ? r : 0);
        const f3 o{L.ray_o[3 * i], L.ray_o[3 * i + 1], L.ray_o[3 * i + 2]};
        const f3 d{L.ray_d[3 * i], L.ray_d[3 * i + 1], L.ray_d[3 * i + 2]};
        float bt;
        int32_t bp, bo;
        RayCounts rc;
        coop_trace_wave<false, ONE_SIDED, true>(L.nodes, L.leaves, L.tris, L.objs, item.target, item.target + 1, has,
                                                o, d, L.cap, L.secondary != 0, bt, bp, bo, rc, wv);
        if (has) {
          L.prim[i] = bp;
          L.t[i] = bt;
        }
      }
    }
    for (int r = lane; item.kind != kItemTri && r < item.count; r += 64) {
      const int64_t i = item.begin + r;
      const f3 o{L.ray_o[3 * i], L.ray_o[3 * i + 1], L.ray_o[3 * i + 2]};
      const f3 d{L.ray_d[3 * i], L.ray_d[3 * i + 1], L.ray_d[3 * i + 2]};
      {
        const BatchTile& tile = L.tiles[item.target];
        float t;
        int32_t n;
        bool h;
        if (item.kind == kItemExact)
          h = exact_march(tile.vol, tile.box_min, tile.box_max, o, d, tile.hp.k_exact, tile.hp.h_min,
                          tile.hp.n_bisect, tile.hp.max_iters, t, n);
        else
          h = plain_march(tile.vol, tile.box_min, tile.box_max, o, d, L.sdf.max_steps, L.sdf.eps, L.sdf.step_scale,
                          L.sdf.step_rule, t, n);
        L.prim[i] = h ? 0 : -1;
        L.t[i] = h ? t : kInf;
      }
    }
    // The wave's one decrement of the target's item count; the wave that
    // takes it to zero stamps the target's completion (every lane stores the
    // same scalar value: no lane-dependent branch inside the loop).
    if (wave_fetch_sub(&L.remaining[item.slot], 1, lane) == 1) L.done_ts[item.slot] = wall_clock64();
  }
}

}  // namespace

int64_t batch_grid(int block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batch_kernel<true>, block,
                                                    (size_t)(block / 64) * sizeof(CoopWave)) != hipSuccess)
    return 0;
  return (int64_t)cus * (per_cu > 0 ? per_cu : 1);
}

hipError_t launch_batch_reset(const BatchLaunch& L, hipStream_t s) {
  const int n = L.n_slots > 1 ? L.n_slots : 1;
  hipLaunchKernelGGL(batch_reset_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L.ctl, L.remaining,
                     L.remaining_init, L.done_ts, L.n_slots);
  return hipGetLastError();
}

hipError_t launch_hit_to_prim(const uint8_t* hit, int32_t* prim, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(hit_to_prim_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, hit, prim, n);
  return hipGetLastError();
}

hipError_t launch_batch_stamp(unsigned long long* done_ts, int32_t slot, int32_t* remaining, hipStream_t s) {
  hipLaunchKernelGGL(batch_stamp_kernel, dim3(1), dim3(64), 0, s, done_ts, slot, remaining);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchLaunch& L, hipStream_t s) {
  if (L.n_items <= 0) return hipSuccess;
  if (L.n_items > kMaxBatchItems) return hipErrorInvalidValue;
  const int block = L.block > 0 ? L.block : 256;
  int64_t grid = L.grid_blocks > 0 ? L.grid_blocks : 1024;
  const int64_t want = (L.n_items * 64 + block - 1) / block;
  if (grid > want) grid = want;
  const size_t lds = (size_t)(block / 64) * sizeof(CoopWave);
  if (L.one_sided)
    hipLaunchKernelGGL(batch_kernel<true>, dim3((unsigned)grid), dim3(block), lds, s, L);
  else
    hipLaunchKernelGGL(batch_kernel<false>, dim3((unsigned)grid), dim3(block), lds, s, L);
  return hipGetLastError();
}

}  // namespace re
