// Persistent step kernel: K BAOAB steps of every system of an engine in one
// cooperative launch, the passes separated by grid-wide barriers.
//
// A step of the graph path is 6 dependent kernel launches (plus the bonded
// branch and, for neighbour-list systems, the list build). For an engine
// with little pair work (one system of a few thousand atoms) the step time is
// then set by launch and dependency latency, not by the work: ~90 us per step
// at 138 or 2,500 atoms against ~10 us of kernels. Here one launch runs K
// steps; each pass is a grid-stride loop over the same per-item and per-tile
// device functions the graph's kernels call (prx_*_dev.h), so every result
// is bit for bit the graph path's.
//
// Grid: kPersistThreads-thread blocks (one tile gather's workgroup), all
// co-resident (hipLaunchCooperativeKernel refuses a grid that is not).
// Barrier: an arrival counter and a generation word in device memory; the
// last block to arrive resets the counter and bumps the generation, the
// others wait on it with s_sleep. Agent-scope fences around it order every
// block's writes of a pass before any block's reads in the next. A wait that
// exceeds kSpinLimit polls sets the error word, and every block leaves the
// kernel at its next barrier, so a fault in the scheme ends the launch with
// an error rather than a hang.
#include "prx_atom_dev.h"
#include "prx_bonded_dev.h"
#include "prx_nl_dev.h"
#include "prx_pair_dev.h"

namespace prx {

namespace {

constexpr int kPersistThreads = 256;
constexpr int kPersistWaves = kPersistThreads / 64;
constexpr unsigned kSpinLimit = 1u << 24;  // polls of ~0.1 us each: seconds, far above any pass

// Waits for every block; false if the launch is being abandoned.
__device__ __forceinline__ bool grid_sync(GridBarrier* bar, unsigned& gen) {
  __shared__ int s_ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    int ok = __hip_atomic_load(&bar->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
    if (ok) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");  // this block's writes, before its arrival
      const unsigned arrived =
          __hip_atomic_fetch_add(&bar->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (arrived == gridDim.x - 1) {
        __hip_atomic_store(&bar->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&bar->generation, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      } else {
        unsigned polls = 0;
        while (__hip_atomic_load(&bar->generation, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
          __builtin_amdgcn_s_sleep(2);
          if (++polls > kSpinLimit ||
              __hip_atomic_load(&bar->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            __hip_atomic_store(&bar->error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // the other blocks' writes, before our reads
    }
    gen += 1;
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

__global__ __launch_bounds__(kPersistThreads) void k_persist_steps(DeviceView v, const int32_t* ranges, int n_slots,
                                                                    int n_steps, int has_nl, GridBarrier* bar) {
  __shared__ PassLds lds[kPersistWaves];
  PassLds& L = lds[threadIdx.x >> 6];
  __shared__ GatherLds G;
  unsigned gen = 0;  // the barrier's generation, tracked by thread 0
  const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * kPersistWaves + (threadIdx.x >> 6)));
  const int n_waves = gridDim.x * kPersistWaves;
  const int tid = blockIdx.x * kPersistThreads + threadIdx.x;
  const int n_threads = gridDim.x * kPersistThreads;
  const int n_bonded = bonded_entries(v);

  // B A O A of the first step from the stored force, then the tile boxes
  for (int g = tid; g < v.atom_hw; g += n_threads) step_begin_atom(v, g);
  if (!grid_sync(bar, gen)) return;
  for (int t = wave; t < v.n_tiles; t += n_waves) tile_box_wave(v, t);
  if (!grid_sync(bar, gen)) return;

  for (int k = 0; k < n_steps; ++k) {
    if (has_nl) {  // the list build (forced at the first step, as in a graph launch)
      for (int s = blockIdx.x; s < n_slots; s += gridDim.x) {
        nl_check_block(v, ranges, s, k == 0);
        __syncthreads();
      }
      if (!grid_sync(bar, gen)) return;
      for (int w = wave; w < v.nl_tp_hw; w += n_waves) nl_mask_wave(v, w);
      if (!grid_sync(bar, gen)) return;
      for (int t = wave; t < v.n_tiles; t += n_waves) nl_rows_wave(v, t);
      if (!grid_sync(bar, gen)) return;
      for (int w = wave; w < v.nl_tp_hw; w += n_waves) nl_fill_wave(v, w);
      if (!grid_sync(bar, gen)) return;
    }
    for (int w = wave; w < v.n_items; w += n_waves) born_item(v, w, L.born);
    for (int e = tid; e < n_bonded; e += n_threads) bonded_entry<false>(v, e);
    if (!grid_sync(bar, gen)) return;
    for (int t = blockIdx.x; t < v.n_tiles; t += gridDim.x) {
      born_gather_tile(v, t, G);
      __syncthreads();
    }
    if (!grid_sync(bar, gen)) return;
    for (int w = wave; w < v.n_items; w += n_waves) force_item<false>(v, w, L.pair);
    if (!grid_sync(bar, gen)) return;
    for (int t = blockIdx.x; t < v.n_tiles; t += gridDim.x) {
      force_gather_tile<false>(v, t, G);
      __syncthreads();
    }
    if (!grid_sync(bar, gen)) return;
    for (int w = wave; w < v.n_items; w += n_waves) chain_item(v, w, L.pair);
    if (!grid_sync(bar, gen)) return;
    const int mode = k + 1 == n_steps ? 2 : 1;  // the last step stops after B(F)
    for (int t = blockIdx.x; t < v.n_tiles; t += gridDim.x) {
      chain_gather_tile(v, t, mode, G);
      __syncthreads();
    }
    if (k + 1 < n_steps && !grid_sync(bar, gen)) return;
  }
}

__global__ void k_barrier_reset(GridBarrier* bar) {
  if (threadIdx.x == 0) {
    bar->count = 0u;
    bar->generation = 0u;
    bar->error = 0;
  }
}

}  // namespace

int persist_max_blocks() {
  int dev = 0, cus = 0, per_cu = 0, coop = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persist_steps, kPersistThreads, 0) != hipSuccess)
    return 0;
  return cus * per_cu;
}

hipError_t launch_persist_steps(const DeviceView& v, const int32_t* slot_ranges, int n_slots, int n_steps,
                                bool has_nl, GridBarrier* bar, int n_blocks, hipStream_t s) {
  hipLaunchKernelGGL(k_barrier_reset, dim3(1), dim3(64), 0, s, bar);
  DeviceView vv = v;
  int hn = has_nl ? 1 : 0;
  void* args[] = {&vv, (void*)&slot_ranges, &n_slots, &n_steps, &hn, &bar};
  return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_persist_steps), dim3(n_blocks),
                                    dim3(kPersistThreads), args, 0, s);
}

}  // namespace prx
