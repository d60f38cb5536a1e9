// Neighbour lists of the gb_cut systems, built on the device at the start
// of an evaluation (after the tile boxes).
//
// The reference's neighbour-list path (physics/neighbor_list.py:155-172, a
// jax_md partition.neighbor_list with a Verlet skin; evaluated by
// generalized_born.py:266-415 and optimization.py:128-194, 299-357) keeps
// per-atom lists of partner atoms. Here the list is per 64-atom tile I: the
// atoms of the later tiles J > I that lie within cutoff + skin of tile I's
// bounding box, in increasing atom order. Pair kernels evaluate it in chunks
// of 64 listed atoms against the tile's 64 atoms with the per-pair cutoff
// test, so a chunk does the work of a tile pair, but only over atoms that can
// be in range: against whole tile pairs selected by box distance, this cuts
// the pair work of a 20,000-atom globule at a 20 A cutoff by ~40%.
//
// Three kernels, no atomics, so every list is a fixed function of the
// positions:
//   k_nl_check  per system: build this evaluation if forced, if skin = 0, or
//               if an atom moved more than skin / 2 since the last build
//               (the Verlet-skin rebuild rule of jax_md's neighbour list);
//   k_nl_mask   per tile pair (I, J > I) without exclusions: a 64-bit mask of
//               the atoms of J within cutoff + skin of tile I's box;
//   k_nl_rows   per tile I: prefix counts over its row's masks, the list
//               entries, the list length, and (with a skin) the reference
//               positions of the tile's atoms.
// Tile pairs that hold exclusions stay ordinary items (their masks are 0).
#include "prx_nl_dev.h"

namespace prx {

namespace {

__device__ __forceinline__ int wave_index() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
}

// One block per listed system.
__global__ __launch_bounds__(kNlThreads) void k_nl_check(DeviceView v, const int32_t* ranges, int force) {
  nl_check_block(v, ranges, blockIdx.x, force);
}

// One wave per tile-pair record.
__global__ __launch_bounds__(kNlThreads) void k_nl_mask(DeviceView v) {
  const int w = wave_index();
  if (w >= v.nl_tp_hw) return;
  nl_mask_wave(v, w);
}

// One wave per tile.
__global__ __launch_bounds__(kNlThreads) void k_nl_rows(DeviceView v) {
  const int t = wave_index();
  if (t >= v.n_tiles) return;
  nl_rows_wave(v, t);
}

}  // namespace

void launch_nl_build(const DeviceView& v, const int32_t* slot_ranges, int n_slots, bool force, hipStream_t s) {
  if (n_slots <= 0) return;
  hipLaunchKernelGGL(k_nl_check, dim3(n_slots), dim3(kNlThreads), 0, s, v, slot_ranges, force ? 1 : 0);
  const int wpb = kNlThreads / 64;
  if (v.nl_tp_hw > 0)
    hipLaunchKernelGGL(k_nl_mask, dim3((v.nl_tp_hw + wpb - 1) / wpb), dim3(kNlThreads), 0, s, v);
  if (v.n_tiles > 0) hipLaunchKernelGGL(k_nl_rows, dim3((v.n_tiles + wpb - 1) / wpb), dim3(kNlThreads), 0, s, v);
}

}  // namespace prx
