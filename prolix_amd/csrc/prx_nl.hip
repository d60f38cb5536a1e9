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
// Four kernels, no atomics, so every list is a fixed function of the
// positions:
//   k_nl_check  per system: build this evaluation if forced, if skin = 0, or
//               if an atom moved more than skin / 2 since the last build
//               (the Verlet-skin rebuild rule of jax_md's neighbour list);
//   k_nl_mask   per tile pair (I, J > I) without exclusions: a 64-bit mask of
//               the atoms of J within cutoff + skin of tile I's box;
//   k_nl_rows   per tile I: prefix counts over its row's masks, the list
//               length, and (with a skin) the reference positions of the
//               tile's atoms;
//   k_nl_fill   per tile pair: its atoms' entries at the row's base plus the
//               pair's prefix (one wave per pair, so a long row is not one
//               wave's serial loop).
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

__global__ __launch_bounds__(kNlThreads) void k_nl_fill(DeviceView v) {
  const int w = wave_index();
  if (w >= v.nl_tp_hw) return;
  nl_fill_wave(v, w);
}

// ---- dense atom lists (the drop-in neighbour-list object) -----------------
// Per atom i, every atom j != i closer than radius, in increasing j, in the
// (N, K) layout of a jax_md Dense neighbour list (sentinel N after the last
// entry); counts[i] = the true number, which may exceed K (overflow).

// Box of tile t of raw (N, 3) positions: (lo, hi) at boxes[2t], [2t + 1].
__global__ __launch_bounds__(kNlThreads) void k_dense_boxes(const float* pos, int n, float4* boxes) {
  const int t = wave_index();
  const int n_tiles = (n + kTile - 1) / kTile;
  if (t >= n_tiles) return;
  const int lane = threadIdx.x & 63;
  const int i = min(t * kTile + lane, n - 1);
  float lx = pos[3 * i], ly = pos[3 * i + 1], lz = pos[3 * i + 2];
  float hx = lx, hy = ly, hz = lz;
  for (int off = 32; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off));
    ly = fminf(ly, __shfl_xor(ly, off));
    lz = fminf(lz, __shfl_xor(lz, off));
    hx = fmaxf(hx, __shfl_xor(hx, off));
    hy = fmaxf(hy, __shfl_xor(hy, off));
    hz = fmaxf(hz, __shfl_xor(hz, off));
  }
  if (lane == 0) boxes[2 * t] = make_float4(lx, ly, lz, 0.0f);
  if (lane == 1) boxes[2 * t + 1] = make_float4(hx, hy, hz, 0.0f);
}

// One wave per tile I; lane = atom i. Tiles J in increasing order, culled by
// box distance; the atoms of a kept J are staged in LDS and tested in order.
__global__ __launch_bounds__(64) void k_dense_list(const float* pos, int n, float r2, int K, const float4* boxes,
                                                   int32_t* idx, int32_t* counts) {
  __shared__ float4 pj[kTile];
  const int I = blockIdx.x;
  const int n_tiles = (n + kTile - 1) / kTile;
  const int lane = threadIdx.x;
  const int i = I * kTile + lane;
  const bool vi = i < n;
  const int ic = min(i, n - 1);
  const float xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
  const float4 li = boxes[2 * I], hi = boxes[2 * I + 1];
  int cnt = 0;
  for (int J = 0; J < n_tiles; ++J) {
    const float4 lj = boxes[2 * J], hj = boxes[2 * J + 1];
    const float bx = fmaxf(0.0f, fmaxf(lj.x - hi.x, li.x - hj.x));
    const float by = fmaxf(0.0f, fmaxf(lj.y - hi.y, li.y - hj.y));
    const float bz = fmaxf(0.0f, fmaxf(lj.z - hi.z, li.z - hj.z));
    if (bx * bx + by * by + bz * bz >= r2 * 1.00001f) continue;  // wave-uniform
    const int nJ = min(kTile, n - J * kTile);
    const int j0 = J * kTile + min(lane, nJ - 1);
    pj[lane] = make_float4(pos[3 * j0], pos[3 * j0 + 1], pos[3 * j0 + 2], 0.0f);
    __syncthreads();
    for (int b = 0; b < nJ; ++b) {
      const float4 q = pj[b];
      const float dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
      const int j = J * kTile + b;
      if (vi && j != i && dx * dx + dy * dy + dz * dz < r2) {
        if (cnt < K) idx[(size_t)i * K + cnt] = j;
        ++cnt;
      }
    }
    __syncthreads();
  }
  if (!vi) return;
  counts[i] = cnt;
  for (int c = cnt; c < K; ++c) idx[(size_t)i * K + c] = n;
}

}  // namespace

void launch_dense_list(const float* pos, int n, float radius, int capacity, int32_t* idx, int32_t* counts,
                       float4* boxes, hipStream_t s) {
  if (n <= 0) return;
  const int n_tiles = (n + kTile - 1) / kTile;
  const int wpb = kNlThreads / 64;
  hipLaunchKernelGGL(k_dense_boxes, dim3((n_tiles + wpb - 1) / wpb), dim3(kNlThreads), 0, s, pos, n, boxes);
  hipLaunchKernelGGL(k_dense_list, dim3(n_tiles), dim3(64), 0, s, pos, n, radius * radius, capacity, boxes, idx,
                     counts);
}

void launch_nl_build(const DeviceView& v, const int32_t* slot_ranges, int n_slots, bool force, hipStream_t s) {
  if (n_slots <= 0) return;
  hipLaunchKernelGGL(k_nl_check, dim3(n_slots), dim3(kNlThreads), 0, s, v, slot_ranges, force ? 1 : 0);
  const int wpb = kNlThreads / 64;
  if (v.nl_tp_hw > 0)
    hipLaunchKernelGGL(k_nl_mask, dim3((v.nl_tp_hw + wpb - 1) / wpb), dim3(kNlThreads), 0, s, v);
  if (v.n_tiles > 0) hipLaunchKernelGGL(k_nl_rows, dim3((v.n_tiles + wpb - 1) / wpb), dim3(kNlThreads), 0, s, v);
  if (v.nl_tp_hw > 0)
    hipLaunchKernelGGL(k_nl_fill, dim3((v.nl_tp_hw + wpb - 1) / wpb), dim3(kNlThreads), 0, s, v);
}

}  // namespace prx
