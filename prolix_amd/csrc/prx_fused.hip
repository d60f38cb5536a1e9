// Fused step kernel for small systems: one workgroup per system runs K BAOAB
// steps of that system, its passes separated by workgroup barriers only.
//
// For a system of a few hundred atoms a step of the graph path is six
// dependent launches of almost empty grids, so its time is launch and
// dependency latency (~90 us at 138 atoms against a few us of work), and the
// persistent kernel pays a grid-wide barrier per pass instead. Here a
// workgroup owns all the work of its system: the system's work items over its
// eight waves, its tiles' gathers two at a time (one group of four waves per
// tile, each group with its own workgroup memory), its bonded entries over its
// threads. (Sixteen waves would cap the registers at 128 and spill.) No other workgroup reads or writes that system's atoms,
// items or records, so __syncthreads (a workgroup-scope fence and barrier;
// all waves of a workgroup share one CU and its vector cache) orders every
// pass. Systems run side by side, one workgroup each, with no barrier
// between them.
//
// A tile gather holds exactly one workgroup barrier (its combine of the four
// waves' sums). In a round of two gathers, a group whose tile index lies
// past the system's last tile executes one barrier in its place, so every
// wave of the workgroup meets the same number of barriers.
//
// Each pass calls the per-item and per-tile device functions of the graph's
// kernels (prx_*_dev.h) on the same items, in the same per-item and per-tile
// order, so every result is bit for bit the graph path's. Systems that take
// their pair work from a neighbour list are not run here.
#include "prx_atom_dev.h"
#include "prx_bonded_dev.h"
#include "prx_nl_dev.h"
#include "prx_pair_dev.h"

namespace prx {

namespace {

constexpr int kFusedGroups = 2;                  // tile gathers side by side
constexpr int kFusedWaves = kGW * kFusedGroups;  // 8
constexpr int kFusedThreads = kFusedWaves * 64;

// The bonded entries of one system (its ranges in each kind's arena).
__device__ __forceinline__ void bonded_slot(const DeviceView& v, const int32_t* R) {
  for (int e = R[4] + (int)threadIdx.x; e < R[5]; e += kFusedThreads) do_bond<false>(v, e);
  for (int e = R[6] + (int)threadIdx.x; e < R[7]; e += kFusedThreads) do_angle<false>(v, e);
  for (int e = R[8] + (int)threadIdx.x; e < R[9]; e += kFusedThreads) do_torsion<false>(v, e);
  for (int e = R[10] + (int)threadIdx.x; e < R[11]; e += kFusedThreads) do_cmap<false>(v, e);
  for (int e = R[12] + (int)threadIdx.x; e < R[13]; e += kFusedThreads) do_exception<false>(v, e);
}

// One round of tile gathers: group g of four waves gathers tile t + g, or,
// past the last tile, takes the one barrier that gather would have held.
// Then the round's barrier, before the workgroup memory is used again.
template <class F>
__device__ __forceinline__ void gather_round(int t, int t1, int grp, F&& gather) {
  if (t + grp < t1)
    gather(t + grp);
  else
    __syncthreads();
  __syncthreads();
}

// Block b: the b-th listed system, ranges row b (item_b, item_e, a0, n, then
// the bonded ranges).
__global__ __launch_bounds__(kFusedThreads) void k_fused_steps(DeviceView v, const int32_t* ranges, int n_steps) {
  __shared__ PassLds lds[kFusedWaves];
  __shared__ GatherLds glds[kFusedGroups];
  PassLds& L = lds[threadIdx.x >> 6];
  const int32_t* R = ranges + (size_t)blockIdx.x * kRangeInts;
  const int item_b = R[0], item_e = R[1], a0 = R[2], n = R[3];
  if (n <= 0) return;
  const int t0 = v.slots[v.slot_of[a0]].tile0;
  const int t1 = t0 + (n + kTile - 1) / kTile;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int grp = wave / kGW;
  GatherLds& G = glds[grp];

  // B A O A of the first step from the stored force, then the tile boxes
  for (int g = a0 + (int)threadIdx.x; g < a0 + n; g += kFusedThreads) step_begin_atom(v, g);
  __syncthreads();
  for (int t = t0 + wave; t < t1; t += kFusedWaves) tile_box_wave(v, t);
  __syncthreads();

  for (int k = 0; k < n_steps; ++k) {
    for (int w = item_b + wave; w < item_e; w += kFusedWaves) born_item(v, w, L.born);
    bonded_slot(v, R);
    __syncthreads();
    for (int t = t0; t < t1; t += kFusedGroups)
      gather_round(t, t1, grp, [&](int tt) { born_gather_tile(v, tt, G); });
    for (int w = item_b + wave; w < item_e; w += kFusedWaves) force_item<false>(v, w, L.pair);
    __syncthreads();
    for (int t = t0; t < t1; t += kFusedGroups)
      gather_round(t, t1, grp, [&](int tt) { force_gather_tile<false>(v, tt, G); });
    for (int w = item_b + wave; w < item_e; w += kFusedWaves) chain_item(v, w, L.pair);
    __syncthreads();
    const int mode = k + 1 == n_steps ? 2 : 1;  // the last step stops after B(F)
    for (int t = t0; t < t1; t += kFusedGroups)
      gather_round(t, t1, grp, [&](int tt) { chain_gather_tile(v, tt, mode, G); });
  }
}

}  // namespace

void launch_fused_steps(const DeviceView& v, const int32_t* slot_ranges, int n_slots, int n_steps, hipStream_t s) {
  // every tile gather must reach its barrier: no stopped-system early return here
  DeviceView fv = v;
  fv.skip_done = 0;
  if (n_slots > 0 && n_steps > 0)
    hipLaunchKernelGGL(k_fused_steps, dim3(n_slots), dim3(kFusedThreads), 0, s, fv, slot_ranges, n_steps);
}

}  // namespace prx
