// All-pairs passes of the GB-OBC2 energy/force evaluation on gfx950.
//
// One wave64 per work item (a 64x64 tile pair I <= J of one system). Lane l
// owns i atom I0 + l. The j tile's atom data is staged in the wave's own LDS
// slice twice over (entries m and m + 64 hold j atom m), so at iteration k
// lane l reads entry l + k -- j atom (l + k) mod 64 -- through an immediate
// LDS offset; the j-side accumulators live in registers and rotate one lane
// per iteration with a DPP wave rotate, so every lane meets every j exactly
// once and no two lanes touch the same j in one iteration. (A DPP rotate
// issues at about two thirds of the FMA rate on gfx950; the LDS reads do not
// take VALU issue slots.) After the loop the wave stores
// its i and j partial sums in the item's own slots of the partial buffer
// (plain coalesced stores, no atomics); the tile gathers of prx_atom.hip add
// them up in a fixed order, so results are bitwise reproducible.
//
// Pass 1 (Born sums, generalized_born.py:211-263 with :32-94):
//     I_i += I(r; rho_i, sr_j),  I_j += I(r; rho_j, sr_i)
// Pass 2 (GB + Coulomb + LJ forces and dE/dB, generalized_born.py:97-208,
//     optimization.py:21-297): GB over every pair, Coulomb/LJ with exclusion
//     scales and the optional cutoff.
// Pass 3 (Born chain rule, generalized_born.py:651-706):
//     F -= (c_i dI_ij/dr + c_j dI_ji/dr) dr/dx, c = dE/dB * dB/dI.
//
// Diagonal tiles (I == J) visit each unordered pair once: offsets k = 1..31
// on every lane and k = 32 on lanes 0..31 only.
//
// Systems with a cutoff: the Born pass classifies each item from its two
// tiles' bounding boxes (near / straddle / far, stored for the later passes
// and the gathers). Far items skip LJ and Coulomb, and with the neighbour-list
// GB form (generalized_born.py:266-415, gb_cut) skip every pass; straddling
// items test r^2 < cutoff^2 per pair.
#include "prx_bonded_dev.h"
#include "prx_pair_dev.h"

namespace prx {
__global__ __launch_bounds__(kPairThreads) void k_born_pairs(DeviceView v) {
  __shared__ BornLds lds[kWavesPerBlock];
  const int w = wave_id();
  if (w >= v.n_items) return;
  born_item(v, w, lds[threadIdx.x >> 6]);
}

template <bool ENERGY>
__global__ __launch_bounds__(kPairThreads) PRX_OCCUPANCY void k_force_pairs(DeviceView v) {
  __shared__ PairLds lds[kWavesPerBlock];
  const int w = wave_id();
  if (w >= v.n_items) return;
  force_item<ENERGY>(v, w, lds[threadIdx.x >> 6]);
}

__global__ __launch_bounds__(kPairThreads) PRX_OCCUPANCY void k_chain_pairs(DeviceView v) {
  __shared__ PairLds lds[kWavesPerBlock];
  const int w = wave_id();
  if (w >= v.n_items) return;
  chain_item(v, w, lds[threadIdx.x >> 6]);
}

// The Born pass with the bonded terms in the same launch: blocks below
// n_pair are the pair kernel's (its XCD-aware item order over those blocks),
// the rest take one bonded entry per thread. Both read positions only and the
// bonded records are read by the chain gather, so the merged launch saves the
// bonded kernel's graph node where that kernel would run in line; results are
// bit for bit those of the two kernels.
__device__ __forceinline__ int wave_id_of(int b, int nb) {
#if PRX_XCD_MAP
  const int per = nb / kXcds, rem = nb % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  b = x * per + min(x, rem) + k;
#endif
  return __builtin_amdgcn_readfirstlane((int)(b * kWavesPerBlock + (threadIdx.x >> 6)));
}

__global__ __launch_bounds__(kPairThreads) void k_born_pairs_bonded(DeviceView v, int n_pair) {
  __shared__ BornLds lds[kWavesPerBlock];
  const int b = (int)blockIdx.x;
  if (b < n_pair) {
    const int w = wave_id_of(b, n_pair);
    if (w >= v.n_items) return;
    born_item(v, w, lds[threadIdx.x >> 6]);
    return;
  }
  const int e = (b - n_pair) * kPairThreads + (int)threadIdx.x;
  if (e < bonded_entries(v)) bonded_entry<false>(v, e);
}

static inline int pair_blocks(int n_items) { return (n_items + kWavesPerBlock - 1) / kWavesPerBlock; }

void launch_born_pairs_bonded(const DeviceView& v, hipStream_t s) {
  const int np = v.n_items ? pair_blocks(v.n_items) : 0;
  const int nb = (bonded_entries(v) + kPairThreads - 1) / kPairThreads;
  if (np + nb > 0) hipLaunchKernelGGL(k_born_pairs_bonded, dim3(np + nb), dim3(kPairThreads), 0, s, v, np);
}

void launch_born_pairs(const DeviceView& v, hipStream_t s) {
  if (v.n_items == 0) return;
  hipLaunchKernelGGL(k_born_pairs, dim3(pair_blocks(v.n_items)), dim3(kPairThreads), 0, s, v);
}

void launch_force_pairs(const DeviceView& v, bool energy, hipStream_t s) {
  if (v.n_items == 0) return;
  if (energy)
    hipLaunchKernelGGL(k_force_pairs<true>, dim3(pair_blocks(v.n_items)), dim3(kPairThreads), 0, s, v);
  else
    hipLaunchKernelGGL(k_force_pairs<false>, dim3(pair_blocks(v.n_items)), dim3(kPairThreads), 0, s, v);
}

void launch_chain_pairs(const DeviceView& v, hipStream_t s) {
  if (v.n_items == 0) return;
  hipLaunchKernelGGL(k_chain_pairs, dim3(pair_blocks(v.n_items)), dim3(kPairThreads), 0, s, v);
}

}  // namespace prx
