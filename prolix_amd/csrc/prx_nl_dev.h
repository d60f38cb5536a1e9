// Device code of the neighbour-list build, shared by prx_nl.hip (three
// kernels) and prx_persist.hip. See prx_nl.hip for the algorithm.
#pragma once

#include "prx_common.h"

namespace prx {

constexpr int kRangeInts = 14;  // slot range record: item_b, item_e, a0, n, then bonded ranges

constexpr int kNlThreads = 256;

__device__ __forceinline__ unsigned long long nl_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Listed system k, by one block of kNlThreads threads.
__device__ __forceinline__ void nl_check_block(const DeviceView& v, const int32_t* ranges, int k, int force) {
  __shared__ float sh[kNlThreads];
  const int32_t* R = ranges + k * kRangeInts;
  const int a0 = R[2], n = R[3];
  if (n <= 0) return;
  const int slot = v.slot_of[a0];
  const SlotConst& S = v.slots[slot];
  if (!S.nl) return;
  if (force || S.skin_half2 <= 0.0f) {
    if (threadIdx.x == 0) v.nl_flag[slot] = 1;
    return;
  }
  float m = 0.0f;
  auto take = [&](const float4& x, const float4& r) {
    const float dx = x.x - r.x, dy = x.y - r.y, dz = x.z - r.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    m = (d2 > m || d2 != d2) ? (d2 != d2 ? INFINITY : d2) : m;  // a non-finite position forces a build
  };
  // four atoms' loads in flight per thread: the scan is bound by load latency
  int i = threadIdx.x;
  for (; i + 3 * kNlThreads < n; i += 4 * kNlThreads) {
    float4 x[4], r[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      x[u] = v.pos[a0 + i + u * kNlThreads];
      r[u] = v.nl_xref[a0 + i + u * kNlThreads];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) take(x[u], r[u]);
  }
  for (; i < n; i += kNlThreads) take(v.pos[a0 + i], v.nl_xref[a0 + i]);
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = kNlThreads / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = fmaxf(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) v.nl_flag[slot] = sh[0] > S.skin_half2 ? 1 : 0;
}

// Tile-pair record w, by one wave.
__device__ __forceinline__ void nl_mask_wave(const DeviceView& v, int w) {
  const TilePairInfo tp = v.nl_tp[w];
  if (tp.ij < 0) return;  // a free record
  const bool excl = tp.slot < 0;
  const int slot = excl ? ~tp.slot : tp.slot;
  if (!v.nl_flag[slot]) return;
  const int lane = threadIdx.x & 63;
  if (excl) {  // an ordinary item covers this tile pair
    if (lane == 0) v.nl_mask[w] = 0ull;
    return;
  }
  const SlotConst& S = v.slots[slot];
  const int I = tp.ij & 0xFFFF, J = tp.ij >> 16;
  const int tI = S.tile0 + I, tJ = S.tile0 + J;
  const float4 li = v.tbox[2 * tI], hi = v.tbox[2 * tI + 1];
  const float4 lj = v.tbox[2 * tJ], hj = v.tbox[2 * tJ + 1];
  const float bx = fmaxf(0.0f, fmaxf(lj.x - hi.x, li.x - hj.x));
  const float by = fmaxf(0.0f, fmaxf(lj.y - hi.y, li.y - hj.y));
  const float bz = fmaxf(0.0f, fmaxf(lj.z - hi.z, li.z - hj.z));
  if (bx * bx + by * by + bz * bz >= S.nl_cut2) {  // wave-uniform: no atom of J is in range
    if (lane == 0) v.nl_mask[w] = 0ull;
    return;
  }
  const int nJ = min(kTile, S.n - J * kTile);
  const float4 p = v.pos[S.a0 + J * kTile + min(lane, nJ - 1)];
  const float gx = fmaxf(0.0f, fmaxf(li.x - p.x, p.x - hi.x));
  const float gy = fmaxf(0.0f, fmaxf(li.y - p.y, p.y - hi.y));
  const float gz = fmaxf(0.0f, fmaxf(li.z - p.z, p.z - hi.z));
  const bool in = lane < nJ && gx * gx + gy * gy + gz * gz < S.nl_cut2;
  const unsigned long long m = __ballot(in);
  if (lane == 0) v.nl_mask[w] = m;
}

// Tile t's row, by one wave: the exclusive prefix of the entry counts of its
// tile pairs (I, J > I) in J order, the row's total, and for a list with a
// skin the reference positions of the tile's atoms.
__device__ __forceinline__ void nl_rows_wave(const DeviceView& v, int t) {
  const int4 tl = v.tiles[t];
  const int slot = tl.z;
  const SlotConst& S = v.slots[slot];
  if (!S.nl || !v.nl_flag[slot]) return;
  const int lane = threadIdx.x & 63;
  const int I = tl.w;
  const int nT = (S.n + kTile - 1) / kTile;
  const int tp0 = S.tp_base + I * (nT - 1) - (I * (I - 1)) / 2;  // record of (I, I + 1)
  int running = 0;
  for (int J0 = I + 1; J0 < nT; J0 += 64) {
    const int J = J0 + lane;
    const bool valid = J < nT;
    const int c = valid ? __popcll(v.nl_mask[tp0 + (J - I - 1)]) : 0;
    int incl = c;  // inclusive prefix over the lanes, in lane order
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(incl, off);
      if (lane >= off) incl += y;
    }
    if (valid) v.nl_pref[tp0 + (J - I - 1)] = running + incl - c;
    running += __builtin_amdgcn_readlane(incl, 63);
  }
  if (lane == 0) v.nl_count[t] = running;
  if (S.skin_half2 > 0.0f && lane < tl.y) v.nl_xref[tl.x + lane] = v.pos[tl.x + lane];
}

// Tile-pair record w, by one wave: the atoms of J in I's list, written at
// I's entry base + the pair's prefix, in lane order -- so each row's entries
// run in increasing atom index.
__device__ __forceinline__ void nl_fill_wave(const DeviceView& v, int w) {
  const TilePairInfo tp = v.nl_tp[w];
  if (tp.ij < 0 || tp.slot < 0) return;  // a free record, or one an ordinary item covers
  if (!v.nl_flag[tp.slot]) return;
  const unsigned long long m = v.nl_mask[w];
  if (m == 0ull) return;
  const SlotConst& S = v.slots[tp.slot];
  const int lane = threadIdx.x & 63;
  const int I = tp.ij & 0xFFFF, J = tp.ij >> 16;
  const int E = v.nl_row[S.tile0 + I].y + v.nl_pref[w];
  if ((m >> lane) & 1ull) v.nl_entries[E + __popcll(m & nl_below(lane))] = S.a0 + J * kTile + lane;
}

}  // namespace prx
