// Device code of the per-atom passes, shared by prx_atom.hip (one kernel per
// pass) and prx_persist.hip (every pass of a step in one persistent kernel):
// Philox noise, the tile gathers, the Born / force / chain finishes, the
// BAOAB halves and the tile boxes. Each *_tile function is the work of one
// 256-thread block for one tile; tile_box_wave is one wave's.
#pragma once

#include "prx_common.h"

namespace prx {


// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller normals.
// ---------------------------------------------------------------------------
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)M0 * c.x;
    const uint64_t p1 = (uint64_t)M1 * c.z;
    U4 n;
    n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
    n.y = (uint32_t)p1;
    n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
    n.w = (uint32_t)p0;
    c = n;
    k0 += W0;
    k1 += W1;
  }
  return c;
}

// Three standard normals for (atom, step) of one system's stream. The
// counter is (local atom index, step, 0, domain); domain 0 is the
// O-step noise and domain 1 the initial momenta.
__device__ __forceinline__ float3 normal3(uint64_t seed, uint32_t step, int local, uint32_t domain) {
  U4 c = {(uint32_t)local, step, 0u, domain};
  const U4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float inv = 2.3283064365386963e-10f;  // 2^-32
  const float u1 = ((float)r.x + 0.5f) * inv;
  const float u2 = ((float)r.y + 0.5f) * inv;
  const float u3 = ((float)r.z + 0.5f) * inv;
  const float u4 = ((float)r.w + 0.5f) * inv;
  const float ra = sqrtf(-2.0f * logf(u1));
  const float rb = sqrtf(-2.0f * logf(u3));
  float s, co;
  sincosf(6.283185307179586f * u2, &s, &co);
  float3 o;
  o.x = ra * co;
  o.y = ra * s;
  o.z = rb * cosf(6.283185307179586f * u4);
  return o;
}

// ---------------------------------------------------------------------------
// Tile gathers: one 256-thread block (four waves) per 64-atom tile T of a
// system with nT tiles. The tile's atoms received i-side partials from the
// row items (T, J), J >= T, and j-side partials from the column items
// (I, T), I <= T; items are stored row-major per system, so item (I, J) =
// base + I nT - I (I - 1) / 2 + (J - I). Wave q sums the row items
// J = T + q + 4n and the column items I = q + 4n in that order, in FP64; the
// four wave sums are then added in the order q = 0, 1, 2, 3 through LDS, so
// every result is a fixed function of the item partials.
// ---------------------------------------------------------------------------
constexpr int kGW = 4;  // waves per tile gather

struct TileCtx {
  int g;       // this lane's atom (arena index)
  bool valid;  // lane < atoms in tile
  int slot, T, nT, base, lane, q;
  int t;          // global tile index
  bool skip_far;  // gb_cut system: far items carry no partials
  bool nl;        // the system's pair work comes from its neighbour list
};

__device__ __forceinline__ TileCtx tile_ctx(const DeviceView& v, int t) {
  const int4 tl = v.tiles[t];
  TileCtx c;
  c.lane = threadIdx.x & 63;
  c.q = __builtin_amdgcn_readfirstlane((int)((threadIdx.x >> 6) & (kGW - 1)));  // wave within its group of kGW
  c.g = tl.x + c.lane;
  c.valid = c.lane < tl.y;
  c.slot = tl.z;
  c.T = tl.w;
  const SlotConst& S = v.slots[c.slot];
  c.nT = (S.n + kTile - 1) / kTile;
  c.base = S.item_base;
  c.skip_far = S.gb_cut != 0;
  c.nl = S.nl != 0;
  c.t = t;
  return c;
}

__device__ __forceinline__ int row_start(const TileCtx& c, int I) { return c.base + I * c.nT - (I * (I - 1)) / 2; }

__device__ __forceinline__ size_t pslot(int w, int side, int lane) { return ((size_t)w * 2 + side) * kTile + lane; }

// Calls add(item, side) for this wave's items in their fixed order. Systems
// whose far items carry no partials test 64 candidates per vector load of
// the item classes and visit only the active ones.
template <class F>
__device__ __forceinline__ void for_tile_items(const DeviceView& v, const TileCtx& c, F&& add) {
  const int r0 = row_start(c, c.T);
  const int n_row = max(0, (c.nT - c.T - c.q + kGW - 1) / kGW);
  const int n_col = max(0, (c.T + 1 - c.q + kGW - 1) / kGW);
  auto col_item = [&](int n) {
    const int I = c.q + kGW * n;
    return row_start(c, I) + (c.T - I);
  };
  if (!c.skip_far) {
    for (int n = 0; n < n_row; ++n) add(r0 + c.q + kGW * n, 0);
    for (int n = 0; n < n_col; ++n) add(col_item(n), 1);
    return;
  }
  for (int g0 = 0; g0 < n_row; g0 += 64) {
    const int n = g0 + c.lane;
    const bool act = n < n_row && (v.item_cls[r0 + c.q + kGW * n] & kClsCutMask) != kClsFar;
    for (unsigned long long m = __ballot(act); m; m &= m - 1) add(r0 + c.q + kGW * (g0 + __builtin_ctzll(m)), 0);
  }
  for (int g0 = 0; g0 < n_col; g0 += 64) {
    const int n = g0 + c.lane;
    const bool act = n < n_col && (v.item_cls[col_item(min(n, n_col - 1))] & kClsCutMask) != kClsFar;
    for (unsigned long long m = __ballot(act); m; m &= m - 1) add(col_item(g0 + __builtin_ctzll(m)), 1);
  }
}

// Like for_tile_items, but with every item's partial loaded by load(item,
// side) and added by add(value): systems without far-item skipping load four
// items' partials before adding them, in the same item order, so each lane
// keeps up to four loads in flight.
__device__ __forceinline__ unsigned long long lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ __forceinline__ unsigned long long shfl64(unsigned long long x, int src) {
  const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)x, src), hi = (uint32_t)__shfl((int)(uint32_t)(x >> 32), src);
  return ((unsigned long long)hi << 32) | lo;
}

// The partials of a neighbour-list system's tile T (prx_nl.hip), for wave q,
// in a fixed order: the row's ordinary items (side 0) and list chunks
// (side 0), then the column's ordinary items (side 1), then the entries of
// T's atoms in the lists of the rows I < T, I = q, q + 4, ... increasing
// (side 1, at the entry's chunk and lane). Four loads in flight per lane.
template <class T, class Load, class Add>
__device__ __forceinline__ void gather_nl(const DeviceView& v, const TileCtx& c, Load&& load, Add&& add) {
  constexpr int kBatch = 4;
  const int4 row = v.nl_row[c.t];  // first chunk item, entry base, first / end ordinary row item
  const int n_reg = row.w - row.z;
  const int n_ch = (v.nl_count[c.t] + kTile - 1) / kTile;
  const int n_row = n_reg + n_ch;
  for (int k = c.q; k < n_row; k += kGW) {
    const int w = k < n_reg ? row.z + k : row.x + (k - n_reg);
    if (k < n_reg && (v.item_cls[w] & kClsCutMask) == kClsFar) continue;  // a far item wrote nothing
    add(load(w, 0, c.lane));
  }
  const int2 col = v.nl_col[c.t];
  for (int k = c.q; k < col.y; k += kGW) {
    const int w = v.nl_colitems[col.x + k];
    if ((v.item_cls[w] & kClsCutMask) == kClsFar) continue;
    add(load(w, 1, c.lane));
  }
  const SlotConst& S = v.slots[c.slot];
  const int nI = max(0, (c.T - c.q + kGW - 1) / kGW);  // rows I = q + 4n < T
  for (int n0 = 0; n0 < nI; n0 += 64) {
    const int n = n0 + c.lane;
    const int I = c.q + kGW * n;
    unsigned long long mk = 0ull;
    int pref = 0, chunk0 = 0;
    if (n < nI) {
      const int tp = S.tp_base + I * (c.nT - 1) - (I * (I - 1)) / 2 + (c.T - I - 1);
      mk = v.nl_mask[tp];
      pref = v.nl_pref[tp];
      chunk0 = v.nl_row[S.tile0 + I].x;
    }
    const int cnt = min(64, nI - n0);
    for (int k = 0; k < cnt; k += kBatch) {
      T val[kBatch];
      bool hit[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int kk = min(k + u, cnt - 1);
        const unsigned long long mu = shfl64(mk, kk);
        const int pu = __shfl(pref, kk), cu = __shfl(chunk0, kk);
        hit[u] = k + u < cnt && ((mu >> c.lane) & 1ull);
        if (hit[u]) {
          const int p = pu + __popcll(mu & lanes_below(c.lane));
          val[u] = load(cu + p / kTile, 1, p % kTile);
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u)
        if (hit[u]) add(val[u]);
    }
  }
}

template <class T, class Load, class Add>
__device__ __forceinline__ void gather_items(const DeviceView& v, const TileCtx& c, Load&& load, Add&& add) {
  if (c.nl) {
    gather_nl<T>(v, c, load, add);
    return;
  }
  if (c.skip_far) {
    for_tile_items(v, c, [&](int w, int side) { add(load(w, side, c.lane)); });
    return;
  }
  constexpr int kBatch = 4;
  const int r0 = row_start(c, c.T);
  const int n_row = max(0, (c.nT - c.T - c.q + kGW - 1) / kGW);
  const int n_col = max(0, (c.T + 1 - c.q + kGW - 1) / kGW);
  for (int n = 0; n < n_row; n += kBatch) {
    T val[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (n + u < n_row) val[u] = load(r0 + c.q + kGW * (n + u), 0, c.lane);
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (n + u < n_row) add(val[u]);
  }
  for (int n = 0; n < n_col; n += kBatch) {
    T val[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (n + u < n_col) {
        const int I = c.q + kGW * (n + u);
        val[u] = load(row_start(c, I) + (c.T - I), 1, c.lane);
      }
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (n + u < n_col) add(val[u]);
  }
}

// Scalar partials (Born sums) of this wave's items for the lane's atom.
__device__ __forceinline__ double gather1(const DeviceView& v, const TileCtx& c) {
  const float* pb = reinterpret_cast<const float*>(v.part);
  double acc = 0.0;
  gather_items<float>(v, c, [&](int w, int side, int lane) { return pb[pslot(w, side, lane)]; },
                      [&](float x) { acc += (double)x; });
  return acc;
}

// Chain partials: three float arrays of 64 per (item, side).
__device__ __forceinline__ double4 gather3(const DeviceView& v, const TileCtx& c) {
  const float* pc = reinterpret_cast<const float*>(v.part);
  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  gather_items<float3>(
      v, c,
      [&](int w, int side, int lane) {
        const float* q = pc + ((size_t)w * 2 + side) * (3 * kTile) + lane;
        return make_float3(q[0], q[kTile], q[2 * kTile]);
      },
      [&](float3 p) { acc.x += p.x; acc.y += p.y; acc.z += p.z; });
  return acc;
}

__device__ __forceinline__ double4 gather4(const DeviceView& v, const TileCtx& c) {
  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  gather_items<float4>(v, c, [&](int w, int side, int lane) { return v.part[pslot(w, side, lane)]; },
                       [&](float4 p) { acc.x += p.x; acc.y += p.y; acc.z += p.z; acc.w += p.w; });
  return acc;
}

// Workgroup memory of one tile gather: the four waves' sums and the bonded
// shares. The kernel that runs the gathers owns it; a workgroup that runs
// several tile gathers side by side gives each group of four waves its own.
struct GatherLds {
  double4 sh4[kGW][kTile];
  float4 shb[kGW][kTile];
};

// The four waves' sums, added in wave order; valid in wave 0 after the call.
// Every thread of the block must call it (it holds a barrier).
__device__ __forceinline__ double combine1(double x, const TileCtx& c, GatherLds& G) {
  double(*sh)[kTile] = reinterpret_cast<double(*)[kTile]>(&G.sh4[0][0]);
  sh[c.q][c.lane] = x;
  __syncthreads();
  return ((sh[0][c.lane] + sh[1][c.lane]) + sh[2][c.lane]) + sh[3][c.lane];
}

__device__ __forceinline__ double4 combine4(double4 x, const TileCtx& c, GatherLds& G) {
  double4(*sh4)[kTile] = G.sh4;
  sh4[c.q][c.lane] = x;
  __syncthreads();
  double4 r = sh4[0][c.lane];
#pragma unroll
  for (int k = 1; k < kGW; ++k) {
    const double4 y = sh4[k][c.lane];
    r.x += y.x; r.y += y.y; r.z += y.z; r.w += y.w;
  }
  return r;
}

// A tile of a system at its step limit, in a step graph
// (DeviceView::skip_done): its pair items were skipped, so its gathers are
// too, and its Born radii, pair force and stored force stay those of its
// final positions. Uniform over the block.
__device__ __forceinline__ bool tile_done(const DeviceView& v, int t) {
  if (!v.skip_done) return false;
  const int4 tl = v.tiles[t];
  const SlotConst& S = v.slots[tl.z];
  return S.step_limit != 0u && slot_done(S, v.atom_step[tl.x]);
}

// B_i from I_i in FP64 (generalized_born.py:81-93); dB/dI for the chain rule.
// Neighbour-list GB systems floor 1/B at 1e-4 (generalized_born.py:333-337).
__device__ __forceinline__ void born_gather_tile(const DeviceView& v, int t, GatherLds& G) {
  if (tile_done(v, t)) return;
  const TileCtx c = tile_ctx(v, t);
  const double I = combine1(gather1(v, c), c, G);
  if (c.q != 0 || !c.valid) return;
  const double4 f = v.fin4[c.g];  // rho, R, q^2, ace K
  const double psi = f.x * I;
  const double tt = psi * (1.0 + psi * (-0.8 + 4.85 * psi));
  const double th = tanh(tt);
  double inv_b = 1.0 / f.x - th / f.y;
  const bool floored = c.skip_far && inv_b < 1e-4;
  if (floored) inv_b = 1e-4;
  const double B = 1.0 / inv_b;
  const double dBdI = floored ? 0.0 : B * B * (1.0 - th * th) * (1.0 + psi * (-1.6 + 14.55 * psi)) * f.x / f.y;
  v.gb4[c.g] = make_float4((float)B, (float)(0.6005612043932249 / B), (float)dBdI, 0.0f);  // sqrt(log2(e)/4)/B
}

// Pair force and dE/dB; c_i = dE/dB_i * dB_i/dI_i with the GB self term and ACE (FP64).
template <bool ENERGY>
__device__ __forceinline__ void force_gather_tile(const DeviceView& v, int t, GatherLds& G) {
  if (tile_done(v, t)) return;
  const TileCtx c = tile_ctx(v, t);
  const double4 a = combine4(gather4(v, c), c, G);
  if (c.q != 0 || !c.valid) return;
  const double4 f = v.fin4[c.g];
  const float4 gb = v.gb4[c.g];
  const double B = (double)gb.x;
  const double pref = v.slots[c.slot].gb_pref;
  const double e_self = pref * f.z / B;
  const double e_ace = f.w / (B * B * B * B * B * B);
  const double dEdB = a.w - e_self / B - 6.0 * e_ace / B;
  v.gb4[c.g].w = (float)(dEdB * (double)gb.z);
  v.fpair[c.g] = make_float4((float)a.x, (float)a.y, (float)a.z, 0.0f);
  if (ENERGY) {
    v.e_atom[(size_t)c.g * 2 + 0] = e_self;
    v.e_atom[(size_t)c.g * 2 + 1] = e_ace;
  }
}

// Bonded records q, q + 4, ... of atom g (wave q's share), added in list
// order. Four records per round: their list indices, then the records, are
// loaded before any is added, so a lane has up to four loads in flight.
__device__ __forceinline__ float3 bonded_force(const DeviceView& v, int g, int q) {
  const int2 lr = v.atom_rec[g];
  float bx = 0.0f, by = 0.0f, bz = 0.0f;
  constexpr int kBatch = 4;
  for (int k = q; k < lr.y; k += kBatch * kGW) {
    int idx[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int kk = k + u * kGW;
      idx[u] = kk < lr.y ? v.rec_list[lr.x + kk] : -1;
    }
    float4 r[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) r[u] = idx[u] >= 0 ? v.rec[idx[u]] : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      if (idx[u] < 0) break;
      bx += r[u].x;
      by += r[u].y;
      bz += r[u].z;
    }
  }
  return make_float3(bx, by, bz);
}

// BAOAB first half (settle.py:2064-2087): B(F_n) A O A for atom g of slot s.
__device__ __forceinline__ float4 begin_step(const DeviceView& v, const SlotConst& S, int s, int g, float4& p,
                                             uint32_t step) {
  const float4 f = v.force[g];
  p.x = fmaf(S.half_dt, f.x, p.x);
  p.y = fmaf(S.half_dt, f.y, p.y);
  p.z = fmaf(S.half_dt, f.z, p.z);
  if (S.caps) {  // velocity limit (batched_simulate.py:273-278)
    const float vm = sqrtf(p.x * p.x + p.y * p.y + p.z * p.z + 1e-30f) * p.w;
    if (vm > kVLimit) {
      const float sc = kVLimit / vm;
      p.x *= sc; p.y *= sc; p.z *= sc;
      atomicAdd(&v.warn[s * kWarnStride + kWarnVLimit], 1u);
    }
  }
  float4 x = v.pos[g];
  const float4 x_old = x;
  const float hv = S.half_dt * p.w;
  x.x = fmaf(hv, p.x, x.x);
  x.y = fmaf(hv, p.y, x.y);
  x.z = fmaf(hv, p.z, x.z);
  const float3 xi = normal3(S.seed, step, g - S.a0, 0u);
  const float amp = S.c2 * S.sqrt_kT * v.sqrt_m[g];
  p.x = fmaf(S.c1, p.x, amp * xi.x);
  p.y = fmaf(S.c1, p.y, amp * xi.y);
  p.z = fmaf(S.c1, p.z, amp * xi.z);
  x.x = fmaf(hv, p.x, x.x);
  x.y = fmaf(hv, p.y, x.y);
  x.z = fmaf(hv, p.z, x.z);
  if (S.caps) {  // displacement cap, momentum rescaled to match (batched_simulate.py:290-303)
    const float dx = x.x - x_old.x, dy = x.y - x_old.y, dz = x.z - x_old.z;
    const float dm = sqrtf(dx * dx + dy * dy + dz * dz + 1e-30f);
    if (dm > kMaxDx) {
      const float sc = kMaxDx / dm;
      x.x = x_old.x + dx * sc; x.y = x_old.y + dy * sc; x.z = x_old.z + dz * sc;
      p.x *= sc; p.y *= sc; p.z *= sc;
      atomicAdd(&v.warn[s * kWarnStride + kWarnDxCap], 1u);
    }
  }
  v.pos[g] = x;
  return x;
}

// Total force = chain partials + pair force + bonded records for atom c.g;
// then mode -1: store it; 2: B(F) and count the step; 1: B(F), then B A O A.
// Returns the atom's position after the call.
__device__ __forceinline__ float4 finish_atom(const DeviceView& v, const TileCtx& c, const double4& a,
                                              const float4 (*shb)[kTile], int mode) {
  float3 fb = make_float3(0.f, 0.f, 0.f);
#pragma unroll
  for (int k = 0; k < kGW; ++k) {
    const float4 y = shb[k][c.lane];
    fb.x += y.x; fb.y += y.y; fb.z += y.z;
  }
  const float4 fp = v.fpair[c.g];
  float3 f = make_float3((float)(a.x + (double)fp.x) + fb.x, (float)(a.y + (double)fp.y) + fb.y,
                         (float)(a.z + (double)fp.z) + fb.z);
  const SlotConst& S = v.slots[c.slot];
  if (S.caps) {
    // non-finite -> 0, then |F| <= kForceCap (batched_simulate.py:323-330). Also in
    // mode -1, so a force recomputed at unchanged positions is the one the step stored.
    if (!isfinite(f.x) || !isfinite(f.y) || !isfinite(f.z)) f = make_float3(0.f, 0.f, 0.f);
    const float fm = sqrtf(f.x * f.x + f.y * f.y + f.z * f.z + 1e-12f);
    if (fm > kForceCap) {
      const float sc = kForceCap / fm;
      f = make_float3(f.x * sc, f.y * sc, f.z * sc);
      if (mode >= 0) atomicAdd(&v.warn[c.slot * kWarnStride + kWarnForceCap], 1u);
    }
  }
  v.force[c.g] = make_float4(f.x, f.y, f.z, 0.0f);
  if (mode < 0 || !S.active) return v.pos[c.g];
  const uint32_t prev = v.atom_step[c.g];  // each lane owns its atom's counter
  if (slot_done(S, prev)) return v.pos[c.g];
  float4 p = v.mom[c.g];  // w = 1/m
  p.x = fmaf(S.half_dt, f.x, p.x);
  p.y = fmaf(S.half_dt, f.y, p.y);
  p.z = fmaf(S.half_dt, f.z, p.z);
  const uint32_t step = prev + 1u;
  v.atom_step[c.g] = step;
  float4 x = mode == 1 && !slot_done(S, step) ? begin_step(v, S, c.slot, c.g, p, step) : v.pos[c.g];
  v.mom[c.g] = p;
  // explosion check: count atom-steps that leave a non-finite position or momentum
  if (!(isfinite(x.x) && isfinite(x.y) && isfinite(x.z) && isfinite(p.x) && isfinite(p.y) && isfinite(p.z)))
    atomicAdd(&v.warn[c.slot * kWarnStride + kWarnNonFinite], 1u);
  return x;
}

// Per tile (one block): the chain partials, bonded records and pair force
// finish each atom's force (finish_atom). In mode 1 the tile's atoms have
// moved to the next step's positions, and wave 0 also writes the tile's box
// for that step (what k_tile_boxes computes), so the step graph needs no box
// kernel after its first step.
__device__ __forceinline__ void chain_gather_tile(const DeviceView& v, int t, int mode, GatherLds& G) {
  if (tile_done(v, t)) return;
  const TileCtx c = tile_ctx(v, t);
  double4 part = gather3(v, c);
  float3 fb = c.valid ? bonded_force(v, c.g, c.q) : make_float3(0.f, 0.f, 0.f);
  // the bonded share travels in the same combine (its float sum is exact in FP64)
  G.shb[c.q][c.lane] = make_float4(fb.x, fb.y, fb.z, 0.0f);
  const double4 a = combine4(part, c, G);
  if (c.q != 0) return;
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  if (c.valid) x = finish_atom(v, c, a, G.shb, mode);
  if (mode != 1 || !v.slots[c.slot].active) return;  // uniform over the tile
  const int last = v.tiles[t].y - 1;
  const float xl = __shfl(x.x, last), yl = __shfl(x.y, last), zl = __shfl(x.z, last);
  if (!c.valid) x = make_float4(xl, yl, zl, 0.f);
  const float4 b = v.born4[c.g - c.lane + min(c.lane, last)];
  float lx = x.x, ly = x.y, lz = x.z, hx = x.x, hy = x.y, hz = x.z, mr = b.x, ms = b.y;
  for (int off = 32; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off));
    ly = fminf(ly, __shfl_xor(ly, off));
    lz = fminf(lz, __shfl_xor(lz, off));
    hx = fmaxf(hx, __shfl_xor(hx, off));
    hy = fmaxf(hy, __shfl_xor(hy, off));
    hz = fmaxf(hz, __shfl_xor(hz, off));
    mr = fmaxf(mr, __shfl_xor(mr, off));
    ms = fmaxf(ms, __shfl_xor(ms, off));
  }
  if (c.lane == 0) v.tbox[2 * t] = make_float4(lx, ly, lz, mr);
  if (c.lane == 1) v.tbox[2 * t + 1] = make_float4(hx, hy, hz, ms);
}

// B A O A from the stored force, for the first step of a launch.
__device__ __forceinline__ void step_begin_atom(const DeviceView& v, int g) {
  if (g >= v.atom_hw) return;
  const int s = v.slot_of[g];
  if (s < 0) return;
  const SlotConst& S = v.slots[s];
  const uint32_t step = v.atom_step[g];
  if (!S.active || slot_done(S, step)) return;
  float4 p = v.mom[g];
  begin_step(v, S, s, g, p, step);
  v.mom[g] = p;
}

// Bounding box of tile t's atoms at the current positions, with the tile's
// largest offset radius rho and descreening radius s; one wave.
__device__ __forceinline__ void tile_box_wave(const DeviceView& v, int t) {
  const int4 tl = v.tiles[t];
  const int lane = threadIdx.x & 63;
  const int g = tl.x + min(lane, tl.y - 1);
  const float4 p = v.pos[g], b = v.born4[g];
  float lx = p.x, ly = p.y, lz = p.z, hx = p.x, hy = p.y, hz = p.z, mr = b.x, ms = b.y;
  for (int off = 32; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off));
    ly = fminf(ly, __shfl_xor(ly, off));
    lz = fminf(lz, __shfl_xor(lz, off));
    hx = fmaxf(hx, __shfl_xor(hx, off));
    hy = fmaxf(hy, __shfl_xor(hy, off));
    hz = fmaxf(hz, __shfl_xor(hz, off));
    mr = fmaxf(mr, __shfl_xor(mr, off));
    ms = fmaxf(ms, __shfl_xor(ms, off));
  }
  if (lane == 0) v.tbox[2 * t] = make_float4(lx, ly, lz, mr);
  if (lane == 1) v.tbox[2 * t + 1] = make_float4(hx, hy, hz, ms);
}


}  // namespace prx
