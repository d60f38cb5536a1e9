// Per-atom kernels: Born-radius finish, chain-rule factors, BAOAB halves,
// Philox noise, copies and deterministic per-system reductions.
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
  bool skip_far;  // neighbour-list GB system: far items carry no partials
};

__device__ __forceinline__ TileCtx tile_ctx(const DeviceView& v, int t) {
  const int4 tl = v.tiles[t];
  TileCtx c;
  c.lane = threadIdx.x & 63;
  c.q = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  c.g = tl.x + c.lane;
  c.valid = c.lane < tl.y;
  c.slot = tl.z;
  c.T = tl.w;
  const SlotConst& S = v.slots[c.slot];
  c.nT = (S.n + kTile - 1) / kTile;
  c.base = S.item_base;
  c.skip_far = S.gb_cut != 0;
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
template <class T, class Load, class Add>
__device__ __forceinline__ void gather_items(const DeviceView& v, const TileCtx& c, Load&& load, Add&& add) {
  if (c.skip_far) {
    for_tile_items(v, c, [&](int w, int side) { add(load(w, side)); });
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
      if (n + u < n_row) val[u] = load(r0 + c.q + kGW * (n + u), 0);
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
        val[u] = load(row_start(c, I) + (c.T - I), 1);
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
  gather_items<float>(v, c, [&](int w, int side) { return pb[pslot(w, side, c.lane)]; },
                      [&](float x) { acc += (double)x; });
  return acc;
}

// Chain partials: three float arrays of 64 per (item, side).
__device__ __forceinline__ double4 gather3(const DeviceView& v, const TileCtx& c) {
  const float* pc = reinterpret_cast<const float*>(v.part);
  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  gather_items<float3>(
      v, c,
      [&](int w, int side) {
        const float* q = pc + ((size_t)w * 2 + side) * (3 * kTile) + c.lane;
        return make_float3(q[0], q[kTile], q[2 * kTile]);
      },
      [&](float3 p) { acc.x += p.x; acc.y += p.y; acc.z += p.z; });
  return acc;
}

__device__ __forceinline__ double4 gather4(const DeviceView& v, const TileCtx& c) {
  double4 acc = make_double4(0.0, 0.0, 0.0, 0.0);
  gather_items<float4>(v, c, [&](int w, int side) { return v.part[pslot(w, side, c.lane)]; },
                       [&](float4 p) { acc.x += p.x; acc.y += p.y; acc.z += p.z; acc.w += p.w; });
  return acc;
}

// The four waves' sums, added in wave order; valid in wave 0 after the call.
// Every thread of the block must call it (it holds a barrier).
__device__ __forceinline__ double combine1(double x, const TileCtx& c) {
  __shared__ double sh[kGW][kTile];
  sh[c.q][c.lane] = x;
  __syncthreads();
  return ((sh[0][c.lane] + sh[1][c.lane]) + sh[2][c.lane]) + sh[3][c.lane];
}

__device__ __forceinline__ double4 combine4(double4 x, const TileCtx& c) {
  __shared__ double4 sh4[kGW][kTile];
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

// B_i from I_i in FP64 (generalized_born.py:81-93); dB/dI for the chain rule.
// Neighbour-list GB systems floor 1/B at 1e-4 (generalized_born.py:333-337).
__global__ __launch_bounds__(256) void k_born_gather(DeviceView v) {
  if ((int)blockIdx.x >= v.n_tiles) return;
  const TileCtx c = tile_ctx(v, blockIdx.x);
  const double I = combine1(gather1(v, c), c);
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
__global__ __launch_bounds__(256) void k_force_gather(DeviceView v) {
  if ((int)blockIdx.x >= v.n_tiles) return;
  const TileCtx c = tile_ctx(v, blockIdx.x);
  const double4 a = combine4(gather4(v, c), c);
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
      atomicAdd(&v.warn[s * 4 + kWarnVLimit], 1u);
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
      atomicAdd(&v.warn[s * 4 + kWarnDxCap], 1u);
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
      if (mode >= 0) atomicAdd(&v.warn[c.slot * 4 + kWarnForceCap], 1u);
    }
  }
  v.force[c.g] = make_float4(f.x, f.y, f.z, 0.0f);
  if (mode < 0 || !S.active) return v.pos[c.g];
  float4 p = v.mom[c.g];  // w = 1/m
  p.x = fmaf(S.half_dt, f.x, p.x);
  p.y = fmaf(S.half_dt, f.y, p.y);
  p.z = fmaf(S.half_dt, f.z, p.z);
  const uint32_t step = v.atom_step[c.g] + 1u;  // each lane owns its atom's counter
  v.atom_step[c.g] = step;
  float4 x = mode == 1 ? begin_step(v, S, c.slot, c.g, p, step) : v.pos[c.g];
  v.mom[c.g] = p;
  return x;
}

// Per tile (one block): the chain partials, bonded records and pair force
// finish each atom's force (finish_atom). In mode 1 the tile's atoms have
// moved to the next step's positions, and wave 0 also writes the tile's box
// for that step (what k_tile_boxes computes), so the step graph needs no box
// kernel after its first step.
__global__ __launch_bounds__(256) void k_chain_gather(DeviceView v, int mode) {
  if ((int)blockIdx.x >= v.n_tiles) return;
  const TileCtx c = tile_ctx(v, blockIdx.x);
  double4 part = gather3(v, c);
  float3 fb = c.valid ? bonded_force(v, c.g, c.q) : make_float3(0.f, 0.f, 0.f);
  // the bonded share travels in the same combine (its float sum is exact in FP64)
  __shared__ float4 shb[kGW][kTile];
  shb[c.q][c.lane] = make_float4(fb.x, fb.y, fb.z, 0.0f);
  const double4 a = combine4(part, c);
  if (c.q != 0) return;
  float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
  if (c.valid) x = finish_atom(v, c, a, shb, mode);
  if (mode != 1 || !v.slots[c.slot].active) return;  // uniform over the tile
  const int last = v.tiles[blockIdx.x].y - 1;
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
  if (c.lane == 0) v.tbox[2 * blockIdx.x] = make_float4(lx, ly, lz, mr);
  if (c.lane == 1) v.tbox[2 * blockIdx.x + 1] = make_float4(hx, hy, hz, ms);
}

// B A O A from the stored force, for the first step of a launch.
__global__ void k_step_begin(DeviceView v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw) return;
  const int s = v.slot_of[g];
  if (s < 0) return;
  const SlotConst& S = v.slots[s];
  if (!S.active) return;
  float4 p = v.mom[g];
  begin_step(v, S, s, g, p, v.atom_step[g]);
  v.mom[g] = p;
}

// ---------------------------------------------------------------------------
// Block reductions (fixed tree: deterministic)
// ---------------------------------------------------------------------------
constexpr int kRedThreads = 256;

__device__ double block_sum(double x, double* sh) {
  const int t = threadIdx.x;
  sh[t] = x;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (t < off) sh[t] += sh[t + off];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Slot range record layout (ints): item_b, item_e, a0, n, bond_b, bond_e, ang_b, ang_e,
// tors_b, tors_e, cmap_b, cmap_e, exc_b, exc_e.
constexpr int kRangeInts = 14;

__global__ __launch_bounds__(kRedThreads) void k_energy_reduce(DeviceView v, const int32_t* ranges,
                                                               double* out) {
  __shared__ double sh[kRedThreads];
  const int32_t* R = ranges + blockIdx.x * kRangeInts;
  const int t = threadIdx.x;
  double e_gbp = 0, e_cl = 0, e_lj = 0;
  for (int i = R[0] + t; i < R[1]; i += kRedThreads) {
    e_gbp += v.e_items[(size_t)i * 3 + 0];
    e_cl += v.e_items[(size_t)i * 3 + 1];
    e_lj += v.e_items[(size_t)i * 3 + 2];
  }
  double e_self = 0, e_ace = 0;
  for (int i = t; i < R[3]; i += kRedThreads) {
    e_self += v.e_atom[(size_t)(R[2] + i) * 2 + 0];
    e_ace += v.e_atom[(size_t)(R[2] + i) * 2 + 1];
  }
  double e_bond = 0, e_ub = 0;
  for (int i = R[4] + t; i < R[5]; i += kRedThreads) {
    if (v.bonds[i].idx.x < 0) continue;
    if (v.bonds[i].idx.z == 0) e_bond += v.e_bond[i]; else e_ub += v.e_bond[i];
  }
  double e_ang = 0;
  for (int i = R[6] + t; i < R[7]; i += kRedThreads)
    if (v.angles[i].idx.x >= 0) e_ang += v.e_angle[i];
  double e_dih = 0, e_imp = 0;
  for (int i = R[8] + t; i < R[9]; i += kRedThreads) {
    if (v.tors[i].idx.x < 0) continue;
    if (v.tors[i].a.w == 0.0f) e_dih += v.e_tors[i]; else e_imp += v.e_tors[i];
  }
  double e_cm = 0;
  for (int i = R[10] + t; i < R[11]; i += kRedThreads)
    if (v.cmaps[i].idx.x >= 0) e_cm += v.e_cmap[i];
  double e_ex = 0;
  for (int i = R[12] + t; i < R[13]; i += kRedThreads)
    if (v.excs[i].idx.x >= 0) e_ex += v.e_exc[i];
  double terms[PRX_N_ENERGY];
  terms[PRX_E_BOND] = block_sum(e_bond, sh);
  terms[PRX_E_ANGLE] = block_sum(e_ang, sh);
  terms[PRX_E_UREY_BRADLEY] = block_sum(e_ub, sh);
  terms[PRX_E_DIHEDRAL] = block_sum(e_dih, sh);
  terms[PRX_E_IMPROPER] = block_sum(e_imp, sh);
  terms[PRX_E_CMAP] = block_sum(e_cm, sh);
  terms[PRX_E_LJ] = block_sum(e_lj, sh);
  terms[PRX_E_EXCEPTION] = block_sum(e_ex, sh);
  terms[PRX_E_GB] = block_sum(e_gbp, sh) + block_sum(e_self, sh);
  terms[PRX_E_COULOMB] = block_sum(e_cl, sh);
  terms[PRX_E_ACE] = block_sum(e_ace, sh);
  double tot = 0.0;
  for (int k = 0; k < PRX_E_TOTAL; ++k) tot += terms[k];
  terms[PRX_E_TOTAL] = tot;
  if (t < PRX_N_ENERGY) out[(size_t)blockIdx.x * PRX_N_ENERGY + t] = terms[t];
}

__global__ __launch_bounds__(kRedThreads) void k_kinetic(DeviceView v, const int32_t* ranges,
                                                         double* out) {
  __shared__ double sh[kRedThreads];
  const int32_t* R = ranges + blockIdx.x * kRangeInts;
  double ke = 0.0;
  for (int i = threadIdx.x; i < R[3]; i += kRedThreads) {
    const float4 p = v.mom[R[2] + i];
    ke += 0.5 * ((double)p.x * p.x + (double)p.y * p.y + (double)p.z * p.z) * (double)p.w;
  }
  ke = block_sum(ke, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = ke;
}

// Positions minus the system's mass-weighted centre of mass (masses from the
// stored 1/m in mom.w), FP64 sums; one block per listed system.
__global__ __launch_bounds__(kRedThreads) void k_recenter_com(DeviceView v, const int32_t* ranges) {
  __shared__ double sh[kRedThreads];
  const int32_t* R = ranges + blockIdx.x * kRangeInts;
  double sm = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
  for (int i = threadIdx.x; i < R[3]; i += kRedThreads) {
    const int g = R[2] + i;
    const double m = 1.0 / (double)v.mom[g].w;
    const float4 x = v.pos[g];
    sm += m; sx += m * x.x; sy += m * x.y; sz += m * x.z;
  }
  const double mt = block_sum(sm, sh) + 1e-30;
  const double cx = block_sum(sx, sh) / mt, cy = block_sum(sy, sh) / mt, cz = block_sum(sz, sh) / mt;
  for (int i = threadIdx.x; i < R[3]; i += kRedThreads) {
    const int g = R[2] + i;
    float4 x = v.pos[g];
    x.x = (float)(x.x - cx); x.y = (float)(x.y - cy); x.z = (float)(x.z - cz);
    v.pos[g] = x;
  }
}

// Maxwell-Boltzmann momenta, centred per system (jax_md initialize_momenta):
// p = sqrt(m kT) xi, then p -= mean(p). One block per listed slot.
__global__ __launch_bounds__(kRedThreads) void k_init_momenta(DeviceView v, const int32_t* slot_list) {
  __shared__ double sh[kRedThreads];
  const int s = slot_list[blockIdx.x];
  const SlotConst& S = v.slots[s];
  double sx = 0, sy = 0, sz = 0;
  for (int i = threadIdx.x; i < S.n; i += kRedThreads) {
    const int g = S.a0 + i;
    const float3 xi = normal3(S.seed, 0u, i, 1u);
    const float a = S.sqrt_kT * v.sqrt_m[g];
    float4 p = v.mom[g];
    p.x = a * xi.x; p.y = a * xi.y; p.z = a * xi.z;
    v.mom[g] = p;
    v.atom_step[g] = 0u;
    sx += p.x; sy += p.y; sz += p.z;
  }
  const double mx = block_sum(sx, sh) / S.n;
  const double my = block_sum(sy, sh) / S.n;
  const double mz = block_sum(sz, sh) / S.n;
  if (S.n > 1) {
    for (int i = threadIdx.x; i < S.n; i += kRedThreads) {
      const int g = S.a0 + i;
      float4 p = v.mom[g];
      p.x = (float)(p.x - mx); p.y = (float)(p.y - my); p.z = (float)(p.z - mz);
      v.mom[g] = p;
    }
  }
}

// map[p] = arena index of packed atom p.
__global__ void k_copy_in_pos(DeviceView v, const float* src, const int32_t* map, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int g = map[p];
  v.pos[g] = make_float4(src[3 * p], src[3 * p + 1], src[3 * p + 2], 0.0f);
}

__global__ void k_copy_in_mom(DeviceView v, const float* src, const int32_t* map, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int g = map[p];
  float4 m = v.mom[g];
  m.x = src[3 * p]; m.y = src[3 * p + 1]; m.z = src[3 * p + 2];
  v.mom[g] = m;
}

__global__ void k_copy_out(DeviceView v, float* pos, float* mom, float* frc, const int32_t* map, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int g = map[p];
  if (pos) { const float4 a = v.pos[g]; pos[3 * p] = a.x; pos[3 * p + 1] = a.y; pos[3 * p + 2] = a.z; }
  if (mom) { const float4 a = v.mom[g]; mom[3 * p] = a.x; mom[3 * p + 1] = a.y; mom[3 * p + 2] = a.z; }
  if (frc) { const float4 a = v.force[g]; frc[3 * p] = a.x; frc[3 * p + 1] = a.y; frc[3 * p + 2] = a.z; }
}

__global__ void k_born_out(DeviceView v, float* out, const int32_t* map, int n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  out[p] = v.gb4[map[p]].x;
}

static inline dim3 atom_grid(int n) { return dim3((n + 255) / 256); }

static inline dim3 tile_grid(int n_tiles) { return dim3((n_tiles + 3) / 4); }  // one wave per tile
static inline dim3 gather_grid(int n_tiles) { return dim3(n_tiles); }          // one block per tile

__device__ __forceinline__ int gather_wave() {
  return __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)));
}

// Bounding box of each tile's atoms at the current positions, with the
// tile's largest offset radius rho and descreening radius s; one wave per tile.
__global__ __launch_bounds__(256) void k_tile_boxes(DeviceView v) {
  const int t = gather_wave();
  if (t >= v.n_tiles) return;
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

void launch_tile_boxes(const DeviceView& v, hipStream_t s) {
  if (v.n_tiles) hipLaunchKernelGGL(k_tile_boxes, tile_grid(v.n_tiles), dim3(256), 0, s, v);
}

void launch_born_gather(const DeviceView& v, hipStream_t s) {
  if (v.n_tiles) hipLaunchKernelGGL(k_born_gather, gather_grid(v.n_tiles), dim3(256), 0, s, v);
}
void launch_force_gather(const DeviceView& v, bool energy, hipStream_t s) {
  if (!v.n_tiles) return;
  if (energy)
    hipLaunchKernelGGL(k_force_gather<true>, gather_grid(v.n_tiles), dim3(256), 0, s, v);
  else
    hipLaunchKernelGGL(k_force_gather<false>, gather_grid(v.n_tiles), dim3(256), 0, s, v);
}
void launch_chain_gather(const DeviceView& v, int mode, hipStream_t s) {
  if (v.n_tiles) hipLaunchKernelGGL(k_chain_gather, gather_grid(v.n_tiles), dim3(256), 0, s, v, mode);
}
void launch_step_begin(const DeviceView& v, hipStream_t s) {
  if (v.atom_hw) hipLaunchKernelGGL(k_step_begin, atom_grid(v.atom_hw), dim3(256), 0, s, v);
}
void launch_energy_reduce(const DeviceView& v, const int32_t* ranges, int n_slots, double* out,
                          hipStream_t s) {
  if (n_slots) hipLaunchKernelGGL(k_energy_reduce, dim3(n_slots), dim3(kRedThreads), 0, s, v, ranges, out);
}
void launch_kinetic(const DeviceView& v, const int32_t* ranges, int n_slots, double* out, hipStream_t s) {
  if (n_slots) hipLaunchKernelGGL(k_kinetic, dim3(n_slots), dim3(kRedThreads), 0, s, v, ranges, out);
}
// One timestamp of the GPU's constant-rate wall clock, stored by lane 0 of a
// one-wave launch; graph dependencies order it between the passes around it.
__global__ void k_tick(unsigned long long* out) {
  const unsigned long long t = wall_clock64();
  if (threadIdx.x == 0) out[threadIdx.x] = t;
}

void launch_tick(unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_tick, dim3(1), dim3(64), 0, s, out);
}

void launch_recenter_com(const DeviceView& v, const int32_t* ranges, int n_slots, hipStream_t s) {
  if (n_slots) hipLaunchKernelGGL(k_recenter_com, dim3(n_slots), dim3(kRedThreads), 0, s, v, ranges);
}
void launch_init_momenta(const DeviceView& v, const int32_t* slot_list, int n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_init_momenta, dim3(n), dim3(kRedThreads), 0, s, v, slot_list);
}
void launch_copy_in_positions(const DeviceView& v, const float* src, const int32_t* map, int n,
                              hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_in_pos, atom_grid(n), dim3(256), 0, s, v, src, map, n);
}
void launch_copy_in_momenta(const DeviceView& v, const float* src, const int32_t* map, int n,
                            hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_in_mom, atom_grid(n), dim3(256), 0, s, v, src, map, n);
}
void launch_copy_out(const DeviceView& v, float* pos, float* mom, float* frc, const int32_t* map, int n,
                     hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_copy_out, atom_grid(n), dim3(256), 0, s, v, pos, mom, frc, map, n);
}
void launch_born_out(const DeviceView& v, float* out, const int32_t* map, int n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_born_out, atom_grid(n), dim3(256), 0, s, v, out, map, n);
}

}  // namespace prx
