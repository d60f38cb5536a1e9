// Per-atom kernels: Born-radius finish, chain-rule factors, BAOAB halves,
// Philox noise, copies and deterministic per-system reductions.
#include "prx_atom_dev.h"

namespace prx {
// One block per tile: the gathers and finishes of prx_atom_dev.h.
__global__ __launch_bounds__(256) void k_born_gather(DeviceView v) {
  __shared__ GatherLds G;
  if ((int)blockIdx.x >= v.n_tiles) return;
  born_gather_tile(v, blockIdx.x, G);
}

template <bool ENERGY>
__global__ __launch_bounds__(256) void k_force_gather(DeviceView v) {
  __shared__ GatherLds G;
  if ((int)blockIdx.x >= v.n_tiles) return;
  force_gather_tile<ENERGY>(v, blockIdx.x, G);
}

__global__ __launch_bounds__(256) void k_chain_gather(DeviceView v, int mode) {
  __shared__ GatherLds G;
  if ((int)blockIdx.x >= v.n_tiles) return;
  chain_gather_tile(v, blockIdx.x, mode, G);
}

__global__ void k_step_begin(DeviceView v) { step_begin_atom(v, blockIdx.x * blockDim.x + threadIdx.x); }

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

// Energy reduction: one 1024-thread block per listed system. Each thread sums
// its strided share of every term in index order (the work-item loop issues
// four iterations' loads before their adds: the loop is bound by load
// latency, not bandwidth), then the lanes of each wave are combined by a
// fixed butterfly and the 16 waves in wave order, so the result is
// deterministic.
constexpr int kEnThreads = 1024;
constexpr int kEnWaves = kEnThreads / 64;

// Elements b + t, b + t + T, ... of [b, e) (T = kEnThreads) in that order:
// acc(kind(k), energy(k)), with four elements' loads issued before their adds.
template <class K, class E, class A>
__device__ __forceinline__ void strided4(int b, int e, int t, K kind, E energy, A acc) {
  int k = b + t;
  for (; k + 3 * kEnThreads < e; k += 4 * kEnThreads) {
    int z[4];
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      z[u] = kind(k + u * kEnThreads);
      x[u] = energy(k + u * kEnThreads);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) acc(z[u], x[u]);
  }
  for (; k < e; k += kEnThreads) acc(kind(k), energy(k));
}

__device__ __forceinline__ double wave_sum_lane0(double x) {
  for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
  return x;  // lane 0's value is used
}

__global__ __launch_bounds__(kEnThreads) void k_energy_reduce(DeviceView v, const int32_t* ranges,
                                                              double* out) {
  constexpr int kTerms = 12;
  __shared__ double part[kTerms][kEnWaves];
  const int32_t* R = ranges + blockIdx.x * kRangeInts;
  const int t = threadIdx.x;
  double e_gbp = 0, e_cl = 0, e_lj = 0;
  int i = R[0] + t;
  for (; i + 3 * kEnThreads < R[1]; i += 4 * kEnThreads) {
    double a[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 3; ++c) a[u][c] = v.e_items[(size_t)(i + u * kEnThreads) * 3 + c];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      e_gbp += a[u][0];
      e_cl += a[u][1];
      e_lj += a[u][2];
    }
  }
  for (; i < R[1]; i += kEnThreads) {
    e_gbp += v.e_items[(size_t)i * 3 + 0];
    e_cl += v.e_items[(size_t)i * 3 + 1];
    e_lj += v.e_items[(size_t)i * 3 + 2];
  }
  double e_self = 0, e_ace = 0;
  for (int k = t; k < R[3]; k += kEnThreads) {
    e_self += v.e_atom[(size_t)(R[2] + k) * 2 + 0];
    e_ace += v.e_atom[(size_t)(R[2] + k) * 2 + 1];
  }
  // A system's bonded ranges hold only its own entries (free entries lie
  // outside every listed range), so no entry is tested for being free; the
  // kind of a bond (bond / Urey-Bradley) and of a torsion entry (proper /
  // improper) is read beside its energy. Four entries' loads per iteration.
  double e_bond = 0, e_ub = 0;
  strided4(R[4], R[5], t, [&](int k) { return v.bonds[k].idx.z; }, [&](int k) { return v.e_bond[k]; },
           [&](int z, double e) { if (z == 0) e_bond += e; else e_ub += e; });
  double e_ang = 0;
  strided4(R[6], R[7], t, [&](int) { return 0; }, [&](int k) { return v.e_angle[k]; },
           [&](int, double e) { e_ang += e; });
  double e_dih = 0, e_imp = 0;
  strided4(R[8], R[9], t, [&](int k) { return v.tors[k].a.w == 0.0f ? 0 : 1; }, [&](int k) { return v.e_tors[k]; },
           [&](int z, double e) { if (z == 0) e_dih += e; else e_imp += e; });
  double e_cm = 0;
  strided4(R[10], R[11], t, [&](int) { return 0; }, [&](int k) { return v.e_cmap[k]; },
           [&](int, double e) { e_cm += e; });
  double e_ex = 0;
  strided4(R[12], R[13], t, [&](int) { return 0; }, [&](int k) { return v.e_exc[k]; },
           [&](int, double e) { e_ex += e; });
  const double mine[kTerms] = {e_bond, e_ang, e_ub, e_dih, e_imp, e_cm, e_lj, e_ex, e_gbp, e_self, e_cl, e_ace};
  const int lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int k = 0; k < kTerms; ++k) {
    const double w = wave_sum_lane0(mine[k]);
    if (lane == 0) part[k][wave] = w;
  }
  __syncthreads();
  if (t != 0) return;
  double sum[kTerms];
  for (int k = 0; k < kTerms; ++k) {
    double acc = 0.0;
    for (int w = 0; w < kEnWaves; ++w) acc += part[k][w];
    sum[k] = acc;
  }
  double terms[PRX_N_ENERGY];
  terms[PRX_E_BOND] = sum[0];
  terms[PRX_E_ANGLE] = sum[1];
  terms[PRX_E_UREY_BRADLEY] = sum[2];
  terms[PRX_E_DIHEDRAL] = sum[3];
  terms[PRX_E_IMPROPER] = sum[4];
  terms[PRX_E_CMAP] = sum[5];
  terms[PRX_E_LJ] = sum[6];
  terms[PRX_E_EXCEPTION] = sum[7];
  terms[PRX_E_GB] = sum[8] + sum[9];
  terms[PRX_E_COULOMB] = sum[10];
  terms[PRX_E_ACE] = sum[11];
  double tot = 0.0;
  for (int k = 0; k < PRX_E_TOTAL; ++k) tot += terms[k];
  terms[PRX_E_TOTAL] = tot;
  for (int k = 0; k < PRX_N_ENERGY; ++k) out[(size_t)blockIdx.x * PRX_N_ENERGY + k] = terms[k];
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

__global__ __launch_bounds__(256) void k_tile_boxes(DeviceView v) {
  const int t = gather_wave();
  if (t >= v.n_tiles) return;
  tile_box_wave(v, t);
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
  if (n_slots) hipLaunchKernelGGL(k_energy_reduce, dim3(n_slots), dim3(kEnThreads), 0, s, v, ranges, out);
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
__global__ void k_fill_u32(uint32_t* dst, uint32_t value, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = value;
}

void launch_fill_u32(uint32_t* dst, uint32_t value, int n, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_fill_u32, atom_grid(n), dim3(256), 0, s, dst, value, n);
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
