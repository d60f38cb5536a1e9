// Per-atom kernels: Born-radius finish, chain-rule factors, BAOAB halves,
// Philox noise, copies and deterministic per-system reductions.
#include "prx_common.h"

namespace prx {

__device__ __forceinline__ double fixed_to_double(unsigned long long v, double inv_scale) {
  return (double)(long long)v * inv_scale;
}

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

__global__ void k_zero_acc(DeviceView v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw) return;
  v.acc[(size_t)g * 4 + 0] = 0ull;
  v.acc[(size_t)g * 4 + 1] = 0ull;
  v.acc[(size_t)g * 4 + 2] = 0ull;
  v.acc[(size_t)g * 4 + 3] = 0ull;
  v.iacc[g] = 0ull;
}

// B_i from I_i in FP64 (generalized_born.py:81-93); dB/dI for the chain rule.
__global__ void k_born_finish(DeviceView v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw || v.slot_of[g] < 0) return;
  const double4 f = v.fin4[g];  // rho, R, q^2, ace K
  const double I = fixed_to_double(v.iacc[g], 1.0 / kBornScale);
  const double psi = f.x * I;
  const double t = psi * (1.0 + psi * (-0.8 + 4.85 * psi));
  const double th = tanh(t);
  const double inv_b = 1.0 / f.x - th / f.y;
  const double B = 1.0 / inv_b;
  const double dBdI = B * B * (1.0 - th * th) * (1.0 + psi * (-1.6 + 14.55 * psi)) * f.x / f.y;
  float4 o;
  o.x = (float)B;
  o.y = (float)(0.6005612043932249 / B);  // sqrt(log2(e) / 4) / B
  o.z = (float)dBdI;
  o.w = 0.0f;
  v.gb4[g] = o;
}

// c_i = dE/dB_i * dB_i/dI_i with the GB self term and ACE added (FP64).
template <bool ENERGY>
__global__ void k_chain_prep(DeviceView v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw) return;
  const int s = v.slot_of[g];
  if (s < 0) return;
  const double4 f = v.fin4[g];
  const float4 gb = v.gb4[g];
  const double B = (double)gb.x;
  const double pref = v.slots[s].gb_pref;
  const double e_self = pref * f.z / B;
  const double e_ace = f.w / (B * B * B * B * B * B);
  const double dEdB = fixed_to_double(v.acc[(size_t)g * 4 + 3], 1.0 / kForceScale) - e_self / B
                      - 6.0 * e_ace / B;
  v.gb4[g].w = (float)(dEdB * (double)gb.z);
  if (ENERGY) {
    v.e_atom[(size_t)g * 2 + 0] = e_self;
    v.e_atom[(size_t)g * 2 + 1] = e_ace;
  }
}

__device__ __forceinline__ float3 acc_force(const DeviceView& v, int g) {
  const double inv = 1.0 / kForceScale;
  float3 f;
  f.x = (float)fixed_to_double(v.acc[(size_t)g * 4 + 0], inv);
  f.y = (float)fixed_to_double(v.acc[(size_t)g * 4 + 1], inv);
  f.z = (float)fixed_to_double(v.acc[(size_t)g * 4 + 2], inv);
  return f;
}

__global__ void k_gather_forces(DeviceView v) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw || v.slot_of[g] < 0) return;
  const float3 f = acc_force(v, g);
  v.force[g] = make_float4(f.x, f.y, f.z, 0.0f);
}

// BAOAB (settle.py:2064-2087, composed as settle.py:1699-1744):
//   mode 0: B(F_n) A O A, zero the accumulators for the next force evaluation
//   mode 2: gather F_{n+1}, B(F_{n+1}), count the step
//   mode 1: mode 2 then mode 0 (the boundary between two steps)
__global__ void k_kick_drift(DeviceView v, int mode) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= v.atom_hw) return;
  const int s = v.slot_of[g];
  if (s < 0) return;
  const SlotConst& S = v.slots[s];
  if (!S.active) return;
  float4 p = v.mom[g];  // w = 1/m
  uint32_t step = v.atom_step[g];  // each thread owns its atom's counter
  if (mode >= 1) {
    const float3 f = acc_force(v, g);
    v.force[g] = make_float4(f.x, f.y, f.z, 0.0f);
    p.x = fmaf(S.half_dt, f.x, p.x);
    p.y = fmaf(S.half_dt, f.y, p.y);
    p.z = fmaf(S.half_dt, f.z, p.z);
    step += 1u;
    v.atom_step[g] = step;
  }
  if (mode <= 1) {
    const float4 f = v.force[g];
    p.x = fmaf(S.half_dt, f.x, p.x);
    p.y = fmaf(S.half_dt, f.y, p.y);
    p.z = fmaf(S.half_dt, f.z, p.z);
    float4 x = v.pos[g];
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
    v.pos[g] = x;
    v.acc[(size_t)g * 4 + 0] = 0ull;
    v.acc[(size_t)g * 4 + 1] = 0ull;
    v.acc[(size_t)g * 4 + 2] = 0ull;
    v.acc[(size_t)g * 4 + 3] = 0ull;
    v.iacc[g] = 0ull;
  }
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
    if (v.tors[i].p.w == 0.0f) e_dih += v.e_tors[i]; else e_imp += v.e_tors[i];
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

void launch_zero_acc(const DeviceView& v, hipStream_t s) {
  if (v.atom_hw) hipLaunchKernelGGL(k_zero_acc, atom_grid(v.atom_hw), dim3(256), 0, s, v);
}
void launch_born_finish(const DeviceView& v, hipStream_t s) {
  if (v.atom_hw) hipLaunchKernelGGL(k_born_finish, atom_grid(v.atom_hw), dim3(256), 0, s, v);
}
void launch_chain_prep(const DeviceView& v, bool energy, hipStream_t s) {
  if (!v.atom_hw) return;
  if (energy)
    hipLaunchKernelGGL(k_chain_prep<true>, atom_grid(v.atom_hw), dim3(256), 0, s, v);
  else
    hipLaunchKernelGGL(k_chain_prep<false>, atom_grid(v.atom_hw), dim3(256), 0, s, v);
}
void launch_gather_forces(const DeviceView& v, hipStream_t s) {
  if (v.atom_hw) hipLaunchKernelGGL(k_gather_forces, atom_grid(v.atom_hw), dim3(256), 0, s, v);
}
void launch_kick_drift(const DeviceView& v, int mode, hipStream_t s) {
  if (v.atom_hw) hipLaunchKernelGGL(k_kick_drift, atom_grid(v.atom_hw), dim3(256), 0, s, v, mode);
}
void launch_energy_reduce(const DeviceView& v, const int32_t* ranges, int n_slots, double* out,
                          hipStream_t s) {
  if (n_slots) hipLaunchKernelGGL(k_energy_reduce, dim3(n_slots), dim3(kRedThreads), 0, s, v, ranges, out);
}
void launch_kinetic(const DeviceView& v, const int32_t* ranges, int n_slots, double* out, hipStream_t s) {
  if (n_slots) hipLaunchKernelGGL(k_kinetic, dim3(n_slots), dim3(kRedThreads), 0, s, v, ranges, out);
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
