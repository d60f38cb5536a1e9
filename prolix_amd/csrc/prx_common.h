// Shared layout of the engine's device buffers and the kernel launch interface.
//
// Every per-atom array is indexed by a global arena index g. A system ("slot")
// owns the contiguous range [a0, a0 + n). Tiles never cross systems: tile k
// of a system covers its atoms [a0 + 64k, a0 + 64k + 64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prolix_gb.h"

namespace prx {

constexpr int kTile = 64;               // wave64: one i atom per lane, 64 j atoms rotate
constexpr int kMaxExclClasses = 3;      // distinct non-unit (vdw, elec) scale pairs
constexpr double kForceScale = 16777216.0;      // 2^24: fixed point of forces and dE/dB
constexpr double kBornScale = 4294967296.0;     // 2^32: fixed point of the Born sums I_i
constexpr float kLog2E = 1.4426950408889634f;
constexpr float kLn2 = 0.6931471805599453f;

// Work item of the pair passes: one 64x64 tile pair (I <= J) of one system.
struct PairItem {
  int32_t i0;    // global index of the i tile's first atom
  int32_t j0;    // global index of the j tile's first atom
  int32_t meta;  // ni (bits 0-6), nj (bits 7-13), flags (bits 14-15), slot (bits 16-31)
  int32_t excl;  // exclusion-mask tile index, or -1
};
constexpr int kFlagDiag = 1;
constexpr int kFlagExcl = 2;

__host__ __device__ inline int32_t item_meta(int ni, int nj, int flags, int slot) {
  return ni | (nj << 7) | (flags << 14) | (slot << 16);
}
__host__ __device__ inline int item_ni(int32_t m) { return m & 127; }
__host__ __device__ inline int item_nj(int32_t m) { return (m >> 7) & 127; }
__host__ __device__ inline int item_flags(int32_t m) { return (m >> 14) & 3; }
__host__ __device__ inline int item_slot(int32_t m) { return (int)((uint32_t)m >> 16); }

// Per-system constants read by kernels.
struct SlotConst {
  int32_t a0, n;        // atom range
  int32_t active;       // 1 if stepping
  int32_t pad0;
  float cutoff2;        // 0 = no cutoff
  float k_coulomb;      // Coulomb constant
  float gb2;            // 2 * GB prefactor, prefactor = -1/2 k_gb (1/eps_in - 1/eps_out)
  float pad1;
  double gb_pref;       // GB prefactor (FP64, per-atom self energy)
  float excl_v[4];      // class -> vdw scale (class 0 = 1)
  float excl_e[4];      // class -> elec scale
  // integrator (AKMA units)
  float dt, half_dt, c1, c2;   // c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2)
  float sqrt_kT, kT, pad2, pad3;
  uint64_t seed;
};

// Bonded entries; idx.x < 0 marks a free arena slot.
struct BondEntry { int4 idx; float4 p; };    // idx.xy atoms, idx.z kind (0 bond, 1 UB); p.xy = r0, k
struct AngleEntry { int4 idx; float4 p; };   // idx.xyz atoms; p.xy = theta0, k
struct TorsionEntry { int4 idx; float4 p; }; // p = (n, phase, k, kind) kind 0 dihedral, 1 periodic
                                             // improper, 2 harmonic improper (p.x = k, p.y = phi0)
struct CmapEntry { int4 idx; int4 idx2; };   // idx = atoms a..d, idx2.x = atom e, .y = coeff base, .z = G
struct ExceptionEntry { int4 idx; float4 p; };  // idx.xy pair; p = (sigma, eps, qq * k_coulomb, 0)

// All device pointers of one engine.
struct DeviceView {
  // per atom (arena)
  float4* pos;         // x, y, z, 0
  float4* mom;         // px, py, pz, 1/m
  float4* force;       // integrator force (float), w unused
  float4* born4;       // rho, sr, sr^2, 0            (static)
  float4* nb4;         // q, sigma/2, sqrt(eps), 0    (static)
  double4* fin4;       // rho, R, q^2, ace_K          (static, FP64)
  float* sqrt_m;       // sqrt(mass)                  (static)
  int32_t* slot_of;    // slot id or -1
  float4* gb4;         // B, exp factor a = sqrt(log2(e)/4)/B, dB/dI, chain factor c
  unsigned long long* acc;   // 4 per atom: fx, fy, fz, dE/dB (fixed point 2^24)
  unsigned long long* iacc;  // Born sums (fixed point 2^32)
  double* e_atom;      // 2 per atom: GB self, ACE
  int32_t atom_hw;     // arena high-water mark
  // pair work
  const PairItem* items;
  int32_t n_items;
  const unsigned long long* excl_masks;  // 128 per exclusion tile: lane * 2 + {bit0, bit1}
  double* e_items;     // 3 per item: GB pairs, Coulomb, LJ
  // bonded
  BondEntry* bonds; int32_t bond_hw;
  AngleEntry* angles; int32_t angle_hw;
  TorsionEntry* tors; int32_t tors_hw;
  CmapEntry* cmaps; int32_t cmap_hw;
  ExceptionEntry* excs; int32_t exc_hw;
  const float* cmap_coeff;
  double* e_bond; double* e_angle; double* e_tors; double* e_cmap; double* e_exc;
  // systems
  SlotConst* slots;
  uint32_t* atom_step;     // per atom: BAOAB steps taken (the noise counter)
  int32_t max_slots;
};

// Launchers (defined in the .hip translation units).
void launch_born_pairs(const DeviceView& v, hipStream_t s);
void launch_born_finish(const DeviceView& v, hipStream_t s);
void launch_force_pairs(const DeviceView& v, bool energy, hipStream_t s);
void launch_chain_prep(const DeviceView& v, bool energy, hipStream_t s);
void launch_chain_pairs(const DeviceView& v, hipStream_t s);
void launch_bonded(const DeviceView& v, bool energy, hipStream_t s);
void launch_zero_acc(const DeviceView& v, hipStream_t s);
void launch_kick_drift(const DeviceView& v, int mode, hipStream_t s);  // 0: first BAOA, 1: B + BAOA, 2: final B
void launch_gather_forces(const DeviceView& v, hipStream_t s);
void launch_energy_reduce(const DeviceView& v, const int32_t* slot_ranges, int n_slots, double* out,
                          hipStream_t s);
void launch_copy_in_positions(const DeviceView& v, const float* packed, const int32_t* map, int n,
                              hipStream_t s);
void launch_copy_out(const DeviceView& v, float* pos, float* mom, float* frc, const int32_t* map, int n,
                     hipStream_t s);
void launch_copy_in_momenta(const DeviceView& v, const float* packed, const int32_t* map, int n,
                            hipStream_t s);
void launch_init_momenta(const DeviceView& v, const int32_t* slot_list, int n_slot_list, hipStream_t s);
void launch_kinetic(const DeviceView& v, const int32_t* slot_ranges, int n_slots, double* out,
                    hipStream_t s);
void launch_born_out(const DeviceView& v, float* out, const int32_t* map, int n, hipStream_t s);

}  // namespace prx
