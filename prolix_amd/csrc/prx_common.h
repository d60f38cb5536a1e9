// Shared layout of the engine's device buffers and the kernel launch interface.
//
// Every per-atom array is indexed by a global arena index g. A system ("slot")
// owns the contiguous range [a0, a0 + n). Tiles never cross systems: tile k
// of a system covers its atoms [a0 + 64k, a0 + 64k + 64).
//
// Reductions are deterministic by construction: every pair work item writes
// its per-lane partial sums to its own slots of a partial buffer (no
// atomics), and one wave per tile sums the partials of the tile's row and
// column items in a fixed order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "prolix_gb.h"

namespace prx {

constexpr int kTile = 64;               // wave64: one i atom per lane, 64 j atoms rotate
constexpr int kMaxExclClasses = 3;      // distinct non-unit (vdw, elec) scale pairs
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

// Class of a work item, decided each evaluation from the bounding boxes of
// its two tiles. Bits 0-1: cutoff class (a margin of 1e-5 in r^2 keeps the
// decision conservative against float rounding of the per-pair test).
constexpr int kClsNear = 0;      // every pair closer than the cutoff (or no cutoff)
constexpr int kClsStraddle = 1;  // per-pair r^2 < cutoff^2 test needed
constexpr int kClsFar = 2;       // every pair at or beyond the cutoff
constexpr int kClsCutMask = 3;
// Bit 2: every pair of the item is in the OBC2 far field in both directions
// (r - s_j > rho_i, s_j <= r/2 and the same with i, j swapped), so the Born
// and chain passes use the far-field polynomials without a per-pair test.
constexpr int kClsObcFar = 4;
// Bit 3: in addition, every pair has s <= r/4 in both directions, where the
// far-field polynomials of degree 4 are as accurate as the wider fits.
constexpr int kClsObcFar4 = 8;

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
  int32_t item_base;    // index of the system's first work item (row-major I <= J)
  float cutoff2;        // 0 = no cutoff
  float k_coulomb;      // Coulomb constant
  float gb2;            // 2 * GB prefactor, prefactor = -1/2 k_gb (1/eps_in - 1/eps_out)
  int32_t gb_cut;       // 1: Born sums and GB pairs also restricted to r < cutoff
  double gb_pref;       // GB prefactor (FP64, per-atom self energy)
  float excl_v[4];      // class -> vdw scale (class 0 = 1)
  float excl_e[4];      // class -> elec scale
  // integrator (AKMA units)
  float dt, half_dt, c1, c2;   // c1 = exp(-gamma dt), c2 = sqrt(1 - c1^2)
  float sqrt_kT, kT;
  int32_t caps;         // 1: batched-step safety caps (velocity, displacement, force)
  int32_t tile0;        // global index of the system's first tile
  uint64_t seed;
  // neighbour-list systems (gb_cut with a cutoff): list build and reuse
  int32_t nl;           // 1: pair work from the list (chunk items), 0: all tile pairs
  int32_t tp_base;      // global index of the system's first tile pair (I < J) record
  float nl_cut2;        // (cutoff + skin)^2 with a 1e-5 relative margin: list membership
  float skin_half2;     // (skin / 2)^2; 0: the list is rebuilt at every evaluation
  // 0: no limit; else the system stops after this many BAOAB steps: its
  // positions, momenta and counter stay as they were after that step while the
  // rest of the batch goes on, and its force is recomputed at those positions
  uint32_t step_limit;
};

// True when a slot with a step limit has taken its steps.
__host__ __device__ __forceinline__ bool slot_done(const SlotConst& S, uint32_t steps) {
  return S.step_limit != 0u && steps >= S.step_limit;
}

// Safety caps of prolix's batched BAOAB step (batched_simulate.py:269-358), in
// the engine's AKMA units: |p/m| <= 20 A/tau, |dx| <= 0.5 A per step,
// |F| <= 1e4 kcal/mol/A per atom.
constexpr float kVLimit = 20.0f;
constexpr float kMaxDx = 0.5f;
constexpr float kForceCap = 10000.0f;
enum { kWarnVLimit = 0, kWarnForceCap = 1, kWarnConstr = 2, kWarnDxCap = 3 };
// Per-slot health counters that are not safety caps: atom-steps that ended
// with a non-finite position or momentum (the explosion check of
// batched_simulate.py:158-205 and simulate.py:1050-1071, kept on the device).
enum { kWarnNonFinite = 4 };
constexpr int kWarnStride = 8;  // counters per slot

// Bonded entries; idx.x < 0 marks a free arena slot.
struct BondEntry { int4 idx; float4 p; };    // idx.xy atoms, idx.z kind (0 bond, 1 UB); p.xy = r0, k
struct AngleEntry { int4 idx; float4 p; };   // idx.xyz atoms; p.xy = theta0, k
// Up to two terms of one torsion: a = (n0, phase0, k0, kind), b = (n1, phase1, k1, n_terms).
// kind 0 dihedral, 1 periodic improper, 2 harmonic improper (a = (k, phi0, 0, 2), one term).
struct TorsionEntry { int4 idx; float4 a; float4 b; };
struct CmapEntry { int4 idx; int4 idx2; };   // idx = atoms a..d, idx2.x = atom e, .y = coeff base, .z = G
struct ExceptionEntry { int4 idx; float4 p; };  // idx.xy pair; p = (sigma, eps, qq * k_coulomb, 0)

// Neighbour list of a gb_cut system (one with a cutoff whose Born sums and
// GB terms are cut too). Built on the device from the tile boxes:
//  - per tile pair (I < J) without exclusions, a 64-bit mask of the atoms of J
//    within cutoff + skin of tile I's box (nl_mask), and the number of listed
//    atoms of the tiles I + 1 .. J - 1 before them (nl_pref);
//  - per tile I, its list: the masked atoms of every J > I in increasing
//    order (nl_entries from the row's entry base, nl_count[I] of them),
//    evaluated 64 at a time by chunk items (PairItem.j0 < 0, see below).
// Tile pairs with exclusions and the diagonal tiles stay ordinary items.
// Chunk item: j0 = -1 - (global entry index of its first j atom); excl = the
// global tile index of its row; ni = atoms of the i tile.
__host__ __device__ inline bool item_is_chunk(const PairItem& it) { return it.j0 < 0; }
// Per-tile-pair record: I | J << 16 (local tiles), and the slot (~slot for a
// pair with exclusions, which the list skips; x = -1 for a free record).
struct TilePairInfo { int32_t ij; int32_t slot; };

// Atoms per bonded entry kind; each entry writes one float4 force record per atom.
constexpr int kRecBond = 2, kRecAngle = 3, kRecTors = 4, kRecCmap = 5, kRecExc = 2;

// All device pointers of one engine.
struct DeviceView {
  // per atom (arena)
  float4* pos;         // x, y, z, 0
  float4* mom;         // px, py, pz, 1/m
  float4* force;       // integrator force (float), w unused
  float4* born4;       // rho, sr, sr^2, sr^3/2       (static)
  float4* nb4;         // q, sigma/2, sqrt(eps), 0    (static)
  double4* fin4;       // rho, R, q^2, ace_K          (static, FP64)
  float* sqrt_m;       // sqrt(mass)                  (static)
  int32_t* slot_of;    // slot id or -1
  float4* gb4;         // B, exp factor a = sqrt(log2(e)/4)/B, dB/dI, chain factor c
  float4* fpair;       // GB + Coulomb + LJ force of the last force pass
  double* e_atom;      // 2 per atom: GB self, ACE
  int32_t atom_hw;     // arena high-water mark
  // pair work
  const PairItem* items;
  int32_t n_items;
  float4* part;        // partial sums: item w, side s (0 = i, 1 = j), lane l -> part[(w * 2 + s) * 64 + l]
  const int4* tiles;   // per tile: (first atom, atoms in tile, slot, tile index T)
  int32_t n_tiles;
  const unsigned long long* excl_masks;  // 128 per exclusion tile: lane * 2 + {bit0, bit1}
  double* e_items;     // 3 per item: GB pairs, Coulomb, LJ
  float4* tbox;        // per tile t: (lo.xyz, max rho) at 2t, (hi.xyz, max s) at 2t + 1
  int32_t* item_cls;   // per item: class bits (kCls*), written by the Born pass
  // bonded
  BondEntry* bonds; int32_t bond_hw;
  AngleEntry* angles; int32_t angle_hw;
  TorsionEntry* tors; int32_t tors_hw;
  CmapEntry* cmaps; int32_t cmap_hw;
  ExceptionEntry* excs; int32_t exc_hw;
  const float* cmap_coeff;
  double* e_bond; double* e_angle; double* e_tors; double* e_cmap; double* e_exc;
  // bonded force records and their per-atom gather lists
  float4* rec;             // records; kind k entry e atom a -> rec_base[k] + e * atoms(k) + a
  int64_t rec_base[5];     // bond, angle, torsion, cmap, exception
  int2* atom_rec;          // per atom: (offset into rec_list, count)
  const int32_t* rec_list; // record indices
  // systems
  SlotConst* slots;
  uint32_t* atom_step;     // per atom: BAOAB steps taken (the noise counter)
  unsigned int* warn;      // per slot: kWarnStride counters (safety caps, non-finite atom-steps)
  int32_t max_slots;
  // neighbour lists (gb_cut systems)
  const TilePairInfo* nl_tp;     // per tile-pair record (static)
  int32_t nl_tp_hw;              // tile-pair arena high-water mark
  unsigned long long* nl_mask;   // per tile pair: listed atoms of J
  int32_t* nl_pref;              // per tile pair: listed atoms of the row before J
  int32_t* nl_entries;           // list entries (atom arena indices)
  const int4* nl_row;            // per tile: (first chunk item, entry base, first, end regular row item)
  int32_t* nl_count;             // per tile: entries in its list
  const int2* nl_col;            // per tile: (offset, count) of its regular column items in nl_colitems
  const int32_t* nl_colitems;    // item indices
  float4* nl_xref;               // per atom: position at the last list build
  int32_t* nl_flag;              // per slot: 1 = build the list at this evaluation
  // 1 in the captured step graphs of a pool: the pair passes and tile
  // gathers skip the systems that have reached their step limit (their
  // positions no longer change, so their Born radii, pair forces and stored
  // forces stand); 0 for evaluations, which compute every item
  int32_t skip_done;
};

// Grid-wide barrier state of the persistent step kernel (device memory).
struct GridBarrier {
  unsigned count;       // blocks arrived at the current barrier
  unsigned generation;  // barriers completed in this launch
  int error;            // 1: a wait timed out; every block leaves the launch
  int pad;
};

// Launchers (defined in the .hip translation units).
void launch_tile_boxes(const DeviceView& v, hipStream_t s);                // -> tbox (every evaluation)
// Neighbour-list build of the gb_cut systems (after the tile boxes): with
// force, every list is rebuilt; otherwise those whose atoms moved more than
// skin / 2 since their last build (every list, when skin = 0).
void launch_nl_build(const DeviceView& v, const int32_t* slot_ranges, int n_slots, bool force, hipStream_t s);
// Dense per-atom neighbour lists of raw (n, 3) positions (prx_neighbor_list_dense).
void launch_dense_list(const float* pos, int n, float radius, int capacity, int32_t* idx, int32_t* counts,
                       float4* boxes, hipStream_t s);
void launch_born_pairs(const DeviceView& v, hipStream_t s);                // also -> item_cls
void launch_born_pairs_bonded(const DeviceView& v, hipStream_t s);         // the same, and launch_bonded (forces only)
void launch_born_gather(const DeviceView& v, hipStream_t s);               // -> B, dB/dI
void launch_force_pairs(const DeviceView& v, bool energy, hipStream_t s);
void launch_force_gather(const DeviceView& v, bool energy, hipStream_t s);  // -> chain factors, pair force
void launch_chain_pairs(const DeviceView& v, hipStream_t s);
void launch_bonded(const DeviceView& v, bool energy, hipStream_t s);
// Sum chain partials + pair force + bonded records into the force, then
//   mode -1: store the force only; 2: B(F), count the step; 1: B(F) then B A O A.
void launch_chain_gather(const DeviceView& v, int mode, hipStream_t s);
void launch_step_begin(const DeviceView& v, hipStream_t s);  // B A O A from the stored force
void launch_energy_reduce(const DeviceView& v, const int32_t* slot_ranges, int n_slots, double* out,
                          hipStream_t s);
void launch_copy_in_positions(const DeviceView& v, const float* packed, const int32_t* map, int n,
                              hipStream_t s);
void launch_copy_out(const DeviceView& v, float* pos, float* mom, float* frc, const int32_t* map, int n,
                     hipStream_t s);
void launch_copy_in_momenta(const DeviceView& v, const float* packed, const int32_t* map, int n,
                            hipStream_t s);
void launch_fill_u32(uint32_t* dst, uint32_t value, int n, hipStream_t s);  // dst[0..n) = value
void launch_init_momenta(const DeviceView& v, const int32_t* slot_list, int n_slot_list, hipStream_t s);
void launch_kinetic(const DeviceView& v, const int32_t* slot_ranges, int n_slots, double* out,
                    hipStream_t s);
void launch_recenter_com(const DeviceView& v, const int32_t* slot_ranges, int n_slots, hipStream_t s);
void launch_tick(unsigned long long* out, hipStream_t s);
// Persistent step kernel (prx_persist.hip): the largest co-resident grid of
// its blocks (0 if cooperative launches are unavailable), and one launch of
// n_steps steps with n_blocks blocks (<= that grid).
int persist_max_blocks();
// Fused step kernel (prx_fused.hip): n_steps steps of every listed system,
// one workgroup per system (slot_ranges rows as for launch_kinetic).
void launch_fused_steps(const DeviceView& v, const int32_t* slot_ranges, int n_slots, int n_steps, hipStream_t s);
hipError_t launch_persist_steps(const DeviceView& v, const int32_t* slot_ranges, int n_slots, int n_steps,
                                bool has_nl, GridBarrier* bar, int n_blocks, hipStream_t s);
void launch_born_out(const DeviceView& v, float* out, const int32_t* map, int n, hipStream_t s);

}  // namespace prx
