// Engine: device arenas, per-system upload, work list, captured BAOAB graphs
// and the C ABI of include/prolix_gb.h.
//
// An engine holds any number of systems ("slots"). prx_system_create admits a
// fixed set sized exactly; a pool admits and releases systems while others
// keep stepping. Per-atom data lives in one arena (a system owns a contiguous
// range), bonded entries in per-kind arenas, so admission never moves another
// system's state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "prolix_gb.h"
#include "prx_common.h"
#include "prx_host.h"

namespace {

thread_local std::string g_error;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct CapacityError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct StateError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Bonded terms on a side stream beside the pair passes (1) or in line (0),
// from this many live atoms on (the default of the engine's runtime setting,
// prx_engine_set_fork_min_atoms): below it the fork's extra graph nodes cost
// more than the bonded kernel they hide (measured: 8% slower steps for one
// 138- or 2,500-atom system, 1.2% faster for the 425,157-atom mix).
#ifndef PRX_FORK_BONDED
#define PRX_FORK_BONDED 1
#endif
#ifndef PRX_FORK_MIN_ATOMS
#define PRX_FORK_MIN_ATOMS 100000
#endif

#define HIPCHECK(x)                                                                          \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess)                                                                    \
      throw HipError(std::string(#x) + ": " + hipGetErrorString(e_));                        \
  } while (0)

constexpr double kAkmaFs = 48.88821291839;  // 1 AKMA time unit in fs (simulate.py:56)
constexpr double kBoltzmann = 0.0019872041; // kcal/mol/K (simulate.py:57)

// First-fit allocator over [0, cap).
class RangeAlloc {
 public:
  explicit RangeAlloc(int64_t cap = 0) : cap_(cap) { if (cap > 0) free_.push_back({0, cap}); }
  int64_t alloc(int64_t n) {
    if (n == 0) return 0;
    for (size_t k = 0; k < free_.size(); ++k) {
      if (free_[k].second - free_[k].first >= n) {
        const int64_t b = free_[k].first;
        free_[k].first += n;
        if (free_[k].first == free_[k].second) free_.erase(free_.begin() + k);
        return b;
      }
    }
    return -1;
  }
  void release(int64_t b, int64_t n) {
    if (n == 0) return;
    free_.push_back({b, b + n});
    std::sort(free_.begin(), free_.end());
    std::vector<std::pair<int64_t, int64_t>> m;
    for (auto& r : free_) {
      if (!m.empty() && m.back().second == r.first) m.back().second = r.second;
      else m.push_back(r);
    }
    free_ = m;
  }
  int64_t high_water() const {
    // end of the last allocated byte: cap minus a free tail, if any
    if (free_.empty()) return cap_;
    if (free_.back().second == cap_) return free_.back().first;
    return cap_;
  }
  int64_t cap() const { return cap_; }

 private:
  int64_t cap_;
  std::vector<std::pair<int64_t, int64_t>> free_;
};

template <class T>
T* dalloc(size_t n) {
  if (n == 0) n = 1;
  void* p = nullptr;
  HIPCHECK(hipMalloc(&p, n * sizeof(T)));
  HIPCHECK(hipMemset(p, 0, n * sizeof(T)));
  return static_cast<T*>(p);
}

struct Slot {
  bool live = false;
  int32_t a0 = 0, n = 0;
  int64_t bond_b = 0, bond_n = 0, ang_b = 0, ang_n = 0, tors_b = 0, tors_n = 0, cmap_b = 0, cmap_n = 0;
  int64_t exc_b = 0, exc_n = 0, coeff_b = 0, coeff_n = 0, xt_b = 0, xt_n = 0, rl_b = 0, rl_n = 0;
  int64_t ent_b = 0, ent_n = 0, tp_b = 0, tp_n = 0;  // neighbour-list entries and tile-pair records
  bool nl = false;
  std::vector<int4> nl_rows;                 // per tile, relative (NlPlan::rows)
  std::vector<std::vector<int32_t>> nl_cols;  // per tile, relative item indices
  std::vector<prx::PairItem> items;  // global atom indices, fixed for the slot's lifetime
  prx::SlotConst c{};
  bool has_params = false;
};

// Torsion entries of one system: the nonzero periodic terms of each torsion
// in groups of two; one entry per nonzero harmonic improper term.
void torsion_entries(const prx_system_desc& d, std::vector<prx::TorsionEntry>* out, int64_t* count,
                     int a0 = 0) {
  int64_t n = 0;
  // the nonzero terms of one torsion, paired in order: one entry per pair
  auto emit_periodic = [&](const int32_t* idx, const double* prm, int T, float kind) {
    const double* first = nullptr;
    auto emit = [&](const double* t0, const double* t1) {
      ++n;
      if (!out) return;
      prx::TorsionEntry e;
      e.idx = make_int4(a0 + idx[0], a0 + idx[1], a0 + idx[2], a0 + idx[3]);
      e.a = make_float4((float)t0[0], (float)t0[1], (float)t0[2], kind);
      e.b = t1 ? make_float4((float)t1[0], (float)t1[1], (float)t1[2], 2.0f) : make_float4(0.f, 0.f, 0.f, 1.0f);
      out->push_back(e);
    };
    for (int k = 0; k < T; ++k) {
      const double* t = prm + (size_t)k * 3;
      if (t[2] == 0.0) continue;
      if (first) {
        emit(first, t);
        first = nullptr;
      } else {
        first = t;
      }
    }
    if (first) emit(first, nullptr);
  };
  for (int t = 0; t < d.n_dihedrals; ++t)
    emit_periodic(&d.dihedral_idx[4 * t], &d.dihedral_params[(size_t)t * d.dihedral_terms * 3], d.dihedral_terms,
                  0.0f);
  const int w = d.improper_param_width;
  for (int t = 0; t < d.n_impropers; ++t) {
    const double* prm = &d.improper_params[(size_t)t * d.improper_terms * w];
    if (w == 3) {
      emit_periodic(&d.improper_idx[4 * t], prm, d.improper_terms, 1.0f);
      continue;
    }
    for (int k = 0; k < d.improper_terms; ++k) {
      if (prm[(size_t)k * 2] == 0.0) continue;
      ++n;
      if (!out) continue;
      prx::TorsionEntry e;
      const int32_t* idx = &d.improper_idx[4 * t];
      e.idx = make_int4(a0 + idx[0], a0 + idx[1], a0 + idx[2], a0 + idx[3]);
      e.a = make_float4((float)prm[(size_t)k * 2], (float)prm[(size_t)k * 2 + 1], 0.0f, 2.0f);
      e.b = make_float4(0.f, 0.f, 0.f, 1.0f);
      out->push_back(e);
    }
  }
  if (count) *count = n;
}

int64_t count_torsion_entries(const prx_system_desc& d) {
  int64_t n = 0;
  torsion_entries(d, nullptr, &n);
  return n;
}

int64_t count_records(const prx_system_desc& d) {
  return (int64_t)prx::kRecBond * (d.n_bonds + d.n_ub) + (int64_t)prx::kRecAngle * d.n_angles +
         (int64_t)prx::kRecTors * count_torsion_entries(d) + (int64_t)prx::kRecCmap * d.n_cmap +
         (int64_t)prx::kRecExc * d.n_exceptions;
}

void check_idx(const int32_t* idx, int64_t rows, int cols, int n, const char* what) {
  for (int64_t r = 0; r < rows * cols; ++r)
    if (idx[r] < 0 || idx[r] >= n)
      throw std::invalid_argument(std::string(what) + " index out of range");
}

// A gb_cut system (all its pair terms cut) takes its pair work from a
// neighbour list built on the device.
bool uses_nl(const prx_system_desc& d) { return d.gb_cutoff && d.cutoff > 0; }

// Work items of one system alone, and its neighbour-list tile pairs and
// entries (0 for a system without a list).
int64_t plan_items(const prx_system_desc& d, const prx::ExclPlan& plan, int64_t* n_tp, int64_t* n_ent) {
  *n_tp = *n_ent = 0;
  const int64_t nT = (d.n_atoms + prx::kTile - 1) / prx::kTile;
  if (!uses_nl(d)) return nT * (nT + 1) / 2;
  int64_t items = 0;
  prx::nl_sizes(d.n_atoms, plan, n_tp, n_ent, &items);
  return items;
}

void validate(const prx_system_desc& d) {
  if (d.n_atoms <= 0) throw std::invalid_argument("n_atoms must be positive (charges is empty)");
  if (!d.charges || !d.sigmas || !d.epsilons || !d.radii)
    throw std::invalid_argument("charges, sigmas, epsilons and radii are required");
  if (d.n_bonds) check_idx(d.bond_idx, d.n_bonds, 2, d.n_atoms, "bond");
  if (d.n_ub) check_idx(d.ub_idx, d.n_ub, 2, d.n_atoms, "urey_bradley");
  if (d.n_angles) check_idx(d.angle_idx, d.n_angles, 3, d.n_atoms, "angle");
  if (d.n_dihedrals) {
    if (d.dihedral_terms <= 0) throw std::invalid_argument("dihedral_terms must be positive");
    check_idx(d.dihedral_idx, d.n_dihedrals, 4, d.n_atoms, "dihedral");
  }
  if (d.n_impropers) {
    if (d.improper_terms <= 0) throw std::invalid_argument("improper_terms must be positive");
    if (d.improper_param_width != 2 && d.improper_param_width != 3)
      throw std::invalid_argument("improper_param_width must be 2 (harmonic) or 3 (periodic)");
    check_idx(d.improper_idx, d.n_impropers, 4, d.n_atoms, "improper");
  }
  if (d.n_cmap) {
    check_idx(d.cmap_torsions, d.n_cmap, 5, d.n_atoms, "cmap torsion");
    if (d.cmap_grid < 2 || d.n_cmap_maps <= 0 || !d.cmap_grids)
      throw std::invalid_argument("cmap needs grids (n_cmap_maps, G, G)");
    for (int k = 0; k < d.n_cmap; ++k)
      if (d.cmap_map[k] < 0 || d.cmap_map[k] >= d.n_cmap_maps) throw std::invalid_argument("cmap map index out of range");
  }
  if (d.n_exceptions) check_idx(d.exc_pairs, d.n_exceptions, 2, d.n_atoms, "exception");
  if (d.cutoff < 0) throw std::invalid_argument("cutoff must be >= 0");
  if (d.implicit_solvent) {
    for (int i = 0; i < d.n_atoms; ++i)
      if (!(d.radii[i] - d.dielectric_offset > 0.0))
        throw std::invalid_argument("GB radii must exceed the dielectric offset");
  }
}

}  // namespace

// ---- allocation-free evaluation (prx_energy_forces_noalloc) ------------------
// Every buffer of the one-shot evaluation is carved from this file-scope
// device array. Each GPU the library runs on holds its own copy (loaded with
// the code object) until the process ends. Host code never passes the
// array's name to a kernel or copy: the device address comes from
// hipGetSymbolAddress, once per device, and is kept in g_scratch_ptr.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int kScratchMaxAtoms = 8192;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The scratch array's device address on the current device.
char* scratch_ptr() {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::invalid_argument("device index out of range for the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<char*>(g_scratch_ptr[dev]);
}
}  // namespace

namespace {
// Float32 patch coefficients of CMAP grids already seen: systems of one
// force field share their maps, so an admission finds them here instead of
// solving the splines again. Keyed by the grid's bytes (compared in full).
struct CmapCache {
  std::mutex mu;
  std::vector<std::pair<std::vector<double>, std::vector<float>>> entries;  // (grid, coefficients)
  static constexpr size_t kMaxEntries = 64;

  void get(const double* grid, int G, float* out) {
    const size_t n = (size_t)G * G;
    {
      std::lock_guard<std::mutex> lock(mu);
      for (const auto& e : entries)
        if (e.first.size() == n && std::memcmp(e.first.data(), grid, n * sizeof(double)) == 0) {
          std::memcpy(out, e.second.data(), e.second.size() * sizeof(float));
          return;
        }
    }
    std::vector<double> c64(n * 16);
    prx::cmap_coefficients(grid, G, c64.data());
    std::vector<float> c32(n * 16);
    for (size_t k = 0; k < c32.size(); ++k) c32[k] = (float)c64[k];
    std::memcpy(out, c32.data(), c32.size() * sizeof(float));
    std::lock_guard<std::mutex> lock(mu);
    if (entries.size() >= kMaxEntries) entries.erase(entries.begin());
    entries.emplace_back(std::vector<double>(grid, grid + n), std::move(c32));
  }
};
CmapCache g_cmap_cache;

// Pinned host memory from which a pool's admissions are copied to the device
// asynchronously on the pool's stream. Reset only once those copies are done
// (a synchronised stream).
struct PinnedStage {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  ~PinnedStage() {
    if (base) hipHostFree(base);
  }
  // Room for bytes more; false if the buffer must first grow.
  bool fits(size_t bytes) const { return used + ((bytes + 255) & ~(size_t)255) <= cap; }
  void grow(size_t at_least) {
    if (base) HIPCHECK(hipHostFree(base));
    base = nullptr;
    cap = std::max<size_t>(at_least, std::max<size_t>(2 * cap, (size_t)4 << 20));
    HIPCHECK(hipHostMalloc((void**)&base, cap, hipHostMallocDefault));
    used = 0;
  }
  char* take(size_t bytes) {
    char* p = base + used;
    used += (bytes + 255) & ~(size_t)255;
    return p;
  }
};
}  // namespace

struct prx_engine {
  // capacities
  int64_t cap_atoms = 0, cap_bond = 0, cap_ang = 0, cap_tors = 0, cap_cmap = 0, cap_exc = 0;
  int64_t cap_coeff = 0, cap_xt = 0, cap_rl = 0, cap_ent = 0, cap_tp = 0;
  int32_t cap_slots = 0;
  bool pool = false;
  RangeAlloc a_atoms, a_bond, a_ang, a_tors, a_cmap, a_exc, a_coeff, a_xt, a_rl, a_ent, a_tp;
  double nl_skin = 0.0;           // neighbour-list skin (A) of systems admitted from now on
  // neighbour-list device arrays
  prx::TilePairInfo* d_nl_tp = nullptr;
  unsigned long long* d_nl_mask = nullptr;
  int32_t* d_nl_pref = nullptr;
  int32_t* d_nl_entries = nullptr;
  float4* d_nl_xref = nullptr;
  int32_t* d_nl_flag = nullptr;
  int4* d_nl_row = nullptr;       // per tile (sized with the tiles)
  int32_t* d_nl_count = nullptr;  // per tile
  int2* d_nl_col = nullptr;       // per tile
  int32_t* d_nl_colitems = nullptr;
  int64_t cap_colitems = 0;
  bool has_nl = false;            // a live system uses its neighbour list
  // step path: 0 = auto (the fused kernel when every system fits in
  // kFusedAutoAtoms atoms and none uses a neighbour list, else graphs: on
  // MI355X the fused kernel beats graphs only there, measured from 138 to
  // 2,500 atoms), 1 = captured graphs, 2 = the persistent kernel, 3 = the
  // fused kernel (one workgroup per system)
  int32_t step_mode = 0;
  static constexpr int32_t kFusedAutoAtoms = 3 * prx::kTile;
  int32_t max_system_atoms = 0;   // largest live system (rebuild)
  prx::GridBarrier* d_bar = nullptr;
  int persist_blocks = -1;        // co-resident grid of the persistent kernel (0: unavailable)
  // pool slots admitted since the last step call: their momenta are drawn and
  // every live system's forces computed once for the whole batch (flush_pending)
  std::vector<int32_t> pending_init;
  // pools: admission uploads go through pinned memory as asynchronous copies
  // on the engine stream (h2d / d_fill); every reader synchronises that
  // stream first
  PinnedStage stage;
  int32_t* d_rec_list = nullptr;
  std::vector<Slot> slots;
  prx::DeviceView v{};
  // device buffers owned here
  prx::PairItem* d_items = nullptr;
  int64_t cap_items = 0;
  float4* d_part = nullptr;   // 2 * 64 partials per item
  // positions and forces of prx_energy_forces: an evaluation at arbitrary
  // positions leaves the integrator's state (v.pos, v.force) untouched
  float4* d_eval_pos = nullptr;
  float4* d_eval_force = nullptr;
  int64_t fork_min_atoms = PRX_FORK_MIN_ATOMS;
  int4* d_tiles = nullptr;
  int64_t cap_tiles = 0;
  float4* d_tbox = nullptr;
  int32_t* d_item_cls = nullptr;
  double* d_e_items = nullptr;
  unsigned long long* d_excl = nullptr;
  float* d_coeff = nullptr;
  int32_t* d_ranges = nullptr;  // 14 ints per listed slot
  int32_t* d_map = nullptr;     // packed atom -> arena index
  int32_t* d_slot_list = nullptr;
  double* d_energy_tmp = nullptr;
  int64_t cap_map = 0;
  std::vector<int32_t> order;   // live slots in packed order
  std::vector<int32_t> offsets; // packed atom offsets of `order`
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // bonded terms, forked from and joined to the pass stream
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipEvent_t ev_order = nullptr;  // orders pool uploads after the caller's stream
  bool dirty = true;            // work list / ranges need rebuilding
  bool initialized = false;
  std::vector<std::pair<int, hipGraphExec_t>> graphs;  // (steps per launch, executable)

  // Arena mode (arena != nullptr): every device buffer is carved, 256-byte
  // aligned and in a fixed order, from one caller-provided region of
  // arena_cap bytes; the engine then never allocates or frees device memory
  // and creates no streams (it runs on the caller's stream only).
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;

  template <class T>
  T* get(size_t n) {
    if (!arena) return dalloc<T>(n);
    if (n == 0) n = 1;
    const size_t off = (arena_used + 255) & ~(size_t)255;
    const size_t bytes = n * sizeof(T);
    if (off + bytes > arena_cap || off + bytes < off)
      throw CapacityError("the system's buffers exceed the " + std::to_string(arena_cap >> 20) +
                          " MiB scratch region");
    arena_used = off + bytes;
    return reinterpret_cast<T*>(arena + off);
  }

  // In arena mode the work list's buffers are sized now, exactly (the engine
  // holds one fixed set of systems); otherwise rebuild() sizes them.
  void alloc_work(int64_t n_items, int64_t n_tiles) {
    cap_items = n_items;
    d_items = get<prx::PairItem>(cap_items);
    d_e_items = get<double>(cap_items * 3);
    d_part = get<float4>(cap_items * 2 * prx::kTile);
    d_item_cls = get<int32_t>(cap_items);
    cap_tiles = n_tiles;
    d_tiles = get<int4>(cap_tiles);
    d_tbox = get<float4>(cap_tiles * 2);
    d_nl_row = get<int4>(cap_tiles);
    d_nl_count = get<int32_t>(cap_tiles);
    d_nl_col = get<int2>(cap_tiles);
    cap_colitems = n_items;
    d_nl_colitems = get<int32_t>(cap_colitems);
  }

  void alloc_device(int64_t arena_items = 0, int64_t arena_tiles = 0) {
    v.pos = get<float4>(cap_atoms);
    v.mom = get<float4>(cap_atoms);
    v.force = get<float4>(cap_atoms);
    v.born4 = get<float4>(cap_atoms);
    v.nb4 = get<float4>(cap_atoms);
    v.fin4 = get<double4>(cap_atoms);
    v.sqrt_m = get<float>(cap_atoms);
    v.slot_of = get<int32_t>(cap_atoms);
    v.gb4 = get<float4>(cap_atoms);
    v.fpair = get<float4>(cap_atoms);
    v.e_atom = get<double>(cap_atoms * 2);
    v.atom_step = get<uint32_t>(cap_atoms);
    v.bonds = get<prx::BondEntry>(cap_bond);
    v.angles = get<prx::AngleEntry>(cap_ang);
    v.tors = get<prx::TorsionEntry>(cap_tors);
    v.cmaps = get<prx::CmapEntry>(cap_cmap);
    v.excs = get<prx::ExceptionEntry>(cap_exc);
    v.e_bond = get<double>(cap_bond);
    v.e_angle = get<double>(cap_ang);
    v.e_tors = get<double>(cap_tors);
    v.e_cmap = get<double>(cap_cmap);
    v.e_exc = get<double>(cap_exc);
    v.rec_base[0] = 0;
    v.rec_base[1] = v.rec_base[0] + (int64_t)prx::kRecBond * cap_bond;
    v.rec_base[2] = v.rec_base[1] + (int64_t)prx::kRecAngle * cap_ang;
    v.rec_base[3] = v.rec_base[2] + (int64_t)prx::kRecTors * cap_tors;
    v.rec_base[4] = v.rec_base[3] + (int64_t)prx::kRecCmap * cap_cmap;
    const int64_t n_rec = v.rec_base[4] + (int64_t)prx::kRecExc * cap_exc;
    v.rec = get<float4>(n_rec);
    v.atom_rec = get<int2>(cap_atoms);
    cap_rl = std::max<int64_t>(cap_rl, 1);
    d_rec_list = get<int32_t>(cap_rl);
    v.rec_list = d_rec_list;
    v.warn = get<unsigned int>((size_t)cap_slots * prx::kWarnStride);
    d_eval_pos = get<float4>(cap_atoms);
    d_eval_force = get<float4>(cap_atoms);
    d_coeff = get<float>(cap_coeff);
    v.cmap_coeff = d_coeff;
    d_excl = get<unsigned long long>(cap_xt * 128);
    v.excl_masks = d_excl;
    v.slots = get<prx::SlotConst>(cap_slots);
    v.max_slots = cap_slots;
    d_ranges = get<int32_t>((size_t)cap_slots * 14);
    d_slot_list = get<int32_t>(cap_slots);
    d_energy_tmp = get<double>((size_t)cap_slots * PRX_N_ENERGY);
    cap_map = cap_atoms;
    d_map = get<int32_t>(cap_map);
    d_nl_tp = get<prx::TilePairInfo>(cap_tp);
    d_nl_mask = get<unsigned long long>(cap_tp);
    d_nl_pref = get<int32_t>(cap_tp);
    d_nl_entries = get<int32_t>(cap_ent);
    d_nl_xref = get<float4>(cap_atoms);
    d_nl_flag = get<int32_t>(cap_slots);
    d_bar = get<prx::GridBarrier>(1);
    v.nl_tp = d_nl_tp;
    v.nl_mask = d_nl_mask;
    v.nl_pref = d_nl_pref;
    v.nl_entries = d_nl_entries;
    v.nl_xref = d_nl_xref;
    v.nl_flag = d_nl_flag;
    if (arena) {
      alloc_work(arena_items, arena_tiles);
      // what the hipMalloc path zeroes per buffer (dalloc), cleared at once
      HIPCHECK(hipMemset(arena, 0, arena_used));
    }
    HIPCHECK(hipMemset(v.slot_of, 0xFF, std::max<int64_t>(cap_atoms, 1) * sizeof(int32_t)));
    mark_free_entries();
    HIPCHECK(hipMemset(d_nl_tp, 0xFF, std::max<int64_t>(cap_tp, 1) * sizeof(prx::TilePairInfo)));
    if (arena) return;
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_order, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  }

  template <class E>
  void free_entries(E* base, int64_t b, int64_t n) {
    if (n <= 0) return;
    std::vector<E> h((size_t)n);
    for (auto& e : h) { std::memset(&e, 0, sizeof(E)); e.idx.x = -1; }
    HIPCHECK(hipMemcpy(base + b, h.data(), n * sizeof(E), hipMemcpyHostToDevice));
  }

  void mark_free_entries() {
    free_entries(v.bonds, 0, cap_bond);
    free_entries(v.angles, 0, cap_ang);
    free_entries(v.tors, 0, cap_tors);
    free_entries(v.cmaps, 0, cap_cmap);
    free_entries(v.excs, 0, cap_exc);
  }

  void release_device() {
    drop_graphs();
    if (arena) return;  // nothing here was allocated
    void* ptrs[] = {v.pos, v.mom, v.force, v.born4, v.nb4, v.fin4, v.sqrt_m, v.slot_of, v.gb4, v.fpair, d_part, d_tiles,
                    v.e_atom, v.atom_step, v.bonds, v.angles, v.tors, v.cmaps, v.excs, v.e_bond, v.e_angle,
                    v.e_tors, v.e_cmap, v.e_exc, d_coeff, d_excl, v.slots, d_ranges, d_slot_list,
                    d_energy_tmp, d_map, d_items, d_e_items, v.rec, v.atom_rec, d_rec_list, v.warn,
                    d_tbox, d_item_cls, d_eval_pos, d_eval_force, d_nl_tp, d_nl_mask, d_nl_pref,
                    d_nl_entries, d_nl_xref, d_nl_flag, d_nl_row, d_nl_count, d_nl_col, d_nl_colitems, d_bar};
    for (void* p : ptrs)
      if (p) hipFree(p);
    if (ev_fork) hipEventDestroy(ev_fork);
    if (ev_order) hipEventDestroy(ev_order);
    if (ev_join) hipEventDestroy(ev_join);
    if (side) hipStreamDestroy(side);
    if (stream) hipStreamDestroy(stream);
  }

  // NULL is the device's legacy default stream (what torch's default stream
  // is), never the engine's own non-blocking stream, so stream-ordered callers
  // see their inputs and outputs ordered with our kernels.
  hipStream_t pick(void* s) const { return (hipStream_t)s; }

  // ---- admission -------------------------------------------------------
  int32_t admit(const prx_system_desc& d) {
    validate(d);
    prx::ExclPlan plan;
    prx::build_exclusions(d, plan);
    int32_t sid = -1;
    for (int32_t k = 0; k < (int32_t)slots.size(); ++k)
      if (!slots[k].live) { sid = k; break; }
    if (sid < 0) {
      if ((int32_t)slots.size() >= cap_slots) throw CapacityError("no free system slot");
      slots.emplace_back();
      sid = (int32_t)slots.size() - 1;
    }
    Slot s;
    s.n = d.n_atoms;
    const int64_t n_bond = d.n_bonds + d.n_ub, n_ang = d.n_angles, n_tors = count_torsion_entries(d);
    const int64_t n_cmap = d.n_cmap, n_exc = d.n_exceptions;
    const int64_t n_coeff = d.n_cmap ? (int64_t)d.n_cmap_maps * d.cmap_grid * d.cmap_grid * 16 : 0;
    const int64_t n_xt = (int64_t)plan.tiles.size();
    const int64_t n_rl = count_records(d);
    // A gb_cut system takes its pair work from a neighbour list when the
    // engine has room for its tile-pair records and entries (pools keep
    // none: their systems sweep all tile pairs, the same results).
    int64_t n_tp = 0, n_ent = 0, n_nl_items = 0;
    if (uses_nl(d)) prx::nl_sizes(d.n_atoms, plan, &n_tp, &n_ent, &n_nl_items);
    s.nl = uses_nl(d) && n_tp <= a_tp.cap() && n_ent <= a_ent.cap() && a_tp.cap() > 0;
    if (!s.nl) n_tp = n_ent = 0;
    struct Want { RangeAlloc* a; int64_t n; int64_t* b; };
    int64_t a0 = 0;
    Want w[] = {{&a_atoms, s.n, &a0}, {&a_bond, n_bond, &s.bond_b}, {&a_ang, n_ang, &s.ang_b},
                {&a_tors, n_tors, &s.tors_b}, {&a_cmap, n_cmap, &s.cmap_b}, {&a_exc, n_exc, &s.exc_b},
                {&a_coeff, n_coeff, &s.coeff_b}, {&a_xt, n_xt, &s.xt_b}, {&a_rl, n_rl, &s.rl_b},
                {&a_tp, n_tp, &s.tp_b}, {&a_ent, n_ent, &s.ent_b}};
    int got = 0;
    for (auto& x : w) {
      const int64_t b = x.a->alloc(x.n);
      if (b < 0) {
        for (int k = 0; k < got; ++k) w[k].a->release(*w[k].b, w[k].n);
        throw CapacityError("pool capacity exhausted");
      }
      *x.b = b;
      ++got;
    }
    s.a0 = (int32_t)a0;
    s.bond_n = n_bond; s.ang_n = n_ang; s.tors_n = n_tors; s.cmap_n = n_cmap; s.exc_n = n_exc;
    s.coeff_n = n_coeff; s.xt_n = n_xt; s.rl_n = n_rl; s.tp_n = n_tp; s.ent_n = n_ent;
    upload(d, sid, s, plan);
    if (s.nl) {
      prx::NlPlan np;
      prx::build_items_nl(s.n, s.a0, sid, plan, (int)s.xt_b, s.ent_b, np);
      s.items = std::move(np.items);
      s.nl_rows = std::move(np.rows);
      s.nl_cols = std::move(np.cols);
      if (!np.tps.empty()) h2d(d_nl_tp + s.tp_b, np.tps.data(), np.tps.size() * sizeof(prx::TilePairInfo));
      prx::SlotConst& c = s.c;
      c.nl = 1;
      c.tp_base = (int32_t)s.tp_b;
      const double rl = d.cutoff + nl_skin;
      c.nl_cut2 = (float)(rl * rl * 1.00001);
      c.skin_half2 = (float)(0.25 * nl_skin * nl_skin);
      push_const(sid, c);
      // the first evaluation builds the list
      const int32_t one = 1;
      h2d(d_nl_flag + sid, &one, sizeof(one));
    } else {
      prx::build_items(s.n, s.a0, sid, plan, (int)s.xt_b, s.items);
    }
    s.live = true;
    slots[sid] = std::move(s);
    order.push_back(sid);
    dirty = true;
    return sid;
  }

  void upload(const prx_system_desc& d, int32_t sid, Slot& s, const prx::ExclPlan& plan) {
    const int n = d.n_atoms;
    const double off = d.dielectric_offset;
    std::vector<float4> born4(n), nb4(n), mom(n, make_float4(0, 0, 0, 0));
    std::vector<double4> fin4(n);
    std::vector<float> sqm(n);
    std::vector<int32_t> slot_of(n, sid);
    // neighbour-list GB systems clamp the radii as generalized_born.py:302, 311, 331
    const bool nl = d.gb_cutoff && d.cutoff > 0;
    for (int i = 0; i < n; ++i) {
      const double R = d.radii[i];
      const double sr = d.scaled_radii ? d.scaled_radii[i] : R;
      const double rho = nl ? std::max(std::max(R, 1e-6) - off, 1e-4) : R - off;
      const double R_tanh = nl ? std::max(R, 1e-4) : R;
      const double q = d.charges[i];
      const double sig = std::max(d.sigmas[i], 1e-6);
      const double m = d.masses ? d.masses[i] : 1.0;
      if (!(m > 0.0)) throw std::invalid_argument("masses must be positive");
      born4[i] = make_float4((float)rho, (float)sr, (float)(sr * sr), (float)(0.5 * sr * sr * sr));
      nb4[i] = make_float4((float)q, (float)(0.5 * sig), (float)std::sqrt(std::max(d.epsilons[i], 0.0)), 0.0f);
      const double rn = (R - off) / 10.0 + 0.009;
      const double aceK = d.implicit_solvent
                              ? 0.239006 * 28.3919551 * (rn + 0.14) * (rn + 0.14) * std::pow(10.0 * rn, 6)
                              : 0.0;
      fin4[i] = make_double4(rho, R_tanh, q * q, aceK);
      sqm[i] = (float)std::sqrt(m);
      mom[i].w = (float)(1.0 / m);
    }
    h2d(v.born4 + s.a0, born4.data(), n * sizeof(float4));
    h2d(v.nb4 + s.a0, nb4.data(), n * sizeof(float4));
    h2d(v.fin4 + s.a0, fin4.data(), n * sizeof(double4));
    h2d(v.sqrt_m + s.a0, sqm.data(), n * sizeof(float));
    h2d(v.mom + s.a0, mom.data(), n * sizeof(float4));
    h2d(v.slot_of + s.a0, slot_of.data(), n * sizeof(int32_t));
    d_fill(v.atom_step + s.a0, 0, n * sizeof(uint32_t));
    d_fill(v.pos + s.a0, 0, n * sizeof(float4));
    d_fill(v.force + s.a0, 0, n * sizeof(float4));

    // bonded entries
    const int a0 = s.a0;
    if (s.bond_n) {
      std::vector<prx::BondEntry> e((size_t)s.bond_n);
      size_t k = 0;
      for (int b = 0; b < d.n_bonds; ++b, ++k) {
        e[k].idx = make_int4(a0 + d.bond_idx[2 * b], a0 + d.bond_idx[2 * b + 1], 0, 0);
        e[k].p = make_float4((float)d.bond_params[2 * b], (float)d.bond_params[2 * b + 1], 0, 0);
      }
      for (int b = 0; b < d.n_ub; ++b, ++k) {
        e[k].idx = make_int4(a0 + d.ub_idx[2 * b], a0 + d.ub_idx[2 * b + 1], 1, 0);
        e[k].p = make_float4((float)d.ub_params[2 * b], (float)d.ub_params[2 * b + 1], 0, 0);
      }
      h2d(v.bonds + s.bond_b, e.data(), e.size() * sizeof(e[0]));
    }
    if (s.ang_n) {
      std::vector<prx::AngleEntry> e((size_t)s.ang_n);
      for (int b = 0; b < d.n_angles; ++b) {
        e[b].idx = make_int4(a0 + d.angle_idx[3 * b], a0 + d.angle_idx[3 * b + 1], a0 + d.angle_idx[3 * b + 2], 0);
        e[b].p = make_float4((float)d.angle_params[2 * b], (float)d.angle_params[2 * b + 1], 0, 0);
      }
      h2d(v.angles + s.ang_b, e.data(), e.size() * sizeof(e[0]));
    }
    if (s.tors_n) {
      std::vector<prx::TorsionEntry> e;
      e.reserve((size_t)s.tors_n);
      torsion_entries(d, &e, nullptr, a0);
      h2d(v.tors + s.tors_b, e.data(), e.size() * sizeof(e[0]));
    }
    if (s.cmap_n) {
      const int G = d.cmap_grid;
      std::vector<float> c32((size_t)s.coeff_n);
      for (int m = 0; m < d.n_cmap_maps; ++m)
        g_cmap_cache.get(d.cmap_grids + (size_t)m * G * G, G, c32.data() + (size_t)m * G * G * 16);
      h2d(d_coeff + s.coeff_b, c32.data(), c32.size() * sizeof(float));
      std::vector<prx::CmapEntry> e((size_t)s.cmap_n);
      for (int k = 0; k < d.n_cmap; ++k) {
        const int32_t* t = &d.cmap_torsions[5 * k];
        e[k].idx = make_int4(a0 + t[0], a0 + t[1], a0 + t[2], a0 + t[3]);
        e[k].idx2 = make_int4(a0 + t[4], (int)(s.coeff_b + (int64_t)d.cmap_map[k] * G * G * 16), G, 0);
      }
      h2d(v.cmaps + s.cmap_b, e.data(), e.size() * sizeof(e[0]));
    }
    if (s.exc_n) {
      std::vector<prx::ExceptionEntry> e((size_t)s.exc_n);
      for (int k = 0; k < d.n_exceptions; ++k) {
        e[k].idx = make_int4(a0 + d.exc_pairs[2 * k], a0 + d.exc_pairs[2 * k + 1], 0, 0);
        e[k].p = make_float4((float)d.exc_sigma[k], (float)d.exc_epsilon[k],
                             (float)(d.exc_chargeprod[k] * d.coulomb_constant), 0.0f);
      }
      h2d(v.excs + s.exc_b, e.data(), e.size() * sizeof(e[0]));
    }
    if (s.xt_n)
      h2d(d_excl + s.xt_b * 128, plan.tile_masks.data(), plan.tile_masks.size() * 8);
    upload_record_lists(d, s);
    // constants
    prx::SlotConst& c = s.c;
    std::memset(&c, 0, sizeof(c));
    c.a0 = s.a0;
    c.n = n;
    c.active = 1;
    c.cutoff2 = d.cutoff > 0 ? (float)(d.cutoff * d.cutoff) : 0.0f;
    c.gb_cut = (d.gb_cutoff && d.cutoff > 0) ? 1 : 0;
    c.k_coulomb = (float)d.coulomb_constant;
    const double pref = d.implicit_solvent
                            ? -0.5 * d.gb_coulomb_constant * (1.0 / d.solute_dielectric - 1.0 / d.solvent_dielectric)
                            : 0.0;
    c.gb_pref = pref;
    c.gb2 = (float)(2.0 * pref);
    c.excl_v[0] = c.excl_e[0] = 1.0f;
    for (size_t k = 0; k < plan.classes.size(); ++k) {
      c.excl_v[k + 1] = (float)plan.classes[k].first;
      c.excl_e[k + 1] = (float)plan.classes[k].second;
    }
    push_const(sid, c);
  }

  // Per-atom lists of the bonded force records that land on each atom, in a
  // fixed order (entry kind, then entry, then position in the entry).
  void upload_record_lists(const prx_system_desc& d, const Slot& s) {
    const int n = s.n;
    // Each atom's records in the order bonds, UB, angles, torsions, CMAP,
    // exceptions (entry order within each), laid out by a counting sort.
    std::vector<prx::TorsionEntry> te;
    torsion_entries(d, &te, nullptr, 0);
    auto visit = [&](auto&& add) {
      int64_t k = 0;
      for (int b = 0; b < d.n_bonds; ++b, ++k)
        for (int a = 0; a < 2; ++a) add(d.bond_idx[2 * b + a], v.rec_base[0] + (s.bond_b + k) * prx::kRecBond + a);
      for (int b = 0; b < d.n_ub; ++b, ++k)
        for (int a = 0; a < 2; ++a) add(d.ub_idx[2 * b + a], v.rec_base[0] + (s.bond_b + k) * prx::kRecBond + a);
      for (int b = 0; b < d.n_angles; ++b)
        for (int a = 0; a < 3; ++a) add(d.angle_idx[3 * b + a], v.rec_base[1] + (s.ang_b + b) * prx::kRecAngle + a);
      for (size_t t = 0; t < te.size(); ++t) {
        const int ids[4] = {te[t].idx.x, te[t].idx.y, te[t].idx.z, te[t].idx.w};
        for (int a = 0; a < 4; ++a) add(ids[a], v.rec_base[2] + (s.tors_b + (int64_t)t) * prx::kRecTors + a);
      }
      for (int c = 0; c < d.n_cmap; ++c)
        for (int a = 0; a < 5; ++a)
          add(d.cmap_torsions[5 * c + a], v.rec_base[3] + (s.cmap_b + c) * prx::kRecCmap + a);
      for (int x = 0; x < d.n_exceptions; ++x)
        for (int a = 0; a < 2; ++a) add(d.exc_pairs[2 * x + a], v.rec_base[4] + (s.exc_b + x) * prx::kRecExc + a);
    };
    std::vector<int2> ar(n, make_int2(0, 0));
    visit([&](int atom, int64_t) { ar[atom].y += 1; });
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
      ar[i].x = (int)(s.rl_b + total);
      total += ar[i].y;
    }
    if (total != s.rl_n) throw std::logic_error("bonded record count mismatch");
    std::vector<int32_t> flat((size_t)total);
    std::vector<int32_t> fill(n, 0);
    visit([&](int atom, int64_t rec) { flat[(size_t)(ar[atom].x - s.rl_b) + fill[atom]++] = (int32_t)rec; });
    h2d(v.atom_rec + s.a0, ar.data(), n * sizeof(int2));
    if (!flat.empty())
      h2d(d_rec_list + s.rl_b, flat.data(), flat.size() * sizeof(int32_t));
  }

  // Host-to-device copy and fill of admission data: asynchronous from pinned
  // memory on the engine stream for pools, synchronous otherwise.
  void h2d(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    if (!pool) {
      HIPCHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
      return;
    }
    if (!stage.fits(bytes)) {
      HIPCHECK(hipStreamSynchronize(stream));  // the copies still reading the buffer
      stage.used = 0;
      if (!stage.fits(bytes)) stage.grow(bytes + 256);
    }
    char* p = stage.take(bytes);
    std::memcpy(p, src, bytes);
    HIPCHECK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, stream));
  }
  void d_fill(void* dst, int byte, size_t bytes) {
    if (!bytes) return;
    if (pool)
      HIPCHECK(hipMemsetAsync(dst, byte, bytes, stream));
    else
      HIPCHECK(hipMemset(dst, byte, bytes));
  }
  void push_const(int32_t sid, const prx::SlotConst& c) { h2d(v.slots + sid, &c, sizeof(c)); }

  void set_params(int32_t sid, const prx_langevin_params& p) {
    if (!(p.dt_fs > 0.0) || p.temperature < 0.0 || p.gamma_ps < 0.0)
      throw std::invalid_argument("integrator needs dt_fs > 0, temperature >= 0, gamma_ps >= 0");
    prx::SlotConst& c = slots[sid].c;
    const double dt = p.dt_fs / kAkmaFs;
    const double gamma = p.gamma_ps * kAkmaFs * 1e-3;
    const double kT = kBoltzmann * p.temperature;
    const double c1 = std::exp(-gamma * dt);
    c.dt = (float)dt;
    c.half_dt = (float)(0.5 * dt);
    c.c1 = (float)c1;
    c.c2 = (float)std::sqrt(std::max(0.0, 1.0 - c1 * c1));
    c.kT = (float)kT;
    c.sqrt_kT = (float)std::sqrt(kT);
    c.caps = p.safety_caps ? 1 : 0;
    c.seed = p.seed;
    d_fill(v.warn + (size_t)sid * prx::kWarnStride, 0, prx::kWarnStride * sizeof(unsigned int));
    slots[sid].has_params = true;
    push_const(sid, c);
  }

  void release(int32_t sid) {
    if (sid < 0 || sid >= (int32_t)slots.size() || !slots[sid].live) throw std::invalid_argument("no such live slot");
    pending_init.erase(std::remove(pending_init.begin(), pending_init.end(), sid), pending_init.end());
    Slot& s = slots[sid];
    // in stream order: after the work already queued that reads the system,
    // before any later admission's uploads into the same ranges
    d_fill(v.slot_of + s.a0, 0xFF, s.n * sizeof(int32_t));
    d_fill(v.atom_rec + s.a0, 0, s.n * sizeof(int2));
    // entries all-ones: idx.x = -1 marks them free, the only field the
    // kernels read of a free entry
    d_fill(v.bonds + s.bond_b, 0xFF, s.bond_n * sizeof(prx::BondEntry));
    d_fill(v.angles + s.ang_b, 0xFF, s.ang_n * sizeof(prx::AngleEntry));
    d_fill(v.tors + s.tors_b, 0xFF, s.tors_n * sizeof(prx::TorsionEntry));
    d_fill(v.cmaps + s.cmap_b, 0xFF, s.cmap_n * sizeof(prx::CmapEntry));
    d_fill(v.excs + s.exc_b, 0xFF, s.exc_n * sizeof(prx::ExceptionEntry));
    a_atoms.release(s.a0, s.n);
    a_bond.release(s.bond_b, s.bond_n);
    a_ang.release(s.ang_b, s.ang_n);
    a_tors.release(s.tors_b, s.tors_n);
    a_cmap.release(s.cmap_b, s.cmap_n);
    a_exc.release(s.exc_b, s.exc_n);
    a_coeff.release(s.coeff_b, s.coeff_n);
    a_xt.release(s.xt_b, s.xt_n);
    a_rl.release(s.rl_b, s.rl_n);
    if (s.tp_n) d_fill(d_nl_tp + s.tp_b, 0xFF, s.tp_n * sizeof(prx::TilePairInfo));
    a_tp.release(s.tp_b, s.tp_n);
    a_ent.release(s.ent_b, s.ent_n);
    s = Slot();
    order.erase(std::remove(order.begin(), order.end(), sid), order.end());
    dirty = true;
  }

  // ---- work list and ranges ---------------------------------------------
  void rebuild() {
    if (!dirty) return;
    // the work lists are rewritten below: no kernel may still read them (the
    // allocation-free entry point has synchronised its stream already)
    if (!arena) HIPCHECK(hipDeviceSynchronize());
    std::vector<prx::PairItem> items;
    std::vector<int32_t> ranges;
    std::vector<int32_t> map;
    std::vector<int4> tiles;
    std::vector<int4> nl_row;
    std::vector<int2> nl_col;
    std::vector<int32_t> nl_colitems;
    offsets.assign(1, 0);
    max_system_atoms = 0;
    has_nl = false;
    for (int32_t sid : order) {
      Slot& s = slots[sid];
      const int32_t ib = (int32_t)items.size();
      s.c.item_base = ib;
      s.c.tile0 = (int32_t)tiles.size();
      push_const(sid, s.c);
      const int nT = (s.n + prx::kTile - 1) / prx::kTile;
      for (int T = 0; T < nT; ++T) {
        tiles.push_back(make_int4(s.a0 + T * prx::kTile, std::min(prx::kTile, s.n - T * prx::kTile), sid, T));
        if (s.nl) {
          const int4 r = s.nl_rows[T];
          nl_row.push_back(make_int4(ib + r.x, (int)s.ent_b + r.y, ib + r.z, ib + r.w));
          nl_col.push_back(make_int2((int)nl_colitems.size(), (int)s.nl_cols[T].size()));
          for (int32_t k : s.nl_cols[T]) nl_colitems.push_back(ib + k);
        } else {
          nl_row.push_back(make_int4(0, 0, 0, 0));
          nl_col.push_back(make_int2(0, 0));
        }
      }
      has_nl = has_nl || s.nl;
      items.insert(items.end(), s.items.begin(), s.items.end());
      if (s.nl)  // chunk items name their row's global tile
        for (size_t k = items.size() - s.items.size(); k < items.size(); ++k)
          if (prx::item_is_chunk(items[k])) items[k].excl += s.c.tile0;
      const int32_t r[14] = {ib, (int32_t)items.size(), s.a0, s.n, (int32_t)s.bond_b, (int32_t)(s.bond_b + s.bond_n),
                             (int32_t)s.ang_b, (int32_t)(s.ang_b + s.ang_n), (int32_t)s.tors_b,
                             (int32_t)(s.tors_b + s.tors_n), (int32_t)s.cmap_b, (int32_t)(s.cmap_b + s.cmap_n),
                             (int32_t)s.exc_b, (int32_t)(s.exc_b + s.exc_n)};
      ranges.insert(ranges.end(), r, r + 14);
      for (int i = 0; i < s.n; ++i) map.push_back(s.a0 + i);
      offsets.push_back(offsets.back() + s.n);
      max_system_atoms = std::max(max_system_atoms, s.n);
    }
    if ((int64_t)items.size() > cap_items) {
      if (arena) throw std::logic_error("arena engine: work list larger than sized");
      if (d_items) HIPCHECK(hipFree(d_items));
      if (d_e_items) HIPCHECK(hipFree(d_e_items));
      if (d_part) HIPCHECK(hipFree(d_part));
      cap_items = std::max<int64_t>((int64_t)items.size(), cap_items * 3 / 2);
      if (d_item_cls) HIPCHECK(hipFree(d_item_cls));
      d_items = dalloc<prx::PairItem>(cap_items);
      d_e_items = dalloc<double>(cap_items * 3);
      d_part = dalloc<float4>(cap_items * 2 * prx::kTile);
      d_item_cls = dalloc<int32_t>(cap_items);
    }
    if ((int64_t)tiles.size() > cap_tiles) {
      if (arena) throw std::logic_error("arena engine: tile list larger than sized");
      void* old[] = {d_tiles, d_tbox, d_nl_row, d_nl_count, d_nl_col};
      for (void* p : old)
        if (p) HIPCHECK(hipFree(p));
      cap_tiles = std::max<int64_t>((int64_t)tiles.size(), cap_tiles * 3 / 2);
      d_tiles = dalloc<int4>(cap_tiles);
      d_tbox = dalloc<float4>(cap_tiles * 2);
      d_nl_row = dalloc<int4>(cap_tiles);
      d_nl_count = dalloc<int32_t>(cap_tiles);
      d_nl_col = dalloc<int2>(cap_tiles);
    }
    if ((int64_t)nl_colitems.size() > cap_colitems) {
      if (arena) throw std::logic_error("arena engine: column lists larger than sized");
      if (d_nl_colitems) HIPCHECK(hipFree(d_nl_colitems));
      cap_colitems = std::max<int64_t>((int64_t)nl_colitems.size(), cap_colitems * 3 / 2);
      d_nl_colitems = dalloc<int32_t>(cap_colitems);
    }
    if (!nl_row.empty()) {
      HIPCHECK(hipMemcpy(d_nl_row, nl_row.data(), nl_row.size() * sizeof(int4), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(d_nl_col, nl_col.data(), nl_col.size() * sizeof(int2), hipMemcpyHostToDevice));
      // every list is rebuilt at the next evaluation (tiles may have moved in the arrays)
      HIPCHECK(hipMemset(d_nl_count, 0, nl_row.size() * sizeof(int32_t)));
    }
    if (!nl_colitems.empty())
      HIPCHECK(hipMemcpy(d_nl_colitems, nl_colitems.data(), nl_colitems.size() * sizeof(int32_t),
                         hipMemcpyHostToDevice));
    if (!tiles.empty())
      HIPCHECK(hipMemcpy(d_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
    if (!items.empty())
      HIPCHECK(hipMemcpy(d_items, items.data(), items.size() * sizeof(prx::PairItem), hipMemcpyHostToDevice));
    if (!ranges.empty())
      HIPCHECK(hipMemcpy(d_ranges, ranges.data(), ranges.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!map.empty()) HIPCHECK(hipMemcpy(d_map, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    v.items = d_items;
    v.n_items = (int32_t)items.size();
    v.e_items = d_e_items;
    v.part = d_part;
    v.tiles = d_tiles;
    v.n_tiles = (int32_t)tiles.size();
    v.tbox = d_tbox;
    v.item_cls = d_item_cls;
    v.nl_row = d_nl_row;
    v.nl_count = d_nl_count;
    v.nl_col = d_nl_col;
    v.nl_colitems = d_nl_colitems;
    v.nl_tp_hw = (int32_t)a_tp.high_water();
    v.atom_hw = (int32_t)a_atoms.high_water();
    v.bond_hw = (int32_t)a_bond.high_water();
    v.angle_hw = (int32_t)a_ang.high_water();
    v.tors_hw = (int32_t)a_tors.high_water();
    v.cmap_hw = (int32_t)a_cmap.high_water();
    v.exc_hw = (int32_t)a_exc.high_water();
    drop_graphs();
    dirty = false;
  }

  int32_t total_atoms() const { return offsets.back(); }
  bool fork_bonded() const { return PRX_FORK_BONDED && total_atoms() >= fork_min_atoms; }

  // The view of the captured steps: systems at their step limit skip the
  // pair passes and the tile gathers, keeping the forces of their final
  // positions. Only pools set step limits.
  prx::DeviceView step_view() const {
    prx::DeviceView g = v;
    g.skip_done = pool ? 1 : 0;
    return g;
  }

  // The engine's view with the evaluation buffers in place of the
  // integrator's positions and forces (prx_energy_forces).
  prx::DeviceView eval_view() const {
    prx::DeviceView e = v;
    e.pos = d_eval_pos;
    e.force = d_eval_force;
    return e;
  }

  // ---- evaluation -------------------------------------------------------
  // Everything of one force evaluation up to the chain gather. The bonded
  // terms read only positions and write their records and term energies,
  // which only the chain gather and the energy reduction read: on engines
  // where fork_bonded() holds they run on the side stream beside the pair
  // passes and join before this returns (in a captured graph, a parallel
  // branch).
  // boxes = false when the previous chain gather (mode 1) already wrote the
  // tile boxes of the current positions.
  void force_passes(const prx::DeviceView& vv, bool energy, hipStream_t s, bool boxes = true) {
    const bool fork = fork_bonded();
    if (fork) {
      HIPCHECK(hipEventRecord(ev_fork, s));
      HIPCHECK(hipStreamWaitEvent(side, ev_fork, 0));
      prx::launch_bonded(vv, energy, side);
      HIPCHECK(hipEventRecord(ev_join, side));
    }
    if (boxes) prx::launch_tile_boxes(vv, s);
    // every list is rebuilt where the boxes are (eager evaluations, the first
    // step of a graph launch); later steps rebuild by the skin rule
    if (has_nl) prx::launch_nl_build(vv, d_ranges, (int)order.size(), boxes, s);
    // in line, the bonded forces ride in the Born pass's launch (one graph
    // node less per step); energy evaluations run the bonded kernel
    const bool merged = !fork && !energy;
    if (merged)
      prx::launch_born_pairs_bonded(vv, s);
    else
      prx::launch_born_pairs(vv, s);
    prx::launch_born_gather(vv, s);
    prx::launch_force_pairs(vv, energy, s);
    prx::launch_force_gather(vv, energy, s);
    if (fork) {
      prx::launch_chain_pairs(vv, s);
      HIPCHECK(hipStreamWaitEvent(s, ev_join, 0));
    } else {
      if (!merged) prx::launch_bonded(vv, energy, s);
      prx::launch_chain_pairs(vv, s);
    }
  }

  void evaluate(const prx::DeviceView& vv, bool energy, hipStream_t s) {
    force_passes(vv, energy, s);
    prx::launch_chain_gather(vv, -1, s);
  }
  void evaluate(bool energy, hipStream_t s) { evaluate(v, energy, s); }

  void drop_graphs() {
    for (auto& g : graphs) hipGraphExecDestroy(g.second);
    graphs.clear();
  }

  hipGraphExec_t capture(int n_steps) {
    for (auto& g : graphs)
      if (g.first == n_steps) return g.second;
    hipGraphExec_t exec = nullptr;
    hipGraph_t g = nullptr;
    HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
    try {
      prx::launch_step_begin(v, stream);
      const prx::DeviceView sv = step_view();
      for (int k = 0; k < n_steps; ++k) {
        // after the first step, the previous chain gather wrote the boxes
        force_passes(sv, false, stream, k == 0);
        prx::launch_chain_gather(sv, k + 1 == n_steps ? 2 : 1, stream);
      }
      HIPCHECK(hipGetLastError());
    } catch (...) {
      // leave the stream out of capture mode before reporting the failure
      if (hipStreamEndCapture(stream, &g) == hipSuccess && g) hipGraphDestroy(g);
      throw;
    }
    HIPCHECK(hipStreamEndCapture(stream, &g));
    HIPCHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    HIPCHECK(hipGraphDestroy(g));
    graphs.push_back({n_steps, exec});
    return exec;
  }

  // Momenta of the pool slots admitted since the last flush, then the forces
  // of every live system at its current positions (a deterministic
  // evaluation: the other systems' stored forces are reproduced bit for bit).
  void flush_pending(hipStream_t s) {
    if (pending_init.empty()) return;
    rebuild();
    // the slot list may still be read by an earlier momentum draw
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipStreamSynchronize(stream));
    stage.used = 0;  // every staged upload has landed
    HIPCHECK(hipMemcpy(d_slot_list, pending_init.data(), pending_init.size() * sizeof(int32_t),
                       hipMemcpyHostToDevice));
    prx::launch_init_momenta(v, d_slot_list, (int)pending_init.size(), s);
    evaluate(false, s);
    HIPCHECK(hipGetLastError());
    pending_init.clear();
  }

  // Steps per persistent or fused launch: long enough to amortise the launch, short
  // enough that one launch stays well under a second on the largest engine.
  static constexpr int kPersistChunk = 1000;

  bool use_persistent() {
    if (step_mode != 2) return false;
    if (persist_blocks < 0) persist_blocks = prx::persist_max_blocks();
    if (persist_blocks <= 0) throw StateError("the persistent step kernel needs cooperative launches");
    return true;
  }

  // The fused kernel: every system in one workgroup of its own. Not for
  // neighbour-list systems (their list build spans workgroups).
  bool use_fused() const {
    if (step_mode == 0) return !has_nl && max_system_atoms <= kFusedAutoAtoms;
    if (step_mode != 3) return false;
    if (has_nl) throw StateError("the fused step kernel does not run neighbour-list systems");
    return true;
  }

  // The persistent kernel: at most one item per wave and one tile per block
  // in each pass, within the co-resident grid.
  void steps_persistent(int n_steps, hipStream_t s) {
    const int64_t want = std::max<int64_t>({(int64_t)v.n_tiles, ((int64_t)v.n_items + 3) / 4, 1});
    const int blocks = (int)std::min<int64_t>(want, persist_blocks);
    for (int left = n_steps; left > 0;) {
      const int k = std::min(left, kPersistChunk);
      HIPCHECK(prx::launch_persist_steps(v, d_ranges, (int)order.size(), k, has_nl, d_bar, blocks, s));
      left -= k;
    }
    // a barrier that timed out leaves the error word set: report it
    prx::GridBarrier b{};
    HIPCHECK(hipMemcpyAsync(&b, d_bar, sizeof(b), hipMemcpyDeviceToHost, s));
    HIPCHECK(hipStreamSynchronize(s));
    if (b.error) throw HipError("persistent step kernel: a grid barrier timed out");
  }

  void steps(int n_steps, hipStream_t s) {
    flush_pending(s);
    rebuild();
    if (order.empty() || n_steps <= 0) return;
    // the steps run on the engine stream, ordered after the caller's
    hipEvent_t ev;
    HIPCHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    try {
      if (s != stream) {
        HIPCHECK(hipEventRecord(ev, s));
        HIPCHECK(hipStreamWaitEvent(stream, ev, 0));
      }
      if (use_fused()) {
        for (int left = n_steps; left > 0;) {
          const int k = std::min(left, kPersistChunk);
          prx::launch_fused_steps(v, d_ranges, (int)order.size(), k, stream);
          HIPCHECK(hipGetLastError());
          left -= k;
        }
      } else if (use_persistent()) {
        steps_persistent(n_steps, stream);
      } else {
        const int chunk = std::min(n_steps, 50);
        for (int left = n_steps; left > 0;) {
          const int k = std::min(chunk, left);
          HIPCHECK(hipGraphLaunch(capture(k), stream));
          left -= k;
        }
      }
      if (s != stream) {
        HIPCHECK(hipEventRecord(ev, stream));
        HIPCHECK(hipStreamWaitEvent(s, ev, 0));
      }
    } catch (...) {
      hipEventDestroy(ev);
      throw;
    }
    HIPCHECK(hipEventDestroy(ev));
  }
};

namespace {

template <class F>
int guard(F&& f) {
  try {
    f();
    g_error.clear();
    return PRX_OK;
  } catch (const CapacityError& e) {
    g_error = e.what();
    return PRX_ERR_CAPACITY;
  } catch (const StateError& e) {
    g_error = e.what();
    return PRX_ERR_STATE;
  } catch (const HipError& e) {
    g_error = e.what();
    return PRX_ERR_HIP;
  } catch (const std::invalid_argument& e) {
    g_error = e.what();
    return PRX_ERR_INVALID;
  } catch (const std::exception& e) {
    g_error = e.what();
    return PRX_ERR_INVALID;
  }
}

void need(bool c, const char* msg) {
  if (!c) throw std::invalid_argument(msg);
}

}  // namespace

extern "C" {

const char* prx_last_error(void) { return g_error.c_str(); }
int32_t prx_abi_version(void) { return PRX_ABI_VERSION; }

int prx_system_create(const prx_system_desc* descs, int32_t n_systems, prx_engine** out) {
  return guard([&] {
    need(descs != nullptr && out != nullptr && n_systems > 0, "prx_system_create: bad arguments");
    // a work item keeps its system's slot in 16 bits (PairItem.meta)
    need(n_systems < 65536, "prx_system_create: at most 65535 systems per engine");
    *out = nullptr;
    auto eng = std::make_unique<prx_engine>();
    for (int k = 0; k < n_systems; ++k) {
      const prx_system_desc& d = descs[k];
      validate(d);
      prx::ExclPlan plan;
      prx::build_exclusions(d, plan);
      eng->cap_atoms += d.n_atoms;
      eng->cap_bond += d.n_bonds + d.n_ub;
      eng->cap_ang += d.n_angles;
      eng->cap_tors += count_torsion_entries(d);
      eng->cap_cmap += d.n_cmap;
      eng->cap_exc += d.n_exceptions;
      eng->cap_coeff += d.n_cmap ? (int64_t)d.n_cmap_maps * d.cmap_grid * d.cmap_grid * 16 : 0;
      eng->cap_xt += (int64_t)plan.tiles.size();
      eng->cap_rl += count_records(d);
      int64_t n_tp = 0, n_ent = 0;
      plan_items(d, plan, &n_tp, &n_ent);
      eng->cap_tp += n_tp;
      eng->cap_ent += n_ent;
    }
    eng->cap_slots = n_systems;
    eng->a_atoms = RangeAlloc(eng->cap_atoms);
    eng->a_bond = RangeAlloc(eng->cap_bond);
    eng->a_ang = RangeAlloc(eng->cap_ang);
    eng->a_tors = RangeAlloc(eng->cap_tors);
    eng->a_cmap = RangeAlloc(eng->cap_cmap);
    eng->a_exc = RangeAlloc(eng->cap_exc);
    eng->a_coeff = RangeAlloc(eng->cap_coeff);
    eng->a_xt = RangeAlloc(eng->cap_xt);
    eng->a_rl = RangeAlloc(std::max<int64_t>(eng->cap_rl, 1));
    eng->a_tp = RangeAlloc(eng->cap_tp);
    eng->a_ent = RangeAlloc(eng->cap_ent);
    try {
      scratch_ptr();  // set-up of the allocation-free entry point's scratch array on this device
      eng->alloc_device();
      for (int k = 0; k < n_systems; ++k) eng->admit(descs[k]);
      eng->rebuild();
    } catch (...) {
      eng->release_device();
      throw;
    }
    *out = eng.release();
  });
}

int prx_pool_create(const prx_pool_limits* lim, prx_engine** out) {
  return guard([&] {
    need(lim != nullptr && out != nullptr, "prx_pool_create: bad arguments");
    need(lim->max_atoms > 0 && lim->max_systems > 0 && lim->max_systems < 65536,
         "prx_pool_create: max_atoms > 0 and 0 < max_systems < 65536 required");
    *out = nullptr;
    auto eng = std::make_unique<prx_engine>();
    eng->pool = true;
    eng->cap_atoms = lim->max_atoms;
    eng->cap_bond = std::max<int64_t>(lim->max_bonded, 1);
    eng->cap_ang = eng->cap_bond;
    eng->cap_tors = eng->cap_bond;
    eng->cap_cmap = std::max<int64_t>(lim->max_bonded / 8, 1);
    eng->cap_exc = std::max<int64_t>(lim->max_bonded / 8, 1);
    eng->cap_coeff = std::max<int64_t>(lim->max_cmap_coeff, 1);
    eng->cap_xt = std::max<int64_t>(lim->max_excl_tiles, 1);
    eng->cap_slots = lim->max_systems;
    eng->cap_rl = (int64_t)prx::kRecBond * eng->cap_bond + (int64_t)prx::kRecAngle * eng->cap_ang +
                  (int64_t)prx::kRecTors * eng->cap_tors + (int64_t)prx::kRecCmap * eng->cap_cmap +
                  (int64_t)prx::kRecExc * eng->cap_exc;
    eng->a_rl = RangeAlloc(eng->cap_rl);
    eng->a_atoms = RangeAlloc(eng->cap_atoms);
    eng->a_bond = RangeAlloc(eng->cap_bond);
    eng->a_ang = RangeAlloc(eng->cap_ang);
    eng->a_tors = RangeAlloc(eng->cap_tors);
    eng->a_cmap = RangeAlloc(eng->cap_cmap);
    eng->a_exc = RangeAlloc(eng->cap_exc);
    eng->a_coeff = RangeAlloc(eng->cap_coeff);
    eng->a_xt = RangeAlloc(eng->cap_xt);
    try {
      scratch_ptr();  // set-up of the allocation-free entry point's scratch array on this device
      eng->alloc_device();
    } catch (...) {
      eng->release_device();
      throw;
    }
    eng->initialized = true;
    *out = eng.release();
  });
}

int prx_engine_destroy(prx_engine* eng) {
  return guard([&] {
    if (!eng) return;
    hipStreamSynchronize(eng->stream);
    eng->release_device();
    delete eng;
  });
}

int prx_engine_layout(const prx_engine* eng, int32_t* offsets, int32_t capacity) {
  return guard([&] {
    need(eng && offsets, "prx_engine_layout: bad arguments");
    need(capacity >= (int32_t)eng->offsets.size(), "prx_engine_layout: capacity too small");
    std::copy(eng->offsets.begin(), eng->offsets.end(), offsets);
  });
}

int prx_energy_forces(prx_engine* eng, const float* positions, float* forces, double* energies,
                      float* born_radii, void* stream) {
  return guard([&] {
    need(eng && positions, "prx_energy_forces: bad arguments");
    eng->rebuild();
    hipStream_t s = eng->pick(stream);
    const int n = eng->total_atoms();
    const prx::DeviceView ev = eng->eval_view();
    prx::launch_copy_in_positions(ev, positions, eng->d_map, n, s);
    eng->evaluate(ev, energies != nullptr, s);
    if (energies)
      prx::launch_energy_reduce(ev, eng->d_ranges, (int)eng->order.size(), energies, s);
    if (forces) prx::launch_copy_out(ev, nullptr, nullptr, forces, eng->d_map, n, s);
    if (born_radii) prx::launch_born_out(ev, born_radii, eng->d_map, n, s);
    HIPCHECK(hipGetLastError());
  });
}

int prx_energy_forces_noalloc(const prx_system_desc* desc, const float* positions, float* forces,
                              double* energies, float* born_radii, void* stream) {
  return guard([&] {
    need(desc && positions, "prx_energy_forces_noalloc: bad arguments");
    const prx_system_desc& d = *desc;
    validate(d);
    need(d.n_atoms <= kScratchMaxAtoms, "prx_energy_forces_noalloc: at most 8192 atoms");
    hipStream_t s = (hipStream_t)stream;
    char* base = scratch_ptr();
    // The previous call on this stream has finished (each call returns after
    // its results are written); its scratch contents are overwritten below.
    HIPCHECK(hipStreamSynchronize(s));
    prx_engine eng;
    eng.arena = base;
    eng.arena_cap = kScratchBytes;
    eng.fork_min_atoms = INT64_MAX;  // one stream: the caller's
    prx::ExclPlan plan;
    prx::build_exclusions(d, plan);
    eng.cap_atoms = d.n_atoms;
    eng.cap_bond = d.n_bonds + d.n_ub;
    eng.cap_ang = d.n_angles;
    eng.cap_tors = count_torsion_entries(d);
    eng.cap_cmap = d.n_cmap;
    eng.cap_exc = d.n_exceptions;
    eng.cap_coeff = d.n_cmap ? (int64_t)d.n_cmap_maps * d.cmap_grid * d.cmap_grid * 16 : 0;
    eng.cap_xt = (int64_t)plan.tiles.size();
    eng.cap_rl = std::max<int64_t>(count_records(d), 1);
    eng.cap_slots = 1;
    int64_t n_tp = 0, n_ent = 0;
    const int64_t n_items = plan_items(d, plan, &n_tp, &n_ent);
    eng.cap_tp = n_tp;
    eng.cap_ent = n_ent;
    eng.a_tp = RangeAlloc(eng.cap_tp);
    eng.a_ent = RangeAlloc(eng.cap_ent);
    eng.a_atoms = RangeAlloc(eng.cap_atoms);
    eng.a_bond = RangeAlloc(eng.cap_bond);
    eng.a_ang = RangeAlloc(eng.cap_ang);
    eng.a_tors = RangeAlloc(eng.cap_tors);
    eng.a_cmap = RangeAlloc(eng.cap_cmap);
    eng.a_exc = RangeAlloc(eng.cap_exc);
    eng.a_coeff = RangeAlloc(eng.cap_coeff);
    eng.a_xt = RangeAlloc(eng.cap_xt);
    eng.a_rl = RangeAlloc(eng.cap_rl);
    const int64_t nT = (d.n_atoms + prx::kTile - 1) / prx::kTile;
    try {
      eng.alloc_device(n_items, nT);  // CapacityError if the system does not fit
      eng.admit(d);
      eng.rebuild();
      const int n = eng.total_atoms();
      const prx::DeviceView ev = eng.eval_view();
      prx::launch_copy_in_positions(ev, positions, eng.d_map, n, s);
      eng.evaluate(ev, energies != nullptr, s);
      if (energies) prx::launch_energy_reduce(ev, eng.d_ranges, 1, energies, s);
      if (forces) prx::launch_copy_out(ev, nullptr, nullptr, forces, eng.d_map, n, s);
      if (born_radii) prx::launch_born_out(ev, born_radii, eng.d_map, n, s);
      HIPCHECK(hipGetLastError());
      HIPCHECK(hipStreamSynchronize(s));
    } catch (...) {
      hipStreamSynchronize(s);
      throw;
    }
  });
}

int prx_integrator_init(prx_engine* eng, const float* positions, const float* momenta,
                        const prx_langevin_params* params, void* stream) {
  return guard([&] {
    need(eng && positions && params, "prx_integrator_init: bad arguments");
    eng->rebuild();
    hipStream_t s = eng->pick(stream);
    const int n = eng->total_atoms();
    // The slot constants are rewritten with synchronous copies: nothing queued
    // on either stream (a step graph, the caller's work) may still read them.
    HIPCHECK(hipStreamSynchronize(eng->stream));
    HIPCHECK(hipStreamSynchronize(s));
    for (size_t k = 0; k < eng->order.size(); ++k) eng->set_params(eng->order[k], params[k]);
    prx::launch_copy_in_positions(eng->v, positions, eng->d_map, n, s);
    if (momenta) {
      prx::launch_copy_in_momenta(eng->v, momenta, eng->d_map, n, s);
      HIPCHECK(hipMemsetAsync(eng->v.atom_step, 0, eng->cap_atoms * sizeof(uint32_t), s));
    } else {
      HIPCHECK(hipMemcpyAsync(eng->d_slot_list, eng->order.data(), eng->order.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, s));
      prx::launch_init_momenta(eng->v, eng->d_slot_list, (int)eng->order.size(), s);
    }
    eng->evaluate(false, s);
    HIPCHECK(hipGetLastError());
    eng->initialized = true;
  });
}

int prx_integrator_steps(prx_engine* eng, int32_t n_steps, void* stream) {
  return guard([&] {
    need(eng != nullptr && n_steps >= 0, "prx_integrator_steps: bad arguments");
    if (!eng->initialized) throw StateError("prx_integrator_steps before prx_integrator_init");
    for (int32_t sid : eng->order)
      if (!eng->slots[sid].has_params) throw StateError("a system has no integrator parameters");
    eng->steps(n_steps, eng->pick(stream));
    HIPCHECK(hipGetLastError());
  });
}

int prx_engine_set_masses(prx_engine* eng, const double* masses) {
  return guard([&] {
    need(eng && masses, "prx_engine_set_masses: bad arguments");
    eng->rebuild();
    HIPCHECK(hipDeviceSynchronize());
    size_t p = 0;
    for (int32_t sid : eng->order) {
      const Slot& s = eng->slots[sid];
      std::vector<float4> mom(s.n);
      std::vector<float> sqm(s.n);
      HIPCHECK(hipMemcpy(mom.data(), eng->v.mom + s.a0, s.n * sizeof(float4), hipMemcpyDeviceToHost));
      for (int i = 0; i < s.n; ++i, ++p) {
        const double m = masses[p];
        need(m > 0.0, "prx_engine_set_masses: masses must be positive");
        mom[i].w = (float)(1.0 / m);
        sqm[i] = (float)std::sqrt(m);
      }
      HIPCHECK(hipMemcpy(eng->v.mom + s.a0, mom.data(), s.n * sizeof(float4), hipMemcpyHostToDevice));
      HIPCHECK(hipMemcpy(eng->v.sqrt_m + s.a0, sqm.data(), s.n * sizeof(float), hipMemcpyHostToDevice));
    }
  });
}

int prx_integrator_state(prx_engine* eng, float* positions, float* momenta, float* forces, void* stream) {
  return guard([&] {
    need(eng != nullptr, "prx_integrator_state: bad arguments");
    eng->flush_pending(eng->pick(stream));
    eng->rebuild();
    prx::launch_copy_out(eng->v, positions, momenta, forces, eng->d_map, eng->total_atoms(), eng->pick(stream));
    HIPCHECK(hipGetLastError());
  });
}

int prx_integrator_kinetic(prx_engine* eng, double* kinetic, void* stream) {
  return guard([&] {
    need(eng && kinetic, "prx_integrator_kinetic: bad arguments");
    eng->flush_pending(eng->pick(stream));
    eng->rebuild();
    prx::launch_kinetic(eng->v, eng->d_ranges, (int)eng->order.size(), kinetic, eng->pick(stream));
    HIPCHECK(hipGetLastError());
  });
}

int prx_recenter_com(prx_engine* eng, void* stream) {
  return guard([&] {
    need(eng, "prx_recenter_com: bad arguments");
    if (!eng->initialized) throw StateError("prx_recenter_com before prx_integrator_init");
    eng->rebuild();
    prx::launch_recenter_com(eng->v, eng->d_ranges, (int)eng->order.size(), eng->pick(stream));
    HIPCHECK(hipGetLastError());
  });
}

int prx_integrator_step_count(prx_engine* eng, int32_t system, int64_t* steps) {
  return guard([&] {
    need(eng && steps && system >= 0 && system < (int32_t)eng->order.size(), "prx_integrator_step_count: bad arguments");
    const Slot& s = eng->slots[eng->order[system]];
    uint32_t c = 0;
    HIPCHECK(hipStreamSynchronize(eng->stream));
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(&c, eng->v.atom_step + s.a0, sizeof(c), hipMemcpyDeviceToHost));
    *steps = c;
  });
}

int prx_integrator_warnings(prx_engine* eng, int32_t system, int64_t* counts) {
  return guard([&] {
    need(eng && counts && system >= 0 && system < (int32_t)eng->order.size(), "prx_integrator_warnings: bad arguments");
    unsigned int w[4] = {0, 0, 0, 0};
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(w, eng->v.warn + (size_t)eng->order[system] * prx::kWarnStride, sizeof(w),
                       hipMemcpyDeviceToHost));
    for (int k = 0; k < 4; ++k) counts[k] = w[k];
  });
}

int prx_integrator_set_step_counts(prx_engine* eng, const int64_t* steps, void* stream) {
  return guard([&] {
    need(eng && steps, "prx_integrator_set_step_counts: bad arguments");
    if (!eng->initialized) throw StateError("prx_integrator_set_step_counts before prx_integrator_init");
    eng->rebuild();
    hipStream_t s = eng->pick(stream);
    for (size_t k = 0; k < eng->order.size(); ++k)
      need(steps[k] >= 0 && steps[k] <= (int64_t)UINT32_MAX,
           "prx_integrator_set_step_counts: step counts must lie in [0, 2^32)");
    for (size_t k = 0; k < eng->order.size(); ++k) {
      const Slot& sl = eng->slots[eng->order[k]];
      prx::launch_fill_u32(eng->v.atom_step + sl.a0, (uint32_t)steps[k], sl.n, s);
    }
    HIPCHECK(hipGetLastError());
  });
}

int prx_integrator_health(prx_engine* eng, int64_t* nonfinite, void* stream) {
  return guard([&] {
    need(eng && nonfinite, "prx_integrator_health: bad arguments");
    hipStream_t s = eng->pick(stream);
    eng->flush_pending(s);
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipStreamSynchronize(eng->stream));
    std::vector<unsigned int> w((size_t)eng->cap_slots * prx::kWarnStride);
    HIPCHECK(hipMemcpy(w.data(), eng->v.warn, w.size() * sizeof(unsigned int), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < eng->order.size(); ++k)
      nonfinite[k] = w[(size_t)eng->order[k] * prx::kWarnStride + prx::kWarnNonFinite];
  });
}

int prx_engine_set_nl_skin(prx_engine* eng, double skin) {
  return guard([&] {
    need(eng && skin >= 0.0 && skin < 1e3, "prx_engine_set_nl_skin: skin must lie in [0, 1000) A");
    HIPCHECK(hipStreamSynchronize(eng->stream));
    eng->nl_skin = skin;
    for (int32_t sid : eng->order) {
      Slot& sl = eng->slots[sid];
      if (!sl.nl) continue;
      const double rl = std::sqrt((double)sl.c.cutoff2) + skin;
      sl.c.nl_cut2 = (float)(rl * rl * 1.00001);
      sl.c.skin_half2 = (float)(0.25 * skin * skin);
      eng->push_const(sid, sl.c);
    }
  });
}

int prx_engine_set_step_mode(prx_engine* eng, int32_t mode) {
  return guard([&] {
    need(eng && mode >= 0 && mode <= 3,
         "prx_engine_set_step_mode: mode is 0 (auto), 1 (graphs), 2 (persistent) or 3 (fused)");
    if (mode == 2) {
      if (eng->persist_blocks < 0) eng->persist_blocks = prx::persist_max_blocks();
      if (eng->persist_blocks <= 0) throw StateError("the persistent step kernel needs cooperative launches");
    }
    HIPCHECK(hipStreamSynchronize(eng->stream));
    eng->step_mode = mode;
  });
}

int prx_engine_set_fork_min_atoms(prx_engine* eng, int64_t min_atoms) {
  return guard([&] {
    need(eng && min_atoms >= 0, "prx_engine_set_fork_min_atoms: bad arguments");
    if (eng->fork_min_atoms == min_atoms) return;
    // captured graphs hold the old shape: wait for them, then drop them
    HIPCHECK(hipStreamSynchronize(eng->stream));
    eng->drop_graphs();
    eng->fork_min_atoms = min_atoms;
  });
}

int prx_pool_admit(prx_engine* pool, const prx_system_desc* desc, const float* positions_host,
                   const prx_langevin_params* params, int32_t* slot, void* stream) {
  return guard([&] {
    need(pool && pool->pool && desc && positions_host && params && slot, "prx_pool_admit: bad arguments");
    // The uploads are asynchronous copies on the pool's stream, after the
    // work already queued there: admission does not wait for the device.
    // Work the caller queued on its own stream is ordered before them.
    hipStream_t s = pool->pick(stream);
    if (s != pool->stream) {
      HIPCHECK(hipEventRecord(pool->ev_order, s));
      HIPCHECK(hipStreamWaitEvent(pool->stream, pool->ev_order, 0));
    }
    const int32_t sid = pool->admit(*desc);
    try {
      pool->set_params(sid, *params);
      const Slot& sl = pool->slots[sid];
      std::vector<float4> p4(sl.n);
      for (int i = 0; i < sl.n; ++i)
        p4[i] = make_float4(positions_host[3 * i], positions_host[3 * i + 1], positions_host[3 * i + 2], 0.f);
      pool->h2d(pool->v.pos + sl.a0, p4.data(), sl.n * sizeof(float4));
      // momenta and forces follow at the next step call or read, once for
      // every system admitted until then
      pool->pending_init.push_back(sid);
    } catch (...) {
      pool->release(sid);
      throw;
    }
    *slot = sid;
  });
}

int prx_pool_set_step_limit(prx_engine* pool, int32_t slot, int64_t limit, void* stream) {
  return guard([&] {
    need(pool && pool->pool && slot >= 0 && slot < (int32_t)pool->slots.size() && pool->slots[slot].live,
         "prx_pool_set_step_limit: no such live slot");
    need(limit >= 0 && limit < (int64_t)UINT32_MAX, "prx_pool_set_step_limit: 0 <= limit < 2^32 - 1 required");
    // kernels in flight read the slot table: the entry is replaced in stream
    // order, after the work queued on the pool's and the caller's streams
    hipStream_t s = pool->pick(stream);
    if (s != pool->stream) {
      HIPCHECK(hipEventRecord(pool->ev_order, s));
      HIPCHECK(hipStreamWaitEvent(pool->stream, pool->ev_order, 0));
    }
    prx::SlotConst& c = pool->slots[slot].c;
    c.step_limit = (uint32_t)limit;
    pool->push_const(slot, c);
  });
}

int prx_pool_release(prx_engine* pool, int32_t slot, void* stream) {
  return guard([&] {
    need(pool && pool->pool, "prx_pool_release: bad arguments");
    // the release's device writes are queued on the pool's stream after the
    // work queued there and on the caller's stream
    hipStream_t s = pool->pick(stream);
    if (s != pool->stream) {
      HIPCHECK(hipEventRecord(pool->ev_order, s));
      HIPCHECK(hipStreamWaitEvent(pool->stream, pool->ev_order, 0));
    }
    pool->release(slot);
  });
}

int prx_pool_read(prx_engine* pool, int32_t slot, float* positions, float* momenta, float* forces,
                  int64_t* steps, void* stream) {
  return guard([&] {
    need(pool && slot >= 0 && slot < (int32_t)pool->slots.size() && pool->slots[slot].live,
         "prx_pool_read: no such live slot");
    hipStream_t s = pool->pick(stream);
    pool->flush_pending(s);
    HIPCHECK(hipStreamSynchronize(s));
    HIPCHECK(hipStreamSynchronize(pool->stream));
    const Slot& sl = pool->slots[slot];
    std::vector<float4> tmp(sl.n);
    auto get = [&](float4* src, float* dst) {
      if (!dst) return;
      HIPCHECK(hipMemcpy(tmp.data(), src + sl.a0, sl.n * sizeof(float4), hipMemcpyDeviceToHost));
      for (int i = 0; i < sl.n; ++i) { dst[3 * i] = tmp[i].x; dst[3 * i + 1] = tmp[i].y; dst[3 * i + 2] = tmp[i].z; }
    };
    get(pool->v.pos, positions);
    get(pool->v.mom, momenta);
    get(pool->v.force, forces);
    if (steps) {
      uint32_t c = 0;
      HIPCHECK(hipMemcpy(&c, pool->v.atom_step + sl.a0, sizeof(c), hipMemcpyDeviceToHost));
      *steps = c;
    }
  });
}

int prx_pool_slots(const prx_engine* pool, int32_t* slots, int32_t capacity, int32_t* n_live) {
  return guard([&] {
    need(pool && n_live, "prx_pool_slots: bad arguments");
    *n_live = (int32_t)pool->order.size();
    if (slots) {
      need(capacity >= (int32_t)pool->order.size(), "prx_pool_slots: capacity too small");
      std::copy(pool->order.begin(), pool->order.end(), slots);
    }
  });
}

int prx_pool_energy(prx_engine* pool, int32_t slot, double* energies_host, void* stream) {
  return guard([&] {
    need(pool && energies_host && slot >= 0 && slot < (int32_t)pool->slots.size() && pool->slots[slot].live,
         "prx_pool_energy: no such live slot");
    hipStream_t s = pool->pick(stream);
    pool->flush_pending(s);
    pool->rebuild();
    HIPCHECK(hipStreamSynchronize(pool->stream));
    // evaluated in the evaluation buffers at a copy of the positions: the
    // integrator's stored forces stay those of its own force-only passes
    const prx::DeviceView ev = pool->eval_view();
    HIPCHECK(hipMemcpyAsync(ev.pos, pool->v.pos, (size_t)pool->v.atom_hw * sizeof(float4), hipMemcpyDeviceToDevice, s));
    pool->evaluate(ev, true, s);
    prx::launch_energy_reduce(ev, pool->d_ranges, (int)pool->order.size(), pool->d_energy_tmp, s);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s));
    const auto it = std::find(pool->order.begin(), pool->order.end(), slot);
    const size_t row = (size_t)(it - pool->order.begin());
    HIPCHECK(hipMemcpy(energies_host, pool->d_energy_tmp + row * PRX_N_ENERGY, PRX_N_ENERGY * sizeof(double),
                       hipMemcpyDeviceToHost));
  });
}

int prx_profile_passes(prx_engine* eng, int32_t n_reps, double* ms) {
  return guard([&] {
    need(eng && ms && n_reps > 0, "prx_profile_passes: bad arguments");
    eng->rebuild();
    HIPCHECK(hipDeviceSynchronize());
    hipStream_t s = eng->stream;
    const int n_pass = 7;
    hipEvent_t ev[n_pass][2];
    for (int k = 0; k < n_pass; ++k) {
      HIPCHECK(hipEventCreate(&ev[k][0]));
      HIPCHECK(hipEventCreate(&ev[k][1]));
      ms[k] = 0.0;
    }
    // Each pass is launched n_reps times back to back between one pair of
    // events (every pass is idempotent on unchanged inputs), so the mean is
    // the kernel's own duration plus the launch gap, comparable with a
    // kernel trace.
    auto timed = [&](int k, auto&& launch) {
      HIPCHECK(hipEventRecord(ev[k][0], s));
      for (int r = 0; r < n_reps; ++r) launch();
      HIPCHECK(hipEventRecord(ev[k][1], s));
      HIPCHECK(hipEventSynchronize(ev[k][1]));
      float t = 0.f;
      HIPCHECK(hipEventElapsedTime(&t, ev[k][0], ev[k][1]));
      ms[k] = t;
    };
    prx::launch_tile_boxes(eng->v, s);
    if (eng->has_nl) prx::launch_nl_build(eng->v, eng->d_ranges, (int)eng->order.size(), true, s);
    timed(0, [&] { prx::launch_born_pairs(eng->v, s); });
    timed(1, [&] { prx::launch_born_gather(eng->v, s); });
    timed(2, [&] { prx::launch_force_pairs(eng->v, false, s); });
    timed(3, [&] { prx::launch_force_gather(eng->v, false, s); });
    timed(4, [&] { prx::launch_chain_pairs(eng->v, s); });
    timed(5, [&] { prx::launch_bonded(eng->v, false, s); });
    timed(6, [&] { prx::launch_chain_gather(eng->v, -1, s); });
    for (int k = 0; k < n_pass; ++k) {
      ms[k] /= n_reps;
      hipEventDestroy(ev[k][0]);
      hipEventDestroy(ev[k][1]);
    }
    HIPCHECK(hipGetLastError());
  });
}

int prx_profile_steps(prx_engine* eng, int32_t n_steps, double* ms) {
  return guard([&] {
    need(eng && ms && n_steps > 0 && n_steps <= 1000, "prx_profile_steps: bad arguments");
    if (!eng->initialized) throw StateError("prx_profile_steps before prx_integrator_init");
    eng->rebuild();
    if (eng->order.empty()) throw std::invalid_argument("prx_profile_steps: no systems");
    HIPCHECK(hipDeviceSynchronize());
    constexpr int n_pass = 7;  // Born pairs, Born gather, force pairs, force gather, chain pairs, bonded, chain gather
    hipStream_t s = eng->stream;
    prx::DeviceView& v = eng->v;
    int dev = 0, khz = 0;
    HIPCHECK(hipGetDevice(&dev));
    HIPCHECK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) throw std::runtime_error("prx_profile_steps: no wall-clock rate");
    const size_t n_ticks = (size_t)n_steps * n_pass * 2;
    unsigned long long* ticks = nullptr;
    hipGraph_t g = nullptr;
    hipGraphExec_t exec = nullptr;
    auto cleanup = [&] {
      if (exec) hipGraphExecDestroy(exec);
      if (g) hipGraphDestroy(g);
      if (ticks) hipFree(ticks);
    };
    try {
      HIPCHECK(hipMalloc(&ticks, n_ticks * sizeof(unsigned long long)));
      auto T = [&](int k, int p, int b) { return ticks + ((size_t)k * n_pass + p) * 2 + b; };
      // The step graph of prx_integrator_steps with a wall-clock tick before
      // and after every pass (bonded and its ticks on the side stream).
      HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
      try {
        prx::launch_step_begin(v, s);
        for (int k = 0; k < n_steps; ++k) {
          auto timed = [&](int p, hipStream_t st, auto&& launch) {
            prx::launch_tick(T(k, p, 0), st);
            launch();
            prx::launch_tick(T(k, p, 1), st);
          };
          const bool fork = eng->fork_bonded();
          if (fork) {
            HIPCHECK(hipEventRecord(eng->ev_fork, s));
            HIPCHECK(hipStreamWaitEvent(eng->side, eng->ev_fork, 0));
            timed(5, eng->side, [&] { prx::launch_bonded(v, false, eng->side); });
            HIPCHECK(hipEventRecord(eng->ev_join, eng->side));
          }
          if (k == 0) prx::launch_tile_boxes(v, s);
          if (eng->has_nl) prx::launch_nl_build(v, eng->d_ranges, (int)eng->order.size(), k == 0, s);
          timed(0, s, [&] { prx::launch_born_pairs(v, s); });
          timed(1, s, [&] { prx::launch_born_gather(v, s); });
          timed(2, s, [&] { prx::launch_force_pairs(v, false, s); });
          timed(3, s, [&] { prx::launch_force_gather(v, false, s); });
          if (!fork) timed(5, s, [&] { prx::launch_bonded(v, false, s); });
          timed(4, s, [&] { prx::launch_chain_pairs(v, s); });
          if (fork) HIPCHECK(hipStreamWaitEvent(s, eng->ev_join, 0));
          timed(6, s, [&] { prx::launch_chain_gather(v, k + 1 == n_steps ? 2 : 1, s); });
        }
        HIPCHECK(hipGetLastError());
      } catch (...) {
        hipGraph_t partial = nullptr;
        if (hipStreamEndCapture(s, &partial) == hipSuccess && partial) hipGraphDestroy(partial);
        throw;
      }
      HIPCHECK(hipStreamEndCapture(s, &g));
      HIPCHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
      HIPCHECK(hipGraphLaunch(exec, s));
      HIPCHECK(hipStreamSynchronize(s));
      std::vector<unsigned long long> h(n_ticks);
      HIPCHECK(hipMemcpy(h.data(), ticks, n_ticks * sizeof(unsigned long long), hipMemcpyDeviceToHost));
      for (int p = 0; p < n_pass; ++p) {
        double sum = 0.0;
        for (int k = 0; k < n_steps; ++k) {
          const size_t i = ((size_t)k * n_pass + p) * 2;
          sum += (double)(h[i + 1] - h[i]) / (double)khz;  // ticks at khz kHz -> ms
        }
        ms[p] = sum / n_steps;
      }
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

int prx_engine_work_items(const prx_engine* eng, int64_t* n_items) {
  return guard([&] {
    need(eng && n_items, "prx_engine_work_items: bad arguments");
    *n_items = eng->v.n_items;
  });
}

int prx_neighbor_list_dense(const float* positions, int32_t n_atoms, double radius, int32_t capacity,
                            int32_t* idx, int32_t* counts, float* workspace, void* stream) {
  return guard([&] {
    need(positions && idx && counts && workspace && n_atoms > 0 && capacity > 0 && radius > 0.0,
         "prx_neighbor_list_dense: bad arguments");
    prx::launch_dense_list(positions, n_atoms, (float)radius, capacity, idx, counts,
                           reinterpret_cast<float4*>(workspace), (hipStream_t)stream);
    HIPCHECK(hipGetLastError());
  });
}

int prx_cmap_coefficients(const double* grid, int32_t G, double* out) {
  return guard([&] {
    need(grid && out && G >= 2, "prx_cmap_coefficients: bad arguments");
    prx::cmap_coefficients(grid, G, out);
  });
}

int prx_plan_tiles(const prx_system_desc* desc, int64_t* n_items, int64_t* n_excl_tiles) {
  return guard([&] {
    need(desc && n_items && n_excl_tiles, "prx_plan_tiles: bad arguments");
    validate(*desc);
    prx::ExclPlan plan;
    prx::build_exclusions(*desc, plan);
    int64_t n_tp = 0, n_ent = 0;
    *n_items = plan_items(*desc, plan, &n_tp, &n_ent);
    *n_excl_tiles = (int64_t)plan.tiles.size();
  });
}

}  // extern "C"
