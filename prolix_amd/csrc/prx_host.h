// Host-side planning shared by the engine and the host-only ABI entry points.
#pragma once

#include <utility>
#include <vector>

#include "prx_common.h"

namespace prx {

struct ExclPlan {
  std::vector<std::pair<double, double>> classes;  // class c + 1 -> (vdw, elec) scale
  std::vector<std::pair<int, int>> tiles;          // (I, J) tile pairs holding exclusions
  std::vector<unsigned long long> tile_masks;      // 128 words per tile
};

// (G, G, 16) coefficients indexed [phi][psi] from a grid stored [psi][phi].
void cmap_coefficients(const double* grid_psi_phi, int G, double* out);
// Exclusion classes and 64x64 tile masks of one system (throws std::invalid_argument).
void build_exclusions(const prx_system_desc& d, ExclPlan& plan);
// Tile pairs I <= J of one system placed at arena offset a0.
void build_items(int n_atoms, int a0, int slot, const ExclPlan& plan, int excl_base,
                 std::vector<PairItem>& out);

// Work layout of a neighbour-list system (relative to the system: item
// indices from its first item, entries from its entry base, tiles local).
struct NlPlan {
  std::vector<PairItem> items;        // regular items (diagonal and exclusion tiles, row-major), then chunks
  std::vector<int4> rows;             // per tile: (first chunk item, entry base, first, end regular row item)
  std::vector<std::vector<int32_t>> cols;  // per tile: its regular column items, in increasing row order
  std::vector<TilePairInfo> tps;      // per tile pair I < J, row-major (slot filled in by the engine)
  int64_t n_entries = 0;              // list capacity: every atom of the exclusion-free J > I, per row I
};
// Tile-pair count and list capacity of a neighbour-list system, without building it.
void nl_sizes(int n_atoms, const ExclPlan& plan, int64_t* n_tp, int64_t* n_entries, int64_t* n_items);
// Items for arena offset a0 and entry base ent_b: chunk items carry
// j0 = -1 - (ent_b + first entry) and excl = local row tile (the engine adds
// the system's first global tile).
void build_items_nl(int n_atoms, int a0, int slot, const ExclPlan& plan, int excl_base, int64_t ent_b,
                    NlPlan& out);

}  // namespace prx
