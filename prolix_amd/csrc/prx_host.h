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

}  // namespace prx
