// Bonded terms: one thread per arena entry. Each entry writes the force on
// each of its atoms to its own float4 record (no atomics); every atom later
// sums its records in a fixed order through a per-atom list (see
// gather_bonded in prx_atom.hip), so the result is deterministic and free of
// same-address atomic contention between neighbouring terms.
//
//   bonds / Urey-Bradley  1/2 k (r - r0)^2             bonded.py:20-44
//   angles                1/2 k (theta - theta0)^2     bonded.py:47-81
//   torsion terms         k (1 + cos(n phi - phase))   bonded.py:84-141
//   harmonic impropers    k wrap(phi - phi0)^2         bonded.py:144-185
//   CMAP bicubic patches                               cmap.py:316-389
//   1-4 exceptions        LJ 12-6 + Coulomb            bonded.py:188-228
//
// The grid is split into five consecutive block ranges, one per entry kind.
#include "prx_bonded_dev.h"

namespace prx {
template <bool ENERGY>
__global__ __launch_bounds__(kBondedThreads) void k_bonded(DeviceView v, int nb_bond, int nb_angle,
                                                           int nb_tors, int nb_cmap) {
  int b = blockIdx.x;
  const int t = threadIdx.x;
  if (b < nb_bond) {
    const int i = b * kBondedThreads + t;
    if (i < v.bond_hw) do_bond<ENERGY>(v, i);
    return;
  }
  b -= nb_bond;
  if (b < nb_angle) {
    const int i = b * kBondedThreads + t;
    if (i < v.angle_hw) do_angle<ENERGY>(v, i);
    return;
  }
  b -= nb_angle;
  if (b < nb_tors) {
    const int i = b * kBondedThreads + t;
    if (i < v.tors_hw) do_torsion<ENERGY>(v, i);
    return;
  }
  b -= nb_tors;
  if (b < nb_cmap) {
    const int i = b * kBondedThreads + t;
    if (i < v.cmap_hw) do_cmap<ENERGY>(v, i);
    return;
  }
  b -= nb_cmap;
  const int i = b * kBondedThreads + t;
  if (i < v.exc_hw) do_exception<ENERGY>(v, i);
}

void launch_bonded(const DeviceView& v, bool energy, hipStream_t s) {
  auto nb = [](int n) { return (n + kBondedThreads - 1) / kBondedThreads; };
  const int b0 = nb(v.bond_hw), b1 = nb(v.angle_hw), b2 = nb(v.tors_hw), b3 = nb(v.cmap_hw), b4 = nb(v.exc_hw);
  const int total = b0 + b1 + b2 + b3 + b4;
  if (total == 0) return;
  if (energy)
    hipLaunchKernelGGL(k_bonded<true>, dim3(total), dim3(kBondedThreads), 0, s, v, b0, b1, b2, b3);
  else
    hipLaunchKernelGGL(k_bonded<false>, dim3(total), dim3(kBondedThreads), 0, s, v, b0, b1, b2, b3);
}

}  // namespace prx
