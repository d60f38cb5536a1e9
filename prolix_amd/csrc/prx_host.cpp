// Host-side planning that needs no GPU: CMAP patch coefficients, exclusion
// classes and tile masks, and the pair work list of one system.
#include "prx_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <unordered_map>
#include <stdexcept>

namespace prx {

// Solve A x = b (n x n, row-major) by Gaussian elimination with partial pivoting.
static void dense_solve(std::vector<double>& A, std::vector<double>& b, int n) {
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::fabs(A[r * n + c]) > std::fabs(A[piv * n + c])) piv = r;
    if (A[piv * n + c] == 0.0) throw std::runtime_error("singular system in CMAP spline");
    if (piv != c) {
      for (int k = 0; k < n; ++k) std::swap(A[c * n + k], A[piv * n + k]);
      std::swap(b[c], b[piv]);
    }
    for (int r = c + 1; r < n; ++r) {
      const double f = A[r * n + c] / A[c * n + c];
      if (f == 0.0) continue;
      for (int k = c; k < n; ++k) A[r * n + k] -= f * A[c * n + k];
      b[r] -= f * b[c];
    }
  }
  for (int r = n - 1; r >= 0; --r) {
    double s = b[r];
    for (int k = r + 1; k < n; ++k) s -= A[r * n + k] * b[k];
    b[r] = s / A[r * n + r];
  }
}

// Inverse of the 16 x 16 bicubic constraint matrix: rows are the values,
// d/da, d/db and d2/dadb at corners (0,0), (1,0), (1,1), (0,1); columns the
// coefficients c[4p + q] of sum c a^p b^q.
static const std::vector<double>& bicubic_inverse() {
  static std::vector<double> inv;
  if (!inv.empty()) return inv;
  const int corners[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
  std::vector<double> M(256, 0.0);
  for (int kind = 0; kind < 4; ++kind)
    for (int ci = 0; ci < 4; ++ci) {
      const double a = corners[ci][0], b = corners[ci][1];
      for (int p = 0; p < 4; ++p)
        for (int q = 0; q < 4; ++q) {
          double v = 0.0;
          const double ap = std::pow(a, p), bq = std::pow(b, q);
          const double apm = p > 0 ? p * std::pow(a, p - 1) : 0.0;
          const double bqm = q > 0 ? q * std::pow(b, q - 1) : 0.0;
          if (kind == 0) v = ap * bq;
          else if (kind == 1) v = apm * bq;
          else if (kind == 2) v = ap * bqm;
          else v = apm * bqm;
          M[(kind * 4 + ci) * 16 + 4 * p + q] = v;
        }
    }
  inv.assign(256, 0.0);
  for (int col = 0; col < 16; ++col) {
    std::vector<double> A = M, e(16, 0.0);
    e[col] = 1.0;
    dense_solve(A, e, 16);
    for (int r = 0; r < 16; ++r) inv[r * 16 + col] = e[r];
  }
  return inv;
}

// Periodic cubic spline derivatives at the nodes x_k = 2 pi k / G of samples y
// (cmap.py:62-163): the cyclic system of cmap.py:84-121, evaluated at t = x_i.
static void spline_node_derivs(const double* y, int G, double* dy) {
  const int n = G;
  std::vector<double> x(n + 1), h(n), yy(n + 1);
  for (int k = 0; k <= n; ++k) x[k] = 2.0 * M_PI * k / n;
  for (int k = 0; k < n; ++k) yy[k] = y[k];
  yy[n] = y[0];
  for (int k = 0; k < n; ++k) h[k] = x[k + 1] - x[k];
  std::vector<double> slope(n), b(n), A((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) slope[k] = (yy[k + 1] - yy[k]) / h[k];
  b[0] = 6.0 * (slope[0] - slope[n - 1]);
  for (int k = 1; k < n; ++k) b[k] = 6.0 * (slope[k] - slope[k - 1]);
  for (int k = 0; k < n; ++k) A[(size_t)k * n + k] = 2.0 * (h[(k - 1 + n) % n] + h[k]);
  for (int k = 1; k < n; ++k) A[(size_t)k * n + k - 1] = h[k];          // lower = h[1:]
  for (int k = 0; k + 1 < n; ++k) A[(size_t)k * n + k + 1] = h[(k + 1) % n];  // upper = roll(h, -1)[:-1]
  A[0 * n + (n - 1)] = h[n - 1];
  A[(size_t)(n - 1) * n + 0] = h[n - 1];
  dense_solve(A, b, n);
  std::vector<double> m(n + 1);
  for (int k = 0; k < n; ++k) m[k] = b[k];
  m[n] = b[0];
  for (int i = 0; i < n; ++i) {
    const double hi = x[i + 1] - x[i];
    dy[i] = (yy[i + 1] - yy[i]) / hi + hi / 6.0 * (-2.0 * m[i] - m[i + 1]);
  }
}

void cmap_coefficients(const double* grid_psi_phi, int G, double* out) {
  // E[phi][psi] = grid[psi][phi] (cmap.py:259)
  std::vector<double> E((size_t)G * G), d1(E.size()), d2(E.size()), d12(E.size());
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < G; ++j) E[(size_t)i * G + j] = grid_psi_phi[(size_t)j * G + i];
  std::vector<double> col(G), dcol(G);
  // d1: along phi (axis 0) for each psi column
  for (int j = 0; j < G; ++j) {
    for (int i = 0; i < G; ++i) col[i] = E[(size_t)i * G + j];
    spline_node_derivs(col.data(), G, dcol.data());
    for (int i = 0; i < G; ++i) d1[(size_t)i * G + j] = dcol[i];
  }
  // d2: along psi (axis 1) for each phi row
  for (int i = 0; i < G; ++i) {
    spline_node_derivs(&E[(size_t)i * G], G, dcol.data());
    for (int j = 0; j < G; ++j) d2[(size_t)i * G + j] = dcol[j];
  }
  // d12: along phi of d2 (cmap.py:204-214)
  for (int j = 0; j < G; ++j) {
    for (int i = 0; i < G; ++i) col[i] = d2[(size_t)i * G + j];
    spline_node_derivs(col.data(), G, dcol.data());
    for (int i = 0; i < G; ++i) d12[(size_t)i * G + j] = dcol[i];
  }
  const std::vector<double>& W = bicubic_inverse();
  const double delta = 2.0 * M_PI / G;
  for (int i = 0; i < G; ++i)
    for (int j = 0; j < G; ++j) {
      const int ni = (i + 1) % G, nj = (j + 1) % G;
      const int ca[4] = {i, ni, ni, i}, cb[4] = {j, j, nj, nj};
      double rhs[16];
      for (int c = 0; c < 4; ++c) {
        const size_t k = (size_t)ca[c] * G + cb[c];
        rhs[c] = E[k];
        rhs[4 + c] = d1[k] * delta;
        rhs[8 + c] = d2[k] * delta;
        rhs[12 + c] = d12[k] * delta * delta;
      }
      double* o = out + ((size_t)i * G + j) * 16;
      for (int r = 0; r < 16; ++r) {
        double s = 0.0;
        for (int c = 0; c < 16; ++c) s += W[r * 16 + c] * rhs[c];
        o[r] = s;
      }
    }
}

void build_exclusions(const prx_system_desc& d, ExclPlan& plan) {
  plan.classes.clear();
  plan.tiles.clear();
  plan.tile_masks.clear();
  const int n = d.n_atoms, C = d.excl_cols;
  if (C <= 0 || d.excl_idx == nullptr) return;
  // S(i, j) from row i: sum of matching entries' scales (optimization.py:55-57).
  // Row i's entries sorted by j (a stable insertion sort of at most C, so a
  // j's entries are summed in column order), duplicates summed, in CSR form.
  struct Ent {
    int j;
    double v, e;
  };
  std::vector<Ent> ent;
  ent.reserve((size_t)n * std::min(C, 16));
  std::vector<int64_t> off((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) {
    const size_t r0 = ent.size();
    for (int k = 0; k < C; ++k) {
      const int j = d.excl_idx[(size_t)i * C + k];
      if (j < 0) continue;
      if (j >= n) throw std::invalid_argument("exclusion index out of range");
      if (j == i) continue;
      Ent x{j, d.excl_scale_vdw ? d.excl_scale_vdw[(size_t)i * C + k] : 0.0,
            d.excl_scale_elec ? d.excl_scale_elec[(size_t)i * C + k] : 0.0};
      size_t p = ent.size();
      ent.push_back(x);
      while (p > r0 && ent[p - 1].j > x.j) {
        ent[p] = ent[p - 1];
        --p;
      }
      ent[p] = x;
    }
    size_t m = r0;
    for (size_t r = r0; r < ent.size(); ++r) {
      if (m > r0 && ent[m - 1].j == ent[r].j) {
        ent[m - 1].v += ent[r].v;
        ent[m - 1].e += ent[r].e;
      } else {
        ent[m++] = ent[r];
      }
    }
    ent.resize(m);
    off[(size_t)i + 1] = (int64_t)m;
  }
  auto find = [&](int i, int j) -> const Ent* {
    for (int64_t r = off[i]; r < off[(size_t)i + 1]; ++r)
      if (ent[r].j >= j) return ent[r].j == j ? &ent[r] : nullptr;
    return nullptr;
  };
  // Symmetrise: the unordered pair (a < b) carries (S(a,b) + S(b,a)) / 2 of
  // each scale, the reference's 1/2 sum over both ordered pairs; a missing
  // ordered pair counts as scale 1. Each unordered pair is formed once, from
  // its (a, b) entry or, when that is absent, from its (b, a) entry.
  struct Pair {
    uint64_t key;  // a << 32 | b
    double v, e;
  };
  std::vector<Pair> pairs;
  pairs.reserve(ent.size());
  for (int i = 0; i < n; ++i)
    for (int64_t r = off[i]; r < off[(size_t)i + 1]; ++r) {
      const int j = ent[r].j;
      const int a = std::min(i, j), b = std::max(i, j);
      const Ent* e1 = find(a, b);
      if (i > j && e1) continue;  // formed from (a, b)
      const Ent* e2 = find(b, a);
      const double v1 = e1 ? e1->v : 1.0, s1 = e1 ? e1->e : 1.0;
      const double v2 = e2 ? e2->v : 1.0, s2 = e2 ? e2->e : 1.0;
      pairs.push_back({((uint64_t)a << 32) | (uint32_t)b, 0.5 * (v1 + v2), 0.5 * (s1 + s2)});
    }
  // classes and tiles numbered in increasing (a, b) order
  std::sort(pairs.begin(), pairs.end(), [](const Pair& a, const Pair& b) { return a.key < b.key; });
  const int64_t nT = (n + kTile - 1) / kTile;
  std::unordered_map<int64_t, int> tile_index;
  for (const Pair& x : pairs) {
    const double sv = x.v, se = x.e;
    if (sv == 1.0 && se == 1.0) continue;
    int cls = -1;
    for (size_t c = 0; c < plan.classes.size(); ++c)
      if ((float)plan.classes[c].first == (float)sv && (float)plan.classes[c].second == (float)se) cls = (int)c;
    if (cls < 0) {
      if ((int)plan.classes.size() >= kMaxExclClasses)
        throw std::invalid_argument("more than 3 distinct non-unit exclusion scale pairs in one system");
      plan.classes.push_back({sv, se});
      cls = (int)plan.classes.size() - 1;
    }
    const int code = cls + 1;  // 1..3
    const int i = (int)(x.key >> 32), j = (int)(uint32_t)x.key;
    const int I = i / kTile, J = j / kTile;
    auto ins = tile_index.emplace((int64_t)I * nT + J, (int)plan.tiles.size());
    if (ins.second) {
      plan.tiles.push_back({I, J});
      plan.tile_masks.resize(plan.tile_masks.size() + 128, 0ull);
    }
    unsigned long long* mk = &plan.tile_masks[(size_t)ins.first->second * 128];
    auto setbit = [&](int lane, int bit) {
      if (code & 1) mk[lane * 2 + 0] |= (1ull << bit);
      if (code & 2) mk[lane * 2 + 1] |= (1ull << bit);
    };
    setbit(i - I * kTile, j - J * kTile);
    if (I == J) setbit(j - J * kTile, i - I * kTile);
  }
}

void build_items(int n_atoms, int a0, int slot, const ExclPlan& plan, int excl_base,
                 std::vector<PairItem>& out) {
  const int nT = (n_atoms + kTile - 1) / kTile;
  std::map<std::pair<int, int>, int> tix;
  for (size_t t = 0; t < plan.tiles.size(); ++t) tix[plan.tiles[t]] = (int)t;
  for (int I = 0; I < nT; ++I)
    for (int J = I; J < nT; ++J) {
      const int ni = std::min(kTile, n_atoms - I * kTile);
      const int nj = std::min(kTile, n_atoms - J * kTile);
      int flags = (I == J) ? kFlagDiag : 0;
      int ex = -1;
      auto it = tix.find({I, J});
      if (it != tix.end()) {
        flags |= kFlagExcl;
        ex = excl_base + it->second;
      }
      PairItem p;
      p.i0 = a0 + I * kTile;
      p.j0 = a0 + J * kTile;
      p.meta = item_meta(ni, nj, flags, slot);
      p.excl = ex;
      out.push_back(p);
    }
}

namespace {
// (I, J) -> exclusion tile index of the plan, for I <= J.
std::map<std::pair<int, int>, int> excl_tile_index(const ExclPlan& plan) {
  std::map<std::pair<int, int>, int> tix;
  for (size_t t = 0; t < plan.tiles.size(); ++t) tix[plan.tiles[t]] = (int)t;
  return tix;
}
}  // namespace

void nl_sizes(int n_atoms, const ExclPlan& plan, int64_t* n_tp, int64_t* n_entries, int64_t* n_items) {
  const int nT = (n_atoms + kTile - 1) / kTile;
  const auto tix = excl_tile_index(plan);
  int64_t ent = 0, items = nT, chunks = 0;
  for (int I = 0; I < nT; ++I) {
    int64_t cap = 0;
    for (int J = I + 1; J < nT; ++J) {
      if (tix.count({I, J})) {
        ++items;
        continue;
      }
      cap += std::min(kTile, n_atoms - J * kTile);
    }
    ent += cap;
    chunks += (cap + kTile - 1) / kTile;
  }
  *n_tp = (int64_t)nT * (nT - 1) / 2;
  *n_entries = ent;
  *n_items = items + chunks;
}

void build_items_nl(int n_atoms, int a0, int slot, const ExclPlan& plan, int excl_base, int64_t ent_b,
                    NlPlan& out) {
  const int nT = (n_atoms + kTile - 1) / kTile;
  const auto tix = excl_tile_index(plan);
  out.items.clear();
  out.rows.assign(nT, make_int4(0, 0, 0, 0));
  out.cols.assign(nT, {});
  out.tps.clear();
  auto n_of = [&](int T) { return std::min(kTile, n_atoms - T * kTile); };
  // regular items, row-major: the diagonal tile and the tile pairs with exclusions
  for (int I = 0; I < nT; ++I) {
    out.rows[I].z = (int)out.items.size();
    for (int J = I; J < nT; ++J) {
      auto it = tix.find({I, J});
      if (J != I && it == tix.end()) continue;
      int flags = (I == J) ? kFlagDiag : 0;
      int ex = -1;
      if (it != tix.end()) {
        flags |= kFlagExcl;
        ex = excl_base + it->second;
      }
      PairItem p;
      p.i0 = a0 + I * kTile;
      p.j0 = a0 + J * kTile;
      p.meta = item_meta(n_of(I), n_of(J), flags, slot);
      p.excl = ex;
      out.cols[J].push_back((int32_t)out.items.size());
      out.items.push_back(p);
    }
    out.rows[I].w = (int)out.items.size();
  }
  // chunk items of each row's list, and the tile-pair records
  int64_t ent = 0;
  for (int I = 0; I < nT; ++I) {
    int64_t cap = 0;
    for (int J = I + 1; J < nT; ++J) {
      const bool ex = tix.count({I, J}) != 0;
      TilePairInfo tp;
      tp.ij = I | (J << 16);
      tp.slot = ex ? ~slot : slot;
      out.tps.push_back(tp);
      if (!ex) cap += n_of(J);
    }
    out.rows[I].x = (int)out.items.size();
    out.rows[I].y = (int)ent;
    for (int64_t c = 0; c < cap; c += kTile) {
      PairItem p;
      p.i0 = a0 + I * kTile;
      p.j0 = (int32_t)(-1 - (ent_b + ent + c));
      p.meta = item_meta(n_of(I), kTile, 0, slot);
      p.excl = I;
      out.items.push_back(p);
    }
    ent += cap;
  }
  out.n_entries = ent;
}

}  // namespace prx
