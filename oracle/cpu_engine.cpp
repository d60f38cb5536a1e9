// CPU restatement (C++/OpenMP) of prolix's GB-OBC2 energy, forces and BAOAB step.
//
// TEST INFRASTRUCTURE ONLY: the large-N checker for the HIP path and the timed
// CPU baseline of the benchmark. The product never links or calls it.
//
// It is the float64/float32 twin of oracle/gb_oracle.py (same formulas, same
// citations): dense O(N^2) Born radii (generalized_born.py:32-94, 211-263), GB
// over all ordered pairs incl. i == j (:97-208), ACE (:600-648), Coulomb and
// LJ with (N, C) exclusion rows (optimization.py:21-297), bonds, angles,
// torsions, harmonic impropers, 1-4 exceptions (bonded.py:20-228), CMAP
// (cmap.py:316-389, coefficients supplied by the caller) and the BAOAB step
// of settle.py:2064-2087 with Philox4x32-10 normals.
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <set>
#include <vector>

#include "prolix_gb.h"

namespace {

struct Sys {
  prx_system_desc d;
  // owned copies
  std::vector<double> q, sig, eps, R, sr, m;
  std::vector<int32_t> bond_idx, ub_idx, angle_idx, dih_idx, imp_idx, cmap_t, cmap_map, exc_pairs;
  std::vector<double> bond_p, ub_p, angle_p, dih_p, imp_p, exc_s, exc_e, exc_q, coeff;
  std::vector<std::vector<std::pair<int, std::pair<double, double>>>> excl;  // per atom (j, (sv, se))
  int G = 0;
  // 1: a pair is inside the cutoff iff the float32 test of the HIP kernels says
  // so, fmaf(dx, dx, fmaf(dy, dy, dz * dz)) < (float)(cutoff^2) with float
  // coordinate differences, so that pairs within float rounding of the cutoff
  // fall on the same side as on the device; 0: the float64 test r^2 < cutoff^2
  int f32_cut = 0;
};

template <class T>
struct Vec3 {
  T x, y, z;
};

template <class T>
inline T dot(const T* a, const T* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <class T>
inline void cross(const T* a, const T* b, T* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

template <class T>
T pair_integral(T r, T rho, T s) {
  const T D = std::fabs(r - s);
  const T L = std::max(std::max(rho, D), (T)1e-4);
  const T U = std::max(r + s, (T)1e-4);
  const T rs = std::max(r, (T)1e-4);
  const T iL = (T)1 / L, iU = (T)1 / U;
  const T v = (T)0.5 * (iL - iU) + (T)0.25 * std::log(L / U) / rs + (T)0.125 * (r - s * s / rs) * (iU * iU - iL * iL);
  return (r + s > rho) ? v : (T)0;
}

template <class T>
T pair_integral_dr(T r, T rho, T s) {
  const T dd = r - s, D = std::fabs(dd);
  const bool useD = D > rho;
  const T Lr = useD ? D : rho;
  T Lp = useD ? (dd >= 0 ? (T)1 : (T)-1) : (T)0;
  const T L = std::max(Lr, (T)1e-4);
  if (!(Lr > (T)1e-4)) Lp = 0;
  const T Ur = r + s, U = std::max(Ur, (T)1e-4);
  const T Up = Ur > (T)1e-4 ? (T)1 : (T)0;
  const T rs = std::max(r, (T)1e-4), rp = r > (T)1e-4 ? (T)1 : (T)0;
  const T iL = (T)1 / L, iU = (T)1 / U, lg = std::log(L / U);
  const T d1 = (T)0.5 * (-Lp * iL * iL + Up * iU * iU);
  const T d2 = (T)0.25 * ((Lp * iL - Up * iU) / rs - lg * rp / (rs * rs));
  const T d3 = (T)0.125 * ((T)1 + s * s * rp / (rs * rs)) * (iU * iU - iL * iL);
  const T d4 = (T)0.125 * (r - s * s / rs) * ((T)-2 * Up * iU * iU * iU + (T)2 * Lp * iL * iL * iL);
  return (r + s > rho) ? d1 + d2 + d3 + d4 : (T)0;
}

template <class T>
T dihedral_grad(const T* p0, const T* p1, const T* p2, const T* p3, T g[4][3]) {
  T b0[3], b1[3], b2[3];
  for (int k = 0; k < 3; ++k) { b0[k] = p0[k] - p1[k]; b1[k] = p2[k] - p1[k]; b2[k] = p3[k] - p2[k]; }
  const T b1n = std::sqrt(dot(b1, b1));
  T u[3], vv[3], ww[3], cu[3];
  for (int k = 0; k < 3; ++k) u[k] = b1[k] / (b1n + (T)1e-8);
  const T bu0 = dot(b0, u), bu2 = dot(b2, u);
  for (int k = 0; k < 3; ++k) { vv[k] = b0[k] - bu0 * u[k]; ww[k] = b2[k] - bu2 * u[k]; }
  cross(u, vv, cu);
  const T phi = std::atan2(dot(cu, ww), dot(vv, ww));
  T F[3], Gv[3], H[3], A[3], B[3];
  for (int k = 0; k < 3; ++k) { F[k] = b0[k]; Gv[k] = -b1[k]; H[k] = b2[k]; }
  cross(F, Gv, A);
  cross(H, Gv, B);
  const T A2 = dot(A, A), B2 = dot(B, B), FG = dot(F, Gv), HG = dot(H, Gv);
  const T ga = b1n / A2, gb = b1n / B2, fa = FG / (A2 * b1n), hb = HG / (B2 * b1n);
  for (int k = 0; k < 3; ++k) {
    g[0][k] = -ga * A[k];
    g[3][k] = gb * B[k];
    g[1][k] = (ga + fa) * A[k] - hb * B[k];
    g[2][k] = -fa * A[k] + (hb - gb) * B[k];
  }
  return phi;
}

template <class T>
void energy_forces(const Sys& S, const T* x, T* f, double* E) {
  const prx_system_desc& d = S.d;
  const int n = d.n_atoms;
  const int nth = omp_get_max_threads();
  std::vector<double> terms(PRX_N_ENERGY, 0.0);
  std::fill(f, f + 3 * n, (T)0);
  const T eps12 = (T)1e-12;
  const T cut2 = d.cutoff > 0 ? (T)(d.cutoff * d.cutoff) : (T)0;
  // neighbour-list GB (generalized_born.py:266-415): pairs closer than the cutoff
  const bool nl = d.gb_cutoff && d.cutoff > 0;
  const float cut2f = (float)(d.cutoff * d.cutoff);
  auto inside = [&](int i, int j, T r2) {
    if (!S.f32_cut) return r2 < cut2;
    const float dx = (float)x[3 * j] - (float)x[3 * i], dy = (float)x[3 * j + 1] - (float)x[3 * i + 1],
                dz = (float)x[3 * j + 2] - (float)x[3 * i + 2];
    return std::fmaf(dx, dx, std::fmaf(dy, dy, dz * dz)) < cut2f;
  };
  auto gb_in = [&](int i, int j, T r2) { return !nl || inside(i, j, r2); };
  auto rho_of = [&](int i) {
    const double r = S.R[i] - d.dielectric_offset;
    return (T)(nl ? std::max(std::max(S.R[i], 1e-6) - d.dielectric_offset, 1e-4) : r);
  };
  // ---- Born sums
  std::vector<T> I(n, 0), B(n), dBdI(n), c(n);
#pragma omp parallel for schedule(dynamic, 8)
  for (int i = 0; i < n; ++i) {
    const T rho = rho_of(i);
    T acc = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      T dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
      const T r2p = dot(dx, dx);
      if (!gb_in(i, j, r2p)) continue;
      const T r = std::sqrt(r2p + eps12);
      acc += pair_integral(r, rho, (T)S.sr[j]);
    }
    I[i] = acc;
  }
  const double pref = d.implicit_solvent
                          ? -0.5 * d.gb_coulomb_constant * (1.0 / d.solute_dielectric - 1.0 / d.solvent_dielectric)
                          : 0.0;
  std::vector<double> dEdB(n, 0.0);
  double e_self = 0, e_ace = 0;
  for (int i = 0; i < n; ++i) {
    const T rho = rho_of(i), Ri = (T)(nl ? std::max(S.R[i], 1e-4) : S.R[i]);
    const T psi = rho * I[i];
    const T t = psi - (T)0.8 * psi * psi + (T)4.85 * psi * psi * psi;
    const T th = std::tanh(t);
    T inv_b = (T)1 / rho - th / Ri;
    const bool clamped = nl && inv_b < (T)1e-4;
    if (clamped) inv_b = (T)1e-4;
    B[i] = (T)1 / inv_b;
    dBdI[i] = clamped ? (T)0 : B[i] * B[i] * ((T)1 - th * th) * ((T)1 - (T)1.6 * psi + (T)14.55 * psi * psi) * rho / Ri;
    const double qi = S.q[i];
    if (nl) {  // self term q^2 / B (generalized_born.py:412)
      e_self += pref * qi * qi / (double)B[i];
      dEdB[i] += -pref * qi * qi / ((double)B[i] * (double)B[i]);
    } else {
      const T e1 = std::exp(-eps12 / ((T)4 * B[i] * B[i]));
      const T fii = std::sqrt(eps12 + B[i] * B[i] * e1);
      e_self += pref * qi * qi / (double)fii;
      dEdB[i] += -pref * qi * qi / (2.0 * (double)fii * fii * fii) * (double)(e1 * ((T)2 * B[i] + eps12 / ((T)2 * B[i])));
    }
    const double rn = (S.R[i] - d.dielectric_offset) / 10.0 + 0.009;
    const double ace = 0.239006 * 28.3919551 * (rn + 0.14) * (rn + 0.14) * std::pow(rn / ((double)B[i] / 10.0), 6);
    if (d.implicit_solvent) {
      e_ace += ace;
      dEdB[i] += -6.0 * ace / (double)B[i];
    }
  }
  // ---- pair pass: GB direct + Coulomb + LJ
  std::vector<T> fth((size_t)nth * 3 * n, (T)0);
  std::vector<double> dth((size_t)nth * n, 0.0);
  std::vector<double> e_gb(nth, 0), e_cl(nth, 0), e_lj(nth, 0);
#pragma omp parallel
  {
    const int tid = omp_get_thread_num();
    T* ft = &fth[(size_t)tid * 3 * n];
    double* dt = &dth[(size_t)tid * n];
    std::vector<double> sv(n, 1.0), se(n, 1.0);
#pragma omp for schedule(dynamic, 8)
    for (int i = 0; i < n; ++i) {
      for (const auto& e : S.excl[i]) { sv[e.first] = e.second.first; se[e.first] = e.second.second; }
      for (int j = i + 1; j < n; ++j) {
        T dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
        const T r2 = dot(dx, dx);
        T fr = 0;  // -dE/dr / r
        if (d.implicit_solvent && gb_in(i, j, r2)) {
          const T r2g = r2 + eps12;
          const T bb = B[i] * B[j];
          const T ex = std::exp(-r2g / ((T)4 * bb));
          const T fg = std::sqrt(r2g + bb * ex);
          const T qq = (T)(S.q[i] * S.q[j]);
          const T eg = (T)(2.0 * pref) * qq / fg;
          e_gb[tid] += (double)eg;
          const T t3 = eg / (fg * fg);
          fr += t3 * ((T)1 - ex / (T)4);
          const T g = (T)-0.5 * t3 * ex * ((T)1 + r2g / ((T)4 * bb));
          dt[i] += (double)(g * B[j]);
          dt[j] += (double)(g * B[i]);
        }
        const T r = std::sqrt(r2);
        const bool in = r > 0 && (cut2 == (T)0 || inside(i, j, r2));
        if (in) {
          const T scv = (T)sv[j], sce = (T)se[j];
          const T qq = (T)(S.q[i] * S.q[j]);
          const T ec = (T)d.coulomb_constant * qq * sce / r;
          const T s = (T)(0.5 * (std::max(S.sig[i], 1e-6) + std::max(S.sig[j], 1e-6)));
          const T ep = (T)std::sqrt(S.eps[i] * S.eps[j]);
          const T s2 = s * s / r2, s6 = s2 * s2 * s2;
          const T el = (T)4 * ep * (s6 * s6 - s6) * scv;
          e_cl[tid] += (double)ec;
          e_lj[tid] += (double)el;
          fr += (ec + (T)4 * ep * ((T)12 * s6 * s6 - (T)6 * s6) * scv) / r2;
        }
        for (int k = 0; k < 3; ++k) { ft[3 * i + k] += fr * dx[k]; ft[3 * j + k] -= fr * dx[k]; }
      }
      for (const auto& e : S.excl[i]) { sv[e.first] = 1.0; se[e.first] = 1.0; }
    }
  }
  for (int t = 0; t < nth; ++t) {
    for (int k = 0; k < 3 * n; ++k) f[k] += fth[(size_t)t * 3 * n + k];
    for (int k = 0; k < n; ++k) dEdB[k] += dth[(size_t)t * n + k];
    terms[PRX_E_GB] += e_gb[t];
    terms[PRX_E_COULOMB] += e_cl[t];
    terms[PRX_E_LJ] += e_lj[t];
  }
  terms[PRX_E_GB] += e_self;
  terms[PRX_E_ACE] = e_ace;
  // ---- chain pass
  if (d.implicit_solvent) {
    for (int i = 0; i < n; ++i) c[i] = (T)(dEdB[i] * (double)dBdI[i]);
    std::fill(fth.begin(), fth.end(), (T)0);
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      T* ft = &fth[(size_t)tid * 3 * n];
#pragma omp for schedule(dynamic, 8)
      for (int i = 0; i < n; ++i) {
        const T rhoi = rho_of(i), si = (T)S.sr[i];
        for (int j = i + 1; j < n; ++j) {
          T dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
          const T r2p = dot(dx, dx);
          if (!gb_in(i, j, r2p)) continue;
          const T r = std::sqrt(r2p + eps12);
          const T rhoj = rho_of(j), sj = (T)S.sr[j];
          const T g = (c[i] * pair_integral_dr(r, rhoi, sj) + c[j] * pair_integral_dr(r, rhoj, si)) / r;
          for (int k = 0; k < 3; ++k) { ft[3 * i + k] -= g * dx[k]; ft[3 * j + k] += g * dx[k]; }
        }
      }
    }
    for (int t = 0; t < nth; ++t)
      for (int k = 0; k < 3 * n; ++k) f[k] += fth[(size_t)t * 3 * n + k];
  }
  // ---- bonded (serial: O(N))
  auto bond = [&](const std::vector<int32_t>& idx, const std::vector<double>& p, int cnt, int term) {
    for (int b = 0; b < cnt; ++b) {
      const int i = idx[2 * b], j = idx[2 * b + 1];
      T dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
      const T r = std::sqrt(dot(dx, dx));
      const T dr = r - (T)p[2 * b];
      terms[term] += 0.5 * p[2 * b + 1] * (double)dr * (double)dr;
      const T s = -(T)p[2 * b + 1] * dr / r;
      for (int k = 0; k < 3; ++k) { f[3 * i + k] += s * dx[k]; f[3 * j + k] -= s * dx[k]; }
    }
  };
  bond(S.bond_idx, S.bond_p, d.n_bonds, PRX_E_BOND);
  bond(S.ub_idx, S.ub_p, d.n_ub, PRX_E_UREY_BRADLEY);
  for (int a = 0; a < d.n_angles; ++a) {
    const int i = S.angle_idx[3 * a], j = S.angle_idx[3 * a + 1], k = S.angle_idx[3 * a + 2];
    T v1[3], v2[3];
    for (int q = 0; q < 3; ++q) { v1[q] = x[3 * i + q] - x[3 * j + q]; v2[q] = x[3 * k + q] - x[3 * j + q]; }
    const T d1 = std::sqrt(dot(v1, v1)), d2 = std::sqrt(dot(v2, v2));
    const T Q = d1 * d2 + (T)1e-8, P = dot(v1, v2), cth = P / Q;
    const T cc = std::min(std::max(cth, (T)-0.999999), (T)0.999999);
    const T th = std::acos(cc), dth = th - (T)S.angle_p[2 * a];
    terms[PRX_E_ANGLE] += 0.5 * S.angle_p[2 * a + 1] * (double)dth * (double)dth;
    const bool inside = cth > (T)-0.999999 && cth < (T)0.999999;
    const T dEdc = inside ? -(T)S.angle_p[2 * a + 1] * dth / std::sqrt((T)1 - cc * cc) : (T)0;
    for (int q = 0; q < 3; ++q) {
      const T g0 = v2[q] / Q - P / (Q * Q) * d2 / d1 * v1[q];
      const T g2 = v1[q] / Q - P / (Q * Q) * d1 / d2 * v2[q];
      f[3 * i + q] -= dEdc * g0;
      f[3 * k + q] -= dEdc * g2;
      f[3 * j + q] += dEdc * (g0 + g2);
    }
  }
  auto torsions = [&](const std::vector<int32_t>& idx, const std::vector<double>& p, int cnt, int T_, int w, int term) {
    for (int t = 0; t < cnt; ++t) {
      const int* a = &idx[4 * t];
      T g[4][3];
      const T phi = dihedral_grad(&x[3 * a[0]], &x[3 * a[1]], &x[3 * a[2]], &x[3 * a[3]], g);
      double e = 0, dE = 0;
      for (int k = 0; k < T_; ++k) {
        const double* q = &p[((size_t)t * T_ + k) * w];
        if (w == 3) {
          e += q[2] * (1.0 + std::cos(q[0] * (double)phi - q[1]));
          dE += -q[2] * q[0] * std::sin(q[0] * (double)phi - q[1]);
        } else {
          double diff = std::fmod((double)phi - q[1] + M_PI, 2 * M_PI);
          if (diff < 0) diff += 2 * M_PI;
          diff -= M_PI;
          e += q[0] * diff * diff;
          dE += 2.0 * q[0] * diff;
        }
      }
      terms[term] += e;
      for (int r = 0; r < 4; ++r)
        for (int k = 0; k < 3; ++k) f[3 * a[r] + k] -= (T)dE * g[r][k];
    }
  };
  torsions(S.dih_idx, S.dih_p, d.n_dihedrals, d.dihedral_terms, 3, PRX_E_DIHEDRAL);
  torsions(S.imp_idx, S.imp_p, d.n_impropers, d.improper_terms, d.improper_param_width, PRX_E_IMPROPER);
  for (int cm = 0; cm < d.n_cmap; ++cm) {
    const int* a = &S.cmap_t[5 * cm];
    T gp[4][3], gs[4][3];
    const T phi = dihedral_grad(&x[3 * a[0]], &x[3 * a[1]], &x[3 * a[2]], &x[3 * a[3]], gp);
    const T psi = dihedral_grad(&x[3 * a[1]], &x[3 * a[2]], &x[3 * a[3]], &x[3 * a[4]], gs);
    const int G = S.G;
    const double delta = 2 * M_PI / G;
    auto loc = [&](double ang, int& i, double& fr) {
      double aa = std::fmod(ang + 2 * M_PI, 2 * M_PI);
      if (aa < 0) aa += 2 * M_PI;
      i = (int)std::floor(aa / delta);
      fr = aa / delta - i;
      i = ((i % G) + G) % G;
    };
    int ia, ib;
    double da, db;
    loc((double)phi, ia, da);
    loc((double)psi, ib, db);
    const double* cf = &S.coeff[(((size_t)S.cmap_map[cm] * G + ia) * G + ib) * 16];
    double e = 0, ea = 0, eb = 0;
    for (int p = 3; p >= 0; --p) {
      const double rv = ((cf[4 * p + 3] * db + cf[4 * p + 2]) * db + cf[4 * p + 1]) * db + cf[4 * p];
      const double rd = (3 * cf[4 * p + 3] * db + 2 * cf[4 * p + 2]) * db + cf[4 * p + 1];
      ea = ea * da + e;
      e = e * da + rv;
      eb = eb * da + rd;
    }
    terms[PRX_E_CMAP] += e;
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < 3; ++k) {
        f[3 * a[r] + k] -= (T)(ea / delta) * gp[r][k];
        f[3 * a[r + 1] + k] -= (T)(eb / delta) * gs[r][k];
      }
  }
  for (int xx = 0; xx < d.n_exceptions; ++xx) {
    const int i = S.exc_pairs[2 * xx], j = S.exc_pairs[2 * xx + 1];
    T dx[3] = {x[3 * i] - x[3 * j], x[3 * i + 1] - x[3 * j + 1], x[3 * i + 2] - x[3 * j + 2]};
    const T r2 = dot(dx, dx) + eps12, r = std::sqrt(r2);
    const T s6 = (T)std::pow((double)S.exc_s[xx] / (double)r, 6);
    const T kq = (T)(d.coulomb_constant * S.exc_q[xx]);
    terms[PRX_E_EXCEPTION] += (double)((T)4 * (T)S.exc_e[xx] * (s6 * s6 - s6) + kq / r);
    const T fr = ((T)4 * (T)S.exc_e[xx] * ((T)12 * s6 * s6 - (T)6 * s6) + kq / r) / r2;
    for (int k = 0; k < 3; ++k) { f[3 * i + k] += fr * dx[k]; f[3 * j + k] -= fr * dx[k]; }
  }
  double tot = 0;
  for (int k = 0; k < PRX_E_TOTAL; ++k) tot += terms[k];
  terms[PRX_E_TOTAL] = tot;
  if (E) std::copy(terms.begin(), terms.end(), E);
}

// Philox4x32-10 and Box-Muller, as the device integrator.
void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
    c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

void normal3(uint64_t seed, uint32_t step, int local, double out[3]) {
  uint32_t c[4] = {(uint32_t)local, step, 0u, 0u};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double u1 = (c[0] + 0.5) * 2.3283064365386963e-10, u2 = (c[1] + 0.5) * 2.3283064365386963e-10;
  const double u3 = (c[2] + 0.5) * 2.3283064365386963e-10, u4 = (c[3] + 0.5) * 2.3283064365386963e-10;
  const double ra = std::sqrt(-2 * std::log(u1)), rb = std::sqrt(-2 * std::log(u3));
  out[0] = ra * std::cos(2 * M_PI * u2);
  out[1] = ra * std::sin(2 * M_PI * u2);
  out[2] = rb * std::cos(2 * M_PI * u4);
}

template <class T>
void baoab(const Sys& S, T* x, T* p, T* f, int steps, double dt_fs, double temperature, double gamma_ps,
           uint64_t seed, uint32_t step0) {
  const int n = S.d.n_atoms;
  const double dt = dt_fs / 48.88821291839, gamma = gamma_ps * 48.88821291839e-3, kT = 0.0019872041 * temperature;
  const double c1 = std::exp(-gamma * dt), c2 = std::sqrt(std::max(0.0, 1 - c1 * c1));
  for (int s = 0; s < steps; ++s) {
#pragma omp parallel for
    for (int i = 0; i < n; ++i) {
      double xi[3];
      normal3(seed, step0 + (uint32_t)s, i, xi);
      const T inv_m = (T)(1.0 / S.m[i]);
      for (int k = 0; k < 3; ++k) {
        T pk = p[3 * i + k] + (T)(0.5 * dt) * f[3 * i + k];
        T xk = x[3 * i + k] + (T)(0.5 * dt) * pk * inv_m;
        pk = (T)c1 * pk + (T)(c2 * std::sqrt(S.m[i] * kT) * xi[k]);
        xk += (T)(0.5 * dt) * pk * inv_m;
        x[3 * i + k] = xk;
        p[3 * i + k] = pk;
      }
    }
    energy_forces<T>(S, x, f, nullptr);
    for (int k = 0; k < 3 * n; ++k) p[k] += (T)(0.5 * dt) * f[k];
  }
}

template <class V>
void copy_in(V& dst, const typename V::value_type* src, size_t n) {
  if (src && n) dst.assign(src, src + n);
}

}  // namespace

extern "C" {

// cmap_coeff: (n_cmap_maps, G, G, 16) float64 patch coefficients, or NULL without CMAP.
void* oracle_cpu_create(const prx_system_desc* d, const double* cmap_coeff) {
  Sys* S = new Sys();
  S->d = *d;
  const int n = d->n_atoms;
  copy_in(S->q, d->charges, n);
  copy_in(S->sig, d->sigmas, n);
  copy_in(S->eps, d->epsilons, n);
  copy_in(S->R, d->radii, n);
  if (d->scaled_radii) copy_in(S->sr, d->scaled_radii, n); else S->sr = S->R;
  if (d->masses) copy_in(S->m, d->masses, n); else S->m.assign(n, 1.0);
  copy_in(S->bond_idx, d->bond_idx, 2 * (size_t)d->n_bonds);
  copy_in(S->bond_p, d->bond_params, 2 * (size_t)d->n_bonds);
  copy_in(S->ub_idx, d->ub_idx, 2 * (size_t)d->n_ub);
  copy_in(S->ub_p, d->ub_params, 2 * (size_t)d->n_ub);
  copy_in(S->angle_idx, d->angle_idx, 3 * (size_t)d->n_angles);
  copy_in(S->angle_p, d->angle_params, 2 * (size_t)d->n_angles);
  copy_in(S->dih_idx, d->dihedral_idx, 4 * (size_t)d->n_dihedrals);
  copy_in(S->dih_p, d->dihedral_params, 3 * (size_t)d->n_dihedrals * std::max(d->dihedral_terms, 0));
  copy_in(S->imp_idx, d->improper_idx, 4 * (size_t)d->n_impropers);
  copy_in(S->imp_p, d->improper_params,
          (size_t)d->improper_param_width * d->n_impropers * std::max(d->improper_terms, 0));
  copy_in(S->cmap_t, d->cmap_torsions, 5 * (size_t)d->n_cmap);
  copy_in(S->cmap_map, d->cmap_map, (size_t)d->n_cmap);
  S->G = d->cmap_grid;
  if (d->n_cmap && cmap_coeff) copy_in(S->coeff, cmap_coeff, (size_t)d->n_cmap_maps * S->G * S->G * 16);
  copy_in(S->exc_pairs, d->exc_pairs, 2 * (size_t)d->n_exceptions);
  copy_in(S->exc_s, d->exc_sigma, (size_t)d->n_exceptions);
  copy_in(S->exc_e, d->exc_epsilon, (size_t)d->n_exceptions);
  copy_in(S->exc_q, d->exc_chargeprod, (size_t)d->n_exceptions);
  // exclusion rows -> symmetrised per-atom partner lists with (S(i,j) + S(j,i)) / 2
  S->excl.assign(n, {});
  const int C = d->excl_cols;
  if (C > 0 && d->excl_idx) {
    std::vector<std::vector<std::pair<int, std::pair<double, double>>>> row(n);
    for (int i = 0; i < n; ++i)
      for (int k = 0; k < C; ++k) {
        const int j = d->excl_idx[(size_t)i * C + k];
        if (j < 0 || j >= n || j == i) continue;
        bool found = false;
        for (auto& e : row[i])
          if (e.first == j) {
            e.second.first += d->excl_scale_vdw[(size_t)i * C + k];
            e.second.second += d->excl_scale_elec[(size_t)i * C + k];
            found = true;
          }
        if (!found) row[i].push_back({j, {d->excl_scale_vdw[(size_t)i * C + k], d->excl_scale_elec[(size_t)i * C + k]}});
      }
    auto get = [&](int i, int j) -> std::pair<double, double> {
      for (auto& e : row[i])
        if (e.first == j) return e.second;
      return {1.0, 1.0};
    };
    std::set<std::pair<int, int>> seen;
    for (int i = 0; i < n; ++i)
      for (auto& e : row[i]) {
        const int a = std::min(i, e.first), b = std::max(i, e.first);
        if (!seen.insert({a, b}).second) continue;
        const auto sa = get(a, b), sb = get(b, a);
        S->excl[a].push_back({b, {0.5 * (sa.first + sb.first), 0.5 * (sa.second + sb.second)}});
      }
  }
  return S;
}

void oracle_cpu_destroy(void* h) { delete static_cast<Sys*>(h); }

void oracle_cpu_set_f32_cutoff(void* h, int on) { static_cast<Sys*>(h)->f32_cut = on ? 1 : 0; }

int oracle_cpu_threads(void) { return omp_get_max_threads(); }

// energies: PRX_N_ENERGY doubles (layout of include/prolix_gb.h)
void oracle_cpu_energy_forces_f64(void* h, const double* x, double* f, double* E) {
  energy_forces<double>(*static_cast<Sys*>(h), x, f, E);
}

void oracle_cpu_energy_forces_f32(void* h, const float* x, float* f, double* E) {
  energy_forces<float>(*static_cast<Sys*>(h), x, f, E);
}

void oracle_cpu_baoab_f64(void* h, double* x, double* p, double* f, int steps, double dt_fs, double temperature,
                          double gamma_ps, uint64_t seed, uint32_t step0) {
  baoab<double>(*static_cast<Sys*>(h), x, p, f, steps, dt_fs, temperature, gamma_ps, seed, step0);
}

void oracle_cpu_baoab_f32(void* h, float* x, float* p, float* f, int steps, double dt_fs, double temperature,
                          double gamma_ps, uint64_t seed, uint32_t step0) {
  baoab<float>(*static_cast<Sys*>(h), x, p, f, steps, dt_fs, temperature, gamma_ps, seed, step0);
}

}  // extern "C"
