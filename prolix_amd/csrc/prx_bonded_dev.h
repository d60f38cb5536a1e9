// Device code of the bonded terms, shared by prx_bonded.hip and
// prx_persist.hip. One thread per arena entry; each entry writes one force
// record per atom (see prx_bonded.hip).
#pragma once

#include "prx_common.h"

namespace prx {

constexpr int kBondedThreads = 256;

__device__ __forceinline__ float3 ld3(const DeviceView& v, int g) {
  const float4 p = v.pos[g];
  return make_float3(p.x, p.y, p.z);
}
__device__ __forceinline__ float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
  return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// a * b - c * d to about one rounding of the result (Kahan's difference of
// products with fused multiply-adds: w = c d rounded, e = w - c d exactly).
// Cross products of nearly parallel vectors keep their relative accuracy,
// where the plain form loses ~|a||b| / |a x b| roundings.
__device__ __forceinline__ float dop(float a, float b, float c, float d) {
  const float w = c * d;
  const float e = fmaf(-c, d, w);
  const float f = fmaf(a, b, -w);
  return f + e;
}
__device__ __forceinline__ float3 cross3_acc(float3 a, float3 b) {
  return make_float3(dop(a.y, b.z, a.z, b.y), dop(a.z, b.x, a.x, b.z), dop(a.x, b.y, a.y, b.x));
}
__device__ __forceinline__ float3 scl3(float s, float3 a) { return make_float3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 add3(float3 a, float3 b) { return make_float3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ void put(float4* r, float3 f) { *r = make_float4(f.x, f.y, f.z, 0.0f); }

// Dihedral angle of bonded.py:84-107 and its gradient with respect to the
// four atoms (Blondel & Karplus, J. Comput. Chem. 17, 1132). With A = b1 x b0
// and B = b1 x b2 (b0 = p0 - p1, b1 = p2 - p1, b2 = p3 - p2), the reference's
// projections satisfy vv.ww = A.B / |b1|^2 and, as A x B = b1 (b1.(b0 x b2)),
// (u x vv).ww = (A x B).b1 / |b1|^3, so phi = atan2((A x B).b1 / |b1|, A.B)
// (its 1e-8 in the norm of b1 changes phi by far less than a rounding). A, B
// and A x B come from cross3_acc, so phi and the gradient keep float32
// accuracy when three of the atoms are nearly collinear (|A| or |B| much
// smaller than the products of the bond lengths); a dot product such as b0.B
// would not, its terms cancelling to sin(theta) of their size.
struct Dihedral {
  float phi;
  float3 g0, g1, g2, g3;
};

__device__ __forceinline__ Dihedral dihedral(float3 p0, float3 p1, float3 p2, float3 p3) {
  const float3 b0 = sub3(p0, p1), b1 = sub3(p2, p1), b2 = sub3(p3, p2);
  const float b1n = sqrtf(dot3(b1, b1));
  // F = p0 - p1, G = p1 - p2, H = p3 - p2
  const float3 F = b0, G = scl3(-1.0f, b1), H = b2;
  const float3 A = cross3_acc(F, G), B = cross3_acc(H, G);
  Dihedral d;
  d.phi = atan2f(dot3(cross3_acc(A, B), b1) / b1n, dot3(A, B));
  const float A2 = dot3(A, A), B2 = dot3(B, B);
  const float FG = dot3(F, G), HG = dot3(H, G);
  const float ga = b1n / A2, gb = b1n / B2;
  const float fa = FG / (A2 * b1n), hb = HG / (B2 * b1n);
  d.g0 = scl3(-ga, A);
  d.g3 = scl3(gb, B);
  d.g1 = add3(scl3(ga + fa, A), scl3(-hb, B));
  d.g2 = add3(scl3(-fa, A), scl3(hb - gb, B));
  return d;
}

__device__ __forceinline__ void torsion_records(float4* r, const Dihedral& d, float dEdphi) {
  put(r + 0, scl3(-dEdphi, d.g0));
  put(r + 1, scl3(-dEdphi, d.g1));
  put(r + 2, scl3(-dEdphi, d.g2));
  put(r + 3, scl3(-dEdphi, d.g3));
}

template <bool ENERGY>
__device__ void do_bond(const DeviceView& v, int i) {
  const BondEntry b = v.bonds[i];
  if (b.idx.x < 0) return;
  const float3 d = sub3(ld3(v, b.idx.x), ld3(v, b.idx.y));
  const float r = sqrtf(dot3(d, d));
  const float dr = r - b.p.x;
  const float s = -b.p.y * dr / r;  // F_0 = s * d
  float4* rec = v.rec + v.rec_base[0] + (int64_t)i * kRecBond;
  put(rec + 0, scl3(s, d));
  put(rec + 1, scl3(-s, d));
  if (ENERGY) v.e_bond[i] = 0.5 * (double)b.p.y * (double)dr * (double)dr;
}

template <bool ENERGY>
__device__ void do_angle(const DeviceView& v, int i) {
  const AngleEntry a = v.angles[i];
  if (a.idx.x < 0) return;
  const float3 pj = ld3(v, a.idx.y);
  const float3 v1 = sub3(ld3(v, a.idx.x), pj), v2 = sub3(ld3(v, a.idx.z), pj);
  const float d1 = sqrtf(dot3(v1, v1)), d2 = sqrtf(dot3(v2, v2));
  const float Q = d1 * d2 + 1e-8f;
  const float P = dot3(v1, v2);
  const float c = P / Q;
  // bonded.py:76-77 clips cos(theta) to +-0.999999: theta is constant beyond
  // it and the force zero
  const bool inside = c > -0.999999f && c < 0.999999f;
  // theta = acos(P / Q) of the reference (Q with its 1e-8) without acos and
  // 1 / sqrt(1 - c^2), which lose ~1e-7 / (1 - c^2) relative near a straight
  // angle (1e-3 at 179.5 degrees): with n = v1 x v2 from cross3_acc,
  // Q^2 - P^2 = |n|^2 + 2e-8 |v1||v2| + 1e-16 = y^2, so theta = atan2(y, P) and
  // dc/dx_i = -(v1 x n) / (Q |v1|^2) + 1e-8 P v1 / (|v1|^2 Q^2), likewise for
  // x_k; the force on i is k (theta - theta0) (Q / y) dc/dx_i.
  const float3 n = cross3_acc(v1, v2);
  const float y = sqrtf(dot3(n, n) + 2e-8f * d1 * d2 + 1e-16f);
  const float th = inside ? atan2f(y, P) : acosf(fminf(fmaxf(c, -0.999999f), 0.999999f));
  const float dth = th - a.p.x;
  const float g = inside ? a.p.y * dth * Q / y : 0.0f;
  const float pe = P * 1e-8f / (Q * Q);
  const float3 fi = scl3(g, add3(scl3(-1.0f / (Q * d1 * d1), cross3(v1, n)), scl3(pe / (d1 * d1), v1)));
  const float3 fk = scl3(g, add3(scl3(1.0f / (Q * d2 * d2), cross3(v2, n)), scl3(pe / (d2 * d2), v2)));
  float4* rec = v.rec + v.rec_base[1] + (int64_t)i * kRecAngle;
  put(rec + 0, fi);
  put(rec + 1, scl3(-1.0f, add3(fi, fk)));
  put(rec + 2, fk);
  if (ENERGY) v.e_angle[i] = 0.5 * (double)a.p.y * (double)dth * (double)dth;
}

__device__ __forceinline__ void periodic_term(float n, float phase, float k, float phi, float& e, float& dE) {
  float sa, ca;
  sincosf(n * phi - phase, &sa, &ca);
  e += k * (1.0f + ca);
  dE -= k * n * sa;
}

template <bool ENERGY>
__device__ void do_torsion(const DeviceView& v, int i) {
  const TorsionEntry t = v.tors[i];
  if (t.idx.x < 0) return;
  const Dihedral d = dihedral(ld3(v, t.idx.x), ld3(v, t.idx.y), ld3(v, t.idx.z), ld3(v, t.idx.w));
  float e = 0.0f, dE = 0.0f;
  if (t.a.w == 2.0f) {  // harmonic improper: a.x = k, a.y = phi0
    const float twopi = 6.283185307179586f;
    float diff = d.phi - t.a.y + 3.141592653589793f;
    diff = diff - twopi * floorf(diff / twopi) - 3.141592653589793f;
    e = t.a.x * diff * diff;
    dE = 2.0f * t.a.x * diff;
  } else {
    periodic_term(t.a.x, t.a.y, t.a.z, d.phi, e, dE);
    if (t.b.w > 1.5f) periodic_term(t.b.x, t.b.y, t.b.z, d.phi, e, dE);
  }
  torsion_records(v.rec + v.rec_base[2] + (int64_t)i * kRecTors, d, dE);
  if (ENERGY) v.e_tors[i] = (double)e;
}

template <bool ENERGY>
__device__ void do_cmap(const DeviceView& v, int i) {
  const CmapEntry c = v.cmaps[i];
  if (c.idx.x < 0) return;
  const float3 p0 = ld3(v, c.idx.x), p1 = ld3(v, c.idx.y), p2 = ld3(v, c.idx.z), p3 = ld3(v, c.idx.w);
  const float3 p4 = ld3(v, c.idx2.x);
  const Dihedral dp = dihedral(p0, p1, p2, p3);
  const Dihedral ds = dihedral(p1, p2, p3, p4);
  const int G = c.idx2.z;
  const float twopi = 6.283185307179586f;
  const float delta = twopi / (float)G;
  float ap = dp.phi + twopi;
  ap = ap - twopi * floorf(ap / twopi);
  float as = ds.phi + twopi;
  as = as - twopi * floorf(as / twopi);
  const float xa = ap / delta, xb = as / delta;
  int ia = (int)floorf(xa), ib = (int)floorf(xb);
  const float da = xa - (float)ia, db = xb - (float)ib;
  ia = ((ia % G) + G) % G;
  ib = ((ib % G) + G) % G;
  const float* cf = v.cmap_coeff + c.idx2.y + ((size_t)ia * G + ib) * 16;
  // E(a, b) = sum_p a^p R_p(b), R_p(b) = sum_q c[4p + q] b^q; Horner in a.
  float e = 0.f, ea = 0.f, eb = 0.f;
  for (int p = 3; p >= 0; --p) {
    const float q0 = cf[4 * p], q1 = cf[4 * p + 1], q2 = cf[4 * p + 2], q3 = cf[4 * p + 3];
    const float rv = ((q3 * db + q2) * db + q1) * db + q0;   // R_p(b)
    const float rd = (3.0f * q3 * db + 2.0f * q2) * db + q1;  // R_p'(b)
    ea = ea * da + e;  // d/da of the running Horner value
    e = e * da + rv;
    eb = eb * da + rd;
  }
  const float dphi = ea / delta, dpsi = eb / delta;
  // atoms a, b, c, d, e: phi uses a..d (gradients dp.g0..g3), psi uses b..e (ds.g0..g3)
  float4* rec = v.rec + v.rec_base[3] + (int64_t)i * kRecCmap;
  put(rec + 0, scl3(-dphi, dp.g0));
  put(rec + 1, add3(scl3(-dphi, dp.g1), scl3(-dpsi, ds.g0)));
  put(rec + 2, add3(scl3(-dphi, dp.g2), scl3(-dpsi, ds.g1)));
  put(rec + 3, add3(scl3(-dphi, dp.g3), scl3(-dpsi, ds.g2)));
  put(rec + 4, scl3(-dpsi, ds.g3));
  if (ENERGY) v.e_cmap[i] = (double)e;
}

template <bool ENERGY>
__device__ void do_exception(const DeviceView& v, int i) {
  const ExceptionEntry x = v.excs[i];
  if (x.idx.x < 0) return;
  const float3 d = sub3(ld3(v, x.idx.x), ld3(v, x.idx.y));
  const float r2 = dot3(d, d) + 1e-12f;
  const float r = sqrtf(r2);
  const float s = x.p.x / r;
  const float s2 = s * s;
  const float s6 = s2 * s2 * s2;
  const float e = 4.0f * x.p.y * (s6 * s6 - s6) + x.p.z / r;
  // -dE/dr / r
  const float fr = (4.0f * x.p.y * (12.0f * s6 * s6 - 6.0f * s6) + x.p.z / r) / r2;
  float4* rec = v.rec + v.rec_base[4] + (int64_t)i * kRecExc;
  put(rec + 0, scl3(fr, d));
  put(rec + 1, scl3(-fr, d));
  if (ENERGY) v.e_exc[i] = (double)e;
}

// The bonded work of flattened entry e over the five kinds in order (bonds,
// angles, torsions, CMAP, exceptions), for a grid-stride loop.
template <bool ENERGY>
__device__ __forceinline__ void bonded_entry(const DeviceView& v, int e) {
  if (e < v.bond_hw) { do_bond<ENERGY>(v, e); return; }
  e -= v.bond_hw;
  if (e < v.angle_hw) { do_angle<ENERGY>(v, e); return; }
  e -= v.angle_hw;
  if (e < v.tors_hw) { do_torsion<ENERGY>(v, e); return; }
  e -= v.tors_hw;
  if (e < v.cmap_hw) { do_cmap<ENERGY>(v, e); return; }
  e -= v.cmap_hw;
  if (e < v.exc_hw) do_exception<ENERGY>(v, e);
}

__host__ __device__ inline int bonded_entries(const DeviceView& v) {
  return v.bond_hw + v.angle_hw + v.tors_hw + v.cmap_hw + v.exc_hw;
}

}  // namespace prx
