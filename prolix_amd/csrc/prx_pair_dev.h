// Device code of the pair passes, shared by prx_pair.hip (one kernel per
// pass) and prx_persist.hip (every pass of a step in one persistent kernel).
// born_item / force_item / chain_item are one wave's work on one item.
#pragma once

#include "prx_common.h"
#include "prx_obc_farfield.h"

namespace prx {

// DPP wave_rol:1 (lane l reads lane l + 1 mod 64). After k rotations lane l
// holds the value that started in lane (l + k) & 63.
#define PRX_DPP_WAVE_ROL1 0x134

// BPERM moves the value through the LDS crossbar (ds_bpermute_b32: lane l
// reads lane (l + 1) & 63), which takes no VALU issue slot; otherwise the DPP
// wave rotate. Chosen per pass (PRX_ROT_BPERM_BORN / _FORCE / _CHAIN below).
template <bool BPERM>
__device__ __forceinline__ float rot(float x) {
  if (BPERM) {
    const int src = ((int)(threadIdx.x + 1) & 63) << 2;
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(x)));
  }
  // Every lane has a valid source under a full-wave rotate, so no "old" value is needed.
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), PRX_DPP_WAVE_ROL1, 0xF, 0xF, true));
}

// Partial-buffer slot of (item w, side, lane); side 0 = i atoms, 1 = j atoms.
__device__ __forceinline__ size_t part_slot(int w, int side, int lane) {
  return ((size_t)w * 2 + side) * kTile + lane;
}

// Chain partials carry three components: stored as three float arrays of 64
// per (item, side), each written and read whole by a wave.
__device__ __forceinline__ void put_part3(const DeviceView& v, int w, int side, int lane, float x, float y,
                                          float z) {
  float* pc = reinterpret_cast<float*>(v.part) + ((size_t)w * 2 + side) * (3 * kTile);
  pc[lane] = x;
  pc[kTile + lane] = y;
  pc[2 * kTile + lane] = z;
}

__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float flog2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ float fexp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Orders this wave's LDS stores before its later LDS loads from other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Settings chosen by the autotuner (prolix_amd/tune.py) for this GPU; the
// defaults below apply to anything it leaves unset.
#include "prx_tuned.h"

// How each pass rotates its j accumulators: 1 = LDS crossbar, 0 = DPP (rot);
// the chain pass's three also split 1 + 2 (value 2) or 2 + 1 (value 3).
#ifndef PRX_ROT_BPERM_BORN
#define PRX_ROT_BPERM_BORN 0
#endif
#ifndef PRX_ROT_BPERM_FORCE
#define PRX_ROT_BPERM_FORCE 0
#endif
#ifndef PRX_ROT_BPERM_CHAIN
#define PRX_ROT_BPERM_CHAIN 0
#endif

// Work items (waves) per workgroup of the pair kernels.
#ifndef PRX_PAIR_WAVES
#define PRX_PAIR_WAVES 4
#endif
constexpr int kWavesPerBlock = PRX_PAIR_WAVES;

// Minimum waves per SIMD the compiler must fit the force and chain kernels'
// registers to (0: no constraint; their unrolled loops otherwise take what
// they need and run at lower occupancy).
#ifndef PRX_MIN_WAVES
#define PRX_MIN_WAVES 0
#endif
#if PRX_MIN_WAVES > 0
#define PRX_OCCUPANCY __attribute__((amdgpu_waves_per_eu(PRX_MIN_WAVES)))
#else
#define PRX_OCCUPANCY
#endif
constexpr int kPairThreads = 64 * kWavesPerBlock;
static_assert(kWavesPerBlock >= 1 && kWavesPerBlock <= 8, "PRX_PAIR_WAVES out of range");

// LDS j accumulators are read and written through volatile pointers: the
// compiler keeps those accesses in program order (the next iteration's read of
// an entry another lane just wrote stays behind that write) while the
// read-only j data loads remain free to move.
typedef float v4f __attribute__((ext_vector_type(4)));

// Build-time tuning knobs (defaults are the measured best on MI355X):
// unroll depth of the three pair loops, and where the j-side accumulators
// live: 0 = registers rotated by DPP; 1 = LDS, stored twice over like the j
// data: lane l adds iteration k's term to entry l + k (an immediate offset;
// no two lanes share an entry in one iteration, and a wave's LDS operations
// complete in issue order, so the read-add-write needs only a compiler
// ordering), and atom j's sum is entry j plus entry j + 64.
#ifndef PRX_UNROLL_BORN
#define PRX_UNROLL_BORN 8
#endif
#ifndef PRX_UNROLL_FORCE
#define PRX_UNROLL_FORCE 4
#endif
#ifndef PRX_UNROLL_CHAIN
#define PRX_UNROLL_CHAIN 8
#endif
#ifndef PRX_LDS_ACC
#define PRX_LDS_ACC 0
#endif
#define PRX_PRAGMA(x) _Pragma(#x)
#define PRX_UNROLL(n) PRX_PRAGMA(unroll n)

// Work item of this wave. Workgroups are dispatched to the 8 XCDs round-robin
// (block b to XCD b mod 8), and each XCD has its own L2. With PRX_XCD_MAP the
// blocks are renumbered so that XCD x runs one contiguous run of items
// (consecutive items share their i tile and neighbouring j tiles, which then
// hit in that XCD's L2): block b = 8k + x becomes x * per + min(x, rem) + k,
// a bijection on [0, n_blocks) with per = n_blocks / 8, rem = n_blocks % 8.
#ifndef PRX_XCD_MAP
#define PRX_XCD_MAP 1
#endif
constexpr int kXcds = 8;

__device__ __forceinline__ int wave_id() {
  int b = (int)blockIdx.x;
#if PRX_XCD_MAP
  const int nb = (int)gridDim.x, per = nb / kXcds, rem = nb % kXcds;
  const int x = b % kXcds, k = b / kXcds;
  b = x * per + min(x, rem) + k;
#endif
  return __builtin_amdgcn_readfirstlane((int)(b * (blockDim.x >> 6) + (threadIdx.x >> 6)));
}

// The j atoms of an item: a tile (gj = j0 + lane), or for a chunk item of a
// neighbour list up to 64 listed atoms (gj = the lane's entry). nj = 0: an
// empty chunk (its row's list is shorter), which does nothing.
struct JSrc {
  int gj;  // this lane's j atom (clamped to the last valid one)
  int nj;  // j atoms of the item
};

__device__ __forceinline__ JSrc j_source(const DeviceView& v, const PairItem& it, int lane) {
  JSrc j;
  if (!item_is_chunk(it)) {
    j.nj = item_nj(it.meta);
    j.gj = it.j0 + min(lane, j.nj - 1);
    return j;
  }
  const int e0 = -1 - it.j0;
  const int tI = it.excl;
  j.nj = max(0, min(kTile, v.nl_count[tI] - (e0 - v.nl_row[tI].y)));
  j.gj = j.nj > 0 ? v.nl_entries[e0 + min(lane, j.nj - 1)] : 0;
  return j;
}

// ---------------------------------------------------------------------------
// OBC2 pair integral pieces
// ---------------------------------------------------------------------------

// Both directions of one pair share r and one reciprocal.
struct BornPair {
  float I1, I2;  // I(r; rho_i, s_j) and I(r; rho_j, s_i)
};

__device__ __forceinline__ BornPair born_pair(float r, float rinv, float rho_i, float s_i, float s_i2,
                                              float rho_j, float s_j, float s_j2) {
  const float U1 = r + s_j, U2 = r + s_i;
  const float L1 = fmaxf(fmaxf(rho_i, fabsf(r - s_j)), 1e-4f);
  const float L2 = fmaxf(fmaxf(rho_j, fabsf(r - s_i)), 1e-4f);
  const float U1c = fmaxf(U1, 1e-4f), U2c = fmaxf(U2, 1e-4f);
  const float p1 = L1 * U1c, p2 = L2 * U2c;
  const float inv = frcp(p1 * p2);
  const float ip1 = inv * p2, ip2 = inv * p1;
  const float iL1 = U1c * ip1, iU1 = L1 * ip1;
  const float iL2 = U2c * ip2, iU2 = L2 * ip2;
  const float rs = fminf(rinv, 1e4f);  // 1 / max(r, 1e-4)
  const float lg1 = flog2(L1 * iU1), lg2 = flog2(L2 * iU2);
  const float a1 = iL1 - iU1, a2 = iL2 - iU2;
  // 1/U^2 - 1/L^2 = -(1/L - 1/U)(1/L + 1/U)
  const float b1 = -a1 * (iL1 + iU1), b2 = -a2 * (iL2 + iU2);
  const float c1 = fmaf(-s_j2, rs, r), c2 = fmaf(-s_i2, rs, r);
  float I1 = fmaf(0.5f, a1, fmaf(0.25f * kLn2 * lg1, rs, 0.125f * c1 * b1));
  float I2 = fmaf(0.5f, a2, fmaf(0.25f * kLn2 * lg2, rs, 0.125f * c2 * b2));
  BornPair o;
  o.I1 = (U1 > rho_i) ? I1 : 0.0f;
  o.I2 = (U2 > rho_j) ? I2 : 0.0f;
  return o;
}

// dI/dr for one direction (rho = rho_i, s = sr_j), the derivative of born_pair's I.
__device__ __forceinline__ float born_dr(float r, float rs, float rho, float s, float s2) {
  const float dd = r - s;
  const float D = fabsf(dd);
  const bool useD = D > rho;
  const float L = useD ? D : rho;
  const float Lc = fmaxf(L, 1e-4f);
  float Lp = useD ? copysignf(1.0f, dd) : 0.0f;
  Lp = (L > 1e-4f) ? Lp : 0.0f;
  const float U = r + s;
  const float Uc = fmaxf(U, 1e-4f);
  const float Up = (U > 1e-4f) ? 1.0f : 0.0f;
  const float rp = (rs < 1e4f) ? 1.0f : 0.0f;
  const float inv = frcp(Lc * Uc);
  const float iL = Uc * inv, iU = Lc * inv;
  const float lg = flog2(Lc * iU) * kLn2;
  const float iL2 = iL * iL, iU2 = iU * iU;
  const float d1 = 0.5f * (Up * iU2 - Lp * iL2);
  const float d2 = 0.25f * rs * ((Lp * iL - Up * iU) - lg * rp * rs);
  const float d3 = 0.125f * fmaf(s2 * rp, rs * rs, 1.0f) * (iU2 - iL2);
  const float d4 = 0.25f * fmaf(-s2, rs, r) * (Lp * iL2 * iL - Up * iU2 * iU);
  // Every intermediate is finite (Lc, Uc >= 1e-4), so a 0/1 factor is exact
  // and keeps the wave free of a divergent branch.
  const float on = (U > rho) ? 1.0f : 0.0f;
  return on * (d1 + d2 + d3 + d4);
}

// Far field (r - s > rho and s <= r/2): L = r - s, U = r + s and
//   I = x^3 h(x^2) / (2r),  dI/dr = -x^3 k(x^2) / r^2,  x = s / r,
// with the polynomials of prx_obc_farfield.h. Same function as the direct
// formula, without its cancellation and without a transcendental.
__device__ __forceinline__ bool obc_far(float r, float rho, float s) {
  return (r - s > rho) && (s <= PRX_OBC_FAR_XMAX * r);
}

// Written with u = 1/r^2 and the per-atom constants s^2 and s^3/2 (born4.z, .w):
//   I = (s^3/2) u^2 h(s^2 u),   dI/dr = -2 (s^3/2) u^2 r^-1 k(s^2 u),
// so each direction costs one product for the argument, the Horner chain and
// two products; u^2 and u^2/r are shared by both directions of a pair.
// F4: the degree-4 fits on y <= 1/16 (s <= r/4), for items whose every pair
// is that far apart (kClsObcFar4); else the fits on y <= 1/4.
template <bool F4 = false>
__device__ __forceinline__ float obc_far_I(float s2, float s3h, float u, float u2) {
  const float y = s2 * u;
  float h;
  if (F4) {
    h = kObcH4[kObcH4Deg];
#pragma unroll
    for (int k = kObcH4Deg - 1; k >= 0; --k) h = fmaf(h, y, kObcH4[k]);
  } else {
    h = kObcH[kObcHDeg];
#pragma unroll
    for (int k = kObcHDeg - 1; k >= 0; --k) h = fmaf(h, y, kObcH[k]);
  }
  return (s3h * u2) * h;
}

// -2 k(y): the factor -2 scales every coefficient exactly, so this is -2 times
// the Horner value of k bit for bit.
template <bool F4 = false>
__device__ __forceinline__ float obc_far_km2(float y) {
  float q;
  if (F4) {
    q = -2.0f * kObcK4[kObcK4Deg];
#pragma unroll
    for (int k = kObcK4Deg - 1; k >= 0; --k) q = fmaf(q, y, -2.0f * kObcK4[k]);
  } else {
    q = -2.0f * kObcK[kObcKDeg];
#pragma unroll
    for (int k = kObcKDeg - 1; k >= 0; --k) q = fmaf(q, y, -2.0f * kObcK[k]);
  }
  return q;
}

__device__ __forceinline__ float obc_far_dI(float s2, float s3h, float u, float u2r) {
  return (s3h * u2r) * obc_far_km2(s2 * u);
}

// Chain-rule scalar of a pair with both directions in the far field:
//   (c_i dI(rho_i, s_j)/dr + c_j dI(rho_j, s_i)/dr) / r
//     = -2 u^3 (c_i s_j^3/2 k(s_j^2 u) + c_j s_i^3/2 k(s_i^2 u)),   u = 1/r^2.
template <bool F4 = false>
__device__ __forceinline__ float chain_far(float u, float sj2, float ci_s3hj, float si2, float cj_s3hi) {
  return (u * u * u) * fmaf(ci_s3hj, obc_far_km2<F4>(sj2 * u), cj_s3hi * obc_far_km2<F4>(si2 * u));
}

// ---------------------------------------------------------------------------
// Pass 1: Born sums
// ---------------------------------------------------------------------------

// LDS slice of one wave: j data staged twice (2 x 64 entries).
struct BornLds {  // (x, y, z, s^2), (s^3/2, rho, s, 0); j accumulators
  float4 a[128];
  float4 b[128];
#if PRX_LDS_ACC
  float acc[128];
#endif
};

// FAR: every pair of the item is in the far field (no per-pair test); F4:
// and at least 4 s apart (the degree-4 fits).
template <bool GENERAL, bool DIAG, bool FAR, bool F4 = false>
__device__ __forceinline__ void born_tile(const DeviceView& v, const PairItem& it, const JSrc& js, int lane, int w,
                                          float cut2, BornLds& L) {
  const int ni = item_ni(it.meta), nj = js.nj;
  const int gi = it.i0 + min(lane, ni - 1);
  const int gj = js.gj;
  const float4 pi = v.pos[gi], bi = v.born4[gi];
  {
    const float4 pj = v.pos[gj], bj = v.born4[gj];
    // the far-field loop reads only a and b.x
    const float4 a = make_float4(pj.x, pj.y, pj.z, bj.z);
    const float4 b = make_float4(bj.w, bj.x, bj.y, 0.0f);
    L.a[lane] = a; L.a[lane + 64] = a;
    L.b[lane] = b; L.b[lane + 64] = b;
#if PRX_LDS_ACC
    L.acc[lane] = 0.0f;
    L.acc[lane + 64] = 0.0f;
#endif
  }
  wave_lds_sync();
  float accI = 0.0f, accJ = 0.0f;
  const bool ivalid = lane < ni;
  constexpr int kbeg = DIAG ? 1 : 0, kend = DIAG ? 32 : 63;
  PRX_UNROLL(PRX_UNROLL_BORN)
  for (int k = kbeg; k <= kend; ++k) {
    const float4 ja = L.a[lane + k];
    const float xj = ja.x, yj = ja.y, zj = ja.z, sj2 = ja.w, sj3h = L.b[lane + k].x;
    const float dx = xj - pi.x, dy = yj - pi.y, dz = zj - pi.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    BornPair bp;
    if (FAR) {
      const float u = frcp(r2), u2 = u * u;
      bp.I1 = obc_far_I<F4>(sj2, sj3h, u, u2);
      bp.I2 = obc_far_I<F4>(bi.z, bi.w, u, u2);
    } else {
      const float rinv = frsq(r2);
      const float u = rinv * rinv, u2 = u * u;
      const float rhoj = L.b[lane + k].y, sj = L.b[lane + k].z;
      const float r = r2 * rinv;
      const bool far1 = obc_far(r, bi.x, sj), far2 = obc_far(r, rhoj, bi.y);
      if (__all(far1 && far2)) {
        bp.I1 = obc_far_I(sj2, sj3h, u, u2);
        bp.I2 = obc_far_I(bi.z, bi.w, u, u2);
      } else {
        bp = born_pair(r, rinv, bi.x, bi.y, bi.z, rhoj, sj, sj2);
        bp.I1 = far1 ? obc_far_I(sj2, sj3h, u, u2) : bp.I1;
        bp.I2 = far2 ? obc_far_I(bi.z, bi.w, u, u2) : bp.I2;
      }
    }
    if (GENERAL) {
      const int orig = (lane + k) & 63;
      bool valid = ivalid && orig < nj;
      if (DIAG && k == 32) valid = valid && lane < 32;
      if (cut2 > 0.0f) valid = valid && r2 < cut2;
      bp.I1 = valid ? bp.I1 : 0.0f;
      bp.I2 = valid ? bp.I2 : 0.0f;
    }
    accI += bp.I1;
#if PRX_LDS_ACC
    {
      volatile float* acc = L.acc;
      acc[lane + k] = acc[lane + k] + bp.I2;
    }
#else
    accJ += bp.I2;
    if (k != kend) accJ = rot<PRX_ROT_BPERM_BORN>(accJ);
#endif
  }
  float* pb = reinterpret_cast<float*>(v.part);  // scalar partials
  pb[part_slot(w, 0, lane)] = accI;
#if PRX_LDS_ACC
  wave_lds_sync();
  pb[part_slot(w, 1, lane)] = L.acc[lane] + L.acc[lane + 64];
#else
  const int home = (lane + kend) & 63;
  pb[part_slot(w, 1, home)] = accJ;
#endif
}

// Class bits of an item from its tiles' bounding boxes (wave-uniform). The
// OBC2 far-field bit needs the box distance above every pair's threshold,
// max(s_j + rho_i, s_i + rho_j, 2 s) over the two tiles, with a 1e-4
// relative margin on the squared distance; the s <= r/4 bit also needs it
// above 4 s.
__device__ __forceinline__ int classify(const DeviceView& v, const SlotConst& S, const PairItem& it) {
  const int tI = S.tile0 + ((it.i0 - S.a0) >> 6);
  const int tJ = S.tile0 + ((it.j0 - S.a0) >> 6);
  const float4 li = v.tbox[2 * tI], hi = v.tbox[2 * tI + 1];
  const float4 lj = v.tbox[2 * tJ], hj = v.tbox[2 * tJ + 1];
  const float gx = fmaxf(0.0f, fmaxf(lj.x - hi.x, li.x - hj.x));
  const float gy = fmaxf(0.0f, fmaxf(lj.y - hi.y, li.y - hj.y));
  const float gz = fmaxf(0.0f, fmaxf(lj.z - hi.z, li.z - hj.z));
  const float ex = fmaxf(hj.x - li.x, hi.x - lj.x);
  const float ey = fmaxf(hj.y - li.y, hi.y - lj.y);
  const float ez = fmaxf(hj.z - li.z, hi.z - lj.z);
  const float dmin2 = gx * gx + gy * gy + gz * gz;
  const float dmax2 = ex * ex + ey * ey + ez * ez;
  const float thr = fmaxf(fmaxf(hj.w + li.w, hi.w + lj.w), fmaxf(hi.w, hj.w) * (1.0f / PRX_OBC_FAR_XMAX));
  int cls = (tI != tJ && dmin2 > thr * thr * 1.0001f) ? kClsObcFar : 0;
  const float thr4 = fmaxf(hi.w, hj.w) * (1.0f / PRX_OBC_FAR4_XMAX);
  if (cls && dmin2 > thr4 * thr4 * 1.0001f) cls |= kClsObcFar4;
  if (S.cutoff2 > 0.0f) {
    if (dmin2 >= S.cutoff2 * 1.00001f)
      cls |= kClsFar;
    else if (!(dmax2 < S.cutoff2 * 0.99999f))
      cls |= kClsStraddle;
  }
  return cls;
}

// Class of a chunk item: every listed pair may straddle the cutoff; the OBC2
// far-field bit comes from the i tile's box and the box of the chunk's j
// atoms (a wave reduction over the lanes' atoms), as for a tile pair.
__device__ __forceinline__ int classify_chunk(const DeviceView& v, const PairItem& it, const JSrc& js, int lane) {
  const float4 p = v.pos[js.gj], b = v.born4[js.gj];
  float lx = p.x, ly = p.y, lz = p.z, hx = p.x, hy = p.y, hz = p.z, mr = b.x, ms = b.y;
  for (int off = 32; off > 0; off >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, off));
    ly = fminf(ly, __shfl_xor(ly, off));
    lz = fminf(lz, __shfl_xor(lz, off));
    hx = fmaxf(hx, __shfl_xor(hx, off));
    hy = fmaxf(hy, __shfl_xor(hy, off));
    hz = fmaxf(hz, __shfl_xor(hz, off));
    mr = fmaxf(mr, __shfl_xor(mr, off));
    ms = fmaxf(ms, __shfl_xor(ms, off));
  }
  const int tI = it.excl;
  const float4 li = v.tbox[2 * tI], hi = v.tbox[2 * tI + 1];
  const float gx = fmaxf(0.0f, fmaxf(lx - hi.x, li.x - hx));
  const float gy = fmaxf(0.0f, fmaxf(ly - hi.y, li.y - hy));
  const float gz = fmaxf(0.0f, fmaxf(lz - hi.z, li.z - hz));
  const float dmin2 = gx * gx + gy * gy + gz * gz;
  const float thr = fmaxf(fmaxf(ms + li.w, hi.w + mr), fmaxf(hi.w, ms) * (1.0f / PRX_OBC_FAR_XMAX));
  return kClsStraddle | (dmin2 > thr * thr * 1.0001f ? kClsObcFar : 0);
}

// A system at its step limit, in a step graph (DeviceView::skip_done).
__device__ __forceinline__ bool skip_item(const DeviceView& v, const SlotConst& S, const PairItem& it) {
  return v.skip_done && S.step_limit != 0u && slot_done(S, v.atom_step[it.i0]);
}

// The Born pass of work item w, by one wave.
__device__ __forceinline__ void born_item(const DeviceView& v, int w, BornLds& L) {
  const PairItem it = v.items[w];
  const int lane = threadIdx.x & 63;
  const int fl = item_flags(it.meta);
  const SlotConst& S = v.slots[item_slot(it.meta)];
  if (skip_item(v, S, it)) return;
  const JSrc js = j_source(v, it, lane);
  if (item_is_chunk(it)) {  // neighbour-list chunk: always the masked path with the cutoff test
    if (js.nj == 0) {
      if (lane == 0) v.item_cls[w] = kClsFar;
      return;
    }
    const int cls = classify_chunk(v, it, js, lane);
    if (lane == 0) v.item_cls[w] = cls;
    if (cls & kClsObcFar)
      born_tile<true, false, true>(v, it, js, lane, w, S.cutoff2, L);
    else
      born_tile<true, false, false>(v, it, js, lane, w, S.cutoff2, L);
    return;
  }
  const bool full = item_ni(it.meta) == kTile && js.nj == kTile;
  const int cls = classify(v, S, it);
  if (lane == 0) v.item_cls[w] = cls;
  const int ccut = cls & kClsCutMask;
  if (S.gb_cut && ccut == kClsFar) return;  // no partials: the gathers skip this item
  const float cut2 = (S.gb_cut && ccut == kClsStraddle) ? S.cutoff2 : 0.0f;
  if (fl & kFlagDiag) {
    born_tile<true, true, false>(v, it, js, lane, w, cut2, L);
  } else if (cls & kClsObcFar4) {
    if (full && cut2 == 0.0f)
      born_tile<false, false, true, true>(v, it, js, lane, w, 0.0f, L);
    else
      born_tile<true, false, true, true>(v, it, js, lane, w, cut2, L);
  } else if (cls & kClsObcFar) {
    if (full && cut2 == 0.0f)
      born_tile<false, false, true>(v, it, js, lane, w, 0.0f, L);
    else
      born_tile<true, false, true>(v, it, js, lane, w, cut2, L);
  } else if (full && cut2 == 0.0f) {
    born_tile<false, false, false>(v, it, js, lane, w, 0.0f, L);
  } else {
    born_tile<true, false, false>(v, it, js, lane, w, cut2, L);
  }
}

// ---------------------------------------------------------------------------
// Pass 2: GB + Coulomb + LJ
// ---------------------------------------------------------------------------

// LJC = false: GB only (items entirely beyond the LJ/Coulomb cutoff).
// GENERAL tiles test r^2 < cut2 per pair when cut2 > 0, on LJ/Coulomb and, if
// gb_cut, on GB too.
struct PairLds {  // two float4 of j data per entry; j accumulators
  float4 a[128];
  float4 b[128];
#if PRX_LDS_ACC
  float4 acc[128];
#endif
};

template <bool GENERAL, bool DIAG, bool ENERGY, bool LJC>
__device__ __forceinline__ void force_tile(const DeviceView& v, const PairItem& it, const JSrc& js, int lane, int w,
                                           float cut2, bool gb_cut, PairLds& L) {
  const int ni = item_ni(it.meta), nj = js.nj;
  const SlotConst& S = v.slots[item_slot(it.meta)];
  const float kc = S.k_coulomb, gb2 = S.gb2;
  const bool has_cut = GENERAL && cut2 > 0.0f;
  const bool gb_has_cut = has_cut && gb_cut;
  const bool has_excl = GENERAL && LJC && !item_is_chunk(it) && it.excl >= 0;  // a chunk's excl is its row tile
  const int gi = it.i0 + min(lane, ni - 1);
  const int gj = js.gj;
  const float4 pi = v.pos[gi], nbi = v.nb4[gi], gbi = v.gb4[gi];
  const float q_kc = kc * nbi.x, q_gb = gb2 * nbi.x;
  {
    const float4 pj = v.pos[gj], nbj = v.nb4[gj], gbj = v.gb4[gj];
    const float4 a = make_float4(pj.x, pj.y, pj.z, nbj.x);    // x, y, z, q
    const float4 b = make_float4(nbj.y, nbj.z, gbj.x, gbj.y);  // sigma/2, sqrt(eps), B, a
    L.a[lane] = a; L.a[lane + 64] = a;
    L.b[lane] = b; L.b[lane + 64] = b;
#if PRX_LDS_ACC
    L.acc[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    L.acc[lane + 64] = make_float4(0.f, 0.f, 0.f, 0.f);
#endif
  }
  wave_lds_sync();
  float fxi = 0.f, fyi = 0.f, fzi = 0.f, dbi = 0.f;
  float fxj = 0.f, fyj = 0.f, fzj = 0.f, dbj = 0.f;
  double egb = 0.0, ecl = 0.0, elj = 0.0;
  unsigned long long m0 = 0, m1 = 0;
  if (has_excl) {
    m0 = v.excl_masks[(size_t)it.excl * 128 + lane * 2];
    m1 = v.excl_masks[(size_t)it.excl * 128 + lane * 2 + 1];
  }
  const bool ivalid = lane < ni;
  constexpr int kbeg = DIAG ? 1 : 0, kend = DIAG ? 32 : 63;
  PRX_UNROLL(PRX_UNROLL_FORCE)
  for (int k = kbeg; k <= kend; ++k) {
    const float4 ja = L.a[lane + k];
    const float4 jb = L.b[lane + k];
    const float xj = ja.x, yj = ja.y, zj = ja.z, qj = ja.w, hsj = jb.x, sej = jb.y, Bj = jb.z, aj = jb.w;
    const float dx = xj - pi.x, dy = yj - pi.y, dz = zj - pi.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float rinv = frsq(r2);
    const float rinv2 = rinv * rinv;
    // Coulomb + LJ (q_kc = k_coulomb q_i, q_gb = gb2 q_i are per-lane constants)
    float ec = 0.0f, el = 0.0f, fl = 0.0f;
    if (LJC) {
      ec = (q_kc * qj) * rinv;
      const float sg = nbi.y + hsj;
      const float ep = nbi.z * sej;
      const float s2 = sg * sg * rinv2;
      const float s6 = s2 * s2 * s2;
      el = 4.0f * ep * (s6 * s6 - s6);
      fl = ep * s6 * fmaf(48.0f, s6, -24.0f);
    }
    // GB over every pair
    const float BB = gbi.x * Bj;
    const float x = r2 * (gbi.y * aj);  // r^2 log2(e) / (4 B_i B_j)
    const float ex = fexp2(-x);
    const float f2 = fmaf(BB, ex, r2);
    const float finv = frsq(f2);
    const float eg = (q_gb * qj) * finv;
    const float t = eg * (finv * finv);                // gb2 qq / f^3
    const float te = t * ex;
    float fgb = fmaf(-0.25f, te, t);                   // -dE/dr / r
    float g = te * fmaf(x, kLn2, 1.0f);                // dE/dB_i = -g B_j / 2, dE/dB_j = -g B_i / 2
    if (GENERAL) {
      const int orig = (lane + k) & 63;
      bool valid = ivalid && orig < nj;
      if (DIAG && k == 32) valid = valid && lane < 32;
      float sv = 1.0f, se = 1.0f;
      if (has_excl) {
        const int cls = (int)((m0 >> orig) & 1ull) | ((int)((m1 >> orig) & 1ull) << 1);
        sv = cls == 0 ? 1.0f : (cls == 1 ? S.excl_v[1] : (cls == 2 ? S.excl_v[2] : S.excl_v[3]));
        se = cls == 0 ? 1.0f : (cls == 1 ? S.excl_e[1] : (cls == 2 ? S.excl_e[2] : S.excl_e[3]));
      }
      bool clv = valid && (sv != 0.0f || se != 0.0f);
      if (has_cut) clv = clv && r2 < cut2;
      const bool gv = gb_has_cut ? (valid && r2 < cut2) : valid;
      ec = clv ? ec * se : 0.0f;
      el = clv ? el * sv : 0.0f;
      fl = clv ? fl * sv : 0.0f;
      fgb = gv ? fgb : 0.0f;
      g = gv ? g : 0.0f;
      if (ENERGY) egb += gv ? (double)eg : 0.0;
    } else {
      if (ENERGY) egb += (double)eg;
    }
    if (ENERGY && LJC) { ecl += (double)ec; elj += (double)el; }
    float fr = LJC ? fmaf(ec + fl, rinv2, fgb) : fgb;
    if (GENERAL) {
      // An idle lane of a ragged tile may pair an atom with itself (r = 0):
      // 0 * inf there is NaN, so the final scalar is selected, not just its parts.
      const int orig = (lane + k) & 63;
      bool valid = ivalid && orig < nj;
      if (DIAG && k == 32) valid = valid && lane < 32;
      fr = valid ? fr : 0.0f;
    }
    // F_i = -fr * (x_j - x_i)
    fxi = fmaf(-fr, dx, fxi); fyi = fmaf(-fr, dy, fyi); fzi = fmaf(-fr, dz, fzi);
    dbi = fmaf(g, Bj, dbi);
#if PRX_LDS_ACC
    {
      volatile v4f* acc = reinterpret_cast<volatile v4f*>(L.acc);
      v4f t = acc[lane + k];
      t.x = fmaf(fr, dx, t.x); t.y = fmaf(fr, dy, t.y); t.z = fmaf(fr, dz, t.z); t.w = fmaf(g, gbi.x, t.w);
      acc[lane + k] = t;
    }
#else
    fxj = fmaf(fr, dx, fxj); fyj = fmaf(fr, dy, fyj); fzj = fmaf(fr, dz, fzj);
    dbj = fmaf(g, gbi.x, dbj);
    if (k != kend) {
      fxj = rot<PRX_ROT_BPERM_FORCE>(fxj); fyj = rot<PRX_ROT_BPERM_FORCE>(fyj);
      fzj = rot<PRX_ROT_BPERM_FORCE>(fzj); dbj = rot<PRX_ROT_BPERM_FORCE>(dbj);
    }
#endif
  }
  v.part[part_slot(w, 0, lane)] = make_float4(fxi, fyi, fzi, -0.5f * dbi);
#if PRX_LDS_ACC
  wave_lds_sync();
  {
    const float4 a0 = L.acc[lane], a1 = L.acc[lane + 64];
    v.part[part_slot(w, 1, lane)] = make_float4(a0.x + a1.x, a0.y + a1.y, a0.z + a1.z, -0.5f * (a0.w + a1.w));
  }
#else
  const int home = (lane + kend) & 63;
  v.part[part_slot(w, 1, home)] = make_float4(fxj, fyj, fzj, -0.5f * dbj);
#endif
  if (ENERGY) {
    for (int off = 32; off > 0; off >>= 1) {
      egb += __shfl_xor(egb, off);
      ecl += __shfl_xor(ecl, off);
      elj += __shfl_xor(elj, off);
    }
    if (lane == 0) {
      v.e_items[(size_t)w * 3 + 0] = egb;
      v.e_items[(size_t)w * 3 + 1] = ecl;
      v.e_items[(size_t)w * 3 + 2] = elj;
    }
  }
}

// The force pass of work item w, by one wave.
template <bool ENERGY>
__device__ __forceinline__ void force_item(const DeviceView& v, int w, PairLds& L) {
  const PairItem it = v.items[w];
  const int lane = threadIdx.x & 63;
  const int fl = item_flags(it.meta);
  const SlotConst& S = v.slots[item_slot(it.meta)];
  if (!ENERGY && skip_item(v, S, it)) return;
  const int cls = S.cutoff2 > 0.0f ? (v.item_cls[w] & kClsCutMask) : kClsNear;
  if (cls == kClsFar) {
    if (S.gb_cut) {
      if (ENERGY && lane == 0) {
        v.e_items[(size_t)w * 3 + 0] = 0.0;
        v.e_items[(size_t)w * 3 + 1] = 0.0;
        v.e_items[(size_t)w * 3 + 2] = 0.0;
      }
      return;
    }
    // GB over every pair, nothing else (a far item is never diagonal)
    const JSrc js = j_source(v, it, lane);
    if (item_ni(it.meta) == kTile && js.nj == kTile)
      force_tile<false, false, ENERGY, false>(v, it, js, lane, w, 0.0f, false, L);
    else
      force_tile<true, false, ENERGY, false>(v, it, js, lane, w, 0.0f, false, L);
    return;
  }
  const JSrc js = j_source(v, it, lane);
  const bool full = item_ni(it.meta) == kTile && js.nj == kTile;
  const float cut2 = cls == kClsStraddle ? S.cutoff2 : 0.0f;
  const bool gb_cut = S.gb_cut != 0;
  if (fl & kFlagDiag) {
    force_tile<true, true, ENERGY, true>(v, it, js, lane, w, cut2, gb_cut, L);
  } else if (full && !(fl & kFlagExcl) && cut2 == 0.0f) {
    force_tile<false, false, ENERGY, true>(v, it, js, lane, w, 0.0f, false, L);
  } else {
    force_tile<true, false, ENERGY, true>(v, it, js, lane, w, cut2, gb_cut, L);
  }
}

// ---------------------------------------------------------------------------
// Pass 3: chain rule through the Born radii
// ---------------------------------------------------------------------------

template <bool GENERAL, bool DIAG, bool FAR, bool F4 = false>
__device__ __forceinline__ void chain_tile(const DeviceView& v, const PairItem& it, const JSrc& js, int lane, int w,
                                           float cut2, PairLds& L) {
  const int ni = item_ni(it.meta), nj = js.nj;
  const int gi = it.i0 + min(lane, ni - 1);
  const int gj = js.gj;
  const float4 pi = v.pos[gi], bi = v.born4[gi];
  const float ci = v.gb4[gi].w;
  {
    const float4 pj = v.pos[gj], bj = v.born4[gj];
    const float4 a = make_float4(pj.x, pj.y, pj.z, bj.z);          // x, y, z, s^2
    const float4 b = make_float4(bj.w, v.gb4[gj].w, bj.x, bj.y);     // s^3/2, c, rho, s
    L.a[lane] = a; L.a[lane + 64] = a;
    L.b[lane] = b; L.b[lane + 64] = b;
#if PRX_LDS_ACC
    L.acc[lane] = make_float4(0.f, 0.f, 0.f, 0.f);
    L.acc[lane + 64] = make_float4(0.f, 0.f, 0.f, 0.f);
#endif
  }
  wave_lds_sync();
  float fxi = 0.f, fyi = 0.f, fzi = 0.f, fxj = 0.f, fyj = 0.f, fzj = 0.f;
  const bool ivalid = lane < ni;
  constexpr int kbeg = DIAG ? 1 : 0, kend = DIAG ? 32 : 63;
  PRX_UNROLL(PRX_UNROLL_CHAIN)
  for (int k = kbeg; k <= kend; ++k) {
    const float4 ja = L.a[lane + k];
    const float2 jb = *reinterpret_cast<const float2*>(&L.b[lane + k]);  // the far field needs only these
    const float xj = ja.x, yj = ja.y, zj = ja.z, sj2 = ja.w, sj3h = jb.x, cj = jb.y;
    const float dx = xj - pi.x, dy = yj - pi.y, dz = zj - pi.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    float gsc;  // (c_i dI(r; rho_i, s_j)/dr + c_j dI(r; rho_j, s_i)/dr) / r; F_i = gsc (x_j - x_i)
    if (FAR) {
      gsc = chain_far<F4>(frcp(r2), sj2, ci * sj3h, bi.z, cj * bi.w);
    } else {
      const float rinv = frsq(r2);
      const float rhoj = L.b[lane + k].z, sj = L.b[lane + k].w;  // near field only
      const float r = r2 * rinv;
      const bool far1 = obc_far(r, bi.x, sj), far2 = obc_far(r, rhoj, bi.y);
      const float u = rinv * rinv;
      if (__all(far1 && far2)) {
        gsc = chain_far(u, sj2, ci * sj3h, bi.z, cj * bi.w);
      } else {
        const float u2r = u * u * rinv, rs = fminf(rinv, 1e4f);
        const float d1 = far1 ? obc_far_dI(sj2, sj3h, u, u2r) : born_dr(r, rs, bi.x, sj, sj2);
        const float d2 = far2 ? obc_far_dI(bi.z, bi.w, u, u2r) : born_dr(r, rs, rhoj, bi.y, bi.z);
        gsc = fmaf(ci, d1, cj * d2) * rinv;
      }
    }
    if (GENERAL) {
      const int orig = (lane + k) & 63;
      bool valid = ivalid && orig < nj;
      if (DIAG && k == 32) valid = valid && lane < 32;
      if (cut2 > 0.0f) valid = valid && r2 < cut2;
      gsc = valid ? gsc : 0.0f;
    }
    fxi = fmaf(gsc, dx, fxi); fyi = fmaf(gsc, dy, fyi); fzi = fmaf(gsc, dz, fzi);
#if PRX_LDS_ACC
    {
      volatile v4f* acc = reinterpret_cast<volatile v4f*>(L.acc);
      v4f t = acc[lane + k];
      t.x = fmaf(-gsc, dx, t.x); t.y = fmaf(-gsc, dy, t.y); t.z = fmaf(-gsc, dz, t.z);
      acc[lane + k] = t;
    }
#else
    fxj = fmaf(-gsc, dx, fxj); fyj = fmaf(-gsc, dy, fyj); fzj = fmaf(-gsc, dz, fzj);
    if (k != kend) {
      // PRX_ROT_BPERM_CHAIN 2 or 3: the first one or two by the crossbar, the rest by DPP
      fxj = rot<(PRX_ROT_BPERM_CHAIN >= 1)>(fxj);
      fyj = rot<(PRX_ROT_BPERM_CHAIN == 1 || PRX_ROT_BPERM_CHAIN == 3)>(fyj);
      fzj = rot<(PRX_ROT_BPERM_CHAIN == 1)>(fzj);
    }
#endif
  }
  put_part3(v, w, 0, lane, fxi, fyi, fzi);
#if PRX_LDS_ACC
  wave_lds_sync();
  {
    const float4 a0 = L.acc[lane], a1 = L.acc[lane + 64];
    put_part3(v, w, 1, lane, a0.x + a1.x, a0.y + a1.y, a0.z + a1.z);
  }
#else
  const int home = (lane + kend) & 63;
  put_part3(v, w, 1, home, fxj, fyj, fzj);
#endif
}

// The chain pass of work item w, by one wave.
__device__ __forceinline__ void chain_item(const DeviceView& v, int w, PairLds& L) {
  const PairItem it = v.items[w];
  const int lane = threadIdx.x & 63;
  const int fl = item_flags(it.meta);
  const SlotConst& S = v.slots[item_slot(it.meta)];
  if (skip_item(v, S, it)) return;
  const int cls = v.item_cls[w];
  float cut2 = 0.0f;
  if (S.gb_cut) {
    if ((cls & kClsCutMask) == kClsFar) return;
    if ((cls & kClsCutMask) == kClsStraddle) cut2 = S.cutoff2;
  }
  const JSrc js = j_source(v, it, lane);
  const bool full = item_ni(it.meta) == kTile && js.nj == kTile;
  if (fl & kFlagDiag) {
    chain_tile<true, true, false>(v, it, js, lane, w, cut2, L);
  } else if (cls & kClsObcFar4) {
    if (full && cut2 == 0.0f)
      chain_tile<false, false, true, true>(v, it, js, lane, w, 0.0f, L);
    else
      chain_tile<true, false, true, true>(v, it, js, lane, w, cut2, L);
  } else if (cls & kClsObcFar) {
    if (full && cut2 == 0.0f)
      chain_tile<false, false, true>(v, it, js, lane, w, 0.0f, L);
    else
      chain_tile<true, false, true>(v, it, js, lane, w, cut2, L);
  } else if (full && cut2 == 0.0f) {
    chain_tile<false, false, false>(v, it, js, lane, w, 0.0f, L);
  } else {
    chain_tile<true, false, false>(v, it, js, lane, w, cut2, L);
  }
}


// One wave's LDS slice in a kernel that runs every pass.
union PassLds {
  BornLds born;
  PairLds pair;
};

}  // namespace prx
