// common.h -- device data layout and helpers shared by the gfx950 kernels.
//
// HBM layout (built by api.hip):
//   receptor  RecPos[P]  {x, y, z fixed point (int32), q (float bits)}   16 B
//             RecAB[P]   {A = sqrt(D) x^6, B = sqrt(D) x^3}               8 B
//   ligands   LigInfo[nlig], LigAtom[sum n] {ke*q, A, B, grid type} (FP64),
//             LigXYZ[sum n] centred conformer (FP64), intra pair list per ligand
//             (packed i | j << 16) with FP64 pair products.
//   poses     FP64 (x, y, z) rows, pose-major, atom_off[p] = first row of pose p.
//
// UFF pair parameters factorise: x_ij = sqrt(x_i x_j), D_ij = sqrt(D_i D_j)
// (src/UFF.ml:39-51), so D_ij (x_ij^12/r^12 - 2 x_ij^6/r^6) = A_i A_j u^6 - 2 B_i B_j u^3
// with u = 1/r^2, A = sqrt(D) x^6, B = sqrt(D) x^3.  A lane owning ligand atom j can
// therefore accumulate receptor-side sums  Sq = sum w q_i / r, S12 = sum w A_i u^6,
// S6 = sum w B_i u^3  and apply j's parameters once per chunk:
//   E_j = ke q_j Sq + A_j S12 - 2 B_j S6.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mmo {

constexpr int kWave = 64;
constexpr double kElecWeight = 332.0637 / 4.0;   // src/UFF.ml:25
constexpr float kCutoff2 = 144.0f;                // src/mol.ml:836
constexpr float kMinR2 = 1.0e-4f;                 // Math.non_zero_dist 0.01 A, src/math.ml:58-62

struct RecPos { int x, y, z; float q; };
struct RecAB { float a, b; };
// quantisation residual of a receptor atom: (x - o) - X 2^-k, in A
struct RecRes { float x, y, z; };

struct LigInfo {
  int n_atoms;
  int atom_base;   // into LigAtom / LigXYZ
  int pair_base;   // into the intra pair list
  int n_pairs;
  int nrb;         // rotatable bonds
  int rb_base;     // into rb_left/rb_right/rg_off (rg_off has nrb+1 entries at rb_base + lig)
  int rg_base;     // first rg_atoms entry of this ligand
  int dof_n;       // 6 + nrb
};

struct LigAtom { double keq, a, b; int type; int anum; };

// fixed-point frame: ix = rint((x - o) * 2^k), saturated to +-lim
struct Frame {
  double ox, oy, oz;
  double scale;      // 2^k
  float inv_s2;      // 2^-2k
  float inv_s;       // 2^-k
  int lim;           // 2^30 - 1
};

// Receptor cell list for culling: atoms sorted by cell (z, y, x order, x fastest);
// cell_start[c]..cell_start[c+1] are the atoms of cell c.  Coordinates in A relative
// to the fixed-point frame origin.
// receptor atoms fetched at once by the rigid scan's row sweep; its receptor arrays carry
// this many inert atoms of padding so that a fetch may run past the last atom
constexpr int kRowChunk = 8;

struct CellGrid {
  const int* cell_start;   // ncell + 1
  float x0, y0, z0;        // corner of cell (0,0,0)
  float c, inv_c;          // cell edge along y and z
  float cx, inv_cx;        // cell edge along x (c / n for the pose sweeps' grid)
  int nx, ny, nz;
  int clip_min;            // rows longer than this are clipped to the sphere's x chord
};

// A pose tile: npose poses of ligand `lig` starting at pose0 (poses of one ligand are
// contiguous after planning); each pose is handled by `split` lanes.
struct Tile { int lig, pose0, npose, split; };

// non-finite coordinates saturate like far-away ones (beyond every cutoff)
__device__ __forceinline__ int to_fixed(double v, double o, double s, int lim, int* sat) {
  double t = (v - o) * s;
  const double dl = (double)lim;
  if (!(t >= -dl && t <= dl)) {   // also catches NaN
    *sat = 1;
    t = (t != t) ? dl : fmin(fmax(t, -dl), dl);
  }
  return __double2int_rn(t);
}

// wave-wide sums over groups of `split` consecutive lanes (split a power of two)
__device__ __forceinline__ double group_sum(double v, int split) {
  for (int off = 1; off < split; off <<= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ int group_or(int v, int split) {
  for (int off = 1; off < split; off <<= 1) v |= __shfl_xor(v, off, kWave);
  return v;
}
// Wavefront scans and sums through DPP (row_shr within rows of 16 lanes, then row_bcast
// across rows): six VALU operations instead of six LDS permutes.  DPP controls (gfx9):
// row_shr:n = 0x110 + n, row_bcast:15 = 0x142, row_bcast:31 = 0x143; lanes without a
// source read 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xf, false));
}
// inclusive prefix sum over the 64 lanes (all lanes active)
__device__ __forceinline__ int wave_incl_scan(int v) {
  v += dpp_i<0x111>(v);
  v += dpp_i<0x112>(v);
  v += dpp_i<0x114>(v);
  v += dpp_i<0x118>(v);
  v += dpp_i<0x142, 0xa>(v);
  v += dpp_i<0x143, 0xc>(v);
  return v;
}
// sum over the 64 lanes (all lanes active), uniform result
__device__ __forceinline__ float wave_sum_f(float v) {
  v += dpp_f<0x111>(v);
  v += dpp_f<0x112>(v);
  v += dpp_f<0x114>(v);
  v += dpp_f<0x118>(v);
  v += dpp_f<0x142, 0xa>(v);
  v += dpp_f<0x143, 0xc>(v);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// FP64 through DPP: both 32-bit halves move under the same control; lanes without a
// source (or outside ROW_MASK) receive `old`
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_d(double v, double old) {
  const long long b = __double_as_longlong(v), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)b, CTRL, ROW_MASK, 0xf, false);
  const int hi =
      __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(b >> 32), CTRL, ROW_MASK, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum / max over the 64 lanes (all lanes active), uniform result: the row scan of
// wave_sum_f in FP64, read from lane 63
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_d<0x111>(v, 0.0);
  v += dpp_d<0x112>(v, 0.0);
  v += dpp_d<0x114>(v, 0.0);
  v += dpp_d<0x118>(v, 0.0);
  v += dpp_d<0x142, 0xa>(v, 0.0);
  v += dpp_d<0x143, 0xc>(v, 0.0);
  return readlane_d(v, 63);
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp_d<0x111>(v, v));
  v = fmax(v, dpp_d<0x112>(v, v));
  v = fmax(v, dpp_d<0x114>(v, v));
  v = fmax(v, dpp_d<0x118>(v, v));
  v = fmax(v, dpp_d<0x142, 0xa>(v, v));
  v = fmax(v, dpp_d<0x143, 0xc>(v, v));
  return readlane_d(v, 63);
}

// orders this wavefront's LDS accesses before the barrier against those after it (LDS
// that only the wavefront itself reads and writes needs no workgroup barrier)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Square root for the culling bounds: the hardware instruction (within 1 ulp) instead
// of the correctly rounded expansion of sqrtf; every bound it feeds carries a margin of
// 0.05 A beyond the 12 A cutoff, so no pair inside the cutoff is culled.
__device__ __forceinline__ float bound_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// t = clamp(1 - r2/144, 0, 1): the cutoff test r2 < 144 and the shift
// S(r) = (1 - (r/12)^2)^2 (src/FF.ml:17-20) in one clamped FMA.
__device__ __forceinline__ float shift_t(float r2c) {
  return __builtin_amdgcn_fmed3f(fmaf(r2c, -1.0f / 144.0f, 1.0f), 0.0f, 1.0f);
}

// Pair-loop constants kept in VGPRs.  On gfx950 a VALU instruction that reads an SGPR
// issues at a quarter of the rate of an all-VGPR one and a literal operand also costs
// issue rate (measured with single-form loops); the pair loops are bound by VALU issue,
// so the constants they use on every pair live in VGPRs.
__device__ __forceinline__ float in_vgpr(float x) {
  asm volatile("" : "+v"(x));
  return x;
}
struct PairK {
  float inv_s2;   // 2^-2k: fixed-point r^2 to A^2
  float min_r2;   // clamp r >= 0.01 A
  float m144;     // -1/144: shift t = 1 - r^2/144
  float near2;    // pairs closer than this (A^2) add the coordinates' quantisation residuals
  float inv_s;    // 2^-k
  const RecRes* rres;   // receptor residuals (global memory, read only for near pairs)
};
__device__ __forceinline__ PairK pair_k(float inv_s2, float inv_s, const RecRes* rres) {
  return {in_vgpr(inv_s2), in_vgpr(kMinR2), in_vgpr(-1.0f / 144.0f), in_vgpr(0.25f), inv_s, rres};
}

// A ligand atom: fixed-point position and its quantisation residual (A).  The 2^-k grid
// (6e-8 A at k = 24) is exact enough for the differences of atoms more than 0.5 A apart;
// closer pairs, whose r^-12 term multiplies a relative error of r by 12, also use the
// residuals.
struct LigPt {
  int x, y, z;
  float rx, ry, rz;
};
__device__ __forceinline__ float quant_residual(double v, double o, double inv_scale, int q) {
  const float r = (float)((v - o) - (double)q * inv_scale);
  return fabsf(r) <= 1.0e-3f ? r : 0.0f;   // saturated or non-finite: far from every atom
}
__device__ __forceinline__ LigPt lig_pt(double x, double y, double z, const Frame& fr, int* sat) {
  LigPt l;
  l.x = to_fixed(x, fr.ox, fr.scale, fr.lim, sat);
  l.y = to_fixed(y, fr.oy, fr.scale, fr.lim, sat);
  l.z = to_fixed(z, fr.oz, fr.scale, fr.lim, sat);
  const double is = 1.0 / fr.scale;
  l.rx = quant_residual(x, fr.ox, is, l.x);
  l.ry = quant_residual(y, fr.oy, is, l.y);
  l.rz = quant_residual(z, fr.oz, is, l.z);
  return l;
}
__device__ __forceinline__ LigPt lig_pt_zero() { return LigPt{0, 0, 0, 0.0f, 0.0f, 0.0f}; }
// an inert lane's atom for the sweeps that skip the 0.01 A clamp on far pairs: at the
// saturation limit on every axis, beyond the cutoff of every receptor atom (the frame
// covers the receptor plus 16 A), so its pairs have finite r and weight exactly 0
__device__ __forceinline__ LigPt lig_pt_far() {
  constexpr int f = (1 << 30) - 1;
  return LigPt{f, f, f, 0.0f, 0.0f, 0.0f};
}

// r^2 (A^2) of receptor atom i and ligand atom l from fixed-point differences d; for
// pairs closer than sqrt(near2) the residuals of both atoms are added (a wavefront
// branch that is rarely taken)
__device__ __forceinline__ float pair_r2(float dx, float dy, float dz, const LigPt& l, int i,
                                         const PairK& pk) {
  float r2 = (dx * dx + dy * dy + dz * dz) * pk.inv_s2;
  if (__builtin_expect(__ballot(r2 < pk.near2) != 0ull, 0)) {
    if (r2 < pk.near2) {
      // 32-bit byte offset from the (wave-uniform) base (a 24-bit multiply, full rate),
      // formed here: the empty asm keeps the compiler from hoisting it onto every pair
      int ii = i;
      asm volatile("" : "+v"(ii));
      const RecRes rr = *reinterpret_cast<const RecRes*>(
          reinterpret_cast<const char*>(pk.rres) + __umul24((uint32_t)ii, (uint32_t)sizeof(RecRes)));
      const float ex = fmaf(dx, pk.inv_s, rr.x - l.rx);
      const float ey = fmaf(dy, pk.inv_s, rr.y - l.ry);
      const float ez = fmaf(dz, pk.inv_s, rr.z - l.rz);
      r2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
    }
  }
  return r2;
}
// max(r2, min_r2) on the bit patterns: both are non-negative finite floats, for which
// integer order is float order (and the float max would canonicalise its operands)
__device__ __forceinline__ float max_r2(float r2, float min_r2) {
  return __int_as_float(max(__float_as_int(r2), __float_as_int(min_r2)));
}

// pair_r2 followed by the 0.01 A clamp (max_r2).  A wavefront none of whose pairs is
// closer than sqrt(near2) skips both: its r2 are above the clamp already, so the result
// is the same and the clamp costs nothing outside the rare branch.
__device__ __forceinline__ float pair_r2_clamped(float dx, float dy, float dz, const LigPt& l,
                                                 int i, const PairK& pk) {
  float r2 = (dx * dx + dy * dy + dz * dz) * pk.inv_s2;
  if (__builtin_expect(__ballot(r2 < pk.near2) != 0ull, 0)) {
    if (r2 < pk.near2) {
      int ii = i;
      asm volatile("" : "+v"(ii));
      const RecRes rr = *reinterpret_cast<const RecRes*>(
          reinterpret_cast<const char*>(pk.rres) + __umul24((uint32_t)ii, (uint32_t)sizeof(RecRes)));
      const float ex = fmaf(dx, pk.inv_s, rr.x - l.rx);
      const float ey = fmaf(dy, pk.inv_s, rr.y - l.ry);
      const float ez = fmaf(dz, pk.inv_s, rr.z - l.rz);
      r2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
    }
    r2 = max_r2(r2, pk.min_r2);
  }
  return r2;
}

// 1/r^2 for the LJ powers.  A Newton step on u = ir^2 was measured to change nothing:
// over 8,192 poses of the rigid scan the largest error against FP64 stays 1.9e-6 with or
// without it (the error of clashing pairs came from the coordinate grid, see LigPt)
__device__ __forceinline__ float inv_r2(float ir, float /*r2c*/) { return ir * ir; }
__device__ __forceinline__ float shift_t(float r2c, const PairK& k) {
  return __builtin_amdgcn_fmed3f(fmaf(r2c, k.m144, 1.0f), 0.0f, 1.0f);
}

}  // namespace mmo
