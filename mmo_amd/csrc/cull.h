// cull.h -- pair arithmetic and the cell-list sweep shared by the scoring, grid and
// Monte Carlo kernels.
#pragma once
#include "common.h"
#include "mmo_ff.h"

namespace mmo {

// --------------------------------------------------------------- pair arithmetic
template <int MODE>
__device__ __forceinline__ void pair_energy(const RecPos& p, const RecAB& ab, int lx, int ly,
                                            int lz, float inv_s2, float& sq, float& s12,
                                            float& s6) {
  const float dx = (float)(p.x - lx), dy = (float)(p.y - ly), dz = (float)(p.z - lz);
  const float r2 = (dx * dx + dy * dy + dz * dz) * inv_s2;
  const float r2c = fmaxf(r2, kMinR2);
  const float ir = __builtin_amdgcn_rsqf(r2c);
  const float u = ir * ir;
  const float u3 = u * u * u;
  const float u6 = u3 * u3;
  if (MODE == MMO_SHIFTED) {
    const float t = shift_t(r2c);
    const float w = t * t;
    sq = fmaf(w * p.q, ir, sq);
    s12 = fmaf(w * ab.a, u6, s12);
    s6 = fmaf(w * ab.b, u3, s6);
  } else {
    sq = fmaf(p.q, ir, sq);
    s12 = fmaf(ab.a, u6, s12);
    s6 = fmaf(ab.b, u3, s6);
  }
}

// ------------------------------------------------------- inter kernel, culled
// Shifted form only.  The whole receptor sits in LDS sorted by 4 A cells.  For each
// block of K atoms per lane the wavefront takes the bounding box of its lanes' atom
// positions and visits only the cell rows that intersect the sphere of radius
// 12 A + half-diagonal around the box centre; every receptor atom within 12 A of any
// of the wavefront's atoms is in a visited cell, and pairs beyond 12 A that are
// visited contribute exactly 0 (the shift weight is clamped to 0).
__device__ __forceinline__ float wave_minf(float v) {
  for (int off = 1; off < kWave; off <<= 1) v = fminf(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
  for (int off = 1; off < kWave; off <<= 1) v = fmaxf(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ float uniformf(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int uniformi(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Visits every receptor atom within R of (qx, qy, qz) (plus some beyond it) through
// the LDS cell list, accumulating the receptor-side sums for K ligand atoms per lane;
// fold(sq, s12, s6) is called once per z slab of cells (FP32 partial sums).
template <int K, class Fold>
__device__ __forceinline__ void cull_sweep(const CellGrid& cg, const int* __restrict__ s_cell,
                                           const RecPos* __restrict__ s_pos,
                                           const RecAB* __restrict__ s_ab, float inv_s2,
                                           float qx, float qy, float qz, float R,
                                           const int (&lx)[K], const int (&ly)[K],
                                           const int (&lz)[K], Fold&& fold) {
  const float R2 = R * R;
  const int cz0 = max(0, (int)floorf((qz - R - cg.z0) * cg.inv_c));
  const int cz1 = min(cg.nz - 1, (int)floorf((qz + R - cg.z0) * cg.inv_c));
  for (int cz = cz0; cz <= cz1; ++cz) {
    const float zlo = cg.z0 + cz * cg.c;
    const float dz = fmaxf(0.0f, fmaxf(zlo - qz, qz - (zlo + cg.c)));
    const float rz2 = R2 - dz * dz;
    if (rz2 < 0.0f) continue;
    const float ry = sqrtf(rz2);
    const int cy0 = max(0, (int)floorf((qy - ry - cg.y0) * cg.inv_c));
    const int cy1 = min(cg.ny - 1, (int)floorf((qy + ry - cg.y0) * cg.inv_c));
    float sq[K], s12[K], s6[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sq[k] = s12[k] = s6[k] = 0.0f;
    for (int cy = cy0; cy <= cy1; ++cy) {
      const float ylo = cg.y0 + cy * cg.c;
      const float dy = fmaxf(0.0f, fmaxf(ylo - qy, qy - (ylo + cg.c)));
      const float rx2 = rz2 - dy * dy;
      if (rx2 < 0.0f) continue;
      const float rx = sqrtf(rx2);
      const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_c));
      const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_c));
      if (cx0 > cx1) continue;
      const int rowb = (cz * cg.ny + cy) * cg.nx;
      const int a0 = uniformi(s_cell[rowb + cx0]);
      const int a1 = uniformi(s_cell[rowb + cx1 + 1]);
#pragma unroll 2
      for (int i = a0; i < a1; ++i) {
        const RecPos p = s_pos[i];
        const RecAB ab = s_ab[i];
#pragma unroll
        for (int k = 0; k < K; ++k)
          pair_energy<MMO_SHIFTED>(p, ab, lx[k], ly[k], lz[k], inv_s2, sq[k], s12[k], s6[k]);
      }
    }
    fold(sq, s12, s6);
  }
}

// FP64 variant of the shifted sweep (Monte Carlo parity mode): exact fixed-point
// differences, FP64 pair arithmetic (sqrt, division) and FP64 sums; returns the sums
// already weighted by ligand parameters cq (ke q_j), ca (A_j), cb (B_j).
template <int K>
__device__ __forceinline__ double cull_sweep_f64(const CellGrid& cg, const int* __restrict__ s_cell,
                                                 const RecPos* __restrict__ s_pos,
                                                 const RecAB* __restrict__ s_ab, double inv_s,
                                                 float qx, float qy, float qz, float R,
                                                 const int (&lx)[K], const int (&ly)[K],
                                                 const int (&lz)[K], const double (&cq)[K],
                                                 const double (&ca)[K], const double (&cb)[K]) {
  double e = 0.0;
  const float R2 = R * R;
  const int cz0 = max(0, (int)floorf((qz - R - cg.z0) * cg.inv_c));
  const int cz1 = min(cg.nz - 1, (int)floorf((qz + R - cg.z0) * cg.inv_c));
  for (int cz = cz0; cz <= cz1; ++cz) {
    const float zlo = cg.z0 + cz * cg.c;
    const float dz = fmaxf(0.0f, fmaxf(zlo - qz, qz - (zlo + cg.c)));
    const float rz2 = R2 - dz * dz;
    if (rz2 < 0.0f) continue;
    const float ry = sqrtf(rz2);
    const int cy0 = max(0, (int)floorf((qy - ry - cg.y0) * cg.inv_c));
    const int cy1 = min(cg.ny - 1, (int)floorf((qy + ry - cg.y0) * cg.inv_c));
    for (int cy = cy0; cy <= cy1; ++cy) {
      const float ylo = cg.y0 + cy * cg.c;
      const float dy = fmaxf(0.0f, fmaxf(ylo - qy, qy - (ylo + cg.c)));
      const float rx2 = rz2 - dy * dy;
      if (rx2 < 0.0f) continue;
      const float rx = sqrtf(rx2);
      const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_c));
      const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_c));
      if (cx0 > cx1) continue;
      const int rowb = (cz * cg.ny + cy) * cg.nx;
      const int a0 = uniformi(s_cell[rowb + cx0]);
      const int a1 = uniformi(s_cell[rowb + cx1 + 1]);
      for (int i = a0; i < a1; ++i) {
        const RecPos p = s_pos[i];
        const RecAB ab = s_ab[i];
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const double dx = (double)(p.x - lx[k]) * inv_s, dy2 = (double)(p.y - ly[k]) * inv_s,
                       dz2 = (double)(p.z - lz[k]) * inv_s;
          const double r2 = dx * dx + dy2 * dy2 + dz2 * dz2;
          if (!(r2 < 144.0)) continue;
          const double rr = sqrt(r2);
          const double r = rr < 0.01 ? 0.01 : rr;     // Math.non_zero_dist
          const double t = 1.0 - (r / 12.0) * (r / 12.0);
          const double w = t * t;                      // FF.shift_12A
          const double ir = 1.0 / r, u = ir * ir, u3 = u * u * u;
          e += w * (cq[k] * (double)p.q * ir + ca[k] * (double)ab.a * u3 * u3 -
                    2.0 * cb[k] * (double)ab.b * u3);
        }
      }
    }
  }
  return e;
}

// whole receptor + cell table into LDS (all threads of the block); receptors too large
// for LDS (LDS = false) are read in place from global memory (L2), by the same
// uniform row ranges
template <bool LDS>
__device__ __forceinline__ void load_receptor(const RecPos* rpos, const RecAB* rab, int nrec,
                                              const CellGrid& cg, const RecPos*& s_pos,
                                              const RecAB*& s_ab, const int*& s_cell, char* smem) {
  if (!LDS) {
    s_pos = rpos;
    s_ab = rab;
    s_cell = cg.cell_start;
    return;
  }
  const int ncell = cg.nx * cg.ny * cg.nz;
  RecPos* p = reinterpret_cast<RecPos*>(smem);
  RecAB* b = reinterpret_cast<RecAB*>(smem + (size_t)nrec * sizeof(RecPos));
  int* c = reinterpret_cast<int*>(smem + (size_t)nrec * (sizeof(RecPos) + sizeof(RecAB)));
  for (int i = threadIdx.x; i < nrec; i += blockDim.x) {
    p[i] = rpos[i];
    b[i] = rab[i];
  }
  for (int i = threadIdx.x; i <= ncell; i += blockDim.x) c[i] = cg.cell_start[i];
  __syncthreads();
  s_pos = p;
  s_ab = b;
  s_cell = c;
}

// bounding box of the wavefront's valid fixed-point positions -> query centre and
// radius 12 A + half-diagonal (+0.05 A for FP32 rounding), uniform across the wave
template <int K>
__device__ __forceinline__ bool wave_query(const int (&lx)[K], const int (&ly)[K],
                                           const int (&lz)[K], const bool (&valid)[K],
                                           float inv_s, float& qx, float& qy, float& qz,
                                           float& R) {
  float mnx = 3.0e38f, mny = 3.0e38f, mnz = 3.0e38f;
  float mxx = -3.0e38f, mxy = -3.0e38f, mxz = -3.0e38f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (valid[k]) {
      const float fx = (float)lx[k] * inv_s, fy = (float)ly[k] * inv_s, fz = (float)lz[k] * inv_s;
      mnx = fminf(mnx, fx); mny = fminf(mny, fy); mnz = fminf(mnz, fz);
      mxx = fmaxf(mxx, fx); mxy = fmaxf(mxy, fy); mxz = fmaxf(mxz, fz);
    }
  }
  mnx = uniformf(wave_minf(mnx)); mny = uniformf(wave_minf(mny)); mnz = uniformf(wave_minf(mnz));
  mxx = uniformf(wave_maxf(mxx)); mxy = uniformf(wave_maxf(mxy)); mxz = uniformf(wave_maxf(mxz));
  if (!(mnx <= mxx)) return false;
  qx = 0.5f * (mnx + mxx); qy = 0.5f * (mny + mxy); qz = 0.5f * (mnz + mxz);
  const float hx = 0.5f * (mxx - mnx), hy = 0.5f * (mxy - mny), hz = 0.5f * (mxz - mnz);
  R = uniformf(12.0f + sqrtf(hx * hx + hy * hy + hz * hz) + 0.05f);
  return true;
}

}  // namespace mmo
