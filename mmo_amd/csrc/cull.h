// cull.h -- pair arithmetic and the cell-list sweep shared by the scoring, grid and
// Monte Carlo kernels.
#pragma once
#include "common.h"
#include "mmo_ff.h"

namespace mmo {

// --------------------------------------------------------------- pair arithmetic
// NEAR = false: the caller knows the pair is at least 0.5 A apart for every lane whose
// sums it keeps, and that an inert lane's atom lies beyond the cutoff (lig_pt_far): the
// residual check and the 0.01 A clamp are skipped, which keeps the loop branch-free
// shift t = 1 - r^2/144 in FP64 from the fixed-point differences and both atoms'
// quantisation residuals.  The FP32 r^2 carries ~1e-7 relative error, so the FP32 t has
// an absolute error of ~1e-7 and the weight w = t^2 a relative error of ~2e-7 / t:
// negligible unless every pair of a sum lies within a few tenths of an A of the cutoff,
// which is the case for a single probe atom on the far side of a grid.
__device__ __forceinline__ float shift_t_exact(const RecPos& p, const LigPt& l, int i,
                                               const PairK& pk) {
  const RecRes rr = pk.rres[i];
  const double s = pk.inv_s;
  const double x = fma((double)(p.x - l.x), s, (double)rr.x - (double)l.rx);
  const double y = fma((double)(p.y - l.y), s, (double)rr.y - (double)l.ry);
  const double z = fma((double)(p.z - l.z), s, (double)rr.z - (double)l.rz);
  const double t = fma(fma(x, x, fma(y, y, z * z)), -1.0 / 144.0, 1.0);
  return (float)fmin(fmax(t, 0.0), 1.0);
}

// EXACT_T: the shift t from shift_t_exact (single-probe sums: the grid precompute)
template <int MODE, bool NEAR = true, bool EXACT_T = false>
__device__ __forceinline__ void pair_energy(const RecPos& p, const RecAB& ab, const LigPt& l,
                                            int i, const PairK& pk, float& sq, float& s12,
                                            float& s6) {
  const float dx = (float)(p.x - l.x), dy = (float)(p.y - l.y), dz = (float)(p.z - l.z);
  // NEAR: the clamp is applied in the near-pair branch (pair_r2_clamped); else not needed
  const float r2c = NEAR ? pair_r2_clamped(dx, dy, dz, l, i, pk)
                         : (dx * dx + dy * dy + dz * dz) * pk.inv_s2;
  const float ir = __builtin_amdgcn_rsqf(r2c);
  const float u = inv_r2(ir, r2c);
  const float u3 = u * (u * u);
  if (MODE == MMO_SHIFTED) {
    // w u^3 serves both LJ terms: 9 multiplies and FMAs after u^3 instead of 10
    const float t = EXACT_T ? shift_t_exact(p, l, i, pk) : shift_t(r2c, pk);
    const float w = t * t;
    const float v = w * u3;
    sq = fmaf(p.q, w * ir, sq);
    s6 = fmaf(ab.b, v, s6);
    s12 = fmaf(ab.a, v * u3, s12);
  } else {
    sq = fmaf(p.q, ir, sq);
    s12 = fmaf(ab.a, u3 * u3, s12);
    s6 = fmaf(ab.b, u3, s6);
  }
}

// shifted-form energy and dE/dr2 of one pair; the gradient is accumulated as
// coef * (x_rec - x_lig) in fixed-point units (the caller scales by -2 * 2^-k).
// F = ke qi qj / r + Aij u^6 - 2 Bij u^3, dF/dr2 = -u (ke qi qj / (2 r) + 6 Aij u^6 - 6 Bij u^3),
// w = t^2 with t = 1 - r2/144, dw/dr2 = -t/72.
__device__ __forceinline__ void pair_grad_shifted(const RecPos& p, const RecAB& ab, const LigPt& l,
                                                  int i, const PairK& pk, float c1, float aj,
                                                  float bj, float& e, float& gx, float& gy,
                                                  float& gz) {
  const float dx = (float)(p.x - l.x), dy = (float)(p.y - l.y), dz = (float)(p.z - l.z);
  const float r2 = pair_r2(dx, dy, dz, l, i, pk);
  const float r2c = max_r2(r2, pk.min_r2);
  const float ir = __builtin_amdgcn_rsqf(r2c);
  const float u = inv_r2(ir, r2c);
  const float u3 = u * u * u;
  const float u6 = u3 * u3;
  const float eq = p.q * ir, e12 = ab.a * u6, e6 = ab.b * u3;
  const float F = fmaf(c1, eq, fmaf(aj, e12, -2.0f * bj * e6));
  const float G = fmaf(0.5f * c1, eq, 6.0f * fmaf(aj, e12, -bj * e6));
  const float t = shift_t(r2c, pk);
  const float w = t * t;
  float coef = fmaf(t * F, -1.0f / 72.0f, -(w * u) * G);
  coef = (r2 >= pk.min_r2) ? coef : 0.0f;   // clamped distance: flat
  e = fmaf(w, F, e);
  gx = fmaf(coef, dx, gx);
  gy = fmaf(coef, dy, gy);
  gz = fmaf(coef, dz, gz);
}

// ------------------------------------------------------- inter kernel, culled
// Shifted form only.  The whole receptor sits in LDS sorted by 4 A cells.  For each
// block of K atoms per lane the wavefront takes the bounding box of its lanes' atom
// positions and visits only the cell rows that intersect the sphere of radius
// 12 A + half-diagonal around the box centre; every receptor atom within 12 A of any
// of the wavefront's atoms is in a visited cell, and pairs beyond 12 A that are
// visited contribute exactly 0 (the shift weight is clamped to 0).
// wavefront min / max through DPP (lanes without a source keep their own value);
// the result is uniform (read from lane 63, which has seen every lane)
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ float dpp_keep_f(float v) {
  const int b = __float_as_int(v);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, CTRL, ROW_MASK, 0xf, false));
}
__device__ __forceinline__ float wave_minf(float v) {
  v = fminf(v, dpp_keep_f<0x111>(v));
  v = fminf(v, dpp_keep_f<0x112>(v));
  v = fminf(v, dpp_keep_f<0x114>(v));
  v = fminf(v, dpp_keep_f<0x118>(v));
  v = fminf(v, dpp_keep_f<0x142, 0xa>(v));
  v = fminf(v, dpp_keep_f<0x143, 0xc>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_maxf(float v) {
  v = fmaxf(v, dpp_keep_f<0x111>(v));
  v = fmaxf(v, dpp_keep_f<0x112>(v));
  v = fmaxf(v, dpp_keep_f<0x114>(v));
  v = fmaxf(v, dpp_keep_f<0x118>(v));
  v = fmaxf(v, dpp_keep_f<0x142, 0xa>(v));
  v = fmaxf(v, dpp_keep_f<0x143, 0xc>(v));
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float uniformf(float v) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ int uniformi(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Visits every receptor atom within 12 A of the box [q - h, q + h] (plus some beyond it)
// through the cell list, accumulating the receptor-side sums for K ligand atoms per
// lane; fold(sq, s12, s6) is called once per z slab of cells (FP32 partial sums).  The
// slabs, rows and row x ranges are those of the box grown by 12 A (the box's Minkowski
// sum with the cutoff ball, ~5% smaller than the ball around the box centre of radius
// 12 A + half-diagonal).  A row of cells holds its atoms sorted by x, so a row longer
// than cg.clip_min atoms is clipped to that x range: one read per lane of the row's
// first (and, past 64 atoms, last) 64 atoms and a ballot give the atoms below and above.
template <int K, bool EXACT_T = false, class Fold>
__device__ __forceinline__ void cull_sweep(const CellGrid& cg, const int* __restrict__ s_cell,
                                           const RecPos* __restrict__ s_pos,
                                           const RecAB* __restrict__ s_ab,
                                           const RecRes* __restrict__ rres, float inv_s,
                                           float inv_s2, float qx, float qy, float qz,
                                           float hx, float hy, float hz,
                                           const LigPt (&l)[K], Fold&& fold) {
  const PairK pk = pair_k(inv_s2, inv_s, rres);
  constexpr float R = 12.05f;   // cutoff + FP32 rounding margin
  constexpr float R2 = R * R;
  const int cz0 = max(0, (int)floorf((qz - hz - R - cg.z0) * cg.inv_c));
  const int cz1 = min(cg.nz - 1, (int)floorf((qz + hz + R - cg.z0) * cg.inv_c));
  for (int cz = cz0; cz <= cz1; ++cz) {
    // distances from the box [q - h, q + h] to the slab, then to the row
    const float zlo = cg.z0 + cz * cg.c;
    const float dz = fmaxf(0.0f, fmaxf(zlo - (qz + hz), (qz - hz) - (zlo + cg.c)));
    const float rz2 = R2 - dz * dz;
    if (rz2 < 0.0f) continue;
    const float ry = bound_sqrt(rz2);
    const int cy0 = max(0, (int)floorf((qy - hy - ry - cg.y0) * cg.inv_c));
    const int cy1 = min(cg.ny - 1, (int)floorf((qy + hy + ry - cg.y0) * cg.inv_c));
    float sq[K], s12[K], s6[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sq[k] = s12[k] = s6[k] = 0.0f;
    for (int cy = cy0; cy <= cy1; ++cy) {
      const float ylo = cg.y0 + cy * cg.c;
      const float dy = fmaxf(0.0f, fmaxf(ylo - (qy + hy), (qy - hy) - (ylo + cg.c)));
      const float rx2 = rz2 - dy * dy;
      if (rx2 < 0.0f) continue;
      const float rx = hx + bound_sqrt(rx2);
      const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_cx));
      const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_cx));
      if (cx0 > cx1) continue;
      // a row whose cells are 0.5 A or more from the box in y or z holds no atom closer
      // than 0.5 A to any lane's atom: its pairs skip the near-pair residual check
      const bool near_row = dz < 0.55f && dy < 0.55f;
      const int rowb = (cz * cg.ny + cy) * cg.nx;
      int a0 = uniformi(s_cell[rowb + cx0]);
      int a1 = uniformi(s_cell[rowb + cx1 + 1]);
      const int len = a1 - a0;
      if (len > cg.clip_min) {
        const int lane = threadIdx.x & (kWave - 1);
        const float xlo = qx - rx, xhi = qx + rx;
        const bool in0 = lane < len;
        const float x0 = in0 ? (float)s_pos[a0 + lane].x * inv_s : 0.0f;
        const int below = __popcll(__ballot(in0 && x0 < xlo));
        int upto;   // atoms with x <= xhi: a prefix of the row
        if (len <= kWave) {
          upto = a0 + __popcll(__ballot(in0 && x0 <= xhi));
        } else {
          const float x1 = (float)s_pos[a1 - kWave + lane].x * inv_s;
          upto = a1 - kWave + __popcll(__ballot(x1 <= xhi));
        }
        a0 = uniformi(a0 + below);
        a1 = uniformi(upto);
      }
      if (near_row) {
        for (int i = a0; i < a1; ++i) {
          const RecPos p = s_pos[i];
          const RecAB ab = s_ab[i];
#pragma unroll
          for (int k = 0; k < K; ++k)
            pair_energy<MMO_SHIFTED, true, EXACT_T>(p, ab, l[k], i, pk, sq[k], s12[k], s6[k]);
        }
      } else {
#pragma unroll 4
        for (int i = a0; i < a1; ++i) {
          const RecPos p = s_pos[i];
          const RecAB ab = s_ab[i];
#pragma unroll
          for (int k = 0; k < K; ++k)
            pair_energy<MMO_SHIFTED, false, EXACT_T>(p, ab, l[k], i, pk, sq[k], s12[k], s6[k]);
        }
      }
    }
    fold(sq, s12, s6);
  }
}

// Rigid-scan sweep over a list of cell rows, K ligand atoms per lane.  The (y, z) rows
// of the grid cg that come within 12 A of the box [q - h, q + h] are enumerated 64 at a
// time across the lanes: each lane takes one candidate row, computes the x chord of the
// box grown by 12 A at that row and the atom range of the cells the chord covers (two
// reads of the cell table, all lanes at once).  The non-empty rows are then swept in
// (z, y) order by the whole wavefront, the lowest set bit of a ballot first, each row's
// range read back with readlane; every receptor atom fetched is paired with the lane's
// K atoms.  FP32 sums are folded once per z slab, after the slab's last row: pairs
// beyond 12 A add exact zeros, so a lane's energy depends only on its own atoms, not on
// the box it was swept with.  fold(sq, s12, s6) receives the K slab sums.
// SMEM: the receptor arrays are read with wavefront-uniform addresses from global memory
// (scalar loads into SGPRs); else from LDS into VGPRs.
// FILT: before the sweep, the lane of each far candidate row tests the row's first 64
// atoms against the box grown by 12.1 A (one gathered position per atom) and keeps a
// bitmask; the atoms of the x cells inside the chord at the row's farthest (y, z) point
// are kept untested.  The sweep then visits only the kept atoms (chunks start at the
// next kept atom).  A dropped atom is more than 12.09 A from every lane's atom, so its pairs would
// add exact zeros: the sums keep their bits and their order.  The row chords alone visit
// about 2x the pairs inside the cutoff; the per-atom test leaves about 1.3x.
template <bool SMEM, int CHUNK, int K, bool FILT, class Fold>
__device__ __forceinline__ void row_sweep(const CellGrid& cg, const RecPos* __restrict__ s_pos,
                                          const RecAB* __restrict__ s_ab,
                                          const RecRes* __restrict__ rres, float inv_s,
                                          float inv_s2, float qx, float qy, float qz, float hx,
                                          float hy, float hz, const LigPt (&l)[K], Fold&& fold) {
  const PairK pk = pair_k(inv_s2, inv_s, rres);
  constexpr float R = 12.05f;   // cutoff + FP32 rounding margin
  constexpr float R2 = R * R;
  const int lane = threadIdx.x & (kWave - 1);
  const int cz0 = max(0, (int)floorf((qz - hz - R - cg.z0) * cg.inv_c));
  const int cz1 = min(cg.nz - 1, (int)floorf((qz + hz + R - cg.z0) * cg.inv_c));
  const int cy0 = max(0, (int)floorf((qy - hy - R - cg.y0) * cg.inv_c));
  const int cy1 = min(cg.ny - 1, (int)floorf((qy + hy + R - cg.y0) * cg.inv_c));
  if (cz0 > cz1 || cy0 > cy1) return;
  const int ncy = cy1 - cy0 + 1;
  const int nc = (cz1 - cz0 + 1) * ncy;
  // candidate c -> (c / ncy, c % ncy) from the float quotient of (c + 1/2) / ncy, which
  // lies at least 1/(2 ncy) from an integer (exact for c < 2^16; corrected by one beyond)
  const float inv_ncy = 1.0f / (float)ncy;
  float sq[K], s12[K], s6[K];
#pragma unroll
  for (int k = 0; k < K; ++k) sq[k] = s12[k] = s6[k] = 0.0f;
  int slab = -1;   // z slab of the running FP32 sums
  for (int base = 0; base < nc; base += kWave) {
    const int cidx = base + lane;
    int a0 = 0, a1 = 0, tag = 0;
    int b0 = 0, b1 = 0;   // FILT: atoms of the row's inner x cells, kept without a test
    if (cidx < nc) {
      int dzi = (int)(((float)cidx + 0.5f) * inv_ncy);
      dzi -= dzi * ncy > cidx ? 1 : 0;
      dzi += (dzi + 1) * ncy <= cidx ? 1 : 0;
      const int cz = cz0 + dzi, cy = cy0 + (cidx - dzi * ncy);
      const float zlo = cg.z0 + cz * cg.c, ylo = cg.y0 + cy * cg.c;
      const float dz = fmaxf(0.0f, fmaxf(zlo - (qz + hz), (qz - hz) - (zlo + cg.c)));
      const float dy = fmaxf(0.0f, fmaxf(ylo - (qy + hy), (qy - hy) - (ylo + cg.c)));
      const float rx2 = R2 - dz * dz - dy * dy;
      if (rx2 >= 0.0f) {
        const float rx = hx + bound_sqrt(rx2);
        const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_cx));
        const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_cx));
        if (cx0 <= cx1) {
          const int rowb = (cz * cg.ny + cy) * cg.nx;
          a0 = cg.cell_start[rowb + cx0];
          a1 = cg.cell_start[rowb + cx1 + 1];
          b0 = b1 = a0;
          if constexpr (FILT) {
            // the chord at the row's farthest (y, z) point: x cells within it hold only
            // atoms within 12.05 A of the box (keeping an atom beyond only visits zeros)
            const float dzf = fmaxf(0.0f, fmaxf((qz - hz) - zlo, (zlo + cg.c) - (qz + hz)));
            const float dyf = fmaxf(0.0f, fmaxf((qy - hy) - ylo, (ylo + cg.c) - (qy + hy)));
            const float ri2 = R2 - dzf * dzf - dyf * dyf;
            if (ri2 > 0.0f) {
              const float ri = hx + bound_sqrt(ri2);
              const int ix0 = max(cx0, (int)ceilf((qx - ri - cg.x0) * cg.inv_cx));
              const int ix1 = min(cx1, (int)floorf((qx + ri - cg.x0) * cg.inv_cx) - 1);
              if (ix0 <= ix1) {
                b0 = cg.cell_start[rowb + ix0];
                b1 = cg.cell_start[rowb + ix1 + 1];
              }
            }
          }
        }
      }
      // a row 0.5 A or more from the box in y or z holds no atom closer than 0.5 A to
      // any lane's atom: its pairs skip the near-pair residual check
      tag = 2 * cz + ((dz < 0.55f && dy < 0.55f) ? 1 : 0);
    }
    uint32_t keep_lo = ~0u, keep_hi = ~0u;   // FILT: kept atoms of this lane's far row
    if constexpr (FILT) {
      constexpr float RF2 = 12.1f * 12.1f;
      // bits (atom - a0) of the first 64 atoms: [0, n0) and [n1, n2) are tested, the
      // inner cells' atoms [n0, n1) kept
      const bool far = a1 > a0 && !(tag & 1);
      const int n2 = min(a1 - a0, 64);
      const int n0 = min(b0 - a0, n2), n1 = min(b1 - a0, n2);
      const int flen = far ? n0 + (n2 - n1) : 0;
      const int maxlen = (int)wave_maxf((float)flen);
      const uint64_t lo_n1 = n1 >= 64 ? ~0ull : ((1ull << n1) - 1ull);
      uint64_t m = n1 > n0 ? (lo_n1 & ~((1ull << n0) - 1ull)) : 0ull;   // then n0 < 64
      constexpr int U = 2;   // gathers in flight (more would spill at 64 VGPRs)
      for (int it = 0; it < maxlen; it += U) {
        // a lane past its tested atoms reads atom 0, and drops it
        int px[U], py[U], pz[U], bit[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          bit[u] = it + u < n0 ? it + u : n1 + (it + u - n0);
          const RecPos p = s_pos[it + u < flen ? a0 + bit[u] : 0];
          px[u] = p.x; py[u] = p.y; pz[u] = p.z;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const float ex = fmaxf(fabsf(fmaf((float)px[u], inv_s, -qx)) - hx, 0.0f);
          const float ey = fmaxf(fabsf(fmaf((float)py[u], inv_s, -qy)) - hy, 0.0f);
          const float ez = fmaxf(fabsf(fmaf((float)pz[u], inv_s, -qz)) - hz, 0.0f);
          const bool in = it + u < flen && fmaf(ex, ex, fmaf(ey, ey, ez * ez)) <= RF2;
          m |= (uint64_t)in << (bit[u] & 63);
        }
      }
      if (far) { keep_lo = (uint32_t)m; keep_hi = (uint32_t)(m >> 32); }
    }
    uint64_t rows = __ballot(a1 > a0);
    while (rows) {
      const int r = __builtin_ctzll(rows);
      rows &= rows - 1;
      const int ra0 = __builtin_amdgcn_readlane(a0, r);
      const int ra1 = __builtin_amdgcn_readlane(a1, r);
      const int rtag = __builtin_amdgcn_readlane(tag, r);
      if ((rtag >> 1) != slab) {
        if (slab >= 0) fold(sq, s12, s6);
#pragma unroll
        for (int k = 0; k < K; ++k) sq[k] = s12[k] = s6[k] = 0.0f;
        slab = rtag >> 1;
      }
      if (rtag & 1) {
        for (int i = ra0; i < ra1; ++i) {
          const RecPos p = s_pos[i];
          const RecAB ab = s_ab[i];
#pragma unroll
          for (int k = 0; k < K; ++k)
            pair_energy<MMO_SHIFTED, true>(p, ab, l[k], i, pk, sq[k], s12[k], s6[k]);
        }
      } else {
        // CHUNK (<= kRowChunk) atoms fetched together (one memory latency per chunk of a
        // short row; the arrays are padded so that a chunk may run past the last atom)
        int tail0 = ra0;   // first atom swept without the filter
        if constexpr (FILT) {
          uint64_t m = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)keep_hi, r) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)keep_lo, r);
          while (m) {
            const int k0 = __builtin_ctzll(m);
            const int c0 = ra0 + k0;
            RecPos pp[CHUNK];
            RecAB aa[CHUNK];
#pragma unroll
            for (int k = 0; k < CHUNK; ++k) {
              pp[k] = s_pos[c0 + k];
              aa[k] = s_ab[c0 + k];
            }
#pragma unroll
            for (int k = 0; k < CHUNK; ++k) {
              if constexpr (SMEM)
                asm volatile("" ::"s"(pp[k].x), "s"(pp[k].y), "s"(pp[k].z), "s"(pp[k].q),
                             "s"(aa[k].a), "s"(aa[k].b));
              else
                asm volatile("" ::"v"(pp[k].x), "v"(pp[k].y), "v"(pp[k].z), "v"(pp[k].q),
                             "v"(aa[k].a), "v"(aa[k].b));
            }
            const uint32_t bits = (uint32_t)(m >> k0);   // bit 0 (atom c0) is set
#pragma unroll
            for (int k = 0; k < CHUNK; ++k)
              if (k == 0 || ((bits >> k) & 1u)) {
#pragma unroll
                for (int mm = 0; mm < K; ++mm)
                  pair_energy<MMO_SHIFTED, false>(pp[k], aa[k], l[mm], c0 + k, pk, sq[mm], s12[mm],
                                                  s6[mm]);
              }
            m &= ~(((1ull << CHUNK) - 1ull) << k0);   // the bits below k0 are clear
          }
          tail0 = ra0 + 64;   // atoms past the first 64 of a long row
        }
        for (int c0 = tail0; c0 < ra1; c0 += CHUNK) {
          RecPos pp[CHUNK];
          RecAB aa[CHUNK];
#pragma unroll
          for (int k = 0; k < CHUNK; ++k) {
            pp[k] = s_pos[c0 + k];
            aa[k] = s_ab[c0 + k];
          }
          // every fetch of the chunk completes here, before the first pair: without this
          // the compiler sinks each atom's fetch (and a wait) into that atom's branch
#pragma unroll
          for (int k = 0; k < CHUNK; ++k) {
            if constexpr (SMEM)
              asm volatile("" ::"s"(pp[k].x), "s"(pp[k].y), "s"(pp[k].z), "s"(pp[k].q), "s"(aa[k].a),
                           "s"(aa[k].b));
            else
              asm volatile("" ::"v"(pp[k].x), "v"(pp[k].y), "v"(pp[k].z), "v"(pp[k].q), "v"(aa[k].a),
                           "v"(aa[k].b));
          }
          const int cnt = ra1 - c0;
#pragma unroll
          for (int k = 0; k < CHUNK; ++k)
            if (k < cnt) {
#pragma unroll
              for (int m = 0; m < K; ++m)
                pair_energy<MMO_SHIFTED, false>(pp[k], aa[k], l[m], c0 + k, pk, sq[m], s12[m], s6[m]);
            }
        }
      }
    }
  }
  if (slab >= 0) fold(sq, s12, s6);
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
    const float ry = bound_sqrt(rz2);
    const int cy0 = max(0, (int)floorf((qy - ry - cg.y0) * cg.inv_c));
    const int cy1 = min(cg.ny - 1, (int)floorf((qy + ry - cg.y0) * cg.inv_c));
    for (int cy = cy0; cy <= cy1; ++cy) {
      const float ylo = cg.y0 + cy * cg.c;
      const float dy = fmaxf(0.0f, fmaxf(ylo - qy, qy - (ylo + cg.c)));
      const float rx2 = rz2 - dy * dy;
      if (rx2 < 0.0f) continue;
      const float rx = bound_sqrt(rx2);
      const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_cx));
      const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_cx));
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

// bytes of LDS that load_receptor<LDS> uses (rounded up to 16)
template <bool LDS>
__host__ __device__ __forceinline__ size_t receptor_lds_bytes(int nrec, const CellGrid& cg) {
  if (!LDS) return 0;
  const size_t b = (size_t)nrec * (sizeof(RecPos) + sizeof(RecAB)) +
                   (size_t)(cg.nx * cg.ny * cg.nz + 1) * sizeof(int);
  return (b + 15) & ~(size_t)15;
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

// positions and charges (RecPos) + cell table into LDS, the LJ factors (RecAB) read in
// place: a wavefront reads the same receptor atom on every lane, so RecAB comes through
// the scalar cache while RecPos (the operands of the coordinate differences) comes from
// LDS into VGPRs, keeping the per-pair SGPR operands to two
__host__ __device__ __forceinline__ size_t receptor_pos_lds_bytes(int nrec, const CellGrid& cg) {
  const size_t b = (size_t)nrec * sizeof(RecPos) + (size_t)(cg.nx * cg.ny * cg.nz + 1) * sizeof(int);
  return (b + 15) & ~(size_t)15;
}
__device__ __forceinline__ void load_receptor_pos(const RecPos* __restrict__ rpos,
                                                  const RecAB* __restrict__ rab, int nrec,
                                                  const CellGrid& cg, const RecPos*& s_pos,
                                                  const RecAB*& s_ab, const int*& s_cell,
                                                  char* smem) {
  const int ncell = cg.nx * cg.ny * cg.nz;
  RecPos* p = reinterpret_cast<RecPos*>(smem);
  int* c = reinterpret_cast<int*>(smem + (size_t)nrec * sizeof(RecPos));
  for (int i = threadIdx.x; i < nrec; i += blockDim.x) p[i] = rpos[i];
  for (int i = threadIdx.x; i <= ncell; i += blockDim.x) c[i] = cg.cell_start[i];
  __syncthreads();
  s_pos = p;
  s_ab = rab;
  s_cell = c;
}

// bounding box of the wavefront's valid fixed-point positions -> centre q and half
// extents h (A), uniform across the wave; false when no lane holds a valid atom
template <int K>
__device__ __forceinline__ bool wave_box(const int (&lx)[K], const int (&ly)[K],
                                         const int (&lz)[K], const bool (&valid)[K],
                                         float inv_s, float& qx, float& qy, float& qz,
                                         float& hx, float& hy, float& hz) {
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
  hx = 0.5f * (mxx - mnx); hy = 0.5f * (mxy - mny); hz = 0.5f * (mxz - mnz);
  return true;
}

template <int K>
__device__ __forceinline__ bool wave_box(const LigPt (&l)[K], const bool (&valid)[K], float inv_s,
                                         float& qx, float& qy, float& qz, float& hx, float& hy,
                                         float& hz) {
  int lx[K], ly[K], lz[K];
#pragma unroll
  for (int k = 0; k < K; ++k) { lx[k] = l[k].x; ly[k] = l[k].y; lz[k] = l[k].z; }
  return wave_box<K>(lx, ly, lz, valid, inv_s, qx, qy, qz, hx, hy, hz);
}

// the same box as a sphere: centre q, radius 12 A + half-diagonal (+0.05 A for FP32
// rounding), uniform across the wave
template <int K>
__device__ __forceinline__ bool wave_query(const int (&lx)[K], const int (&ly)[K],
                                           const int (&lz)[K], const bool (&valid)[K],
                                           float inv_s, float& qx, float& qy, float& qz,
                                           float& R) {
  float hx, hy, hz;
  if (!wave_box<K>(lx, ly, lz, valid, inv_s, qx, qy, qz, hx, hy, hz)) return false;
  R = uniformf(12.0f + sqrtf(hx * hx + hy * hy + hz * hz) + 0.05f);
  return true;
}

// ------------------------------------------------- one pose per wavefront, culled
// The atoms of one pose are taken kSub = 8 at a time; lane = a + 8 r holds atom slot a
// of the sub-block and walks every 8th receptor atom (slot r) of the cell rows that
// meet the sphere of radius 12 A + half-diagonal around the sub-block's bounding box.
// Eight neighbouring atoms have a small box, so the sweep stays close to the pairs
// inside the cutoff even though a whole pose spans up to 24 A.  The (z, y) cell rows of
// the sphere's box are enumerated across the lanes, which fetch their atom ranges in
// parallel (one memory latency per 64 rows, not one per row); the rows are then swept
// in order, each lane reading the ranges back with readlane.  atom(j, l, cq, ca, cb)
// loads atom j (LigPt); with GRAD, gout(j, gx, gy, gz) receives dE/dx_j (A^-1 units,
// FP64) once per atom.  Sums: FP32 per 64 rows, FP64 across them.
constexpr int kSub = 8;

template <bool GRAD, class AtomFn, class GradOut>
__device__ __forceinline__ void pose_sweep(const CellGrid& cg, const int* __restrict__ s_cell,
                                           const RecPos* __restrict__ s_pos,
                                           const RecAB* __restrict__ s_ab,
                                           const RecRes* __restrict__ rres, const Frame& fr,
                                           int n_atoms, AtomFn&& atom, GradOut&& gout,
                                           double& e_el, double& e_vd) {
  const PairK pk = pair_k(fr.inv_s2, fr.inv_s, rres);
  const int lane = threadIdx.x & (kWave - 1);
  const int aslot = lane & (kSub - 1), rslot = lane >> 3;
  for (int j0 = 0; j0 < n_atoms; j0 += kSub) {
    const int j = j0 + aslot;
    LigPt l[1];
    bool valid[1];
    double cq = 0.0, ca = 0.0, cb = 0.0;
    valid[0] = j < n_atoms;
    l[0] = lig_pt_zero();
    if (valid[0]) atom(j, l[0], cq, ca, cb);
    int lx[1] = {l[0].x}, ly[1] = {l[0].y}, lz[1] = {l[0].z};
    float qx, qy, qz, R;
    if (!wave_query<1>(lx, ly, lz, valid, fr.inv_s, qx, qy, qz, R)) continue;
    const float c1 = (float)cq, fa = (float)ca, fb = (float)cb;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    const float R2 = R * R;
    const int cz0 = max(0, (int)floorf((qz - R - cg.z0) * cg.inv_c));
    const int cz1 = min(cg.nz - 1, (int)floorf((qz + R - cg.z0) * cg.inv_c));
    const int cy0 = max(0, (int)floorf((qy - R - cg.y0) * cg.inv_c));
    const int cy1 = min(cg.ny - 1, (int)floorf((qy + R - cg.y0) * cg.inv_c));
    const int nyc = cy1 - cy0 + 1;
    const int nrows = (cz1 >= cz0 && nyc > 0) ? (cz1 - cz0 + 1) * nyc : 0;
    for (int base = 0; base < nrows; base += kWave) {
      // lane k: row base + k -> its atom range [r0, r1) (empty if the row misses the sphere)
      int r0 = 0, r1 = 0;
      const int idx = base + lane;
      if (idx < nrows) {
        const int cz = cz0 + idx / nyc, cy = cy0 + idx % nyc;
        const float zlo = cg.z0 + cz * cg.c, ylo = cg.y0 + cy * cg.c;
        const float dzb = fmaxf(0.0f, fmaxf(zlo - qz, qz - (zlo + cg.c)));
        const float dyb = fmaxf(0.0f, fmaxf(ylo - qy, qy - (ylo + cg.c)));
        const float rx2 = R2 - dzb * dzb - dyb * dyb;
        if (rx2 >= 0.0f) {
          const float rx = bound_sqrt(rx2);
          const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_cx));
          const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_cx));
          if (cx0 <= cx1) {
            const int rowb = (cz * cg.ny + cy) * cg.nx;
            r0 = s_cell[rowb + cx0];
            r1 = s_cell[rowb + cx1 + 1];
          }
        }
      }
      float sq = 0.0f, s12 = 0.0f, s6 = 0.0f, fe = 0.0f, fx = 0.0f, fy = 0.0f, fz = 0.0f;
      const int nk = min(kWave, nrows - base);
      for (int k = 0; k < nk; ++k) {
        const int a0 = __builtin_amdgcn_readlane(r0, k);
        const int a1 = __builtin_amdgcn_readlane(r1, k);
        for (int i = a0 + rslot; i < a1; i += kSub) {
          const RecPos p = s_pos[i];
          const RecAB ab = s_ab[i];
          if (GRAD)
            pair_grad_shifted(p, ab, l[0], i, pk, c1, fa, fb, fe, fx, fy, fz);
          else
            pair_energy<MMO_SHIFTED>(p, ab, l[0], i, pk, sq, s12, s6);
        }
      }
      if (GRAD) {
        e_vd += (double)fe;
        gx += (double)fx; gy += (double)fy; gz += (double)fz;
      } else {
        e_el = fma((double)sq, cq, e_el);
        e_vd = fma((double)s12, ca, e_vd);
        e_vd = fma(-2.0 * (double)s6, cb, e_vd);
      }
    }
    if (GRAD) {
      // sum over the receptor slots of each atom slot
      for (int off = kSub; off < kWave; off <<= 1) {
        gx += __shfl_xor(gx, off, kWave);
        gy += __shfl_xor(gy, off, kWave);
        gz += __shfl_xor(gz, off, kWave);
      }
      // dE/dx_lig = -2 * sum coef * (x_rec - x_lig), differences in fixed-point units
      const double gs = -2.0 * (double)fr.inv_s;
      if (rslot == 0 && valid[0]) gout(j, gs * gx, gs * gy, gs * gz);
    }
  }
}

// ------------------------------------------ one pose per wavefront, lanes on receptor
// For each atom j of the pose in turn, the wavefront visits the receptor atoms of the
// cell rows that meet the 12 A sphere around atom j itself (no bounding-box growth):
// lane k fetches the atom range of row k, the non-empty rows are compacted into LDS
// with a wavefront prefix sum that lays their ranges end to end, and the concatenation
// (`total` pairs) is cut into 64 contiguous segments of ceil(total / 64) positions, one
// per lane.  A lane finds the row of its first position from a count of rows per
// first lane (LDS atomics) and a prefix sum across the lanes, and then walks its
// segment, stepping to the next row where one ends, so the bookkeeping is a few LDS
// round trips per atom and a compare per pair.
// wl: kAtomSweepInts ints of this wavefront's LDS.  atom(j, l, cq, ca, cb) gives
// atom j (LigPt); with PRELOAD it is called by lane j % 64 for its chunk's atoms (so it
// may depend on the lane's j), else by all lanes for one j at a time.  NEAR = false
// (energy form only) is for callers whose pairs are all at least 0.5 A apart: it skips
// the residual correction and the 0.01 A clamp (pair_energy); with
// GRAD, gout(j, gx, gy, gz) receives dE/dx_j (A^-1, FP64) from the lane holding atom j
// (PRELOAD) or from lane 0.  Sums: FP32 per lane and atom, FP64 across atoms.
constexpr int kAtomSweepInts = 3 * kWave;

// ceil(a / m) for 0 <= a < 64 m, given inv_m within 1 ulp of 1/m: the float quotient
// (below 64) is within 1 of floor(a / m) and is corrected in integers
__device__ __forceinline__ int ceil_div_small(int a, int m, float inv_m) {
  int q = (int)((float)a * inv_m);
  q -= q * m > a ? 1 : 0;
  q += (q + 1) * m <= a ? 1 : 0;
  return q + (q * m < a ? 1 : 0);
}

__device__ __forceinline__ int in_vgpr_i(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <bool GRAD, class AtomFn, class GradOut, bool NEAR = true, bool PRELOAD = true,
          int NA = 1>
__device__ __forceinline__ void atom_sweep(const CellGrid& cg, const int* __restrict__ s_cell,
                                           const RecPos* __restrict__ s_pos,
                                           const RecAB* __restrict__ s_ab,
                                           const RecRes* __restrict__ rres, const Frame& fr,
                                           int n_atoms, int* wl, AtomFn&& atom, GradOut&& gout,
                                           double& e_el, double& e_vd) {
  static_assert(NA == 1 || (NA >= 2 && NA <= 4 && PRELOAD), "several atoms per pass need the preload");
  const PairK pk = pair_k(fr.inv_s2, fr.inv_s, rres);
  const int lane = threadIdx.x & (kWave - 1);
  int* s_r0 = wl;                 // compacted rows: first receptor atom
  int* s_beg = wl + kWave;        //                 first position; s_beg[nc] = total
  int* s_cnt = wl + 2 * kWave;    // [1..63]: rows whose first segment lane is this one
  const uint64_t below = lane ? ((1ull << lane) - 1ull) : 0ull;
  constexpr float R = 12.05f;     // cutoff + FP32 rounding margin
  constexpr float R2 = R * R;
  // PRELOAD: the atoms of a 64-atom chunk are fetched together, one per lane, and
  // broadcast one at a time; in GRAD the lane of an atom keeps its gradient until the
  // chunk ends, so no atom waits on its own load or gradient write.  NA > 1: NA
  // consecutive atoms per pass share one row listing (the rows of the union of their
  // cutoff spheres, each row's x range the hull of theirs) and every receptor atom read
  // is paired with each of them; a short last group repeats its first atom, whose copies
  // are dropped.
  for (int jb = 0; jb < n_atoms; jb += (PRELOAD ? kWave : 1)) {
    LigPt mine = lig_pt_zero();
    double mq = 0.0, ma = 0.0, mb = 0.0;
    if (PRELOAD && jb + lane < n_atoms) atom(jb + lane, mine, mq, ma, mb);
    float gxm = 0.0f, gym = 0.0f, gzm = 0.0f;
    const int nj = PRELOAD ? min(kWave, n_atoms - jb) : 1;
    for (int jj = 0; jj < nj; jj += NA) {
      LigPt lu[NA];
      double cq[NA], ca[NA], cb[NA];
      bool live[NA];
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        live[u] = jj + u < nj;
        const int ju = live[u] ? jj + u : jj;
        if (PRELOAD) {
          lu[u].x = __builtin_amdgcn_readlane(mine.x, ju);
          lu[u].y = __builtin_amdgcn_readlane(mine.y, ju);
          lu[u].z = __builtin_amdgcn_readlane(mine.z, ju);
          lu[u].rx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.rx), ju));
          lu[u].ry = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.ry), ju));
          lu[u].rz = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine.rz), ju));
          cq[u] = readlane_d(mq, ju);
          ca[u] = readlane_d(ma, ju);
          cb[u] = readlane_d(mb, ju);
        } else {
          atom(jb + jj, lu[u], cq[u], ca[u], cb[u]);
        }
      }
      float qx[NA], qy[NA], qz[NA], c1[NA], fa[NA], fb[NA];
      LigPt lv[NA];
      float zlo_q = 3.0e38f, zhi_q = -3.0e38f, ylo_q = 3.0e38f, yhi_q = -3.0e38f;
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        qx[u] = uniformf((float)lu[u].x * fr.inv_s);
        qy[u] = uniformf((float)lu[u].y * fr.inv_s);
        qz[u] = uniformf((float)lu[u].z * fr.inv_s);
        // the pair loop reads the (wave-uniform) ligand atoms from VGPRs: an SGPR operand
        // would cost issue rate on every pair
        lv[u] = LigPt{in_vgpr_i(lu[u].x), in_vgpr_i(lu[u].y), in_vgpr_i(lu[u].z), lu[u].rx,
                      lu[u].ry, lu[u].rz};
        c1[u] = (float)cq[u]; fa[u] = (float)ca[u]; fb[u] = (float)cb[u];
        zlo_q = fminf(zlo_q, qz[u]); zhi_q = fmaxf(zhi_q, qz[u]);
        ylo_q = fminf(ylo_q, qy[u]); yhi_q = fmaxf(yhi_q, qy[u]);
      }
      const int cz0 = max(0, (int)floorf((zlo_q - R - cg.z0) * cg.inv_c));
      const int cz1 = min(cg.nz - 1, (int)floorf((zhi_q + R - cg.z0) * cg.inv_c));
      const int cy0 = max(0, (int)floorf((ylo_q - R - cg.y0) * cg.inv_c));
      const int cy1 = min(cg.ny - 1, (int)floorf((yhi_q + R - cg.y0) * cg.inv_c));
      const int nyc = cy1 - cy0 + 1, nzc = cz1 - cz0 + 1;
      const int nrows = (nzc > 0 && nyc > 0) ? nzc * nyc : 0;
      // up to 8 x 8 rows (cells of 3.5 A or more): lane = 8 z + y, no division
      const bool grid8 = nyc <= 8 && nzc <= 8;
      float sq[NA], s12[NA], s6[NA], fe[NA], fx[NA], fy[NA], fz[NA];
#pragma unroll
      for (int u = 0; u < NA; ++u) sq[u] = s12[u] = s6[u] = fe[u] = fx[u] = fy[u] = fz[u] = 0.0f;
      for (int base_row = 0; base_row < nrows; base_row += kWave) {
        int r0 = 0, len = 0;
        const int idx = base_row + lane;
        const bool in = grid8 ? (base_row == 0 && (lane >> 3) < nzc && (lane & 7) < nyc) : idx < nrows;
        if (in) {
          const int cz = grid8 ? cz0 + (lane >> 3) : cz0 + idx / nyc;
          const int cy = grid8 ? cy0 + (lane & 7) : cy0 + idx % nyc;
          const float zlo = cg.z0 + cz * cg.c, ylo = cg.y0 + cy * cg.c;
          float xlo = 3.0e38f, xhi = -3.0e38f;
#pragma unroll
          for (int u = 0; u < NA; ++u) {
            const float dzb = fmaxf(0.0f, fmaxf(zlo - qz[u], qz[u] - (zlo + cg.c)));
            const float dyb = fmaxf(0.0f, fmaxf(ylo - qy[u], qy[u] - (ylo + cg.c)));
            const float rx2 = R2 - dzb * dzb - dyb * dyb;
            if (rx2 >= 0.0f) {
              const float rx = bound_sqrt(rx2);
              xlo = fminf(xlo, qx[u] - rx);
              xhi = fmaxf(xhi, qx[u] + rx);
            }
          }
          if (xlo <= xhi) {
            const int cx0 = max(0, (int)floorf((xlo - cg.x0) * cg.inv_cx));
            const int cx1 = min(cg.nx - 1, (int)floorf((xhi - cg.x0) * cg.inv_cx));
            if (cx0 <= cx1) {
              const int rowb = (cz * cg.ny + cy) * cg.nx;
              r0 = s_cell[rowb + cx0];
              len = s_cell[rowb + cx1 + 1] - r0;
            }
          }
        }
        // inclusive prefix sum of the row lengths: row k covers [incl - len, incl)
        const int incl = wave_incl_scan(len);
        const int total = __builtin_amdgcn_readlane(incl, kWave - 1);
        if (total == 0) {
          if (grid8) break;
          continue;
        }
        const uint64_t nonempty = __ballot(len > 0);
        const int nc = __popcll(nonempty);
        // this lane's segment [t, t_end) of the concatenation
        const int m = (total + kWave - 1) / kWave;
        int t = lane * m;
        const int t_end = min(total, t + m);
        // first lane whose segment starts at or after this row's start; only the first
        // non-empty row starts at 0 (lane 0), and a row that no segment starts in (first
        // lane 64) is reached by walking
        const int first_lane = ceil_div_small(incl - len, m, __builtin_amdgcn_rcpf((float)m));
        wave_lds_sync();            // the previous rows' readers are done
        if (len > 0) {
          const int c = __popcll(nonempty & below);
          s_r0[c] = r0;
          s_beg[c] = incl - len;
        }
        if (lane == 0) s_beg[nc] = total;
        if (lane > 0) s_cnt[lane] = 0;
        wave_lds_sync();
        if (len > 0 && first_lane > 0 && first_lane < kWave) atomicAdd(&s_cnt[first_lane], 1);
        wave_lds_sync();
        // the row holding t: k = (rows starting at or before t) - 1, a prefix sum over
        // the lanes of how many rows first reach each lane (1 at lane 0: row 0)
        const int k = wave_incl_scan(lane > 0 ? s_cnt[lane] : 1) - 1;
        int i = s_r0[k] + (t - s_beg[k]);
        int row_end = s_beg[k + 1];
        int kw = k;
        for (int it = 0; it < m; ++it, ++t) {
          if (t < t_end) {
            if (t == row_end) {       // the next compacted row starts here
              ++kw;
              i = s_r0[kw];
              row_end = s_beg[kw + 1];
            }
            const RecPos p = s_pos[i];
            const RecAB ab = s_ab[i];
#pragma unroll
            for (int u = 0; u < NA; ++u) {
              if (GRAD)
                pair_grad_shifted(p, ab, lv[u], i, pk, c1[u], fa[u], fb[u], fe[u], fx[u], fy[u],
                                  fz[u]);
              else
                pair_energy<MMO_SHIFTED, NEAR>(p, ab, lv[u], i, pk, sq[u], s12[u], s6[u]);
            }
            ++i;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < NA; ++u) {
        if (!live[u]) continue;
        if (GRAD) {
          e_vd += (double)fe[u];
          const float gx = wave_sum_f(fx[u]), gy = wave_sum_f(fy[u]), gz = wave_sum_f(fz[u]);
          if (PRELOAD) {
            gxm = lane == jj + u ? gx : gxm;
            gym = lane == jj + u ? gy : gym;
            gzm = lane == jj + u ? gz : gzm;
          } else if (lane == 0) {
            const double gs = -2.0 * (double)fr.inv_s;
            gout(jb + jj, gs * (double)gx, gs * (double)gy, gs * (double)gz);
          }
        } else {
          e_el = fma((double)sq[u], cq[u], e_el);
          e_vd = fma((double)s12[u], ca[u], e_vd);
          e_vd = fma(-2.0 * (double)s6[u], cb[u], e_vd);
        }
      }
    }
    if (GRAD && PRELOAD && jb + lane < n_atoms) {
      const double gs = -2.0 * (double)fr.inv_s;
      gout(jb + lane, gs * (double)gxm, gs * (double)gym, gs * (double)gzm);
    }
  }
}

}  // namespace mmo
