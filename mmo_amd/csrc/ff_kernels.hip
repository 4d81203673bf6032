// ff_kernels.hip -- gfx950 kernels for the UFF nonbonded scoring path.
//
//   k_inter   ligand-receptor LJ 12-6 + Coulomb, direct sum over the receptor
//             (MODE 1: shifted, src/mol.ml:822-849 / :928-960; MODE 0: global,
//             src/mol.ml:796-818).  The receptor is staged through LDS in tiles and
//             read by broadcast; ligand atoms of a wavefront's pose tile sit in VGPRs
//             (K per lane) as 32-bit fixed point, so x_rec - x_lig is exact in integer
//             arithmetic.  Pair math is FP32, chunks of kChunk receptor atoms are
//             folded into FP64 per-pose sums.
//   k_intra   intra-ligand nonbonded energy, src/mol.ml:881-903, FP64, one workgroup
//             per pose over the precomputed list of pairs with topo distance >= 3.
//   k_interp  trilinear interpolation in per-type FP32 grids, src/G3D.ml:97-157 and
//             src/mol.ml:1012-1020, FP64 weights.
// Every kernel has an adjoint (GRAD) variant giving dE/dxyz of the ligand atoms.
#include <algorithm>

#include "common.h"
#include "cull.h"
#include "dof.h"
#include "kernels.h"

namespace mmo {

constexpr int kChunk = 64;   // receptor atoms per FP32 partial sum

// ------------------------------------------------------------------ pose sources
struct CoordSrc {
  const double* xyz;      // rows (x, y, z), pose-major
  const int* atom_off;    // first row of each pose
  const Tile* tiles;
  const int* perm;        // tile slot -> pose index (poses grouped by ligand)
  __device__ Tile tile(int t) const { return tiles[t]; }
  __device__ int pose_id(int slot) const { return perm[slot]; }
  __device__ void pos(int pose, int /*pslot*/, int j, const LigInfo&, const LigAtom*,
                      const double* /*lig_xyz*/, double& x, double& y, double& z) const {
    const double* r = xyz + 3 * (size_t)(atom_off[pose] + j);
    x = r[0]; y = r[1]; z = r[2];
  }
  __device__ int row(int pose, int j) const { return atom_off[pose] + j; }
};

// rigid scan: pose = rot * ntrans + trans, x = R c + T (src/lds.ml:1081-1094).
// Tiles walk the translations in the order tperm (Morton order of their positions),
// so the 64 poses of a wavefront sit in a compact brick of space.
struct ScanSrc {
  const double* rots;     // nrot x 9
  const double* trans;    // ntrans x 3
  const int* tperm;       // tile slot -> translation index
  int ntrans, tiles_per_rot, lig;
  __device__ Tile tile(int t) const {
    const int r = t / tiles_per_rot, tb = t - r * tiles_per_rot;
    Tile o;
    o.lig = lig;
    o.pose0 = r * ntrans + tb * kWave;
    o.npose = min(kWave, ntrans - tb * kWave);
    o.split = 1;
    return o;
  }
  __device__ int pose_id(int slot) const {
    const int r = slot / ntrans;
    return r * ntrans + tperm[slot - r * ntrans];
  }
  __device__ void pos(int pose, int, int j, const LigInfo& li, const LigAtom*,
                      const double* lig_xyz, double& x, double& y, double& z) const {
    const int r = pose / ntrans, t = pose - r * ntrans;
    const double* R = rots + 9 * r;
    const double* T = trans + 3 * t;
    const double* c = lig_xyz + 3 * (size_t)(li.atom_base + j);
    // Rot.rotate (src/rot.ml:97-100) then translate
    x = R[0] * c[0] + R[1] * c[1] + R[2] * c[2] + T[0];
    y = R[3] * c[0] + R[4] * c[1] + R[5] * c[2] + T[1];
    z = R[6] * c[0] + R[7] * c[1] + R[8] * c[2] + T[2];
  }
  __device__ int row(int, int) const { return 0; }
};

// the kept poses of a filtered scan: entry i of `list` is a slot of the scan's pose order
// (rotation * ntrans + Morton slot of the translation), so consecutive entries keep the
// scan's compact bricks; tiles hold 64 consecutive entries
struct ScanListSrc : ScanSrc {
  const int* list;
  int nlist;
  __device__ Tile tile(int t) const {
    Tile o;
    o.lig = lig;
    o.pose0 = t * kWave;
    o.npose = min(kWave, nlist - t * kWave);
    o.split = 1;
    return o;
  }
  __device__ int pose_id(int i) const { return ScanSrc::pose_id(list[i]); }
};

template <class Src>
struct InterArgs {
  Src src;
  int ntiles;
  Frame fr;
  const RecPos* rpos;
  const RecAB* rab;
  const RecRes* rres;
  int nrec, rtile;
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;
  double* e_out;      // e_out[pose * e_stride + 0] = E_inter
  int e_stride;
  double* e_comp;     // nullable: [pose*e_stride + 2] elec, [+3] vdW (same array as e_out)
  double* grad;       // GRAD: rows (x, y, z)
  int write_e;        // GRAD kernel: also write E
  int rslice;         // receptor atoms per gridDim.y slice
  int npose_total;    // for partial outputs when gridDim.y > 1
  const int* atom_order;   // nullable: per ligand at atom_base, the sweeps' atom order
};

// --------------------------------------------------------------- pair arithmetic
// energy + dE/dr2 contribution; gradient accumulated as coef * (x_rec - x_lig) in
// fixed-point units (caller scales by -2 * 2^-k).
template <int MODE>
__device__ __forceinline__ void pair_grad(const RecPos& p, const RecAB& ab, const LigPt& l,
                                          int i, const PairK& pk, float c1, float aj, float bj,
                                          float& e, float& gx, float& gy, float& gz) {
  const float dx = (float)(p.x - l.x), dy = (float)(p.y - l.y), dz = (float)(p.z - l.z);
  const float r2 = pair_r2(dx, dy, dz, l, i, pk);
  const float r2c = max_r2(r2, pk.min_r2);
  const float ir = __builtin_amdgcn_rsqf(r2c);
  const float u = inv_r2(ir, r2c);
  const float u3 = u * u * u;
  const float u6 = u3 * u3;
  const float eq = p.q * ir, e12 = ab.a * u6, e6 = ab.b * u3;
  // F = ke qi qj / r + Aij u^6 - 2 Bij u^3 ;  dF/dr2 = -u (ke qi qj/(2r) + 6 Aij u^6 - 6 Bij u^3)
  const float F = fmaf(c1, eq, fmaf(aj, e12, -2.0f * bj * e6));
  const float G = fmaf(0.5f * c1, eq, 6.0f * fmaf(aj, e12, -bj * e6));
  float w, coef;
  if (MODE == MMO_SHIFTED) {
    const float t = shift_t(r2c, pk);
    w = t * t;
    coef = fmaf(t * F, -1.0f / 72.0f, -(w * u) * G);   // dw/dr2 = -t/72
  } else {
    w = 1.0f;
    coef = -u * G;
  }
  coef = (r2 >= pk.min_r2) ? coef : 0.0f;   // clamped distance: flat
  e = fmaf(w, F, e);
  gx = fmaf(coef, dx, gx);
  gy = fmaf(coef, dy, gy);
  gz = fmaf(coef, dz, gz);
}

// Global form for a ligand atom outside the fixed-point window (more than ~2^(30-k) A
// from the frame origin): FP64 pairs over the receptor tile from the atom's caller-frame
// position relative to the frame origin, like the reference's all-pairs loop
// (src/mol.ml:796-818, no cutoff, r = max(r, 0.01)).  Such an atom is far from every
// receptor atom, so the pair terms are smooth and need no fixed-point differences.
// With GRAD, (gx, gy, gz) receive dE/dx of the atom in A^-1.
template <bool GRAD>
__device__ void far_atom_pairs(const RecPos* __restrict__ s_pos, const RecAB* __restrict__ s_ab,
                               int nt, double inv_scale, double xr, double yr, double zr,
                               double keq, double aj, double bj, double& e_el, double& e_vd,
                               double& gx, double& gy, double& gz) {
  double el = 0.0, vd = 0.0;
  for (int i = 0; i < nt; ++i) {
    const RecPos p = s_pos[i];
    const RecAB ab = s_ab[i];
    const double dx = (double)p.x * inv_scale - xr, dy = (double)p.y * inv_scale - yr,
                 dz = (double)p.z * inv_scale - zr;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rr = sqrt(r2);
    const bool apart = rr >= 0.01;
    const double r = apart ? rr : 0.01;
    const double ir = 1.0 / r, u = ir * ir, u3 = u * u * u;
    const double eq = (double)p.q * ir, e12 = aj * (double)ab.a * u3 * u3,
                 e6 = bj * (double)ab.b * u3;
    el += eq;
    vd += e12 - 2.0 * e6;
    if (GRAD && apart) {
      // dE/dr / r, times (x_lig - x_rec) = -d
      const double c = -(keq * eq + 12.0 * e12 - 12.0 * e6) * u;
      gx -= c * dx; gy -= c * dy; gz -= c * dz;
    }
  }
  e_el = fma(keq, el, e_el);
  e_vd += vd;
}

// ------------------------------------------------------------------ inter kernel
template <int MODE, int K, bool GRAD, class Src>
__global__ void __launch_bounds__(256) k_inter(InterArgs<Src> a) {
  extern __shared__ __align__(16) char smem[];
  RecPos* s_pos = reinterpret_cast<RecPos*>(smem);
  RecAB* s_ab = reinterpret_cast<RecAB*>(smem + (size_t)a.rtile * sizeof(RecPos));
  const int lane = threadIdx.x & (kWave - 1);
  const int nw = blockDim.x / kWave;
  const int tidx = blockIdx.x * nw + (threadIdx.x / kWave);
  const bool tile_ok = tidx < a.ntiles;
  Tile t;
  if (tile_ok) t = a.src.tile(tidx);
  else { t.lig = 0; t.pose0 = 0; t.npose = 0; t.split = kWave; }
  const int split = t.split;
  const int pslot = lane / split, aslot = lane - pslot * split;
  const bool pose_ok = tile_ok && pslot < t.npose;
  const int pose = pose_ok ? a.src.pose_id(t.pose0 + pslot) : 0;
  const LigInfo li = a.ligs[t.lig];
  const int per_lane = (li.n_atoms + split - 1) / split;
  const int nblk = (per_lane + K - 1) / K;
  const Frame fr = a.fr;
  const PairK pk = pair_k(fr.inv_s2, fr.inv_s, a.rres);
  double e_el = 0.0, e_vd = 0.0;
  const int r_begin = blockIdx.y * a.rslice;
  const int r_end = min(a.nrec, r_begin + a.rslice);
  const bool split_rec = gridDim.y > 1;

  for (int rb = r_begin; rb < r_end; rb += a.rtile) {
    const int nt = min(a.rtile, r_end - rb);
    __syncthreads();
    for (int i = threadIdx.x; i < nt; i += blockDim.x) {
      s_pos[i] = a.rpos[rb + i];
      s_ab[i] = a.rab[rb + i];
    }
    __syncthreads();
    if (!tile_ok) continue;
    for (int blk = 0; blk < nblk; ++blk) {
      LigPt l[K];
      double cq[K], ca[K], cb[K];
      float c1[K], fa[K], fb[K];
      int row[K];
      bool far[K];      // GLOBAL: outside the fixed-point window, summed by far_atom_pairs
      bool any_far = false;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const int j = aslot + split * (blk * K + k);
        far[k] = false;
        if (pose_ok && j < li.n_atoms) {
          double x, y, z;
          a.src.pos(pose, pslot, j, li, a.latoms, a.lig_xyz, x, y, z);
          int satk = 0;
          l[k] = lig_pt(x, y, z, fr, &satk);
          const LigAtom la = a.latoms[li.atom_base + j];
          cq[k] = la.keq; ca[k] = la.a; cb[k] = la.b;
          row[k] = GRAD ? a.src.row(pose, j) : 0;
          if (MODE == MMO_GLOBAL && satk) {
            // the saturated position takes no part in the fixed-point pairs
            far[k] = any_far = true;
            cq[k] = ca[k] = cb[k] = 0.0;
          }
        } else {
          l[k] = lig_pt_zero();
          cq[k] = ca[k] = cb[k] = 0.0;
          row[k] = -1;
        }
        c1[k] = (float)cq[k]; fa[k] = (float)ca[k]; fb[k] = (float)cb[k];
      }
      // GLOBAL: FP64 pairs of the far atoms over this receptor tile (wave-uniform branch)
      double far_gx[K], far_gy[K], far_gz[K];
#pragma unroll
      for (int k = 0; k < K; ++k) far_gx[k] = far_gy[k] = far_gz[k] = 0.0;
      if (MODE == MMO_GLOBAL && __ballot(any_far) != 0ull) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!far[k]) continue;
          const int j = aslot + split * (blk * K + k);
          double x, y, z;
          a.src.pos(pose, pslot, j, li, a.latoms, a.lig_xyz, x, y, z);
          const LigAtom la = a.latoms[li.atom_base + j];
          far_atom_pairs<GRAD>(s_pos, s_ab, nt, 1.0 / fr.scale, x - fr.ox, y - fr.oy, z - fr.oz,
                               la.keq, la.a, la.b, e_el, e_vd, far_gx[k], far_gy[k], far_gz[k]);
        }
      }
      if (!GRAD) {
        for (int c0 = 0; c0 < nt; c0 += kChunk) {
          const int c1e = min(nt, c0 + kChunk);
          float sq[K], s12[K], s6[K];
#pragma unroll
          for (int k = 0; k < K; ++k) sq[k] = s12[k] = s6[k] = 0.0f;
#pragma unroll 2
          for (int i = c0; i < c1e; ++i) {
            const RecPos p = s_pos[i];
            const RecAB ab = s_ab[i];
#pragma unroll
            for (int k = 0; k < K; ++k)
              pair_energy<MODE>(p, ab, l[k], rb + i, pk, sq[k], s12[k], s6[k]);
          }
#pragma unroll
          for (int k = 0; k < K; ++k) {
            e_el = fma((double)sq[k], cq[k], e_el);
            e_vd = fma((double)s12[k], ca[k], e_vd);
            e_vd = fma(-2.0 * (double)s6[k], cb[k], e_vd);
          }
        }
      } else {
        double gx[K], gy[K], gz[K];
#pragma unroll
        for (int k = 0; k < K; ++k) gx[k] = gy[k] = gz[k] = 0.0;
        for (int c0 = 0; c0 < nt; c0 += kChunk) {
          const int c1e = min(nt, c0 + kChunk);
          float fe[K], fx[K], fy[K], fz[K];
#pragma unroll
          for (int k = 0; k < K; ++k) fe[k] = fx[k] = fy[k] = fz[k] = 0.0f;
          for (int i = c0; i < c1e; ++i) {
            const RecPos p = s_pos[i];
            const RecAB ab = s_ab[i];
#pragma unroll
            for (int k = 0; k < K; ++k)
              pair_grad<MODE>(p, ab, l[k], rb + i, pk, c1[k], fa[k], fb[k], fe[k],
                              fx[k], fy[k], fz[k]);
          }
#pragma unroll
          for (int k = 0; k < K; ++k) {
            e_vd += (double)fe[k];
            gx[k] += (double)fx[k]; gy[k] += (double)fy[k]; gz[k] += (double)fz[k];
          }
        }
        // dE/dx_lig = sum coef * d(r2)/dx_lig = -2 * sum coef * (x_rec - x_lig)
        const double gs = -2.0 * (double)fr.inv_s;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (MODE == MMO_GLOBAL && far[k]) {   // A^-1 -> the units scaled by gs below
            gx[k] += far_gx[k] / gs; gy[k] += far_gy[k] / gs; gz[k] += far_gz[k] / gs;
          }
          if (row[k] >= 0) {
            double* g = a.grad + 3 * (size_t)row[k];
            if (split_rec) {   // grad pre-zeroed, slices add concurrently
              unsafeAtomicAdd(&g[0], gs * gx[k]);
              unsafeAtomicAdd(&g[1], gs * gy[k]);
              unsafeAtomicAdd(&g[2], gs * gz[k]);
            } else if (rb == r_begin) {
              g[0] = gs * gx[k]; g[1] = gs * gy[k]; g[2] = gs * gz[k];
            } else {
              g[0] += gs * gx[k]; g[1] += gs * gy[k]; g[2] += gs * gz[k];
            }
          }
        }
      }
    }
  }
  if (!tile_ok) return;
  e_el = group_sum(e_el, split);
  e_vd = group_sum(e_vd, split);
  if (pose_ok && aslot == 0) {
    // saturated atoms stay beyond the cutoff of every receptor atom in the shifted form;
    // the global form sums them through far_atom_pairs
    const double e = e_el + e_vd;
    // receptor slices write partial sums to slice-major buffers reduced afterwards
    const size_t o = (size_t)blockIdx.y * a.npose_total * a.e_stride + (size_t)pose * a.e_stride;
    if (!GRAD || a.write_e) a.e_out[o] = e;
    if (!GRAD && a.e_comp) {
      a.e_comp[o + 2] = e_el;
      a.e_comp[o + 3] = e_vd;
    }
  }
}

// RMODE: 0 receptor read in place (scalar cache), 1 receptor + cells in LDS, 2 positions
// and cells in LDS with the LJ factors read in place
template <int K, int RMODE, class Src>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(K == 1 ? 8 : 1)))
k_inter_cull(InterArgs<Src> a, CellGrid cg) {
  extern __shared__ __align__(16) char smem[];
  const RecPos* s_pos;
  const RecAB* s_ab;
  const int* s_cell;
  if constexpr (RMODE == 2)
    load_receptor_pos(a.rpos, a.rab, a.nrec, cg, s_pos, s_ab, s_cell, smem);
  else
    load_receptor<RMODE == 1>(a.rpos, a.rab, a.nrec, cg, s_pos, s_ab, s_cell, smem);

  const int lane = threadIdx.x & (kWave - 1);
  const int tidx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
  if (tidx >= a.ntiles) return;
  const Tile t = a.src.tile(tidx);
  const int split = t.split;
  const int pslot = lane / split, aslot = lane - pslot * split;
  const bool pose_ok = pslot < t.npose;
  const int pose = pose_ok ? a.src.pose_id(t.pose0 + pslot) : 0;
  const LigInfo li = a.ligs[t.lig];
  const int per_lane = (li.n_atoms + split - 1) / split;
  const int nblk = (per_lane + K - 1) / K;
  const Frame fr = a.fr;
  double e_el = 0.0, e_vd = 0.0;
  int sat = 0;

  for (int blk = 0; blk < nblk; ++blk) {
    LigPt l[K];
    bool valid[K];
    double cq[K], ca[K], cb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = aslot + split * (blk * K + k);
      valid[k] = pose_ok && j < li.n_atoms;
      if (valid[k]) {
        double x, y, z;
        a.src.pos(pose, pslot, j, li, a.latoms, a.lig_xyz, x, y, z);
        l[k] = lig_pt(x, y, z, fr, &sat);
        const LigAtom la = a.latoms[li.atom_base + j];
        cq[k] = la.keq; ca[k] = la.a; cb[k] = la.b;
      } else {   // inert: zero parameters
        l[k] = lig_pt_far();
        cq[k] = ca[k] = cb[k] = 0.0;
      }
    }
    float qx, qy, qz, hx, hy, hz;
    if (!wave_box<K>(l, valid, fr.inv_s, qx, qy, qz, hx, hy, hz)) continue;
    cull_sweep<K>(cg, s_cell, s_pos, s_ab, a.rres, fr.inv_s, fr.inv_s2, qx, qy, qz, hx, hy, hz, l,
                  [&](const float (&sq)[K], const float (&s12)[K], const float (&s6)[K]) {
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                      e_el = fma((double)sq[k], cq[k], e_el);
                      e_vd = fma((double)s12[k], ca[k], e_vd);
                      e_vd = fma(-2.0 * (double)s6[k], cb[k], e_vd);
                    }
                  });
  }
  e_el = group_sum(e_el, split);
  e_vd = group_sum(e_vd, split);
  (void)sat;   // saturated atoms stay beyond the cutoff of every receptor atom
  if (pose_ok && aslot == 0) {
    const size_t o = (size_t)pose * a.e_stride;
    a.e_out[o] = e_el + e_vd;
    if (a.e_comp) {
      a.e_comp[o + 2] = e_el;
      a.e_comp[o + 3] = e_vd;
    }
  }
}

// Rigid scan over the row lists of a finer cell grid (row_sweep): one pose per lane, K
// ligand atoms at a time (consecutive atoms of the ligand's sweep order, a walk of its
// bond graph, so the K atoms are close and the wavefront's box grows little).  RMODE bit
// 0 clear: receptor read in place (uniform addresses, scalar cache); set: receptor
// positions and LJ factors staged in LDS.  Bit 1: the far rows' atoms are tested against
// the box grown by the cutoff before the sweep (row_sweep's FILT).  The cell table stays
// in global memory (it is read per lane, once per candidate row).
template <int RMODE, int CHUNK, int K, class Src>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8)))
k_scan_rows(InterArgs<Src> a, CellGrid cg) {
  extern __shared__ __align__(16) char smem[];
  const RecPos* s_pos = a.rpos;
  const RecAB* s_ab = a.rab;
  if constexpr ((RMODE & 1) == 1) {
    // with the padding atoms (the row sweep's chunks may read past the last atom)
    const int nr = a.nrec + kRowChunk;
    RecPos* p = reinterpret_cast<RecPos*>(smem);
    RecAB* b = reinterpret_cast<RecAB*>(smem + (size_t)nr * sizeof(RecPos));
    for (int i = threadIdx.x; i < nr; i += blockDim.x) {
      p[i] = a.rpos[i];
      b[i] = a.rab[i];
    }
    __syncthreads();
    s_pos = p;
    s_ab = b;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int tidx = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
  if (tidx >= a.ntiles) return;
  const Tile t = a.src.tile(tidx);
  const bool pose_ok = lane < t.npose;
  const int pose = pose_ok ? a.src.pose_id(t.pose0 + lane) : 0;
  const LigInfo li = a.ligs[t.lig];
  const Frame fr = a.fr;
  double e_el = 0.0, e_vd = 0.0;
  int sat = 0;
  for (int j0 = 0; j0 < li.n_atoms; j0 += K) {
    LigPt l[K];
    bool valid[K];
    double cq[K], ca[K], cb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      valid[k] = pose_ok && j0 + k < li.n_atoms;
      cq[k] = ca[k] = cb[k] = 0.0;
      l[k] = lig_pt_far();
      if (j0 + k < li.n_atoms) {
        const int j = (K > 1 && a.atom_order) ? a.atom_order[li.atom_base + j0 + k] : j0 + k;
        const LigAtom la = a.latoms[li.atom_base + j];
        if (pose_ok) {
          double x, y, z;
          a.src.pos(pose, lane, j, li, a.latoms, a.lig_xyz, x, y, z);
          l[k] = lig_pt(x, y, z, fr, &sat);
          cq[k] = la.keq; ca[k] = la.a; cb[k] = la.b;
        }
      }
    }
    float qx, qy, qz, hx, hy, hz;
    if (!wave_box<K>(l, valid, fr.inv_s, qx, qy, qz, hx, hy, hz)) continue;
    row_sweep<(RMODE & 1) == 0, CHUNK, K, (RMODE & 2) != 0>(cg, s_pos, s_ab, a.rres, fr.inv_s, fr.inv_s2, qx, qy, qz, hx,
                                    hy, hz, l,
                                    [&](const float (&sq)[K], const float (&s12)[K],
                                        const float (&s6)[K]) {
#pragma unroll
                                      for (int k = 0; k < K; ++k) {
                                        e_el = fma((double)sq[k], cq[k], e_el);
                                        e_vd = fma((double)s12[k], ca[k], e_vd);
                                        e_vd = fma(-2.0 * (double)s6[k], cb[k], e_vd);
                                      }
                                    });
  }
  (void)sat;   // saturated atoms stay beyond the cutoff of every receptor atom
  if (pose_ok) {
    const size_t o = (size_t)pose * a.e_stride;
    a.e_out[o] = e_el + e_vd;
    if (a.e_comp) {
      a.e_comp[o + 2] = e_el;
      a.e_comp[o + 3] = e_vd;
    }
  }
}

// The cell-packed copy of the grids (GridDev::cells): one thread per (type, cell) gathers
// the cell's 8 corners and writes them as two 16-byte stores.
__global__ void __launch_bounds__(256) k_pack_cells(GridDev g, int ntypes, float* cells) {
  const size_t cx = g.nx - 1, cy = g.ny - 1, cz = g.nz - 1;
  const size_t ncell = cx * cy * cz;
  const size_t n = ncell * ntypes;
  for (size_t u = blockIdx.x * (size_t)blockDim.x + threadIdx.x; u < n;
       u += (size_t)gridDim.x * blockDim.x) {
    const size_t t = u / ncell, c = u - t * ncell;
    const int k = (int)(c / (cx * cy)), j = (int)((c / cx) % cy), i = (int)(c % cx);
    float v[8];
    grid_corners<false>(g, (int)t, i, j, k, v);
    float4* q = reinterpret_cast<float4*>(cells + 8 * u);
    q[0] = make_float4(v[0], v[1], v[2], v[3]);
    q[1] = make_float4(v[4], v[5], v[6], v[7]);
  }
}

hipError_t launch_pack_cells(const GridDev& g, int ntypes, float* cells, hipStream_t s) {
  const size_t n = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1) * ntypes;
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_pack_cells, dim3(blocks), dim3(256), 0, s, g, ntypes, cells);
  return hipGetLastError();
}

// ----------------------------------------------------- inter kernel, pose-culled
// ligand atoms per pass of the per-atom sweep
constexpr int kPoseSweepAtoms = 4;

// Shifted form for arbitrary poses (batches, gradients, minimisation): one wavefront per
// pose; the receptor is swept per 8-atom sub-block of the pose (pose_sweep) or per atom
// (atom_sweep), so the cost follows the pairs inside the cutoff wherever the poses are.
// With an order, slot s of the launch (workgroup s / waves, wavefront s % waves) scores
// pose order[s]: the host deals the poses, largest ligands first, so that the SIMDs
// receive equal shares of ligand atoms instead of whatever the batch order gives them.
template <bool GRAD, bool LDS, bool ATOMWISE>
__global__ void __launch_bounds__(1024) k_pose_cull(PoseCullArgs a, Frame fr, const RecPos* rpos,
                                                    const RecAB* rab, int nrec, CellGrid cg) {
  extern __shared__ __align__(16) char smem[];
  const RecPos* s_pos;
  const RecAB* s_ab;
  const int* s_cell;
  load_receptor<LDS>(rpos, rab, nrec, cg, s_pos, s_ab, s_cell, smem);
  const int lane = threadIdx.x & (kWave - 1);
  // this wavefront's scratch for atom_sweep, after the receptor copy
  int* wl = reinterpret_cast<int*>(smem + receptor_lds_bytes<LDS>(nrec, cg)) +
            (threadIdx.x / kWave) * kAtomSweepInts;
  auto score = [&](int p) {
    const LigInfo li = a.ligs[a.pose_lig[p]];
    const size_t row0 = (size_t)a.atom_off[p];
    const double* x = a.xyz + 3 * row0;
    int sat = 0;
    double e_el = 0.0, e_vd = 0.0;
    // sweep position j -> ligand atom (the per-atom sweep's order, if given)
    auto at = [&](int j) { return (ATOMWISE && a.atom_order) ? a.atom_order[li.atom_base + j] : j; };
    auto atom = [&](int j, LigPt& l, double& cq, double& ca, double& cb) {
      const int ja = at(j);
      l = lig_pt(x[3 * ja], x[3 * ja + 1], x[3 * ja + 2], fr, &sat);
      const LigAtom la = a.latoms[li.atom_base + ja];
      cq = la.keq; ca = la.a; cb = la.b;
    };
    auto gout = [&](int j, double gx, double gy, double gz) {
      double* g = a.grad + 3 * (row0 + at(j));
      if (a.grad_assign) {
        g[0] = gx; g[1] = gy; g[2] = gz;
      } else {
        g[0] += gx; g[1] += gy; g[2] += gz;
      }
    };
    if (ATOMWISE)
      atom_sweep<GRAD, decltype(atom)&, decltype(gout)&, true, true, kPoseSweepAtoms>(
          cg, s_cell, s_pos, s_ab, a.rres, fr, li.n_atoms, wl, atom, gout, e_el, e_vd);
    else
      pose_sweep<GRAD>(cg, s_cell, s_pos, s_ab, a.rres, fr, li.n_atoms, atom, gout, e_el, e_vd);
    (void)sat;   // saturated atoms stay beyond the cutoff of every receptor atom
    // every lane of an atom slot holds part of the sums: the wavefront sum counts each
    // pair once (lanes of one slot share the atom, split the receptor atoms)
    e_el = wave_sum(e_el);
    e_vd = wave_sum(e_vd);
    if (lane == 0) {
      const size_t o = (size_t)p * a.e_stride;
      if (!GRAD || a.write_e) a.e_out[o] = e_el + e_vd;
      if (!GRAD && a.e_comp) {
        a.e_comp[o + 2] = e_el;
        a.e_comp[o + 3] = e_vd;
      }
    }
  };
  const int slot = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
  if (slot < a.npose) score(a.order ? a.order[slot] : slot);
}

hipError_t launch_pose_cull(const InterLaunch& L, const PoseCullArgs& a, bool grad, int waves,
                            bool atomwise, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  waves = std::max(1, std::min(16, waves));
  const int blocks = (a.npose + waves - 1) / waves;
  const bool in_lds = L.cull == 1;
  const size_t lds = (in_lds ? receptor_lds_bytes<true>(L.nrec, L.cg) : 0) +
                     (size_t)waves * kAtomSweepInts * sizeof(int);
#define MMO_POSE_CULL(G, A)                                                                    \
  if (in_lds) {                                                                                \
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pose_cull<G, true, A>), \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);  \
    if (e != hipSuccess) return e;                                                             \
    hipLaunchKernelGGL((k_pose_cull<G, true, A>), dim3(blocks), dim3(waves * kWave), lds, s,   \
                       a, L.fr, L.rpos, L.rab, L.nrec, L.cg);                                  \
  } else {                                                                                     \
    hipLaunchKernelGGL((k_pose_cull<G, false, A>), dim3(blocks), dim3(waves * kWave), lds, s,  \
                       a, L.fr, L.rpos, L.rab, L.nrec, L.cg);                                  \
  }
  if (grad && atomwise) {
    MMO_POSE_CULL(true, true)
  } else if (grad) {
    MMO_POSE_CULL(true, false)
  } else if (atomwise) {
    MMO_POSE_CULL(false, true)
  } else {
    MMO_POSE_CULL(false, false)
  }
#undef MMO_POSE_CULL
  return hipGetLastError();
}

// ------------------------------------------------------------- grid precompute
// Per-type FP32 grids for single-atom probes (anum_t, q_t) at every grid point:
// E_t = min(max_E, ke q_t Sq + A_t S12 - 2 B_t S6) with the receptor sums of the
// shifted form (Mol.ene_inter_UFF_shifted_grid src/mol.ml:964-989, clamp
// src/lds.ml:466).  Points outside the ROI mask stay 0 (src/G3D.ml:47-51,
// src/lds.ml:269-305).  A wavefront handles a 4x4x4 brick of points.
template <bool LDS>
__global__ void __launch_bounds__(1024) k_grid(GridArgs a, CellGrid cg) {
  extern __shared__ __align__(16) char smem[];
  const RecPos* s_pos;
  const RecAB* s_ab;
  const int* s_cell;
  load_receptor<LDS>(a.rpos, a.rab, a.nrec, cg, s_pos, s_ab, s_cell, smem);
  const int lane = threadIdx.x & (kWave - 1);
  const int brick = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
  const int bx = (a.nx + 3) / 4, by = (a.ny + 3) / 4, bz = (a.nz + 3) / 4;
  if (brick >= bx * by * bz) return;
  const int ib = brick % bx, jb = (brick / bx) % by, kb = brick / (bx * by);
  const int i = ib * 4 + (lane & 3), j = jb * 4 + ((lane >> 2) & 3), k = kb * 4 + (lane >> 4);
  const bool inside = i < a.nx && j < a.ny && k < a.nz;
  // node position (Grid.t: node (i,j,k) at origin + (i,j,k) step, src/grid.ml:40-52)
  const double x = a.ox + i * a.step, y = a.oy + j * a.step, z = a.oz + k * a.step;
  bool valid[1];
  valid[0] = inside;
  if (inside && a.roi_r2 > 0.0) {
    const double dx = a.roi_x - x, dy = a.roi_y - y, dz = a.roi_z - z;
    valid[0] = (dx * dx + dy * dy + dz * dz) < a.roi_r2;
  }
  const Frame fr = a.fr;
  int sat = 0;
  LigPt l[1];
  l[0] = valid[0] ? lig_pt(x, y, z, fr, &sat) : lig_pt_far();
  double Sq = 0.0, S12 = 0.0, S6 = 0.0;
  float qx, qy, qz, hx, hy, hz;
  if (wave_box<1>(l, valid, fr.inv_s, qx, qy, qz, hx, hy, hz)) {
    cull_sweep<1, true>(cg, s_cell, s_pos, s_ab, a.rres, fr.inv_s, fr.inv_s2, qx, qy, qz, hx, hy, hz, l,
                        [&](const float (&sq)[1], const float (&s12)[1], const float (&s6)[1]) {
                    Sq += (double)sq[0];
                    S12 += (double)s12[0];
                    S6 += (double)s6[0];
                  });
  }
  if (!inside) return;
  const size_t idx = (size_t)i + (size_t)j * a.nx + (size_t)k * a.nx * a.ny;
  for (int t = 0; t < a.ntypes; ++t) {
    float v = 0.0f;
    if (valid[0]) {
      const double4 p = a.probe[t];   // ke q_t, A_t, B_t
      const double e = p.x * Sq + p.y * S12 - 2.0 * p.z * S6;
      v = (float)(e < a.max_e ? e : a.max_e);
    }
    a.out[(size_t)t * a.nvox + idx] = v;
  }
}

// ------------------------------------------------------------------ intra kernel
// One workgroup of kIntraWaves wavefronts per pose; the pairs of the ligand's list are
// spread over all its lanes, and each lane loads its next pair while it evaluates the
// current one, so a large ligand's list is not walked by one wavefront one dependent
// global load at a time.  Each wavefront adds its gradient terms into its own LDS copy
// (the lanes of one LDS atomic are applied in lane order) and the copies are summed in
// wavefront order, so the result does not depend on how the wavefronts interleave.
constexpr int kIntraWaves = 4;

template <bool GRAD>
__global__ void __launch_bounds__(kIntraWaves * kWave) k_intra(IntraArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double s_part[kIntraWaves];
  double* sx = reinterpret_cast<double*>(smem);
  const int pose = blockIdx.x;
  if (pose >= a.npose) return;
  const int tid = threadIdx.x, nt = blockDim.x;
  const LigInfo li = a.ligs[a.pose_lig[pose]];
  const int L = li.n_atoms;
  double* sy = sx + L;
  double* sz = sy + L;
  double* g0 = sz + L;                        // kIntraWaves copies of (gx, gy, gz)
  double* gx = g0 + 3 * L * (tid / kWave);    // this wavefront's copy
  double* gy = gx + L;
  double* gz = gy + L;
  const int base = a.atom_off[pose];
  for (int j = tid; j < L; j += nt) {
    const double* r = a.xyz + 3 * (size_t)(base + j);
    sx[j] = r[0]; sy[j] = r[1]; sz[j] = r[2];
  }
  if (GRAD)
    for (int j = tid; j < 3 * L * (nt / kWave); j += nt) g0[j] = 0.0;
  const uint32_t* pij = a.pair_ij + li.pair_base;
  const double4* ppar = a.pair_par + li.pair_base;
  const int n = li.n_pairs;
  uint32_t ij_next = tid < n ? pij[tid] : 0u;
  double4 pp_next = tid < n ? ppar[tid] : make_double4(0.0, 0.0, 0.0, 0.0);
  __syncthreads();
  double e = 0.0;
  for (int p = tid; p < n; p += nt) {
    const uint32_t ij = ij_next;
    const double4 pp = pp_next;   // ke qi qj, Ai Aj, 2 Bi Bj
    if (p + nt < n) {
      ij_next = pij[p + nt];
      pp_next = ppar[p + nt];
    }
    const int i = ij & 0xffff, j = ij >> 16;
    const double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const bool apart = r2 >= 1.0e-4;              // r >= 0.01 (Math.non_zero_dist)
    const double ir = apart ? rsqrt(r2) : 100.0;  // 1 / max(r, 0.01)
    const double u = ir * ir;
    const double u3 = u * u * u;
    const double u6 = u3 * u3;
    e += pp.x * ir + pp.y * u6 - pp.z * u3;
    if (GRAD && apart) {
      // dE/dr / r = -(ke qq / r^3) - 12 AA u^7 + 6 BB2 u^4
      const double c = (-pp.x * ir * u - 12.0 * pp.y * u6 * u + 6.0 * pp.z * u3 * u);
      atomicAdd(&gx[i], c * dx); atomicAdd(&gy[i], c * dy); atomicAdd(&gz[i], c * dz);
      atomicAdd(&gx[j], -c * dx); atomicAdd(&gy[j], -c * dy); atomicAdd(&gz[j], -c * dz);
    }
  }
  e = wave_sum(e);
  if ((tid & (kWave - 1)) == 0) s_part[tid / kWave] = e;
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int w = 0; w < nt / kWave; ++w) t += s_part[w];
    a.e_out[(size_t)pose * a.e_stride] = t;
  }
  if (GRAD) {
    for (int j = tid; j < L; j += nt) {
      double sum[3] = {0.0, 0.0, 0.0};
      for (int w = 0; w < nt / kWave; ++w)
        for (int c = 0; c < 3; ++c) sum[c] += g0[3 * L * w + c * L + j];
      double* g = a.grad + 3 * (size_t)(base + j);
      if (a.grad_accumulate) { g[0] += sum[0]; g[1] += sum[1]; g[2] += sum[2]; }
      else { g[0] = sum[0]; g[1] = sum[1]; g[2] = sum[2]; }
    }
  }
}

// ----------------------------------------------------------------- interp kernel
template <class Src>
struct InterpArgs {
  Src src;
  int ntiles;
  GridDev g;
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;
  const int* type_override;   // nullable: per-atom types of the plain form
  double* e_out;
  int e_stride;
  double* grad;
  int grad_accumulate;
};

// CELLS: read the cell-packed grids when there are any (batches: a pose's atoms across the
// lanes); the scans' lanes hold one atom of 64 neighbouring poses, which share grid nodes,
// and read the node arrays (the packed copy measured 15.1 against 10.4 ms on the
// 36.9M-pose scan)
template <bool GRAD, class Src, bool CELLS>
__global__ void __launch_bounds__(256) k_interp(InterpArgs<Src> a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int tidx = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (tidx >= a.ntiles) return;
  const Tile t = a.src.tile(tidx);
  const int split = t.split;
  const int pslot = lane / split, aslot = lane - pslot * split;
  const bool pose_ok = pslot < t.npose;
  const int pose = pose_ok ? a.src.pose_id(t.pose0 + pslot) : 0;
  const LigInfo li = a.ligs[t.lig];
  double e = 0.0;
  int bad = 0;
  if (pose_ok) {
    for (int j = aslot; j < li.n_atoms; j += split) {
      double x, y, z;
      a.src.pos(pose, pslot, j, li, a.latoms, a.lig_xyz, x, y, z);
      const int ty = a.type_override ? a.type_override[j] : a.latoms[li.atom_base + j].type;
      bool ok = true;
      double dg[3] = {0.0, 0.0, 0.0};
      e += trilin<GRAD, CELLS>(a.g, ty, x - a.g.ox, y - a.g.oy, z - a.g.oz, ok, dg);
      if (!ok) bad = 1;
      if (GRAD) {
        double* gr = a.grad + 3 * (size_t)a.src.row(pose, j);
        if (!ok) dg[0] = dg[1] = dg[2] = __builtin_nan("");
        if (a.grad_accumulate) { gr[0] += dg[0]; gr[1] += dg[1]; gr[2] += dg[2]; }
        else { gr[0] = dg[0]; gr[1] = dg[1]; gr[2] = dg[2]; }
      }
    }
  }
  e = group_sum(e, split);
  bad = group_or(bad, split);
  if (pose_ok && aslot == 0) a.e_out[(size_t)pose * a.e_stride] = bad ? __builtin_nan("") : e;
}

__global__ void k_fill_int(int* p, size_t n, int v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill(double* p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// ------------------------------------------------------------------------- launchers
template <int MODE, int K, bool GRAD, class Src>
static hipError_t launch_inter_t(const InterArgs<Src>& a, int waves, hipStream_t s) {
  const int blocks = (a.ntiles + waves - 1) / waves;
  if (blocks == 0) return hipSuccess;
  const int ny = (a.nrec + a.rslice - 1) / a.rslice;
  const size_t lds = (size_t)a.rtile * (sizeof(RecPos) + sizeof(RecAB));
  hipLaunchKernelGGL((k_inter<MODE, K, GRAD, Src>), dim3(blocks, ny), dim3(waves * kWave), lds, s, a);
  return hipGetLastError();
}

// sums the gridDim.y partial outputs (slice-major) into out, in slice order
__global__ void k_reduce_slices(const double* part, int nslice, size_t n, int stride, int ncomp,
                                double* out) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    for (int c = 0; c < ncomp; ++c) {
      if (ncomp == 1 && c > 0) break;
      const int cc = (ncomp == 1) ? 0 : (c == 0 ? 0 : c + 1);   // 0, 2, 3
      double acc = 0.0;
      for (int y = 0; y < nslice; ++y) acc += part[(size_t)y * n * stride + i * stride + cc];
      out[i * stride + cc] = acc;
    }
  }
}

template <class Src>
static hipError_t launch_inter_any(int mode, int k, bool grad, const InterArgs<Src>& a, int waves,
                                   hipStream_t s) {
#define MMO_K(KK)                                                                       \
  if (k == KK) {                                                                        \
    if (grad) return mode == MMO_SHIFTED ? launch_inter_t<MMO_SHIFTED, KK, true>(a, waves, s) \
                                         : launch_inter_t<MMO_GLOBAL, KK, true>(a, waves, s);  \
    return mode == MMO_SHIFTED ? launch_inter_t<MMO_SHIFTED, KK, false>(a, waves, s)           \
                               : launch_inter_t<MMO_GLOBAL, KK, false>(a, waves, s);           \
  }
  MMO_K(1) MMO_K(2) MMO_K(4) MMO_K(8)
#undef MMO_K
  return hipErrorInvalidValue;
}

template <class Src>
static hipError_t launch_cull(const InterLaunch& L, const InterArgs<Src>& a, hipStream_t s);

hipError_t launch_inter_coords(const InterLaunch& L, const Tile* tiles, const int* perm,
                               int ntiles, int npose, const double* xyz, const int* atom_off,
                               double* e_out, int e_stride, bool comp, double* grad, int write_e,
                               double* part, hipStream_t s) {
  InterArgs<CoordSrc> a;
  a.src.xyz = xyz; a.src.atom_off = atom_off; a.src.tiles = tiles; a.src.perm = perm;
  a.ntiles = ntiles; a.fr = L.fr; a.rpos = L.rpos; a.rab = L.rab; a.rres = L.rres; a.nrec = L.nrec;
  a.rtile = L.rtile; a.ligs = L.ligs; a.latoms = L.latoms; a.lig_xyz = L.lig_xyz;
  a.e_stride = e_stride; a.grad = grad; a.write_e = write_e; a.atom_order = nullptr;
  a.rslice = L.rslice > 0 ? L.rslice : L.nrec;
  a.npose_total = npose;
  if (L.cull && !grad) {
    a.e_out = e_out;
    a.e_comp = comp ? e_out : nullptr;
    return launch_cull(L, a, s);
  }
  const int ny = (L.nrec + a.rslice - 1) / a.rslice;
  a.e_out = ny > 1 ? part : e_out;
  a.e_comp = comp ? a.e_out : nullptr;
  hipError_t err = launch_inter_any(L.mode, L.k, grad != nullptr, a, L.waves, s);
  if (err != hipSuccess || ny == 1 || (grad && !write_e)) return err;
  const int blocks = std::min(4096, (npose + 255) / 256);
  hipLaunchKernelGGL(k_reduce_slices, dim3(blocks), dim3(256), 0, s, part, ny, (size_t)npose,
                     e_stride, comp ? 3 : 1, e_out);
  return hipGetLastError();
}

template <class Src>
static hipError_t launch_rows(const InterLaunch& L, const InterArgs<Src>& a, hipStream_t s) {
  if constexpr (std::is_same<Src, CoordSrc>::value) {
    return hipErrorInvalidValue;   // row lists serve the rigid scan only
  } else {
    const int waves = L.cull_waves;
    const int blocks = (a.ntiles + waves - 1) / waves;
    if (blocks == 0) return hipSuccess;
    const size_t lds =
        L.cull == 5 ? (size_t)(a.nrec + kRowChunk) * (sizeof(RecPos) + sizeof(RecAB)) : 0;
#define MMO_ROWS(RM, CH, KK)                                                                   \
  {                                                                                            \
    if (lds > 0) {                                                                             \
      hipError_t e = hipFuncSetAttribute(                                                      \
          reinterpret_cast<const void*>(&k_scan_rows<RM, CH, KK, Src>),                        \
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                               \
      if (e != hipSuccess) return e;                                                           \
    }                                                                                          \
    hipLaunchKernelGGL((k_scan_rows<RM, CH, KK, Src>), dim3(blocks), dim3(waves * kWave), lds, \
                       s, a, L.cg);                                                            \
  }
    // one ligand atom per lane: two (bonded neighbours of the sweep order) were measured
    // 15-20% slower on the 36.9M-pose scan, whichever receptor path
    const int rm = (L.cull == 5 ? 1 : 0) | (L.row_filter ? 2 : 0);
    if (rm == 0) {
      if (L.row_chunk == 4) MMO_ROWS(0, 4, 1) else MMO_ROWS(0, 8, 1)
    } else if (rm == 1) {
      if (L.row_chunk == 4) MMO_ROWS(1, 4, 1) else MMO_ROWS(1, 8, 1)
    } else if (rm == 2) {
      if (L.row_chunk == 4) MMO_ROWS(2, 4, 1) else MMO_ROWS(2, 8, 1)
    } else {
      if (L.row_chunk == 4) MMO_ROWS(3, 4, 1) else MMO_ROWS(3, 8, 1)
    }
#undef MMO_ROWS
    return hipGetLastError();
  }
}

template <class Src>
static hipError_t launch_cull(const InterLaunch& L, const InterArgs<Src>& a, hipStream_t s) {
  if (L.cull >= 4) return launch_rows(L, a, s);
  const int waves = L.cull_waves;
  const int blocks = (a.ntiles + waves - 1) / waves;
  if (blocks == 0) return hipSuccess;
  // L.cull: 1 receptor in LDS, 2 read in place, 3 positions in LDS + LJ factors in place
  const int rmode = L.cull == 1 ? 1 : (L.cull == 3 ? 2 : 0);
  const size_t lds = rmode == 1 ? receptor_lds_bytes<true>(a.nrec, L.cg)
                   : rmode == 2 ? receptor_pos_lds_bytes(a.nrec, L.cg) : 0;
  // dynamic LDS above 64 KiB must be allowed per kernel (gfx950 has 160 KiB per CU)
#define MMO_CULL_MODE(KK, RM)                                                                  \
  {                                                                                            \
    if (lds > 0) {                                                                             \
      hipError_t e = hipFuncSetAttribute(                                                      \
          reinterpret_cast<const void*>(&k_inter_cull<KK, RM, Src>),                           \
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);                               \
      if (e != hipSuccess) return e;                                                           \
    }                                                                                          \
    hipLaunchKernelGGL((k_inter_cull<KK, RM, Src>), dim3(blocks), dim3(waves * kWave), lds, s, \
                       a, L.cg);                                                               \
  }
#define MMO_CULL(KK)                                                                           \
  case KK: {                                                                                   \
    if (rmode == 1) MMO_CULL_MODE(KK, 1)                                                       \
    else if (rmode == 2) MMO_CULL_MODE(KK, 2)                                                  \
    else MMO_CULL_MODE(KK, 0)                                                                  \
    break;                                                                                     \
  }
  switch (L.cull_k) {
    MMO_CULL(1)
    MMO_CULL(2)
    MMO_CULL(4)
    default: return hipErrorInvalidValue;
  }
#undef MMO_CULL
#undef MMO_CULL_MODE
  return hipGetLastError();
}

hipError_t launch_inter_scan(const InterLaunch& L, int lig, const double* rots, int nrot,
                             const double* trans, const int* tperm, int ntrans, double* e_out,
                             hipStream_t s) {
  InterArgs<ScanSrc> a;
  a.src.rots = rots; a.src.trans = trans; a.src.tperm = tperm; a.src.ntrans = ntrans;
  a.src.tiles_per_rot = (ntrans + kWave - 1) / kWave; a.src.lig = lig;
  a.ntiles = nrot * a.src.tiles_per_rot; a.fr = L.fr; a.rpos = L.rpos; a.rab = L.rab; a.rres = L.rres;
  a.nrec = L.nrec; a.rtile = L.rtile; a.ligs = L.ligs; a.latoms = L.latoms;
  a.lig_xyz = L.lig_xyz; a.e_out = e_out; a.e_stride = 1; a.e_comp = nullptr; a.grad = nullptr;
  a.write_e = 1; a.rslice = L.nrec; a.npose_total = nrot * ntrans; a.atom_order = L.atom_order;
  if (L.cull) return launch_cull(L, a, s);
  return launch_inter_any(L.mode, L.k, false, a, L.waves, s);
}

hipError_t launch_inter_scan_list(const InterLaunch& L, int lig, const double* rots,
                                  const double* trans, const int* tperm, int ntrans,
                                  const int* list, int nlist, double* e_out, hipStream_t s) {
  if (!L.cull) return hipErrorInvalidValue;   // the list form has only the culled kernel
  InterArgs<ScanListSrc> a;
  a.src.rots = rots; a.src.trans = trans; a.src.tperm = tperm; a.src.ntrans = ntrans;
  a.src.tiles_per_rot = 0; a.src.lig = lig; a.src.list = list; a.src.nlist = nlist;
  a.ntiles = (nlist + kWave - 1) / kWave; a.fr = L.fr; a.rpos = L.rpos; a.rab = L.rab;
  a.rres = L.rres; a.nrec = L.nrec; a.rtile = L.rtile; a.ligs = L.ligs; a.latoms = L.latoms;
  a.lig_xyz = L.lig_xyz; a.e_out = e_out; a.e_stride = 1; a.e_comp = nullptr; a.grad = nullptr;
  a.write_e = 1; a.rslice = L.nrec; a.npose_total = nlist; a.atom_order = L.atom_order;
  return launch_cull(L, a, s);
}

hipError_t launch_intra(const IntraArgs& a, bool grad, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  // coordinates, then one gradient copy per wavefront
  const size_t lds = (size_t)a.max_atoms * (grad ? 3 + 3 * kIntraWaves : 3) * sizeof(double);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(grad ? reinterpret_cast<const void*>(&k_intra<true>)
                                            : reinterpret_cast<const void*>(&k_intra<false>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  const dim3 block(kIntraWaves * kWave);
  if (grad) hipLaunchKernelGGL(k_intra<true>, dim3(a.npose), block, lds, s, a);
  else hipLaunchKernelGGL(k_intra<false>, dim3(a.npose), block, lds, s, a);
  return hipGetLastError();
}

hipError_t launch_interp_coords(const InterpLaunch& L, const Tile* tiles, const int* perm,
                                int ntiles, const double* xyz, const int* atom_off,
                                const int* type_override, double* e_out, int e_stride,
                                double* grad, int grad_accumulate, hipStream_t s) {
  InterpArgs<CoordSrc> a;
  a.src.xyz = xyz; a.src.atom_off = atom_off; a.src.tiles = tiles; a.src.perm = perm;
  a.ntiles = ntiles; a.g = L.g; a.ligs = L.ligs; a.latoms = L.latoms; a.lig_xyz = L.lig_xyz;
  a.type_override = type_override; a.e_out = e_out; a.e_stride = e_stride; a.grad = grad;
  a.grad_accumulate = grad_accumulate;
  const int waves = 4;
  const int blocks = (ntiles + waves - 1) / waves;
  if (blocks == 0) return hipSuccess;
  if (grad) hipLaunchKernelGGL((k_interp<true, CoordSrc, true>), dim3(blocks), dim3(waves * kWave), 0, s, a);
  else hipLaunchKernelGGL((k_interp<false, CoordSrc, true>), dim3(blocks), dim3(waves * kWave), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_interp_scan(const InterpLaunch& L, int lig, const double* rots, int nrot,
                              const double* trans, const int* tperm, int ntrans, double* e_out,
                              hipStream_t s) {
  InterpArgs<ScanSrc> a;
  a.src.rots = rots; a.src.trans = trans; a.src.tperm = tperm; a.src.ntrans = ntrans;
  a.src.tiles_per_rot = (ntrans + kWave - 1) / kWave; a.src.lig = lig;
  a.ntiles = nrot * a.src.tiles_per_rot; a.g = L.g; a.ligs = L.ligs; a.latoms = L.latoms;
  a.lig_xyz = L.lig_xyz; a.type_override = nullptr; a.e_out = e_out; a.e_stride = 1;
  a.grad = nullptr; a.grad_accumulate = 0;
  const int waves = 4;
  const int blocks = (a.ntiles + waves - 1) / waves;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_interp<false, ScanSrc, false>), dim3(blocks), dim3(waves * kWave), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_interp_scan_list(const InterpLaunch& L, int lig, const double* rots,
                                   const double* trans, const int* tperm, int ntrans,
                                   const int* list, int nlist, double* e_out, hipStream_t s) {
  InterpArgs<ScanListSrc> a;
  a.src.rots = rots; a.src.trans = trans; a.src.tperm = tperm; a.src.ntrans = ntrans;
  a.src.tiles_per_rot = 0; a.src.lig = lig; a.src.list = list; a.src.nlist = nlist;
  a.ntiles = (nlist + kWave - 1) / kWave; a.g = L.g; a.ligs = L.ligs; a.latoms = L.latoms;
  a.lig_xyz = L.lig_xyz; a.type_override = nullptr; a.e_out = e_out; a.e_stride = 1;
  a.grad = nullptr; a.grad_accumulate = 0;
  const int waves = 4;
  const int blocks = (a.ntiles + waves - 1) / waves;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_interp<false, ScanListSrc, false>), dim3(blocks), dim3(waves * kWave), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_grid(const GridArgs& a, const CellGrid& cg, bool in_lds, hipStream_t s) {
  const int bricks = ((a.nx + 3) / 4) * ((a.ny + 3) / 4) * ((a.nz + 3) / 4);
  const int waves = 16;
  const int blocks = (bricks + waves - 1) / waves;
  if (in_lds) {
    const size_t lds = (size_t)a.nrec * (sizeof(RecPos) + sizeof(RecAB)) +
                       (size_t)(cg.nx * cg.ny * cg.nz + 1) * sizeof(int);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_grid<true>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_grid<true>, dim3(blocks), dim3(waves * kWave), lds, s, a, cg);
  } else {
    hipLaunchKernelGGL(k_grid<false>, dim3(blocks), dim3(waves * kWave), 0, s, a, cg);
  }
  return hipGetLastError();
}

hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_fill_int(int* p, size_t n, int v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill_int, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace mmo
