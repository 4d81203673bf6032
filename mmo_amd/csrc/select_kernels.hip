// select_kernels.hip -- rigid-scan clash filters and top-k selection of scan energies.
//
// Clash filters (Lds.exhaustive_rigid_ligand_docking, src/lds.ml:1040-1114): a pose is
// skipped when its translation lies outside the ROI sphere or, for ligands whose
// centre is inside one of their atoms, on a voxel whose 8 corners are all protein vdW
// volume (G3D.vdW_clash_AND, src/G3D.ml:188-213), or when any atom has a protein-volume
// voxel among its 8 corners (Mol.protein_ligand_clash + G3D.vdW_clash_OR,
// src/mol.ml:1195-1206, src/G3D.ml:159-185).  Skipped poses get +inf.
//
// Top-k (the TopKeeper of src/lds.ml:1055-1064, :1110-1113, lowest energies first):
// an MSB-first radix select over order-preserving 64-bit keys of the energies finds the
// k-th smallest key T (8 histogram passes of one byte each); keys < T are gathered,
// keys == T are taken in index order (so ties resolve to the lowest pose indices), and
// one workgroup sorts the k (energy, index) pairs.  NaN sorts as the largest value.
#include <cstdint>

#include <hipcub/hipcub.hpp>

#include "common.h"
#include "kernels.h"

namespace mmo {

// ------------------------------------------------------------------ clash filter
// whether an atom of pose (rotation r, translation t) has a protein-volume voxel among its
// 8 corners (Mol.protein_ligand_clash); s_lig: the centred conformer rows
__device__ bool pose_atom_clash(const ScanFilterArgs& a, const double* s_lig, int r, int t) {
  const double* R = a.rots + 9 * (size_t)r;
  const double* T = a.trans + 3 * (size_t)t;
  const size_t nxy = (size_t)a.nx * a.ny;
  for (int j = 0; j < a.n_atoms; ++j) {
    const double* c = s_lig + 3 * j;
    // the pose's atom position as the scan kernel computes it
    const double x = R[0] * c[0] + R[1] * c[1] + R[2] * c[2] + T[0];
    const double y = R[3] * c[0] + R[4] * c[1] + R[5] * c[2] + T[1];
    const double z = R[6] * c[0] + R[7] * c[1] + R[8] * c[2] + T[2];
    const double fx = (x - a.ox) * a.inv_step, fy = (y - a.oy) * a.inv_step,
                 fz = (z - a.oz) * a.inv_step;
    // corners outside the grid count as free (bounds checked in floating point first)
    if (!(fx > -1.0 && fy > -1.0 && fz > -1.0 && fx < (double)(a.nx - 1) &&
          fy < (double)(a.ny - 1) && fz < (double)(a.nz - 1)))
      continue;
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;   // truncation, as int_of_float
    bool hit = false;
#pragma unroll
    for (int cidx = 0; cidx < 8; ++cidx) {
      const size_t v = (size_t)(i0 + (cidx & 1)) + (size_t)(j0 + ((cidx >> 1) & 1)) * a.nx +
                       (size_t)(k0 + (cidx >> 2)) * nxy;
      hit |= ((a.mask[v >> 5] >> (v & 31)) & 1u) != 0u;
    }
    if (hit) return true;
  }
  return false;
}

__global__ void __launch_bounds__(256) k_scan_filter(ScanFilterArgs a) {
  extern __shared__ double s_lig[];
  for (int i = threadIdx.x; i < 3 * a.n_atoms; i += blockDim.x) s_lig[i] = a.lig[i];
  __syncthreads();
  const size_t pose = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (pose >= (size_t)a.nrot * a.ntrans) return;
  const int r = (int)(pose / a.ntrans), t = (int)(pose - (size_t)r * a.ntrans);
  if (a.skip[t] || (a.atom_clash && pose_atom_clash(a, s_lig, r, t))) a.e[pose] = __builtin_inf();
}

// filter before scoring: the same tests, as keep flags over the scan's pose order
__global__ void __launch_bounds__(256) k_scan_keep(ScanFilterArgs a, const int* tperm, uint8_t* keep) {
  extern __shared__ double s_lig[];
  for (int i = threadIdx.x; i < 3 * a.n_atoms; i += blockDim.x) s_lig[i] = a.lig[i];
  __syncthreads();
  const size_t slot = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= (size_t)a.nrot * a.ntrans) return;
  const int r = (int)(slot / a.ntrans);
  const int t = tperm[slot - (size_t)r * a.ntrans];
  keep[slot] = (a.skip[t] || (a.atom_clash && pose_atom_clash(a, s_lig, r, t))) ? 0 : 1;
}

hipError_t launch_scan_keep(const ScanFilterArgs& a, const int* tperm, uint8_t* keep, hipStream_t s) {
  const size_t n = (size_t)a.nrot * a.ntrans;
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_scan_keep, dim3((unsigned)blocks), dim3(256),
                     (size_t)3 * a.n_atoms * sizeof(double), s, a, tperm, keep);
  return hipGetLastError();
}

hipError_t select_flagged_slots(const uint8_t* flags, int n, int* out, int* n_out, void* temp,
                                size_t* temp_bytes, hipStream_t s) {
  return hipcub::DeviceSelect::Flagged(temp, *temp_bytes, hipcub::CountingInputIterator<int>(0),
                                       flags, out, n_out, n, s);
}

hipError_t launch_scan_filter(const ScanFilterArgs& a, hipStream_t s) {
  const size_t n = (size_t)a.nrot * a.ntrans;
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_scan_filter, dim3((unsigned)blocks), dim3(256),
                     (size_t)3 * a.n_atoms * sizeof(double), s, a);
  return hipGetLastError();
}

// -------------------------------------------------------------------- top-k
__device__ __forceinline__ uint64_t order_key(double v) {
  if (v != v) return ~0ull;
  const uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// st[0] = prefix (selected high bytes), st[1] = mask of selected bytes, st[2] = how many
// keys equal to the final threshold still have to be taken
__global__ void __launch_bounds__(256) k_topk_hist(const double* e, size_t n, const uint64_t* st,
                                                   int shift, unsigned* hist) {
  __shared__ unsigned h[256];
  h[threadIdx.x] = 0u;
  __syncthreads();
  const uint64_t prefix = st[0], mask = st[1];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint64_t k = order_key(e[i]);
    if ((k & mask) == prefix) atomicAdd(&h[(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

__global__ void k_topk_digit(unsigned* hist, uint64_t* st, int shift) {
  if (threadIdx.x != 0) return;
  const uint64_t rem = st[2];
  uint64_t cum = 0;
  int b = 255;
  for (int i = 0; i < 256; ++i) {
    if (cum + hist[i] >= rem) {
      b = i;
      break;
    }
    cum += hist[i];
  }
  st[0] |= (uint64_t)b << shift;
  st[1] |= 0xFFull << shift;
  st[2] = rem - cum;
  for (int i = 0; i < 256; ++i) hist[i] = 0u;
}

__global__ void k_topk_init(uint64_t* st, unsigned* hist, unsigned* cnt, uint64_t k) {
  if (threadIdx.x == 0) {
    st[0] = 0ull;
    st[1] = 0ull;
    st[2] = k;
    cnt[0] = 0u;
  }
  hist[threadIdx.x] = 0u;
}

// keys below the threshold: k - st[2] of them, in any order
__global__ void __launch_bounds__(256) k_topk_less(const double* e, size_t n, const uint64_t* st,
                                                   unsigned* cnt, double* out_e, int64_t* out_i) {
  const uint64_t T = st[0];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = e[i];
    if (order_key(v) < T) {
      const unsigned p = atomicAdd(cnt, 1u);
      out_e[p] = v;
      out_i[p] = (int64_t)i;
    }
  }
}

// keys equal to the threshold, counted per chunk of `chunk` consecutive indices
__global__ void __launch_bounds__(64) k_topk_eq_count(const double* e, size_t n, size_t chunk,
                                                      const uint64_t* st, unsigned* bcnt) {
  const uint64_t T = st[0];
  const size_t b0 = (size_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  unsigned c = 0;
  for (size_t i = b0 + threadIdx.x; i < b1; i += 64) c += order_key(e[i]) == T ? 1u : 0u;
  for (int off = 1; off < 64; off <<= 1) c += __shfl_xor(c, off, 64);
  if (threadIdx.x == 0) bcnt[blockIdx.x] = c;
}

// exclusive prefix sum of the per-chunk counts: each thread scans a contiguous run of
// chunks, the run totals are scanned across the workgroup
__global__ void __launch_bounds__(1024) k_topk_eq_scan(unsigned* bcnt, int nb) {
  __shared__ unsigned part[1024];
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  unsigned sum = 0;
  for (int b = b0; b < b1; ++b) sum += bcnt[b];
  part[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned v = t >= off ? part[t - off] : 0u;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  unsigned run = part[t] - sum;
  for (int b = b0; b < b1; ++b) {
    const unsigned c = bcnt[b];
    bcnt[b] = run;
    run += c;
  }
}

// the first st[2] equal keys in index order go after the k - st[2] smaller ones
__global__ void __launch_bounds__(64) k_topk_eq_write(const double* e, size_t n, size_t chunk,
                                                      const uint64_t* st, const unsigned* bcnt,
                                                      uint64_t k, double* out_e, int64_t* out_i) {
  const uint64_t T = st[0], take = st[2];
  const size_t b0 = (size_t)blockIdx.x * chunk, b1 = min(n, b0 + chunk);
  uint64_t base = bcnt[blockIdx.x];
  if (base >= take) return;
  const int lane = threadIdx.x;
  for (size_t s0 = b0; s0 < b1 && base < take; s0 += 64) {
    const size_t i = s0 + lane;
    const bool f = i < b1 && order_key(e[i]) == T;
    const uint64_t m = __ballot(f);
    const uint64_t below = lane ? (m & ((1ull << lane) - 1ull)) : 0ull;
    const uint64_t rank = base + (uint64_t)__popcll(below);
    if (f && rank < take) {
      const uint64_t p = (k - take) + rank;
      out_e[p] = e[i];
      out_i[p] = (int64_t)i;
    }
    base += (uint64_t)__popcll(m);
  }
}

// one workgroup: bitonic sort of the k pairs by (energy key, index), k <= kTopkMax
__global__ void __launch_bounds__(1024) k_topk_sort(double* e, int64_t* idx, int k) {
  __shared__ uint64_t sk[kTopkMax];
  __shared__ int64_t si[kTopkMax];
  int n2 = 1;
  while (n2 < k) n2 <<= 1;
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    sk[i] = i < k ? order_key(e[i]) : ~0ull;
    si[i] = i < k ? idx[i] : INT64_MAX;
  }
  __syncthreads();
  for (int size = 2; size <= n2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < n2; i += blockDim.x) {
        const int j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const bool gt = sk[i] > sk[j] || (sk[i] == sk[j] && si[i] > si[j]);
          if (gt == up) {
            const uint64_t tk = sk[i]; sk[i] = sk[j]; sk[j] = tk;
            const int64_t ti = si[i]; si[i] = si[j]; si[j] = ti;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const uint64_t kk = sk[i];
    const uint64_t b = (kk >> 63) ? (kk & 0x7FFFFFFFFFFFFFFFull) : ~kk;
    e[i] = kk == ~0ull ? __builtin_nan("") : __longlong_as_double((long long)b);
    idx[i] = si[i];
  }
}

hipError_t launch_topk(const double* e, size_t n, int k, double* out_e, int64_t* out_i,
                       TopkScratch w, hipStream_t s) {
  if (k <= 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_topk_init, dim3(1), dim3(256), 0, s, w.st, w.hist, w.cnt, (uint64_t)k);
  const int blocks = (int)std::min<size_t>(4096, (n + 255) / 256);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(k_topk_hist, dim3(blocks), dim3(256), 0, s, e, n, w.st, shift, w.hist);
    hipLaunchKernelGGL(k_topk_digit, dim3(1), dim3(64), 0, s, w.hist, w.st, shift);
  }
  hipLaunchKernelGGL(k_topk_less, dim3(blocks), dim3(256), 0, s, e, n, w.st, w.cnt, out_e, out_i);
  const int nb = (int)((n + w.chunk - 1) / w.chunk);
  hipLaunchKernelGGL(k_topk_eq_count, dim3(nb), dim3(64), 0, s, e, n, w.chunk, w.st, w.bcnt);
  hipLaunchKernelGGL(k_topk_eq_scan, dim3(1), dim3(1024), 0, s, w.bcnt, nb);
  hipLaunchKernelGGL(k_topk_eq_write, dim3(nb), dim3(64), 0, s, e, n, w.chunk, w.st, w.bcnt,
                     (uint64_t)k, out_e, out_i);
  hipLaunchKernelGGL(k_topk_sort, dim3(1), dim3(1024), 0, s, out_e, out_i, k);
  return hipGetLastError();
}

}  // namespace mmo
