// desolv_kernels.hip -- desolvation sums of Lds.protein_desolv / Lds.desolvation_penalty
// (src/lds.ml:201-267; the reference defines them but never calls them).
//
// Per listed grid node, the sum over atoms within 12 A of (q_j / d^2)^2 in FP64 (one
// thread per node, the atoms staged through LDS in tiles).  k_desolv_nodes writes
// scale x sum per node (the protein contributions); k_desolv_pose reduces, per
// workgroup, the nodes' sums and the protein contributions of the same nodes (one
// ligand pose's penalty), so the host adds a few partial sums in workgroup order.
#include "common.h"
#include "kernels.h"

namespace mmo {

constexpr int kDesolvThreads = kDesolvNodes;
constexpr int kDesolvTile = 1024;

__device__ __forceinline__ double desolv_node_sum(const DesolvArgs& a, int v, double4* s_at) {
  double x = 0.0, y = 0.0, z = 0.0;
  if (v >= 0) {
    const long k = v / ((long)a.nx * a.ny), j = (v - k * a.nx * a.ny) / a.nx,
               i = v - (k * a.nx * a.ny + j * a.nx);
    x = a.ox + i * a.step; y = a.oy + j * a.step; z = a.oz + k * a.step;
  }
  double s = 0.0;
  for (int t0 = 0; t0 < a.natoms; t0 += kDesolvTile) {
    const int nt = min(kDesolvTile, a.natoms - t0);
    __syncthreads();
    for (int q = threadIdx.x; q < nt; q += blockDim.x) s_at[q] = a.atoms[t0 + q];
    __syncthreads();
    if (v >= 0)
      for (int q = 0; q < nt; ++q) {
        const double4 p = s_at[q];
        const double dx = x - p.x, dy = y - p.y, dz = z - p.z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < 144.0) {   // BST.neighbours within Const.charged_cutoff
          const double w = p.w / d2;
          s += w * w;
        }
      }
  }
  return s;
}

__global__ void __launch_bounds__(kDesolvThreads) k_desolv_nodes(DesolvArgs a) {
  __shared__ double4 s_at[kDesolvTile];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = n < a.nvox ? a.vox[n] : -1;
  const double s = desolv_node_sum(a, v, s_at);
  if (v >= 0) a.out[v] = a.scale * s;
}

__global__ void __launch_bounds__(kDesolvThreads) k_desolv_pose(DesolvArgs a) {
  __shared__ double4 s_at[kDesolvTile];
  __shared__ double s_red[2][kDesolvThreads / kWave];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  const int v = n < a.nvox ? a.vox[n] : -1;
  double sl = desolv_node_sum(a, v, s_at);
  double sp = v >= 0 ? a.prot[v] : 0.0;
  sl = wave_sum(sl);
  sp = wave_sum(sp);
  const int w = threadIdx.x / kWave;
  if ((threadIdx.x & (kWave - 1)) == 0) { s_red[0][w] = sl; s_red[1][w] = sp; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tl = 0.0, tp = 0.0;
    for (int q = 0; q < kDesolvThreads / kWave; ++q) { tl += s_red[0][q]; tp += s_red[1][q]; }
    a.part[2 * blockIdx.x] = tl;
    a.part[2 * blockIdx.x + 1] = tp;
  }
}

hipError_t launch_desolv(const DesolvArgs& a, bool pose, hipStream_t s) {
  if (a.nvox == 0) return hipSuccess;
  const int blocks = (a.nvox + kDesolvThreads - 1) / kDesolvThreads;
  if (pose) hipLaunchKernelGGL(k_desolv_pose, dim3(blocks), dim3(kDesolvThreads), 0, s, a);
  else hipLaunchKernelGGL(k_desolv_nodes, dim3(blocks), dim3(kDesolvThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mmo
