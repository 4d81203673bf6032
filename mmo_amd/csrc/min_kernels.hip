// min_kernels.hip -- batched L-BFGS over DOF vectors, one wavefront per pose.
//
// The reference minimises E_intra + E_inter over the DOF vector
// [x, y, z, alpha, beta, gamma, tau_0..] of Optim.apply_config with NLopt inside box
// bounds (ROI cube for the position, [-pi, pi] for angles), with gradients from
// central differences (Lds.minimize_ligand / nlopt_minimize / approx_local_gradient,
// src/lds.ml:1014-1036, :1118-1297, src/optim.ml:34-80).  Here every pose of a batch
// runs limited-memory BFGS on the analytic adjoint gradient:
//
//   iteration = one energy + gradient evaluation of every pose's trial point (the
//   batched DOF kernels) followed by k_lbfgs, which per pose either accepts the trial
//   (Armijo sufficient decrease), updates the (s, y) history and computes the next
//   direction with the two-loop recursion, or halves the step (backtracking);
//   trial points are projected onto the box bounds.  Poses whose torsions make the
//   ligand longer than 12 A (Mol.Too_long, src/optim.ml:75) or whose energy is not
//   finite count as failed trials.
//
// Lanes hold the DOF components (at most 64 per pose); dot products are wavefront
// reductions in FP64.
#include "common.h"
#include "dofmin.h"
#include "kernels.h"

namespace mmo {

__global__ void __launch_bounds__(256) k_lbfgs(MinArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int p = blockIdx.x * (blockDim.x / kWave) + (threadIdx.x / kWave);
  if (p >= a.npose) return;
  MinPose s = a.st[p];
  if (s.status != kMinRunning) return;
  const int o = a.dof_off[p];
  const int n = a.dof_off[p + 1] - o;
  const bool act = lane < n;
  const size_t i = (size_t)o + (act ? lane : 0);
  const double ft = a.et[4 * (size_t)p] + a.et[4 * (size_t)p + 1];
  const double gti = act ? a.gt[i] : 0.0;
  const double xti = act ? a.xt[i] : 0.0;
  const double lo = (act && a.lo) ? a.lo[i] : -__builtin_inf();
  const double hi = (act && a.hi) ? a.hi[i] : __builtin_inf();
  double x = 0.0, g = 0.0, d = 0.0;
  if (s.init && act) {
    x = a.x[i];
    g = a.g[i];
    d = a.d[i];
  }
  const MinCfg cfg{a.m, a.max_step, a.min_step, a.gtol, a.c1};
  const double xn = lbfgs_wave(s, x, g, d, xti, ft, gti, a.rad[p], lo, hi, act, a.S + i,
                               a.Y + i, (size_t)a.ndof_total, a.rho + p, (size_t)a.npose, cfg);
  if (act) {
    a.x[i] = x;
    a.g[i] = g;
    a.d[i] = d;
    a.xt[i] = xn;
  }
  if (lane == 0) a.st[p] = s;
}

// best energies out (the accepted point is the best point: accepted steps decrease E)
__global__ void k_min_energy(const MinPose* st, int npose, double* e_out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npose) e_out[p] = st[p].f;
}

__global__ void k_min_reset(MinPose* st, int npose) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < npose) {
    MinPose z{};
    st[p] = z;
  }
}

// ---------------------------------------------------------- persistent minimiser
// LDS per wavefront: the conformer and the adjoint (6 x max_atoms FP64) and the
// atom-sweep scratch; before them, with REC_LDS, the receptor and its cell table.
__host__ __device__ inline size_t min_wave_lds_bytes(int max_atoms) {
  return ((size_t)6 * max_atoms * sizeof(double) + kAtomSweepInts * sizeof(int) + 15) & ~(size_t)15;
}

constexpr int kMinPersistWaves = 16;
// Wavefronts per workgroup each form is compiled for.  With the receptor in LDS at most
// 8-9 wavefronts fit in a workgroup's LDS anyway, so that form is compiled for 8 (up to
// 256 VGPRs, 2 wavefronts per SIMD): at 16 (128 VGPRs) it spilled 236 B per lane.
__host__ __device__ constexpr int min_persist_waves(bool rec_lds) {
  return rec_lds ? 8 : kMinPersistWaves;
}

template <bool REC_LDS, int INTER>
__global__ void __launch_bounds__(min_persist_waves(REC_LDS) * 64) k_min_persist(MinPersistArgs a) {
  extern __shared__ __align__(16) char smem[];
  const RecPos* s_pos;
  const RecAB* s_ab;
  const int* s_cell;
  load_receptor<REC_LDS>(a.rpos, a.rab, a.nrec, a.cg, s_pos, s_ab, s_cell, smem);
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  char* mine = smem + receptor_lds_bytes<REC_LDS>(a.nrec, a.cg) + wv * min_wave_lds_bytes(a.max_atoms);
  double* tx = reinterpret_cast<double*>(mine);
  double* ty = tx + a.max_atoms;
  double* tz = ty + a.max_atoms;
  double* hx = tz + a.max_atoms;
  double* hy = hx + a.max_atoms;
  double* hz = hy + a.max_atoms;
  int* wl = reinterpret_cast<int*>(hz + a.max_atoms);
  const MinCfg cfg{a.m, a.max_step, a.min_step, a.gtol, a.c1};
  for (;;) {
    // the next pose: lane 0 takes it from the queue, readlane makes it a scalar, so the
    // loop exit is uniform (with a vector broadcast the compiler treats the exit as
    // divergent and may send the other lanes round the loop without lane 0's atomic)
    int qv = 0;
    if (lane == 0) qv = atomicAdd(a.queue, 1);
    const int q = __builtin_amdgcn_readlane(qv, 0);
    if (q >= a.npose) break;
    const int p = a.order ? a.order[q] : q;
    const LigInfo li = a.ligs[a.pose_lig[p]];
    const int o = a.dof_off[p];
    const int n = a.dof_off[p + 1] - o;
    const bool act = lane < n;
    const size_t i = (size_t)o + (act ? lane : 0);
    DofEval ev;
    ev.li = li; ev.latoms = a.latoms; ev.atom_order = a.atom_order; ev.rb_left = a.rb_left;
    ev.rb_right = a.rb_right; ev.rg_off = a.rg_off; ev.rg_atoms = a.rg_atoms;
    ev.pair_ij = a.pair_ij; ev.pair_par = a.pair_par; ev.ntors = n - 6; ev.intra = a.intra != 0;
    ev.fr = a.fr; ev.s_pos = s_pos; ev.s_ab = s_ab; ev.s_cell = s_cell; ev.rres = a.rres;
    ev.cg = a.cg; ev.wl = wl; ev.g = a.g;
    const double* base = a.lig_xyz + 3 * (size_t)li.atom_base;
    const double lo = (act && a.lo) ? a.lo[i] : -__builtin_inf();
    const double hi = (act && a.hi) ? a.hi[i] : __builtin_inf();
    MinPose s{};
    double xt = act ? a.x[i] : 0.0, x = 0.0, g = 0.0, d = 0.0;
    for (int it = 0; it < a.iters && __builtin_amdgcn_readfirstlane(s.status) == kMinRunning; ++it) {
      double gt, ei, ea, rad;
      const double f = dof_eval<INTER>(ev, base, xt, tx, ty, tz, hx, hy, hz, gt, ei, ea, rad);
      xt = lbfgs_wave(s, x, g, d, xt, f, act ? gt : 0.0, rad, lo, hi, act, a.S + i, a.Y + i,
                      (size_t)a.ndof_total, a.rho + p, (size_t)a.npose, cfg);
    }
    if (a.iters > 0) {
      if (act) a.x[i] = x;
      if (lane == 0) {
        a.st[p] = s;
        if (a.e_out) a.e_out[p] = s.f;
      }
    }
    // all lanes meet here before the next pose (a convergent point: without it the
    // compiler joins this lane-0 block to the next queue fetch, and the other lanes
    // would start the next pose without lane 0)
    wave_lds_sync();
  }
}

size_t min_persist_lds_bytes(const MinPersistArgs& a, int waves, bool rec_lds) {
  return (rec_lds ? receptor_lds_bytes<true>(a.nrec, a.cg) : 0) +
         (size_t)waves * min_wave_lds_bytes(a.max_atoms);
}

template <bool R, int I>
static hipError_t min_persist_go(const MinPersistArgs& a, int nblocks, int waves, hipStream_t s) {
  const size_t lds = min_persist_lds_bytes(a, waves, R);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_min_persist<R, I>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_min_persist<R, I>), dim3(nblocks), dim3(waves * kWave), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_min_persist(const MinPersistArgs& a, int inter, int nblocks, int waves,
                              bool rec_lds, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  if (waves < 1 || waves > min_persist_waves(rec_lds)) return hipErrorInvalidValue;
  switch (inter + (rec_lds ? 8 : 0)) {
    case kDofInterAtom + 8: return min_persist_go<true, kDofInterAtom>(a, nblocks, waves, s);
    case kDofInterF64 + 8: return min_persist_go<true, kDofInterF64>(a, nblocks, waves, s);
    case kDofInterAtom: return min_persist_go<false, kDofInterAtom>(a, nblocks, waves, s);
    case kDofInterF64: return min_persist_go<false, kDofInterF64>(a, nblocks, waves, s);
    case kDofInterGrid: return min_persist_go<false, kDofInterGrid>(a, nblocks, waves, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_lbfgs(const MinArgs& a, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  constexpr int kWaves = 4;
  hipLaunchKernelGGL(k_lbfgs, dim3((a.npose + kWaves - 1) / kWaves), dim3(kWaves * kWave), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_min_reset(MinPose* st, int npose, hipStream_t s) {
  if (npose == 0) return hipSuccess;
  hipLaunchKernelGGL(k_min_reset, dim3((npose + 255) / 256), dim3(256), 0, s, st, npose);
  return hipGetLastError();
}

hipError_t launch_min_energy(const MinPose* st, int npose, double* e_out, hipStream_t s) {
  if (npose == 0) return hipSuccess;
  hipLaunchKernelGGL(k_min_energy, dim3((npose + 255) / 256), dim3(256), 0, s, st, npose, e_out);
  return hipGetLastError();
}

}  // namespace mmo
