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
#include "kernels.h"

namespace mmo {

namespace {

__device__ __forceinline__ double dot(double a, double b) { return wave_sum(a * b); }

__device__ __forceinline__ double wave_absmax(double v) {
  v = fabs(v);
  return wave_max(v);
}

}  // namespace

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
  const size_t nt = (size_t)a.ndof_total;

  double ft = a.et[4 * (size_t)p] + a.et[4 * (size_t)p + 1];
  if (!(a.rad[p] <= 12.0) || !isfinite(ft)) ft = __builtin_inf();
  const double gti = act ? a.gt[i] : 0.0;
  const double xti = act ? a.xt[i] : 0.0;
  const double lo = (act && a.lo) ? a.lo[i] : -__builtin_inf();
  const double hi = (act && a.hi) ? a.hi[i] : __builtin_inf();
  s.n_eval += 1;

  double x, g, d;
  if (!s.init) {
    // first evaluation: the start point itself
    x = xti;
    g = gti;
    s.f = ft;
    s.f0 = ft;
    s.init = 1;
    s.hist = 0;
    s.head = 0;
    if (!isfinite(ft)) s.status = kMinFailed;
    d = -g;
    const double dm = wave_absmax(act ? d : 0.0);
    s.alpha = dm > 0.0 ? fmin(1.0, a.max_step / dm) : 0.0;
    if (dm <= a.gtol) s.status = kMinConverged;
  } else {
    x = act ? a.x[i] : 0.0;
    g = act ? a.g[i] : 0.0;
    d = act ? a.d[i] : 0.0;
    const double sv = xti - x;                 // the step actually taken (after projection)
    const double gs = dot(act ? g : 0.0, act ? sv : 0.0);
    if (gs < 0.0 && ft <= s.f + a.c1 * gs) {
      // accept; curvature pair
      const double yv = gti - g;
      const double sy = dot(act ? sv : 0.0, act ? yv : 0.0);
      const double yy = dot(act ? yv : 0.0, act ? yv : 0.0);
      if (sy > 1e-12 * yy && sy > 0.0) {
        const int slot = s.head;
        if (act) {
          a.S[(size_t)slot * nt + i] = sv;
          a.Y[(size_t)slot * nt + i] = yv;
        }
        if (lane == 0) a.rho[(size_t)slot * a.npose + p] = 1.0 / sy;
        s.head = (s.head + 1) % a.m;
        s.hist = min(s.hist + 1, a.m);
        s.gamma = sy / yy;
      }
      x = xti;
      g = gti;
      s.f = ft;
      s.n_acc += 1;
      // two-loop recursion, newest pair first
      double q = act ? g : 0.0;
      double al[kMinMaxHist];
      for (int k = 0; k < s.hist; ++k) {
        const int idx = (s.head - 1 - k + 2 * a.m) % a.m;
        const double sk = act ? a.S[(size_t)idx * nt + i] : 0.0;
        const double yk = act ? a.Y[(size_t)idx * nt + i] : 0.0;
        const double rk = a.rho[(size_t)idx * a.npose + p];
        al[k] = rk * dot(sk, q);
        q -= al[k] * yk;
      }
      double r = (s.hist > 0 ? s.gamma : 1.0) * q;
      for (int k = s.hist - 1; k >= 0; --k) {
        const int idx = (s.head - 1 - k + 2 * a.m) % a.m;
        const double sk = act ? a.S[(size_t)idx * nt + i] : 0.0;
        const double yk = act ? a.Y[(size_t)idx * nt + i] : 0.0;
        const double rk = a.rho[(size_t)idx * a.npose + p];
        const double b = rk * dot(yk, r);
        r += sk * (al[k] - b);
      }
      d = act ? -r : 0.0;
      if (dot(d, act ? g : 0.0) >= 0.0) {   // not a descent direction: restart
        s.hist = 0;
        s.n_reset += 1;
        d = act ? -g : 0.0;
      }
      const double dm = wave_absmax(d);
      const double gm = wave_absmax(act ? g : 0.0);
      s.alpha = dm > 0.0 ? fmin(1.0, a.max_step / dm) : 0.0;
      if (gm <= a.gtol) s.status = kMinConverged;
    } else {
      // backtrack
      s.alpha *= 0.5;
      const double dm = wave_absmax(act ? d : 0.0);
      if (s.alpha * dm < a.min_step) {
        if (s.hist == 0) {
          s.status = kMinConverged;        // no decrease along steepest descent
        } else {                           // drop the history and restart
          s.hist = 0;
          s.n_reset += 1;
          d = act ? -g : 0.0;
          const double dm2 = wave_absmax(d);
          s.alpha = dm2 > 0.0 ? fmin(1.0, a.max_step / dm2) : 0.0;
        }
      }
    }
  }
  const double xn = (s.status == kMinRunning) ? fmin(hi, fmax(lo, x + s.alpha * d)) : x;
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
