// pose_kernels.hip -- DOFs -> coordinates (Optim.apply_config, src/optim.ml:64-80) and
// its adjoint (dE/dxyz -> dE/dDOF), one workgroup per pose (one wavefront; kDofWaves
// when the adjoint also evaluates E_intra), FP64 throughout.
//
// Forward: the registered centred conformer c0; for k = 0..R-1 the atoms of rotation
// group k turn by tau_k about the axis right_k - left_k through right_k
// (Mol.rotate_bond, src/mol.ml:610-631), in rbond order, each about the current
// positions; then x = r_xyz(alpha, beta, gamma) c + t (Rot.r_xyz src/rot.ml:52-66,
// Mol.rotate_then_translate_copy src/mol.ml:669-672).
//
// Adjoint: reverse pass over k.  Neither `left` nor `right` of rbond k moves under
// rotation k (the group excludes both ends, src/pqrs.ml:80-87), so c^k is recovered
// from c^{k+1} by the inverse rotation and no intermediate conformers are stored.
// For y_i = p + R(a, tau)(x_i - p), with upstream adjoint h_i:
//   dE/dtau  = sum_i h_i . (a x (y_i - p))
//   x_i_bar  = R^T h_i
//   p_bar    = sum_i (h_i - R^T h_i)                       (pivot = right)
//   a_bar    = sum_i (z_i x h_i) sin + h_i (a.z_i)(1-cos) + z_i (a.h_i)(1-cos),  z_i = x_i - p
//   v_bar    = (I - a a^T) a_bar / |v|,  v = x_right - x_left
//   right_bar += v_bar + p_bar,  left_bar -= v_bar.
#include "common.h"
#include "dof.h"
#include "kernels.h"

namespace mmo {

// wavefronts per pose of the adjoint when it also evaluates E_intra
constexpr int kDofWaves = 2;

// LDS of the DOF kernels: `arrays` FP64 coordinate arrays of max_atoms entries, then the
// staged torsion tree (TorsionLds)
__host__ __device__ constexpr size_t dof_coord_bytes(int max_atoms, int arrays) {
  return (size_t)arrays * max_atoms * sizeof(double);
}
constexpr size_t torsion_lds_bytes(int max_rb, int max_rg) {
  return 2 * (size_t)(max_rb + 3) * sizeof(double) + (size_t)(3 * max_rb + 1 + max_rg) * sizeof(int);
}

// The pose's torsion tree staged in LDS before the serial torsion loop, so that the
// loop reads nothing from global memory: rbond ends, rotation groups (offsets relative
// to the ligand's first group atom) and the sine and cosine of every torsion angle
// (entries 0..nrb-1) and of the three rigid-body angles (entries nrb..nrb+2), evaluated
// by all lanes at once.
struct TorsionLds {
  double* sn;
  double* cs;
  int* left;
  int* right;
  int* off;
  int* atoms;
};

__device__ TorsionLds stage_torsions(const DofArgs& a, const LigInfo& li, const double* d,
                                     char* base) {
  TorsionLds t;
  t.sn = reinterpret_cast<double*>(base);
  t.cs = t.sn + a.max_rb + 3;
  t.left = reinterpret_cast<int*>(t.cs + a.max_rb + 3);
  t.right = t.left + a.max_rb;
  t.off = t.right + a.max_rb;
  t.atoms = t.off + a.max_rb + 1;
  const int nrb = li.nrb;
  for (int k = threadIdx.x; k < nrb + 3; k += blockDim.x) {
    double s, c;
    sincos(k < nrb ? d[6 + k] : d[3 + (k - nrb)], &s, &c);
    t.sn[k] = s;
    t.cs[k] = c;
    if (k < nrb) {
      t.left[k] = a.rb_left[li.rb_base + k];
      t.right[k] = a.rb_right[li.rb_base + k];
    }
  }
  const int g0 = a.rg_off[li.rg_base];
  for (int k = threadIdx.x; k <= nrb; k += blockDim.x) t.off[k] = a.rg_off[li.rg_base + k] - g0;
  const int ng = a.rg_off[li.rg_base + nrb] - g0;
  for (int m = threadIdx.x; m < ng; m += blockDim.x) t.atoms[m] = a.rg_atoms[g0 + m];
  return t;
}

// torsions in rbond order on the conformer held in LDS (all threads participate)
__device__ void forward_torsions(const LigInfo& li, const TorsionLds& t, double* sx, double* sy,
                                 double* sz) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = 0; k < li.nrb; ++k) {
    __syncthreads();
    const int left = t.left[k], right = t.right[k];
    const double px = sx[right], py = sy[right], pz = sz[right];
    const double vx = px - sx[left], vy = py - sy[left], vz = pz - sz[left];
    const double im = rsqrt(vx * vx + vy * vy + vz * vz);   // V3.normalize: v / |v|
    const Rod R = rodrigues(vx * im, vy * im, vz * im, t.sn[k], t.cs[k]);
    const int g0 = t.off[k], g1 = t.off[k + 1];
    __syncthreads();
    for (int m = g0 + tid; m < g1; m += nt) {
      const int i = t.atoms[m];
      const double zx = sx[i] - px, zy = sy[i] - py, zz = sz[i] - pz;
      sx[i] = (R.m[0] * zx + R.m[1] * zy + R.m[2] * zz) + px;
      sy[i] = (R.m[3] * zx + R.m[4] * zy + R.m[5] * zz) + py;
      sz[i] = (R.m[6] * zx + R.m[7] * zy + R.m[8] * zz) + pz;
    }
  }
  __syncthreads();
}

// src/mol.ml:881-903 over the ligand's pair list (topo distance >= 3) for the conformer
// in LDS, FP64 as in k_intra, every thread of the workgroup taking pairs p = tid,
// tid + blockDim.x, ... (four per pass, their loads issued together); gradient terms go
// to this wavefront's LDS copy g*.  Returns the wavefront's energy sum (uniform).
__device__ double intra_pairs(const uint32_t* pij, const double4* ppar, int n, const double* sx,
                              const double* sy, const double* sz, double* gx, double* gy,
                              double* gz) {
  const int tid = threadIdx.x, nt = blockDim.x;
  double e = 0.0;
  for (int p0 = tid; p0 < n; p0 += 4 * nt) {
    uint32_t ijs[4];
    double4 pps[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int p = p0 + q * nt;
      ijs[q] = p < n ? pij[p] : 0u;
      pps[q] = p < n ? ppar[p] : make_double4(0.0, 0.0, 0.0, 0.0);   // ke qi qj, Ai Aj, 2 Bi Bj
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (p0 + q * nt >= n) break;
      const double4 pp = pps[q];
      const int i = ijs[q] & 0xffff, j = ijs[q] >> 16;
      const double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const bool apart = r2 >= 1.0e-4;              // r >= 0.01 (Math.non_zero_dist)
      const double irr = rsqrt(fmax(r2, 1.0e-4));   // no branch around the refinement
      const double ir = apart ? irr : 100.0;        // 1 / max(r, 0.01)
      const double u = ir * ir;
      const double u3 = u * u * u;
      const double u6 = u3 * u3;
      e += pp.x * ir + pp.y * u6 - pp.z * u3;
      if (apart) {
        // dE/dr / r = -(ke qq / r^3) - 12 AA u^7 + 6 BB2 u^4
        const double c = (-pp.x * ir * u - 12.0 * pp.y * u6 * u + 6.0 * pp.z * u3 * u);
        atomicAdd(&gx[i], c * dx); atomicAdd(&gy[i], c * dy); atomicAdd(&gz[i], c * dz);
        atomicAdd(&gx[j], -c * dx); atomicAdd(&gy[j], -c * dy); atomicAdd(&gz[j], -c * dz);
      }
    }
  }
  return wave_sum(e);
}

__device__ void load_conformer(const DofArgs& a, const LigInfo& li, double* sx, double* sy,
                               double* sz) {
  for (int j = threadIdx.x; j < li.n_atoms; j += blockDim.x) {
    const double* c = a.lig_xyz + 3 * (size_t)(li.atom_base + j);
    sx[j] = c[0]; sy[j] = c[1]; sz[j] = c[2];
  }
}

__global__ void __launch_bounds__(64) k_apply_config(DofArgs a) {
  extern __shared__ __align__(16) char smem[];
  const int pose = blockIdx.x;
  if (pose >= a.npose) return;
  const LigInfo li = a.ligs[a.pose_lig[pose]];
  const int L = li.n_atoms, lane = threadIdx.x;
  double* sx = reinterpret_cast<double*>(smem);
  double* sy = sx + L;
  double* sz = sy + L;
  const double* d = a.dofs + a.dof_off[pose];
  const TorsionLds t = stage_torsions(a, li, d, smem + dof_coord_bytes(a.max_atoms, 3));
  load_conformer(a, li, sx, sy, sz);
  forward_torsions(li, t, sx, sy, sz);   // its barriers publish the staged tree
  // centre: the registered conformer's (origin) until a torsion updates it to the mean
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if (li.nrb > 0) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int j = lane; j < L; j += kWave) { s0 += sx[j]; s1 += sy[j]; s2 += sz[j]; }
    cx = wave_sum(s0) / L; cy = wave_sum(s1) / L; cz = wave_sum(s2) / L;
  }
  double rmax = 0.0;   // Mol.radius, src/mol.ml:576-583
  for (int j = lane; j < L; j += kWave) {
    const double dx = cx - sx[j], dy = cy - sy[j], dz = cz - sz[j];
    rmax = fmax(rmax, 0.01 + sqrt(dx * dx + dy * dy + dz * dz));
  }
  rmax = wave_max(rmax);
  if (lane == 0 && a.radius_out) a.radius_out[pose] = rmax;
  double M[9];
  rot_xyz(t.sn + li.nrb, t.cs + li.nrb, M, nullptr, nullptr, nullptr);
  const int base = a.atom_off[pose];
  if (a.conf)
    for (int j = lane; j < L; j += kWave) {
      double* c = a.conf + 3 * (size_t)(base + j);
      c[0] = sx[j]; c[1] = sy[j]; c[2] = sz[j];
    }
  for (int j = lane; j < L; j += kWave) {
    double* o = a.xyz_out + 3 * (size_t)(base + j);
    o[0] = (M[0] * sx[j] + M[1] * sy[j] + M[2] * sz[j]) + d[0];
    o[1] = (M[3] * sx[j] + M[4] * sy[j] + M[5] * sz[j]) + d[1];
    o[2] = (M[6] * sx[j] + M[7] * sy[j] + M[8] * sz[j]) + d[2];
  }
}

// INTRA: E_intra and its gradient are evaluated here too, on the conformer in LDS
// (E_intra does not depend on the rigid-body DOFs): all kDofWaves wavefronts share the
// pair list as in k_intra, each into its own LDS gradient copy, summed in wavefront
// order into the conformer-frame adjoint; E_intra goes to e_intra[pose * e_stride].
// The rigid-body sums and the reverse torsion pass then run on wavefront 0 alone.
template <bool INTRA>
__global__ void __launch_bounds__(INTRA ? kDofWaves * kWave : kWave) k_dof_grad(DofArgs a) {
  extern __shared__ __align__(16) char smem[];
  __shared__ double s_part[kDofWaves];
  const int pose = blockIdx.x;
  if (pose >= a.npose) return;
  const LigInfo li = a.ligs[a.pose_lig[pose]];
  const int L = li.n_atoms, tid = threadIdx.x, lane = tid & (kWave - 1);
  double* sx = reinterpret_cast<double*>(smem);
  double* sy = sx + L;
  double* sz = sy + L;
  double* hx = sz + L;
  double* hy = hx + L;
  double* hz = hy + L;
  double* gi = hz + L;   // INTRA: kDofWaves copies of (gx, gy, gz)
  const double* d = a.dofs + a.dof_off[pose];
  double* gd = a.gdof + a.dof_off[pose];
  const TorsionLds t =
      stage_torsions(a, li, d, smem + dof_coord_bytes(a.max_atoms, INTRA ? 6 + 3 * kDofWaves : 6));
  if (INTRA)
    for (int j = tid; j < 3 * L * kDofWaves; j += blockDim.x) gi[j] = 0.0;
  if (a.conf) {
    // the conformer apply_config left after the torsions
    const int base0 = a.atom_off[pose];
    for (int j = tid; j < L; j += blockDim.x) {
      const double* c = a.conf + 3 * (size_t)(base0 + j);
      sx[j] = c[0]; sy[j] = c[1]; sz[j] = c[2];
    }
    __syncthreads();
  } else {
    load_conformer(a, li, sx, sy, sz);
    forward_torsions(li, t, sx, sy, sz);   // ends with a barrier
  }
  if (INTRA) {
    double* gx = gi + 3 * L * (tid / kWave);
    double* gy = gx + L;
    double* gz = gy + L;
    const double e = intra_pairs(a.pair_ij + li.pair_base, a.pair_par + li.pair_base, li.n_pairs,
                                 sx, sy, sz, gx, gy, gz);
    if (lane == 0) s_part[tid / kWave] = e;
    __syncthreads();
    if (tid == 0) {
      double sum = 0.0;
      for (int w = 0; w < kDofWaves; ++w) sum += s_part[w];
      a.e_intra[(size_t)pose * a.e_stride] = sum;
    }
    if (tid >= kWave) return;   // wavefront 0 goes on alone: no workgroup barrier below
  }
  double M[9], dA[9], dB[9], dG[9];
  rot_xyz(t.sn + li.nrb, t.cs + li.nrb, M, dA, dB, dG);
  const int base = a.atom_off[pose];
  double tx = 0.0, ty = 0.0, tz = 0.0, S[9];
  for (int q = 0; q < 9; ++q) S[q] = 0.0;
  for (int j = lane; j < L; j += kWave) {
    const double* g = a.grad_xyz + 3 * (size_t)(base + j);
    const double g0 = g[0], g1 = g[1], g2 = g[2];
    tx += g0; ty += g1; tz += g2;
    const double c[3] = {sx[j], sy[j], sz[j]};
    for (int r = 0; r < 3; ++r)
      for (int cc = 0; cc < 3; ++cc) S[3 * r + cc] += (r == 0 ? g0 : r == 1 ? g1 : g2) * c[cc];
    // h = M^T g (+ the conformer-frame intra gradient)
    double i0 = 0.0, i1 = 0.0, i2 = 0.0;
    if (INTRA)
      for (int w = 0; w < kDofWaves; ++w) {
        i0 += gi[3 * L * w + j];
        i1 += gi[3 * L * w + L + j];
        i2 += gi[3 * L * w + 2 * L + j];
      }
    hx[j] = (M[0] * g0 + M[3] * g1 + M[6] * g2) + i0;
    hy[j] = (M[1] * g0 + M[4] * g1 + M[7] * g2) + i1;
    hz[j] = (M[2] * g0 + M[5] * g1 + M[8] * g2) + i2;
  }
  tx = wave_sum(tx); ty = wave_sum(ty); tz = wave_sum(tz);
  double ga = 0.0, gb = 0.0, gg = 0.0;
  for (int q = 0; q < 9; ++q) {
    const double sq = wave_sum(S[q]);
    ga += dA[q] * sq; gb += dB[q] * sq; gg += dG[q] * sq;
  }
  if (lane == 0) { gd[0] = tx; gd[1] = ty; gd[2] = tz; gd[3] = ga; gd[4] = gb; gd[5] = gg; }
  wave_lds_sync();
  for (int k = li.nrb - 1; k >= 0; --k) {
    const int left = t.left[k], right = t.right[k];
    const double px = sx[right], py = sy[right], pz = sz[right];
    const double vx = px - sx[left], vy = py - sy[left], vz = pz - sz[left];
    const double im = rsqrt(vx * vx + vy * vy + vz * vz);   // 1 / |v|
    const double ax = vx * im, ay = vy * im, az = vz * im;
    const double sn = t.sn[k], cs = t.cs[k];
    const Rod R = rodrigues(ax, ay, az, sn, cs);
    const double omc = 1.0 - cs;
    const int g0 = t.off[k], g1 = t.off[k + 1];
    double tb = 0.0, abx = 0.0, aby = 0.0, abz = 0.0, pbx = 0.0, pby = 0.0, pbz = 0.0;
    for (int m = g0 + lane; m < g1; m += kWave) {
      const int i = t.atoms[m];
      const double yx = sx[i] - px, yy = sy[i] - py, yz = sz[i] - pz;   // rotated offset
      const double h0 = hx[i], h1 = hy[i], h2 = hz[i];
      // dE/dtau: h . (a x y)
      tb += h0 * (ay * yz - az * yy) + h1 * (az * yx - ax * yz) + h2 * (ax * yy - ay * yx);
      // pre-rotation offset z = R^T y, and R^T h
      const double zx = R.m[0] * yx + R.m[3] * yy + R.m[6] * yz;
      const double zy = R.m[1] * yx + R.m[4] * yy + R.m[7] * yz;
      const double zz = R.m[2] * yx + R.m[5] * yy + R.m[8] * yz;
      const double rx = R.m[0] * h0 + R.m[3] * h1 + R.m[6] * h2;
      const double ry = R.m[1] * h0 + R.m[4] * h1 + R.m[7] * h2;
      const double rz = R.m[2] * h0 + R.m[5] * h1 + R.m[8] * h2;
      const double az_ = ax * zx + ay * zy + az * zz, ah = ax * h0 + ay * h1 + az * h2;
      abx += (zy * h2 - zz * h1) * sn + h0 * az_ * omc + zx * ah * omc;
      aby += (zz * h0 - zx * h2) * sn + h1 * az_ * omc + zy * ah * omc;
      abz += (zx * h1 - zy * h0) * sn + h2 * az_ * omc + zz * ah * omc;
      pbx += h0 - rx; pby += h1 - ry; pbz += h2 - rz;
      hx[i] = rx; hy[i] = ry; hz[i] = rz;
      sx[i] = zx + px; sy[i] = zy + py; sz[i] = zz + pz;   // back to c^k
    }
    tb = wave_sum(tb);
    abx = wave_sum(abx); aby = wave_sum(aby); abz = wave_sum(abz);
    pbx = wave_sum(pbx); pby = wave_sum(pby); pbz = wave_sum(pbz);
    wave_lds_sync();
    if (lane == 0) {
      gd[6 + k] = tb;
      const double adot = ax * abx + ay * aby + az * abz;
      const double vbx = (abx - ax * adot) * im, vby = (aby - ay * adot) * im,
                   vbz = (abz - az * adot) * im;
      hx[right] += vbx + pbx; hy[right] += vby + pby; hz[right] += vbz + pbz;
      hx[left] -= vbx; hy[left] -= vby; hz[left] -= vbz;
    }
    wave_lds_sync();
  }
}

size_t dof_lds_bytes(const DofArgs& a, bool grad) {
  const int arrays = grad ? (a.e_intra ? 6 + 3 * kDofWaves : 6) : 3;
  return dof_coord_bytes(a.max_atoms, arrays) + torsion_lds_bytes(a.max_rb, a.max_rg);
}

// the launches above 64 KiB of LDS raise the kernel's dynamic-LDS limit (160 KiB per
// workgroup on gfx950); callers check dof_lds_bytes against kMaxLds first
template <class K>
static hipError_t lds_limit(K* kernel, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_apply_config(const DofArgs& a, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  const size_t lds = dof_lds_bytes(a, false);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  hipError_t e = lds_limit(&k_apply_config, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_apply_config, dim3(a.npose), dim3(kWave), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_dof_grad(const DofArgs& a, hipStream_t s) {
  if (a.npose == 0) return hipSuccess;
  const size_t lds = dof_lds_bytes(a, true);
  if (lds > kMaxLds) return hipErrorInvalidConfiguration;
  if (a.e_intra) {
    hipError_t e = lds_limit(&k_dof_grad<true>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dof_grad<true>, dim3(a.npose), dim3(kDofWaves * kWave), lds, s, a);
  } else {
    hipError_t e = lds_limit(&k_dof_grad<false>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dof_grad<false>, dim3(a.npose), dim3(kWave), lds, s, a);
  }
  return hipGetLastError();
}

}  // namespace mmo
