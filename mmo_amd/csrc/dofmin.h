// dofmin.h -- one wavefront's DOF-space energy + gradient and its L-BFGS update, shared
// by the batched minimiser kernels (min_kernels.hip) and the Monte Carlo chains with
// local minimisation (mc_kernels.hip).
//
// A pose is Optim.apply_config (src/optim.ml:64-80) of a base conformer: the torsions
// tau_0..tau_{T-1} turn the rotation groups in rbond order (Mol.rotate_bond,
// src/mol.ml:610-631), then x = r_xyz(alpha, beta, gamma) c + t.  The wavefront holds the
// DOF vector in its lanes (lane i = DOF i, at most 64), the posed conformer and the
// adjoint in LDS, and evaluates
//   E = E_inter(x) + E_intra(c)  and  dE/dDOF
// with E_inter the shifted direct sum (FP32 pairs through the per-atom sweep, or FP64
// pairs in the parity mode) or trilinear interpolation, E_intra the FP64 UFF_NB pairs,
// and the gradient chained through the rigid motion and the reverse torsion pass (the
// same adjoint as k_dof_grad).  Poses longer than 12 A after the torsions (Mol.Too_long)
// evaluate to +inf.
#pragma once
#include "common.h"
#include "cull.h"
#include "dof.h"
#include "kernels.h"

namespace mmo {

// L-BFGS constants of a run (set by the host)
struct MinCfg {
  int m;                 // history pairs (1..kMinMaxHist)
  double max_step;       // first trial step per component
  double min_step;       // step below which the history is dropped / the run converges
  double gtol;           // convergence at max |g| <= gtol
  double c1;             // Armijo constant
};

// One L-BFGS update of a pose after an evaluation at the trial point: lane i holds DOF i
// (act = i < ndof); x, g, d are the accepted point, its gradient and the direction, xt
// the trial point just evaluated (energy ft, gradient gt, radius rad).  S, Y point at
// this lane's entry of history slot 0 (slot k at + k * slot_stride), rho at the pose's
// slot 0 (slot k at + k * rho_stride).  Returns the next trial point (projected onto
// [lo, hi]) or x once the pose has finished.  The control flow is the batched minimiser's:
// Armijo acceptance, curvature pairs with s.y > 0, two-loop recursion newest first,
// restart on a non-descent direction, step halving, restart or convergence below min_step.
__device__ __forceinline__ double lbfgs_wave(MinPose& s, double& x, double& g, double& d,
                                             double xt, double ft, double gt, double rad,
                                             double lo, double hi, bool act, double* S, double* Y,
                                             size_t slot_stride, double* rho, size_t rho_stride,
                                             const MinCfg& c) {
  const int lane = threadIdx.x & (kWave - 1);
  if (!(rad <= 12.0) || !isfinite(ft)) ft = __builtin_inf();
  s.n_eval += 1;
  auto dot = [](double a, double b) { return wave_sum(a * b); };
  auto absmax = [](double v) { return wave_max(fabs(v)); };
  if (!s.init) {
    // first evaluation: the start point itself
    x = xt;
    g = gt;
    s.f = ft;
    s.f0 = ft;
    s.init = 1;
    s.hist = 0;
    s.head = 0;
    if (!isfinite(ft)) s.status = kMinFailed;
    d = -g;
    const double dm = absmax(act ? d : 0.0);
    s.alpha = dm > 0.0 ? fmin(1.0, c.max_step / dm) : 0.0;
    if (dm <= c.gtol) s.status = kMinConverged;
  } else {
    const double sv = xt - x;                  // the step actually taken (after projection)
    const double gs = dot(act ? g : 0.0, act ? sv : 0.0);
    if (gs < 0.0 && ft <= s.f + c.c1 * gs) {
      // accept; curvature pair
      const double yv = gt - g;
      const double sy = dot(act ? sv : 0.0, act ? yv : 0.0);
      const double yy = dot(act ? yv : 0.0, act ? yv : 0.0);
      if (sy > 1e-12 * yy && sy > 0.0) {
        const int slot = s.head;
        if (act) {
          S[(size_t)slot * slot_stride] = sv;
          Y[(size_t)slot * slot_stride] = yv;
        }
        if (lane == 0) rho[(size_t)slot * rho_stride] = 1.0 / sy;
        s.head = (s.head + 1) % c.m;
        s.hist = min(s.hist + 1, c.m);
        s.gamma = sy / yy;
      }
      x = xt;
      g = gt;
      s.f = ft;
      s.n_acc += 1;
      // two-loop recursion, newest pair first; alpha_k is kept in lane k of al_l (a
      // register, where an indexed array would live in scratch memory)
      double q = act ? g : 0.0;
      double al_l = 0.0;
      for (int k = 0; k < s.hist; ++k) {
        const int idx = (s.head - 1 - k + 2 * c.m) % c.m;
        const double sk = act ? S[(size_t)idx * slot_stride] : 0.0;
        const double yk = act ? Y[(size_t)idx * slot_stride] : 0.0;
        const double rk = rho[(size_t)idx * rho_stride];
        const double alk = rk * dot(sk, q);
        al_l = lane == k ? alk : al_l;
        q -= alk * yk;
      }
      double r = (s.hist > 0 ? s.gamma : 1.0) * q;
      for (int k = s.hist - 1; k >= 0; --k) {
        const int idx = (s.head - 1 - k + 2 * c.m) % c.m;
        const double sk = act ? S[(size_t)idx * slot_stride] : 0.0;
        const double yk = act ? Y[(size_t)idx * slot_stride] : 0.0;
        const double rk = rho[(size_t)idx * rho_stride];
        const double b = rk * dot(yk, r);
        r += sk * (readlane_d(al_l, k) - b);
      }
      d = act ? -r : 0.0;
      if (dot(d, act ? g : 0.0) >= 0.0) {   // not a descent direction: restart
        s.hist = 0;
        s.n_reset += 1;
        d = act ? -g : 0.0;
      }
      const double dm = absmax(d);
      const double gm = absmax(act ? g : 0.0);
      s.alpha = dm > 0.0 ? fmin(1.0, c.max_step / dm) : 0.0;
      if (gm <= c.gtol) s.status = kMinConverged;
    } else {
      // backtrack
      s.alpha *= 0.5;
      const double dm = absmax(act ? d : 0.0);
      if (s.alpha * dm < c.min_step) {
        if (s.hist == 0) {
          s.status = kMinConverged;        // no decrease along steepest descent
        } else {                           // drop the history and restart
          s.hist = 0;
          s.n_reset += 1;
          d = act ? -g : 0.0;
          const double dm2 = absmax(d);
          s.alpha = dm2 > 0.0 ? fmin(1.0, c.max_step / dm2) : 0.0;
        }
      }
    }
  }
  return (s.status == kMinRunning) ? fmin(hi, fmax(lo, x + s.alpha * d)) : x;
}

// ------------------------------------------------------------------ evaluation
// What one wavefront needs to evaluate a pose of one ligand.
struct DofEval {
  LigInfo li;
  const LigAtom* latoms;
  const int* atom_order;        // nullable: per ligand at atom_base, the sweep's atom order
  const int* rb_left;           // per ligand at rb_base
  const int* rb_right;
  const int* rg_off;            // per ligand at rg_base (nrb + 1 absolute offsets)
  const int* rg_atoms;
  const uint32_t* pair_ij;      // intra pair list (i | j << 16) and FP64 products
  const double4* pair_par;
  int ntors;                    // torsion DOFs (0 for rigid docking), <= li.nrb
  bool intra;                   // add E_intra (UFF_NB)
  // receptor (direct forms)
  Frame fr;
  const RecPos* s_pos;
  const RecAB* s_ab;
  const int* s_cell;
  const RecRes* rres;
  CellGrid cg;
  int* wl;                      // this wavefront's atom-sweep scratch (LDS, kAtomSweepInts)
  // grids (interpolated form)
  GridDev g;
};

// src/mol.ml:881-903 over the ligand's pair list for the conformer in LDS, FP64, the
// lanes of one wavefront taking pairs lane, lane + 64, ...; the gradient goes into the
// wavefront's own LDS arrays (LDS atomics, so that two lanes may hit one atom).
__device__ __forceinline__ double intra_pairs_wave(const uint32_t* pij, const double4* ppar, int n,
                                                   const double* sx, const double* sy,
                                                   const double* sz, double* gx, double* gy,
                                                   double* gz, bool grad) {
  const int lane = threadIdx.x & (kWave - 1);
  double e = 0.0;
  // four pairs per lane per pass, their loads issued together (one memory latency per
  // four pairs); each lane still takes its pairs lane, lane + 64, ... in that order
  for (int p0 = lane; p0 < n; p0 += 4 * kWave) {
    uint32_t ijs[4];
    double4 pps[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int p = p0 + q * kWave;
      ijs[q] = p < n ? pij[p] : 0u;
      pps[q] = p < n ? ppar[p] : make_double4(0.0, 0.0, 0.0, 0.0);   // ke qi qj, Ai Aj, 2 Bi Bj
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (p0 + q * kWave >= n) break;
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
      if (grad && apart) {
        const double c = (-pp.x * ir * u - 12.0 * pp.y * u6 * u + 6.0 * pp.z * u3 * u);
        atomicAdd(&gx[i], c * dx); atomicAdd(&gy[i], c * dy); atomicAdd(&gz[i], c * dz);
        atomicAdd(&gx[j], -c * dx); atomicAdd(&gy[j], -c * dy); atomicAdd(&gz[j], -c * dz);
      }
    }
  }
  return wave_sum(e);
}

// Shifted direct sum with FP64 pair arithmetic and its gradient (the parity mode): two
// atoms per lane, the cell rows within 12 A of the wavefront's bounding sphere; exact
// fixed-point differences, r = max(sqrt(r2), 0.01), w = (1 - r^2/144)^2 (src/mol.ml:822-849).
// pos(j, x, y, z) gives atom j; gout(j, gx, gy, gz) receives dE/dx_j (A^-1) from the
// lane holding atom j.  Returns the wavefront's energy (uniform).
template <class PosFn, class GradOut>
__device__ double f64_sweep_grad(const DofEval& ev, int n_atoms, PosFn&& pos, GradOut&& gout) {
  constexpr int K = 2;
  const int lane = threadIdx.x & (kWave - 1);
  const Frame& fr = ev.fr;
  const CellGrid& cg = ev.cg;
  const double inv_s = 1.0 / fr.scale;
  double e = 0.0;
  int sat = 0;
  const int nblk = (n_atoms + kWave * K - 1) / (kWave * K);
  for (int blk = 0; blk < nblk; ++blk) {
    int lx[K], ly[K], lz[K];
    bool valid[K];
    double cq[K], ca[K], cb[K], gx[K], gy[K], gz[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = lane + kWave * (blk * K + k);
      valid[k] = j < n_atoms;
      gx[k] = gy[k] = gz[k] = 0.0;
      lx[k] = ly[k] = lz[k] = 0;
      cq[k] = ca[k] = cb[k] = 0.0;
      if (valid[k]) {
        double x, y, z;
        pos(j, x, y, z);
        lx[k] = to_fixed(x, fr.ox, fr.scale, fr.lim, &sat);
        ly[k] = to_fixed(y, fr.oy, fr.scale, fr.lim, &sat);
        lz[k] = to_fixed(z, fr.oz, fr.scale, fr.lim, &sat);
        const LigAtom la = ev.latoms[ev.li.atom_base + j];
        cq[k] = la.keq; ca[k] = la.a; cb[k] = la.b;
      }
    }
    float qx, qy, qz, R;
    if (wave_query<K>(lx, ly, lz, valid, fr.inv_s, qx, qy, qz, R)) {
      const float R2 = R * R;
      const int cz0 = max(0, (int)floorf((qz - R - cg.z0) * cg.inv_c));
      const int cz1 = min(cg.nz - 1, (int)floorf((qz + R - cg.z0) * cg.inv_c));
      for (int cz = cz0; cz <= cz1; ++cz) {
        const float zlo = cg.z0 + cz * cg.c;
        const float dzb = fmaxf(0.0f, fmaxf(zlo - qz, qz - (zlo + cg.c)));
        const float rz2 = R2 - dzb * dzb;
        if (rz2 < 0.0f) continue;
        const float ry = bound_sqrt(rz2);
        const int cy0 = max(0, (int)floorf((qy - ry - cg.y0) * cg.inv_c));
        const int cy1 = min(cg.ny - 1, (int)floorf((qy + ry - cg.y0) * cg.inv_c));
        for (int cy = cy0; cy <= cy1; ++cy) {
          const float ylo = cg.y0 + cy * cg.c;
          const float dyb = fmaxf(0.0f, fmaxf(ylo - qy, qy - (ylo + cg.c)));
          const float rx2 = rz2 - dyb * dyb;
          if (rx2 < 0.0f) continue;
          const float rx = bound_sqrt(rx2);
          const int cx0 = max(0, (int)floorf((qx - rx - cg.x0) * cg.inv_cx));
          const int cx1 = min(cg.nx - 1, (int)floorf((qx + rx - cg.x0) * cg.inv_cx));
          if (cx0 > cx1) continue;
          const int rowb = (cz * cg.ny + cy) * cg.nx;
          const int a0 = uniformi(ev.s_cell[rowb + cx0]);
          const int a1 = uniformi(ev.s_cell[rowb + cx1 + 1]);
          for (int i = a0; i < a1; ++i) {
            const RecPos p = ev.s_pos[i];
            const RecAB ab = ev.s_ab[i];
#pragma unroll
            for (int k = 0; k < K; ++k) {
              // d = x_rec - x_lig
              const double dx = (double)(p.x - lx[k]) * inv_s, dy = (double)(p.y - ly[k]) * inv_s,
                           dz = (double)(p.z - lz[k]) * inv_s;
              const double r2 = dx * dx + dy * dy + dz * dz;
              if (!(r2 < 144.0)) continue;
              const double rr = sqrt(r2);
              const bool apart = rr >= 0.01;
              const double r = apart ? rr : 0.01;        // Math.non_zero_dist
              const double t = 1.0 - (r / 12.0) * (r / 12.0);
              const double w = t * t;                     // FF.shift_12A
              const double ir = 1.0 / r, u = ir * ir, u3 = u * u * u;
              const double eq = cq[k] * (double)p.q * ir, e12 = ca[k] * (double)ab.a * u3 * u3,
                           e6 = cb[k] * (double)ab.b * u3;
              const double F = eq + e12 - 2.0 * e6;
              e += w * F;
              if (apart) {
                // dE/dr2 = -t/72 F - w u (eq / 2 + 6 e12 - 6 e6); dE/dx_lig = -2 dE/dr2 d
                const double coef = -t / 72.0 * F - w * u * (0.5 * eq + 6.0 * e12 - 6.0 * e6);
                gx[k] -= 2.0 * coef * dx; gy[k] -= 2.0 * coef * dy; gz[k] -= 2.0 * coef * dz;
              }
            }
          }
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = lane + kWave * (blk * K + k);
      if (valid[k]) gout(j, gx[k], gy[k], gz[k]);
    }
  }
  (void)sat;   // saturated atoms stay beyond the cutoff of every receptor atom
  return wave_sum(e);
}

// Torsions x_tors (lane 6 + k holds tau_k; sn, cs its sine and cosine) on the conformer
// in LDS, in rbond order, about the current positions; then the enclosing radius about
// the mean (about the origin without torsions, like k_apply_config).
__device__ __forceinline__ double dof_torsions(const DofEval& ev, double sn, double cs, double* tx,
                                              double* ty, double* tz) {
  const int lane = threadIdx.x & (kWave - 1);
  const LigInfo& li = ev.li;
  const int L = li.n_atoms;
  for (int k = 0; k < ev.ntors; ++k) {
    const int left = ev.rb_left[li.rb_base + k], right = ev.rb_right[li.rb_base + k];
    const double px = tx[right], py = ty[right], pz = tz[right];
    const double vx = px - tx[left], vy = py - ty[left], vz = pz - tz[left];
    const double im = rsqrt(vx * vx + vy * vy + vz * vz);   // V3.normalize: v / |v|
    const Rod R = rodrigues(vx * im, vy * im, vz * im, readlane_d(sn, 6 + k), readlane_d(cs, 6 + k));
    const int g0 = ev.rg_off[li.rg_base + k], g1 = ev.rg_off[li.rg_base + k + 1];
    // the axis atoms are not in the group: no lane's move changes what the others read
    for (int m = g0 + lane; m < g1; m += kWave) {
      const int i = ev.rg_atoms[m];
      const double zx = tx[i] - px, zy = ty[i] - py, zz = tz[i] - pz;
      tx[i] = (R.m[0] * zx + R.m[1] * zy + R.m[2] * zz) + px;
      ty[i] = (R.m[3] * zx + R.m[4] * zy + R.m[5] * zz) + py;
      tz[i] = (R.m[6] * zx + R.m[7] * zy + R.m[8] * zz) + pz;
    }
    wave_lds_sync();
  }
  double cx = 0.0, cy = 0.0, cz = 0.0;
  if (ev.ntors > 0) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int j = lane; j < L; j += kWave) { s0 += tx[j]; s1 += ty[j]; s2 += tz[j]; }
    cx = wave_sum(s0) / L; cy = wave_sum(s1) / L; cz = wave_sum(s2) / L;
  }
  double rmax = 0.0;   // Mol.radius, src/mol.ml:576-583
  for (int j = lane; j < L; j += kWave) {
    const double dx = cx - tx[j], dy = cy - ty[j], dz = cz - tz[j];
    rmax = fmax(rmax, 0.01 + sqrt(dx * dx + dy * dy + dz * dz));
  }
  return wave_max(rmax);
}

// The pose of DOF vector x (lane i = DOF i): base rows (HBM, lane j reads the rows it
// owns) -> torsions -> conformer t (LDS); M, T the rigid motion.  Returns the radius.
__device__ __forceinline__ double dof_pose(const DofEval& ev, const double* base, double x,
                                           double* tx, double* ty, double* tz, double* M,
                                           double* dA, double* dB, double* dG, double* T,
                                           double& sn, double& cs) {
  const int lane = threadIdx.x & (kWave - 1);
  const int L = ev.li.n_atoms;
  for (int j = lane; j < L; j += kWave) {
    tx[j] = base[3 * j]; ty[j] = base[3 * j + 1]; tz[j] = base[3 * j + 2];
  }
  sincos(x, &sn, &cs);
  wave_lds_sync();
  const double rad = dof_torsions(ev, sn, cs, tx, ty, tz);
  double s3[3], c3[3];
  for (int k = 0; k < 3; ++k) {
    s3[k] = readlane_d(sn, 3 + k);
    c3[k] = readlane_d(cs, 3 + k);
    T[k] = readlane_d(x, k);
  }
  rot_xyz(s3, c3, M, dA, dB, dG);   // dA, dB, dG nullable
  return rad;
}

// E and dE/dDOF (lane i) of DOF vector x for one wavefront.  tx.. (LDS, 3 x max_atoms)
// hold the conformer during the evaluation (the reverse pass returns them to the base
// rows' values), hx.. (LDS) serve as the adjoint.  e_inter / e_intra receive the two
// parts.  Returns E (+inf for a pose longer than 12 A after its torsions, NaN for an
// interpolated pose outside the grid).
template <int INTER>
__device__ double dof_eval(const DofEval& ev, const double* base, double x, double* tx, double* ty,
                           double* tz, double* hx, double* hy, double* hz, double& gdof,
                           double& e_inter, double& e_intra, double& rad) {
  const int lane = threadIdx.x & (kWave - 1);
  const LigInfo& li = ev.li;
  const int L = li.n_atoms;
  double M[9], T[3], sn, cs;
  rad = dof_pose(ev, base, x, tx, ty, tz, M, nullptr, nullptr, nullptr, T, sn, cs);
  gdof = 0.0;
  e_inter = e_intra = 0.0;
  if (!(rad <= 12.0)) return __builtin_inf();   // Mol.Too_long (src/optim.ml:72-75)
  auto posed = [&](int j, double& x0, double& y0, double& z0) {
    const double a = tx[j], b = ty[j], c = tz[j];
    x0 = (M[0] * a + M[1] * b + M[2] * c) + T[0];
    y0 = (M[3] * a + M[4] * b + M[5] * c) + T[1];
    z0 = (M[6] * a + M[7] * b + M[8] * c) + T[2];
  };
  // E_inter and dE/dx in the pose frame, into h
  if (INTER == kDofInterAtom) {
    int sat = 0;
    auto at = [&](int j) { return ev.atom_order ? ev.atom_order[li.atom_base + j] : j; };
    auto atom = [&](int j, LigPt& l, double& cq, double& ca, double& cb) {
      const int ja = at(j);
      double px, py, pz;
      posed(ja, px, py, pz);
      l = lig_pt(px, py, pz, ev.fr, &sat);
      const LigAtom la = ev.latoms[li.atom_base + ja];
      cq = la.keq; ca = la.a; cb = la.b;
    };
    auto gout = [&](int j, double gx, double gy, double gz) {
      const int ja = at(j);
      hx[ja] = gx; hy[ja] = gy; hz[ja] = gz;
    };
    double e_el = 0.0, e_vd = 0.0;
    atom_sweep<true, decltype(atom)&, decltype(gout)&, true, true, 4>(
        ev.cg, ev.s_cell, ev.s_pos, ev.s_ab, ev.rres, ev.fr, L, ev.wl, atom, gout, e_el, e_vd);
    (void)sat;   // saturated atoms stay beyond the cutoff of every receptor atom
    e_inter = wave_sum(e_el + e_vd);
  } else if (INTER == kDofInterF64) {
    auto gout = [&](int j, double gx, double gy, double gz) { hx[j] = gx; hy[j] = gy; hz[j] = gz; };
    e_inter = f64_sweep_grad(ev, L, posed, gout);
  } else {
    double e = 0.0;
    int bad = 0;
    for (int j = lane; j < L; j += kWave) {
      double px, py, pz;
      posed(j, px, py, pz);
      bool ok = true;
      double dg[3] = {0.0, 0.0, 0.0};
      e += trilin<true>(ev.g, ev.latoms[li.atom_base + j].type, px - ev.g.ox, py - ev.g.oy,
                        pz - ev.g.oz, ok, dg);
      if (!ok) bad = 1;
      hx[j] = dg[0]; hy[j] = dg[1]; hz[j] = dg[2];
    }
    e_inter = wave_sum(e);
    if (__ballot(bad) != 0ull) e_inter = __builtin_nan("");
  }
  wave_lds_sync();
  // rigid-body part: dE/dt = sum g, dE/dangle = sum_q dM_q S_q with S = sum g c^T; then the
  // conformer-frame adjoint h = M^T g
  double tsx = 0.0, tsy = 0.0, tsz = 0.0, S[9];
  for (int q = 0; q < 9; ++q) S[q] = 0.0;
  for (int j = lane; j < L; j += kWave) {
    const double g0 = hx[j], g1 = hy[j], g2 = hz[j];
    const double c0 = tx[j], c1 = ty[j], c2 = tz[j];
    tsx += g0; tsy += g1; tsz += g2;
    S[0] += g0 * c0; S[1] += g0 * c1; S[2] += g0 * c2;
    S[3] += g1 * c0; S[4] += g1 * c1; S[5] += g1 * c2;
    S[6] += g2 * c0; S[7] += g2 * c1; S[8] += g2 * c2;
    hx[j] = M[0] * g0 + M[3] * g1 + M[6] * g2;
    hy[j] = M[1] * g0 + M[4] * g1 + M[7] * g2;
    hz[j] = M[2] * g0 + M[5] * g1 + M[8] * g2;
  }
  double ga = 0.0, gb = 0.0, gg = 0.0;
  {
    // the partial derivatives of r_xyz, only now (they are not live across the sweep)
    double s3[3], c3[3], Mr[9], dA[9], dB[9], dG[9];
    for (int k = 0; k < 3; ++k) {
      s3[k] = readlane_d(sn, 3 + k);
      c3[k] = readlane_d(cs, 3 + k);
    }
    rot_xyz(s3, c3, Mr, dA, dB, dG);
    for (int q = 0; q < 9; ++q) {
      const double sq = wave_sum(S[q]);
      ga += dA[q] * sq; gb += dB[q] * sq; gg += dG[q] * sq;
    }
  }
  tsx = wave_sum(tsx); tsy = wave_sum(tsy); tsz = wave_sum(tsz);
  gdof = lane == 0 ? tsx : lane == 1 ? tsy : lane == 2 ? tsz : lane == 3 ? ga : lane == 4 ? gb
       : lane == 5 ? gg : 0.0;
  wave_lds_sync();
  if (ev.intra)
    e_intra = intra_pairs_wave(ev.pair_ij + li.pair_base, ev.pair_par + li.pair_base, li.n_pairs,
                               tx, ty, tz, hx, hy, hz, true);
  wave_lds_sync();
  // reverse torsion pass (the adjoint of dof_torsions, as in k_dof_grad)
  for (int k = ev.ntors - 1; k >= 0; --k) {
    const int left = ev.rb_left[li.rb_base + k], right = ev.rb_right[li.rb_base + k];
    const double px = tx[right], py = ty[right], pz = tz[right];
    const double vx = px - tx[left], vy = py - ty[left], vz = pz - tz[left];
    const double im = rsqrt(vx * vx + vy * vy + vz * vz);   // 1 / |v|
    const double ax = vx * im, ay = vy * im, az = vz * im;
    const double s = readlane_d(sn, 6 + k), c = readlane_d(cs, 6 + k);
    const Rod R = rodrigues(ax, ay, az, s, c);
    const double omc = 1.0 - c;
    const int g0 = ev.rg_off[li.rg_base + k], g1 = ev.rg_off[li.rg_base + k + 1];
    double tb = 0.0, abx = 0.0, aby = 0.0, abz = 0.0, pbx = 0.0, pby = 0.0, pbz = 0.0;
    for (int m = g0 + lane; m < g1; m += kWave) {
      const int i = ev.rg_atoms[m];
      const double yx = tx[i] - px, yy = ty[i] - py, yz = tz[i] - pz;   // rotated offset
      const double h0 = hx[i], h1 = hy[i], h2 = hz[i];
      tb += h0 * (ay * yz - az * yy) + h1 * (az * yx - ax * yz) + h2 * (ax * yy - ay * yx);
      const double zx = R.m[0] * yx + R.m[3] * yy + R.m[6] * yz;
      const double zy = R.m[1] * yx + R.m[4] * yy + R.m[7] * yz;
      const double zz = R.m[2] * yx + R.m[5] * yy + R.m[8] * yz;
      const double rx = R.m[0] * h0 + R.m[3] * h1 + R.m[6] * h2;
      const double ry = R.m[1] * h0 + R.m[4] * h1 + R.m[7] * h2;
      const double rz = R.m[2] * h0 + R.m[5] * h1 + R.m[8] * h2;
      const double az_ = ax * zx + ay * zy + az * zz, ah = ax * h0 + ay * h1 + az * h2;
      abx += (zy * h2 - zz * h1) * s + h0 * az_ * omc + zx * ah * omc;
      aby += (zz * h0 - zx * h2) * s + h1 * az_ * omc + zy * ah * omc;
      abz += (zx * h1 - zy * h0) * s + h2 * az_ * omc + zz * ah * omc;
      pbx += h0 - rx; pby += h1 - ry; pbz += h2 - rz;
      hx[i] = rx; hy[i] = ry; hz[i] = rz;
      tx[i] = zx + px; ty[i] = zy + py; tz[i] = zz + pz;   // back to c^k
    }
    tb = wave_sum(tb);
    abx = wave_sum(abx); aby = wave_sum(aby); abz = wave_sum(abz);
    pbx = wave_sum(pbx); pby = wave_sum(pby); pbz = wave_sum(pbz);
    gdof = lane == 6 + k ? tb : gdof;
    wave_lds_sync();
    if (lane == 0) {
      const double adot = ax * abx + ay * aby + az * abz;
      const double vbx = (abx - ax * adot) * im, vby = (aby - ay * adot) * im,
                   vbz = (abz - az * adot) * im;
      hx[right] += vbx + pbx; hy[right] += vby + pby; hz[right] += vbz + pbz;
      hx[left] -= vbx; hy[left] -= vby; hz[left] -= vbz;
    }
    wave_lds_sync();
  }
  return e_inter + e_intra;
}

}  // namespace mmo
