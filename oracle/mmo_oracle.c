/*
 * mmo_oracle.c -- FP64 CPU restatement of MMO's UFF nonbonded scoring path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded only by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product path
 * never uses it.  Each function names the reference function it restates
 * (UnixJunkie/MMO, src/<file>:<line>).  All arithmetic is FP64 with the
 * reference's loop order and clamps; grids are FP32 storage as in
 * src/G3D.ml:11 / src/BA1.ml:7.
 *
 * Parity status (details in the test files):
 *   pair terms, BrG, BrL, intra      pinned: closed forms (LJ minimum at x_ij, zero at
 *                                    x_ij/2^(1/6), Coulomb), shift values, clamp,
 *                                    mpmath brute force, invariants
 *   trilinear / interp               pinned: exact on nodes and affine fields, corner order
 *   grid energies                    pinned: node value == direct shifted sum, clamp, mask
 *   rotations / SO3 / dihedral       pinned: known matrices, unit quaternions, the 8
 *                                    dihedral known answers of src/tests.ml:47-54
 *   rotate_bond / apply_config       pinned: bond lengths preserved, +tau then -tau identity,
 *                                    apply_config with zero DOFs is the identity
 *   gradients                        pinned: central finite differences of the energies
 *   last-ulp of L.frange / Vector3.normalize: parity unpinned (ulp level only)
 */
#include "mmo_oracle.h"
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/UFF.ml:10-22 : (anum, (x_i [A], D_i [kcal/mol])) */
static const struct { int anum; double x, d; } UFF_TABLE[12] = {
  { 0, 0.0,   0.0  }, { 1, 2.886, 0.044}, { 6, 3.851, 0.105}, { 7, 3.660, 0.069},
  { 8, 3.500, 0.060}, { 9, 3.364, 0.050}, {12, 3.021, 0.111}, {15, 4.147, 0.305},
  {16, 4.035, 0.274}, {17, 3.947, 0.227}, {35, 4.189, 0.251}, {53, 4.500, 0.339}};

#define CUTOFF2 144.0          /* src/mol.ml:836 (12*12) */
#define MAX_E   100000.0       /* src/params.ml:26 */

double orc_elec_weight(void) { return 332.0637 / 4.0; }   /* src/UFF.ml:25 */

int orc_uff_params(int anum, double* x, double* d) {
  for (int k = 0; k < 12; ++k)
    if (UFF_TABLE[k].anum == anum) { *x = UFF_TABLE[k].x; *d = UFF_TABLE[k].d; return 0; }
  *x = NAN; *d = NAN;  /* xidi_params is initialised to nan, src/UFF.ml:37 */
  return -1;
}

/* src/UFF.ml:39-51 with FF.geo_mean src/FF.ml:14-15 */
void orc_uff_pair(int a1, int a2, double* x_ij, double* d_ij) {
  double x1, d1, x2, d2;
  int ok1 = orc_uff_params(a1, &x1, &d1), ok2 = orc_uff_params(a2, &x2, &d2);
  if (ok1 || ok2) { *x_ij = NAN; *d_ij = NAN; return; }
  *x_ij = sqrt(x1 * x2);
  *d_ij = sqrt(d1 * d2);
}

double orc_non_zero_dist(double r) { return r < 0.01 ? 0.01 : r; }   /* src/math.ml:58-62 */
double orc_pow6(double x) { double x2 = x * x; return x2 * x2 * x2; }  /* src/FF.ml:8-12: pow3 (x*x) */
double orc_shift_12A(double d) {                                        /* src/FF.ml:17-20 */
  if (d < 12.0) { double s = d / 12.0; double t = 1.0 - s * s; return t * t; }
  return 0.0;
}

/* pair parameter cache (119x119 in the reference, src/UFF.ml:32-40) */
typedef struct { double x, d; } xd_t;
static void pair_params(int a1, int a2, xd_t* p) { orc_uff_pair(a1, a2, &p->x, &p->d); }

/* ------------------------------------------------------------------------ */
/* src/mol.ml:796-818  ene_inter_UFF_global_brute: protein outer, ligand inner */
double orc_ene_inter_global(const double* pxs, const double* pys, const double* pzs,
                            const double* pq, const int32_t* panum, int np,
                            const double* lxs, const double* lys, const double* lzs,
                            const double* lq, const int32_t* lanum, int nl,
                            double* e_elec, double* e_vdw) {
  double sum_elec = 0.0, sum_vdw = 0.0;
  for (int i = 0; i < np; ++i) {
    const double qi = pq[i];
    for (int j = 0; j < nl; ++j) {
      const double dx = pxs[i] - lxs[j], dy = pys[i] - lys[j], dz = pzs[i] - lzs[j];
      const double r = orc_non_zero_dist(sqrt(dx * dx + dy * dy + dz * dz));
      xd_t v; pair_params(panum[i], lanum[j], &v);
      const double p6 = orc_pow6(v.x / r);
      sum_elec = sum_elec + (qi * lq[j]) / r;
      sum_vdw = sum_vdw + v.d * ((-2.0 * p6) + (p6 * p6));
    }
  }
  const double ke = orc_elec_weight();
  if (e_elec) *e_elec = ke * sum_elec;
  if (e_vdw) *e_vdw = sum_vdw;
  return (ke * sum_elec) + sum_vdw;
}

/* src/mol.ml:822-849  ene_inter_UFF_shifted_brute (BrL); components as in
 * ene_inter_UFF_shifted_bst_components src/mol.ml:928-956 (elec pre-multiplied) */
double orc_ene_inter_shifted(const double* pxs, const double* pys, const double* pzs,
                             const double* pq, const int32_t* panum, int np,
                             const double* lxs, const double* lys, const double* lzs,
                             const double* lq, const int32_t* lanum, int nl,
                             double* e_elec, double* e_vdw) {
  double sum_elec = 0.0, sum_vdw = 0.0;
  for (int i = 0; i < np; ++i) {
    const double qi = pq[i];
    for (int j = 0; j < nl; ++j) {
      const double dx = pxs[i] - lxs[j], dy = pys[i] - lys[j], dz = pzs[i] - lzs[j];
      const double r2 = dx * dx + dy * dy + dz * dz;   /* V3.dist2, src/V3.ml:11-16 */
      if (r2 < CUTOFF2) {
        const double r = orc_non_zero_dist(sqrt(r2));
        const double w = orc_shift_12A(r);
        xd_t v; pair_params(panum[i], lanum[j], &v);
        const double p6 = orc_pow6(v.x / r);
        sum_elec = sum_elec + w * ((qi * lq[j]) / r);
        sum_vdw = sum_vdw + w * (v.d * ((-2.0 * p6) + (p6 * p6)));
      }
    }
  }
  const double ke = orc_elec_weight();
  if (e_elec) *e_elec = ke * sum_elec;
  if (e_vdw) *e_vdw = sum_vdw;
  return (ke * sum_elec) + sum_vdw;
}

/* src/mol.ml:881-903 ene_intra_UFFNB_brute; interacting = dists[i + j*n] >= 3
 * (create_interacting_a src/mol.ml:203-208, get_topo_dist src/mol.ml:151-152) */
double orc_ene_intra(const double* xs, const double* ys, const double* zs,
                     const double* q, const int32_t* anum, const int32_t* dists, int n) {
  double sum_elec = 0.0, sum_vdw = 0.0;
  for (int i = 0; i < n - 1; ++i) {
    for (int j = i + 1; j < n; ++j) {
      if (dists[i + j * n] >= 3) {
        const double dx = xs[i] - xs[j], dy = ys[i] - ys[j], dz = zs[i] - zs[j];
        const double r = orc_non_zero_dist(sqrt(dx * dx + dy * dy + dz * dz));
        xd_t v; pair_params(anum[i], anum[j], &v);
        const double p6 = orc_pow6(v.x / r);
        sum_elec = sum_elec + (q[i] * q[j]) / r;
        sum_vdw = sum_vdw + v.d * ((-2.0 * p6) + (p6 * p6));
      }
    }
  }
  return (orc_elec_weight() * sum_elec) + sum_vdw;
}

/* ------------------------------------------------------------------------ */
/* src/G3D.ml:97-157 trilin; grid xs[i] = i*step (src/grid.ml:49-51) */
static int cell_of(const orc_grid_t* g, double px, double py, double pz,
                   int* i0, int* j0, int* k0) {
  /* int_of_float truncates toward zero, src/G3D.ml:101-103 */
  *i0 = (int)(px * g->one_div_step);
  *j0 = (int)(py * g->one_div_step);
  *k0 = (int)(pz * g->one_div_step);
  return (px >= 0.0 && py >= 0.0 && pz >= 0.0 &&
          *i0 >= 0 && *j0 >= 0 && *k0 >= 0 &&
          *i0 + 1 < g->x_dim && *j0 + 1 < g->y_dim && *k0 + 1 < g->z_dim);
}

double orc_trilin(const orc_grid_t* g, const float* arr, double px, double py, double pz) {
  int i0, j0, k0;
  cell_of(g, px, py, pz, &i0, &j0, &k0);
  const int i1 = i0 + 1, j1 = j0 + 1, k1 = k0 + 1;
  const long xy = (long)g->x_dim * g->y_dim;
  const long j0x = (long)j0 * g->x_dim, j1x = (long)j1 * g->x_dim;
  const long k0xy = (long)k0 * xy, k1xy = (long)k1 * xy;
  const double lx = i0 * g->step, ly = j0 * g->step, lz = k0 * g->step;
  const double wlx = (px - lx) * g->one_div_step, wly = (py - ly) * g->one_div_step,
               wlz = (pz - lz) * g->one_div_step;
  const double whx = 1.0 - wlx, why = 1.0 - wly, whz = 1.0 - wlz;
  /* corner order a,b,c,d,e,f,g,h and left-to-right sum, src/G3D.ml:144-151 */
  return (double)arr[i0 + j0x + k0xy] * (whx * why * whz) +
         (double)arr[i1 + j0x + k0xy] * (wlx * why * whz) +
         (double)arr[i1 + j1x + k0xy] * (wlx * wly * whz) +
         (double)arr[i0 + j1x + k0xy] * (whx * wly * whz) +
         (double)arr[i0 + j0x + k1xy] * (whx * why * wlz) +
         (double)arr[i1 + j0x + k1xy] * (wlx * why * wlz) +
         (double)arr[i1 + j1x + k1xy] * (wlx * wly * wlz) +
         (double)arr[i0 + j1x + k1xy] * (whx * wly * wlz);
}

/* src/mol.ml:1012-1020 */
double orc_ene_inter_interp(const orc_grid_t* g, const float* const* grids,
                            const double* lxs, const double* lys, const double* lzs,
                            const int32_t* ltype, int nl) {
  double res = 0.0;
  for (int j = 0; j < nl; ++j) {
    int i0, j0, k0;
    if (!cell_of(g, lxs[j], lys[j], lzs[j], &i0, &j0, &k0)) return NAN;
    res = res + orc_trilin(g, grids[ltype[j]], lxs[j], lys[j], lzs[j]);
  }
  return res;
}

/* src/mol.ml:964-989 + src/lds.ml:452-469 */
void orc_grid_energies(const double* pxs, const double* pys, const double* pzs,
                       const double* pq, const int32_t* panum, int np,
                       const orc_grid_t* g, const uint8_t* mask,
                       const int32_t* probe_anum, const double* probe_q, int ntypes,
                       float* const* out) {
  const long nx = g->x_dim, ny = g->y_dim, nz = g->z_dim;
  const double ke = orc_elec_weight();
  #pragma omp parallel
  {
    double* se = (double*)malloc(sizeof(double) * ntypes);
    double* sv = (double*)malloc(sizeof(double) * ntypes);
    #pragma omp for schedule(dynamic, 64)
    for (long idx = 0; idx < nx * ny * nz; ++idx) {
      if (mask && !mask[idx]) continue;
      /* Grid.ijk_of_idx, src/grid.ml:101-105 */
      const long k = idx / (nx * ny), j = (idx - k * nx * ny) / nx, i = idx - (k * nx * ny + j * nx);
      const double x = i * g->step, y = j * g->step, z = k * g->step;
      for (int t = 0; t < ntypes; ++t) { se[t] = 0.0; sv[t] = 0.0; }
      for (int a = 0; a < np; ++a) {
        const double dx = pxs[a] - x, dy = pys[a] - y, dz = pzs[a] - z;
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (!(d2 < CUTOFF2)) continue;   /* BST.neighbors .. 12.0; S(r>=12) = 0 */
        const double r = orc_non_zero_dist(sqrt(d2));
        const double w = orc_shift_12A(r);
        for (int t = 0; t < ntypes; ++t) {
          xd_t v; pair_params(probe_anum[t], panum[a], &v);
          const double p6 = orc_pow6(v.x / r);
          se[t] = se[t] + w * ((pq[a] * probe_q[t]) / r);
          sv[t] = sv[t] + w * (v.d * ((-2.0 * p6) + (p6 * p6)));
        }
      }
      for (int t = 0; t < ntypes; ++t) {
        const double e = ke * se[t] + sv[t];
        out[t][idx] = (float)(e < MAX_E ? e : MAX_E);   /* min Params.max_E, src/lds.ml:466 */
      }
    }
    free(se); free(sv);
  }
}

/* src/lds.ml:269-305 */
void orc_bitmask_roi(const orc_grid_t* g, double cx, double cy, double cz, double out_r,
                     uint8_t* mask) {
  const double r = out_r + 12.0 * 2.0;
  const double r2 = r * r;
  for (long k = 0; k < g->z_dim; ++k)
    for (long j = 0; j < g->y_dim; ++j)
      for (long i = 0; i < g->x_dim; ++i) {
        const double dx = cx - i * g->step, dy = cy - j * g->step, dz = cz - k * g->step;
        mask[i + j * g->x_dim + k * (long)g->x_dim * g->y_dim] = (dx * dx + dy * dy + dz * dz) < r2;
      }
}

/* ------------------------------------------------------------------------ */
/* rotations, src/rot.ml */
void orc_rot_x(double t, double* m) {
  const double c = cos(t), s = sin(t);
  double r[9] = {1.0, 0.0, 0.0, 0.0, c, s, 0.0, -s, c};
  memcpy(m, r, sizeof r);
}
void orc_rot_y(double t, double* m) {
  const double c = cos(t), s = sin(t);
  double r[9] = {c, 0.0, -s, 0.0, 1.0, 0.0, s, 0.0, c};
  memcpy(m, r, sizeof r);
}
void orc_rot_z(double t, double* m) {
  const double c = cos(t), s = sin(t);
  double r[9] = {c, s, 0.0, -s, c, 0.0, 0.0, 0.0, 1.0};
  memcpy(m, r, sizeof r);
}
void orc_rot_xyz(double a_, double b_, double g_, double* m) {
  const double ac = cos(a_), as = sin(a_), bc = cos(b_), bs = sin(b_), gc = cos(g_), gs = sin(g_);
  m[0] = bc * gc;
  m[1] = gc * as * bs - ac * gs;
  m[2] = as * gs + ac * gc * bs;
  m[3] = bc * gs;
  m[4] = ac * gc + as * bs * gs;
  m[5] = ac * bs * gs - gc * as;
  m[6] = -bs;
  m[7] = bc * as;
  m[8] = ac * bc;
}
void orc_rot_mult(const double* r1, const double* r2, double* m) {
  double t[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      t[3 * r + c] = r1[3 * r] * r2[c] + r1[3 * r + 1] * r2[3 + c] + r1[3 * r + 2] * r2[6 + c];
  memcpy(m, t, sizeof t);
}
void orc_rot_of_quat(double w, double x, double y, double z, double* m) {
  const double px = x * x, py = y * y, pz = z * z;
  m[0] = 1. - 2. * (py + pz);
  m[1] = 2. * (x * y - z * w);
  m[2] = 2. * (x * z + y * w);
  m[3] = 2. * (x * y + z * w);
  m[4] = 1. - 2. * (px + pz);
  m[5] = 2. * (y * z - x * w);
  m[6] = 2. * (x * z - y * w);
  m[7] = 2. * (y * z + x * w);
  m[8] = 1. - 2. * (px + py);
}
void orc_rot_of_axis_angle(double x, double y, double z, double theta, double* m) {
  const double c = cos(theta), s = sin(theta), omc = 1.0 - c;
  m[0] = c + x * x * omc;
  m[1] = x * y * omc - z * s;
  m[2] = x * z * omc + y * s;
  m[3] = x * y * omc + z * s;
  m[4] = c + y * y * omc;
  m[5] = y * z * omc - x * s;
  m[6] = x * z * omc - y * s;
  m[7] = y * z * omc + x * s;
  m[8] = c + z * z * omc;
}
/* src/quat.ml:32-36 */
void orc_quat_to_axis_angle(const double* q, double* axis, double* theta) {
  const double mag = sqrt(q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  axis[0] = q[1] / mag; axis[1] = q[2] / mag; axis[2] = q[3] / mag;
  *theta = 2.0 * atan2(mag, q[0]);
}
/* src/SO3.ml:18-28 (constants phi = sqrt 2, psi from p4 of Alexa CVPR 2022) */
void orc_so3_quat(int n, int i, double* q) {
  const double pi = 4.0 * atan(1.0), two_pi = 2.0 * pi;   /* src/math.ml:13-15 */
  const double phi = sqrt(2.0), psi = 1.533751168755204288118041;
  const double s = (double)i + 0.5, t = s / (double)n, d = two_pi * s;
  const double cr = sqrt(t), cR = sqrt(1.0 - t), al = d / phi, be = d / psi;
  q[0] = cr * sin(al); q[1] = cr * cos(al); q[2] = cR * sin(be); q[3] = cR * cos(be);
}
void orc_so3_rotations(int n, double* mats) {
  for (int i = 0; i < n; ++i) {
    double q[4], ax[3], th;
    orc_so3_quat(n, i, q);
    orc_quat_to_axis_angle(q, ax, &th);
    orc_rot_of_axis_angle(ax[0], ax[1], ax[2], th, mats + 9 * i);
  }
}
/* src/math.ml:46-52 */
double orc_dihedral(const double* i, const double* j, const double* k, const double* l) {
  const double u1[3] = {j[0] - i[0], j[1] - i[1], j[2] - i[2]};
  const double u2[3] = {k[0] - j[0], k[1] - j[1], k[2] - j[2]};
  const double u3[3] = {l[0] - k[0], l[1] - k[1], l[2] - k[2]};
  const double u23[3] = {u2[1] * u3[2] - u2[2] * u3[1], u2[2] * u3[0] - u2[0] * u3[2],
                         u2[0] * u3[1] - u2[1] * u3[0]};
  const double m2 = sqrt(u2[0] * u2[0] + u2[1] * u2[1] + u2[2] * u2[2]);
  const double u12[3] = {u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                         u1[0] * u2[1] - u1[1] * u2[0]};
  const double y = (u1[0] * m2) * u23[0] + (u1[1] * m2) * u23[1] + (u1[2] * m2) * u23[2];
  const double x = u12[0] * u23[0] + u12[1] * u23[1] + u12[2] * u23[2];
  return atan2(y, x);
}

/* ------------------------------------------------------------------------ */
static void rotate_v(const double* r, double x, double y, double z, double* o) {
  /* Rot.rotate, src/rot.ml:97-100 */
  o[0] = r[0] * x + r[1] * y + r[2] * z;
  o[1] = r[3] * x + r[4] * y + r[5] * z;
  o[2] = r[6] * x + r[7] * y + r[8] * z;
}

void orc_center(const double* xs, const double* ys, const double* zs, int n, double* c) {
  double sx = 0.0, sy = 0.0, sz = 0.0;   /* A.favg: sum then divide */
  for (int i = 0; i < n; ++i) { sx += xs[i]; sy += ys[i]; sz += zs[i]; }
  c[0] = sx / n; c[1] = sy / n; c[2] = sz / n;
}

double orc_radius(const double* xs, const double* ys, const double* zs, int n, const double* c) {
  double maxi = 0.0;   /* src/mol.ml:576-583 */
  for (int i = 0; i < n; ++i) {
    const double dx = c[0] - xs[i], dy = c[1] - ys[i], dz = c[2] - zs[i];
    const double r = 0.01 + sqrt(dx * dx + dy * dy + dz * dz);
    if (r > maxi) maxi = r;
  }
  return maxi;
}

/* src/mol.ml:610-631 */
void orc_rotate_bond(double* xs, double* ys, double* zs, int left, int right,
                     const int32_t* rgroup, int ngroup, double alpha) {
  const double px = xs[right], py = ys[right], pz = zs[right];
  double ax = px - xs[left], ay = py - ys[left], az = pz - zs[left];
  const double mag = sqrt(ax * ax + ay * ay + az * az);   /* V3.normalize */
  ax /= mag; ay /= mag; az /= mag;
  double R[9];
  orc_rot_of_axis_angle(ax, ay, az, alpha, R);
  for (int k = 0; k < ngroup; ++k) {
    const int i = rgroup[k];
    double o[3];
    rotate_v(R, xs[i] - px, ys[i] - py, zs[i] - pz, o);
    xs[i] = o[0] + px; ys[i] = o[1] + py; zs[i] = o[2] + pz;
  }
}

/* src/mol.ml:699-710 */
void orc_center_rotate_translate(const double* xs, const double* ys, const double* zs, int n,
                                 const double* center, const double* rot, const double* pos,
                                 double* oxs, double* oys, double* ozs) {
  for (int i = 0; i < n; ++i) {
    double o[3];
    rotate_v(rot, xs[i] - center[0], ys[i] - center[1], zs[i] - center[2], o);
    oxs[i] = o[0] + pos[0]; oys[i] = o[1] + pos[1]; ozs[i] = o[2] + pos[2];
  }
}

/* src/optim.ml:64-80 (and Mol.rotate_then_translate_copy src/mol.ml:669-672) */
double orc_apply_config(const double* cxs, const double* cys, const double* czs, int n,
                        const int32_t* rb_left, const int32_t* rb_right,
                        const int32_t* rg_off, const int32_t* rg_atoms, int nrb,
                        const double* dofs,
                        double* oxs, double* oys, double* ozs) {
  double* xs = (double*)malloc(sizeof(double) * 3 * n);
  double *ys = xs + n, *zs = xs + 2 * n;
  memcpy(xs, cxs, sizeof(double) * n); memcpy(ys, cys, sizeof(double) * n);
  memcpy(zs, czs, sizeof(double) * n);
  double c[3] = {0.0, 0.0, 0.0};   /* centered_lig's stored center is the origin */
  for (int k = 0; k < nrb; ++k) {
    orc_rotate_bond(xs, ys, zs, rb_left[k], rb_right[k], rg_atoms + rg_off[k],
                    rg_off[k + 1] - rg_off[k], dofs[6 + k]);
    orc_center(xs, ys, zs, n, c);   /* update_center */
  }
  const double rad = orc_radius(xs, ys, zs, n, c);
  double R[9];
  orc_rot_xyz(dofs[3], dofs[4], dofs[5], R);
  for (int i = 0; i < n; ++i) {
    double o[3];
    rotate_v(R, xs[i], ys[i], zs[i], o);
    oxs[i] = o[0] + dofs[0]; oys[i] = o[1] + dofs[1]; ozs[i] = o[2] + dofs[2];
  }
  free(xs);
  return rad;
}

/* Forward-mode (tangent) derivative of orc_apply_config: tan[(i*n + j)*3 + c] is
 * d out_j[c] / d dofs[i] for every DOF i (6 + nrb of them).  Each torsion
 * y = p + R(a, tau)(x - p), p = x_right, a = (x_right - x_left)/|.|, carries the tangents
 * dy = dp + R(dx - dp) + dtau dR/dtau w + s [da]x w + (1 - c)(da (a.w) + a (da.w)),
 * w = x - p, da = (I - a a^T) dv / |v|, dR/dtau w = -s w + c a x w + s a (a.w)
 * (Rodrigues: R = c I + s [a]x + (1 - c) a a^T); the rigid part r_xyz = Rz(g) Ry(b) Rx(a)
 * (src/rot.ml:48-66) is differentiated factor by factor.  This is this build's exact
 * DOF gradient for the CPU checks (the reference differentiates numerically,
 * src/lds.ml:1118-1136); tests pin it against central differences. */
static void cross3(const double* a, const double* b, double* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static void mat3_mult(const double* A, const double* B, double* C) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c)
      C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}
/* active axis rotations and their derivatives */
static void act_rx(double t, double* m, double* d) {
  const double c = cos(t), s = sin(t);
  const double M[9] = {1, 0, 0, 0, c, -s, 0, s, c}, D[9] = {0, 0, 0, 0, -s, -c, 0, c, -s};
  memcpy(m, M, sizeof M); memcpy(d, D, sizeof D);
}
static void act_ry(double t, double* m, double* d) {
  const double c = cos(t), s = sin(t);
  const double M[9] = {c, 0, s, 0, 1, 0, -s, 0, c}, D[9] = {-s, 0, c, 0, 0, 0, -c, 0, -s};
  memcpy(m, M, sizeof M); memcpy(d, D, sizeof D);
}
static void act_rz(double t, double* m, double* d) {
  const double c = cos(t), s = sin(t);
  const double M[9] = {c, -s, 0, s, c, 0, 0, 0, 1}, D[9] = {-s, -c, 0, c, -s, 0, 0, 0, 0};
  memcpy(m, M, sizeof M); memcpy(d, D, sizeof D);
}
double orc_apply_config_jvp(const double* cxs, const double* cys, const double* czs, int n,
                            const int32_t* rb_left, const int32_t* rb_right,
                            const int32_t* rg_off, const int32_t* rg_atoms, int nrb,
                            const double* dofs,
                            double* oxs, double* oys, double* ozs, double* tan) {
  const int nd = 6 + nrb;
  double* x = (double*)malloc(sizeof(double) * 3 * n);
  double* dx = (double*)calloc((size_t)nd * n * 3, sizeof(double));
  double* dp = (double*)malloc(sizeof(double) * 3 * nd);
  double* da = (double*)malloc(sizeof(double) * 3 * nd);
  for (int j = 0; j < n; ++j) { x[3 * j] = cxs[j]; x[3 * j + 1] = cys[j]; x[3 * j + 2] = czs[j]; }
  double c[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < nrb; ++k) {
    const int L = rb_left[k], Rr = rb_right[k];
    const double* p = x + 3 * Rr;
    const double v[3] = {p[0] - x[3 * L], p[1] - x[3 * L + 1], p[2] - x[3 * L + 2]};
    const double nv = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double a[3] = {v[0] / nv, v[1] / nv, v[2] / nv};
    const double tau = dofs[6 + k], cs = cos(tau), sn = sin(tau);
    double R[9];
    orc_rot_of_axis_angle(a[0], a[1], a[2], tau, R);
    for (int i = 0; i < nd; ++i) {
      const double* dxi = dx + (size_t)i * n * 3;
      double dv[3];
      for (int q = 0; q < 3; ++q) {
        dp[3 * i + q] = dxi[3 * Rr + q];
        dv[q] = dxi[3 * Rr + q] - dxi[3 * L + q];
      }
      const double adv = a[0] * dv[0] + a[1] * dv[1] + a[2] * dv[2];
      for (int q = 0; q < 3; ++q) da[3 * i + q] = (dv[q] - a[q] * adv) / nv;
    }
    const double px = p[0], py = p[1], pz = p[2];
    for (int m = rg_off[k]; m < rg_off[k + 1]; ++m) {
      const int j = rg_atoms[m];
      const double w[3] = {x[3 * j] - px, x[3 * j + 1] - py, x[3 * j + 2] - pz};
      const double aw = a[0] * w[0] + a[1] * w[1] + a[2] * w[2];
      double axw[3];
      cross3(a, w, axw);
      for (int i = 0; i < nd; ++i) {
        double* dxi = dx + (size_t)i * n * 3;
        const double* dpi = dp + 3 * i;
        const double* dai = da + 3 * i;
        const double dw[3] = {dxi[3 * j] - dpi[0], dxi[3 * j + 1] - dpi[1], dxi[3 * j + 2] - dpi[2]};
        double Rdw[3], daxw[3];
        rotate_v(R, dw[0], dw[1], dw[2], Rdw);
        cross3(dai, w, daxw);
        const double daw = dai[0] * w[0] + dai[1] * w[1] + dai[2] * w[2];
        const double dt = (i == 6 + k) ? 1.0 : 0.0;
        for (int q = 0; q < 3; ++q)
          dxi[3 * j + q] = dpi[q] + Rdw[q] + dt * (-sn * w[q] + cs * axw[q] + sn * a[q] * aw) +
                           sn * daxw[q] + (1.0 - cs) * (dai[q] * aw + a[q] * daw);
      }
      double y[3];
      rotate_v(R, w[0], w[1], w[2], y);
      x[3 * j] = y[0] + px; x[3 * j + 1] = y[1] + py; x[3 * j + 2] = y[2] + pz;
    }
  }
  if (nrb > 0) {
    c[0] = c[1] = c[2] = 0.0;
    for (int j = 0; j < n; ++j) { c[0] += x[3 * j]; c[1] += x[3 * j + 1]; c[2] += x[3 * j + 2]; }
    c[0] /= n; c[1] /= n; c[2] /= n;
  }
  double rad = 0.0;
  for (int j = 0; j < n; ++j) {
    const double ex = c[0] - x[3 * j], ey = c[1] - x[3 * j + 1], ez = c[2] - x[3 * j + 2];
    const double r = 0.01 + sqrt(ex * ex + ey * ey + ez * ez);
    if (r > rad) rad = r;
  }
  /* r_xyz = Rz(g) Ry(b) Rx(a) and its three partial derivatives */
  double X[9], dX[9], Y[9], dY[9], Z[9], dZ[9], ZY[9], M[9], D[3][9], tmp[9];
  act_rx(dofs[3], X, dX); act_ry(dofs[4], Y, dY); act_rz(dofs[5], Z, dZ);
  mat3_mult(Z, Y, ZY); mat3_mult(ZY, X, M);
  mat3_mult(ZY, dX, D[0]);
  mat3_mult(Z, dY, tmp); mat3_mult(tmp, X, D[1]);
  mat3_mult(dZ, Y, tmp); mat3_mult(tmp, X, D[2]);
  for (int j = 0; j < n; ++j) {
    double o[3];
    rotate_v(M, x[3 * j], x[3 * j + 1], x[3 * j + 2], o);
    oxs[j] = o[0] + dofs[0]; oys[j] = o[1] + dofs[1]; ozs[j] = o[2] + dofs[2];
    for (int i = 0; i < nd; ++i) {
      double* ti = tan + ((size_t)i * n + j) * 3;
      if (i < 3) {
        ti[0] = ti[1] = ti[2] = 0.0;
        ti[i] = 1.0;
      } else if (i < 6) {
        rotate_v(D[i - 3], x[3 * j], x[3 * j + 1], x[3 * j + 2], ti);
      } else {
        const double* dxi = dx + ((size_t)i * n + j) * 3;
        rotate_v(M, dxi[0], dxi[1], dxi[2], ti);
      }
    }
  }
  free(x); free(dx); free(dp); free(da);
  return rad;
}

/* ------------------------------------------------------------------------ */
/* analytic gradients.  Pair energy with d = x_lig - x_prot, r = max(|d|, 0.01):
 *   E = w(r) [ke qi qj / r + D (p6^2 - 2 p6)],  p6 = (x/r)^6,
 *   dE/dr = w' F + w F',  F' = -ke qi qj / r^2 - 12 D (p6^2 - p6) / r,
 *   w = (1 - r^2/144)^2 (shifted) -> w' = -r (1 - r^2/144) / 36;  w = 1 (global).
 *   clamped pairs (|d| < 0.01) have zero gradient. */
static double pair_grad(int shifted, double ke, double qq, const xd_t* v, double dx, double dy,
                        double dz, double* g) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  g[0] = g[1] = g[2] = 0.0;
  if (shifted && !(r2 < CUTOFF2)) return 0.0;
  const double rr = sqrt(r2);
  const double r = orc_non_zero_dist(rr);
  const double p6 = orc_pow6(v->x / r);
  const double F = ke * qq / r + v->d * (p6 * p6 - 2.0 * p6);
  const double Fp = -ke * qq / (r * r) - 12.0 * v->d * (p6 * p6 - p6) / r;
  double w = 1.0, wp = 0.0;
  if (shifted) {
    const double t = 1.0 - r * r / 144.0;
    w = t * t;
    wp = -r * t / 36.0;
  }
  if (rr >= 0.01) {
    const double dEdr = wp * F + w * Fp;
    g[0] = dEdr * dx / rr; g[1] = dEdr * dy / rr; g[2] = dEdr * dz / rr;
  }
  return w * F;
}

double orc_grad_inter(int shifted,
                      const double* pxs, const double* pys, const double* pzs,
                      const double* pq, const int32_t* panum, int np,
                      const double* lxs, const double* lys, const double* lzs,
                      const double* lq, const int32_t* lanum, int nl,
                      double* gx, double* gy, double* gz) {
  const double ke = orc_elec_weight();
  double e = 0.0;
  for (int j = 0; j < nl; ++j) {
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int i = 0; i < np; ++i) {
      xd_t v; pair_params(panum[i], lanum[j], &v);
      double g[3];
      e += pair_grad(shifted, ke, pq[i] * lq[j], &v, lxs[j] - pxs[i], lys[j] - pys[i],
                     lzs[j] - pzs[i], g);
      sx += g[0]; sy += g[1]; sz += g[2];
    }
    gx[j] = sx; gy[j] = sy; gz[j] = sz;
  }
  return e;
}

double orc_grad_intra(const double* xs, const double* ys, const double* zs,
                      const double* q, const int32_t* anum, const int32_t* dists, int n,
                      double* gx, double* gy, double* gz) {
  const double ke = orc_elec_weight();
  double e = 0.0;
  for (int i = 0; i < n; ++i) gx[i] = gy[i] = gz[i] = 0.0;
  for (int i = 0; i < n - 1; ++i)
    for (int j = i + 1; j < n; ++j) {
      if (dists[i + j * n] < 3) continue;
      xd_t v; pair_params(anum[i], anum[j], &v);
      double g[3];
      e += pair_grad(0, ke, q[i] * q[j], &v, xs[i] - xs[j], ys[i] - ys[j], zs[i] - zs[j], g);
      gx[i] += g[0]; gy[i] += g[1]; gz[i] += g[2];
      gx[j] -= g[0]; gy[j] -= g[1]; gz[j] -= g[2];
    }
  return e;
}

/* d trilin / d p inside the cell (src/G3D.ml:97-157 differentiated). */
double orc_grad_interp(const orc_grid_t* g, const float* const* grids,
                       const double* lxs, const double* lys, const double* lzs,
                       const int32_t* ltype, int nl,
                       double* gx, double* gy, double* gz) {
  double e = 0.0;
  const long xy = (long)g->x_dim * g->y_dim;
  for (int j = 0; j < nl; ++j) {
    int i0, j0, k0;
    if (!cell_of(g, lxs[j], lys[j], lzs[j], &i0, &j0, &k0)) {
      gx[j] = gy[j] = gz[j] = NAN; e = NAN; continue;
    }
    const float* a = grids[ltype[j]];
    const long b = i0 + (long)j0 * g->x_dim + (long)k0 * xy;
    const double A = a[b], B = a[b + 1], C = a[b + 1 + g->x_dim], D = a[b + g->x_dim];
    const double E = a[b + xy], F = a[b + 1 + xy], G = a[b + 1 + g->x_dim + xy],
                 H = a[b + g->x_dim + xy];
    const double o = g->one_div_step;
    const double wlx = (lxs[j] - i0 * g->step) * o, wly = (lys[j] - j0 * g->step) * o,
                 wlz = (lzs[j] - k0 * g->step) * o;
    const double whx = 1.0 - wlx, why = 1.0 - wly, whz = 1.0 - wlz;
    e += orc_trilin(g, a, lxs[j], lys[j], lzs[j]);
    gx[j] = o * ((B - A) * why * whz + (C - D) * wly * whz + (F - E) * why * wlz + (G - H) * wly * wlz);
    gy[j] = o * ((D - A) * whx * whz + (C - B) * wlx * whz + (H - E) * whx * wlz + (G - F) * wlx * wlz);
    gz[j] = o * ((E - A) * whx * why + (F - B) * wlx * why + (G - C) * wlx * wly + (H - D) * whx * wly);
  }
  return e;
}

/* ------------------------------------------------------------------------ */
/* rigid-scan batches for the CPU baseline (x = R c + t, src/lds.ml:1081-1096) */
void orc_scan_shifted(const double* pxs, const double* pys, const double* pzs,
                      const double* pq, const int32_t* panum, int np,
                      const double* cxs, const double* cys, const double* czs,
                      const double* lq, const int32_t* lanum, int nl,
                      const double* rots, const int32_t* pose_rot,
                      const double* trans, const int32_t* pose_trans, int npose,
                      int nthreads, double* e_out) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  #pragma omp parallel num_threads(nthreads)
#endif
  {
    double* buf = (double*)malloc(sizeof(double) * 3 * nl);
#ifdef _OPENMP
    #pragma omp for schedule(dynamic, 4)
#endif
    for (int p = 0; p < npose; ++p) {
      const double* R = rots + 9 * pose_rot[p];
      const double* t = trans + 3 * pose_trans[p];
      for (int j = 0; j < nl; ++j) {
        double o[3];
        rotate_v(R, cxs[j], cys[j], czs[j], o);
        buf[j] = o[0] + t[0]; buf[nl + j] = o[1] + t[1]; buf[2 * nl + j] = o[2] + t[2];
      }
      e_out[p] = orc_ene_inter_shifted(pxs, pys, pzs, pq, panum, np, buf, buf + nl, buf + 2 * nl,
                                       lq, lanum, nl, NULL, NULL);
    }
    free(buf);
  }
}

/* The reference's default form, -ff Bst (src/lds.ml:1790; ene_inter_UFF_shifted_bst_components,
 * src/mol.ml:928-956): per ligand atom, only the receptor atoms that a neighbour query
 * returns within 12 A are visited (the reference queries a vantage-point tree of the
 * third-party `bst` library).  Restated with a grid of 4 A cells over the receptor: for
 * each (y, z) row of cells that meets the atom's 12 A sphere, the cells the sphere's x
 * chord covers.  Same FP64 terms as orc_ene_inter_shifted, summed ligand atom by ligand
 * atom; the energies agree with orc_scan_shifted to FP64 rounding.  The CPU timing
 * baseline of the scan. */
void orc_scan_shifted_cells(const double* pxs, const double* pys, const double* pzs,
                            const double* pq, const int32_t* panum, int np,
                            const double* cxs, const double* cys, const double* czs,
                            const double* lq, const int32_t* lanum, int nl,
                            const double* rots, const int32_t* pose_rot,
                            const double* trans, const int32_t* pose_trans, int npose,
                            int nthreads, double* e_out) {
  const double cs = 4.0;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < np; ++i) {
    const double v[3] = {pxs[i], pys[i], pzs[i]};
    for (int a = 0; a < 3; ++a) { if (v[a] < lo[a]) lo[a] = v[a]; if (v[a] > hi[a]) hi[a] = v[a]; }
  }
  int n[3];
  for (int a = 0; a < 3; ++a) n[a] = (int)floor((hi[a] - lo[a]) / cs) + 1;
  const long ncell = (long)n[0] * n[1] * n[2];
  int* start = (int*)calloc((size_t)ncell + 1, sizeof(int));
  int* cell = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  int* order = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
  for (int i = 0; i < np; ++i) {
    int c[3];
    const double v[3] = {pxs[i], pys[i], pzs[i]};
    for (int a = 0; a < 3; ++a) {
      c[a] = (int)floor((v[a] - lo[a]) / cs);
      if (c[a] >= n[a]) c[a] = n[a] - 1;
    }
    cell[i] = (c[2] * n[1] + c[1]) * n[0] + c[0];
    start[cell[i] + 1]++;
  }
  for (long k = 0; k < ncell; ++k) start[k + 1] += start[k];
  {
    int* fill = (int*)malloc(sizeof(int) * (size_t)(ncell > 0 ? ncell : 1));
    for (long k = 0; k < ncell; ++k) fill[k] = start[k];
    for (int i = 0; i < np; ++i) order[fill[cell[i]]++] = i;
    free(fill);
  }
  /* the receptor in cell order */
  double* sx = (double*)malloc(sizeof(double) * 4 * (size_t)(np > 0 ? np : 1));
  double *sy = sx + np, *sz = sy + np, *sq = sz + np;
  int32_t* sa = (int32_t*)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
  for (int k = 0; k < np; ++k) {
    const int i = order[k];
    sx[k] = pxs[i]; sy[k] = pys[i]; sz[k] = pzs[i]; sq[k] = pq[i]; sa[k] = panum[i];
  }
  const double ke = orc_elec_weight();
  /* pair parameters looked up, as the reference's 119 x 119 table (UFF.vdW_xiDi,
   * src/UFF.ml:42-51), for the element pairs present */
  double* tx = (double*)calloc(119 * 119, sizeof(double));
  double* td = (double*)calloc(119 * 119, sizeof(double));
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < nl; ++j) {
      const int a1 = sa[k], a2 = lanum[j];
      if (a1 < 0 || a1 > 118 || a2 < 0 || a2 > 118) continue;
      if (tx[a1 * 119 + a2] == 0.0) {
        xd_t v; pair_params(a1, a2, &v);
        tx[a1 * 119 + a2] = v.x; td[a1 * 119 + a2] = v.d;
      }
    }
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  #pragma omp parallel num_threads(nthreads)
#endif
  {
#ifdef _OPENMP
    #pragma omp for schedule(dynamic, 4)
#endif
    for (int p = 0; p < npose; ++p) {
      const double* R = rots + 9 * pose_rot[p];
      const double* t = trans + 3 * pose_trans[p];
      double sum_elec = 0.0, sum_vdw = 0.0;
      for (int j = 0; j < nl; ++j) {
        double o[3];
        rotate_v(R, cxs[j], cys[j], czs[j], o);
        const double x = o[0] + t[0], y = o[1] + t[1], z = o[2] + t[2];
        const int z0 = (int)floor((z - 12.0 - lo[2]) / cs), z1 = (int)floor((z + 12.0 - lo[2]) / cs);
        const int y0 = (int)floor((y - 12.0 - lo[1]) / cs), y1 = (int)floor((y + 12.0 - lo[1]) / cs);
        for (int cz = (z0 < 0 ? 0 : z0); cz <= z1 && cz < n[2]; ++cz) {
          const double zl = lo[2] + cz * cs;
          const double dz = z < zl ? zl - z : (z > zl + cs ? z - (zl + cs) : 0.0);
          for (int cy = (y0 < 0 ? 0 : y0); cy <= y1 && cy < n[1]; ++cy) {
            const double yl = lo[1] + cy * cs;
            const double dy = y < yl ? yl - y : (y > yl + cs ? y - (yl + cs) : 0.0);
            const double rx2 = 144.0 - dz * dz - dy * dy;
            if (rx2 < 0.0) continue;
            const double rx = sqrt(rx2);
            int x0 = (int)floor((x - rx - lo[0]) / cs), x1 = (int)floor((x + rx - lo[0]) / cs);
            if (x0 < 0) x0 = 0;
            if (x1 >= n[0]) x1 = n[0] - 1;
            if (x0 > x1) continue;
            const long rowb = ((long)cz * n[1] + cy) * n[0];
            for (int k = start[rowb + x0]; k < start[rowb + x1 + 1]; ++k) {
              const double ddx = sx[k] - x, ddy = sy[k] - y, ddz = sz[k] - z;
              const double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
              if (!(r2 < CUTOFF2)) continue;
              const double r = orc_non_zero_dist(sqrt(r2));
              const double w = orc_shift_12A(r);
              const int pi = sa[k] * 119 + lanum[j];
              const double p6 = orc_pow6(tx[pi] / r);
              sum_elec = sum_elec + w * ((sq[k] * lq[j]) / r);
              sum_vdw = sum_vdw + w * (td[pi] * ((-2.0 * p6) + (p6 * p6)));
            }
          }
        }
      }
      e_out[p] = (ke * sum_elec) + sum_vdw;
    }
  }
  free(start); free(cell); free(order); free(sx); free(sa); free(tx); free(td);
}

void orc_scan_interp(const orc_grid_t* g, const float* const* grids,
                     const double* cxs, const double* cys, const double* czs,
                     const int32_t* ltype, int nl,
                     const double* rots, const int32_t* pose_rot,
                     const double* trans, const int32_t* pose_trans, int npose,
                     int nthreads, double* e_out) {
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
  #pragma omp parallel num_threads(nthreads)
#endif
  {
    double* buf = (double*)malloc(sizeof(double) * 3 * nl);
#ifdef _OPENMP
    #pragma omp for schedule(static)
#endif
    for (int p = 0; p < npose; ++p) {
      const double* R = rots + 9 * pose_rot[p];
      const double* t = trans + 3 * pose_trans[p];
      for (int j = 0; j < nl; ++j) {
        double o[3];
        rotate_v(R, cxs[j], cys[j], czs[j], o);
        buf[j] = o[0] + t[0]; buf[nl + j] = o[1] + t[1]; buf[2 * nl + j] = o[2] + t[2];
      }
      e_out[p] = orc_ene_inter_interp(g, grids, buf, buf + nl, buf + 2 * nl, ltype, nl);
    }
    free(buf);
  }
}

/* ------------------------------------------------------------------ desolvation */
/* Const.desolvation, src/const.ml:31 (eps_prot 4, eps_H2O 78.5, Const.r_H2O 1.4) */
double orc_desolvation_const(void) {
  const double pi = 4.0 * atan(1.0);
  return (1.0 / 4.0 - 1.0 / 78.5) / (8.0 * pi);
}

/* Lds.atom_bitmask_set, src/lds.ml:145-169: the nodes of the cube of half-side
 * ceil(r / step) cells around Grid.coord_of_point (truncation, src/grid.ml:87-90) whose
 * distance to p is below r; indices outside the grid are skipped (the reference reads
 * past its node arrays there) */
static void atom_bitmask_set(const orc_grid_t* g, uint8_t* mask, double px, double py, double pz,
                             double r, uint8_t b) {
  const long nx = g->x_dim, ny = g->y_dim, nz = g->z_dim;
  const long ci = (long)(px / g->step), cj = (long)(py / g->step), ck = (long)(pz / g->step);
  const long rs = (long)ceil(r / g->step);
  const double r2 = r * r;
  for (long i = ci - rs; i <= ci + rs; ++i) {
    if (i < 0 || i >= nx) continue;
    for (long j = cj - rs; j <= cj + rs; ++j) {
      if (j < 0 || j >= ny) continue;
      for (long k = ck - rs; k <= ck + rs; ++k) {
        if (k < 0 || k >= nz) continue;
        const double dx = px - i * g->step, dy = py - j * g->step, dz = pz - k * g->step;
        if (dx * dx + dy * dy + dz * dz < r2) mask[i + j * nx + k * nx * ny] = b;
      }
    }
  }
}

/* Lds.first_solvent_shell, src/lds.ml:171-184: set within r_vdW + r_H2O of every atom,
 * then unset within r_vdW of every atom */
void orc_first_solvent_shell(const double* xs, const double* ys, const double* zs,
                             const double* radii, int n, const orc_grid_t* g, uint8_t* mask) {
  memset(mask, 0, (size_t)g->x_dim * g->y_dim * g->z_dim);
  for (int a = 0; a < n; ++a) atom_bitmask_set(g, mask, xs[a], ys[a], zs[a], radii[a] + 1.4, 1);
  for (int a = 0; a < n; ++a) atom_bitmask_set(g, mask, xs[a], ys[a], zs[a], radii[a], 0);
}

/* Lds.protein_desolv, src/lds.ml:201-234: for every protein solvent-shell node inside the
 * ROI sphere, desolvation x voxel volume x sum over protein atoms within 12 A (BST
 * neighbours) of (q_j / d^2)^2; 0 elsewhere */
void orc_protein_desolv(const double* pxs, const double* pys, const double* pzs,
                        const double* pq, int np, const orc_grid_t* g, const uint8_t* shell,
                        double cx, double cy, double cz, double roi_r, double* out) {
  const long nx = g->x_dim, ny = g->y_dim, nz = g->z_dim;
  const double vol = g->step * g->step * g->step, kd = orc_desolvation_const();
  for (long idx = 0; idx < nx * ny * nz; ++idx) {
    out[idx] = 0.0;
    if (!shell[idx]) continue;
    const long k = idx / (nx * ny), j = (idx - k * nx * ny) / nx, i = idx - (k * nx * ny + j * nx);
    const double x = i * g->step, y = j * g->step, z = k * g->step;
    const double rx = cx - x, ry = cy - y, rz = cz - z;
    if (!(rx * rx + ry * ry + rz * rz < roi_r * roi_r)) continue;   /* ROI.is_inside */
    double s = 0.0;
    for (int a = 0; a < np; ++a) {
      const double dx = x - pxs[a], dy = y - pys[a], dz = z - pzs[a];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (!(d2 < CUTOFF2)) continue;
      const double v = pq[a] / d2;
      s += v * v;
    }
    out[idx] = kd * (vol * s);
  }
}

/* Lds.desolvation_penalty, src/lds.ml:236-267: over the nodes in both the protein's and
 * the ligand's first solvent shells, the protein contributions (prot) and desolvation x
 * voxel volume x the ligand atoms' (q_j / d^2)^2 within 12 A (lig) */
void orc_desolvation_penalty(const orc_grid_t* g, const double* prot_contribs,
                             const uint8_t* prot_shell, const double* xs, const double* ys,
                             const double* zs, const double* q, const double* radii, int n,
                             double* prot, double* lig) {
  const long nx = g->x_dim, ny = g->y_dim, nz = g->z_dim, nv = nx * ny * nz;
  uint8_t* ls = (uint8_t*)malloc((size_t)nv);
  orc_first_solvent_shell(xs, ys, zs, radii, n, g, ls);
  double sp = 0.0, sl = 0.0;
  for (long idx = 0; idx < nv; ++idx) {
    if (!(ls[idx] && prot_shell[idx])) continue;
    const long k = idx / (nx * ny), j = (idx - k * nx * ny) / nx, i = idx - (k * nx * ny + j * nx);
    const double x = i * g->step, y = j * g->step, z = k * g->step;
    for (int a = 0; a < n; ++a) {
      const double dx = x - xs[a], dy = y - ys[a], dz = z - zs[a];
      const double d2 = dx * dx + dy * dy + dz * dz;
      if (d2 < CUTOFF2) {
        const double v = q[a] / d2;
        sl += v * v;
      }
    }
    sp += prot_contribs[idx];
  }
  free(ls);
  *prot = sp;
  *lig = orc_desolvation_const() * (g->step * g->step * g->step * sl);
}
