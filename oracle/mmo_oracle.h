/*
 * mmo_oracle.h -- FP64 CPU restatement of MMO's UFF nonbonded scoring path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded only
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
 * product path (mmo_amd/) never links, imports or calls it.
 *
 * Every function follows one reference function (cited as src/<file>:<line> of
 * UnixJunkie/MMO) in FP64, with the same loop order and the same clamps.  The
 * reference is OCaml and cannot be built in this image, so pinning is done by
 * closed forms, invariants and brute-force checks in tests/test_oracle*.py.
 *
 * Conventions (mirroring Mol.t, src/mol.ml:17-35):
 *   molecules are SoA: xs, ys, zs, q (charges), anum (atomic numbers), n atoms;
 *   topological distance matrix is column-major  dists[i + j*n]  (src/mol.ml:151-152);
 *   rotation matrices are row-major {a,b,c,d,e,f,g,h,i} (src/rot.ml:5-7).
 */
#ifndef MMO_ORACLE_H
#define MMO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* constants -------------------------------------------------------------- */
double orc_elec_weight(void);          /* 332.0637 / 4.0          src/UFF.ml:25, src/const.ml:18 */
int    orc_uff_params(int anum, double* x, double* d);  /* src/UFF.ml:10-22; -1 if unsupported */
void   orc_uff_pair(int a1, int a2, double* x_ij, double* d_ij); /* src/UFF.ml:37-51 (NaN if unsupported) */

/* scalar helpers ---------------------------------------------------------- */
double orc_non_zero_dist(double r);    /* src/math.ml:58-62 */
double orc_shift_12A(double d);        /* src/FF.ml:17-20   */
double orc_pow6(double x);             /* src/FF.ml:11-12   */

/* energies ----------------------------------------------------------------- */
/* Mol.ene_inter_UFF_global_brute, src/mol.ml:796-818 (BrG) */
double orc_ene_inter_global(const double* pxs, const double* pys, const double* pzs,
                            const double* pq, const int32_t* panum, int np,
                            const double* lxs, const double* lys, const double* lzs,
                            const double* lq, const int32_t* lanum, int nl,
                            double* e_elec, double* e_vdw);
/* Mol.ene_inter_UFF_shifted_brute, src/mol.ml:822-849 (BrL; == Bst src/mol.ml:928-960) */
double orc_ene_inter_shifted(const double* pxs, const double* pys, const double* pzs,
                             const double* pq, const int32_t* panum, int np,
                             const double* lxs, const double* lys, const double* lzs,
                             const double* lq, const int32_t* lanum, int nl,
                             double* e_elec, double* e_vdw);
/* Mol.ene_intra_UFFNB_brute, src/mol.ml:881-903, mask src/mol.ml:203-208 */
double orc_ene_intra(const double* xs, const double* ys, const double* zs,
                     const double* q, const int32_t* anum, const int32_t* dists, int n);

/* grid ---------------------------------------------------------------------- */
/* Grid.t with origin (0,0,0) (src/grid.ml:9-19); xs[i] = i*step. */
typedef struct {
  double step, one_div_step;
  int x_dim, y_dim, z_dim;
} orc_grid_t;
/* G3D.trilin, src/G3D.ml:97-157 (no bounds check in the reference; here the
 * caller must guarantee 0 <= i0 < x_dim-1 etc.) */
double orc_trilin(const orc_grid_t* g, const float* arr, double px, double py, double pz);
/* Mol.ene_inter_UFF_interp, src/mol.ml:1012-1020; grids[t] is the grid of type t.
 * Returns NaN if an atom's cell is outside the grid (reference: undefined). */
double orc_ene_inter_interp(const orc_grid_t* g, const float* const* grids,
                            const double* lxs, const double* lys, const double* lzs,
                            const int32_t* ltype, int nl);
/* Grid-point energies for single-atom probes (anum, q), clamped at max_E:
 * Mol.ene_inter_UFF_shifted_grid src/mol.ml:964-989 called from
 * pre_calculate_FF_components_grid src/lds.ml:452-469 (min Params.max_E ...).
 * Only points with mask[idx] != 0 are computed (others stay 0, src/G3D.ml:47-51);
 * mask may be NULL (= all points).  out[t][idx], idx = i + j*x_dim + k*x_dim*y_dim. */
void orc_grid_energies(const double* pxs, const double* pys, const double* pzs,
                       const double* pq, const int32_t* panum, int np,
                       const orc_grid_t* g, const uint8_t* mask,
                       const int32_t* probe_anum, const double* probe_q, int ntypes,
                       float* const* out);
/* bitmask_ROI_only, src/lds.ml:269-305: point inside sphere(c, out_r + 24). */
void orc_bitmask_roi(const orc_grid_t* g, double cx, double cy, double cz, double out_r,
                     uint8_t* mask);

/* rotations (row-major 3x3) ---------------------------------------------------- */
void orc_rot_x(double t, double* m);                         /* src/rot.ml:22-28  */
void orc_rot_y(double t, double* m);                         /* src/rot.ml:31-37  */
void orc_rot_z(double t, double* m);                         /* src/rot.ml:40-46  */
void orc_rot_xyz(double a, double b, double g, double* m);   /* src/rot.ml:52-66  */
void orc_rot_mult(const double* r1, const double* r2, double* m); /* src/rot.ml:77-94 */
void orc_rot_of_quat(double w, double x, double y, double z, double* m); /* src/rot.ml:104-114 */
void orc_rot_of_axis_angle(double x, double y, double z, double theta, double* m); /* src/rot.ml:121-146 */
void orc_quat_to_axis_angle(const double* q, double* axis, double* theta); /* src/quat.ml:32-36 */
void orc_so3_quat(int n, int i, double* q);                  /* src/SO3.ml:18-28  */
void orc_so3_rotations(int n, double* mats);                 /* src/SO3.ml:35-39, n*9 doubles */
double orc_dihedral(const double* i, const double* j, const double* k, const double* l); /* src/math.ml:46-52 */

/* ligand pose construction (coordinates modified in place) ----------------------- */
/* Mol.rotate_bond, src/mol.ml:610-631; rgroup excludes `right` (src/pqrs.ml:80-87). */
void orc_rotate_bond(double* xs, double* ys, double* zs, int left, int right,
                     const int32_t* rgroup, int ngroup, double alpha);
/* mean of coordinates (Mol.update_center, src/mol.ml:353-356) */
void orc_center(const double* xs, const double* ys, const double* zs, int n, double* c);
/* Mol.radius, src/mol.ml:576-583 (about the given center) */
double orc_radius(const double* xs, const double* ys, const double* zs, int n, const double* c);
/* Mol.center_rotate_translate_copy, src/mol.ml:705-710: x' = R (x - center) + pos */
void orc_center_rotate_translate(const double* xs, const double* ys, const double* zs, int n,
                                 const double* center, const double* rot, const double* pos,
                                 double* oxs, double* oys, double* ozs);
/* Optim.apply_config, src/optim.ml:64-80: dofs = [x,y,z,a,b,g,tau_0..tau_{R-1}].
 * rb_left/rb_right[k] is rbond k; rgroup k is rg_atoms[rg_off[k]..rg_off[k+1]).
 * Returns the enclosing radius after the torsions (caller checks > 12 -> Too_long). */
double orc_apply_config(const double* cxs, const double* cys, const double* czs, int n,
                        const int32_t* rb_left, const int32_t* rb_right,
                        const int32_t* rg_off, const int32_t* rg_atoms, int nrb,
                        const double* dofs,
                        double* oxs, double* oys, double* ozs);

/* Forward-mode derivative of orc_apply_config (this build's own, for the CPU checks of
 * the DOF gradients): also writes tan[(i*n + j)*3 + c] = d out_j[c] / d dofs[i] for the
 * 6 + nrb DOFs. */
double orc_apply_config_jvp(const double* cxs, const double* cys, const double* czs, int n,
                            const int32_t* rb_left, const int32_t* rb_right,
                            const int32_t* rg_off, const int32_t* rg_atoms, int nrb,
                            const double* dofs,
                            double* oxs, double* oys, double* ozs, double* tan);

/* analytic FP64 gradients (dE/dxyz of ligand atoms) -------------------------------
 * Derived from the energy formulas above (not present in the reference, which only
 * has central finite differences, src/lds.ml:1118-1136); checked by central FD of
 * the energies in tests. */
double orc_grad_inter(int shifted,
                      const double* pxs, const double* pys, const double* pzs,
                      const double* pq, const int32_t* panum, int np,
                      const double* lxs, const double* lys, const double* lzs,
                      const double* lq, const int32_t* lanum, int nl,
                      double* gx, double* gy, double* gz);
double orc_grad_intra(const double* xs, const double* ys, const double* zs,
                      const double* q, const int32_t* anum, const int32_t* dists, int n,
                      double* gx, double* gy, double* gz);
double orc_grad_interp(const orc_grid_t* g, const float* const* grids,
                       const double* lxs, const double* lys, const double* lzs,
                       const int32_t* ltype, int nl,
                       double* gx, double* gy, double* gz);

/* batched CPU baseline: rigid-scan poses (rotation r, translation t):
 * x = R_r c + t_t, energy = shifted direct E_inter (BrL).  nthreads<=1: serial. */
void orc_scan_shifted(const double* pxs, const double* pys, const double* pzs,
                      const double* pq, const int32_t* panum, int np,
                      const double* cxs, const double* cys, const double* czs,
                      const double* lq, const int32_t* lanum, int nl,
                      const double* rots, const int32_t* pose_rot,
                      const double* trans, const int32_t* pose_trans, int npose,
                      int nthreads, double* e_out);
/* -ff Bst restated with a cell grid (the CPU timing baseline; same terms) */
void orc_scan_shifted_cells(const double* pxs, const double* pys, const double* pzs,
                            const double* pq, const int32_t* panum, int np,
                            const double* cxs, const double* cys, const double* czs,
                            const double* lq, const int32_t* lanum, int nl,
                            const double* rots, const int32_t* pose_rot,
                            const double* trans, const int32_t* pose_trans, int npose,
                            int nthreads, double* e_out);
void orc_scan_interp(const orc_grid_t* g, const float* const* grids,
                     const double* cxs, const double* cys, const double* czs,
                     const int32_t* ltype, int nl,
                     const double* rots, const int32_t* pose_rot,
                     const double* trans, const int32_t* pose_trans, int npose,
                     int nthreads, double* e_out);

#ifdef __cplusplus
}
#endif
/* desolvation (dead code in the reference, src/lds.ml:145-267, src/const.ml:31);
 * coordinates relative to the grid's lowest node */
double orc_desolvation_const(void);
void orc_first_solvent_shell(const double* xs, const double* ys, const double* zs,
                             const double* radii, int n, const orc_grid_t* g, uint8_t* mask);
void orc_protein_desolv(const double* pxs, const double* pys, const double* pzs,
                        const double* pq, int np, const orc_grid_t* g, const uint8_t* shell,
                        double cx, double cy, double cz, double roi_r, double* out);
void orc_desolvation_penalty(const orc_grid_t* g, const double* prot_contribs,
                             const uint8_t* prot_shell, const double* xs, const double* ys,
                             const double* zs, const double* q, const double* radii, int n,
                             double* prot, double* lig);

#endif
