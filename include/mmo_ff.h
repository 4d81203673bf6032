/*
 * mmo_ff.h -- C ABI of the MI355X (gfx950) scoring library libmmo_ff.so.
 *
 * Drop-in boundary for MMO's UFF nonbonded scoring path (UnixJunkie/MMO).  The
 * reference is OCaml: its scoring functions are pure functions over Mol.t SoA
 * arrays (src/mol.ml:17-35), consumed as closures `ene_inter : Mol.t -> float` and
 * `ene_intra : Mol.t -> float` (src/lds.ml:702-739, :1952-1983).  Each entry point
 * below names the reference function it replaces.  An OCaml binding is a thin
 * `external` stub per entry point over these symbols.
 *
 * Conventions
 *   - All functions return MMO_OK (0) or a negative MMO_ERR_*; mmo_last_error()
 *     gives a message.  Scoring itself never fails on finite input.
 *   - Host pointers are caller-owned and only read/written during the call.
 *     `_dev` entry points take device pointers and a HIP stream (`void*`, NULL =
 *     the null stream); they enqueue work and return without synchronising.
 *   - Coordinates are FP64 in the caller's frame.  The library re-origins them (in
 *     FP64) and keeps receptor/ligand positions as 32-bit fixed point, so pair
 *     differences are exact; pair arithmetic is FP32, per-pose sums are FP64.
 *     Atoms beyond the fixed-point window (~64 A from the receptor centre) take no part
 *     in the shifted form (they are beyond every cutoff) and are summed with FP64
 *     pairs in the global form.  Host-buffer entry points reject non-finite ligand
 *     input (MMO_ERR_BAD_ARG); in device batches a non-finite atom contributes nothing
 *     to the shifted form and makes the global energy NaN.
 *   - A context belongs to the process that created it: calls from a forked child
 *     fail with MMO_ERR_FORK (HIP state does not survive fork()).
 *   - A context must be used by one thread at a time.
 *   - Batches, Monte Carlo runs and minimisers may be destroyed before or after their
 *     context (or batch): a destroy call reads nothing but its own handle.
 */
#ifndef MMO_FF_H
#define MMO_FF_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mmo_ctx mmo_ctx;
typedef struct mmo_batch mmo_batch;

enum {
  MMO_OK = 0,
  MMO_ERR_BAD_ARG = -1,      /* NULL pointer, bad size, unsupported element ...      */
  MMO_ERR_DEVICE = -2,       /* a HIP call failed                                     */
  MMO_ERR_NOMEM = -3,        /* device or host allocation failed                      */
  MMO_ERR_OUT_OF_GRID = -4,  /* an atom's interpolation cell is outside the grid      */
  MMO_ERR_NO_RECEPTOR = -5,  /* mmo_set_receptor() has not been called                */
  MMO_ERR_NO_GRID = -6,      /* mmo_set_grids() has not been called                   */
  MMO_ERR_FORK = -7,         /* context used from a process other than its creator    */
  MMO_ERR_RANGE = -8         /* reserved (no entry point returns it)                  */
};

/* energy forms */
enum {
  MMO_GLOBAL = 0,    /* Mol.ene_inter_UFF_global_brute  src/mol.ml:796-818 (BrG)          */
  MMO_SHIFTED = 1,   /* Mol.ene_inter_UFF_shifted_brute src/mol.ml:822-849 (BrL), and the
                        BST form it equals, src/mol.ml:928-960 (Bst, the default -ff)     */
  MMO_INTERP = 2     /* Mol.ene_inter_UFF_interp        src/mol.ml:1012-1020              */
};

/* ---------------------------------------------------------------- context */
int  mmo_ctx_create(int device, mmo_ctx** out);
void mmo_ctx_destroy(mmo_ctx* ctx);
const char* mmo_last_error(const mmo_ctx* ctx);   /* ctx may be NULL: last global error */
/* library version string, e.g. "mmo_ff 1.0 gfx950" */
const char* mmo_version(void);

/* UFF per-element nonbond parameters.  Replaces UFF.anum_xi_Di (src/UFF.ml:10-22);
 * pair parameters follow the same geometric-mean rule (src/UFF.ml:39-51).  The
 * default table (set at creation) is the reference's: anum 0 (virtual), H, C, N, O,
 * F, Mg, P, S, Cl, Br, I.  Elements outside the table are rejected (the reference
 * reads NaN parameters for them, src/UFF.ml:37). */
int mmo_set_uff_table(mmo_ctx* ctx, const int32_t* anum, const double* x, const double* d, int n);

/* Receptor (the `prot` Mol.t of src/mol.ml:796-960).  Uploaded once; replaces
 * the previous receptor.  origin (nullable): fixed-point frame centre; default is
 * the centre of the receptor's bounding box. */
int mmo_set_receptor(mmo_ctx* ctx, const double* xs, const double* ys, const double* zs,
                     const double* q, const int32_t* anum, int n, const double* origin);

/* Per-type FP32 grids (G3D.t array, src/G3D.ml:11; the `ff_comps` of
 * src/mol.ml:1012).  grids[t] has nx*ny*nz floats, index i + j*nx + k*nx*ny
 * (src/grid.ml:98-99); origin is the position of node (0,0,0) (the reference's
 * grid starts at (0,0,0), src/grid.ml:9). */
int mmo_set_grids(mmo_ctx* ctx, const float* const* grids, int ntypes, double step,
                  int nx, int ny, int nz, const double* origin);

/* GPU grid precompute.  Per-type FP32 grids for single-atom probes (probe_anum[t],
 * probe_q[t]) at the nodes origin + (i, j, k) * step: the shifted E_inter of the
 * probe, clamped at max_e (Params.max_E = 1e5): Mol.ene_inter_UFF_shifted_grid
 * (src/mol.ml:964-989) via pre_calculate_FF_components_grid (src/lds.ml:452-469).
 * Nodes at distance >= roi_r from roi_center stay 0 (bitmask_ROI_only,
 * src/lds.ml:269-305, uses roi_r = ROI radius + 24 A); roi_center NULL: every node.
 * The result becomes the context's interpolation grids (like mmo_set_grids); out
 * (host, nullable) receives ntypes * nx*ny*nz floats, grid t at out + t*nx*ny*nz. */
int mmo_grid_compute(mmo_ctx* ctx, const int32_t* probe_anum, const double* probe_q, int ntypes,
                     double step, int nx, int ny, int nz, const double* origin,
                     const double* roi_center, double roi_r, double max_e, float* out);

/* ------------------------------------------------- plain forms (one ligand) */
/* Mol.ene_inter_UFF_global_brute / _shifted_brute / _shifted_bst_components
 * (src/mol.ml:796-960).  e_elec/e_vdw nullable; e_elec already includes the
 * 332.0637/4 factor (src/mol.ml:955). */
int mmo_ene_inter(mmo_ctx* ctx, int mode, const double* xs, const double* ys, const double* zs,
                  const double* q, const int32_t* anum, int n,
                  double* e, double* e_elec, double* e_vdw);

/* Mol.ene_inter_UFF_interp (src/mol.ml:1012-1020) with G3D.trilin (src/G3D.ml:97-157);
 * type[j] selects the grid (Mol.t.t_a, src/mol.ml:455-461). */
int mmo_ene_inter_interp(mmo_ctx* ctx, const double* xs, const double* ys, const double* zs,
                         const int32_t* type, int n, double* e);

/* Mol.ene_intra_UFFNB_brute (src/mol.ml:881-903); dists is the topological distance
 * matrix as stored in Mol.t.dists_a: dists[i + j*n] (src/mol.ml:151-152). */
int mmo_ene_intra(mmo_ctx* ctx, const double* xs, const double* ys, const double* zs,
                  const double* q, const int32_t* anum, const int32_t* dists, int n, double* e);

/* Gradient forms of the three plain functions: dE/dxyz for each ligand atom
 * (the reference only has central finite differences, src/lds.ml:1118-1136). */
int mmo_ene_inter_grad(mmo_ctx* ctx, int mode, const double* xs, const double* ys,
                       const double* zs, const double* q, const int32_t* anum, int n,
                       double* e, double* gx, double* gy, double* gz);
int mmo_ene_intra_grad(mmo_ctx* ctx, const double* xs, const double* ys, const double* zs,
                       const double* q, const int32_t* anum, const int32_t* dists, int n,
                       double* e, double* gx, double* gy, double* gz);
int mmo_ene_inter_interp_grad(mmo_ctx* ctx, const double* xs, const double* ys,
                              const double* zs, const int32_t* type, int n,
                              double* e, double* gx, double* gy, double* gz);

/* --------------------------------------------------------- ligand registry */
/* Registers a ligand topology for the batched, DOF, scan and MC forms: the
 * centred conformer (Lds.preprocess_ligand, src/lds.ml:45-51), charges, atomic
 * numbers, the dists matrix (dists[i + j*n]), rotatable bonds with the axis
 * left->right and rotation groups rg_atoms[rg_off[k]..rg_off[k+1]) excluding
 * `right` (src/mol.ml:385-402, src/pqrs.ml:80-87), and grid types (nullable).  Returns an id >= 0. */
int mmo_ligand_add(mmo_ctx* ctx, const double* xs, const double* ys, const double* zs,
                   const double* q, const int32_t* anum, const int32_t* dists, int n,
                   const int32_t* rb_left, const int32_t* rb_right, const int32_t* rg_off,
                   const int32_t* rg_atoms, int nrb, const int32_t* types, int* lig_id);
int mmo_ligand_count(const mmo_ctx* ctx);
int mmo_ligand_natoms(const mmo_ctx* ctx, int lig_id);

/* ------------------------------------------------------------ batched forms */
/* A batch is a fixed list of poses (pose p is a conformation of ligand
 * pose_lig[p]); its atoms are stored pose-major: atom j of pose p is row
 * atom_off[p] + j of an (n_atoms_total x 3) FP64 array.  Planning (grouping poses
 * of one ligand into wavefront tiles, tile shape from the autotuner) happens here. */
int mmo_batch_create(mmo_ctx* ctx, const int32_t* pose_lig, int npose, mmo_batch** out);
void mmo_batch_destroy(mmo_batch* b);
int mmo_batch_natoms(const mmo_batch* b);           /* total atoms (rows)      */
int mmo_batch_atom_offsets(const mmo_batch* b, int32_t* off /* npose+1 */);

/* flags for batched scoring */
enum {
  MMO_F_INTER_SHIFTED = 1,   /* e_out[p*4+0] += shifted direct E_inter                 */
  MMO_F_INTER_GLOBAL = 2,    /* e_out[p*4+0] += global direct E_inter                  */
  MMO_F_INTER_INTERP = 4,    /* e_out[p*4+0] += interpolated E_inter                   */
  MMO_F_INTRA = 8,           /* e_out[p*4+1]  = E_intra                                */
  MMO_F_GRAD = 16,           /* also dE/dxyz (inter + intra) into grad (n_atoms x 3)   */
  MMO_F_COMPONENTS = 32      /* e_out[p*4+2], [p*4+3] = elec, vdW parts of direct E_inter */
};

/* Scores every pose of the batch from coordinates.  xyz: device FP64
 * (n_atoms_total x 3); e_out: device FP64 (npose x 4: inter, intra, elec, vdW);
 * grad: device FP64 (n_atoms_total x 3) or NULL.  Out-of-grid poses (INTERP) get
 * NaN in e_out[p*4+0].  GLOBAL sums atoms anywhere (those outside the fixed-point
 * window, ~64 A from the receptor centre, through FP64 pairs); non-finite atoms give
 * NaN. */
int mmo_batch_score_dev(mmo_batch* b, int flags, const double* xyz, double* e_out,
                        double* grad, void* stream);
/* Host-pointer variant (copies in and out; synchronous). */
int mmo_batch_score(mmo_batch* b, int flags, const double* xyz, double* e_out, double* grad);

/* DOF form.  Optim.apply_config (src/optim.ml:64-80): for pose p the DOF vector is
 * dofs[dof_off[p] .. dof_off[p+1]) = [x, y, z, alpha, beta, gamma, tau_0..tau_{R-1}]
 * (src/optim.ml:19-22), torsions applied in rbond order to the registered centred
 * conformer (Mol.rotate_bond, src/mol.ml:610-631), then Rot.r_xyz and translation.
 * Writes coordinates to xyz_out (n_atoms_total x 3, nullable), energies to e_out,
 * and with MMO_F_GRAD dE/dDOF to gdof (same layout as dofs).  radius_out (nullable,
 * npose) gets the enclosing radius after the torsions (Mol.check_elongation_exn uses
 * > 12 A as Too_long, src/optim.ml:75).  All pointers are device pointers. */
int mmo_batch_dof_offsets(const mmo_batch* b, int32_t* dof_off /* npose+1 */);
int mmo_batch_score_dofs_dev(mmo_batch* b, int flags, const double* dofs, double* xyz_out,
                             double* e_out, double* gdof, double* radius_out, void* stream);

/* --------------------------------------------------------------- rigid scan */
/* Exhaustive rigid scan (Lds.exhaustive_rigid_ligand_docking, src/lds.ml:1040-1114,
 * without its clash prefilters): pose (r, t) places the registered centred ligand at
 * R_r c + T_t (Mol.centered_rotate_copy + translate_copy_to, src/lds.ml:1081-1094);
 * e_out[r * ntrans + t] = E_inter of that pose in `mode` (MMO_SHIFTED, MMO_GLOBAL or
 * MMO_INTERP).  rots: host (nrot x 9, row major); trans: host (ntrans x 3);
 * e_out: device (nrot * ntrans).  Enqueued on `stream`. */
int mmo_scan_dev(mmo_ctx* ctx, int lig_id, int mode, const double* rots, int nrot,
                 const double* trans, int ntrans, double* e_out, void* stream);

/* ------------------------------------------- rigid-scan filters and top-k */
/* Protein vdW volume bitmask (Lds.vdW_volume / atom_bitmask_set, src/lds.ml:164-197):
 * voxel (i, j, k) of the grid origin + (i, j, k) * step is set when it lies strictly
 * inside the vdW sphere of a receptor atom.  radii (nullable) per receptor atom; NULL:
 * Ptable.vdW_radii by element (src/ptable.ml:43-53).  origin NULL = (0, 0, 0), the
 * reference's grid corner.  Needs the receptor; replaced by the next receptor. */
int mmo_set_vdw_mask(mmo_ctx* ctx, const double* radii, double step, int nx, int ny, int nz,
                     const double* origin);
/* Copies the mask's 32-bit words (bit v of word v/32 = voxel v, v = i + j*nx + k*nx*ny)
 * to words (nullable) and returns the number of words. */
int mmo_get_vdw_mask(const mmo_ctx* ctx, uint32_t* words);
enum {
  MMO_SCAN_ROI = 1,           /* skip translations outside the ROI sphere (ROI.is_inside) */
  MMO_SCAN_CENTER_CLASH = 2,  /* ligands whose centre lies inside one of their atoms: skip
                                 translations whose 8 voxel corners are all protein
                                 volume (G3D.vdW_clash_AND, src/G3D.ml:188-213)        */
  MMO_SCAN_ATOM_CLASH = 4     /* skip poses with an atom that has a protein-volume corner
                                 (Mol.protein_ligand_clash, src/mol.ml:1195-1206)     */
};
/* The clash filters of Lds.exhaustive_rigid_ligand_docking (src/lds.ml:1040-1114) on the
 * energies of a finished mmo_scan_dev over the same rotations and translations: the
 * skipped poses' energies become +inf (they are never best and never in the top-k).
 * lig_radii: the ligand's atom radii (for MMO_SCAN_CENTER_CLASH).  Corners outside the
 * mask's grid count as free.  e_inout: device (nrot * ntrans).  Returns after the
 * filter is enqueued and its host inputs are consumed. */
int mmo_scan_filter_dev(mmo_ctx* ctx, int lig_id, const double* rots, int nrot, const double* trans,
                        int ntrans, const double* roi_center, double roi_r,
                        const double* lig_radii, int flags, double* e_inout, void* stream);
/* Filter before scoring: the exhaustive scan of mmo_scan_dev restricted to the poses the
 * clash filters of mmo_scan_filter_dev keep, as Lds.exhaustive_rigid_ligand_docking
 * scores only the poses it does not skip (src/lds.ml:1082-1092).  The filters run first
 * (per translation: ROI sphere and centre clash; per pose: atom clashes), the kept poses
 * are compacted in the scan's pose order and only they are scored (mode MMO_SHIFTED, the
 * culled kernel, or MMO_INTERP); e_out (device, nrot * ntrans) holds their energies and
 * +inf for every skipped pose, bit for bit what mmo_scan_dev followed by
 * mmo_scan_filter_dev gives, so the top-k is the same.  n_kept (host, nullable) receives
 * the number of scored poses.  Returns when e_out is complete. */
int mmo_scan_filtered_dev(mmo_ctx* ctx, int lig_id, int mode, const double* rots, int nrot,
                          const double* trans, int ntrans, const double* roi_center, double roi_r,
                          const double* lig_radii, int flags, double* e_out, int64_t* n_kept,
                          void* stream);
/* The k lowest energies of e (device, n) with their indices, ascending, ties by index
 * (the TopKeeper of src/lds.ml:1055-1064, :1110-1113); NaN sorts last; k <= 2048
 * (k > n gives n).  e_out, idx_out: device.  Enqueued on stream. */
int mmo_topk_dev(mmo_ctx* ctx, const double* e, int64_t n, int k, double* e_out, int64_t* idx_out,
                 void* stream);
/* mmo_topk_dev without any allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops).  The
 * kernels' scratch (radix state, histogram, counters, per-chunk counts) is a static
 * device array of 132 KiB per GPU, set up by mmo_ctx_create, so n <= 134,217,728
 * (larger n -> MMO_ERR_BAD_ARG); same results as mmo_topk_dev.  All calls on one GPU
 * share that array: calls that overlap on one GPU (other streams, other contexts,
 * other threads) race and must be ordered by the caller. */
int mmo_topk_dev_noalloc(mmo_ctx* ctx, const double* e, int64_t n, int k, double* e_out,
                         int64_t* idx_out, void* stream);

/* ------------------------------------------------------------- desolvation */
/* Protein desolvation (Lds.first_solvent_shell and Lds.protein_desolv, src/lds.ml:171-234;
 * formula (2) of Majeux, Scarsi and Caflisch, Proteins 2001): on the grid origin + (i, j,
 * k) * step (origin NULL = (0, 0, 0), the reference's grid corner), the receptor's first
 * solvent shell (nodes within r_vdW + 1.4 A of an atom and not within r_vdW of any;
 * radii nullable: Ptable.vdW_radii by element), and per shell node strictly inside the
 * ROI sphere (roi_center NULL: every shell node) the contribution
 * Const.desolvation (src/const.ml:31) x step^3 x sum over receptor atoms closer than
 * 12 A of (q / d^2)^2; 0 at every other node.  Needs the receptor; replaced by the
 * next receptor.  Returns the number of nodes given a contribution (>= 0) or an error. */
int mmo_desolv_setup(mmo_ctx* ctx, const double* radii, double step, int nx, int ny, int nz,
                     const double* origin, const double* roi_center, double roi_r);
/* The receptor's first solvent shell of the last mmo_desolv_setup as 32-bit words (bit v
 * of word v/32 = node v, v = i + j*nx + k*nx*ny) into words (nullable); returns the
 * number of words. */
int mmo_get_solvent_shell(const mmo_ctx* ctx, uint32_t* words);
/* The per-node protein contributions of the last mmo_desolv_setup into out (host,
 * nullable, nx*ny*nz doubles); returns the number of nodes. */
int mmo_get_desolv_contribs(mmo_ctx* ctx, double* out);
/* Lds.desolvation_penalty (src/lds.ml:236-267) of one ligand pose: over the nodes in both
 * the receptor's and the ligand's first solvent shells, *prot = the sum of the protein
 * contributions and *lig = Const.desolvation x step^3 x the sum of the ligand atoms'
 * (q / d^2)^2 closer than 12 A.  xyz: host (n x 3 rows, the receptor's frame); q, radii:
 * host (n).  Sums in FP64 (the node order of the sums differs from the reference's, so
 * results agree to rounding).  Needs mmo_desolv_setup. */
int mmo_desolv_penalty(mmo_ctx* ctx, const double* xyz, const double* q, const double* radii,
                       int n, double* prot, double* lig);

/* -------------------------------------------------------------- Monte Carlo */
/* Batched MC chains (Lds.simulate_lig, src/lds.ml:741-1000) on the GPU: chain k
 * docks ligand chain_lig[k] from the start pose (rot0[9k..], pos0[3k..]) with rigid
 * moves on even frames and torsion moves on odd frames, E = E_inter (MMO_SHIFTED:
 * direct sum; MMO_INTERP: the context's grids) + E_intra (UFF_NB when
 * MMO_MC_INTRA_UFF), Metropolis at `temperature` K, adaptive step sizes, best pose.
 * Random numbers are Philox4x32-10 (seed; counter = chain, frame, draw), so runs are
 * reproducible and independent of launch shape.  Every chain runs total_frames
 * frames in all.  The reference has one process per chain (src/lds.ml:2050); here all
 * chains of all ligands run in one persistent launch. */
typedef struct mmo_mc mmo_mc;
enum {
  MMO_MC_TWEAK_RBONDS = 1,       /* torsion moves (the reference's tweak_rbonds)       */
  MMO_MC_HARD_ROI = 2,           /* --hard-ROI: reset outside the ROI (src/lds.ml:908) */
  MMO_MC_INTRA_UFF = 4,          /* E_intra = Mol.ene_intra_UFFNB_brute                */
  MMO_MC_METROPOLIS_ALWAYS = 8,  /* accept/reject also without hard ROI (see below)    */
  MMO_MC_NO_FLIP = 16,           /* --no-flip-move (src/lds.ml:1828)                   */
  MMO_MC_FP64 = 32               /* FP64 pair arithmetic for MMO_SHIFTED (parity mode)  */
};
/* Without MMO_MC_HARD_ROI the reference never reaches its acceptance test: OCaml binds
 * the `else` of src/lds.ml:917 to the inner `if dist_roi > out_radius`.  The default
 * reproduces that; MMO_MC_METROPOLIS_ALWAYS is the evident intent (not the reference). */
int mmo_mc_create(mmo_ctx* ctx, const int32_t* chain_lig, int nchain, const double* rot0,
                  const double* pos0, int inter_mode, int flags, const double* roi_center,
                  double roi_r, double temperature, uint64_t seed, int total_frames,
                  mmo_mc** out);
/* Advances every unfinished chain by nsteps frames in one persistent launch. */
int mmo_mc_run(mmo_mc* mc, int nsteps, void* stream);
/* Same launch captured once in a hipGraph (queue reset + kernel) and replayed
 * nreplay times: nreplay * nsteps frames per chain. */
int mmo_mc_run_graph(mmo_mc* mc, int nsteps, int nreplay, void* stream);
/* Plain form for comparison: one launch per ligand (its chains only). */
int mmo_mc_run_per_ligand(mmo_mc* mc, int nsteps, void* stream);
/* Monte Carlo-plus-minimisation: after every proposal (rigid or torsion move) the chain
 * runs `iters` L-BFGS energy + gradient evaluations (the batched minimiser's algorithm,
 * history pairs `history`, first steps capped at max_step, convergence at gtol) over the
 * proposal's DOF vector [t, alpha, beta, gamma, tau..] (torsions only with
 * MMO_MC_TWEAK_RBONDS), starting from the proposal itself, and the Metropolis test, the
 * E_inter == 0 and hard-ROI resets, the best pose and the step adaptation then apply to
 * the minimised pose (the ROI test to its centre).  The reference has no such mode: its
 * MC (Lds.simulate_lig, src/lds.ml:741-1000) and its minimiser (Lds.minimize_ligand,
 * src/lds.ml:1138-1297) are separate; this composes them per frame in the persistent
 * launch.  iters = 0 switches it off.  The start pose is scored, not minimised. */
int mmo_mc_set_minimizer(mmo_mc* mc, int iters, int history, double max_step, double gtol);
/* Energy + gradient evaluations of the local minimisations so far: per chain (nullable)
 * and in total (nullable). */
int mmo_mc_min_evals(mmo_mc* mc, int32_t* evals, int64_t* total);
int mmo_mc_natoms(const mmo_mc* mc);
/* Per chain (host arrays, each nullable): best_e[k]; cur_e[3k..] = E, E_inter,
 * E_intra of the current state; best pose rotation and position; info[8k..] = frame,
 * status (0 running, 1 done, 2 stopped by Mol.Too_long), accepted, resets, out-of-ROI
 * resets, rigid steps, conformer steps, initialised; steps[2k..] = max_rot, max_trans. */
int mmo_mc_get(mmo_mc* mc, double* best_e, double* cur_e, double* best_rot, double* best_pos,
               int32_t* info, double* steps);
/* Current conformers, best poses (rows of all chains, chain order) and torsion steps. */
int mmo_mc_coords(mmo_mc* mc, double* conf, double* best_xyz, double* rb_dr);
void mmo_mc_destroy(mmo_mc* mc);

/* ------------------------------------------------------- local minimisation */
/* Batched L-BFGS in DOF space.  The reference minimises E_intra + E_inter of
 * Optim.apply_config(dofs) with NLopt inside box bounds, gradients by central
 * differences (Lds.minimize_ligand / nlopt_minimize / approx_local_gradient,
 * src/lds.ml:1014-1036, :1118-1297; bounds src/optim.ml:34-60).  Here every pose of
 * the batch runs limited-memory BFGS (history 1..16 pairs) on the analytic gradient,
 * with Armijo backtracking, trial points projected onto [lo, hi] (host arrays laid out
 * like the batch's DOFs, nullable = unbounded), first trial steps capped at max_step
 * per component (A or rad; <= 0: 1.0) and convergence at max |dE/dDOF| <= gtol
 * (<= 0: 1e-6).  flags: one of MMO_F_INTER_SHIFTED / _GLOBAL / _INTERP, and/or
 * MMO_F_INTRA.  Poses longer than 12 A after their torsions (Mol.Too_long) or with a
 * non-finite energy count as failed trials (the reference abandons the ligand). */
typedef struct mmo_min mmo_min;
int mmo_min_create(mmo_batch* b, int history, int flags, const double* lo, const double* hi,
                   double max_step, double gtol, mmo_min** out);
/* Minimises from dofs (device, in/out: start point in, best point out); each of the
 * iters iterations is one energy + gradient evaluation of every pose followed by one
 * L-BFGS update.  e_out (device, npose, nullable): best energies.  Enqueued on stream. */
int mmo_min_run_dev(mmo_min* mz, double* dofs, int iters, double* e_out, void* stream);
/* After a run (synchronises): info[4p..] = status (0 running, 1 converged, 2 failed
 * start), evaluations, accepted steps, history restarts; e_start[p]: start energy. */
int mmo_min_get(mmo_min* mz, int32_t* info, double* e_start);
void mmo_min_destroy(mmo_min* mz);

/* ------------------------------------------------------------------ tuning */
/* Kernel time of the last launch of `kernel` ("inter", "interp", "intra", "grid") measured
 * with HIP events on the launch stream, in milliseconds (-1 if none). */
double mmo_last_kernel_ms(mmo_ctx* ctx, const char* kernel);
/* ------------------------------------------------------------- ligand input
 * Host-only (no GPU needed).  The reference prepares ligands with mol2pqrs and reads
 * them with Pqrs/Mol (src/mol2.ml:99-323, src/mol_graph.ml:53-192, src/mol2pqrs.ml:9-43,
 * src/pqrs.ml:17-87, src/mol.ml:420-440): a .mol2 file is parsed (lone pairs removed),
 * its topology derived (BFS topological distances; rotatable bonds: single, both ends of
 * degree > 1, not in a ring; rotation group: the smaller side), written as .pqrs text
 * and read back, so .mol2 and .pqrs inputs give the same molecules.  A molset holds the
 * molecules; the accessors copy them into caller buffers sized by mmo_molset_dims.
 * Malformed input -> MMO_ERR_BAD_ARG with the reason in mmo_last_error(NULL). */
typedef struct mmo_molset mmo_molset;
int mmo_molset_from_mol2(const char* text, size_t len, mmo_molset** out);
/* ligands != 0: ligand records (header n:nrb:name, rot-bond lines, distance matrix);
 * 0: one receptor record (n:name, atom lines only; Mol.receptor_of_pqrs_file) */
int mmo_molset_from_pqrs(const char* text, size_t len, int ligands, mmo_molset** out);
int mmo_molset_count(const mmo_molset* s);
const char* mmo_molset_name(const mmo_molset* s, int i);
/* atoms, rotatable bonds (-1 for a receptor record), atoms of all rotation groups */
int mmo_molset_dims(const mmo_molset* s, int i, int* n_atoms, int* n_rbonds, int* n_rg_atoms);
/* coordinates, partial charges, vdW radii (NaN where the element has none), elements;
 * any pointer may be NULL */
int mmo_molset_atoms(const mmo_molset* s, int i, double* xs, double* ys, double* zs,
                     double* q, double* r, int32_t* anum);
/* Mol.t topology: dists[i + j n] (n x n), rotatable bonds as (left, right) with right on
 * the moving side, rotation groups in CSR (rg_off: n_rbonds + 1; rg_atoms: the moving
 * atoms without `right`); any pointer may be NULL; receptor records have none */
int mmo_molset_topology(const mmo_molset* s, int i, int32_t* dists, int32_t* rb_left,
                        int32_t* rb_right, int32_t* rg_off, int32_t* rg_atoms);
/* the .pqrs text of the set (the converted text for .mol2 input, the input itself for
 * .pqrs); *len receives its size; buf may be NULL to query the size */
int mmo_molset_pqrs(const mmo_molset* s, char* buf, size_t cap, size_t* len);
void mmo_molset_destroy(mmo_molset* s);

/* Fixed-point scale exponent k (coordinates are stored as round(x * 2^k)). */
int mmo_fixed_point_bits(const mmo_ctx* ctx);
/* Kernel shape parameters, set by the autotuner (mmo_amd/tune.py): "inter_k",
 * "inter_waves", "inter_rtile", "cull_k", "cull_waves" (rigid scan), "scan_cull" (rigid
 * scan receptor path: 1 LDS, 2 scalar cache, 3 scalar cache on x-cut cells, 4 positions
 * in LDS; 5 row lists of the scan grid through the scalar cache, 6 the same with the
 * receptor in LDS), "scan_cell" and "scan_xsub" (the row lists' cell edge in y and z and
 * its x cut; next mmo_set_receptor), "scan_chunk" (receptor atoms the row lists fetch at
 * once: 4 or 8), "row_filter" (1: the row lists test each far row's atoms against the
 * wavefront's box grown by the cutoff and sweep only those within it; same energies), "cell_size" (applies to the next mmo_set_receptor), "mc_waves_per_cu",
 * "pose_waves" (batches: wavefronts per workgroup), "pose_cull" (1: batches in
 * the shifted form run one wavefront per pose with a culled receptor sweep; 0: tiles of
 * poses over the whole receptor), "pose_sweep" (that sweep: 1 lanes over the receptor
 * atoms near each ligand atom, 0 per 8-atom sub-block), "mc_sweep" (the same choice for
 * the Monte Carlo direct sum), "row_clip" (scan and grid sweeps clip cell rows longer
 * than this many atoms to the cutoff sphere's x chord), "cell_xsub" (the pose and Monte
 * Carlo sweeps use cells up to this many times narrower along x, fewer when the receptor
 * and that cell table would pass 96 KiB; applies to the next mmo_set_receptor),
 * "mc_wg_waves" (Monte Carlo chains per workgroup).  No counterpart in the reference. */
int mmo_tune_set(mmo_ctx* ctx, const char* key, double value);
double mmo_tune_get(mmo_ctx* ctx, const char* key);

#ifdef __cplusplus
}
#endif
#endif
