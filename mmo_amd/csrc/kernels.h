// kernels.h -- host-side launchers of the gfx950 kernels (implemented in *.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"
#include "mmo_ff.h"

namespace mmo {

struct InterLaunch {
  int mode;            // MMO_SHIFTED / MMO_GLOBAL
  int k;               // ligand atoms per lane (1, 2, 4, 8)
  int waves;           // wavefronts per block
  int rtile;           // receptor atoms per LDS tile
  int rslice;          // receptor atoms per block row (gridDim.y > 1 splits the receptor)
  int cull;            // culled shifted form: 1 receptor cell list in LDS, 2 in global memory,
                       // 3 positions in LDS; 4 row lists (k_scan_rows), receptor in place,
                       // 5 row lists with the receptor in LDS
  int cull_waves;      // wavefronts per block of the culled kernel
  int cull_k;          // ligand atoms per lane in the culled kernel (1, 2, 4)
  int row_chunk;       // k_scan_rows: receptor atoms fetched at once (4 or 8)
  int row_filter;      // k_scan_rows: far rows' atoms tested against the box first (0 / 1)
  const int* atom_order;   // nullable: per ligand at atom_base, the sweeps' atom order
  CellGrid cg;
  Frame fr;
  const RecPos* rpos;
  const RecAB* rab;
  const RecRes* rres;
  int nrec;
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;
};

struct GridDev {
  const float* data;   // ntypes * nvox
  int nx, ny, nz;
  size_t nvox;
  double step, ods;    // step and 1/step
  double ox, oy, oz;   // position of node (0,0,0)
  // nullable: the same values by cell, the 8 corners of cell (i, j, k) (i < nx - 1, ...)
  // contiguous in the trilinear corner order a..h, 32 B per cell and type:
  // cells[((t * (nz-1) + k) * (ny-1) + j) * (nx-1) + i) * 8 + corner]
  const float* cells;
};
hipError_t launch_pack_cells(const GridDev& g, int ntypes, float* cells, hipStream_t s);

struct InterpLaunch {
  GridDev g;
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;
};

struct IntraArgs {
  const LigInfo* ligs;
  const uint32_t* pair_ij;
  const double4* pair_par;
  const double* xyz;
  const int* atom_off;
  const int* pose_lig;
  int npose;
  int max_atoms;
  double* e_out;
  int e_stride;
  double* grad;
  int grad_accumulate;
};

struct GridArgs {
  Frame fr;
  const RecPos* rpos;
  const RecAB* rab;
  const RecRes* rres;
  int nrec;
  int nx, ny, nz;
  size_t nvox;
  double ox, oy, oz, step;       // node (0,0,0) and spacing, caller frame
  double roi_x, roi_y, roi_z;    // mask: points with |p - roi| < sqrt(roi_r2); roi_r2 <= 0: all
  double roi_r2;
  const double4* probe;          // per type: ke q, A, B
  int ntypes;
  double max_e;                  // clamp (Params.max_E = 1e5, src/params.ml:26)
  float* out;                    // ntypes * nvox, index i + j nx + k nx ny
};

struct DofArgs {
  const LigInfo* ligs;
  const double* lig_xyz;        // centred conformers
  const int* rb_left;           // per ligand at rb_base
  const int* rb_right;
  const int* rg_off;            // per ligand at rb_base + lig (nrb+1 entries)
  const int* rg_atoms;
  const int* pose_lig;
  const int* atom_off;
  const int* dof_off;
  int npose, max_atoms;
  int max_rb, max_rg;           // most rbonds and rotation-group atoms of a ligand in the batch
  const double* dofs;
  double* xyz_out;
  double* radius_out;           // nullable
  double* conf;                 // nullable: the posed conformers before the rigid-body motion
                                // (rows), written by apply_config and read by the adjoint
  const double* grad_xyz;       // dE/dxyz (for the adjoint)
  double* gdof;                 // dE/dDOF
  // nullable: the adjoint also evaluates E_intra and its gradient (pair list below) and
  // writes E_intra to e_intra[pose * e_stride]
  double* e_intra;
  int e_stride;
  const uint32_t* pair_ij;
  const double4* pair_par;      // ke qi qj, Ai Aj, 2 Bi Bj
};

// part: scratch of (nslices * npose * e_stride) doubles when the receptor is split
hipError_t launch_inter_coords(const InterLaunch& L, const Tile* tiles, const int* perm,
                               int ntiles, int npose, const double* xyz, const int* atom_off,
                               double* e_out, int e_stride, bool comp, double* grad, int write_e,
                               double* part, hipStream_t s);
hipError_t launch_inter_scan(const InterLaunch& L, int lig, const double* rots, int nrot,
                             const double* trans, const int* tperm, int ntrans, double* e_out,
                             hipStream_t s);
hipError_t launch_intra(const IntraArgs& a, bool grad, hipStream_t s);
hipError_t launch_interp_coords(const InterpLaunch& L, const Tile* tiles, const int* perm,
                                int ntiles, const double* xyz, const int* atom_off,
                                const int* type_override, double* e_out, int e_stride,
                                double* grad, int grad_accumulate, hipStream_t s);
hipError_t launch_interp_scan(const InterpLaunch& L, int lig, const double* rots, int nrot,
                              const double* trans, const int* tperm, int ntrans, double* e_out,
                              hipStream_t s);
// filtered scan: score only the listed slots of the scan's pose order (the culled kernel)
hipError_t launch_inter_scan_list(const InterLaunch& L, int lig, const double* rots,
                                  const double* trans, const int* tperm, int ntrans,
                                  const int* list, int nlist, double* e_out, hipStream_t s);
hipError_t launch_interp_scan_list(const InterpLaunch& L, int lig, const double* rots,
                                   const double* trans, const int* tperm, int ntrans,
                                   const int* list, int nlist, double* e_out, hipStream_t s);
// desolvation sums (src/lds.ml:201-267): per listed node v (index i + j nx + k nx ny of
// the grid origin + (i, j, k) step), sum over atoms (x, y, z, q) within 12 A of (q / d^2)^2
struct DesolvArgs {
  const int* vox;             // nodes
  int nvox;
  const double4* atoms;       // x, y, z, q
  int natoms;
  double ox, oy, oz, step;
  int nx, ny;
  double scale;               // nodes: out[v] = scale * sum
  double* out;                // nodes: per grid node
  const double* prot;         // pose: the protein contributions per grid node
  double* part;               // pose: per workgroup (sum, sum of prot[v])
};
// k_desolv_pose writes part[2b], part[2b+1] for workgroup b of kDesolvNodes nodes
constexpr int kDesolvNodes = 256;
hipError_t launch_desolv(const DesolvArgs& a, bool pose, hipStream_t s);
hipError_t launch_fill(double* p, size_t n, double v, hipStream_t s);
hipError_t launch_fill_int(int* p, size_t n, int v, hipStream_t s);
hipError_t launch_grid(const GridArgs& a, const CellGrid& cg, bool in_lds, hipStream_t s);
// one pose per wavefront, shifted form, receptor culled per 8-atom sub-block
struct PoseCullArgs {
  const double* xyz;        // rows (x, y, z), pose-major
  const int* atom_off;
  const int* pose_lig;
  int npose;
  const LigInfo* ligs;
  const LigAtom* latoms;
  double* e_out;            // e_out[pose * e_stride] = E_inter
  int e_stride;
  double* e_comp;           // nullable: [pose*e_stride + 2] elec, [+3] vdW
  double* grad;             // GRAD: dE/dxyz rows, added to what the buffer holds
  int grad_assign;          // GRAD: write the rows instead (each atom's row is written once)
  int write_e;
  const RecRes* rres;       // receptor quantisation residuals
  // nullable: the pose of each launch slot (slot = workgroup * waves + wavefront), a
  // permutation of 0..npose-1
  const int* order;
  // nullable: per ligand at atom_base, the order in which the per-atom sweep takes the
  // atoms (a permutation of 0..n_atoms-1)
  const int* atom_order;
};
// atomwise: lanes over the receptor atoms near each ligand atom (atom_sweep) instead of
// 8-atom sub-blocks (pose_sweep)
hipError_t launch_pose_cull(const InterLaunch& L, const PoseCullArgs& a, bool grad, int waves,
                            bool atomwise, hipStream_t s);

// LDS a workgroup may hold on gfx950
constexpr size_t kMaxLds = 160 * 1024;
// LDS bytes of one pose of the DOF kernels (coordinates plus the staged torsion tree)
size_t dof_lds_bytes(const DofArgs& a, bool grad);
hipError_t launch_apply_config(const DofArgs& a, hipStream_t s);
hipError_t launch_dof_grad(const DofArgs& a, hipStream_t s);

// -------------------------------------------------- rigid-scan filters, top-k
struct ScanFilterArgs {
  const double* rots;         // nrot x 9
  const double* trans;        // ntrans x 3
  int nrot, ntrans;
  const double* lig;          // centred conformer rows (x, y, z)
  int n_atoms;
  const uint32_t* mask;       // protein vdW volume bitmask, voxel i + j*nx + k*nx*ny
  double ox, oy, oz, inv_step;
  int nx, ny, nz;
  const uint8_t* skip;        // per translation: 1 = skipped (ROI / centre clash)
  int atom_clash;             // test every atom's 8 corners
  double* e;                  // scan energies, set to +inf for skipped poses
};
hipError_t launch_scan_filter(const ScanFilterArgs& a, hipStream_t s);
// keep[slot] = 1 for the poses the filters keep, over the scan's pose order (rotation r,
// translation tperm[slot - r * ntrans])
hipError_t launch_scan_keep(const ScanFilterArgs& a, const int* tperm, uint8_t* keep, hipStream_t s);
// out[0 .. *n_out) = the slots whose flag is set, in order (n_out: device); temp: device
// scratch of *temp_bytes (with temp == nullptr only *temp_bytes is set)
hipError_t select_flagged_slots(const uint8_t* flags, int n, int* out, int* n_out, void* temp,
                                size_t* temp_bytes, hipStream_t s);

constexpr int kTopkMax = 2048;
struct TopkScratch {
  uint64_t* st;               // 3
  unsigned* hist;             // 256
  unsigned* cnt;              // 1
  unsigned* bcnt;             // one per chunk
  size_t chunk;
};
hipError_t launch_topk(const double* e, size_t n, int k, double* out_e, int64_t* out_i,
                       TopkScratch w, hipStream_t s);

// ------------------------------------------------------------ L-BFGS minimiser
constexpr int kMinMaxHist = 16;
enum { kMinRunning = 0, kMinConverged = 1, kMinFailed = 2 };
struct MinPose {
  double f, f0, alpha, gamma;   // energy at x, at the start, step length, H0 scale
  int init, hist, head, status;
  int n_eval, n_acc, n_reset, pad;
};
struct MinArgs {
  int npose, m, ndof_total;
  const int* dof_off;
  double* x;            // accepted point (the best point)
  double* g;            // gradient at x
  double* d;            // search direction
  double* xt;           // trial point: the input of the next evaluation
  const double* gt;     // gradient at the trial point
  const double* et;     // energies at the trial point (npose x 4: inter, intra, ...)
  const double* rad;    // enclosing radius after the torsions
  double* S;            // history (m x ndof_total)
  double* Y;
  double* rho;          // m x npose
  MinPose* st;
  const double* lo;     // box bounds (ndof_total) or null
  const double* hi;
  double max_step, min_step, gtol, c1;
};
hipError_t launch_lbfgs(const MinArgs& a, hipStream_t s);

// energy forms of the DOF evaluator (dofmin.h; the Monte Carlo kernel's numbering)
enum { kDofInterGrid = 0, kDofInterAtom = 1, kDofInterF64 = 3 };

// The whole minimisation of a batch in one persistent launch: each wavefront takes poses
// from a queue and runs all iterations of one pose (evaluation and update) without
// leaving the kernel.  inter: kDofInterAtom (shifted, FP32 per-atom sweep), kDofInterF64
// or kDofInterGrid (dofmin.h); the batch's DOF vectors apply to the registered conformers.
struct MinPersistArgs {
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;        // registered centred conformers (the base of every pose)
  const int* atom_order;
  const int* rb_left;
  const int* rb_right;
  const int* rg_off;
  const int* rg_atoms;
  const uint32_t* pair_ij;
  const double4* pair_par;
  int npose, max_atoms, ndof_total, iters, m;
  const int* pose_lig;
  const int* dof_off;
  const int* order;             // nullable: queue position -> pose (largest ligands first)
  double* x;                    // DOF vectors: start in, best point out
  const double* lo;             // box bounds (ndof_total) or null
  const double* hi;
  double* S;                    // history: S[slot * ndof_total + dof], rho[slot * npose + pose]
  double* Y;
  double* rho;
  MinPose* st;
  double* e_out;                // nullable: best energy per pose
  int* queue;
  int intra;
  double max_step, min_step, gtol, c1;
  Frame fr;
  const RecPos* rpos;
  const RecAB* rab;
  const RecRes* rres;
  int nrec;
  CellGrid cg;
  GridDev g;
};
size_t min_persist_lds_bytes(const MinPersistArgs& a, int waves, bool rec_lds);
hipError_t launch_min_persist(const MinPersistArgs& a, int inter, int nblocks, int waves,
                              bool rec_lds, hipStream_t s);
hipError_t launch_min_reset(MinPose* st, int npose, hipStream_t s);
hipError_t launch_min_energy(const MinPose* st, int npose, double* e_out, hipStream_t s);

}  // namespace mmo
