// mc.h -- device state of the batched Monte Carlo chains (mc_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"
#include "kernels.h"

namespace mmo {

constexpr int kMcBlock = 100;    // Params.block_size, src/params.ml:29
// energy form of a chain launch (one kernel instantiation each)
enum { kMcInterGrid = 0, kMcInterAtom = 1, kMcInterSub = 2, kMcInterF64 = 3 };
constexpr int kMcMaxRb = 32;     // rotatable bonds per ligand
enum { kMcRunning = 0, kMcDone = 1, kMcTooLong = 2 };
enum {
  kMcTweak = 1,               // torsion moves (tweak_rbonds)
  kMcHardRoi = 2,             // --hard-ROI
  kMcIntraUff = 4,            // UFF_NB E_intra (else the NULL intra force field: 0)
  kMcMetropolisAlways = 8,    // Metropolis without hard ROI (not the reference)
  kMcNoFlip = 16,             // --no-flip-move (src/lds.ml:1828)
  kMcFp64 = 32                // FP64 pair arithmetic for the direct E_inter (parity mode)
};

// SW.t of 100 events as a bit ring
struct SwState {
  uint32_t bits[4];
  int head, count, acc;
};

struct McState {
  double rot[9], pos[3];            // current pose
  double rot0[9], pos0[3];          // start pose (resets return here)
  double best_rot[9], best_pos[3];
  double center[3];                 // centre field of the conformer
  double max_rot, max_trans;
  double prev_e, prev_e_inter, prev_e_intra, best_e;
  SwState sw;                       // rigid-move acceptance window
  int frame, rigid_step, conf_step;
  int status, initialized;
  int accepted, resets, ooroi;
  int min_evals;                    // local minimisation: energy + gradient evaluations
};
static_assert(sizeof(McState) % 8 == 0, "McState is copied as 32-bit words into 8-aligned LDS slots");

struct McArgs {
  // ligands
  const LigInfo* ligs;
  const LigAtom* latoms;
  const double* lig_xyz;        // registered centred conformers (rows, at atom_base)
  const int* atom_order;        // nullable: per ligand at atom_base, the sweep's atom order
  const int* rb_left;
  const int* rb_right;
  const int* rg_off;
  const int* rg_atoms;
  const uint32_t* pair_ij;
  const double4* pair_par;
  int max_atoms;
  // chains
  int nwork;                 // chains to process in this launch
  const int* work;           // their indices, most expensive first
  const int* chain_lig;
  const int* chain_atom;     // first conformer row of each chain
  const int* chain_rb;       // first rbond slot of each chain
  McState* st;
  double* conf;              // current conformers (rows)
  double* best_xyz;          // best poses (rows)
  double* undo;              // per chain, like conf: a torsion proposal's moved atoms
  double* rb_dr;             // per chain and rbond: torsion step
  SwState* rb_sw;            // per chain and rbond: acceptance window
  int* queue;
  // moves
  int nsteps, total_frames, flags;
  uint32_t seed_lo, seed_hi;
  double beta, max_rot0, max_trans0;
  double roi_x, roi_y, roi_z, roi_r;
  // scoring
  int inter_mode;            // MMO_SHIFTED (direct, culled) or MMO_INTERP
  Frame fr;
  const RecPos* rpos;
  const RecAB* rab;
  const RecRes* rres;       // receptor quantisation residuals
  int nrec;
  int sweep;                 // 1: atom_sweep, 0: pose_sweep (direct sum, FP32)
  CellGrid cg;
  GridDev g;
  // local minimisation after every proposal (Monte Carlo-plus-minimisation): min_iters
  // L-BFGS evaluations over the proposal's DOFs; per chain the base rows (like conf),
  // history S, Y (min_m slots of 64 lanes) and rho (min_m)
  int min_iters, min_m;
  double min_max_step, min_min_step, min_gtol, min_c1;
  double* min_base;
  double* min_S;
  double* min_Y;
  double* min_rho;
};

// waves: chains per workgroup (one wavefront each); rec_lds: the workgroup first copies
// the receptor and its cell table to LDS (they must fit next to the conformers)
// wavefronts per workgroup the Monte Carlo kernel is compiled for (the launch bound)
int mc_max_waves(const McArgs& a);
hipError_t launch_mc(const McArgs& a, int nblocks, int waves, bool rec_lds, hipStream_t s);
size_t mc_lds_bytes(const McArgs& a, int waves, bool rec_lds);
hipError_t mc_prepare();

}  // namespace mmo
