// mc_kernels.hip -- batched Monte Carlo chains on gfx950, one wavefront per chain.
//
// Restates Lds.simulate_lig (src/lds.ml:741-1000) for many ligands and chains in one
// persistent launch: wavefronts take chains from a work queue (most expensive first)
// and advance each by `nsteps` frames; the chain state lives in HBM between launches,
// so a launch (or a hipGraph replay of it) continues where the last one stopped.
//
// Per frame (src/lds.ml:882-995):
//   even frames: rigid move, rot' = R_axis(theta) rot (Move.rand_rot src/move.ml:20-31,
//     passive axis matrices src/rot.ml:22-46), pos' = pos + dt U(-1,1)^3 (src/move.ml:46-54);
//   odd frames: torsion move of the conformer (turned in place, the moved atoms restored
//     from the chain's undo rows unless the move is kept), tweak U(-dr_i, dr_i) or, every
//     100th conformer step, flip U(-pi, pi) (src/lds.ml:781-798, src/mol.ml:635-647);
//     enclosing radius > 12 A (Mol.check_elongation_exn) ends the chain (the exception
//     leaves the frame loop, src/lds.ml:996-998);
//   pose x = rot' (conf' - centre(conf')) + pos' (src/mol.ml:705-710);
//   E_intra (odd frames) and E_inter are stored before the acceptance test
//   (src/lds.ml:900-904);
//   hard ROI: a pose centre beyond the ROI radius resets the chain; otherwise
//   E_inter == 0 resets it and Metropolis at beta = 1/(kB T) decides (src/lds.ml:906-954).
//   OCaml attaches the `else` of src/lds.ml:917 to the inner `if dist_roi > out_radius`,
//   so without hard ROI no move is ever accepted; MMO_MC_METROPOLIS_ALWAYS runs the
//   E == 0 / Metropolis branch in that case as well (not the reference's behaviour).
//   Best pose, adaptive steps (x0.95 / x1.05 outside the [0.45, 0.55] acceptance band
//   over 100-event windows, src/lds.ml:586-621, :971-988).
// Random numbers: Philox4x32-10 keyed by the seed, counter (chain, frame, draw).
#include "common.h"
#include "cull.h"
#include "dofmin.h"
#include "kernels.h"
#include "mc.h"

namespace mmo {

// the chains' atom sweep fetches its ligand atoms 64 at a time (one per lane) and
// broadcasts them, rather than one dependent fetch per atom
constexpr bool kMcPreload = true;
// ligand atoms per pass of the chains' per-atom sweep
constexpr int kMcSweepAtoms = 4;

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC 2011).
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}
// two uniform doubles in [0, 1) with 53 random bits each
__device__ __forceinline__ void uniform2(uint4 r, double& a, double& b) {
  a = ((double)(r.x >> 5) * 67108864.0 + (double)(r.y >> 6)) * (1.0 / 9007199254740992.0);
  b = ((double)(r.z >> 5) * 67108864.0 + (double)(r.w >> 6)) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------------ sliding window
// SW.t (src/SW.ml:7-42): last 100 accept/reject events; ratio = accepts / events.
__device__ __forceinline__ void sw_process(SwState& w, bool ev) {
  const int word = w.head >> 5, bit = w.head & 31;
  if (w.count == kMcBlock) {
    if ((w.bits[word] >> bit) & 1u) --w.acc;
  } else {
    ++w.count;
  }
  if (ev) { w.bits[word] |= 1u << bit; ++w.acc; }
  else w.bits[word] &= ~(1u << bit);
  w.head = (w.head + 1 == kMcBlock) ? 0 : w.head + 1;
}
__device__ __forceinline__ double sw_ratio(const SwState& w) {
  return (double)w.acc / (double)w.count;   // 0/0 = NaN: no adaptation (comparisons false)
}
__device__ __forceinline__ void sw_reset(SwState& w) {
  w.bits[0] = w.bits[1] = w.bits[2] = w.bits[3] = 0u;
  w.head = w.count = w.acc = 0;
}

// Rot.of_axis_angle (src/rot.ml:121-146)
__device__ __forceinline__ void axis_angle(double x, double y, double z, double th, double* m) {
  double s, c;
  sincos(th, &s, &c);
  const double o = 1.0 - c;
  m[0] = c + x * x * o;     m[1] = x * y * o - z * s; m[2] = x * z * o + y * s;
  m[3] = x * y * o + z * s; m[4] = c + y * y * o;     m[5] = y * z * o - x * s;
  m[6] = x * z * o - y * s; m[7] = y * z * o + x * s; m[8] = c + z * z * o;
}

// --------------------------------------------------------------- energies of a pose
// The pose being scored (rotation, conformer centre, position) and the frame's Metropolis
// uniform, in the chain's LDS slot: the energy sweeps read it from there, so none of it
// is held in registers (and spilled to scratch) while a sweep runs.
struct McPose {
  double m[9], c[3], p[3], u;
};
static_assert(sizeof(McPose) % 16 == 0, "McPose slots keep 16-byte alignment");

// pose atom j: rot (conf_j - centre) + pos
struct PoseView {
  const double* sx;
  const double* sy;
  const double* sz;
  const McPose* q;
  __device__ void at(int j, double& x, double& y, double& z) const {
    const double a = sx[j] - q->c[0], b = sy[j] - q->c[1], c = sz[j] - q->c[2];
    x = (q->m[0] * a + q->m[1] * b + q->m[2] * c) + q->p[0];
    y = (q->m[3] * a + q->m[4] * b + q->m[5] * c) + q->p[1];
    z = (q->m[6] * a + q->m[7] * b + q->m[8] * c) + q->p[2];
  }
};

__device__ double mc_inter_interp(const McArgs& a, const LigInfo& li, const PoseView& v) {
  const int lane = threadIdx.x & (kWave - 1);
  double e = 0.0;
  int bad = 0;
  const GridDev& g = a.g;
  for (int j = lane; j < li.n_atoms; j += kWave) {
    double x, y, z;
    v.at(j, x, y, z);
    const double px = x - g.ox, py = y - g.oy, pz = z - g.oz;
    const double fx = px * g.ods, fy = py * g.ods, fz = pz * g.ods;
    if (!(px >= 0.0 && py >= 0.0 && pz >= 0.0 && fx < (double)(g.nx - 1) &&
          fy < (double)(g.ny - 1) && fz < (double)(g.nz - 1))) {
      bad = 1;
      continue;
    }
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    float cv[8];
    grid_corners(g, a.latoms[li.atom_base + j].type, i0, j0, k0, cv);
    const double wlx = (px - i0 * g.step) * g.ods, wly = (py - j0 * g.step) * g.ods,
                 wlz = (pz - k0 * g.step) * g.ods;
    const double whx = 1.0 - wlx, why = 1.0 - wly, whz = 1.0 - wlz;
    // corner order of G3D.trilin, src/G3D.ml:144-151
    e += (double)cv[0] * (whx * why * whz) + (double)cv[1] * (wlx * why * whz) +
         (double)cv[2] * (wlx * wly * whz) + (double)cv[3] * (whx * wly * whz) +
         (double)cv[4] * (whx * why * wlz) + (double)cv[5] * (wlx * why * wlz) +
         (double)cv[6] * (wlx * wly * wlz) + (double)cv[7] * (whx * wly * wlz);
  }
  e = wave_sum(e);
  for (int off = 1; off < kWave; off <<= 1) bad |= __shfl_xor(bad, off, kWave);
  return bad ? __builtin_nan("") : e;
}

// shifted direct sum of the pose: FP32 pairs swept per 8-atom sub-block (pose_sweep);
// the FP64 parity mode sweeps with FP64 pair arithmetic, two atoms per lane
// INTER: kMcInterAtom / kMcInterSub (FP32, atom_sweep / pose_sweep) or kMcInterF64
template <int INTER>
__device__ double mc_inter_direct(const McArgs& a, const LigInfo& li, const PoseView& v,
                                  const RecPos* s_pos, const RecAB* s_ab, const int* s_cell,
                                  int* wl) {
  const Frame fr = a.fr;
  int sat = 0;
  if (INTER != kMcInterF64) {
    double e_el = 0.0, e_vd = 0.0;
    auto atom = [&](int j, LigPt& l, double& cq, double& ca, double& cb) {
      // sweep position j -> ligand atom (the per-atom sweep's order, if given)
      const int ja = (INTER == kMcInterAtom && a.atom_order) ? a.atom_order[li.atom_base + j] : j;
      double x, y, z;
      v.at(ja, x, y, z);
      l = lig_pt(x, y, z, fr, &sat);
      const LigAtom la = a.latoms[li.atom_base + ja];
      cq = la.keq; ca = la.a; cb = la.b;
    };
    auto none = [](int, double, double, double) {};
    if (INTER == kMcInterAtom)
      // pairs closer than 0.5 A add both atoms' quantisation residuals (without them a
      // clashing pose's energy, above 1e7 kcal/mol, is off by up to ~2e-5 relative)
      atom_sweep<false, decltype(atom)&, decltype(none)&, true, kMcPreload, kMcSweepAtoms>(
          a.cg, s_cell, s_pos, s_ab, a.rres, fr, li.n_atoms, wl, atom, none, e_el, e_vd);
    else
      pose_sweep<false>(a.cg, s_cell, s_pos, s_ab, a.rres, fr, li.n_atoms, atom, none, e_el, e_vd);
    return wave_sum(e_el + e_vd);
  }
  constexpr int K = 2;
  const int lane = threadIdx.x & (kWave - 1);
  double e_vd = 0.0;
  const int nblk = (li.n_atoms + kWave * K - 1) / (kWave * K);
  for (int blk = 0; blk < nblk; ++blk) {
    int lx[K], ly[K], lz[K];
    bool valid[K];
    double cq[K], ca[K], cb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int j = lane + kWave * (blk * K + k);
      valid[k] = j < li.n_atoms;
      if (valid[k]) {
        double x, y, z;
        v.at(j, x, y, z);
        lx[k] = to_fixed(x, fr.ox, fr.scale, fr.lim, &sat);
        ly[k] = to_fixed(y, fr.oy, fr.scale, fr.lim, &sat);
        lz[k] = to_fixed(z, fr.oz, fr.scale, fr.lim, &sat);
        const LigAtom la = a.latoms[li.atom_base + j];
        cq[k] = la.keq; ca[k] = la.a; cb[k] = la.b;
      } else {
        lx[k] = ly[k] = lz[k] = 0;
        cq[k] = ca[k] = cb[k] = 0.0;
      }
    }
    float qx, qy, qz, R;
    if (!wave_query<K>(lx, ly, lz, valid, fr.inv_s, qx, qy, qz, R)) continue;
    e_vd += cull_sweep_f64<K>(a.cg, s_cell, s_pos, s_ab, 1.0 / fr.scale, qx, qy, qz, R, lx, ly,
                              lz, cq, ca, cb);
  }
  return wave_sum(e_vd);
}

// E_intra of the conformer in LDS (rigid-motion invariant, so the pose is not needed);
// src/mol.ml:881-903 over the precomputed pair list, FP64
__device__ double mc_intra(const McArgs& a, const LigInfo& li, const double* sx, const double* sy,
                           const double* sz) {
  const int lane = threadIdx.x & (kWave - 1);
  const int n = li.n_pairs;
  const uint32_t* pij = a.pair_ij + li.pair_base;
  const double4* ppar = a.pair_par + li.pair_base;
  double e = 0.0;
  // four pairs per lane per pass, their loads issued together (one memory latency per
  // four pairs); each lane still sums its pairs lane, lane + 64, ... in that order
  for (int p0 = lane; p0 < n; p0 += 4 * kWave) {
    uint32_t ij[4];
    double4 pp[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int p = p0 + u * kWave;
      ij[u] = p < n ? pij[p] : 0u;
      pp[u] = p < n ? ppar[p] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (p0 + u * kWave >= n) break;
      const int i = ij[u] & 0xffff, j = ij[u] >> 16;
      const double dx = sx[i] - sx[j], dy = sy[i] - sy[j], dz = sz[i] - sz[j];
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double irr = rsqrt(fmax(r2, 1.0e-4));   // no branch around the refinement
      const double ir = r2 >= 1.0e-4 ? irr : 100.0;   // 1 / max(r, 0.01)
      const double uu = ir * ir, u3 = uu * uu * uu, u6 = u3 * u3;
      e += pp[u].x * ir + pp[u].y * u6 - pp[u].z * u3;
    }
  }
  return wave_sum(e);
}

struct RecView {
  const RecPos* pos;
  const RecAB* ab;
  const int* cell;
  int* wl;                        // this wavefront's atom_sweep scratch (LDS)
};

template <int INTER>
__device__ double mc_inter(const McArgs& a, const LigInfo& li, const PoseView& v, const RecView& r) {
  if (INTER == kMcInterGrid) return mc_inter_interp(a, li, v);
  return mc_inter_direct<INTER>(a, li, v, r.pos, r.ab, r.cell, r.wl);
}

// LDS writes of one wavefront made visible to its other lanes (the chain's conformer
// buffers belong to one wavefront; the workgroup's other wavefronts run other chains)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ------------------------------------------------- Monte Carlo-plus-minimisation
// LDS of one chain's minimiser (after the chain slots): the minimised conformer and the
// adjoint, 6 x max_atoms FP64
struct MinLds {
  double *tx, *ty, *tz, *hx, *hy, *hz;
};

// Local minimisation of the proposal (conformer sx.. with centre c, rotation nrot,
// position npos): the base rows are the proposal pose centred at the origin,
// b_j = nrot (x_j - c), so that the DOF vector [npos, 0, 0, 0, 0..] is the proposal
// itself; min_iters L-BFGS evaluations of E_inter + E_intra (the chain's energy form)
// over [t, alpha, beta, gamma, tau_0..] (torsions only for flexible ligands).  Returns
// the energy of the best point and leaves its conformer (before the rigid motion) in
// m.tx.., its rigid motion in M, T and the pose centre (M mean(conformer) + T) in cpos;
// E_inter / E_intra of that point go to the chain state.
template <int INTER>
__device__ double mc_minimize(const McArgs& a, const LigInfo& li, int chain, bool flex,
                              bool use_intra, const PoseView& v, const RecView& rv, const MinLds& m,
                              McState& st, double* M, double* T, double* cpos) {
  const int lane = threadIdx.x & (kWave - 1);
  const int L = li.n_atoms;
  double* gbase = a.min_base + 3 * (size_t)a.chain_atom[chain];
  for (int j = lane; j < L; j += kWave) {   // lane j later reads back the rows it wrote
    double x, y, z;
    v.at(j, x, y, z);
    gbase[3 * j] = x - v.q->p[0]; gbase[3 * j + 1] = y - v.q->p[1]; gbase[3 * j + 2] = z - v.q->p[2];
  }
  DofEval ev;
  ev.li = li; ev.latoms = a.latoms; ev.atom_order = a.atom_order; ev.rb_left = a.rb_left;
  ev.rb_right = a.rb_right; ev.rg_off = a.rg_off; ev.rg_atoms = a.rg_atoms;
  ev.pair_ij = a.pair_ij; ev.pair_par = a.pair_par; ev.ntors = flex ? li.nrb : 0;
  ev.intra = use_intra; ev.fr = a.fr; ev.s_pos = rv.pos; ev.s_ab = rv.ab; ev.s_cell = rv.cell;
  ev.rres = a.rres; ev.cg = a.cg; ev.wl = rv.wl; ev.g = a.g;
  const int n = 6 + ev.ntors;
  const bool act = lane < n;
  const MinCfg cfg{a.min_m, a.min_max_step, a.min_min_step, a.min_gtol, a.min_c1};
  double* S = a.min_S + (size_t)chain * a.min_m * kWave + lane;
  double* Y = a.min_Y + (size_t)chain * a.min_m * kWave + lane;
  double* rho = a.min_rho + (size_t)chain * a.min_m;
  MinPose s{};
  double xt = lane == 0 ? v.q->p[0] : lane == 1 ? v.q->p[1] : lane == 2 ? v.q->p[2] : 0.0;
  double x = 0.0, g = 0.0, d = 0.0, e_inter = 0.0, e_intra = 0.0;
  for (int it = 0; it < a.min_iters && __builtin_amdgcn_readfirstlane(s.status) == kMinRunning; ++it) {
    double gt, ei, ea, rad;
    const double f = dof_eval<INTER>(ev, gbase, xt, m.tx, m.ty, m.tz, m.hx, m.hy, m.hz, gt, ei, ea, rad);
    const int acc0 = s.n_acc, init0 = s.init;
    xt = lbfgs_wave(s, x, g, d, xt, f, act ? gt : 0.0, rad, -__builtin_inf(), __builtin_inf(), act,
                    S, Y, (size_t)kWave, rho, 1, cfg);
    if (!init0 || s.n_acc != acc0) { e_inter = ei; e_intra = ea; }   // the accepted point
  }
  st.min_evals += s.n_eval;
  st.prev_e_inter = e_inter;
  st.prev_e_intra = e_intra;
  // the best point's conformer and rigid motion
  double sn, cs;
  dof_pose(ev, gbase, x, m.tx, m.ty, m.tz, M, nullptr, nullptr, nullptr, T, sn, cs);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int j = lane; j < L; j += kWave) { s0 += m.tx[j]; s1 += m.ty[j]; s2 += m.tz[j]; }
  const double cx = wave_sum(s0) / L, cy = wave_sum(s1) / L, cz = wave_sum(s2) / L;
  cpos[0] = (M[0] * cx + M[1] * cy + M[2] * cz) + T[0];
  cpos[1] = (M[3] * cx + M[4] * cy + M[5] * cz) + T[1];
  cpos[2] = (M[6] * cx + M[7] * cy + M[8] * cz) + T[2];
  return s.f;
}

// --------------------------------------------------------------------- the kernel
// A workgroup of `waves` wavefronts, each running its own chains; with REC_LDS the
// workgroup first stages the receptor and its cell table in LDS, shared by all chains.
// one instantiation per energy form, so that each carries only its own registers
// Wavefronts per workgroup each form is compiled for.  16 caps every form at 128 VGPRs
// (4 wavefronts per SIMD; the chain's control code then spills ~260-340 B per lane to
// scratch): measured faster than 8 (about 200 VGPRs, 2 per SIMD) for the direct forms
// (15.4 M against 14.2 M steps/s) and for the grid form (225 M against 203 M).
__host__ __device__ constexpr int mc_form_waves(int, bool min) { return min ? 8 : 16; }

template <bool REC_LDS, int INTER, bool MIN>
__global__ void __launch_bounds__(mc_form_waves(INTER, MIN) * 64) k_mc(McArgs a, int rec_bytes) {
  extern __shared__ __align__(16) char mc_lds[];
  RecView rv;
  load_receptor<REC_LDS>(a.rpos, a.rab, a.nrec, a.cg, rv.pos, rv.ab, rv.cell, mc_lds);
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  // the chain's conformer; a torsion proposal turns it in place, with the moved atoms'
  // previous coordinates in the chain's undo rows (HBM) until the decision
  double* cx_ = reinterpret_cast<double*>(mc_lds + rec_bytes) + (size_t)wv * 3 * a.max_atoms;
  double* cy_ = cx_ + a.max_atoms;
  double* cz_ = cy_ + a.max_atoms;
  const int nwv = blockDim.x / kWave;
  McState* s_st = reinterpret_cast<McState*>(
      mc_lds + rec_bytes + (((size_t)nwv * 3 * a.max_atoms * sizeof(double) + 15) & ~(size_t)15));
  McPose* pose = reinterpret_cast<McPose*>(s_st + nwv) + wv;   // this chain's pose slot
  rv.wl = reinterpret_cast<int*>(reinterpret_cast<McPose*>(s_st + nwv) + nwv) + wv * kAtomSweepInts;
  MinLds ml{};
  if (MIN) {
    // after the scratch of every wavefront: 6 x max_atoms FP64 per wavefront
    double* b = reinterpret_cast<double*>(
                    reinterpret_cast<int*>(reinterpret_cast<McPose*>(s_st + nwv) + nwv) +
                    nwv * kAtomSweepInts) +
                (size_t)wv * 6 * a.max_atoms;
    ml = MinLds{b, b + a.max_atoms, b + 2 * a.max_atoms, b + 3 * a.max_atoms,
                b + 4 * a.max_atoms, b + 5 * a.max_atoms};
  }
  const double pi = 3.14159265358979311600;       // 4 atan 1 (src/math.ml:13)

  for (;;) {
    // the next chain: lane 0 takes it from the queue; readlane makes it a scalar, so the
    // loop exit is uniform
    int qv = 0;
    if (lane == 0) qv = atomicAdd(a.queue, 1);
    const int q = __builtin_amdgcn_readlane(qv, 0);
    if (q >= a.nwork) break;
    const int chain = a.work[q];
    // the chain's state in this wavefront's LDS slot (wave-uniform values)
    McState& st = s_st[wv];
    {
      const int* src = reinterpret_cast<const int*>(a.st + chain);
      int* dst = reinterpret_cast<int*>(&st);
      for (int k = lane; k < (int)(sizeof(McState) / 4); k += kWave) dst[k] = src[k];
    }
    wave_sync();
    if (st.status != kMcRunning) continue;
    const int lig = a.chain_lig[chain];
    const LigInfo li = a.ligs[lig];
    const int L = li.n_atoms, nrb = li.nrb;
    double* gconf = a.conf + 3 * (size_t)a.chain_atom[chain];
    double* gbest = a.best_xyz + 3 * (size_t)a.chain_atom[chain];
    double* gundo = a.undo + 3 * (size_t)a.chain_atom[chain];
    // group [g0, g1) of a torsion back to the coordinates saved in the undo rows (each
    // lane restores the atoms it saved)
    auto restore = [&](int g0, int g1) {
      for (int m = g0 + lane; m < g1; m += kWave) {
        const int at = a.rg_atoms[m];
        const double* u = gundo + 3 * (size_t)(m - g0);
        cx_[at] = u[0]; cy_[at] = u[1]; cz_[at] = u[2];
      }
      wave_sync();
    };
    double* dr = a.rb_dr + a.chain_rb[chain];
    SwState* rsw = a.rb_sw + a.chain_rb[chain];
    for (int j = lane; j < L; j += kWave) {
      cx_[j] = gconf[3 * j]; cy_[j] = gconf[3 * j + 1]; cz_[j] = gconf[3 * j + 2];
    }
    wave_sync();
    const bool flex = (a.flags & kMcTweak) && nrb > 0;
    const bool use_intra = (a.flags & kMcIntraUff) != 0;

    if (!st.initialized) {
      // start pose: Mol.rotate_then_translate_copy centered_lig rot0 pos0, no
      // re-centring (src/lds.ml:759, src/mol.ml:669-672); E of the start pose
      if (lane == 0) {
        for (int k = 0; k < 9; ++k) pose->m[k] = st.rot0[k];
        for (int k = 0; k < 3; ++k) { pose->c[k] = 0.0; pose->p[k] = st.pos0[k]; }
      }
      wave_sync();
      const PoseView v{cx_, cy_, cz_, pose};
      st.prev_e_intra = use_intra ? mc_intra(a, li, cx_, cy_, cz_) : 0.0;
      st.prev_e_inter = mc_inter<INTER>(a, li, v, rv);
      st.prev_e = st.prev_e_inter + st.prev_e_intra;
      st.best_e = st.prev_e;
      for (int k = 0; k < 9; ++k) { st.rot[k] = st.rot0[k]; st.best_rot[k] = st.rot0[k]; }
      for (int k = 0; k < 3; ++k) { st.pos[k] = st.pos0[k]; st.best_pos[k] = st.pos0[k]; }
      for (int j = lane; j < L; j += kWave) {
        double x, y, z;
        v.at(j, x, y, z);
        gbest[3 * j] = x; gbest[3 * j + 1] = y; gbest[3 * j + 2] = z;
      }
      st.initialized = 1;
    }

    const int frame_end = min(st.frame + a.nsteps, a.total_frames);
    for (; st.frame < frame_end; ++st.frame) {
      const bool rigid = (st.frame & 1) == 0;
      const uint4 r0 = philox(make_uint4((uint32_t)chain, (uint32_t)st.frame, 0u, 0x4D4D4F31u), a.seed_lo, a.seed_hi);
      const uint4 r1 = philox(make_uint4((uint32_t)chain, (uint32_t)st.frame, 1u, 0x4D4D4F31u), a.seed_lo, a.seed_hi);
      const uint4 r2 = philox(make_uint4((uint32_t)chain, (uint32_t)st.frame, 2u, 0x4D4D4F31u), a.seed_lo, a.seed_hi);
      double u0, u1, u2, u3, u4, u5;
      uniform2(r0, u0, u1);
      uniform2(r1, u2, u3);
      uniform2(r2, u4, u5);

      // ---- proposal
      double nrot[9], npos[3], ccx, ccy, ccz;
      const double* sx = cx_;
      const double* sy = cy_;
      const double* sz = cz_;
      int just_rotated = -1;
      int mg0 = 0, mg1 = 0;   // torsion group turned in place by this proposal
      ccx = st.center[0]; ccy = st.center[1]; ccz = st.center[2];
      if (!rigid) {
        for (int k = 0; k < 9; ++k) nrot[k] = st.rot[k];
        for (int k = 0; k < 3; ++k) npos[k] = st.pos[k];
        if (flex) {
          // Mol.flip_rbond / tweak_rbond on a copy (src/lds.ml:784-795, src/mol.ml:635-647)
          int i = (int)(u0 * nrb);
          i = i < nrb ? i : nrb - 1;
          const bool flip = st.conf_step > 0 && st.conf_step % kMcBlock == 0 && !(a.flags & kMcNoFlip);
          const double amp = flip ? pi : dr[i];
          const double ang = u1 * (2.0 * amp) - amp;   // Move.rand_angle, src/move.ml:14-16
          const int left = a.rb_left[li.rb_base + i], right = a.rb_right[li.rb_base + i];
          const double ox = cx_[right], oy = cy_[right], oz = cz_[right];
          const double vx = ox - cx_[left], vy = oy - cy_[left], vz = oz - cz_[left];
          const double mag = sqrt(vx * vx + vy * vy + vz * vz);
          double Rm[9];
          axis_angle(vx / mag, vy / mag, vz / mag, ang, Rm);
          wave_sync();
          const int g0 = a.rg_off[li.rg_base + i], g1 = a.rg_off[li.rg_base + i + 1];
          for (int m = g0 + lane; m < g1; m += kWave) {
            const int at = a.rg_atoms[m];
            const double x0 = cx_[at], y0 = cy_[at], z0 = cz_[at];
            double* u = gundo + 3 * (size_t)(m - g0);
            u[0] = x0; u[1] = y0; u[2] = z0;
            const double zx = x0 - ox, zy = y0 - oy, zz = z0 - oz;
            cx_[at] = (Rm[0] * zx + Rm[1] * zy + Rm[2] * zz) + ox;
            cy_[at] = (Rm[3] * zx + Rm[4] * zy + Rm[5] * zz) + oy;
            cz_[at] = (Rm[6] * zx + Rm[7] * zy + Rm[8] * zz) + oz;
          }
          wave_sync();
          // update_center (A.favg) and enclosing radius (src/mol.ml:353-356, :576-591)
          double s0 = 0.0, s1 = 0.0, s2 = 0.0;
          for (int j = lane; j < L; j += kWave) { s0 += cx_[j]; s1 += cy_[j]; s2 += cz_[j]; }
          ccx = wave_sum(s0) / L; ccy = wave_sum(s1) / L; ccz = wave_sum(s2) / L;
          double rmax = 0.0;
          for (int j = lane; j < L; j += kWave) {
            const double dx = ccx - cx_[j], dy = ccy - cy_[j], dz = ccz - cz_[j];
            rmax = fmax(rmax, 0.01 + sqrt(dx * dx + dy * dy + dz * dz));
          }
          rmax = wave_max(rmax);
          if (rmax > 12.0) {   // Mol.Too_long: the chain ends (src/lds.ml:794, :997)
            restore(g0, g1);   // with its last accepted conformer
            st.status = kMcTooLong;
            break;
          }
          mg0 = g0;
          mg1 = g1;
          just_rotated = i;
        }
      } else {
        // Move.rand_rot: left-multiply the passive axis rotation Rot.rx / ry / rz
        // (src/move.ml:20-31, src/rot.ml:22-46).  Two rows of rot mix,
        // p' = c p + s q and q' = c q - s p, with (p, q) = rows (1, 2) for x, (2, 0) for y,
        // (0, 1) for z.  Written with selects: a switch that built the 3x3 matrix per axis
        // was mis-lowered by hipcc (the -s entry of the z matrix came out 0).
        const double th = u0 * (2.0 * st.max_rot) - st.max_rot;
        int ax = (int)(u1 * 3.0);
        ax = ax < 3 ? ax : 2;
        double s, c;
        sincos(th, &s, &c);
        const bool ax0 = ax == 0, ax1 = ax == 1;
        for (int cc = 0; cc < 3; ++cc) {
          const double r0 = st.rot[cc], r1 = st.rot[3 + cc], r2 = st.rot[6 + cc];
          const double p = ax0 ? r1 : (ax1 ? r2 : r0);
          const double q = ax0 ? r2 : (ax1 ? r0 : r1);
          const double pn = c * p + s * q, qn = c * q - s * p;
          nrot[cc] = ax0 ? r0 : (ax1 ? qn : pn);
          nrot[3 + cc] = ax0 ? pn : (ax1 ? r1 : qn);
          nrot[6 + cc] = ax0 ? qn : (ax1 ? pn : r2);
        }
        // Move.rand_trans (src/move.ml:46-54)
        npos[0] = st.pos[0] + (u2 * 2.0 - 1.0) * st.max_trans;
        npos[1] = st.pos[1] + (u3 * 2.0 - 1.0) * st.max_trans;
        npos[2] = st.pos[2] + (u4 * 2.0 - 1.0) * st.max_trans;
      }

      // the proposal's pose and the frame's Metropolis uniform go to the chain's LDS slot;
      // from here on they are read from there (the memory clobber below keeps the
      // compiler from carrying them in registers across the energy sweep)
      if (lane == 0) {
        for (int k = 0; k < 9; ++k) pose->m[k] = nrot[k];
        pose->c[0] = ccx; pose->c[1] = ccy; pose->c[2] = ccz;
        for (int k = 0; k < 3; ++k) pose->p[k] = npos[k];
        pose->u = u5;
      }
      wave_sync();
      asm volatile("" ::: "memory");

      // ---- energies (stored before the acceptance test, src/lds.ml:900-904)
      const PoseView v{sx, sy, sz, pose};
      double curr_e, cpos[3];
      double mM[9], mT[3];   // MIN: rigid motion of the minimised pose
      if constexpr (MIN) {
        // the proposal, locally minimised; its centre is the pose centre of the ROI test
        curr_e = mc_minimize<INTER>(a, li, chain, flex, use_intra, v, rv, ml, st, mM, mT, cpos);
      } else {
        if (!rigid && use_intra && flex) st.prev_e_intra = mc_intra(a, li, sx, sy, sz);
        st.prev_e_inter = mc_inter<INTER>(a, li, v, rv);
        curr_e = st.prev_e_inter + st.prev_e_intra;
      }
      asm volatile("" ::: "memory");
      const McPose& P = *pose;
      if constexpr (!MIN)
        for (int k = 0; k < 3; ++k) cpos[k] = P.p[k];
      const double dxr = a.roi_x - cpos[0], dyr = a.roi_y - cpos[1], dzr = a.roi_z - cpos[2];
      const double dist_roi = sqrt(dxr * dxr + dyr * dyr + dzr * dzr);
      const bool hard = (a.flags & kMcHardRoi) != 0;
      bool reset = false, metropolis = false;
      if (hard) {
        if (dist_roi > a.roi_r) { reset = true; ++st.ooroi; }
        else metropolis = true;
      } else {
        metropolis = (a.flags & kMcMetropolisAlways) != 0;
      }
      if (metropolis && st.prev_e_inter == 0.0) { reset = true; metropolis = false; }
      if (reset) {
        // reset_run_params (src/lds.ml:632-648): the conformer is kept; the best pose
        // becomes the registered conformer at the start pose (best_lig := prev_lig :=
        // rotate_then_translate_copy centered_lig rot0 pos0, src/lds.ml:645-646)
        st.max_rot = a.max_rot0;
        st.max_trans = a.max_trans0;
        for (int k = 0; k < 9; ++k) { st.rot[k] = st.rot0[k]; st.best_rot[k] = (k % 4 == 0) ? 1.0 : 0.0; }
        for (int k = 0; k < 3; ++k) { st.pos[k] = st.pos0[k]; st.best_pos[k] = 0.0; }
        {
          const double* c0 = a.lig_xyz + 3 * (size_t)li.atom_base;
          const double* R0 = st.rot0;
          const double* p0 = st.pos0;
          for (int j = lane; j < L; j += kWave) {
            const double x0 = c0[3 * j], y0 = c0[3 * j + 1], z0 = c0[3 * j + 2];
            gbest[3 * j] = (R0[0] * x0 + R0[1] * y0 + R0[2] * z0) + p0[0];
            gbest[3 * j + 1] = (R0[3] * x0 + R0[4] * y0 + R0[5] * z0) + p0[1];
            gbest[3 * j + 2] = (R0[6] * x0 + R0[7] * y0 + R0[8] * z0) + p0[2];
          }
        }
        st.prev_e = __builtin_inf();
        st.best_e = __builtin_inf();
        sw_reset(st.sw);
        ++st.resets;
      } else if (metropolis) {
        // u5 is a multiple of 2^-53, so once positive it exceeds exp(-x) for every x >= 40
        // (exp(-40) < 2^-53): those proposals are rejected without evaluating exp
        const double x = (curr_e - st.prev_e) * a.beta;
        const double u5p = P.u;
        const bool accept =
            (curr_e <= st.prev_e) || (!(x >= 40.0 && u5p > 0.0) && u5p < exp(-x));
        if (rigid) sw_process(st.sw, accept);
        else if (just_rotated > -1) sw_process(rsw[just_rotated], accept);
        if (accept) {
          ++st.accepted;
          st.prev_e = curr_e;
          if constexpr (MIN) {
            // the minimised pose: conformer nrot^T c* (the chain's frame), centre nrot^T
            // mean(c*), rotation M* nrot, position its pose centre
            double cm[3] = {0.0, 0.0, 0.0};
            for (int j = lane; j < L; j += kWave) {
              const double x0 = ml.tx[j], y0 = ml.ty[j], z0 = ml.tz[j];
              cx_[j] = P.m[0] * x0 + P.m[3] * y0 + P.m[6] * z0;
              cy_[j] = P.m[1] * x0 + P.m[4] * y0 + P.m[7] * z0;
              cz_[j] = P.m[2] * x0 + P.m[5] * y0 + P.m[8] * z0;
              cm[0] += x0; cm[1] += y0; cm[2] += z0;
            }
            for (int k = 0; k < 3; ++k) cm[k] = wave_sum(cm[k]) / L;
            for (int k = 0; k < 3; ++k)
              st.center[k] = P.m[k] * cm[0] + P.m[3 + k] * cm[1] + P.m[6 + k] * cm[2];
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc)
                st.rot[3 * r + cc] = mM[3 * r] * P.m[cc] + mM[3 * r + 1] * P.m[3 + cc] +
                                     mM[3 * r + 2] * P.m[6 + cc];
            for (int k = 0; k < 3; ++k) st.pos[k] = cpos[k];
            mg0 = mg1 = 0;   // conf holds the minimised conformer: nothing to restore
            wave_sync();
          } else {
            for (int k = 0; k < 9; ++k) st.rot[k] = P.m[k];
            for (int k = 0; k < 3; ++k) st.pos[k] = P.p[k];
            if (mg1 > mg0) {   // conf := conf' (already in place): keep it
              for (int k = 0; k < 3; ++k) st.center[k] = P.c[k];
              mg0 = mg1 = 0;
            }
          }
        }
        if (curr_e < st.best_e) {   // best pose, accepted or not (src/lds.ml:959-969)
          st.best_e = curr_e;
          if constexpr (MIN) {
            for (int r = 0; r < 3; ++r)
              for (int cc = 0; cc < 3; ++cc)
                st.best_rot[3 * r + cc] = mM[3 * r] * P.m[cc] + mM[3 * r + 1] * P.m[3 + cc] +
                                          mM[3 * r + 2] * P.m[6 + cc];
            for (int k = 0; k < 3; ++k) st.best_pos[k] = cpos[k];
            for (int j = lane; j < L; j += kWave) {
              const double x0 = ml.tx[j], y0 = ml.ty[j], z0 = ml.tz[j];
              gbest[3 * j] = (mM[0] * x0 + mM[1] * y0 + mM[2] * z0) + mT[0];
              gbest[3 * j + 1] = (mM[3] * x0 + mM[4] * y0 + mM[5] * z0) + mT[1];
              gbest[3 * j + 2] = (mM[6] * x0 + mM[7] * y0 + mM[8] * z0) + mT[2];
            }
          } else {
            for (int k = 0; k < 9; ++k) st.best_rot[k] = P.m[k];
            for (int k = 0; k < 3; ++k) st.best_pos[k] = P.p[k];
            for (int j = lane; j < L; j += kWave) {
              double x, y, z;
              v.at(j, x, y, z);
              gbest[3 * j] = x; gbest[3 * j + 1] = y; gbest[3 * j + 2] = z;
            }
          }
        }
        // adaptive rigid steps every 100 rigid frames (src/lds.ml:971-974, :586-600)
        if (rigid && st.rigid_step > 0 && st.rigid_step % kMcBlock == 0) {
          const double ar = sw_ratio(st.sw);
          if (ar <= 0.45) { st.max_trans *= 0.95; st.max_rot *= 0.95; }
          else if (ar >= 0.55) { st.max_trans *= 1.05; st.max_rot = fmin(pi, 1.05 * st.max_rot); }
        }
        // adaptive torsion steps on the proposal's copy (src/lds.ml:976-979, :603-621):
        // kept only when the proposal was accepted (conf := conf')
        if (accept && flex && !rigid && st.conf_step > 0 && st.conf_step % (kMcBlock * nrb) == 0) {
          for (int k = 0; k < nrb; ++k) {
            const double ar = sw_ratio(rsw[k]);
            if (ar <= 0.45) dr[k] = 0.95 * dr[k];
            else if (ar >= 0.55) dr[k] = fmin(pi, 1.05 * dr[k]);
          }
        }
      }
      if (mg1 > mg0) restore(mg0, mg1);   // the proposal was not taken: conf stays
      if (rigid) ++st.rigid_step;
      else ++st.conf_step;
    }
    if (st.status == kMcRunning && st.frame >= a.total_frames) st.status = kMcDone;
    // write the chain back (lane 0 for the scalars, all lanes for the conformer)
    for (int j = lane; j < L; j += kWave) {
      gconf[3 * j] = cx_[j]; gconf[3 * j + 1] = cy_[j]; gconf[3 * j + 2] = cz_[j];
    }
    wave_sync();
    {
      const int* src = reinterpret_cast<const int*>(&st);
      int* dst = reinterpret_cast<int*>(a.st + chain);
      for (int k = lane; k < (int)(sizeof(McState) / 4); k += kWave) dst[k] = src[k];
    }
    wave_sync();
  }
}

__global__ void k_mc_reset_queue(int* q) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *q = 0;
}

size_t mc_lds_bytes(const McArgs& a, int waves, bool rec_lds) {
  size_t rec = 0;
  if (rec_lds)
    rec = ((size_t)a.nrec * (sizeof(RecPos) + sizeof(RecAB)) +
           (size_t)(a.cg.nx * a.cg.ny * a.cg.nz + 1) * sizeof(int) + 15) & ~(size_t)15;
  const size_t conf = ((size_t)waves * 3 * a.max_atoms * sizeof(double) + 15) & ~(size_t)15;
  const size_t mins = a.min_iters > 0 ? (size_t)waves * 6 * a.max_atoms * sizeof(double) : 0;
  return rec + conf + (size_t)waves * (sizeof(McState) + sizeof(McPose) + kAtomSweepInts * sizeof(int)) +
         mins;
}

template <bool R, int I>
static hipError_t mc_allow_lds(int cap) {
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mc<R, I, false>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  if (e == hipSuccess && I != kMcInterSub)
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mc<R, I, I != kMcInterSub>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, cap);
  return e;
}

hipError_t mc_prepare() {
  // dynamic LDS above 64 KiB must be allowed per kernel (gfx950: 160 KiB per CU); set
  // once, outside any stream capture
  const int cap = 160 * 1024;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = mc_allow_lds<true, kMcInterAtom>(cap);
  if (e == hipSuccess) e = mc_allow_lds<true, kMcInterSub>(cap);
  if (e == hipSuccess) e = mc_allow_lds<true, kMcInterF64>(cap);
  if (e == hipSuccess) e = mc_allow_lds<false, kMcInterAtom>(cap);
  if (e == hipSuccess) e = mc_allow_lds<false, kMcInterSub>(cap);
  if (e == hipSuccess) e = mc_allow_lds<false, kMcInterF64>(cap);
  if (e == hipSuccess) e = mc_allow_lds<false, kMcInterGrid>(cap);
  return e;
}

int mc_inter_kind(const McArgs& a) {
  if (a.inter_mode == MMO_INTERP) return kMcInterGrid;
  if (a.flags & kMcFp64) return kMcInterF64;
  // the minimiser's gradients come from the per-atom sweep
  return (a.sweep == 1 || a.min_iters > 0) ? kMcInterAtom : kMcInterSub;
}

int mc_max_waves(const McArgs& a) { return mc_form_waves(mc_inter_kind(a), a.min_iters > 0); }

hipError_t launch_mc(const McArgs& a, int nblocks, int waves, bool rec_lds, hipStream_t s) {
  if (waves < 1 || waves > mc_max_waves(a)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_mc_reset_queue, dim3(1), dim3(1), 0, s, a.queue);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t lds = mc_lds_bytes(a, waves, rec_lds);
  const int rec_bytes = (int)(mc_lds_bytes(a, waves, rec_lds) - mc_lds_bytes(a, waves, false));
  const dim3 g(nblocks), b(waves * kWave);
  const int key = mc_inter_kind(a) + (rec_lds ? 8 : 0) + (a.min_iters > 0 ? 16 : 0);
  switch (key) {
    case kMcInterAtom + 8: hipLaunchKernelGGL((k_mc<true, kMcInterAtom, false>), g, b, lds, s, a, rec_bytes); break;
    case kMcInterSub + 8: hipLaunchKernelGGL((k_mc<true, kMcInterSub, false>), g, b, lds, s, a, rec_bytes); break;
    case kMcInterF64 + 8: hipLaunchKernelGGL((k_mc<true, kMcInterF64, false>), g, b, lds, s, a, rec_bytes); break;
    case kMcInterAtom: hipLaunchKernelGGL((k_mc<false, kMcInterAtom, false>), g, b, lds, s, a, 0); break;
    case kMcInterSub: hipLaunchKernelGGL((k_mc<false, kMcInterSub, false>), g, b, lds, s, a, 0); break;
    case kMcInterF64: hipLaunchKernelGGL((k_mc<false, kMcInterF64, false>), g, b, lds, s, a, 0); break;
    case kMcInterGrid: hipLaunchKernelGGL((k_mc<false, kMcInterGrid, false>), g, b, lds, s, a, 0); break;
    // with local minimisation after every proposal
    case kMcInterAtom + 24: hipLaunchKernelGGL((k_mc<true, kMcInterAtom, true>), g, b, lds, s, a, rec_bytes); break;
    case kMcInterF64 + 24: hipLaunchKernelGGL((k_mc<true, kMcInterF64, true>), g, b, lds, s, a, rec_bytes); break;
    case kMcInterAtom + 16: hipLaunchKernelGGL((k_mc<false, kMcInterAtom, true>), g, b, lds, s, a, 0); break;
    case kMcInterF64 + 16: hipLaunchKernelGGL((k_mc<false, kMcInterF64, true>), g, b, lds, s, a, 0); break;
    case kMcInterGrid + 16: hipLaunchKernelGGL((k_mc<false, kMcInterGrid, true>), g, b, lds, s, a, 0); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace mmo
