// api.hip -- the C ABI of libmmo_ff.so (declared in include/mmo_ff.h).
//
// Host responsibilities: validating inputs, the UFF per-element table, uploading the
// receptor as fixed-point positions + factorised LJ parameters, the ligand registry
// (topology, intra pair lists, rotation groups), planning pose tiles for batches, and
// launching the kernels of ff_kernels.hip / pose_kernels.hip on a HIP stream.
#include <hip/hip_runtime.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "mc.h"
#include "mmo_ff.h"

using namespace mmo;

namespace {

std::string g_last_error;

// ----------------------------------------------- no-allocation top-k (mmo_topk_dev_noalloc)
// Scratch of the top-k kernels in a file-scope device array instead of library
// allocations: radix state (3 x u64) at byte 0, histogram (256 x u32) at 256, counter
// (u32) at 1280, per-chunk counts (u32) from 1536.  Each GPU holds its own copy for the
// life of the process; host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.  Every call on one GPU reuses its copy: the kernels clear the
// state, histogram and counter themselves and write every chunk count before reading it.
constexpr size_t kTopkChunk = 4096;                       // energies per chunk count
constexpr size_t kTopkNoallocChunks = 32768;              // -> at most 134,217,728 energies
constexpr size_t kTopkNoallocBytes = 1536 + 4 * kTopkNoallocChunks;
static __device__ __align__(256) unsigned char scratch[kTopkNoallocBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};

// the calling thread's current device must be `device`
void scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  if (e == hipSuccess) g_scratch_ptr[device] = p;
}

struct Uff { double x, d; };

// src/UFF.ml:10-22
const std::map<int, Uff> kDefaultUff = {
    {0, {0.0, 0.0}},     {1, {2.886, 0.044}},  {6, {3.851, 0.105}},  {7, {3.660, 0.069}},
    {8, {3.500, 0.060}}, {9, {3.364, 0.050}},  {12, {3.021, 0.111}}, {15, {4.147, 0.305}},
    {16, {4.035, 0.274}}, {17, {3.947, 0.227}}, {35, {4.189, 0.251}}, {53, {4.500, 0.339}}};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t count) {
    if (count <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    const size_t c = std::max<size_t>(count, 1);
    hipError_t e = hipMalloc(&p, c * sizeof(T));
    if (e == hipSuccess) n = c;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

struct LigHost {
  int n;
  std::vector<double> xyz;   // centred conformer, rows
  std::vector<double> q;
  std::vector<int> anum, type;
  std::vector<uint32_t> pair_ij;
  std::vector<int> rb_left, rb_right, rg_off, rg_atoms;
  std::vector<int> sweep;   // order in which the per-atom sweeps take the atoms
};

}  // namespace

struct mmo_ctx {
  int device = 0;
  pid_t pid = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<int, Uff> uff = kDefaultUff;
  // receptor
  bool has_rec = false;
  int nrec = 0;
  Frame fr{};
  int fp_bits = 24;
  DevBuf<RecPos> rpos;
  DevBuf<RecAB> rab;
  DevBuf<RecRes> rres;
  // receptor cell list (culled shifted kernel)
  float cell_size = 4.0f;
  int row_clip = 4;           // scan/grid sweeps: clip rows longer than this to the x chord
  int cell_xsub = 3;          // pose/MC sweeps: cells cut n-fold along x (rows clipped finer)
  int pose_cull = 1;          // batches in the shifted form: one wavefront per pose, culled
  int pose_sweep = 1;         // 1: lanes over receptor atoms per ligand atom; 0: 8-atom sub-blocks
  int mc_sweep = 1;           // the same choice for the Monte Carlo direct sum
  bool cull_ok = false;
  int scan_cull = 5;
  int cull_k = 1;
  int mc_waves_per_cu = 16;
  int mc_wg_waves = 16;       // Monte Carlo chains (wavefronts) per workgroup
  int min_fused = 1;          // minimiser: all iterations in one persistent launch
  int min_f64 = 0;            // fused minimiser: FP64 pair arithmetic (parity mode)
  int min_waves = 16;         // fused minimiser: wavefronts per workgroup
  int min_rec_lds = 1;        // fused minimiser: receptor in LDS when >= 8 wavefronts fit
  int mc_rec_lds = 1;         // Monte Carlo direct form: receptor in LDS when it fits
  int cull_waves = 8;          // rigid scan: wavefronts per workgroup
  int pose_waves = 16;         // batches (one wavefront per pose): wavefronts per workgroup
  CellGrid cg{};              // scan and grid kernels: cubic cells
  DevBuf<int> d_cell;
  CellGrid cgp{};             // pose and Monte Carlo sweeps: cells cut along x, same atom order
  DevBuf<int> d_cellp;
  // rigid scan row lists (scan_cull 5, 6): cells of scan_cell A in y and z, cut scan_xsub
  // times along x, with a copy of the receptor sorted by that grid
  float scan_cell = 5.0f;
  int scan_xsub = 8;
  int scan_chunk = 4;
  int row_filter = 1;
  CellGrid cgs{};
  DevBuf<int> d_cells;
  DevBuf<RecPos> rpos_s;
  DevBuf<RecAB> rab_s;
  DevBuf<RecRes> rres_s;
  bool cullp_ok = false;
  // grids
  bool has_grid = false;
  int ntypes = 0;
  GridDev g{};
  DevBuf<float> grid;
  DevBuf<float> grid_cells;   // the grids packed by cell (GridDev::cells)
  int grid_pack = 1;          // pack the grids by cell when they are set (up to 4 GiB)
  // ligand registry
  std::vector<LigHost> ligs;
  bool lig_dirty = true;
  int lig_version = 0;         // bumped by every registry upload (device pointers may move)
  int max_atoms = 0;
  DevBuf<LigInfo> d_ligs;
  DevBuf<LigAtom> d_latoms;
  DevBuf<double> d_lig_xyz;
  DevBuf<uint32_t> d_pair_ij;
  DevBuf<double4> d_pair_par;
  DevBuf<int> d_sweep;          // per ligand at atom_base: the sweeps' atom order
  DevBuf<int> d_rb_left, d_rb_right, d_rg_off, d_rg_atoms;
  // scratch for the plain (one ligand) forms
  DevBuf<double> s_xyz, s_e, s_g, s_part;
  DevBuf<LigInfo> s_lig;
  DevBuf<LigAtom> s_latom;
  DevBuf<uint32_t> s_pij;
  DevBuf<double4> s_ppar;
  DevBuf<int> s_int;
  DevBuf<Tile> s_tile;
  DevBuf<double> s_scan;   // rotations + translations of the last scan
  DevBuf<int> s_tperm;
  // receptor on the host (rows x, y, z; atomic numbers) for host-side set-ups
  std::vector<double> rec_host;
  std::vector<int> rec_anum;
  std::vector<double> rec_q;
  // protein vdW volume bitmask (Lds.vdW_volume) for the rigid scan's clash filters
  bool has_mask = false;
  double mask_step = 0.0, mask_o[3] = {0.0, 0.0, 0.0};
  int mask_n[3] = {0, 0, 0};
  std::vector<uint32_t> mask_host;
  std::vector<uint8_t> skip_host;   // the last scan filter's per-translation skips
  // desolvation: the receptor's first solvent shell on a grid, its nodes' contributions
  bool has_desolv = false;
  double desolv_step = 0.0, desolv_o[3] = {0.0, 0.0, 0.0};
  int desolv_n[3] = {0, 0, 0};
  std::vector<uint32_t> desolv_shell;
  DevBuf<double> d_desolv;      // per grid node
  DevBuf<double4> d_rec64;      // receptor x, y, z, q (FP64)
  DevBuf<int> s_dvox;
  DevBuf<double4> s_datoms;
  DevBuf<double> s_dpart;
  DevBuf<uint32_t> d_mask;
  DevBuf<uint8_t> s_skip;
  DevBuf<uint64_t> s_tk_st;
  DevBuf<unsigned> s_tk_u;
  // filtered scan: keep flags and compacted slots of the scan's pose order, the kept
  // count, the compaction's scratch
  DevBuf<uint8_t> s_keep, s_select;
  DevBuf<int> s_list, s_nkept;
  // tuning
  int inter_k = 4, inter_waves = 4, inter_rtile = 2048;
  std::map<std::string, std::pair<hipEvent_t, hipEvent_t>> ev;
  std::map<std::string, float> last_ms;   // -2: pending
};

struct mmo_mc {
  mmo_ctx* ctx = nullptr;
  int device = 0;      // the context's, kept so that destruction never reads the context
  pid_t pid = 0;
  int nchain = 0, natoms = 0, nrb = 0, max_atoms = 0, inter_mode = MMO_SHIFTED, flags = 0;
  int total_frames = 0;
  uint64_t seed = 0;
  double beta = 0.0, roi[3] = {0.0, 0.0, 0.0}, roi_r = 0.0;
  std::vector<int> chain_lig, chain_atom, chain_rb, work, work_by_lig;
  std::vector<std::pair<int, int>> groups;   // (first, count) in work_by_lig per ligand
  DevBuf<int> d_chain_lig, d_chain_atom, d_chain_rb, d_work, d_work_by_lig, d_queue;
  DevBuf<McState> d_st;
  DevBuf<double> d_conf, d_best, d_dr, d_undo;
  DevBuf<SwState> d_sw;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  int graph_version = -1;      // ligand registry version the graph was captured with
  // local minimisation after every proposal (mmo_mc_set_minimizer)
  int min_iters = 0, min_m = 8;
  double min_max_step = 0.2, min_gtol = 1e-6;
  DevBuf<double> d_min_base, d_min_S, d_min_Y, d_min_rho;
};

struct mmo_batch {
  mmo_ctx* ctx = nullptr;
  int device = 0;      // the context's, kept so that destruction never reads the context
  pid_t pid = 0;
  int npose = 0, natoms = 0, ndof = 0, ntiles = 0, max_atoms = 0, max_rb = 0, max_rg = 0;
  std::vector<int> pose_lig, atom_off, dof_off, perm;
  std::vector<Tile> tiles;
  // poses by ligand size, largest first, and the pose kernel's slot order dealt from it
  // for workgroups of order_waves wavefronts (0: not dealt yet)
  std::vector<int> by_size, order;
  int order_waves = 0;
  DevBuf<int> d_pose_lig, d_atom_off, d_dof_off, d_perm, d_order, d_by_size;
  DevBuf<Tile> d_tiles;
  DevBuf<double> d_xyz, d_e, d_g, d_part, d_conf;
};

namespace {

int fail(mmo_ctx* c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  if (c) c->err = buf;
  return code;
}

#define HIPCHK(c, expr)                                                                     \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess)                                                                   \
      return fail((c), _e == hipErrorOutOfMemory ? MMO_ERR_NOMEM : MMO_ERR_DEVICE,         \
                  "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__);     \
  } while (0)

int check_ctx(mmo_ctx* c) {
  if (!c) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL context");
  if (getpid() != c->pid)
    return fail(c, MMO_ERR_FORK, "context created by pid %d used from pid %d", (int)c->pid,
                (int)getpid());
  hipError_t e = hipSetDevice(c->device);
  if (e != hipSuccess) return fail(c, MMO_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
  return MMO_OK;
}

// A = sqrt(D) x^6, B = sqrt(D) x^3 (factorised geometric-mean pair parameters)
int atom_ab(mmo_ctx* c, int anum, double* A, double* B) {
  auto it = c->uff.find(anum);
  if (it == c->uff.end()) return fail(c, MMO_ERR_BAD_ARG, "unsupported atomic number %d", anum);
  const double x = it->second.x, sd = std::sqrt(it->second.d);
  const double x3 = x * x * x;
  *A = sd * x3 * x3;
  *B = sd * x3;
  return MMO_OK;
}

template <class T>
int upload(mmo_ctx* c, DevBuf<T>& b, const T* src, size_t n, hipStream_t s) {
  HIPCHK(c, b.ensure(n));
  if (n) HIPCHK(c, hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return MMO_OK;
}

// Order of the atoms for the per-atom sweeps, which take consecutive atoms in groups:
// a depth-first walk of the bond graph (topological distance 1) that lists each atom's
// terminal neighbours (hydrogens, halogens) right after it, so that a group of
// consecutive atoms is a compact piece of the molecule.  Fragments follow one another.
std::vector<int> sweep_order(const int32_t* dists, int n) {
  std::vector<std::vector<int>> nb(n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      if (i != j && dists[i + (size_t)j * n] == 1) nb[i].push_back(j);
  std::vector<int> order;
  std::vector<char> seen(n, 0);
  std::vector<int> stack;
  for (int root = 0; root < n; ++root) {
    if (seen[root]) continue;
    seen[root] = 1;
    stack.push_back(root);
    while (!stack.empty()) {
      const int a = stack.back();
      stack.pop_back();
      order.push_back(a);
      for (int b : nb[a])   // terminal neighbours now
        if (!seen[b] && nb[b].size() <= 1) {
          seen[b] = 1;
          order.push_back(b);
        }
      for (auto it = nb[a].rbegin(); it != nb[a].rend(); ++it)   // branches next, lowest first
        if (!seen[*it]) {
          seen[*it] = 1;
          stack.push_back(*it);
        }
    }
  }
  return order;
}

std::vector<uint32_t> intra_pairs(const int32_t* dists, int n) {
  // src/mol.ml:203-208 (interacting = dists[i + j*n] >= 3), pairs i < j (src/mol.ml:885-890).
  // Listed by diagonal (d = j - i, then i) rather than by row: consecutive entries then
  // hold distinct i and distinct j, so the lanes of one wavefront add their gradient
  // terms to distinct LDS words instead of all to atom i's.
  std::vector<uint32_t> out;
  for (int d = 1; d < n; ++d)
    for (int i = 0; i + d < n; ++i) {
      const int j = i + d;
      if (dists[i + (size_t)j * n] >= 3) out.push_back((uint32_t)i | ((uint32_t)j << 16));
    }
  return out;
}

int pair_params(mmo_ctx* c, const std::vector<uint32_t>& pij, const double* q,
                const int32_t* anum, std::vector<double4>& out) {
  out.resize(pij.size());
  for (size_t p = 0; p < pij.size(); ++p) {
    const int i = pij[p] & 0xffff, j = pij[p] >> 16;
    double Ai, Bi, Aj, Bj;
    int r = atom_ab(c, anum[i], &Ai, &Bi);
    if (r) return r;
    r = atom_ab(c, anum[j], &Aj, &Bj);
    if (r) return r;
    out[p] = make_double4(kElecWeight * q[i] * q[j], Ai * Aj, 2.0 * Bi * Bj, 0.0);
  }
  return MMO_OK;
}

// Slot order of the pose kernel (slot = workgroup * waves + wavefront) for workgroups of
// `waves` wavefronts: wavefront w of a workgroup runs on SIMD w % 4 of its CU, so the
// wavefronts of one workgroup with equal w % 4 form a bin that shares one SIMD.  The
// poses, largest ligand first, are dealt over the bins in snake order (forward, then
// backward, ...), which evens out the ligand atoms per SIMD.
std::vector<int> deal_poses(const std::vector<int>& by_size, int waves) {
  const int npose = (int)by_size.size();
  const int simds = std::min(waves, 4), rounds = (waves + simds - 1) / simds;
  const int nbin = ((npose + waves - 1) / waves) * simds;
  std::vector<int> order(npose, -1);
  size_t next = 0;
  for (int t = 0; t < rounds; ++t)
    for (int k = 0; k < nbin; ++k) {
      const int bin = (t & 1) ? nbin - 1 - k : k;
      const int w = bin % simds + simds * t;
      const long long slot = (long long)(bin / simds) * waves + w;
      if (w < waves && slot < npose) order[slot] = by_size[next++];
    }
  return order;   // every slot < npose is visited once: next == npose
}

int sync_ligands(mmo_ctx* c) {
  if (!c->lig_dirty) return MMO_OK;
  std::vector<LigInfo> info;
  std::vector<LigAtom> atoms;
  std::vector<double> xyz;
  std::vector<uint32_t> pij;
  std::vector<double4> ppar;
  std::vector<int> rbl, rbr, rgo, rga, swp;
  c->max_atoms = 0;
  for (const LigHost& h : c->ligs) {
    LigInfo li{};
    li.n_atoms = h.n;
    li.atom_base = (int)atoms.size();
    li.pair_base = (int)pij.size();
    li.n_pairs = (int)h.pair_ij.size();
    li.nrb = (int)h.rb_left.size();
    li.rb_base = (int)rbl.size();
    li.rg_base = (int)rgo.size();
    li.dof_n = 6 + li.nrb;
    for (int j = 0; j < h.n; ++j) {
      double A, B;
      int r = atom_ab(c, h.anum[j], &A, &B);
      if (r) return r;
      LigAtom la{kElecWeight * h.q[j], A, B, h.type[j], h.anum[j]};
      atoms.push_back(la);
    }
    xyz.insert(xyz.end(), h.xyz.begin(), h.xyz.end());
    swp.insert(swp.end(), h.sweep.begin(), h.sweep.end());
    std::vector<double4> pp;
    int r = pair_params(c, h.pair_ij, h.q.data(), h.anum.data(), pp);
    if (r) return r;
    pij.insert(pij.end(), h.pair_ij.begin(), h.pair_ij.end());
    ppar.insert(ppar.end(), pp.begin(), pp.end());
    rbl.insert(rbl.end(), h.rb_left.begin(), h.rb_left.end());
    rbr.insert(rbr.end(), h.rb_right.begin(), h.rb_right.end());
    const int ga = (int)rga.size();
    for (int v : h.rg_off) rgo.push_back(v + ga);
    rga.insert(rga.end(), h.rg_atoms.begin(), h.rg_atoms.end());
    info.push_back(li);
    c->max_atoms = std::max(c->max_atoms, h.n);
  }
  hipStream_t s = c->stream;
  int r;
  if ((r = upload(c, c->d_ligs, info.data(), info.size(), s))) return r;
  if ((r = upload(c, c->d_latoms, atoms.data(), atoms.size(), s))) return r;
  if ((r = upload(c, c->d_lig_xyz, xyz.data(), xyz.size(), s))) return r;
  if ((r = upload(c, c->d_pair_ij, pij.data(), pij.size(), s))) return r;
  if ((r = upload(c, c->d_pair_par, ppar.data(), ppar.size(), s))) return r;
  if ((r = upload(c, c->d_rb_left, rbl.data(), rbl.size(), s))) return r;
  if ((r = upload(c, c->d_rb_right, rbr.data(), rbr.size(), s))) return r;
  if ((r = upload(c, c->d_rg_off, rgo.data(), rgo.size(), s))) return r;
  if ((r = upload(c, c->d_rg_atoms, rga.data(), rga.size(), s))) return r;
  if ((r = upload(c, c->d_sweep, swp.data(), swp.size(), s))) return r;
  HIPCHK(c, hipStreamSynchronize(s));
  c->lig_dirty = false;
  ++c->lig_version;
  return MMO_OK;
}

InterLaunch inter_launch(mmo_ctx* c, int mode, const LigInfo* ligs, const LigAtom* latoms,
                         const double* lig_xyz) {
  InterLaunch L;
  L.mode = mode;
  L.k = c->inter_k;
  L.waves = c->inter_waves;
  L.rtile = std::min(c->inter_rtile, std::max(c->nrec, 1));
  L.rslice = c->nrec;
  L.cull = 0;
  L.cull_waves = c->cull_waves;
  L.cull_k = c->cull_k;
  L.row_chunk = c->scan_chunk;
  L.row_filter = c->row_filter;
  L.atom_order = nullptr;
  L.cg = c->cg;
  L.fr = c->fr;
  L.rpos = c->rpos.p;
  L.rab = c->rab.p;
  L.rres = c->rres.p;
  L.nrec = c->nrec;
  L.ligs = ligs;
  L.latoms = latoms;
  L.lig_xyz = lig_xyz;
  return L;
}

int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Pose tiles: poses grouped by ligand; `split` lanes per pose so that a wavefront
// holds min(64, #poses) poses of one ligand (split = 64 / poses per tile).
void plan_tiles(const std::vector<int>& pose_lig, const std::vector<LigHost>* ligs,
                const std::vector<int>* natoms, int k, std::vector<int>& perm,
                std::vector<Tile>& tiles) {
  const int npose = (int)pose_lig.size();
  perm.resize(npose);
  for (int p = 0; p < npose; ++p) perm[p] = p;
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int a, int b) { return pose_lig[a] < pose_lig[b]; });
  tiles.clear();
  int p = 0;
  while (p < npose) {
    int q = p;
    while (q < npose && pose_lig[perm[q]] == pose_lig[perm[p]]) ++q;
    const int lig = pose_lig[perm[p]];
    const int m = q - p;
    const int n_at = ligs ? (*ligs)[lig].n : (*natoms)[lig];
    (void)n_at;
    (void)k;
    const int per = std::min(kWave, next_pow2(m));   // poses per tile
    const int split = kWave / per;                   // lanes per pose
    for (int s = p; s < q; s += per) tiles.push_back(Tile{lig, s, std::min(per, q - s), split});
    p = q;
  }
}

// HIP events bracketing the last launch of each kernel family, on its stream
void timer_begin(mmo_ctx* c, const char* name, hipStream_t s) {
  (void)hipEventRecord(c->ev[name].first, s);
}
void timer_end(mmo_ctx* c, const char* name, hipStream_t s) {
  (void)hipEventRecord(c->ev[name].second, s);
  c->last_ms[name] = -2.0f;   // resolved by mmo_last_kernel_ms
}

// plain one-ligand helper: uploads coordinates and a one-entry ligand table
int plain_setup(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                const double* q, const int32_t* anum, int n, bool need_params) {
  std::vector<double> xyz(3 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    if (!std::isfinite(xs[j]) || !std::isfinite(ys[j]) || !std::isfinite(zs[j]) ||
        (q && !std::isfinite(q[j])))
      return fail(c, MMO_ERR_BAD_ARG, "ligand atom %d: coordinates or charge not finite", j);
    xyz[3 * j] = xs[j]; xyz[3 * j + 1] = ys[j]; xyz[3 * j + 2] = zs[j];
  }
  std::vector<LigAtom> la(n);
  for (int j = 0; j < n; ++j) {
    double A = 0.0, B = 0.0;
    if (need_params) {
      int r = atom_ab(c, anum[j], &A, &B);
      if (r) return r;
    }
    la[j] = LigAtom{need_params ? kElecWeight * q[j] : 0.0, A, B, 0, need_params ? anum[j] : 0};
  }
  LigInfo li{};
  li.n_atoms = n;
  li.dof_n = 6;
  int r;
  hipStream_t s = c->stream;
  if ((r = upload(c, c->s_xyz, xyz.data(), xyz.size(), s))) return r;
  if ((r = upload(c, c->s_latom, la.data(), la.size(), s))) return r;
  if ((r = upload(c, c->s_lig, &li, 1, s))) return r;
  return MMO_OK;
}

// the scan kernel's configuration.  scan_cull 1: receptor staged in LDS when it fits;
// 2: read through the scalar cache; 3: read through the scalar cache, rows cut along x by
// the pose sweeps' cell table; 4: positions in LDS, LJ factors through the scalar cache;
// 5: row lists over the scan grid (k_scan_rows), receptor through the scalar cache;
// 6: row lists, receptor in LDS when it fits; 0: the unculled kernel
InterLaunch scan_launch(mmo_ctx* c, int mode, int scan_cull) {
  InterLaunch L = inter_launch(c, mode, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p);
  L.cull = 0;
  if (mode == MMO_SHIFTED && scan_cull >= 5) {
    L.cull = (scan_cull == 6 && (size_t)(c->nrec + kRowChunk) * (sizeof(RecPos) + sizeof(RecAB)) <=
                                    150 * 1024) ? 5 : 4;
    L.cg = c->cgs;
    L.rpos = c->rpos_s.p; L.rab = c->rab_s.p; L.rres = c->rres_s.p;
    L.row_chunk = c->scan_chunk;
    L.row_filter = c->row_filter;
    L.atom_order = c->d_sweep.p;
    return L;
  }
  if (mode == MMO_SHIFTED && scan_cull) {
    L.cull = 2;
    if (scan_cull == 1 && c->cull_ok) L.cull = 1;
    if (scan_cull == 4 &&
        (size_t)c->nrec * sizeof(RecPos) + (size_t)(c->cg.nx * c->cg.ny * c->cg.nz + 1) * sizeof(int) <=
            150 * 1024)
      L.cull = 3;
  }
  if (L.cull && scan_cull == 3) {
    L.cg = c->cgp;
    L.cg.clip_min = 1 << 30;
  }
  return L;
}

// translations in Morton order of their per-axis ranks: 64 consecutive slots form a
// compact brick (a 4x4x4 block of a lattice)
std::vector<int> morton_translations(const double* trans, int ntrans) {
  std::vector<int> tperm(ntrans);
  std::vector<uint32_t> rank[3];
  for (int a = 0; a < 3; ++a) {
    std::vector<double> v(ntrans);
    for (int t = 0; t < ntrans; ++t) v[t] = trans[3 * t + a];
    std::vector<double> u = v;
    std::sort(u.begin(), u.end());
    std::vector<double> uq;
    for (double x : u)
      if (uq.empty() || x - uq.back() > 1e-9) uq.push_back(x);
    rank[a].resize(ntrans);
    for (int t = 0; t < ntrans; ++t)
      rank[a][t] = (uint32_t)(std::lower_bound(uq.begin(), uq.end(), v[t] - 1e-9) - uq.begin());
  }
  auto spread = [](uint32_t x) {
    uint64_t r = 0;
    for (int b = 0; b < 21; ++b) r |= (uint64_t)((x >> b) & 1u) << (3 * b);
    return r;
  };
  std::vector<uint64_t> key(ntrans);
  for (int t = 0; t < ntrans; ++t)
    key[t] = spread(rank[0][t]) | (spread(rank[1][t]) << 1) | (spread(rank[2][t]) << 2);
  for (int t = 0; t < ntrans; ++t) tperm[t] = t;
  std::stable_sort(tperm.begin(), tperm.end(), [&](int a, int b) { return key[a] < key[b]; });
  return tperm;
}

}  // namespace

namespace mmo {
// errors of the host-only entry points (ligand input), read back by mmo_last_error(NULL)
int host_fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace mmo

// =========================================================================== C ABI
extern "C" {

const char* mmo_version(void) { return "mmo_ff 1.0 gfx950"; }

const char* mmo_last_error(const mmo_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_last_error.c_str();
}

int mmo_ctx_create(int device, mmo_ctx** out) {
  if (!out) return fail(nullptr, MMO_ERR_BAD_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0)
    return fail(nullptr, MMO_ERR_DEVICE, "no HIP device (%s)", hipGetErrorString(e));
  if (device < 0 || device >= ndev)
    return fail(nullptr, MMO_ERR_BAD_ARG, "device %d out of range (%d devices)", device, ndev);
  e = hipSetDevice(device);
  if (e != hipSuccess) return fail(nullptr, MMO_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
  scratch_setup(device);
  mmo_ctx* c = new mmo_ctx();
  c->device = device;
  c->pid = getpid();
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return fail(nullptr, MMO_ERR_DEVICE, "stream creation failed");
  }
  for (const char* k : {"inter", "interp", "intra", "grid", "mc", "min"}) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      delete c;
      return fail(nullptr, MMO_ERR_DEVICE, "event creation failed");
    }
    c->ev[k] = {a, b};
  }
  *out = c;
  return MMO_OK;
}

void mmo_ctx_destroy(mmo_ctx* c) {
  if (!c) return;
  if (getpid() == c->pid) {
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    c->rpos.release(); c->rab.release(); c->rres.release(); c->grid.release();
    c->grid_cells.release();
    c->rpos_s.release(); c->rab_s.release(); c->rres_s.release(); c->d_cells.release();
    c->d_ligs.release(); c->d_latoms.release(); c->d_lig_xyz.release();
    c->d_pair_ij.release(); c->d_pair_par.release(); c->d_sweep.release();
    c->d_rb_left.release(); c->d_rb_right.release(); c->d_rg_off.release(); c->d_rg_atoms.release();
    c->s_xyz.release(); c->s_e.release(); c->s_g.release(); c->s_part.release();
    c->s_lig.release(); c->s_latom.release(); c->s_pij.release(); c->s_ppar.release();
    c->s_int.release(); c->s_tile.release(); c->s_scan.release(); c->s_tperm.release();
    c->d_cell.release(); c->d_cellp.release(); c->d_mask.release(); c->s_skip.release();
    c->s_tk_st.release(); c->s_tk_u.release(); c->s_keep.release(); c->s_select.release();
    c->s_list.release(); c->s_nkept.release(); c->d_desolv.release(); c->d_rec64.release();
    c->s_dvox.release(); c->s_datoms.release(); c->s_dpart.release();
    for (auto& kv : c->ev) {
      (void)hipEventDestroy(kv.second.first);
      (void)hipEventDestroy(kv.second.second);
    }
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
}

int mmo_set_uff_table(mmo_ctx* c, const int32_t* anum, const double* x, const double* d, int n) {
  int r = check_ctx(c);
  if (r) return r;
  if (!anum || !x || !d || n <= 0) return fail(c, MMO_ERR_BAD_ARG, "bad UFF table arguments");
  std::map<int, Uff> t;
  for (int i = 0; i < n; ++i) {
    if (!(x[i] >= 0.0) || !(d[i] >= 0.0) || !std::isfinite(x[i]) || !std::isfinite(d[i]))
      return fail(c, MMO_ERR_BAD_ARG, "UFF entry %d: x and D must be finite and >= 0", i);
    t[anum[i]] = Uff{x[i], d[i]};
  }
  c->uff = t;
  c->lig_dirty = true;
  c->has_rec = false;   // receptor parameters depend on the table: upload again
  return MMO_OK;
}

int mmo_set_receptor(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                     const double* q, const int32_t* anum, int n, const double* origin) {
  int r = check_ctx(c);
  if (r) return r;
  if (!xs || !ys || !zs || !q || !anum || n <= 0)
    return fail(c, MMO_ERR_BAD_ARG, "bad receptor arguments (n=%d)", n);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; ++i) {
    const double v[3] = {xs[i], ys[i], zs[i]};
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(v[a])) return fail(c, MMO_ERR_BAD_ARG, "receptor atom %d not finite", i);
      lo[a] = std::min(lo[a], v[a]);
      hi[a] = std::max(hi[a], v[a]);
    }
    if (!std::isfinite(q[i])) return fail(c, MMO_ERR_BAD_ARG, "receptor charge %d not finite", i);
  }
  Frame fr{};
  if (origin) { fr.ox = origin[0]; fr.oy = origin[1]; fr.oz = origin[2]; }
  else { fr.ox = 0.5 * (lo[0] + hi[0]); fr.oy = 0.5 * (lo[1] + hi[1]); fr.oz = 0.5 * (lo[2] + hi[2]); }
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double o = a == 0 ? fr.ox : a == 1 ? fr.oy : fr.oz;
    ext = std::max(ext, std::max(std::fabs(hi[a] - o), std::fabs(lo[a] - o)));
  }
  // fixed point: |x| <= 2^(30-k) must cover the receptor plus 16 A, so that ligand
  // atoms saturated at the range boundary are beyond the 12 A cutoff
  int k = 26;
  while (k > 8 && std::ldexp(1.0, 30 - k) < ext + 16.0) --k;
  fr.scale = std::ldexp(1.0, k);
  fr.inv_s = (float)std::ldexp(1.0, -k);
  fr.inv_s2 = (float)std::ldexp(1.0, -2 * k);
  fr.lim = (1 << 30) - 1;
  // cell list: atoms sorted by cell (x fastest), cell_start as exclusive prefix sums;
  // coordinates in A relative to the frame origin
  const double cs = c->cell_size;
  const double gx0 = lo[0] - fr.ox - 1e-3, gy0 = lo[1] - fr.oy - 1e-3, gz0 = lo[2] - fr.oz - 1e-3;
  const int gnx = (int)std::floor((hi[0] - fr.ox - gx0) / cs) + 1;
  const int gny = (int)std::floor((hi[1] - fr.oy - gy0) / cs) + 1;
  const int gnz = (int)std::floor((hi[2] - fr.oz - gz0) / cs) + 1;
  const size_t ncell = (size_t)gnx * gny * gnz;
  std::vector<int> cell(n), order(n);
  for (int i = 0; i < n; ++i) {
    const int cx = std::min(gnx - 1, std::max(0, (int)std::floor((xs[i] - fr.ox - gx0) / cs)));
    const int cy = std::min(gny - 1, std::max(0, (int)std::floor((ys[i] - fr.oy - gy0) / cs)));
    const int cz = std::min(gnz - 1, std::max(0, (int)std::floor((zs[i] - fr.oz - gz0) / cs)));
    cell[i] = (cz * gny + cy) * gnx + cx;
    order[i] = i;
  }
  // within a cell by x, so that every row of cells (cz, cy) is sorted by x: the sweeps
  // clip a row to an x interval with one ballot
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return cell[a] < cell[b] || (cell[a] == cell[b] && xs[a] < xs[b]);
  });
  std::vector<int> cstart(ncell + 1, 0);
  for (int i = 0; i < n; ++i) cstart[cell[i] + 1]++;
  for (size_t k = 0; k < ncell; ++k) cstart[k + 1] += cstart[k];
  // the pose sweeps' grid: the same rows, cells cell_xsub times narrower along x; the
  // atom order (by row, then x) is sorted for both grids
  // the cut is reduced until the receptor and this table fit in 96 KiB of LDS, leaving
  // room for the Monte Carlo workgroup's conformers and chain states
  int xsub = c->cell_xsub;
  auto ncell_x = [&](int n) { return (int)std::floor((hi[0] - fr.ox - gx0) / (cs / n)) + 1; };
  while (xsub > 1 && (size_t)n * (sizeof(RecPos) + sizeof(RecAB)) +
                             ((size_t)ncell_x(xsub) * gny * gnz + 1) * sizeof(int) > 96 * 1024)
    --xsub;
  const double csx = cs / xsub;
  const int gnxp = ncell_x(xsub);
  const size_t ncellp = (size_t)gnxp * gny * gnz;
  std::vector<int> cstartp(ncellp + 1, 0);
  size_t prev = 0;
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    const int cxp = std::min(gnxp - 1, std::max(0, (int)std::floor((xs[i] - fr.ox - gx0) / csx)));
    const size_t cp = (size_t)(cell[i] / gnx) * gnxp + cxp;
    if (cp < prev) return fail(c, MMO_ERR_DEVICE, "receptor cell order inconsistent along x");
    prev = cp;
    cstartp[cp + 1]++;
  }
  for (size_t k = 0; k < ncellp; ++k) cstartp[k + 1] += cstartp[k];
  std::vector<RecPos> pos(n);
  std::vector<RecAB> ab(n);
  std::vector<RecRes> res(n);
  for (int k = 0; k < n; ++k) {
    const int i = order[k];
    double A, B;
    r = atom_ab(c, anum[i], &A, &B);
    if (r) return r;
    pos[k].x = (int)std::llrint((xs[i] - fr.ox) * fr.scale);
    pos[k].y = (int)std::llrint((ys[i] - fr.oy) * fr.scale);
    pos[k].z = (int)std::llrint((zs[i] - fr.oz) * fr.scale);
    pos[k].q = (float)q[i];
    // quantisation residuals (A): the pair kernels add them for atoms closer than 0.5 A
    res[k].x = (float)((xs[i] - fr.ox) - pos[k].x / fr.scale);
    res[k].y = (float)((ys[i] - fr.oy) - pos[k].y / fr.scale);
    res[k].z = (float)((zs[i] - fr.oz) - pos[k].z / fr.scale);
    ab[k].a = (float)A;
    ab[k].b = (float)B;
  }
  if ((r = upload(c, c->rpos, pos.data(), pos.size(), c->stream))) return r;
  if ((r = upload(c, c->rab, ab.data(), ab.size(), c->stream))) return r;
  if ((r = upload(c, c->rres, res.data(), res.size(), c->stream))) return r;
  if ((r = upload(c, c->d_cell, cstart.data(), cstart.size(), c->stream))) return r;
  if ((r = upload(c, c->d_cellp, cstartp.data(), cstartp.size(), c->stream))) return r;
  // the rigid scan's row-list grid: cells of scan_cell A in y and z and scan_cell /
  // scan_xsub A along x, atoms sorted by cell and by x within a cell (so every row of
  // cells is sorted by x), with its own copy of the receptor arrays in that order
  const double css = c->scan_cell, cssx = css / c->scan_xsub;
  const int snx = (int)std::floor((hi[0] - fr.ox - gx0) / cssx) + 1;
  const int sny = (int)std::floor((hi[1] - fr.oy - gy0) / css) + 1;
  const int snz = (int)std::floor((hi[2] - fr.oz - gz0) / css) + 1;
  if ((double)snx * sny * snz > 6.4e7)
    return fail(c, MMO_ERR_BAD_ARG, "scan cell grid too large (%d x %d x %d)", snx, sny, snz);
  std::vector<int> rank(n), cells(n), orders(n);
  for (int k2 = 0; k2 < n; ++k2) rank[order[k2]] = k2;
  for (int i = 0; i < n; ++i) {
    const int cx = std::min(snx - 1, std::max(0, (int)std::floor((xs[i] - fr.ox - gx0) / cssx)));
    const int cy = std::min(sny - 1, std::max(0, (int)std::floor((ys[i] - fr.oy - gy0) / css)));
    const int cz = std::min(snz - 1, std::max(0, (int)std::floor((zs[i] - fr.oz - gz0) / css)));
    cells[i] = (cz * sny + cy) * snx + cx;
    orders[i] = i;
  }
  std::stable_sort(orders.begin(), orders.end(), [&](int a, int b) {
    return cells[a] < cells[b] || (cells[a] == cells[b] && xs[a] < xs[b]);
  });
  std::vector<int> cstarts((size_t)snx * sny * snz + 1, 0);
  for (int i = 0; i < n; ++i) cstarts[cells[i] + 1]++;
  for (size_t k2 = 0; k2 + 1 < cstarts.size(); ++k2) cstarts[k2 + 1] += cstarts[k2];
  // kRowChunk - 1 inert atoms of padding: row_sweep fetches whole chunks
  std::vector<RecPos> pos_s(n + kRowChunk);
  std::vector<RecAB> ab_s(n + kRowChunk);
  std::vector<RecRes> res_s(n + kRowChunk);
  for (int k2 = 0; k2 < n; ++k2) {
    const int from = rank[orders[k2]];
    pos_s[k2] = pos[from]; ab_s[k2] = ab[from]; res_s[k2] = res[from];
  }
  if ((r = upload(c, c->rpos_s, pos_s.data(), pos_s.size(), c->stream))) return r;
  if ((r = upload(c, c->rab_s, ab_s.data(), ab_s.size(), c->stream))) return r;
  if ((r = upload(c, c->rres_s, res_s.data(), res_s.size(), c->stream))) return r;
  if ((r = upload(c, c->d_cells, cstarts.data(), cstarts.size(), c->stream))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  CellGrid cg{};
  cg.cell_start = c->d_cell.p;
  cg.x0 = (float)gx0; cg.y0 = (float)gy0; cg.z0 = (float)gz0;
  cg.c = (float)cs; cg.inv_c = (float)(1.0 / cs);
  cg.cx = cg.c; cg.inv_cx = cg.inv_c;
  cg.nx = gnx; cg.ny = gny; cg.nz = gnz;
  cg.clip_min = c->row_clip;
  c->cg = cg;
  CellGrid cgp = cg;
  cgp.cell_start = c->d_cellp.p;
  cgp.cx = (float)csx; cgp.inv_cx = (float)(1.0 / csx);
  cgp.nx = gnxp;
  c->cgp = cgp;
  CellGrid cgs = cg;
  cgs.cell_start = c->d_cells.p;
  cgs.c = (float)css; cgs.inv_c = (float)(1.0 / css);
  cgs.cx = (float)cssx; cgs.inv_cx = (float)(1.0 / cssx);
  cgs.nx = snx; cgs.ny = sny; cgs.nz = snz;
  c->cgs = cgs;
  // the culled kernel keeps the whole receptor and the cell table in LDS (160 KiB/CU)
  const size_t lds = (size_t)n * (sizeof(RecPos) + sizeof(RecAB)) + (ncell + 1) * sizeof(int);
  c->cull_ok = lds <= 150 * 1024;
  c->cullp_ok = (size_t)n * (sizeof(RecPos) + sizeof(RecAB)) + (ncellp + 1) * sizeof(int) <= 150 * 1024;
  c->fr = fr;
  c->fp_bits = k;
  c->nrec = n;
  c->has_rec = true;
  c->rec_host.resize(3 * (size_t)n);
  c->rec_anum.assign(anum, anum + n);
  c->rec_q.assign(q, q + n);
  c->has_desolv = false;   // desolvation data belong to the receptor they came from
  for (int i = 0; i < n; ++i) {
    c->rec_host[3 * i] = xs[i]; c->rec_host[3 * i + 1] = ys[i]; c->rec_host[3 * i + 2] = zs[i];
  }
  c->has_mask = false;   // a mask belongs to the receptor it was computed from
  return MMO_OK;
}

namespace {
// The cell-packed copy of the context's grids (8 corners per cell contiguous, 32 B): the
// interpolating kernels then fetch a cell with two 16-byte loads instead of 8 gathers.
// Skipped (the kernels gather from the node arrays) when grid_pack is 0 or the copy
// would pass 4 GiB.
int pack_grid_cells(mmo_ctx* c, GridDev& g, int ntypes, hipStream_t s) {
  g.cells = nullptr;
  const size_t n = (size_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1) * ntypes * 8;
  if (!c->grid_pack || n * sizeof(float) > ((size_t)4 << 30)) {
    c->grid_cells.release();
    return MMO_OK;
  }
  HIPCHK(c, c->grid_cells.ensure(n));
  HIPCHK(c, launch_pack_cells(g, ntypes, c->grid_cells.p, s));
  HIPCHK(c, hipStreamSynchronize(s));
  g.cells = c->grid_cells.p;
  return MMO_OK;
}
}  // namespace

int mmo_set_grids(mmo_ctx* c, const float* const* grids, int ntypes, double step, int nx, int ny,
                  int nz, const double* origin) {
  int r = check_ctx(c);
  if (r) return r;
  if (!grids || ntypes <= 0 || !(step > 0.0) || nx < 2 || ny < 2 || nz < 2)
    return fail(c, MMO_ERR_BAD_ARG, "bad grid arguments");
  const size_t nvox = (size_t)nx * ny * nz;
  HIPCHK(c, c->grid.ensure(nvox * ntypes));
  for (int t = 0; t < ntypes; ++t) {
    if (!grids[t]) return fail(c, MMO_ERR_BAD_ARG, "grid %d is NULL", t);
    HIPCHK(c, hipMemcpyAsync(c->grid.p + t * nvox, grids[t], nvox * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  GridDev g{};
  g.data = c->grid.p;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nvox = nvox;
  g.step = step;
  g.ods = 1.0 / step;   // Grid.one_div_step, src/grid.ml:41
  g.ox = origin ? origin[0] : 0.0;
  g.oy = origin ? origin[1] : 0.0;
  g.oz = origin ? origin[2] : 0.0;
  if ((r = pack_grid_cells(c, g, ntypes, c->stream))) return r;
  c->g = g;
  c->ntypes = ntypes;
  c->has_grid = true;
  return MMO_OK;
}

// ------------------------------------------------------------------ plain forms
static int plain_inter(mmo_ctx* c, int mode, const double* xs, const double* ys, const double* zs,
                       const double* q, const int32_t* anum, int n, double* e, double* e_elec,
                       double* e_vdw, double* gx, double* gy, double* gz) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (mode != MMO_GLOBAL && mode != MMO_SHIFTED) return fail(c, MMO_ERR_BAD_ARG, "bad mode %d", mode);
  if (!xs || !ys || !zs || !q || !anum || n <= 0 || n > 65535 || !e)
    return fail(c, MMO_ERR_BAD_ARG, "bad ligand arguments (n=%d)", n);
  if ((r = plain_setup(c, xs, ys, zs, q, anum, n, true))) return r;
  hipStream_t s = c->stream;
  const bool grad = gx != nullptr;
  // one pose: split the receptor over block rows so the whole chip works on it
  InterLaunch L = inter_launch(c, mode, c->s_lig.p, c->s_latom.p, nullptr);
  L.rslice = std::max(64, std::min(c->nrec, 256));
  L.rtile = std::min(L.rtile, L.rslice);
  L.waves = 1;
  const int nslice = (c->nrec + L.rslice - 1) / L.rslice;
  int zero = 0;
  Tile t{0, 0, 1, kWave};
  if ((r = upload(c, c->s_int, &zero, 1, s))) return r;    // atom_off = perm = {0}
  if ((r = upload(c, c->s_tile, &t, 1, s))) return r;
  HIPCHK(c, c->s_e.ensure(4));
  HIPCHK(c, c->s_part.ensure((size_t)nslice * 4));
  if (grad) {
    HIPCHK(c, c->s_g.ensure(3 * (size_t)n));
    HIPCHK(c, launch_fill(c->s_g.p, 3 * (size_t)n, 0.0, s));
  }
  HIPCHK(c, launch_inter_coords(L, c->s_tile.p, c->s_int.p, 1, 1, c->s_xyz.p, c->s_int.p,
                                c->s_e.p, 4, !grad, grad ? c->s_g.p : nullptr, 1,
                                c->s_part.p, s));
  double h[4];
  HIPCHK(c, hipMemcpyAsync(h, c->s_e.p, sizeof h, hipMemcpyDeviceToHost, s));
  std::vector<double> g;
  if (grad) {
    g.resize(3 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(g.data(), c->s_g.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(c, hipStreamSynchronize(s));
  *e = h[0];
  if (e_elec) *e_elec = h[2];
  if (e_vdw) *e_vdw = h[3];
  if (grad)
    for (int j = 0; j < n; ++j) { gx[j] = g[3 * j]; gy[j] = g[3 * j + 1]; gz[j] = g[3 * j + 2]; }
  return MMO_OK;
}

int mmo_ene_inter(mmo_ctx* c, int mode, const double* xs, const double* ys, const double* zs,
                  const double* q, const int32_t* anum, int n, double* e, double* e_elec,
                  double* e_vdw) {
  return plain_inter(c, mode, xs, ys, zs, q, anum, n, e, e_elec, e_vdw, nullptr, nullptr, nullptr);
}

int mmo_ene_inter_grad(mmo_ctx* c, int mode, const double* xs, const double* ys, const double* zs,
                       const double* q, const int32_t* anum, int n, double* e, double* gx,
                       double* gy, double* gz) {
  if (!gx || !gy || !gz) return fail(c, MMO_ERR_BAD_ARG, "gradient outputs are NULL");
  return plain_inter(c, mode, xs, ys, zs, q, anum, n, e, nullptr, nullptr, gx, gy, gz);
}

static int plain_interp(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                        const int32_t* type, int n, double* e, double* gx, double* gy, double* gz) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_grid) return fail(c, MMO_ERR_NO_GRID, "no grids set");
  if (!xs || !ys || !zs || !type || n <= 0 || !e) return fail(c, MMO_ERR_BAD_ARG, "bad arguments");
  for (int j = 0; j < n; ++j)
    if (type[j] < 0 || type[j] >= c->ntypes)
      return fail(c, MMO_ERR_BAD_ARG, "atom %d: type %d not in [0, %d)", j, type[j], c->ntypes);
  if ((r = plain_setup(c, xs, ys, zs, nullptr, nullptr, n, false))) return r;
  hipStream_t s = c->stream;
  std::vector<int> ints(2 + n);
  ints[0] = 0; ints[1] = 0;
  for (int j = 0; j < n; ++j) ints[2 + j] = type[j];
  Tile t{0, 0, 1, kWave};
  if ((r = upload(c, c->s_int, ints.data(), ints.size(), s))) return r;
  if ((r = upload(c, c->s_tile, &t, 1, s))) return r;
  HIPCHK(c, c->s_e.ensure(4));
  const bool grad = gx != nullptr;
  if (grad) HIPCHK(c, c->s_g.ensure(3 * (size_t)n));
  InterpLaunch L{c->g, c->s_lig.p, c->s_latom.p, nullptr};
  HIPCHK(c, launch_interp_coords(L, c->s_tile.p, c->s_int.p, 1, c->s_xyz.p, c->s_int.p + 1,
                                 c->s_int.p + 2, c->s_e.p, 4, grad ? c->s_g.p : nullptr, 0, s));
  double h = 0.0;
  HIPCHK(c, hipMemcpyAsync(&h, c->s_e.p, sizeof h, hipMemcpyDeviceToHost, s));
  std::vector<double> g;
  if (grad) {
    g.resize(3 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(g.data(), c->s_g.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(c, hipStreamSynchronize(s));
  if (std::isnan(h)) return fail(c, MMO_ERR_OUT_OF_GRID, "an atom is outside the grid");
  *e = h;
  if (grad)
    for (int j = 0; j < n; ++j) { gx[j] = g[3 * j]; gy[j] = g[3 * j + 1]; gz[j] = g[3 * j + 2]; }
  return MMO_OK;
}

int mmo_ene_inter_interp(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                         const int32_t* type, int n, double* e) {
  return plain_interp(c, xs, ys, zs, type, n, e, nullptr, nullptr, nullptr);
}

int mmo_ene_inter_interp_grad(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                              const int32_t* type, int n, double* e, double* gx, double* gy,
                              double* gz) {
  if (!gx || !gy || !gz) return fail(c, MMO_ERR_BAD_ARG, "gradient outputs are NULL");
  return plain_interp(c, xs, ys, zs, type, n, e, gx, gy, gz);
}

static int plain_intra(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                       const double* q, const int32_t* anum, const int32_t* dists, int n, double* e,
                       double* gx, double* gy, double* gz) {
  int r = check_ctx(c);
  if (r) return r;
  if (!xs || !ys || !zs || !q || !anum || !dists || n <= 0 || n > 65535 || !e)
    return fail(c, MMO_ERR_BAD_ARG, "bad ligand arguments (n=%d)", n);
  if ((r = plain_setup(c, xs, ys, zs, q, anum, n, true))) return r;
  std::vector<uint32_t> pij = intra_pairs(dists, n);
  std::vector<double4> ppar;
  if ((r = pair_params(c, pij, q, anum, ppar))) return r;
  LigInfo li{};
  li.n_atoms = n;
  li.n_pairs = (int)pij.size();
  hipStream_t s = c->stream;
  if ((r = upload(c, c->s_lig, &li, 1, s))) return r;
  if ((r = upload(c, c->s_pij, pij.data(), pij.size(), s))) return r;
  if ((r = upload(c, c->s_ppar, ppar.data(), ppar.size(), s))) return r;
  int ints[2] = {0, 0};   // atom_off, pose_lig
  if ((r = upload(c, c->s_int, ints, 2, s))) return r;
  HIPCHK(c, c->s_e.ensure(4));
  const bool grad = gx != nullptr;
  if (grad) HIPCHK(c, c->s_g.ensure(3 * (size_t)n));
  IntraArgs a{};
  a.ligs = c->s_lig.p; a.pair_ij = c->s_pij.p; a.pair_par = c->s_ppar.p; a.xyz = c->s_xyz.p;
  a.atom_off = c->s_int.p; a.pose_lig = c->s_int.p + 1; a.npose = 1; a.max_atoms = n;
  a.e_out = c->s_e.p; a.e_stride = 1; a.grad = grad ? c->s_g.p : nullptr; a.grad_accumulate = 0;
  HIPCHK(c, launch_intra(a, grad, s));
  double h = 0.0;
  HIPCHK(c, hipMemcpyAsync(&h, c->s_e.p, sizeof h, hipMemcpyDeviceToHost, s));
  std::vector<double> g;
  if (grad) {
    g.resize(3 * (size_t)n);
    HIPCHK(c, hipMemcpyAsync(g.data(), c->s_g.p, g.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  HIPCHK(c, hipStreamSynchronize(s));
  *e = h;
  if (grad)
    for (int j = 0; j < n; ++j) { gx[j] = g[3 * j]; gy[j] = g[3 * j + 1]; gz[j] = g[3 * j + 2]; }
  return MMO_OK;
}

int mmo_ene_intra(mmo_ctx* c, const double* xs, const double* ys, const double* zs, const double* q,
                  const int32_t* anum, const int32_t* dists, int n, double* e) {
  return plain_intra(c, xs, ys, zs, q, anum, dists, n, e, nullptr, nullptr, nullptr);
}

int mmo_ene_intra_grad(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                       const double* q, const int32_t* anum, const int32_t* dists, int n, double* e,
                       double* gx, double* gy, double* gz) {
  if (!gx || !gy || !gz) return fail(c, MMO_ERR_BAD_ARG, "gradient outputs are NULL");
  return plain_intra(c, xs, ys, zs, q, anum, dists, n, e, gx, gy, gz);
}

// ------------------------------------------------------------------ ligand registry
int mmo_ligand_add(mmo_ctx* c, const double* xs, const double* ys, const double* zs,
                   const double* q, const int32_t* anum, const int32_t* dists, int n,
                   const int32_t* rb_left, const int32_t* rb_right, const int32_t* rg_off,
                   const int32_t* rg_atoms, int nrb, const int32_t* types, int* lig_id) {
  int r = check_ctx(c);
  if (r) return r;
  if (!xs || !ys || !zs || !q || !anum || !dists || n <= 0 || n > 65535 || nrb < 0 || !lig_id)
    return fail(c, MMO_ERR_BAD_ARG, "bad ligand arguments (n=%d, nrb=%d)", n, nrb);
  if (nrb > 0 && (!rb_left || !rb_right || !rg_off || !rg_atoms))
    return fail(c, MMO_ERR_BAD_ARG, "rotatable bonds given without their arrays");
  LigHost h;
  h.n = n;
  h.xyz.resize(3 * (size_t)n);
  for (int j = 0; j < n; ++j) {
    h.xyz[3 * j] = xs[j]; h.xyz[3 * j + 1] = ys[j]; h.xyz[3 * j + 2] = zs[j];
    double A, B;
    if ((r = atom_ab(c, anum[j], &A, &B))) return r;
  }
  h.q.assign(q, q + n);
  h.anum.assign(anum, anum + n);
  h.type.assign(n, 0);
  if (types) {
    for (int j = 0; j < n; ++j)
      if (types[j] < 0) return fail(c, MMO_ERR_BAD_ARG, "atom %d: negative grid type", j);
    h.type.assign(types, types + n);
  }
  h.pair_ij = intra_pairs(dists, n);
  h.sweep = sweep_order(dists, n);
  for (int k = 0; k < nrb; ++k) {
    if (rb_left[k] < 0 || rb_left[k] >= n || rb_right[k] < 0 || rb_right[k] >= n ||
        rb_left[k] == rb_right[k])
      return fail(c, MMO_ERR_BAD_ARG, "rbond %d: bad atoms %d-%d", k, rb_left[k], rb_right[k]);
    if (rg_off[k] < 0 || rg_off[k + 1] < rg_off[k])
      return fail(c, MMO_ERR_BAD_ARG, "rbond %d: bad rotation group offsets", k);
    for (int m = rg_off[k]; m < rg_off[k + 1]; ++m)
      if (rg_atoms[m] < 0 || rg_atoms[m] >= n || rg_atoms[m] == rb_left[k] ||
          rg_atoms[m] == rb_right[k])
        return fail(c, MMO_ERR_BAD_ARG, "rbond %d: rotation group holds atom %d", k, rg_atoms[m]);
  }
  if (nrb > 0) {
    h.rb_left.assign(rb_left, rb_left + nrb);
    h.rb_right.assign(rb_right, rb_right + nrb);
    h.rg_off.assign(rg_off, rg_off + nrb + 1);
    for (int& v : h.rg_off) v -= rg_off[0];
    h.rg_atoms.assign(rg_atoms + rg_off[0], rg_atoms + rg_off[nrb]);
  } else {
    h.rg_off.assign(1, 0);
  }
  c->ligs.push_back(std::move(h));
  c->lig_dirty = true;
  *lig_id = (int)c->ligs.size() - 1;
  return MMO_OK;
}

int mmo_ligand_count(const mmo_ctx* c) { return c ? (int)c->ligs.size() : 0; }
int mmo_ligand_natoms(const mmo_ctx* c, int id) {
  if (!c || id < 0 || id >= (int)c->ligs.size()) return MMO_ERR_BAD_ARG;
  return c->ligs[id].n;
}

// ------------------------------------------------------------------ batches
// wavefronts (poses) per workgroup of the pose kernel
static int pose_cull_waves(const mmo_ctx* c, int npose) {
  return std::max(1, std::min(c->pose_waves, (npose + 255) / 256));
}

// Deals the batch's slot order for workgroups of `waves` wavefronts (at creation, and
// again only if the workgroup shape changes).  The host copy stays in the batch while
// the upload may be in flight; a re-deal first waits for the previous upload.
static int deal_batch(mmo_batch* b, int waves, hipStream_t s) {
  if (b->order_waves == waves) return MMO_OK;
  mmo_ctx* c = b->ctx;
  if (b->order_waves != 0) HIPCHK(c, hipStreamSynchronize(s));
  b->order = deal_poses(b->by_size, waves);
  int r = upload(c, b->d_order, b->order.data(), b->order.size(), s);
  if (r) return r;
  b->order_waves = waves;
  return MMO_OK;
}

int mmo_batch_create(mmo_ctx* c, const int32_t* pose_lig, int npose, mmo_batch** out) {
  int r = check_ctx(c);
  if (r) return r;
  if (!pose_lig || npose <= 0 || !out) return fail(c, MMO_ERR_BAD_ARG, "bad batch arguments");
  *out = nullptr;
  mmo_batch* b = new mmo_batch();
  b->ctx = c;
  b->device = c->device;
  b->pid = c->pid;
  b->npose = npose;
  b->pose_lig.assign(pose_lig, pose_lig + npose);
  b->atom_off.resize(npose + 1);
  b->dof_off.resize(npose + 1);
  long long na = 0, nd = 0;
  for (int p = 0; p < npose; ++p) {
    const int l = pose_lig[p];
    if (l < 0 || l >= (int)c->ligs.size()) {
      delete b;
      return fail(c, MMO_ERR_BAD_ARG, "pose %d: unknown ligand %d", p, l);
    }
    b->atom_off[p] = (int)na;
    b->dof_off[p] = (int)nd;
    na += c->ligs[l].n;
    nd += 6 + (long long)c->ligs[l].rb_left.size();
    if (na > 2147483647LL || nd > 2147483647LL) {
      delete b;
      return fail(c, MMO_ERR_BAD_ARG, "batch too large: more than 2^31 - 1 atom rows");
    }
    b->max_atoms = std::max(b->max_atoms, c->ligs[l].n);
    b->max_rb = std::max(b->max_rb, (int)c->ligs[l].rb_left.size());
    b->max_rg = std::max(b->max_rg, c->ligs[l].rg_off.back());
  }
  b->atom_off[npose] = (int)na;
  b->dof_off[npose] = (int)nd;
  b->natoms = (int)na;
  b->ndof = (int)nd;
  plan_tiles(b->pose_lig, &c->ligs, nullptr, c->inter_k, b->perm, b->tiles);
  b->ntiles = (int)b->tiles.size();
  b->by_size.resize(npose);
  for (int p = 0; p < npose; ++p) b->by_size[p] = p;
  std::stable_sort(b->by_size.begin(), b->by_size.end(), [&](int u, int v) {
    return c->ligs[pose_lig[u]].n > c->ligs[pose_lig[v]].n;
  });
  hipStream_t s = c->stream;
  if ((r = deal_batch(b, pose_cull_waves(c, npose), s)) ||
      (r = upload(c, b->d_pose_lig, b->pose_lig.data(), npose, s)) ||
      (r = upload(c, b->d_atom_off, b->atom_off.data(), npose + 1, s)) ||
      (r = upload(c, b->d_dof_off, b->dof_off.data(), npose + 1, s)) ||
      (r = upload(c, b->d_perm, b->perm.data(), npose, s)) ||
      (r = upload(c, b->d_tiles, b->tiles.data(), b->tiles.size(), s))) {
    delete b;
    return r;
  }
  hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete b;
    return fail(c, MMO_ERR_DEVICE, "batch upload: %s", hipGetErrorString(e));
  }
  *out = b;
  return MMO_OK;
}

void mmo_batch_destroy(mmo_batch* b) {
  if (!b) return;
  if (getpid() == b->pid) {
    (void)hipSetDevice(b->device);
    b->d_pose_lig.release(); b->d_atom_off.release(); b->d_dof_off.release(); b->d_perm.release();
    b->d_order.release(); b->d_by_size.release();
    b->d_tiles.release(); b->d_xyz.release(); b->d_e.release(); b->d_g.release(); b->d_part.release();
    b->d_conf.release();
  }
  delete b;
}

int mmo_batch_natoms(const mmo_batch* b) { return b ? b->natoms : MMO_ERR_BAD_ARG; }

int mmo_batch_atom_offsets(const mmo_batch* b, int32_t* off) {
  if (!b || !off) return MMO_ERR_BAD_ARG;
  std::copy(b->atom_off.begin(), b->atom_off.end(), off);
  return MMO_OK;
}

int mmo_batch_dof_offsets(const mmo_batch* b, int32_t* off) {
  if (!b || !off) return MMO_ERR_BAD_ARG;
  std::copy(b->dof_off.begin(), b->dof_off.end(), off);
  return MMO_OK;
}

static int batch_score(mmo_batch* b, int flags, const double* xyz, double* e_out, double* grad,
                       hipStream_t s) {
  mmo_ctx* c = b->ctx;
  const int inter = flags & (MMO_F_INTER_SHIFTED | MMO_F_INTER_GLOBAL | MMO_F_INTER_INTERP);
  if (inter && (inter & (inter - 1)))
    return fail(c, MMO_ERR_BAD_ARG, "at most one E_inter form per call");
  const bool want_grad = (flags & MMO_F_GRAD) != 0;
  if (!xyz || !e_out || (want_grad && !grad)) return fail(c, MMO_ERR_BAD_ARG, "NULL buffer");
  if ((inter & (MMO_F_INTER_SHIFTED | MMO_F_INTER_GLOBAL)) && !c->has_rec)
    return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if ((inter & MMO_F_INTER_INTERP) && !c->has_grid) return fail(c, MMO_ERR_NO_GRID, "no grids set");
  int r = sync_ligands(c);
  if (r) return r;
  if (inter & MMO_F_INTER_INTERP) {
    for (const LigHost& h : c->ligs)
      for (int t : h.type)
        if (t >= c->ntypes) return fail(c, MMO_ERR_BAD_ARG, "ligand grid type %d >= %d", t, c->ntypes);
  }
  // every gradient producer adds to the buffer, except the pose kernel: when it runs it
  // is the first producer and writes each atom's row once, so the buffer is not cleared
  const bool pose_first = want_grad && (inter & MMO_F_INTER_SHIFTED) && c->pose_cull;
  if (want_grad && !pose_first) HIPCHK(c, launch_fill(grad, 3 * (size_t)b->natoms, 0.0, s));
  HIPCHK(c, launch_fill(e_out, 4 * (size_t)b->npose, 0.0, s));
  if ((inter & MMO_F_INTER_SHIFTED) && c->pose_cull) {
    // one wavefront per pose over the culled receptor (per-atom or 8-atom sub-block sweep)
    InterLaunch L = inter_launch(c, MMO_SHIFTED, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p);
    L.cg = c->cgp;
    // the receptor goes to LDS when enough poses per workgroup amortise the copy
    L.cull = (c->cullp_ok && b->npose >= 1024) ? 1 : 2;
    PoseCullArgs a{};
    a.xyz = xyz; a.atom_off = b->d_atom_off.p; a.pose_lig = b->d_pose_lig.p; a.npose = b->npose;
    a.rres = c->rres.p;
    a.ligs = c->d_ligs.p; a.latoms = c->d_latoms.p; a.atom_order = c->d_sweep.p;
    a.e_out = e_out; a.e_stride = 4;
    const bool comp = (flags & MMO_F_COMPONENTS) != 0;
    const int waves = pose_cull_waves(c, b->npose);
    if ((r = deal_batch(b, waves, s))) return r;
    a.order = b->d_order.p;
    timer_begin(c, "inter", s);
    if (!want_grad || comp) {
      a.e_comp = comp ? e_out : nullptr;
      HIPCHK(c, launch_pose_cull(L, a, false, waves, c->pose_sweep == 1, s));
    }
    if (want_grad) {
      a.e_comp = nullptr;
      a.grad = grad;
      a.grad_assign = 1;
      a.write_e = comp ? 0 : 1;
      HIPCHK(c, launch_pose_cull(L, a, true, waves, c->pose_sweep == 1, s));
    }
    timer_end(c, "inter", s);
  } else if (inter & (MMO_F_INTER_SHIFTED | MMO_F_INTER_GLOBAL)) {
    const int mode = (inter & MMO_F_INTER_SHIFTED) ? MMO_SHIFTED : MMO_GLOBAL;
    InterLaunch L = inter_launch(c, mode, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p);
    if (mode == MMO_SHIFTED && b->ntiles >= 1024) {
      L.cull = c->cull_ok ? 1 : 2;
      L.cull_waves = std::max(1, std::min(c->cull_waves, b->ntiles / 512));
    }
    // few tiles: split the receptor over block rows so that the chip fills up
    const int blocks = (b->ntiles + L.waves - 1) / L.waves;
    int nslice = 1;
    while (!L.cull && blocks * nslice < 512 && c->nrec / (nslice * 2) >= 128) nslice *= 2;
    L.rslice = (c->nrec + nslice - 1) / nslice;
    L.rtile = std::min(L.rtile, L.rslice);
    nslice = (c->nrec + L.rslice - 1) / L.rslice;
    if (nslice > 1) HIPCHK(c, b->d_part.ensure((size_t)nslice * b->npose * 4));
    const bool comp = (flags & MMO_F_COMPONENTS) != 0;
    timer_begin(c, "inter", s);
    if (!want_grad || comp)
      HIPCHK(c, launch_inter_coords(L, b->d_tiles.p, b->d_perm.p, b->ntiles, b->npose, xyz,
                                    b->d_atom_off.p, e_out, 4, comp, nullptr, 1, b->d_part.p, s));
    if (want_grad)
      HIPCHK(c, launch_inter_coords(L, b->d_tiles.p, b->d_perm.p, b->ntiles, b->npose, xyz,
                                    b->d_atom_off.p, e_out, 4, false, grad, comp ? 0 : 1,
                                    b->d_part.p, s));
    timer_end(c, "inter", s);
  }
  if (inter & MMO_F_INTER_INTERP) {
    InterpLaunch L{c->g, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p};
    timer_begin(c, "interp", s);
    HIPCHK(c, launch_interp_coords(L, b->d_tiles.p, b->d_perm.p, b->ntiles, xyz, b->d_atom_off.p,
                                   nullptr, e_out, 4, want_grad ? grad : nullptr, 1, s));
    timer_end(c, "interp", s);
  }
  if (flags & MMO_F_INTRA) {
    IntraArgs a{};
    a.ligs = c->d_ligs.p; a.pair_ij = c->d_pair_ij.p; a.pair_par = c->d_pair_par.p; a.xyz = xyz;
    a.atom_off = b->d_atom_off.p; a.pose_lig = b->d_pose_lig.p; a.npose = b->npose;
    a.max_atoms = b->max_atoms; a.e_out = e_out + 1; a.e_stride = 4;
    a.grad = want_grad ? grad : nullptr; a.grad_accumulate = 1;
    timer_begin(c, "intra", s);
    HIPCHK(c, launch_intra(a, want_grad, s));
    timer_end(c, "intra", s);
  }
  return MMO_OK;
}

int mmo_batch_score_dev(mmo_batch* b, int flags, const double* xyz, double* e_out, double* grad,
                        void* stream) {
  if (!b) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL batch");
  int r = check_ctx(b->ctx);
  if (r) return r;
  return batch_score(b, flags, xyz, e_out, grad, (hipStream_t)stream);
}

int mmo_batch_score(mmo_batch* b, int flags, const double* xyz, double* e_out, double* grad) {
  if (!b) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL batch");
  mmo_ctx* c = b->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (!xyz || !e_out) return fail(c, MMO_ERR_BAD_ARG, "NULL buffer");
  hipStream_t s = c->stream;
  HIPCHK(c, b->d_xyz.ensure(3 * (size_t)b->natoms));
  HIPCHK(c, b->d_e.ensure(4 * (size_t)b->npose));
  const bool g = (flags & MMO_F_GRAD) != 0;
  if (g) {
    if (!grad) return fail(c, MMO_ERR_BAD_ARG, "NULL gradient buffer");
    HIPCHK(c, b->d_g.ensure(3 * (size_t)b->natoms));
  }
  HIPCHK(c, hipMemcpyAsync(b->d_xyz.p, xyz, 3 * (size_t)b->natoms * sizeof(double),
                           hipMemcpyHostToDevice, s));
  r = batch_score(b, flags, b->d_xyz.p, b->d_e.p, g ? b->d_g.p : nullptr, s);
  if (r) return r;
  HIPCHK(c, hipMemcpyAsync(e_out, b->d_e.p, 4 * (size_t)b->npose * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  if (g)
    HIPCHK(c, hipMemcpyAsync(grad, b->d_g.p, 3 * (size_t)b->natoms * sizeof(double),
                             hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  return MMO_OK;
}

int mmo_batch_score_dofs_dev(mmo_batch* b, int flags, const double* dofs, double* xyz_out,
                             double* e_out, double* gdof, double* radius_out, void* stream) {
  if (!b) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL batch");
  mmo_ctx* c = b->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (!dofs) return fail(c, MMO_ERR_BAD_ARG, "NULL dofs");
  const bool g = (flags & MMO_F_GRAD) != 0;
  if (g && !gdof) return fail(c, MMO_ERR_BAD_ARG, "NULL dof gradient");
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = (hipStream_t)stream;
  double* xyz = xyz_out;
  if (!xyz) {
    HIPCHK(c, b->d_xyz.ensure(3 * (size_t)b->natoms));
    xyz = b->d_xyz.p;
  }
  DofArgs a{};
  a.ligs = c->d_ligs.p; a.lig_xyz = c->d_lig_xyz.p; a.rb_left = c->d_rb_left.p;
  a.rb_right = c->d_rb_right.p; a.rg_off = c->d_rg_off.p; a.rg_atoms = c->d_rg_atoms.p;
  a.pose_lig = b->d_pose_lig.p; a.atom_off = b->d_atom_off.p; a.dof_off = b->d_dof_off.p;
  a.npose = b->npose; a.max_atoms = b->max_atoms; a.max_rb = b->max_rb; a.max_rg = b->max_rg;
  a.dofs = dofs; a.xyz_out = xyz;
  a.radius_out = radius_out;
  if (g) {   // the adjoint reads the conformers apply_config leaves
    HIPCHK(c, b->d_conf.ensure(3 * (size_t)b->natoms));
    a.conf = b->d_conf.p;
  }
  if (dof_lds_bytes(a, g) > kMaxLds)
    return fail(c, MMO_ERR_BAD_ARG,
                "ligands too large for the DOF kernels: %zu bytes of LDS per pose (limit %zu)",
                dof_lds_bytes(a, g), kMaxLds);
  HIPCHK(c, launch_apply_config(a, s));
  const int score_flags = flags & ~MMO_F_COMPONENTS;
  if (score_flags & ~MMO_F_GRAD) {
    if (!e_out) return fail(c, MMO_ERR_BAD_ARG, "NULL energy buffer");
    double* gx = nullptr;
    int bflags = score_flags;
    if (g) {
      HIPCHK(c, b->d_g.ensure(3 * (size_t)b->natoms));
      gx = b->d_g.p;
      // with the gradient, E_intra is evaluated by the adjoint on its conformer (when
      // the larger LDS of that form fits)
      if (score_flags & MMO_F_INTRA) {
        DofArgs ai = a;
        ai.e_intra = e_out + 1;
        ai.e_stride = 4;
        ai.pair_ij = c->d_pair_ij.p;
        ai.pair_par = c->d_pair_par.p;
        if (dof_lds_bytes(ai, true) <= kMaxLds) {
          a = ai;
          bflags &= ~MMO_F_INTRA;
        }
      }
    }
    if ((r = batch_score(b, bflags, xyz, e_out, gx, s))) return r;
    if (g) {
      a.grad_xyz = gx;
      a.gdof = gdof;
      HIPCHK(c, launch_dof_grad(a, s));
    }
  }
  return MMO_OK;
}

// ------------------------------------------------------------------ rigid scan
int mmo_scan_dev(mmo_ctx* c, int lig_id, int mode, const double* rots, int nrot,
                 const double* trans, int ntrans, double* e_out, void* stream) {
  int r = check_ctx(c);
  if (r) return r;
  if (lig_id < 0 || lig_id >= (int)c->ligs.size())
    return fail(c, MMO_ERR_BAD_ARG, "unknown ligand %d", lig_id);
  if (!rots || !trans || nrot <= 0 || ntrans <= 0 || !e_out)
    return fail(c, MMO_ERR_BAD_ARG, "bad scan arguments");
  if ((long long)nrot * ntrans > 2147483647LL)
    return fail(c, MMO_ERR_BAD_ARG, "too many poses (%d x %d)", nrot, ntrans);
  if ((mode == MMO_SHIFTED || mode == MMO_GLOBAL) && !c->has_rec)
    return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (mode == MMO_INTERP && !c->has_grid) return fail(c, MMO_ERR_NO_GRID, "no grids set");
  if (mode != MMO_SHIFTED && mode != MMO_GLOBAL && mode != MMO_INTERP)
    return fail(c, MMO_ERR_BAD_ARG, "bad mode %d", mode);
  if ((r = sync_ligands(c))) return r;
  if (mode == MMO_INTERP)
    for (int t : c->ligs[lig_id].type)
      if (t >= c->ntypes) return fail(c, MMO_ERR_BAD_ARG, "ligand grid type %d >= %d", t, c->ntypes);
  hipStream_t s = (hipStream_t)stream;
  const std::vector<int> tperm = morton_translations(trans, ntrans);
  if ((r = upload(c, c->s_tperm, tperm.data(), tperm.size(), s))) return r;
  HIPCHK(c, c->s_scan.ensure(9 * (size_t)nrot + 3 * (size_t)ntrans));
  HIPCHK(c, hipMemcpyAsync(c->s_scan.p, rots, 9 * (size_t)nrot * sizeof(double),
                           hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->s_scan.p + 9 * (size_t)nrot, trans, 3 * (size_t)ntrans * sizeof(double),
                           hipMemcpyHostToDevice, s));
  const double* d_rot = c->s_scan.p;
  const double* d_trans = c->s_scan.p + 9 * (size_t)nrot;
  if (mode == MMO_INTERP) {
    InterpLaunch L{c->g, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p};
    timer_begin(c, "interp", s);
    HIPCHK(c, launch_interp_scan(L, lig_id, d_rot, nrot, d_trans, c->s_tperm.p, ntrans, e_out, s));
    timer_end(c, "interp", s);
  } else {
    InterLaunch L = scan_launch(c, mode, c->scan_cull);
    timer_begin(c, "inter", s);
    HIPCHK(c, launch_inter_scan(L, lig_id, d_rot, nrot, d_trans, c->s_tperm.p, ntrans, e_out, s));
    timer_end(c, "inter", s);
  }
  // the host arrays were copied asynchronously from pageable memory; wait so that the
  // caller may reuse them
  HIPCHK(c, hipStreamSynchronize(s));
  return MMO_OK;
}

// ------------------------------------------------------------------ grid precompute
int mmo_grid_compute(mmo_ctx* c, const int32_t* probe_anum, const double* probe_q, int ntypes,
                     double step, int nx, int ny, int nz, const double* origin,
                     const double* roi_center, double roi_r, double max_e, float* out) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (!probe_anum || !probe_q || ntypes <= 0 || !(step > 0.0) || nx < 2 || ny < 2 || nz < 2 ||
      !origin || !(max_e > 0.0))
    return fail(c, MMO_ERR_BAD_ARG, "bad grid arguments");
  const size_t nvox = (size_t)nx * ny * nz;
  std::vector<double4> probe(ntypes);
  for (int t = 0; t < ntypes; ++t) {
    double A, B;
    if ((r = atom_ab(c, probe_anum[t], &A, &B))) return r;
    if (!std::isfinite(probe_q[t])) return fail(c, MMO_ERR_BAD_ARG, "probe %d: charge not finite", t);
    probe[t] = make_double4(kElecWeight * probe_q[t], A, B, 0.0);
  }
  hipStream_t s = c->stream;
  if ((r = upload(c, c->s_ppar, probe.data(), probe.size(), s))) return r;
  HIPCHK(c, c->grid.ensure(nvox * ntypes));
  GridArgs a{};
  a.fr = c->fr; a.rpos = c->rpos.p; a.rab = c->rab.p; a.rres = c->rres.p; a.nrec = c->nrec;
  a.nx = nx; a.ny = ny; a.nz = nz; a.nvox = nvox;
  a.ox = origin[0]; a.oy = origin[1]; a.oz = origin[2]; a.step = step;
  if (roi_center) {
    a.roi_x = roi_center[0]; a.roi_y = roi_center[1]; a.roi_z = roi_center[2];
    a.roi_r2 = roi_r * roi_r;
  } else {
    a.roi_r2 = -1.0;
  }
  a.probe = c->s_ppar.p; a.ntypes = ntypes; a.max_e = max_e; a.out = c->grid.p;
  timer_begin(c, "grid", s);
  HIPCHK(c, launch_grid(a, c->cg, c->cull_ok, s));
  timer_end(c, "grid", s);
  if (out) HIPCHK(c, hipMemcpyAsync(out, c->grid.p, nvox * ntypes * sizeof(float),
                                    hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  GridDev g{};
  g.data = c->grid.p;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nvox = nvox;
  g.step = step;
  g.ods = 1.0 / step;
  g.ox = origin[0]; g.oy = origin[1]; g.oz = origin[2];
  if ((r = pack_grid_cells(c, g, ntypes, s))) return r;
  c->g = g;
  c->ntypes = ntypes;
  c->has_grid = true;
  return MMO_OK;
}

// ------------------------------------------------------------------ Monte Carlo
int mmo_mc_create(mmo_ctx* c, const int32_t* chain_lig, int nchain, const double* rot0,
                  const double* pos0, int inter_mode, int flags, const double* roi_center,
                  double roi_r, double temperature, uint64_t seed, int total_frames,
                  mmo_mc** out) {
  int r = check_ctx(c);
  if (r) return r;
  if (!chain_lig || nchain <= 0 || !rot0 || !pos0 || !out || total_frames < 0 ||
      !(temperature > 0.0))
    return fail(c, MMO_ERR_BAD_ARG, "bad Monte Carlo arguments");
  if (inter_mode == MMO_SHIFTED) {
    if (!c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  } else if (inter_mode == MMO_INTERP) {
    if (!c->has_grid) return fail(c, MMO_ERR_NO_GRID, "no grids set");
  } else {
    return fail(c, MMO_ERR_BAD_ARG, "Monte Carlo E_inter form must be SHIFTED or INTERP");
  }
  if ((flags & kMcHardRoi) && (!roi_center || !(roi_r > 0.0)))
    return fail(c, MMO_ERR_BAD_ARG, "hard ROI needs a centre and a radius");
  if ((r = sync_ligands(c))) return r;
  *out = nullptr;
  mmo_mc* m = new mmo_mc();
  m->ctx = c;
  m->device = c->device;
  m->pid = c->pid;
  m->nchain = nchain;
  m->inter_mode = inter_mode;
  m->flags = flags;
  m->total_frames = total_frames;
  m->seed = seed;
  m->beta = 1.0 / (0.0019872041 * temperature);   // src/lds.ml:66-68, src/const.ml:23-28
  if (roi_center) for (int k = 0; k < 3; ++k) m->roi[k] = roi_center[k];
  m->roi_r = roi_r;
  m->chain_lig.assign(chain_lig, chain_lig + nchain);
  m->chain_atom.resize(nchain);
  m->chain_rb.resize(nchain);
  std::vector<double> conf;
  std::vector<McState> st(nchain);
  const double deg = (4.0 * std::atan(1.0)) / 180.0;   // Math.to_radian
  for (int k = 0; k < nchain; ++k) {
    const int l = chain_lig[k];
    if (l < 0 || l >= (int)c->ligs.size()) {
      delete m;
      return fail(c, MMO_ERR_BAD_ARG, "chain %d: unknown ligand %d", k, l);
    }
    const LigHost& h = c->ligs[l];
    if ((int)h.rb_left.size() > kMcMaxRb) {
      delete m;
      return fail(c, MMO_ERR_BAD_ARG, "ligand %d has more than %d rotatable bonds", l, kMcMaxRb);
    }
    if (inter_mode == MMO_INTERP)
      for (int t : h.type)
        if (t >= c->ntypes) {
          delete m;
          return fail(c, MMO_ERR_BAD_ARG, "ligand %d: grid type %d >= %d", l, t, c->ntypes);
        }
    m->chain_atom[k] = m->natoms;
    m->chain_rb[k] = m->nrb;
    m->natoms += h.n;
    m->nrb += (int)h.rb_left.size();
    m->max_atoms = std::max(m->max_atoms, h.n);
    conf.insert(conf.end(), h.xyz.begin(), h.xyz.end());
    McState& s0 = st[k];
    std::memset(&s0, 0, sizeof(McState));
    for (int q = 0; q < 9; ++q) s0.rot0[q] = s0.rot[q] = rot0[9 * k + q];
    for (int q = 0; q < 3; ++q) s0.pos0[q] = s0.pos[q] = pos0[3 * k + q];
    s0.max_rot = 15.0 * deg;     // Params.max_rot, src/params.ml:10-11
    s0.max_trans = 0.15;         // Params.max_trans, src/params.ml:13-14
    s0.status = kMcRunning;
  }
  {
    // one chain per workgroup must fit: its conformer, undo rows and chain state in LDS
    McArgs a{};
    a.max_atoms = m->max_atoms;
    const size_t need = mc_lds_bytes(a, 1, false);
    if (need > kMaxLds) {
      const int na = m->max_atoms;
      delete m;
      return fail(c, MMO_ERR_BAD_ARG,
                  "ligands of %d atoms need %zu bytes of LDS per chain (at most %zu)", na, need,
                  (size_t)kMaxLds);
    }
  }
  std::vector<double> dr(std::max(m->nrb, 1), 5.0 * deg);   // Params.max_rbond_rot
  std::vector<SwState> sw(std::max(m->nrb, 1));
  std::memset(sw.data(), 0, sw.size() * sizeof(SwState));
  // processing order: most expensive chains first (atoms and intra pairs)
  m->work.resize(nchain);
  for (int k = 0; k < nchain; ++k) m->work[k] = k;
  auto cost = [&](int k) {
    const LigHost& h = c->ligs[chain_lig[k]];
    return (double)h.n * 64.0 + (double)h.pair_ij.size();
  };
  std::stable_sort(m->work.begin(), m->work.end(), [&](int a, int b) { return cost(a) > cost(b); });
  // per-ligand groups (the one-launch-per-ligand form)
  m->work_by_lig.resize(nchain);
  for (int k = 0; k < nchain; ++k) m->work_by_lig[k] = k;
  std::stable_sort(m->work_by_lig.begin(), m->work_by_lig.end(),
                   [&](int a, int b) { return chain_lig[a] < chain_lig[b]; });
  for (int k = 0; k < nchain;) {
    int e = k;
    while (e < nchain && chain_lig[m->work_by_lig[e]] == chain_lig[m->work_by_lig[k]]) ++e;
    m->groups.push_back({k, e - k});
    k = e;
  }
  hipStream_t s = c->stream;
  if ((r = upload(c, m->d_chain_lig, m->chain_lig.data(), nchain, s)) ||
      (r = upload(c, m->d_chain_atom, m->chain_atom.data(), nchain, s)) ||
      (r = upload(c, m->d_chain_rb, m->chain_rb.data(), nchain, s)) ||
      (r = upload(c, m->d_work, m->work.data(), nchain, s)) ||
      (r = upload(c, m->d_work_by_lig, m->work_by_lig.data(), nchain, s)) ||
      (r = upload(c, m->d_st, st.data(), st.size(), s)) ||
      (r = upload(c, m->d_conf, conf.data(), conf.size(), s)) ||
      (r = upload(c, m->d_best, conf.data(), conf.size(), s)) ||
      (r = upload(c, m->d_undo, conf.data(), conf.size(), s)) ||
      (r = upload(c, m->d_dr, dr.data(), dr.size(), s)) ||
      (r = upload(c, m->d_sw, sw.data(), sw.size(), s))) {
    delete m;
    return r;
  }
  hipError_t e = m->d_queue.ensure(1);
  if (e == hipSuccess) e = mc_prepare();
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete m;
    return fail(c, MMO_ERR_DEVICE, "Monte Carlo setup: %s", hipGetErrorString(e));
  }
  *out = m;
  return MMO_OK;
}

static McArgs mc_args(mmo_mc* m, int nsteps, int nwork, const int* work) {
  mmo_ctx* c = m->ctx;
  McArgs a{};
  a.ligs = c->d_ligs.p; a.latoms = c->d_latoms.p; a.lig_xyz = c->d_lig_xyz.p;
  a.atom_order = c->d_sweep.p; a.rb_left = c->d_rb_left.p;
  a.rb_right = c->d_rb_right.p; a.rg_off = c->d_rg_off.p; a.rg_atoms = c->d_rg_atoms.p;
  a.pair_ij = c->d_pair_ij.p; a.pair_par = c->d_pair_par.p; a.max_atoms = m->max_atoms;
  a.nwork = nwork; a.work = work; a.chain_lig = m->d_chain_lig.p; a.chain_atom = m->d_chain_atom.p;
  a.chain_rb = m->d_chain_rb.p; a.st = m->d_st.p; a.conf = m->d_conf.p; a.best_xyz = m->d_best.p;
  a.undo = m->d_undo.p;
  a.rb_dr = m->d_dr.p; a.rb_sw = m->d_sw.p; a.queue = m->d_queue.p;
  a.nsteps = nsteps; a.total_frames = m->total_frames; a.flags = m->flags;
  a.seed_lo = (uint32_t)m->seed; a.seed_hi = (uint32_t)(m->seed >> 32);
  a.beta = m->beta;
  const double deg = (4.0 * std::atan(1.0)) / 180.0;
  a.max_rot0 = 15.0 * deg; a.max_trans0 = 0.15;
  a.roi_x = m->roi[0]; a.roi_y = m->roi[1]; a.roi_z = m->roi[2]; a.roi_r = m->roi_r;
  a.inter_mode = m->inter_mode; a.fr = c->fr; a.rpos = c->rpos.p; a.rab = c->rab.p;
  a.rres = c->rres.p;
  a.nrec = c->nrec; a.sweep = c->mc_sweep;
  a.cg = c->cgp; a.g = c->g;
  a.min_iters = m->min_iters; a.min_m = m->min_m; a.min_max_step = m->min_max_step;
  a.min_min_step = 1e-10; a.min_gtol = m->min_gtol; a.min_c1 = 1e-4;
  a.min_base = m->d_min_base.p; a.min_S = m->d_min_S.p; a.min_Y = m->d_min_Y.p;
  a.min_rho = m->d_min_rho.p;
  return a;
}

// launch shape: up to 8 chains (wavefronts) per workgroup; in the direct form the
// workgroup stages the receptor in LDS when it fits next to the conformers
static hipError_t mc_launch(mmo_mc* m, const McArgs& a, int nwork, hipStream_t s) {
  int dev = 0, ncu = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  const size_t cap = 160 * 1024;
  bool rec_lds = m->inter_mode != MMO_INTERP && m->ctx->cullp_ok && m->ctx->mc_rec_lds;
  const int wmax = std::min(mc_max_waves(a), m->ctx->mc_wg_waves);
  int waves = wmax;
  while (waves > 1 && mc_lds_bytes(a, waves, rec_lds) > cap) --waves;
  if (rec_lds && (waves < std::min(4, wmax) || mc_lds_bytes(a, waves, rec_lds) > cap)) {
    rec_lds = false;
    waves = wmax;
    while (waves > 1 && mc_lds_bytes(a, waves, false) > cap) --waves;
  }
  // at least one workgroup per CU; more when the residency setting allows
  const int resident = ncu * std::max(1, m->ctx->mc_waves_per_cu / waves);
  const int blocks = std::max(1, std::min((nwork + waves - 1) / waves, resident));
  return launch_mc(a, blocks, waves, rec_lds, s);
}

int mmo_mc_run(mmo_mc* m, int nsteps, void* stream) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (nsteps <= 0) return fail(c, MMO_ERR_BAD_ARG, "nsteps must be positive");
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  McArgs a = mc_args(m, nsteps, m->nchain, m->d_work.p);
  timer_begin(c, "mc", s);
  HIPCHK(c, mc_launch(m, a, m->nchain, s));
  timer_end(c, "mc", s);
  return MMO_OK;
}

int mmo_mc_run_graph(mmo_mc* m, int nsteps, int nreplay, void* stream) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (nsteps <= 0 || nreplay <= 0) return fail(c, MMO_ERR_BAD_ARG, "bad graph arguments");
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // a graph holds the registry's device pointers: capture again when they may have moved
  if (m->gexec && (m->graph_steps != nsteps || m->graph_version != c->lig_version)) {
    (void)hipGraphExecDestroy(m->gexec);
    (void)hipGraphDestroy(m->graph);
    m->gexec = nullptr;
    m->graph = nullptr;
  }
  if (!m->gexec) {
    // capture queue reset + persistent launch on the context's own stream
    McArgs a = mc_args(m, nsteps, m->nchain, m->d_work.p);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = mc_launch(m, a, m->nchain, c->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess)
      return fail(c, MMO_ERR_DEVICE, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    m->graph = g;
    HIPCHK(c, hipGraphInstantiate(&m->gexec, g, nullptr, nullptr, 0));
    m->graph_steps = nsteps;
    m->graph_version = c->lig_version;
  }
  if (s != c->stream) {
    // order the replays after the caller's prior work
    hipEvent_t ev;
    HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(ev, s));
    HIPCHK(c, hipStreamWaitEvent(c->stream, ev, 0));
    (void)hipEventDestroy(ev);
  }
  timer_begin(c, "mc", c->stream);
  for (int k = 0; k < nreplay; ++k) HIPCHK(c, hipGraphLaunch(m->gexec, c->stream));
  timer_end(c, "mc", c->stream);
  if (s != c->stream) {
    hipEvent_t ev;
    HIPCHK(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(c, hipEventRecord(ev, c->stream));
    HIPCHK(c, hipStreamWaitEvent(s, ev, 0));
    (void)hipEventDestroy(ev);
  }
  return MMO_OK;
}

int mmo_mc_run_per_ligand(mmo_mc* m, int nsteps, void* stream) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (nsteps <= 0) return fail(c, MMO_ERR_BAD_ARG, "nsteps must be positive");
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  timer_begin(c, "mc", s);
  for (const auto& gr : m->groups) {
    McArgs a = mc_args(m, nsteps, gr.second, m->d_work_by_lig.p + gr.first);
    HIPCHK(c, mc_launch(m, a, gr.second, s));
  }
  timer_end(c, "mc", s);
  return MMO_OK;
}

int mmo_mc_set_minimizer(mmo_mc* m, int iters, int history, double max_step, double gtol) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (iters < 0 || history < 1 || history > kMinMaxHist)
    return fail(c, MMO_ERR_BAD_ARG, "minimiser: iters %d, history %d outside 1..%d", iters,
                history, kMinMaxHist);
  // the minimising chains of the direct form on a receptor too large for LDS faulted
  // once at scale (131,072 chains inside a 15,000-atom receptor) with a cause not yet
  // found: that form is refused
  if (iters > 0 && m->inter_mode == MMO_SHIFTED && !c->cullp_ok)
    return fail(c, MMO_ERR_BAD_ARG,
                "minimising chains in the direct form need a receptor that fits in LDS");
  if (iters > 0) {
    McArgs a{};
    a.max_atoms = m->max_atoms;
    a.min_iters = iters;
    if (mc_lds_bytes(a, 1, false) > kMaxLds)
      return fail(c, MMO_ERR_BAD_ARG, "ligands of %d atoms exceed the minimising chains' LDS",
                  m->max_atoms);
    const size_t nc = std::max(m->nchain, 1);
    hipError_t e = m->d_min_base.ensure(3 * (size_t)std::max(m->natoms, 1));
    if (e == hipSuccess) e = m->d_min_S.ensure(nc * history * kWave);
    if (e == hipSuccess) e = m->d_min_Y.ensure(nc * history * kWave);
    if (e == hipSuccess) e = m->d_min_rho.ensure(nc * history);
    if (e != hipSuccess) return fail(c, MMO_ERR_NOMEM, "minimiser buffers: %s", hipGetErrorString(e));
  }
  m->min_iters = iters;
  m->min_m = history;
  if (max_step > 0.0) m->min_max_step = max_step;
  if (gtol > 0.0) m->min_gtol = gtol;
  if (m->gexec) {   // the captured launch no longer matches
    (void)hipGraphExecDestroy(m->gexec);
    (void)hipGraphDestroy(m->graph);
    m->gexec = nullptr;
    m->graph = nullptr;
  }
  return MMO_OK;
}

int mmo_mc_min_evals(mmo_mc* m, int32_t* evals, int64_t* total) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  std::vector<McState> st(m->nchain);
  HIPCHK(c, hipDeviceSynchronize());
  HIPCHK(c, hipMemcpy(st.data(), m->d_st.p, st.size() * sizeof(McState), hipMemcpyDeviceToHost));
  int64_t sum = 0;
  for (int k = 0; k < m->nchain; ++k) {
    if (evals) evals[k] = st[k].min_evals;
    sum += st[k].min_evals;
  }
  if (total) *total = sum;
  return MMO_OK;
}

int mmo_mc_natoms(const mmo_mc* m) { return m ? m->natoms : MMO_ERR_BAD_ARG; }

int mmo_mc_get(mmo_mc* m, double* best_e, double* cur_e, double* best_rot, double* best_pos,
               int32_t* info, double* steps) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  std::vector<McState> st(m->nchain);
  HIPCHK(c, hipDeviceSynchronize());
  HIPCHK(c, hipMemcpy(st.data(), m->d_st.p, st.size() * sizeof(McState), hipMemcpyDeviceToHost));
  for (int k = 0; k < m->nchain; ++k) {
    const McState& x = st[k];
    if (best_e) best_e[k] = x.best_e;
    if (cur_e) { cur_e[3 * k] = x.prev_e; cur_e[3 * k + 1] = x.prev_e_inter; cur_e[3 * k + 2] = x.prev_e_intra; }
    if (best_rot) for (int q = 0; q < 9; ++q) best_rot[9 * k + q] = x.best_rot[q];
    if (best_pos) for (int q = 0; q < 3; ++q) best_pos[3 * k + q] = x.best_pos[q];
    if (info) {
      int32_t* o = info + 8 * k;
      o[0] = x.frame; o[1] = x.status; o[2] = x.accepted; o[3] = x.resets; o[4] = x.ooroi;
      o[5] = x.rigid_step; o[6] = x.conf_step; o[7] = x.initialized;
    }
    if (steps) { steps[2 * k] = x.max_rot; steps[2 * k + 1] = x.max_trans; }
  }
  return MMO_OK;
}

int mmo_mc_coords(mmo_mc* m, double* conf, double* best_xyz, double* rb_dr) {
  if (!m) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL Monte Carlo handle");
  mmo_ctx* c = m->ctx;
  int r = check_ctx(c);
  if (r) return r;
  HIPCHK(c, hipDeviceSynchronize());
  if (conf) HIPCHK(c, hipMemcpy(conf, m->d_conf.p, 3 * (size_t)m->natoms * sizeof(double), hipMemcpyDeviceToHost));
  if (best_xyz) HIPCHK(c, hipMemcpy(best_xyz, m->d_best.p, 3 * (size_t)m->natoms * sizeof(double), hipMemcpyDeviceToHost));
  if (rb_dr && m->nrb) HIPCHK(c, hipMemcpy(rb_dr, m->d_dr.p, (size_t)m->nrb * sizeof(double), hipMemcpyDeviceToHost));
  return MMO_OK;
}

void mmo_mc_destroy(mmo_mc* m) {
  if (!m) return;
  if (getpid() == m->pid) {
    (void)hipSetDevice(m->device);
    (void)hipDeviceSynchronize();
    if (m->gexec) (void)hipGraphExecDestroy(m->gexec);
    if (m->graph) (void)hipGraphDestroy(m->graph);
    m->d_chain_lig.release(); m->d_chain_atom.release(); m->d_chain_rb.release();
    m->d_work.release(); m->d_work_by_lig.release(); m->d_queue.release(); m->d_st.release();
    m->d_conf.release(); m->d_best.release(); m->d_dr.release(); m->d_undo.release(); m->d_sw.release();
    m->d_min_base.release(); m->d_min_S.release(); m->d_min_Y.release(); m->d_min_rho.release();
  }
  delete m;
}

// ------------------------------------------------- rigid-scan filters and top-k
namespace {
// Ptable.vdW_radii (src/ptable.ml:43-53)
double bondi_radius(int anum) {
  switch (anum) {
    case 1: return 1.2;
    case 6: return 1.7;
    case 7: return 1.6;
    case 8: return 1.55;
    case 9: return 1.5;
    case 12: return 2.2;
    case 15: return 1.95;
    case 16: return 1.8;
    case 17: return 1.8;
    case 35: return 1.9;
    case 53: return 2.1;
    default: return std::nan("");
  }
}

bool mask_get(const mmo_ctx* c, long i, long j, long k) {
  if (i < 0 || j < 0 || k < 0 || i >= c->mask_n[0] || j >= c->mask_n[1] || k >= c->mask_n[2])
    return false;
  const size_t v = (size_t)i + (size_t)j * c->mask_n[0] + (size_t)k * c->mask_n[0] * c->mask_n[1];
  return (c->mask_host[v >> 5] >> (v & 31)) & 1u;
}

// Lds.atom_bitmask_set (src/lds.ml:145-169) on the grid o + (i, j, k) step: the nodes of
// the cube of half-side ceil(r / step) cells around Grid.coord_of_point (truncation of
// (x - x0) / step, src/grid.ml:87-90) closer than r to p get bit b; nodes outside the grid
// are skipped (the reference reads past its node arrays there)
void bitmask_atom(std::vector<uint32_t>& mask, const int n[3], const double o[3], double step,
                  const double* p, double r, bool b) {
  const long ci = (long)((p[0] - o[0]) / step), cj = (long)((p[1] - o[1]) / step),
             ck = (long)((p[2] - o[2]) / step);
  const long rs = (long)std::ceil(r / step);
  const double r2 = r * r;
  for (long i = ci - rs; i <= ci + rs; ++i) {
    if (i < 0 || i >= n[0]) continue;
    const double dx = p[0] - (o[0] + i * step);
    for (long j = cj - rs; j <= cj + rs; ++j) {
      if (j < 0 || j >= n[1]) continue;
      const double dy = p[1] - (o[1] + j * step);
      for (long k = ck - rs; k <= ck + rs; ++k) {
        if (k < 0 || k >= n[2]) continue;
        const double dz = p[2] - (o[2] + k * step);
        if (dx * dx + dy * dy + dz * dz < r2) {
          const size_t v = (size_t)i + (size_t)j * n[0] + (size_t)k * n[0] * n[1];
          if (b) mask[v >> 5] |= 1u << (v & 31);
          else mask[v >> 5] &= ~(1u << (v & 31));
        }
      }
    }
  }
}

// Lds.first_solvent_shell (src/lds.ml:171-184): set within r_vdW + r_H2O (1.4 A,
// src/const.ml:10) of every atom, then unset within r_vdW of every atom
void solvent_shell(std::vector<uint32_t>& mask, const int n[3], const double o[3], double step,
                   const double* xyz, const double* radii, int natoms) {
  mask.assign(((size_t)n[0] * n[1] * n[2] + 31) / 32, 0u);
  for (int a = 0; a < natoms; ++a) bitmask_atom(mask, n, o, step, xyz + 3 * (size_t)a, radii[a] + 1.4, true);
  for (int a = 0; a < natoms; ++a) bitmask_atom(mask, n, o, step, xyz + 3 * (size_t)a, radii[a], false);
}

// Const.desolvation (src/const.ml:31): (1/eps_prot - 1/eps_H2O) / (8 pi), eps 4 and 78.5
double desolvation_const() {
  return (1.0 / 4.0 - 1.0 / 78.5) / (8.0 * (4.0 * std::atan(1.0)));
}

// The scan filters' inputs on the device: the per-translation skips (ROI sphere, centre
// clash; computed here on the host), the rotations and translations, the conformer and
// the vdW mask; `a.e` is left to the caller.  Enqueued on s from host arrays that the
// caller waits for before returning.
int scan_filter_setup(mmo_ctx* c, int lig_id, const double* rots, int nrot, const double* trans,
                      int ntrans, const double* roi_center, double roi_r, const double* lig_radii,
                      int flags, hipStream_t s, ScanFilterArgs& a) {
  const LigHost& h = c->ligs[lig_id];
  // Mol.is_ligand_center_vdW_occuppied (src/mol.ml:1208-1218)
  bool centre_occupied = false;
  if (flags & MMO_SCAN_CENTER_CLASH) {
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int i = 0; i < h.n; ++i) { cx += h.xyz[3 * i]; cy += h.xyz[3 * i + 1]; cz += h.xyz[3 * i + 2]; }
    cx /= h.n; cy /= h.n; cz /= h.n;
    for (int i = 0; i < h.n && !centre_occupied; ++i) {
      const double dx = cx - h.xyz[3 * i], dy = cy - h.xyz[3 * i + 1], dz = cz - h.xyz[3 * i + 2];
      centre_occupied = std::sqrt(dx * dx + dy * dy + dz * dz) < lig_radii[i];
    }
  }
  std::vector<uint8_t>& skip = c->skip_host;
  skip.assign(ntrans, 0);
  for (int t = 0; t < ntrans; ++t) {
    const double* p = trans + 3 * (size_t)t;
    if (flags & MMO_SCAN_ROI) {   // ROI.is_inside: dist2 < out_r^2 (src/ROI.ml:65-66)
      const double dx = roi_center[0] - p[0], dy = roi_center[1] - p[1], dz = roi_center[2] - p[2];
      if (!(dx * dx + dy * dy + dz * dz < roi_r * roi_r)) { skip[t] = 1; continue; }
    }
    if (centre_occupied) {   // G3D.vdW_clash_AND: all 8 corners in the protein volume
      const double inv = 1.0 / c->mask_step;
      const long i0 = (long)((p[0] - c->mask_o[0]) * inv), j0 = (long)((p[1] - c->mask_o[1]) * inv),
                 k0 = (long)((p[2] - c->mask_o[2]) * inv);
      bool all = true;
      for (int q = 0; q < 8 && all; ++q)
        all = mask_get(c, i0 + (q & 1), j0 + ((q >> 1) & 1), k0 + (q >> 2));
      if (all) skip[t] = 1;
    }
  }
  HIPCHK(c, c->s_skip.ensure(ntrans));
  HIPCHK(c, c->s_scan.ensure(9 * (size_t)nrot + 3 * (size_t)ntrans));
  HIPCHK(c, hipMemcpyAsync(c->s_skip.p, skip.data(), ntrans, hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->s_scan.p, rots, 9 * (size_t)nrot * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(c, hipMemcpyAsync(c->s_scan.p + 9 * (size_t)nrot, trans, 3 * (size_t)ntrans * sizeof(double),
                           hipMemcpyHostToDevice, s));
  a = ScanFilterArgs{};
  a.rots = c->s_scan.p; a.trans = c->s_scan.p + 9 * (size_t)nrot; a.nrot = nrot; a.ntrans = ntrans;
  size_t base = 0;
  for (int l = 0; l < lig_id; ++l) base += (size_t)c->ligs[l].n;   // LigInfo.atom_base
  a.lig = c->d_lig_xyz.p + 3 * base; a.n_atoms = h.n;
  a.mask = c->has_mask ? c->d_mask.p : nullptr;
  a.ox = c->mask_o[0]; a.oy = c->mask_o[1]; a.oz = c->mask_o[2];
  a.inv_step = c->has_mask ? 1.0 / c->mask_step : 0.0;
  a.nx = c->mask_n[0]; a.ny = c->mask_n[1]; a.nz = c->mask_n[2];
  a.skip = c->s_skip.p; a.atom_clash = (flags & MMO_SCAN_ATOM_CLASH) ? 1 : 0; a.e = nullptr;
  return MMO_OK;
}
}  // namespace

int mmo_set_vdw_mask(mmo_ctx* c, const double* radii, double step, int nx, int ny, int nz,
                     const double* origin) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (!(step > 0.0) || nx < 2 || ny < 2 || nz < 2 || (double)nx * ny * nz > 2147483647.0)
    return fail(c, MMO_ERR_BAD_ARG, "bad mask grid (step %g, %d x %d x %d)", step, nx, ny, nz);
  const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
  const size_t nvox = (size_t)nx * ny * nz;
  std::vector<uint32_t> mask((nvox + 31) / 32, 0u);
  const int dims[3] = {nx, ny, nz};
  // atom_bitmask_set grid mask xyz r_vdW true (src/lds.ml:164-184) for every atom
  for (int a = 0; a < c->nrec; ++a) {
    const double rad = radii ? radii[a] : bondi_radius(c->rec_anum[a]);
    if (!(rad > 0.0) || !std::isfinite(rad))
      return fail(c, MMO_ERR_BAD_ARG, "receptor atom %d: no vdW radius (element %d)", a,
                  c->rec_anum[a]);
    bitmask_atom(mask, dims, o, step, &c->rec_host[3 * (size_t)a], rad, true);
  }
  if ((r = upload(c, c->d_mask, mask.data(), mask.size(), c->stream))) return r;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mask_host.swap(mask);
  c->mask_step = step;
  c->mask_o[0] = o[0]; c->mask_o[1] = o[1]; c->mask_o[2] = o[2];
  c->mask_n[0] = nx; c->mask_n[1] = ny; c->mask_n[2] = nz;
  c->has_mask = true;
  return MMO_OK;
}

int mmo_get_vdw_mask(const mmo_ctx* c, uint32_t* words) {
  if (!c || !c->has_mask) return MMO_ERR_BAD_ARG;
  if (words) std::memcpy(words, c->mask_host.data(), c->mask_host.size() * sizeof(uint32_t));
  return (int)c->mask_host.size();
}

// Desolvation (src/lds.ml:171-267, formula (2) of Majeux, Scarsi and Caflisch, Proteins
// 2001): the receptor's first solvent shell on the grid origin + (i, j, k) step, and per
// shell node inside the ROI sphere the protein contribution desolvation x voxel volume x
// the sum over receptor atoms within 12 A of (q / d^2)^2 (k_desolv_nodes, FP64).
int mmo_desolv_setup(mmo_ctx* c, const double* radii, double step, int nx, int ny, int nz,
                     const double* origin, const double* roi_center, double roi_r) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (!(step > 0.0) || nx < 2 || ny < 2 || nz < 2 || (double)nx * ny * nz > 2147483647.0)
    return fail(c, MMO_ERR_BAD_ARG, "bad desolvation grid (step %g, %d x %d x %d)", step, nx,
                ny, nz);
  if (roi_center && !(roi_r > 0.0)) return fail(c, MMO_ERR_BAD_ARG, "ROI radius %g", roi_r);
  const double o[3] = {origin ? origin[0] : 0.0, origin ? origin[1] : 0.0, origin ? origin[2] : 0.0};
  const int dims[3] = {nx, ny, nz};
  const size_t nvox = (size_t)nx * ny * nz;
  std::vector<double> rad(c->nrec);
  for (int a = 0; a < c->nrec; ++a) {
    rad[a] = radii ? radii[a] : bondi_radius(c->rec_anum[a]);
    if (!(rad[a] > 0.0) || !std::isfinite(rad[a]))
      return fail(c, MMO_ERR_BAD_ARG, "receptor atom %d: no vdW radius (element %d)", a,
                  c->rec_anum[a]);
  }
  std::vector<uint32_t> shell;
  solvent_shell(shell, dims, o, step, c->rec_host.data(), rad.data(), c->nrec);
  // the shell nodes inside the ROI (ROI.is_inside: squared distance strictly below r^2)
  std::vector<int> vox;
  for (size_t w = 0; w < shell.size(); ++w)
    for (uint32_t bits = shell[w]; bits; bits &= bits - 1) {
      const size_t v = 32 * w + __builtin_ctz(bits);
      if (roi_center) {
        const size_t k = v / ((size_t)nx * ny), j = (v / nx) % ny, i = v % nx;
        const double dx = roi_center[0] - (o[0] + i * step), dy = roi_center[1] - (o[1] + j * step),
                     dz = roi_center[2] - (o[2] + k * step);
        if (!(dx * dx + dy * dy + dz * dz < roi_r * roi_r)) continue;
      }
      vox.push_back((int)v);
    }
  std::vector<double4> at(c->nrec);
  for (int a = 0; a < c->nrec; ++a)
    at[a] = make_double4(c->rec_host[3 * (size_t)a], c->rec_host[3 * (size_t)a + 1],
                         c->rec_host[3 * (size_t)a + 2], c->rec_q[a]);
  hipStream_t s = c->stream;
  HIPCHK(c, c->d_desolv.ensure(nvox));
  HIPCHK(c, launch_fill(c->d_desolv.p, nvox, 0.0, s));
  if ((r = upload(c, c->d_rec64, at.data(), at.size(), s))) return r;
  if ((r = upload(c, c->s_dvox, vox.data(), vox.size(), s))) return r;
  DesolvArgs a{};
  a.vox = c->s_dvox.p; a.nvox = (int)vox.size();
  a.atoms = c->d_rec64.p; a.natoms = c->nrec;
  a.ox = o[0]; a.oy = o[1]; a.oz = o[2]; a.step = step; a.nx = nx; a.ny = ny;
  a.scale = desolvation_const() * (step * step * step);
  a.out = c->d_desolv.p;
  HIPCHK(c, launch_desolv(a, false, s));
  HIPCHK(c, hipStreamSynchronize(s));
  c->desolv_shell.swap(shell);
  c->desolv_step = step;
  c->desolv_o[0] = o[0]; c->desolv_o[1] = o[1]; c->desolv_o[2] = o[2];
  c->desolv_n[0] = nx; c->desolv_n[1] = ny; c->desolv_n[2] = nz;
  c->has_desolv = true;
  return (int)vox.size();
}

int mmo_get_solvent_shell(const mmo_ctx* c, uint32_t* words) {
  if (!c || !c->has_desolv) return MMO_ERR_BAD_ARG;
  if (words) std::memcpy(words, c->desolv_shell.data(), c->desolv_shell.size() * sizeof(uint32_t));
  return (int)c->desolv_shell.size();
}

int mmo_get_desolv_contribs(mmo_ctx* c, double* out) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_desolv) return fail(c, MMO_ERR_BAD_ARG, "no desolvation data (mmo_desolv_setup)");
  const size_t nvox = (size_t)c->desolv_n[0] * c->desolv_n[1] * c->desolv_n[2];
  if (out) {
    HIPCHK(c, hipMemcpyAsync(out, c->d_desolv.p, nvox * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return (int)nvox;
}

// Lds.desolvation_penalty (src/lds.ml:236-267) of one ligand pose: over the nodes of both
// the receptor's and the ligand's first solvent shells, the sum of the protein
// contributions (*prot) and desolvation x voxel volume x the ligand atoms' (q / d^2)^2
// within 12 A (*lig); k_desolv_pose sums the nodes, the host adds the workgroups' partials.
int mmo_desolv_penalty(mmo_ctx* c, const double* xyz, const double* q, const double* radii,
                       int n, double* prot, double* lig) {
  int r = check_ctx(c);
  if (r) return r;
  if (!c->has_desolv) return fail(c, MMO_ERR_BAD_ARG, "no desolvation data (mmo_desolv_setup)");
  if (!xyz || !q || !radii || n <= 0 || !prot || !lig)
    return fail(c, MMO_ERR_BAD_ARG, "bad desolvation penalty arguments (n=%d)", n);
  for (int a = 0; a < n; ++a) {
    if (!std::isfinite(xyz[3 * (size_t)a]) || !std::isfinite(xyz[3 * (size_t)a + 1]) ||
        !std::isfinite(xyz[3 * (size_t)a + 2]) || !std::isfinite(q[a]))
      return fail(c, MMO_ERR_BAD_ARG, "ligand atom %d not finite", a);
    if (!(radii[a] > 0.0) || !std::isfinite(radii[a]))
      return fail(c, MMO_ERR_BAD_ARG, "ligand atom %d: radius %g", a, radii[a]);
  }
  std::vector<uint32_t> ls;
  solvent_shell(ls, c->desolv_n, c->desolv_o, c->desolv_step, xyz, radii, n);
  std::vector<int> vox;
  for (size_t w = 0; w < ls.size(); ++w)
    for (uint32_t bits = ls[w] & c->desolv_shell[w]; bits; bits &= bits - 1)
      vox.push_back((int)(32 * w + __builtin_ctz(bits)));
  *prot = 0.0;
  *lig = 0.0;
  if (vox.empty()) return MMO_OK;
  std::vector<double4> at(n);
  for (int a = 0; a < n; ++a)
    at[a] = make_double4(xyz[3 * (size_t)a], xyz[3 * (size_t)a + 1], xyz[3 * (size_t)a + 2], q[a]);
  const int blocks = ((int)vox.size() + kDesolvNodes - 1) / kDesolvNodes;
  hipStream_t s = c->stream;
  if ((r = upload(c, c->s_datoms, at.data(), at.size(), s))) return r;
  if ((r = upload(c, c->s_dvox, vox.data(), vox.size(), s))) return r;
  HIPCHK(c, c->s_dpart.ensure(2 * (size_t)blocks));
  DesolvArgs a{};
  a.vox = c->s_dvox.p; a.nvox = (int)vox.size();
  a.atoms = c->s_datoms.p; a.natoms = n;
  a.ox = c->desolv_o[0]; a.oy = c->desolv_o[1]; a.oz = c->desolv_o[2];
  a.step = c->desolv_step; a.nx = c->desolv_n[0]; a.ny = c->desolv_n[1];
  a.prot = c->d_desolv.p;
  a.part = c->s_dpart.p;
  HIPCHK(c, launch_desolv(a, true, s));
  std::vector<double> part(2 * (size_t)blocks);
  HIPCHK(c, hipMemcpyAsync(part.data(), c->s_dpart.p, part.size() * sizeof(double),
                           hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  double sl = 0.0, sp = 0.0;
  for (int b = 0; b < blocks; ++b) { sl += part[2 * b]; sp += part[2 * b + 1]; }
  *prot = sp;
  *lig = desolvation_const() * (c->desolv_step * c->desolv_step * c->desolv_step * sl);
  return MMO_OK;
}

int mmo_scan_filter_dev(mmo_ctx* c, int lig_id, const double* rots, int nrot, const double* trans,
                        int ntrans, const double* roi_center, double roi_r,
                        const double* lig_radii, int flags, double* e_inout, void* stream) {
  int r = check_ctx(c);
  if (r) return r;
  if (lig_id < 0 || lig_id >= (int)c->ligs.size())
    return fail(c, MMO_ERR_BAD_ARG, "unknown ligand %d", lig_id);
  if (!rots || !trans || nrot <= 0 || ntrans <= 0 || !e_inout)
    return fail(c, MMO_ERR_BAD_ARG, "bad scan filter arguments");
  if ((long long)nrot * ntrans > 2147483647LL)
    return fail(c, MMO_ERR_BAD_ARG, "too many poses (%d x %d)", nrot, ntrans);
  if ((flags & MMO_SCAN_ROI) && (!roi_center || !(roi_r > 0.0)))
    return fail(c, MMO_ERR_BAD_ARG, "ROI filter without an ROI");
  if ((flags & (MMO_SCAN_CENTER_CLASH | MMO_SCAN_ATOM_CLASH)) && !c->has_mask)
    return fail(c, MMO_ERR_BAD_ARG, "clash filters need mmo_set_vdw_mask first");
  if ((flags & MMO_SCAN_CENTER_CLASH) && !lig_radii)
    return fail(c, MMO_ERR_BAD_ARG, "the centre-clash filter needs the ligand's radii");
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = (hipStream_t)stream;
  ScanFilterArgs a{};
  if ((r = scan_filter_setup(c, lig_id, rots, nrot, trans, ntrans, roi_center, roi_r, lig_radii,
                             flags, s, a)))
    return r;
  a.e = e_inout;
  HIPCHK(c, launch_scan_filter(a, s));
  // the host arrays are read by the copies above: wait before returning
  HIPCHK(c, hipStreamSynchronize(s));
  return MMO_OK;
}

int mmo_scan_filtered_dev(mmo_ctx* c, int lig_id, int mode, const double* rots, int nrot,
                          const double* trans, int ntrans, const double* roi_center, double roi_r,
                          const double* lig_radii, int flags, double* e_out, int64_t* n_kept,
                          void* stream) {
  int r = check_ctx(c);
  if (r) return r;
  if (lig_id < 0 || lig_id >= (int)c->ligs.size())
    return fail(c, MMO_ERR_BAD_ARG, "unknown ligand %d", lig_id);
  if (!rots || !trans || nrot <= 0 || ntrans <= 0 || !e_out)
    return fail(c, MMO_ERR_BAD_ARG, "bad filtered scan arguments");
  if ((long long)nrot * ntrans > 2147483647LL)
    return fail(c, MMO_ERR_BAD_ARG, "too many poses (%d x %d)", nrot, ntrans);
  if (mode != MMO_SHIFTED && mode != MMO_INTERP)
    return fail(c, MMO_ERR_BAD_ARG, "the filtered scan scores with MMO_SHIFTED or MMO_INTERP");
  if (mode == MMO_SHIFTED && !c->has_rec) return fail(c, MMO_ERR_NO_RECEPTOR, "no receptor set");
  if (mode == MMO_INTERP && !c->has_grid) return fail(c, MMO_ERR_NO_GRID, "no grids set");
  if ((flags & MMO_SCAN_ROI) && (!roi_center || !(roi_r > 0.0)))
    return fail(c, MMO_ERR_BAD_ARG, "ROI filter without an ROI");
  if ((flags & (MMO_SCAN_CENTER_CLASH | MMO_SCAN_ATOM_CLASH)) && !c->has_mask)
    return fail(c, MMO_ERR_BAD_ARG, "clash filters need mmo_set_vdw_mask first");
  if ((flags & MMO_SCAN_CENTER_CLASH) && !lig_radii)
    return fail(c, MMO_ERR_BAD_ARG, "the centre-clash filter needs the ligand's radii");
  if ((r = sync_ligands(c))) return r;
  if (mode == MMO_INTERP)
    for (int t : c->ligs[lig_id].type)
      if (t >= c->ntypes) return fail(c, MMO_ERR_BAD_ARG, "ligand grid type %d >= %d", t, c->ntypes);
  hipStream_t s = (hipStream_t)stream;
  ScanFilterArgs a{};
  if ((r = scan_filter_setup(c, lig_id, rots, nrot, trans, ntrans, roi_center, roi_r, lig_radii,
                             flags, s, a)))
    return r;
  const std::vector<int> tperm = morton_translations(trans, ntrans);
  if ((r = upload(c, c->s_tperm, tperm.data(), tperm.size(), s))) return r;
  const int n = nrot * ntrans;
  HIPCHK(c, c->s_keep.ensure(n));
  HIPCHK(c, c->s_list.ensure(n));
  HIPCHK(c, c->s_nkept.ensure(1));
  // keep flags in the scan's pose order, then the kept slots compacted in that order (so
  // that consecutive kept poses stay in the scan's bricks)
  HIPCHK(c, launch_scan_keep(a, c->s_tperm.p, c->s_keep.p, s));
  size_t tb = 0;
  HIPCHK(c, select_flagged_slots(c->s_keep.p, n, c->s_list.p, c->s_nkept.p, nullptr, &tb, s));
  HIPCHK(c, c->s_select.ensure(tb));
  HIPCHK(c, select_flagged_slots(c->s_keep.p, n, c->s_list.p, c->s_nkept.p, c->s_select.p, &tb, s));
  HIPCHK(c, launch_fill(e_out, (size_t)n, __builtin_inf(), s));
  int nk = 0;
  HIPCHK(c, hipMemcpyAsync(&nk, c->s_nkept.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (nk > 0) {
    if (mode == MMO_INTERP) {
      InterpLaunch L{c->g, c->d_ligs.p, c->d_latoms.p, c->d_lig_xyz.p};
      timer_begin(c, "interp", s);
      HIPCHK(c, launch_interp_scan_list(L, lig_id, a.rots, a.trans, c->s_tperm.p, ntrans, c->s_list.p,
                                        nk, e_out, s));
      timer_end(c, "interp", s);
    } else {
      // the scan's own culled kernel (the list form has no unculled one): the kept poses'
      // energies are those of mmo_scan_dev, bit for bit
      InterLaunch L = scan_launch(c, MMO_SHIFTED, c->scan_cull ? c->scan_cull : 2);
      timer_begin(c, "inter", s);
      HIPCHK(c, launch_inter_scan_list(L, lig_id, a.rots, a.trans, c->s_tperm.p, ntrans, c->s_list.p,
                                       nk, e_out, s));
      timer_end(c, "inter", s);
    }
  }
  if (n_kept) *n_kept = nk;
  HIPCHK(c, hipStreamSynchronize(s));
  return MMO_OK;
}

int mmo_topk_dev(mmo_ctx* c, const double* e, int64_t n, int k, double* e_out, int64_t* idx_out,
                 void* stream) {
  int r = check_ctx(c);
  if (r) return r;
  if (!e || n <= 0 || k <= 0 || k > kTopkMax || !e_out || !idx_out)
    return fail(c, MMO_ERR_BAD_ARG, "bad top-k arguments (n %lld, k %d, k <= %d)", (long long)n,
                k, kTopkMax);
  const int kk = (int)std::min<int64_t>(k, n);
  const size_t chunk = kTopkChunk;
  const size_t nb = ((size_t)n + chunk - 1) / chunk;
  HIPCHK(c, c->s_tk_st.ensure(3));
  HIPCHK(c, c->s_tk_u.ensure(256 + 1 + nb));
  TopkScratch w{c->s_tk_st.p, c->s_tk_u.p, c->s_tk_u.p + 256, c->s_tk_u.p + 257, chunk};
  HIPCHK(c, launch_topk(e, (size_t)n, kk, e_out, idx_out, w, (hipStream_t)stream));
  return MMO_OK;
}

int mmo_topk_dev_noalloc(mmo_ctx* c, const double* e, int64_t n, int k, double* e_out,
                         int64_t* idx_out, void* stream) {
  int r = check_ctx(c);
  if (r) return r;
  if (!e || n <= 0 || k <= 0 || k > kTopkMax || !e_out || !idx_out)
    return fail(c, MMO_ERR_BAD_ARG, "bad top-k arguments (n %lld, k %d, k <= %d)", (long long)n,
                k, kTopkMax);
  const uint64_t max_n = (uint64_t)kTopkNoallocChunks * kTopkChunk;
  if ((uint64_t)n > max_n)
    return fail(c, MMO_ERR_BAD_ARG,
                "top-k without allocation: n = %lld exceeds the scratch limit of %llu energies",
                (long long)n, (unsigned long long)max_n);
  void* base = nullptr;
  hipError_t se = hipErrorInvalidDevice;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (c->device >= 0 && c->device < kMaxDevices) {
      base = g_scratch_ptr[c->device];
      se = g_scratch_err[c->device];
    }
  }
  if (!base)
    return fail(c, MMO_ERR_DEVICE, "top-k scratch of device %d unavailable: %s", c->device,
                hipGetErrorString(se));
  unsigned char* b = static_cast<unsigned char*>(base);
  const int kk = (int)std::min<int64_t>(k, n);
  TopkScratch w{reinterpret_cast<uint64_t*>(b), reinterpret_cast<unsigned*>(b + 256),
                reinterpret_cast<unsigned*>(b + 1280), reinterpret_cast<unsigned*>(b + 1536),
                kTopkChunk};
  HIPCHK(c, launch_topk(e, (size_t)n, kk, e_out, idx_out, w, (hipStream_t)stream));
  return MMO_OK;
}

// ------------------------------------------------------------- L-BFGS minimiser
struct mmo_min {
  mmo_batch* b = nullptr;
  int device = 0;      // kept so that destruction never reads the batch or the context
  pid_t pid = 0;
  int m = 8, flags = 0;
  double max_step = 1.0, gtol = 1e-6;
  DevBuf<double> x, g, d, xt, gt, e, rad, S, Y, rho, lo, hi;
  DevBuf<MinPose> st;
  DevBuf<int> queue;
};

int mmo_min_create(mmo_batch* b, int history, int flags, const double* lo, const double* hi,
                   double max_step, double gtol, mmo_min** out) {
  if (!b || !out) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL batch or output");
  mmo_ctx* c = b->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (history < 1 || history > kMinMaxHist)
    return fail(c, MMO_ERR_BAD_ARG, "history %d outside 1..%d", history, kMinMaxHist);
  const int inter = flags & (MMO_F_INTER_SHIFTED | MMO_F_INTER_GLOBAL | MMO_F_INTER_INTERP);
  if ((inter & (inter - 1)) || (flags & ~(inter | MMO_F_INTRA)) || !(flags & (inter | MMO_F_INTRA)))
    return fail(c, MMO_ERR_BAD_ARG, "flags: at most one E_inter form, optionally MMO_F_INTRA");
  for (int p = 0; p < b->npose; ++p)
    if (b->dof_off[p + 1] - b->dof_off[p] > kWave)
      return fail(c, MMO_ERR_BAD_ARG, "pose %d has more than %d DOFs", p, kWave);
  mmo_min* z = new mmo_min;
  z->b = b;
  z->device = c->device;
  z->pid = c->pid;
  z->m = history;
  z->flags = flags;
  if (max_step > 0.0) z->max_step = max_step;
  if (gtol > 0.0) z->gtol = gtol;
  const size_t nd = std::max(b->ndof, 1), np = std::max(b->npose, 1);
  hipStream_t s = c->stream;
  hipError_t e = hipSuccess;
  for (DevBuf<double>* v : {&z->x, &z->g, &z->d, &z->xt, &z->gt})
    if (e == hipSuccess) e = v->ensure(nd);
  if (e == hipSuccess) e = z->e.ensure(4 * np);
  if (e == hipSuccess) e = z->rad.ensure(np);
  if (e == hipSuccess) e = z->S.ensure((size_t)history * nd);
  if (e == hipSuccess) e = z->Y.ensure((size_t)history * nd);
  if (e == hipSuccess) e = z->rho.ensure((size_t)history * np);
  if (e == hipSuccess) e = z->st.ensure(np);
  if (e == hipSuccess) e = z->queue.ensure(1);
  if (e == hipSuccess && b->d_by_size.n < (size_t)b->npose)
    e = b->d_by_size.ensure(b->npose) == hipSuccess
            ? hipMemcpyAsync(b->d_by_size.p, b->by_size.data(), b->npose * sizeof(int),
                             hipMemcpyHostToDevice, s)
            : hipErrorOutOfMemory;
  if (e == hipSuccess && lo) e = z->lo.ensure(nd);
  if (e == hipSuccess && hi) e = z->hi.ensure(nd);
  if (e == hipSuccess && lo) e = hipMemcpyAsync(z->lo.p, lo, (size_t)b->ndof * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && hi) e = hipMemcpyAsync(z->hi.p, hi, (size_t)b->ndof * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    mmo_min_destroy(z);
    return fail(c, MMO_ERR_NOMEM, "minimiser buffers: %s", hipGetErrorString(e));
  }
  *out = z;
  return MMO_OK;
}

int mmo_min_run_dev(mmo_min* z, double* dofs, int iters, double* e_out, void* stream) {
  if (!z) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL minimiser");
  mmo_batch* b = z->b;
  mmo_ctx* c = b->ctx;
  int r = check_ctx(c);
  if (r) return r;
  if (!dofs || iters < 0) return fail(c, MMO_ERR_BAD_ARG, "bad minimiser arguments");
  // the kernels read the ligand registry: upload it if ligands were added since
  if ((r = sync_ligands(c))) return r;
  hipStream_t s = (hipStream_t)stream;
  HIPCHK(c, launch_min_reset(z->st.p, b->npose, s));
  MinArgs a{};
  a.npose = b->npose; a.m = z->m; a.ndof_total = b->ndof; a.dof_off = b->d_dof_off.p;
  a.x = z->x.p; a.g = z->g.p; a.d = z->d.p; a.xt = z->xt.p; a.gt = z->gt.p; a.et = z->e.p;
  a.rad = z->rad.p; a.S = z->S.p; a.Y = z->Y.p; a.rho = z->rho.p; a.st = z->st.p;
  a.lo = z->lo.p; a.hi = z->hi.p;
  a.max_step = z->max_step; a.min_step = 1e-10; a.gtol = z->gtol; a.c1 = 1e-4;
  const int inter = z->flags & (MMO_F_INTER_SHIFTED | MMO_F_INTER_GLOBAL | MMO_F_INTER_INTERP);
  // the one-launch form reads a receptor too large for LDS (cell layout beyond 150 KiB)
  // only when a caller forces it (min_rec_lds = 0 on a receptor that fits): run at scale
  // on such a receptor, with poses inside it, it faulted once with a cause not yet
  // found, so large receptors take the launch sequence
  if (c->min_fused && inter != 0 && inter != MMO_F_INTER_GLOBAL &&
      (inter != MMO_F_INTER_SHIFTED || c->cullp_ok) &&
      (inter != MMO_F_INTER_INTERP || c->has_grid)) {
    // every iteration of every pose in one persistent launch (one wavefront per pose)
    if (iters == 0) return MMO_OK;
    MinPersistArgs p{};
    p.ligs = c->d_ligs.p; p.latoms = c->d_latoms.p; p.lig_xyz = c->d_lig_xyz.p;
    p.atom_order = c->d_sweep.p; p.rb_left = c->d_rb_left.p; p.rb_right = c->d_rb_right.p;
    p.rg_off = c->d_rg_off.p; p.rg_atoms = c->d_rg_atoms.p; p.pair_ij = c->d_pair_ij.p;
    p.pair_par = c->d_pair_par.p;
    p.npose = b->npose; p.max_atoms = b->max_atoms; p.ndof_total = b->ndof; p.iters = iters;
    p.m = z->m; p.pose_lig = b->d_pose_lig.p; p.dof_off = b->d_dof_off.p; p.order = b->d_by_size.p;
    p.x = dofs; p.lo = z->lo.p; p.hi = z->hi.p; p.S = z->S.p; p.Y = z->Y.p; p.rho = z->rho.p;
    p.st = z->st.p; p.e_out = e_out; p.queue = z->queue.p; p.intra = (z->flags & MMO_F_INTRA) != 0;
    p.max_step = a.max_step; p.min_step = a.min_step; p.gtol = a.gtol; p.c1 = a.c1;
    p.fr = c->fr; p.rpos = c->rpos.p; p.rab = c->rab.p; p.rres = c->rres.p; p.nrec = c->nrec;
    p.cg = c->cgp; p.g = c->g;
    const int kind = inter == MMO_F_INTER_INTERP ? kDofInterGrid
                     : c->min_f64 ? kDofInterF64 : kDofInterAtom;
    int dev = 0, ncu = 256;
    if (hipGetDevice(&dev) == hipSuccess)
      (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
    int waves = c->min_waves;
    while (waves > 1 && min_persist_lds_bytes(p, waves, false) > kMaxLds) --waves;
    if (min_persist_lds_bytes(p, 1, false) > kMaxLds)
      return fail(c, MMO_ERR_BAD_ARG, "ligands of %d atoms exceed the minimiser's LDS", b->max_atoms);
    bool rec_lds = false;
    if (kind != kDofInterGrid && c->min_rec_lds) {
      int wl = std::min(waves, 8);   // the receptor-in-LDS form is compiled for 8 wavefronts
      while (wl > 1 && min_persist_lds_bytes(p, wl, true) > kMaxLds) --wl;
      if (wl >= 8 && min_persist_lds_bytes(p, wl, true) <= kMaxLds) {
        rec_lds = true;
        waves = wl;
      }
    }
    const int blocks = std::max(1, std::min((b->npose + waves - 1) / waves, ncu * std::max(1, 16 / waves)));
    HIPCHK(c, launch_fill_int(z->queue.p, 1, 0, s));
    timer_begin(c, "min", s);
    HIPCHK(c, launch_min_persist(p, kind, blocks, waves, rec_lds, s));
    timer_end(c, "min", s);
    return MMO_OK;
  }
  // one launch sequence per iteration: batched evaluation, then the update
  HIPCHK(c, hipMemcpyAsync(z->xt.p, dofs, (size_t)b->ndof * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
  HIPCHK(c, hipMemcpyAsync(z->x.p, dofs, (size_t)b->ndof * sizeof(double),
                           hipMemcpyDeviceToDevice, s));
  for (int it = 0; it < iters; ++it) {
    r = mmo_batch_score_dofs_dev(b, z->flags | MMO_F_GRAD, z->xt.p, nullptr, z->e.p, z->gt.p,
                                 z->rad.p, stream);
    if (r) return r;
    HIPCHK(c, launch_lbfgs(a, s));
  }
  if (iters > 0) {
    HIPCHK(c, hipMemcpyAsync(dofs, z->x.p, (size_t)b->ndof * sizeof(double),
                             hipMemcpyDeviceToDevice, s));
    if (e_out) HIPCHK(c, launch_min_energy(z->st.p, b->npose, e_out, s));
  }
  return MMO_OK;
}

int mmo_min_get(mmo_min* z, int32_t* info, double* e_start) {
  if (!z) return fail(nullptr, MMO_ERR_BAD_ARG, "NULL minimiser");
  mmo_ctx* c = z->b->ctx;
  int r = check_ctx(c);
  if (r) return r;
  const int np = z->b->npose;
  std::vector<MinPose> st(np);
  HIPCHK(c, hipDeviceSynchronize());
  if (np) HIPCHK(c, hipMemcpy(st.data(), z->st.p, np * sizeof(MinPose), hipMemcpyDeviceToHost));
  for (int p = 0; p < np; ++p) {
    if (info) {
      info[4 * p] = st[p].status;
      info[4 * p + 1] = st[p].n_eval;
      info[4 * p + 2] = st[p].n_acc;
      info[4 * p + 3] = st[p].n_reset;
    }
    if (e_start) e_start[p] = st[p].f0;
  }
  return MMO_OK;
}

void mmo_min_destroy(mmo_min* z) {
  if (!z) return;
  if (getpid() == z->pid) {
    (void)hipSetDevice(z->device);
    (void)hipDeviceSynchronize();
    for (DevBuf<double>* v : {&z->x, &z->g, &z->d, &z->xt, &z->gt, &z->e, &z->rad, &z->S, &z->Y,
                              &z->rho, &z->lo, &z->hi})
      v->release();
    z->st.release();
    z->queue.release();
  }
  delete z;
}

double mmo_last_kernel_ms(mmo_ctx* c, const char* kernel) {
  if (!c || !kernel) return -1.0;
  auto it = c->last_ms.find(kernel);
  if (it == c->last_ms.end()) return -1.0;
  if (it->second == -2.0f) {
    auto& e = c->ev[kernel];
    if (hipEventSynchronize(e.second) != hipSuccess) return -1.0;
    float ms = -1.0f;
    if (hipEventElapsedTime(&ms, e.first, e.second) != hipSuccess) return -1.0;
    it->second = ms;
  }
  return it->second;
}

int mmo_fixed_point_bits(const mmo_ctx* c) { return c ? c->fp_bits : MMO_ERR_BAD_ARG; }

// kernel shape parameters (the autotuner sets these); cell_size applies to the next
// mmo_set_receptor call
int mmo_tune_set(mmo_ctx* c, const char* key, double v) {
  int r = check_ctx(c);
  if (r) return r;
  if (!key) return fail(c, MMO_ERR_BAD_ARG, "NULL key");
  const std::string k = key;
  const int iv = (int)v;
  auto pow2_in = [](int x, int lo, int hi) { return x >= lo && x <= hi && (x & (x - 1)) == 0; };
  if (k == "inter_k" && pow2_in(iv, 1, 8)) c->inter_k = iv;
  else if (k == "inter_waves" && iv >= 1 && iv <= 4) c->inter_waves = iv;
  else if (k == "inter_rtile" && iv >= 64 && iv <= 6144) c->inter_rtile = iv;
  else if (k == "cull_k" && pow2_in(iv, 1, 4)) c->cull_k = iv;
  else if (k == "cull_waves" && iv >= 1 && iv <= 16) c->cull_waves = iv;
  else if (k == "pose_waves" && iv >= 1 && iv <= 16) c->pose_waves = iv;
  else if (k == "scan_cull" && iv >= 0 && iv <= 6) c->scan_cull = iv;
  else if (k == "cell_size" && v >= 1.0 && v <= 24.0) c->cell_size = (float)v;
  else if (k == "scan_cell" && v >= 1.0 && v <= 24.0) c->scan_cell = (float)v;
  else if (k == "scan_xsub" && iv >= 1 && iv <= 16) c->scan_xsub = iv;
  else if (k == "scan_chunk" && (iv == 4 || iv == 8)) c->scan_chunk = iv;
  else if (k == "row_filter" && (iv == 0 || iv == 1)) c->row_filter = iv;
  else if (k == "grid_pack" && (iv == 0 || iv == 1)) c->grid_pack = iv;
  else if (k == "mc_waves_per_cu" && iv >= 1 && iv <= 32) c->mc_waves_per_cu = iv;
  else if (k == "mc_wg_waves" && iv >= 1 && iv <= 16) c->mc_wg_waves = iv;
  else if (k == "pose_cull" && (iv == 0 || iv == 1)) c->pose_cull = iv;
  else if (k == "pose_sweep" && (iv == 0 || iv == 1)) c->pose_sweep = iv;
  else if (k == "mc_sweep" && (iv == 0 || iv == 1)) c->mc_sweep = iv;
  else if (k == "row_clip" && iv >= 0 && iv <= (1 << 20)) c->row_clip = c->cg.clip_min = iv;
  else if (k == "cell_xsub" && iv >= 1 && iv <= 16) c->cell_xsub = iv;
  else if (k == "min_fused" && (iv == 0 || iv == 1)) c->min_fused = iv;
  else if (k == "min_f64" && (iv == 0 || iv == 1)) c->min_f64 = iv;
  else if (k == "min_waves" && iv >= 1 && iv <= 16) c->min_waves = iv;
  else if (k == "min_rec_lds" && (iv == 0 || iv == 1)) c->min_rec_lds = iv;
  else if (k == "mc_rec_lds" && (iv == 0 || iv == 1)) c->mc_rec_lds = iv;
  else return fail(c, MMO_ERR_BAD_ARG, "bad tuning key/value %s=%g", key, v);
  return MMO_OK;
}

double mmo_tune_get(mmo_ctx* c, const char* key) {
  if (!c || !key) return -1.0;
  const std::string k = key;
  if (k == "inter_k") return c->inter_k;
  if (k == "inter_waves") return c->inter_waves;
  if (k == "inter_rtile") return c->inter_rtile;
  if (k == "cull_k") return c->cull_k;
  if (k == "cull_waves") return c->cull_waves;
  if (k == "scan_cell") return c->scan_cell;
  if (k == "scan_xsub") return c->scan_xsub;
  if (k == "scan_chunk") return c->scan_chunk;
  if (k == "row_filter") return c->row_filter;
  if (k == "grid_pack") return c->grid_pack;
  if (k == "pose_waves") return c->pose_waves;
  if (k == "scan_cull") return c->scan_cull;
  if (k == "cell_size") return c->cell_size;
  if (k == "cull_ok") return c->cull_ok ? 1.0 : 0.0;
  if (k == "cullp_ok") return c->cullp_ok ? 1.0 : 0.0;
  if (k == "mc_waves_per_cu") return c->mc_waves_per_cu;
  if (k == "mc_wg_waves") return c->mc_wg_waves;
  if (k == "pose_cull") return c->pose_cull;
  if (k == "pose_sweep") return c->pose_sweep;
  if (k == "mc_sweep") return c->mc_sweep;
  if (k == "row_clip") return c->row_clip;
  if (k == "cell_xsub") return c->cell_xsub;
  if (k == "min_fused") return c->min_fused;
  if (k == "min_f64") return c->min_f64;
  if (k == "min_waves") return c->min_waves;
  if (k == "min_rec_lds") return c->min_rec_lds;
  if (k == "mc_rec_lds") return c->mc_rec_lds;
  return -1.0;
}

}  // extern "C"
