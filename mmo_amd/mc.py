"""Batched Monte Carlo docking chains on the GPU (Lds.simulate_lig, src/lds.ml:741-1000).

    mc = MonteCarlo(ctx, chain_lig, rot0, pos0, mode=ff.SHIFTED, hard_roi=(center, r))
    mc.run(1000)                 # one persistent launch, all chains advance 1000 frames
    best_e, info = mc.best_energies(), mc.info()

All chains of all ligands run in one launch; ``run_graph`` replays that launch from a
captured hipGraph and ``run_per_ligand`` launches once per ligand (the plain form).
"""
import ctypes as ct
import sys
import math

import numpy as np

from . import _lib
from ._lib import lib
from .ff import _f64, _i32, _ptr, _stream_ptr

TWEAK_RBONDS = _lib.MC_TWEAK_RBONDS
HARD_ROI = _lib.MC_HARD_ROI
INTRA_UFF = _lib.MC_INTRA_UFF
METROPOLIS_ALWAYS = _lib.MC_METROPOLIS_ALWAYS
NO_FLIP = _lib.MC_NO_FLIP
FP64 = _lib.MC_FP64
ROOM_TEMP_K = 293.15            # Const.room_temp_K, src/const.ml:23


class MonteCarlo:
    def __init__(self, ctx, chain_lig, rot0, pos0, mode, total_frames, flags=TWEAK_RBONDS | INTRA_UFF,
                 roi_center=None, roi_r=0.0, temperature=ROOM_TEMP_K, seed=12345):
        self.ctx = ctx
        cl = _i32(chain_lig)
        r0 = _f64(np.asarray(rot0).reshape(len(cl), 9))
        p0 = _f64(np.asarray(pos0).reshape(len(cl), 3))
        rc = _f64(roi_center) if roi_center is not None else None
        h = ct.c_void_p()
        ctx._check(lib.mmo_mc_create(ctx.h, _ptr(cl), len(cl), _ptr(r0), _ptr(p0), int(mode), int(flags),
                                     _ptr(rc), float(roi_r), float(temperature),
                                     ct.c_uint64(int(seed)), int(total_frames), ct.byref(h)))
        self.h = h
        self.nchain = len(cl)
        self.chain_lig = cl
        self.natoms = lib.mmo_mc_natoms(h)
        sizes = np.array([ctx.ligand_natoms(int(l)) for l in cl], dtype=np.int64)
        self.chain_atom = np.concatenate([[0], np.cumsum(sizes)])
        nrb = np.array([ctx.ligand_nrb(int(l)) for l in cl], dtype=np.int64)
        self.chain_rb = np.concatenate([[0], np.cumsum(nrb)])

    def __del__(self):
        h = getattr(self, "h", None)
        # at interpreter exit the HIP runtime may already be gone: the process frees it all
        if h and not sys.is_finalizing():
            lib.mmo_mc_destroy(h)
            self.h = None

    def set_minimizer(self, iters, history=8, max_step=0.2, gtol=1e-6):
        """Monte Carlo-plus-minimisation: `iters` L-BFGS evaluations after every proposal
        (0 switches it off)."""
        self.ctx._check(lib.mmo_mc_set_minimizer(self.h, int(iters), int(history),
                                                 float(max_step), float(gtol)))

    def min_evals(self):
        """Energy + gradient evaluations of the local minimisations: (per chain, total)."""
        per = np.zeros(self.nchain, dtype=np.int32)
        tot = ct.c_int64()
        self.ctx._check(lib.mmo_mc_min_evals(self.h, _ptr(per), ct.byref(tot)))
        return per, tot.value

    def run(self, nsteps, stream=None):
        self.ctx._check(lib.mmo_mc_run(self.h, int(nsteps), _stream_ptr(stream) if stream else None))

    def run_graph(self, nsteps, nreplay=1, stream=None):
        self.ctx._check(lib.mmo_mc_run_graph(self.h, int(nsteps), int(nreplay),
                                             _stream_ptr(stream) if stream else None))

    def run_per_ligand(self, nsteps, stream=None):
        self.ctx._check(lib.mmo_mc_run_per_ligand(self.h, int(nsteps),
                                                  _stream_ptr(stream) if stream else None))

    def state(self):
        n = self.nchain
        best_e = np.zeros(n)
        cur = np.zeros((n, 3))
        brot = np.zeros((n, 9))
        bpos = np.zeros((n, 3))
        info = np.zeros((n, 8), dtype=np.int32)
        steps = np.zeros((n, 2))
        self.ctx._check(lib.mmo_mc_get(self.h, _ptr(best_e), _ptr(cur), _ptr(brot), _ptr(bpos),
                                       _ptr(info), _ptr(steps)))
        return {"best_e": best_e, "e": cur[:, 0], "e_inter": cur[:, 1], "e_intra": cur[:, 2],
                "best_rot": brot.reshape(n, 3, 3), "best_pos": bpos, "frame": info[:, 0],
                "status": info[:, 1], "accepted": info[:, 2], "resets": info[:, 3],
                "ooroi": info[:, 4], "rigid_step": info[:, 5], "conf_step": info[:, 6],
                "max_rot": steps[:, 0], "max_trans": steps[:, 1]}

    def coords(self):
        conf = np.zeros((self.natoms, 3))
        best = np.zeros((self.natoms, 3))
        self.ctx._check(lib.mmo_mc_coords(self.h, _ptr(conf), _ptr(best), None))
        split = lambda a: [a[self.chain_atom[k]:self.chain_atom[k + 1]] for k in range(self.nchain)]
        return split(conf), split(best)

    def bond_steps(self):
        """Each chain's per-rotatable-bond torsion step sizes (radians)."""
        dr = np.zeros(max(int(self.chain_rb[-1]), 1))
        self.ctx._check(lib.mmo_mc_coords(self.h, None, None, _ptr(dr)))
        return [dr[self.chain_rb[k]:self.chain_rb[k + 1]] for k in range(self.nchain)]

    def rescore(self):
        """E_inter of every chain's best pose by the direct shifted sum, whatever form the
        chains ran with: the reference's final-pose rescoring, whose ene_rescore is
        Mol.ene_inter_UFF_shifted_bst (src/lds.ml:681-687, :1983).  One batched launch."""
        from .ff import F_INTER_SHIFTED
        conf = np.zeros((self.natoms, 3))
        best = np.zeros((self.natoms, 3))
        self.ctx._check(lib.mmo_mc_coords(self.h, _ptr(conf), _ptr(best), None))
        e, _ = self.ctx.batch(self.chain_lig).score(best, F_INTER_SHIFTED)
        return e[:, 0].copy()


DETAILED_SCORES_HEADER = "#name\tE_inter\tRB\tHA\n"


def detailed_scores_line(name, e_inter, num_rbonds, num_heavy_atoms):
    """One line of the reference's detailed_scores.tsv (Lds.rescore, src/lds.ml:687; header
    src/lds.ml:1806): name, exact E_inter (%g), rotatable bonds, heavy atoms."""
    return "%s\t%g\t%d\t%d\n" % (name, e_inter, num_rbonds, num_heavy_atoms)


def write_detailed_scores(fn, names, e_inter, ligs):
    """detailed_scores.tsv for chains whose ligands are `ligs` (dicts or Mol records with
    anum and rb_left): the header, then one line per chain."""
    with open(fn, "w") as f:
        f.write(DETAILED_SCORES_HEADER)
        for name, e, lig in zip(names, e_inter, ligs):
            anum = lig["anum"] if isinstance(lig, dict) else lig.anum
            rb = lig["rb_left"] if isinstance(lig, dict) else lig.rb_left
            f.write(detailed_scores_line(name, float(e), len(rb), int(np.sum(np.asarray(anum) > 1))))


def philox4x32(counter, key):
    """Philox4x32-10 (Salmon et al., SC 2011) on the host, the same stream as the kernel:
    counter and key are uint32 sequences of length 4 and 2."""
    c = [int(v) & 0xFFFFFFFF for v in counter]
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = 0xD2511F53 * c[0]
        p1 = 0xCD9E8D57 * c[2]
        hi0, lo0 = (p0 >> 32) & 0xFFFFFFFF, p0 & 0xFFFFFFFF
        hi1, lo1 = (p1 >> 32) & 0xFFFFFFFF, p1 & 0xFFFFFFFF
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0 = (k0 + 0x9E3779B9) & 0xFFFFFFFF
        k1 = (k1 + 0xBB67AE85) & 0xFFFFFFFF
    return c


def frame_uniforms(seed, chain, frame):
    """The six uniforms [0, 1) the kernel draws for (chain, frame)."""
    out = []
    for draw in range(3):
        r = philox([chain, frame, draw, 0x4D4D4F31], [seed & 0xFFFFFFFF, seed >> 32])
        out.append(((r[0] >> 5) * 67108864.0 + (r[1] >> 6)) / 9007199254740992.0)
        out.append(((r[2] >> 5) * 67108864.0 + (r[3] >> 6)) / 9007199254740992.0)
    return out


philox = philox4x32
