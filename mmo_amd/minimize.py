"""Batched L-BFGS local minimisation in DOF space on the GPU.

    mz = Minimizer(batch, flags=ff.F_INTER_SHIFTED | ff.F_INTRA, bounds=(lo, hi))
    e_best = mz.run(dofs, iters=200)      # dofs: CUDA FP64 tensor, overwritten in place
    info = mz.info()                       # status, evaluations, accepted steps, restarts

The objective is E_inter + E_intra of Optim.apply_config(dofs) as in
Lds.minimize_ligand (src/lds.ml:1118-1297); the reference drives NLopt with
central-difference gradients inside the bounds of src/optim.ml:34-60, this runs L-BFGS
on the analytic adjoint gradient for every pose of the batch at once.
"""
import ctypes as ct
import sys
import math

import numpy as np

from . import ff
from ._lib import lib
from .ff import _f64, _ptr, _stream_ptr

RUNNING, CONVERGED, FAILED = 0, 1, 2


def roi_bounds(batch, center, radius, nrb_of_pose):
    """Optim.get_min_bounds / get_max_bounds (src/optim.ml:34-60) laid out like the
    batch's DOFs: position inside the ROI cube, every angle in [-pi, pi]."""
    lo = np.empty(int(batch.dof_off[-1]))
    hi = np.empty_like(lo)
    c = np.asarray(center, dtype=np.float64)
    for p in range(batch.npose):
        o = int(batch.dof_off[p])
        n = int(batch.dof_off[p + 1]) - o
        assert n == 6 + int(nrb_of_pose[p])
        lo[o:o + 3], hi[o:o + 3] = c - radius, c + radius
        lo[o + 3:o + n], hi[o + 3:o + n] = -math.pi, math.pi
    return lo, hi


class Minimizer:
    def __init__(self, batch, flags=ff.F_INTER_SHIFTED | ff.F_INTRA, history=8, bounds=None,
                 max_step=1.0, gtol=1e-6):
        self.batch = batch
        self.ctx = batch.ctx
        lo = hi = None
        if bounds is not None:
            lo, hi = (_f64(v) for v in bounds)
            assert lo.size == hi.size == int(batch.dof_off[-1])
        h = ct.c_void_p()
        self.ctx._check(lib.mmo_min_create(batch.h, int(history), int(flags), _ptr(lo), _ptr(hi),
                                           float(max_step), float(gtol), ct.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        # at interpreter exit the HIP runtime may already be gone: the process frees it all
        if h and not sys.is_finalizing():
            lib.mmo_min_destroy(h)
            self.h = None

    def run(self, dofs, iters, e_out=None, stream=None):
        """Minimises in place; returns the best energies (CUDA FP64, npose)."""
        import torch
        assert dofs.is_cuda and dofs.dtype == torch.float64 and dofs.is_contiguous()
        if e_out is None:
            e_out = torch.empty(self.batch.npose, dtype=torch.float64, device=dofs.device)
        self.ctx._check(lib.mmo_min_run_dev(self.h, ct.c_void_p(dofs.data_ptr()), int(iters),
                                            ct.c_void_p(e_out.data_ptr()), _stream_ptr(stream)))
        return e_out

    def info(self):
        n = self.batch.npose
        info = np.zeros((n, 4), dtype=np.int32)
        e0 = np.zeros(n)
        self.ctx._check(lib.mmo_min_get(self.h, _ptr(info), _ptr(e0)))
        return {"status": info[:, 0], "evals": info[:, 1], "accepted": info[:, 2],
                "restarts": info[:, 3], "e_start": e0}
