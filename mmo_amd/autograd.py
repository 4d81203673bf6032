"""torch.autograd wrappers of the batched energy + adjoint kernels.

    E = pose_energy(batch, xyz)          # (npose,) E_inter + E_intra, xyz (n_atoms, 3)
    E = dof_energy(batch, dofs)          # (npose,) from Optim.apply_config DOF vectors
    E.sum().backward()                   # xyz.grad / dofs.grad from the adjoint kernels

Both run the forward and the adjoint in one launch sequence (the adjoint kernels need
only the forward's inputs) and keep dE/dinput for the backward, which scales it by the
incoming per-pose gradient.  The energy is the scoring used by the Monte Carlo loop:
`Mol.ene_inter_UFF_shifted_*` (src/mol.ml:822-960) or the grid interpolation
(src/mol.ml:1012-1020), plus `Mol.ene_intra_UFFNB_brute` (src/mol.ml:881-903);
DOF vectors follow src/optim.ml:19-22 and :64-80.  Inputs are FP64 CUDA tensors; the
product path is the HIP library only (there is no CPU fallback).
"""
import torch

from . import ff

# Mol.check_elongation_exn with Const.charged_cutoff (src/mol.ml:589-591, src/const.ml:12)
TOO_LONG_RADIUS = 12.0

def _pose_index(offsets, device):
    off = torch.as_tensor(offsets, dtype=torch.int64)
    counts = off[1:] - off[:-1]
    return torch.repeat_interleave(torch.arange(len(counts)), counts).to(device)


def _row_pose(batch, which, device):
    """Row -> pose map of `batch` (``which`` = "atom" rows or "dof" entries) on `device`,
    kept on the batch object itself so that it lives and dies with the batch."""
    maps = batch.__dict__.setdefault("_row_pose", {})
    key = (which, str(device))
    if key not in maps:
        maps[key] = _pose_index(batch.atom_off if which == "atom" else batch.dof_off, device)
    return maps[key]


def _check_input(t, rows, what):
    if not t.is_cuda or t.dtype != torch.float64:
        raise TypeError("%s must be a CUDA float64 tensor" % what)
    if t.numel() != rows:
        raise ValueError("%s has %d values, the batch needs %d" % (what, t.numel(), rows))


class PoseEnergy(torch.autograd.Function):
    """E_p = E_inter(x_p) + E_intra(x_p) for every pose p of `batch`."""

    @staticmethod
    def forward(ctx, xyz, batch, flags):
        _check_input(xyz, 3 * batch.natoms, "xyz")
        x = xyz.detach().contiguous()
        e = torch.empty(batch.npose, 4, dtype=torch.float64, device=x.device)
        g = torch.empty(batch.natoms, 3, dtype=torch.float64, device=x.device)
        need = ctx.needs_input_grad[0]
        batch.score_dev(x, e, g if need else None, flags | (ff.F_GRAD if need else 0))
        ctx.batch = batch
        ctx.save_for_backward(g if need else None)
        return e[:, 0] + e[:, 1]

    @staticmethod
    def backward(ctx, gout):
        (g,) = ctx.saved_tensors
        if g is None:
            return None, None, None
        w = gout[_row_pose(ctx.batch, "atom", g.device)]
        return (w[:, None] * g).view(-1, 3), None, None


class DofEnergy(torch.autograd.Function):
    """E_p of the pose built from DOF vector p (rigid body + torsions).

    Optim.apply_config raises Mol.Too_long when the torsions make the ligand longer
    than 12 A (src/optim.ml:72-75) and its callers then score the worst value; here
    such a pose gets E = +inf and a zero gradient."""

    @staticmethod
    def forward(ctx, dofs, batch, flags):
        _check_input(dofs, int(batch.dof_off[-1]), "dofs")
        d = dofs.detach().contiguous()
        e = torch.empty(batch.npose, 4, dtype=torch.float64, device=d.device)
        rad = torch.empty(batch.npose, dtype=torch.float64, device=d.device)
        need = ctx.needs_input_grad[0]
        gd = torch.empty_like(d) if need else None
        batch.score_dofs_dev(d, e, gdof=gd, radius=rad, flags=flags | (ff.F_GRAD if need else 0))
        too_long = rad > TOO_LONG_RADIUS
        if need:
            gd.masked_fill_(too_long[_row_pose(batch, "dof", gd.device)], 0.0)
        ctx.batch = batch
        ctx.save_for_backward(gd)
        return torch.where(too_long, torch.full_like(rad, float("inf")), e[:, 0] + e[:, 1])

    @staticmethod
    def backward(ctx, gout):
        (gd,) = ctx.saved_tensors
        if gd is None:
            return None, None, None
        return gout[_row_pose(ctx.batch, "dof", gd.device)] * gd, None, None


def pose_energy(batch, xyz, flags=ff.F_INTER_SHIFTED | ff.F_INTRA):
    return PoseEnergy.apply(xyz, batch, flags)


def dof_energy(batch, dofs, flags=ff.F_INTER_SHIFTED | ff.F_INTRA):
    return DofEnergy.apply(dofs, batch, flags)
