"""Python handle on libmmo_ff.so: contexts, ligand registry, batches, rigid scans.

``Context`` owns one GPU context (``mmo_ctx``): the receptor, the UFF table, grids and
registered ligands.  Everything computes on the GPU through the C ABI; there is no
CPU path.  Molecules are passed as ``mmo_amd.mol.Mol`` objects or plain dicts with the
``Mol.t`` fields (``xs, ys, zs, q, anum``; ``dists`` for E_intra).
"""
import ctypes as ct
import sys

import numpy as np

from . import _lib
from ._lib import lib, MMOError

GLOBAL, SHIFTED, INTERP = _lib.GLOBAL, _lib.SHIFTED, _lib.INTERP
F_INTER_SHIFTED, F_INTER_GLOBAL, F_INTER_INTERP = (_lib.F_INTER_SHIFTED, _lib.F_INTER_GLOBAL,
                                                   _lib.F_INTER_INTERP)
F_INTRA, F_GRAD, F_COMPONENTS = _lib.F_INTRA, _lib.F_GRAD, _lib.F_COMPONENTS
_MODE_FLAG = {SHIFTED: F_INTER_SHIFTED, GLOBAL: F_INTER_GLOBAL, INTERP: F_INTER_INTERP}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _ptr(a):
    return ct.c_void_p(a.ctypes.data) if a is not None else None


def _field(m, *names):
    for n in names:
        if isinstance(m, dict):
            if n in m:
                return m[n]
        elif hasattr(m, n):
            return getattr(m, n)
    raise KeyError("molecule has none of %s" % (names,))


def mol_arrays(m):
    """(xs, ys, zs, q, anum) of a Mol or dict as contiguous arrays."""
    return (_f64(_field(m, "xs")), _f64(_field(m, "ys")), _f64(_field(m, "zs")),
            _f64(_field(m, "q", "q_a")), _i32(_field(m, "anum", "elt_a")))


def dists_colmajor(m):
    """Topological distances as stored in Mol.t.dists_a: flat[i + j*n] = row i, column j
    of the .pqrs matrix (src/pqrs.ml:62-77)."""
    d = _field(m, "dists", "dists_a")
    n = len(_field(m, "xs"))
    d = np.asarray(d, dtype=np.int32)
    if d.ndim == 2:
        return _i32(d.T.reshape(-1))
    return _i32(d.reshape(-1))


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ct.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))


class Context:
    """One GPU scoring context (receptor, UFF table, grids, ligand registry)."""

    def __init__(self, device=0):
        h = ct.c_void_p()
        self._check(lib.mmo_ctx_create(int(device), ct.byref(h)), None)
        self.h = h
        self.device = device
        self.nrec = 0
        self.grid_shape = None
        self._ligs = []
        self._nrb = []

    def __del__(self):
        h = getattr(self, "h", None)
        # at interpreter exit the HIP runtime may already be gone: the process frees it all
        if h and not sys.is_finalizing():
            lib.mmo_ctx_destroy(h)
            self.h = None

    def _check(self, rc, h=-1):
        if rc != 0:
            msg = lib.mmo_last_error(self.h if h == -1 else h)
            raise MMOError(rc, msg.decode() if msg else "")
        return rc

    # ------------------------------------------------------------------ setup
    def set_uff_table(self, table):
        """table: {anum: (x_i, D_i)}; replaces UFF.anum_xi_Di (src/UFF.ml:10-22)."""
        an = _i32(list(table))
        x = _f64([table[a][0] for a in table])
        d = _f64([table[a][1] for a in table])
        self._check(lib.mmo_set_uff_table(self.h, _ptr(an), _ptr(x), _ptr(d), len(an)))

    def set_receptor(self, prot, origin=None):
        xs, ys, zs, q, an = mol_arrays(prot)
        o = _f64(origin) if origin is not None else None
        self._check(lib.mmo_set_receptor(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), _ptr(an),
                                         len(xs), _ptr(o)))
        self.nrec = len(xs)

    def set_grids(self, grids, step, origin):
        """grids: (ntypes, nz, ny, nx) FP32 (x fastest, src/grid.ml:98-99)."""
        g = np.ascontiguousarray(grids, dtype=np.float32)
        nt, nz, ny, nx = g.shape
        ptrs = (ct.c_void_p * nt)(*[g[t].ctypes.data for t in range(nt)])
        o = _f64(origin)
        self._check(lib.mmo_set_grids(self.h, ptrs, nt, float(step), nx, ny, nz, _ptr(o)))
        self.grid_shape = (nt, nz, ny, nx)
        self._grids_keep = g

    def compute_grids(self, probe_anum, probe_q, step, dims, origin, roi_center=None,
                      roi_r=0.0, max_e=1e5, copy_out=True):
        """GPU grid precompute for probe types (anum, q); installs the grids for the
        interpolated form and returns them as (ntypes, nz, ny, nx) FP32 if copy_out."""
        an = _i32(probe_anum)
        q = _f64(probe_q)
        nx, ny, nz = (int(v) for v in dims)
        o = _f64(origin)
        rc = _f64(roi_center) if roi_center is not None else None
        out = np.zeros((len(an), nz, ny, nx), dtype=np.float32) if copy_out else None
        self._check(lib.mmo_grid_compute(self.h, _ptr(an), _ptr(q), len(an), float(step), nx, ny, nz,
                                         _ptr(o), _ptr(rc), float(roi_r), float(max_e), _ptr(out)))
        self.grid_shape = (len(an), nz, ny, nx)
        return out

    @property
    def fixed_point_bits(self):
        return lib.mmo_fixed_point_bits(self.h)

    # ------------------------------------------------------------ plain forms
    def ene_inter(self, lig, mode=SHIFTED, components=False):
        xs, ys, zs, q, an = mol_arrays(lig)
        e, el, vd = ct.c_double(), ct.c_double(), ct.c_double()
        self._check(lib.mmo_ene_inter(self.h, int(mode), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q),
                                      _ptr(an), len(xs), ct.byref(e), ct.byref(el), ct.byref(vd)))
        return (e.value, el.value, vd.value) if components else e.value

    def ene_inter_grad(self, lig, mode=SHIFTED):
        xs, ys, zs, q, an = mol_arrays(lig)
        n = len(xs)
        g = np.zeros((3, n))
        e = ct.c_double()
        self._check(lib.mmo_ene_inter_grad(self.h, int(mode), _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q),
                                           _ptr(an), n, ct.byref(e), _ptr(g[0]), _ptr(g[1]),
                                           _ptr(g[2])))
        return e.value, g.T.copy()

    def ene_intra(self, lig):
        xs, ys, zs, q, an = mol_arrays(lig)
        d = dists_colmajor(lig)
        e = ct.c_double()
        self._check(lib.mmo_ene_intra(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), _ptr(an),
                                      _ptr(d), len(xs), ct.byref(e)))
        return e.value

    def ene_intra_grad(self, lig):
        xs, ys, zs, q, an = mol_arrays(lig)
        d = dists_colmajor(lig)
        n = len(xs)
        g = np.zeros((3, n))
        e = ct.c_double()
        self._check(lib.mmo_ene_intra_grad(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), _ptr(an),
                                           _ptr(d), n, ct.byref(e), _ptr(g[0]), _ptr(g[1]),
                                           _ptr(g[2])))
        return e.value, g.T.copy()

    def ene_inter_interp(self, lig, types=None):
        xs, ys, zs = (_f64(_field(lig, k)) for k in ("xs", "ys", "zs"))
        t = _i32(types if types is not None else _field(lig, "types_id", "t_a"))
        e = ct.c_double()
        self._check(lib.mmo_ene_inter_interp(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(t), len(xs),
                                             ct.byref(e)))
        return e.value

    def ene_inter_interp_grad(self, lig, types=None):
        xs, ys, zs = (_f64(_field(lig, k)) for k in ("xs", "ys", "zs"))
        t = _i32(types if types is not None else _field(lig, "types_id", "t_a"))
        n = len(xs)
        g = np.zeros((3, n))
        e = ct.c_double()
        self._check(lib.mmo_ene_inter_interp_grad(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(t), n,
                                                  ct.byref(e), _ptr(g[0]), _ptr(g[1]), _ptr(g[2])))
        return e.value, g.T.copy()

    # --------------------------------------------------------- ligand registry
    def add_ligand(self, lig, types=None):
        """Registers a centred ligand (topology, rotatable bonds, grid types); returns its id."""
        xs, ys, zs, q, an = mol_arrays(lig)
        d = dists_colmajor(lig)
        n = len(xs)
        rbl = _i32(_field(lig, "rb_left")) if _has(lig, "rb_left") else _i32([])
        rbr = _i32(_field(lig, "rb_right")) if _has(lig, "rb_right") else _i32([])
        rgo = _i32(_field(lig, "rg_off")) if _has(lig, "rg_off") else _i32([0])
        rga = _i32(_field(lig, "rg_atoms")) if _has(lig, "rg_atoms") else _i32([])
        if types is None and _has(lig, "types_id"):
            types = _field(lig, "types_id")
        t = _i32(types) if types is not None else None
        out = ct.c_int()
        self._check(lib.mmo_ligand_add(self.h, _ptr(xs), _ptr(ys), _ptr(zs), _ptr(q), _ptr(an),
                                       _ptr(d), n, _ptr(rbl), _ptr(rbr), _ptr(rgo),
                                       _ptr(rga if len(rga) else None), len(rbl), _ptr(t),
                                       ct.byref(out)))
        self._ligs.append(n)
        self._nrb.append(len(rbl))
        return out.value

    def ligand_natoms(self, lig_id):
        return self._ligs[lig_id]

    def ligand_nrb(self, lig_id):
        return self._nrb[lig_id]

    def batch(self, pose_lig):
        return Batch(self, pose_lig)

    # --------------------------------------------------------------- scans
    def scan(self, lig_id, mode, rots, trans, out=None, stream=None):
        """Rigid scan on the GPU; returns a torch FP64 tensor (nrot * ntrans) on the device."""
        import torch
        r = _f64(np.asarray(rots).reshape(-1, 9))
        t = _f64(np.asarray(trans).reshape(-1, 3))
        n = len(r) * len(t)
        if out is None:
            out = torch.empty(n, dtype=torch.float64, device="cuda:%d" % self.device)
        assert out.is_cuda and out.dtype == torch.float64 and out.numel() >= n and out.is_contiguous()
        self._check(lib.mmo_scan_dev(self.h, int(lig_id), int(mode), _ptr(r), len(r), _ptr(t), len(t),
                                     ct.c_void_p(out.data_ptr()), _stream_ptr(stream)))
        return out

    def set_vdw_mask(self, step, dims, origin=None, radii=None):
        """Protein vdW volume bitmask on the grid origin + (i, j, k) * step (Lds.vdW_volume);
        radii default to the reference's per-element vdW radii."""
        r = _f64(radii) if radii is not None else None
        o = _f64(origin) if origin is not None else None
        nx, ny, nz = (int(v) for v in dims)
        self._check(lib.mmo_set_vdw_mask(self.h, _ptr(r), float(step), nx, ny, nz, _ptr(o)))

    def vdw_mask(self):
        """The mask's 32-bit words (bit v of word v // 32 is voxel v = i + j*nx + k*nx*ny)."""
        n = lib.mmo_get_vdw_mask(self.h, None)
        if n < 0:
            raise MMOError(n, "no vdW mask")
        w = np.zeros(n, dtype=np.uint32)
        lib.mmo_get_vdw_mask(self.h, _ptr(w))
        return w

    def desolv_setup(self, step, dims, origin=None, roi=None, radii=None):
        """The receptor's first solvent shell and per-node protein desolvation contributions
        on the grid origin + (i, j, k) * step (Lds.first_solvent_shell and
        Lds.protein_desolv, src/lds.ml:171-234); roi = (center, radius) limits the
        contributions to the shell nodes inside the sphere.  Returns the number of nodes
        given a contribution."""
        r = _f64(radii) if radii is not None else None
        o = _f64(origin) if origin is not None else None
        c = _f64(roi[0]) if roi is not None else None
        nx, ny, nz = (int(v) for v in dims)
        rc = lib.mmo_desolv_setup(self.h, _ptr(r), float(step), nx, ny, nz, _ptr(o), _ptr(c),
                                  float(roi[1]) if roi is not None else 0.0)
        if rc < 0:
            self._check(rc)
        return rc

    def solvent_shell(self):
        """The receptor's first solvent shell as 32-bit words (bit v of word v // 32 is node
        v = i + j*nx + k*nx*ny)."""
        n = lib.mmo_get_solvent_shell(self.h, None)
        if n < 0:
            raise MMOError(n, "no desolvation data")
        w = np.zeros(n, dtype=np.uint32)
        lib.mmo_get_solvent_shell(self.h, _ptr(w))
        return w

    def desolv_contribs(self):
        """The per-node protein desolvation contributions (flat, node i + j*nx + k*nx*ny)."""
        n = lib.mmo_get_desolv_contribs(self.h, None)
        if n < 0:
            self._check(n)
        out = np.zeros(n, dtype=np.float64)
        self._check(min(0, lib.mmo_get_desolv_contribs(self.h, _ptr(out))))
        return out

    def desolvation_penalty(self, xyz, q, radii):
        """Lds.desolvation_penalty (src/lds.ml:236-267) of one ligand pose: (prot, lig)."""
        x = _f64(np.asarray(xyz).reshape(-1, 3))
        qq = _f64(q)
        rr = _f64(radii)
        if len(qq) != len(x) or len(rr) != len(x):
            raise ValueError("xyz, q and radii disagree on the number of atoms")
        prot = ct.c_double()
        lig = ct.c_double()
        self._check(lib.mmo_desolv_penalty(self.h, _ptr(x), _ptr(qq), _ptr(rr), len(x),
                                           ct.byref(prot), ct.byref(lig)))
        return prot.value, lig.value

    def scan_filter(self, lig_id, rots, trans, e, roi=None, lig_radii=None, flags=None,
                    stream=None):
        """Clash filters of the exhaustive scan on the energies `e` of ctx.scan(lig_id, ...,
        rots, trans): skipped poses become +inf.  roi = (center, radius)."""
        r = _f64(np.asarray(rots).reshape(-1, 9))
        t = _f64(np.asarray(trans).reshape(-1, 3))
        if flags is None:
            flags = (_lib.SCAN_ROI if roi is not None else 0) | _lib.SCAN_ATOM_CLASH | \
                (_lib.SCAN_CENTER_CLASH if lig_radii is not None else 0)
        c = _f64(roi[0]) if roi is not None else None
        rad = _f64(lig_radii) if lig_radii is not None else None
        self._check(lib.mmo_scan_filter_dev(self.h, int(lig_id), _ptr(r), len(r), _ptr(t), len(t),
                                            _ptr(c), float(roi[1]) if roi is not None else 0.0,
                                            _ptr(rad), int(flags), ct.c_void_p(e.data_ptr()),
                                            _stream_ptr(stream)))
        return e

    def scan_filtered(self, lig_id, mode, rots, trans, roi=None, lig_radii=None, flags=None,
                      out=None, stream=None):
        """Filter before scoring (Lds.exhaustive_rigid_ligand_docking, src/lds.ml:1040-1114):
        the clash filters of scan_filter first, then only the kept poses scored; returns
        (energies, number of poses scored) with +inf for every skipped pose, the same as
        scan() followed by scan_filter()."""
        import torch
        r = _f64(np.asarray(rots).reshape(-1, 9))
        t = _f64(np.asarray(trans).reshape(-1, 3))
        if flags is None:
            flags = (_lib.SCAN_ROI if roi is not None else 0) | _lib.SCAN_ATOM_CLASH | \
                (_lib.SCAN_CENTER_CLASH if lig_radii is not None else 0)
        if out is None:
            out = torch.empty(len(r) * len(t), dtype=torch.float64, device="cuda:%d" % self.device)
        c = _f64(roi[0]) if roi is not None else None
        rad = _f64(lig_radii) if lig_radii is not None else None
        nk = ct.c_int64()
        self._check(lib.mmo_scan_filtered_dev(self.h, int(lig_id), int(mode), _ptr(r), len(r), _ptr(t),
                                              len(t), _ptr(c), float(roi[1]) if roi is not None else 0.0,
                                              _ptr(rad), int(flags), ct.c_void_p(out.data_ptr()),
                                              ct.byref(nk), _stream_ptr(stream)))
        return out, nk.value

    def topk(self, e, k, stream=None):
        """The k lowest values of the CUDA FP64 tensor e and their indices (ascending)."""
        import torch
        n = e.numel()
        kk = min(int(k), n)
        vals = torch.empty(kk, dtype=torch.float64, device=e.device)
        idx = torch.empty(kk, dtype=torch.int64, device=e.device)
        self._check(lib.mmo_topk_dev(self.h, ct.c_void_p(e.data_ptr()), ct.c_int64(n), int(k),
                                     ct.c_void_p(vals.data_ptr()), ct.c_void_p(idx.data_ptr()),
                                     _stream_ptr(stream)))
        return vals, idx

    def topk_noalloc(self, e, k, vals, idx, stream=None):
        """topk into caller tensors (vals FP64, idx int64, min(k, n) each) through the
        entry point that allocates nothing (per-GPU static scratch; calls that overlap
        on one GPU race)."""
        n = e.numel()
        self._check(lib.mmo_topk_dev_noalloc(self.h, ct.c_void_p(e.data_ptr()), ct.c_int64(n),
                                             int(k), ct.c_void_p(vals.data_ptr()),
                                             ct.c_void_p(idx.data_ptr()), _stream_ptr(stream)))
        return vals, idx

    def tune(self, **kv):
        for k, v in kv.items():
            self._check(lib.mmo_tune_set(self.h, k.encode(), float(v)))

    def tuning(self, key):
        return lib.mmo_tune_get(self.h, key.encode())

    def last_kernel_ms(self, name):
        return lib.mmo_last_kernel_ms(self.h, name.encode())


def _has(m, name):
    return (name in m) if isinstance(m, dict) else hasattr(m, name)


class Batch:
    """A fixed list of poses (pose p is a conformation of ligand pose_lig[p])."""

    def __init__(self, ctx, pose_lig):
        self.ctx = ctx
        pl = _i32(pose_lig)
        h = ct.c_void_p()
        ctx._check(lib.mmo_batch_create(ctx.h, _ptr(pl), len(pl), ct.byref(h)))
        self.h = h
        self.npose = len(pl)
        self.natoms = lib.mmo_batch_natoms(h)
        self.atom_off = np.zeros(self.npose + 1, dtype=np.int32)
        lib.mmo_batch_atom_offsets(h, _ptr(self.atom_off))
        self.dof_off = np.zeros(self.npose + 1, dtype=np.int32)
        lib.mmo_batch_dof_offsets(h, _ptr(self.dof_off))

    def __del__(self):
        h = getattr(self, "h", None)
        # at interpreter exit the HIP runtime may already be gone: the process frees it all
        if h and not sys.is_finalizing():
            lib.mmo_batch_destroy(h)
            self.h = None

    def score(self, xyz, flags):
        """Host form: xyz (natoms, 3) FP64 -> (e (npose, 4), grad (natoms, 3) or None)."""
        x = _f64(xyz).reshape(-1)
        assert x.size == 3 * self.natoms
        e = np.zeros((self.npose, 4))
        g = np.zeros((self.natoms, 3)) if flags & F_GRAD else None
        self.ctx._check(lib.mmo_batch_score(self.h, int(flags), _ptr(x), _ptr(e), _ptr(g)))
        return e, g

    def score_dev(self, xyz, e_out, grad=None, flags=F_INTER_SHIFTED, stream=None):
        """Device form on torch CUDA FP64 tensors (enqueued on `stream`)."""
        self.ctx._check(lib.mmo_batch_score_dev(
            self.h, int(flags), ct.c_void_p(xyz.data_ptr()), ct.c_void_p(e_out.data_ptr()),
            ct.c_void_p(grad.data_ptr()) if grad is not None else None, _stream_ptr(stream)))

    def score_dofs_dev(self, dofs, e_out, xyz_out=None, gdof=None, radius=None,
                       flags=F_INTER_SHIFTED | F_INTRA, stream=None):
        """Optim.apply_config form on torch CUDA FP64 tensors."""
        def p(t):
            return ct.c_void_p(t.data_ptr()) if t is not None else None
        self.ctx._check(lib.mmo_batch_score_dofs_dev(self.h, int(flags), p(dofs), p(xyz_out),
                                                     p(e_out), p(gdof), p(radius),
                                                     _stream_ptr(stream)))
