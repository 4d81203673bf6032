"""FP64 CPU oracle for MMO's UFF nonbonded scoring path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg may
import this package.  It is the checker, never the thing measured or shipped:
``mmo_amd`` never imports it.

``oracle.lib`` is the C restatement (``mmo_oracle.c``) loaded with ctypes; the
functions below wrap it with numpy arrays.  ``oracle.twin`` is an independent numpy
restatement used to cross-check the C code.  Each wrapped function follows the
reference function named in ``mmo_oracle.h`` (UnixJunkie/MMO ``src/<file>:<line>``).
"""
import ctypes as _ct
import os as _os
import subprocess as _sp

import numpy as _np

_HERE = _os.path.dirname(_os.path.abspath(__file__))
_SO = _os.path.join(_HERE, "_build", "libmmo_oracle.so")


def build():
    """Compile the C oracle (gcc, -O3 -fno-fast-math -fopenmp)."""
    _sp.run(["make", "-s", "-C", _HERE], check=True)


if not _os.path.exists(_SO):
    build()
lib = _ct.CDLL(_SO)

_d = _ct.c_double
_i = _ct.c_int
_p = _ct.c_void_p

ELEC_WEIGHT = 332.0637 / 4.0          # src/UFF.ml:25
UFF_TABLE = {0: (0.0, 0.0), 1: (2.886, 0.044), 6: (3.851, 0.105), 7: (3.660, 0.069),
             8: (3.500, 0.060), 9: (3.364, 0.050), 12: (3.021, 0.111), 15: (4.147, 0.305),
             16: (4.035, 0.274), 17: (3.947, 0.227), 35: (4.189, 0.251), 53: (4.500, 0.339)}


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


for _n in ("orc_non_zero_dist", "orc_shift_12A", "orc_pow6"):
    _sig(_n, _d, [_d])
_sig("orc_elec_weight", _d, [])
_sig("orc_uff_pair", None, [_i, _i, _p, _p])
_mol = [_p, _p, _p, _p, _p, _i]
_sig("orc_ene_inter_global", _d, _mol + _mol + [_p, _p])
_sig("orc_ene_inter_shifted", _d, _mol + _mol + [_p, _p])
_sig("orc_ene_intra", _d, [_p, _p, _p, _p, _p, _p, _i])


class Grid(_ct.Structure):
    """Grid.t with origin (0,0,0), src/grid.ml:9-19."""
    _fields_ = [("step", _d), ("one_div_step", _d), ("x_dim", _i), ("y_dim", _i), ("z_dim", _i)]

    @classmethod
    def make(cls, step, dims):
        return cls(step, 1.0 / step, int(dims[0]), int(dims[1]), int(dims[2]))


_sig("orc_trilin", _d, [_p, _p, _d, _d, _d])
_sig("orc_ene_inter_interp", _d, [_p, _p, _p, _p, _p, _p, _i])
_sig("orc_grid_energies", None, _mol + [_p, _p, _p, _p, _i, _p])
_sig("orc_bitmask_roi", None, [_p, _d, _d, _d, _d, _p])
for _n in ("orc_rot_x", "orc_rot_y", "orc_rot_z"):
    _sig(_n, None, [_d, _p])
_sig("orc_rot_xyz", None, [_d, _d, _d, _p])
_sig("orc_rot_mult", None, [_p, _p, _p])
_sig("orc_rot_of_quat", None, [_d, _d, _d, _d, _p])
_sig("orc_rot_of_axis_angle", None, [_d, _d, _d, _d, _p])
_sig("orc_quat_to_axis_angle", None, [_p, _p, _p])
_sig("orc_so3_quat", None, [_i, _i, _p])
_sig("orc_so3_rotations", None, [_i, _p])
_sig("orc_dihedral", _d, [_p, _p, _p, _p])
_sig("orc_rotate_bond", None, [_p, _p, _p, _i, _i, _p, _i, _d])
_sig("orc_center", None, [_p, _p, _p, _i, _p])
_sig("orc_radius", _d, [_p, _p, _p, _i, _p])
_sig("orc_center_rotate_translate", None, [_p, _p, _p, _i, _p, _p, _p, _p, _p, _p])
_sig("orc_apply_config", _d, [_p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p])
_sig("orc_apply_config_jvp", _d, [_p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p])
_sig("orc_grad_inter", _d, [_i] + _mol + _mol + [_p, _p, _p])
_sig("orc_grad_intra", _d, [_p, _p, _p, _p, _p, _p, _i, _p, _p, _p])
_sig("orc_grad_interp", _d, [_p, _p, _p, _p, _p, _p, _i, _p, _p, _p])
_sig("orc_scan_shifted", None, _mol + [_p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _i, _p])
_sig("orc_scan_shifted_cells", None, _mol + [_p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _i, _p])
_sig("orc_scan_interp", None, [_p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _i, _p])


def _f64(a):
    return _np.ascontiguousarray(a, dtype=_np.float64)


def _i32(a):
    return _np.ascontiguousarray(a, dtype=_np.int32)


def _ptr(a):
    return a.ctypes.data_as(_p)


class _Keep:
    """Holds converted arrays alive for the duration of a call."""

    def __init__(self):
        self.keep = []

    def f(self, a):
        a = _f64(a)
        self.keep.append(a)
        return _ptr(a)

    def i(self, a):
        a = _i32(a)
        self.keep.append(a)
        return _ptr(a)


def _molargs(k, m):
    return [k.f(m["xs"]), k.f(m["ys"]), k.f(m["zs"]), k.f(m["q"]), k.i(m["anum"]), len(m["xs"])]


def non_zero_dist(r):
    return lib.orc_non_zero_dist(r)


def shift_12A(d):
    return lib.orc_shift_12A(d)


def uff_pair(a1, a2):
    x = _np.zeros(1)
    d = _np.zeros(1)
    lib.orc_uff_pair(a1, a2, _ptr(x), _ptr(d))
    return float(x[0]), float(d[0])


def ene_inter(prot, lig, shifted=True, components=False):
    """Mol.ene_inter_UFF_shifted_brute (shifted=True) or _global_brute (False)."""
    k = _Keep()
    el = _np.zeros(1)
    vd = _np.zeros(1)
    fn = lib.orc_ene_inter_shifted if shifted else lib.orc_ene_inter_global
    e = fn(*_molargs(k, prot), *_molargs(k, lig), _ptr(el), _ptr(vd))
    return (e, float(el[0]), float(vd[0])) if components else e


def ene_intra(lig):
    """Mol.ene_intra_UFFNB_brute; lig['dists'] is the n x n matrix with dists[i, j] =
    row i column j of the .pqrs matrix (stored column-major i + j*n, src/pqrs.ml:62-77)."""
    k = _Keep()
    n = len(lig["xs"])
    dm = _np.asarray(lig["dists"], dtype=_np.int32).reshape(n, n)
    flat = _i32(dm.T.reshape(-1))   # flat[i + j*n] = dm[i, j]
    k.keep.append(flat)
    return lib.orc_ene_intra(k.f(lig["xs"]), k.f(lig["ys"]), k.f(lig["zs"]), k.f(lig["q"]),
                             k.i(lig["anum"]), _ptr(flat), n)


def _grid_ptrs(grids):
    arrs = [_np.ascontiguousarray(g, dtype=_np.float32).reshape(-1) for g in grids]
    ptrs = (_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def trilin(grid, arr, p):
    a = _np.ascontiguousarray(arr, dtype=_np.float32).reshape(-1)
    return lib.orc_trilin(_ct.byref(grid), _ptr(a), float(p[0]), float(p[1]), float(p[2]))


def ene_inter_interp(grid, grids, lig_xyz, ltype):
    k = _Keep()
    arrs, ptrs = _grid_ptrs(grids)
    x = _np.asarray(lig_xyz, dtype=_np.float64)
    return lib.orc_ene_inter_interp(_ct.byref(grid), ptrs, k.f(x[:, 0]), k.f(x[:, 1]),
                                    k.f(x[:, 2]), k.i(ltype), len(x))


def grid_energies(prot, grid, probe_anum, probe_q, mask=None):
    """Per-type FP32 grids (ntypes, z, y, x) computed as src/lds.ml:452-469."""
    k = _Keep()
    nt = len(probe_anum)
    nvox = grid.x_dim * grid.y_dim * grid.z_dim
    out = [_np.zeros(nvox, dtype=_np.float32) for _ in range(nt)]
    ptrs = (_p * nt)(*[o.ctypes.data for o in out])
    mptr = None
    if mask is not None:
        m = _np.ascontiguousarray(mask, dtype=_np.uint8).reshape(-1)
        k.keep.append(m)
        mptr = _ptr(m)
    lib.orc_grid_energies(*_molargs(k, prot), _ct.byref(grid), mptr, k.i(probe_anum),
                          k.f(probe_q), nt, ptrs)
    return _np.stack([o.reshape(grid.z_dim, grid.y_dim, grid.x_dim) for o in out])


def bitmask_roi(grid, center, out_r):
    m = _np.zeros(grid.x_dim * grid.y_dim * grid.z_dim, dtype=_np.uint8)
    lib.orc_bitmask_roi(_ct.byref(grid), float(center[0]), float(center[1]), float(center[2]),
                        float(out_r), _ptr(m))
    return m


_sig("orc_desolvation_const", _d, [])
_sig("orc_first_solvent_shell", None, [_p, _p, _p, _p, _i, _p, _p])
_sig("orc_protein_desolv", None, [_p, _p, _p, _p, _i, _p, _p, _d, _d, _d, _d, _p])
_sig("orc_desolvation_penalty", None, [_p, _p, _p, _p, _p, _p, _p, _p, _i, _p, _p])


def desolvation_const():
    return lib.orc_desolvation_const()


def first_solvent_shell(grid, xyz, radii):
    """Lds.first_solvent_shell on the grid with lowest node (0, 0, 0): uint8 per node."""
    k = _Keep()
    x = _np.asarray(xyz, dtype=_np.float64)
    m = _np.zeros(grid.x_dim * grid.y_dim * grid.z_dim, dtype=_np.uint8)
    lib.orc_first_solvent_shell(k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]), k.f(radii), len(x),
                                _ct.byref(grid), _ptr(m))
    return m


def protein_desolv(grid, prot, shell, roi_center, roi_r):
    """Lds.protein_desolv: per node (FP64, 0 outside the shell and the ROI)."""
    k = _Keep()
    out = _np.zeros(grid.x_dim * grid.y_dim * grid.z_dim)
    sh = _np.ascontiguousarray(shell, dtype=_np.uint8)
    k.keep.append(sh)
    lib.orc_protein_desolv(k.f(prot["xs"]), k.f(prot["ys"]), k.f(prot["zs"]), k.f(prot["q"]),
                           len(prot["xs"]), _ct.byref(grid), _ptr(sh), float(roi_center[0]),
                           float(roi_center[1]), float(roi_center[2]), float(roi_r), _ptr(out))
    return out


def desolvation_penalty(grid, prot_contribs, prot_shell, xyz, q, radii):
    """Lds.desolvation_penalty: (prot, lig) of one ligand pose."""
    k = _Keep()
    x = _np.asarray(xyz, dtype=_np.float64)
    pc = _np.ascontiguousarray(prot_contribs, dtype=_np.float64)
    sh = _np.ascontiguousarray(prot_shell, dtype=_np.uint8)
    k.keep += [pc, sh]
    p, l = _ct.c_double(), _ct.c_double()
    lib.orc_desolvation_penalty(_ct.byref(grid), _ptr(pc), _ptr(sh), k.f(x[:, 0]), k.f(x[:, 1]),
                                k.f(x[:, 2]), k.f(q), k.f(radii), len(x), _ct.byref(p), _ct.byref(l))
    return p.value, l.value


def _mat(fn, *args):
    m = _np.zeros(9)
    fn(*args, _ptr(m))
    return m.reshape(3, 3)


def rot_x(t):
    return _mat(lib.orc_rot_x, t)


def rot_y(t):
    return _mat(lib.orc_rot_y, t)


def rot_z(t):
    return _mat(lib.orc_rot_z, t)


def rot_xyz(a, b, g):
    return _mat(lib.orc_rot_xyz, a, b, g)


def rot_mult(r1, r2):
    a = _f64(r1).reshape(-1)
    b = _f64(r2).reshape(-1)
    m = _np.zeros(9)
    lib.orc_rot_mult(_ptr(a), _ptr(b), _ptr(m))
    return m.reshape(3, 3)


def rot_of_quat(w, x, y, z):
    return _mat(lib.orc_rot_of_quat, w, x, y, z)


def rot_of_axis_angle(axis, theta):
    return _mat(lib.orc_rot_of_axis_angle, float(axis[0]), float(axis[1]), float(axis[2]), theta)


def so3_quat(n, i):
    q = _np.zeros(4)
    lib.orc_so3_quat(n, i, _ptr(q))
    return q


def so3_rotations(n):
    m = _np.zeros(9 * n)
    lib.orc_so3_rotations(n, _ptr(m))
    return m.reshape(n, 3, 3)


def dihedral(i, j, k, l):
    a = [_f64(v) for v in (i, j, k, l)]
    return lib.orc_dihedral(*[_ptr(v) for v in a])


def rotate_bond(xyz, left, right, rgroup, alpha):
    """Returns a rotated copy (src/mol.ml:610-631)."""
    x = _np.array(xyz, dtype=_np.float64)
    xs, ys, zs = _f64(x[:, 0]), _f64(x[:, 1]), _f64(x[:, 2])
    rg = _i32(rgroup)
    lib.orc_rotate_bond(_ptr(xs), _ptr(ys), _ptr(zs), left, right, _ptr(rg), len(rg), alpha)
    return _np.stack([xs, ys, zs], axis=1)


def center_rotate_translate(xyz, center, rot, pos):
    k = _Keep()
    x = _np.asarray(xyz, dtype=_np.float64)
    n = len(x)
    o = _np.zeros((3, n))
    lib.orc_center_rotate_translate(k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]), n, k.f(center),
                                    k.f(_np.asarray(rot).reshape(-1)), k.f(pos),
                                    _ptr(o[0]), _ptr(o[1]), _ptr(o[2]))
    return o.T.copy()


def apply_config(centered_xyz, rb_left, rb_right, rg_off, rg_atoms, dofs):
    """Optim.apply_config; returns (coords (n,3), enclosing radius after torsions)."""
    k = _Keep()
    x = _np.asarray(centered_xyz, dtype=_np.float64)
    n = len(x)
    o = _np.zeros((3, n))
    rad = lib.orc_apply_config(k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]), n, k.i(rb_left),
                               k.i(rb_right), k.i(rg_off), k.i(rg_atoms), len(rb_left),
                               k.f(dofs), _ptr(o[0]), _ptr(o[1]), _ptr(o[2]))
    return o.T.copy(), rad


def apply_config_jvp(centered_xyz, rb_left, rb_right, rg_off, rg_atoms, dofs):
    """apply_config with forward-mode derivatives: (coords (n,3), radius, tangents
    (6 + nrb, n, 3)) where tangents[i, j] = d coords[j] / d dofs[i]."""
    k = _Keep()
    x = _np.asarray(centered_xyz, dtype=_np.float64)
    n = len(x)
    nrb = len(rb_left)
    o = _np.zeros((3, n))
    tan = _np.zeros((6 + nrb, n, 3))
    rad = lib.orc_apply_config_jvp(k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]), n, k.i(rb_left),
                                   k.i(rb_right), k.i(rg_off), k.i(rg_atoms), nrb, k.f(dofs),
                                   _ptr(o[0]), _ptr(o[1]), _ptr(o[2]), _ptr(tan))
    return o.T.copy(), rad, tan


def dof_energy_grad(base_xyz, lig, dofs, inter_grad, intra=True):
    """E_inter + E_intra of the pose Optim.apply_config(base_xyz, dofs) and dE/dDOF,
    FP64: dE/dxyz from inter_grad(x) -> (E_inter, (n,3)) and grad_intra, chained through
    the tangent pass (gdof_i = sum_j g_j . d x_j / d dof_i).  dofs has 6 + R entries:
    the first R rotatable bonds of `lig` are turned.  Returns (e_inter, e_intra, gdof,
    radius after the torsions)."""
    dofs = _np.asarray(dofs, dtype=_np.float64)
    r = len(dofs) - 6
    rg_off = _np.asarray(lig["rg_off"])[:r + 1]
    x, rad, tan = apply_config_jvp(base_xyz, _np.asarray(lig["rb_left"])[:r],
                                   _np.asarray(lig["rb_right"])[:r], rg_off,
                                   _np.asarray(lig["rg_atoms"])[:int(rg_off[-1])], dofs)
    e_inter, g = inter_grad(x)
    e_intra = 0.0
    if intra:
        e_intra, gi = grad_intra(dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2]))
        g = g + gi
    return e_inter, e_intra, _np.einsum("ijc,jc->i", tan, g), rad


def grad_inter(prot, lig, shifted=True):
    k = _Keep()
    n = len(lig["xs"])
    g = _np.zeros((3, n))
    e = lib.orc_grad_inter(int(bool(shifted)), *_molargs(k, prot), *_molargs(k, lig),
                           _ptr(g[0]), _ptr(g[1]), _ptr(g[2]))
    return e, g.T.copy()


def grad_intra(lig):
    k = _Keep()
    n = len(lig["xs"])
    dm = _np.asarray(lig["dists"], dtype=_np.int32).reshape(n, n)
    flat = _i32(dm.T.reshape(-1))
    k.keep.append(flat)
    g = _np.zeros((3, n))
    e = lib.orc_grad_intra(k.f(lig["xs"]), k.f(lig["ys"]), k.f(lig["zs"]), k.f(lig["q"]),
                           k.i(lig["anum"]), _ptr(flat), n, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]))
    return e, g.T.copy()


def grad_interp(grid, grids, lig_xyz, ltype):
    k = _Keep()
    arrs, ptrs = _grid_ptrs(grids)
    x = _np.asarray(lig_xyz, dtype=_np.float64)
    n = len(x)
    g = _np.zeros((3, n))
    e = lib.orc_grad_interp(_ct.byref(grid), ptrs, k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]),
                            k.i(ltype), n, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]))
    return e, g.T.copy()


def scan_shifted(prot, centered_lig, rots, pose_rot, trans, pose_trans, nthreads=1):
    """FP64 BrL energies of rigid-scan poses x = R[pose_rot] c + T[pose_trans]."""
    k = _Keep()
    npose = len(pose_rot)
    out = _np.zeros(npose)
    lig = centered_lig
    lib.orc_scan_shifted(*_molargs(k, prot), k.f(lig["xs"]), k.f(lig["ys"]), k.f(lig["zs"]),
                         k.f(lig["q"]), k.i(lig["anum"]), len(lig["xs"]),
                         k.f(_np.asarray(rots).reshape(-1)), k.i(pose_rot),
                         k.f(_np.asarray(trans).reshape(-1)), k.i(pose_trans), npose,
                         int(nthreads), _ptr(out))
    return out


def scan_shifted_cells(prot, centered_lig, rots, pose_rot, trans, pose_trans, nthreads=1):
    """FP64 shifted energies of rigid-scan poses through a neighbour search (the reference's
    default Bst form), the same terms as scan_shifted in another order."""
    k = _Keep()
    npose = len(pose_rot)
    out = _np.zeros(npose)
    lig = centered_lig
    lib.orc_scan_shifted_cells(*_molargs(k, prot), k.f(lig["xs"]), k.f(lig["ys"]), k.f(lig["zs"]),
                               k.f(lig["q"]), k.i(lig["anum"]), len(lig["xs"]),
                               k.f(_np.asarray(rots).reshape(-1)), k.i(pose_rot),
                               k.f(_np.asarray(trans).reshape(-1)), k.i(pose_trans), npose,
                               int(nthreads), _ptr(out))
    return out


def scan_interp(grid, grids, centered_xyz, ltype, rots, pose_rot, trans, pose_trans, nthreads=1):
    k = _Keep()
    arrs, ptrs = _grid_ptrs(grids)
    x = _np.asarray(centered_xyz, dtype=_np.float64)
    npose = len(pose_rot)
    out = _np.zeros(npose)
    lib.orc_scan_interp(_ct.byref(grid), ptrs, k.f(x[:, 0]), k.f(x[:, 1]), k.f(x[:, 2]),
                        k.i(ltype), len(x), k.f(_np.asarray(rots).reshape(-1)), k.i(pose_rot),
                        k.f(_np.asarray(trans).reshape(-1)), k.i(pose_trans), npose,
                        int(nthreads), _ptr(out))
    return out
