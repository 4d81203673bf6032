"""CPU restatement of the exhaustive scan's clash filters and top-k -- TEST
INFRASTRUCTURE ONLY (the checker; the product path is libmmo_ff.so).

Follows Lds.vdW_volume / atom_bitmask_set (src/lds.ml:164-197), G3D.vdW_clash_OR /
vdW_clash_AND (src/G3D.ml:159-213), Mol.protein_ligand_clash and
Mol.is_ligand_center_vdW_occuppied (src/mol.ml:1195-1218), ROI.is_inside
(src/ROI.ml:65-66) and the TopKeeper ordering of src/lds.ml:1055-1064, :1110-1113.
Grid nodes are origin + i * step (the reference's L.frange may differ by an ulp:
parity unpinned at that level); corners outside the grid count as free.
"""
import math

import numpy as np

# Ptable.vdW_radii, src/ptable.ml:43-53
VDW_RADII = {1: 1.2, 6: 1.7, 7: 1.6, 8: 1.55, 9: 1.5, 12: 2.2, 15: 1.95, 16: 1.8, 17: 1.8,
             35: 1.9, 53: 2.1}


def vdw_mask(xyz, radii, step, dims, origin=(0.0, 0.0, 0.0)):
    """Boolean (nx, ny, nz) array: node strictly inside some atom's vdW sphere."""
    nx, ny, nz = dims
    o = np.asarray(origin, dtype=np.float64)
    m = np.zeros((nx, ny, nz), dtype=bool)
    for p, r in zip(np.asarray(xyz, dtype=np.float64), radii):
        ci, cj, ck = (math.trunc((p[a] - o[a]) / step) for a in range(3))   # coord_of_point
        rs = int(math.ceil(r / step))
        for i in range(max(ci - rs, 0), min(ci + rs, nx - 1) + 1):
            dx = p[0] - (o[0] + i * step)
            for j in range(max(cj - rs, 0), min(cj + rs, ny - 1) + 1):
                dy = p[1] - (o[1] + j * step)
                for k in range(max(ck - rs, 0), min(ck + rs, nz - 1) + 1):
                    dz = p[2] - (o[2] + k * step)
                    if dx * dx + dy * dy + dz * dz < r * r:
                        m[i, j, k] = True
    return m


def mask_words(m):
    """The (nx, ny, nz) mask as 32-bit words, voxel v = i + j*nx + k*nx*ny at bit v % 32."""
    flat = np.transpose(m, (2, 1, 0)).reshape(-1)          # x fastest
    n = len(flat)
    padded = np.zeros(((n + 31) // 32) * 32, dtype=np.uint64)
    padded[:n] = flat
    bits = padded.reshape(-1, 32) << np.arange(32, dtype=np.uint64)
    return bits.sum(axis=1).astype(np.uint32)


def _corners(m, step, origin, p):
    nx, ny, nz = m.shape
    f = (np.asarray(p, dtype=np.float64) - np.asarray(origin)) * (1.0 / step)
    if not (f[0] > -1.0 and f[1] > -1.0 and f[2] > -1.0 and f[0] < nx - 1 and f[1] < ny - 1
            and f[2] < nz - 1):
        return None
    i0, j0, k0 = (math.trunc(v) for v in f)
    return [m[i0 + (c & 1), j0 + ((c >> 1) & 1), k0 + (c >> 2)] for c in range(8)]


def clash_or(m, step, origin, p):
    c = _corners(m, step, origin, p)
    return c is not None and any(c)


def clash_and(m, step, origin, p):
    c = _corners(m, step, origin, p)
    return c is not None and all(c)


def center_occupied(xyz, radii):
    xyz = np.asarray(xyz, dtype=np.float64)
    c = xyz.mean(axis=0)
    return bool(np.any(np.linalg.norm(xyz - c, axis=1) < np.asarray(radii)))


def pose_skipped(m, step, origin, centred_xyz, lig_radii, R, T, roi=None, atom_clash=True,
                 centre_clash=True):
    """Whether the exhaustive scan skips pose (R, T) (src/lds.ml:1088-1095)."""
    T = np.asarray(T, dtype=np.float64)
    if roi is not None and not (np.sum((np.asarray(roi[0]) - T) ** 2) < roi[1] ** 2):
        return True
    if centre_clash and center_occupied(centred_xyz, lig_radii) and clash_and(m, step, origin, T):
        return True
    if atom_clash:
        x = np.asarray(centred_xyz) @ np.asarray(R).T + T
        return any(clash_or(m, step, origin, p) for p in x)
    return False


def topk(e, k):
    """(values, indices) of the k lowest energies, ascending, ties by index, NaN last."""
    e = np.asarray(e, dtype=np.float64)
    key = np.where(np.isnan(e), np.inf, e)
    nan_last = np.isnan(e).astype(np.int8)
    order = np.lexsort((np.arange(len(e)), key, nan_last))[:k]
    return e[order], order
