"""Independent numpy FP64 restatement of the scoring formulas -- TEST INFRASTRUCTURE ONLY.

Written as vectorised pair matrices (not loops) so that it shares no code with the
C oracle; tests cross-check the two.  Formulas: src/mol.ml:796-903 (BrG, BrL, intra),
src/G3D.ml:97-157 (trilinear), src/UFF.ml:10-51 (pair parameters).
"""
import numpy as np

KE = 332.0637 / 4.0
_XD = {0: (0.0, 0.0), 1: (2.886, 0.044), 6: (3.851, 0.105), 7: (3.660, 0.069),
       8: (3.500, 0.060), 9: (3.364, 0.050), 12: (3.021, 0.111), 15: (4.147, 0.305),
       16: (4.035, 0.274), 17: (3.947, 0.227), 35: (4.189, 0.251), 53: (4.500, 0.339)}


def _params(anum):
    x = np.array([_XD[int(a)][0] for a in anum])
    d = np.array([_XD[int(a)][1] for a in anum])
    return x, d


def _pair_terms(pxyz, pq, pa, lxyz, lq, la):
    d = pxyz[:, None, :] - lxyz[None, :, :]
    r2 = np.einsum("ijk,ijk->ij", d, d)
    r = np.maximum(np.sqrt(r2), 0.01)
    px, pd = _params(pa)
    lx, ld = _params(la)
    xij = np.sqrt(np.outer(px, lx))
    dij = np.sqrt(np.outer(pd, ld))
    p6 = (xij / r) ** 6
    elec = np.outer(pq, lq) / r
    vdw = dij * (p6 * p6 - 2.0 * p6)
    return r2, r, elec, vdw


def ene_inter(prot, lig, shifted=True):
    pxyz = np.stack([prot["xs"], prot["ys"], prot["zs"]], 1)
    lxyz = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    r2, r, elec, vdw = _pair_terms(pxyz, np.asarray(prot["q"]), prot["anum"], lxyz,
                                   np.asarray(lig["q"]), lig["anum"])
    if shifted:
        w = np.where(r2 < 144.0, (1.0 - (r / 12.0) ** 2) ** 2, 0.0)
        w = np.where(r < 12.0, w, 0.0)
    else:
        w = np.ones_like(r)
    return KE * np.sum(w * elec) + np.sum(w * vdw)


def ene_intra(lig):
    xyz = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    n = len(xyz)
    _, _, elec, vdw = _pair_terms(xyz, np.asarray(lig["q"]), lig["anum"], xyz,
                                  np.asarray(lig["q"]), lig["anum"])
    dm = np.asarray(lig["dists"]).reshape(n, n)
    iu = np.triu(np.ones((n, n), dtype=bool), 1) & (dm >= 3)
    return KE * np.sum(elec[iu]) + np.sum(vdw[iu])


def trilin(step, arr_zyx, p):
    """Trilinear value as the weighted average over the 8 corners of the cell
    containing p (grid origin 0, index x fastest)."""
    arr = np.asarray(arr_zyx, dtype=np.float64)
    f = np.asarray(p, dtype=np.float64) / step
    i0 = np.trunc(f).astype(int)
    t = np.asarray(p, dtype=np.float64) / step - i0
    val = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                wgt = ((t[0] if dx else 1 - t[0]) * (t[1] if dy else 1 - t[1]) *
                       (t[2] if dz else 1 - t[2]))
                val += wgt * arr[i0[2] + dz, i0[1] + dy, i0[0] + dx]
    return val
