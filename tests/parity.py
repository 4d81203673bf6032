"""Parity measures against the FP64 oracle (test helpers).

Energy tolerance 1e-5, gradient 1e-4, defined as:
  rel  = |E - E64| / |E64|   on poses with |E64| >= 1e-3 * sum|e_pair|
  cond = |E - E64| / sum|e_pair|  on all poses
and for gradients component-wise |g - g64| / max|g64| (per pose).
"""
import numpy as np

from oracle import twin

E_TOL = 1e-5
G_TOL = 1e-4


def inter_abs_sum(prot, xyz, lig, shifted=True):
    """sum over pairs of |w (ke qi qj / r + vdW)| (numpy, FP64)."""
    pxyz = np.stack([prot["xs"], prot["ys"], prot["zs"]], 1)
    r2, r, elec, vdw = twin._pair_terms(pxyz, np.asarray(prot["q"]), prot["anum"], np.asarray(xyz),
                                        np.asarray(lig["q"]), lig["anum"])
    w = np.where(r2 < 144.0, (1.0 - (r / 12.0) ** 2) ** 2, 0.0) if shifted else 1.0
    return float(np.sum(np.abs(w * (twin.KE * elec + vdw))))


def intra_abs_sum(lig, xyz):
    xyz = np.asarray(xyz)
    n = len(xyz)
    _, _, elec, vdw = twin._pair_terms(xyz, np.asarray(lig["q"]), lig["anum"], xyz,
                                       np.asarray(lig["q"]), lig["anum"])
    dm = np.asarray(lig["dists"]).reshape(n, n)
    iu = np.triu(np.ones((n, n), dtype=bool), 1) & (dm >= 3)
    return float(np.sum(np.abs(twin.KE * elec[iu] + vdw[iu])))


def energy_errors(e, e64, abs_sum):
    e, e64, abs_sum = (np.atleast_1d(np.asarray(v, dtype=np.float64)) for v in (e, e64, abs_sum))
    err = np.abs(e - e64)
    cond = err / np.maximum(abs_sum, 1e-300)
    big = np.abs(e64) >= 1e-3 * abs_sum
    rel = np.where(big, err / np.maximum(np.abs(e64), 1e-300), 0.0)
    return float(rel.max()), float(cond.max())


def assert_energy(e, e64, abs_sum, tol=E_TOL, what=""):
    rel, cond = energy_errors(e, e64, abs_sum)
    assert rel <= tol and cond <= tol, "%s: rel %.3g cond %.3g (tol %g)" % (what, rel, cond, tol)
    return rel, cond


def assert_grad(g, g64, tol=G_TOL, what=""):
    g, g64 = np.asarray(g), np.asarray(g64)
    scale = max(np.abs(g64).max(), 1e-300)
    err = np.abs(g - g64).max() / scale
    assert err <= tol, "%s: grad err %.3g (tol %g)" % (what, err, tol)
    return err
