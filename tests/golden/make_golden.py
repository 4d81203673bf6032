"""Writes tests/golden/golden_v1.npz: FP64 reference values of the scoring path on
seeded inputs, from the C oracle, each checked against the independent numpy
restatement (oracle/twin.py) before it is written.

    python tests/golden/make_golden.py

Contents (all inputs are stored, so the fixture does not depend on the generators):
  mini_*      8 small systems (receptor P = 6..40 atoms, ligand L = 4..12 atoms):
              E_inter shifted and global with elec / vdW parts, E_intra, and the
              analytic dE/dxyz of the shifted E_inter and of E_intra;
  c0_*        the 75-atom / 3000-atom workload: 256 rigid poses (rotation,
              translation) and their shifted E_inter;
  interp_*    two FP32 grids of 9 x 8 x 7 points and 200 probe points with their
              trilinear values (points inside the grid);
  dof_*       a flexible library ligand, 6 DOF vectors and the coordinates
              Optim.apply_config gives for them.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle                      # noqa: E402
from oracle import twin            # noqa: E402
from mmo_amd import synth          # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_v1.npz")
ELEMENTS = np.array([1, 6, 7, 8, 9, 15, 16, 17, 35, 53])


def mini_system(rng, P, L):
    rec = {"xs": rng.uniform(-6, 6, P), "ys": rng.uniform(-6, 6, P), "zs": rng.uniform(-6, 6, P),
           "q": rng.normal(0, 0.3, P), "anum": rng.choice(ELEMENTS, P).astype(np.int32)}
    # a chain ligand: topological distance |i - j|
    steps = rng.normal(size=(L, 3))
    steps = 1.4 * steps / np.linalg.norm(steps, axis=1)[:, None]
    x = np.cumsum(steps, 0) + rng.uniform(-3, 3, 3)
    idx = np.arange(L)
    lig = {"xs": x[:, 0], "ys": x[:, 1], "zs": x[:, 2], "q": rng.normal(0, 0.25, L),
           "anum": rng.choice(ELEMENTS, L).astype(np.int32),
           "dists": np.abs(idx[:, None] - idx[None, :]).astype(np.int32)}
    return rec, lig


def close(a, b, rtol):
    return np.allclose(a, b, rtol=rtol, atol=rtol * max(1.0, float(np.max(np.abs(b)))))


def main():
    rng = np.random.default_rng(20260)
    out = {}
    # ---- mini systems
    for s in range(8):
        P, L = int(rng.integers(6, 41)), int(rng.integers(4, 13))
        rec, lig = mini_system(rng, P, L)
        es, els, vds = oracle.ene_inter(rec, lig, shifted=True, components=True)
        eg, elg, vdg = oracle.ene_inter(rec, lig, shifted=False, components=True)
        ei = oracle.ene_intra(lig)
        _, gs = oracle.grad_inter(rec, lig, shifted=True)
        _, gi = oracle.grad_intra(lig)
        assert close(es, twin.ene_inter(rec, lig, True), 1e-12)
        assert close(eg, twin.ene_inter(rec, lig, False), 1e-12)
        assert close(ei, twin.ene_intra(lig), 1e-12)
        for k in ("xs", "ys", "zs", "q", "anum"):
            out["mini%d_rec_%s" % (s, k)] = np.asarray(rec[k])
            out["mini%d_lig_%s" % (s, k)] = np.asarray(lig[k])
        out["mini%d_lig_dists" % s] = lig["dists"]
        out["mini%d_e" % s] = np.array([es, els, vds, eg, elg, vdg, ei])
        out["mini%d_grad_inter" % s] = gs
        out["mini%d_grad_intra" % s] = gi
    # ---- config-0 workload, 256 rigid poses
    lig, rec = synth.config0()
    rots = synth.random_rotations(rng, 256)
    trans = rng.uniform(-4.75, 4.75, (256, 3))
    e = oracle.scan_shifted(rec, lig, rots, np.arange(256, dtype=np.int32), trans,
                            np.arange(256, dtype=np.int32), nthreads=8)
    for p in range(0, 256, 32):
        x = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1) @ rots[p].T + trans[p]
        m = dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
        assert close(e[p], twin.ene_inter(rec, m, True), 1e-11)
    for k in ("xs", "ys", "zs", "q", "anum"):
        out["c0_rec_%s" % k] = np.asarray(rec[k])
        out["c0_lig_%s" % k] = np.asarray(lig[k])
    out["c0_rots"] = rots
    out["c0_trans"] = trans
    out["c0_e"] = np.asarray(e)
    # ---- trilinear interpolation
    dims, step = (9, 8, 7), 0.5
    grids = rng.normal(0, 10, (2, dims[2], dims[1], dims[0])).astype(np.float32)
    g = oracle.Grid.make(step, dims)
    pts = rng.uniform(0, 1, (200, 3)) * (np.array(dims) - 1.0001) * step
    vals = np.array([[oracle.trilin(g, grids[t], p) for p in pts] for t in range(2)])
    for t in range(2):
        for p in range(0, 200, 40):
            assert abs(vals[t, p] - twin.trilin(step, grids[t], pts[p])) <= 1e-9
    out["interp_grids"] = grids
    out["interp_step"] = np.array(step)
    out["interp_points"] = pts
    out["interp_values"] = vals
    # ---- apply_config
    lib = [l for l in synth.library(8, seed=4242) if len(l["rb_left"]) >= 3][0]
    c = np.stack([lib["xs"], lib["ys"], lib["zs"]], 1)
    nrb = len(lib["rb_left"])
    dofs = np.concatenate([rng.uniform(-5, 5, (6, 3)), rng.uniform(-np.pi, np.pi, (6, 3 + nrb))], 1)
    coords = np.stack([oracle.apply_config(c, lib["rb_left"], lib["rb_right"], lib["rg_off"],
                                           lib["rg_atoms"], d)[0] for d in dofs])
    for k in ("xs", "ys", "zs", "q", "anum", "rb_left", "rb_right", "rg_off", "rg_atoms"):
        out["dof_lig_%s" % k] = np.asarray(lib[k])
    out["dof_lig_dists"] = np.asarray(lib["dists"])
    out["dof_dofs"] = dofs
    out["dof_coords"] = coords
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
