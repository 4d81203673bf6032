"""Edge cases and error paths through the C ABI: empty and single-atom inputs, rigid
ligands (no rotatable bonds) in every caller, saturated and out-of-grid atoms, size
limits, and calls made in the wrong order.  Errors must come back as status codes with
a message, never as a fault."""
import math

import numpy as np
import pytest

import oracle
from mmo_amd import ff, synth
from tests.parity import assert_energy, inter_abs_sum
from mmo_amd._lib import (ERR_BAD_ARG, ERR_NO_GRID, ERR_NO_RECEPTOR, ERR_OUT_OF_GRID,
                          MMOError)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    lig, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    return ctx, lig, rec


def _code(fn):
    with pytest.raises(MMOError) as e:
        fn()
    return e.value.code


def test_empty_and_wrong_order_calls():
    ctx = ff.Context(0)
    lig, _ = synth.config0()
    print("edge: calls before set-up", flush=True)
    assert _code(lambda: ctx.ene_inter(lig)) == ERR_NO_RECEPTOR
    assert _code(lambda: ctx.ene_inter_interp(lig, types=np.zeros(75, dtype=np.int32))) == ERR_NO_GRID
    empty = {"xs": np.zeros(0), "ys": np.zeros(0), "zs": np.zeros(0), "q": np.zeros(0),
             "anum": np.zeros(0, dtype=np.int32)}
    assert _code(lambda: ctx.set_receptor(empty)) == ERR_BAD_ARG
    assert _code(lambda: ctx.batch([])) == ERR_BAD_ARG
    assert _code(lambda: ctx.batch([0])) == ERR_BAD_ARG          # no ligand registered
    assert _code(lambda: ctx.scan(0, ff.SHIFTED, np.eye(3)[None], np.zeros((1, 3)))) == ERR_BAD_ARG


def test_single_atom_ligand_every_form(system):
    ctx, _, rec = system
    one = {"xs": np.array([30.0]), "ys": np.zeros(1), "zs": np.zeros(1), "q": np.array([0.3]),
           "anum": np.array([8], dtype=np.int32), "dists": np.zeros((1, 1), dtype=np.int32)}
    print("edge: one-atom ligand", flush=True)
    for mode, shifted in ((ff.SHIFTED, True), (ff.GLOBAL, False)):
        assert ctx.ene_inter(one, mode) == pytest.approx(oracle.ene_inter(rec, one, shifted),
                                                         rel=1e-5, abs=1e-7)
    assert ctx.ene_intra(one) == 0.0
    lid = ctx.add_ligand(dict(one, xs=np.zeros(1)))
    b = ctx.batch([lid, lid, lid])
    xyz = np.array([[30.0, 0, 0], [0.0, 29.0, 0], [200.0, 0, 0]])
    e, g = b.score(xyz, ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD)
    for p in range(3):
        m = dict(one, xs=xyz[p:p + 1, 0], ys=xyz[p:p + 1, 1], zs=xyz[p:p + 1, 2])
        assert e[p, 0] == pytest.approx(oracle.ene_inter(rec, m), rel=1e-5, abs=1e-7)
        assert e[p, 1] == 0.0
    assert e[2, 0] == 0.0 and np.all(g[2] == 0.0)                 # beyond every cutoff


def test_saturated_atoms(system):
    """Atoms beyond the fixed-point range: exactly 0 in the shifted form; the global
    form (no cutoff) sums them like the oracle."""
    ctx, lig, rec = system
    far = dict(lig, xs=np.asarray(lig["xs"]) + 1e6)
    print("edge: saturated coordinates", flush=True)
    assert ctx.ene_inter(far, ff.SHIFTED) == 0.0
    e = ctx.ene_inter(far, ff.GLOBAL)
    ref = oracle.ene_inter(rec, far, shifted=False)
    # the net-neutral ligand's pair terms cancel at 1e6 A: the conditioned measure
    xyz = np.stack([far["xs"], far["ys"], far["zs"]], 1)
    assert_energy(e, ref, inter_abs_sum(rec, xyz, far, False), what="global, 1e6 A")
    # non-finite input: rejected by the plain forms; in batches it counts as out of range
    nan = dict(lig, xs=np.full(len(lig["xs"]), np.nan))
    assert _code(lambda: ctx.ene_inter(nan, ff.SHIFTED)) == ERR_BAD_ARG
    lid = ctx.add_ligand(lig)
    b = ctx.batch([lid])
    e, _ = b.score(np.full((len(lig["xs"]), 3), np.nan), ff.F_INTER_SHIFTED)
    assert e[0, 0] == 0.0


def test_out_of_grid(system):
    ctx, lig, _ = system
    rng = np.random.default_rng(0)
    grids = rng.normal(size=(2, 8, 8, 8)).astype(np.float32)
    ctx.set_grids(grids, 0.5, (0.0, 0.0, 0.0))
    probe = {"xs": np.array([10.0]), "ys": np.array([1.0]), "zs": np.array([1.0])}
    print("edge: out of grid", flush=True)
    assert _code(lambda: ctx.ene_inter_interp(probe, types=[0])) == ERR_OUT_OF_GRID
    assert _code(lambda: ctx.ene_inter_interp(dict(probe, xs=np.array([1.0])), types=[5])) == ERR_BAD_ARG


def test_rigid_ligand_in_dof_mc_and_minimiser(system):
    """A ligand with no rotatable bond: 6 DOFs, torsion frames are no-ops in MC."""
    import torch
    from mmo_amd import mc as mcmod
    from mmo_amd.minimize import Minimizer
    ctx, lig, rec = system
    rigid = dict(lig, rb_left=[], rb_right=[], rg_off=[0], rg_atoms=[])
    lid = ctx.add_ligand(rigid)
    b = ctx.batch([lid] * 4)
    assert int(b.dof_off[-1]) == 24
    d = torch.tensor(np.tile([30.0, 0, 0, 0.1, 0.2, 0.3], 4), device="cuda:0")
    print("edge: rigid ligand minimiser", flush=True)
    e = Minimizer(b, max_step=0.2).run(d, 20).cpu().numpy()
    assert np.all(np.isfinite(e))
    print("edge: rigid ligand MC", flush=True)
    m = mcmod.MonteCarlo(ctx, [lid] * 3, np.tile(np.eye(3), (3, 1, 1)),
                         np.array([[30.0, 0, 0]] * 3), ff.SHIFTED, total_frames=20,
                         flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI,
                         roi_center=np.array([30.0, 0, 0]), roi_r=4.0, seed=2)
    m.run(20)
    st = m.state()
    assert (st["frame"] == 20).all() and (st["status"] == 1).all()
    # torsion frames leave the conformer as registered
    conf, _ = m.coords()
    c = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    assert np.array_equal(conf[0], c)


def test_size_limits(system):
    import torch
    from mmo_amd import mc as mcmod
    from mmo_amd.minimize import Minimizer
    ctx, lig, _ = system
    print("edge: size limits", flush=True)
    # more than 32 rotatable bonds: rejected by the chains, accepted by the batches
    n = 40
    chain = {"xs": np.arange(n) * 1.5, "ys": np.zeros(n), "zs": np.zeros(n), "q": np.zeros(n),
             "anum": np.full(n, 6, dtype=np.int32),
             "dists": np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]).astype(np.int32),
             "rb_left": list(range(0, 36)), "rb_right": list(range(1, 37)),
             "rg_off": [0], "rg_atoms": []}
    groups = [list(range(k + 2, n)) for k in range(36)]
    chain["rg_off"] = np.concatenate([[0], np.cumsum([len(g) for g in groups])])
    chain["rg_atoms"] = np.concatenate(groups)
    chain["xs"] = chain["xs"] - chain["xs"].mean()
    lid = ctx.add_ligand(chain)
    assert _code(lambda: mcmod.MonteCarlo(ctx, [lid], np.eye(3)[None], np.zeros((1, 3)),
                                          ff.SHIFTED, total_frames=1)) == ERR_BAD_ARG
    b = ctx.batch([lid])
    assert int(b.dof_off[-1]) == 42
    d = torch.zeros(42, dtype=torch.float64, device="cuda:0")
    e = torch.zeros((1, 4), dtype=torch.float64, device="cuda:0")
    b.score_dofs_dev(d, e, flags=ff.F_INTER_SHIFTED | ff.F_INTRA)
    # a ligand whose chain state does not fit in one workgroup's LDS (7,000 atoms: 24 B of
    # conformer per atom): rejected when the chains are created, before any launch
    n = 7000
    dists = np.full((n, n), 2, dtype=np.int32)
    np.fill_diagonal(dists, 0)
    big = {"xs": np.linspace(-5.0, 5.0, n), "ys": np.zeros(n), "zs": np.zeros(n),
           "q": np.zeros(n), "anum": np.full(n, 6, dtype=np.int32), "dists": dists}
    bid = ctx.add_ligand(big)
    assert _code(lambda: mcmod.MonteCarlo(ctx, [bid], np.eye(3)[None], np.zeros((1, 3)),
                                          ff.SHIFTED, total_frames=1)) == ERR_BAD_ARG
    # top-k bounds
    with pytest.raises(MMOError):
        ctx.topk(torch.zeros(10, dtype=torch.float64, device="cuda:0"), 5000)
    # minimiser history bounds
    assert _code(lambda: Minimizer(b, history=0)) == ERR_BAD_ARG
    assert _code(lambda: Minimizer(b, history=17)) == ERR_BAD_ARG
    torch.cuda.synchronize()


def _zigzag(n, nrb):
    """Planar zig-zag chain of n carbons (1.5 A bonds, 120 degree angles) centred at the
    origin, with rotatable bonds 1-2, 2-3, ... (nrb of them) turning the atoms beyond."""
    k = np.arange(n)
    xs = 1.5 * math.cos(math.pi / 6) * k
    ys = np.where(k % 2 == 0, 0.0, 0.75)
    c = np.stack([xs, ys, np.zeros(n)], 1)
    c -= c.mean(0)
    groups = [list(range(b + 2, n)) for b in range(1, nrb + 1)]
    return {"xs": c[:, 0], "ys": c[:, 1], "zs": c[:, 2],
            "q": np.where(k % 2 == 0, 0.1, -0.1), "anum": np.full(n, 6, dtype=np.int32),
            "dists": np.abs(k[:, None] - k[None, :]).astype(np.int32),
            "rb_left": list(range(1, nrb + 1)), "rb_right": list(range(2, nrb + 2)),
            "rg_off": np.concatenate([[0], np.cumsum([len(g) for g in groups])]),
            "rg_atoms": np.concatenate(groups).astype(np.int32)}


def test_dof_gradient_intra_routes(system):
    """With the gradient, E_intra comes from the DOF adjoint (on the conformer, before
    the rigid-body motion), without it from the intra kernel (on the posed coordinates):
    the two agree to the energy tolerance, also for a 1200-atom chain whose adjoint
    needs more than 64 KiB of LDS."""
    import torch
    from tests.parity import assert_energy, intra_abs_sum
    ctx, _, _ = system
    rng = np.random.default_rng(11)
    for n in (60, 1200):
        print("edge: intra route, %d-atom chain" % n, flush=True)
        lig = _zigzag(n, 3)
        lid = ctx.add_ligand(lig)
        b = ctx.batch([lid, lid])
        d = torch.tensor(np.concatenate([np.concatenate([[40.0, 40.0, 40.0], rng.uniform(-1, 1, 3),
                                                         rng.uniform(-0.5, 0.5, 3)])
                                         for _ in range(2)]), device="cuda:0")
        e_g = torch.zeros((2, 4), dtype=torch.float64, device="cuda:0")
        e_0 = torch.zeros_like(e_g)
        x_t = torch.zeros((b.natoms, 3), dtype=torch.float64, device="cuda:0")
        g = torch.zeros_like(d)
        b.score_dofs_dev(d, e_g, xyz_out=x_t, gdof=g, flags=ff.F_INTRA | ff.F_GRAD)
        b.score_dofs_dev(d, e_0, flags=ff.F_INTRA)
        torch.cuda.synchronize()
        eg, e0 = e_g.cpu().numpy()[:, 1], e_0.cpu().numpy()[:, 1]
        assert np.isfinite(g.cpu().numpy()).all()
        x = x_t.cpu().numpy()
        sums = [intra_abs_sum(lig, x[n * p:n * (p + 1)]) for p in range(2)]
        assert_energy(eg, e0, sums, what="adjoint intra, %d atoms" % n)


def _grazing_poses(lig, rec, n, seed):
    """Poses of `lig` pushed out of the receptor along random directions until their
    nearest receptor atom lies at a chosen distance in [10.5, 11.99] A (so that only a
    few pairs, all near the 12 A cutoff, contribute).  Returns [(R, T, x, d_min)]."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.stack([rec["xs"], rec["ys"], rec["zs"]], 1))
    c = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        R = synth.random_rotations(rng, 1)[0]
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        target = rng.uniform(10.5, 11.99)
        lo, hi = 0.0, 100.0
        for _ in range(60):
            m = 0.5 * (lo + hi)
            if tree.query(c @ R.T + m * u)[0].min() < target:
                lo = m
            else:
                hi = m
        x = c @ R.T + hi * u
        out.append((R, hi * u, x, float(tree.query(x)[0].min())))
    return out


def test_grazing_poses_carry_the_fp32_shift_weight_error(system):
    """Poses whose every pair inside 12 A lies within ~1.5 A of the cutoff.  The FP32
    kernels form the shift weight w = t^2, t = 1 - r^2/144, from an FP32 r^2 (~1e-7
    relative), so t carries an absolute error of ~1.2e-7 and w a relative one of
    ~2.4e-7 / t, which only matters when the pairs with t -> 0 dominate the sum.  Against
    the FP64 oracle: within 1e-5 sum|w e| + 3e-7 sum t|e| for every pose (scan kernel
    and batched pose kernel), and within 1e-5 (the parity bar) for every pose with a
    pair inside 11.7 A (t >= 0.05).  Such poses score below 0.1 kcal/mol."""
    from oracle import twin
    ctx, lig, rec = system
    lid = ctx.add_ligand(lig)
    poses = _grazing_poses(lig, rec, 48, seed=12)
    xyz = np.concatenate([p[2] for p in poses])
    print("grazing poses: batch and scan", flush=True)
    eb, _ = ctx.batch([lid] * len(poses)).score(xyz, ff.F_INTER_SHIFTED)
    pxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    near_bar = 0
    for k, (R, T, x, dmin) in enumerate(poses):
        es = float(ctx.scan(lid, ff.SHIFTED, R[None], T[None]).cpu().numpy()[0])
        m = dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
        ref = oracle.ene_inter(rec, m, shifted=True)
        r2, r, elec, vdw = twin._pair_terms(pxyz, np.asarray(rec["q"]), rec["anum"], x,
                                            np.asarray(lig["q"]), lig["anum"])
        t = np.where(r2 < 144.0, 1.0 - r2 / 144.0, 0.0)
        e_abs = np.abs(twin.KE * elec + vdw)
        s1, s2 = float(np.sum(t * t * e_abs)), float(np.sum(t * e_abs))
        assert abs(ref) < 0.1
        for name, e in (("batch", eb[k, 0]), ("scan", es)):
            assert abs(e - ref) <= 1e-5 * s1 + 3e-7 * s2, (k, name, e, ref, dmin)
            if dmin <= 11.7:
                assert_energy(e, ref, s1, what="pose %d %s (nearest %.2f A)" % (k, name, dmin))
        near_bar += dmin <= 11.7
    assert 10 <= near_bar <= 40
