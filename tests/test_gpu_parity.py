"""GPU parity of the gfx950 kernels against the FP64 oracle, through the C ABI.

Tolerances (tests/parity.py): energies 1e-5 (rel on well-conditioned poses and
conditioned on all), gradients 1e-4 of the largest component; the interpolated form
uses identical FP32 grids on both sides and is held to 1e-12."""
import math

import numpy as np
import pytest

import oracle
from mmo_amd import _lib, ff, synth
from mmo_amd._lib import MMOError
from tests.parity import assert_energy, assert_grad, inter_abs_sum, intra_abs_sum

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sys0():
    lig, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    return ctx, lig, rec


def _place(lig, R, t):
    xyz = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1) @ np.asarray(R).T + np.asarray(t)
    return dict(lig, xs=xyz[:, 0].copy(), ys=xyz[:, 1].copy(), zs=xyz[:, 2].copy()), xyz


def _poses(n, seed=0, spread=4.75):
    rng = np.random.default_rng(seed)
    return synth.random_rotations(rng, n), rng.uniform(-spread, spread, (n, 3))


def test_plain_inter_shifted_and_global(sys0):
    ctx, lig, rec = sys0
    rots, trans = _poses(12)
    for mode, shifted in ((ff.SHIFTED, True), (ff.GLOBAL, False)):
        es, refs, sums = [], [], []
        for R, t in zip(rots, trans):
            l, xyz = _place(lig, R, t)
            print("plain inter mode", mode, flush=True)
            e, el, vd = ctx.ene_inter(l, mode, components=True)
            r, rel_, rvd = oracle.ene_inter(rec, l, shifted=shifted, components=True)
            es.append(e)
            refs.append(r)
            sums.append(inter_abs_sum(rec, xyz, l, shifted))
            assert el + vd == pytest.approx(e, rel=1e-12, abs=1e-9)
        assert_energy(es, refs, sums, what="inter mode %d" % mode)


def test_plain_inter_components(sys0):
    ctx, lig, rec = sys0
    l, xyz = _place(lig, np.eye(3), [0.5, -1.0, 2.0])
    e, el, vd = ctx.ene_inter(l, ff.SHIFTED, components=True)
    r, rel_, rvd = oracle.ene_inter(rec, l, shifted=True, components=True)
    s = inter_abs_sum(rec, xyz, l, True)
    assert abs(el - rel_) / s < 1e-5 and abs(vd - rvd) / s < 1e-5


def test_plain_inter_far_away(sys0):
    """A ligand 200 A away (beyond the fixed-point window too): exactly 0 in the shifted
    form; the global form (no cutoff, src/mol.ml:796-818) sums every pair like the
    oracle, energy and gradient, also in a batch next to an in-window pose."""
    ctx, lig, rec = sys0
    l, xyz = _place(lig, np.eye(3), [200.0, 0.0, 0.0])
    print("far away: plain forms", flush=True)
    assert ctx.ene_inter(l, ff.SHIFTED) == 0.0
    e = ctx.ene_inter(l, ff.GLOBAL)
    ref = oracle.ene_inter(rec, l, shifted=False)
    s = inter_abs_sum(rec, xyz, l, False)
    assert math.isfinite(e)
    assert_energy(e, ref, s, what="global, 200 A")
    eg, g = ctx.ene_inter_grad(l, ff.GLOBAL)
    e64, g64 = oracle.grad_inter(rec, l, shifted=False)
    assert_energy(eg, e64, s, what="global, 200 A, gradient call")
    assert_grad(g, g64, what="global grad, 200 A")
    print("far away: batch of a near and a far pose", flush=True)
    lid = ctx.add_ligand(lig)
    b = ctx.batch([lid, lid])
    near, nxyz = _place(lig, np.eye(3), [0.5, -1.0, 2.0])
    eb, gb = b.score(np.concatenate([nxyz, xyz]), ff.F_INTER_GLOBAL | ff.F_GRAD)
    assert_energy(eb[0, 0], oracle.ene_inter(rec, near, shifted=False),
                  inter_abs_sum(rec, nxyz, near, False), what="global batch, near")
    assert_energy(eb[1, 0], ref, s, what="global batch, far")
    n = len(lig["xs"])
    assert_grad(gb[n:], g64, what="global batch grad, far")


def test_plain_intra(sys0):
    ctx, lig, rec = sys0
    e = ctx.ene_intra(lig)
    xyz = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    assert_energy(e, oracle.ene_intra(lig), intra_abs_sum(lig, xyz), tol=1e-10, what="intra")


def test_unsupported_element_rejected(sys0):
    ctx, lig, rec = sys0
    bad = dict(lig, anum=np.where(np.arange(len(lig["xs"])) == 3, 5, lig["anum"]))
    with pytest.raises(MMOError) as e:
        ctx.ene_inter(bad, ff.SHIFTED)
    assert e.value.code == -1


def test_translated_frame_gives_same_energies(sys0):
    """The reference works in a box frame at +36..+80 A (src/lds.ml:20-41): moving
    receptor and ligand together leaves the GPU energies unchanged to 1e-5."""
    ctx, lig, rec = sys0
    off = np.array([55.0, 61.0, 48.0])
    rec2 = dict(rec, xs=rec["xs"] + off[0], ys=rec["ys"] + off[1], zs=rec["zs"] + off[2])
    c2 = ff.Context(0)
    c2.set_receptor(rec2)
    rots, trans = _poses(4, seed=3)
    for R, t in zip(rots, trans):
        l, xyz = _place(lig, R, t)
        l2, xyz2 = _place(lig, R, t + off)
        s = inter_abs_sum(rec, xyz, l, True)
        assert abs(ctx.ene_inter(l) - c2.ene_inter(l2)) / s < 1e-5


def _grids(rng, nt=3, n=24, step=0.5, origin=(-6.0, -6.0, -6.0)):
    return rng.normal(0, 5, (nt, n, n, n)).astype(np.float32), step, np.array(origin)


def test_plain_interp_matches_oracle_exactly():
    rng = np.random.default_rng(7)
    grids, step, origin = _grids(rng)
    ctx = ff.Context(0)
    ctx.set_grids(grids, step, origin)
    g = oracle.Grid.make(step, (24, 24, 24))
    for _ in range(5):
        xyz = rng.uniform(-5.5, 5.0, (30, 3))
        types = rng.integers(0, 3, 30)
        l = {"xs": xyz[:, 0], "ys": xyz[:, 1], "zs": xyz[:, 2]}
        e = ctx.ene_inter_interp(l, types)
        r = oracle.ene_inter_interp(g, list(grids), xyz - origin, types)
        assert e == pytest.approx(r, rel=1e-12, abs=1e-12)
        e2, gr = ctx.ene_inter_interp_grad(l, types)
        r2, grr = oracle.grad_interp(g, list(grids), xyz - origin, types)
        assert e2 == pytest.approx(r2, rel=1e-12, abs=1e-12)
        assert np.allclose(gr, grr, rtol=1e-10, atol=1e-10)
    out = {"xs": np.array([7.0]), "ys": np.array([0.0]), "zs": np.array([0.0])}
    with pytest.raises(MMOError) as e:
        ctx.ene_inter_interp(out, [0])
    assert e.value.code == -4


def test_plain_gradients(sys0):
    ctx, lig, rec = sys0
    rots, trans = _poses(3, seed=11)
    for R, t in zip(rots, trans):
        l, xyz = _place(lig, R, t)
        for mode, shifted in ((ff.SHIFTED, True), (ff.GLOBAL, False)):
            e, g = ctx.ene_inter_grad(l, mode)
            r, gr = oracle.grad_inter(rec, l, shifted=shifted)
            assert_energy(e, r, inter_abs_sum(rec, xyz, l, shifted), what="grad-E")
            assert_grad(g, gr, what="inter grad mode %d" % mode)
    e, g = ctx.ene_intra_grad(lig)
    r, gr = oracle.grad_intra(lig)
    assert e == pytest.approx(r, rel=1e-10)
    assert_grad(g, gr, tol=1e-10, what="intra grad")


# ----------------------------------------------------------------- batches
@pytest.fixture(scope="module")
def lib_sys():
    ligs = synth.library(12, seed=99)
    _, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    rng = np.random.default_rng(0)
    grids, step, origin = _grids(rng, nt=4, n=40, step=0.5, origin=(-10.0, -10.0, -10.0))
    ctx.set_grids(grids, step, origin)
    ids = []
    for k, l in enumerate(ligs):
        l["types_id"] = (np.arange(len(l["xs"])) + k) % 4
        ids.append(ctx.add_ligand(l))
    return ctx, ligs, ids, rec, (grids, step, origin)


@pytest.mark.parametrize("pose_sweep", [1, 0])
def test_batch_gradient_rows_fully_written(lib_sys, pose_sweep):
    """The pose kernel writes every atom's gradient row (the buffer is not cleared
    first): a device gradient buffer filled with NaN comes back finite, and a second call
    on the same batch gives bit for bit what a fresh batch gives; under both pose sweeps
    (per atom, and per 8-atom sub-block)."""
    ctx, ligs, ids, rec, _ = lib_sys
    keep = ctx.tuning("pose_sweep")
    ctx.tune(pose_sweep=pose_sweep)
    try:
        _gradient_rows_fully_written(ctx, ligs, ids)
    finally:
        ctx.tune(pose_sweep=keep)


def _gradient_rows_fully_written(ctx, ligs, ids):
    import torch
    rng = np.random.default_rng(12)
    pose_lig = np.asarray([ids[k % len(ids)] for k in range(37)])
    b = ctx.batch(pose_lig)

    def poses():
        x = np.zeros((b.natoms, 3))
        for p, l in enumerate(pose_lig):
            _, xyz = _place(ligs[l], synth.random_rotations(rng, 1)[0], rng.uniform(-3, 3, 3))
            x[b.atom_off[p]:b.atom_off[p + 1]] = xyz
        return x
    x1, x2 = poses(), poses()
    flags = ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD
    print("gradient rows", b.npose, flush=True)
    dev = torch.device("cuda:0")
    xt = torch.tensor(x2, device=dev)
    et = torch.zeros((b.npose, 4), dtype=torch.float64, device=dev)
    gt = torch.full((b.natoms, 3), float("nan"), dtype=torch.float64, device=dev)
    b.score_dev(xt, et, grad=gt, flags=flags)
    torch.cuda.synchronize()
    assert torch.isfinite(gt).all()
    b.score(x1, flags)
    e_again, g_again = b.score(x2, flags)
    e_fresh, g_fresh = ctx.batch(pose_lig).score(x2, flags)
    assert np.array_equal(g_again, g_fresh) and np.array_equal(e_again, e_fresh)
    assert np.array_equal(gt.cpu().numpy(), g_fresh)


def test_batch_ragged_poses(lib_sys):
    ctx, ligs, ids, rec, (grids, step, origin) = lib_sys
    rng = np.random.default_rng(1)
    # uneven pose counts, interleaved ligands: 1, 3, 70, 5, ... poses
    counts = [1, 3, 70, 5, 2, 9, 1, 17, 4, 1, 65, 6]
    pose_lig = np.concatenate([[ids[k]] * c for k, c in enumerate(counts)])
    rng.shuffle(pose_lig)
    b = ctx.batch(pose_lig)
    xyz_all = np.zeros((b.natoms, 3))
    poses = []
    for p, l in enumerate(pose_lig):
        R = synth.random_rotations(rng, 1)[0]
        t = rng.uniform(-3, 3, 3)
        lg, xyz = _place(ligs[l], R, t)
        xyz_all[b.atom_off[p]:b.atom_off[p + 1]] = xyz
        poses.append((lg, xyz))
    print("batch shifted+intra npose", len(pose_lig), flush=True)
    e, _ = b.score(xyz_all, ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_COMPONENTS)
    sample = rng.choice(len(pose_lig), 40, replace=False)
    ref = [oracle.ene_inter(rec, poses[p][0]) for p in sample]
    sums = [inter_abs_sum(rec, poses[p][1], poses[p][0]) for p in sample]
    assert_energy(e[sample, 0], ref, sums, what="batch inter")
    assert np.allclose(e[sample, 2] + e[sample, 3], e[sample, 0], rtol=1e-12, atol=1e-9)
    refi = [oracle.ene_intra(poses[p][0]) for p in sample]
    sumi = [intra_abs_sum(poses[p][0], poses[p][1]) for p in sample]
    assert_energy(e[sample, 1], refi, sumi, tol=1e-10, what="batch intra")
    # interpolated form on the same poses (atoms inside the 19.5 A grid)
    print("batch interp", flush=True)
    e2, _ = b.score(xyz_all, ff.F_INTER_INTERP)
    g = oracle.Grid.make(step, (40, 40, 40))
    for p in sample[:15]:
        lg, xyz = poses[p]
        r = oracle.ene_inter_interp(g, list(grids), xyz - origin, lg["types_id"])
        if math.isnan(r):
            assert math.isnan(e2[p, 0])
        else:
            assert e2[p, 0] == pytest.approx(r, rel=1e-12, abs=1e-12)
    # gradients through the batch (inter + intra)
    print("batch grad", flush=True)
    e3, gb = b.score(xyz_all, ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD)
    for p in sample[:6]:
        lg, xyz = poses[p]
        _, g1 = oracle.grad_inter(rec, lg)
        _, g2 = oracle.grad_intra(lg)
        assert_grad(gb[b.atom_off[p]:b.atom_off[p + 1]], g1 + g2, what="batch grad")


# ----------------------------------------------------------------- rigid scan
def test_rigid_scan_subset_matches_oracle(sys0):
    import torch
    ctx, lig, rec = sys0
    lid = ctx.add_ligand(lig)
    rots = synth.so3_rotations(4608)[::577]           # 8 rotations
    trans = synth.lattice((0, 0, 0), 20, 0.5)[::125]  # 64 translations
    print("scan subset", len(rots) * len(trans), flush=True)
    for mode, shifted in ((ff.SHIFTED, True), (ff.GLOBAL, False)):
        e = ctx.scan(lid, mode, rots, trans).cpu().numpy()
        pr = np.repeat(np.arange(len(rots)), len(trans))
        pt = np.tile(np.arange(len(trans)), len(rots))
        sample = np.arange(0, len(e), 7)
        if shifted:
            ref = oracle.scan_shifted(rec, lig, rots, pr[sample], trans, pt[sample], nthreads=8)
        else:
            ref = []
            for k in sample:
                l, _ = _place(lig, rots[pr[k]], trans[pt[k]])
                ref.append(oracle.ene_inter(rec, l, shifted=False))
        sums = [inter_abs_sum(rec, _place(lig, rots[pr[k]], trans[pt[k]])[1], lig, shifted)
                for k in sample]
        assert_energy(e[sample], ref, sums, what="scan mode %d" % mode)
    torch.cuda.synchronize()


def test_rigid_scan_full_size_sample_and_invariants(sys0):
    """The scan at its full size (4,608 rotations x 20^3 translations = 36,864,000 poses):
    a random sample of 1,024 poses against the oracle, under the library's tuned shape
    and every receptor path (LDS, scalar cache, x-cut cells, positions in LDS, row lists
    through the scalar cache or LDS), which agree with one another to FP32 rounding (the
    paths fold their FP32 partial sums over different groups of receptor atoms)."""
    import torch
    ctx, lig, rec = sys0
    lid = ctx.add_ligand(lig)
    rots = synth.so3_rotations(4608)
    trans = synth.lattice((0, 0, 0), 20, 0.5)
    n = len(rots) * len(trans)
    out = torch.empty(n, dtype=torch.float64, device="cuda:0")
    print("full scan", n, flush=True)
    ctx.scan(lid, ff.SHIFTED, rots, trans, out=out)
    e = out.cpu().numpy()
    assert np.isfinite(e).all()
    rng = np.random.default_rng(5)
    sample = rng.choice(n, 1024, replace=False)
    pr, pt = (sample // len(trans)).astype(np.int32), (sample % len(trans)).astype(np.int32)
    ref = oracle.scan_shifted(rec, lig, rots, pr, trans, pt, nthreads=8)
    sums = [inter_abs_sum(rec, _place(lig, rots[a], trans[b])[1], lig, True) for a, b in zip(pr, pt)]
    assert_energy(e[sample], ref, sums, what="full scan sample")
    keep = ctx.tuning("scan_cull")
    try:
        for sc in (1, 2, 4, 5, 6):
            ctx.tune(scan_cull=sc)
            print("full scan, receptor path", sc, flush=True)
            ctx.scan(lid, ff.SHIFTED, rots, trans, out=out)
            e2 = out.cpu().numpy()
            assert_energy(e2[sample], ref, sums, what="full scan path %d" % sc)
            rel = np.abs(e2 - e) / np.maximum(np.abs(e), 1e-3)
            assert rel.max() < 1e-6, ("path", sc, rel.max())
    finally:
        ctx.tune(scan_cull=keep)
    torch.cuda.synchronize()


def test_rigid_scan_interp_matches_oracle():
    import torch
    lig, rec = synth.config0()
    rng = np.random.default_rng(4)
    n = 64
    grids = rng.normal(0, 3, (len(lig["xs"]), n, n, n)).astype(np.float32)
    origin, _, step = synth.grid_spec(n=n, step=0.375)
    ctx = ff.Context(0)
    ctx.set_grids(grids, step, origin)
    lid = ctx.add_ligand(dict(lig, types_id=np.arange(len(lig["xs"]))))
    rots = synth.so3_rotations(64)
    trans = synth.lattice((0, 0, 0), 4, 0.5)
    e = ctx.scan(lid, ff.INTERP, rots, trans).cpu().numpy()
    g = oracle.Grid.make(step, (n, n, n))
    pr = np.repeat(np.arange(len(rots)), len(trans))
    pt = np.tile(np.arange(len(trans)), len(rots))
    ref = oracle.scan_interp(g, list(grids), np.stack([lig["xs"], lig["ys"], lig["zs"]], 1),
                             np.arange(len(lig["xs"])), rots, pr, trans - origin, pt, nthreads=4)
    ok = ~np.isnan(ref)
    assert ok.sum() > 100
    assert np.array_equal(np.isnan(e), ~ok)
    assert np.allclose(e[ok], ref[ok], rtol=1e-12, atol=1e-12)
    torch.cuda.synchronize()


# ----------------------------------------------------------------- DOF form
def test_dof_coordinates_and_gradient(lib_sys):
    import torch
    ctx, ligs, ids, rec, _ = lib_sys
    pick = [k for k in range(len(ligs)) if len(ligs[k]["rb_left"]) >= 2][:4]
    pose_lig = np.array([ids[k] for k in pick for _ in range(3)])
    b = ctx.batch(pose_lig)
    rng = np.random.default_rng(5)
    dofs = []
    for p, l in enumerate(pose_lig):
        nrb = len(ligs[l]["rb_left"])
        dofs.append(np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-math.pi, math.pi, 3),
                                    rng.uniform(-math.pi, math.pi, nrb)]))
    dflat = np.concatenate(dofs)
    dev = torch.device("cuda:0")
    d_t = torch.tensor(dflat, device=dev)
    xyz_t = torch.zeros((b.natoms, 3), dtype=torch.float64, device=dev)
    e_t = torch.zeros((b.npose, 4), dtype=torch.float64, device=dev)
    g_t = torch.zeros_like(d_t)
    rad_t = torch.zeros(b.npose, dtype=torch.float64, device=dev)
    print("dof batch", b.npose, flush=True)
    b.score_dofs_dev(d_t, e_t, xyz_out=xyz_t, gdof=g_t, radius=rad_t,
                     flags=ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD)
    torch.cuda.synchronize()
    xyz, e, gd, rad = (t.cpu().numpy() for t in (xyz_t, e_t, g_t, rad_t))
    for p, l in enumerate(pose_lig):
        lg = ligs[l]
        c = np.stack([lg["xs"], lg["ys"], lg["zs"]], 1)
        args = (lg["rb_left"], lg["rb_right"], lg["rg_off"], lg["rg_atoms"])
        x_ref, r_ref = oracle.apply_config(c, *args, dofs[p])
        assert np.allclose(xyz[b.atom_off[p]:b.atom_off[p + 1]], x_ref, atol=1e-10)
        assert rad[p] == pytest.approx(r_ref, rel=1e-12)

        def E(d):
            x, _ = oracle.apply_config(c, *args, d)
            m = dict(lg, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
            return oracle.ene_inter(rec, m) + oracle.ene_intra(m)
        assert e[p, 0] + e[p, 1] == pytest.approx(E(dofs[p]), rel=1e-5, abs=1e-6)
        fd = np.zeros(len(dofs[p]))
        for k in range(len(dofs[p])):
            h = 1e-5
            dp, dm = dofs[p].copy(), dofs[p].copy()
            dp[k] += h
            dm[k] -= h
            fd[k] = (E(dp) - E(dm)) / (2 * h)
        got = gd[b.dof_off[p]:b.dof_off[p + 1]]
        assert_grad(got, fd, tol=1e-4, what="dof grad pose %d" % p)


def test_batched_energy_gradient_full_size():
    """The batched energy + gradient workload at its full size: 4 DOF poses of each of
    1,024 library ligands (4,096 poses, 0-15 torsions).  A sample of poses against the
    oracle (energies) and central differences of the oracle energy (DOF gradients), and
    the batched results equal those of one launch sequence per ligand."""
    import torch
    lib = synth.library(1024, seed=3, processes=8)
    _, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    ids = [ctx.add_ligand(l) for l in lib]
    rng = np.random.default_rng(3)
    pose_lig = np.repeat(np.asarray(ids), 4)
    dofs = [np.concatenate([rng.uniform(-1.5, 1.5, 3), rng.uniform(-math.pi, math.pi, 3),
                            rng.uniform(-math.pi, math.pi, len(lib[l]["rb_left"]))]) for l in pose_lig]
    dev = torch.device("cuda:0")
    flags = ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD
    b = ctx.batch(pose_lig)
    d_t = torch.tensor(np.concatenate(dofs), device=dev)
    e_t = torch.zeros((b.npose, 4), dtype=torch.float64, device=dev)
    g_t = torch.zeros_like(d_t)
    print("full egrad batch", b.npose, flush=True)
    b.score_dofs_dev(d_t, e_t, gdof=g_t, flags=flags)
    torch.cuda.synchronize()
    e, gd = e_t.cpu().numpy(), g_t.cpu().numpy()
    assert np.isfinite(e).all() and np.isfinite(gd).all()
    sample = rng.choice(b.npose, 24, replace=False)
    for p in sample:
        lg = lib[pose_lig[p]]
        c = np.stack([lg["xs"], lg["ys"], lg["zs"]], 1)
        args = (lg["rb_left"], lg["rb_right"], lg["rg_off"], lg["rg_atoms"])

        def E(d):
            x, _ = oracle.apply_config(c, *args, d)
            m = dict(lg, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
            return oracle.ene_inter(rec, m) + oracle.ene_intra(m), m, x
        e64, m, x = E(dofs[p])
        assert_energy(e[p, 0] + e[p, 1], e64,
                      inter_abs_sum(rec, x, m) + intra_abs_sum(m, x), what="pose %d" % p)
        fd = np.zeros(len(dofs[p]))
        for k in range(len(dofs[p])):
            dp, dm = dofs[p].copy(), dofs[p].copy()
            dp[k] += 1e-5
            dm[k] -= 1e-5
            fd[k] = (E(dp)[0] - E(dm)[0]) / 2e-5
        assert_grad(gd[b.dof_off[p]:b.dof_off[p + 1]], fd, tol=1e-4, what="dof grad pose %d" % p)
    # one launch sequence per ligand (receptor read from L2 instead of LDS): same sums
    for l in np.unique(pose_lig[sample])[:8]:
        rows = np.flatnonzero(pose_lig == l)
        bl = ctx.batch(pose_lig[rows])
        dl = torch.tensor(np.concatenate([dofs[r] for r in rows]), device=dev)
        el = torch.zeros((bl.npose, 4), dtype=torch.float64, device=dev)
        gl = torch.zeros_like(dl)
        bl.score_dofs_dev(dl, el, gdof=gl, flags=flags)
        torch.cuda.synchronize()
        assert np.array_equal(el.cpu().numpy(), e[rows])
        assert np.array_equal(gl.cpu().numpy(), np.concatenate([gd[b.dof_off[r]:b.dof_off[r + 1]] for r in rows]))


def test_autograd_pose_and_dof_energy(lib_sys):
    """torch.autograd wrappers: xyz.grad against the FP64 analytic gradient of the
    oracle, dofs.grad against central differences of the oracle energy (h = 1e-5: the
    random torsions clash, and at h = 1e-4 the FD truncation error alone reaches 3e-5
    of the largest component), both scaled by the per-pose weights of the backward."""
    import torch
    from mmo_amd.autograd import dof_energy, pose_energy
    ctx, ligs, ids, rec, _ = lib_sys
    pick = [k for k in range(len(ligs)) if len(ligs[k]["rb_left"]) >= 1][:3]
    pose_lig = np.array([ids[k] for k in pick for _ in range(2)])
    b = ctx.batch(pose_lig)
    rng = np.random.default_rng(12)
    dev = torch.device("cuda:0")
    # coordinates form
    xyz = np.zeros((b.natoms, 3))
    mols = []
    for p, l in enumerate(pose_lig):
        lg, x = _place(ligs[l], synth.random_rotations(rng, 1)[0], rng.uniform(-3, 3, 3))
        xyz[b.atom_off[p]:b.atom_off[p + 1]] = x
        mols.append(lg)
    w = rng.uniform(0.5, 2.0, len(pose_lig))
    x_t = torch.tensor(xyz, device=dev, requires_grad=True)
    print("autograd pose energy", len(pose_lig), flush=True)
    E = pose_energy(b, x_t)
    (E * torch.tensor(w, device=dev)).sum().backward()
    torch.cuda.synchronize()
    e, gx = E.detach().cpu().numpy(), x_t.grad.cpu().numpy()
    for p, lg in enumerate(mols):
        ei, gi = oracle.grad_inter(rec, lg)
        ea, ga = oracle.grad_intra(lg)
        assert e[p] == pytest.approx(ei + ea, rel=1e-5, abs=1e-6)
        assert_grad(gx[b.atom_off[p]:b.atom_off[p + 1]], w[p] * (gi + ga), tol=1e-4,
                    what="autograd xyz pose %d" % p)
    # DOF form
    dofs = []
    for l in pose_lig:
        nrb = len(ligs[l]["rb_left"])
        dofs.append(np.concatenate([rng.uniform(-2, 2, 3), rng.uniform(-math.pi, math.pi, 3 + nrb)]))
    d_t = torch.tensor(np.concatenate(dofs), device=dev, requires_grad=True)
    print("autograd dof energy", flush=True)
    E = dof_energy(b, d_t)
    (E * torch.tensor(w, device=dev)).sum().backward()
    torch.cuda.synchronize()
    e, gd = E.detach().cpu().numpy(), d_t.grad.cpu().numpy()
    for p, l in enumerate(pose_lig):
        lg = ligs[l]
        c = np.stack([lg["xs"], lg["ys"], lg["zs"]], 1)
        args = (lg["rb_left"], lg["rb_right"], lg["rg_off"], lg["rg_atoms"])

        def Ef(d):
            x, _ = oracle.apply_config(c, *args, d)
            m = dict(lg, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
            return oracle.ene_inter(rec, m) + oracle.ene_intra(m)
        assert e[p] == pytest.approx(Ef(dofs[p]), rel=1e-5, abs=1e-6)
        fd = np.zeros(len(dofs[p]))
        for k in range(len(dofs[p])):
            h = 1e-5
            dp, dm = dofs[p].copy(), dofs[p].copy()
            dp[k] += h
            dm[k] -= h
            fd[k] = (Ef(dp) - Ef(dm)) / (2 * h)
        assert_grad(gd[b.dof_off[p]:b.dof_off[p + 1]], w[p] * fd, tol=1e-4,
                    what="autograd dof pose %d" % p)


def test_pose_cull_matches_unculled(lib_sys):
    """The pose-culled batch kernel (one wavefront per pose, receptor swept per 8-atom
    sub-block) drops only pairs beyond 12 A: energies, components and gradients agree
    with the unculled kernel, on poses inside, at and away from the receptor."""
    ctx, ligs, ids, rec, _ = lib_sys
    rng = np.random.default_rng(17)
    pose_lig = np.array([ids[k % len(ids)] for k in range(300)])
    b = ctx.batch(pose_lig)
    xyz = np.zeros((b.natoms, 3))
    for p, l in enumerate(pose_lig):
        r = rng.uniform(0, 40)
        d = rng.normal(size=3)
        _, x = _place(ligs[l], synth.random_rotations(rng, 1)[0], r * d / np.linalg.norm(d))
        xyz[b.atom_off[p]:b.atom_off[p + 1]] = x
    flags = ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD
    out = {}
    for pc, sweep in ((0, 1), (1, 1), (1, 0)):
        print("pose cull", pc, "sweep", sweep, b.npose, flush=True)
        ctx.tune(pose_cull=pc, pose_sweep=sweep)
        e, g = b.score(xyz, flags)
        ec, _ = b.score(xyz, ff.F_INTER_SHIFTED | ff.F_COMPONENTS)
        out[(pc, sweep)] = (e, g, ec)
    ctx.tune(pose_cull=1, pose_sweep=1)
    # the kernels sum FP32 pair terms in different orders: errors scale with sum |e_pair|
    scale = np.array([inter_abs_sum(rec, xyz[b.atom_off[p]:b.atom_off[p + 1]], ligs[l])
                      for p, l in enumerate(pose_lig)]) + 1e-3
    e0, g0, c0 = out[(0, 1)]
    for key in ((1, 1), (1, 0)):
        e1, g1, c1 = out[key]
        assert np.all(np.abs(e1[:, 0] - e0[:, 0]) <= 2e-6 * scale), key
        assert np.all(np.abs(c1[:, 2:] - c0[:, 2:]) <= 2e-6 * scale[:, None]), key
        assert np.array_equal(e1[:, 1], e0[:, 1])
        for p in range(b.npose):
            ga, gb = g1[b.atom_off[p]:b.atom_off[p + 1]], g0[b.atom_off[p]:b.atom_off[p + 1]]
            assert np.abs(ga - gb).max() <= 2e-5 * max(np.abs(gb).max(), 1e-3), (key, p)


def test_rigid_scan_culled_equals_unculled_and_oracle(sys0):
    """The culled kernel (cell list in LDS) drops only pairs beyond 12 A."""
    import torch
    ctx, lig, rec = sys0
    lid = ctx.add_ligand(lig)
    rots = synth.so3_rotations(4608)[::288]          # 16 rotations
    trans = synth.lattice((0, 0, 0), 20, 0.5)        # 8000 translations (Morton-ordered bricks)
    print("culled scan", len(rots) * len(trans), flush=True)
    ctx.tune(scan_cull=1)
    ec = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
    ctx.tune(scan_cull=0)
    eu = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
    ctx.tune(scan_cull=1)
    assert np.all(np.isfinite(ec))
    rel = np.abs(ec - eu) / np.maximum(np.abs(eu), 1e-9)
    assert np.percentile(rel, 99.9) < 2e-6, rel.max()
    rng = np.random.default_rng(8)
    sample = rng.choice(len(ec), 48, replace=False)
    pr, pt = sample // len(trans), sample % len(trans)
    ref = oracle.scan_shifted(rec, lig, rots, pr, trans, pt, nthreads=8)
    sums = [inter_abs_sum(rec, _place(lig, rots[a], trans[b])[1], lig, True) for a, b in zip(pr, pt)]
    assert_energy(ec[sample], ref, sums, what="culled scan")
    for k in (2, 4):
        ctx.tune(cull_k=k)
        e2 = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
        assert np.allclose(e2, ec, rtol=2e-6, atol=1e-6)
    ctx.tune(cull_k=1)
    torch.cuda.synchronize()


def test_batch_culled_path(sys0):
    """Batches with >= 1024 wavefront tiles take the culled kernel."""
    ctx, lig, rec = sys0
    lib2 = synth.library(2, seed=5)
    ids = [ctx.add_ligand(l) for l in lib2]
    rng = np.random.default_rng(9)
    npl = 33000
    pose_lig = np.array([ids[0]] * npl + [ids[1]] * npl)
    b = ctx.batch(pose_lig)
    xyz = np.zeros((b.natoms, 3))
    rots = synth.random_rotations(rng, len(pose_lig))
    trans = rng.uniform(-4, 4, (len(pose_lig), 3))
    for p, l in enumerate(pose_lig):
        lg = lib2[ids.index(l)]
        c = np.stack([lg["xs"], lg["ys"], lg["zs"]], 1)
        xyz[b.atom_off[p]:b.atom_off[p + 1]] = c @ rots[p].T + trans[p]
    print("culled batch", len(pose_lig), flush=True)
    e, _ = b.score(xyz, ff.F_INTER_SHIFTED)
    sample = rng.choice(len(pose_lig), 24, replace=False)
    refs, sums = [], []
    for p in sample:
        lg = lib2[ids.index(pose_lig[p])]
        x = xyz[b.atom_off[p]:b.atom_off[p + 1]]
        m = dict(lg, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
        refs.append(oracle.ene_inter(rec, m))
        sums.append(inter_abs_sum(rec, x, m))
    assert_energy(e[sample, 0], refs, sums, what="culled batch")


# ----------------------------------------------------------------- grid precompute
def _abs_sum_at(rec, pts, anum, q):
    """sum_i |w e_pair| of a single-atom probe at each point (conditioning scale)."""
    from oracle import twin
    pxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    out = []
    for p in pts:
        r2, r, elec, vdw = twin._pair_terms(pxyz, np.asarray(rec["q"]), rec["anum"], p[None, :],
                                            np.array([q]), [anum])
        w = np.where(r2 < 144.0, (1.0 - (r / 12.0) ** 2) ** 2, 0.0)
        out.append(float(np.sum(np.abs(w * (twin.KE * elec + vdw)))))
    return np.array(out)


def test_grid_precompute_matches_oracle(sys0):
    ctx, lig, rec = sys0
    dims = (20, 18, 16)
    step = 0.75
    origin = np.array([-7.0, -6.5, -6.0])
    probe_a = np.array([6, 1, 8, 16])
    probe_q = np.array([0.1, -0.3, 0.0, 0.25])
    roi_c, roi_r = np.array([0.5, 0.0, -0.5]), 9.0
    print("grid precompute", dims, flush=True)
    g = ctx.compute_grids(probe_a, probe_q, step, dims, origin, roi_center=roi_c, roi_r=roi_r)
    og = oracle.Grid.make(step, dims)
    mask = oracle.bitmask_roi(og, roi_c - origin, roi_r - 24.0)
    shifted = dict(rec, xs=rec["xs"] - origin[0], ys=rec["ys"] - origin[1], zs=rec["zs"] - origin[2])
    ref = oracle.grid_energies(shifted, og, probe_a, probe_q, mask=mask)
    m3 = mask.reshape(dims[2], dims[1], dims[0]).astype(bool)
    assert np.all(g[:, ~m3] == 0.0)
    assert m3.sum() > 1000 and (~m3).sum() > 100
    rng = np.random.default_rng(3)
    kk, jj, ii = np.nonzero(m3)
    pick = rng.choice(len(ii), 200, replace=False)
    pts = np.stack([ii[pick], jj[pick], kk[pick]], 1) * step + origin
    for t in range(len(probe_a)):
        a = g[t, kk[pick], jj[pick], ii[pick]].astype(np.float64)
        b = ref[t, kk[pick], jj[pick], ii[pick]].astype(np.float64)
        clamped = b >= 1e5
        assert np.array_equal(a >= 1e5, clamped)
        scale = _abs_sum_at(rec, pts, probe_a[t], probe_q[t])
        err = np.abs(a - b)[~clamped] / np.maximum(scale[~clamped], 1e-30)
        assert err.max() < 1e-5, (t, err.max())
    # the computed grids are installed for the interpolated form
    xyz = rng.uniform(-4.0, 4.0, (25, 3))
    types = rng.integers(0, 4, 25)
    e = ctx.ene_inter_interp({"xs": xyz[:, 0], "ys": xyz[:, 1], "zs": xyz[:, 2]}, types)
    r = oracle.ene_inter_interp(og, list(g), xyz - origin, types)
    assert e == pytest.approx(r, rel=1e-12, abs=1e-12)


def test_large_receptor_global_memory_cull():
    """A receptor too large for LDS (15,000 atoms in a 60 A box) takes the culled
    kernels with the cell list read from global memory."""
    import torch
    rng = np.random.default_rng(21)
    rec = synth.make_receptor(rng, 15000, min_sep=2.0, box=60.0)
    lig, _ = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    assert ctx.tuning("cull_ok") == 0.0
    lid = ctx.add_ligand(lig)
    rots = synth.so3_rotations(4608)[::1152]
    trans = synth.lattice((0, 0, 0), 8, 1.0)
    print("global-memory cull scan", len(rots) * len(trans), flush=True)
    ec = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
    sample = rng.choice(len(ec), 16, replace=False)
    pr, pt = sample // len(trans), sample % len(trans)
    ref = oracle.scan_shifted(rec, lig, rots, pr, trans, pt, nthreads=8)
    sums = [inter_abs_sum(rec, _place(lig, rots[a], trans[b])[1], lig, True) for a, b in zip(pr, pt)]
    assert_energy(ec[sample], ref, sums, what="global cull scan")
    g = ctx.compute_grids([6, 8], [0.2, -0.4], 1.0, (12, 12, 12), (-6.0, -6.0, -6.0))
    og = oracle.Grid.make(1.0, (12, 12, 12))
    shifted = dict(rec, xs=rec["xs"] + 6.0, ys=rec["ys"] + 6.0, zs=rec["zs"] + 6.0)
    ref = oracle.grid_energies(shifted, og, [6, 8], [0.2, -0.4])
    ok = ref < 1e5
    assert np.allclose(g[ok], ref[ok], rtol=1e-4, atol=1e-3)
    torch.cuda.synchronize()


# ----------------------------------------------------------------- Monte Carlo
def _mc_setup(seed=17):
    from mmo_amd import mc as mcmod
    _, rec = synth.config0()
    ligs = [l for l in synth.library(10, seed=seed) if 2 <= len(l["rb_left"]) <= 8][:2]
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    ids = [ctx.add_ligand(l) for l in ligs]
    rng = np.random.default_rng(seed)
    chain_lig = np.array([ids[0], ids[1], ids[0], ids[1], ids[0]])
    rot0 = synth.random_rotations(rng, len(chain_lig))
    # on the receptor's outer surface (r_out ~ 28 A): moderate energies, so that no
    # accept/reject decision hinges on the last digits of E
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(chain_lig), 3))
    return mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0


SURFACE = np.array([36.0, 0.0, 0.0])


def test_mc_matches_cpu_restatement():
    from tests.mc_ref import simulate
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup()
    roi = (SURFACE, 4.0)
    # FP64 pair arithmetic; receptor charges and LJ factors stay FP32 and coordinates
    # 2^-24 A fixed point, so energies agree to ~1e-7 relative, not to the last bit
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI | mcmod.FP64
    m = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, total_frames=80, flags=flags,
                         roi_center=roi[0], roi_r=roi[1], seed=99)
    print("mc run 2 x 40 frames", flush=True)
    m.run(40)
    m.run(40)
    st = m.state()
    conf, best = m.coords()
    assert (st["frame"] == 80).all() and (st["status"] != 0).all()
    for k, l in enumerate(chain_lig):
        ref = simulate(ligs[ids.index(l)], rec, rot0[k], pos0[k], 80, seed=99, chain=k,
                       hard_roi=roi)
        print("chain", k, "gpu", st["accepted"][k], st["best_e"][k], st["e"][k], "cpu",
              ref["accepted"], ref["best_e"], ref["e"], flush=True)
        assert st["status"][k] == (1 if ref["status"] == 0 else 2)
        assert st["accepted"][k] == ref["accepted"], (k, st["accepted"][k], ref["accepted"])
        assert st["resets"][k] == ref["resets"] and st["ooroi"][k] == ref["ooroi"]
        for key in ("best_e", "e"):
            if math.isinf(ref[key]):
                assert math.isinf(st[key][k])
            else:
                assert st[key][k] == pytest.approx(ref[key], rel=1e-6, abs=1e-4), key
        assert np.allclose(st["max_rot"][k], ref["max_rot"]) and np.allclose(st["max_trans"][k], ref["max_trans"])
        assert np.allclose(conf[k], ref["conf"], atol=1e-9)


def test_mc_fp32_matches_cpu_restatement():
    """Default (FP32 pair arithmetic) chains follow the FP64 restatement: same
    decisions, energies within the FP32 scoring tolerance."""
    from tests.mc_ref import simulate
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup(seed=19)
    roi = (SURFACE, 4.0)
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    refs = [simulate(ligs[ids.index(l)], rec, rot0[k], pos0[k], 60, seed=3, chain=k, hard_roi=roi)
            for k, l in enumerate(chain_lig)]
    for sweep in (0, 1):          # 8-atom sub-blocks, lanes over receptor atoms
        ctx.tune(mc_sweep=sweep)
        print("mc fp32 sweep", sweep, flush=True)
        m = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, total_frames=60, flags=flags,
                             roi_center=roi[0], roi_r=roi[1], seed=3)
        m.run(60)
        st = m.state()
        for k, ref in enumerate(refs):
            assert st["accepted"][k] == ref["accepted"]
            assert st["resets"][k] == ref["resets"]
            assert st["best_e"][k] == pytest.approx(ref["best_e"], rel=1e-5, abs=1e-3)
    ctx.tune(mc_sweep=1)


def test_mc_config0_full_length_fp64():
    """The CPU-runnable workload at its full length: the 75-atom ligand in the 3000-atom
    receptor, one chain, 10,000 MC steps with FP64 pair arithmetic, against the CPU
    restatement (decisions, step sizes, energies, final conformer)."""
    from tests.mc_ref import simulate
    from mmo_amd import mc as mcmod
    lig, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    lid = ctx.add_ligand(lig)
    rot0 = synth.random_rotations(np.random.default_rng(7), 1)
    roi = (np.zeros(3), 6.0)
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI | mcmod.FP64
    m = mcmod.MonteCarlo(ctx, [lid], rot0, np.zeros((1, 3)), ff.SHIFTED, total_frames=10000,
                         flags=flags, roi_center=roi[0], roi_r=roi[1], seed=11)
    print("config0 mc 10000 frames", flush=True)
    m.run(10000)
    st = m.state()
    conf, _ = m.coords()
    ref = simulate(lig, rec, rot0[0], np.zeros(3), 10000, seed=11, chain=0, hard_roi=roi)
    assert st["frame"][0] == 10000
    assert st["accepted"][0] == ref["accepted"] and st["resets"][0] == ref["resets"]
    assert st["ooroi"][0] == ref["ooroi"]
    for key in ("best_e", "e"):
        assert st[key][0] == pytest.approx(ref[key], rel=1e-6, abs=1e-4), key
    assert np.allclose(st["max_rot"][0], ref["max_rot"]) and np.allclose(st["max_trans"][0], ref["max_trans"])
    assert np.allclose(conf[0], ref["conf"], atol=1e-8)


def _long_chain_system(nlig=6, chains=2, seed=41, max_rb=4):
    """Library ligands with 1..max_rb rotatable bonds (so that 1,000 frames pass conformer
    step 100 nrb), `chains` chains each, starting on the receptor's outer surface, where
    energies are moderate and few decisions hinge on the last digits of E."""
    lib = [l for l in synth.library(64, seed=seed) if 1 <= len(l["rb_left"]) <= max_rb][:nlig]
    _, rec = synth.config0()
    rng = np.random.default_rng(seed)
    ncl = len(lib) * chains
    rot0 = synth.random_rotations(rng, ncl)
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (ncl, 3))
    return lib, rec, rot0, pos0


LONG_ROI = (SURFACE, 4.0)


def _compare_long_chains(m, lig_of_chain, refs, min_margin, e_rel):
    """GPU chains against restatement runs (with traces): chains whose every decision in
    the restatement lies farther than min_margin (relative) from its boundary must take
    the same decisions, so their counters, step sizes, per-bond steps, conformers and
    best poses agree; returns the indices of the chains compared."""
    st = m.state()
    conf, best = m.coords()
    dr = m.bond_steps()
    compared = []
    for k, (ref, tr) in enumerate(refs):
        if min(t["margin"] for t in tr) <= min_margin:
            continue
        compared.append(k)
        lig = lig_of_chain[k]
        assert st["status"][k] == (1 if ref["status"] == 0 else 2), k
        assert st["frame"][k] == len(tr) or ref["status"] != 0, k
        assert st["accepted"][k] == ref["accepted"], (k, st["accepted"][k], ref["accepted"])
        assert st["resets"][k] == ref["resets"] and st["ooroi"][k] == ref["ooroi"], k
        assert st["rigid_step"][k] == ref["rigid_step"] and st["conf_step"][k] == ref["conf_step"]
        assert st["max_rot"][k] == ref["max_rot"] and st["max_trans"][k] == ref["max_trans"], k
        assert np.array_equal(dr[k], np.asarray(ref["dr"])), (k, dr[k], ref["dr"])
        assert np.allclose(conf[k], ref["conf"], rtol=0, atol=1e-9), k
        assert np.allclose(best[k], ref["best_xyz"], rtol=0, atol=1e-8), k
        for key in ("best_e", "e"):
            if math.isinf(ref[key]):
                assert math.isinf(st[key][k]), (k, key)
            else:
                assert st[key][k] == pytest.approx(ref[key], rel=e_rel, abs=1e-6), (k, key)
    return compared


def _assert_rules_exercised(refs, compared, lig_of_chain):
    """The compared chains went past the adaptive rigid steps (rigid step 100), the
    flips (conformer step 100) and the per-bond steps (conformer step 100 nrb)."""
    assert len(compared) >= len(refs) // 2, "too many chains with a close decision"
    full = [k for k in compared if refs[k][0]["status"] == 0]
    assert len(full) >= 2, "too few compared chains ran all their frames"
    assert all(refs[k][0]["rigid_step"] >= 200 for k in full)
    assert all(refs[k][0]["conf_step"] >= 100 for k in full)
    assert any(refs[k][0]["conf_step"] >= 100 * len(lig_of_chain[k]["rb_left"]) for k in full)
    assert any(refs[k][0]["accepted"] > 0 and refs[k][0]["max_trans"] != 0.15 for k in full)


def test_mc_atom_sweep_1000_frames_follow_restatement():
    """The benchmarked chains (FP32 per-atom sweep, receptor in LDS) for the workload's
    full 1,000 frames against the FP64 restatement, past every adaptive rule."""
    from tests.mc_ref import simulate
    from mmo_amd import mc as mcmod
    lib, rec, rot0, pos0 = _long_chain_system()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    ids = [ctx.add_ligand(l) for l in lib]
    cl = np.repeat(np.asarray(ids), 2)
    lig_of = [lib[ids.index(l)] for l in cl]
    roi = LONG_ROI
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    m = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, total_frames=1000, flags=flags,
                         roi_center=roi[0], roi_r=roi[1], seed=1234)
    print("mc atom sweep 1000 frames, %d chains" % len(cl), flush=True)
    m.run_graph(250, nreplay=4)
    refs = []
    for k in range(len(cl)):
        tr = []
        refs.append((simulate(lig_of[k], rec, rot0[k], pos0[k], 1000, seed=1234, chain=k,
                              hard_roi=roi, trace=tr), tr))
        print("  restatement chain %d done" % k, flush=True)
    compared = _compare_long_chains(m, lig_of, refs, min_margin=1e-5, e_rel=1e-5)
    print("compared chains", compared, flush=True)
    _assert_rules_exercised(refs, compared, lig_of)


def test_mc_grid_1000_frames_follow_restatement():
    """The grid-interpolation chains (10 per-type FP32 grids of 100^3 at 0.375 A computed
    on the GPU) for 1,000 frames against the restatement on the same grids."""
    from tests.mc_ref import simulate
    from mmo_amd import mc as mcmod
    lib, rec, rot0, pos0 = _long_chain_system(seed=43)
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    probe_a, probe_q, types = synth.grid_types(lib, 10)
    origin, dims, step = synth.grid_spec(SURFACE, 100, 0.375)
    grids = ctx.compute_grids(probe_a, probe_q, step, dims, origin)
    ids = [ctx.add_ligand(l, types=t) for l, t in zip(lib, types)]
    cl = np.repeat(np.asarray(ids), 2)
    lig_of = [lib[ids.index(l)] for l in cl]
    typ_of = [types[ids.index(l)] for l in cl]
    roi = LONG_ROI
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    m = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.INTERP, total_frames=1000, flags=flags,
                         roi_center=roi[0], roi_r=roi[1], seed=4321)
    print("mc grid 1000 frames, %d chains" % len(cl), flush=True)
    m.run_graph(250, nreplay=4)
    g = oracle.Grid.make(step, dims)
    gl = list(grids)
    refs = []
    for k in range(len(cl)):
        tr = []

        def inter(x, k=k):
            return oracle.ene_inter_interp(g, gl, x - origin, typ_of[k])
        refs.append((simulate(lig_of[k], rec, rot0[k], pos0[k], 1000, seed=4321, chain=k,
                              hard_roi=roi, inter_fn=inter, trace=tr), tr))
    compared = _compare_long_chains(m, lig_of, refs, min_margin=1e-9, e_rel=1e-9)
    print("compared chains", compared, flush=True)
    _assert_rules_exercised(refs, compared, lig_of)


def test_mc_chains_in_contact_keep_energies_within_tolerance():
    """Chains started inside the receptor's atoms (clashing poses, energies far above
    1e7 kcal/mol): every proposal's E_inter of the FP32 sweep stays within 1e-5 of the
    FP64 restatement's (pairs closer than 0.5 A add the quantisation residuals)."""
    from tests.mc_ref import simulate
    from mmo_amd import mc as mcmod
    lib, rec, _, _ = _long_chain_system(nlig=3, chains=1, seed=45)
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    ids = [ctx.add_ligand(l) for l in lib]
    rxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    pos0 = rxyz[[5, 500, 1500]] + 0.05          # ligand centres on receptor atoms
    rot0 = synth.random_rotations(np.random.default_rng(45), 3)
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    nf = 24
    m = mcmod.MonteCarlo(ctx, ids, rot0, pos0, ff.SHIFTED, total_frames=nf, flags=flags,
                         roi_center=np.zeros(3), roi_r=40.0, seed=77)
    got = []
    for f in range(nf):
        m.run(1)
        got.append(m.state()["e_inter"].copy())
    print("contact chains: E_inter of the first proposals", got[0], flush=True)
    for k in range(3):
        tr = []
        simulate(lib[k], rec, rot0[k], pos0[k], nf, seed=77, chain=k,
                 hard_roi=(np.zeros(3), 40.0), trace=tr)
        ref = np.array([t["e_inter"] for t in tr])
        gpu = np.array([g[k] for g in got])
        assert ref.max() > 1e7, "chain %d did not start in contact" % k
        assert np.all(np.abs(gpu - ref) <= 1e-5 * np.abs(ref)), (k, np.abs(gpu / ref - 1).max())


def test_mc_launch_forms_agree():
    """One persistent launch, hipGraph replays and one launch per ligand give the same
    chains bit for bit (each chain depends only on its own counter-based stream)."""
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup(seed=23)
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    kw = dict(total_frames=60, flags=flags, roi_center=SURFACE, roi_r=4.0, seed=5)
    a = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, **kw)
    b = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, **kw)
    c = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, **kw)
    print("mc launch forms", flush=True)
    a.run(60)
    b.run_graph(20, nreplay=3)
    c.run_per_ligand(30)
    c.run_per_ligand(30)
    sa, sb, sc = a.state(), b.state(), c.state()
    for key in ("best_e", "e", "accepted", "resets", "frame", "max_rot"):
        assert np.array_equal(sa[key], sb[key]), key
        assert np.array_equal(sa[key], sc[key]), key
    assert np.array_equal(a.coords()[0][0], b.coords()[0][0])


def test_mc_full_chain_count_forms_agree_and_follow_restatement():
    """The batched Monte Carlo at its full chain count (1,024 library ligands x 64 chains
    = 65,536 chains), 40 frames: one persistent launch, hipGraph replays and one launch
    per ligand give the same chains bit for bit, and sampled chains follow the CPU
    restatement (decisions, resets, best energies within the FP32 tolerance)."""
    from tests.mc_ref import simulate
    from mmo_amd import mc as mcmod
    lib = synth.library(1024, seed=4, processes=8)
    _, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    ids = [ctx.add_ligand(l) for l in lib]
    rng = np.random.default_rng(4)
    cl = np.repeat(np.asarray(ids), 64)
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = rng.uniform(-1.0, 1.0, (len(cl), 3))
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    kw = dict(total_frames=40, flags=flags, roi_center=np.zeros(3), roi_r=6.0, seed=77)
    a = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, **kw)
    b = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, **kw)
    c = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, **kw)
    print("mc 65536 chains", flush=True)
    a.run(40)
    b.run_graph(10, nreplay=4)
    c.run_per_ligand(40)
    sa, sb, sc = a.state(), b.state(), c.state()
    for key in ("best_e", "e", "accepted", "resets", "frame", "max_rot", "max_trans"):
        assert np.array_equal(sa[key], sb[key]), key
        assert np.array_equal(sa[key], sc[key]), key
    assert ((sa["frame"] == 40) | (sa["status"] == 2)).all()
    # conformers (turned in place by torsion moves, restored unless kept) agree bit for bit
    ca, cb, cc = a.coords()[0], b.coords()[0], c.coords()[0]
    assert all(np.array_equal(x, y) and np.array_equal(x, z) for x, y, z in zip(ca, cb, cc))
    # sampled chains, among them chains that ended on a too-long conformer
    too_long = np.flatnonzero(sa["status"] == 2)
    sample = list(rng.choice(len(cl), 6, replace=False)) + list(too_long[:2])
    for k in sample:
        ref = simulate(lib[ids.index(cl[k])], rec, rot0[k], pos0[k], 40, seed=77, chain=int(k),
                       hard_roi=(np.zeros(3), 6.0))
        # the last kept conformer, also for a chain that stopped on a too-long proposal
        assert np.allclose(ca[k], ref["conf"], rtol=0, atol=1e-9), k
        if ref["status"] != 0:
            assert sa["status"][k] == 2
            continue
        assert sa["accepted"][k] == ref["accepted"] and sa["resets"][k] == ref["resets"]
        assert sa["best_e"][k] == pytest.approx(ref["best_e"], rel=1e-5, abs=1e-3)


def test_mc_without_hard_roi_never_accepts():
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup(seed=29)
    m = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, total_frames=40,
                         flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF, seed=1)
    m.run(40)
    st = m.state()
    assert (st["accepted"] == 0).all()
    m2 = mcmod.MonteCarlo(ctx, chain_lig, rot0, pos0, ff.SHIFTED, total_frames=40,
                          flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.METROPOLIS_ALWAYS,
                          seed=1)
    m2.run(40)
    assert (m2.state()["accepted"] > 0).any()


def test_mc_interp_mode_matches_cpu_restatement():
    from tests.mc_ref import simulate
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup(seed=31)
    # per-type grids for each ligand atom's (anum, q), computed on the GPU
    types, probe_a, probe_q = [], [], []
    for l in ligs:
        t = []
        for a, q in zip(l["anum"], l["q"]):
            probe_a.append(int(a))
            probe_q.append(float(q))
            t.append(len(probe_a) - 1)
        types.append(np.array(t))
    origin, dims, step = synth.grid_spec(center=SURFACE, n=64, step=0.5)
    grids = ctx.compute_grids(probe_a, probe_q, step, dims, origin)
    ids2 = [ctx.add_ligand(l, types=t) for l, t in zip(ligs, types)]
    cl = np.array([ids2[ids.index(l)] for l in chain_lig])
    roi = (SURFACE, 3.0)
    m = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.INTERP, total_frames=60,
                         flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI,
                         roi_center=roi[0], roi_r=roi[1], seed=8)
    m.run(60)
    st = m.state()
    g = oracle.Grid.make(step, dims)
    for k, l in enumerate(chain_lig):
        li = ids.index(l)

        def inter(x, li=li):
            return oracle.ene_inter_interp(g, list(grids), x - origin, types[li])
        ref = simulate(ligs[li], rec, rot0[k], pos0[k], 60, seed=8, chain=k, hard_roi=roi,
                       inter_fn=inter)
        assert st["accepted"][k] == ref["accepted"]
        assert st["best_e"][k] == pytest.approx(ref["best_e"], rel=1e-9, abs=1e-9)
    # final-pose rescoring: the direct shifted E_inter of each chain's best pose
    e_rs = m.rescore()
    _, best = m.coords()
    for k, l in enumerate(chain_lig):
        lg = dict(ligs[ids.index(l)])
        lg["xs"], lg["ys"], lg["zs"] = (np.ascontiguousarray(best[k][:, a]) for a in range(3))
        assert_energy(e_rs[k], oracle.ene_inter(rec, lg, shifted=True),
                      inter_abs_sum(rec, best[k], lg), what="rescore chain %d" % k)


# ------------------------------------------------------------------ L-BFGS minimiser
def _one_atom_system(anum=6):
    ctx = ff.Context(0)
    rec = {"xs": np.zeros(1), "ys": np.zeros(1), "zs": np.zeros(1), "q": np.zeros(1),
           "anum": np.array([anum], dtype=np.int32)}
    ctx.set_receptor(rec)
    lig = {"xs": np.zeros(1), "ys": np.zeros(1), "zs": np.zeros(1), "q": np.zeros(1),
           "anum": np.array([anum], dtype=np.int32), "dists": np.zeros((1, 1), dtype=np.int32),
           "rb_left": [], "rb_right": [], "rg_off": [0], "rg_atoms": []}
    return ctx, rec, lig


def test_minimizer_lj_closed_form():
    """One C atom against one C atom: the global form's minimum sits at r = x_CC = 3.851 A
    with E = -D_CC = -0.105 kcal/mol (UFF puts the LJ minimum at x_ij); the shifted
    form's minimum is the 1-D minimum of S(r) LJ(r), found here by bounded Brent on the
    oracle's pair energy."""
    import torch
    from scipy.optimize import minimize_scalar
    from mmo_amd.minimize import CONVERGED, RUNNING, Minimizer
    ctx, rec, lig = _one_atom_system()
    lid = ctx.add_ligand(lig)
    npose = 64
    b = ctx.batch([lid] * npose)
    rng = np.random.default_rng(2)
    dirs = rng.normal(size=(npose, 3))
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    r0 = rng.uniform(3.3, 6.0, npose)
    start = np.concatenate([np.concatenate([r0[p] * dirs[p], rng.uniform(-1, 1, 3)]) for p in range(npose)])

    def shifted_pair(r):
        m = dict(lig, xs=np.array([r]), ys=np.zeros(1), zs=np.zeros(1))
        return oracle.ene_inter(rec, m, shifted=True)
    rs = minimize_scalar(shifted_pair, bounds=(3.0, 5.0), method="bounded",
                         options={"xatol": 1e-10}).x
    for flags, r_min, e_min in ((ff.F_INTER_GLOBAL, 3.851, -0.105),
                                (ff.F_INTER_SHIFTED, rs, shifted_pair(rs))):
        d = torch.tensor(start, device="cuda:0")
        mz = Minimizer(b, flags=flags, history=8, max_step=0.5, gtol=1e-9)
        print("minimizer one-atom flags", flags, flush=True)
        e = mz.run(d, 200).cpu().numpy()
        x = d.cpu().numpy().reshape(npose, 6)
        info = mz.info()
        r = np.linalg.norm(x[:, :3], axis=1)
        # FP32 pair energies resolve E to ~1e-8 kcal/mol near the minimum, where
        # E'' = 72 D / x^2 = 0.51 kcal/mol/A^2: positions to ~3e-4 A
        assert np.all(np.abs(r - r_min) < 1e-3), np.abs(r - r_min).max()
        assert np.allclose(e, e_min, rtol=1e-5, atol=1e-9)
        assert np.all(e <= info["e_start"])
        # near the minimum FP32 energy noise keeps some poses backtracking until the
        # iteration limit; none may fail
        assert np.all((info["status"] == CONVERGED) | (info["status"] == RUNNING))


def test_minimizer_flexible_ligands(lib_sys):
    """Library ligands near the receptor: every pose ends at or below its start energy,
    the reported energy is the oracle's energy of the returned DOFs, and the returned
    DOFs stay inside the ROI / angle bounds."""
    import torch
    from mmo_amd.minimize import Minimizer, roi_bounds
    ctx, ligs, ids, rec, _ = lib_sys
    pick = [k for k in range(len(ligs)) if 1 <= len(ligs[k]["rb_left"]) <= 8][:4]
    pose_lig = np.array([ids[k] for k in pick for _ in range(4)])
    b = ctx.batch(pose_lig)
    center, radius = np.array([30.0, 0.0, 0.0]), 3.0
    nrb = [len(ligs[l]["rb_left"]) for l in pose_lig]
    lo, hi = roi_bounds(b, center, radius, nrb)
    rng = np.random.default_rng(21)
    start = []
    for p, l in enumerate(pose_lig):
        start.append(np.concatenate([center + rng.uniform(-2, 2, 3), rng.uniform(-3, 3, 3),
                                     rng.uniform(-0.3, 0.3, nrb[p])]))
    d = torch.tensor(np.concatenate(start), device="cuda:0")
    mz = Minimizer(b, flags=ff.F_INTER_SHIFTED | ff.F_INTRA, history=8, bounds=(lo, hi),
                   max_step=0.2)
    print("minimizer flexible", b.npose, flush=True)
    e = mz.run(d, 150).cpu().numpy()
    x = d.cpu().numpy()
    info = mz.info()
    assert np.all(x >= lo - 1e-12) and np.all(x <= hi + 1e-12)
    assert np.all(e <= info["e_start"] + 1e-12)
    assert np.mean(info["e_start"] - e) > 0.0
    refs, sums = [], []
    for p, l in enumerate(pose_lig):
        lg = ligs[l]
        c = np.stack([lg["xs"], lg["ys"], lg["zs"]], 1)
        dof = x[b.dof_off[p]:b.dof_off[p + 1]]
        xyz, rad = oracle.apply_config(c, lg["rb_left"], lg["rb_right"], lg["rg_off"], lg["rg_atoms"], dof)
        m = dict(lg, xs=xyz[:, 0], ys=xyz[:, 1], zs=xyz[:, 2])
        refs.append(oracle.ene_inter(rec, m) + oracle.ene_intra(m))
        sums.append(inter_abs_sum(rec, xyz, m) + intra_abs_sum(m, xyz))
        assert rad <= 12.0
    assert_energy(e, refs, sums, what="minimised poses")


# ------------------------------------------------ exhaustive-scan filters, top-k
def _bondi(anum):
    from oracle.scan_ref import VDW_RADII
    return np.array([VDW_RADII[int(a)] for a in anum])


def test_vdw_mask_and_scan_filters(sys0):
    """Protein vdW bitmask bit-exact against the CPU restatement; the scan's ROI,
    centre-clash and atom-clash filters skip exactly the poses the restatement skips."""
    import torch
    from oracle import scan_ref
    ctx, lig, rec = sys0
    rxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    step, dims = 0.5, (48, 44, 40)
    origin = rxyz.min(axis=0) - 1.0
    print("vdw mask", dims, flush=True)
    ctx.set_vdw_mask(step, dims, origin)
    ref = scan_ref.vdw_mask(rxyz, _bondi(rec["anum"]), step, dims, origin)
    assert np.array_equal(ctx.vdw_mask(), scan_ref.mask_words(ref))
    # a ligand whose centre sits inside one of its atoms (a 3-atom line), and the
    # 75-atom ligand; a 16 x 8000 scan around the cavity, ROI radius 4 A
    line = {"xs": np.array([-1.5, 0.0, 1.5]), "ys": np.zeros(3), "zs": np.zeros(3),
            "q": np.zeros(3), "anum": np.array([6, 6, 6], dtype=np.int32),
            "dists": np.array([[0, 1, 2], [1, 0, 1], [2, 1, 0]], dtype=np.int32)}
    rots = synth.so3_rotations(4608)[::288]
    trans = synth.lattice((0.0, 0.0, 0.0), 20, 0.5)
    roi = (np.array([0.3, -0.2, 0.1]), 4.0)
    rng = np.random.default_rng(3)
    for L in (line, lig):
        lid = ctx.add_ligand(L)
        radii = _bondi(L["anum"])
        e = ctx.scan(lid, ff.SHIFTED, rots, trans)
        print("scan filter atoms", len(L["xs"]), flush=True)
        ctx.scan_filter(lid, rots, trans, e, roi=roi, lig_radii=radii)
        got = np.isinf(e.cpu().numpy())
        c = np.stack([L["xs"], L["ys"], L["zs"]], 1)
        sample = rng.choice(len(got), 600, replace=False)
        for s in sample:
            r, t = divmod(int(s), len(trans))
            want = scan_ref.pose_skipped(ref, step, origin, c, radii, rots[r], trans[t], roi=roi)
            assert got[s] == want, (len(L["xs"]), r, t)
        assert 0 < got.sum() < len(got)
    torch.cuda.synchronize()


def test_filtered_scan_equals_scan_then_filter(sys0):
    """Filter before scoring, as the reference's exhaustive docking does: the filters
    first, then only the kept poses scored.  At the headline scan's full size (4,608
    rotations x 20^3 translations at 0.5 A), against the full scan followed by the
    filters, in the direct and the interpolated form, with an ROI that keeps a few poses,
    one that keeps a share of them, and the ROI alone around the whole lattice (every pose
    kept, so every slot goes through the list form): every energy bit for bit (+inf for
    skipped poses) and the same top-100."""
    import torch
    ctx, lig, rec = sys0
    rxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    step = 0.375
    origin = rxyz.min(axis=0) - 2.0
    dims = tuple(int(v) for v in np.ceil((rxyz.max(axis=0) + 2.0 - origin) / step) + 1)
    ctx.set_vdw_mask(step, dims, origin)
    probe_a, probe_q, types = synth.grid_types([lig], 10)
    gorigin, gdims, gstep = synth.grid_spec((0.0, 0.0, 0.0), 64, 0.375)
    ctx.compute_grids(probe_a, probe_q, gstep, gdims, gorigin, copy_out=False)
    lid = ctx.add_ligand(lig, types=types[0])
    rots = synth.so3_rotations(4608)
    trans = synth.lattice((0.0, 0.0, 0.0), 20, 0.5)
    radii = _bondi(lig["anum"])
    cases = [("roi 4.75 + clashes", (np.zeros(3), 4.75 + 1e-9), radii, None),
             ("roi 6 + clashes", (np.array([0.5, -0.5, 0.25]), 6.0), radii, None),
             ("roi only, all kept", (np.zeros(3), 20.0), None, _lib.SCAN_ROI)]
    for mode in (ff.SHIFTED, ff.INTERP):
        full = ctx.scan(lid, mode, rots, trans)
        for name, roi, rad, flags in cases:
            print("filtered scan", mode, name, flush=True)
            e_ref = full.clone()
            ctx.scan_filter(lid, rots, trans, e_ref, roi=roi, lig_radii=rad, flags=flags)
            e, nk = ctx.scan_filtered(lid, mode, rots, trans, roi=roi, lig_radii=rad, flags=flags)
            a, b = e.cpu().numpy(), e_ref.cpu().numpy()
            assert nk == int((~np.isinf(b)).sum()), (name, nk)
            assert np.array_equal(a, b, equal_nan=True), (mode, name)
            if flags == _lib.SCAN_ROI:
                assert nk == len(a)
            else:
                assert 0 < nk < len(a)
            v1, i1 = ctx.topk(e, 100)
            v2, i2 = ctx.topk(e_ref, 100)
            assert torch.equal(i1, i2) and np.array_equal(v1.cpu().numpy(), v2.cpu().numpy(),
                                                          equal_nan=True)
        del full
    torch.cuda.synchronize()


def test_topk_matches_numpy(sys0):
    import torch
    from oracle import scan_ref
    ctx, _, _ = sys0
    rng = np.random.default_rng(11)
    for n, k in ((1_000_003, 100), (1_000_003, 2048), (500, 2048), (70_000, 1)):
        e = rng.normal(0, 10, n)
        e[rng.choice(n, n // 10, replace=False)] = np.round(e[: n // 10], 1)   # ties
        e[rng.choice(n, n // 50, replace=False)] = np.inf
        e[rng.choice(n, 3, replace=False)] = np.nan
        print("topk", n, k, flush=True)
        v, i = ctx.topk(torch.tensor(e, device="cuda:0"), k)
        rv, ri = scan_ref.topk(e, k)
        assert np.array_equal(i.cpu().numpy(), ri)
        assert np.array_equal(v.cpu().numpy(), rv, equal_nan=True)
    # all equal: the lowest indices win
    v, i = ctx.topk(torch.full((10_000,), 2.5, dtype=torch.float64, device="cuda:0"), 7)
    assert i.cpu().numpy().tolist() == list(range(7))


# ------------------------------------------------------------ .mol2 input path
def test_mol2_ligands_score_like_the_oracle(sys0, tmp_path):
    """Ligands read from .mol2 (topology, rotatable bonds and groups computed here)
    flow into the DOF kernels: coordinates and energies agree with the oracle."""
    import torch
    from mmo_amd import mol as molmod
    from tests.test_mol2 import BUTANE, RING
    ctx, _, rec = sys0
    fn = str(tmp_path / "ligs.mol2")
    with open(fn, "w") as f:
        f.write(BUTANE + RING)
    ligs = []
    for m in molmod.ligands_of_mol2_file(fn):
        m = m.with_coords(m.coords() - m.coords().mean(axis=0))   # Lds.preprocess_ligand
        n = m.num_atoms()
        d = {"xs": m.xs, "ys": m.ys, "zs": m.zs, "q": m.q_a, "anum": m.elt_a,
             "dists": np.asarray(m.dists_a).reshape(n, n).T, "rb_left": m.rb_left,
             "rb_right": m.rb_right, "rg_off": m.rg_off, "rg_atoms": m.rg_atoms}
        ligs.append((ctx.add_ligand(m), d))
    rng = np.random.default_rng(2)
    pose_lig = [l for l, _ in ligs for _ in range(3)]
    b = ctx.batch(pose_lig)
    dofs = []
    for lid in pose_lig:
        d = dict(ligs)[lid]
        dofs.append(np.concatenate([[26.0, 0, 0] + rng.uniform(-2, 2, 3),
                                    rng.uniform(-np.pi, np.pi, 3 + len(d["rb_left"]))]))
    dt = torch.tensor(np.concatenate(dofs), device="cuda:0")
    x = torch.zeros((b.natoms, 3), dtype=torch.float64, device="cuda:0")
    e = torch.zeros((b.npose, 4), dtype=torch.float64, device="cuda:0")
    print("mol2 ligands dof batch", b.npose, flush=True)
    b.score_dofs_dev(dt, e, xyz_out=x, flags=ff.F_INTER_SHIFTED | ff.F_INTRA)
    torch.cuda.synchronize()
    x, e = x.cpu().numpy(), e.cpu().numpy()
    refs, sums = [], []
    for p, lid in enumerate(pose_lig):
        d = dict(ligs)[lid]
        c = np.stack([d["xs"], d["ys"], d["zs"]], 1)
        xyz, _ = oracle.apply_config(c, d["rb_left"], d["rb_right"], d["rg_off"], d["rg_atoms"], dofs[p])
        assert np.allclose(x[b.atom_off[p]:b.atom_off[p + 1]], xyz, atol=1e-10)
        m = dict(d, xs=xyz[:, 0], ys=xyz[:, 1], zs=xyz[:, 2])
        refs.append(oracle.ene_inter(rec, m) + oracle.ene_intra(m))
        sums.append(inter_abs_sum(rec, xyz, m) + intra_abs_sum(m, xyz))
    assert_energy(e[:, 0] + e[:, 1], refs, sums, what="mol2 ligands")


def test_cell_packed_grids_bit_identical():
    """The cell-packed copy of the grids (8 corners per cell, two 16-byte loads per lookup)
    holds the node arrays' values: grid-form Monte Carlo chains, with and without the
    minimiser, and batched interpolated energies and gradients are bit for bit those of
    the node arrays."""
    import torch
    mcmod, ctx, ligs, ids, rec, chain_lig, rot0, pos0 = _mc_setup(seed=31)
    types, probe_a, probe_q = [], [], []
    for l in ligs:
        t = []
        for a, q in zip(l["anum"], l["q"]):
            probe_a.append(int(a))
            probe_q.append(float(q))
            t.append(len(probe_a) - 1)
        types.append(np.array(t))
    origin, dims, step = synth.grid_spec(center=SURFACE, n=64, step=0.5)
    ids2 = [ctx.add_ligand(l, types=t) for l, t in zip(ligs, types)]
    cl = np.array([ids2[ids.index(l)] for l in chain_lig])
    roi = (SURFACE, 3.0)
    rng = np.random.default_rng(2)
    pose_lig = np.repeat(ids2, 2)
    xyz = []
    for l in pose_lig:
        lig = ligs[ids2.index(l)]
        _, x = _place(lig, synth.random_rotations(rng, 1)[0], SURFACE + rng.uniform(-1, 1, 3))
        xyz.append(x)
    xyz = np.concatenate(xyz)
    out = {}
    for pack in (1, 0):
        ctx.tune(grid_pack=pack)
        ctx.compute_grids(probe_a, probe_q, step, dims, origin, copy_out=False)
        print("grid_pack", pack, flush=True)
        res = []
        for iters in (0, 8):
            m = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.INTERP, total_frames=40,
                                 flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI,
                                 roi_center=roi[0], roi_r=roi[1], seed=8)
            if iters:
                m.set_minimizer(iters, history=8, max_step=0.2)
            m.run(40)
            st = m.state()
            res.append((st["best_e"].copy(), st["accepted"].copy(), st["e"].copy()))
        b = ctx.batch(pose_lig)
        e, g = b.score(xyz, ff.F_INTER_INTERP | ff.F_GRAD)
        res.append((e, g))
        out[pack] = res
    ctx.tune(grid_pack=1)
    for a, b in zip(out[1], out[0]):
        for u, v in zip(a, b):
            assert np.array_equal(u, v, equal_nan=True)
    torch.cuda.synchronize()


def test_row_list_scan_grids_agree():
    """The row-list scan under several scan grids (1.5 A rows: several passes of 64
    candidate rows per atom; 6 A rows; x cuts 1, 2, 16; 10 A rows longer than 64 atoms)
    and both receptor paths gives the
    energies of the default grid to FP32 rounding, and the oracle's on a sample: a row
    grid only changes which pairs beyond the cutoff are visited.  With and without the
    per-atom box test of the far rows (row_filter) the energies are bit-identical: the
    atoms it drops would add exact zeros."""
    import torch
    lig, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    lid = ctx.add_ligand(lig)
    rots = synth.so3_rotations(4608)[::288]
    trans = synth.lattice((0.0, 0.0, 0.0), 20, 0.5)
    ctx.tune(scan_cull=5)
    ref = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
    rng = np.random.default_rng(11)
    sample = rng.choice(len(ref), 64, replace=False)
    pr, pt = (sample // len(trans)).astype(np.int32), (sample % len(trans)).astype(np.int32)
    want = oracle.scan_shifted(rec, lig, rots, pr, trans, pt, nthreads=8)
    sums = [inter_abs_sum(rec, _place(lig, rots[a], trans[b])[1], lig, True) for a, b in zip(pr, pt)]
    # 10 A rows uncut along x hold more than 64 atoms: the atom test covers the first 64
    for cell, xsub, chunk in ((1.5, 1, 4), (6.0, 2, 8), (3.5, 16, 4), (10.0, 1, 4)):
        ctx.tune(scan_cell=cell, scan_xsub=xsub, scan_chunk=chunk)
        ctx.set_receptor(rec)
        for sc in (5, 6):
            ctx.tune(scan_cull=sc)
            print("row grid", cell, xsub, chunk, sc, flush=True)
            e = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
            rel = np.abs(e - ref) / np.maximum(np.abs(ref), 1e-3)
            assert rel.max() < 1e-6, (cell, xsub, sc, rel.max())
            assert_energy(e[sample], want, sums, what="row grid %g/%d" % (cell, xsub))
            ctx.tune(row_filter=0)
            e0 = ctx.scan(lid, ff.SHIFTED, rots, trans).cpu().numpy()
            ctx.tune(row_filter=1)
            assert np.array_equal(e0, e), (cell, xsub, sc, np.abs(e0 - e).max())
    torch.cuda.synchronize()
