"""GPU parity of the persistent minimiser and of Monte Carlo-plus-minimisation (all
through the C ABI).

The grid form evaluates FP64 trilinear interpolation of the same FP32 grids on both
sides, so the GPU runs follow the CPU restatements (tests/lbfgs_ref.py, tests/mc_ref.py
over oracle.dof_energy_grad) step by step.  The direct forms (FP32 pairs, and FP64 pairs
with FP32-rounded receptor parameters) are checked by the energies of the points they
return against the FP64 oracle, by never ending above the start, and by the agreement
of the one-launch and the launch-sequence minimisers."""
import math

import numpy as np
import pytest

import oracle
from mmo_amd import ff, synth
from tests.parity import assert_energy, inter_abs_sum, intra_abs_sum

pytestmark = pytest.mark.gpu

SURFACE = np.array([36.0, 0.0, 0.0])


@pytest.fixture(scope="module")
def system():
    # ligands of moderate internal strain (the synthetic conformers' 1-4 contacts put
    # E_intra at 4e2 to 1e6 kcal/mol; the steepest make rounding differences between
    # two exact implementations grow along a minimisation)
    lib = [l for l in synth.library(96, seed=51)
           if 1 <= len(l["rb_left"]) <= 6 and oracle.ene_intra(l) < 2000.0][:8]
    assert len(lib) == 8
    _, rec = synth.config0()
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    probe_a, probe_q, types = synth.grid_types(lib, 10)
    origin, dims, step = synth.grid_spec(SURFACE, 100, 0.375)
    grids = ctx.compute_grids(probe_a, probe_q, step, dims, origin)
    ids = [ctx.add_ligand(l, types=t) for l, t in zip(lib, types)]
    return {"ctx": ctx, "lib": lib, "rec": rec, "ids": ids, "types": types, "grids": list(grids),
            "origin": origin, "dims": dims, "step": step}


def _grid_inter_grad(s, k):
    g = oracle.Grid.make(s["step"], s["dims"])

    def f(x):
        return oracle.grad_interp(g, s["grids"], x - s["origin"], s["types"][k])
    return f


def _start_dofs(s, rng, per_lig=2):
    pose_lig, d0 = [], []
    for k, l in enumerate(s["lib"]):
        for _ in range(per_lig):
            pose_lig.append(s["ids"][k])
            d0.append(np.concatenate([SURFACE + rng.uniform(-1.5, 1.5, 3), rng.uniform(-0.4, 0.4, 3),
                                      rng.uniform(-0.3, 0.3, len(l["rb_left"]))]))
    return np.asarray(pose_lig), d0


def _energy_at(s, lig, d, intra=True):
    """FP64 oracle: shifted direct E_inter + E_intra of apply_config(registered, d)."""
    base = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    x, _ = oracle.apply_config(base, lig["rb_left"], lig["rb_right"], lig["rg_off"], lig["rg_atoms"], d)
    m = dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
    e = oracle.ene_inter(s["rec"], m, shifted=True) + (oracle.ene_intra(m) if intra else 0.0)
    return e, inter_abs_sum(s["rec"], x, m, True) + (intra_abs_sum(m, x) if intra else 0.0)


@pytest.mark.parametrize("iters", [12, 60])
def test_fused_minimiser_grid_form_follows_restatement(system, iters):
    """One persistent launch of L-BFGS iterations per pose in the grid form against the
    CPU restatement: evaluations, accepted steps, status, energies and DOF vectors.
    Both sides evaluate the same FP64 trilinear field and differ only in the order of
    their sums.  Over 12 iterations the two follow the same points; over 60, on the
    steep clash energies some start poses have (up to ~1e6 kcal/mol), those rounding
    differences are amplified along the trajectory (inputs perturbed at 1e-16 move a
    pose's minimum by up to ~1e-5 relative), so there the decisions must still agree
    exactly and the values to within that amplification."""
    import torch
    from mmo_amd.minimize import Minimizer
    from tests.lbfgs_ref import minimize
    s = system
    ctx = s["ctx"]
    ctx.tune(min_fused=1)
    rng = np.random.default_rng(3)
    pose_lig, d0 = _start_dofs(s, rng)
    b = ctx.batch(pose_lig)
    d = torch.tensor(np.concatenate(d0), device="cuda:0")
    mz = Minimizer(b, flags=ff.F_INTER_INTERP | ff.F_INTRA, history=8, max_step=0.2)
    print("fused minimiser, grid form, %d poses, %d iterations" % (b.npose, iters), flush=True)
    e = mz.run(d, iters).cpu().numpy()
    info = mz.info()
    x = d.cpu().numpy()
    close = 0
    for p, l in enumerate(pose_lig):
        k = s["ids"].index(l)
        lig = s["lib"][k]
        base = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
        ig = _grid_inter_grad(s, k)

        def fg(v, lig=lig, base=base, ig=ig):
            ei, ea, g, rad = oracle.dof_energy_grad(base, lig, v, ig)
            return ei + ea, g, rad
        ref = minimize(fg, d0[p], iters, m=8, max_step=0.2, gtol=1e-6)
        o, n = int(b.dof_off[p]), int(b.dof_off[p + 1])
        assert info["evals"][p] == ref.n_eval and info["accepted"][p] == ref.n_acc, p
        assert info["status"][p] == ref.status, p
        assert e[p] <= info["e_start"][p]
        if iters <= 12:
            assert e[p] == pytest.approx(ref.f, rel=1e-9, abs=1e-9), p
            assert np.allclose(x[o:n], ref.x, rtol=0, atol=1e-6), p
        else:
            assert e[p] == pytest.approx(ref.f, rel=1e-4, abs=1e-6), p
            assert np.allclose(x[o:n], ref.x, rtol=0, atol=5e-2), p
            close += abs(e[p] - ref.f) <= 1e-6 * max(1.0, abs(ref.f))
    if iters > 12:
        assert close >= 0.8 * len(pose_lig), close


@pytest.mark.parametrize("f64", [0, 1])
def test_fused_minimiser_direct_form(system, f64):
    """The shifted direct form (FP32 pairs, or FP64 pairs): the best energy of every pose
    equals the FP64 oracle's energy at the returned DOF vector (<= 1e-5), never exceeds
    the start, and the one-launch minimiser agrees with the launch sequence (FP32)."""
    import torch
    from mmo_amd.minimize import Minimizer
    s = system
    ctx = s["ctx"]
    rng = np.random.default_rng(4 + f64)
    pose_lig, d0 = _start_dofs(s, rng, per_lig=3)
    b = ctx.batch(pose_lig)
    flags = ff.F_INTER_SHIFTED | ff.F_INTRA
    out = {}
    for fused in (1, 0):
        if fused == 0 and f64:
            continue
        ctx.tune(min_fused=fused, min_f64=f64)
        d = torch.tensor(np.concatenate(d0), device="cuda:0")
        mz = Minimizer(b, flags=flags, history=8, max_step=0.2)
        print("minimiser fused=%d f64=%d" % (fused, f64), flush=True)
        e = mz.run(d, 80).cpu().numpy()
        out[fused] = (e, d.cpu().numpy(), mz.info())
    ctx.tune(min_fused=1, min_f64=0)
    e, x, info = out[1]
    for p, l in enumerate(pose_lig):
        lig = s["lib"][s["ids"].index(l)]
        o, n = int(b.dof_off[p]), int(b.dof_off[p + 1])
        ref, sabs = _energy_at(s, lig, x[o:n])
        assert_energy(e[p], ref, sabs, what="pose %d" % p)
        assert e[p] <= info["e_start"][p]
        assert info["accepted"][p] > 0
    if 0 in out:
        # two FP32 implementations of the evaluation (one launch / launch sequence) on
        # chaotic trajectories: the minima they reach agree to 1e-3 for most poses
        e0 = out[0][0]
        close = np.abs(e - e0) <= 1e-3 * np.maximum(1.0, np.abs(e0))
        assert close.mean() >= 0.8, (e, e0)


def _mc_min(s, cl, rot0, pos0, mode, frames, iters, seed, flags=None, roi_r=5.0):
    from mmo_amd import mc as mcmod
    if flags is None:
        flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    m = mcmod.MonteCarlo(s["ctx"], cl, rot0, pos0, mode, total_frames=frames, flags=flags,
                         roi_center=SURFACE, roi_r=roi_r, seed=seed)
    m.set_minimizer(iters, history=8, max_step=0.2)
    return m


def test_mc_min_grid_form_follows_restatement(system):
    """Monte Carlo-plus-minimisation in the grid form (12 L-BFGS evaluations after every
    proposal, 8 frames) against the CPU restatement: decisions, evaluations, energies,
    step sizes, final conformers and best poses.  (Short minimisations stay in the
    well-conditioned part of the search, where both sides follow the same points.)"""
    from tests.mc_ref import simulate
    s = system
    rng = np.random.default_rng(5)
    cl = np.array(s["ids"][:6])
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(cl), 3))
    m = _mc_min(s, cl, rot0, pos0, ff.INTERP, 8, 12, seed=21)
    print("mc+min grid form", flush=True)
    m.run(8)
    st = m.state()
    conf, best = m.coords()
    per, tot = m.min_evals()
    g = oracle.Grid.make(s["step"], s["dims"])
    for k in range(len(cl)):
        lig = s["lib"][k]
        ig = _grid_inter_grad(s, k)

        def inter(x, k=k):
            return oracle.ene_inter_interp(g, s["grids"], x - s["origin"], s["types"][k])
        ref = simulate(lig, s["rec"], rot0[k], pos0[k], 8, seed=21, chain=k,
                       hard_roi=(SURFACE, 5.0), inter_fn=inter, inter_grad=ig,
                       minimize={"iters": 12, "m": 8, "max_step": 0.2})
        assert st["accepted"][k] == ref["accepted"] and st["resets"][k] == ref["resets"], k
        assert per[k] == ref["min_evals"], k
        for key in ("e", "best_e"):
            assert st[key][k] == pytest.approx(ref[key], rel=1e-6, abs=1e-6), (k, key)
        assert st["max_rot"][k] == ref["max_rot"] and st["max_trans"][k] == ref["max_trans"]
        assert np.allclose(conf[k] - conf[k].mean(0), ref["conf"] - ref["conf"].mean(0), atol=1e-5), k
        assert np.allclose(best[k], ref["best_xyz"], atol=1e-5), k
    assert tot == per.sum() > 8 * len(cl)


def test_mc_min_direct_form_best_poses_score_like_the_oracle(system):
    """The benchmarked direct form (FP32 per-atom sweep) with minimisation: every chain's
    best energy is the FP64 oracle energy of its best pose (<= 1e-5), and the best
    energies are lower than those of the same chains without minimisation."""
    from mmo_amd import mc as mcmod
    s = system
    rng = np.random.default_rng(6)
    cl = np.repeat(np.array(s["ids"]), 2)
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(cl), 3))
    m = _mc_min(s, cl, rot0, pos0, ff.SHIFTED, 10, 30, seed=22)
    print("mc+min direct form", flush=True)
    m.run(10)
    st = m.state()
    _, best = m.coords()
    plain = mcmod.MonteCarlo(s["ctx"], cl, rot0, pos0, ff.SHIFTED, total_frames=10,
                             flags=mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI,
                             roi_center=SURFACE, roi_r=5.0, seed=22)
    plain.run(10)
    sp = plain.state()
    scored = 0
    for k, l in enumerate(cl):
        lig = s["lib"][s["ids"].index(l)]
        x = best[k]
        if math.isinf(st["best_e"][k]):
            # the chain ended on a reset: best pose = the registered conformer at the
            # start pose, best energy +inf (src/lds.ml:632-648)
            c0 = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
            assert np.allclose(x, c0 @ rot0[k].T + pos0[k], atol=1e-12), k
            continue
        mol = dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
        ref = oracle.ene_inter(s["rec"], mol, shifted=True) + oracle.ene_intra(mol)
        sabs = inter_abs_sum(s["rec"], x, mol, True) + intra_abs_sum(mol, x)
        assert_energy(st["best_e"][k], ref, sabs, what="chain %d best" % k)
        scored += 1
    assert scored >= len(cl) // 2
    fin = np.isfinite(st["best_e"]) & np.isfinite(sp["best_e"])
    assert (st["best_e"][fin] <= sp["best_e"][fin] + 1e-9).mean() >= 0.9
    per, _ = m.min_evals()
    assert (per > 10).all()


def test_mc_min_launch_forms_agree(system):
    """Minimising chains in one persistent launch, hipGraph replays and one launch per
    ligand: bit-identical results."""
    s = system
    rng = np.random.default_rng(7)
    cl = np.repeat(np.array(s["ids"][:4]), 3)
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(cl), 3))
    a = _mc_min(s, cl, rot0, pos0, ff.SHIFTED, 6, 20, seed=23)
    b = _mc_min(s, cl, rot0, pos0, ff.SHIFTED, 6, 20, seed=23)
    c = _mc_min(s, cl, rot0, pos0, ff.SHIFTED, 6, 20, seed=23)
    print("mc+min launch forms", flush=True)
    a.run(6)
    b.run_graph(2, nreplay=3)
    c.run_per_ligand(3)
    c.run_per_ligand(3)
    sa, sb, sc = a.state(), b.state(), c.state()
    for key in ("best_e", "e", "accepted", "resets", "frame", "max_rot"):
        assert np.array_equal(sa[key], sb[key]), key
        assert np.array_equal(sa[key], sc[key]), key
    assert all(np.array_equal(x, y) for x, y in zip(a.coords()[1], b.coords()[1]))
    assert np.array_equal(a.min_evals()[0], c.min_evals()[0])


def test_mc_graph_recaptured_after_the_registry_grows(system):
    """A hipGraph of the chains holds the ligand registry's device pointers: registering
    more ligands between replays (the registry's buffers grow and move) makes the next
    replay capture again, and the chains equal those of plain launches."""
    from mmo_amd import mc as mcmod
    s = system
    ctx = ff.Context(0)
    ctx.set_receptor(s["rec"])
    ids = [ctx.add_ligand(l) for l in s["lib"][:3]]
    rng = np.random.default_rng(8)
    cl = np.repeat(np.array(ids), 2)
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(cl), 3))
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    kw = dict(total_frames=40, flags=flags, roi_center=SURFACE, roi_r=5.0, seed=24)
    a = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, **kw)
    b = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, **kw)
    print("graph, registry growth, graph", flush=True)
    a.run_graph(10, nreplay=2)
    for l in synth.library(40, seed=77):        # the registry's buffers grow
        ctx.add_ligand(l)
    a.run_graph(10, nreplay=2)
    b.run(40)
    sa, sb = a.state(), b.state()
    for key in ("best_e", "e", "accepted", "resets", "frame"):
        assert np.array_equal(sa[key], sb[key]), key


def test_receptor_in_lds_or_global_memory_bit_identical(system):
    """The direct forms read the receptor and its cell table from LDS when they fit and
    from global memory when they do not (large receptors): both reads give the same
    numbers, so the minimiser, plain chains and minimising chains are bit-identical."""
    import torch
    from mmo_amd import mc as mcmod
    from mmo_amd.minimize import Minimizer
    s = system
    ctx = s["ctx"]
    rng = np.random.default_rng(9)
    pose_lig, d0 = _start_dofs(s, rng)
    b = ctx.batch(pose_lig)
    cl = np.repeat(np.array(s["ids"][:4]), 2)
    rot0 = synth.random_rotations(rng, len(cl))
    pos0 = SURFACE + rng.uniform(-1.0, 1.0, (len(cl), 3))
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    out = {}
    for lds in (1, 0):
        ctx.tune(min_rec_lds=lds, mc_rec_lds=lds, min_fused=1)
        d = torch.tensor(np.concatenate(d0), device="cuda:0")
        mz = Minimizer(b, flags=ff.F_INTER_SHIFTED | ff.F_INTRA, history=8, max_step=0.2)
        print("receptor in LDS %d: minimiser" % lds, flush=True)
        e = mz.run(d, 40).cpu().numpy()
        plain = mcmod.MonteCarlo(ctx, cl, rot0, pos0, ff.SHIFTED, total_frames=60, flags=flags,
                                 roi_center=SURFACE, roi_r=5.0, seed=31)
        print("receptor in LDS %d: chains" % lds, flush=True)
        plain.run(60)
        mm = _mc_min(s, cl, rot0, pos0, ff.SHIFTED, 6, 20, seed=32)
        print("receptor in LDS %d: minimising chains" % lds, flush=True)
        mm.run(6)
        out[lds] = (e, d.cpu().numpy(), plain.state(), plain.coords()[1], mm.state(), mm.coords()[1])
    ctx.tune(min_rec_lds=1, mc_rec_lds=1)
    a, g = out[1], out[0]
    assert np.array_equal(a[0], g[0]) and np.array_equal(a[1], g[1])
    for sa, sg, ba, bg in ((a[2], g[2], a[3], g[3]), (a[4], g[4], a[5], g[5])):
        for key in ("best_e", "e", "accepted", "resets", "frame", "max_rot", "max_trans"):
            assert np.array_equal(sa[key], sg[key]), key
        assert all(np.array_equal(x, y) for x, y in zip(ba, bg))
