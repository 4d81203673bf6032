"""CPU checks of the Monte Carlo random stream and the CPU restatement used by the GPU
tests."""
import math

import numpy as np
import pytest

from mmo_amd.mc import frame_uniforms, philox4x32


def test_philox_known_answers():
    """Random123's Philox4x32-10 known-answer vectors."""
    assert philox4x32([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert philox4x32([0xffffffff] * 4, [0xffffffff, 0xffffffff]) == \
        [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert philox4x32([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                      [0xa4093822, 0x299f31d0]) == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_frame_uniforms_range_and_independence():
    u = np.array([frame_uniforms(7, c, f) for c in range(8) for f in range(64)])
    assert u.shape == (512, 6)
    assert (u >= 0).all() and (u < 1).all()
    assert abs(u.mean() - 0.5) < 0.02
    assert abs(np.corrcoef(u[:, 0], u[:, 1])[0, 1]) < 0.1


def test_cpu_restatement_runs_and_accepts_with_hard_roi(config0):
    from tests.mc_ref import simulate
    from mmo_amd import synth
    lig, rec = config0
    sub = {k: np.asarray(rec[k])[:400] for k in ("xs", "ys", "zs", "q", "anum")}
    r = simulate(lig, sub, np.eye(3), np.array([9.0, 0.0, 0.0]), 30, seed=3, chain=0,
                 hard_roi=(np.array([9.0, 0.0, 0.0]), 5.0))
    assert r["accepted"] > 0 and r["status"] == 0
    r2 = simulate(lig, sub, np.eye(3), np.array([9.0, 0.0, 0.0]), 30, seed=3, chain=0)
    assert r2["accepted"] == 0   # no acceptance test without hard ROI (src/lds.ml:908-917)
    assert math.isfinite(r2["best_e"])


def _small_system(config0):
    lig, rec = config0
    sub = {k: np.asarray(rec[k])[:400] for k in ("xs", "ys", "zs", "q", "anum")}
    return lig, sub


def _energy_of_state(lig, rec, r):
    import oracle
    x = (r["conf"] - r["center"]) @ r["rot"].T + r["pos"]
    m = dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2])
    c = dict(lig, xs=r["conf"][:, 0], ys=r["conf"][:, 1], zs=r["conf"][:, 2])
    return oracle.ene_inter(rec, m) + oracle.ene_intra(c)


def test_restatement_state_energy_and_best(config0):
    """The chain's current energy is the energy of its current pose and conformer, and
    the best energy is the minimum over the start pose and every proposal scored."""
    from tests.mc_ref import simulate
    lig, rec = _small_system(config0)
    tr = []
    start = np.array([9.0, 0.0, 0.0])
    r = simulate(lig, rec, np.eye(3), start, 40, seed=11, chain=2, hard_roi=(start, 50.0), trace=tr)
    assert r["resets"] == 0 and r["accepted"] > 0
    assert r["e"] == pytest.approx(_energy_of_state(lig, rec, r), rel=1e-12)
    r0 = simulate(lig, rec, np.eye(3), start, 0, seed=11, chain=2, hard_roi=(start, 50.0))
    assert r["best_e"] == min([r0["e"]] + [t["e_prop"] for t in tr])


def test_restatement_metropolis_limits(config0):
    """beta -> infinity: accepted energies never increase; beta = 0: every proposal
    inside the ROI is accepted."""
    from tests.mc_ref import simulate
    lig, rec = _small_system(config0)
    start = np.array([9.0, 0.0, 0.0])
    tr = []
    simulate(lig, rec, np.eye(3), start, 40, seed=5, chain=0, hard_roi=(start, 50.0),
             beta=1e300, trace=tr)
    e = [t["e"] for t in tr]
    assert all(b <= a for a, b in zip(e, e[1:]))
    r = simulate(lig, rec, np.eye(3), start, 40, seed=5, chain=0, hard_roi=(start, 50.0), beta=0.0)
    assert r["accepted"] == 40 and r["resets"] == 0


def test_detailed_scores_lines(tmp_path):
    """Lds.rescore's line (src/lds.ml:687: "%s\t%g\t%d\t%d\n") and header (src/lds.ml:1806);
    heavy atoms are anum > 1 (Mol.num_heavy_atoms, src/mol.ml:67-68)."""
    from mmo_amd import mc as mcmod
    assert mcmod.detailed_scores_line("lig_1", -12.3456789, 3, 20) == "lig_1\t-12.3457\t3\t20\n"
    assert mcmod.detailed_scores_line("x", 1.0e-5, 0, 1) == "x\t1e-05\t0\t1\n"
    assert mcmod.detailed_scores_line("y", 123456789.0, 15, 70) == "y\t1.23457e+08\t15\t70\n"
    lig = {"anum": np.array([6, 1, 8, 1, 0, 17]), "rb_left": [0, 2]}
    fn = tmp_path / "detailed_scores.tsv"
    mcmod.write_detailed_scores(str(fn), ["a"], [-1.5], [lig])
    assert fn.read_text() == "#name\tE_inter\tRB\tHA\na\t-1.5\t2\t3\n"


# ---------------------------------------------------------------------------------
# Control rules of Lds.simulate_lig pinned by closed-form outcomes past the first
# adaptation windows (src/lds.ml:586-621, :632-648, :781-798, :971-988;
# src/params.ml:10-33).  Each scenario forces every Metropolis decision (beta = 0
# accepts everything, beta -> inf with a rising energy rejects everything), so the
# step sizes after each 100-event window are known exactly.

DEG = math.pi / 180.0
START = np.array([9.0, 0.0, 0.0])


class _Rising:
    """E_inter that rises at every evaluation (never 0, so never an E = 0 reset)."""

    def __init__(self):
        self.k = 0

    def __call__(self, x):
        self.k += 1
        return 1.0 + 1e-3 * self.k


def _trace(config0, nframes, **kw):
    from tests.mc_ref import simulate
    lig, rec = _small_system(config0)
    tr = []
    kw.setdefault("hard_roi", (START, 50.0))
    r = simulate(lig, rec, np.eye(3), START, nframes, trace=tr, **kw)
    return lig, r, tr


def test_rigid_steps_grow_by_5_percent_when_everything_is_accepted(config0):
    """beta = 0: the rigid window holds only accepts (ratio 1 >= 0.55), so at rigid
    steps 100 and 200 (frames 200 and 400) both steps grow by 1.05."""
    lig, r, tr = _trace(config0, 402, seed=21, chain=1, beta=0.0, tweak=False)
    assert r["resets"] == 0 and r["status"] == 0 and r["accepted"] == 402
    assert tr[198]["max_trans"] == 0.15 and tr[198]["max_rot"] == 15.0 * DEG
    assert tr[199]["max_trans"] == 0.15                       # conformer frame: no update
    assert tr[200]["max_trans"] == 1.05 * 0.15 and tr[200]["max_rot"] == 1.05 * (15.0 * DEG)
    assert tr[398]["max_trans"] == 1.05 * 0.15
    assert tr[400]["max_trans"] == 1.05 * (1.05 * 0.15)
    assert tr[400]["max_rot"] == min(math.pi, 1.05 * (1.05 * (15.0 * DEG)))


def test_rigid_steps_shrink_by_5_percent_when_everything_is_rejected(config0):
    """beta -> inf and a rising energy: every proposal is rejected (ratio 0 <= 0.45),
    so at rigid step 100 both steps shrink by 0.95; nothing is accepted."""
    lig, r, tr = _trace(config0, 202, seed=22, chain=0, beta=1e300, inter_fn=_Rising(),
                        intra=False, tweak=False)
    assert r["accepted"] == 0 and r["resets"] == 0
    assert tr[198]["max_trans"] == 0.15
    assert tr[200]["max_trans"] == 0.95 * 0.15 and tr[200]["max_rot"] == 0.95 * (15.0 * DEG)


def _dihedral_across(lig, conf, left, right, group):
    """Dihedral a-left-right-b with a bonded to left (off the moving side) and b bonded
    to right on the moving side (topological distance 1 in dists)."""
    import oracle
    n = len(lig["xs"])
    d = np.asarray(lig["dists"]).reshape(n, n)
    gset = set(int(g) for g in group)
    a = next(i for i in range(n) if d[i, left] == 1 and i != right and i not in gset)
    b = next(i for i in gset if d[i, right] == 1)
    return oracle.dihedral(conf[a], conf[left], conf[right], conf[b])


def _torsion_change(lig, tr, frame):
    """|change of the dihedral across the bond turned at `frame`| (radians, in [0, pi])."""
    k = tr[frame]["bond"]
    left, right = int(lig["rb_left"][k]), int(lig["rb_right"][k])
    g = lig["rg_atoms"][lig["rg_off"][k]:lig["rg_off"][k + 1]]
    d0 = _dihedral_across(lig, tr[frame - 1]["conf"], left, right, g)
    d1 = _dihedral_across(lig, tr[frame]["conf"], left, right, g)
    dd = abs(d1 - d0) % (2.0 * math.pi)
    return min(dd, 2.0 * math.pi - dd)


def test_flip_at_every_100th_conformer_step(config0):
    """Conformer step 100 (frame 201) turns its bond by U(-pi, pi) (Mol.flip_rbond);
    the conformer steps around it tweak by at most the bond's step (5 degrees).  The
    change is measured as the dihedral across the bond."""
    lig, r, tr = _trace(config0, 204, seed=3, chain=5, beta=0.0)
    assert r["status"] == 0 and r["resets"] == 0
    assert tr[201]["bond"] >= 0
    flip = _torsion_change(lig, tr, 201)
    assert flip == pytest.approx(abs(tr[201]["angle"]), abs=1e-9)
    assert flip > 5.0 * DEG                   # this seed's draw is outside the tweak range
    for f in (199, 203):
        assert _torsion_change(lig, tr, f) <= 5.0 * DEG + 1e-12
    # --no-flip-move: frame 201 is an ordinary tweak
    lig, r2, tr2 = _trace(config0, 204, seed=3, chain=5, beta=0.0, no_flip=True)
    assert _torsion_change(lig, tr2, 201) <= 5.0 * DEG + 1e-12


def test_bond_steps_update_every_100_nrb_conformer_steps_and_only_when_kept(config0):
    """Per-bond steps adapt at conformer step 100 nrb (frame 200 nrb + 1) on the
    proposal's copy (src/lds.ml:976-979): with everything accepted each bond that has
    events grows by 1.05; with everything rejected the shrunk copy is discarded."""
    lig, _ = _small_system(config0)
    nrb = len(lig["rb_left"])
    fr = 200 * nrb + 1
    lig, r, tr = _trace(config0, fr + 3, seed=8, chain=2, beta=0.0, no_flip=True)
    assert r["status"] == 0 and r["resets"] == 0
    assert tr[fr - 2]["dr"] == [5.0 * DEG] * nrb
    used = set(t["bond"] for t in tr[:fr + 1] if t["bond"] >= 0)
    assert len(used) == nrb
    assert tr[fr]["dr"] == [1.05 * (5.0 * DEG)] * nrb
    lig, r, tr = _trace(config0, fr + 3, seed=8, chain=2, beta=1e300, inter_fn=_Rising(),
                        intra=False, no_flip=True)
    assert r["accepted"] == 0
    assert tr[fr]["dr"] == [5.0 * DEG] * nrb


def test_reset_restores_the_start_pose_of_the_registered_conformer(config0):
    """Leaving the hard ROI resets the chain (src/lds.ml:908-913): steps back to their
    defaults, best energy +inf and best pose = the registered conformer at rot0, pos0
    (reset_run_params, src/lds.ml:632-648), whatever the current conformer is."""
    lig, r, tr = _trace(config0, 400, seed=5, chain=3, beta=0.0, hard_roi=(START, 0.6))
    resets = [t["frame"] for t in tr if t["reset"]]
    assert resets, "the chain never left the 0.6 A ROI"
    c0 = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    changed = False
    for f in resets:
        t = tr[f]
        assert t["best_e"] == math.inf and t["e"] == math.inf
        assert t["max_trans"] == 0.15 and t["max_rot"] == 15.0 * DEG
        assert np.array_equal(t["best_xyz"], c0 @ np.eye(3).T + START)
        changed |= not np.allclose(t["conf"], c0)
    assert changed, "no reset happened with a turned conformer"


# ------------------------------------------------------------- MC plus minimisation
def test_mc_min_with_one_evaluation_is_plain_mc(config0):
    """One L-BFGS evaluation per frame scores the proposal itself: the chain takes the
    plain Monte Carlo's decisions and energies.  (Without E_intra: the plain chain keeps
    a rejected torsion proposal's E_intra for the next frames, src/lds.ml:900-903, while
    the minimising chain scores the pose it holds.)"""
    lig, r0, tr0 = _trace(config0, 40, seed=31, chain=4, intra=False)
    lig, r1, tr1 = _trace(config0, 40, seed=31, chain=4, intra=False, minimize={"iters": 1})
    assert r1["accepted"] == r0["accepted"] > 0 and r1["resets"] == r0["resets"]
    for a, b in zip(tr0, tr1):
        assert b["e"] == pytest.approx(a["e"], rel=1e-10, abs=1e-9)
        assert b["e_prop"] == pytest.approx(a["e_prop"], rel=1e-10, abs=1e-9)
    # the minimising chain stores its conformer centred: equal up to that translation
    assert np.allclose(r1["conf"] - r1["center"], r0["conf"] - r0["center"], atol=1e-10)
    assert np.allclose(r1["rot"], r0["rot"], atol=1e-12) and np.allclose(r1["pos"], r0["pos"], atol=1e-12)
    assert r1["min_evals"] == 40


def test_mc_min_frames_lower_the_proposal_and_keep_a_consistent_state(config0):
    """With 30 evaluations per frame: every minimised energy is at most the proposal's,
    most frames go lower, and the chain's current energy is the oracle energy of its pose
    (conformer, centre, rotation, position)."""
    lig, r, tr = _trace(config0, 12, seed=32, chain=1, minimize={"iters": 30, "max_step": 0.2})
    assert r["status"] == 0 and r["accepted"] > 0
    lower = 0
    for t in tr:
        assert t["e_prop"] <= t["e_start"]
        lower += t["e_prop"] < t["e_start"] - 1e-6
    assert lower >= len(tr) // 2
    _, rec = _small_system(config0)
    assert r["e"] == pytest.approx(_energy_of_state(lig, rec, r), rel=1e-9)
    assert 12 < r["min_evals"] <= 12 * 30
