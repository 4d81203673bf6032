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
