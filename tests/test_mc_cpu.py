"""CPU checks of the Monte Carlo random stream and the CPU restatement used by the GPU
tests."""
import math

import numpy as np

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
