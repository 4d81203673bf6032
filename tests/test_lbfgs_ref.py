"""The CPU restatement of the library's L-BFGS (tests/lbfgs_ref.py) pinned by closed
forms: the exact minimum of a convex quadratic, Rosenbrock's minimum, the projection
onto box bounds, and failed (too long) trials treated as +inf."""
import math

import numpy as np
import pytest

from tests.lbfgs_ref import CONVERGED, LBFGS, minimize


def test_convex_quadratic_reaches_the_exact_minimum():
    rng = np.random.default_rng(1)
    Q = rng.normal(size=(8, 8))
    A = Q @ Q.T + 8 * np.eye(8)
    b = rng.normal(size=8)

    def fg(x):
        return 0.5 * x @ A @ x - b @ x, A @ x - b, 0.0
    opt = minimize(fg, np.zeros(8), 200, m=8, gtol=1e-10)
    assert opt.status == CONVERGED
    assert np.allclose(opt.x, np.linalg.solve(A, b), atol=1e-9)
    assert opt.f <= 0.5 * opt.x @ A @ opt.x - b @ opt.x + 1e-15


def test_rosenbrock_minimum():
    def fg(x):
        a, c = x
        f = (1 - a) ** 2 + 100 * (c - a * a) ** 2
        g = np.array([-2 * (1 - a) - 400 * a * (c - a * a), 200 * (c - a * a)])
        return f, g, 0.0
    opt = minimize(fg, np.array([-1.2, 1.0]), 400, m=8, gtol=1e-9)
    assert np.allclose(opt.x, [1.0, 1.0], atol=1e-6)
    assert opt.n_acc > 10


def test_accepted_energies_never_increase():
    def fg(x):
        return float(np.sum(np.cos(3 * x) + 0.1 * x * x)), -3 * np.sin(3 * x) + 0.2 * x, 0.0
    opt = LBFGS(5, m=4, max_step=0.3)
    xt = np.linspace(-2, 2, 5)
    fs = []
    for _ in range(60):
        f, g, r = fg(xt)
        xt = opt.update(xt, f, g, r)
        fs.append(opt.f)
    assert all(b <= a for a, b in zip(fs, fs[1:]))


def test_box_bounds_project_the_trial_points():
    c = np.array([3.0, -3.0, 0.5])

    def fg(x):
        return float(np.sum((x - c) ** 2)), 2 * (x - c), 0.0
    opt = minimize(fg, np.zeros(3), 100, lo=[-1, -1, -1], hi=[1, 1, 1], gtol=1e-12)
    assert np.allclose(opt.x, [1.0, -1.0, 0.5], atol=1e-9)


def test_too_long_trials_count_as_infinite_energy():
    """A trial past radius 12 is rejected and the step halved, although its energy
    would satisfy the sufficient-decrease test."""
    opt = LBFGS(1, max_step=0.4)
    x0 = np.array([0.45])
    t1 = opt.update(x0, float(x0 @ x0), 2 * x0, 1.0)      # first trial: 0.45 - 0.4
    assert t1[0] == pytest.approx(0.05)
    t2 = opt.update(t1, float(t1 @ t1), 2 * t1, 13.0)     # too long
    assert opt.x[0] == 0.45 and opt.n_acc == 0
    assert t2[0] == pytest.approx(0.25)
    t3 = opt.update(t2, float(t2 @ t2), 2 * t2, 1.0)      # accepted
    assert opt.n_acc == 1 and opt.x[0] == pytest.approx(0.25)
