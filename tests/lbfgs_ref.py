"""CPU restatement of the library's batched L-BFGS update (one pose), for checking the
GPU minimiser and the Monte Carlo-plus-minimisation chains.  It follows the kernel's
control flow step by step: one energy + gradient evaluation per iteration, Armijo
sufficient decrease (c1) on the trial point, curvature pairs kept when s.y > 0, the
two-loop recursion over at most m pairs (newest first), restart on a non-descent
direction, step halving on a failed trial and a restart (or convergence) once the step
falls below min_step, first trial steps capped at max_step per component, box bounds on
trial points, convergence at max |g| <= gtol.  Poses whose torsions make the ligand
longer than 12 A (Mol.Too_long, src/optim.ml:72-75) count as failed trials.

The reference minimises with NLopt on central differences (src/lds.ml:1014-1036,
:1118-1297); this minimiser is the library's own algorithm, so the restatement pins the
GPU implementation, not the reference's search path."""
import math

import numpy as np

RUNNING, CONVERGED, FAILED = 0, 1, 2


class LBFGS:
    def __init__(self, n, m=8, max_step=1.0, gtol=1e-6, min_step=1e-10, c1=1e-4, lo=None, hi=None):
        self.n, self.m = n, m
        self.max_step, self.gtol, self.min_step, self.c1 = max_step, gtol, min_step, c1
        self.lo = np.full(n, -np.inf) if lo is None else np.asarray(lo, float)
        self.hi = np.full(n, np.inf) if hi is None else np.asarray(hi, float)
        self.S = np.zeros((m, n))
        self.Y = np.zeros((m, n))
        self.rho = np.zeros(m)
        self.init = False
        self.hist = self.head = 0
        self.status = RUNNING
        self.f = self.f0 = self.alpha = self.gamma = 0.0
        self.n_eval = self.n_acc = self.n_reset = 0
        self.x = self.g = self.d = None

    def update(self, xt, ft, gt, rad):
        """One evaluation at the trial point xt; returns the next trial point."""
        if self.init and self.status != RUNNING:
            return self.x.copy()          # a finished pose is no longer updated
        xt, gt = np.asarray(xt, float), np.asarray(gt, float)
        if not (rad <= 12.0) or not math.isfinite(ft):
            ft = math.inf
        self.n_eval += 1
        if not self.init:
            self.x, self.g = xt.copy(), gt.copy()
            self.f = self.f0 = ft
            self.init = True
            self.hist = self.head = 0
            if not math.isfinite(ft):
                self.status = FAILED
            self.d = -self.g
            dm = np.abs(self.d).max()
            self.alpha = min(1.0, self.max_step / dm) if dm > 0.0 else 0.0
            if dm <= self.gtol:
                self.status = CONVERGED
        else:
            sv = xt - self.x
            gs = float(self.g @ sv)
            if gs < 0.0 and ft <= self.f + self.c1 * gs:
                yv = gt - self.g
                sy, yy = float(sv @ yv), float(yv @ yv)
                if sy > 1e-12 * yy and sy > 0.0:
                    self.S[self.head], self.Y[self.head] = sv, yv
                    self.rho[self.head] = 1.0 / sy
                    self.head = (self.head + 1) % self.m
                    self.hist = min(self.hist + 1, self.m)
                    self.gamma = sy / yy
                self.x, self.g, self.f = xt.copy(), gt.copy(), ft
                self.n_acc += 1
                q = self.g.copy()
                al = []
                for k in range(self.hist):
                    i = (self.head - 1 - k + 2 * self.m) % self.m
                    al.append(self.rho[i] * float(self.S[i] @ q))
                    q -= al[k] * self.Y[i]
                r = (self.gamma if self.hist > 0 else 1.0) * q
                for k in range(self.hist - 1, -1, -1):
                    i = (self.head - 1 - k + 2 * self.m) % self.m
                    b = self.rho[i] * float(self.Y[i] @ r)
                    r += self.S[i] * (al[k] - b)
                self.d = -r
                if float(self.d @ self.g) >= 0.0:
                    self.hist = 0
                    self.n_reset += 1
                    self.d = -self.g
                dm, gm = np.abs(self.d).max(), np.abs(self.g).max()
                self.alpha = min(1.0, self.max_step / dm) if dm > 0.0 else 0.0
                if gm <= self.gtol:
                    self.status = CONVERGED
            else:
                self.alpha *= 0.5
                dm = np.abs(self.d).max()
                if self.alpha * dm < self.min_step:
                    if self.hist == 0:
                        self.status = CONVERGED
                    else:
                        self.hist = 0
                        self.n_reset += 1
                        self.d = -self.g
                        dm2 = np.abs(self.d).max()
                        self.alpha = min(1.0, self.max_step / dm2) if dm2 > 0.0 else 0.0
        if self.status == RUNNING:
            return np.minimum(self.hi, np.maximum(self.lo, self.x + self.alpha * self.d))
        return self.x.copy()


def minimize(fg, x0, iters, **kw):
    """Runs `iters` evaluations of fg(x) -> (f, g, radius) from x0 (the first evaluation is
    x0 itself); returns the LBFGS object (x: best point, f: its energy)."""
    opt = LBFGS(len(x0), **kw)
    xt = np.asarray(x0, float).copy()
    for _ in range(iters):
        f, g, rad = fg(xt)
        xt = opt.update(xt, f, g, rad)
    return opt
