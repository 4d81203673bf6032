"""CPU restatement of Lds.simulate_lig (src/lds.ml:741-1000) for checking the GPU chains:
FP64 energies from the oracle, the same Philox uniforms as the kernel (one stream per
chain and frame), the reference's control flow including the dangling `else` of
src/lds.ml:917 (no acceptance test without hard ROI) and its adaptive step rules."""
import math

import numpy as np

import oracle
from mmo_amd.mc import frame_uniforms

PI = 4.0 * math.atan(1.0)
DEG = PI / 180.0


class SW:
    """SW.t, src/SW.ml:7-42."""

    def __init__(self, size=100):
        self.size, self.events = size, []

    def process(self, ev):
        self.events.append(bool(ev))
        if len(self.events) > self.size:
            self.events.pop(0)

    def ratio(self):
        n = len(self.events)
        return sum(self.events) / n if n else float("nan")

    def reset(self):
        self.events = []


def _decision_margin(cur, prev_e, best_e, u, beta, scale):
    """Distance of this frame's two energy comparisons from their boundaries, relative
    to `scale` (the magnitudes of the E_inter values involved: only E_inter carries the
    GPU's FP32 pair error; E_intra is FP64 on both sides).  The proposal is accepted iff
    cur - prev < T = -ln(u) / beta (which covers cur <= prev) and becomes the best pose
    iff cur < best; an E_inter error below this fraction of `scale` changes neither."""
    m = math.inf
    if math.isfinite(prev_e) and math.isfinite(cur):
        T = math.inf if (beta == 0.0 or u == 0.0) else -math.log(u) / beta
        if math.isfinite(T):
            m = abs((cur - prev_e) - T) / scale
    if math.isfinite(best_e) and math.isfinite(cur):
        m = min(m, abs(cur - best_e) / scale)
    return m


def minimise_proposal(lig, conf2, c2, rot2, pos2, ntors, intra, inter_grad, mz):
    """The library's Monte Carlo-plus-minimisation step on the CPU: base rows
    b = rot2 (conf2 - c2) (the proposal centred at the origin), L-BFGS (tests/lbfgs_ref.py)
    from [pos2, 0, 0, 0, 0..] over the FP64 energy and its exact DOF gradient
    (oracle.dof_energy_grad).  Returns (E, E_inter, E_intra, conformer of the best point
    before the rigid motion, M, T, evaluations)."""
    from tests.lbfgs_ref import LBFGS
    base = (conf2 - c2) @ rot2.T
    n = 6 + ntors
    opt = LBFGS(n, m=mz.get("m", 8), max_step=mz.get("max_step", 0.2), gtol=mz.get("gtol", 1e-6))
    xt = np.zeros(n)
    xt[:3] = pos2
    parts = (0.0, 0.0)
    for _ in range(mz["iters"]):
        if opt.init and opt.status != 0:
            break
        ei, ea, g, rad = oracle.dof_energy_grad(base, lig, xt, inter_grad, intra=intra)
        init0, acc0 = opt.init, opt.n_acc
        xt = opt.update(xt, ei + ea, g, rad)
        if not init0 or opt.n_acc != acc0:
            parts = (ei, ea)
    x = opt.x
    d0 = np.concatenate([np.zeros(6), x[6:]])
    r = len(x) - 6
    rg_off = np.asarray(lig["rg_off"])[:r + 1]
    tconf, _ = oracle.apply_config(base, np.asarray(lig["rb_left"])[:r], np.asarray(lig["rb_right"])[:r],
                                   rg_off, np.asarray(lig["rg_atoms"])[:int(rg_off[-1])], d0)
    M = oracle.rot_xyz(x[3], x[4], x[5])
    return opt.f, parts[0], parts[1], tconf, M, x[:3].copy(), opt.n_eval, opt.f0


def simulate(lig, rec, rot0, pos0, nframes, seed, chain, tweak=True, intra=True, hard_roi=None,
             metropolis_always=False, beta=1.0 / (0.0019872041 * 293.15), inter_fn=None,
             trace=None, no_flip=False, minimize=None, inter_grad=None):
    """Runs one chain for `nframes`; `trace`, when a list, receives one dict per frame.
    best_xyz is best_lig: the pose of the best proposal, or after a reset the registered
    conformer at the start pose (src/lds.ml:645-646).

    minimize = {"iters": N, "m": .., "max_step": .., "gtol": ..} runs the library's
    Monte Carlo-plus-minimisation (not the reference's: its MC and its minimiser are
    separate): every proposal is minimised first and the frame's decisions apply to the
    minimised pose; inter_grad(x) -> (E_inter, dE/dx) is the energy form's gradient
    (default: the shifted direct sum against `rec`)."""
    n = len(lig["xs"])
    conf = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1).astype(np.float64)
    center = np.zeros(3)
    nrb = len(lig["rb_left"])
    flex = tweak and nrb > 0
    groups = [np.asarray(lig["rg_atoms"][lig["rg_off"][k]:lig["rg_off"][k + 1]]) for k in range(nrb)]
    rot, pos = np.array(rot0, dtype=float).reshape(3, 3), np.array(pos0, dtype=float)
    max_rot0, max_trans0 = 15.0 * DEG, 0.15
    max_rot, max_trans = max_rot0, max_trans0
    dr = [5.0 * DEG] * nrb
    sw = SW()
    rsw = [SW() for _ in range(nrb)]

    def e_intra(c):
        return oracle.ene_intra(dict(lig, xs=c[:, 0], ys=c[:, 1], zs=c[:, 2])) if intra else 0.0

    def e_inter(x):
        if inter_fn is not None:
            return inter_fn(x)
        return oracle.ene_inter(rec, dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2]), shifted=True)

    if inter_grad is None:
        def inter_grad(x):
            return oracle.grad_inter(rec, dict(lig, xs=x[:, 0], ys=x[:, 1], zs=x[:, 2]), shifted=True)
    min_evals = 0

    conf0 = conf.copy()
    start = conf @ rot.T + pos
    best_xyz = start.copy()
    prev_intra = e_intra(conf)
    prev_inter = e_inter(start)
    prev_e = prev_inter + prev_intra
    best_e = prev_e
    prev_e_inter = best_inter = prev_inter   # E_inter of the current state / best pose
    best_rot, best_pos = rot.copy(), pos.copy()
    rigid_step = conf_step = 0
    accepted = resets = ooroi = 0
    status = 0
    for frame in range(nframes):
        u = frame_uniforms(seed, chain, frame)
        rigid = frame % 2 == 0
        conf2, c2, jr, ndr = conf, center, -1, list(dr)
        rot2, pos2 = rot, pos
        ang = None
        if not rigid:
            if flex:
                i = min(int(u[0] * nrb), nrb - 1)
                flip = conf_step > 0 and conf_step % 100 == 0 and not no_flip
                amp = PI if flip else ndr[i]
                ang = u[1] * (2.0 * amp) - amp
                conf2 = oracle.rotate_bond(conf, int(lig["rb_left"][i]), int(lig["rb_right"][i]),
                                           groups[i], ang)
                c2 = conf2.sum(axis=0) / n
                rad = np.max(0.01 + np.linalg.norm(conf2 - c2, axis=1))
                if rad > 12.0:
                    status = 2
                    break
                jr = i
        else:
            th = u[0] * (2.0 * max_rot) - max_rot
            ax = min(int(u[1] * 3.0), 2)
            A = (oracle.rot_x, oracle.rot_y, oracle.rot_z)[ax](th)
            rot2 = A @ rot
            pos2 = pos + (np.array(u[2:5]) * 2.0 - 1.0) * max_trans
        pose = (conf2 - c2) @ rot2.T + pos2
        cpos = pos2
        e_start = None
        if minimize is not None:
            cur, prev_inter, prev_intra, tconf, mM, mT, ne, e_start = minimise_proposal(
                lig, conf2, c2, rot2, pos2, nrb if flex else 0, intra, inter_grad, minimize)
            min_evals += ne
            pose = tconf @ mM.T + mT
            cpos = mM @ tconf.mean(axis=0) + mT
        else:
            if not rigid and flex and intra:
                prev_intra = e_intra(conf2)
            prev_inter = e_inter(pose)
            cur = prev_inter + prev_intra
        metropolis = False
        reset = False
        margin = math.inf
        if hard_roi is not None:
            if np.linalg.norm(np.asarray(hard_roi[0]) - cpos) > hard_roi[1]:
                reset = True
                ooroi += 1
            else:
                metropolis = True
        else:
            metropolis = metropolis_always
        if metropolis and prev_inter == 0.0:
            reset, metropolis = True, False
        if reset:
            max_rot, max_trans = max_rot0, max_trans0
            rot, pos = np.array(rot0, dtype=float).reshape(3, 3), np.array(pos0, dtype=float)
            best_rot, best_pos = np.eye(3), np.zeros(3)
            best_xyz = conf0 @ rot.T + pos
            prev_e = math.inf
            best_e = math.inf
            sw.reset()
            resets += 1
        elif metropolis:
            acc = (cur <= prev_e) or (u[5] < math.exp(-(cur - prev_e) * beta))
            margin = _decision_margin(cur, prev_e, best_e, u[5], beta,
                                      abs(prev_inter) + abs(prev_e_inter) + abs(best_inter) + 1e-300)
            if rigid:
                sw.process(acc)
            elif jr > -1:
                rsw[jr].process(acc)
            if acc:
                accepted += 1
                rot, pos, prev_e = rot2, pos2, cur
                prev_e_inter = prev_inter
                conf, center = conf2, c2
                if minimize is not None:
                    # the minimised pose in the chain's frame: conformer rot2^T c*, its
                    # centre, rotation M* rot2, position the pose centre
                    conf = tconf @ rot2
                    center = tconf.mean(axis=0) @ rot2
                    rot, pos = mM @ rot2, cpos
            if cur < best_e:
                best_e, best_rot, best_pos = cur, rot2, pos2
                if minimize is not None:
                    best_rot, best_pos = mM @ rot2, cpos
                best_inter = prev_inter
                best_xyz = pose.copy()
            if rigid and rigid_step > 0 and rigid_step % 100 == 0:
                ar = sw.ratio()
                if ar <= 0.45:
                    max_trans *= 0.95
                    max_rot *= 0.95
                elif ar >= 0.55:
                    max_trans *= 1.05
                    max_rot = min(PI, 1.05 * max_rot)
            if flex and not rigid and conf_step > 0 and conf_step % (100 * nrb) == 0:
                for k in range(nrb):
                    ar = rsw[k].ratio()
                    if ar <= 0.45:
                        ndr[k] = 0.95 * ndr[k]
                    elif ar >= 0.55:
                        ndr[k] = min(PI, 1.05 * ndr[k])
                if acc:
                    dr = ndr
        if rigid:
            rigid_step += 1
        else:
            conf_step += 1
        if trace is not None:
            trace.append({"frame": frame, "e_inter": prev_inter, "e_intra": prev_intra, "e_prop": cur,
                          "e": prev_e, "best_e": best_e, "accepted": accepted,
                          "resets": resets, "pos": pos.copy(), "rot": rot.copy(),
                          "conf": conf.copy(), "center": np.array(center), "pose": pose.copy(),
                          "bond": jr, "angle": ang, "max_rot": max_rot, "max_trans": max_trans,
                          "dr": list(dr), "best_xyz": best_xyz.copy(), "reset": reset,
                          "margin": margin, "e_start": e_start})
    return {"best_e": best_e, "e": prev_e, "e_inter": prev_inter, "e_intra": prev_intra,
            "rot": rot, "pos": pos, "conf": conf, "center": np.array(center), "accepted": accepted, "resets": resets,
            "ooroi": ooroi, "status": status, "max_rot": max_rot, "max_trans": max_trans,
            "best_rot": best_rot, "best_pos": best_pos, "best_xyz": best_xyz, "dr": dr,
            "rigid_step": rigid_step, "conf_step": conf_step, "min_evals": min_evals}
