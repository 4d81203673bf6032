"""Benchmark of the UFF scoring path on one GPU.

Headline (the JSON line's "value"): exhaustive rigid-pose scan with the plain
direct-sum method.  The 75-atom synthetic ligand (40 heavy + 35 H, 6 rotatable bonds)
and the 3000-atom Poisson-disc receptor; 4,608 super-Fibonacci SO(3) rotations x 20^3
lattice translations at 0.5 A = 36,864,000 poses per step, each scored with the shifted
LJ 12-6 + Coulomb direct sum (Mol.ene_inter_UFF_shifted_brute semantics).  Inputs are
resident on the GPU before timing starts; outputs stay on the GPU.

Secondary figures (the "extra" object), each on its own workload:
  scan_interp  the same scan through trilinear interpolation in 10 per-type FP32 grids of
               64^3 at 0.375 A, precomputed on the GPU (precompute time reported);
  egrad        energy + adjoint gradient in DOF space for 4,096 poses of 1,024 library
               ligands: one batched launch sequence vs one per ligand;
  min          200 batched L-BFGS iterations (energy + gradient + update) on those poses;
  mc           Monte Carlo chains, 64 per ligand for 1,024 library ligands, 1,000 frames,
               batched persistent launch replayed from a hipGraph vs one launch per ligand,
               direct sum and grid interpolation;
  scan_select  the scan's clash filters (ROI, centre and atom clashes against the protein
               vdW bitmask) and the top-100 of the filtered energies.

Prints one JSON line with "roofline" (the scan kernel's fraction of the FP32 peak) and
"cpu_baseline" (the FP64 C restatement timed on this host's cores).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3        # MI355X FP32 vector peak (AMD spec)
PEAK_HBM_GBS = 8000.0
# L2 read bandwidth: 16 channels of 128 B per clock in each XCD's L2, 8 XCDs, 2.4 GHz
# (the L2 structure AMD documents for CDNA 3, kept on CDNA 4; no MI355X figure is published)
PEAK_L2_GBS = 8 * 16 * 128 * 2.4
N_ROT, N_TRANS_SIDE, TRANS_STEP = 4608, 20, 0.5
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print("bench: " + msg, file=sys.stderr, flush=True)


def algorithmic_flops_per_pose(lig, rec, rots, trans, sample=4096, seed=0):
    """8 FLOP per (lig, rec) pair + 19 per pair inside 12 A (per-pose work constants);
    the in-cutoff count is the mean over a random sample of the scan's poses."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    tree = cKDTree(np.stack([rec["xs"], rec["ys"], rec["zs"]], 1))
    c = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
    n_in = 0
    for _ in range(sample):
        x = c @ rots[rng.integers(len(rots))].T + trans[rng.integers(len(trans))]
        n_in += sum(len(v) for v in tree.query_ball_point(x, 12.0 - 1e-9))
    L, P = len(c), len(rec["xs"])
    mean_in = n_in / sample
    return 8.0 * L * P + 19.0 * mean_in, mean_in


def host_cores():
    """Host cores this process may run on (its affinity set)."""
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()


def cpu_threads():
    """Threads for the CPU legs: the host cores, at most 16 (one GPU's share of a shared
    host, whatever the host shows)."""
    return max(1, min(host_cores(), 16))


def in_cutoff_pairs(rec, poses, cutoff=12.0):
    """(ligand atom, receptor atom) pairs closer than `cutoff` in each pose ((n, 3) arrays)."""
    from scipy.spatial import cKDTree
    tree = cKDTree(np.stack([rec["xs"], rec["ys"], rec["zs"]], 1))
    return np.array([int(tree.query_ball_point(x, cutoff - 1e-9, return_length=True).sum())
                     for x in poses], dtype=np.int64)


def valu_roofline(flop, kernel_ms, kernel, **more):
    tf = flop / (kernel_ms * 1e-3) / 1e12
    r = {"bound": "fp32-valu", "achieved": tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
         "frac": tf / PEAK_FP32_TFLOPS, "kernel": kernel, "kernel_ms": kernel_ms}
    r.update(more)
    return r


def gather_roofline(nbytes, kernel_ms, kernel, traffic=None, **more):
    """Trilinear gathers: the grids stay in L2 and the Infinity Cache, so the corner reads
    are measured against the L2 read bandwidth; `traffic` (bytes per launch beyond L2,
    from the FETCH_SIZE / WRITE_SIZE counters) is also given as a rate against HBM."""
    gbs = nbytes / (kernel_ms * 1e-3) / 1e9
    r = {"bound": "l2", "achieved": gbs, "peak": PEAK_L2_GBS, "unit": "GB/s",
         "frac": gbs / PEAK_L2_GBS, "kernel": kernel, "kernel_ms": kernel_ms,
         "traffic": traffic,
         "note": "grid corner reads (32 B per ligand atom) over the kernel time, served by "
                 "L2 and the Infinity Cache"}
    if traffic:
        r["beyond_l2_GBs"] = traffic / (kernel_ms * 1e-3) / 1e9
        r["beyond_l2_frac_of_hbm"] = r["beyond_l2_GBs"] / PEAK_HBM_GBS
    r.update(more)
    return r


def timed_threads(work, items, threads, target_s):
    """Run work(item) over `items` (cycled) on `threads` threads for about target_s
    seconds; returns (items done, seconds).  The oracle's C calls release the GIL."""
    from concurrent.futures import ThreadPoolExecutor
    n = max(threads, 4)
    while True:
        batch = [items[i % len(items)] for i in range(n)]
        t0 = time.perf_counter()
        if threads == 1:
            for it in batch:
                work(it)
        else:
            with ThreadPoolExecutor(threads) as ex:
                list(ex.map(work, batch))
        dt = time.perf_counter() - t0
        if dt >= target_s or n >= 1 << 22:
            return n, dt
        n = int(n * min(16.0, max(2.0, 1.1 * target_s / max(dt, 1e-3))))


def cpu_pair(work, items, unit, what, target_s=6.0, single_s=3.0):
    """A CPU figure on all CPU-leg threads and on one thread."""
    threads = cpu_threads()
    n, dt = timed_threads(work, items, threads, target_s)
    n1, dt1 = timed_threads(work, items, 1, single_s)
    return {"value": n / dt, "unit": unit, "cores": threads, "host_cores": host_cores(),
            "kind": "port", "sample": "%d %s, %.1f s" % (n, what, dt),
            "single_thread": {"value": n1 / dt1, "unit": unit, "cores": 1,
                              "sample": "%d %s, %.1f s" % (n1, what, dt1)}}


def cpu_baseline(lig, rec, rots, trans, target_s=12.0, single_s=4.0, brute_s=6.0):
    """FP64 C restatement of the reference's default scoring path (-ff Bst: per ligand atom
    the receptor atoms within 12 A from a neighbour search, pair parameters from a table;
    oracle, OpenMP over poses) on a random sample of the same scan's poses: all host
    threads (the reported value) and one thread (the reference scores one pose at a time
    per process); beside it the brute-force form (BrL: all 225,000 pairs tested)."""
    import oracle
    rng = np.random.default_rng(1)

    def run(fn, threads, target):
        n = 16
        while True:
            pr = rng.integers(0, len(rots), n).astype(np.int32)
            pt = rng.integers(0, len(trans), n).astype(np.int32)
            t0 = time.perf_counter()
            fn(rec, lig, rots, pr, trans, pt, nthreads=threads)
            dt = time.perf_counter() - t0
            if dt >= target or n >= 1 << 20:
                return n, dt
            n = int(n * min(16.0, max(2.0, 1.1 * target / max(dt, 1e-3))))

    threads = cpu_threads()
    n, dt = run(oracle.scan_shifted_cells, threads, target_s)
    n1, dt1 = run(oracle.scan_shifted_cells, 1, single_s)
    nb, dtb = run(oracle.scan_shifted, threads, brute_s)
    return {"value": n / dt, "unit": "poses/s", "cores": threads, "host_cores": host_cores(),
            "kind": "port",
            "sample": "%d random poses of the scan (shifted E_inter over the receptor atoms "
                      "within 12 A of each ligand atom, FP64, OpenMP), %.1f s" % (n, dt),
            "single_thread": {"value": n1 / dt1, "unit": "poses/s", "cores": 1,
                              "sample": "%d random poses, %.1f s" % (n1, dt1)},
            "brute_force": {"value": nb / dtb, "unit": "poses/s", "cores": threads,
                            "sample": "%d random poses, all 225,000 pairs tested, %.1f s"
                                      % (nb, dtb)}}


def counter_record(kernel):
    """The committed rocprofv3 counter figures of one kernel on one workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f).get(kernel) or {}
    except (OSError, ValueError):
        return {}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled)."""
    return counter_record(kernel).get("bytes_per_launch")


def executed_tflops(kernel, units, kernel_ms):
    """FP32 operations the kernel executes per unit of work (pose, MC step, evaluation:
    SQ_INSTS_VALU_{FMA,MUL,ADD,TRANS}_F32 x 64 lanes, FMA = 2, over the units of the
    profiled run) x this run's units, over this run's kernel time."""
    ops = counter_record(kernel).get("fp32_ops_per_unit")
    return ops * units / (kernel_ms * 1e-3) / 1e12 if ops else None


def timed_scan(ctx, lid, mode, rots, trans, out, steps, warmup, kernel):
    import torch
    from mmo_amd import ff
    for _ in range(warmup):
        ctx.scan(lid, mode, rots, trans, out=out)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(steps):
        ctx.scan(lid, mode, rots, trans, out=out)
        kernel_ms.append(ctx.last_kernel_ms(kernel))
        log("%s scan step %d kernel %.1f ms" % ("direct" if mode == ff.SHIFTED else "interp",
                                                k, kernel_ms[-1]))
    torch.cuda.synchronize()
    return time.perf_counter() - t0, kernel_ms


def bench_interp(ctx, lig, rots, trans, steps, warmup):
    """Scan through 10 per-type grids (64^3 at 0.375 A) computed on the GPU."""
    import torch
    from mmo_amd import ff, synth
    probe_a, probe_q, types = synth.grid_types([lig], 10)
    origin, dims, step = synth.grid_spec((0.0, 0.0, 0.0), 64, 0.375)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.compute_grids(probe_a, probe_q, step, dims, origin, copy_out=False)
    torch.cuda.synchronize()
    grid_s = time.perf_counter() - t0
    grid_ms = ctx.last_kernel_ms("grid")
    lid = ctx.add_ligand(lig, types=types[0])
    npose = len(rots) * len(trans)
    out = torch.empty(npose, dtype=torch.float64, device="cuda:0")
    elapsed, kms = timed_scan(ctx, lid, ff.INTERP, rots, trans, out, steps, warmup, "interp")
    e = out.cpu().numpy()
    inside = int(np.isfinite(e).sum())
    avg = float(np.mean(kms))
    lookups = inside * len(lig["xs"])
    ops = 43.0 * lookups
    return {"value": npose * steps / elapsed, "unit": "poses/s", "kernel_ms": avg,
            "grids": "10 x 64^3 FP32 at 0.375 A (%.1f MB)" % (10 * 64 ** 3 * 4 / 1e6),
            "grid_precompute_ms": grid_ms, "grid_precompute_wall_s": grid_s,
            "poses_inside_grid": inside, "poses_out_of_grid": npose - inside,
            "gather_GBs": 32.0 * lookups / (avg * 1e-3) / 1e9,
            "roofline": gather_roofline(32.0 * lookups, avg, "k_interp (scan)",
                                        traffic=measured_traffic("k_interp scan"),
                                        algorithmic_TFLOPs=ops / (avg * 1e-3) / 1e12,
                                        executed=executed_tflops("k_interp scan", npose, avg)),
            "note": "poses with an atom outside the grid return NaN and are counted"}


def bench_select(ctx, lig, rec, lid, rots, trans, out, k=100):
    """The exhaustive scan's clash filters (ROI sphere, centre and atom clashes against
    the protein vdW bitmask) and the top-k of the filtered energies, on the energies of
    the direct scan still in `out`."""
    import torch
    from mmo_amd import mol as molmod
    rxyz = np.stack([rec["xs"], rec["ys"], rec["zs"]], 1)
    step = 0.375
    origin = rxyz.min(axis=0) - 2.0
    dims = tuple(int(v) for v in np.ceil((rxyz.max(axis=0) + 2.0 - origin) / step) + 1)
    t0 = time.perf_counter()
    ctx.set_vdw_mask(step, dims, origin)
    mask_s = time.perf_counter() - t0
    radii = [molmod.VDW_RADII[int(a)] for a in lig["anum"]]
    roi = (np.zeros(3), 4.75 + 1e-9)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.scan_filter(lid, rots, trans, out, roi=roi, lig_radii=radii)
    torch.cuda.synchronize()
    filt_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    vals, idx = ctx.topk(out, k)
    torch.cuda.synchronize()
    topk_s = time.perf_counter() - t0
    kept = int(torch.isfinite(out).sum().item())
    best = int(idx[0].item())
    # filter before scoring (the reference's order): filters, compaction, then only the
    # kept poses scored; energies and top-k are those of the path above
    from mmo_amd import ff
    e2 = torch.empty_like(out)
    ctx.scan_filtered(lid, ff.SHIFTED, rots, trans, roi=roi, lig_radii=radii, out=e2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e2, nk = ctx.scan_filtered(lid, ff.SHIFTED, rots, trans, roi=roi, lig_radii=radii, out=e2)
    v2, i2 = ctx.topk(e2, k)
    torch.cuda.synchronize()
    fs_s = time.perf_counter() - t0
    same = bool(torch.equal(i2, idx)) and nk == kept
    return {"poses": out.numel(), "kept_after_filters": kept, "filter_s": filt_s,
            "topk_s": topk_s, "k": k, "best_e": float(vals[0].item()),
            "filtered_scan_s": fs_s, "filtered_scan_same_topk": same,
            "filtered_scan_note": "filters first, kept poses compacted and scored, then top-k "
                                  "(wall clock, rotations and translations from the host)",
            "best_rotation": best // len(trans), "best_translation": best % len(trans),
            "mask": "%d x %d x %d at %.3f A, built in %.2f s on the host" % (dims + (step, mask_s)),
            "filters": "ROI sphere r = 4.75 A, centre clash (AND), atom clash (OR)"}


def library_system(n_lig):
    from mmo_amd import ff, synth
    ligs = synth.library(n_lig, processes=min(16, cpu_threads()))
    _, rec = synth.config0()
    return ligs, rec


def bench_egrad(ctx, ligs, ids, rec, poses_per_lig=4, reps=20, cpu=True):
    """Energy + gradient in DOF space: 4 poses of each library ligand, batched vs one
    launch sequence per ligand; the pose kernel's roofline (8 FLOP per ligand-receptor
    pair, 19 + 18 per pair inside 12 A) and the oracle's FP64 E + dE/dDOF on the host."""
    import torch
    from mmo_amd import ff
    rng = np.random.default_rng(3)
    pose_lig = np.repeat(np.asarray(ids), poses_per_lig)
    flags = ff.F_INTER_SHIFTED | ff.F_INTRA | ff.F_GRAD
    lig_of = {l: ligs[k] for k, l in enumerate(ids)}

    def dofs_for(pl):
        d = []
        for l in pl:
            nrb = len(lig_of[l]["rb_left"])
            d.append(np.concatenate([rng.uniform(-1.5, 1.5, 3), rng.uniform(-math.pi, math.pi, 3),
                                     rng.uniform(-math.pi, math.pi, nrb)]))
        return d

    b = ctx.batch(pose_lig)
    d_host = dofs_for(pose_lig)
    d = torch.tensor(np.concatenate(d_host), device="cuda:0")
    e = torch.empty(b.npose, 4, dtype=torch.float64, device="cuda:0")
    g = torch.empty_like(d)
    xyz = torch.empty(b.natoms, 3, dtype=torch.float64, device="cuda:0")
    b.score_dofs_dev(d, e, xyz_out=xyz, gdof=g, flags=flags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.score_dofs_dev(d, e, gdof=g, flags=flags)
    torch.cuda.synchronize()
    tb = (time.perf_counter() - t0) / reps
    kms = []
    for _ in range(reps):
        b.score_dofs_dev(d, e, gdof=g, flags=flags)
        kms.append(ctx.last_kernel_ms("inter"))
    kernel_ms = float(np.mean(kms))
    # one launch sequence per ligand
    per = []
    for l in ids:
        bl = ctx.batch(np.full(poses_per_lig, l))
        dl = torch.tensor(np.concatenate(dofs_for([l] * poses_per_lig)), device="cuda:0")
        per.append((bl, dl, torch.empty(poses_per_lig, 4, dtype=torch.float64, device="cuda:0"),
                    torch.empty_like(dl)))
    for bl, dl, el, gl in per:
        bl.score_dofs_dev(dl, el, gdof=gl, flags=flags)
    torch.cuda.synchronize()
    preps = max(1, reps // 5)
    t0 = time.perf_counter()
    for _ in range(preps):
        for bl, dl, el, gl in per:
            bl.score_dofs_dev(dl, el, gdof=gl, flags=flags)
    torch.cuda.synchronize()
    tp = (time.perf_counter() - t0) / preps
    n = len(pose_lig)
    # algorithmic work of the batched poses
    x = xyz.cpu().numpy()
    off = np.asarray(b.atom_off)
    poses = [x[off[p]:off[p + 1]] for p in range(n)]
    n_in = in_cutoff_pairs(rec, poses)
    L = np.diff(off)
    flop = float(np.sum(8.0 * L * len(rec["xs"]) + (19.0 + 18.0) * n_in))
    out = {"batched": n / tb, "per_ligand": n / tp, "speedup": tp / tb, "unit": "poses/s",
           "poses": n, "ligands": len(ids), "ms_batched": tb * 1e3, "ms_per_ligand": tp * 1e3,
           "dtype": "fp32 pairs, fp64 per-pose sums and DOF chain",
           "roofline": valu_roofline(flop, kernel_ms, "k_pose_cull<grad> (per-atom sweep)",
                                     flop_per_pose=flop / n,
                                     in_cutoff_pairs_per_pose=float(n_in.mean()),
                                     executed=executed_tflops("k_pose_cull egrad", n, kernel_ms),
                                     note="the pose kernel of the batched step; the step also "
                                          "runs the DOF chain and the FP64 intra kernels")}
    if cpu:
        import oracle
        items = []
        for p in range(min(n, 512)):
            lig = lig_of[int(pose_lig[p])]
            base = np.stack([lig["xs"], lig["ys"], lig["zs"]], 1)
            items.append((lig, base, d_host[p]))

        def work(it):
            lig, base, dof = it
            oracle.dof_energy_grad(
                base, lig, dof,
                lambda xx: oracle.grad_inter(rec, dict(lig, xs=xx[:, 0], ys=xx[:, 1], zs=xx[:, 2])))
        out["cpu"] = cpu_pair(work, items, "poses/s",
                              "poses (FP64 shifted E_inter + E_intra and dE/dDOF, the oracle)")
    return out


def bench_min(ctx, ligs, ids, rec, poses_per_lig=4, iters=200):
    """200 L-BFGS iterations (one energy + gradient evaluation and one update each) for
    4 poses of each library ligand, started near the receptor inside a 4 A ROI box: one
    persistent launch (the default) and the launch sequence of the batched kernels."""
    import torch
    from mmo_amd import ff
    from mmo_amd.minimize import Minimizer, roi_bounds
    rng = np.random.default_rng(6)
    pose_lig = np.repeat(np.asarray(ids), poses_per_lig)
    lig_of = {l: ligs[k] for k, l in enumerate(ids)}
    nrb = [len(lig_of[l]["rb_left"]) for l in pose_lig]
    b = ctx.batch(pose_lig)
    center = np.array([30.0, 0.0, 0.0])
    lo, hi = roi_bounds(b, center, 4.0, nrb)
    d0 = np.concatenate([np.concatenate([center + rng.uniform(-2, 2, 3), rng.uniform(-3, 3, 3),
                                         rng.uniform(-0.2, 0.2, nrb[p])]) for p in range(len(pose_lig))])
    n = len(pose_lig)
    out = {"unit": "pose energy+gradient evaluations/s", "poses": n, "iterations": iters}
    for name, fused in (("sequence", 0), ("fused", 1)):
        ctx.tune(min_fused=fused)
        d = torch.tensor(d0, device="cuda:0")
        mz = Minimizer(b, flags=ff.F_INTER_SHIFTED | ff.F_INTRA, history=8, bounds=(lo, hi),
                       max_step=0.2)
        mz.run(d, 2)                                     # warm-up
        d.copy_(torch.tensor(d0, device="cuda:0"))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = mz.run(d, iters)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        info = mz.info()
        e = e.cpu().numpy()
        ok = np.isfinite(info["e_start"])
        evals = float(np.sum(info["evals"])) if "evals" in info else float(n * iters)
        r = {"pose_evals_per_s": n * iters / dt, "evaluations_run_per_s": evals / dt, "s": dt,
             "median_e_start": float(np.median(info["e_start"][ok])),
             "median_e_end": float(np.median(e[ok])),
             "failed_start": int((~ok).sum()), "converged": int((info["status"] == 1).sum()),
             "mean_accepted_steps": float(np.mean(info["accepted"]))}
        if fused:
            kms = ctx.last_kernel_ms("min")
            xyz = torch.empty(b.natoms, 3, dtype=torch.float64, device="cuda:0")
            ee = torch.empty(n, 4, dtype=torch.float64, device="cuda:0")
            b.score_dofs_dev(d, ee, xyz_out=xyz, flags=ff.F_INTER_SHIFTED | ff.F_INTRA)
            x = xyz.cpu().numpy()
            off = np.asarray(b.atom_off)
            n_in = in_cutoff_pairs(rec, [x[off[p]:off[p + 1]] for p in range(n)])
            flop_eval = float(np.mean(8.0 * np.diff(off) * len(rec["xs"]) + 37.0 * n_in))
            r["roofline"] = valu_roofline(
                flop_eval * evals, kms, "k_min_persist (one launch)",
                evaluations=evals, flop_per_evaluation=flop_eval,
                executed=executed_tflops("k_min_persist min", evals, kms),
                note="evaluations actually run (converged poses stop early); in-cutoff pairs "
                     "counted at the final poses")
        out[name] = r
    out["pose_evals_per_s"] = out["fused"]["pose_evals_per_s"]
    return out


def mc_chain_poses(m):
    """Best-pose coordinates of every chain ((n, 3) arrays)."""
    _, best = m.coords()
    return best


def mc_rooflines(m, rec, st, kernel_ms, mode, kernel, sample=2048, seed=9, counter_key=None):
    """Roofline of one MC run: direct sum, 8 FLOP per ligand-receptor pair + 19 per pair
    inside 12 A for each step's E_inter (in-cutoff pairs counted at the best poses of a
    sample of chains); grid form, 32 B of corner reads per ligand atom and step."""
    from mmo_amd import ff
    L = np.diff(m.chain_atom)
    frames = st["frame"].astype(np.float64)
    if mode == ff.INTERP:
        return gather_roofline(32.0 * float(np.sum(frames * L)), kernel_ms, kernel,
                               algorithmic_TFLOPs=43.0 * float(np.sum(frames * L)) /
                               (kernel_ms * 1e-3) / 1e12)
    rng = np.random.default_rng(seed)
    pick = rng.choice(m.nchain, min(sample, m.nchain), replace=False)
    best = mc_chain_poses(m)
    n_in = in_cutoff_pairs(rec, [best[k] for k in pick])
    per_step = 8.0 * L[pick] * len(rec["xs"]) + 19.0 * n_in
    flop = float(np.sum(frames[pick] * per_step)) * m.nchain / len(pick)
    return valu_roofline(flop, kernel_ms, kernel,
                         executed=executed_tflops(counter_key, float(frames.sum()), kernel_ms)
                         if counter_key else None,
                         flop_per_step=float(np.sum(frames[pick] * per_step) / np.sum(frames[pick])),
                         in_cutoff_pairs_per_step=float(n_in.mean()),
                         note="E_inter of every step (in-cutoff pairs at the best poses of %d "
                              "chains); the FP64 intra pairs of torsion steps are not counted"
                              % len(pick))


def cpu_mc(ligs_of_chain, best, rec, nsample=256, seed=8):
    """The CPU port's MC-step rate bound: one FP64 shifted direct E_inter per step (the
    oracle's loop), on chains' best poses; proposals, E_intra and Metropolis not counted."""
    import oracle
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(best), min(nsample, len(best)), replace=False)
    items = [dict(ligs_of_chain[k], xs=best[k][:, 0].copy(), ys=best[k][:, 1].copy(),
                  zs=best[k][:, 2].copy()) for k in pick]

    def work(lig):
        oracle.ene_inter(rec, lig, shifted=True)
    r = cpu_pair(work, items, "MC steps/s (E_inter per step, an upper bound)",
                 "E_inter evaluations (FP64 shifted direct sum, the oracle)")
    return r


def bench_mc(ctx, ligs, ids, tids, rec, chains_per_lig, frames, per_lig_sample, autotune, cpu=True):
    """MC steps/s: all chains of all ligands in one launch (hipGraph replays) vs one
    launch per ligand (the latter on the first `per_lig_sample` ligands); direct sum and
    grid interpolation (ligands `tids`, registered with grid types)."""
    import torch
    from mmo_amd import ff, tune
    from mmo_amd import mc as mcmod
    from mmo_amd import synth
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    roi = (np.zeros(3), 6.0)

    def make(lig_ids, seed=4, mode=ff.SHIFTED):
        rng = np.random.default_rng(seed)
        cl = np.repeat(np.asarray(lig_ids), chains_per_lig)
        rot0 = synth.random_rotations(rng, len(cl))
        pos0 = rng.uniform(-1.0, 1.0, (len(cl), 3))
        return mcmod.MonteCarlo(ctx, cl, rot0, pos0, mode, total_frames=frames, flags=flags,
                                roi_center=roi[0], roi_r=roi[1], seed=2024)

    key = tune.key_for("mc", ligands=len(ids), chains=chains_per_lig)
    if autotune:
        log("autotune MC residency")
        tuned = tune.tune_mc(ctx, lambda: make(ids[:256], seed=5), frames=50,
                             key_sig={"ligands": len(ids), "chains": chains_per_lig})["best"]
    else:
        tuned = tune.apply(ctx, key)
    m = make(ids)
    nrep = 4 if frames % 4 == 0 else 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.run_graph(frames // nrep, nreplay=nrep)
    torch.cuda.synchronize()
    tb = time.perf_counter() - t0
    kms_b = ctx.last_kernel_ms("mc")
    st = m.state()
    done = int((st["frame"] == frames).sum())
    steps_b = float(st["frame"].sum())
    roof_b = mc_rooflines(m, rec, st, kms_b, ff.SHIFTED, "k_mc direct (persistent, receptor in LDS)",
                          counter_key="k_mc direct")
    cpu_r = None
    if cpu:
        lig_of = {l: ligs[k] for k, l in enumerate(ids)}
        cpu_r = cpu_mc([lig_of[int(l)] for l in m.chain_lig], mc_chain_poses(m), rec)
    sub = ids[:per_lig_sample]
    m2 = make(sub)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m2.run_per_ligand(frames)
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    steps_p = float(m2.state()["frame"].sum())
    # the reference's default energy: trilinear interpolation in per-type grids
    mi = make(tids, mode=ff.INTERP)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mi.run_graph(frames // nrep, nreplay=nrep)
    torch.cuda.synchronize()
    ti = time.perf_counter() - t0
    kms_i = ctx.last_kernel_ms("mc")
    sti = mi.state()
    steps_i = float(sti["frame"].sum())
    out = {"batched": steps_b / tb, "per_ligand": steps_p / tp,
           "batched_interp": steps_i / ti, "steps": steps_b, "steps_interp": steps_i,
           "interp_grids": "10 x 100^3 FP32 at 0.375 A (%.0f MB)" % (10 * 100 ** 3 * 4 / 1e6),
           "speedup": (steps_b / tb) / (steps_p / tp), "unit": "MC steps/s",
           "chains": m.nchain, "frames": frames, "chains_finished": done,
           "too_long_chains": int((st["status"] == 2).sum()),
           "accept_rate": float(st["accepted"].sum() / max(steps_b, 1.0)),
           "s_batched": tb, "s_per_ligand": tp,
           "per_ligand_sample": "%d ligands x %d chains" % (len(sub), chains_per_lig),
           "roofline": roof_b,
           "roofline_interp": mc_rooflines(mi, rec, sti, kms_i, ff.INTERP, "k_mc grid form"),
           "tuned": tuned, "mc_waves_per_cu": ctx.tuning("mc_waves_per_cu"),
           "mc_wg_waves": ctx.tuning("mc_wg_waves"),
           "launch": "one persistent launch per hipGraph replay, %d replays of %d frames"
                     % (nrep, frames // nrep)}
    if cpu_r:
        out["cpu"] = cpu_r
    return out


def bench_mc_min(ctx, ligs, ids, tids, rec, chains_per_lig=4, frames=10, iters=200):
    """Monte Carlo-plus-minimisation: 4 chains of each library ligand (4,096 chains), every
    proposal followed by up to 200 L-BFGS evaluations in DOF space inside the same
    persistent launch; direct sum (receptor in LDS) and grid interpolation."""
    import torch
    from mmo_amd import ff, synth
    from mmo_amd import mc as mcmod
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI
    roi = (np.zeros(3), 6.0)
    out = {"unit": "MC+minimisation steps/s", "chains": len(ids) * chains_per_lig,
           "frames": frames, "min_iterations": iters}
    for name, lig_ids, mode in (("direct", ids, ff.SHIFTED), ("interp", tids, ff.INTERP)):
        rng = np.random.default_rng(14)
        cl = np.repeat(np.asarray(lig_ids), chains_per_lig)
        m = mcmod.MonteCarlo(ctx, cl, synth.random_rotations(rng, len(cl)),
                             rng.uniform(-1.0, 1.0, (len(cl), 3)), mode, total_frames=frames,
                             flags=flags, roi_center=roi[0], roi_r=roi[1], seed=77)
        m.set_minimizer(iters, history=8, max_step=0.2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.run_graph(frames, nreplay=1)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = m.state()
        _, evals = m.min_evals()
        steps = float(st["frame"].sum())
        out[name] = {"steps_per_s": steps / dt, "evaluations_per_s": evals / dt,
                     "steps": steps, "evaluations": evals,
                     "evaluations_per_step": evals / max(steps, 1.0), "s": dt,
                     "kernel_ms": ctx.last_kernel_ms("mc"),
                     "accept_rate": float(st["accepted"].sum() / max(steps, 1.0)),
                     "median_best_e": float(np.median(st["best_e"]))}
    return out


def bench_config4(ligs, c4_frames=500, c4_direct_frames=50, c4_min_frames=2, chains=32,
                  min_iters=200):
    """At scale: the 15,000-atom receptor in a 60 A box, 10 per-type grids of 128^3 at
    0.375 A computed on the GPU, 4,096 library ligands x 32 chains (131,072 chains):
    plain Monte Carlo in the grid form (one persistent launch replayed from a hipGraph)
    and in the direct form (receptor read from global memory); Monte Carlo-plus-
    minimisation (up to 200 L-BFGS evaluations per step) in the grid form on a bounded
    number of frames.  (The direct-form gradient sweeps on this receptor are not run at
    this scale: the minimising chains of that form faulted once here.)"""
    import torch
    from mmo_amd import ff, synth
    from mmo_amd import mc as mcmod
    t0 = time.perf_counter()
    rec = synth.config4_receptor()
    setup_s = time.perf_counter() - t0
    ctx = ff.Context(0)
    ctx.set_receptor(rec)
    probe_a, probe_q, types = synth.grid_types(ligs, 10)
    origin, dims, step = synth.grid_spec((0.0, 0.0, 0.0), 128, 0.375)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.compute_grids(probe_a, probe_q, step, dims, origin, copy_out=False)
    torch.cuda.synchronize()
    grid_s = time.perf_counter() - t0
    tids = [ctx.add_ligand(l, types=t) for l, t in zip(ligs, types)]
    cl = np.repeat(np.asarray(tids), chains)
    roi = (np.zeros(3), 10.0)
    flags = mcmod.TWEAK_RBONDS | mcmod.INTRA_UFF | mcmod.HARD_ROI

    def make(mode, frames, seed=44):
        rng = np.random.default_rng(seed)
        return mcmod.MonteCarlo(ctx, cl, synth.random_rotations(rng, len(cl)),
                                rng.uniform(-8.0, 8.0, (len(cl), 3)), mode, total_frames=frames,
                                flags=flags, roi_center=roi[0], roi_r=roi[1], seed=4)

    def run(m, frames, nrep):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.run_graph(frames // nrep, nreplay=nrep)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, ctx.last_kernel_ms("mc"), m.state()

    out = {"receptor_atoms": len(rec["xs"]), "receptor": "60 A box, min separation 2.0 A",
           "ligands": len(ligs), "chains": len(cl),
           "grids": "10 x 128^3 FP32 at 0.375 A (%.0f MB)" % (10 * 128 ** 3 * 4 / 1e6),
           "grid_precompute_s": grid_s, "host_receptor_s": setup_s}
    log("config4: MC grid form, %d frames" % c4_frames)
    m = make(ff.INTERP, c4_frames)
    dt, kms, st = run(m, c4_frames, 4 if c4_frames % 4 == 0 else 1)
    steps = float(st["frame"].sum())
    out["mc_interp"] = {"steps_per_s": steps / dt, "frames": c4_frames, "s": dt,
                        "roofline": mc_rooflines(m, rec, st, kms, ff.INTERP, "k_mc grid form")}
    del m
    log("config4: MC direct form, %d frames" % c4_direct_frames)
    m = make(ff.SHIFTED, c4_direct_frames)
    dt, kms, st = run(m, c4_direct_frames, 1)
    steps = float(st["frame"].sum())
    out["mc_direct"] = {"steps_per_s": steps / dt, "frames": c4_direct_frames, "s": dt,
                        "roofline": mc_rooflines(m, rec, st, kms, ff.SHIFTED,
                                                 "k_mc direct (receptor in global memory)")}
    del m
    for name, mode in (("mc_min_interp", ff.INTERP),):
        log("config4: MC+minimisation %s, %d frames" % (name, c4_min_frames))
        m = make(mode, c4_min_frames)
        m.set_minimizer(min_iters, history=8, max_step=0.2)
        dt, kms, st = run(m, c4_min_frames, 1)
        _, evals = m.min_evals()
        steps = float(st["frame"].sum())
        out[name] = {"steps_per_s": steps / dt, "evaluations_per_s": evals / dt,
                     "evaluations_per_step": evals / max(steps, 1.0), "frames": c4_min_frames,
                     "min_iterations": min_iters, "s": dt, "kernel_ms": kms}
        del m
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rotations", type=int, default=N_ROT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--only-scan", action="store_true", help="headline scan only (no extras)")
    ap.add_argument("--extras", default="select,interp,egrad,min,mc,mc_min,config4",
                    help="comma list of secondary figures to run: select, interp, egrad, min, mc,"
                         " mc_min, config4 (the 15,000-atom workload)")
    ap.add_argument("--ligands", type=int, default=1024)
    ap.add_argument("--mc-chains", type=int, default=64)
    ap.add_argument("--mc-frames", type=int, default=1000)
    ap.add_argument("--mc-per-ligand-sample", type=int, default=32)
    ap.add_argument("--mc-min-frames", type=int, default=10,
                    help="frames of the Monte Carlo-plus-minimisation figure")
    ap.add_argument("--c4-ligands", type=int, default=4096)
    ap.add_argument("--c4-frames", type=int, default=500, help="config4 grid-form MC frames")
    ap.add_argument("--c4-direct-frames", type=int, default=50,
                    help="config4 direct-form MC frames")
    ap.add_argument("--c4-min-frames", type=int, default=2,
                    help="config4 MC-plus-minimisation frames")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one device copy of this size first (counter scale reference)")
    ap.add_argument("--tune", default="", help="kernel shape, e.g. cull_k=1,cull_waves=16")
    ap.add_argument("--autotune", action="store_true",
                    help="sweep kernel shapes first (untimed) and record them in mmo_amd/tuned.json;"
                         " without it the recorded shapes are applied")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the scoring path is built and measured on a single "
              "MI355X; multi-GPU runs are not part of this build (asked for %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mmo_amd import ff, synth

    lig, rec = synth.config0()
    rots = synth.so3_rotations(args.rotations)
    trans = synth.lattice((0.0, 0.0, 0.0), N_TRANS_SIDE, TRANS_STEP)
    npose = len(rots) * len(trans)
    from mmo_amd import tune as tuner
    ctx = ff.Context(0)
    scan_key = tuner.key_for("scan", lig_atoms=len(lig["xs"]), rec_atoms=len(rec["xs"]))
    tuned = None if args.autotune else tuner.apply(ctx, scan_key)
    if tuned is None and not args.autotune:
        log("no recorded tuning for %s: library defaults" % scan_key)
    tune = dict(kv.split("=") for kv in args.tune.split(",") if kv)
    for key in ("cell_size", "cell_xsub", "scan_cell", "scan_xsub"):   # grids built by set_receptor
        if key in tune:
            ctx.tune(**{key: float(tune.pop(key))})
    ctx.set_receptor(rec)
    ctx.tune(**{k: float(v) for k, v in tune.items()})
    lid = ctx.add_ligand(lig)
    if args.autotune:
        log("autotune scan shapes on %d rotations" % min(256, len(rots)))
        tuned = tuner.tune_scan(ctx, rec, lid, rots[:256], trans)["best"]
    out = torch.empty(npose, dtype=torch.float64, device="cuda:0")
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: %d MiB read + %d MiB written" % (args.copy_probe_mib, args.copy_probe_mib))
        del src, dst
    log("%d poses per step, %d warmup, %d timed" % (npose, args.warmup, args.steps))
    elapsed, kernel_ms = timed_scan(ctx, lid, ff.SHIFTED, rots, trans, out, args.steps,
                                    args.warmup, "inter")
    e = out.cpu().numpy()
    assert np.isfinite(e).all()

    flops_pose, mean_in = algorithmic_flops_per_pose(lig, rec, rots, trans)
    avg_ms = float(np.mean(kernel_ms))
    achieved = flops_pose * npose / (avg_ms * 1e-3) / 1e12
    sc = ctx.tuning("scan_cull")
    kernel = ("k_scan_rows" if sc >= 5 else
              "k_inter_cull" if sc > 0 and ctx.tuning("cull_ok") > 0 else "k_inter")
    line = {
        "metric": "poses/sec (energy, energy+grad) on 1 MI355X at <=1e-5 rel err vs FP64; MC steps/s",
        "value": npose * args.steps / elapsed,
        "unit": "poses/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32", "accumulate": "fp32 chunks folded into fp64 per-pose sums",
        "data": "synthetic",
        "config": {"workload": "exhaustive rigid scan, direct shifted LJ+Coulomb",
                   "ligand_atoms": len(lig["xs"]), "receptor_atoms": len(rec["xs"]),
                   "rotations": len(rots), "translations": len(trans), "poses": npose,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "fp32-valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": measured_traffic(kernel + " scan"),
                     "executed": executed_tflops(kernel + " scan", npose, avg_ms),
                     "kernel": kernel + " (scan)", "kernel_ms": avg_ms,
                     "flop_per_pose": flops_pose, "in_cutoff_pairs_per_pose": mean_in,
                     "note": "achieved prices every (ligand, receptor) pair of a pose at 8 FLOP "
                             "plus 19 per pair within 12 A (the brute-force sum's work); the "
                             "culled kernels visit a fraction of the pairs, so achieved can "
                             "exceed the FP32 peak; executed is the FP32 rate the SQ counters "
                             "measured on this kernel"},
        "tuning": {"key": scan_key, "applied": tuned, "cell_size": ctx.tuning("cell_size"),
                   "scan_cull": ctx.tuning("scan_cull"),
                   "cull_k": ctx.tuning("cull_k"), "cull_waves": ctx.tuning("cull_waves"),
                   "scan_cell": ctx.tuning("scan_cell"), "scan_xsub": ctx.tuning("scan_xsub"),
                   "scan_chunk": ctx.tuning("scan_chunk"),
                   "row_filter": ctx.tuning("row_filter")},
    }
    want = set(x for x in args.extras.split(",") if x) if not args.only_scan else set()
    if want:
        extra = {}
        if "select" in want:
            log("scan filters and top-k")
            extra["scan_select"] = bench_select(ctx, lig, rec, lid, rots, trans, out)
        if "interp" in want:
            log("interp scan")
            extra["scan_interp"] = bench_interp(ctx, lig, rots, trans, args.steps, args.warmup)
        big = None
        if want & {"egrad", "min", "mc", "mc_min", "config4"}:
            # ligand k of a library depends only on k: the first 1,024 of the 4,096-ligand
            # library are the 1,024-ligand library
            n_lib = max(args.ligands, args.c4_ligands if "config4" in want else 0)
            log("library of %d ligands" % n_lib)
            big, _ = library_system(n_lib)
            ligs = big[:args.ligands]
            ids = [ctx.add_ligand(l) for l in ligs]
        if "egrad" in want:
            log("energy+gradient, batched vs per ligand")
            extra["egrad"] = bench_egrad(ctx, ligs, ids, rec, cpu=not args.no_cpu)
        if "min" in want:
            log("L-BFGS minimisation")
            extra["min"] = bench_min(ctx, ligs, ids, rec)
        if want & {"mc", "mc_min"}:
            # the reference's default energy: trilinear interpolation in per-type grids
            probe_a, probe_q, types = synth.grid_types(ligs, 10)
            origin, dims, gstep = synth.grid_spec((0.0, 0.0, 0.0), 100, 0.375)
            ctx.compute_grids(probe_a, probe_q, gstep, dims, origin, copy_out=False)
            tids = [ctx.add_ligand(l, types=t) for l, t in zip(ligs, types)]
        if "mc" in want:
            log("Monte Carlo, batched vs per ligand")
            extra["mc"] = bench_mc(ctx, ligs, ids, tids, rec, args.mc_chains, args.mc_frames,
                                   args.mc_per_ligand_sample, args.autotune, cpu=not args.no_cpu)
        if "mc_min" in want:
            log("Monte Carlo-plus-minimisation")
            extra["mc_min"] = bench_mc_min(ctx, ligs, ids, tids, rec, frames=args.mc_min_frames)
        if "config4" in want:
            log("at scale: 15,000-atom receptor, 128^3 grids, %d chains" % (32 * args.c4_ligands))
            extra["config4"] = bench_config4(big[:args.c4_ligands], c4_frames=args.c4_frames,
                                             c4_direct_frames=args.c4_direct_frames,
                                             c4_min_frames=args.c4_min_frames)
        line["extra"] = extra
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lig, rec, rots, trans)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
