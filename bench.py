"""Benchmark: exhaustive rigid-pose scan with the plain (direct-sum) method on one GPU.

Workload: the 75-atom synthetic ligand (40 heavy + 35 H, 6 rotatable bonds) and the
3000-atom Poisson-disc receptor; 4,608 super-Fibonacci SO(3) rotations x 20^3 lattice
translations at 0.5 A = 36,864,000 poses.  One step scores every pose with the shifted
LJ 12-6 + Coulomb direct sum (Mol.ene_inter_UFF_shifted_brute semantics).  Inputs are
resident on the GPU before timing starts; outputs stay on the GPU.

Prints one JSON line (see the "roofline" and "cpu_baseline" objects for the kernel's
fraction of the FP32 VALU peak and the FP64 CPU restatement timed on this host).
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
N_ROT, N_TRANS_SIDE, TRANS_STEP = 4608, 20, 0.5


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


def cpu_baseline(lig, rec, rots, trans, target_s=12.0):
    """FP64 C restatement of the reference loop (oracle, OpenMP over poses) on a random
    sample of the same scan's poses."""
    import oracle
    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = max(1, min(threads, 16))
    rng = np.random.default_rng(1)
    n = 64
    while True:
        pr = rng.integers(0, len(rots), n).astype(np.int32)
        pt = rng.integers(0, len(trans), n).astype(np.int32)
        t0 = time.perf_counter()
        oracle.scan_shifted(rec, lig, rots, pr, trans, pt, nthreads=threads)
        dt = time.perf_counter() - t0
        if dt >= target_s or n >= 1 << 20:
            break
        n = int(n * min(16.0, max(2.0, 1.1 * target_s / max(dt, 1e-3))))
    return {"value": n / dt, "unit": "poses/s", "cores": threads, "kind": "port",
            "sample": "%d random poses of the scan (shifted direct E_inter, FP64, OpenMP), %.1f s"
                      % (n, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--rotations", type=int, default=N_ROT)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--tune", default="", help="kernel shape, e.g. cull_k=1,cull_waves=16")
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
    ctx = ff.Context(0)
    tune = dict(kv.split("=") for kv in args.tune.split(",") if kv)
    if "cell_size" in tune:
        ctx.tune(cell_size=float(tune.pop("cell_size")))
    ctx.set_receptor(rec)
    ctx.tune(**{k: float(v) for k, v in tune.items()})
    lid = ctx.add_ligand(lig)
    out = torch.empty(npose, dtype=torch.float64, device="cuda:0")
    print("bench: %d poses per step, %d warmup, %d timed" % (npose, args.warmup, args.steps),
          file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        ctx.scan(lid, ff.SHIFTED, rots, trans, out=out)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        ctx.scan(lid, ff.SHIFTED, rots, trans, out=out)
        kernel_ms.append(ctx.last_kernel_ms("inter"))
        print("bench: step %d kernel %.1f ms" % (k, kernel_ms[-1]), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    e = out.cpu().numpy()
    assert np.isfinite(e).all()

    flops_pose, mean_in = algorithmic_flops_per_pose(lig, rec, rots, trans)
    avg_ms = float(np.mean(kernel_ms))
    achieved = flops_pose * npose / (avg_ms * 1e-3) / 1e12
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
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_inter_cull (scan)" if ctx.tuning("scan_cull") > 0 and
                     ctx.tuning("cull_ok") > 0 else "k_inter (scan)", "kernel_ms": avg_ms,
                     "flop_per_pose": flops_pose, "in_cutoff_pairs_per_pose": mean_in},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(lig, rec, rots, trans)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
