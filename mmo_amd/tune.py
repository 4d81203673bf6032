"""Autotuner for the kernel shapes of the scoring library.

The library exposes its launch shapes as context settings (`Context.tune`):

  scan / culled batches   cull_k (ligand atoms per lane: 1, 2, 4), cull_waves
                          (wavefronts per workgroup, 1-16), cell_size (receptor cell
                          edge in A; takes effect at the next set_receptor), scan_cull
                          (1 receptor in LDS, 2 read through the scalar cache, 3 the
                          same on the cells cut along x, 4 positions in LDS, 5 row
                          lists of the scan grid, 6 row lists with the receptor in
                          LDS), scan_cell / scan_xsub (the row lists' cell edge in y
                          and z and its x cut; next set_receptor), scan_chunk
                          (receptor atoms fetched at once by the row lists: 4 or 8),
                          cell_xsub (x cut of the pose and Monte Carlo sweeps' cells;
                          next set_receptor)
  plain batches           inter_k (1, 2, 4, 8), inter_waves (1-4), inter_rtile (receptor
                          atoms per LDS tile)
  batches (pose culled)   pose_waves (wavefronts per workgroup, 1-16)
  Monte Carlo             mc_wg_waves (chains per workgroup, 1-16), mc_waves_per_cu
                          (resident chains per CU, 1-32)

Each sweep times the real kernel on the caller's workload (HIP-event kernel times read
back through `Context.last_kernel_ms`, or wall time around a synchronised launch for
the MC), checks every candidate's energies against the current setting's, keeps the
fastest, and records the choice in a JSON cache keyed by device and workload, so a
later run applies it without sweeping (`apply`).
"""
import itertools
import json
import os
import time

import numpy as np

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned.json")


def _device_name():
    """Architecture and CU count (the marketing name varies between driver setups)."""
    import torch
    if not torch.cuda.is_available():
        return "cpu"
    p = torch.cuda.get_device_properties(0)
    arch = (getattr(p, "gcnArchName", "") or p.name).split(":")[0]
    return "%s-%dcu" % (arch, p.multi_processor_count)


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, entry):
    data = _load()
    data[key] = entry
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def key_for(kind, **sig):
    return "%s|%s|%s" % (_device_name(), kind, ",".join("%s=%s" % kv for kv in sorted(sig.items())))


def apply(ctx, key):
    """Applies a cached choice; returns it (or None when the workload was never tuned).
    cell_size is returned but takes effect only at the next set_receptor."""
    entry = _load().get(key)
    if not entry:
        return None
    ctx.tune(**{k: v for k, v in entry["best"].items()})
    return entry["best"]


def _time(fn, reps):
    import torch
    fn()
    torch.cuda.synchronize()
    t = []
    for _ in range(reps):
        t.append(fn())
    torch.cuda.synchronize()
    return float(np.median(t))


def tune_scan(ctx, rec, lig_id, rots, trans, cell_sizes=(4.0, 5.0), ks=(1,),
              waves=(4, 8, 16), scan_culls=(1, 2, 3),
              row_grids=((3.5, 8), (4.0, 8), (5.0, 8), (6.0, 8), (5.0, 4)), row_chunks=(4, 8),
              row_waves=(4, 8), row_filters=(1,), reps=2, store=True):
    """Culled rigid scan: sweeps cell size x receptor path (LDS or scalar cache) x atoms
    per lane x wavefronts per workgroup of the cell-sweep kernel, and row-list grid (cell
    edge in y and z, x cut) x receptor atoms fetched at once x wavefronts per workgroup of
    the row-list kernel (scan_cull 5) x its far-row atom test (row_filter), on a slice of
    the caller's scan (rots, trans), and leaves the best setting applied.  The pose sweeps' x cut (cell_xsub) stays as set: it
    is the Monte Carlo's choice."""
    from . import ff
    import torch
    n = len(rots) * len(trans)
    out = torch.empty(n, dtype=torch.float64, device="cuda:%d" % ctx.device)
    ctx.scan(lig_id, ff.SHIFTED, rots, trans, out=out)
    ref = out.clone()
    results = []

    def measure(setting):
        def run():
            ctx.scan(lig_id, ff.SHIFTED, rots, trans, out=out)
            return ctx.last_kernel_ms("inter")
        ms = _time(run, reps)
        ok = bool(torch.allclose(out, ref, rtol=2e-6, atol=1e-6))
        results.append(dict(setting, ms=ms, ok=ok))

    keep_cell, keep_scell, keep_xsub = (ctx.tuning("cell_size"), ctx.tuning("scan_cell"),
                                        ctx.tuning("scan_xsub"))
    for cs in cell_sizes:
        ctx.tune(cell_size=cs)
        ctx.set_receptor(rec)
        for sc, k, w in itertools.product(scan_culls, ks, waves):
            ctx.tune(scan_cull=sc, cull_k=k, cull_waves=w)
            measure({"cell_size": cs, "scan_cull": sc, "cull_k": k, "cull_waves": w})
    ctx.tune(cell_size=keep_cell)
    for (cell, xsub) in row_grids:
        ctx.tune(scan_cell=cell, scan_xsub=xsub)
        ctx.set_receptor(rec)
        for ch, w, f in itertools.product(row_chunks, row_waves, row_filters):
            ctx.tune(scan_cull=5, scan_chunk=ch, cull_waves=w, cull_k=1, row_filter=f)
            measure({"cell_size": keep_cell, "scan_cull": 5, "cull_k": 1, "cull_waves": w,
                     "scan_cell": cell, "scan_xsub": xsub, "scan_chunk": ch, "row_filter": f})
    good = [r for r in results if r["ok"]]
    best = min(good, key=lambda r: r["ms"])
    chosen = {k: v for k, v in best.items() if k not in ("ms", "ok")}
    chosen.setdefault("scan_cell", keep_scell)
    chosen.setdefault("scan_xsub", keep_xsub)
    ctx.tune(cell_size=chosen["cell_size"], scan_cell=chosen["scan_cell"],
             scan_xsub=chosen["scan_xsub"])
    ctx.set_receptor(rec)
    ctx.tune(**{k: v for k, v in chosen.items()
                if k not in ("cell_size", "scan_cell", "scan_xsub")})
    entry = {"best": chosen, "ms": best["ms"], "poses": n, "sweep": results}
    if store:
        _store(key_for("scan", lig_atoms=ctx.ligand_natoms(lig_id), rec_atoms=len(rec["xs"])), entry)
    return entry


def tune_batch(ctx, batch, xyz, flags, ks=(1, 2, 4), waves=(1, 2, 4), rtiles=(1024, 2048, 4096),
               reps=3, store=True):
    """Plain (unculled) batched scoring: atoms per lane x wavefronts x receptor tile."""
    import torch
    e = torch.empty(batch.npose, 4, dtype=torch.float64, device=xyz.device)
    batch.score_dev(xyz, e, flags=flags)
    ref = e.clone()
    results = []
    for k, w, rt in itertools.product(ks, waves, rtiles):
        ctx.tune(inter_k=k, inter_waves=w, inter_rtile=rt)

        def run():
            batch.score_dev(xyz, e, flags=flags)
            return ctx.last_kernel_ms("inter")
        ms = _time(run, reps)
        ok = bool(torch.allclose(e, ref, rtol=2e-6, atol=1e-6))
        results.append({"inter_k": k, "inter_waves": w, "inter_rtile": rt, "ms": ms, "ok": ok})
    best = min((r for r in results if r["ok"]), key=lambda r: r["ms"])
    ctx.tune(inter_k=best["inter_k"], inter_waves=best["inter_waves"], inter_rtile=best["inter_rtile"])
    entry = {"best": {k: best[k] for k in ("inter_k", "inter_waves", "inter_rtile")},
             "ms": best["ms"], "poses": batch.npose, "sweep": results}
    if store:
        _store(key_for("batch", poses=batch.npose, atoms=batch.natoms, flags=flags), entry)
    return entry


def tune_mc(ctx, make_mc, frames=50, shapes=((8, 8), (8, 16), (12, 12), (16, 16), (16, 32)),
            store=True, key_sig=None):
    """Monte Carlo: chains (wavefronts) per workgroup x resident chains per CU.  make_mc()
    builds a fresh MonteCarlo of the caller's workload; each candidate advances it by
    `frames` frames (same results for every setting: chains depend only on their own
    random streams)."""
    import torch
    results = []
    ref = None
    for wg, w in shapes:
        ctx.tune(mc_wg_waves=wg, mc_waves_per_cu=w)
        m = make_mc()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.run(frames)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        st = m.state()
        if ref is None:
            ref = st["e"].copy()
        ok = bool(np.array_equal(st["e"], ref))
        results.append({"mc_wg_waves": wg, "mc_waves_per_cu": w, "s": dt, "ok": ok,
                        "steps_per_s": float(st["frame"].sum()) / dt})
        del m
    best = min((r for r in results if r["ok"]), key=lambda r: r["s"])
    ctx.tune(mc_wg_waves=best["mc_wg_waves"], mc_waves_per_cu=best["mc_waves_per_cu"])
    entry = {"best": {"mc_wg_waves": best["mc_wg_waves"], "mc_waves_per_cu": best["mc_waves_per_cu"]},
             "s": best["s"], "sweep": results}
    if store:
        _store(key_for("mc", **(key_sig or {})), entry)
    return entry
