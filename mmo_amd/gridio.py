"""Per-type grid files in the reference's on-disk format, and its atom-type table.

Reference layout (src/G3D.ml:14-64, src/lds.ml:471-555, src/mol.ml:280-304):

- one grid per ligand atom type t, stored raw as FP32 in native byte order, x fastest
  (index i + j*x_dim + k*x_dim*y_dim, src/grid.ml:98-99), in "<lig_fn>.t<t>.ba1";
- next to it "<lig_fn>.t<t>.ba1.dims", four text lines "step: %g", "x_dim: %d",
  "y_dim: %d", "z_dim: %d";
- the grid's lowest corner is (0, 0, 0) (src/grid.ml:9);
- atom types are the distinct (anum, partial charge) pairs in order of first
  appearance over the ligands (Mol.get_type2id_ht), listed in "atom_types.txt" as
  "%d (%d, %g)" lines (src/lds.ml:1600-1606); --less-charges first rounds charges to
  0.01 (Utls.reduce_precision, src/utls.ml:129-132).

- unless run with --no-compress, the reference keeps each grid file compressed with the
  zstd tool as "<file>.ba1.zst" (the raw file removed, its .dims file kept) and reads
  that form when present (src/lds.ml:471-555, src/utls.ml:22-45); `write_ba1(...,
  compress=True)` writes the same zstd frame and `read_ba1` / `load_grids` read either.

Grids written here load in the reference and the other way round; `load_grids` returns
the (ntypes, z, y, x) FP32 array that `Context.set_grids` takes.
"""
import math
import os

import numpy as np


def reduce_precision(x):
    """Utls.reduce_precision: round half away from zero to 0.01 (via truncation)."""
    return float(math.trunc(x * 100.0 + (0.5 if x >= 0.0 else -0.5))) / 100.0


def type_table(ligs, less_charges=False):
    """Distinct (anum, q) pairs in order of first appearance and, per ligand, the type
    id of every atom (Mol.get_type2id_ht + Mol.assign_FF_types)."""
    ids, types, per_lig = {}, [], []
    for lig in ligs:
        q = [reduce_precision(float(v)) if less_charges else float(v) for v in lig["q"]]
        t = np.empty(len(q), dtype=np.int32)
        for i, (a, c) in enumerate(zip(lig["anum"], q)):
            key = (int(a), c)
            if key not in ids:
                ids[key] = len(types)
                types.append(key)
            t[i] = ids[key]
        per_lig.append(t)
    return types, per_lig


def write_atom_types(fn, types):
    with open(fn, "w") as f:
        for i, (anum, q) in enumerate(types):
            f.write("%d (%d, %g)\n" % (i, anum, q))


def read_atom_types(fn):
    out = []
    with open(fn) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            i, rest = line.split(" ", 1)
            a, q = rest.strip("()").split(",")
            assert int(i) == len(out), "atom types out of order in %s" % fn
            out.append((int(a), float(q)))
    return out


def grid_file(lig_fn, t):
    return "%s.t%d.ba1" % (lig_fn, t)


def write_dims(fn, step, dims):
    nx, ny, nz = (int(v) for v in dims)
    with open(fn, "w") as f:
        f.write("step: %g\nx_dim: %d\ny_dim: %d\nz_dim: %d\n" % (step, nx, ny, nz))


def read_dims(fn):
    vals = {}
    with open(fn) as f:
        for key in ("step", "x_dim", "y_dim", "z_dim"):
            k, v = f.readline().split(":")
            if k.strip() != key:
                raise ValueError("%s: expected '%s:' line" % (fn, key))
            vals[key] = float(v) if key == "step" else int(v)
    return vals["step"], (vals["x_dim"], vals["y_dim"], vals["z_dim"])


def write_ba1(fn, grid, step, compress=False):
    """grid: (z, y, x) array; written as FP32 with its .dims file.  compress: replace the
    raw file by "<fn>.zst" (as `zstd --rm` does); returns the data file's name."""
    g = np.ascontiguousarray(grid, dtype=np.float32)
    nz, ny, nx = g.shape
    write_dims(fn + ".dims", step, (nx, ny, nz))
    if not compress:
        g.tofile(fn)
        return fn
    from . import zstd
    with open(fn + ".zst", "wb") as f:
        f.write(zstd.compress(g.tobytes()))
    if os.path.exists(fn):
        os.remove(fn)
    return fn + ".zst"


def read_ba1(fn, dims=None):
    """Returns (grid (z, y, x) FP32, step) of "<x>.ba1" or "<x>.ba1.zst"; dims from the
    raw file's .dims file unless given."""
    raw = fn[:-4] if fn.endswith(".zst") else fn
    step = None
    if dims is None:
        step, dims = read_dims(raw + ".dims")
    nx, ny, nz = dims
    if fn.endswith(".zst"):
        from . import zstd
        with open(fn, "rb") as f:
            a = np.frombuffer(zstd.decompress(f.read(), size=4 * nx * ny * nz), dtype=np.float32)
    else:
        a = np.fromfile(fn, dtype=np.float32)
    if a.size != nx * ny * nz:
        raise ValueError("%s holds %d floats, the dims give %d" % (fn, a.size, nx * ny * nz))
    return a.reshape(nz, ny, nx), step


def save_grids(lig_fn, grids, step, compress=False):
    """grids: (ntypes, z, y, x); writes <lig_fn>.t<t>.ba1 (+ .dims) for every type, or
    <lig_fn>.t<t>.ba1.zst with compress."""
    for t in range(len(grids)):
        write_ba1(grid_file(lig_fn, t), grids[t], step, compress=compress)


def load_grids(lig_fn, ntypes):
    """Reads the ntypes grids of <lig_fn> (each from its .ba1 file, or from .ba1.zst when
    only that exists); all must share step and dimensions."""
    out, step0, dims0 = [], None, None
    for t in range(ntypes):
        raw = grid_file(lig_fn, t)
        fn = raw if os.path.exists(raw) or not os.path.exists(raw + ".zst") else raw + ".zst"
        step, dims = read_dims(raw + ".dims")
        if step0 is None:
            step0, dims0 = step, dims
        elif step != step0 or dims != dims0:
            raise ValueError("%s: step/dims differ from type 0" % fn)
        g, _ = read_ba1(fn, dims)
        out.append(g)
    return np.stack(out), step0
