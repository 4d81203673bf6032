"""Seeded synthetic workloads for the scoring path (ligands, receptor, pose sets).

The benchmark configurations are described in words only (a random-tree ligand with
UFF types {C_3,C_R,N_3,N_R,O_2,O_3,S_3,H_}, a Poisson-disc receptor around a cavity,
SO(3) x lattice rigid scans, ligand libraries with lognormal sizes), so this module
builds them deterministically from numpy's PCG64 with fixed seeds.

Ligand records mirror ``Mol.t`` (src/mol.ml:17-35) as plain numpy arrays:
``xs, ys, zs, q, anum`` (SoA), ``dists`` (n x n topological distances,
src/pqrs.ml:62-77), rotatable bonds ``rb_left, rb_right`` with the axis pointing
from the fixed atom to the moving one and rotation groups ``rg_off/rg_atoms``
that exclude the ``right`` atom (src/mol.ml:385-402, src/pqrs.ml:80-87), as produced by mol2pqrs
(smaller side of the cut, src/mol_graph.ml:157-192).

Choices the configuration text leaves open, recorded here:
  * UFF types map to elements for LJ (C_3,C_R->C, N_3,N_R->N, O_2,O_3->O, S_3->S,
    H_->H): the reference looks parameters up by atomic number only
    (src/UFF.ml:39-40, src/mol.ml:810-812).
  * Conformers are grown depth-first with backtracking and rejected with the form of
    the reference's clash rule d^2 < (a (r_vdW,i + r_vdW,j))^2 (src/ptable.ml:62-80)
    on pairs four or more bonds apart, with a = 0.65 (0.5 for pairs with a hydrogen)
    instead of 0.8: with 1.0 A bonds every 1-4 pair would fail the topo >= 3 form, and
    large random trees do not fit at 0.8.
  * A receptor of 3000 atoms with a 2.6 A minimum separation cannot be packed in a
    6-14 A shell (random sequential adsorption jams near 630 atoms there), so the
    shell's outer radius is grown until 3000 atoms fit (``receptor_outer_radius``).
"""
import math
import os

import numpy as np

# UFF type -> (atomic number, max neighbours, sp2?)
UFF_TYPES = {
    "C_3": (6, 4, False), "C_R": (6, 3, True), "N_3": (7, 3, False), "N_R": (7, 3, True),
    "O_2": (8, 1, True), "O_3": (8, 2, False), "S_3": (16, 2, False), "H_": (1, 1, False),
}
HEAVY_FRACTIONS = {"C_3": 0.30, "C_R": 0.35, "N_3": 0.08, "N_R": 0.08, "O_2": 0.08,
                   "O_3": 0.07, "S_3": 0.04}
RECEPTOR_FRACTIONS = {"C_3": 0.25, "C_R": 0.15, "N_3": 0.08, "N_R": 0.05, "O_2": 0.08,
                      "O_3": 0.06, "S_3": 0.03, "H_": 0.30}
# van der Waals radii used by the reference's clash rule (src/ptable.ml:32-54, :62-80)
VDW_RADII = {1: 1.2, 6: 1.7, 7: 1.6, 8: 1.55, 9: 1.5, 12: 2.2, 15: 1.95, 16: 1.8, 17: 1.8,
             35: 1.9, 53: 2.1}

# conformer-growth clash threshold d < alpha (r_vdW,i + r_vdW,j); the reference's docking
# filter uses 0.8 (src/ptable.ml:62-80), growth uses 0.65 so that large random trees fit
CLASH_ALPHA = 0.65
CLASH_ALPHA_H = 0.5   # pairs involving a hydrogen

SEEDS = {"config0": 20240, "config2": 20242, "config3": 20243, "config4": 20244}


# ----------------------------------------------------------------------------- ligands
def _draw_types(rng, n, fractions):
    names = list(fractions)
    p = np.array([fractions[k] for k in names])
    counts = np.floor(p / p.sum() * n).astype(int)
    # largest remainders so that the fractions are met as closely as possible
    rem = p / p.sum() * n - counts
    for k in np.argsort(-rem)[: n - counts.sum()]:
        counts[k] += 1
    out = []
    for k, c in zip(names, counts):
        out += [k] * int(c)
    rng.shuffle(out)
    return out


def _bfs_dists(adj):
    n = len(adj)
    d = np.full((n, n), -1, dtype=np.int32)
    for s in range(n):
        d[s, s] = 0
        frontier = [s]
        while frontier:
            nxt = []
            for a in frontier:
                for b in adj[a]:
                    if d[s, b] < 0:
                        d[s, b] = d[s, a] + 1
                        nxt.append(b)
            frontier = nxt
    return d


def _side(adj, a, b):
    """Atoms reachable from b once bond a-b is cut."""
    seen = {b}
    stack = [b]
    while stack:
        u = stack.pop()
        for v in adj[u]:
            if v == a and u == b:
                continue
            if v not in seen:
                seen.add(v)
                stack.append(v)
    return seen


def _nerf(a, b, c, length, theta, phi):
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    nn = np.linalg.norm(n)
    if nn < 1e-8:
        n = np.cross(bc, np.array([0.0, 0.0, 1.0]))
        if np.linalg.norm(n) < 1e-8:
            n = np.cross(bc, np.array([0.0, 1.0, 0.0]))
        nn = np.linalg.norm(n)
    n /= nn
    m = np.cross(n, bc)
    d = np.array([-length * math.cos(theta), length * math.sin(theta) * math.cos(phi),
                  length * math.sin(theta) * math.sin(phi)])
    return c + d[0] * bc + d[1] * m + d[2] * n


def _build_tree(rng, n_heavy, n_h):
    """Random heavy-atom tree respecting max neighbour counts, then hydrogens."""
    types = _draw_types(rng, n_heavy, HEAVY_FRACTIONS)
    # make sure the first atom can branch (O_2 is terminal)
    for k, t in enumerate(types):
        if UFF_TYPES[t][1] >= 3:
            types[0], types[k] = types[k], types[0]
            break
    # terminal O_2 atoms are attached last so that the chain can always grow
    types = [t for t in types if t != "O_2"] + [t for t in types if t == "O_2"]
    adj = [[] for _ in range(n_heavy)]
    order = [0]
    for i in range(1, n_heavy):
        # at most 3 heavy neighbours, chain extension preferred over branching:
        # drug-like skeletons, not sterically impossible dendrimers
        cand = [a for a in order if len(adj[a]) < min(3, UFF_TYPES[types[a]][1])]
        if not cand:
            raise RuntimeError("tree cannot grow")
        # branching grows with size so that large ligands stay within 12 A of their centre
        wts = np.array([{0: 1.0, 1: 1.0, 2: 0.15 + n_heavy / 300.0}[len(adj[a])] for a in cand])
        a = cand[rng.choice(len(cand), p=wts / wts.sum())]
        adj[a].append(i)
        adj[i].append(a)
        order.append(i)
    # hydrogens on free valences
    free = [UFF_TYPES[types[a]][1] - len(adj[a]) for a in range(n_heavy)]
    slots = [a for a in range(n_heavy) if types[a] != "O_2" for _ in range(max(free[a], 0))]
    rng.shuffle(slots)
    n_h = min(n_h, len(slots))
    for k in range(n_h):
        a = slots[k]
        h = len(types)
        types.append("H_")
        adj.append([a])
        adj[a].append(h)
    return types, adj


def _place(rng, types, adj, tries=8, budget=6000):
    """Depth-first conformer growth with backtracking: each atom's children are placed
    at tetrahedral/trigonal angles with a random dihedral (extended ones first), the
    subtree below is grown before the next sibling, and a subtree that cannot be
    grown clash-free re-draws the dihedrals one level up."""
    n = len(types)
    anum = [UFF_TYPES[t][0] for t in types]
    vdw = np.array([VDW_RADII[a] for a in anum])
    anum_a = np.array(anum)
    topo = _bfs_dists(adj)
    pos = np.zeros((n, 3))
    placed = np.zeros(n, dtype=bool)
    parent = [-1] * n
    log = []
    work = [0]

    def clash_free(c, p):
        pidx = np.nonzero(placed)[0]
        far = pidx[topo[c, pidx] >= 4]
        if not len(far):
            return True
        alpha = np.where((anum_a[c] == 1) | (anum_a[far] == 1), CLASH_ALPHA_H, CLASH_ALPHA)
        lim = alpha * (vdw[c] + vdw[far])
        return not np.any(np.sum((pos[far] - p) ** 2, axis=1) < lim * lim)

    def put(c, p, b):
        pos[c] = p
        placed[c] = True
        parent[c] = b
        log.append(c)

    def rollback(mark):
        while len(log) > mark:
            placed[log.pop()] = False

    def refs(b):
        a = parent[b]
        if a < 0:
            return None
        za = parent[a]
        if za < 0:
            sib = [s for s in adj[a] if placed[s] and s != b]
            return (pos[sib[0]] if sib else pos[a] + np.array([0.0, 1.0, 0.3])), pos[a]
        return pos[za], pos[a]

    def grow(b):
        kids = [c for c in adj[b] if not placed[c]]
        if not kids:
            return True
        sp2 = UFF_TYPES[types[b]][2]
        theta = math.radians(120.0 if sp2 else 109.5)
        step = 2.0 * math.pi / (2 if sp2 else 3)
        for attempt in range(tries):
            work[0] += 1
            if work[0] > budget:
                return False
            mark = len(log)
            r = refs(b)
            first = []
            if r is None:
                # root: first neighbour along a random direction, the others around it
                d = rng.normal(size=3)
                d /= np.linalg.norm(d)
                first = [(kids[0], pos[b] + d * rng.uniform(1.0, 1.54))]
                ref_a, ref_b = pos[b] + np.cross(d, [0.3, 0.7, 0.2]) - d, first[0][1]
                rest = kids[1:]
                # the angle is measured at the root between the first bond and the others
                trial = first + [(c, _nerf(ref_a, ref_b, pos[b], rng.uniform(1.0, 1.54), theta,
                                           rng.uniform(-math.pi, math.pi) + m * step))
                                 for m, c in enumerate(rest)]
            else:
                ref_a, ref_b = r
                phi0 = (math.pi + rng.uniform(-math.pi / 3, math.pi / 3)) if attempt < tries // 2 \
                    else rng.uniform(-math.pi, math.pi)
                trial = [(c, _nerf(ref_a, ref_b, pos[b], rng.uniform(1.0, 1.54), theta,
                                   phi0 + m * step)) for m, c in enumerate(kids)]
            ok = True
            for c, p in trial:
                if not clash_free(c, p):
                    ok = False
                    break
                put(c, p, b)
            if ok and all(grow(c) for c, _ in trial):
                return True
            rollback(mark)
        return False

    put(0, np.zeros(3), -1)
    return pos if grow(0) else None


def make_ligand(rng, n_heavy, n_h, n_rot, name="lig"):
    """A seeded random-tree ligand (see module docstring); charges ~N(0, 0.25) with
    net charge 0; n_rot rotatable bonds among internal heavy-heavy bonds."""
    for _attempt in range(400):
        types, adj = _build_tree(rng, n_heavy, n_h)
        pos = _place(rng, types, adj)
        # the reference drops ligands whose enclosing radius exceeds 12 A
        # (Mol.check_elongation_exn at load, src/lds.ml:1875-1885)
        if pos is not None and 0.01 + np.max(np.linalg.norm(pos - pos.mean(0), axis=1)) <= 12.0:
            break
    else:
        raise RuntimeError("could not build a clash-free ligand")
    n = len(types)
    anum = np.array([UFF_TYPES[t][0] for t in types], dtype=np.int32)
    q = rng.normal(0.0, 0.25, n)
    q -= q.mean()
    dists = _bfs_dists(adj)
    # rotatable candidates: heavy-heavy bonds whose both ends have degree > 1
    # (src/mol_graph.ml:128-138) and whose smaller side holds a heavy atom
    cands = []
    for a in range(n):
        for b in adj[a]:
            if a < b and anum[a] > 1 and anum[b] > 1 and len(adj[a]) > 1 and len(adj[b]) > 1:
                hv_a = sum(1 for c in adj[a] if anum[c] > 1)
                hv_b = sum(1 for c in adj[b] if anum[c] > 1)
                if hv_a > 1 and hv_b > 1:
                    cands.append((a, b))
    n_rot = min(n_rot, len(cands))
    pick = sorted(rng.choice(len(cands), n_rot, replace=False)) if n_rot else []
    rb_left, rb_right, rg_off, rg_atoms = [], [], [0], []
    for k in pick:
        a, b = cands[k]
        side_b = _side(adj, a, b)
        side_a = set(range(n)) - side_b
        # movable = smallest group (src/mol_graph.ml:140-192)
        if len(side_b) <= len(side_a):
            left, right, group = a, b, side_b
        else:
            left, right, group = b, a, side_a
        rb_left.append(left)
        rb_right.append(right)
        members = sorted(i for i in group if i != right)   # src/pqrs.ml:80-87
        rg_atoms += members
        rg_off.append(len(rg_atoms))
    pos = pos - pos.mean(axis=0)
    return {
        "name": name, "types": types,
        "xs": pos[:, 0].copy(), "ys": pos[:, 1].copy(), "zs": pos[:, 2].copy(),
        "q": q, "anum": anum, "dists": dists,
        "rb_left": np.array(rb_left, dtype=np.int32), "rb_right": np.array(rb_right, dtype=np.int32),
        "rg_off": np.array(rg_off, dtype=np.int32), "rg_atoms": np.array(rg_atoms, dtype=np.int32),
    }


# ----------------------------------------------------------------------------- receptor
def receptor_outer_radius(n_atoms, r_cavity, min_sep, fill=0.80):
    """Outer radius such that n_atoms fill `fill` of the RSA jamming density
    (packing fraction 0.384 for hard spheres of diameter min_sep)."""
    rho = 0.384 / (math.pi / 6.0 * min_sep ** 3) * fill
    vol = n_atoms / rho
    return (vol * 3.0 / (4.0 * math.pi) + r_cavity ** 3) ** (1.0 / 3.0)


def make_receptor(rng, n_atoms=3000, r_cavity=6.0, min_sep=2.6, r_out=None, box=None,
                  name="rec"):
    """Poisson-disc (dart throwing) atoms in the shell r_cavity < |x| < r_out around the
    origin, or in a cube of side `box` centred on the origin when `box` is given."""
    from scipy.spatial import cKDTree

    if box is None and r_out is None:
        r_out = receptor_outer_radius(n_atoms, r_cavity, min_sep)
    pts = np.zeros((0, 3))
    batch = 512
    min2 = min_sep * min_sep
    for _ in range(200000):
        if box is None:
            c = rng.uniform(-r_out, r_out, (batch, 3))
            rr = np.einsum("ij,ij->i", c, c)
            c = c[(rr > r_cavity ** 2) & (rr < r_out ** 2)]
        else:
            c = rng.uniform(-box / 2.0, box / 2.0, (batch, 3))
        if len(pts):
            d, _ = cKDTree(pts).query(c, k=1)
            c = c[d >= min_sep]
        if len(c) == 0:
            continue
        # greedy sequential acceptance inside the batch (dart order preserved)
        d2 = np.sum((c[:, None, :] - c[None, :, :]) ** 2, axis=2)
        keep = np.zeros(len(c), dtype=bool)
        for i in range(len(c)):
            if not np.any(keep[:i] & (d2[i, :i] < min2)):
                keep[i] = True
        pts = np.vstack([pts, c[keep][: n_atoms - len(pts)]])
        if len(pts) >= n_atoms:
            break
    else:
        raise RuntimeError("receptor packing did not converge")
    pts = pts[:n_atoms]
    types = _draw_types(rng, n_atoms, RECEPTOR_FRACTIONS)
    anum = np.array([UFF_TYPES[t][0] for t in types], dtype=np.int32)
    q = rng.normal(0.0, 0.3, n_atoms)
    return {"name": name, "types": types, "xs": pts[:, 0].copy(), "ys": pts[:, 1].copy(),
            "zs": pts[:, 2].copy(), "q": q, "anum": anum, "r_cavity": r_cavity,
            "r_out": r_out, "min_sep": min_sep}


# ----------------------------------------------------------------------------- rotations
def so3_rotations(n):
    """Super-Fibonacci SO(3) samples as rotation matrices (src/SO3.ml:18-39):
    quaternion -> axis-angle (src/quat.ml:32-36) -> Rodrigues (src/rot.ml:121-146)."""
    i = np.arange(n, dtype=np.float64)
    s = i + 0.5
    t = s / n
    d = 2.0 * (4.0 * math.atan(1.0)) * s
    cr, cR = np.sqrt(t), np.sqrt(1.0 - t)
    al, be = d / math.sqrt(2.0), d / 1.533751168755204288118041
    qw, qx, qy, qz = cr * np.sin(al), cr * np.cos(al), cR * np.sin(be), cR * np.cos(be)
    mag = np.sqrt(qx * qx + qy * qy + qz * qz)
    x, y, z = qx / mag, qy / mag, qz / mag
    th = 2.0 * np.arctan2(mag, qw)
    c, sn = np.cos(th), np.sin(th)
    omc = 1.0 - c
    m = np.empty((n, 3, 3))
    m[:, 0, 0] = c + x * x * omc
    m[:, 0, 1] = x * y * omc - z * sn
    m[:, 0, 2] = x * z * omc + y * sn
    m[:, 1, 0] = x * y * omc + z * sn
    m[:, 1, 1] = c + y * y * omc
    m[:, 1, 2] = y * z * omc - x * sn
    m[:, 2, 0] = x * z * omc - y * sn
    m[:, 2, 1] = y * z * omc + x * sn
    m[:, 2, 2] = c + z * z * omc
    return m


def random_rotations(rng, n):
    """Uniform random rotations from Shoemake's quaternion sampling."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    w, x, y, z = a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2), \
        b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)
    # quaternion -> matrix (src/rot.ml:104-114)
    m = np.empty((n, 3, 3))
    m[:, 0, 0] = 1 - 2 * (y * y + z * z)
    m[:, 0, 1] = 2 * (x * y - z * w)
    m[:, 0, 2] = 2 * (x * z + y * w)
    m[:, 1, 0] = 2 * (x * y + z * w)
    m[:, 1, 1] = 1 - 2 * (x * x + z * z)
    m[:, 1, 2] = 2 * (y * z - x * w)
    m[:, 2, 0] = 2 * (x * z - y * w)
    m[:, 2, 1] = 2 * (y * z + x * w)
    m[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return m


def lattice(center, n, step):
    """n^3 translations at `step` spacing centred on `center`, x fastest."""
    off = (np.arange(n) - (n - 1) / 2.0) * step
    z, y, x = np.meshgrid(off, off, off, indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1) + np.asarray(center, dtype=np.float64)


# ----------------------------------------------------------------------------- configs
_CACHE = {}


def config0():
    """One ligand (40 heavy + 35 H, 6 rotatable bonds) and the 3000-atom receptor."""
    if "config0" not in _CACHE:
        rng = np.random.Generator(np.random.PCG64(SEEDS["config0"]))
        lig = make_ligand(rng, 40, 35, 6, name="config0_lig")
        rec = make_receptor(rng, 3000, r_cavity=6.0, min_sep=2.6, name="config0_rec")
        _CACHE["config0"] = (lig, rec)
    return _CACHE["config0"]


def config1_scan(n_rot=4608, n_trans=20, trans_step=0.5):
    """Rigid scan: n_rot super-Fibonacci rotations x n_trans^3 lattice translations
    around the cavity centre (pose index = rot * n_trans^3 + trans)."""
    return so3_rotations(n_rot), lattice((0.0, 0.0, 0.0), n_trans, trans_step)


def config4_receptor():
    """15,000 atoms in a 60 A box.  Random sequential adsorption jams near 2.45 A at
    this density, so the minimum separation is 2.0 A (recorded with the workload)."""
    if "config4" not in _CACHE:
        rng = np.random.Generator(np.random.PCG64(SEEDS["config4"]))
        _CACHE["config4"] = make_receptor(rng, 15000, box=60.0, min_sep=2.0, name="config4_rec")
    return _CACHE["config4"]


def grid_types(ligs, ntypes=10):
    """Probe types for the per-type grids (src/lds.ml:452-469 builds one grid per
    distinct (anum, q)).  Atoms of `ligs` are grouped by element and charge sign, and
    the group with the widest charge range is split at its median until there are
    `ntypes` groups; each group's probe charge is its mean.  Returns
    (probe_anum, probe_q, [type id per atom of each ligand])."""
    an = np.concatenate([np.asarray(l["anum"]) for l in ligs])
    q = np.concatenate([np.asarray(l["q"], dtype=np.float64) for l in ligs])
    groups = {}
    for i, (a, v) in enumerate(zip(an, q)):
        groups.setdefault((int(a), v >= 0.0), []).append(i)
    groups = [np.array(g) for _, g in sorted(groups.items())]
    while len(groups) < ntypes:
        k = max(range(len(groups)), key=lambda j: np.ptp(q[groups[j]]) if len(groups[j]) > 1 else -1)
        g = groups[k]
        if len(g) < 2:
            break
        order = g[np.argsort(q[g], kind="stable")]
        groups[k:k + 1] = [order[:len(order) // 2], order[len(order) // 2:]]
    tid = np.zeros(len(an), dtype=np.int32)
    for t, g in enumerate(groups):
        tid[g] = t
    probe_a = [int(an[g[0]]) for g in groups]
    probe_q = [float(q[g].mean()) for g in groups]
    out, off = [], 0
    for l in ligs:
        n = len(l["anum"])
        out.append(tid[off:off + n].copy())
        off += n
    return probe_a, probe_q, out


def grid_spec(center=(0.0, 0.0, 0.0), n=64, step=0.375):
    """A cubic grid of n^3 points at `step` centred on `center` -> (origin, dims, step)."""
    half = step * (n - 1) / 2.0
    origin = np.asarray(center, dtype=np.float64) - half
    return origin, (n, n, n), step


def _lib_one(args):
    seed, k, nh, n_h, nr = args
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence([seed, k])))
    return make_ligand(rng, nh, n_h, nr, name="lib%04d" % k)


def library(n_ligands=1024, seed=None, median_heavy=30, sigma=0.35, h_ratio=0.875,
            max_rot=15, processes=1):
    """Ligand library: heavy ~ round(lognormal(ln median, sigma)) clipped to [12, 70],
    H ~ h_ratio * heavy, rotatable bonds ~ U{0..max_rot}.  Ligand k is grown from its
    own seed (seed, k), so the library is the same for any number of processes."""
    seed = SEEDS["config2"] if seed is None else seed
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    for k in range(n_ligands):
        nh = int(np.clip(round(rng.lognormal(math.log(median_heavy), sigma)), 12, 70))
        nr = int(rng.integers(0, max_rot + 1))
        jobs.append((seed, k, nh, int(round(h_ratio * nh)), nr))
    # worker processes are started fresh ("spawn"): nothing GPU-related is inherited
    if processes and processes > 1 and n_ligands >= 32:
        import multiprocessing as mp
        # close() + join() let the workers exit on their own; the context manager's
        # terminate() would end them with SIGTERM
        pool = mp.get_context("spawn").Pool(min(processes, 8))
        try:
            out = pool.map(_lib_one, jobs, chunksize=4)
            pool.close()
            pool.join()
            return out
        except BaseException:
            pool.terminate()
            raise
    return [_lib_one(j) for j in jobs]
