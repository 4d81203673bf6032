"""Oracle (test infrastructure only): a Python restatement of the reference's ligand
input path, the checker of the library's C++ implementation (mmo_molset_*).

Follows Mol2 parsing (src/mol2.ml:150-323: atom lines "id name x y z type subst_id
subst_name charge", bond lines "id src dst type", lone pairs removed), Ptable's mol2
type -> element map and vdW radii (src/ptable.ml:43-53, :169-225), and Mol_graph
(src/mol_graph.ml:53-192):

- topological distances: all-pairs shortest paths of the bond graph (unit weights;
  a disconnected atom is an error);
- rotatable bonds: bond order 1 (types "1" and "am"), both atoms of degree > 1, and not
  in a ring (removing the bond leaves its atoms connected), in bond-file order;
- rotation group of a rotatable bond: the smaller of the two parts the bond separates
  (on a tie the part without the bond's first atom; the reference's choice on ties
  follows hash-table order: parity unpinned there);

writes the .pqrs block of Mol2pqrs.process_ligand (src/mol2pqrs.ml:9-43): "n:nrb:name",
one "x y z q r elt" line per atom (%g), one "src-dst= b_0 .. b_{n-1}" line per rotatable
bond, then the distance matrix rows; and reads .pqrs text back (src/pqrs.ml:17-87,
:204-250) into plain dict records.
"""
from collections import deque

import numpy as np


MOLECULE_TAG, ATOM_TAG, BOND_TAG = "@<TRIPOS>MOLECULE", "@<TRIPOS>ATOM", "@<TRIPOS>BOND"
UNITY_TAG = "@<TRIPOS>UNITY_ATOM_ATTR"

# Ptable.anum_of_mol2_type (src/ptable.ml:169-225)
MOL2_TYPE_ANUM = {
    "H": 1, "H.t3p": 1, "H.spc": 1, "Li": 3,
    "C.3": 6, "C.2": 6, "C.1": 6, "C.ar": 6, "C.cat": 6,
    "N.3": 7, "N.2": 7, "N.1": 7, "N.ar": 7, "N.am": 7, "N.pl3": 7, "N.4": 7,
    "O.3": 8, "O.2": 8, "O.co2": 8, "O.spc": 8, "O.t3p": 8,
    "F": 9, "Na": 11, "Mg": 12, "Al": 13, "Si": 14, "P.3": 15,
    "S.3": 16, "S.2": 16, "S.O": 16, "S.o": 16, "S.O2": 16, "S.o2": 16,
    "Cl": 17, "K": 19, "Ca": 20, "Cr.th": 24, "Cr.oh": 24, "Mn": 25, "Fe": 26,
    "Co.oh": 27, "Cu": 29, "Zn": 30, "Se": 34, "Br": 35, "Mo": 42, "Sn": 50, "I": 53,
}
# Ptable.vdW_radii (src/ptable.ml:43-53)
VDW_RADII = {1: 1.2, 6: 1.7, 7: 1.6, 8: 1.55, 9: 1.5, 12: 2.2, 15: 1.95, 16: 1.8, 17: 1.8,
             35: 1.9, 53: 2.1}
# Mol2.bond_order (src/mol2.ml:140-150)
BOND_ORDER = {"1": 1.0, "2": 2.0, "3": 3.0, "ar": 1.5, "am": 1.0}


class Mol2:
    def __init__(self, name, atoms, bonds):
        self.name = name
        self.atoms = atoms      # dicts: name, x, y, z, type, q, anum, r, symb
        self.bonds = bonds      # (src, dst, order, type), 0-based

    def num_atoms(self):
        return len(self.atoms)


def _blocks(text):
    block = None
    for line in text.splitlines():
        if line.strip() == MOLECULE_TAG:
            if block:
                yield block
            block = [line]
        elif block is not None:
            block.append(line)
    if block:
        yield block


def _parse_block(block):
    name = block[1].strip()
    n_atoms, n_bonds = (int(v) for v in block[2].split()[:2])
    i = 3
    while block[i].strip() != ATOM_TAG:
        i += 1
    i += 1
    atoms = []
    for _ in range(n_atoms):
        f = block[i].split()
        i += 1
        typ = f[5]
        q = float(f[8]) if len(f) > 8 else 0.0
        if typ == "LP":
            atoms.append({"name": f[1], "type": typ, "lp": True})
            continue
        if typ not in MOL2_TYPE_ANUM:
            raise ValueError("%s: unsupported mol2 atom type %s" % (name, typ))
        anum = MOL2_TYPE_ANUM[typ]
        atoms.append({"name": f[1], "x": float(f[2]), "y": float(f[3]), "z": float(f[4]),
                      "type": typ, "q": q, "anum": anum, "r": VDW_RADII.get(anum, float("nan")),
                      "symb": typ.split(".")[0], "lp": False})
    while block[i].strip() != BOND_TAG:
        i += 1
    i += 1
    bonds = []
    for _ in range(n_bonds):
        f = block[i].split()
        i += 1
        if f[3] not in BOND_ORDER:
            raise ValueError("%s: unsupported bond type %s" % (name, f[3]))
        bonds.append((int(f[1]) - 1, int(f[2]) - 1, BOND_ORDER[f[3]], f[3]))
    # Mol2.remove_lone_pairs (src/mol2.ml:99-113): drop LPs and their bonds, renumber
    keep = [k for k, a in enumerate(atoms) if not a["lp"]]
    new = {k: j for j, k in enumerate(keep)}
    bonds = [(new[s], new[d], o, t) for s, d, o, t in bonds if s in new and d in new]
    return Mol2(name, [atoms[k] for k in keep], bonds)


def read_mol2(fn):
    with open(fn) as f:
        return [_parse_block(b) for b in _blocks(f.read())]


# ------------------------------------------------------------------ topology
def _adjacency(n, bonds, skip=None):
    adj = [[] for _ in range(n)]
    for k, (s, d, _, _) in enumerate(bonds):
        if k != skip:
            adj[s].append(d)
            adj[d].append(s)
    return adj


def _bfs(adj, src):
    dist = [-1] * len(adj)
    dist[src] = 0
    q = deque([src])
    while q:
        u = q.popleft()
        for v in adj[u]:
            if dist[v] < 0:
                dist[v] = dist[u] + 1
                q.append(v)
    return dist


def dist_matrix(m):
    """Topological distances, n x n int32 (Mol_graph.dist_matrix)."""
    n = m.num_atoms()
    adj = _adjacency(n, m.bonds)
    out = np.zeros((n, n), dtype=np.int32)
    for i in range(n):
        d = _bfs(adj, i)
        if min(d) < 0:
            raise ValueError("disconnected atom in %s" % m.name)
        out[i] = d
    return out


def degrees(m):
    deg = np.zeros(m.num_atoms(), dtype=np.int32)
    for s, d, _, _ in m.bonds:
        deg[s] += 1
        deg[d] += 1
    return deg


def rotatable_bonds(m):
    """Indices (into m.bonds) of the rotatable bonds (Mol_graph.list_rotatable_bonds)."""
    n, deg, out = m.num_atoms(), degrees(m), []
    for k, (s, d, order, _) in enumerate(m.bonds):
        if order == 1.0 and deg[s] > 1 and deg[d] > 1:
            if _bfs(_adjacency(n, m.bonds, skip=k), s)[d] < 0:   # not a ring bond
                out.append(k)
    return out


def rot_group(m, k):
    """The atoms moved by rotatable bond k: the smaller part (Mol_graph.compute_rot_group)."""
    s, d, _, _ = m.bonds[k]
    side_d = np.array(_bfs(_adjacency(m.num_atoms(), m.bonds, skip=k), d)) >= 0
    n_d = int(side_d.sum())
    n_s = m.num_atoms() - n_d
    return side_d if n_d <= n_s else ~side_d


def pqrs_text(m):
    """The .pqrs block Mol2pqrs.process_ligand writes for this molecule."""
    n = m.num_atoms()
    rb = rotatable_bonds(m)
    lines = ["%d:%d:%s" % (n, len(rb), m.name)]
    for a in m.atoms:
        lines.append("%g %g %g %g %g %s" % (a["x"], a["y"], a["z"], a["q"], a["r"], a["symb"]))
    for k in rb:
        s, d, _, _ = m.bonds[k]
        g = rot_group(m, k)
        lines.append("%d-%d=%s" % (s, d, "".join(" 1" if b else " 0" for b in g)))
    dm = dist_matrix(m)
    for i in range(n):
        lines.append(" ".join(str(int(v)) for v in dm[i]))   # symmetric: rows = columns
    return "\n".join(lines) + "\n"


def write_pqrs_from_mol2(mol2_fn, pqrs_fn):
    with open(pqrs_fn, "w") as f:
        for m in read_mol2(mol2_fn):
            f.write(pqrs_text(m))


SYM2ANUM = {"C": 6, "H": 1, "N": 7, "O": 8, "P": 15, "S": 16, "F": 9, "Cl": 17, "Br": 35,
            "I": 53, "Mg": 12}            # src/ptable.ml:86-107


def _parse_header(line):
    """src/pqrs.ml:17-29: 'N:R:name' (ligand) or 'N:name' (receptor)."""
    parts = line.strip().split(":")
    if len(parts) >= 3:
        try:
            return int(parts[0]), int(parts[1]), ":".join(parts[2:])
        except ValueError:
            pass
    return int(parts[0]), -1, ":".join(parts[1:])


def _read_one(lines, pos, ligand):
    n, nrb, name = _parse_header(lines[pos])
    pos += 1
    rec = {"name": name, "xs": [], "ys": [], "zs": [], "q": [], "r": [], "anum": []}
    for _ in range(n):
        f = lines[pos].split()
        pos += 1
        for key, v in zip(("xs", "ys", "zs", "q", "r"), f[:5]):
            rec[key].append(float(v))
        if f[5] not in SYM2ANUM:
            raise ValueError("unsupported element: %s" % f[5])   # src/ptable.ml:103-107
        rec["anum"].append(SYM2ANUM[f[5]])
    if not ligand:
        return rec, pos
    rec.update(rb_left=[], rb_right=[], groups=[])
    for _ in range(max(nrb, 0)):
        head, bits = lines[pos].split("=", 1)
        pos += 1
        beg, end = (int(v) for v in head.split("-"))
        b = [int(v) for v in bits.split()]
        if len(b) != n or b[beg] == b[end]:
            raise ValueError("bad rot-bond line: %s" % lines[pos - 1])
        # axis from the fixed atom to the moving one (src/pqrs.ml:41-60)
        left, right = (beg, end) if b[end] else (end, beg)
        rec["rb_left"].append(left)
        rec["rb_right"].append(right)
        rec["groups"].append([i for i in range(n) if b[i] and i != right])   # src/pqrs.ml:80-87
    dm = np.zeros((n, n), dtype=np.int32)
    for i in range(n):
        row = [int(v) for v in lines[pos].split()]
        pos += 1
        if len(row) != n or row[i] != 0:
            raise ValueError("bad distance matrix row %d of %s" % (i, name))
        dm[i] = row
    rec["dists_colmajor"] = dm.T.reshape(-1)        # flat[i + j n] = row i, column j
    return rec, pos


def read_pqrs_text(text, ligand=True):
    lines = [l for l in text.splitlines() if l.strip()]
    out, pos, seen = [], 0, set()
    while pos < len(lines):
        m, pos = _read_one(lines, pos, ligand)
        if m["name"] in seen:
            raise ValueError("duplicate molecule name %s" % m["name"])
        seen.add(m["name"])
        out.append(m)
        if not ligand:
            break
    return out


def mol2_to_pqrs_text(fn):
    return "".join(pqrs_text(m) for m in read_mol2(fn))
