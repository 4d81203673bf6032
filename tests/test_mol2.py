"""The .mol2 -> topology -> .pqrs input path (src/mol2.ml, src/mol_graph.ml,
src/mol2pqrs.ml, src/pqrs.ml).  The oracle's restatement (oracle/mol2_ref.py) is pinned
on hand-written molecules whose topology is known; the library's C++ reader
(mmo_molset_*) must then give the same .pqrs text and the same molecules, on those and
on the .mol2 files the reference ships (tests/golden/*.mol2, copied from its data/)."""
import os

import numpy as np
import pytest

from mmo_amd import mol
from oracle import mol2_ref as ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
REF_MOL2 = ("docked.mol2", "minimized.mol2", "ligdecs.mol2")

# n-butane with its 10 hydrogens: C1-C2-C3-C4 chain
BUTANE = """@<TRIPOS>MOLECULE
butane
   14    13     0     0     0
SMALL
USER_CHARGES

@<TRIPOS>ATOM
      1 C1          0.0000    0.0000    0.0000 C.3       1 <0>        -0.1800
      2 C2          1.5300    0.0000    0.0000 C.3       1 <0>        -0.1200
      3 C3          2.0400    1.4400    0.0000 C.3       1 <0>        -0.1200
      4 C4          3.5700    1.4400    0.0000 C.3       1 <0>        -0.1800
      5 H1         -0.3600   -1.0200    0.0000 H         1 <0>         0.0600
      6 H2         -0.3600    0.5100    0.8800 H         1 <0>         0.0600
      7 H3         -0.3600    0.5100   -0.8800 H         1 <0>         0.0600
      8 H4          1.8900   -0.5100    0.8800 H         1 <0>         0.0600
      9 H5          1.8900   -0.5100   -0.8800 H         1 <0>         0.0600
     10 H6          1.6800    1.9500    0.8800 H         1 <0>         0.0600
     11 H7          1.6800    1.9500   -0.8800 H         1 <0>         0.0600
     12 H8          3.9300    2.4600    0.0000 H         1 <0>         0.0600
     13 H9          3.9300    0.9300    0.8800 H         1 <0>         0.0600
     14 H10         3.9300    0.9300   -0.8800 H         1 <0>         0.0600
@<TRIPOS>BOND
     1     1     2 1
     2     2     3 1
     3     3     4 1
     4     1     5 1
     5     1     6 1
     6     1     7 1
     7     2     8 1
     8     2     9 1
     9     3    10 1
    10     3    11 1
    11     4    12 1
    12     4    13 1
    13     4    14 1
"""

# methylcyclopropane with an amide-bonded N and a lone pair to be dropped
RING = """@<TRIPOS>MOLECULE
ring_lp
    7     7     0     0     0
SMALL
USER_CHARGES

@<TRIPOS>ATOM
      1 C1          0.0000    0.0000    0.0000 C.3       1 <0>        -0.1000
      2 C2          1.5000    0.0000    0.0000 C.3       1 <0>        -0.1000
      3 C3          0.7500    1.3000    0.0000 C.3       1 <0>        -0.1000
      4 C4         -1.3000   -0.7500    0.0000 C.2       1 <0>         0.5000
      5 LP1        -1.0000    1.0000    1.0000 LP        1 <0>         0.0000
      6 N1         -2.6000    0.0000    0.0000 N.am      1 <0>        -0.4000
      7 O1         -1.3000   -2.0000    0.0000 O.2       1 <0>         0.2000
@<TRIPOS>BOND
     1     1     2 1
     2     2     3 1
     3     3     1 1
     4     1     4 1
     5     4     6 am
     6     4     7 2
     7     6     5 1
"""


def _write(tmp_path, text, name):
    fn = str(tmp_path / name)
    with open(fn, "w") as f:
        f.write(text)
    return fn


def test_butane_topology(tmp_path):
    (m,) = ref.read_mol2(_write(tmp_path, BUTANE, "b.mol2"))
    assert m.num_atoms() == 14 and len(m.bonds) == 13
    dm = ref.dist_matrix(m)
    assert dm[0, 3] == 3 and dm[4, 11] == 5 and np.array_equal(dm, dm.T)
    # every C-C bond joins atoms of degree 4: all three are rotatable; C-H bonds are not
    assert ref.rotatable_bonds(m) == [0, 1, 2]
    g = ref.rot_group(m, 0)                 # C1-C2: the methyl end moves
    assert np.flatnonzero(g).tolist() == [0, 4, 5, 6]
    g = ref.rot_group(m, 1)                 # C2-C3: 7 vs 7 atoms, the C3 side moves
    assert np.flatnonzero(g).tolist() == [2, 3, 9, 10, 11, 12, 13]


def test_ring_bonds_lone_pairs_and_bond_orders(tmp_path):
    (m,) = ref.read_mol2(_write(tmp_path, RING, "r.mol2"))
    assert m.num_atoms() == 6 and len(m.bonds) == 6          # LP and its bond dropped
    assert [a["anum"] for a in m.atoms] == [6, 6, 6, 6, 7, 8]
    # ring bonds are not rotatable; C1-C4 is single and both ends have degree > 1;
    # the amide C4-N1 is single ("am") but N1 is terminal once the LP is gone
    assert ref.rotatable_bonds(m) == [3]
    assert np.flatnonzero(ref.rot_group(m, 3)).tolist() == [3, 4, 5]


def test_pqrs_text_round_trip(tmp_path):
    fn = _write(tmp_path, BUTANE + RING, "two.mol2")
    ligs = mol.ligands_of_mol2_file(fn)
    assert [l.name for l in ligs] == ["butane", "ring_lp"]
    b = ligs[0]
    assert b.num_rbonds() == 3
    # axis from the fixed atom to the moving one; the group excludes the axis atom
    assert list(b.rbonds_a[1]) == [1, 2]
    assert sorted(b.rgroups_a[1]) == [3, 9, 10, 11, 12, 13]
    n = b.num_atoms()
    d = np.asarray(b.dists_a).reshape(n, n)
    assert d[0, 3] == 3 and d[3, 0] == 3
    assert np.allclose(b.r_a[:4], 1.7) and np.allclose(b.r_a[4:], 1.2)
    text = mol.pqrs_text_of_mol2(fn)
    assert text.splitlines()[0] == "14:3:butane"
    assert text.splitlines()[1] == "0 0 0 -0.18 1.7 C"


def _same_molecules(got, want):
    assert [m.name for m in got] == [w["name"] for w in want]
    for m, w in zip(got, want):
        for key, arr in (("xs", m.xs), ("ys", m.ys), ("zs", m.zs), ("q", m.q_a)):
            assert np.array_equal(arr, np.asarray(w[key])), (m.name, key)
        assert np.array_equal(m.r_a, np.asarray(w["r"]), equal_nan=True)
        assert m.elt_a.tolist() == w["anum"]
        if "rb_left" in w:
            assert [tuple(b) for b in m.rbonds_a] == list(zip(w["rb_left"], w["rb_right"]))
            assert [g.tolist() for g in m.rgroups_a] == w["groups"]
            assert np.array_equal(m.dists_a, w["dists_colmajor"])


@pytest.mark.parametrize("name", REF_MOL2)
def test_library_reader_matches_oracle_on_reference_files(name):
    fn = os.path.join(GOLDEN, name)
    text = mol.pqrs_text_of_mol2(fn)
    assert text == ref.mol2_to_pqrs_text(fn)
    _same_molecules(mol.ligands_of_mol2_file(fn), ref.read_pqrs_text(text))
    # the .pqrs route gives the same molecules
    _same_molecules(mol.ligands_of_pqrs_text(text), ref.read_pqrs_text(text))


def test_library_reader_matches_oracle_on_known_topologies(tmp_path):
    fn = _write(tmp_path, BUTANE + RING, "two.mol2")
    text = mol.pqrs_text_of_mol2(fn)
    assert text == ref.mol2_to_pqrs_text(fn)
    _same_molecules(mol.ligands_of_mol2_file(fn), ref.read_pqrs_text(text))


def test_reference_mol2_files_topology_invariants():
    for name in REF_MOL2:
        for m in ref.read_mol2(os.path.join(GOLDEN, name)):
            dm = ref.dist_matrix(m)
            assert np.array_equal(dm, dm.T) and (np.diag(dm) == 0).all()
            deg = ref.degrees(m)
            assert deg.sum() == 2 * len(m.bonds)
            for k in ref.rotatable_bonds(m):
                s, d, order, _ = m.bonds[k]
                assert order == 1.0 and deg[s] > 1 and deg[d] > 1
                g = ref.rot_group(m, k)
                assert g[s] != g[d] and g.sum() <= m.num_atoms() - g.sum()


def test_receptor_pqrs_and_errors(tmp_path):
    rec = "3:prot\n0 0 0 0.1 1.7 C\n1.5 0 0 -0.1 1.55 O\n3 0 0 0 1.2 H\n"
    fn = _write(tmp_path, rec, "rec.pqrs")
    r = mol.receptor_of_pqrs_file(fn)
    assert r.name == "prot" and r.elt_a.tolist() == [6, 8, 1] and r.dists_a is None
    assert np.array_equal(r.xs, [0.0, 1.5, 3.0]) and np.array_equal(r.q_a, [0.1, -0.1, 0.0])
    good = ref.mol2_to_pqrs_text(_write(tmp_path, BUTANE, "b.mol2"))
    bad_elt = good.replace(" C\n", " Xx\n", 1)
    with pytest.raises(ValueError, match="unsupported element"):
        mol.ligands_of_pqrs_text(bad_elt)
    with pytest.raises(ValueError, match="duplicate molecule name"):
        mol.ligands_of_pqrs_text(good + good)
    lines = good.splitlines()
    lines[15] = lines[15].replace("=", "")               # a rot-bond line without '='
    with pytest.raises(ValueError, match="rot-bond"):
        mol.ligands_of_pqrs_text("\n".join(lines) + "\n")
    with pytest.raises(ValueError, match="unsupported mol2 atom type"):
        mol.molecules_of_text(BUTANE.replace(" C.3 ", " Q.9 ", 1), "mol2")
    with pytest.raises(ValueError, match="disconnected atom"):
        mol.molecules_of_text(BUTANE.replace("    13     4    14 1\n", "").replace("   14    13 ", "   14    12 "), "mol2")
