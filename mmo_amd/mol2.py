"""Ligand input from .mol2 files (the reference's mol2pqrs step), parsed by the library.

`mmo_molset_from_mol2` (C++) reads the Tripos blocks (lone pairs removed, src/mol2.ml:
99-323), derives the topology (topological distances, rotatable bonds, rotation groups,
src/mol_graph.ml:53-192), writes the .pqrs block of Mol2pqrs.process_ligand
(src/mol2pqrs.ml:9-43) and reads it back, so ligands from .mol2 and from the .pqrs file
the reference would write are identical Mol records.
"""
from . import mol as molmod

# Ptable.vdW_radii (src/ptable.ml:43-53), by atomic number
VDW_RADII = {1: 1.2, 6: 1.7, 7: 1.6, 8: 1.55, 9: 1.5, 12: 2.2, 15: 1.95, 16: 1.8, 17: 1.8,
             35: 1.9, 53: 2.1}


def ligands_of_mol2_text(text):
    return molmod.molecules_of_text(text, "mol2")[0]


def ligands_of_mol2_file(fn):
    with open(fn) as f:
        return ligands_of_mol2_text(f.read())


def pqrs_text_of_mol2(fn):
    """The .pqrs text mol2pqrs writes for every molecule of the file."""
    with open(fn) as f:
        return molmod.molecules_of_text(f.read(), "mol2")[1]


def write_pqrs_from_mol2(mol2_fn, pqrs_fn):
    with open(pqrs_fn, "w") as f:
        f.write(pqrs_text_of_mol2(mol2_fn))
