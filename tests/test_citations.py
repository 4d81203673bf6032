"""Reference citations (``src/<file>.ml:<a>-<b>``) in the product, the oracle and the
design notes point at lines that exist, and the anchor citations below point at the
functions they name.  Skipped when the reference tree is not mounted."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"

CITE = re.compile(r"((?:src|bin|data|doc)/[A-Za-z0-9_./-]+\.(?:ml|py|sh|txt|bild|mol2|pdb))"
                  r"((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")
SCANNED_DIRS = ("include", "oracle", "mmo_amd", "tests")
SCANNED_FILES = ("DESIGN.md", "INTEGRATION.md", "bench.py", "__graft_entry__.py")
EXTS = (".py", ".hip", ".h", ".c", ".cpp", ".md")

# (citation, text the cited lines must contain)
ANCHORS = [
    ("src/UFF.ml:10-22", "anum_xi_Di"),
    ("src/UFF.ml:25", "332.0637"),
    ("src/UFF.ml:37", "nan"),
    ("src/UFF.ml:39-51", "geo_mean"),
    ("src/FF.ml:17-20", "shift_12A"),
    ("src/math.ml:58-62", "non_zero_dist"),
    ("src/const.ml:23-28", "room_temp_K"),
    ("src/params.ml:10-33", "max_rot"),
    ("src/mol.ml:796-818", "ene_inter_UFF_global_brute"),
    ("src/mol.ml:822-849", "ene_inter_UFF_shifted_brute"),
    ("src/mol.ml:881-903", "ene_intra_UFFNB_brute"),
    ("src/mol.ml:928-960", "ene_inter_UFF_shifted_bst"),
    ("src/mol.ml:1012-1020", "ene_inter_UFF_interp"),
    ("src/mol.ml:610-631", "rotate_bond"),
    ("src/mol.ml:368-402", "pqrs_read_header_coords"),
    ("src/pqrs.ml:17-87", "parse_header"),
    ("src/pqrs.ml:62-77", "parse_dist_matrix"),
    ("src/pqrs.ml:80-87", "rgroup_of_bools"),
    ("src/ptable.ml:43-53", "vdW_radii"),
    ("src/ptable.ml:169-225", "anum_of_mol2_type"),
    ("src/G3D.ml:97-157", "trilin"),
    ("src/optim.ml:64-80", "apply_config"),
    ("src/rot.ml:52-66", "r_xyz"),
    ("src/SO3.ml:18-39", "super_fibonacci"),
    ("src/lds.ml:586-621", "update_rot_trans_params"),
    ("src/lds.ml:632-648", "reset_run_params"),
    ("src/lds.ml:741-1000", "simulate_lig"),
    ("src/lds.ml:1040-1114", "exhaustive_rigid_ligand_docking"),
    ("src/lds.ml:1118-1136", "approx_local_gradient"),
    ("src/lds.ml:1138-1297", "minimize_ligand"),
    ("src/lds.ml:198-267", "desolvation"),
    ("src/ROI.ml:22-32", "from_bild"),
]

needs_ref = pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src")),
                               reason="reference tree not mounted")


def _sources():
    for d in SCANNED_DIRS:
        for base, _, files in os.walk(os.path.join(ROOT, d)):
            if "__pycache__" in base or "_build" in base:
                continue
            for f in files:
                if f.endswith(EXTS):
                    yield os.path.join(base, f)
    for f in SCANNED_FILES:
        p = os.path.join(ROOT, f)
        if os.path.exists(p):
            yield p


def _ref_lines(rel, cache={}):
    p = os.path.join(REF, rel)
    if p not in cache:
        with open(p, errors="replace") as f:
            cache[p] = f.read().splitlines()
    return cache[p]


def _ranges(spec):
    for a, b in re.findall(r"(\d+)(?:-(\d+))?", spec):
        yield int(a), int(b) if b else int(a)


@needs_ref
def test_every_citation_is_in_range():
    bad = []
    n = 0
    for path in _sources():
        with open(path, errors="replace") as f:
            for ln, line in enumerate(f, 1):
                for m in CITE.finditer(line):
                    rel = m.group(1)
                    where = "%s:%d" % (os.path.relpath(path, ROOT), ln)
                    if not os.path.exists(os.path.join(REF, rel)):
                        bad.append("%s cites missing file %s" % (where, rel))
                        continue
                    nl = len(_ref_lines(rel))
                    for a, b in _ranges(m.group(2)):
                        n += 1
                        if not (1 <= a <= b <= nl):
                            bad.append("%s cites %s:%d-%d (%d lines)" % (where, rel, a, b, nl))
    assert n > 200, "too few citations found (%d): scanner broken?" % n
    assert not bad, "\n".join(bad)


@needs_ref
@pytest.mark.parametrize("cite,text", ANCHORS)
def test_anchor_citations_name_the_right_code(cite, text):
    rel, spec = cite.split(":", 1)
    (a, b), = list(_ranges(spec))
    body = "\n".join(_ref_lines(rel)[a - 1:b])
    assert text in body, "%s does not contain %r" % (cite, text)
