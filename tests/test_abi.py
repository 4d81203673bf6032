"""CPU checks of the C-ABI boundary: libmmo_ff.so loads (no GPU needed to load it),
exports every function include/mmo_ff.h declares, and the Python signature table
covers exactly those functions.  No compute calls are made here."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmo_ff.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mmo_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for f in ("mmo_ctx_create", "mmo_set_receptor", "mmo_ene_inter", "mmo_ene_intra",
              "mmo_ene_inter_interp", "mmo_batch_score_dev", "mmo_scan_dev"):
        assert f in names


def test_library_exports_every_declared_symbol():
    from mmo_amd import _lib
    so = ct.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(so, n)]
    assert not missing, missing


def test_python_signatures_match_header():
    from mmo_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


def test_version_string_without_gpu():
    from mmo_amd._lib import lib
    assert lib.mmo_version().decode().startswith("mmo_ff")


def test_ctx_create_reports_errors_cleanly():
    """Without a GPU (this container) context creation fails with MMO_ERR_DEVICE and a
    message; with a GPU it succeeds.  Either way there is no crash and no fallback."""
    import torch
    from mmo_amd import ff
    from mmo_amd._lib import MMOError
    if torch.cuda.is_available():
        ff.Context(0)
    else:
        with pytest.raises(MMOError) as e:
            ff.Context(0)
        assert e.value.code == -2


def test_flags_and_modes_match_header():
    from mmo_amd import _lib
    text = open(HEADER).read()
    for name, val in (("MMO_GLOBAL", _lib.GLOBAL), ("MMO_SHIFTED", _lib.SHIFTED),
                      ("MMO_INTERP", _lib.INTERP), ("MMO_F_INTRA", _lib.F_INTRA),
                      ("MMO_F_GRAD", _lib.F_GRAD), ("MMO_ERR_OUT_OF_GRID", _lib.ERR_OUT_OF_GRID),
                      ("MMO_ERR_RANGE", _lib.ERR_RANGE)):
        m = re.search(r"\b%s\s*=\s*(-?\d+)" % name, text)
        assert m and int(m.group(1)) == val, name
