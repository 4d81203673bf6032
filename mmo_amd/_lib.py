"""ctypes loader for libmmo_ff.so (the gfx950 scoring library).

The library is built in-tree by ``python -m mmo_amd.build`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared object is missing
or a symbol is absent, importing fails with an error.
"""
import ctypes as ct
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmmo_ff.so")

OK = 0
ERR_BAD_ARG, ERR_DEVICE, ERR_NOMEM, ERR_OUT_OF_GRID = -1, -2, -3, -4
ERR_NO_RECEPTOR, ERR_NO_GRID, ERR_FORK, ERR_RANGE = -5, -6, -7, -8
GLOBAL, SHIFTED, INTERP = 0, 1, 2
F_INTER_SHIFTED, F_INTER_GLOBAL, F_INTER_INTERP = 1, 2, 4
F_INTRA, F_GRAD, F_COMPONENTS = 8, 16, 32
SCAN_ROI, SCAN_CENTER_CLASH, SCAN_ATOM_CLASH = 1, 2, 4
MC_TWEAK_RBONDS, MC_HARD_ROI, MC_INTRA_UFF, MC_METROPOLIS_ALWAYS, MC_NO_FLIP, MC_FP64 = 1, 2, 4, 8, 16, 32

_d, _i, _p = ct.c_double, ct.c_int, ct.c_void_p
_pp = ct.POINTER(ct.c_void_p)

# name -> (restype, argtypes); mirrors include/mmo_ff.h
SIGNATURES = {
    "mmo_version": (ct.c_char_p, []),
    "mmo_last_error": (ct.c_char_p, [_p]),
    "mmo_ctx_create": (_i, [_i, _pp]),
    "mmo_ctx_destroy": (None, [_p]),
    "mmo_set_uff_table": (_i, [_p, _p, _p, _p, _i]),
    "mmo_set_receptor": (_i, [_p, _p, _p, _p, _p, _p, _i, _p]),
    "mmo_set_grids": (_i, [_p, _p, _i, _d, _i, _i, _i, _p]),
    "mmo_grid_compute": (_i, [_p, _p, _p, _i, _d, _i, _i, _i, _p, _p, _d, _d, _p]),
    "mmo_ene_inter": (_i, [_p, _i, _p, _p, _p, _p, _p, _i, _p, _p, _p]),
    "mmo_ene_inter_interp": (_i, [_p, _p, _p, _p, _p, _i, _p]),
    "mmo_ene_intra": (_i, [_p, _p, _p, _p, _p, _p, _p, _i, _p]),
    "mmo_ene_inter_grad": (_i, [_p, _i, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p]),
    "mmo_ene_intra_grad": (_i, [_p, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p]),
    "mmo_ene_inter_interp_grad": (_i, [_p, _p, _p, _p, _p, _i, _p, _p, _p, _p]),
    "mmo_ligand_add": (_i, [_p, _p, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _i, _p, _p]),
    "mmo_ligand_count": (_i, [_p]),
    "mmo_ligand_natoms": (_i, [_p, _i]),
    "mmo_batch_create": (_i, [_p, _p, _i, _pp]),
    "mmo_batch_destroy": (None, [_p]),
    "mmo_batch_natoms": (_i, [_p]),
    "mmo_batch_atom_offsets": (_i, [_p, _p]),
    "mmo_batch_dof_offsets": (_i, [_p, _p]),
    "mmo_batch_score_dev": (_i, [_p, _i, _p, _p, _p, _p]),
    "mmo_batch_score": (_i, [_p, _i, _p, _p, _p]),
    "mmo_batch_score_dofs_dev": (_i, [_p, _i, _p, _p, _p, _p, _p, _p]),
    "mmo_scan_dev": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _p]),
    "mmo_set_vdw_mask": (_i, [_p, _p, _d, _i, _i, _i, _p]),
    "mmo_get_vdw_mask": (_i, [_p, _p]),
    "mmo_desolv_setup": (_i, [_p, _p, _d, _i, _i, _i, _p, _p, _d]),
    "mmo_get_solvent_shell": (_i, [_p, _p]),
    "mmo_get_desolv_contribs": (_i, [_p, _p]),
    "mmo_desolv_penalty": (_i, [_p, _p, _p, _p, _i, _p, _p]),
    "mmo_scan_filter_dev": (_i, [_p, _i, _p, _i, _p, _i, _p, _d, _p, _i, _p, _p]),
    "mmo_scan_filtered_dev": (_i, [_p, _i, _i, _p, _i, _p, _i, _p, _d, _p, _i, _p, _p, _p]),
    "mmo_topk_dev": (_i, [_p, _p, ct.c_int64, _i, _p, _p, _p]),
    "mmo_topk_dev_noalloc": (_i, [_p, _p, ct.c_int64, _i, _p, _p, _p]),
    "mmo_mc_create": (_i, [_p, _p, _i, _p, _p, _i, _i, _p, _d, _d, ct.c_uint64, _i, _pp]),
    "mmo_mc_run": (_i, [_p, _i, _p]),
    "mmo_mc_run_graph": (_i, [_p, _i, _i, _p]),
    "mmo_mc_run_per_ligand": (_i, [_p, _i, _p]),
    "mmo_mc_set_minimizer": (_i, [_p, _i, _i, _d, _d]),
    "mmo_mc_min_evals": (_i, [_p, _p, _p]),
    "mmo_mc_natoms": (_i, [_p]),
    "mmo_mc_get": (_i, [_p, _p, _p, _p, _p, _p, _p]),
    "mmo_mc_coords": (_i, [_p, _p, _p, _p]),
    "mmo_mc_destroy": (None, [_p]),
    "mmo_min_create": (_i, [_p, _i, _i, _p, _p, _d, _d, _pp]),
    "mmo_min_run_dev": (_i, [_p, _p, _i, _p, _p]),
    "mmo_min_get": (_i, [_p, _p, _p]),
    "mmo_min_destroy": (None, [_p]),
    "mmo_last_kernel_ms": (_d, [_p, ct.c_char_p]),
    "mmo_fixed_point_bits": (_i, [_p]),
    "mmo_tune_set": (_i, [_p, ct.c_char_p, _d]),
    "mmo_tune_get": (_d, [_p, ct.c_char_p]),
    "mmo_molset_from_mol2": (_i, [ct.c_char_p, ct.c_size_t, _pp]),
    "mmo_molset_from_pqrs": (_i, [ct.c_char_p, ct.c_size_t, _i, _pp]),
    "mmo_molset_count": (_i, [_p]),
    "mmo_molset_name": (ct.c_char_p, [_p, _i]),
    "mmo_molset_dims": (_i, [_p, _i, _p, _p, _p]),
    "mmo_molset_atoms": (_i, [_p, _i, _p, _p, _p, _p, _p, _p]),
    "mmo_molset_topology": (_i, [_p, _i, _p, _p, _p, _p, _p]),
    "mmo_molset_pqrs": (_i, [_p, ct.c_char_p, ct.c_size_t, ct.POINTER(ct.c_size_t)]),
    "mmo_molset_destroy": (None, [_p]),
}


class MMOError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mmo_ff error %d: %s" % (code, msg))
        self.code = code


def load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libmmo_ff.so is not built (%s); run python -m mmo_amd.build" % LIB_PATH)
    # PyTorch ships its own HIP runtime (libamdhip64.so.7).  Loading torch first makes
    # the dynamic linker bind this library's libamdhip64.so.7 dependency to that same
    # runtime: one HIP runtime per process, so device pointers and streams are shared.
    # Loaded the other way round, torch would bring up a second runtime that finds no
    # GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ct.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)   # AttributeError if the symbol is missing
        f.restype = res
        f.argtypes = args
    return lib


lib = load()
