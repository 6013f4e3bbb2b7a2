"""ctypes declarations of the C ABI in include/lepic.h (liblepic_hip.so).

The library is loaded from the package directory; if it is missing the
loader raises, so no caller can fall back to anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblepic_hip.so")

LEPIC_OK, LEPIC_ERR_ARG, LEPIC_ERR_UNSUPPORTED, LEPIC_ERR_HIP = 0, -1, -2, -3
LEPIC_ERR_CAPACITY, LEPIC_ERR_DIVERGED, LEPIC_ERR_MCC, LEPIC_ERR_STATE = -4, -5, -6, -7
FP64, MIXED = 0, 1
MCC_CHECK, MCC_PROD = 0, 1
START_PREV, START_COLD, START_GIVEN, START_NET = 0, 1, 2, 3
GRID_PHI, GRID_RHS, GRID_E, GRID_NP, GRID_NPRED, GRID_SINK, GRID_SRC = range(7)

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_bp = C.POINTER(C.c_int8)


class GridDesc(C.Structure):
    _fields_ = [("n1", C.c_int32), ("n2", C.c_int32), ("h1", C.c_double), ("h2", C.c_double),
                ("zmax", C.c_double), ("dt", C.c_double), ("eps0", C.c_double), ("qe", C.c_double),
                ("ng", C.c_int32), ("ngrid", C.c_int32), ("nspecies", C.c_int32),
                ("precision", C.c_int32), ("flag_B_pos", C.c_int32), ("tag_neg", C.c_int32),
                ("track_z", C.c_int32), ("mg_tail_nodes", C.c_int32)]


class SpeciesDesc(C.Structure):
    _fields_ = [("charge", C.c_double), ("mass", C.c_double), ("Nm", C.c_double),
                ("capacity", C.c_int64)]


class SolveInfo(C.Structure):
    _fields_ = [("vcycles", C.c_int32), ("status", C.c_int32), ("res", C.c_double),
                ("ktot", C.c_double), ("res0", C.c_double), ("start_used", C.c_int32),
                ("nlog", C.c_int32), ("res0_alt", C.c_double)]


class MccInfo(C.Structure):
    _fields_ = [("Nc", C.c_int64 * 8), ("nu_max", C.c_double * 8), ("events", C.c_int64 * 64),
                ("created", C.c_int64 * 8), ("missed", C.c_int64), ("seed", C.c_int32),
                ("reserved", C.c_int32)]


class Reactions(C.Structure):
    _fields_ = [("ntype_all", C.c_int32), ("ncol", C.c_int32), ("npt", C.c_int32),
                ("ns_coll", C.c_int32), ("n_cell", C.c_int32), ("n_rg", C.c_int32),
                ("iyl_rg", C.c_int32), ("iyr_rg", C.c_int32), ("ixl_rg", C.c_int32 * 5),
                ("ixr_rg", C.c_int32 * 5), ("flag_ionrep", C.c_int32), ("reserved", C.c_int32),
                ("sig_Er", _dp), ("sig", _dp), ("Eth", _dp), ("sigv_mx", _dp),
                ("col_info", _ip), ("p_ncol", _ip), ("sig_list", _ip),
                ("charge", _dp), ("mass", _dp), ("Nm", _dp), ("vt0", _dp), ("np_mx", _dp),
                ("nu_uplim", _dp)]


class Heating(C.Structure):
    _fields_ = [("flag_c", C.c_int32), ("flag_gpp", C.c_int32), ("eheat_type", C.c_int32),
                ("n_icp", C.c_int32), ("xa", C.c_double), ("dr", C.c_double), ("yl_pow", C.c_double),
                ("yr_pow", C.c_double), ("nudt", C.c_double), ("Pabs_cor", C.c_double), ("vt", _dp)]


class Emitter(C.Structure):
    _fields_ = [("jne", C.c_double), ("Lg", C.c_double), ("Lh", C.c_double), ("xg1", C.c_double),
                ("THm", C.c_double), ("chmfrd_array", C.c_double * 4), ("n_holes", C.c_int32),
                ("n_icp", C.c_int32), ("ptype_fx", C.c_int32), ("tag_b", C.c_int32), ("flag_chmfrd", C.c_int32),
                ("ns_Hm", C.c_int32)]


class EmitterInfo(C.Structure):
    _fields_ = [("Nh", C.c_int64), ("created", C.c_int64 * 8), ("cnt_b", C.c_int64 * 2),
                ("P_loss4", C.c_double * 8), ("Lgrd", C.c_double), ("dNh", C.c_double), ("seed", C.c_int32),
                ("reserved", C.c_int32)]


class Injection(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("opt_inj", "flag_c", "flag_nmn", "flag_Tp", "flag_diffsrc", "ns_inj")] + \
               [(k, C.c_double) for k in ("I_inj", "xl_pow", "xr_pow", "yl_pow", "yr_pow", "xa", "dr", "THm",
                                          "th_B")] + [("ni0", _dp), ("vt0", _dp)]


class InjectionInfo(C.Structure):
    _fields_ = [("injected", C.c_int64 * 8), ("P_loss4", C.c_double * 8), ("P_loss1", C.c_double * 8),
                ("seed", C.c_int32), ("reserved", C.c_int32)]


class StepArgs(C.Structure):
    _fields_ = [("eps", C.c_double), ("ncycl", C.c_int32), ("nsort", C.c_int32),
                ("ns_coll", C.c_int32), ("mcc_mode", C.c_int32), ("seed", C.c_uint64),
                ("iseed", C.c_int32), ("start", C.c_int32), ("solve", SolveInfo),
                ("ms_rhs_poisson_e", C.c_double), ("ms_sort", C.c_double), ("ms_mcc", C.c_double),
                ("ms_push", C.c_double), ("heat", C.POINTER(Heating)), ("ns_heat", C.c_int32),
                ("heat_mode", C.c_int32), ("P_heat", C.c_double), ("emit", C.POINTER(Emitter)),
                ("emitted", C.c_int64), ("inj", C.POINTER(Injection)), ("injected", C.c_int64)]


_ctxp = C.c_void_p

# name -> (restype, argtypes); every symbol declared in include/lepic.h
SIGNATURES = {
    "lepic_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(SpeciesDesc), C.POINTER(_ctxp)]),
    "lepic_destroy": (C.c_int, [_ctxp]),
    "lepic_last_error": (C.c_char_p, [_ctxp]),
    "lepic_sync": (C.c_int, [_ctxp]),
    "lepic_stream": (C.c_int, [_ctxp, C.POINTER(C.c_void_p)]),
    "lepic_kernel_timing": (C.c_int, [_ctxp, C.c_int32]),
    "lepic_net_create": (C.c_int, [_ctxp, C.c_int32, C.c_uint64]),
    "lepic_net_recording": (C.c_int, [_ctxp, C.c_int32]),
    "lepic_net_samples": (C.c_int, [_ctxp, _ip]),
    "lepic_net_copy_samples": (C.c_int, [_ctxp, _ctxp, C.POINTER(C.c_int32)]),
    "lepic_net_copy": (C.c_int, [_ctxp, _ctxp]),
    "lepic_net_policy": (C.c_int, [_ctxp, C.c_float, C.c_int32]),
    "lepic_backup_write": (C.c_int, [C.c_char_p, C.c_double, C.c_int32, C.c_double, C.c_int32, C.c_int32, _ip,
                                     C.POINTER(_dp)]),
    "lepic_backup_header": (C.c_int, [C.c_char_p, C.c_int32, _dp, _ip, _dp, _ip, _ip, C.c_int32]),
    "lepic_backup_read_species": (C.c_int, [C.c_char_p, C.c_int32, C.c_int32, _dp, C.c_int64,
                                            C.POINTER(C.c_int64)]),
    "lepic_moments_accumulate": (C.c_int, [_ctxp, C.c_int32]),
    "lepic_moments_get": (C.c_int, [_ctxp, _dp, _dp, _ip]),
    "lepic_moments_reset": (C.c_int, [_ctxp]),
    "lepic_species_totals": (C.c_int, [_ctxp, C.c_int32, _dp]),
    "lepic_field_energy": (C.c_int, [_ctxp, _dp]),
    "lepic_save_particles": (C.c_int, [_ctxp, C.c_char_p, C.c_double, C.c_int32, C.c_double]),
    "lepic_load_particles": (C.c_int, [_ctxp, C.c_char_p, _dp, _ip, _dp]),
    "lepic_net_train": (C.c_int, [_ctxp, C.c_int32, C.c_int32, C.c_float, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float)]),
    "lepic_net_set_params": (C.c_int, [_ctxp, C.POINTER(C.c_float), C.c_int64]),
    "lepic_net_get_params": (C.c_int, [_ctxp, C.POINTER(C.c_float), C.c_int64]),
    "lepic_net_forward": (C.c_int, [_ctxp, C.POINTER(C.c_float), C.c_int32, C.c_int32, C.c_int32,
                                    C.POINTER(C.c_float)]),
    "lepic_net_loss_grad": (C.c_int, [_ctxp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float, _dp,
                                      C.POINTER(C.c_float), C.c_int64]),
    "lepic_get_kernel_ms": (C.c_int, [_ctxp, C.c_int32, C.POINTER(C.c_float)]),
    "lepic_set_boundary": (C.c_int, [_ctxp, _ip, _dp]),
    "lepic_set_bfield": (C.c_int, [_ctxp, C.c_int32, C.c_int32, C.c_double, C.c_double, _dp]),
    "lepic_upload_particles": (C.c_int, [_ctxp, C.c_int32, _dp, C.c_int64]),
    "lepic_download_particles": (C.c_int, [_ctxp, C.c_int32, _dp, C.c_int64, C.POINTER(C.c_int64)]),
    "lepic_download_particles_range": (C.c_int, [_ctxp, C.c_int32, C.c_int64, C.c_int64, _dp]),
    "lepic_get_count": (C.c_int, [_ctxp, C.c_int32, C.POINTER(C.c_int64)]),
    "lepic_set_flag_dead": (C.c_int, [_ctxp, C.c_int32, _bp, C.c_int64]),
    "lepic_get_flag_dead": (C.c_int, [_ctxp, C.c_int32, _bp, C.c_int64]),
    "lepic_get_grid": (C.c_int, [_ctxp, C.c_int32, C.c_int32, _dp]),
    "lepic_set_grid": (C.c_int, [_ctxp, C.c_int32, C.c_int32, _dp]),
    "lepic_part_mover": (C.c_int, [_ctxp, C.c_int32, C.c_int32]),
    "lepic_charge_deposition": (C.c_int, [_ctxp, C.c_int32]),
    "lepic_get_mover_tally": (C.c_int, [_ctxp, C.c_int32, _dp, _dp, _dp]),
    "lepic_reset_tallies": (C.c_int, [_ctxp]),
    "lepic_calc_rhs": (C.c_int, [_ctxp]),
    "lepic_calc_efield": (C.c_int, [_ctxp]),
    "lepic_iterative_solver": (C.c_int, [_ctxp, C.c_double, C.c_int32, C.c_int32,
                                         C.POINTER(SolveInfo), _ip, C.c_int32]),
    "lepic_part_sorting": (C.c_int, [_ctxp, C.c_int32, _ip]),
    "lepic_part_sorting_noalloc": (C.c_int, [_ctxp, C.c_int32, _ip]),
    "lepic_dens_red": (C.c_int, [_ctxp, _dp]),
    "lepic_set_reactions": (C.c_int, [_ctxp, C.POINTER(Reactions)]),
    "lepic_collisions": (C.c_int, [_ctxp, C.c_int32, C.c_int32, C.c_uint64, _ip,
                                   C.POINTER(MccInfo)]),
    "lepic_get_cnt_col": (C.c_int, [_ctxp, _dp]),
    "lepic_get_mcc_ploss": (C.c_int, [_ctxp, _dp]),
    "lepic_flux_emitter": (C.c_int, [_ctxp, C.POINTER(Emitter), C.c_int32, C.c_int32, C.c_uint64, _ip,
                                     C.POINTER(EmitterInfo)]),
    "lepic_part_injection": (C.c_int, [_ctxp, C.POINTER(Injection), C.POINTER(C.c_int64), C.c_int32, C.c_int32,
                                       C.c_uint64, _ip, C.c_int32, C.POINTER(InjectionInfo)]),
    "lepic_eheating": (C.c_int, [_ctxp, C.POINTER(Heating), C.c_int32, C.c_uint64, C.c_int32, _ip, _dp]),
    "lepic_load_uniform": (C.c_int, [_ctxp, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.c_double, C.c_uint64]),
    "lepic_step": (C.c_int, [_ctxp, C.c_int32, C.POINTER(StepArgs)]),
}

_lib = None


def load(path: str = LIB_PATH):
    """Load liblepic_hip.so and declare every entry point; raises if absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"liblepic_hip.so is not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib
