"""ctypes binding of libbamrex_hip.so (include/bamrex_hip.h).

The product path is the HIP library; there is no CPU fallback.  Loading fails
loudly when the library has not been built.  Importing torch in the same
process before this library is loaded would bind it to torch's bundled HIP
runtime (same soname), so this package never imports torch.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BAMREX_LIB") or os.path.join(_HERE, "libbamrex_hip.so")

BX_OK = 0
METHOD = {"muscl_hancock": 0, "hllc": 1, "slic": 2, "force": 3, "lax_friedrichs": 4}
LIMITER = {"minbee": 0, "van_leer": 1, "van_albada": 2, "superbee": 3}
BC = {"transmissive": 0, "periodic": 1, "reflective": 2}


class bx_params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("num_method", ctypes.c_int),
        ("adiabatic", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("limiter", ctypes.c_int),
        ("limiter_indices", ctypes.c_int * 5),
        ("limit_on_primv", ctypes.c_int * 5),
    ]


class bx_box(ctypes.Structure):
    _fields_ = [
        ("lo", ctypes.c_int * 3),
        ("hi", ctypes.c_int * 3),
        ("offset", ctypes.c_int64),
    ]


class bx_plot_level(ctypes.Structure):
    _fields_ = [("boxes", ctypes.c_void_p), ("nbox", ctypes.c_int), ("ng", ctypes.c_int),
                ("U", ctypes.c_void_p), ("domain_lo", ctypes.c_int * 3), ("domain_hi", ctypes.c_int * 3),
                ("dx", ctypes.c_double * 3), ("level_steps", ctypes.c_int)]


class bx_chk_level(ctypes.Structure):
    _fields_ = [("boxes", ctypes.c_void_p), ("nbox", ctypes.c_int), ("ng", ctypes.c_int),
                ("U", ctypes.c_void_p), ("domain_lo", ctypes.c_int * 3), ("domain_hi", ctypes.c_int * 3),
                ("dx", ctypes.c_double * 3), ("level_steps", ctypes.c_int), ("dt", ctypes.c_double)]


class bx_chk_info(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("nlev", ctypes.c_int), ("ref_ratio", ctypes.c_int),
                ("time", ctypes.c_double), ("prob_lo", ctypes.c_double * 3),
                ("prob_hi", ctypes.c_double * 3)]


class bx_chk_level_info(ctypes.Structure):
    _fields_ = [("nbox", ctypes.c_int), ("ng", ctypes.c_int), ("level_steps", ctypes.c_int),
                ("dt", ctypes.c_double), ("domain_lo", ctypes.c_int * 3),
                ("domain_hi", ctypes.c_int * 3), ("dx", ctypes.c_double * 3),
                ("ndoubles", ctypes.c_int64), ("checksum", ctypes.c_uint64)]


class bx_geometry(ctypes.Structure):
    _fields_ = [
        ("domain_lo", ctypes.c_int * 3),
        ("domain_hi", ctypes.c_int * 3),
        ("bc_lo", ctypes.c_int * 3),
        ("bc_hi", ctypes.c_int * 3),
        ("ng", ctypes.c_int),
    ]


class bx_tuning(ctypes.Structure):
    _fields_ = [
        ("block_x", ctypes.c_int),
        ("block_t", ctypes.c_int),
        ("seg_k", ctypes.c_int),
        ("grid_cap", ctypes.c_int),
    ]


class BamrexError(RuntimeError):
    pass


_lib = None

P = ctypes.c_void_p
DP = ctypes.POINTER(ctypes.c_double)
_SIGS = {
    "bx_create": (ctypes.c_int, [ctypes.POINTER(bx_params), ctypes.POINTER(P)]),
    "bx_destroy": (None, [P]),
    "bx_set_level": (ctypes.c_int, [P, ctypes.POINTER(bx_geometry), ctypes.POINTER(bx_box), ctypes.c_int]),
    "bx_pack_offsets": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.POINTER(bx_box), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "bx_sweep": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_double, DP, P, P, P]),
    "bx_flux_box": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, DP, P, P, ctypes.c_int64, P]),
    "bx_update_box": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, DP, P, P, ctypes.c_int64,
                                     P, P]),
    "bx_fill_boundary": (ctypes.c_int, [P, P, P]),
    "bx_max_wave_speed": (ctypes.c_int, [P, P, DP, P]),
    "bx_est_dt": (ctypes.c_int, [P, P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_double, DP, P]),
    "bx_advance": (ctypes.c_int, [P, ctypes.c_double, DP, P, P, P]),
    "bx_run": (ctypes.c_int, [P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, P, P,
                              ctypes.c_int, P, DP, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "bx_get_tuning": (ctypes.c_int, [P, ctypes.POINTER(bx_tuning)]),
    "bx_set_tuning": (ctypes.c_int, [P, ctypes.POINTER(bx_tuning)]),
    "bx_autotune": (ctypes.c_int, [P, DP, ctypes.c_double, P, P, P, DP]),
    "bx_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    "bx_malloc": (ctypes.c_int, [ctypes.POINTER(P), ctypes.c_size_t]),
    "bx_free": (ctypes.c_int, [P]),
    "bx_memcpy_htod": (ctypes.c_int, [P, P, ctypes.c_size_t]),
    "bx_memcpy_dtoh": (ctypes.c_int, [P, P, ctypes.c_size_t]),
    "bx_memcpy_dtod": (ctypes.c_int, [P, P, ctypes.c_size_t]),
    "bx_memset": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_size_t]),
    "bx_stream_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "bx_stream_destroy": (ctypes.c_int, [P]),
    "bx_stream_sync": (ctypes.c_int, [P]),
    "bx_event_create": (ctypes.c_int, [ctypes.POINTER(P)]),
    "bx_event_destroy": (ctypes.c_int, [P]),
    "bx_event_record": (ctypes.c_int, [P, P]),
    "bx_event_elapsed_ms": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_float)]),
    "bx_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "bx_copy_kernel": (ctypes.c_int, [P, P, ctypes.c_size_t, P]),
    "bx_last_error": (ctypes.c_char_p, []),
    # pressure Helmholtz solve + learned preconditioner
    "bx_helm_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.POINTER(P)]),
    "bx_helm_destroy": (None, [P]),
    "bx_helm_set_coeffs": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                          P, P, P, P, P, P]),
    "bx_helm_set_options": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int]),
    "bx_helm_apply": (ctypes.c_int, [P, P, P, P]),
    "bx_helm_residual": (ctypes.c_int, [P, P, P, P, P]),
    "bx_helm_smooth": (ctypes.c_int, [P, P, P, ctypes.c_int, P]),
    "bx_helm_precond": (ctypes.c_int, [P, P, P, P]),
    "bx_helm_attach_cnn": (ctypes.c_int, [P, P, ctypes.c_double, ctypes.c_int]),
    "bx_helm_set_graph": (ctypes.c_int, [P, ctypes.c_int]),
    "bx_helm_set_fused": (ctypes.c_int, [P, ctypes.c_int]),
    "bx_helm_solve": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.c_double, ctypes.c_int, P, P]),
    "bx_write_checkpoint": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                           DP, DP, ctypes.c_double, P]),
    "bx_read_checkpoint_info": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(bx_chk_info), P,
                                               ctypes.c_int]),
    "bx_read_checkpoint_boxes": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P, ctypes.c_int]),
    "bx_read_checkpoint_level": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, P, ctypes.c_int64, P]),
    "bx_write_plotfile": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P, ctypes.c_int, DP, DP,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "bx_amr_create": (ctypes.c_int, [P, P, ctypes.POINTER(P)]),
    "bx_amr_destroy": (None, [P]),
    "bx_amr_info": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]),
    "bx_amr_avg_down": (ctypes.c_int, [P, P, P, P]),
    "bx_amr_fill_fine": (ctypes.c_int, [P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        P]),
    "bx_amr_advance": (ctypes.c_int, [P, ctypes.c_double, DP, ctypes.c_double, P, P, P, P, P, P]),
    "bx_amr_est_dt": (ctypes.c_int, [P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_double, P, P,
                                     ctypes.POINTER(ctypes.c_double), P]),
    "bx_amr_run": (ctypes.c_int, [P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_int, P, P, P, P, P,
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), P]),
    "bx_amr_create_levels": (ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(P)]),
    "bx_amr_advance_levels": (ctypes.c_int, [P, ctypes.c_double, DP, ctypes.c_double, P, P, P, P]),
    "bx_amr_est_dt_levels": (ctypes.c_int, [P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_double, P,
                                            ctypes.POINTER(ctypes.c_double), P]),
    "bx_amr_run_levels": (ctypes.c_int, [P, DP, ctypes.c_double, ctypes.c_double, ctypes.c_int, P, P, P,
                                         ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), P]),
    "bx_amr_set_graph": (ctypes.c_int, [P, ctypes.c_int]),
    "bx_helm_last_error": (ctypes.c_char_p, []),
    "bx_imex_pressure": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, P,
                                        P]),
    "bx_imex_advective_dt": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.POINTER(ctypes.c_double), P]),
    "bx_imex_peek_assemble": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                             P, P, P, P, P, P, P]),
    "bx_imex_peek_finish": (ctypes.c_int, [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           P, P]),
    "bx_imex_stage_assemble": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, P, P, P, P, P, P, P, P, P]),
    "bx_imex_stage_fluxes": (ctypes.c_int, [P, P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, P, P, P]),
    "bx_imex_rk_combine": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int, DP, DP,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, P, P,
                                          P]),
    "bx_imex_rk_pressure": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                           ctypes.c_int, P, P]),
    "bx_imex_rk_update": (ctypes.c_int, [P, ctypes.POINTER(P), ctypes.POINTER(P), ctypes.c_int, DP,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, P, P]),
    "bx_imex_shock_indicator": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, P, P, ctypes.POINTER(ctypes.c_double), P]),
    "bx_imex_stab_blend": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.c_int, P, P, P]),
    "bx_imex_picard_update": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                             ctypes.c_double, P, P, P, P, P]),
    "bx_imex_acoustic_dt": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.POINTER(ctypes.c_double), P]),
    "bx_imex_acoustic_dt_noalloc": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double, P, P]),
    "bx_cnn_create": (ctypes.c_int, [ctypes.c_uint, ctypes.POINTER(P)]),
    "bx_cnn_destroy": (None, [P]),
    "bx_cnn_set_params": (ctypes.c_int, [P, P, ctypes.c_int]),
    "bx_cnn_get_params": (ctypes.c_int, [P, P, ctypes.c_int]),
    "bx_cnn_save": (ctypes.c_int, [P, ctypes.c_char_p]),
    "bx_cnn_load": (ctypes.c_int, [P, ctypes.c_char_p]),
    "bx_cnn_set_operator": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_double, P, P, P, P]),
    "bx_cnn_infer": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_double, P]),
    "bx_cnn_net_forward": (ctypes.c_int, [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]),
    "bx_cnn_loss_grad": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_int,
                                        ctypes.c_int, DP, P, P]),
    "bx_cnn_train": (ctypes.c_int, [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_float, ctypes.c_uint, P, P]),
    "bx_cnn_last_error": (ctypes.c_char_p, []),
}

BX_SOLVE = {"mg": 0, "fgmres_mg": 1, "fgmres_cnn": 2, "mg_fused": 3}
BX_IMEX_ADVECTION = {"rusanov": 0, "muscl_rusanov": 1}
BX_IMEX_MAX_STAGES = 4
BX_IMEX_KE = {"ex": 0, "split": 1, "fi": 2, "conservative_kinetic": 3}
BX_IMEX_ENTHALPY = {"reuse_pressure": 0, "conservative": 1, "conservative_ex": 2}
BX_CNN_MODE = {"smoother": 0, "corrector": 1}
BX_CNN_NPARAMS = 4896


class bx_solve_stats(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("fell_back", ctypes.c_int),
        ("precond_applies", ctypes.c_int),
        ("rel_resid", ctypes.c_double),
        ("seconds", ctypes.c_double),
    ]


def exported_symbols():
    return list(_SIGS)


def lib():
    """Load libbamrex_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BamrexError(
            "libbamrex_hip.so is not built (%s); run __graft_entry__.build() "
            "or python bamrex_amd/build.py" % LIB_PATH)
    l = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    _lib = l
    return l


def check(rc, what=""):
    if rc != BX_OK:
        msg = lib().bx_last_error()
        raise BamrexError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))


def dvec3(v):
    a = (ctypes.c_double * 3)(*([float(x) for x in v] + [0.0] * (3 - len(v))))
    return a
