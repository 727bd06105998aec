"""ctypes binding of libjanc_amd.so (declarations: include/janc_step.h).

The library is the only compute path: there is no CPU fallback.  Loading it
fails loudly when it has not been built (``python -m janc_amd.build``).
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjanc_amd.so")

JH_OK, JH_ERR_ARG, JH_ERR_MECH, JH_ERR_HIP = 0, 1, 2, 3
PREC_FP64, PREC_MIXED = 0, 1
BC_KINDS = {"periodic": 0, "zero_gradient": 1, "slip_wall": 2, "no_slip_wall": 3,
            "pressure_outlet": 4, "user": 5, "injection_inlet": 6}
SIDES = ("left", "right", "bottom", "top")
# flux_config values (enum jh_solver_type / jh_reconstruction / jh_riemann_solver)
SOLVER_TYPES = {"flux_splitting": 0, "godunov": 1}
SPLIT_METHODS = {"LF": 0, "SW": 1}
RECONSTRUCTIONS = {"WENO5_JS": 0, "MUSCL": 1}
RIEMANN_SOLVERS = {"HLLC": 0}
KERNEL_KINDS = ("reduce", "bc", "stage", "chem")
# temperature-solve stop (enum jh_halley_stop)
HALLEY_STOP = {"per_cell": 0, "grid": 1}
# health flag bits (enum jh_health_flag)
HEALTH_FLAGS = {"nonfinite": 1, "negative_rho": 2, "halley_noconv": 4, "singular_pivot": 8,
                "halley_regress": 16, "nonpositive_T": 32}


def health_names(flags):
    """Names of the health flag bits set in ``flags``."""
    return [k for k, b in HEALTH_FLAGS.items() if flags & b]

_c_int, _c_double, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
# name -> (restype, argtypes); every symbol the public header declares
SIGNATURES = {
    "jh_mech_const_count": (_c_int, []),
    "jh_mech_digest": (ctypes.c_char_p, []),
    "jh_num_species": (_c_int, []),
    "jh_ctx_create": (_c_int, [_c_int, _c_int, _c_double, _c_double, ctypes.POINTER(_c_int),
                               ctypes.POINTER(_c_double), _c_int, ctypes.POINTER(_c_double),
                               _c_int, _vp, ctypes.POINTER(_vp)]),
    "jh_ctx_destroy": (_c_int, [_vp]),
    "jh_ctx_set_bc_param": (_c_int, [_vp, _c_int, _c_double]),
    "jh_ctx_set_stream": (_c_int, [_vp, _vp]),
    "jh_ctx_set_inlet": (_c_int, [_vp, _c_double, ctypes.POINTER(_c_double), _c_int]),
    "jh_ctx_set_tile": (_c_int, [_vp, _c_int]),
    "jh_ctx_set_flux": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int]),
    "jh_ctx_get_tile": (_c_int, [_vp]),
    "jh_ctx_set_halley_stop": (_c_int, [_vp, _c_int]),
    "jh_ctx_get_halley_stop": (_c_int, [_vp]),
    "jh_ctx_set_trace": (_c_int, [_vp, _vp, _c_int]),
    "jh_ctx_set_clip_branches": (_c_int, [_vp, _vp, _c_int, _c_int]),
    "jh_tile_variants": (_c_int, []),
    "jh_tile_info": (_c_int, [_c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                              ctypes.POINTER(_c_int)]),
    "jh_last_error": (ctypes.c_char_p, [_vp]),
    "jh_advance": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_double, _c_double, _c_int, _vp]),
    "jh_rhs": (_c_int, [_vp, _vp, _vp, _c_double, _vp]),
    "jh_update_aux": (_c_int, [_vp, _vp, _vp, _vp]),
    "jh_reaction_source": (_c_int, [_vp, _vp, _vp, _c_double, _vp]),
    "jh_cfl_dt": (_c_int, [_vp, _vp, _vp, _c_double, _vp]),
    "jh_lf_alpha": (_c_int, [_vp, _vp, _vp, _vp]),
    "jh_step_vjp": (_c_int, [_vp, _vp, _vp, _c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jh_step_vjp_known": (_c_int, [_vp, _vp, _vp, _c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jh_step_res_count": (_c_int, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "jh_step_keep": (_c_int, [_vp, _vp, _vp, _c_double, _vp, _vp, _vp]),
    "jh_step_vjp_res": (_c_int, [_vp, _vp, _vp, _c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "jh_ghost_strip": (_c_int, [_vp, _c_int, ctypes.POINTER(_vp), ctypes.POINTER(_c_int)]),
    "jh_set_ghost_strip": (_c_int, [_vp, _c_int, _vp]),
    "jh_step_begin": (_c_int, [_vp, _vp, _vp, _c_double, _c_double]),
    "jh_step_stage": (_c_int, [_vp, _c_int, _vp, _vp, _vp, _vp, _vp, _c_double, _c_double]),
    "jh_step_finish": (_c_int, [_vp, _vp, _vp, _c_double, _c_double, _vp]),
    "jh_health": (_c_int, [_vp, ctypes.POINTER(ctypes.c_uint)]),
    "jh_timing_enable": (_c_int, [_vp, _c_int]),
    "jh_timing_read": (_c_int, [_vp, ctypes.POINTER(_c_double), ctypes.POINTER(_c_int)]),
    "jh_batch_create": (_c_int, [_c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                 ctypes.POINTER(_c_double), ctypes.POINTER(_c_double),
                                 ctypes.POINTER(_c_int), ctypes.POINTER(_c_double), _c_int,
                                 ctypes.POINTER(_c_double), _c_int, _vp, ctypes.POINTER(_vp)]),
    "jh_batch_destroy": (_c_int, [_vp]),
    "jh_batch_bind": (_c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "jh_batch_set_stream": (_c_int, [_vp, _vp]),
    "jh_batch_set_tile": (_c_int, [_vp, _c_int]),
    "jh_batch_set_inlet": (_c_int, [_vp, _c_int, _c_double, ctypes.POINTER(_c_double), _c_int]),
    "jh_batch_get_tile": (_c_int, [_vp]),
    "jh_batch_advance": (_c_int, [_vp, _c_double, _c_double, _c_int, _vp, _c_int]),
    "jh_batch_health": (_c_int, [_vp, ctypes.POINTER(ctypes.c_uint)]),
    "jh_batch_last_error": (ctypes.c_char_p, [_vp]),
    "jh_blocks_create": (_c_int, [_c_int, _c_int, _c_double, _c_double, _c_int,
                                  ctypes.POINTER(_c_double), _c_int, _vp, ctypes.POINTER(_vp)]),
    "jh_blocks_destroy": (_c_int, [_vp]),
    "jh_blocks_fill_ghosts": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp,
                                       ctypes.c_longlong, ctypes.c_longlong, ctypes.c_longlong,
                                       _vp, _vp, _vp, _vp]),
    "jh_blocks_set_stream": (_c_int, [_vp, _vp]),
    "jh_blocks_last_error": (ctypes.c_char_p, [_vp]),
    "jh_blocks_flux": (_c_int, [_vp, _c_int, _vp, _c_double, _vp]),
    "jh_blocks_stage": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_double, _vp]),
    "jh_blocks_stage_noalloc": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _c_double, _vp]),
    "jh_blocks_chem": (_c_int, [_vp, _c_int, _vp, _c_double]),
    "jh_blocks_health": (_c_int, [_vp, ctypes.POINTER(ctypes.c_uint)]),
}

_lib = None


def load():
    """Load the step library once; raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("janc_amd: %s is not built; run `python -m janc_amd.build`"
                           % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class JancError(RuntimeError):
    pass


def check(rc, ctx=None):
    if rc != JH_OK:
        msg = load().jh_last_error(ctx)
        raise JancError("janc_amd error %d: %s" % (rc, msg.decode() if msg else ""))
