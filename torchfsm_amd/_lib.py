"""ctypes binding of libtfsm.so (the C ABI declared in include/tfsm.h).

The library is mandatory: importing this module raises if libtfsm.so is missing or cannot be
loaded, and every engine call raises on a non-zero status.  There is no CPU or torch fallback.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtfsm.so")

TFSM_OK = 0
TFSM_E_INVALID = -1
TFSM_E_STATE = -2
TFSM_E_DEVICE = -3
TFSM_E_OOM = -4

TERM_NONE, TERM_CONV, TERM_VC, TERM_NSP, TERM_KS, TERM_KS_KEEP_MEAN, TERM_CONS = 0, 1, 2, 3, 4, 5, 6
LIN_LAP, LIN_BIHARM, LIN_UNIT, LIN_DERIV = 1, 2, 3, 4
MAP_DEALIAS, MAP_DIV, MAP_GRAD, MAP_CURL, MAP_V2V, MAP_PRESSURE, MAP_CONSDIV = 0, 1, 2, 3, 4, 5, 6
ETDRK0, ETDRK1, ETDRK2, SETDRK1, SETDRK2, SETDRK3, SETDRK4 = range(7)
NOALLOC_SCRATCH_BYTES = 64 << 20  # TFSM_NOALLOC_SCRATCH_BYTES


class Geometry(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("length", ctypes.c_double * 3),
        ("batch", ctypes.c_int32),
        ("channels", ctypes.c_int32),
        ("length_per_sample", ctypes.POINTER(ctypes.c_double)),
    ]


class Terms(ctypes.Structure):
    _fields_ = [
        ("nonlinear", ctypes.c_int32),
        ("nonlinear_coef", ctypes.c_float),
        ("dealias_rate", ctypes.c_double),
        ("n_linear", ctypes.c_int32),
        ("linear_kind", ctypes.c_int32 * 4),
        ("linear_coef", ctypes.POINTER(ctypes.c_float)),
        ("coef_batch", ctypes.c_int32),
        ("linear_dim", ctypes.c_int32 * 4),
        ("linear_order", ctypes.c_int32 * 4),
    ]


VP = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64

# name -> (restype, argtypes); every symbol include/tfsm.h declares
SIGNATURES = {
    "tfsm_last_error": (ctypes.c_char_p, []),
    "tfsm_version": (ctypes.c_int, []),
    "tfsm_plan_create": (ctypes.c_int, [ctypes.POINTER(Geometry), ctypes.POINTER(Terms), VP, ctypes.POINTER(VP)]),
    "tfsm_plan_destroy": (ctypes.c_int, [VP]),
    "tfsm_plan_pitch": (I64, [VP]),
    "tfsm_plan_state_elems": (I64, [VP]),
    "tfsm_plan_workspace_bytes": (I64, [VP]),
    "tfsm_plan_inverse_scratch_bytes": (I64, [VP, I32]),
    "tfsm_plan_set_tuning": (ctypes.c_int, [VP, I32, I32, I32, I32, I32]),
    "tfsm_plan_set_streams": (ctypes.c_int, [VP, I32]),
    "tfsm_plan_set_persistent": (ctypes.c_int, [VP, I32]),
    "tfsm_plan_persistent": (I32, [VP]),
    "tfsm_plan_set_general": (ctypes.c_int, [VP, I32]),
    "tfsm_plan_general": (I32, [VP]),
    "tfsm_plan_get_tuning": (ctypes.c_int, [VP, VP]),
    "tfsm_plan_set_scheme": (ctypes.c_int, [VP, I32, ctypes.c_double, I32, ctypes.c_double, VP]),
    "tfsm_plan_coefficients": (ctypes.c_int, [VP, I32, VP, VP]),
    "tfsm_plan_set_sample": (ctypes.c_int, [VP, I32, VP, ctypes.c_double, VP]),
    "tfsm_plan_add_source": (ctypes.c_int, [VP, VP, I32, I32, ctypes.c_float, VP]),
    "tfsm_plan_add_force": (ctypes.c_int, [VP, VP, I32, I32, ctypes.c_float, VP]),
    "tfsm_forward": (ctypes.c_int, [VP, VP, VP, I32, VP]),
    "tfsm_inverse": (ctypes.c_int, [VP, VP, VP, I32, VP, VP]),
    "tfsm_c2c": (ctypes.c_int, [VP, VP, VP, I32, I32, VP]),
    "tfsm_half_to_full": (ctypes.c_int, [VP, VP, VP, I32, VP]),
    "tfsm_full_to_half": (ctypes.c_int, [VP, VP, VP, I32, VP]),
    "tfsm_step": (ctypes.c_int, [VP, VP, I32, VP, VP, I32, VP, VP]),
    "tfsm_step_noalloc": (ctypes.c_int, [VP, VP, I32, VP, VP, I32, VP]),
    "tfsm_step_varying": (ctypes.c_int, [VP, VP, VP, I32, VP, VP]),
    "tfsm_apply": (ctypes.c_int, [VP, VP, VP, VP, VP]),
    "tfsm_solve": (ctypes.c_int, [VP, VP, VP, VP]),
    "tfsm_rk_combine": (ctypes.c_int, [VP, VP, VP, VP, VP, I32, ctypes.c_float, VP]),
    "tfsm_spectral_map": (ctypes.c_int, [VP, I32, ctypes.c_float, VP, VP, VP]),
    "tfsm_pointwise_products": (ctypes.c_int, [VP, VP, VP, VP]),
    "tfsm_launch_pass": (ctypes.c_int, [VP, VP, VP, I32, I32, I32, VP]),
    "tfsm_launch_pass_timed": (ctypes.c_int, [VP, VP, VP, I32, I32, I32, VP, VP, VP]),
    "tfsm_plan_adjoint_workspace_bytes": (I64, [VP]),
    "tfsm_adjoint_prepare": (ctypes.c_int, [VP, I32, VP]),
    "tfsm_adjoint_step": (ctypes.c_int, [VP, VP, VP, VP, VP, VP, VP, VP]),
    "tfsm_graph_create": (ctypes.c_int, [VP, VP, I32, VP, VP, ctypes.POINTER(VP)]),
    "tfsm_graph_create_varying": (ctypes.c_int, [VP, VP, VP, I32, VP, VP, ctypes.POINTER(VP)]),
    "tfsm_graph_launch": (ctypes.c_int, [VP, VP]),
    "tfsm_graph_destroy": (ctypes.c_int, [VP]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libtfsm.so not found at {LIB_PATH}: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python torchfsm_amd/_build.py` (hipcc for gfx950 is required; there is no CPU fallback)"
        )
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


class EngineError(RuntimeError):
    pass


def check(rc, what=""):
    """Raise the exception type the reference raises for the same failure class."""
    if rc == TFSM_OK:
        return
    msg = lib.tfsm_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if rc == TFSM_E_INVALID:
        raise ValueError(msg)
    if rc == TFSM_E_OOM:
        raise RuntimeError(
            "Out of device memory in the HIP engine. " + msg + " Please try to use a smaller mesh or a low-order integrator."
        )
    raise EngineError(msg)
