"""ctypes bindings of libtealeaf_hip (include/tealeaf_hip.h).

The library is loaded from the package tree (``tealeaf.jl_amd/lib``); a
missing or unloadable library raises immediately -- there is no fallback
implementation of any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(ROOT, "lib", "libtealeaf_hip.so")

(TL_OK, TL_E_INVALID, TL_E_HIP, TL_E_NOMEM, TL_E_BREAKDOWN, TL_E_EIGEN, TL_E_PARSE, TL_E_STATE, TL_E_IO,
 TL_E_CHECK) = range(10)
TL_CG, TL_CHEBY, TL_PPCG, TL_JACOBI = 0, 1, 2, 3
TL_FP64, TL_FP32_IR = 0, 1
TL_STOP_REFERENCE, TL_STOP_RELATIVE = 0, 1
TL_WARM_NONE, TL_WARM_LEARNED, TL_WARM_GUESS = 0, 1, 2
TL_LAUNCH_STREAM, TL_LAUNCH_GRAPH = 0, 1
TL_COEF_FP16, TL_COEF_BF16, TL_COEF_FP32 = 1, 2, 3
FIELDS = ["density0", "density", "energy0", "energy", "u", "u0", "p", "r", "mi", "w", "kx", "ky", "sd"]
FID = {n: i for i, n in enumerate(FIELDS)}
X_DENSITY, X_P, X_ENERGY0, X_ENERGY, X_U, X_SD = 1, 2, 4, 8, 16, 32
MAX_STATES = 128

# every symbol include/tealeaf_hip.h declares
EXPORTS = [
    "tl_default_params", "tl_create", "tl_destroy", "tl_last_error", "tl_dims", "tl_set_field", "tl_get_field",
    "tl_set_fields", "tl_set_guess", "tl_initialise", "tl_halo_update", "tl_solve", "tl_solve_finished",
    "tl_field_summary", "tl_step", "tl_deck_parse", "tl_deck_create", "tl_run", "tl_eigen_bounds",
    "tl_cheby_coef", "tl_cheby_calciter", "tl_apply_operator", "tl_get_cg_history", "tl_unet_param_count",
    "tl_unet_set_weights", "tl_unet_get_weights", "tl_unet_record", "tl_unet_dataset_size", "tl_unet_train",
    "tl_unet_predict", "tl_profile_begin", "tl_profile_read", "tl_profile_end",
    "tl_kernel_repeat", "tl_unet_eval", "tl_checking_value", "tl_qa_check", "tl_check_solution",
    "tl_field_summary_noalloc",
]
TL_START_PREVIOUS, TL_START_WARM = 0, 1
(TL_K_CG_W64, TL_K_CG_UR64, TL_K_CG_P64, TL_K_CG_W32, TL_K_CG_UR32, TL_K_CG_P32, TL_K_CHEBY64, TL_K_PPCG64,
 TL_K_JACOBI64, TL_K_RESID64, TL_K_CGF32, TL_K_PPCG2_64, TL_K_CHEBY2_64, TL_K_CGP_G32, TL_K_CGP_U32) = range(15)
TL_K_COUNT = 15
TL_FP32_KERNELS_AUTO, TL_FP32_KERNELS_TWO, TL_FP32_KERNELS_FUSED, TL_FP32_KERNELS_PAIRS = 0, 1, 2, 3
TL_WARM_POLICY_BACKOFF, TL_WARM_POLICY_EVERY = 0, 1
TL_PPCG_PAIRS_AUTO, TL_PPCG_PAIRS_OFF = 0, 1
TL_CHEBY_PAIRS_AUTO, TL_CHEBY_PAIRS_OFF = 0, 1
KERNEL_NAMES = {0: "k_cg_pwv<double, 2, 4, true>", 1: "k_cg_urv<double, 2, 4, false>", 2: "k_cg_p<double>",
                3: "k_cgg<4, false, 2, true, 1>", 4: "k_cgg<4, true, 4, true, 1>",
                5: "(fused into k_cgg<4, false, 2, true, 1>)",
                6: "k_vs64<0, 4, true>", 7: "k_vs64<1, 4, true>", 8: "k_vs64<2, 4, true>", 9: "k_residual",
                10: "k_cgf<1, 1, 0>", 11: "k_ppcg2", 12: "k_cheby2", 13: "k_cgp<1, false>", 14: "k_cgp<1, true>",
                -1: "k_copy_full", -2: "k_copy_vec"}


def kernel_name(kid, coef_format=None):
    """Symbol of profiled kernel `kid`.  The FP32 kernels' last template
    argument is the cell coefficient's 16-bit format (1 = scaled half,
    0 = bfloat16; the fused kernel's second, before its LDS staging depth,
    0 in the default build; the pair kernels' first), which the solver picks per solve
    (`SolveStats.coef_format`); without a format the half instantiation is
    named."""
    name = KERNEL_NAMES[kid]
    if coef_format == TL_COEF_BF16 and kid == TL_K_CGF32:
        name = name.replace("<1, 1,", "<1, 0,")
    elif coef_format == TL_COEF_BF16 and kid in (TL_K_CGP_G32, TL_K_CGP_U32):
        name = name.replace("<1,", "<0,")
    elif coef_format == TL_COEF_BF16 and kid in (TL_K_CG_W32, TL_K_CG_UR32, TL_K_CG_P32):
        name = name.replace(", 1>", ", 0>")
    return name


class SolveParams(C.Structure):
    _fields_ = [
        ("solver", C.c_int), ("coefficient", C.c_int), ("max_iters", C.c_int), ("presteps", C.c_int),
        ("ppcg_inner_steps", C.c_int), ("error_switch", C.c_int), ("eps", C.c_double), ("epslim", C.c_double),
        ("precision", C.c_int), ("stop_rule", C.c_int), ("rel_tol", C.c_double), ("warm_start", C.c_int),
        ("check_result", C.c_int), ("ir_delta", C.c_double), ("launch", C.c_int), ("fp32_kernels", C.c_int),
        ("warm_policy", C.c_int), ("ppcg_pairs", C.c_int), ("cheby_pairs", C.c_int), ("reserved", C.c_int * 3),
    ]


class SolveStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int), ("cg_iters", C.c_int), ("cheby_iters", C.c_int), ("est_iters", C.c_int),
        ("inner_iters", C.c_int), ("refreshes", C.c_int), ("converged", C.c_int), ("fallback", C.c_int),
        ("error", C.c_double), ("eigmin", C.c_double), ("eigmax", C.c_double), ("rel_resid", C.c_double),
        ("init_rel_resid", C.c_double), ("solve_ms", C.c_double), ("start", C.c_int), ("coef_format", C.c_int),
        ("warm_attempted", C.c_int), ("reserved_i", C.c_int), ("warm_ms", C.c_double), ("reserved", C.c_double * 1),
    ]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if not n.startswith("reserved")}


class State(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("density", "energy", "xmin", "ymin", "xmax", "ymax", "radius")] + [
        ("geometry", C.c_int), ("defined", C.c_int)
    ]


class Deck(C.Structure):
    _fields_ = [
        ("x_cells", C.c_int), ("y_cells", C.c_int), ("halo_depth", C.c_int),
        ("xmin", C.c_double), ("ymin", C.c_double), ("xmax", C.c_double), ("ymax", C.c_double),
        ("dx", C.c_double), ("dy", C.c_double), ("end_step", C.c_longlong), ("summary_frequency", C.c_int),
        ("dt_init", C.c_double), ("end_time", C.c_double), ("check_result", C.c_int),
        ("params", SolveParams), ("nstates", C.c_int), ("states", State * MAX_STATES), ("n_warnings", C.c_int),
    ]


class TrainParams(C.Structure):
    _fields_ = [
        ("steps", C.c_int), ("batch", C.c_int), ("crop", C.c_int), ("lr", C.c_float), ("seed", C.c_int),
        ("reserved", C.c_int * 8),
    ]


STEP_CB = C.CFUNCTYPE(None, C.c_int, C.POINTER(C.c_double), C.POINTER(SolveStats), C.c_void_p)

_lib = None


class TeaLeafError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load():
    """Load libtealeaf_hip.so; raises if it is missing (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libtealeaf_hip.so not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip, fp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_float)
    L.tl_default_params.argtypes = [C.POINTER(SolveParams)]
    L.tl_default_params.restype = None
    L.tl_create.argtypes = [C.POINTER(vp), C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_int]
    L.tl_destroy.argtypes = [vp]
    L.tl_destroy.restype = None
    L.tl_last_error.argtypes = [vp]
    L.tl_last_error.restype = C.c_char_p
    L.tl_dims.argtypes = [vp, ip, ip]
    L.tl_set_field.argtypes = [vp, C.c_int, dp]
    L.tl_get_field.argtypes = [vp, C.c_int, dp]
    L.tl_set_fields.argtypes = [vp, dp, dp]
    L.tl_set_guess.argtypes = [vp, dp]
    L.tl_initialise.argtypes = [vp, C.POINTER(State), C.c_int]
    L.tl_halo_update.argtypes = [vp, C.c_uint, C.c_int]
    L.tl_solve.argtypes = [vp, C.POINTER(SolveParams), C.c_double, C.c_double, C.POINTER(SolveStats)]
    L.tl_solve_finished.argtypes = [vp, C.c_int]
    L.tl_field_summary.argtypes = [vp, dp]
    L.tl_step.argtypes = [vp, C.POINTER(SolveParams), C.c_double, C.POINTER(SolveStats)]
    L.tl_deck_parse.argtypes = [C.c_char_p, C.POINTER(Deck)]
    L.tl_deck_create.argtypes = [C.POINTER(vp), C.POINTER(Deck), C.c_int]
    L.tl_run.argtypes = [vp, C.POINTER(Deck), C.POINTER(SolveParams), C.c_int, STEP_CB, C.c_void_p]
    L.tl_eigen_bounds.argtypes = [dp, dp, C.c_int, dp, dp]
    L.tl_cheby_coef.argtypes = [C.c_double, C.c_double, C.c_int, dp, dp, dp]
    L.tl_cheby_coef.restype = None
    L.tl_cheby_calciter.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double]
    L.tl_apply_operator.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int]
    L.tl_get_cg_history.argtypes = [vp, dp, dp, C.c_int]
    L.tl_unet_param_count.argtypes = []
    L.tl_unet_set_weights.argtypes = [vp, fp]
    L.tl_unet_get_weights.argtypes = [vp, fp]
    L.tl_unet_record.argtypes = [vp, C.c_int, C.c_int]
    L.tl_unet_dataset_size.argtypes = [vp]
    L.tl_unet_train.argtypes = [vp, C.POINTER(TrainParams), fp]
    L.tl_unet_predict.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.tl_unet_eval.argtypes = [vp, fp, fp, C.c_int, C.c_int, C.c_int, fp, fp, fp]
    L.tl_profile_begin.argtypes = [vp, C.c_uint, C.c_int]
    L.tl_profile_read.argtypes = [vp, C.c_int, dp, C.c_int, ip]
    L.tl_profile_end.argtypes = [vp]
    L.tl_kernel_repeat.argtypes = [vp, C.c_int, C.c_int]
    L.tl_checking_value.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, dp, ip]
    L.tl_qa_check.argtypes = [C.c_double, C.c_double, dp]
    L.tl_field_summary_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                           C.c_double, C.c_double, C.c_void_p, C.c_void_p]
    L.tl_check_solution.argtypes = [C.POINTER(SolveParams), C.POINTER(SolveStats), C.c_double, C.c_double,
                                    C.c_double]
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = ""):
    if rc != TL_OK:
        msg = load().tl_last_error(ctx)
        raise TeaLeafError(rc, f"{what}: {msg.decode() if msg else ''}")
