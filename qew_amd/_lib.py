"""ctypes declaration of the C ABI in include/qew_amd.h.

Loading fails loudly when libqew_amd.so has not been built: there is no CPU
or Python fallback for any entry point.
"""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libqew_amd.so")

QEW_OK, QEW_ERR_INVALID_ARGUMENT, QEW_ERR_HIP, QEW_ERR_OUT_OF_MEMORY, QEW_ERR_UNSUPPORTED = range(5)
F64, F32, BF16 = 0, 1, 2  # BF16: table storage only (read with F32 line state)
# table storage only: FP32 values contracted with the B-spline x-weights at
# fraction 1/2 (read with F32 line state, lines with spacing = table dx, even n)
F32_MID = 3
# the same contraction kept in FP64 (read by the batched L-BFGS with F64 state)
F64_MID = 4
LINEAR, ARCLENGTH = 0, 1
TABLE, SINCOS, CONST, ZERO, HARMONIC, WASHBOARD = range(6)


class Params(C.Structure):
    _fields_ = [("physical_size", C.c_double), ("line_tension", C.c_double),
                ("driving_force", C.c_double), ("model", C.c_int)]


class LbfgsParams(C.Structure):
    _fields_ = [("m", C.c_int), ("ftol", C.c_double), ("max_iter", C.c_int),
                ("c1", C.c_double), ("ls_shrink", C.c_double), ("max_ls", C.c_int),
                ("curvature_eps", C.c_double), ("precondition", C.c_int),
                ("precond_shift", C.c_double)]


class LbfgsResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_force", C.c_double),
                ("converged", C.c_int), ("n_func_evals", C.c_long),
                ("n_curvature_skips", C.c_long), ("n_ls_restarts", C.c_long)]


class FireParams(C.Structure):
    _fields_ = [("dt_max", C.c_double), ("dt_init", C.c_double), ("ftol", C.c_double),
                ("max_iter", C.c_int), ("n_min", C.c_int), ("f_inc", C.c_double),
                ("f_dec", C.c_double), ("alpha_start", C.c_double), ("f_alpha", C.c_double)]


class FireResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("max_force", C.c_double), ("converged", C.c_int)]


class RkParams(C.Structure):
    _fields_ = [("dstep", C.c_double), ("max_advance", C.c_double),
                ("root_tol", C.c_double), ("sweep_tol", C.c_double),
                ("runaway", C.c_double), ("max_sweeps", C.c_int)]


class RkResult(C.Structure):
    _fields_ = [("blocked", C.c_int), ("runaway", C.c_int), ("sweeps", C.c_int),
                ("total_advance", C.c_double), ("mean_h", C.c_double),
                ("n_velocity_evals", C.c_long), ("n_velocity_evals_done", C.c_long)]


VP = C.c_void_p
DP = C.POINTER(C.c_double)
I = C.c_int
D = C.c_double
U64 = C.c_uint64

# (name, argtypes); every function returns qew_status (int) unless listed in _RESTYPE
SIGNATURES = [
    ("qew_ctx_create", [I, VP, C.POINTER(VP)]),
    ("qew_ctx_destroy", [VP]),
    ("qew_ctx_synchronize", [VP]),
    ("qew_device_alloc", [VP, C.c_size_t, C.POINTER(VP)]),
    ("qew_device_free", [VP, VP]),
    ("qew_copy_to_device", [VP, VP, VP, C.c_size_t]),
    ("qew_copy_to_host", [VP, VP, VP, C.c_size_t]),
    ("qew_noise_generate", [VP, I, I, D, D, D, D, D, U64, I, C.POINTER(VP)]),
    ("qew_noise_generate_batch", [VP, I, C.POINTER(I), C.POINTER(I), DP, DP, DP, DP, DP,
                                  C.POINTER(U64), I, C.POINTER(VP)]),
    ("qew_noise_convert", [VP, I, C.POINTER(VP)]),
    ("qew_noise_from_field", [VP, DP, I, I, D, D, I, C.POINTER(VP)]),
    ("qew_noise_from_coefficients", [VP, DP, I, I, D, D, I, C.POINTER(VP)]),
    ("qew_noise_analytic", [VP, I, DP, C.POINTER(VP)]),
    ("qew_noise_destroy", [VP]),
    ("qew_noise_info", [VP, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(I), DP, DP]),
    ("qew_noise_field", [VP, DP]),
    ("qew_noise_coefficients", [VP, DP]),
    ("qew_noise_sample", [VP, VP, I, VP, VP, VP, VP, VP]),
    ("qew_noise_sample_value", [VP, VP, I, VP, VP, VP]),
    ("qew_noise_sample_dy", [VP, VP, I, VP, VP, VP]),
    ("qew_objective", [VP, VP, C.POINTER(Params), I, VP, DP]),
    ("qew_objective_trial", [VP, VP, C.POINTER(Params), I, VP, VP, D, DP]),
    ("qew_gradient", [VP, VP, C.POINTER(Params), I, VP, VP]),
    ("qew_eval", [VP, VP, C.POINTER(Params), I, VP, VP, DP]),
    ("qew_gradient_edges", [VP, VP, C.POINTER(Params), I, VP, VP, VP, VP]),
    ("qew_eval_noalloc", [VP, VP, C.POINTER(Params), I, VP, VP, VP]),
    ("qew_batch_create", [VP, I, C.POINTER(I), C.POINTER(VP), C.POINTER(Params), C.POINTER(VP)]),
    ("qew_batch_destroy", [VP]),
    ("qew_batch_offsets", [VP, C.POINTER(C.c_int64)]),
    ("qew_batch_set_forces", [VP, DP]),
    ("qew_batch_set_precond_shifts", [VP, DP]),
    ("qew_batch_eval", [VP, I, VP, VP, VP, VP]),
    ("qew_lbfgs_minimize", [VP, VP, C.POINTER(Params), C.POINTER(LbfgsParams), I, VP,
                            C.POINTER(LbfgsResult)]),
    ("qew_batch_lbfgs", [VP, I, C.POINTER(LbfgsParams), VP, C.POINTER(LbfgsResult)]),
    ("qew_batch_lbfgs_plan", [VP, C.POINTER(I), C.POINTER(I)]),
    ("qew_roughness", [VP, I, I, D, VP, VP, VP, VP]),
    ("qew_fire_minimize", [VP, VP, C.POINTER(Params), C.POINTER(FireParams), I, VP,
                           C.POINTER(FireResult)]),
    ("qew_batch_fire", [VP, C.POINTER(FireParams), VP, C.POINTER(FireResult)]),
    ("qew_langevin_step", [VP, VP, C.POINTER(Params), I, VP, D, D, U64, U64]),
    ("qew_batch_langevin", [VP, I, VP, VP, I, DP, DP, U64, U64]),
    ("qew_rk_block", [VP, VP, C.POINTER(Params), C.POINTER(RkParams), I, VP, C.POINTER(RkResult)]),
    ("qew_batch_rk_block", [VP, C.POINTER(RkParams), VP, C.POINTER(RkResult)]),
    ("qew_rk_site_velocity", [VP, VP, C.POINTER(Params), I, I, VP, VP, VP, VP, VP]),
    ("qew_rk_forward_zero", [VP, VP, C.POINTER(Params), C.POINTER(RkParams), I, I, VP, VP, VP,
                             VP, VP, VP, VP]),
]

_RESTYPE = {
    "qew_last_error": C.c_char_p,
    "qew_version": C.c_char_p,
    "qew_lbfgs_default_params": LbfgsParams,
    "qew_fire_default_params": FireParams,
}

_lib = None


class QewError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"qew status {status}: {msg}")
        self.status = status


class QewInvalidArgument(QewError, ValueError):
    pass


def load():
    """Loads libqew_amd.so (once). Raises ImportError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m qew_amd.build` "
                          "(there is no fallback implementation)")
    lib = C.CDLL(LIB_PATH)
    for name, argtypes in SIGNATURES:
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int
    for name, rt in _RESTYPE.items():
        fn = getattr(lib, name)
        fn.argtypes = []
        fn.restype = rt
    _lib = lib
    return lib


def check(status):
    if status == QEW_OK:
        return
    msg = load().qew_last_error().decode(errors="replace")
    if status == QEW_ERR_INVALID_ARGUMENT:
        raise QewInvalidArgument(status, msg)
    raise QewError(status, msg)


def call(name, *args):
    check(getattr(load(), name)(*args))
