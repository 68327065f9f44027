"""ctypes binding of the C-ABI in include/mpsft.h (libmpsft_hip.so).

The HIP library is the only compute path: loading fails with an error when
it has not been built (``python -m mpsft_amd.build``); there is no fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpsft_hip.so")

MPSFT_FP64 = 0
MPSFT_MIXED = 1

STATUS = {
    0: "ok", 1: "invalid argument", 2: "shape mismatch", 3: "unsupported",
    4: "HIP runtime error", 5: "device allocation failed",
    6: "found-mode capacity exceeded", 7: "iteration cap reached",
}

# Symbols of include/mpsft.h; tests check the library exports all of them.
EXPORTS = [
    "mpsft_status_string", "mpsft_version", "mpsft_device_count", "mpsft_malloc",
    "mpsft_free", "mpsft_memcpy_htod", "mpsft_memcpy_dtoh", "mpsft_memset",
    "mpsft_stream_create", "mpsft_stream_destroy", "mpsft_stream_synchronize",
    "mpsft_device_synchronize", "mpsft_transform_make", "mpsft_num_bits",
    "mpsft_window_p", "mpsft_window_delta", "mpsft_generate_signal", "mpsft_plan_create", "mpsft_plan_create_window",
    "mpsft_plan_delta", "mpsft_plan_destroy", "mpsft_bin_in_time",
    "mpsft_bin_in_freq", "mpsft_identify_estimate", "mpsft_mat_pencil", "mpsft_solver_create",
    "mpsft_solver_destroy", "mpsft_solve", "mpsft_solve_profiled", "mpsft_iterate", "mpsft_solver_modes_count",
    "mpsft_solver_modes_get", "mpsft_solver_modes_set", "mpsft_solver_post_analyze",
    "mpsft_solver_post_analyze_noalloc",
    "mpsft_solver_bins_get",
    "mpsft_solver_log_get", "mpsft_solver_transforms_get", "mpsft_solver_set_tile_residues",
    "mpsft_solver_set_record", "mpsft_solver_info", "mpsft_solver_replay",
    "mpsft_solver_adjoint",
    "mpsft_batch_create", "mpsft_batch_destroy", "mpsft_batch_solve", "mpsft_batch_autotune",
    "mpsft_batch_solve_profiled",
    "mpsft_batch_modes_count", "mpsft_batch_modes_get", "mpsft_batch_log_get",
    "mpsft_batch_set_record", "mpsft_batch_info", "mpsft_batch_replay", "mpsft_batch_adjoint",
    "mpsft_batch_modes_counts", "mpsft_batch_modes_device",
]


class MpsftError(RuntimeError):
    def __init__(self, what, status):
        super().__init__(f"{what}: {STATUS.get(status, status)} (status {status})")
        self.status = status


class Transform(C.Structure):
    """mpsft_transform = Transform{a, b, c, a_inv} (src/integer.h:75-82)."""

    _fields_ = [("a", C.c_int64), ("b", C.c_int64), ("c", C.c_int64), ("a_inv", C.c_int64)]

    def __repr__(self):
        return f"Transform(a={self.a}, b={self.b}, c={self.c}, a_inv={self.a_inv})"


class IterateOptions(C.Structure):
    """mpsft_iterate_options = IterateOptions (src/iterate.h:26-59)."""

    _fields_ = [
        ("bins", C.c_int32), ("window_delta", C.c_double), ("trials", C.c_int32),
        ("bin_threshold", C.c_double), ("window_threshold", C.c_double),
        ("sv_threshold", C.c_double), ("bin_in_time_type", C.c_int32),
        ("bin_in_freq_type", C.c_int32),
    ]


class SolveOptions(C.Structure):
    """mpsft_solve_options = Demo1Options (src/demo1.h:25-33) + k, sigma."""

    _fields_ = [
        ("trials", C.c_int32), ("min_bins", C.c_int32), ("window_delta", C.c_double),
        ("window_threshold", C.c_double), ("max_stale_iter", C.c_int32), ("k", C.c_int32),
        ("sigma", C.c_double), ("max_iter", C.c_int32),
    ]


_lib = None


def lib():
    """Loads libmpsft_hip.so (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m mpsft_amd.build` "
            "(the HIP library is required, there is no CPU fallback)")
    _start_torch_hip()
    _lib = declare(C.CDLL(LIB_PATH))
    return _lib


def _start_torch_hip():
    """PyTorch for ROCm carries its own HIP runtime beside the system one this
    library links.  In one process the two coexist (they share the device's
    address space, so tensors and this library's buffers mix) only when
    torch's runtime opens the device first: started after ours, torch finds
    no GPU.  So torch, when installed, is started before this library."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def declare(L):
    """Sets the ctypes prototype of every include/mpsft.h symbol on L."""
    vp, i32, i64, f64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_size_t
    pi32, pi64, pf64 = C.POINTER(i32), C.POINTER(i64), C.POINTER(f64)
    sig = {
        "mpsft_status_string": (C.c_char_p, [C.c_int]),
        "mpsft_version": (C.c_int, []),
        "mpsft_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "mpsft_malloc": (C.c_int, [C.POINTER(vp), sz]),
        "mpsft_free": (C.c_int, [vp]),
        "mpsft_memcpy_htod": (C.c_int, [vp, vp, sz]),
        "mpsft_memcpy_dtoh": (C.c_int, [vp, vp, sz]),
        "mpsft_memset": (C.c_int, [vp, C.c_int, sz]),
        "mpsft_stream_create": (C.c_int, [C.POINTER(vp)]),
        "mpsft_stream_destroy": (C.c_int, [vp]),
        "mpsft_stream_synchronize": (C.c_int, [vp]),
        "mpsft_device_synchronize": (C.c_int, []),
        "mpsft_transform_make": (C.c_int, [i64, i64, i64, i64, C.POINTER(Transform)]),
        "mpsft_num_bits": (i32, [i64, i32]),
        "mpsft_window_p": (i32, [i64, i32, f64]),
        "mpsft_window_delta": (C.c_int, [i64, i32, i32, pf64, pf64]),
        "mpsft_generate_signal": (C.c_int, [i64, i32, f64, C.c_uint32, vp, pi64, pf64, vp]),
        "mpsft_plan_create": (C.c_int, [i64, i32, f64, i32, C.c_int, C.POINTER(vp)]),
        "mpsft_plan_create_window": (C.c_int, [i64, i32, i32, pf64, i32, C.c_int,
                                               C.POINTER(vp)]),
        "mpsft_plan_delta": (C.c_int, [vp, pf64]),
        "mpsft_identify_estimate": (C.c_int, [vp, vp, i32, C.POINTER(Transform), pi64, f64, f64,
                                              f64, vp, vp, vp, vp, vp, vp]),
        "mpsft_plan_destroy": (C.c_int, [vp]),
        "mpsft_bin_in_time": (C.c_int, [vp, vp, i64, C.POINTER(Transform), i64, vp, vp]),
        "mpsft_bin_in_freq": (C.c_int, [vp, vp, vp, i64, C.POINTER(Transform), i64, vp, vp]),
        "mpsft_mat_pencil": (C.c_int, [vp, i64, vp, vp, vp]),
        "mpsft_solver_create": (C.c_int, [i64, C.POINTER(SolveOptions), C.c_int, i64,
                                          C.POINTER(vp)]),
        "mpsft_solver_destroy": (C.c_int, [vp]),
        "mpsft_solve": (C.c_int, [vp, vp, C.c_uint32, pi32, vp]),
        "mpsft_solve_profiled": (C.c_int, [vp, vp, C.c_uint32, pi32, pf64, vp]),
        "mpsft_iterate": (C.c_int, [vp, vp, C.POINTER(IterateOptions), C.POINTER(Transform),
                                    pi64, pi32, vp]),
        "mpsft_solver_modes_count": (C.c_int, [vp, pi64]),
        "mpsft_solver_modes_get": (C.c_int, [vp, pi64, pf64, i64, pi64]),
        "mpsft_solver_modes_set": (C.c_int, [vp, pi64, pf64, i64]),
        "mpsft_solver_post_analyze": (C.c_int, [vp, vp, i64, vp, pf64, vp]),
        "mpsft_solver_post_analyze_noalloc": (C.c_int, [vp, vp, i64, vp, vp, vp]),
        "mpsft_solver_bins_get": (C.c_int, [vp, pf64, i64, pi32, pi32, pi32]),
        "mpsft_solver_log_get": (C.c_int, [vp, pi32, pi32, i32, pi32]),
        "mpsft_solver_transforms_get": (C.c_int, [vp, C.POINTER(Transform), pi64, i32, pi32]),
        "mpsft_solver_set_tile_residues": (C.c_int, [vp, i32]),
        "mpsft_solver_set_record": (C.c_int, [vp, i32]),
        "mpsft_solver_info": (C.c_int, [vp, pi64, pi32, pi64]),
        "mpsft_solver_replay": (C.c_int, [vp, vp, C.c_int, vp]),
        "mpsft_solver_adjoint": (C.c_int, [vp, vp, vp, i64, vp, C.c_int, vp]),
        "mpsft_batch_create": (C.c_int, [i32, pi64, C.POINTER(SolveOptions), C.c_int,
                                         C.POINTER(vp)]),
        "mpsft_batch_destroy": (C.c_int, [vp]),
        "mpsft_batch_solve": (C.c_int, [vp, C.POINTER(vp), C.POINTER(C.c_uint32), pi32, pi32,
                                        vp]),
        "mpsft_batch_autotune": (C.c_int, [vp, C.POINTER(vp), C.POINTER(C.c_uint32), pi32, vp]),
        "mpsft_batch_solve_profiled": (C.c_int, [vp, C.POINTER(vp), C.POINTER(C.c_uint32), pi32,
                                                 pi32, C.POINTER(C.c_double), pi32, vp]),
        "mpsft_batch_modes_count": (C.c_int, [vp, i32, pi64]),
        "mpsft_batch_modes_get": (C.c_int, [vp, i32, pi64, pf64, i64, pi64]),
        "mpsft_batch_log_get": (C.c_int, [vp, i32, pi32, pi32, i32, pi32]),
        "mpsft_batch_set_record": (C.c_int, [vp, i32]),
        "mpsft_batch_modes_counts": (C.c_int, [vp, pi64]),
        "mpsft_batch_modes_device": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), vp]),
        "mpsft_batch_info": (C.c_int, [vp, i32, pi64, pi32, pi64]),
        "mpsft_batch_replay": (C.c_int, [vp, i32, vp, C.c_int, vp]),
        "mpsft_batch_adjoint": (C.c_int, [vp, C.POINTER(vp), C.POINTER(vp), pi64, C.POINTER(vp),
                                          C.c_int, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def check(status, what):
    if status != 0:
        raise MpsftError(what, status)
