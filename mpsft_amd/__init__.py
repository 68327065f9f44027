"""mpsft_amd -- MI355X-native MPSFT recovery loop (HIP kernels for gfx950
behind a C-ABI, include/mpsft.h).

The compute path is libmpsft_hip.so only; build it with
``python -m mpsft_amd.build``.  Importing this package does not touch the
GPU; the library is loaded on first use.
"""
from ._lib import (MPSFT_FP64, MPSFT_MIXED, EXPORTS, LIB_PATH, IterateOptions, MpsftError,
                   SolveOptions, Transform, lib)
from .api import (BIN_TYPE_FP64, BIN_TYPE_MIXED, Batch, BinInFreq, BinInTime, DeviceArray, Plan,
                  Solver, Window, device_count, iterate_options, mat_pencil, num_bits,
                  solve_options, synchronize, transform_make, window_delta)

__all__ = [
    "MPSFT_FP64", "MPSFT_MIXED", "EXPORTS", "LIB_PATH", "IterateOptions", "MpsftError",
    "SolveOptions", "Transform", "lib", "BIN_TYPE_FP64", "BIN_TYPE_MIXED", "Batch", "BinInFreq",
    "BinInTime", "DeviceArray", "Plan", "Solver", "Window", "device_count",
    "iterate_options", "mat_pencil", "num_bits", "solve_options", "synchronize",
    "transform_make", "window_delta",
]
