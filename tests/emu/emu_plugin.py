"""pytest plugin: binds the C-ABI to the host emulator build
(tests/emu/_build/libmpsft_emu.so) instead of the HIP library, so the GPU
tests run their kernels on the CPU under AddressSanitizer.  Used as
  LD_PRELOAD=$(g++ -print-file-name=libasan.so) ASAN_OPTIONS=detect_leaks=0 \
  PYTHONPATH=tests/emu python -m pytest -p emu_plugin -m gpu tests/test_gpu_parity.py
Test infrastructure only."""
import ctypes as C
import os

from mpsft_amd import _lib

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib._lib = _lib.declare(C.CDLL(os.path.join(_HERE, "_build", "libmpsft_emu.so")))
