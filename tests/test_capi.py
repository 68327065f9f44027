"""C-ABI checks that need no GPU: the library loads, exports every function
include/mpsft.h declares, and its host-only helpers and argument checks
behave like the reference (restated by the oracle)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mpsft_amd
from mpsft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", fn)) as f:
            src = f.read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(mpsft_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_header_symbol():
    names = header_functions()
    assert len(names) >= 29
    L = C.CDLL(_lib.LIB_PATH)
    for name in sorted(names):
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) == names


def test_version_and_status_strings():
    L = mpsft_amd.lib()
    assert L.mpsft_version() == 1
    for st, text in _lib.STATUS.items():
        assert L.mpsft_status_string(st).decode() == text


@pytest.mark.parametrize("n,bins", [(2053, 11), (1 << 22, 2001), (1 << 22, 101), (1048573, 2001),
                                    (5, 11), (1 << 26, 8001), (1109, 1)])
def test_num_bits_matches_oracle(orc, n, bins):
    assert mpsft_amd.num_bits(n, bins) == orc.num_bits(n, bins)


@pytest.mark.parametrize("n,bins,delta", [(1109, 5, 1e-6), (1 << 22, 2001, 1e-6),
                                          (1 << 22, 101, 1e-5), (1 << 20, 1, 1e-6),
                                          (1 << 26, 8001, 1e-6)])
def test_window_p_matches_oracle(orc, n, bins, delta):
    assert mpsft_amd.lib().mpsft_window_p(n, bins, delta) == orc.window_params(n, bins, delta)["p"]


def test_transform_make_matches_oracle(orc, kats):
    for case in ("time_freq_match", "time_freq_match_bigger"):
        k = kats[case]
        tf = mpsft_amd.transform_make(k["n"], *k["tf"])
        ref = orc.transform_make(k["n"], *k["tf"])
        assert (tf.a, tf.b, tf.c, tf.a_inv) == tuple(int(v) for v in ref)
    tf = mpsft_amd.transform_make(1 << 22, 12345, 7, 9)
    assert (tf.a * tf.a_inv) % (1 << 22) == 1


def test_inverse_mod_power_of_two():
    """The inverse of an odd multiplier modulo n = 2^k (Newton's iteration on
    the host) equals Python's modular inverse for k = 1..62 (five steps are
    needed above 2^48); other n (extended Euclid) alongside."""
    rng = np.random.default_rng(5)
    for k in range(1, 63):
        n = 1 << k
        for a in [int(v) | 1 for v in rng.integers(0, n, 40, dtype=np.uint64)] + [1, n - 1]:
            tf = mpsft_amd.transform_make(n, a, 0, 0)
            assert tf.a_inv == pow(a, -1, n), (n, a)
    for n in (3, 1109, 1048573, 3 * (1 << 20)):
        for a in rng.integers(1, n, 20):
            if np.gcd(int(a), n) == 1:
                assert mpsft_amd.transform_make(n, int(a), 0, 0).a_inv == pow(int(a), -1, n)


def test_argument_errors_need_no_gpu():
    L = mpsft_amd.lib()
    tf = mpsft_amd.Transform()
    # even a has no inverse modulo 2^k
    assert L.mpsft_transform_make(1 << 20, 4, 1, 1, C.byref(tf)) == 1
    assert L.mpsft_window_p(1109, 4, 1e-6) == 0           # even bins (src/window.cc:33)
    assert L.mpsft_window_p(1109, 5, 0.0) == 0            # delta must be > 0
    h = C.c_void_p()
    assert L.mpsft_plan_create(1109, 4, 1e-6, 2, 0, C.byref(h)) == 1   # even bins
    assert L.mpsft_plan_create(1109, 5, 1e-6, 0, 0, C.byref(h)) == 1   # bits < 1
    assert L.mpsft_plan_create(1109, 5, 1e-6, 2, 9, C.byref(h)) == 3   # unknown precision
    assert L.mpsft_plan_create(1 << 20, 32769, 1e-6, 2, 0, C.byref(h)) == 3  # B > 32767
    opt = mpsft_amd.solve_options(10, 1e-6)
    assert L.mpsft_solver_create(1 << 16, C.byref(opt), 9, 0, C.byref(h)) == 3
    assert L.mpsft_solve(None, None, 0, None, None) == 1
    assert L.mpsft_iterate(None, None, None, None, None, None, None) == 1
    assert L.mpsft_solver_set_record(None, 1) == 1
    assert L.mpsft_solver_info(None, None, None, None) == 1
    assert L.mpsft_solver_replay(None, None, 0, None) == 1
    assert L.mpsft_solver_adjoint(None, None, None, 0, None, 0, None) == 1


def test_torch_operator_registers_and_checks_arguments():
    """libmpsft_torch.so (the differentiable recovery) loads without a GPU
    and registers torch.ops.mpsft.recover with its schema."""
    import torch

    from mpsft_amd import autograd as ag
    ops = ag.load()
    assert str(ops.recover.default._schema) == (
        "mpsft::recover(Tensor x, int solver, int seed, int stream) -> (Tensor, Tensor)")
    with pytest.raises(RuntimeError, match="null solver"):
        ops.recover(torch.zeros(8, dtype=torch.complex128), 0, 1, 0)


def test_binner_factory_rejects_reference_cpu_types():
    win_ok = True
    try:
        w = mpsft_amd.Window(1109, 5, 1e-6)
    except Exception:
        win_ok = False
    assert win_ok and w.p() == 185
    for t in (0, 1, 2, 3, 4, 7):
        with pytest.raises(mpsft_amd.MpsftError):
            mpsft_amd.BinInTime.create(t, w, 2)
    with pytest.raises(mpsft_amd.MpsftError):
        mpsft_amd.Window(1109, 4, 1e-6)


@pytest.mark.parametrize("n,bins,delta", [(1 << 20, 101, 1e-6), (1 << 20, 101, 1e-5),
                                          (1 << 22, 2001, 1e-6), (1 << 22, 2001, 1e-5),
                                          (4099, 11, 1e-3), (1 << 26, 8001, 1e-6)])
def test_window_delta_from_taps(orc, n, bins, delta):
    """A reference Window keeps delta private (src/window.h:40) but exposes
    its taps: delta recovered from the oracle's taps (Window::wt(),
    src/window.cc:56-75) to 1e-12, and taps that are no window's refused."""
    wt, t = orc.window_taps(n, bins, delta)
    got = mpsft_amd.window_delta(n, bins, wt)
    assert abs(got - delta) <= 1e-12 * delta
    for bad in (wt * 1.001, np.concatenate([wt, wt[:bins]]), wt[: wt.size - 2],
                np.where(np.arange(wt.size) == wt.size // 3, wt * 1.01, wt)):
        with pytest.raises(mpsft_amd.MpsftError):
            mpsft_amd.window_delta(n, bins, bad)


def test_batch_seed_array_takes_low_32_bits():
    """Batch solve seeds reach the C-ABI as uint32: every kind of integer
    (Python, numpy, negative, wider than 32 bits) and integral floats keep
    their low 32 bits; a wrong count is refused."""
    from mpsft_amd.api import _u32_array
    vals = [5, (1 << 40) + 5, -1, np.int64(7), np.uint64((1 << 63) + 9), 2.0]
    assert list(_u32_array(vals, len(vals))) == [5, 5, 0xFFFFFFFF, 7, 9, 2]
    assert list(_u32_array(np.arange(4, dtype=np.int64), 4)) == [0, 1, 2, 3]
    with pytest.raises(ValueError):
        _u32_array([1, 2, 3], 4)
