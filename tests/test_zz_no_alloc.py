"""mpsft_solver_post_analyze_noalloc, the PostAnalyze entry point that
allocates nothing (its partial sums sit in a static device array), called
through the C-ABI and checked against mpsft_solver_post_analyze on two
calls in a row, the second on a smaller signal: the figures must be equal
bit for bit (same kernels, same workgroup count, same summation order)."""
import ctypes as C

import numpy as np
import pytest

from mpsft_amd.signals import make_signal

pytestmark = pytest.mark.gpu


def _noalloc(gpu, solver, truth, xh):
    L = gpu.lib()
    keys = gpu.DeviceArray.from_numpy(np.sort(np.array(list(truth), np.int64)))
    xd = gpu.DeviceArray.from_numpy(np.ascontiguousarray(xh, np.complex128))
    out = gpu.DeviceArray(8 * 10)
    rc = L.mpsft_solver_post_analyze_noalloc(solver.h, keys.ptr, len(truth), xd.ptr, out.ptr,
                                             None)
    assert rc == 0
    gpu.synchronize()
    return out.to_numpy(np.float64, (10,))


def test_post_analyze_noalloc_matches_allocating_form(gpu):
    names = ("hit", "hit_l1", "hit_l2", "miss", "miss_l1", "miss_l2", "l1", "l2", "max")
    for n, k, sigma, seed in ((1 << 20, 200, 1e-7, 91), (1 << 16, 30, 1e-6, 92)):
        print(f"[no_alloc] n={n} k={k}", flush=True)
        x, truth, xh = make_signal(n, k, sigma, seed=seed, with_xhat=True)
        s = gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_FP64)
        s.solve(s.upload(x), seed)
        want = s.post_analyze(list(truth), xh)
        got = _noalloc(gpu, s, truth, xh)
        assert got[9] == 0.0
        for i, name in enumerate(names):
            assert got[i] == want[name] or (np.isnan(got[i]) and np.isnan(want[name])), name
        assert want["hit"] == k
