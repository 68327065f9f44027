"""The C-ABI from C++: tests/capi_client.cpp, compiled with g++ against
include/mpsft.h alone and linked to libmpsft_hip.so (built by build()),
runs plan -> BinInTime -> BinInFreq -> identify/estimate -> solve on a
signal this test writes; its outputs are checked against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from mpsft_amd.signals import make_signal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLIENT = os.path.join(ROOT, "tests", "capi_client")
TOL = 1e-12


def test_cpp_client_against_oracle(orc, tmp_path):
    assert os.path.exists(CLIENT), "tests/capi_client is built by __graft_entry__.build()"
    n, bins, delta, k, sigma, seed = 1 << 16, 101, 1e-6, 40, 1e-6, 9
    x, truth = make_signal(n, k, sigma, seed=80)
    bits = orc.num_bits(n, bins)
    rows = 1 + 2 * bits
    orc.seed(81)
    tf = orc.transform_draw(n)
    q = int(orc.rand_int32() % n)
    keys = np.array(sorted(truth)[:20], np.int64)  # half the modes already known
    coefs = np.array([truth[int(key)] for key in keys], np.complex128)
    bin_thr, sv_thr, win_thr = 3 * sigma / np.sqrt(bins), sigma / np.sqrt(bins), 0.1

    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(struct.pack("<10q", n, bins, bits, q, int(tf[0]), int(tf[1]), int(tf[2]),
                            keys.size, k, seed))
        f.write(struct.pack("<5d", delta, bin_thr, sv_thr, win_thr, sigma))
        f.write(np.ascontiguousarray(x, np.complex128).tobytes())
        f.write(keys.tobytes())
        f.write(coefs.tobytes())
    print(f"[capi_client] n={n} bins={bins} bits={bits} q={q}", flush=True)
    r = subprocess.run([CLIENT, str(inp), str(out)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr

    buf = open(out, "rb").read()
    pos = 0

    def take(dtype, count):
        nonlocal pos
        a = np.frombuffer(buf, dtype=dtype, count=count, offset=pos)
        pos += a.nbytes
        return a

    a_time = take(np.complex128, rows * bins).reshape(rows, bins)
    a_freq = take(np.complex128, rows * bins).reshape(rows, bins)
    decide, k1 = take(np.int32, bins), take(np.int64, bins)
    k0, coef = take(np.int64, bins), take(np.complex128, bins)
    iters = int(take(np.int32, 1)[0])
    count = int(take(np.int64, 1)[0])
    fkeys, fcoefs = take(np.int64, count), take(np.complex128, count)
    log_bins = take(np.int32, iters)
    assert pos == len(buf)

    # BinInTime (definition V0) and BinInFreq (V1)
    ref = orc.bin_in_time(x, bins, delta, bits, tf, q, version=0)
    scale = np.abs(ref).max()
    assert np.abs(a_time - ref).max() < TOL * scale
    ref_f = orc.bin_in_freq(keys.astype(np.int32), coefs, n, bins, delta, bits, tf, q, ref.copy())
    assert np.abs(a_freq - ref_f).max() < TOL * scale

    # Iterate's per-bin body on those bins: the oracle's Iterate with the
    # known modes as its map runs the same BinInTime + BinInFreq
    mm = orc.ModeMap({int(key): complex(v) for key, v in zip(keys, coefs)})
    opt = orc.IterateOptions(bins=bins, window_delta=delta, trials=1, bin_threshold=bin_thr,
                             window_threshold=win_thr, sv_threshold=sv_thr)
    _, tr = orc.iterate(x, opt, mm, tf=tf, qs=np.array([q], np.int32))
    assert np.array_equal(decide, tr["decide"])
    assert np.array_equal(k1, tr["k1"].astype(np.int64))
    found = decide == 3
    assert found.any()
    after = mm.to_dict()
    for b in np.flatnonzero(found):
        key = int(k0[b])
        added = after[key] - dict(zip(keys.tolist(), coefs)).get(key, 0)
        assert abs(coef[b] - added) <= TOL * max(abs(added), 1e-2)

    # Demo1::Run
    orc.seed(seed)
    ref_mm, ref_bins, _ = orc.demo1_run(x, k, sigma, orc.Demo1Options(window_delta=delta))
    assert iters == len(ref_bins) and np.array_equal(log_bins, ref_bins)
    got = dict(zip(fkeys.tolist(), fcoefs))
    want = ref_mm.to_dict()
    assert set(got) == set(want)
    for key, v in want.items():
        assert abs(got[key] - v) <= TOL * max(abs(v), 1e-2)
