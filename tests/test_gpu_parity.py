"""Parity of the HIP path with the CPU oracle, through the C-ABI.

Bars: FP64 bin matrices within 1e-12 of the oracle (relative to the row's
largest magnitude); the recovered support (integer locations) identical;
coefficients within 1e-12 relative for FP64 and for MIXED alike: MIXED widens
each complex64 sample exactly to double and computes in double from there, so
the oracle run on the same complex64-rounded input is its exact reference.
"""
import numpy as np
import pytest

from demo1_cases import STALE_THEN_FOUND, check_demo1_schedule
from mpsft_amd.signals import make_signal, quantization_sigma

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-12  # FP64 arithmetic on the exactly widened complex64 samples


def rel_err(got, ref):
    scale = np.abs(ref).max()
    return np.abs(got - ref).max() / (scale if scale > 0 else 1.0)


def assert_coefs_close(got, ref, tol, where=""):
    """|got - ref| <= tol * max(|ref|, 1e-2 * largest |ref|).  A coefficient
    is estimated from bins that also held (and had subtracted) the largest
    modes, so its rounding error is set by their scale: for FP64 about
    1e-16 * |c_max| per term, e.g. 1.03e-15 absolute on a 2.4e-9 leakage
    mode beside unit modes after 357 iterations.  Every coefficient above
    1e-2 of the largest is held to tol relative to itself; smaller ones to
    the absolute bound tol * 1e-2 * |c_max| (1e-14 |c_max| for FP64)."""
    assert set(got) == set(ref), where
    if not ref:
        return
    floor = 1e-2 * max(abs(v) for v in ref.values())
    for key, v in ref.items():
        assert abs(got[key] - v) <= tol * max(abs(v), floor), (where, key, got[key], v)


def _tf(gpu, orc_tf, n):
    return gpu.transform_make(n, int(orc_tf[0]), int(orc_tf[1]), int(orc_tf[2]))


def _modes(orc, n, modes):
    mm = orc.ModeMap({int(k): complex(re, im) for k, re, im in modes})
    return mm, orc.evaluate_modes(n, mm)


# ---------------------------------------------------------------- BinInTime
@pytest.mark.parametrize("case", ["time_freq_match", "time_freq_match_bigger", "bin_in_time_match"])
def test_bin_in_time_reference_cases(gpu, orc, kats, case):
    k = kats[case]
    n, bins, bits, q = k["n"], k["bins"], k["bits"], k["q"]
    mm, x = _modes(orc, n, k["modes"])
    tf_o = orc.transform_make(n, *k["tf"])
    ref = orc.bin_in_time(x, bins, k["delta"], bits, tf_o, q, version=0)
    plan = gpu.Plan(n, bins, k["delta"], bits)
    got = plan.bin_in_time(x, _tf(gpu, tf_o, n), q)
    assert rel_err(got, ref) < FP64_TOL
    if "row0_bin1" in k:
        assert abs(got[0, 1] - complex(*k["row0_bin1"])) < 1.19e-5 * 2.2
    # BinInFreq on the same modes cancels BinInTime (src/binner_test.cc:98-102)
    keys, coefs = mm.items()
    fr = np.zeros_like(got)
    plan.bin_in_freq(keys.astype(np.int64), coefs, _tf(gpu, tf_o, n), q, fr)
    ref_fr = np.zeros_like(ref)
    orc.bin_in_freq(keys, coefs, n, bins, k["delta"], bits, tf_o, q, ref_fr, version=1)
    assert rel_err(fr, ref_fr) < FP64_TOL
    assert np.all(np.abs(got[0] + fr[0]) < 1.19e-5)


@pytest.mark.parametrize("n,bins,delta", [(1 << 16, 101, 1e-6), (1 << 18, 2001, 1e-6),
                                          (1 << 14, 4095, 1e-6), (1048573, 2001, 1e-6),
                                          (1 << 20, 101, 1e-5), (1 << 12, 1, 1e-6),
                                          (4099, 3, 1e-6),
                                          # B > 4096: the four-step bin FFT
                                          (1 << 16, 4097, 1e-6), (1 << 18, 8001, 1e-6),
                                          (1 << 17, 12001, 1e-5), (1 << 18, 16385, 1e-6)])
def test_bin_in_time_random(gpu, orc, n, bins, delta):
    rng = np.random.default_rng(n ^ bins)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    bits = orc.num_bits(n, bins)
    orc.seed(bins)
    tf_o = orc.transform_draw(n)
    q = int(rng.integers(0, n))
    ref = orc.bin_in_time(x, bins, delta, bits, tf_o, q, version=0)
    got = gpu.Plan(n, bins, delta, bits).bin_in_time(x, _tf(gpu, tf_o, n), q)
    for r in range(ref.shape[0]):
        assert rel_err(got[r], ref[r]) < FP64_TOL, r


def test_bin_in_time_mixed_input(gpu, orc):
    n, bins, delta = 1 << 18, 2001, 1e-6
    rng = np.random.default_rng(9)
    x64 = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    bits = orc.num_bits(n, bins)
    orc.seed(3)
    tf_o = orc.transform_draw(n)
    ref = orc.bin_in_time(x64.astype(np.complex128), bins, delta, bits, tf_o, 12345, version=0)
    got = gpu.Plan(n, bins, delta, bits, gpu.MPSFT_MIXED).bin_in_time(x64, _tf(gpu, tf_o, n), 12345)
    assert rel_err(got, ref) < MIXED_TOL


def test_bin_in_freq_many_modes(gpu, orc):
    n, bins, delta = 1 << 16, 101, 1e-6
    bits = orc.num_bits(n, bins)
    rng = np.random.default_rng(2)
    keys = rng.choice(n, 500, replace=False).astype(np.int64)
    coefs = rng.standard_normal(500) + 1j * rng.standard_normal(500)
    orc.seed(8)
    tf_o = orc.transform_draw(n)
    base = rng.standard_normal((1 + 2 * bits, bins)) + 1j * rng.standard_normal((1 + 2 * bits, bins))
    ref = base.copy()
    orc.bin_in_freq(keys.astype(np.int32), coefs, n, bins, delta, bits, tf_o, 77, ref, version=1)
    got = base.copy()
    gpu.Plan(n, bins, delta, bits).bin_in_freq(keys, coefs, _tf(gpu, tf_o, n), 77, got)
    assert rel_err(got, ref) < FP64_TOL


# ---------------------------------------------------------------- MatPencil
@pytest.mark.parametrize("delta", [1e-5, 1e-6])
def test_plan_from_window_taps(gpu, orc, delta):
    """A plan built from a reference Window's taps alone (its delta is
    private, src/window.h:40): BinInTime with the caller's taps and BinInFreq
    with the recovered delta match the delta-built plan and the oracle."""
    n, bins, bits, k = 1 << 16, 101, 5, 40
    x, truth = make_signal(n, k, 1e-6, seed=60)
    wt, _ = orc.window_taps(n, bins, delta)
    pw = gpu.Plan.from_window(n, bins, wt, bits)
    pd = gpu.Plan(n, bins, delta, bits)
    assert abs(pw.delta - delta) <= 1e-12 * delta
    orc.seed(61)
    for _ in range(3):
        tf_o = orc.transform_draw(n)
        q = orc.rand_int32() % n
        tf = _tf(gpu, tf_o, n)
        a_w, a_d = pw.bin_in_time(x, tf, q), pd.bin_in_time(x, tf, q)
        ref = orc.bin_in_time(x, bins, delta, bits, tf_o, q, version=0)
        assert rel_err(a_w, a_d) < FP64_TOL and rel_err(a_w, ref) < FP64_TOL
        keys = np.array(list(truth), np.int64)
        coefs = np.array([truth[int(key)] for key in keys], np.complex128)
        f_w = pw.bin_in_freq(keys, coefs, tf, q, a_w.copy())
        f_d = pd.bin_in_freq(keys, coefs, tf, q, a_d.copy())
        f_o = orc.bin_in_freq(keys.astype(np.int32), coefs, n, bins, delta, bits, tf_o, q,
                              ref.copy())
        scale = np.abs(a_d).max()
        assert np.abs(f_w - f_d).max() < FP64_TOL * scale
        assert np.abs(f_w - f_o).max() < FP64_TOL * scale


@pytest.mark.parametrize("trials", [1, 3])
def test_identify_estimate_matches_oracle_trace(gpu, orc, trials):
    """mpsft_identify_estimate on the oracle's own bin matrices (after
    BinInFreq) of a real iteration: every bin's decision and k1 equal the
    oracle's Iterate trace, found coefficients within 1e-12 of what it adds
    to the map, and the decision margins agree (for the bins whose every
    comparison both evaluate)."""
    n, k, sigma = 1 << 16, 300, 1e-6
    x, _ = make_signal(n, k, sigma, seed=21 + trials, magnitude="loguniform")
    bins = 2 * k + 1
    kw = dict(bins=bins, window_delta=1e-6, trials=trials, bin_threshold=3 * sigma / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sigma / np.sqrt(bins))
    orc.seed(70 + trials)
    mm = orc.ModeMap()
    for step in range(2):  # the second step runs with the first step's modes subtracted
        before = mm.to_dict()
        tf_o = orc.transform_draw(n)
        qs = np.array([orc.rand_int32() % n for _ in range(trials)], np.int32)
        _, tr = orc.iterate(x, orc.IterateOptions(**kw), mm, tf=tf_o, qs=qs, want_a=True)
        plan = gpu.Plan(n, bins, 1e-6, tr["bits"])
        got = plan.identify_estimate(tr["a"], _tf(gpu, tf_o, n), qs.astype(np.int64),
                                     kw["bin_threshold"], kw["sv_threshold"],
                                     kw["window_threshold"])
        assert np.array_equal(got["decide"], tr["decide"]), step
        assert np.array_equal(got["k1"], tr["k1"].astype(np.int64)), step
        found = got["decide"] == 3
        assert found.sum() == tr["found"] and found.any()
        after = mm.to_dict()
        for b in np.flatnonzero(found):
            key = int(got["k0"][b])
            added = after[key] - before.get(key, 0)
            assert abs(got["coef"][b] - added) <= FP64_TOL * max(abs(added), 1e-2), (step, b)
        full = got["decide"] != 1  # the oracle stops at the first sv rejection
        assert np.abs(got["margin"][full] - tr["margin"][full]).max() < 1e-9
        assert np.all(got["margin"][~full] <= tr["margin"][~full] + 1e-12)


def test_mat_pencil_kats_and_random(gpu, orc, kats):
    cases = kats["pencil_kats"]["cases"]
    u0 = np.array([complex(*c["u0"]) for c in cases])
    u1 = np.array([complex(*c["u1"]) for c in cases])
    u2 = np.array([complex(*c["u2"]) for c in cases])
    votes, sig = gpu.mat_pencil(u0, u1, u2)
    for i, c in enumerate(cases):
        assert votes[i] == c["vote"]
        assert np.allclose(sig[i], c["sigma"], rtol=1e-8)
    rng = np.random.default_rng(4)
    m = 4096
    u = (rng.standard_normal((3, m)) + 1j * rng.standard_normal((3, m))) * 10.0 ** rng.uniform(-8, 2, m)
    votes, sig = gpu.mat_pencil(u[0], u[1], u[2])
    for i in range(m):
        v, s = orc.mat_pencil(u[0, i], u[1, i], u[2, i])
        assert v == votes[i]
        assert np.allclose(s, sig[i], rtol=1e-13, atol=0)


# ---------------------------------------------------------------- Iterate
def _iterate_both(gpu, orc, x, n, iopts, orc_opt, steps, seed, precision=0, x_orc=None):
    """Runs `steps` Iterate calls on the GPU and on the oracle with the same
    transforms (drawn from the oracle's RNG) and compares them."""
    orc.seed(seed)
    solver = gpu.Solver(n, gpu.solve_options(1, 0.0), precision)
    xd = solver.upload(x)
    mm = orc.ModeMap()
    for step in range(steps):
        tf_o = orc.transform_draw(n)
        qs = [orc.rand_int32() % n for _ in range(orc_opt.trials)]
        f_ref, tr = orc.iterate(x if x_orc is None else x_orc, orc_opt, mm, tf=tf_o,
                                qs=np.array(qs, np.int32), want_a=True)
        f_got = solver.iterate(xd, iopts, _tf(gpu, tf_o, n), qs)
        a_got = solver.bins_matrix(orc_opt.trials, tr["a"].shape[1], orc_opt.bins)
        tol = FP64_TOL if precision == 0 else MIXED_TOL
        for t in range(orc_opt.trials):
            for r in range(a_got.shape[1]):
                assert rel_err(a_got[t, r], tr["a"][t, r]) < tol, (step, t, r)
        assert f_got == f_ref
        assert_coefs_close(solver.modes_dict(), mm.to_dict(), tol, f"step {step}")
    return mm


@pytest.mark.parametrize("trials", [1, 3])
def test_iterate_parity(gpu, orc, trials):
    n, k, sigma = 1 << 16, 300, 1e-6
    x, truth = make_signal(n, k, sigma, seed=21, magnitude="loguniform")
    bins = 2 * k + 1
    kw = dict(bins=bins, window_delta=1e-6, trials=trials, bin_threshold=3 * sigma / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sigma / np.sqrt(bins))
    _iterate_both(gpu, orc, x, n, gpu.iterate_options(**kw), orc.IterateOptions(**kw), 3,
                  seed=5 + trials)


def test_iterate_parity_large_bins(gpu, orc):
    """B = 5001 (M = 16384, the four-step bin FFT); the second step
    subtracts the modes the first found (BinInFreq inside k_big_out)."""
    n, k, sigma = 1 << 18, 2500, 1e-6
    x, _ = make_signal(n, k, sigma, seed=23)
    bins = 2 * k + 1
    kw = dict(bins=bins, window_delta=1e-6, trials=1, bin_threshold=3 * sigma / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sigma / np.sqrt(bins))
    mm = _iterate_both(gpu, orc, x, n, gpu.iterate_options(**kw), orc.IterateOptions(**kw), 2,
                       seed=41)
    assert len(mm.to_dict()) > k // 2


def test_reference_iterate_once_on_gpu(gpu, orc, kats):
    """IterateOnce (src/iterate_test.cc:47-81) through the HIP path: prime n,
    B=11, trials=5; identical to the oracle and within the test's bounds."""
    k = kats["iterate_once"]
    orc.seed(k["seed"])
    modes = {int(m[0]): complex(m[1], m[2]) for m in k["modes"]}
    xh = orc.generate_xhat(k["n"], orc.ModeMap(modes), k["sigma"])
    x = np.fft.ifft(xh) * k["n"]
    kw = dict(bins=k["bins"], window_delta=k["delta"], trials=k["trials"],
              bin_threshold=k["bin_threshold"], window_threshold=k["window_threshold"])
    mm = _iterate_both(gpu, orc, x, k["n"], gpu.iterate_options(**kw), orc.IterateOptions(**kw),
                       1, seed=k["seed"] + 1)
    d = mm.to_dict()
    assert len(d) == 2
    for key, want in modes.items():
        assert abs(d[key] - want) < k["tolerance"]


def test_reference_iterate_full_on_gpu(gpu, orc, kats):
    """IterateFull (src/iterate_test.cc:126-175): n=1048573, 1000 modes,
    B=2001, trials=3, four iterations, identical to the oracle."""
    k = kats["iterate_full"]
    orc.seed(k["seed"])
    truth = orc.generate_mode_map(k["n"], k["num_modes"])
    xh = orc.generate_xhat(k["n"], truth, k["sigma"])
    x = np.fft.ifft(xh) * k["n"]
    s = k["sigma"]
    kw = dict(bins=k["bins"], window_delta=k["delta"], trials=k["trials"],
              bin_threshold=3.0 * s / np.sqrt(k["bins"]), sv_threshold=1.0 * s / np.sqrt(k["bins"]),
              window_threshold=k["window_threshold"])
    mm = _iterate_both(gpu, orc, x, k["n"], gpu.iterate_options(**kw), orc.IterateOptions(**kw),
                       4, seed=k["seed"] + 2)
    td = truth.to_dict()
    good = sum(1 for key in mm.to_dict() if key in td)
    assert good >= k["min_good"][3]


def test_iterate_mixed_parity(gpu, orc):
    n, k, sigma = 1 << 16, 300, 1e-6
    x, _ = make_signal(n, k, sigma, seed=22, magnitude="loguniform")
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    bins = 2 * k + 1
    kw = dict(bins=bins, window_delta=1e-6, trials=1, bin_threshold=3 * sig_eff / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sig_eff / np.sqrt(bins))
    _iterate_both(gpu, orc, x64, n, gpu.iterate_options(precision=1, **kw),
                  orc.IterateOptions(**kw), 2, seed=31, precision=1,
                  x_orc=x64.astype(np.complex128))


# ---------------------------------------------------------------- Demo1::Run
def _solve_both(gpu, orc, x_gpu, x_orc, n, k, sigma, seed, precision=0, **kw):
    solver = gpu.Solver(n, gpu.solve_options(k, sigma, **kw), precision)
    xd = solver.upload(x_gpu)
    iters = solver.solve(xd, seed)
    orc.seed(seed)
    found, bins_log, found_log = orc.demo1_run(x_orc, k, sigma, orc.Demo1Options(**kw))
    g_bins, g_found = solver.log()
    assert iters == len(bins_log)
    assert np.array_equal(g_bins, bins_log)
    assert np.array_equal(g_found, found_log)
    ref = found.to_dict()
    got = solver.modes_dict()
    assert set(got) == set(ref)
    return got, ref, solver, xd


def test_solve_parity_config1(gpu, orc):
    """N=2^20, S=50 unit modes, sigma=1e-8, Demo1Options defaults."""
    n, k, sigma = 1 << 20, 50, 1e-8
    x, truth = make_signal(n, k, sigma, seed=1)
    got, ref, solver, xd = _solve_both(gpu, orc, x, x, n, k, sigma, seed=7)
    assert_coefs_close(got, ref, FP64_TOL)
    xh = np.fft.fft(x) / n  # global oracle: the dense FFT
    for key in truth:
        assert key in got and abs(got[key] - xh[key]) <= 1e-6 * abs(xh[key])
    # a second solve on the same solver reproduces the first
    assert solver.solve(xd, 7) == len(solver.log()[0])
    assert solver.modes_dict().keys() == got.keys()


def test_solve_parity_config2_fp64(gpu, orc):
    """N=2^22, S=1000, |c| log-uniform in [0.1, 10], sigma=1e-7, complex128."""
    n, k, sigma = 1 << 22, 1000, 1e-7
    x, truth = make_signal(n, k, sigma, seed=2, magnitude="loguniform")
    got, ref, _, _ = _solve_both(gpu, orc, x, x, n, k, sigma, seed=11)
    assert_coefs_close(got, ref, FP64_TOL)


def test_post_analyze_config2_fp64(gpu, orc):
    """PostAnalyze (src/demo1.cc:61-112) on the device after the N=2^22,
    S=1000 FP64 solve, against the oracle's on the same found modes and
    x-hat: counts exact, errors within 1e-12; the same figures bitwise on a
    second call (fixed summation order)."""
    n, k, sigma = 1 << 22, 1000, 1e-7
    x, truth, xh = make_signal(n, k, sigma, seed=2, magnitude="loguniform", with_xhat=True)
    s = gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_FP64)
    s.solve(s.upload(x), 11)
    got = s.post_analyze(list(truth), xh)
    found = orc.ModeMap(s.modes_dict())
    ref = orc.post_analyze(n, orc.ModeMap(truth), xh, found)
    print(f"[post_analyze] gpu {got}\n[post_analyze] orc {ref}", flush=True)
    assert got["hit"] == ref["hit"] and got["miss"] == ref["miss"] and got["miss"] > 0
    for key in ("hit_l1", "hit_l2", "miss_l1", "miss_l2", "l1", "l2", "max"):
        assert abs(got[key] - ref[key]) <= 1e-12 * abs(ref[key]), key
    assert s.post_analyze(list(truth), xh) == got
    with pytest.raises(gpu.MpsftError):  # truth keys must be distinct
        s.post_analyze(list(truth) + [next(iter(truth))], xh)


def test_solve_parity_config2_mixed(gpu, orc):
    """The same signal stored as complex64; thresholds calibrated to
    sigma_eff = sqrt(sigma^2 + sigma_q^2); compared with the oracle on the
    same complex64-rounded input."""
    n, k, sigma = 1 << 22, 1000, 1e-7
    x, truth = make_signal(n, k, sigma, seed=2, magnitude="loguniform")
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    got, ref, _, _ = _solve_both(gpu, orc, x64, x64.astype(np.complex128), n, k, sig_eff,
                                 seed=11, precision=1)
    assert_coefs_close(got, ref, MIXED_TOL)
    for key in truth:  # every recovered true mode within 1e-6 of itself
        if key in got:
            assert abs(got[key] - ref[key]) <= MIXED_TOL * abs(ref[key])
    hits = [key for key in truth if key in got]
    assert len(hits) >= 0.95 * k


@pytest.mark.parametrize("precision", [0, 1])
def test_solve_back_to_back(gpu, precision):
    """Ten solves of the N=2^22, S=1000 complex64-rounded signal queued back
    to back on one solver with nothing read in between: each starts while
    launches the previous one queued past its stop are still on the stream
    (they must neither report progress nor touch the loop state).  Every
    solve reports the first one's iteration count, and the last leaves the
    first one's iteration log and support."""
    n, k, sigma = 1 << 22, 1000, 1e-7
    x, _ = make_signal(n, k, sigma, seed=2, magnitude="loguniform")
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    s = gpu.Solver(n, gpu.solve_options(k, sig_eff, window_delta=1e-6), precision)
    xd = s.upload(x64 if precision else x64.astype(np.complex128))
    it0 = s.solve(xd, 11)
    m0, b0 = s.modes_dict(), s.log()[0].tolist()
    its = [s.solve(xd, 11) for _ in range(10)]
    assert its == [it0] * 10
    m1 = s.modes_dict()
    assert s.log()[0].tolist() == b0
    assert_coefs_close(m1, m0, 1e-9)


def test_solve_trials3_prime_n(gpu, orc):
    n, k, sigma = 1048573, 100, 1e-6
    x, _ = make_signal(n, k, sigma, seed=3)
    got, ref, _, _ = _solve_both(gpu, orc, x, x, n, k, sigma, seed=13, trials=3)
    assert_coefs_close(got, ref, FP64_TOL)


def test_solve_parity_large_bins(gpu, orc):
    """S = 2200: the first B is 4401 (four-step bin FFT), later ones fit
    k_bin_fft; both kinds run in one captured graph."""
    n, k, sigma = 1 << 18, 2200, 1e-6
    x, truth = make_signal(n, k, sigma, seed=4)
    got, ref, solver, _ = _solve_both(gpu, orc, x, x, n, k, sigma, seed=17)
    assert_coefs_close(got, ref, FP64_TOL)
    assert solver.log()[0][0] == 2 * k + 1
    assert len([key for key in truth if key in got]) >= 0.95 * k


@pytest.mark.timeout(400)  # 20-40 s of host-side signal synthesis at n = 2^26
def test_solve_parity_config3_mixed(gpu, orc):
    """N=2^26 complex64, S=4000 (1000 in 10 clusters of +-64 bins), unit
    magnitudes, sigma=1e-6: first B = 8001, bits = 14, 29 rows."""
    n, k, sigma = 1 << 26, 4000, 1e-6
    x, truth = make_signal(n, k, sigma, seed=3, clusters=(10, 0.25, 64))
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    del x
    x_orc = x64.astype(np.complex128)
    got, ref, solver, _ = _solve_both(gpu, orc, x64, x_orc, n, k, sig_eff, seed=19, precision=1)
    assert_coefs_close(got, ref, MIXED_TOL)
    assert solver.log()[0][0] == 8001
    assert len([key for key in truth if key in got]) >= 0.95 * k
    # PostAnalyze on the device against the oracle's, with x-hat the dense
    # FFT of the complex64 samples the solve saw (the global oracle)
    xh = np.fft.fft(x_orc) / n
    st = solver.post_analyze(list(truth), xh)
    st_o = orc.post_analyze(n, orc.ModeMap(truth), xh, orc.ModeMap(got))
    assert st["hit"] == st_o["hit"] and st["miss"] == st_o["miss"]
    for key in ("hit_l1", "hit_l2", "l1", "l2", "max"):
        assert abs(st[key] - st_o[key]) <= 1e-12 * abs(st_o[key]), key


# ---------------------------------------------------------------- batched form
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_matches_single_and_oracle(gpu, orc, precision):
    """12 signals, n = 2^14..2^18, S = 10..300: the batched solve gives each
    signal the result of its own single-signal solve, and the oracle's
    support and iteration log."""
    rng = np.random.default_rng(40 + precision)
    ns = [1 << int(v) for v in rng.integers(14, 19, 12)]
    ks = [int(v) for v in rng.integers(10, 301, 12)]
    seeds = [int(v) for v in rng.integers(1, 1000, 12)]
    sigma = 1e-6
    xs, sig = [], []
    for i, (n, k) in enumerate(zip(ns, ks)):
        x, _ = make_signal(n, k, sigma, seed=100 + i)
        if precision:
            x64 = x.astype(np.complex64)
            sig.append(float(np.hypot(sigma, quantization_sigma(x, x64))))
            x = x64
        else:
            sig.append(sigma)
        xs.append(x)
    opts = [gpu.solve_options(k, s_) for k, s_ in zip(ks, sig)]
    b = gpu.Batch(ns, opts, precision)
    xd = b.upload(xs)
    iters, stats = b.solve(xd, seeds)
    assert stats == [0] * 12
    before = [b.modes_dict(i) for i in range(12)]
    # every signal's modes copied on the device in one launch, same as read one by one
    for i, (k_, c_) in enumerate(b.modes_all()):
        assert {int(q): complex(v) for q, v in zip(k_, c_)} == before[i]
    tol = FP64_TOL if precision == 0 else MIXED_TOL
    for i in range(12):
        got = b.modes_dict(i)
        s = gpu.Solver(ns[i], opts[i], precision)
        it1 = s.solve(s.upload(xs[i]), seeds[i])
        assert it1 == iters[i]
        assert_coefs_close(got, s.modes_dict(), 1e-12, f"single {i}")
        orc.seed(seeds[i])
        ref, rb, _ = orc.demo1_run(xs[i].astype(np.complex128), ks[i], sig[i], orc.Demo1Options())
        assert np.array_equal(b.log(i)[0], rb)
        assert_coefs_close(got, ref.to_dict(), tol, f"oracle {i}")
    # the tile width chosen by the autotuner changes nothing but speed
    r = b.autotune(xd, seeds)
    assert r in (0, 1, 2, 4, 8)
    iters2, _ = b.solve(xd, seeds)
    assert iters2 == iters
    for i in range(12):  # a different tile width only re-groups the fold's sums
        assert_coefs_close(b.modes_dict(i), before[i], 1e-12, f"retuned {i}")


@pytest.mark.timeout(900)
def test_batch_config4_full_workload(gpu, orc):
    """The batched workload at its full size: 256 complex64 signals,
    n = 2^U{16..22}, S = U{10..500}, unit |c|, sigma = 1e-6, synthesised on
    the GPU; tile width from the autotuner; one batched solve.  Every
    signal's iteration log and support equal its own single-signal solve;
    16 signals spanning n = 2^16..2^22 equal the oracle run on the same
    complex64 samples."""
    import torch
    from mpsft_amd.signals import make_signal_device
    count, sigma = 256, 1e-6
    rng = np.random.default_rng(4)
    ns = [1 << int(v) for v in rng.integers(16, 23, count)]
    ks = [int(v) for v in rng.integers(10, 501, count)]
    seeds = [int(v) for v in rng.integers(1, 1 << 30, count)]
    print(f"[config4] synthesising {count} signals, {sum(ns) / 1e6:.0f} M samples", flush=True)
    xs, opts = [], []
    for i, (n, k) in enumerate(zip(ns, ks)):
        x, sq, _ = make_signal_device(n, k, sigma, seed=7000 + i)
        xs.append(x)
        opts.append(gpu.solve_options(k, float(np.hypot(sigma, sq)), window_delta=1e-6))
    torch.cuda.synchronize()
    b = gpu.Batch(ns, opts, gpu.MPSFT_MIXED)
    tile = b.autotune(xs, seeds)
    assert tile in (0, 1, 2, 4, 8)
    iters, stats = b.solve(xs, seeds)
    assert stats == [0] * count
    print(f"[config4] batched solve done (tile {tile}); single-signal solves", flush=True)
    got = [b.modes_dict(i) for i in range(count)]
    for i in range(count):
        s = gpu.Solver(ns[i], opts[i], gpu.MPSFT_MIXED)
        assert s.solve(xs[i], seeds[i]) == iters[i], i
        assert np.array_equal(s.log()[0], b.log(i)[0]), i
        assert_coefs_close(got[i], s.modes_dict(), 1e-12, f"single {i}")
        del s
    # the oracle on 16 signals: the two smallest S at every n, 2^16 .. 2^22
    by_n = {}
    for i in sorted(range(count), key=lambda j: ks[j]):
        by_n.setdefault(ns[i], []).append(i)
    subset = [i for n in sorted(by_n) for i in by_n[n][:2]]
    subset += [i for i in by_n[max(by_n)][2:2 + 16 - len(subset)]]
    assert len(subset) >= 16 and {ns[i] for i in subset} == {1 << e for e in range(16, 23)}
    for i in subset:
        print(f"[config4] oracle signal {i}: n={ns[i]} k={ks[i]}", flush=True)
        x_orc = xs[i].cpu().numpy().astype(np.complex128)
        orc.seed(seeds[i])
        ref, rb, _ = orc.demo1_run(x_orc, ks[i], opts[i].sigma, orc.Demo1Options())
        assert np.array_equal(b.log(i)[0], rb), i
        assert_coefs_close(got[i], ref.to_dict(), MIXED_TOL, f"oracle {i}")


def test_batch_more_signals_than_plan_threads(gpu):
    """1100 small complex64 signals in one batched solve: more signals than
    the batched plan kernel's threads (1024), so each plan thread starts the
    iterations of two signals and the work maps are filled from device
    memory.  Every signal stops normally, and a spread of 24 signals equal
    their single-signal solves."""
    import torch
    from mpsft_amd.signals import make_signal_device
    count, sigma = 1100, 1e-6
    rng = np.random.default_rng(11)
    ns = [1 << int(v) for v in rng.integers(12, 15, count)]
    ks = [int(v) for v in rng.integers(1, 21, count)]
    seeds = [int(v) for v in rng.integers(1, 1 << 30, count)]
    xs, opts = [], []
    for i, (n, k) in enumerate(zip(ns, ks)):
        x, sq, _ = make_signal_device(n, k, sigma, seed=9000 + i)
        xs.append(x)
        opts.append(gpu.solve_options(k, float(np.hypot(sigma, sq))))
    torch.cuda.synchronize()
    b = gpu.Batch(ns, opts, gpu.MPSFT_MIXED)
    iters, stats = b.solve(xs, seeds)
    assert stats == [0] * count
    for i in np.linspace(0, count - 1, 24).astype(int):
        s = gpu.Solver(ns[i], opts[i], gpu.MPSFT_MIXED)
        assert s.solve(xs[i], seeds[i]) == iters[i], i
        assert np.array_equal(s.log()[0], b.log(i)[0]), i
        assert_coefs_close(b.modes_dict(i), s.modes_dict(), 1e-12, f"single {i}")
        del s


# ---------------------------------------------------------------- edge cases
def test_solve_zero_modes_and_errors(gpu):
    n = 1 << 12
    x = np.zeros(n, np.complex128)
    s = gpu.Solver(n, gpu.solve_options(0, 1e-6))
    xd = s.upload(x)
    it = s.solve(xd, 1)
    assert it == 3 and s.modes()[0].size == 0  # three stale iterations at B=101
    with pytest.raises(gpu.MpsftError):
        gpu.Solver(n, gpu.solve_options(5, 1e-6, trials=2)).solve(xd, 1)  # even trials
    with pytest.raises(gpu.MpsftError):
        gpu.Solver(n, gpu.solve_options(5, 1e-6, window_threshold=0.0)).solve(xd, 1)
    with pytest.raises(gpu.MpsftError):
        gpu.Solver(n, gpu.solve_options(5, 1e-6, min_bins=100)).solve(xd, 1)
    bad = gpu.iterate_options(bins=11)
    bad.bin_in_time_type = 4  # a reference CPU binner type
    with pytest.raises(gpu.MpsftError):
        s.iterate(xd, bad, gpu.transform_make(n, 3, 0, 0), [0])


def test_capacity_overflow_is_reported(gpu):
    n, k = 1 << 14, 50
    x, _ = make_signal(n, k, 1e-6, seed=5)
    s = gpu.Solver(n, gpu.solve_options(k, 1e-6), capacity=8)
    with pytest.raises(gpu.MpsftError) as e:
        s.solve(s.upload(x), 3)
    assert e.value.status == 6


def test_modes_set_get_roundtrip(gpu):
    n = 1 << 16
    s = gpu.Solver(n, gpu.solve_options(10, 1e-6))
    d = {5: 1 + 2j, 70000 % n: -0.5j, n - 1: 3.0 + 0j}
    s.set_modes(d)
    assert s.modes_dict() == d
    s.set_modes({})
    assert s.modes_dict() == {}
    # keys equal modulo n are one ModeMap entry: refused, and the map is kept
    s.set_modes(d)
    import ctypes as C
    keys = np.array([3, 3 + n], np.int64)
    coefs = np.array([1.0, 0.0, 2.0, 0.0])
    rc = gpu.lib().mpsft_solver_modes_set(s.h, keys.ctypes.data_as(C.POINTER(C.c_int64)),
                                          coefs.ctypes.data_as(C.POINTER(C.c_double)), 2)
    assert rc == 1  # MPSFT_ERR_INVALID_ARG
    assert s.modes_dict() == d


# ---------------------------------------------------------------- Demo1::Run schedule
@pytest.mark.parametrize("case", range(len(STALE_THEN_FOUND)))
def test_solve_schedule_stale_then_found(gpu, orc, case):
    """Solves whose schedule finds modes after a stale iteration: the device
    loop's per-iteration log (bins, |found|) equals the oracle's, satisfies
    the Demo1::Run recurrences on its own, and ends where the oracle's
    never-reset stale count ends it."""
    n, k, sigma, min_bins, max_stale, sseed, seed = STALE_THEN_FOUND[case]
    x, _ = make_signal(n, k, sigma, seed=sseed)
    kw = dict(min_bins=min_bins, max_stale_iter=max_stale)
    got, ref, solver, _ = _solve_both(gpu, orc, x, x, n, k, sigma, seed=seed, **kw)
    assert_coefs_close(got, ref, FP64_TOL)
    g_bins, g_found = solver.log()
    check_demo1_schedule(k, sigma, min_bins, max_stale, g_bins, g_found)
    orc.seed(seed)
    _, lg = orc.demo1_run_log(x, k, sigma, orc.Demo1Options(**kw))
    res = lg["res"].astype(np.int64)
    assert res[int(np.argmin(res)):].any()  # a finding iteration after a stale one
    check_demo1_schedule(k, sigma, min_bins, max_stale, g_bins, g_found, res)


# ---------------------------------------------------------------- determinism
@pytest.mark.parametrize("precision", [0, 1])
def test_solve_is_bitwise_reproducible(gpu, precision):
    """Two solvers solving the N=2^22, S=1000 signal with the same seed give
    the same iteration log and bit-identical coefficients: each bin's
    BinInFreq terms are summed exactly (integer limbs), so neither the order
    the terms reach a bin nor the order new keys join the map changes a
    result."""
    n, k, sigma = 1 << 22, 1000, 1e-7
    x, _ = make_signal(n, k, sigma, seed=2, magnitude="loguniform")
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    xs = x64 if precision else x64.astype(np.complex128)
    runs = []
    for _ in range(2):
        s = gpu.Solver(n, gpu.solve_options(k, sig_eff), precision)
        s.solve(s.upload(xs), 11)
        runs.append((s.log()[0].tolist(), s.modes_dict()))
    s.solve(s.upload(xs), 11)  # the same solver again
    runs.append((s.log()[0].tolist(), s.modes_dict()))
    assert runs[0][0] == runs[1][0] == runs[2][0]
    assert runs[0][1] == runs[1][1] == runs[2][1]  # exact equality of every coefficient
