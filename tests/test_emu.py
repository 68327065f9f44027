"""The library's own HIP sources run on the CPU through the host emulator in
tests/emu (workgroups as fibers, HIP runtime calls as host calls) under
AddressSanitizer: every kernel access to device buffers and LDS is bounds-
and lifetime-checked, and the results are compared with the oracle.

Each case runs in a child process with libasan preloaded.  The emulator's
sincospi is the host's sin/cos of pi*x, so results agree with the oracle to
rounding (not bit for bit as on the GPU); the checks here are memory safety,
identical support / iteration logs, and coefficients to 1e-10.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")


@pytest.fixture(scope="module")
def emu_env():
    if shutil.which("g++") is None or shutil.which("make") is None:
        pytest.skip("g++/make not available")
    r = subprocess.run(["make", "-C", EMU], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("emulator build failed:\n" + r.stderr[-3000:])
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1")
    return env


def run_case(env, timeout=600, **case):
    r = subprocess.run([sys.executable, os.path.join(EMU, "run_case.py"), json.dumps(case)],
                       capture_output=True, text=True, env=env, timeout=timeout)
    assert "AddressSanitizer" not in r.stderr, r.stderr[-6000:]
    assert r.returncode == 0, r.stderr[-3000:]
    line = [x for x in r.stdout.splitlines() if x.startswith("RESULT ")]
    assert line, r.stdout[-2000:]
    return json.loads(line[-1][7:])


# Every case of this module, run as child processes side by side (each is
# single-threaded) the first time a test asks for one.
CASES = {
    "plan": dict(kind="plan", n=4096, bins=101, bits=5, seed=1),
    "plan_four_step": dict(kind="plan", n=1 << 16, bins=4097, bits=4, seed=2),
    "plan_freq": dict(kind="plan_freq", n=1 << 14, bins=301, bits=6, m=900, seed=4),
    "iterate_four_step": dict(kind="iterate", n=1 << 16, k=2100, sigma=1e-6, seed=3, bins=4201),
    "solve_0": dict(kind="solve", n=1 << 14, k=20, sigma=1e-6, seed=7, precision=0, repeat=2),
    "solve_1": dict(kind="solve", n=1 << 14, k=20, sigma=3e-6, seed=7, precision=1, repeat=2),
    "solve_2": dict(kind="solve", n=1 << 16, k=300, sigma=1e-6, seed=5, precision=0, repeat=1),
    "batch": dict(kind="batch", ns=[4096, 8192, 16384, 4096], ks=[5, 12, 20, 8], sigma=1e-6,
                  seeds=[3, 4, 5, 6]),
    # the gather-fold split over the 8 XCD slices of x, forced at small n
    "solve_sliced": dict(kind="solve", n=1 << 14, k=40, sigma=3e-6, seed=7, precision=1,
                         repeat=1, setenv={"MPSFT_SLICES": "8"}),
    "batch_sliced": dict(kind="batch", ns=[4096, 8192, 16384, 4096], ks=[5, 12, 20, 8],
                         sigma=1e-6, seeds=[3, 4, 5, 6], setenv={"MPSFT_SLICES": "8"}),
    "plan_sliced": dict(kind="plan", n=4096, bins=101, bits=5, seed=1,
                        setenv={"MPSFT_SLICES": "8"}),
    "adjoint_16384_20_1": dict(kind="adjoint", n=1 << 14, k=20, sigma=1e-6, seed=7, trials=1),
    "adjoint_16384_20_3": dict(kind="adjoint", n=1 << 14, k=20, sigma=1e-6, seed=7, trials=3),
    "adjoint_65536_300_1": dict(kind="adjoint", n=1 << 16, k=300, sigma=1e-6, seed=7, trials=1),
    "adjoint_10007_10_1": dict(kind="adjoint", n=10007, k=10, sigma=1e-6, seed=7, trials=1),
    "batch_adjoint": dict(kind="batch_adjoint", ns=[4096, 8192, 16384], ks=[5, 12, 20],
                          sigma=1e-6, seeds=[3, 4, 5]),
    "gen_4099": dict(kind="gen", n=4099, k=7, sigma=0.5, seed=77),
    "gen_65536": dict(kind="gen", n=1 << 16, k=30, sigma=1e-6, seed=123534),
    "analyze": dict(kind="analyze", n=1 << 14, k=40, sigma=1e-5, seed=9),
    "identify_1": dict(kind="identify", n=1 << 14, k=60, sigma=1e-6, seed=11, trials=1),
    "identify_3": dict(kind="identify", n=1 << 14, k=60, sigma=1e-6, seed=12, trials=3),
}


@pytest.fixture(scope="module")
def emu_runs(emu_env):
    import concurrent.futures as cf

    def one(case):
        try:
            return run_case(emu_env, **case)
        except AssertionError as e:  # reported by the test that reads the case
            return e
    workers = max(1, min(len(CASES), (os.cpu_count() or 2) - 1))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        futs = {name: ex.submit(one, case) for name, case in CASES.items()}
        return {name: f.result() for name, f in futs.items()}


def result(runs, name):
    r = runs[name]
    if isinstance(r, Exception):
        raise r
    return r


def test_emu_plan_bin_in_time(emu_runs):
    assert result(emu_runs, "plan")["rel_err"] < 1e-12


def test_emu_plan_four_step(emu_runs):
    """B = 4097: M = 16384, the four-step bin FFT (k_big_cols/rows/out)."""
    assert result(emu_runs, "plan_four_step")["rel_err"] < 1e-12


def test_emu_plan_bin_in_freq_exact_sums(emu_runs):
    """A plan's BinInFreq sums each bin's terms exactly: it agrees with the
    oracle to rounding and is bit for bit the same when the modes come in
    another order."""
    r = result(emu_runs, "plan_freq")
    assert r["abs_err"] < 1e-12 * r["scale"] and r["identical"], r


def test_emu_iterate_four_step_with_found_modes(emu_runs):
    """Two Iterate steps at B = 4201: the second subtracts the first's
    modes inside k_big_out."""
    r = result(emu_runs, "iterate_four_step")
    assert r["same_support"] and r["n_found"] > 1000 and r["max_abs_err"] < 1e-10


@pytest.mark.parametrize("case", ["solve_0", "solve_1", "solve_2"])
def test_emu_solve_graph(emu_runs, case):
    """Demo1::Run through the captured two-iteration graphs, on one solver
    (twice for the first two cases), then three times back to back, each
    solve starting behind the launches the previous one queued past its
    stop.  solve_2 (k = 300) starts at B = 601 and goes over from the
    three-launch iterations to the two-launch late form at B = min_bins."""
    r = result(emu_runs, case)
    assert all(r[f"iters{i}"] == r["ref_iters"] for i in range(2) if f"iters{i}" in r)
    assert r["bins"] == r["ref_bins"] and r["same_support"]
    assert r["max_rel_err"] < 1e-10
    assert r["back_to_back_iters"] == [r["iters0"]] * 3 and r["back_to_back_same"]
    assert r["bins"] == r["ref_bins"] and r["same_support"]
    assert r["max_rel_err"] < 1e-10


def test_emu_sliced_gather_fold(emu_runs):
    """The gather-fold split over the 8 XCD slices of x (forced): a plan's
    BinInTime, a single solve through the graphs and the late form, and the
    batched form all agree with the oracle as the unsplit form does."""
    assert result(emu_runs, "plan_sliced")["rel_err"] < 1e-12
    r = result(emu_runs, "solve_sliced")
    assert r["bins"] == r["ref_bins"] and r["same_support"] and r["max_rel_err"] < 1e-10
    r = result(emu_runs, "batch_sliced")
    assert r["status"] == [0, 0, 0, 0]
    assert all(r["same_as_single"]) and all(r["same_as_oracle"])


def test_emu_batch_matches_single_and_oracle(emu_runs):
    """The batched form on 4 signals of different n and S: each signal's
    result equals its own single-signal solve and the oracle."""
    r = result(emu_runs, "batch")
    assert r["status"] == [0, 0, 0, 0]
    assert all(r["same_as_single"]) and all(r["same_as_oracle"])
    assert r["modes_all_match"]  # the device-side read-back of every signal's modes


@pytest.mark.parametrize("n,k,trials", [(1 << 14, 20, 1), (1 << 14, 20, 3), (1 << 16, 300, 1),
                                         (10007, 10, 1)])
def test_emu_replay_and_adjoint(emu_runs, n, k, trials):
    """A recorded solve replayed on its own input gives the solve's modes;
    the replayed map is linear; the adjoint kernels satisfy
    <F d, g> = <d, F^H g>.  k = 300 takes B = 601 through the register
    Stockham bin FFT; the prime n takes the 64-bit index path of the
    gather."""
    r = result(emu_runs, f"adjoint_{n}_{k}_{trials}")
    assert r["n_found"] >= k and r["same_keys"] and r["replay_err"] == 0.0
    assert r["linear_err"] < 1e-13 and r["dot_rel"] < 1e-12


def test_emu_batched_adjoint(emu_runs):
    """The batched adjoint (one launch per stage over 3 signals of different
    n and S): the dot-product identity per signal, and the same output as
    each signal's single-signal adjoint."""
    r = result(emu_runs, "batch_adjoint")
    assert r["status"] == [0, 0, 0]
    assert max(r["dot_rel"]) < 1e-12 and max(r["vs_single"]) < 1e-12


@pytest.mark.parametrize("case", ["gen_4099", "gen_65536"])
def test_emu_generate_signal(emu_runs, case):
    """Demo1's signal from the reference's RNG stream (GenerateModeMap on the
    host, GenerateXhatAlt on the device) equals the oracle's bit for bit."""
    r = result(emu_runs, case)
    assert r["rc"] == 0 and r["same_modes"] and r["same_xhat"]


def test_emu_post_analyze(emu_runs):
    """PostAnalyze on the device against the oracle's, and its static-scratch
    form bit for bit against the allocating one."""
    r = result(emu_runs, "analyze")
    assert r["hit"] == r["hit_o"] == 40 and r["miss"] == r["miss_o"]
    assert r["rel"] < 1e-12
    assert r["noalloc_rc"] == 0 and r["noalloc_same"] and r["noalloc_flag"] == 0.0


@pytest.mark.parametrize("case", ["identify_1", "identify_3"])
def test_emu_identify_estimate(emu_runs, case):
    """mpsft_identify_estimate on the oracle's bin matrices: decisions and
    k1 equal, coefficients and decision margins to rounding."""
    r = result(emu_runs, case)
    assert r["same_decide"] and r["same_k1"] and r["n_found"] > 10
    assert r["coef_err"] < 1e-12 and r["margin_err"] < 1e-12


@pytest.mark.skipif(os.environ.get("MPSFT_SLOW_TESTS") != "1",
                    reason="about 15 minutes; set MPSFT_SLOW_TESTS=1")
def test_emu_solve_past_initial_transform_draw(emu_env, tmp_path):
    """A solve longer than the 256 transforms drawn up front: the arrays grow
    while the graph runs.  Before the fix the graph kept the freed arrays'
    addresses (a device use-after-free); this case ran 357 iterations."""
    from mpsft_amd.signals import make_signal
    x, _ = make_signal(1 << 20, 50, 1e-8, seed=1)
    path = str(tmp_path / "x.npy")
    np.save(path, x)
    r = run_case(emu_env, timeout=3600, kind="solve", n=1 << 20, k=50, sigma=1e-8, seed=7,
                 x_path=path, repeat=2)
    assert r["iters0"] == r["iters1"] == r["ref_iters"] == 357
    assert r["bins"] == r["ref_bins"] and r["same_support"]
