#!/usr/bin/env python3
"""MPSFT recovery on one MI355X: microseconds per full solve (Demo1::Run to
its stale stop) at N=2^22, S=1000 modes with log-uniform magnitudes in
[0.1, 10] and noise sigma=1e-7.  The timed form is MIXED (complex64 input,
thresholds calibrated to sigma_eff = sqrt(sigma^2 + sigma_q^2)); the FP64
form on the same complex64-rounded input, and the plain FP64 form on the
complex128 input with sigma, are timed beside it.

A step is one solve; inputs are resident in device memory before timing.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

N, K, SIGMA, DELTA = 1 << 22, 1000, 1e-7, 1e-6
SIGNAL_SEED, SOLVE_SEED = 2, 11
BATCH_SEED0 = 100  # signal seeds of the batched form
HBM_PEAK_GBS = 8000.0
METRIC = "µs per MPSFT solve N=2^22 S=1000: mixed vs FP64 GPU at 1e-6 coef err"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def time_solves(M, solver, xd, steps, warmup):
    for _ in range(warmup):
        solver.solve(xd, SOLVE_SEED)
    M.synchronize()
    t0 = time.perf_counter()
    iters = 0
    for _ in range(steps):
        iters = solver.solve(xd, SOLVE_SEED)
    M.synchronize()
    return (time.perf_counter() - t0) / steps * 1e6, iters


# Bytes an HBM read request of a random gather moves on MI355X: every 8-byte
# gather is served as one 128-byte line (rocprofv3 TCC_EA0_RDREQ_128B, with
# plain, non-temporal and sc0/sc1 loads alike).
LINE = 128


def gather_bytes(M, solver, elem_bytes, n=N, delta=DELTA, per_launch=None):
    """Algorithmic bytes of every k_bin_time launch of the last solve: LINE
    bytes per distinct LINE-byte line of x its gathers touch, plus the fold
    it writes (rows x B double2).  Also the reference's logical gather
    bytes.  per_launch (a list) receives each launch's algorithmic bytes."""
    bins_log, _ = solver.log()
    tfs, qs = solver.transforms()
    line_bytes = logical = 0
    per_line = LINE // elem_bytes
    seen = np.zeros(n // per_line, dtype=bool)
    for it, B in enumerate(bins_log):
        B = int(B)
        before = line_bytes
        P = M.lib().mpsft_window_p(n, B, delta)
        p2 = (P - 1) // 2
        bits = M.num_bits(n, B)
        a, c = int(tfs[it, 0]), int(tfs[it, 2])
        rows = 1 + 2 * bits
        for q in qs[it]:
            offs = [0] + [s * B * (1 << k) for k in range(bits) for s in (1, -1)]
            seen[:] = False
            for o in offs:
                u = np.arange(o - p2, o + p2 + 1, dtype=np.int64) + int(q)
                seen[((a * (u % n) + c) % n) // per_line] = True
            line_bytes += LINE * int(seen.sum()) + rows * B * 16
            logical += rows * P * elem_bytes
        if per_launch is not None:
            per_launch.append(line_bytes - before)
    return line_bytes, logical, len(bins_log)


LDS_PEAK_GBS = 256 * 128 * 2.4  # 256 CUs x 128 B/clk x 2.4 GHz (CDNA LDS read/write port)
FP64_PEAK_TFS = 78.6


def fft_lds_bytes(B):
    """LDS bytes one k_bin_fft workgroup moves for one row's Bluestein DFT
    of B points over M = 2^ceil(log2(2B-1)) (the count of its loads and
    stores of 16-byte elements).  M <= 1024: twiddles staged (M/2), the
    bit-reversed input (M), radix-4 passes (each M loads + M stores + 3M/4
    twiddle loads) forward and inverse, the filter product (2M), the result
    (2B) and its read-out (B).  M = 2048, 4096: three Stockham passes each
    way (the forward one starts from registers: 6 stores, 5 loads of M), the
    chirp (B stored, B read), the result (2B) and its read-out (B)."""
    logM = max(1, int(np.ceil(np.log2(2 * B - 1))))
    Mm = 1 << logM
    if logM <= 10:
        passes = (logM + 1) // 2
        el = Mm / 2 + Mm + 2 * passes * (2 * Mm + 0.75 * Mm) + 2 * Mm + 3 * B
    else:
        el = 11 * Mm + 5 * B
    return el * 16, 2 * 5 * Mm * logM  # bytes, FFT flops (two M-point FFTs)


def fft_throughput(M, solver, fft_ms):
    """The bin FFT of the last (instrumented) solve against the LDS and FP64
    peaks: model LDS bytes and FFT flops of every launch (rows x trials
    Bluestein DFTs) over the event-timed k_bin_fft time."""
    bins_log, _ = solver.log()
    lds = flops = 0.0
    for B in bins_log:
        B = int(B)
        rows = 1 + 2 * M.num_bits(N, B)
        b, f = fft_lds_bytes(B)
        lds += rows * b
        flops += rows * f
    t = fft_ms / 1000.0
    busy_cus = 1 + 2 * M.num_bits(N, int(bins_log[0]))
    return {"launches": len(bins_log), "avg_launch_us": round(fft_ms * 1000.0 / len(bins_log), 2),
            "lds_model_bytes_per_solve": int(lds), "lds_GBs": round(lds / t / 1e9, 1),
            "lds_peak_GBs": LDS_PEAK_GBS, "lds_frac": round(lds / t / 1e9 / LDS_PEAK_GBS, 4),
            "lds_frac_of_busy_cus": round(lds / t / 1e9 / (LDS_PEAK_GBS * busy_cus / 256), 4),
            "fft_GFLOPs": round(flops / t / 1e9, 1),
            "fft_frac_of_fp64_peak": round(flops / t / 1e12 / FP64_PEAK_TFS, 5)}


def large_workload(M, steps, warmup):
    """N=2^26 complex64, S=4000 unit modes (1000 within +-64 bins of 10
    centres), sigma=1e-6, MIXED: us per solve and the gather-fold's bytes/s."""
    from mpsft_amd.signals import make_signal, quantization_sigma
    n, k, sigma = 1 << 26, 4000, 1e-6
    log("generating the N=2^26, S=4000 signal")
    x, truth = make_signal(n, k, sigma, seed=3, clusters=(10, 0.25, 64))
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(sigma, quantization_sigma(x, x64)))
    del x
    s = M.Solver(n, M.solve_options(k, sig_eff, window_delta=DELTA), M.MPSFT_MIXED)
    xd = s.upload(x64)
    M.synchronize()
    log("timing the N=2^26 solve")
    us, iters = time_solves(M, s, xd, steps, warmup)
    got = s.modes_dict()
    hits = sum(1 for key in truth if key in got)
    err = max(abs(got[key] - v) / abs(v) for key, v in truth.items() if key in got)
    _, ms = s.solve_profiled(xd, SOLVE_SEED)
    bins_log, _ = s.log()
    per_launch = []
    line_bytes, logical, n_launch = gather_bytes(M, s, 8, n=n, per_launch=per_launch)
    # the FP64 form on the same complex64-rounded samples
    del xd
    s64 = M.Solver(n, M.solve_options(k, sig_eff, window_delta=DELTA), M.MPSFT_FP64)
    xd64 = s64.upload(x64.astype(np.complex128))
    M.synchronize()
    log("timing the N=2^26 FP64 solve on the same input")
    us64, iters64 = time_solves(M, s64, xd64, steps, warmup)
    got64 = s64.modes_dict()
    rel64 = max((abs(got[q] - v) / abs(v) for q, v in got64.items() if q in got), default=None)
    t = ms[0] / 1000.0
    return {
        "fp64_same_input_us": round(us64, 1), "fp64_same_input_iterations": iters64,
        "mixed_vs_fp64_same_input": round(us64 / us, 3),
        "support_identical_mixed_vs_fp64": set(got) == set(got64),
        "max_rel_coef_diff_mixed_vs_fp64": rel64,
        "workload": "MPSFT solve N=2^26 S=4000 (25% in 10 clusters of +-64) unit |c| sigma=1e-6, "
                    "complex64 MIXED",
        "us_per_solve": round(us, 1), "solves_per_sec": round(1e6 / us, 1), "iterations": iters,
        "first_bins": int(bins_log[0]), "sigma_eff": sig_eff, "true_modes_recovered": hits,
        "max_rel_coef_err_vs_truth": err,
        "kernel_ms_per_solve": {"k_bin_time": round(ms[0], 4), "bin_fft": round(ms[1], 4),
                                "k_identify": round(ms[2], 4)},
        "bin_in_time": {"algorithmic_bytes_per_solve": line_bytes,
                        "effective_GBs": round(line_bytes / t / 1e9, 1),
                        "frac_of_8TBs": round(line_bytes / t / 1e9 / HBM_PEAK_GBS, 4),
                        "logical_gather_GBs": round(logical / t / 1e9, 1),
                        "launches": n_launch,
                        "bins_per_launch": [int(b) for b in bins_log],
                        "algorithmic_bytes_per_launch": per_launch},
    }


def batch_workload(M, steps, warmup, count=256):
    """Batched uneven instances: `count` complex64 signals, n = 2^k with k
    uniform in [16, 22], S uniform in [10, 500], unit coefficients with
    uniform phases, sigma = 1e-6, synthesised on the GPU.  Solves/s of one
    batched solve (tile width from the autotuner) against a serial loop of
    single-signal solves of the same signals; default_ms is the same batched
    solve before autotuning (default tile widths, 16 bins per identify unit)."""
    import torch
    from mpsft_amd.signals import make_signal_device
    rng = np.random.default_rng(4)
    ns = [1 << int(v) for v in rng.integers(16, 23, count)]
    ks = [int(v) for v in rng.integers(10, 501, count)]
    seeds = [int(v) for v in rng.integers(1, 1 << 30, count)]
    sigma = 1e-6
    log(f"synthesising {count} signals on the GPU (total {sum(ns) / 1e6:.0f} M samples)")
    xs, opts, truths = [], [], []
    for i, (n, k) in enumerate(zip(ns, ks)):
        x, sq, truth = make_signal_device(n, k, sigma, seed=7000 + i)
        xs.append(x)
        truths.append(truth)
        opts.append(M.solve_options(k, float(np.hypot(sigma, sq)), window_delta=DELTA))
    torch.cuda.synchronize()
    b = M.Batch(ns, opts, M.MPSFT_MIXED)

    def timed():
        for _ in range(warmup):
            b.solve(xs, seeds)
        M.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = b.solve(xs, seeds)
        M.synchronize()
        return (time.perf_counter() - t0) / steps, r

    log("batched solves at the default tile widths and identify bins")
    t_default, _ = timed()
    log("autotuning the batch")
    tile = b.autotune(xs, seeds)
    t_batch, (iters, stats) = timed()
    log("serial loop of single-signal solves")
    solvers = [M.Solver(n, o, M.MPSFT_MIXED) for n, o in zip(ns, opts)]
    for s_, x, sd in zip(solvers, xs, seeds):  # warm: tables, graphs
        s_.solve(x, sd)
    M.synchronize()
    t0 = time.perf_counter()
    for s_, x, sd in zip(solvers, xs, seeds):
        s_.solve(x, sd)
    M.synchronize()
    t_serial = time.perf_counter() - t0
    got = [b.modes_dict(i) for i in range(count)]
    same = sum(1 for i in range(count) if set(got[i]) == set(solvers[i].modes_dict()))
    hits = sum(sum(1 for key in truths[i] if key in got[i]) for i in range(count))
    return {
        "workload": f"{count} complex64 signals, n=2^U[16,22], S=U[10,500], unit |c|, "
                    "sigma=1e-6, MIXED, one batched solve",
        "signals": count, "total_samples": int(sum(ns)), "total_modes": int(sum(ks)),
        "batched_ms": round(t_batch * 1e3, 3), "batched_solves_per_sec": round(count / t_batch, 1),
        "default_ms": round(t_default * 1e3, 3),
        "serial_ms": round(t_serial * 1e3, 3), "serial_solves_per_sec": round(count / t_serial, 1),
        "speedup_vs_serial": round(t_serial / t_batch, 2), "tile_residues": tile,
        "max_iterations": int(max(iters)), "status_ok": int(sum(1 for v in stats if v == 0)),
        "signals_identical_to_serial": same, "true_modes_recovered": hits,
    }


def adjoint_workload(M, steps, warmup, count=32):
    """Differentiable recovery: `count` complex64 signals, n = 2^20, S = 100
    unit modes, sigma = 1e-6; loss = sum |c_hat - c_target|^2 over the
    recovered coefficients with c_target = truth (1 + 0.01 CN(0,1)).  A step
    is the forward (one batched recorded solve of every signal,
    torch.ops.mpsft.recover_batch), the loss, and the backward (one batched
    adjoint of every signal's recovery)."""
    import torch
    from mpsft_amd import autograd as ag
    from mpsft_amd.signals import loss_targets, make_signal
    n, k, sigma = 1 << 20, 100, 1e-6
    dev = torch.device("cuda:0")
    log(f"synthesising {count} complex64 signals n=2^20 S={k}")
    xs, dense_t = [], []
    for i in range(count):
        x, truth = make_signal(n, k, sigma, seed=100 + i, dtype=np.complex64)
        t = loss_targets(truth, seed=200 + i)
        td = torch.zeros(n, dtype=torch.complex128)
        td[torch.tensor(list(t.keys()))] = torch.tensor(list(t.values()), dtype=torch.complex128)
        dense_t.append(td)
        xs.append(torch.from_numpy(x).to(dev).requires_grad_(True))
    # every signal's targets in one vector: signal i's frequency k at i n + k
    dense_t = torch.cat(dense_t).to(dev)
    offs = [i * n for i in range(count)]
    seeds = [300 + i for i in range(count)]
    batch = M.Batch([n] * count, [M.solve_options(k, sigma, window_delta=DELTA)] * count,
                    M.MPSFT_MIXED)
    ag.load()

    def step():
        t0 = time.perf_counter()
        keys, coefs = ag.recover_batch(xs, batch, seeds)
        q = torch.cat([kk + o for kk, o in zip(keys, offs)])
        loss = (torch.cat(coefs) - dense_t[q]).abs().pow(2).sum()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1, float(loss), sum(int(q.numel()) for q in keys)

    for _ in range(warmup):
        step()
        for x in xs:
            x.grad = None
    fw = bw = 0.0
    for _ in range(steps):
        a, b, loss, found = step()
        fw += a
        bw += b
        for x in xs:
            x.grad = None
    fw, bw = fw / steps, bw / steps
    return {
        "metric": "ms per differentiable-recovery step (forward + backward), "
                  f"{count} x complex64 N=2^20 S=100",
        "value": round((fw + bw) * 1e3, 3), "unit": "ms", "higher_is_better": False,
        "steps": steps, "warmup": warmup, "dtype": "c64 input, f64 arithmetic",
        "data": "synthetic",
        "config": {"workload": f"{count} complex64 signals, n=2^20, S=100, unit |c|, sigma=1e-6, "
                               "loss sum |c_hat - c_target|^2, c_target = truth (1 + 0.01 CN(0,1)), "
                               "batched forward and backward"},
        "forward_ms": round(fw * 1e3, 3), "backward_ms": round(bw * 1e3, 3),
        "signals_per_sec": round(count / (fw + bw), 1), "loss": loss, "modes_found": int(found),
    }


NEAR_MARGIN = 1e-9


def first_iteration_decisions(M, x64, sig, seed=SOLVE_SEED):
    """The first iteration's per-bin decisions on the device (B = 2K + 1,
    thresholds 3 sigma/sqrt(B) and sigma/sqrt(B) from sigma_eff, one trial,
    a fixed transform and q) for the MIXED form on x64 and the FP64 form on
    the same complex64-rounded input: each form's decisions and relative
    margins (distance from the nearest threshold a decision compared with)."""
    rng = np.random.default_rng(seed)
    a = int(rng.integers(0, N // 2)) * 2 + 1
    b, c, q = (int(v) for v in rng.integers(0, N, 3))
    B = 2 * K + 1
    bits = M.num_bits(N, B)
    tf = M.transform_make(N, a, b, c)
    out = {}
    for name, prec, xin in (("mixed", M.MPSFT_MIXED, x64),
                            ("fp64", M.MPSFT_FP64, x64.astype(np.complex128))):
        p = M.Plan(N, B, DELTA, bits, prec)
        A = p.bin_in_time(xin, tf, q)
        out[name] = p.identify_estimate(A[None], tf, [q], 3 * sig / np.sqrt(B),
                                        sig / np.sqrt(B), 0.1)
        del p
    return dict(tf=(a, b, c), q=q, B=B, **out)


def margin_summary(d):
    m = np.asarray(d["margin"])
    return {"near_margin_decisions": int(np.sum(m < NEAR_MARGIN)),
            "min_margin": float(m.min()), "found": int(np.sum(d["decide"] == 3))}


def decision_flips(d0, d1):
    """Bins whose decision or located frequency differ between two records."""
    diff = (np.asarray(d0["decide"]) != np.asarray(d1["decide"]))
    found = (np.asarray(d0["decide"]) == 3) & (np.asarray(d1["decide"]) == 3)
    diff |= found & (np.asarray(d0["k1"]) != np.asarray(d1["k1"]))
    return int(np.sum(diff))


def cpu_model():
    """The host CPU's model name (lscpu) and logical core count."""
    import subprocess
    model = "unknown"
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            if ln.startswith("Model name:"):
                model = ln.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return model, os.cpu_count()


def host_threads():
    """Threads for the all-core CPU baseline: the CPU share this process is
    given (OMP_NUM_THREADS where set), at most the logical cores."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_solve_time(x, k, sigma, threads, solves, seed):
    """Mean wall time of `solves` oracle Demo1::Run solves (the FP64 CPU
    restatement of the reference: BinInTime V1 order, BinInFreq V1,
    closed-form pencil) on `threads` threads (OpenMP over bin rows).
    Returns (us per solve, iterations, the last solve's modes)."""
    from oracle import pyoracle as O
    O.set_threads(threads)
    t0 = time.perf_counter()
    iters, modes = 0, {}
    for _ in range(solves):
        O.seed(seed)
        mm, bl, _ = O.demo1_run(x, k, sigma)
        iters = len(bl)
        modes = mm.to_dict()
    dt = (time.perf_counter() - t0) / solves
    O.set_threads(0)
    return dt * 1e6, iters, modes


def oracle_first_iteration(xq, sig, dec):
    """The oracle's Iterate (one trial) on the complex64-rounded input with
    the transform and q of first_iteration_decisions: its decisions and
    margins, for counting flips against the device's."""
    from oracle import pyoracle as O
    B = dec["B"]
    opt = O.IterateOptions(bins=B, window_delta=DELTA, trials=1,
                           bin_threshold=3 * sig / np.sqrt(B), window_threshold=0.1,
                           sv_threshold=sig / np.sqrt(B))
    tf = O.transform_make(N, *dec["tf"])
    _, tr = O.iterate(xq, opt, O.ModeMap(), tf=tf, qs=np.array([dec["q"]], np.int32))
    return tr


def copy_calibration():
    """A plain copy of known size: 256 MiB read and 256 MiB written, for
    scaling the FETCH_SIZE / WRITE_SIZE counters of other kernels."""
    import torch
    src = torch.ones(1 << 26, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4}))


def traffic_from_profile():
    """HBM bytes per k_bin_time_batch<float2> launch measured by rocprofv3
    --pmc (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration), as
    committed in profiles/; None when absent."""
    path = os.path.join(ROOT, "profiles",
                        "k_bin_time_batch_traffic_mpsft_batched_64x_n22_s1000_mixed.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.startswith("traffic_bytes_per_launch:"):
                return int(float(ln.split(":", 1)[1]))
    return None


def time_batched(M, b, xd, seeds_for, steps, warmup):
    """Mean seconds per batched solve; seeds_for(step) gives the step's
    solve seeds (fresh ones per step)."""
    for w in range(warmup):
        b.solve(xd, seeds_for(-1 - w))
    M.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        b.solve(xd, seeds_for(i))
    M.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--autotune", action="store_true",
                    help="autotune the timed form's batch (tile width x identify bins) before "
                         "timing it, on solve seeds the timed steps do not use")
    ap.add_argument("--batch-count", type=int, default=64,
                    help="signals per batched solve (the timed form)")
    ap.add_argument("--cpu-solves", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--roofline-only", action="store_true",
                    help="run only the instrumented MIXED solve (for rocprofv3 runs)")
    ap.add_argument("--solve-only", action="store_true",
                    help="only timed single-signal MIXED solves (for rocprofv3 kernel statistics "
                         "of the solve loop)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the timed batched MIXED solves (for rocprofv3 kernel statistics)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="one 256 MiB device-to-device copy (scales rocprofv3 byte counters)")
    ap.add_argument("--config3", action="store_true",
                    help="only the N=2^26 S=4000 clustered MIXED workload (its own JSON line)")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched workload: 256 uneven signals (its own JSON line)")
    ap.add_argument("--adjoint", action="store_true",
                    help="only the differentiable recovery: 32 signals forward + backward "
                         "(its own JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (every box here has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import mpsft_amd as M
    from mpsft_amd.signals import make_signal, quantization_sigma

    if args.config3:
        print(json.dumps(large_workload(M, args.steps, args.warmup)), flush=True)
        return
    if args.batch:
        print(json.dumps(batch_workload(M, args.steps, args.warmup)), flush=True)
        return
    if args.adjoint:
        print(json.dumps(adjoint_workload(M, args.steps, args.warmup)), flush=True)
        return

    log("generating the N=2^22, S=1000 signal")
    x, truth = make_signal(N, K, SIGMA, seed=SIGNAL_SEED, magnitude="loguniform")
    x64 = x.astype(np.complex64)
    sig_eff = float(np.hypot(SIGMA, quantization_sigma(x, x64)))

    mixed = M.Solver(N, M.solve_options(K, sig_eff, window_delta=DELTA), M.MPSFT_MIXED)
    if args.roofline_only:
        xd64 = mixed.upload(x64)
        it_p, ms = mixed.solve_profiled(xd64, SOLVE_SEED)
        line_bytes, logical, n_launch = gather_bytes(M, mixed, 8)
        print(json.dumps({"k_bin_time_avg_launch_us": ms[0] * 1000.0 / n_launch,
                          "launches": n_launch, "timed_repeats_per_iteration": 8,
                          "algorithmic_bytes_per_launch": line_bytes / n_launch,
                          "kernel_ms_per_solve": ms}), flush=True)
        return
    if args.solve_only:
        xd64 = mixed.upload(x64)
        M.synchronize()
        us, it = time_solves(M, mixed, xd64, args.steps, args.warmup)
        print(json.dumps({"us_per_solve": round(us, 2), "iterations": it, "steps": args.steps,
                          "warmup": args.warmup}), flush=True)
        return

    # ---- the timed form: one batched solve of G config-2 signals per step
    G = args.batch_count
    log(f"generating {G} config-2 signals for the batched form")
    xs64, xs128q, sigs = [], [], []
    for g in range(G):
        xg, _ = make_signal(N, K, SIGMA, seed=BATCH_SEED0 + g, magnitude="loguniform")
        xg64 = xg.astype(np.complex64)
        sigs.append(float(np.hypot(SIGMA, quantization_sigma(xg, xg64))))
        xs64.append(xg64)
        xs128q.append(xg64.astype(np.complex128))
    opts = [M.solve_options(K, s_, window_delta=DELTA) for s_ in sigs]

    def seeds_for(step):  # fresh solve seeds every step (warm-up steps negative)
        return [(SOLVE_SEED + 7919 * (step + 10) + g) & 0x7fffffff for g in range(G)]

    bm = M.Batch([N] * G, opts, M.MPSFT_MIXED)
    xdb = bm.upload(xs64)
    tuned = None
    if args.autotune:
        log("autotuning the batch")
        tuned = bm.autotune(xdb, seeds_for(-1000))
    log("timing the batched MIXED form")
    t_bm = time_batched(M, bm, xdb, seeds_for, args.steps, args.warmup)
    if args.batched_only:
        print(json.dumps({"us_per_solve": round(t_bm / G * 1e6, 2), "signals": G,
                          "steps": args.steps, "warmup": args.warmup,
                          "tile_residues": tuned}), flush=True)
        return
    # identical support against the FP64 single-signal form on the same
    # complex64-rounded input, for every signal of the last step's seeds
    last = seeds_for(args.steps - 1)
    bm_modes = [bm.modes_dict(g) for g in range(G)]
    if not args.no_roofline:
        log("instrumented batched solve (event-timed stages, the last step's seeds)")
        _, _, stage_ms, n_batched = bm.solve_profiled(xdb, last)
    del xdb
    bf = M.Batch([N] * G, opts, M.MPSFT_FP64)
    xdf = bf.upload(xs128q)
    log("timing the batched FP64 form (same inputs)")
    t_bf = time_batched(M, bf, xdf, seeds_for, args.steps, args.warmup)
    del xdf, bf
    same_support, max_rel = 0, 0.0
    line_total = logical_total = 0
    for g in range(G):
        s_chk = M.Solver(N, opts[g], M.MPSFT_FP64)
        s_chk.solve(s_chk.upload(xs128q[g]), last[g])
        ref = s_chk.modes_dict()
        del s_chk
        if set(ref) == set(bm_modes[g]):
            same_support += 1
            max_rel = max([max_rel] + [abs(bm_modes[g][q] - v) / abs(v) for q, v in ref.items()])
        if not args.no_roofline:
            # the signal's gathers in the batched solve are its single MIXED
            # solve's (same seed, same iterations): their line bytes
            s_mx = M.Solver(N, opts[g], M.MPSFT_MIXED)
            s_mx.solve(s_mx.upload(xs64[g]), last[g])
            lb, lg, _ = gather_bytes(M, s_mx, 8)
            line_total += lb
            logical_total += lg
            del s_mx

    # ---- single-signal forms on the headline signal
    fp64q = M.Solver(N, M.solve_options(K, sig_eff, window_delta=DELTA), M.MPSFT_FP64)
    fp64 = M.Solver(N, M.solve_options(K, SIGMA, window_delta=DELTA), M.MPSFT_FP64)
    xd64 = mixed.upload(x64)
    xd128q = fp64q.upload(x64.astype(np.complex128))
    xd128 = fp64.upload(x)
    M.synchronize()
    log("timing the single-signal forms")

    def time_fresh(solver, xd):
        for w in range(args.warmup):
            solver.solve(xd, seeds_for(-1 - w)[0])
        M.synchronize()
        t0 = time.perf_counter()
        iters = []
        for i in range(args.steps):
            iters.append(solver.solve(xd, seeds_for(i)[0]))
        M.synchronize()
        return (time.perf_counter() - t0) / args.steps * 1e6, float(np.mean(iters))

    us_mixed, it_mixed = time_fresh(mixed, xd64)
    us_mixed_fixed, it_mixed_fixed = time_solves(M, mixed, xd64, args.steps, args.warmup)
    us_fp64q, it_fp64q = time_fresh(fp64q, xd128q)
    us_fp64, it_fp64 = time_fresh(fp64, xd128)

    # accuracy at the fixed seed: MIXED vs FP64 on the same input, and truth
    mixed.solve(xd64, SOLVE_SEED)
    fp64q.solve(xd128q, SOLVE_SEED)
    fp64.solve(xd128, SOLVE_SEED)
    gm, gq, g64 = mixed.modes_dict(), fp64q.modes_dict(), fp64.modes_dict()
    rel = [abs(gm[q] - gq[q]) / abs(gq[q]) for q in gq if q in gm]
    hits = sum(1 for q in truth if q in gm)
    truth_err = max(abs(gm[q] - v) / abs(v) for q, v in truth.items() if q in gm)
    us_batch = t_bm / G * 1e6

    line = {
        "metric": METRIC,
        "value": round(us_batch, 2),
        "unit": "us/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_bm * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 input; f64 arithmetic",
        "data": "synthetic",
        "config": {"workload": f"MPSFT solves N=2^22 S=1000 |c| log-uniform [0.1,10] sigma=1e-7, "
                               f"complex64 MIXED, {G} signals per batched solve (a step), fresh "
                               "solve seeds per step",
                   "n": N, "k": K, "sigma": SIGMA, "window_delta": DELTA, "signals_per_step": G,
                   "signal_seeds": [BATCH_SEED0, BATCH_SEED0 + G - 1]},
        "solves_per_sec": round(1e6 / us_batch, 1),
        "batched_fp64_same_input_us": round(t_bf / G * 1e6, 2),
        "mixed_vs_fp64_batched": round(t_bf / t_bm, 3),
        "batched_support_identical_to_single_fp64": f"{same_support}/{G}",
        "batched_max_rel_coef_diff_vs_single_fp64": max_rel,
        "single_solve": {
            "mixed_us_fresh_seeds": round(us_mixed, 2), "mixed_mean_iterations": it_mixed,
            "mixed_us_fixed_seed": round(us_mixed_fixed, 2), "fixed_seed_iterations": it_mixed_fixed,
            "fp64_same_input_us": round(us_fp64q, 2), "fp64_same_input_mean_iterations": it_fp64q,
            "fp64_c128_sigma_us": round(us_fp64, 2), "fp64_c128_sigma_mean_iterations": it_fp64,
            "sigma_eff": sig_eff, "signal_seed": SIGNAL_SEED,
        },
        "mixed_batched_speedup_vs_single_fp64_same_input": round(us_fp64q / us_batch, 3),
        "support_identical_mixed_vs_fp64_single": set(gm) == set(gq),
        "max_rel_coef_diff_mixed_vs_fp64_single": max(rel) if rel else None,
        "true_modes_recovered": hits,
        "max_rel_coef_err_vs_truth": truth_err,
        "true_modes_recovered_plain_fp64": sum(1 for q in truth if q in g64),
    }

    log("first-iteration decision margins (MIXED and FP64 on the same input)")
    dec = first_iteration_decisions(M, x64, sig_eff)
    line["decisions"] = {
        "what": f"first iteration, B={dec['B']}, one trial, fixed transform; near = relative "
                f"margin below {NEAR_MARGIN:g}",
        "mixed": margin_summary(dec["mixed"]), "fp64_same_input": margin_summary(dec["fp64"]),
        "flips_mixed_vs_fp64": decision_flips(dec["mixed"], dec["fp64"]),
    }

    if not args.no_roofline:
        # the timed step's dominant kernel: the batched gather-fold
        t_bt = stage_ms[1] / 1000.0
        achieved = line_total / t_bt / 1e9
        line["roofline"] = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_from_profile(),
            "kernel": "k_bin_time_batch<float2>", "launches": n_batched,
            "avg_launch_us": round(stage_ms[1] * 1000.0 / n_batched, 2),
            "algorithmic_bytes_per_launch": int(line_total / n_batched),
            "line_bytes": LINE,
            "logical_gather_GBs": round(logical_total / t_bt / 1e9, 1),
        }
        line["batched_stage_ms_per_step"] = {
            "plan": round(stage_ms[0], 4), "k_bin_time_batch": round(stage_ms[1], 4),
            "k_bin_fft_batch": round(stage_ms[2], 4), "k_identify_batch": round(stage_ms[3], 4),
            "launches": n_batched}
        log("profiled single solve")
        it_p, ms = mixed.solve_profiled(xd64, SOLVE_SEED)
        line_bytes, logical, n_launch = gather_bytes(M, mixed, 8)
        line["single_solve"]["kernel_ms"] = {
            "k_bin_time": round(ms[0], 4), "k_bin_fft": round(ms[1], 4),
            "k_identify": round(ms[2], 4)}
        line["single_solve"]["k_bin_time_GBs"] = round(line_bytes / (ms[0] / 1000.0) / 1e9, 1)
        line["bin_fft"] = fft_throughput(M, mixed, ms[1])

    if not args.no_cpu_baseline:
        threads = host_threads()
        model, ncpu = cpu_model()
        xq = x64.astype(np.complex128)
        log(f"CPU baseline: oracle solves of the headline workload on {threads} and 1 threads")
        us_all, it_all, orc_modes = cpu_solve_time(xq, K, sig_eff, threads, args.cpu_solves,
                                                   SOLVE_SEED)
        us_one, _, _ = cpu_solve_time(xq, K, sig_eff, 1, max(1, args.cpu_solves - 1), SOLVE_SEED)
        us_plain, it_plain, _ = cpu_solve_time(x, K, SIGMA, 1, 1, SOLVE_SEED)
        # the same solves' results against the device's (same seed, same input)
        sym = {f: len(set(orc_modes) ^ set(g)) for f, g in (("mixed", gm), ("fp64", gq))}
        rel_o = [abs(gq[q] - v) / abs(v) for q, v in orc_modes.items() if q in gq]
        tr = oracle_first_iteration(xq, sig_eff, dec)
        line["decisions"]["oracle"] = {
            "near_margin_decisions": int(np.sum(tr["margin"] < NEAR_MARGIN)),
            "min_margin": float(tr["margin"].min()),
            "flips_vs_mixed": decision_flips(tr, dec["mixed"]),
            "flips_vs_fp64": decision_flips(tr, dec["fp64"]),
            "solve_support_differences": sym,
            "solve_max_rel_coef_diff_fp64": max(rel_o) if rel_o else None,
        }
        line["cpu_baseline"] = {
            "value": round(us_all, 1), "unit": "us/solve", "cores": threads, "kind": "port",
            "sample": f"{args.cpu_solves} solves of the headline signal (complex64-rounded, "
                      f"sigma_eff, solve seed {SOLVE_SEED}, {it_all} iterations) by the oracle, "
                      "OpenMP over bin rows",
            "single_core_us": round(us_one, 1),
            "plain_fp64_single_core_us": round(us_plain, 1),
            "plain_fp64_iterations": it_plain,
            "cpu_model": model, "logical_cpus": ncpu,
        }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
