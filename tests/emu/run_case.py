"""Runs one scenario of the library on the host emulator (tests/emu) and
prints a JSON summary.  Started by tests/test_emu.py in a child process
with AddressSanitizer preloaded; the kernels' memory accesses are checked
by ASan, the results against the oracle."""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from mpsft_amd import _lib  # noqa: E402

_lib._lib = _lib.declare(C.CDLL(os.path.join(HERE, "_build", "libmpsft_emu.so")))

import mpsft_amd as gpu  # noqa: E402
from oracle import pyoracle as orc  # noqa: E402


def signal(n, k, sigma, seed):
    rng = np.random.default_rng(seed)
    locs = rng.choice(n, size=k, replace=False)
    xh = np.zeros(n, np.complex128)
    xh[locs] = np.exp(2j * np.pi * rng.uniform(0, 1, k))
    xh += (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * sigma / np.sqrt(2 * n)
    return np.fft.ifft(xh) * n


def solve(n, k, sigma, seed, precision=0, repeat=1, x_path=None, **kw):
    x = np.load(x_path) if x_path else signal(n, k, sigma, seed)
    xg = x.astype(np.complex64) if precision else x
    xo = xg.astype(np.complex128)
    s = gpu.Solver(n, gpu.solve_options(k, sigma, **kw), precision)
    xd = s.upload(xg)
    out = {}
    for r in range(repeat):
        it = s.solve(xd, seed)
        bins, found = s.log()
        got = s.modes_dict()
        out[f"iters{r}"] = it
    # Back to back, nothing read in between, the host launching up to 6
    # iterations ahead and never falling back to waiting on the stream: each
    # solve starts while launches the previous one queued past its stop
    # still wait on the device, and must neither see nor be disturbed by them.
    os.environ.update(MPSFT_LOOKAHEAD="6", MPSFT_STALL_MS="600000")
    out["back_to_back_iters"] = [s.solve(xd, seed) for _ in range(3)]
    del os.environ["MPSFT_LOOKAHEAD"], os.environ["MPSFT_STALL_MS"]
    out["back_to_back_same"] = s.modes_dict() == got and s.log()[0].tolist() == bins.tolist()
    orc.seed(seed)
    ref, rb, rf = orc.demo1_run(xo, k, sigma, orc.Demo1Options(**kw))
    ref = ref.to_dict()
    err = max((abs(got[key] - ref[key]) / max(abs(ref[key]), 1e-3) for key in ref if key in got),
              default=0.0)
    out.update(iters=it, ref_iters=len(rb), bins=bins.tolist(), ref_bins=list(map(int, rb)),
               same_support=set(got) == set(ref), n_found=len(got), max_rel_err=err)
    return out


def iterate(n, k, sigma, seed, bins, steps=2):
    x = signal(n, k, sigma, seed)
    kw = dict(bins=bins, window_delta=1e-6, trials=1, bin_threshold=3 * sigma / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sigma / np.sqrt(bins))
    s = gpu.Solver(n, gpu.solve_options(1, 0.0))
    xd = s.upload(x)
    mm = orc.ModeMap()
    orc.seed(seed)
    errs = []
    for _ in range(steps):
        tf = orc.transform_draw(n)
        q = orc.rand_int32() % n
        orc.iterate(x, orc.IterateOptions(**kw), mm, tf=tf, qs=np.array([q], np.int32))
        s.iterate(xd, gpu.iterate_options(**kw), gpu.transform_make(n, *map(int, tf[:3])), [q])
        got, ref = s.modes_dict(), mm.to_dict()
        errs.append(max((abs(got[key] - ref[key]) for key in ref if key in got), default=0.0))
    return dict(same_support=set(got) == set(ref), n_found=len(got), max_abs_err=max(errs))


def plan(n, bins, bits, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    orc.seed(seed)
    tf = orc.transform_draw(n)
    q = int(rng.integers(0, n))
    ref = orc.bin_in_time(x, bins, 1e-6, bits, tf, q, version=0)
    got = gpu.Plan(n, bins, 1e-6, bits).bin_in_time(x, gpu.transform_make(n, *map(int, tf[:3])), q)
    return dict(rel_err=float(np.abs(got - ref).max() / np.abs(ref).max()))


def plan_freq(n, bins, bits, m, seed):
    """BinInFreq on a plan: against the oracle, and bit for bit the same
    after the modes are permuted (the per-bin sums are exact)."""
    rng = np.random.default_rng(seed)
    keys = rng.choice(n, m, replace=False).astype(np.int64)
    coefs = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    coefs *= np.exp(rng.uniform(-8, 0, m))
    base = rng.standard_normal((1 + 2 * bits, bins)) + 1j * rng.standard_normal((1 + 2 * bits, bins))
    orc.seed(seed)
    tf = orc.transform_draw(n)
    q = int(rng.integers(0, n))
    ref = orc.bin_in_freq(keys, coefs, n, bins, 1e-6, bits, tf, q, base.copy(), version=0)
    p = gpu.Plan(n, bins, 1e-6, bits)
    gtf = gpu.transform_make(n, *map(int, tf[:3]))
    got = p.bin_in_freq(keys, coefs, gtf, q, base.copy())
    perm = rng.permutation(m)
    got2 = p.bin_in_freq(keys[perm], coefs[perm], gtf, q, base.copy())
    return dict(abs_err=float(np.abs(got - ref).max()), scale=float(np.abs(ref).max()),
                identical=bool(np.array_equal(got.view(np.float64), got2.view(np.float64))))


def batch(ns, ks, sigma, seeds, precision=0, autotune=False):
    xs = [signal(n, k, sigma, sd) for n, k, sd in zip(ns, ks, seeds)]
    xg = [x.astype(np.complex64) if precision else x for x in xs]
    opts = [gpu.solve_options(k, sigma) for k in ks]
    b = gpu.Batch(ns, opts, precision)
    xd = b.upload(xg)
    tuned = b.autotune(xd, seeds) if autotune else None
    iters, stats = b.solve(xd, seeds)
    out = {"iters": iters, "status": stats, "tuned": tuned, "same_as_single": [], "same_as_oracle": []}
    allm = b.modes_all()
    out["modes_all_match"] = all(
        {int(k): complex(c) for k, c in zip(*allm[i])} == b.modes_dict(i) for i in range(len(ns)))
    for i, (n, k, sd) in enumerate(zip(ns, ks, seeds)):
        got = b.modes_dict(i)
        s = gpu.Solver(n, opts[i], precision)
        xs1 = s.upload(xg[i])
        it1 = s.solve(xs1, sd)
        one = s.modes_dict()
        orc.seed(sd)
        ref, rb, _ = orc.demo1_run(xg[i].astype(np.complex128), k, sigma, orc.Demo1Options())
        ref = ref.to_dict()
        e1 = max((abs(got[key] - one[key]) for key in one if key in got), default=0.0)
        out["same_as_single"].append(bool(set(got) == set(one) and it1 == iters[i] and e1 < 1e-12))
        out["same_as_oracle"].append(bool(set(got) == set(ref) and len(rb) == iters[i]))
    return out


def adjoint(n, k, sigma, seed, trials=1):
    """Record a solve, replay it, and check the adjoint by the dot-product
    identity <F dx, g> = <dx, F^H g>."""
    x = signal(n, k, sigma, seed)
    s = gpu.Solver(n, gpu.solve_options(k, sigma, trials=trials))
    s.set_record(True)
    xd = s.upload(x)
    iters = s.solve(xd, seed)
    keys, c0 = s.modes()
    s.replay(xd)
    r0 = s.modes_dict()
    rng = np.random.default_rng(seed + 1)
    dx = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    s.replay(s.upload(dx))
    fdx = s.modes_dict()
    s.replay(s.upload(x + dx))
    fsum = s.modes_dict()
    g = rng.standard_normal(keys.size) + 1j * rng.standard_normal(keys.size)
    adj = s.adjoint(keys, g)
    fdx_v = np.array([fdx[int(q)] for q in keys])
    lhs, rhs = np.vdot(g, fdx_v), np.vdot(adj, dx)
    return dict(iters=iters, n_found=int(keys.size), same_keys=set(r0) == set(map(int, keys)),
                replay_err=max(abs(r0[int(q)] - v) for q, v in zip(keys, c0)),
                linear_err=max(abs(fsum[int(q)] - r0[int(q)] - fdx[int(q)]) for q in keys),
                dot_rel=float(abs(lhs - rhs) / abs(lhs)), dot=abs(lhs))


def batch_adjoint(ns, ks, sigma, seeds, precision=0):
    """Batched record + adjoint: per signal the dot-product identity with the
    replayed map, and agreement with the single-signal adjoint."""
    xs = [signal(n, k, sigma, sd) for n, k, sd in zip(ns, ks, seeds)]
    xg = [x.astype(np.complex64) if precision else x for x in xs]
    opts = [gpu.solve_options(k, sigma) for k in ks]
    b = gpu.Batch(ns, opts, precision)
    b.set_record(True)
    xd = b.upload(xg)
    iters, stats = b.solve(xd, seeds)
    rng = np.random.default_rng(5)
    keys_l, g_l, fd_l, d_l = [], [], [], []
    for i, n in enumerate(ns):
        keys = np.array(sorted(b.modes_dict(i)), np.int64)
        d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        b.replay(i, gpu.DeviceArray.from_numpy(d), precision=gpu.MPSFT_FP64)
        fd = b.modes_dict(i)
        keys_l.append(keys)
        g_l.append(rng.standard_normal(keys.size) + 1j * rng.standard_normal(keys.size))
        fd_l.append(np.array([fd[int(q)] for q in keys]))
        d_l.append(d)
    adj = b.adjoint(keys_l, g_l)
    dot, single = [], []
    for i, n in enumerate(ns):
        lhs, rhs = np.vdot(g_l[i], fd_l[i]), np.vdot(adj[i], d_l[i])
        dot.append(float(abs(lhs - rhs) / abs(lhs)))
        s = gpu.Solver(n, opts[i], precision)
        s.set_record(True)
        s.solve(s.upload(xg[i]), seeds[i])
        a1 = s.adjoint(keys_l[i], g_l[i])
        single.append(float(np.abs(a1 - adj[i]).max() / np.abs(a1).max()))
    return dict(iters=iters, status=stats, dot_rel=dot, vs_single=single)


def gen(n, k, sigma, seed):
    """mpsft_generate_signal against the oracle's GenerateModeMap /
    GenerateXhatAlt on the same seed (the emulator's log and sqrt are the
    host's, so x-hat must agree bit for bit)."""
    xd = gpu.DeviceArray(16 * n)
    keys = np.zeros(k, np.int64)
    coefs = np.zeros(2 * k)
    rc = gpu.lib().mpsft_generate_signal(n, k, sigma, seed, xd.ptr,
                                         keys.ctypes.data_as(C.POINTER(C.c_int64)),
                                         coefs.ctypes.data_as(C.POINTER(C.c_double)), None)
    got = xd.to_numpy(np.complex128, (n,))
    orc.seed(seed)
    mm = orc.generate_mode_map(n, k)
    xh = orc.generate_xhat_alt(n, mm, sigma)
    return dict(rc=rc, same_modes=dict(zip(keys.tolist(), coefs.view(np.complex128).tolist()))
                == mm.to_dict(), same_xhat=bool(np.array_equal(got, xh)))


def analyze(n, k, sigma, seed):
    """A solve, then PostAnalyze on the device (allocating and static-scratch
    forms) against the oracle's on the same found modes and x-hat."""
    rng = np.random.default_rng(seed)
    locs = rng.choice(n, size=k, replace=False)
    xh = np.zeros(n, np.complex128)
    xh[locs] = np.exp(2j * np.pi * rng.uniform(0, 1, k))
    truth = {int(q): complex(xh[q]) for q in locs}
    xh += (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * sigma / np.sqrt(2 * n)
    x = np.fft.ifft(xh) * n
    s = gpu.Solver(n, gpu.solve_options(k, sigma), 0)
    s.solve(s.upload(x), seed)
    st = s.post_analyze(list(truth), xh)
    st_o = orc.post_analyze(n, orc.ModeMap(truth), xh, orc.ModeMap(s.modes_dict()))
    kd = gpu.DeviceArray.from_numpy(np.sort(np.array(list(truth), np.int64)))
    xd = gpu.DeviceArray.from_numpy(xh)
    out = gpu.DeviceArray(80)
    rc = gpu.lib().mpsft_solver_post_analyze_noalloc(s.h, kd.ptr, k, xd.ptr, out.ptr, None)
    gpu.synchronize()
    na = out.to_numpy(np.float64, (10,))
    names = ("hit", "hit_l1", "hit_l2", "miss", "miss_l1", "miss_l2", "l1", "l2", "max")
    rel = max(abs(st[m] - st_o[m]) / abs(st_o[m]) for m in names if st_o[m] and st_o[m] == st_o[m])
    return dict(hit=st["hit"], hit_o=st_o["hit"], miss=st["miss"], miss_o=st_o["miss"], rel=rel,
                noalloc_rc=rc, noalloc_same=all(na[i] == st[m] or (na[i] != na[i] and st[m] != st[m])
                                                for i, m in enumerate(names)),
                noalloc_flag=float(na[9]))


def identify(n, k, sigma, seed, trials=1):
    """mpsft_identify_estimate on the oracle's bin matrices of one Iterate:
    decisions, k1 and margins against the oracle's trace."""
    x = signal(n, k, sigma, seed)
    bins = 2 * k + 1
    kw = dict(bins=bins, window_delta=1e-6, trials=trials, bin_threshold=3 * sigma / np.sqrt(bins),
              window_threshold=0.1, sv_threshold=sigma / np.sqrt(bins))
    orc.seed(seed)
    mm = orc.ModeMap()
    tf_o = orc.transform_draw(n)
    qs = np.array([orc.rand_int32() % n for _ in range(trials)], np.int32)
    _, tr = orc.iterate(x, orc.IterateOptions(**kw), mm, tf=tf_o, qs=qs, want_a=True)
    plan = gpu.Plan(n, bins, 1e-6, tr["bits"])
    tf = gpu.transform_make(n, int(tf_o[0]), int(tf_o[1]), int(tf_o[2]))
    got = plan.identify_estimate(tr["a"], tf, qs.astype(np.int64), kw["bin_threshold"],
                                 kw["sv_threshold"], kw["window_threshold"])
    found = got["decide"] == 3
    full = got["decide"] != 1
    ref = mm.to_dict()
    return dict(same_decide=bool(np.array_equal(got["decide"], tr["decide"])),
                same_k1=bool(np.array_equal(got["k1"], tr["k1"].astype(np.int64))),
                n_found=int(found.sum()),
                coef_err=max(abs(got["coef"][b] - ref[int(got["k0"][b])]) for b in np.flatnonzero(found)),
                margin_err=float(np.abs(got["margin"][full] - tr["margin"][full]).max()))


if __name__ == "__main__":
    case = json.loads(sys.argv[1])
    os.environ.update(case.pop("setenv", {}))  # read by the library at its first table build
    fn = {"solve": solve, "iterate": iterate, "plan": plan, "batch": batch,
          "adjoint": adjoint, "batch_adjoint": batch_adjoint, "gen": gen, "analyze": analyze,
          "identify": identify, "plan_freq": plan_freq}[case.pop("kind")]
    print("RESULT " + json.dumps(fn(**case)), flush=True)
