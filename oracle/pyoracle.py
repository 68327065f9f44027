"""ctypes wrapper over the CPU FP64 oracle (``libmpsft_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, never by the product package.
The functions restate the reference (jiaweichiu/mpsft); see
``mpsft_oracle.h`` for the file:line each one follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpsft_oracle.so")

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class IterateOptions(C.Structure):
    """orc_iterate_options; defaults of IterateOptions (src/iterate.h:26-59)."""

    _fields_ = [
        ("bins", C.c_int32),
        ("window_delta", C.c_double),
        ("trials", C.c_int32),
        ("bin_threshold", C.c_double),
        ("window_threshold", C.c_double),
        ("sv_threshold", C.c_double),
        ("bin_in_time_type", C.c_int32),
        ("bin_in_freq_type", C.c_int32),
    ]

    def __init__(self, **kw):
        super().__init__()
        self.bins = 0
        self.window_delta = 1e-6
        self.trials = 1
        self.bin_threshold = 0.0
        self.window_threshold = 0.0
        self.sv_threshold = 0.0
        self.bin_in_time_type = 1
        self.bin_in_freq_type = 1
        for k, v in kw.items():
            setattr(self, k, v)


class IterateTrace(C.Structure):
    _fields_ = [
        ("tf", C.c_int32 * 4),
        ("q_out", C.c_void_p),
        ("a_out", C.c_void_p),
        ("decide_out", C.c_void_p),
        ("k1_out", C.c_void_p),
        ("margin_out", C.c_void_p),
        ("bits", C.c_int32),
        ("found", C.c_int32),
    ]


class Demo1Options(C.Structure):
    """orc_demo1_options; defaults of Demo1Options (src/demo1.h:25-33)."""

    _fields_ = [
        ("trials", C.c_int32),
        ("min_bins", C.c_int32),
        ("window_delta", C.c_double),
        ("window_threshold", C.c_double),
        ("max_stale_iter", C.c_int32),
    ]

    def __init__(self, **kw):
        super().__init__()
        self.trials = 1
        self.min_bins = 101
        self.window_delta = 1e-6
        self.window_threshold = 0.1
        self.max_stale_iter = 3
        for k, v in kw.items():
            setattr(self, k, v)


class Demo1Log(C.Structure):
    """orc_demo1_log: per-iteration record of the Demo1::Run loop."""

    _fields_ = [
        ("bins", C.c_void_p),
        ("found", C.c_void_p),
        ("res", C.c_void_p),
        ("stale", C.c_void_p),
        ("bin_thr", C.c_void_p),
        ("sv_thr", C.c_void_p),
        ("cap", C.c_int32),
    ]


def _demo1_log(cap):
    """A Demo1Log over fresh numpy arrays of cap entries (kept alive in .arrays)."""
    arrays = dict(bins=np.zeros(cap, np.int32), found=np.zeros(cap, np.int32),
                  res=np.zeros(cap, np.uint8), stale=np.zeros(cap, np.int32),
                  bin_thr=np.zeros(cap, np.float64), sv_thr=np.zeros(cap, np.float64))
    lg = Demo1Log()
    for name, arr in arrays.items():
        setattr(lg, name, arr.ctypes.data)
    lg.cap = cap
    return lg, arrays


def build():
    """Compile the oracle with its Makefile (g++; no GPU involved)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = C.CDLL(LIB_PATH)
    L.orc_set_threads.argtypes = [C.c_int32]
    L.orc_get_threads.restype = C.c_int32
    L.orc_rng_seed.argtypes = [C.c_uint32]
    L.orc_rng_int32.restype = C.c_int32
    L.orc_rng_int64.restype = C.c_int64
    L.orc_rng_normal.restype = C.c_double
    L.orc_rng_double.restype = C.c_double
    for f in ("orc_mul_mod", "orc_mul_pos_mod"):
        getattr(L, f).argtypes = [C.c_int64] * 3
        getattr(L, f).restype = C.c_int32
    L.orc_mul_add_pos_mod.argtypes = [C.c_int64] * 4
    L.orc_mul_add_pos_mod.restype = C.c_int32
    L.orc_transform_draw.argtypes = [C.c_int32, _i32p]
    L.orc_transform_make.argtypes = [C.c_int32] * 4 + [_i32p]
    L.orc_num_bits.argtypes = [C.c_int32, C.c_int32]
    L.orc_num_bits.restype = C.c_int32
    L.orc_window_params.argtypes = [C.c_int32, C.c_int32, C.c_double,
                                    C.POINTER(C.c_int32)] + [C.POINTER(C.c_double)] * 3
    L.orc_window_taps.argtypes = [C.c_int32, C.c_int32, C.c_double, _f64p, _i32p]
    L.orc_window_sample_in_time.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_int32]
    L.orc_window_sample_in_time.restype = C.c_double
    L.orc_window_sample_in_freq.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double]
    L.orc_window_sample_in_freq.restype = C.c_double
    L.orc_fft.argtypes = [_f64p, _f64p, C.c_int32, C.c_int]
    L.orc_dft_naive.argtypes = [_f64p, _f64p, C.c_int32, C.c_int]
    L.orc_bin_in_time.argtypes = [C.c_int, _f64p, C.c_int32, C.c_int32, C.c_double,
                                  C.c_int32, _i32p, C.c_int32, _f64p]
    L.orc_bin_in_freq.argtypes = [C.c_int, _i32p, _f64p, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_double, C.c_int32, _i32p, C.c_int32, _f64p]
    L.orc_mat_pencil.argtypes = [_f64p, _f64p]
    L.orc_mat_pencil.restype = C.c_int
    L.orc_modemap_create.restype = C.c_void_p
    L.orc_modemap_destroy.argtypes = [C.c_void_p]
    L.orc_modemap_size.argtypes = [C.c_void_p]
    L.orc_modemap_size.restype = C.c_int32
    L.orc_modemap_get.argtypes = [C.c_void_p, _i32p, _f64p]
    L.orc_modemap_set.argtypes = [C.c_void_p, _i32p, _f64p, C.c_int32]
    L.orc_iterate.argtypes = [_f64p, C.c_int32, C.POINTER(IterateOptions), C.c_void_p,
                              C.POINTER(IterateTrace)]
    L.orc_iterate_fixed.argtypes = [_f64p, C.c_int32, C.POINTER(IterateOptions), _i32p,
                                    _i32p, C.c_void_p, C.POINTER(IterateTrace)]
    L.orc_identify.argtypes = [_f64p, C.c_int32, C.c_int32, C.POINTER(IterateOptions), _i32p,
                               _i32p, C.c_void_p, C.POINTER(IterateTrace)]
    L.orc_demo1_run.argtypes = [_f64p, C.c_int32, C.c_int32, C.c_double,
                                C.POINTER(Demo1Options), C.c_void_p, _i32p, _i32p, C.c_int32]
    L.orc_demo1_run.restype = C.c_int32
    L.orc_demo1_run_log.argtypes = [_f64p, C.c_int32, C.c_int32, C.c_double,
                                    C.POINTER(Demo1Options), C.c_void_p, C.POINTER(Demo1Log)]
    L.orc_demo1_run_log.restype = C.c_int32
    L.orc_demo1_schedule.argtypes = [C.c_int32, C.c_double, C.POINTER(Demo1Options),
                                     C.POINTER(C.c_uint8), _i32p, C.c_int32,
                                     C.POINTER(Demo1Log)]
    L.orc_demo1_schedule.restype = C.c_int32
    L.orc_generate_mode_map.argtypes = [C.c_int32, C.c_int32, C.c_void_p]
    L.orc_evaluate_modes.argtypes = [C.c_int32, C.c_void_p, _f64p]
    L.orc_generate_xhat.argtypes = [C.c_int32, C.c_void_p, C.c_double, _f64p]
    L.orc_generate_xhat_alt.argtypes = [C.c_int32, C.c_void_p, C.c_double, _f64p]
    L.orc_post_analyze.argtypes = [C.c_int32, C.c_void_p, _f64p, C.c_void_p, _f64p]
    _lib = L
    return L


def _c128(a):
    """complex array -> contiguous interleaved float64 view."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


# ---- RNG (src/base.cc:32-56) ----
def seed(s):
    lib().orc_rng_seed(s)


def rand_int32():
    return lib().orc_rng_int32()


def rand_normal():
    return lib().orc_rng_normal()


def set_threads(t):
    lib().orc_set_threads(t)


def get_threads():
    return lib().orc_get_threads()


# ---- integer (src/integer.{h,cc}) ----
def mul_mod(a, b, d):
    return lib().orc_mul_mod(a, b, d)


def mul_pos_mod(a, b, d):
    return lib().orc_mul_pos_mod(a, b, d)


def transform_draw(n):
    tf = np.zeros(4, np.int32)
    _check(lib().orc_transform_draw(n, tf), "transform_draw")
    return tf


def transform_make(n, a, b, c):
    tf = np.zeros(4, np.int32)
    _check(lib().orc_transform_make(n, a, b, c, tf), "transform_make")
    return tf


def num_bits(n, bins):
    return lib().orc_num_bits(n, bins)


# ---- window (src/window.cc) ----
def window_params(n, bins, delta):
    p = C.c_int32()
    w, sf, st = C.c_double(), C.c_double(), C.c_double()
    _check(lib().orc_window_params(n, bins, delta, C.byref(p), C.byref(w), C.byref(sf),
                                   C.byref(st)), "window_params")
    return dict(p=p.value, width=w.value, sigma_f=sf.value, sigma_t=st.value)


def window_taps(n, bins, delta):
    p = window_params(n, bins, delta)["p"]
    wt = np.zeros(p, np.float64)
    t = np.zeros(p, np.int32)
    _check(lib().orc_window_taps(n, bins, delta, wt, t), "window_taps")
    return wt, t


def sample_in_time(n, bins, delta, i):
    return lib().orc_window_sample_in_time(n, bins, delta, i)


def sample_in_freq(n, bins, delta, xi):
    return lib().orc_window_sample_in_freq(n, bins, delta, xi)


# ---- FFT ----
def fft(x, sign=-1):
    x = _c128(x)
    n = x.size // 2
    out = np.zeros(2 * n, np.float64)
    _check(lib().orc_fft(x, out, n, sign), "fft")
    return out.view(np.complex128)


def dft_naive(x, sign=-1):
    x = _c128(x)
    n = x.size // 2
    out = np.zeros(2 * n, np.float64)
    _check(lib().orc_dft_naive(x, out, n, sign), "dft_naive")
    return out.view(np.complex128)


# ---- binners ----
def bin_in_time(x, bins, delta, bits, tf, q, version=0):
    x = _c128(x)
    n = x.size // 2
    rows = 1 + 2 * bits
    out = np.zeros(2 * rows * bins, np.float64)
    _check(lib().orc_bin_in_time(version, x, n, bins, delta, bits,
                                 np.ascontiguousarray(tf, np.int32), q, out), "bin_in_time")
    return out.view(np.complex128).reshape(rows, bins)


def bin_in_freq(keys, coefs, n, bins, delta, bits, tf, q, out, version=1):
    """Subtracts in place from ``out`` (rows x bins complex128)."""
    keys = np.ascontiguousarray(keys, np.int32)
    coefs = _c128(coefs)
    buf = _c128(out).copy()
    _check(lib().orc_bin_in_freq(version, keys, coefs, keys.size, n, bins, delta, bits,
                                 np.ascontiguousarray(tf, np.int32), q, buf), "bin_in_freq")
    out[...] = buf.view(np.complex128).reshape(out.shape)
    return out


def mat_pencil(u0, u1, u2):
    u = np.array([u0.real, u0.imag, u1.real, u1.imag, u2.real, u2.imag], np.float64)
    s = np.zeros(2, np.float64)
    vote = lib().orc_mat_pencil(u, s)
    return bool(vote), s


# ---- mode maps ----
class ModeMap:
    """Owning handle over the oracle's std::unordered_map<int32, complex>."""

    def __init__(self, d=None):
        self.h = C.c_void_p(lib().orc_modemap_create())
        if d:
            self.set(d)

    def __del__(self):
        try:
            if self.h:
                lib().orc_modemap_destroy(self.h)
        except Exception:
            pass

    def __len__(self):
        return lib().orc_modemap_size(self.h)

    def set(self, d):
        keys = np.array(list(d.keys()), np.int32)
        vals = np.array([complex(v) for v in d.values()], np.complex128)
        lib().orc_modemap_set(self.h, keys, _c128(vals), keys.size)

    def items(self):
        m = len(self)
        keys = np.zeros(m, np.int32)
        coefs = np.zeros(2 * m, np.float64)
        if m:
            lib().orc_modemap_get(self.h, keys, coefs)
        return keys, coefs.view(np.complex128)

    def to_dict(self):
        k, c = self.items()
        return {int(a): complex(b) for a, b in zip(k, c)}


def iterate(x, opt, mm, tf=None, qs=None, want_a=False):
    """One Iterate (src/iterate.cc:80-177).  Returns (found, trace dict)."""
    x = _c128(x)
    n = x.size // 2
    bits = num_bits(n, opt.bins)
    rows = 1 + 2 * bits
    q_out = np.zeros(opt.trials, np.int32)
    a_out = np.zeros(2 * opt.trials * rows * opt.bins, np.float64) if want_a else None
    decide = np.zeros(opt.bins, np.int32)
    k1 = np.zeros(opt.bins, np.int32)
    margin = np.zeros(opt.bins, np.float64)
    tr = IterateTrace()
    tr.q_out = q_out.ctypes.data
    tr.a_out = a_out.ctypes.data if want_a else None
    tr.decide_out = decide.ctypes.data
    tr.k1_out = k1.ctypes.data
    tr.margin_out = margin.ctypes.data
    if tf is None:
        rc = lib().orc_iterate(x, n, C.byref(opt), mm.h, C.byref(tr))
    else:
        rc = lib().orc_iterate_fixed(x, n, C.byref(opt), np.ascontiguousarray(tf, np.int32),
                                     np.ascontiguousarray(qs, np.int32), mm.h, C.byref(tr))
    if rc < 0:
        raise RuntimeError(f"iterate failed with status {-rc}")
    trace = dict(tf=np.array(tr.tf[:], np.int32), q=q_out, bits=tr.bits, found=tr.found,
                 decide=decide, k1=k1, margin=margin)
    if want_a:
        trace["a"] = a_out.view(np.complex128).reshape(opt.trials, rows, opt.bins)
    return bool(rc), trace


def identify(a, n, opt, tf, qs, mm):
    """Iterate's per-bin body on given bin matrices a (trials, rows, bins)
    complex128 (src/iterate.cc:113-175).  Returns (found, trace dict)."""
    a = np.ascontiguousarray(a, np.complex128)
    decide = np.zeros(opt.bins, np.int32)
    k1 = np.zeros(opt.bins, np.int32)
    margin = np.zeros(opt.bins, np.float64)
    tr = IterateTrace()
    tr.decide_out = decide.ctypes.data
    tr.k1_out = k1.ctypes.data
    tr.margin_out = margin.ctypes.data
    rc = lib().orc_identify(a.view(np.float64).reshape(-1), a.shape[0], n, C.byref(opt),
                            np.ascontiguousarray(tf, np.int32), np.ascontiguousarray(qs, np.int32),
                            mm.h, C.byref(tr))
    if rc < 0:
        raise RuntimeError(f"identify failed with status {-rc}")
    return bool(rc), dict(decide=decide, k1=k1, margin=margin, found=tr.found, bits=tr.bits)


def demo1_run(x, k, sigma, opt=None, log_cap=4096):
    """Demo1::Run (src/demo1.cc:36-59).  Returns (ModeMap, bins_log, found_log)."""
    opt = opt or Demo1Options()
    x = _c128(x)
    n = x.size // 2
    mm = ModeMap()
    bl = np.zeros(log_cap, np.int32)
    fl = np.zeros(log_cap, np.int32)
    it = lib().orc_demo1_run(x, n, k, sigma, C.byref(opt), mm.h, bl, fl, log_cap)
    if it < 0:
        raise RuntimeError(f"demo1_run failed with status {-it}")
    return mm, bl[:it].copy(), fl[:it].copy()


def demo1_run_log(x, k, sigma, opt=None, log_cap=4096):
    """Demo1::Run with its full per-iteration record.  Returns (ModeMap, log)
    where log maps bins / found / res / stale / bin_thr / sv_thr to arrays
    of one entry per iteration."""
    opt = opt or Demo1Options()
    x = _c128(x)
    n = x.size // 2
    mm = ModeMap()
    lg, arrays = _demo1_log(log_cap)
    it = lib().orc_demo1_run_log(x, n, k, sigma, C.byref(opt), mm.h, C.byref(lg))
    if it < 0:
        raise RuntimeError(f"demo1_run failed with status {-it}")
    return mm, {name: arr[:it].copy() for name, arr in arrays.items()}


def demo1_schedule(k, sigma, script_res, script_size, opt=None):
    """The Demo1::Run loop with Iterate replaced by a script (iteration i
    returns script_res[i] and leaves script_size[i] modes).  Returns the
    per-iteration record as demo1_run_log does."""
    opt = opt or Demo1Options()
    res = np.ascontiguousarray(script_res, np.uint8)
    size = np.ascontiguousarray(script_size, np.int32)
    lg, arrays = _demo1_log(max(1, res.size))
    it = lib().orc_demo1_schedule(k, sigma, C.byref(opt),
                                  res.ctypes.data_as(C.POINTER(C.c_uint8)), size,
                                  res.size, C.byref(lg))
    if it < 0:
        raise RuntimeError(f"demo1_schedule failed with status {-it}")
    return {name: arr[:it].copy() for name, arr in arrays.items()}


# ---- gen (src/gen.cc) ----
def generate_mode_map(n, k):
    mm = ModeMap()
    _check(lib().orc_generate_mode_map(n, k, mm.h), "generate_mode_map")
    return mm


def evaluate_modes(n, mm):
    out = np.zeros(2 * n, np.float64)
    _check(lib().orc_evaluate_modes(n, mm.h, out), "evaluate_modes")
    return out.view(np.complex128)


def generate_xhat(n, mm, sigma):
    out = np.zeros(2 * n, np.float64)
    _check(lib().orc_generate_xhat(n, mm.h, sigma, out), "generate_xhat")
    return out.view(np.complex128)


def generate_xhat_alt(n, mm, sigma):
    out = np.zeros(2 * n, np.float64)
    _check(lib().orc_generate_xhat_alt(n, mm.h, sigma, out), "generate_xhat_alt")
    return out.view(np.complex128)


def post_analyze(n, truth, xh, found):
    st = np.zeros(9, np.float64)
    _check(lib().orc_post_analyze(n, truth.h, _c128(xh), found.h, st), "post_analyze")
    keys = ["hit", "hit_l1", "hit_l2", "miss", "miss_l1", "miss_l2", "l1", "l2", "max"]
    return dict(zip(keys, st.tolist()))
