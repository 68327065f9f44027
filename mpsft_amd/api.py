"""Python mirror of the reference's binner / window / freqid / iterate /
demo1 interfaces, over the HIP library's C-ABI.

Reference interfaces (jiaweichiu/mpsft, paths under src/):
  Window(n, bins, delta)                      window.h:25-55
  Transform(n, a, b, c)                       integer.h:75-82
  BinInTime::Create(type, win, bits)->Run     binner.h:30-37, binner.cc:58-76
  BinInFreq::Create(type, win, bits)->Run     binner.h:123-131, binner.cc:526-535
  MatPencil(u0, u1, u2, sigma)                freqid.h:28
  IterateOptions / Iterate(x, opt, mm)        iterate.h:26-65
  Demo1Options / Demo1::Run                   demo1.h:25-53, demo1.cc:36-59
The HIP forms are new binner types: 5 = FP64, 6 = MIXED (complex64 input).
Types 0-4 / 0-1 are the reference's CPU binners and are rejected here, as
an unknown type is by the reference (binner.cc:74,533), but with an error
instead of an abort.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (MPSFT_FP64, MPSFT_MIXED, IterateOptions, MpsftError, SolveOptions,
                   Transform, check, lib)

BIN_TYPE_FP64 = 5
BIN_TYPE_MIXED = 6
_TYPE_PRECISION = {BIN_TYPE_FP64: MPSFT_FP64, BIN_TYPE_MIXED: MPSFT_MIXED}


def device_count():
    n = C.c_int(0)
    st = lib().mpsft_device_count(C.byref(n))
    return n.value if st == 0 else 0


class DeviceArray:
    """A device allocation made through the C-ABI (hipMalloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().mpsft_malloc(C.byref(self.ptr), max(1, self.nbytes)), "mpsft_malloc")

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        if a.nbytes:
            check(lib().mpsft_memcpy_htod(d.ptr, a.ctypes.data, a.nbytes), "mpsft_memcpy_htod")
        d.dtype, d.shape = a.dtype, a.shape
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes:
            check(lib().mpsft_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes),
                  "mpsft_memcpy_dtoh")
        return out

    def zero(self):
        check(lib().mpsft_memset(self.ptr, 0, self.nbytes), "mpsft_memset")

    def free(self):
        if self.ptr:
            lib().mpsft_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib().mpsft_device_synchronize(), "mpsft_device_synchronize")


def transform_make(n, a, b, c):
    """Transform(n, a, b, c) (src/integer.cc:51-56)."""
    tf = Transform()
    check(lib().mpsft_transform_make(n, a, b, c, C.byref(tf)), "mpsft_transform_make")
    return tf


def num_bits(n, bins):
    """NumBits (src/iterate.cc:30-42)."""
    return lib().mpsft_num_bits(n, bins)


class Window:
    """Window(n, bins, delta) (src/window.cc:28-81): host-side parameters.
    The taps themselves are computed on the device inside BinInTime."""

    def __init__(self, n, bins, delta):
        if bins <= 0 or bins % 2 != 1:
            raise MpsftError("Window", 1)
        self.n, self.bins_, self.delta = int(n), int(bins), float(delta)
        self.p_ = lib().mpsft_window_p(n, bins, delta)
        if self.p_ <= 0:
            raise MpsftError("Window", 1)

    def bins(self):
        return self.bins_

    def p(self):
        return self.p_


def _ptr(d):
    """Device address of a DeviceArray, a torch tensor or a plain int."""
    if isinstance(d, DeviceArray):
        return d.ptr
    if hasattr(d, "data_ptr"):
        return C.c_void_p(d.data_ptr())
    return C.c_void_p(int(d))


def _as_device(x, precision):
    """numpy complex array -> (DeviceArray, keepalive)."""
    if isinstance(x, DeviceArray):
        return x
    dt = np.complex128 if precision == MPSFT_FP64 else np.complex64
    return DeviceArray.from_numpy(np.ascontiguousarray(x, dtype=dt))


def window_delta(n, bins, wt):
    """The delta of Window(n, bins, delta) recovered from its taps wt
    (Window::wt(), src/window.h:36)."""
    wt = np.ascontiguousarray(wt, dtype=np.float64)
    d = C.c_double()
    check(lib().mpsft_window_delta(n, bins, wt.size, wt.ctypes.data_as(C.POINTER(C.c_double)),
                                   C.byref(d)), "mpsft_window_delta")
    return d.value


class Plan:
    """mpsft_plan: one window + binner pair on the device."""

    def __init__(self, n, bins, delta, bits, precision=MPSFT_FP64, taps=None):
        """delta: the window's delta; or taps: the p taps Window::wt() of a
        window whose delta the caller does not hold (delta is then None)."""
        self.n, self.bins, self.bits = int(n), int(bins), int(bits)
        self.rows = 1 + 2 * self.bits
        self.precision = precision
        self.h = C.c_void_p()
        if taps is None:
            check(lib().mpsft_plan_create(n, bins, delta, bits, precision, C.byref(self.h)),
                  "mpsft_plan_create")
        else:
            self._taps = np.ascontiguousarray(taps, dtype=np.float64)
            check(lib().mpsft_plan_create_window(
                n, bins, self._taps.size, self._taps.ctypes.data_as(C.POINTER(C.c_double)), bits,
                precision, C.byref(self.h)), "mpsft_plan_create_window")
        d = C.c_double()
        check(lib().mpsft_plan_delta(self.h, C.byref(d)), "mpsft_plan_delta")
        self.delta = d.value

    @classmethod
    def from_window(cls, n, bins, taps, bits, precision=MPSFT_FP64):
        """A plan from a Window's taps alone (mpsft_plan_create_window)."""
        return cls(n, bins, None, bits, precision, taps=taps)

    def __del__(self):
        try:
            if self.h:
                lib().mpsft_plan_destroy(self.h)
        except Exception:
            pass

    def bin_in_time_dev(self, x_dev, tf, q, out_dev, stream=None):
        check(lib().mpsft_bin_in_time(self.h, x_dev.ptr, self.n, C.byref(tf), q, out_dev.ptr,
                                      stream), "mpsft_bin_in_time")

    def bin_in_freq_dev(self, k0_dev, coef_dev, m, tf, q, out_dev, stream=None):
        check(lib().mpsft_bin_in_freq(self.h, k0_dev.ptr if m else None,
                                      coef_dev.ptr if m else None, m, C.byref(tf), q,
                                      out_dev.ptr, stream), "mpsft_bin_in_freq")

    def bin_in_time(self, x, tf, q):
        """BinInTime::Run on host data: returns (rows, bins) complex128."""
        xd = _as_device(x, self.precision)
        out = DeviceArray(16 * self.rows * self.bins)
        self.bin_in_time_dev(xd, tf, q, out)
        synchronize()
        return out.to_numpy(np.complex128, (self.rows, self.bins))

    def bin_in_freq(self, keys, coefs, tf, q, out):
        """BinInFreq::Run on host data: subtracts from out (rows, bins)."""
        keys = np.ascontiguousarray(keys, dtype=np.int64)
        coefs = np.ascontiguousarray(coefs, dtype=np.complex128)
        od = DeviceArray.from_numpy(np.ascontiguousarray(out, dtype=np.complex128))
        kd = DeviceArray.from_numpy(keys) if keys.size else None
        cd = DeviceArray.from_numpy(coefs) if keys.size else None
        self.bin_in_freq_dev(kd, cd, keys.size, tf, q, od)
        synchronize()
        out[...] = od.to_numpy(np.complex128, out.shape)
        return out


    def identify_estimate(self, A, tf, qs, bin_threshold, sv_threshold, window_threshold):
        """Iterate's per-bin body (src/iterate.cc:113-175) over bin matrices
        A (trials, rows, bins) complex128.  Returns a dict of per-bin arrays:
        decide (0 energy, 1 sv, 2 window, 3 found), k1, k0, coef, margin."""
        A = np.ascontiguousarray(A, dtype=np.complex128)
        trials = A.shape[0]
        if A.shape != (trials, self.rows, self.bins):
            raise MpsftError("identify_estimate: A shape", 2)
        q = np.ascontiguousarray(qs, dtype=np.int64)
        if q.size != trials:
            raise MpsftError("identify_estimate: one q per trial", 2)
        ad = DeviceArray.from_numpy(A)
        outs = dict(decide=DeviceArray(4 * self.bins), k1=DeviceArray(8 * self.bins),
                    k0=DeviceArray(8 * self.bins), coef=DeviceArray(16 * self.bins),
                    margin=DeviceArray(8 * self.bins))
        check(lib().mpsft_identify_estimate(
            self.h, ad.ptr, trials, C.byref(tf), q.ctypes.data_as(C.POINTER(C.c_int64)),
            bin_threshold, sv_threshold, window_threshold, outs["decide"].ptr, outs["k1"].ptr,
            outs["k0"].ptr, outs["coef"].ptr, outs["margin"].ptr, None),
            "mpsft_identify_estimate")
        synchronize()
        types = dict(decide=np.int32, k1=np.int64, k0=np.int64, coef=np.complex128,
                     margin=np.float64)
        return {k: d.to_numpy(types[k], (self.bins,)) for k, d in outs.items()}


class BinInTime:
    """BinInTime::Create(type, win, bits) (src/binner.cc:58-76)."""

    @staticmethod
    def create(binner_type, win, bits):
        if binner_type not in _TYPE_PRECISION:
            raise MpsftError(f"BinInTime::Create(type={binner_type})", 3)
        return _BinInTimeHip(Plan(win.n, win.bins_, win.delta, bits,
                                  _TYPE_PRECISION[binner_type]))


class _BinInTimeHip:
    def __init__(self, plan):
        self.plan = plan

    def run(self, x, tf, q, out=None):
        """Overwrites and returns the (1+2*bits, bins) bin matrix."""
        r = self.plan.bin_in_time(x, tf, q)
        if out is not None:
            if out.shape != r.shape:
                raise MpsftError("BinInTime::Run rows", 2)
            out[...] = r
            return out
        return r


class BinInFreq:
    """BinInFreq::Create(type, win, bits) (src/binner.cc:526-535)."""

    @staticmethod
    def create(binner_type, win, bits):
        if binner_type not in _TYPE_PRECISION:
            raise MpsftError(f"BinInFreq::Create(type={binner_type})", 3)
        return _BinInFreqHip(Plan(win.n, win.bins_, win.delta, bits, MPSFT_FP64))


class _BinInFreqHip:
    def __init__(self, plan):
        self.plan = plan

    def run(self, mm, tf, q, out):
        """Subtracts the modes of mm ({k0: coef}) from out in place."""
        if out.shape != (self.plan.rows, self.plan.bins):
            raise MpsftError("BinInFreq::Run rows", 2)
        keys = np.array(list(mm.keys()), dtype=np.int64)
        coefs = np.array([complex(v) for v in mm.values()], dtype=np.complex128)
        return self.plan.bin_in_freq(keys, coefs, tf, q, out)


def mat_pencil(u0, u1, u2):
    """MatPencil (src/freqid.cc:27-48), batched on the device.  Returns
    (votes bool[m], sigma float64[m, 2])."""
    u0, u1, u2 = (np.atleast_1d(np.asarray(v, dtype=np.complex128)) for v in (u0, u1, u2))
    m = u0.size
    u = np.stack([u0, u1, u2], axis=1).reshape(-1)
    ud = DeviceArray.from_numpy(u)
    sd = DeviceArray(16 * m)
    vd = DeviceArray(m)
    check(lib().mpsft_mat_pencil(ud.ptr, m, sd.ptr, vd.ptr, None), "mpsft_mat_pencil")
    synchronize()
    return vd.to_numpy(np.uint8, (m,)).astype(bool), sd.to_numpy(np.float64, (m, 2))


def iterate_options(bins, window_delta=1e-6, trials=1, bin_threshold=0.0, window_threshold=0.0,
                    sv_threshold=0.0, precision=MPSFT_FP64):
    """IterateOptions (src/iterate.h:26-59) with the HIP binner types."""
    t = BIN_TYPE_FP64 if precision == MPSFT_FP64 else BIN_TYPE_MIXED
    return IterateOptions(bins, window_delta, trials, bin_threshold, window_threshold,
                          sv_threshold, t, t)


def solve_options(k, sigma, trials=1, min_bins=101, window_delta=1e-6, window_threshold=0.1,
                  max_stale_iter=3, max_iter=0):
    """Demo1Options (src/demo1.h:25-33) plus Demo1's k and sigma."""
    return SolveOptions(trials, min_bins, window_delta, window_threshold, max_stale_iter, k,
                        sigma, max_iter)


class Solver:
    """Device state for Iterate / Demo1::Run on signals of length n."""

    def __init__(self, n, opt, precision=MPSFT_FP64, capacity=0):
        self.n, self.opt, self.precision = int(n), opt, precision
        self.h = C.c_void_p()
        check(lib().mpsft_solver_create(n, C.byref(opt), precision, capacity, C.byref(self.h)),
              "mpsft_solver_create")

    def __del__(self):
        try:
            if self.h:
                lib().mpsft_solver_destroy(self.h)
        except Exception:
            pass

    def upload(self, x):
        return _as_device(x, self.precision)

    def set_tile_residues(self, residues):
        """Gather-fold tile width (1, 2, 4, 8; 0 = default)."""
        check(lib().mpsft_solver_set_tile_residues(self.h, residues), "set_tile_residues")

    def solve(self, x_dev, seed, stream=None):
        """Demo1::Run; returns the iteration count."""
        it = C.c_int32(0)
        check(lib().mpsft_solve(self.h, _ptr(x_dev), seed, C.byref(it), stream), "mpsft_solve")
        return it.value

    def solve_profiled(self, x_dev, seed, stream=None):
        """The same solve with per-kernel HIP event timing; returns
        (iterations, [ms gather-fold, ms bin FFT, ms identify])."""
        it = C.c_int32(0)
        ms = (C.c_double * 3)()
        check(lib().mpsft_solve_profiled(self.h, _ptr(x_dev), seed, C.byref(it), ms, stream),
              "mpsft_solve_profiled")
        return it.value, list(ms)

    def iterate(self, x_dev, iopt, tf, qs, stream=None):
        """Iterate with explicit transform and q's; returns Iterate's bool."""
        q = (C.c_int64 * len(qs))(*[int(v) for v in qs])
        found = C.c_int32(0)
        check(lib().mpsft_iterate(self.h, x_dev.ptr, C.byref(iopt), C.byref(tf), q,
                                  C.byref(found), stream), "mpsft_iterate")
        return bool(found.value)

    def set_record(self, on=True):
        """While on, solves keep their decisions for replay / adjoint."""
        check(lib().mpsft_solver_set_record(self.h, 1 if on else 0), "set_record")

    def replay(self, x_dev, precision=None, stream=None):
        """The last recorded solve's map (decisions fixed, linear in x) on
        x_dev; the found modes become its output."""
        prec = self.precision if precision is None else precision
        check(lib().mpsft_solver_replay(self.h, _ptr(x_dev), prec, stream), "mpsft_solver_replay")

    def adjoint_dev(self, keys_dev, grad_dev, m, out_dev, precision=None, stream=None):
        """out_dev = F^H g for the last recorded solve's map F; keys_dev
        (int64) and grad_dev (complex128) hold m device entries."""
        prec = self.precision if precision is None else precision
        check(lib().mpsft_solver_adjoint(self.h, _ptr(keys_dev), _ptr(grad_dev), int(m),
                                         _ptr(out_dev), prec, stream), "mpsft_solver_adjoint")

    def adjoint(self, keys, grad, precision=MPSFT_FP64):
        """Host form of adjoint_dev: numpy keys / complex gradients in,
        numpy complex array of n out."""
        keys = np.ascontiguousarray(keys, np.int64)
        grad = np.ascontiguousarray(grad, np.complex128)
        kd, gd = DeviceArray.from_numpy(keys), DeviceArray.from_numpy(grad)
        esz = 16 if precision == MPSFT_FP64 else 8
        out = DeviceArray(esz * self.n)
        self.adjoint_dev(kd, gd, keys.size, out, precision)
        return out.to_numpy(np.complex128 if precision == MPSFT_FP64 else np.complex64, (self.n,))

    def modes(self):
        """Found-mode map as (keys int64[m], coefs complex128[m])."""
        cnt = C.c_int64(0)
        check(lib().mpsft_solver_modes_count(self.h, C.byref(cnt)), "modes_count")
        m = cnt.value
        keys = np.zeros(max(m, 1), np.int64)
        coefs = np.zeros(2 * max(m, 1), np.float64)
        check(lib().mpsft_solver_modes_get(
            self.h, keys.ctypes.data_as(C.POINTER(C.c_int64)),
            coefs.ctypes.data_as(C.POINTER(C.c_double)), m, C.byref(cnt)), "modes_get")
        return keys[:m], coefs.view(np.complex128)[:m]

    def modes_dict(self):
        k, c = self.modes()
        return {int(a): complex(b) for a, b in zip(k, c)}

    def set_modes(self, d):
        keys = np.array(list(d.keys()), np.int64)
        coefs = np.array([complex(v) for v in d.values()], np.complex128).view(np.float64)
        check(lib().mpsft_solver_modes_set(
            self.h, keys.ctypes.data_as(C.POINTER(C.c_int64)),
            coefs.ctypes.data_as(C.POINTER(C.c_double)), keys.size), "modes_set")

    def post_analyze(self, truth_keys, xhat):
        """PostAnalyze (src/demo1.cc:61-112) of the last solve on the device.
        truth_keys: the true locations; xhat: the signal's x-hat (n complex,
        a host array or a device array / tensor of complex128).  Returns the
        figures by name."""
        keys = np.ascontiguousarray(np.sort(np.asarray(truth_keys, np.int64)))
        kd = DeviceArray.from_numpy(keys) if keys.size else None
        xd = xhat if not isinstance(xhat, np.ndarray) else \
            DeviceArray.from_numpy(np.ascontiguousarray(xhat, np.complex128))
        st = (C.c_double * 9)()
        check(lib().mpsft_solver_post_analyze(self.h, _ptr(kd) if kd else None, keys.size,
                                              _ptr(xd), st, None), "mpsft_solver_post_analyze")
        names = ("hit", "hit_l1", "hit_l2", "miss", "miss_l1", "miss_l2", "l1", "l2", "max")
        return dict(zip(names, list(st)))

    def bins_matrix(self, trials, rows, bins):
        out = np.zeros(2 * trials * rows * bins, np.float64)
        t, r, b = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().mpsft_solver_bins_get(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                          out.size, C.byref(t), C.byref(r), C.byref(b)),
              "bins_get")
        return out.view(np.complex128).reshape(t.value, r.value, b.value)

    def log(self):
        cap = 1 << 16
        bins = np.zeros(cap, np.int32)
        found = np.zeros(cap, np.int32)
        n_it = C.c_int32()
        check(lib().mpsft_solver_log_get(self.h, bins.ctypes.data_as(C.POINTER(C.c_int32)),
                                         found.ctypes.data_as(C.POINTER(C.c_int32)), cap,
                                         C.byref(n_it)), "log_get")
        return bins[:n_it.value].copy(), found[:n_it.value].copy()

    def transforms(self):
        """(a, b, c, a_inv) rows and q's of the last solve's iterations."""
        cap = 1 << 16
        tfs = (Transform * cap)()
        qs = np.zeros(cap * self.opt.trials, np.int64)
        cnt = C.c_int32()
        check(lib().mpsft_solver_transforms_get(self.h, tfs, qs.ctypes.data_as(C.POINTER(C.c_int64)),
                                                cap, C.byref(cnt)), "transforms_get")
        m = cnt.value
        tf = np.array([(t.a, t.b, t.c, t.a_inv) for t in tfs[:m]], np.int64).reshape(m, 4)
        return tf, qs[:m * self.opt.trials].reshape(m, self.opt.trials)


def _u32_array(values, count):
    """`count` integers as a uint32 array, each taken modulo 2^32 (ctypes
    stores an int's low 32 bits); other numbers go through int()."""
    out = (C.c_uint32 * count)()
    try:
        out[:] = values
    except TypeError:
        out[:] = [int(v) & 0xFFFFFFFF for v in values]
    return out


class Batch:
    """The batched form of Demo1::Run (no reference counterpart): many
    independent signals, each with its own n, options and seed, recovered
    together with one launch per stage over all of them."""

    def __init__(self, ns, opts, precision=MPSFT_FP64):
        self.ns = [int(n) for n in ns]
        self.opts = list(opts)
        self.precision = precision
        self.count = len(self.ns)
        narr = (C.c_int64 * self.count)(*self.ns)
        oarr = (SolveOptions * self.count)(*self.opts)
        self.h = C.c_void_p()
        check(lib().mpsft_batch_create(self.count, narr, oarr, precision, C.byref(self.h)),
              "mpsft_batch_create")

    def __del__(self):
        try:
            if self.h:
                lib().mpsft_batch_destroy(self.h)
        except Exception:
            pass

    def upload(self, xs):
        return [_as_device(x, self.precision) for x in xs]

    def _ptrs(self, xds, seeds):
        # A list of DeviceArrays (whose addresses never change) keeps its
        # pointer array between calls; it is held, so the ids stay its own.
        # Other inputs (torch tensors, raw addresses) are read every call.
        if len(xds) != self.count:
            raise ValueError(f"expected {self.count} signals, got {len(xds)}")
        key = tuple(map(id, xds))
        cached = getattr(self, "_ptr_cache", None)
        if cached is not None and cached[0] == key:
            ptrs = cached[1]
        else:
            ptrs = (C.c_void_p * self.count)(*[_ptr(d).value for d in xds])
            if all(isinstance(d, DeviceArray) for d in xds):
                self._ptr_cache = (key, ptrs, list(xds))
        return ptrs, _u32_array(seeds, self.count)

    def solve(self, xds, seeds, stream=None):
        """Returns (iterations, statuses) per signal; raises on a failure
        other than a per-signal status."""
        ptrs, sd = self._ptrs(xds, seeds)
        it = (C.c_int32 * self.count)()
        stv = (C.c_int32 * self.count)()
        rc = lib().mpsft_batch_solve(self.h, ptrs, sd, it, stv, stream)
        if rc not in (0, 6, 7):  # per-signal capacity / iteration-cap statuses are reported
            check(rc, "mpsft_batch_solve")
        return it[:], stv[:]

    def solve_profiled(self, xds, seeds, stream=None):
        """The batched solve one iteration at a time with event-timed
        stages: returns (iterations, statuses, kernel_ms, launches) with
        kernel_ms the summed ms of the plan, gather-fold, bin-FFT and
        identify kernels."""
        ptrs, sd = self._ptrs(xds, seeds)
        it = (C.c_int32 * self.count)()
        stv = (C.c_int32 * self.count)()
        ms = (C.c_double * 4)()
        nl = C.c_int32(0)
        rc = lib().mpsft_batch_solve_profiled(self.h, ptrs, sd, it, stv, ms, C.byref(nl), stream)
        if rc not in (0, 6, 7):
            check(rc, "mpsft_batch_solve_profiled")
        return list(it), list(stv), list(ms), nl.value

    def autotune(self, xds, seeds, stream=None):
        ptrs, sd = self._ptrs(xds, seeds)
        r = C.c_int32(0)
        check(lib().mpsft_batch_autotune(self.h, ptrs, sd, C.byref(r), stream),
              "mpsft_batch_autotune")
        return r.value

    def set_record(self, on=True):
        """While on, batched solves keep every signal's decisions."""
        check(lib().mpsft_batch_set_record(self.h, 1 if on else 0), "batch_set_record")

    def replay(self, i, x_dev, precision=None, stream=None):
        """Signal i's recorded map on x_dev (see Solver.replay)."""
        prec = self.precision if precision is None else precision
        check(lib().mpsft_batch_replay(self.h, i, _ptr(x_dev), prec, stream), "batch_replay")

    def adjoint_dev(self, keys_devs, grad_devs, ms, out_devs, precision=None, stream=None):
        """Every signal's adjoint in one batched pass (device buffers)."""
        prec = self.precision if precision is None else precision
        K = self.count
        kp = (C.c_void_p * K)(*[_ptr(d).value for d in keys_devs])
        gp = (C.c_void_p * K)(*[_ptr(d).value for d in grad_devs])
        op = (C.c_void_p * K)(*[_ptr(d).value for d in out_devs])
        mm = (C.c_int64 * K)(*[int(v) for v in ms])
        check(lib().mpsft_batch_adjoint(self.h, kp, gp, mm, op, prec, stream), "batch_adjoint")

    def adjoint(self, keys_list, grad_list, precision=MPSFT_FP64):
        """Host form: lists of numpy keys / complex gradients in, list of
        numpy complex arrays (n[i] each) out."""
        kd = [DeviceArray.from_numpy(np.ascontiguousarray(k, np.int64)) for k in keys_list]
        gd = [DeviceArray.from_numpy(np.ascontiguousarray(g, np.complex128)) for g in grad_list]
        esz = 16 if precision == MPSFT_FP64 else 8
        od = [DeviceArray(esz * n) for n in self.ns]
        self.adjoint_dev(kd, gd, [len(k) for k in keys_list], od, precision)
        dt = np.complex128 if precision == MPSFT_FP64 else np.complex64
        return [o.to_numpy(dt, (n,)) for o, n in zip(od, self.ns)]

    def modes_all(self):
        """Every signal's (keys, coefs) of the last batched solve, copied on
        the device in one launch (mpsft_batch_modes_counts / _device)."""
        K = self.count
        counts = (C.c_int64 * K)()
        check(lib().mpsft_batch_modes_counts(self.h, counts), "batch_modes_counts")
        kd = [DeviceArray(8 * max(int(m), 1)) for m in counts]
        cd = [DeviceArray(16 * max(int(m), 1)) for m in counts]
        kp = (C.c_void_p * K)(*[d.ptr.value for d in kd])
        cp = (C.c_void_p * K)(*[d.ptr.value for d in cd])
        check(lib().mpsft_batch_modes_device(self.h, kp, cp, None), "batch_modes_device")
        return [(k.to_numpy(np.int64, (int(m),)), c.to_numpy(np.complex128, (int(m),)))
                for k, c, m in zip(kd, cd, counts)]

    def modes_dict(self, i):
        cnt = C.c_int64(0)
        check(lib().mpsft_batch_modes_count(self.h, i, C.byref(cnt)), "batch_modes_count")
        m = cnt.value
        keys = np.zeros(max(m, 1), np.int64)
        coefs = np.zeros(2 * max(m, 1), np.float64)
        check(lib().mpsft_batch_modes_get(self.h, i, keys.ctypes.data_as(C.POINTER(C.c_int64)),
                                          coefs.ctypes.data_as(C.POINTER(C.c_double)), m,
                                          C.byref(cnt)), "batch_modes_get")
        return {int(a): complex(b) for a, b in zip(keys[:m], coefs.view(np.complex128)[:m])}

    def log(self, i):
        cap = 1 << 16
        bins = np.zeros(cap, np.int32)
        found = np.zeros(cap, np.int32)
        n_it = C.c_int32()
        check(lib().mpsft_batch_log_get(self.h, i, bins.ctypes.data_as(C.POINTER(C.c_int32)),
                                        found.ctypes.data_as(C.POINTER(C.c_int32)), cap,
                                        C.byref(n_it)), "batch_log_get")
        return bins[:n_it.value].copy(), found[:n_it.value].copy()

