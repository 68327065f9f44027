"""Synthetic test signals (host, numpy), after the reference's gen.cc.

The reference builds x as x-hat (sparse modes plus complex Gaussian noise of
per-component std sigma/sqrt(2n), GenerateXhatAlt, src/gen.cc:78-89) followed
by an unnormalised backward DFT (src/demo1.cc:30-33), so each time sample
carries noise of E|e|^2 = sigma^2.  These generators keep that construction
and add the mode layouts of the benchmark workloads: log-uniform
magnitudes, clustered locations.  numpy's own random generator is used; the
reference's mt19937 stream is only needed for its RNG-driven transforms,
which the solver draws itself.  The n-point backward DFT runs through
torch's multithreaded CPU FFT (numpy's is single-threaded: 18 s at n = 2^26).
"""
import numpy as np
import torch


def make_signal(n, k, sigma, seed, magnitude="unit", clusters=None, dtype=np.complex128,
                with_xhat=False):
    """Returns (x, xhat_modes) with x a length-n signal of k modes, and with
    with_xhat also its x-hat (modes plus noise, complex128) last.

    magnitude: "unit" (|c| = 1) or "loguniform" (|c| = 10^U[-1, 1]).
    clusters: None, or (n_centres, fraction, halfwidth): that fraction of
    the modes lies within +-halfwidth bins of n_centres random centres.
    The phases are uniform.  xhat_modes maps location -> coefficient.
    """
    rng = np.random.default_rng(seed)
    if clusters is None:
        locs = rng.choice(n, size=k, replace=False)
    else:
        n_c, frac, hw = clusters
        k_c = int(round(k * frac))
        centres = rng.choice(n, size=n_c, replace=False)
        chosen = set()
        while len(chosen) < k_c:
            c = centres[rng.integers(n_c)]
            chosen.add(int((c + rng.integers(-hw, hw + 1)) % n))
        while len(chosen) < k:
            chosen.add(int(rng.integers(n)))
        locs = np.array(sorted(chosen), dtype=np.int64)
        rng.shuffle(locs)
    if magnitude == "unit":
        mag = np.ones(k)
    elif magnitude == "loguniform":
        mag = 10.0 ** rng.uniform(-1.0, 1.0, size=k)
    else:
        raise ValueError(magnitude)
    coef = mag * np.exp(2j * np.pi * rng.uniform(0.0, 1.0, size=k))
    xh = np.zeros(n, np.complex128)
    xh[locs] = coef
    scale = sigma / np.sqrt(2.0 * n)
    xh += (rng.standard_normal(n) + 1j * rng.standard_normal(n)) * scale
    x = torch.fft.ifft(torch.from_numpy(xh)).numpy() * n  # unnormalised backward DFT
    if with_xhat:
        return x.astype(dtype), dict(zip(locs.tolist(), coef.tolist())), xh
    return x.astype(dtype), dict(zip(locs.tolist(), coef.tolist()))


def make_signal_device(n, k, sigma, seed):
    """make_signal's construction (unit magnitudes, uniform phases, distinct
    uniform locations) computed on the GPU with torch: returns the signal as
    a complex64 torch tensor on cuda, the rms rounding error of storing it
    in complex64 (sigma_q), and the modes {location: coefficient}."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(int(seed))
    locs = torch.randperm(n, generator=g, device="cuda")[:k]
    phase = torch.rand(k, generator=g, device="cuda", dtype=torch.float64)
    coef = torch.polar(torch.ones(k, device="cuda", dtype=torch.float64), 2 * np.pi * phase)
    scale = sigma / np.sqrt(2.0 * n)
    xh = torch.complex(torch.randn(n, generator=g, device="cuda", dtype=torch.float64),
                       torch.randn(n, generator=g, device="cuda", dtype=torch.float64)) * scale
    xh[locs] += coef
    x = torch.fft.ifft(xh) * n
    x64 = x.to(torch.complex64)
    sigma_q = float((x - x64.to(torch.complex128)).abs().pow(2).mean().sqrt())
    truth = dict(zip(locs.cpu().tolist(), coef.cpu().numpy().tolist()))
    return x64, sigma_q, truth


def demo1_signal_device(n, k, sigma, seed):
    """The signal of Demo1(opt, n, k, sigma) (src/demo1.cc:27-34) on the GPU
    from the reference's RNG stream seeded with seed: GenerateModeMap and
    GenerateXhatAlt by mpsft_generate_signal, then x = the unnormalised
    backward DFT of x-hat with torch's GPU FFT.  Returns (x, xhat, truth):
    complex128 cuda tensors and {location: coefficient}."""
    import ctypes as C
    from ._lib import check, lib
    xh = torch.empty(n, dtype=torch.complex128, device="cuda")
    keys = np.zeros(k, np.int64)
    coefs = np.zeros(2 * k, np.float64)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    check(lib().mpsft_generate_signal(n, k, sigma, seed, C.c_void_p(xh.data_ptr()),
                                      keys.ctypes.data_as(C.POINTER(C.c_int64)),
                                      coefs.ctypes.data_as(C.POINTER(C.c_double)), stream),
          "mpsft_generate_signal")
    x = torch.fft.ifft(xh) * n
    return x, xh, dict(zip(keys.tolist(), coefs.view(np.complex128).tolist()))


def quantization_sigma(x128, x64):
    """RMS per-sample error of storing x in complex64: the noise the mixed
    form sees on top of sigma."""
    d = x64.astype(np.complex128) - x128
    return float(np.sqrt(np.mean(np.abs(d) ** 2)))


def quantization_sigma_c64(x64):
    """quantization_sigma from the complex64 samples alone, for a caller
    that never held the complex128 signal: rounding to nearest leaves each
    component an error uniform over one float32 spacing of the stored value
    (variance spacing^2 / 12), so sigma_q^2 = mean(spacing(re)^2 +
    spacing(im)^2) / 12."""
    x64 = np.asarray(x64, np.complex64)
    sr = np.spacing(np.abs(x64.real)).astype(np.float64)
    si = np.spacing(np.abs(x64.imag)).astype(np.float64)
    return float(np.sqrt(np.mean(sr * sr + si * si) / 12.0))


def loss_targets(truth, seed, rel=0.01):
    """Targets of the differentiable-recovery loss: each true coefficient
    times (1 + rel * z), z complex Gaussian with E|z|^2 = 1."""
    rng = np.random.default_rng(seed)
    keys = sorted(truth)
    z = (rng.standard_normal(len(keys)) + 1j * rng.standard_normal(len(keys))) / np.sqrt(2.0)
    return {key: truth[key] * (1.0 + rel * zz) for key, zz in zip(keys, z)}
