"""Signal synthesis on the device (mpsft_generate_signal): Demo1's signal
(src/demo1.cc:27-34) from the reference's RNG stream, checked against the
oracle's GenerateModeMap / GenerateXhatAlt on the same seed; and a solve of
such a signal checked with PostAnalyze on the device against the x-hat it
was made from (the dense global oracle)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k,sigma,seed", [(1 << 14, 20, 1e-2, 5), (1 << 20, 50, 1e-8, 123534),
                                            (4099, 7, 0.5, 77)])
def test_generate_signal_matches_oracle(gpu, orc, n, k, sigma, seed):
    from mpsft_amd.signals import demo1_signal_device
    x, xh, truth = demo1_signal_device(n, k, sigma, seed)
    orc.seed(seed)
    mm = orc.generate_mode_map(n, k)
    xh_o = orc.generate_xhat_alt(n, mm, sigma)
    # the mode map: the host part, bit for bit
    assert truth == mm.to_dict()
    # x-hat: the same stream; the device's log and sqrt (m = sqrt(-2 ln r2 /
    # r2)) may differ from glibc's by an ulp or two, so each sample agrees
    # to a few ulps, and most bit for bit
    got = xh.cpu().numpy()
    rel = np.abs(got - xh_o) / np.abs(xh_o)
    same = float(np.mean(got == xh_o))
    print(f"[gen] n={n}: max rel {rel.max():.2e}, identical samples {same:.3f}", flush=True)
    assert rel.max() <= 2e-15 and same > 0.5
    noise = np.delete(got, list(truth))
    # per-component std sigma/sqrt(2n): E|noise|^2 = sigma^2 / n per location
    assert abs(np.sqrt(np.mean(np.abs(noise) ** 2) * n) / sigma - 1) < 0.05
    # x = the backward DFT of x-hat
    x_np = np.fft.ifft(xh_o) * n
    assert np.abs(x.cpu().numpy() - x_np).max() <= 1e-12 * np.abs(x_np).max()


def test_solve_generated_signal_post_analyze(gpu, orc):
    """Config 1's shape (n = 2^20, S = 50, sigma = 1e-8) made on the device:
    the solve equals the oracle's on the same samples, and PostAnalyze on
    the device against the generated x-hat equals the oracle's."""
    from mpsft_amd.signals import demo1_signal_device
    n, k, sigma, seed = 1 << 20, 50, 1e-8, 3
    x, xh, truth = demo1_signal_device(n, k, sigma, seed)
    s = gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_FP64)
    iters = s.solve(x, 7)
    got = s.modes_dict()
    x_np = x.cpu().numpy()
    orc.seed(7)
    ref, bins_log, _ = orc.demo1_run(x_np, k, sigma)
    assert iters == len(bins_log) and set(got) == set(ref.to_dict())
    st = s.post_analyze(list(truth), xh)
    xh_np = xh.cpu().numpy()
    st_o = orc.post_analyze(n, orc.ModeMap(truth), xh_np, orc.ModeMap(got))
    assert st["hit"] == st_o["hit"] == k
    for key in ("hit_l1", "hit_l2", "l1", "l2", "max"):
        assert abs(st[key] - st_o[key]) <= 1e-12 * abs(st_o[key]), key
    # every true coefficient recovered to the noise level of its bin
    assert st["hit_l2"] < 1e-6
