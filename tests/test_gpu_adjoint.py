"""The differentiable recovery on the GPU: replay of a recorded solve, the
hand-written adjoint, and the torch operator (torch.ops.mpsft.recover).

Bars:
  replay of the recorded solve on its own input: the solve's support, and
    coefficients within 1e-12 (FP64) of the solve's;
  linearity of the replayed map: |F(x+d) - F x - F d| <= 1e-12 max|F x|;
  adjoint: <F d, g> = <d, F^H g> within 1e-12 relative (FP64);
  gradient of the loss sum |c - c_target|^2 (torch convention, x.grad)
    against FP64 central finite differences of the same loss (step 1e-6,
    decisions held fixed, FP64 replays of the complex64 signals) along 64
    random directions: within 1e-5 relative to the largest directional
    derivative; against the gradient of the FP64 forward on the same
    (complex64-valued) signals: within 1e-6 relative.
"""
import numpy as np
import pytest

from mpsft_amd.signals import loss_targets, make_signal

pytestmark = pytest.mark.gpu


def _recorded(gpu, n, k, sigma, seed, trials=1, precision=0, sig_seed=3):
    x, truth = make_signal(n, k, sigma, seed=sig_seed)
    if precision:
        x = x.astype(np.complex64)
    s = gpu.Solver(n, gpu.solve_options(k, sigma, trials=trials), precision)
    s.set_record(True)
    xd = s.upload(x)
    s.solve(xd, seed)
    return s, x, xd, truth


@pytest.mark.parametrize("precision", [0, 1])
def test_replay_reproduces_solve_and_is_linear(gpu, precision):
    n, k = 1 << 16, 60
    s, x, xd, _ = _recorded(gpu, n, k, 1e-6, seed=7, precision=precision)
    solved = s.modes_dict()
    s.replay(xd)
    rep = s.modes_dict()
    assert set(rep) == set(solved)
    scale = max(abs(v) for v in solved.values())
    assert max(abs(rep[q] - v) for q, v in solved.items()) <= 1e-12 * scale
    rng = np.random.default_rng(11)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x64 = x.astype(np.complex128)
    fx, fd, fs = [], [], []
    for v, out in ((x64, fx), (d, fd), (x64 + d, fs)):
        s.replay(gpu.DeviceArray.from_numpy(v), precision=gpu.MPSFT_FP64)
        out.append(s.modes_dict())
    fx, fd, fs = fx[0], fd[0], fs[0]
    assert max(abs(fs[q] - fx[q] - fd[q]) for q in solved) <= 1e-12 * scale


@pytest.mark.parametrize("trials", [1, 3])
def test_adjoint_dot_product(gpu, trials):
    n, k = 1 << 18, 200
    s, x, xd, _ = _recorded(gpu, n, k, 1e-6, seed=5, trials=trials)
    keys, _ = s.modes()
    rng = np.random.default_rng(13)
    d = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    s.replay(gpu.DeviceArray.from_numpy(d))
    fd = s.modes_dict()
    g = rng.standard_normal(keys.size) + 1j * rng.standard_normal(keys.size)
    adj = s.adjoint(keys, g)
    lhs = np.vdot(g, np.array([fd[int(q)] for q in keys]))
    rhs = np.vdot(adj, d)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)
    # keys repeated add; keys outside [0, n) carry nothing
    keys2 = np.concatenate([keys, keys[:5], np.array([-1, n, n + 7], np.int64)])
    g2 = np.concatenate([g, g[:5], np.ones(3, np.complex128)])
    adj2 = s.adjoint(keys2, g2)
    g3 = g.copy()
    g3[:5] *= 2
    adj3 = s.adjoint(keys, g3)
    assert np.abs(adj2 - adj3).max() <= 1e-12 * np.abs(adj3).max()


def test_adjoint_needs_a_recording(gpu):
    n, k = 1 << 14, 10
    x, _ = make_signal(n, k, 1e-6, seed=1)
    s = gpu.Solver(n, gpu.solve_options(k, 1e-6))
    xd = s.upload(x)
    s.solve(xd, 1)
    with pytest.raises(gpu.MpsftError):  # not recorded
        s.replay(xd)
    s.set_record(True)
    s.solve(xd, 1)
    keys, coefs = s.modes()
    s.adjoint(keys, coefs)
    s.set_modes({int(keys[0]): 1.0})
    with pytest.raises(gpu.MpsftError):  # the map changed since
        s.adjoint(keys, coefs)


def _loss(got, target):
    return sum(abs(v - target.get(q, 0.0)) ** 2 for q, v in got.items())


def _targets_dense(t, keys, dev):
    import torch
    return torch.tensor([complex(t.get(int(q), 0.0)) for q in keys.cpu().numpy()],
                        dtype=torch.complex128, device=dev)


@pytest.mark.timeout(600)
def test_autograd_config5_against_finite_differences(gpu):
    """32 complex64 signals, n = 2^20, S = 100, sigma = 1e-6; loss
    sum |c_hat - c_target|^2 with c_target = truth (1 + 0.01 CN(0,1)),
    through the batched operator (one batched solve, one batched adjoint)."""
    import torch

    from mpsft_amd import autograd as ag

    count, n, k, sigma = 32, 1 << 20, 100, 1e-6
    dev = torch.device("cuda:0")
    xs, targets = [], []
    for i in range(count):
        x, truth = make_signal(n, k, sigma, seed=100 + i, dtype=np.complex64)
        targets.append(loss_targets(truth, seed=200 + i))
        xs.append(torch.from_numpy(x).to(dev).requires_grad_(True))
    seeds = [300 + i for i in range(count)]
    batch = gpu.Batch([n] * count, [gpu.solve_options(k, sigma)] * count, gpu.MPSFT_MIXED)
    keys, coefs = ag.recover_batch(xs, batch, seeds)
    loss = sum((c - _targets_dense(t, q, dev)).abs().pow(2).sum()
               for q, c, t in zip(keys, coefs, targets))
    loss.backward()
    grads = [x.grad.detach().to(torch.complex128) for x in xs]
    assert all(bool(torch.isfinite(torch.view_as_real(g)).all()) for g in grads)

    # The same loss through the FP64 forward (single-signal operator) on the
    # same signal values.
    scale, diff = 0.0, 0.0
    for i, (x, t, sd) in enumerate(zip(xs, targets, seeds)):
        x64 = x.detach().to(torch.complex128).requires_grad_(True)
        s64 = gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_FP64)
        kk, cc = ag.recover(x64, s64, sd)
        assert torch.equal(kk.sort().values, keys[i].sort().values)
        (cc - _targets_dense(t, kk, dev)).abs().pow(2).sum().backward()
        scale = max(scale, float(x64.grad.abs().max()))
        diff = max(diff, float((x64.grad - grads[i]).abs().max()))
    assert diff <= 1e-6 * scale, (diff, scale)

    # FP64 central differences of the loss with the decisions held (FP64
    # replays of each signal's recorded map).
    h = 1e-6
    base = [x.detach().to(torch.complex128) for x in xs]
    gen = torch.Generator(device=dev)
    gen.manual_seed(17)
    fds, ans = [], []
    for _ in range(64):
        fd = 0.0
        an = 0.0
        for i, (x0, g, t) in enumerate(zip(base, grads, targets)):
            d = torch.randn(n, generator=gen, device=dev, dtype=torch.complex128)
            vals = []
            for sgn in (1.0, -1.0):
                xp = (x0 + (sgn * h) * d).contiguous()
                batch.replay(i, xp, precision=gpu.MPSFT_FP64)
                vals.append(_loss(batch.modes_dict(i), t))
            fd += (vals[0] - vals[1]) / (2 * h)
            an += float(torch.vdot(g, d).real)
        fds.append(fd)
        ans.append(an)
    fds, ans = np.array(fds), np.array(ans)
    assert np.abs(fds - ans).max() <= 1e-5 * np.abs(fds).max(), (fds[:4], ans[:4])


def test_batched_adjoint_matches_single(gpu):
    """mpsft_batch_adjoint over 6 signals of different n and S equals each
    signal's single-signal adjoint of the same recorded solve."""
    ns = [1 << 14, 1 << 16, 1 << 18, 1 << 15, 1 << 17, 1 << 16]
    ks = [10, 40, 150, 25, 90, 60]
    seeds = [5, 6, 7, 8, 9, 10]
    xs = [make_signal(n, k, 1e-6, seed=40 + i)[0] for i, (n, k) in enumerate(zip(ns, ks))]
    opts = [gpu.solve_options(k, 1e-6) for k in ks]
    b = gpu.Batch(ns, opts)
    b.set_record(True)
    _, stats = b.solve(b.upload(xs), seeds)
    assert stats == [0] * len(ns)
    rng = np.random.default_rng(3)
    keys_l = [np.array(sorted(b.modes_dict(i)), np.int64) for i in range(len(ns))]
    g_l = [rng.standard_normal(q.size) + 1j * rng.standard_normal(q.size) for q in keys_l]
    adj = b.adjoint(keys_l, g_l)
    for i, (n, x) in enumerate(zip(ns, xs)):
        s = gpu.Solver(n, opts[i])
        s.set_record(True)
        s.solve(s.upload(x), seeds[i])
        a1 = s.adjoint(keys_l[i], g_l[i])
        assert np.abs(a1 - adj[i]).max() <= 1e-12 * np.abs(a1).max(), i
