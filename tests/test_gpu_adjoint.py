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


@pytest.mark.timeout(600)
def test_autograd_config5_against_finite_differences(gpu):
    """32 complex64 signals, n = 2^20, S = 100, sigma = 1e-6; loss
    sum |c_hat - c_target|^2 with c_target = truth (1 + 0.01 CN(0,1))."""
    import torch

    from mpsft_amd import autograd as ag

    count, n, k, sigma = 32, 1 << 20, 100, 1e-6
    dev = torch.device("cuda:0")
    solvers, xs, targets, seeds = [], [], [], []
    for i in range(count):
        x, truth = make_signal(n, k, sigma, seed=100 + i, dtype=np.complex64)
        targets.append(loss_targets(truth, seed=200 + i))
        xs.append(torch.from_numpy(x).to(dev).requires_grad_(True))
        solvers.append(gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_MIXED))
        seeds.append(300 + i)
    loss = 0.0
    for x, s, t, sd in zip(xs, solvers, targets, seeds):
        keys, coefs = ag.recover(x, s, sd)
        tk = torch.tensor([complex(t.get(int(q), 0.0)) for q in keys.cpu().numpy()],
                          dtype=torch.complex128, device=dev)
        loss = loss + (coefs - tk).abs().pow(2).sum()
    loss.backward()
    grads = [x.grad.detach().cpu().numpy().astype(np.complex128) for x in xs]
    assert all(np.isfinite(g).all() for g in grads)

    # The same loss through the FP64 forward on the same signal values.
    grads64 = []
    for x, t, sd in zip(xs, targets, seeds):
        x64 = x.detach().to(torch.complex128).requires_grad_(True)
        s64 = gpu.Solver(n, gpu.solve_options(k, sigma), gpu.MPSFT_FP64)
        keys, coefs = ag.recover(x64, s64, sd)
        tk = torch.tensor([complex(t.get(int(q), 0.0)) for q in keys.cpu().numpy()],
                          dtype=torch.complex128, device=dev)
        (coefs - tk).abs().pow(2).sum().backward()
        grads64.append(x64.grad.cpu().numpy())
    scale = max(np.abs(g).max() for g in grads64)
    assert max(np.abs(a - b).max() for a, b in zip(grads, grads64)) <= 1e-6 * scale

    # FP64 central differences of the loss with the decisions held.
    h = 1e-6
    base = [x.detach().to(torch.complex128) for x in xs]
    gdev = [torch.from_numpy(g).to(dev) for g in grads]
    gen = torch.Generator(device=dev)
    gen.manual_seed(17)
    fds, ans = [], []
    for _ in range(64):
        fd = 0.0
        an = 0.0
        for s, x0, g, t in zip(solvers, base, gdev, targets):
            d = torch.randn(n, generator=gen, device=dev, dtype=torch.complex128)
            vals = []
            for sgn in (1.0, -1.0):
                xp = (x0 + (sgn * h) * d).contiguous()
                s.replay(xp, precision=gpu.MPSFT_FP64)
                vals.append(_loss(s.modes_dict(), t))
            fd += (vals[0] - vals[1]) / (2 * h)
            an += float(torch.vdot(g, d).real)
        fds.append(fd)
        ans.append(an)
    fds, ans = np.array(fds), np.array(ans)
    assert np.abs(fds - ans).max() <= 1e-5 * np.abs(fds).max(), (fds[:4], ans[:4])
