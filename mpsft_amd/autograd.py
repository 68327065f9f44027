"""The differentiable recovery: torch.ops.mpsft.recover, a C++
torch::autograd::Function (csrc/torch_recover.cpp, built into
libmpsft_torch.so) over the HIP library's record / adjoint entry points.

Forward: Demo1::Run (src/demo1.cc:36-59) on a device signal, recording its
decisions; the coefficients are then c = F x with F linear.  Backward:
x.grad = F^H grad_c, computed by the hand-written adjoint kernels.  The
reference has no differentiable form.
"""
import os

from ._lib import HERE, MPSFT_FP64, MPSFT_MIXED

TORCH_LIB = os.path.join(HERE, "libmpsft_torch.so")
_loaded = False


def load():
    """Registers torch.ops.mpsft (fails if libmpsft_torch.so is not built)."""
    global _loaded
    if not _loaded:
        import torch
        if not os.path.exists(TORCH_LIB):
            raise RuntimeError(f"{TORCH_LIB} is not built: run python -m mpsft_amd.build")
        torch.ops.load_library(TORCH_LIB)
        _loaded = True
    import torch
    return torch.ops.mpsft


def recover(x, solver, seed, stream=None):
    """Recovers the modes of x (a 1-D complex128 / complex64 cuda tensor
    matching the solver's n and precision) and returns (keys, coefs) as
    device tensors; coefs (complex128) is differentiable with respect to x.
    The solver must not solve again before the backward pass."""
    import torch
    ops = load()
    if stream is None:
        stream = torch.cuda.current_stream(x.device).cuda_stream
    return ops.recover(x, int(solver.h.value), int(seed), int(stream))


def recover_batch(xs, batch, seeds, stream=None):
    """recover over a Batch: xs a list of 1-D device tensors (signal i of
    n[i] samples in the batch's precision); returns (keys list, coefs list)
    with every coefs[i] differentiable with respect to xs[i].  One batched
    solve forward, one batched adjoint backward."""
    import torch
    ops = load()
    if stream is None:
        stream = torch.cuda.current_stream(xs[0].device).cuda_stream
    return ops.recover_batch(list(xs), int(batch.h.value), [int(v) for v in seeds], int(stream))


def precision_of(dtype):
    import torch
    return MPSFT_MIXED if dtype == torch.complex64 else MPSFT_FP64
