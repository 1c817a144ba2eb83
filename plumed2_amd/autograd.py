"""COORDINATION as a differentiable torch function (learned-CV training).

When a gradient is wanted, the forward runs the pair kernels once for the
value and the derivatives and keeps the derivatives; the backward is then
grad * dCV/dx, with no second pass over the pairs.  Without a gradient
(torch.no_grad, or positions that do not require one) the forward runs the
cheaper value-only kernel.  Every forward keeps its own derivatives, so two
forwards followed by the backward of both give each its own gradient.
Positions may be float32 or float64 CUDA tensors; the gradient comes back in
the input dtype.  Both passes run on torch's current stream.
"""
from __future__ import annotations

import ctypes as C  # noqa: F401  (kept for callers that bind the C ABI through this module)

import torch

from . import _capi as A  # noqa: F401
from .coordination import Coordination, _box9


class CoordinationFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, action: Coordination, box, step: int = 0):
        if not positions.is_cuda:
            raise ValueError("positions must be a CUDA tensor")
        pos64 = positions.detach().to(torch.float64).contiguous()
        if pos64.shape != (action.natoms, 3):
            raise ValueError(f"expected positions of shape ({action.natoms}, 3)")
        val = torch.empty(1, dtype=torch.float64, device=positions.device)
        action.set_stream(torch.cuda.current_stream(positions.device).cuda_stream)
        want_grad = ctx.needs_input_grad[0]
        der = torch.empty_like(pos64) if want_grad else None
        # CoordinationBase::calculate: value and derivatives in one pass
        action.calculate_device(pos64.data_ptr(), box, step, val.data_ptr(), der.data_ptr() if want_grad else 0, 0)
        ctx.in_dtype = positions.dtype
        if want_grad:
            ctx.save_for_backward(der)
        return val.to(positions.dtype).reshape(())

    @staticmethod
    def backward(ctx, grad):
        (der,) = ctx.saved_tensors
        # one elementwise kernel: grad * der, written in the input dtype
        g = torch.empty(der.shape, dtype=ctx.in_dtype, device=der.device)
        torch.mul(der, grad.to(torch.float64), out=g)
        return g, None, None, None


def coordination(positions, action: Coordination, box, step: int = 0):
    """Differentiable COORDINATION value of ``positions`` (natoms x 3)."""
    return CoordinationFunction.apply(positions, action, box, step)


class CoordinationModule(torch.nn.Module):
    """torch.nn wrapper around one COORDINATION action line."""

    def __init__(self, line: str, box, device: int = 0):
        super().__init__()
        self.action = Coordination(line, device=device)
        self.box = box
        self.step = 0

    def forward(self, positions):
        out = coordination(positions, self.action, self.box, self.step)
        self.step += 1
        return out
