"""COORDINATION as a differentiable torch function (learned-CV training).

The forward runs the value-only pair kernel; the backward is the library's
backward entry point, which recomputes the pair terms with derivatives on the
neighbour structure the forward built, so no per-pair data is kept between
the two.  The forward records which build of the structure it used; when
another forward on the same action has replaced it before this backward (two
forwards, then the backward of both), the backward first rebuilds the
structure at its own positions.  Positions may be float32 or float64 CUDA tensors; the gradient comes
back in the input dtype.  Both passes run on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _capi as A
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
        action.calculate_device(pos64.data_ptr(), box, step, val.data_ptr(), 0, 0)
        ctx.generation = action.list_generation()
        ctx.action = action
        ctx.box = _box9(box)
        ctx.in_dtype = positions.dtype
        ctx.save_for_backward(pos64)
        return val.to(positions.dtype).reshape(())

    @staticmethod
    def backward(ctx, grad):
        (pos64,) = ctx.saved_tensors
        action = ctx.action
        der = torch.empty_like(pos64)
        action.set_stream(torch.cuda.current_stream(pos64.device).cuda_stream)
        if action.list_generation() != ctx.generation:
            A.check(A.lib().p2a_coord_rebuild_device(action._ctx, ctx.box.ctypes.data, C.c_void_p(pos64.data_ptr())))
        A.check(A.lib().p2a_coord_backward_device(action._ctx, ctx.box.ctypes.data, C.c_void_p(pos64.data_ptr()),
                                                  C.c_void_p(der.data_ptr()), None))
        g = grad.to(torch.float64) * der
        return g.to(ctx.in_dtype), None, None, None


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
