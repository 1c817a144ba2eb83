"""Batched COORDINATION: many independent single-group systems in one step.

Each system k has its own atom count n_k and periodic cell; all share one
switching function and list radius.  This replaces running one

    c: COORDINATION GROUPA=1-<n_k> SWITCH={...} NLISTCELLS NL_CUTOFF=<r> NL_STRIDE=1

action per system (src/colvar/CoordinationBase.cpp:137-232) and collecting
the per-system outputs: values[nsys], derivatives of the concatenated
positions, virials[nsys, 3, 3].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A
from .coordination import parse_switch


class CoordinationBatch:
    def __init__(self, natoms, switch: str, rlist: float, device: int = 0):
        self.natoms = np.ascontiguousarray(np.asarray(natoms, dtype=np.int64))
        self.nsys = len(self.natoms)
        self.offsets = np.concatenate([[0], np.cumsum(self.natoms)])
        self.switch = parse_switch(switch)
        self._b = C.c_void_p()
        A.check(A.lib().p2a_batch_create(C.byref(self._b), int(device), self.nsys, self.natoms.ctypes.data,
                                         C.byref(self.switch), float(rlist)))

    def close(self):
        if getattr(self, "_b", None) is not None and self._b.value:
            A.lib().p2a_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_boxes(self, boxes):
        b = np.ascontiguousarray(np.asarray(boxes, dtype=np.float64).reshape(self.nsys, 9))
        A.check(A.lib().p2a_batch_set_boxes(self._b, b.ctypes.data))

    def set_stream(self, handle):
        A.check(A.lib().p2a_batch_set_stream(self._b, C.c_void_p(handle or 0)))

    def set_graph(self, on: bool):
        A.check(A.lib().p2a_batch_set_graph(self._b, int(bool(on))))

    def set_tuning(self, chunk: int = 64, blocks_per_cu: int = 0):
        A.check(A.lib().p2a_batch_set_tuning(self._b, int(chunk), int(blocks_per_cu)))

    def autotune(self, pos_ptr: int, reps: int = 3):
        ch, bpc = C.c_int(), C.c_int()
        A.check(A.lib().p2a_batch_autotune(self._b, C.c_void_p(pos_ptr), int(reps), C.byref(ch), C.byref(bpc)))
        return ch.value, bpc.value

    def calculate_device(self, pos_ptr: int, values_ptr: int, deriv_ptr: int = 0, virials_ptr: int = 0):
        """Device pointers; returns when the step has finished."""
        A.check(A.lib().p2a_batch_calculate_device(self._b, C.c_void_p(pos_ptr), C.c_void_p(values_ptr),
                                                   C.c_void_p(deriv_ptr or None), C.c_void_p(virials_ptr or None)))

    def calculate_torch(self, pos, out=None):
        """pos: contiguous float64 CUDA tensor [sum natoms, 3] ->
        (values[nsys], deriv[sum natoms, 3], virials[nsys, 3, 3])."""
        import torch

        if not (pos.is_cuda and pos.dtype == torch.float64 and pos.is_contiguous()):
            raise ValueError("positions must be a contiguous float64 CUDA tensor")
        if pos.shape != (int(self.offsets[-1]), 3):
            raise ValueError(f"expected positions of shape ({int(self.offsets[-1])}, 3)")
        if out is None:
            out = (torch.empty(self.nsys, dtype=torch.float64, device=pos.device),
                   torch.empty_like(pos),
                   torch.empty((self.nsys, 3, 3), dtype=torch.float64, device=pos.device))
        v, d, w = out
        torch.cuda.current_stream(pos.device).synchronize()
        self.calculate_device(pos.data_ptr(), v.data_ptr(), d.data_ptr(), w.data_ptr())
        return v, d, w

    def pair_counts(self):
        out = np.zeros(self.nsys, dtype=np.int64)
        A.check(A.lib().p2a_batch_get_pair_counts(self._b, out.ctypes.data))
        return out

    def stats(self):
        st = A.p2a_batch_stats()
        A.check(A.lib().p2a_batch_get_stats(self._b, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}
