"""A METAD step on device-resident CVs with the bias force put back on the
atoms, every kernel on one stream.

This is the loop PlumedMain::calc runs (src/core/PlumedMain.cpp:1148-1403)
for an input such as

    c:  COORDINATION GROUPA=1-100000 SWITCH={RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2}
        NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10
    c1: COM ATOMS=<group 1>
    c2: COM ATOMS=<group 2>
    d:  DISTANCE ATOMS=c1,c2
    m:  METAD ARG=c,d SIGMA=20,0.05 HEIGHT=1.2 PACE=500 BIASFACTOR=10 TEMP=300
        GRID_MIN=... GRID_MAX=... GRID_BIN=512,512

with the positions already on the GPU:

  1. justCalculate: COORDINATION (value, derivatives, box derivative), then
     the centres and their DISTANCE, then MetaD::calculate (bias and the
     forces -dV/ds on the CVs, src/bias/MetaD.cpp:2315-2375);
  2. backwardPropagate: Bias::apply (src/bias/Bias.cpp:59-70) hands each CV
     its force, Colvar::apply / ActionWithVirtualAtom::apply and
     ActionWithValue::checkForForces (src/core/ActionWithValue.cpp:337-413)
     turn it into forces on the atoms and on the cell;
  3. update: MetaD::update deposits a hill every PACE steps, never on the
     first step (MetaD.cpp:2377-2470).

The CV values land in one device array [ncv], which the METAD step reads
directly: nothing goes back to the host inside a step.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A
from .coordination import Coordination, _box9
from .metad import MetaD


class CenterDistance:
    """c1: COM ATOMS=group1, c2: COM ATOMS=group2, d: DISTANCE ATOMS=c1,c2
    (src/vatom/Center.cpp, src/colvar/Distance.h).  Indices are 0-based rows
    of the positions array (PLUMED's atom numbers minus one); masses
    [n1+n2] (group 1 then group 2) or None for equal masses."""

    def __init__(self, natoms: int, group1, group2, masses=None, pbc: bool = True, device: int = 0):
        g1 = np.ascontiguousarray(group1, dtype=np.int64)
        g2 = np.ascontiguousarray(group2, dtype=np.int64)
        m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
        if m is not None and m.size != g1.size + g2.size:
            raise ValueError("one mass per listed atom (group 1 then group 2)")
        self.n1, self.n2, self.natoms = int(g1.size), int(g2.size), int(natoms)
        self._ctx = C.c_void_p()
        A.check(A.lib().p2a_cdist_create(C.byref(self._ctx), int(device), self.natoms, self.n1, g1.ctypes.data,
                                         self.n2, g2.ctypes.data, None if m is None else m.ctypes.data,
                                         1 if pbc else 0))

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            A.lib().p2a_cdist_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def calculate_device(self, stream: int, box, pos_ptr: int, value_ptr: int, deriv_ptr: int, virial_ptr: int):
        b = _box9(box)
        A.check(A.lib().p2a_cdist_calculate_device(self._ctx, C.c_void_p(stream), b.ctypes.data, C.c_void_p(pos_ptr),
                                                   C.c_void_p(value_ptr), C.c_void_p(deriv_ptr),
                                                   C.c_void_p(virial_ptr)))

    def apply_force_device(self, stream: int, f_ptr: int, deriv_ptr: int, virial_cv_ptr: int, forces_ptr: int,
                           virial_ptr: int):
        A.check(A.lib().p2a_cdist_apply_force_device(self._ctx, C.c_void_p(stream), C.c_void_p(f_ptr),
                                                     C.c_void_p(deriv_ptr), C.c_void_p(virial_cv_ptr),
                                                     C.c_void_p(forces_ptr), C.c_void_p(virial_ptr)))

    def calculate_torch(self, pos, box):
        """(value[1], deriv[n1+n2,3], virial[3,3]) as CUDA tensors, enqueued on torch's current stream."""
        import torch

        dev = pos.device
        out = (torch.empty(1, dtype=torch.float64, device=dev),
               torch.empty((self.n1 + self.n2, 3), dtype=torch.float64, device=dev),
               torch.empty((3, 3), dtype=torch.float64, device=dev))
        self.calculate_device(torch.cuda.current_stream(dev).cuda_stream, box, pos.data_ptr(),
                              *(t.data_ptr() for t in out))
        return out


class MetadStep:
    """One MD step of METAD on COORDINATION (and optionally a DISTANCE between
    two centres) with forces on the atoms, enqueued on torch's current stream.

    ``coord`` is a Coordination action, ``cdist`` a CenterDistance or None,
    ``metad`` a MetaD whose ARG lists the coordination first (and the
    distance second).  ``step(step, pos, box)`` returns the device tensors of
    that step: cv [ncv], bias [1], cv_force [ncv] (= -dV/ds), forces
    [natoms,3] (the METAD forces on the atoms) and virial [3,3] (the forces on
    the cell), plus height [1] of a hill deposited at this step (0 if none)."""

    def __init__(self, coord: Coordination, metad: MetaD, cdist: CenterDistance | None = None):
        import torch

        ncv = 1 + (cdist is not None)
        if metad.ncv != ncv:
            raise ValueError(f"the METAD takes {metad.ncv} arguments, the chain provides {ncv}")
        self.coord, self.cdist, self.metad = coord, cdist, metad
        self.natoms = coord.natoms
        dev = torch.device("cuda", coord.device)
        f64 = dict(dtype=torch.float64, device=dev)
        self.cv = torch.zeros(ncv, **f64)
        self.bias = torch.zeros(1, **f64)
        self.cv_force = torch.zeros(ncv, **f64)
        self.height = torch.zeros(1, **f64)
        self.d_coord = torch.empty((self.natoms, 3), **f64)
        self.w_coord = torch.empty((3, 3), **f64)
        self.forces = torch.zeros((self.natoms, 3), **f64)
        self.virial = torch.zeros((3, 3), **f64)
        if cdist is not None:
            self.d_dist = torch.empty((cdist.n1 + cdist.n2, 3), **f64)
            self.w_dist = torch.empty((3, 3), **f64)

    def step(self, step: int, pos, box):
        import torch

        if not (pos.is_cuda and pos.dtype == torch.float64 and pos.is_contiguous() and pos.shape[0] == self.natoms):
            raise ValueError("positions must be a contiguous float64 CUDA tensor [natoms,3]")
        stream = torch.cuda.current_stream(pos.device).cuda_stream
        cvp = self.cv.data_ptr()
        # 1. CVs and bias
        self.coord.set_stream(stream)
        self.coord.calculate_device(pos.data_ptr(), box, step, cvp, self.d_coord.data_ptr(), self.w_coord.data_ptr())
        if self.cdist is not None:
            self.cdist.calculate_device(stream, box, pos.data_ptr(), cvp + 8, self.d_dist.data_ptr(),
                                        self.w_dist.data_ptr())
        self.metad.step_device(step, cvp, self.bias.data_ptr(), self.cv_force.data_ptr(), self.height.data_ptr(),
                               stream=stream)
        # 2. forces on atoms and cell (this step's METAD forces only)
        self.forces.zero_()
        self.virial.zero_()
        fp = self.cv_force.data_ptr()
        A.check(A.lib().p2a_apply_force_device(C.c_void_p(stream), self.natoms, C.c_void_p(fp),
                                               C.c_void_p(self.d_coord.data_ptr()),
                                               C.c_void_p(self.w_coord.data_ptr()),
                                               C.c_void_p(self.forces.data_ptr()), C.c_void_p(self.virial.data_ptr())))
        if self.cdist is not None:
            self.cdist.apply_force_device(stream, fp + 8, self.d_dist.data_ptr(), self.w_dist.data_ptr(),
                                          self.forces.data_ptr(), self.virial.data_ptr())
        return {"cv": self.cv, "bias": self.bias, "cv_force": self.cv_force, "forces": self.forces,
                "virial": self.virial, "height": self.height}
