"""COORDINATION action, host side.

Mirrors the keyword interface of the reference action
(src/colvar/Coordination.cpp:113-157, src/colvar/CoordinationBase.cpp:30-232):
an input line such as

    c: COORDINATION GROUPA=1-1000 GROUPB=1001-2000 SWITCH={RATIONAL R_0=0.3 D_MAX=1.0}

builds the action; ``calculate(positions, box, step)`` returns the value, the
derivatives with respect to the requested atoms (GROUPA then GROUPB, the order
of CoordinationBase's requestAtoms) and the virial (box derivative).  All
arithmetic runs in the HIP library; errors raise Plumed2AmdError with the
reference's messages.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as A


def parse_switch(definition: str) -> A.p2a_switch:
    """SWITCH={...} body -> resolved parameters (SwitchingFunction::set(string))."""
    sw = A.p2a_switch()
    A.check(A.lib().p2a_switch_parse(definition.encode(), C.byref(sw)))
    return sw


def rational_switch(nn: int, mm: int, r0: float, d0: float = 0.0) -> A.p2a_switch:
    """NN/MM/R_0/D_0 form (SwitchingFunction::set(nn,mm,r0,d0))."""
    sw = A.p2a_switch()
    A.check(A.lib().p2a_switch_rational(int(nn), int(mm), float(r0), float(d0), C.byref(sw)))
    return sw


def describe_switch(sw: A.p2a_switch) -> str:
    buf = C.create_string_buffer(256)
    A.check(A.lib().p2a_switch_describe(C.byref(sw), buf, 256))
    return buf.value.decode()


class CoordinationSpec:
    """Parsed keywords of one COORDINATION line."""

    def __init__(self, line: str):
        self._spec = A.p2a_coord_spec()
        A.check(A.lib().p2a_coord_parse(line.encode(), C.byref(self._spec)))
        s = self._spec
        self.groupa = np.ctypeslib.as_array(s.groupa, shape=(s.na,)).copy() if s.na else np.zeros(0, np.int64)
        self.groupb = np.ctypeslib.as_array(s.groupb, shape=(s.nb,)).copy() if s.nb else np.zeros(0, np.int64)
        self.pbc = bool(s.pbc)
        self.pair = bool(s.pair)
        self.nl_mode = int(s.nl_mode)
        self.nl_cutoff = float(s.nl_cutoff)
        self.nl_stride = int(s.nl_stride)
        self.switch = s.sw

    def __del__(self):
        try:
            A.lib().p2a_coord_spec_free(C.byref(self._spec))
        except Exception:
            pass

    @property
    def natoms(self):
        return len(self.groupa) + len(self.groupb)


def _box9(box):
    b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(-1))
    if b.size == 3:
        b = np.diag(b).reshape(-1).copy()
    if b.size != 9:
        raise ValueError("box must have 3 (orthorhombic) or 9 entries")
    return b


class Coordination:
    """GPU COORDINATION action built from a reference-syntax input line."""

    def __init__(self, line: str, device: int = 0):
        self.spec = CoordinationSpec(line)
        self.device = device
        self._ctx = C.c_void_p()
        A.check(A.lib().p2a_coord_create(C.byref(self._ctx), int(device), C.byref(self.spec._spec)))

    def close(self):
        if getattr(self, "_ctx", None) is not None and self._ctx.value:
            A.lib().p2a_coord_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def natoms(self):
        return self.spec.natoms

    def calculate(self, positions, box, step: int = 0):
        """Host arrays in, host arrays out: (value, deriv[natoms,3], virial[3,3])."""
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).reshape(-1, 3))
        if pos.shape[0] != self.natoms:
            raise ValueError(f"expected {self.natoms} positions, got {pos.shape[0]}")
        b = _box9(box)
        val = C.c_double()
        der = np.empty((self.natoms, 3))
        vir = np.empty(9)
        A.check(A.lib().p2a_coord_calculate(self._ctx, int(step), b.ctypes.data, pos.ctypes.data, C.addressof(val),
                                            der.ctypes.data, vir.ctypes.data))
        return val.value, der, vir.reshape(3, 3)

    def set_stream(self, stream_handle: int | None):
        A.check(A.lib().p2a_coord_set_stream(self._ctx, C.c_void_p(stream_handle or 0)))

    def calculate_device(self, pos_ptr: int, box, step: int, value_ptr: int, deriv_ptr: int, virial_ptr: int):
        """Device pointers (ints) in and out; asynchronous on the context stream."""
        b = _box9(box)
        A.check(A.lib().p2a_coord_calculate_device(self._ctx, int(step), b.ctypes.data, C.c_void_p(pos_ptr),
                                                   C.c_void_p(value_ptr), C.c_void_p(deriv_ptr),
                                                   C.c_void_p(virial_ptr)))

    def calculate_torch(self, pos, box, step: int = 0, out=None):
        """torch.float64 CUDA tensor [natoms,3] -> (value[1], deriv[natoms,3], virial[3,3]) CUDA tensors,
        enqueued on torch's current stream."""
        import torch

        if not (pos.is_cuda and pos.dtype == torch.float64 and pos.is_contiguous()):
            raise ValueError("positions must be a contiguous float64 CUDA tensor")
        if out is None:
            out = (torch.empty(1, dtype=torch.float64, device=pos.device),
                   torch.empty((self.natoms, 3), dtype=torch.float64, device=pos.device),
                   torch.empty((3, 3), dtype=torch.float64, device=pos.device))
        val, der, vir = out
        self.set_stream(torch.cuda.current_stream(pos.device).cuda_stream)
        self.calculate_device(pos.data_ptr(), box, step, val.data_ptr(), der.data_ptr(), vir.data_ptr())
        return val, der, vir

    def list_generation(self) -> int:
        """Builds of the neighbour structure enqueued so far (no synchronisation)."""
        g = C.c_int64()
        A.check(A.lib().p2a_coord_list_generation(self._ctx, C.byref(g)))
        return g.value

    def rebuild_device(self, pos_ptr: int, box):
        """Build the neighbour structure now at these device positions."""
        b = _box9(box)
        A.check(A.lib().p2a_coord_rebuild_device(self._ctx, b.ctypes.data, C.c_void_p(pos_ptr)))

    def stats(self):
        st = A.p2a_coord_stats()
        A.check(A.lib().p2a_coord_get_stats(self._ctx, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "pad_"}

    def set_profiling(self, on: bool, every: int = 1):
        """Time the pair kernels of every `every`-th call with HIP events."""
        A.check(A.lib().p2a_coord_set_profiling(self._ctx, int(every) if on else 0))

    def set_tuning(self, chunk: int = 64, blocks_per_cu: int = 4):
        """Entries per wave (32/64) and persistent grid of the pair kernel."""
        A.check(A.lib().p2a_coord_set_tuning(self._ctx, int(chunk), int(blocks_per_cu)))

    def set_skin(self, skin: float):
        """Without NLIST/NLISTCELLS: keep the pair list built at D_MAX + skin
        while no atom has moved more than skin/2 (0: rebuild every call)."""
        A.check(A.lib().p2a_coord_set_skin(self._ctx, float(skin)))


def coordination_numbers_torch(switch, pos, box=None, species=None, species_a=None, species_b=None):
    """COORDINATIONNUMBER (src/symfunc/CoordinationNumbers.cpp:326-372) values on
    the GPU: cn[k] = sum over environment atoms j != row k of s(r_kj).

    ``switch`` is a SWITCH definition string or a p2a_switch; ``pos`` a contiguous
    float64 CUDA tensor [natoms,3]; ``box`` None for NOPBC, else 3x3 rows =
    lattice vectors.  SPECIES=list gives rows = environment = list; SPECIESA and
    SPECIESB give rows = A, environment = B.  Indices are 0-based rows of ``pos``
    (PLUMED's atom numbers minus one).  Returns a float64 CUDA tensor [nrows],
    enqueued on torch's current stream."""
    import torch

    if not (pos.is_cuda and pos.dtype == torch.float64 and pos.is_contiguous() and pos.dim() == 2
            and pos.shape[1] == 3):
        raise ValueError("positions must be a contiguous float64 CUDA tensor [natoms,3]")
    if species is not None:
        if species_a is not None or species_b is not None:
            raise ValueError("give SPECIES or SPECIESA/SPECIESB, not both")
        rows = env = list(species)
    elif species_a is not None and species_b is not None:
        rows, env = list(species_a), list(species_b)
    else:
        raise ValueError("missing atoms input use SPECIES or SPECIESA/SPECIESB")
    n = int(pos.shape[0])
    for idx in (rows, env):
        if any(int(i) < 0 or int(i) >= n for i in idx):
            raise ValueError("atom index outside the positions array")
    sw = parse_switch(switch) if isinstance(switch, str) else switch
    dev = pos.device
    rows_t = torch.tensor(rows, dtype=torch.int32, device=dev)
    env_t = torch.tensor(env, dtype=torch.int32, device=dev)
    cn = torch.empty(len(rows), dtype=torch.float64, device=dev)
    b9 = None if box is None else _box9(box)
    stream = torch.cuda.current_stream(dev).cuda_stream
    A.check(A.lib().p2a_coordination_numbers_device(
        stream, C.byref(sw), None if b9 is None else b9.ctypes.data, 0 if box is None else 1, n, len(rows),
        rows_t.data_ptr() if rows else None, len(env), env_t.data_ptr() if env else None, pos.data_ptr(),
        cn.data_ptr()))
    return cn


def _atoms_arg(idx):
    """0-based indices, in the given order, as a PLUMED atom list (1-based,
    consecutive runs written as ranges)."""
    idx = [int(i) for i in idx]
    out = []
    k = 0
    while k < len(idx):
        j = k
        while j + 1 < len(idx) and idx[j + 1] == idx[j] + 1:
            j += 1
        out.append(f"{idx[k] + 1}" if j == k else f"{idx[k] + 1}-{idx[j] + 1}")
        k = j + 1
    return ",".join(out)


class CoordinationNumbers:
    """COORDINATIONNUMBER (src/symfunc/CoordinationNumbers.cpp:326-372) on the
    cluster pair list: cn[k] = sum over the environment atoms j of s(r_kj)
    (ContactMatrix rows times ONES), with the vector-Jacobian product
    sum_k w_k dcn_k/dx for given row weights (what a function of the vector,
    e.g. MORE_THAN summed, sends back as forces).

    SPECIES=list gives rows = environment = list; SPECIESA/SPECIESB give rows =
    A, environment = B.  Indices are 0-based rows of the positions passed to
    the call.  It runs as a COORDINATION context whose GROUPA is the rows and
    GROUPB the environment (none for SPECIES), so the same list rules apply
    (``nl`` takes e.g. "NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10")."""

    def __init__(self, switch: str, species=None, species_a=None, species_b=None, *, pbc: bool = True,
                 nl: str = "", device: int = 0):
        if species is not None:
            if species_a is not None or species_b is not None:
                raise ValueError("give SPECIES or SPECIESA/SPECIESB, not both")
            rows, env = [int(i) for i in species], None
        elif species_a is not None and species_b is not None:
            rows, env = [int(i) for i in species_a], [int(i) for i in species_b]
        else:
            raise ValueError("missing atoms input use SPECIES or SPECIESA/SPECIESB")
        if not rows or (env is not None and not env):
            raise ValueError("empty species list")
        if min(rows + (env or [])) < 0:
            raise ValueError("atom index outside the positions array")
        line = f"cn: COORDINATION GROUPA={_atoms_arg(rows)} "
        if env is not None:
            line += f"GROUPB={_atoms_arg(env)} "
        line += f"SWITCH={{{switch}}}" + ("" if pbc else " NOPBC") + (f" {nl}" if nl else "")
        self.action = Coordination(line, device=device)
        self.rows = rows
        self.env = env
        self.pbc = pbc
        self._idx = None

    def _req(self, device):
        import torch

        if self._idx is None or self._idx.device != device:
            self._idx = torch.tensor(self.rows + (self.env or []), dtype=torch.long, device=device)
        return self._idx

    def __call__(self, pos, box=None, step: int = 0, weights=None):
        """pos: float64 CUDA tensor [natoms,3]; box 3x3 rows = lattice vectors
        (ignored with pbc=False).  Returns cn [nrows]; with weights [nrows]
        also (cn, vjp [natoms,3], box derivative [3,3]).  Enqueued on torch's
        current stream."""
        import torch

        if not (pos.is_cuda and pos.dtype == torch.float64 and pos.dim() == 2 and pos.shape[1] == 3):
            raise ValueError("positions must be a float64 CUDA tensor [natoms,3]")
        idx = self._req(pos.device)
        if int(idx.max()) >= pos.shape[0]:
            raise ValueError("atom index outside the positions array")
        req = pos.index_select(0, idx).contiguous()
        b9 = _box9(box if box is not None else np.eye(3))
        self.action.set_stream(torch.cuda.current_stream(pos.device).cuda_stream)
        cn = torch.empty(len(self.rows), dtype=torch.float64, device=pos.device)
        if weights is None:
            A.check(A.lib().p2a_coord_rows_device(self.action._ctx, int(step), b9.ctypes.data,
                                                  C.c_void_p(req.data_ptr()), C.c_void_p(cn.data_ptr()), None, None,
                                                  None))
            return cn
        w = torch.as_tensor(weights, dtype=torch.float64, device=pos.device).contiguous()
        if w.shape != (len(self.rows),):
            raise ValueError(f"expected {len(self.rows)} weights")
        vjp_req = torch.empty((len(idx), 3), dtype=torch.float64, device=pos.device)
        vir = torch.empty((3, 3), dtype=torch.float64, device=pos.device)
        A.check(A.lib().p2a_coord_rows_device(self.action._ctx, int(step), b9.ctypes.data,
                                              C.c_void_p(req.data_ptr()), C.c_void_p(cn.data_ptr()),
                                              C.c_void_p(w.data_ptr()), C.c_void_p(vjp_req.data_ptr()),
                                              C.c_void_p(vir.data_ptr())))
        vjp = torch.zeros_like(pos)
        vjp.index_add_(0, idx, vjp_req)
        return cn, vjp, vir

    def close(self):
        self.action.close()
