"""The allocation-free entry point p2a_coord_calculate_noalloc, through the C
ABI, against the existing entry point p2a_coord_calculate_device on the same
device-resident positions: two calls in a row on one device, the second with
fewer atoms (the static scratch keeps the first call's partials).  Bar: CV to
1e-6 relative, derivatives to 1e-5 max|F|, virial to 1e-5 max|W|; both are
also held against the FP64 oracle (the allocation-free path is FP64, 1e-10)."""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _close(v, d, w, ov, od, ow, cv_tol, f_tol):
    assert abs(v - ov) <= cv_tol * abs(ov), (v, ov)
    assert np.max(np.abs(d - od)) <= f_tol * np.max(np.abs(od))
    assert np.max(np.abs(w - ow)) <= f_tol * np.max(np.abs(ow))


def _noalloc(A, torch, sw, box, xt, na, nb, absidx=None):
    dev = xt.device
    val = torch.empty(1, dtype=torch.float64, device=dev)
    der = torch.empty_like(xt)
    vir = torch.empty((3, 3), dtype=torch.float64, device=dev)
    b9 = np.ascontiguousarray(box, dtype=np.float64).reshape(-1)
    ab = None if absidx is None else torch.tensor(absidx, dtype=torch.int64, device=dev)
    A.check(A.lib().p2a_coord_calculate_noalloc(
        C.c_void_p(torch.cuda.current_stream(dev).cuda_stream), C.byref(sw), b9.ctypes.data, 1, na, nb,
        None if ab is None else C.c_void_p(ab.data_ptr()), C.c_void_p(xt.data_ptr()), C.c_void_p(val.data_ptr()),
        C.c_void_p(der.data_ptr()), C.c_void_p(vir.data_ptr())))
    torch.cuda.synchronize()
    return float(val.item()), der.cpu().numpy(), vir.cpu().numpy()


def _existing(P, torch, line, box, xt):
    act = P.Coordination(line)
    val, der, vir = act.calculate_torch(xt, box, step=0)
    torch.cuda.synchronize()
    out = float(val.item()), der.cpu().numpy(), vir.cpu().numpy()
    act.close()
    return out


def test_noalloc_matches_existing_entry_point_twice():
    import torch

    import plumed2_amd as P
    from plumed2_amd import _capi as A

    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0"
    sw = P.parse_switch(defn)
    osw = O.Switch(defn)
    # call 1: GROUPA=1000 x GROUPB=1000 in a 5 nm cube
    x1 = np.random.default_rng(21).uniform(0.0, 5.0, (2000, 3))
    box1 = np.diag([5.0, 5.0, 5.0])
    t1 = torch.tensor(x1, dtype=torch.float64, device="cuda")
    got1 = _noalloc(A, torch, sw, box1, t1, 1000, 1000)
    ref1 = _existing(P, torch, f"c: COORDINATION GROUPA=1-1000 GROUPB=1001-2000 SWITCH={{{defn}}}", box1, t1)
    # call 2: fewer atoms, one group, a triclinic cell
    x2 = np.random.default_rng(22).uniform(0.0, 4.0, (700, 3))
    box2 = np.array([[4.4, 0.0, 0.0], [0.3, 4.2, 0.0], [-0.2, 0.4, 4.3]])
    t2 = torch.tensor(x2, dtype=torch.float64, device="cuda")
    got2 = _noalloc(A, torch, sw, box2, t2, 700, 0)
    ref2 = _existing(P, torch, f"c: COORDINATION GROUPA=1-700 SWITCH={{{defn}}}", box2, t2)
    for got, ref, x, box, na, nb in ((got1, ref1, x1, box1, 1000, 1000), (got2, ref2, x2, box2, 700, 0)):
        _close(*got, *ref, 1e-6, 1e-5)
        ov, od, ow, _ = O.coordination(osw, box, x, na, nb)
        _close(*got, ov, od, ow, 1e-10, 1e-10)


def test_noalloc_overlapping_groups_and_size_limit():
    """Absolute indices skip the pairs of one atom with itself; more atoms
    than the scratch covers is an error raised before any launch."""
    import torch

    import plumed2_amd as P
    from plumed2_amd import _capi as A

    defn = "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0"
    sw = P.parse_switch(defn)
    rng = np.random.default_rng(23)
    xa = rng.uniform(0.0, 4.0, (300, 3))
    xb = np.concatenate([xa[200:], rng.uniform(0.0, 4.0, (200, 3))])  # atoms 201-300 in both groups
    x = np.concatenate([xa, xb])
    absidx = np.concatenate([np.arange(300), np.arange(200, 500)])
    box = np.diag([4.0, 4.0, 4.0])
    t = torch.tensor(x, dtype=torch.float64, device="cuda")
    got = _noalloc(A, torch, sw, box, t, 300, 300, absidx)
    ref = _existing(P, torch, f"c: COORDINATION GROUPA=1-300 GROUPB=201-500 SWITCH={{{defn}}}", box, t)
    _close(*got, *ref, 1e-6, 1e-5)
    ov, od, ow, _ = O.coordination(O.Switch(defn), box, x, 300, 300, absidx=absidx)
    _close(*got, ov, od, ow, 1e-10, 1e-10)
    with pytest.raises(P.Plumed2AmdError, match="at most 262144 atoms"):
        _noalloc(A, torch, sw, box, t, 200000, 100000)
