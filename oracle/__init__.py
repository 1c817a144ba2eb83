"""ctypes bindings of the FP64 CPU oracle (liboracle.so) and of the compiled
reference subset (_ref/libplumedref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke`` as a
checker and by the CPU-baseline leg of ``bench.py``.  Nothing in the product
package imports this module.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import shlex

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

# switch type codes of coordination_oracle.c
RATIONAL_FIXED, RATIONAL, EXP, GAUSSIAN, FASTGAUSSIAN, SMAP, CUBIC, TANH, COSINUS, Q = range(10)
_NAMES = {"RATIONAL": RATIONAL, "EXP": EXP, "GAUSSIAN": GAUSSIAN, "SMAP": SMAP, "CUBIC": CUBIC,
          "TANH": TANH, "COSINUS": COSINUS, "Q": Q}

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        L.osw_init.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                               C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        L.osw_init_nnmm.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double]
        L.osw_sizeof.restype = C.c_size_t
        L.osw_calculate_sqr.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.osw_calculate_sqr.restype = C.c_double
        L.osw_calculate.argtypes = [C.c_void_p, C.c_double, C.POINTER(C.c_double)]
        L.osw_calculate.restype = C.c_double
        L.osw_get.argtypes = [C.c_void_p, C.c_int]
        L.osw_get.restype = C.c_double
        L.opbc_sizeof.restype = C.c_size_t
        L.opbc_set_box.argtypes = [C.c_void_p, _dp]
        L.opbc_distance.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.opbc_get_reduced.argtypes = [C.c_void_p, _dp]
        L.opbc_type.argtypes = [C.c_void_p]
        L.ocells_count.argtypes = [C.c_void_p, C.c_double, _i32p]
        L.onl_cells.argtypes = [C.c_void_p, C.c_double, _dp, C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.onl_cells.restype = C.c_int64
        L.onl_classic.argtypes = [C.c_void_p, C.c_int, C.c_double, _dp, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int64]
        L.onl_classic.restype = C.c_int64
        L.ocoord_calculate.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _dp, _i64p, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p, C.c_int64, C.POINTER(C.c_double), _dp, _dp]
        L.ocoord_calculate.restype = C.c_int64
        L.ocn_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _dp, C.c_int, C.c_int, C.c_void_p, C.c_int,
                               C.c_void_p, C.c_void_p, _dp, _dp, _dp]
        L.ocn_rows.restype = None
        L.ocenter_distance.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                       C.c_void_p, _dp, _dp, _dp, _dp]
        L.ocenter_distance.restype = C.c_double
        L.ogrid_sizeof.restype = C.c_size_t
        L.ogrid_init.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _u32p, _i32p]
        L.ogrid_free.argtypes = [C.c_void_p]
        L.ogrid_size.argtypes = [C.c_void_p]
        L.ogrid_size.restype = C.c_size_t
        L.ogrid_values.argtypes = [C.c_void_p]
        L.ogrid_values.restype = C.POINTER(C.c_double)
        L.ogrid_derivs.argtypes = [C.c_void_p]
        L.ogrid_derivs.restype = C.POINTER(C.c_double)
        L.ogrid_add_gaussian.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_double]
        L.ogrid_add_gaussian.restype = C.c_int64
        L.ogrid_spline.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double), _dp]
        L.ogrid_deposit_wt.argtypes = [C.c_void_p, C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_void_p]
        L.ometad_bias_hills.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.ometad_bias_hills.restype = C.c_double
        _LIB = L
    return _LIB


# --------------------------------------------------------------------------
# switching functions
# --------------------------------------------------------------------------

def _split_words(definition: str):
    # Tools::getWords splits on blanks, keeping {...} groups; SWITCH bodies have none
    return shlex.split(definition.replace("{", " ").replace("}", " "))


class Switch:
    """FP64 switching function, from a SWITCH={...} body or from NN/MM/R_0/D_0."""

    def __init__(self, definition: str | None = None, *, nn=6, mm=0, r0=None, d0=0.0):
        self._buf = C.create_string_buffer(lib().osw_sizeof())
        if definition is None:
            lib().osw_init_nnmm(self._buf, int(nn), int(mm), float(r0), float(d0))
            return
        words = _split_words(definition)
        name = words[0].upper()
        kv = {}
        flags = set()
        for w in words[1:]:
            if "=" in w:
                k, v = w.split("=", 1)
                kv[k.upper()] = v
            else:
                flags.add(w.upper())
        d0 = float(kv.get("D_0", 0.0))
        dmax = float(kv["D_MAX"]) if "D_MAX" in kv else -1.0
        nostretch = 1 if "NOSTRETCH" in flags else 0
        r0 = float(kv.get("R_0", 1.0))
        if name not in _NAMES:
            raise ValueError(f"switch type {name} not in the oracle")
        t = _NAMES[name]
        if t == RATIONAL:
            nn = int(kv.get("NN", 6))
            mm = int(kv.get("MM", 0))
        a = int(kv.get("A", 0))
        b = int(kv.get("B", 0))
        beta = float(kv.get("BETA", 50.0))
        lam = float(kv.get("LAMBDA", 1.8))
        ref = float(kv.get("REF", 0.0))
        rc = lib().osw_init(self._buf, t, d0, dmax, r0, int(nn), int(mm), a, b, beta, lam, ref, nostretch)
        if rc != 0:
            raise ValueError("bad switch")

    @property
    def ptr(self):
        return self._buf

    def sqr(self, d2):
        d2 = np.atleast_1d(np.asarray(d2, dtype=np.float64))
        v = np.empty_like(d2)
        df = np.empty_like(d2)
        tmp = C.c_double()
        L = lib()
        for i, x in enumerate(d2):
            v[i] = L.osw_calculate_sqr(self._buf, float(x), C.byref(tmp))
            df[i] = tmp.value
        return v, df

    def dist(self, d):
        d = np.atleast_1d(np.asarray(d, dtype=np.float64))
        v = np.empty_like(d)
        df = np.empty_like(d)
        tmp = C.c_double()
        L = lib()
        for i, x in enumerate(d):
            v[i] = L.osw_calculate(self._buf, float(x), C.byref(tmp))
            df[i] = tmp.value
        return v, df

    @property
    def dmax(self):
        return lib().osw_get(self._buf, 0)

    @property
    def stretch(self):
        return lib().osw_get(self._buf, 1)

    @property
    def shift(self):
        return lib().osw_get(self._buf, 2)

    @property
    def type(self):
        return int(lib().osw_get(self._buf, 5))


# --------------------------------------------------------------------------
# periodic boundaries
# --------------------------------------------------------------------------

class Pbc:
    def __init__(self, box):
        self._buf = C.create_string_buffer(lib().opbc_sizeof())
        b = np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9))
        lib().opbc_set_box(self._buf, b)

    @property
    def ptr(self):
        return self._buf

    @property
    def type(self):
        return lib().opbc_type(self._buf)

    def distance(self, v1, v2):
        v1 = np.ascontiguousarray(np.asarray(v1, dtype=np.float64).reshape(-1, 3))
        v2 = np.ascontiguousarray(np.asarray(v2, dtype=np.float64).reshape(-1, 3))
        out = np.empty_like(v1)
        o = np.empty(3)
        L = lib()
        for i in range(len(v1)):
            L.opbc_distance(self._buf, v1[i].copy(), v2[i].copy(), o)
            out[i] = o
        return out

    def reduced(self):
        out = np.empty(9)
        lib().opbc_get_reduced(self._buf, out)
        return out.reshape(3, 3)

    def cells(self, cutoff):
        nc = np.zeros(3, dtype=np.int32)
        lib().ocells_count(self._buf, float(cutoff), nc)
        return nc


def nl_cells(pbc: Pbc, cutoff, pos, nA, nB):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    n = lib().onl_cells(pbc.ptr, float(cutoff), pos, int(nA), int(nB), None, 0)
    pairs = np.empty((max(n, 1), 2), dtype=np.int32)
    lib().onl_cells(pbc.ptr, float(cutoff), pos, int(nA), int(nB), pairs.ctypes.data, n)
    return pairs[:n]


def nl_classic(pbc: Pbc, use_pbc, cutoff, pos, nA, nB, dopair=False):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    n = lib().onl_classic(pbc.ptr, int(use_pbc), float(cutoff), pos, int(nA), int(nB), int(dopair), None, 0)
    pairs = np.empty((max(n, 1), 2), dtype=np.int32)
    lib().onl_classic(pbc.ptr, int(use_pbc), float(cutoff), pos, int(nA), int(nB), int(dopair), pairs.ctypes.data, n)
    return pairs[:n]


def coordination(sw: Switch, box, pos, nA, nB=0, *, absidx=None, use_pbc=True, dopair=False, pairs=None):
    """CoordinationBase::calculate restated in FP64.  Returns
    (value, deriv[nA+nB,3], virial[3,3], contributing_pairs)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    nat = nA + nB
    assert pos.shape[0] == nat
    if absidx is None:
        absidx = np.arange(nat, dtype=np.int64)
    absidx = np.ascontiguousarray(absidx, dtype=np.int64)
    pbc = Pbc(box) if not isinstance(box, Pbc) else box
    deriv = np.zeros(3 * nat)
    vir = np.zeros(9)
    val = C.c_double()
    if pairs is not None:
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        pp, npairs = pairs.ctypes.data, len(pairs)
    else:
        pp, npairs = None, 0
    nc = lib().ocoord_calculate(sw.ptr, pbc.ptr, int(use_pbc), pos.reshape(-1), absidx, int(nA), int(nB),
                                int(dopair), pp, npairs, C.byref(val), deriv, vir)
    return val.value, deriv.reshape(nat, 3), vir.reshape(3, 3), int(nc)


def coordination_numbers(sw: Switch, box, pos, species=None, species_a=None, species_b=None, *, use_pbc=True,
                         weights=None):
    """COORDINATIONNUMBER restated in FP64: the vector cn_k = sum_j w_kj.

    The reference builds it as a shortcut (symfunc/CoordinationNumbers.cpp:326-372):
    CONTACT_MATRIX rows times a vector of ONES.  SPECIES=list gives one row per
    listed atom over the listed atoms; SPECIESA/SPECIESB give one row per A atom
    over the B atoms.  The matrix element is ContactMatrix::calculateWeight
    (adjmat/ContactMatrix.h:95-121): nothing for |r|^2 < DBL_EPSILON (the
    diagonal, or two atoms at one place), s(r) zeroed below DBL_EPSILON.
    Indices are 0-based rows of `pos`.  Returns cn[len(rows)]; with `weights`
    (one per row) also the vector-Jacobian product sum_k weights[k] d cn_k/dx
    [natoms,3] and the matching box derivative [3,3]."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    if species is not None:
        rows = env = [int(i) for i in species]
    else:
        rows, env = [int(i) for i in species_a], [int(j) for j in species_b]
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    env = np.ascontiguousarray(env, dtype=np.int32)
    pbc = Pbc(box) if not isinstance(box, Pbc) else box
    n = pos.shape[0]
    cn = np.zeros(len(rows))
    der = np.zeros(3 * n)
    vir = np.zeros(9)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    lib().ocn_rows(sw.ptr, pbc.ptr, int(use_pbc), pos.reshape(-1), n, len(rows), rows.ctypes.data, len(env),
                   env.ctypes.data, None if w is None else w.ctypes.data, cn, der, vir)
    if weights is None:
        return cn
    return cn, der.reshape(n, 3), vir.reshape(3, 3)


def center_distance(box, pos, group1, group2, masses=None, *, use_pbc=True):
    """DISTANCE between the centres of mass of two groups (c1: COM ATOMS=...,
    c2: COM ATOMS=..., d: DISTANCE ATOMS=c1,c2), restated in FP64 with the
    reference's order of operations.  Returns (value, deriv[n1+n2,3] for the
    listed atoms, group 1 then group 2, virial[3,3], c1, c2)."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    g1 = np.ascontiguousarray(group1, dtype=np.int32)
    g2 = np.ascontiguousarray(group2, dtype=np.int32)
    pbc = Pbc(box) if not isinstance(box, Pbc) else box
    m = None if masses is None else np.ascontiguousarray(masses, dtype=np.float64)
    der = np.zeros(3 * (len(g1) + len(g2)))
    vir = np.zeros(9)
    c1, c2 = np.zeros(3), np.zeros(3)
    v = lib().ocenter_distance(pbc.ptr, int(use_pbc), pos.reshape(-1), len(g1), g1.ctypes.data, len(g2),
                               g2.ctypes.data, None if m is None else m.ctypes.data, der, vir, c1, c2)
    return v, der.reshape(-1, 3), vir.reshape(3, 3), c1, c2


# --------------------------------------------------------------------------
# METAD grid
# --------------------------------------------------------------------------

DP2CUTOFF_A = 1.00193418799744762399
DP2CUTOFF_B = -.00193418799744762399
KBOLTZMANN = 0.0083144621


class Grid:
    def __init__(self, gmin, gmax, nbin, pbc=None):
        dim = len(gmin)
        self.dim = dim
        self._buf = C.create_string_buffer(lib().ogrid_sizeof())
        pbc = [0] * dim if pbc is None else [int(bool(p)) for p in pbc]
        rc = lib().ogrid_init(self._buf, dim, np.asarray(gmin, dtype=np.float64), np.asarray(gmax, dtype=np.float64),
                              np.asarray(nbin, dtype=np.uint32), np.asarray(pbc, dtype=np.int32))
        if rc:
            raise ValueError("bad grid")
        self.size = lib().ogrid_size(self._buf)

    def __del__(self):
        try:
            lib().ogrid_free(self._buf)
        except Exception:
            pass

    def values(self):
        p = lib().ogrid_values(self._buf)
        return np.ctypeslib.as_array(p, shape=(self.size,)).copy()

    def derivs(self):
        p = lib().ogrid_derivs(self._buf)
        return np.ctypeslib.as_array(p, shape=(self.size, self.dim)).copy()

    def set(self, values, derivs):
        pv = np.ctypeslib.as_array(lib().ogrid_values(self._buf), shape=(self.size,))
        pd = np.ctypeslib.as_array(lib().ogrid_derivs(self._buf), shape=(self.size, self.dim))
        pv[:] = values
        pd[:] = derivs

    def add_gaussian(self, center, sigma, height, stretchA=DP2CUTOFF_A, stretchB=DP2CUTOFF_B):
        return lib().ogrid_add_gaussian(self._buf, np.asarray(center, dtype=np.float64),
                                        np.asarray(sigma, dtype=np.float64), float(height), stretchA, stretchB)

    def spline(self, x):
        v = C.c_double()
        der = np.zeros(self.dim)
        rc = lib().ogrid_spline(self._buf, np.asarray(x, dtype=np.float64), C.byref(v), der)
        if rc:
            raise ValueError("outside grid")
        return v.value, der

    def deposit_wt(self, centers, sigma, height0, biasf, kbt, stretchA=DP2CUTOFF_A, stretchB=DP2CUTOFF_B):
        centers = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, self.dim)
        hts = np.zeros(len(centers))
        rc = lib().ogrid_deposit_wt(self._buf, len(centers), centers.reshape(-1), np.asarray(sigma, dtype=np.float64),
                                    float(height0), float(biasf), float(kbt), stretchA, stretchB, hts.ctypes.data)
        if rc:
            raise ValueError(f"hill {-1 - rc} outside grid")
        return hts


def bias_hills(centers, sigma, heights, cv, stretchA=DP2CUTOFF_A, stretchB=DP2CUTOFF_B):
    centers = np.ascontiguousarray(centers, dtype=np.float64)
    ncv = centers.shape[1]
    der = np.zeros(ncv)
    b = lib().ometad_bias_hills(ncv, len(centers), centers.reshape(-1), np.asarray(sigma, dtype=np.float64),
                                np.ascontiguousarray(heights, dtype=np.float64), np.asarray(cv, dtype=np.float64),
                                stretchA, stretchB, der)
    return b, der


# --------------------------------------------------------------------------
# compiled reference subset (present only where the reference tree was)
# --------------------------------------------------------------------------

def ref():
    """The reference's own tools compiled by oracle/Makefile, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libplumedref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_last_error.restype = C.c_char_p
        R.ref_switch_create.argtypes = [C.c_char_p]
        R.ref_switch_create.restype = C.c_void_p
        R.ref_switch_create_nnmm.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        R.ref_switch_create_nnmm.restype = C.c_void_p
        R.ref_switch_destroy.argtypes = [C.c_void_p]
        R.ref_switch_eval_sqr.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        R.ref_switch_eval.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        R.ref_switch_dmax.argtypes = [C.c_void_p]
        R.ref_switch_dmax.restype = C.c_double
        R.ref_pbc_distance.argtypes = [_dp, C.c_int, _dp, _dp, _dp]
        R.ref_coordination.argtypes = [C.c_void_p, _dp, C.c_int, _dp, _i64p, C.c_int, C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.POINTER(C.c_double), _dp, _dp]
        R.ref_coordination.restype = C.c_int64
        R.ref_grid_create.argtypes = [C.c_int, _dp, _dp, _u32p, _i32p]
        R.ref_grid_create.restype = C.c_void_p
        R.ref_grid_destroy.argtypes = [C.c_void_p]
        R.ref_grid_size.argtypes = [C.c_void_p]
        R.ref_grid_size.restype = C.c_uint64
        R.ref_grid_set.argtypes = [C.c_void_p, C.c_uint64, C.c_double, _dp]
        R.ref_grid_spline.argtypes = [C.c_void_p, _dp, C.POINTER(C.c_double), _dp]
        R.ref_grid_neighbors.argtypes = [C.c_void_p, _dp, _u32p, _u64p, C.c_int64]
        R.ref_grid_neighbors.restype = C.c_int64
        R.ref_grid_point.argtypes = [C.c_void_p, C.c_uint64, _dp]
        R.ref_coordctx_create.argtypes = [C.c_void_p, _dp, _i64p, C.c_int, C.c_int, C.c_int, C.c_double]
        R.ref_coordctx_create.restype = C.c_void_p
        R.ref_coordctx_destroy.argtypes = [C.c_void_p]
        R.ref_coordctx_update.argtypes = [C.c_void_p, _dp]
        R.ref_coordctx_update.restype = C.c_int64
        R.ref_coordctx_reduce.argtypes = [C.c_void_p]
        R.ref_coordctx_reduce.restype = C.c_int64
        R.ref_coordctx_calc.argtypes = [C.c_void_p, _dp, C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        R.ref_coordctx_calc.restype = C.c_int64
        _REF = R
    return _REF


class RefSwitch:
    def __init__(self, definition=None, *, nn=6, mm=0, r0=None, d0=0.0):
        R = ref()
        if definition is None:
            self.h = R.ref_switch_create_nnmm(int(nn), int(mm), float(r0), float(d0))
        else:
            self.h = R.ref_switch_create(definition.encode())
            if not self.h:
                raise ValueError(R.ref_last_error().decode())

    def __del__(self):
        try:
            ref().ref_switch_destroy(self.h)
        except Exception:
            pass

    def sqr(self, d2):
        d2 = np.ascontiguousarray(np.atleast_1d(d2), dtype=np.float64)
        v, df = np.empty_like(d2), np.empty_like(d2)
        ref().ref_switch_eval_sqr(self.h, len(d2), d2, v, df)
        return v, df

    def dist(self, d):
        d = np.ascontiguousarray(np.atleast_1d(d), dtype=np.float64)
        v, df = np.empty_like(d), np.empty_like(d)
        ref().ref_switch_eval(self.h, len(d), d, v, df)
        return v, df

    @property
    def dmax(self):
        return ref().ref_switch_dmax(self.h)


def ref_pbc_distance(box, v1, v2):
    v1 = np.ascontiguousarray(v1, dtype=np.float64).reshape(-1, 3)
    v2 = np.ascontiguousarray(v2, dtype=np.float64).reshape(-1, 3)
    out = np.empty_like(v1)
    ref().ref_pbc_distance(np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9)), len(v1),
                           v1.reshape(-1), v2.reshape(-1), out.reshape(-1))
    return out


def ref_coordination(sw: RefSwitch, box, pos, nA, nB=0, *, absidx=None, use_pbc=True, dopair=False,
                     nlmode=0, nl_cut=1.0e30):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    nat = nA + nB
    if absidx is None:
        absidx = np.arange(nat, dtype=np.int64)
    absidx = np.ascontiguousarray(absidx, dtype=np.int64)
    deriv = np.zeros(3 * nat)
    vir = np.zeros(9)
    val = C.c_double()
    n = ref().ref_coordination(sw.h, np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9)), int(use_pbc),
                               pos.reshape(-1), absidx, int(nA), int(nB), int(dopair), int(nlmode), float(nl_cut),
                               C.byref(val), deriv, vir)
    if n < 0:
        raise RuntimeError(ref().ref_last_error().decode())
    return val.value, deriv.reshape(nat, 3), vir.reshape(3, 3), int(n)


class RefCoordination:
    """Reference NeighborList (cells/classic) + the CoordinationBase pair loop
    with OpenMP, for timing the CPU baseline."""

    def __init__(self, sw: RefSwitch, box, nA, nB=0, *, nlmode=2, nl_cut=1.5, absidx=None):
        nat = nA + nB
        absidx = np.arange(nat, dtype=np.int64) if absidx is None else np.ascontiguousarray(absidx, dtype=np.int64)
        self.sw = sw
        self.nat = nat
        self.h = ref().ref_coordctx_create(sw.h, np.ascontiguousarray(np.asarray(box, dtype=np.float64).reshape(9)),
                                           absidx, int(nA), int(nB), int(nlmode), float(nl_cut))
        if not self.h:
            raise RuntimeError(ref().ref_last_error().decode())

    def __del__(self):
        try:
            ref().ref_coordctx_destroy(self.h)
        except Exception:
            pass

    def update(self, pos):
        return ref().ref_coordctx_update(self.h, np.ascontiguousarray(pos, dtype=np.float64).reshape(-1))

    def reduce(self):
        """NeighborList::getReducedAtomList after a rebuild; returns the
        number of requested atoms."""
        return ref().ref_coordctx_reduce(self.h)

    def calc(self, pos, nthreads=1, want_deriv=False):
        val = C.c_double()
        der = np.zeros(3 * self.nat) if want_deriv else None
        vir = np.zeros(9) if want_deriv else None
        n = ref().ref_coordctx_calc(self.h, np.ascontiguousarray(pos, dtype=np.float64).reshape(-1), int(nthreads),
                                    C.byref(val), None if der is None else der.ctypes.data,
                                    None if vir is None else vir.ctypes.data)
        return val.value, der, vir, int(n)


class RefGrid:
    def __init__(self, gmin, gmax, nbin, pbc=None):
        dim = len(gmin)
        self.dim = dim
        pbc = [0] * dim if pbc is None else [int(bool(p)) for p in pbc]
        self.h = ref().ref_grid_create(dim, np.asarray(gmin, dtype=np.float64), np.asarray(gmax, dtype=np.float64),
                                       np.asarray(nbin, dtype=np.uint32), np.asarray(pbc, dtype=np.int32))
        self.size = ref().ref_grid_size(self.h)

    def __del__(self):
        try:
            ref().ref_grid_destroy(self.h)
        except Exception:
            pass

    def set(self, values, derivs):
        R = ref()
        for i in range(self.size):
            R.ref_grid_set(self.h, i, float(values[i]), np.ascontiguousarray(derivs[i], dtype=np.float64))

    def spline(self, x):
        v = C.c_double()
        der = np.zeros(self.dim)
        if ref().ref_grid_spline(self.h, np.asarray(x, dtype=np.float64), C.byref(v), der):
            raise ValueError(ref().ref_last_error().decode())
        return v.value, der

    def neighbors(self, x, nneigh):
        cap = 1
        for n in nneigh:
            cap *= 2 * n + 1
        out = np.zeros(cap, dtype=np.uint64)
        n = ref().ref_grid_neighbors(self.h, np.asarray(x, dtype=np.float64), np.asarray(nneigh, dtype=np.uint32),
                                     out, cap)
        if n < 0:
            raise ValueError(ref().ref_last_error().decode())
        return out[:n]

    def point(self, index):
        x = np.zeros(self.dim)
        ref().ref_grid_point(self.h, int(index), x)
        return x


def support_bins(sigma, dx):
    """getGaussianSupport (MetaD.cpp:1939-1994), diagonal hills."""
    return [int(math.ceil(math.sqrt(2.0 * 6.25) * s / d)) for s, d in zip(sigma, dx)]
