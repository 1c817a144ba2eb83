"""ctypes declarations of the C ABI in include/plumed2_amd.h.

The shared library is the in-tree ``plumed2_amd/libplumed2amd.so``.  There is
no fallback: importing the bindings without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libplumed2amd.so")

P2A_OK = 0
P2A_ERR_ARG = -1
P2A_ERR_HIP = -2
P2A_ERR_NOMEM = -3
P2A_ERR_UNSUPPORTED = -4
P2A_ERR_GRID = -5

NL_NONE, NL_CLASSIC, NL_CELLS = 0, 1, 2


class p2a_switch(C.Structure):
    _fields_ = [("type", C.c_int32), ("fast", C.c_int32), ("simplified", C.c_int32),
                ("nn", C.c_int32), ("mm", C.c_int32), ("nnf", C.c_int32), ("mmf", C.c_int32),
                ("a", C.c_int32), ("b", C.c_int32), ("pad_", C.c_int32),
                ("d0", C.c_double), ("dmax", C.c_double), ("dmax_2", C.c_double), ("invr0", C.c_double),
                ("invr0_2", C.c_double), ("stretch", C.c_double), ("shift", C.c_double),
                ("preRes", C.c_double), ("preDfunc", C.c_double), ("preSecDev", C.c_double),
                ("preDfuncF", C.c_double), ("preSecDevF", C.c_double),
                ("c", C.c_double), ("d", C.c_double),
                ("beta", C.c_double), ("lambda_", C.c_double), ("ref", C.c_double)]


class p2a_coord_spec(C.Structure):
    _fields_ = [("groupa", C.POINTER(C.c_int64)), ("na", C.c_int64),
                ("groupb", C.POINTER(C.c_int64)), ("nb", C.c_int64),
                ("pbc", C.c_int32), ("pair", C.c_int32), ("serial", C.c_int32), ("nl_mode", C.c_int32),
                ("nl_cutoff", C.c_double), ("nl_stride", C.c_int32), ("pad_", C.c_int32),
                ("sw", p2a_switch)]


class p2a_coord_stats(C.Structure):
    _fields_ = [("natoms", C.c_int64), ("contributing_pairs", C.c_int64), ("clusters_i", C.c_int64),
                ("clusters_j", C.c_int64), ("list_entries", C.c_int64), ("rebuilds", C.c_int64),
                ("path", C.c_int32), ("pad_", C.c_int32), ("pair_kernel_ms", C.c_double),
                ("pair_kernel_launches", C.c_int64), ("total_contributing_pairs", C.c_int64)]


class p2a_batch_stats(C.Structure):
    _fields_ = [("nsys", C.c_int64), ("natoms", C.c_int64), ("clusters", C.c_int64), ("list_entries", C.c_int64),
                ("items", C.c_int64), ("contributing_pairs", C.c_int64), ("chunk", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("graph", C.c_int32), ("regrowths", C.c_int32),
                ("last_step_ms", C.c_double)]


_lib = None

_SIGS = {
    "p2a_version": ([], C.c_char_p),
    "p2a_last_error": ([], C.c_char_p),
    "p2a_device_count": ([], C.c_int),
    "p2a_switch_parse": ([C.c_char_p, C.POINTER(p2a_switch)], C.c_int),
    "p2a_switch_rational": ([C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(p2a_switch)], C.c_int),
    "p2a_switch_describe": ([C.POINTER(p2a_switch), C.c_char_p, C.c_size_t], C.c_int),
    "p2a_coord_parse": ([C.c_char_p, C.POINTER(p2a_coord_spec)], C.c_int),
    "p2a_coord_spec_free": ([C.POINTER(p2a_coord_spec)], None),
    "p2a_coord_create": ([C.POINTER(C.c_void_p), C.c_int, C.POINTER(p2a_coord_spec)], C.c_int),
    "p2a_coord_destroy": ([C.c_void_p], C.c_int),
    "p2a_coord_set_stream": ([C.c_void_p, C.c_void_p], C.c_int),
    "p2a_coord_calculate": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                            C.c_int),
    "p2a_coord_calculate_device": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p], C.c_int),
    "p2a_coord_backward_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "p2a_coord_calculate_noalloc": ([C.c_void_p, C.POINTER(p2a_switch), C.c_void_p, C.c_int, C.c_int64, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "p2a_coord_list_generation": ([C.c_void_p, C.POINTER(C.c_int64)], C.c_int),
    "p2a_coord_rebuild_device": ([C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "p2a_coord_get_stats": ([C.c_void_p, C.POINTER(p2a_coord_stats)], C.c_int),
    "p2a_coord_rows_device": ([C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p], C.c_int),
    "p2a_coordination_numbers_device": ([C.c_void_p, C.POINTER(p2a_switch), C.c_void_p, C.c_int, C.c_int64,
                                         C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p],
                                        C.c_int),
    "p2a_coord_set_profiling": ([C.c_void_p, C.c_int], C.c_int),
    "p2a_coord_set_tuning": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "p2a_coord_set_skin": ([C.c_void_p, C.c_double], C.c_int),
    "p2a_cdist_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                          C.c_void_p, C.c_int], C.c_int),
    "p2a_cdist_destroy": ([C.c_void_p], C.c_int),
    "p2a_cdist_calculate_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p], C.c_int),
    "p2a_cdist_apply_force_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p], C.c_int),
    "p2a_batch_create": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_void_p, C.POINTER(p2a_switch), C.c_double],
                         C.c_int),
    "p2a_batch_destroy": ([C.c_void_p], C.c_int),
    "p2a_batch_set_stream": ([C.c_void_p, C.c_void_p], C.c_int),
    "p2a_batch_set_boxes": ([C.c_void_p, C.c_void_p], C.c_int),
    "p2a_batch_calculate_device": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "p2a_batch_get_pair_counts": ([C.c_void_p, C.c_void_p], C.c_int),
    "p2a_batch_set_graph": ([C.c_void_p, C.c_int], C.c_int),
    "p2a_batch_set_tuning": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "p2a_batch_autotune": ([C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "p2a_batch_get_stats": ([C.c_void_p, C.c_void_p], C.c_int),
}

# symbols declared in the header that later sections bind lazily (METAD, batch)
OPTIONAL_SIGS: dict = {}


class Plumed2AmdError(RuntimeError):
    """Raised for a non-zero status of the C ABI (the PLMD::Exception analogue)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is not built: run __graft_entry__.build() or python plumed2_amd/build_lib.py")
        # torch ships its own HIP runtime under the soname libamdhip64.so.7 but
        # links it by file name; loaded after ours it would bring a second
        # runtime into the process.  Loading torch first lets this library bind
        # to the runtime already mapped, so device pointers are shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIBPATH)
        for name, (args, res) in {**_SIGS, **OPTIONAL_SIGS}.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(rc):
    if rc != P2A_OK:
        msg = lib().p2a_last_error()
        raise Plumed2AmdError(rc, msg.decode() if msg else f"error {rc}")
    return rc
