"""The C-ABI library loads, exports every function include/*.h declares, and
its host-only entry points (keyword parsing) behave like the reference. No GPU
calls are made here."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import oracle as O
import plumed2_amd as P
from plumed2_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(p2a_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(_capi.LIBPATH)
    names = declared_functions()
    assert len(names) > 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_version_and_no_device_here():
    assert b"gfx950" in P.lib().p2a_version()
    assert P.lib().p2a_device_count() >= 0


@pytest.mark.parametrize("definition", [
    "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.0", "RATIONAL R_0=1.3 NN=5 MM=11 D_MAX=2.6 NOSTRETCH",
    "RATIONAL R_0=1 NN=7 MM=14 D_0=0.1 D_MAX=3", "EXP R_0=0.8 D_0=0.5 D_MAX=2.6", "GAUSSIAN R_0=1.0 D_MAX=2.6",
    "GAUSSIAN R_0=1.0 D_0=0.3 D_MAX=2.6", "SMAP R_0=1.3 A=3 B=2 D_MAX=2.6", "CUBIC D_0=0 D_MAX=2.6",
    "TANH R_0=1.3 D_MAX=2.6", "COSINUS R_0=2.6", "Q R_0=1.0 D_0=0.3 BETA=1.0 LAMBDA=1.0 REF=1.3 D_MAX=2.6",
    "RATIONAL R_0=1"])
def test_switch_parse_matches_oracle(definition):
    s = P.parse_switch(definition)
    o = O.Switch(definition)
    assert s.type == o.type
    assert s.dmax == o.dmax and s.stretch == o.stretch and s.shift == o.shift


def test_switch_nnmm_matches_oracle():
    for nn, mm, r0, d0 in [(6, 0, 1.0, 0.0), (7, 0, 1.0, 0.1), (8, 16, 0.35, 0.0), (5, 11, 1.3, 0.0)]:
        s = P.rational_switch(nn, mm, r0, d0)
        o = O.Switch(nn=nn, mm=mm, r0=r0, d0=d0)
        assert (s.dmax, s.stretch, s.shift) == (o.dmax, o.stretch, o.shift)


@pytest.mark.parametrize("bad,msg", [
    ("RATIONAL NN=6", "R_0 is required for RATIONAL"),
    ("RATIONAL R_0=1 FOO=3", "found the following rogue keywords in switching function input : FOO=3"),
    ("BLAH R_0=1", "cannot understand switching function type 'BLAH'"),
    ("RATIONAL R_0=1 D_MAX=x", "could not parse D_MAX"),
    ("SMAP R_0=1 A=2", "B is required for SMAP"),
    ("", "missing all input for switching function"),
])
def test_switch_parse_errors(bad, msg):
    with pytest.raises(P.Plumed2AmdError) as e:
        P.parse_switch(bad)
    assert msg in str(e.value)


def test_custom_switch_is_reported_unsupported():
    with pytest.raises(P.Plumed2AmdError) as e:
        P.parse_switch("CUSTOM FUNC=1/(1+x^6) R_0=1")
    assert e.value.code == _capi.P2A_ERR_UNSUPPORTED


def test_action_line_parse():
    sp = P.CoordinationSpec("c: COORDINATION GROUPA=1-10,15,20-30:5 GROUPB=100-102 "
                            "SWITCH={RATIONAL R_0=0.3 D_MAX=1.0} NLISTCELLS NL_CUTOFF=1.5 NL_STRIDE=10")
    np.testing.assert_array_equal(sp.groupa, [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 14, 19, 24, 29])
    np.testing.assert_array_equal(sp.groupb, [99, 100, 101])
    assert (sp.nl_mode, sp.nl_cutoff, sp.nl_stride, sp.pbc, sp.pair) == (2, 1.5, 10, True, False)
    sp = P.CoordinationSpec("COORDINATION GROUPA=1-4 GROUPB=5-8 PAIR NOPBC NN=8 MM=16 R_0=0.35")
    assert sp.pair and not sp.pbc and sp.switch.nn == 8 and sp.switch.type == 0


@pytest.mark.parametrize("bad,msg", [
    ("COORDINATION GROUPA=1-10 NN=6", "R_0 should be explicitly specified and positive"),
    ("COORDINATION GROUPA=1-10 R_0=1 NLIST", "NL_CUTOFF should be explicitly specified and positive"),
    ("COORDINATION GROUPA=1-10 R_0=1 NLIST NL_CUTOFF=1", "NL_STRIDE should be explicitly specified and positive"),
    ("COORDINATION GROUPA=1-10 R_0=1 NLIST NLISTCELLS NL_CUTOFF=1 NL_STRIDE=2",
     "Please activate only one of the two version of the NL"),
    ("COORDINATION GROUPA=1-10 GROUPB=11-20 R_0=1 PAIR NLISTCELLS NL_CUTOFF=1 NL_STRIDE=2",
     "Pair is not compatible with the CELLS implementation of the NL"),
    ("COORDINATION GROUPA=1-10 R_0=1 FOO=2", "cannot understand the following words from the input line : FOO=2"),
    ("COORDINATION GROUPA=1-10 SWITCH={RATIONAL NN=6}", "problem reading SWITCH keyword : R_0 is required for RATIONAL"),
])
def test_action_line_errors(bad, msg):
    with pytest.raises(P.Plumed2AmdError) as e:
        P.CoordinationSpec(bad)
    assert msg in str(e.value)


def test_context_creation_fails_loudly_without_gpu():
    if P.lib().p2a_device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(P.Plumed2AmdError, match="no HIP device"):
        P.Coordination("c: COORDINATION GROUPA=1-10 R_0=1")


def test_batch_validates_arguments_and_needs_a_device():
    import ctypes as C

    import numpy as np

    import plumed2_amd as P
    from plumed2_amd import _capi as A
    sw = P.parse_switch("RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2")
    b = C.c_void_p()
    nat = np.array([3000, 0], dtype=np.int64)
    assert A.lib().p2a_batch_create(C.byref(b), 0, 2, nat.ctypes.data, C.byref(sw), 1.2) == A.P2A_ERR_ARG
    assert b"system 1 has no atoms" in A.lib().p2a_last_error()
    nat = np.array([3000, 4000], dtype=np.int64)
    assert A.lib().p2a_batch_create(C.byref(b), 0, 2, nat.ctypes.data, C.byref(sw), 1.0) == A.P2A_ERR_UNSUPPORTED
    assert b"smaller than D_MAX" in A.lib().p2a_last_error()
    if P.lib().p2a_device_count() > 0:
        return
    with pytest.raises(P.Plumed2AmdError, match="no HIP device"):
        P.CoordinationBatch([3000, 4000], "RATIONAL R_0=0.3 NN=6 MM=12 D_MAX=1.2", 1.2)


def test_coordination_numbers_argument_checks_without_device():
    """The COORDINATIONNUMBER entry point settles empty and invalid calls on the
    host, before any HIP call: no rows is a no-op even with null buffers."""
    import ctypes as C

    import plumed2_amd as P
    from plumed2_amd import _capi as A

    sw = P.parse_switch("RATIONAL R_0=0.3")
    f = A.lib().p2a_coordination_numbers_device
    assert f(None, C.byref(sw), None, 0, 10, 0, None, 5, None, None, None) == 0
    assert f(None, None, None, 0, 10, 1, None, 5, None, None, None) != 0
    assert f(None, C.byref(sw), None, 0, -1, 1, None, 5, None, None, None) != 0
    assert f(None, C.byref(sw), None, 0, 10, 3, None, 5, None, None, None) != 0
