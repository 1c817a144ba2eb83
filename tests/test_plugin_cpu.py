"""The PLUMED plugin actions (COORDINATION_HIP, METAD_HIP) compile against the
reference's own headers (core/ActionRegister.h, core/Colvar.h, bias/Bias.h,
tools/OFile.h).  Skipped where the reference source tree is absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/src"
PLUGINS = ["CoordinationHIP.cpp", "MetaDHIP.cpp"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference source tree not present")
@pytest.mark.parametrize("src", PLUGINS)
def test_plugin_compiles_against_reference_headers(src):
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + REF, "-I" + os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "plumed2_amd", "plugin", src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
