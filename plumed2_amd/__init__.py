"""plumed2_amd: MI355X-native COORDINATION collective variable and METAD grid.

The arithmetic lives in the HIP library ``libplumed2amd.so`` (C ABI in
``include/plumed2_amd.h``); this package is the host-side mirror of the
reference's action interface.
"""
from ._capi import Plumed2AmdError, lib  # noqa: F401
from .coordination import (Coordination, CoordinationNumbers, CoordinationSpec,  # noqa: F401
                           coordination_numbers_torch, describe_switch, parse_switch, rational_switch)
from .batch import CoordinationBatch  # noqa: F401
from .metad import MetaD  # noqa: F401
from .chain import CenterDistance, MetadStep  # noqa: F401

__version__ = "0.1.0"
