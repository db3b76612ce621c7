"""bamrex_amd -- MI355X-native implementation of BAMReX's explicit hot path
(MUSCL-Hancock + HLLC sweeps, conservative update, ghost fill, CFL dt) behind
the reference's level-integrator interface.

The compute path is libbamrex_hip.so (hand-written HIP for gfx950, C ABI in
include/bamrex_hip.h); this package is its Python host mirror.  There is no
CPU fallback: every call raises if the library is missing.
"""
from ._lib import BamrexError, lib  # noqa: F401
from .level import Layout, Level, chop_domain, make_params  # noqa: F401
from . import problems  # noqa: F401
from .inputs import ParmParse, read_inputs  # noqa: F401

__all__ = ["BamrexError", "Layout", "Level", "chop_domain", "make_params", "problems",
           "ParmParse", "read_inputs", "lib"]
