"""AMReX plotfiles of one or more levels (bx_write_plotfile): the on-disk
format the reference writes through AmrLevel::writePlotFile
(Source/AmrLevelAdv.cpp:116-136), readable by VisIt and yt."""
import ctypes

import numpy as np

from . import _lib as L
from .level import check


def write_plotfile(path, levels, prob_lo, prob_hi, time, adiabatic, epsilon=1.0, derived=False,
                   ref_ratio=2):
    """levels: list of dicts with layout, U (host array of the packed
    level), domain_lo, domain_hi, dx and steps, coarsest first."""
    dim = levels[0]["layout"].dim
    arr = (L.bx_plot_level * len(levels))()
    keep = []
    for l, lv in enumerate(levels):
        lay = lv["layout"]
        U = np.ascontiguousarray(lv["U"], dtype=np.float64)
        keep.append(U)
        e = arr[l]
        e.boxes = ctypes.cast(lay.c_boxes, ctypes.c_void_p)
        e.nbox = len(lay.boxes)
        e.ng = lay.ng
        e.U = U.ctypes.data
        for d in range(3):
            e.domain_lo[d] = lv["domain_lo"][d] if d < dim else 0
            e.domain_hi[d] = lv["domain_hi"][d] if d < dim else 0
            e.dx[d] = lv["dx"][d] if d < dim else 0.0
        e.level_steps = int(lv.get("steps", 0))
    plo = L.dvec3(prob_lo)
    phi = L.dvec3(prob_hi)
    check(L.lib().bx_write_plotfile(str(path).encode(), dim, len(levels), ctypes.cast(arr, ctypes.c_void_p),
                                    ref_ratio, plo, phi, time, adiabatic, epsilon, int(bool(derived))),
          "bx_write_plotfile")
