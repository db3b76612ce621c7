"""Parity of the configurations the benchmark times, at their full sizes.

The benchmark autotunes the launch shapes (block sizes, y-segment length,
persistent-grid cap) and replays each step as a captured hipGraph.  These
tests run exactly that -- the 4096^2 quadrant level (1024 boxes of 128^2)
and the 3000-box uneven level -- under the autotuned shape and under fixed
shapes that exercise the partial-segment, small-block and capped-grid code
paths, and compare every valid cell bit for bit with the oracle built on the
shared portable pow (liboracle_bxpow.so).

Reference path: compute_MUSCL_hancock_flux (Source/Fluxes/MUSCL_Hancock.cpp:
15-52) and the split update / ghost refill of advance_explicit
(Source/AmrLevelAdv_advance.cpp:261-360), stepped by the main.cpp:76-94 loop.
"""
import contextlib
import ctypes

import numpy as np
import pytest

import pyoracle
from bamrex_amd import _lib, problems
from bamrex_amd.level import Level, make_params

from conftest import announce

pytestmark = pytest.mark.gpu

STEPS = 3

# (label, block_x, block_t, seg_k, grid_cap); None = the library's default
# shape.  43 does not divide 128 (a partial last segment per column), 7 and
# 128 are the short and whole-column extremes, 64-thread blocks and caps of
# 1 and 4 blocks per CU change the persistent grid and the XCD remapping.
FIXED_SHAPES = [
    ("default", None),
    ("256-256-k43-cap4", (256, 256, 43, 4)),
    ("64-64-k7-cap1", (64, 64, 7, 1)),
    ("128-128-k128-cap0", (128, 128, 128, 0)),
    ("256-64-k16-cap2", (256, 64, 16, 2)),
]


@contextlib.contextmanager
def created_stream():
    """A created HIP stream: bx_run captures and replays its step graphs only
    on one (the null stream runs the host loop)."""
    st = ctypes.c_void_p()
    _lib.check(_lib.lib().bx_stream_create(ctypes.byref(st)), "bx_stream_create")
    try:
        yield st
    finally:
        _lib.lib().bx_stream_destroy(st)


def valid_mask(layout):
    m = np.zeros(layout.total, dtype=bool)
    for b in range(len(layout.boxes)):
        layout.valid(m, b)[...] = True
    return m


def first_difference(got, ref, layout):
    for b in range(len(layout.boxes)):
        va, vb = layout.valid(got, b), layout.valid(ref, b)
        if not np.array_equal(va, vb):
            d = np.argwhere(va != vb)
            return "box %d: %d cells differ, first %s gpu %r oracle %r" % (
                b, len(d), tuple(d[0]), va[tuple(d[0])], vb[tuple(d[0])])
    return "no difference"


@pytest.fixture(scope="module")
def quadrant_ref():
    """The 4096^2 quadrant level and the oracle's state after STEPS steps."""
    s = problems.quadrant_setup(ncell=4096, max_grid_size=128)
    op = pyoracle.params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    ref = s["U"].copy()
    t, n = pyoracle.Oracle(bxpow=True).run(op, s["layout"], ref, s["domain_lo"], s["domain_hi"],
                                          s["bc"], s["bc"], s["dx"], s["cfl"], max_step=STEPS)
    return s, ref, (t, n), valid_mask(s["layout"])


def run_quadrant(s, shape, stream):
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    lev = Level(gp, s["layout"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    try:
        lev.set_state(s["U"])
        if shape == "autotune":
            # as the benchmark does: tune on the initial state, then reload it
            lev.fill_boundary()
            dt0 = lev.estTimeStep(s["dx"], s["cfl"])
            lev.autotune(s["dx"], dt0)
            lev.set_state(s["U"])
        elif shape is not None:
            bx_, bt_, k_, cap_ = shape
            lev.set_tuning(block_x=bx_, block_t=bt_, seg_k=k_, grid_cap=cap_)
        tuning = lev.get_tuning()
        tn = lev.run(s["dx"], s["cfl"], max_step=STEPS, use_graph=True)
        return lev.get_state(), tn, tuning
    finally:
        lev.close()


SHAPES = [("autotuned", "autotune")] + FIXED_SHAPES


@pytest.mark.parametrize("label,shape", SHAPES, ids=[sh[0] for sh in SHAPES])
def test_quadrant_4096_graph_steps_bit_exact(quadrant_ref, label, shape):
    """Config 2 at full size: 3 graph-replayed steps under the given launch
    shape equal the oracle's 3 steps in every valid cell, with the same time."""
    s, ref, tn_ref, mask = quadrant_ref
    announce("quadrant 4096^2 mgs 128, %d graph steps, shape %s" % (STEPS, label))
    with created_stream() as st:
        got, tn, tuning = run_quadrant(s, shape, st)
    announce("quadrant 4096^2 shape %s ran as %s" % (label, tuning))
    assert tn == tn_ref
    if not np.array_equal(got[mask], ref[mask]):
        raise AssertionError("shape %s (%s): %s" % (label, tuning,
                                                    first_difference(got, ref, s["layout"])))


@pytest.fixture(scope="module")
def uneven_level():
    s = problems.uneven_setup()
    return s, valid_mask(s["layout"])


def test_uneven_3000_boxes_autotuned_sweeps_and_graph_steps(uneven_level):
    """Config 4 at full size (3000 boxes of 8..96 cells per side): under the
    autotuned shape, one sweep per direction over all boxes equals the
    oracle's sweep, and 3 graph-replayed steps equal the oracle's 3 steps."""
    s, mask = uneven_level
    lay = s["layout"]
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    op = pyoracle.params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    orc = pyoracle.Oracle(bxpow=True)
    with created_stream() as st:
        got, tn, tuning = _uneven_gpu(s, mask, gp, op, orc, st)
    ref = s["U"].copy()
    tn_ref = orc.run(op, lay, ref, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], s["dx"],
                     s["cfl"], max_step=STEPS)
    assert tn == tn_ref
    if not np.array_equal(got[mask], ref[mask]):
        raise AssertionError("graph steps (%s): %s" % (tuning, first_difference(got, ref, lay)))


def _uneven_gpu(s, mask, gp, op, orc, stream):
    lay = s["layout"]
    lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    try:
        lev.set_state(s["U"])
        lev.set_tmp(s["U"])
        announce("uneven 3000 boxes: autotune")
        dt = lev.estTimeStep(s["dx"], s["cfl"])
        lev.autotune(s["dx"], dt)
        tuning = lev.get_tuning()
        announce("uneven 3000 boxes: tuned shape %s" % tuning)
        for d in range(2):
            lev.set_state(s["U"])
            lev.set_tmp(s["U"])
            announce("uneven 3000 boxes: sweep dir %d" % d)
            lev.sweep(d, dt, s["dx"])
            _lib.check(_lib.lib().bx_stream_sync(stream), "bx_stream_sync")
            got = lev.get_state("tmp")
            ref = s["U"].copy()
            orc.sweep(op, lay, s["U"], ref, d, dt, s["dx"], s["domain_lo"], s["domain_hi"],
                      s["bc"], s["bc"])
            if not np.array_equal(got[mask], ref[mask]):
                raise AssertionError("sweep %d (%s): %s" % (d, tuning,
                                                            first_difference(got, ref, lay)))
        lev.set_state(s["U"])
        announce("uneven 3000 boxes: %d graph steps" % STEPS)
        tn = lev.run(s["dx"], s["cfl"], max_step=STEPS, use_graph=True)
        return lev.get_state(), tn, tuning
    finally:
        lev.close()
