"""GPU parity of the two-level driver (bx_amr_*, through the C ABI) against
the oracle's restatement (or_amr_*): bit-exact fine ghost fill and whole
runs -- coarse advance, two subcycled fine advances with coarse-fine ghost
interpolation, flux registers, reflux, avgDown and the two-level dt."""
import numpy as np
import pytest

from bamrex_amd import amr, problems
from bamrex_amd.level import make_params

from conftest import announce
from test_amr_oracle import oparams, uniform_setup, vortex2d, vortex3d

pytestmark = pytest.mark.gpu


def gparams(s, method="muscl_hancock"):
    return make_params(s["layout"].dim, s["adiabatic"], s["epsilon"], num_method=method,
                       limiter=s["limiter"], limiter_indices=s.get("limiter_indices"),
                       limit_on_primv=s.get("limit_on_primv"))


def assert_equal_valid(lay, got, ref, tag):
    for b in range(len(lay.boxes)):
        g, r = lay.valid(got, b), lay.valid(ref, b)
        if not np.array_equal(g, r):
            d = np.argwhere(g != r)
            raise AssertionError("%s box %d: %d cells differ, first %s gpu %r oracle %r" % (
                tag, b, len(d), d[0], g[tuple(d[0])], r[tuple(d[0])]))


def quadrant(ncell, max_grid_size):
    return problems.quadrant_setup(ncell=ncell, max_grid_size=max_grid_size)


CASES = {
    # name: (setup, coarse ncell, mgs, region lo, region hi, fine mgs, bc override, steps, method)
    "quadrant_inner": (quadrant, 48, 16, (12, 8), (35, 31), 16, None, 8, "muscl_hancock"),
    "quadrant_reflective_wall": (quadrant, 48, 16, (0, 10), (23, 37), 16, "reflective", 8,
                                 "muscl_hancock"),
    "vortex_periodic_edge": (vortex2d, 32, 16, (0, 4), (11, 27), 8, None, 10, "muscl_hancock"),
    "vortex_force": (vortex2d, 32, 16, (6, 10), (21, 25), 8, None, 6, "force"),
    "vortex3d": (vortex3d, 12, 6, (2, 3, 4), (9, 8, 9), 6, None, 4, "muscl_hancock"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_amr_run_parity(oracle_bxpow, name):
    make, n, mgs, rlo, rhi, fmgs, bc, steps, method = CASES[name]
    s = amr.refine_setup(make, n, mgs, rlo, rhi, fine_max_grid_size=fmgs)
    if bc:
        s["bc"] = [bc] * s["layout"].dim
    cfl = s.get("cfl", 0.8)
    announce("amr run %s (%d steps)" % (name, steps))
    h = amr.AmrHierarchy(gparams(s, method), s["layout"], s["fine_layout"], s["domain_lo"],
                         s["domain_hi"], s["bc"], s["bc"])
    try:
        info = h.info()
        assert info["register_faces"] > 0 and info["cf_segments"] > 0
        h.set_state(s["U"], s["Uf"])
        tg, ng = h.run(s["dx"], cfl, max_step=steps)
        Uc, Uf = h.get_state()
    finally:
        h.close()
    Uco, Ufo = s["U"].copy(), s["Uf"].copy()
    to, no = oracle_bxpow.amr_run(oparams(s, method), s["layout"], Uco, s["fine_layout"], Ufo,
                                  s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], s["dx"], cfl,
                                  max_step=steps)
    assert (tg, ng) == (to, no)
    assert_equal_valid(s["layout"], Uc, Uco, name + " coarse")
    assert_equal_valid(s["fine_layout"], Uf, Ufo, name + " fine")


def face_mask(lay, b):
    """Valid cells and face ghost cells (outside the box in one direction)
    of box b's FAB: the cells a sweep reads.  Edge and corner ghost cells
    next to the coarse-fine boundary are left unfilled by design."""
    lo, hi = lay.boxes[b]
    dim, ng = lay.dim, lay.ng
    idx = np.indices(tuple(hi[d] - lo[d] + 1 + 2 * ng for d in reversed(range(dim))))
    nout = np.zeros(idx.shape[1:], dtype=int)
    for d in range(dim):
        x = idx[dim - 1 - d] - ng
        nout += (x < 0) | (x > hi[d] - lo[d])
    return nout <= 1


@pytest.mark.parametrize("bc", ["transmissive", "periodic", "reflective"])
def test_amr_fill_fine_parity(oracle_bxpow, bc):
    """Fine ghost fill at three times between two random coarse states:
    valid and face ghost cells equal the oracle's."""
    rng = np.random.default_rng(3)
    s = amr.refine_setup(lambda **k: uniform_setup(bc=bc, **k), 16, 8, (0, 3), (9, 12),
                         fine_max_grid_size=6)
    lay, flay = s["layout"], s["fine_layout"]
    Uo = 1.0 + rng.random(lay.total)
    Un = 1.0 + rng.random(lay.total)
    Uf0 = 1.0 + rng.random(flay.total)
    h = amr.AmrHierarchy(gparams(s), lay, flay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        for t in (0.0, 0.3, 1.0):
            announce("amr fill_fine %s t=%g" % (bc, t))
            h.set_coarse_old(Uo)
            h.set_state(Un, Uf0)
            h.fill_fine(0.0, 1.0, t)
            got = h.fine.get_state()
            ref = Uf0.copy()
            oracle_bxpow.amr_fill_fine(oparams(s), lay, Uo, Un, flay, ref, s["domain_lo"],
                                       s["domain_hi"], s["bc"], s["bc"], 0.0, 1.0, t)
            for b in range(len(flay.boxes)):
                m = face_mask(flay, b)
                g, r = flay.fab(got, b)[:, m], flay.fab(ref, b)[:, m]
                assert np.array_equal(g, r), (t, b, np.argwhere(g != r)[:4])
    finally:
        h.close()


def test_amr_conservation_on_gpu():
    """The GPU run conserves the composite totals to rounding (periodic
    vortex, fine patch across the periodic boundary)."""
    s = amr.refine_setup(vortex2d, 64, 16, (0, 8), (23, 55), fine_max_grid_size=16)
    announce("amr conservation 64^2")
    h = amr.AmrHierarchy(gparams(s), s["layout"], s["fine_layout"], s["domain_lo"], s["domain_hi"],
                         s["bc"], s["bc"])
    try:
        h.set_state(s["U"], s["Uf"])
        h.run(s["dx"], 0.9, max_step=0)
        Uc, Uf = h.get_state()
        tot0 = amr.composite_totals(s["layout"], Uc, s["fine_layout"], Uf, s["dx"])
        h.run(s["dx"], 0.9, max_step=20)
        Uc, Uf = h.get_state()
        tot1 = amr.composite_totals(s["layout"], Uc, s["fine_layout"], Uf, s["dx"])
    finally:
        h.close()
    assert np.all(np.abs(tot1 - tot0) <= 2e-14 * np.abs(tot0).max())


def test_amr_plotfile_after_run(tmp_path):
    """A two-level plotfile written after a few steps holds both levels'
    valid cells as the device state has them."""
    from plotfile_reader import read_plotfile
    s = amr.refine_setup(quadrant, 32, 16, (8, 8), (23, 23), fine_max_grid_size=8)
    announce("amr plotfile 32^2")
    h = amr.AmrHierarchy(gparams(s), s["layout"], s["fine_layout"], s["domain_lo"], s["domain_hi"],
                         s["bc"], s["bc"])
    try:
        h.set_state(s["U"], s["Uf"])
        t, n = h.run(s["dx"], 0.8, max_step=3)
        h.write_plotfile(str(tmp_path / "plt00003"), t, s["dx"], (0.0, 0.0), (1.0, 1.0), steps=n)
        Uc, Uf = h.get_state()
    finally:
        h.close()
    pf = read_plotfile(str(tmp_path / "plt00003"))
    assert pf["time"] == t and pf["steps"] == [3, 6]
    for lay, U, lv in ((s["layout"], Uc, pf["levels"][0]), (s["fine_layout"], Uf, pf["levels"][1])):
        for b, fab in enumerate(lv["fabs"]):
            assert np.array_equal(fab["data"], lay.valid(U, b))


def test_amr_create_rejects_bad_hierarchies():
    """Overlapping fine boxes and boxes not aligned to the ratio are refused
    at set-up (no kernel runs)."""
    from bamrex_amd._lib import BamrexError
    from bamrex_amd.level import Layout
    s = uniform_setup(16, 8)
    for boxes in ([((4, 4), (11, 11)), ((10, 4), (17, 11))],   # overlap
                  [((3, 4), (10, 11))]):                       # odd lo
        flay = Layout(2, boxes, 2)
        with pytest.raises(BamrexError):
            amr.AmrHierarchy(gparams(s), s["layout"], flay, s["domain_lo"], s["domain_hi"],
                             s["bc"], s["bc"]).close()



@pytest.mark.parametrize("case", ["2d", "2d_periodic_edge"])
def test_three_level_run_parity(oracle_bxpow, case):
    """Three levels, recursively subcycled: bit-exact against the oracle."""
    from test_amr_oracle import THREE_LEVELS
    c = THREE_LEVELS[case]
    s = amr.refine_levels(c["make"], c["ncell"], c["mgs"], c["regions"], c["mgss"])
    announce("amr 3 levels %s" % case)
    h = amr.AmrLevels(gparams(s), s["layouts"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        h.set_state(s["Us"])
        tg, ng = h.run(s["dx"], 0.9, max_step=5)
        got = h.get_state()
    finally:
        h.close()
    Us = [U.copy() for U in s["Us"]]
    to, no = oracle_bxpow.amr_run_levels(oparams(s), s["layouts"], Us, s["domain_lo"],
                                         s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9,
                                         max_step=5)
    assert (tg, ng) == (to, no)
    for l, lay in enumerate(s["layouts"]):
        assert_equal_valid(lay, got[l], Us[l], "%s level %d" % (case, l))


@pytest.mark.parametrize("name", ["quadrant_inner", "vortex_periodic_edge", "vortex3d"])
def test_amr_graph_replay_equals_host_steps(name):
    """Runs of a fixed number of steps replay one captured coarse step whose
    dt, time, register coefficients and coarse-fine weights come from the
    device schedule: the same bits and times as host-driven steps, also for
    a second run that continues from the first's time (the graph reused)
    and after a launch-shape change (the graph recaptured)."""
    make, n, mgs, rlo, rhi, fmgs, bc, steps, method = CASES[name]
    s = amr.refine_setup(make, n, mgs, rlo, rhi, fine_max_grid_size=fmgs)
    cfl = s.get("cfl", 0.8)
    out = {}
    for graph in (False, True):
        announce("amr %s graph=%s" % (name, graph))
        h = amr.AmrHierarchy(gparams(s, method), s["layout"], s["fine_layout"], s["domain_lo"],
                             s["domain_hi"], s["bc"], s["bc"])
        try:
            h.set_graph(graph)
            h.set_state(s["U"], s["Uf"])
            t1, n1 = h.run(s["dx"], cfl, max_step=3)
            t2, n2 = h.run(s["dx"], cfl, max_step=steps - 3, start_time=t1)
            h.coarse.set_tuning(block_x=128, block_t=64, seg_k=8)
            t3, n3 = h.run(s["dx"], cfl, max_step=2, start_time=t2)
            out[graph] = ((t1, n1, t2, n2, t3, n3), h.get_state())
        finally:
            h.close()
    assert out[True][0] == out[False][0]
    assert_equal_valid(s["layout"], out[True][1][0], out[False][1][0], name + " coarse")
    assert_equal_valid(s["fine_layout"], out[True][1][1], out[False][1][1], name + " fine")


def test_three_level_graph_replay_equals_host_steps():
    from test_amr_oracle import THREE_LEVELS
    c = THREE_LEVELS["2d"]
    s = amr.refine_levels(c["make"], c["ncell"], c["mgs"], c["regions"], c["mgss"])
    out = {}
    for graph in (False, True):
        announce("amr 3 levels graph=%s" % graph)
        h = amr.AmrLevels(gparams(s), s["layouts"], s["domain_lo"], s["domain_hi"], s["bc"],
                          s["bc"])
        try:
            h.set_graph(graph)
            h.set_state(s["Us"])
            out[graph] = (h.run(s["dx"], 0.9, max_step=5), h.get_state())
        finally:
            h.close()
    assert out[True][0] == out[False][0]
    for l, lay in enumerate(s["layouts"]):
        assert_equal_valid(lay, out[True][1][l], out[False][1][l], "level %d" % l)


def test_improperly_nested_level_is_refused():
    """A level-2 box reaching within 2 level-1 cells of the level-1 region's
    edge (inside the domain) is refused at set-up."""
    from bamrex_amd._lib import BamrexError
    from test_amr_oracle import THREE_LEVELS
    c = dict(THREE_LEVELS["2d"])
    c["regions"] = [((8, 12), (55, 51)), ((18, 40), (79, 71))]
    s = amr.refine_levels(c["make"], c["ncell"], c["mgs"], c["regions"], c["mgss"])
    with pytest.raises(BamrexError):
        amr.AmrLevels(gparams(s), s["layouts"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
