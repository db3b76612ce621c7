"""CPU tests of the oracle's two-level (ratio 2, subcycled) driver, the
checker of bx_amr_*.  What AMReX itself computes cannot be run here; these
pin the properties its algorithm guarantees: free streams stay free,
refluxing makes the composite grid conservative (and removing it breaks
that), the coarse-fine interpolation reproduces linear data, conserves the
coarse cell average and stays within the neighbourhood's range, and the
time interpolation weights are linear."""
import numpy as np
import pytest

import pyoracle
from bamrex_amd import amr, problems
from bamrex_amd.level import Layout, chop_domain


def uniform_setup(ncell, max_grid_size, dim=2, bc="periodic", prim=(1.0, 0.5, -0.25, 1.0)):
    lo, hi = (0,) * dim, (ncell - 1,) * dim
    boxes = chop_domain(lo, hi, max_grid_size)
    lay = Layout(dim, boxes, 2)
    U = lay.new_array()
    q = [np.float64(prim[0])] + [np.float64(prim[1 + d]) if d < 2 else np.float64(0.125)
                                 for d in range(dim)] + [np.float64(prim[-1])]
    c = problems.consv_from_primv(q, 1.4, 1.0, dim)
    for b in range(len(boxes)):
        for n in range(dim + 2):
            lay.valid(U, b)[n][...] = c[n]
    return dict(layout=lay, U=U, dx=(1.0 / ncell,) * dim, domain_lo=lo, domain_hi=hi,
                bc=[bc] * dim, adiabatic=1.4, epsilon=1.0, limiter="minbee")


def oparams(s, method="muscl_hancock"):
    return pyoracle.params(s["layout"].dim, s["adiabatic"], s["epsilon"], num_method=method,
                           limiter=s["limiter"], limiter_indices=s.get("limiter_indices"),
                           limit_on_primv=s.get("limit_on_primv"))


def max_diff(lay, A, B):
    return max(float(np.max(np.abs(lay.valid(A, b) - lay.valid(B, b)))) for b in range(len(lay.boxes)))


@pytest.mark.parametrize("bc", ["periodic", "transmissive", "reflective"])
def test_free_stream_two_levels(oracle_bxpow, bc):
    """Uniform flow with a fine patch: every flux difference, interpolation
    slope and register sum is zero; what remains is the rounding of the time
    interpolation (alpha u + beta u, alpha + beta = 1 only to rounding) and
    of the 4-cell average -- a few ulps, no truncation error."""
    prim = (1.0, 0.5, -0.25, 1.0) if bc != "reflective" else (1.0, 0.0, 0.0, 1.0)
    s = amr.refine_setup(lambda **k: uniform_setup(bc=bc, prim=prim, **k), 32, 16, (8, 8), (19, 23),
                         fine_max_grid_size=8)
    Uc, Uf = s["U"].copy(), s["Uf"].copy()
    t, n = oracle_bxpow.amr_run(oparams(s), s["layout"], Uc, s["fine_layout"], Uf, s["domain_lo"],
                                s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.8, max_step=5)
    assert n == 5 and t > 0
    assert max_diff(s["fine_layout"], Uf, s["Uf"]) <= 1e-15 * np.max(np.abs(s["Uf"]))
    assert max_diff(s["layout"], Uc, s["U"]) <= 1e-15 * np.max(np.abs(s["U"]))


def vortex2d(ncell, max_grid_size):
    return problems.vortex_setup(ncell=ncell, dim=2, max_grid_size=max_grid_size)


def vortex3d(ncell, max_grid_size):
    return problems.vortex_setup(ncell=ncell, dim=3, max_grid_size=max_grid_size)


@pytest.mark.parametrize("case", ["2d_inner", "2d_periodic_edge", "3d"])
def test_composite_conservation_needs_reflux(oracle_bxpow, case):
    """Periodic isentropic vortex with a fine patch (in the interior, or
    touching the periodic boundary so the coarse-fine cells wrap): the
    composite totals of every conserved variable stay constant to rounding
    with refluxing; without it they drift by orders of magnitude more."""
    if case == "2d_inner":
        s = amr.refine_setup(vortex2d, 32, 16, (6, 10), (21, 25), fine_max_grid_size=8)
    elif case == "2d_periodic_edge":
        s = amr.refine_setup(vortex2d, 32, 16, (0, 4), (11, 27), fine_max_grid_size=8)
    else:
        s = amr.refine_setup(vortex3d, 12, 6, (2, 3, 4), (9, 8, 9), fine_max_grid_size=6)
    op = oparams(s)
    steps = 6 if case == "3d" else 10
    drift = {}
    for rf in (True, False):
        Uc, Uf = s["U"].copy(), s["Uf"].copy()
        oracle_bxpow.amr_run(op, s["layout"], Uc, s["fine_layout"], Uf, s["domain_lo"],
                             s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9, max_step=0)
        tot0 = amr.composite_totals(s["layout"], Uc, s["fine_layout"], Uf, s["dx"])
        oracle_bxpow.amr_run(op, s["layout"], Uc, s["fine_layout"], Uf, s["domain_lo"],
                             s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9, max_step=steps,
                             do_reflux=rf)
        tot1 = amr.composite_totals(s["layout"], Uc, s["fine_layout"], Uf, s["dx"])
        scale = np.abs(tot0).max()
        drift[rf] = np.abs(tot1 - tot0) / scale
    assert np.all(drift[True] <= 2e-14), drift
    assert np.max(drift[False]) > 1e-8, drift


def fill_case(oracle, coarse_fn, dim=2, bc="transmissive", t=(0.0, 1.0, 0.0), new_fn=None):
    """Coarse data from coarse_fn(x) (x = coarse cell index, ghosts
    included), a fine patch, and the oracle's fine ghost fill."""
    n = 16
    s = uniform_setup(n, 8, dim=dim, bc=bc)
    lay = s["layout"]
    Uo = lay.new_array()
    Un = lay.new_array()
    for b in range(len(lay.boxes)):
        fab_o = lay.fab(Uo, b)
        fab_n = lay.fab(Un, b)
        lo, _ = lay.boxes[b]
        idx = np.indices(fab_o.shape[1:])
        coords = [idx[dim - 1 - d] + lo[d] - lay.ng for d in range(dim)]
        for m in range(dim + 2):
            fab_o[m] = coarse_fn(coords, m)
            fab_n[m] = (new_fn or coarse_fn)(coords, m)
    flo, fhi = (8,) * dim, (19,) * dim
    flay = Layout(dim, chop_domain(flo, fhi, 6), 2)
    Uf = flay.new_array()
    oracle.amr_fill_fine(oparams(s), lay, Uo, Un, flay, Uf, s["domain_lo"], s["domain_hi"],
                         s["bc"], s["bc"], *t)
    return flay, Uf


def face_ghosts(flay, b):
    """(index tuple into the fab, global fine cell) of the face ghost cells
    of fine box b that lie outside the fine level."""
    dim = flay.dim
    lo, hi = flay.boxes[b]
    ng = flay.ng
    out = []
    shape = [hi[d] - lo[d] + 1 + 2 * ng for d in range(dim)]
    for flat in range(int(np.prod(shape))):
        x = []
        r = flat
        for d in range(dim):
            x.append(lo[d] - ng + r % shape[d])
            r //= shape[d]
        nout = sum(1 for d in range(dim) if x[d] < lo[d] or x[d] > hi[d])
        if nout != 1:
            continue
        if any(all(blo[d] <= x[d] <= bhi[d] for d in range(dim)) for blo, bhi in flay.boxes):
            continue
        idx = tuple(x[d] - lo[d] + ng for d in reversed(range(dim)))
        out.append((idx, x))
    return out


@pytest.mark.parametrize("dim", [2, 3])
def test_cf_interpolation_reproduces_linear_data(oracle_bxpow, dim):
    """Coarse data linear in the cell index: MC slopes equal the gradient,
    the limiter does not act, and each fine ghost value is the linear
    function at the fine cell centre, (x_fine + 1/2)/2 - 1/2 in coarse units."""
    g = [0.75, -0.5, 0.25]

    def lin(c, m):
        return 2.0 + m + sum(g[d] * c[d] for d in range(dim))

    flay, Uf = fill_case(oracle_bxpow, lin, dim=dim)
    worst = 0.0
    count = 0
    for b in range(len(flay.boxes)):
        fab = flay.fab(Uf, b)
        for idx, x in face_ghosts(flay, b):
            xc = [(x[d] + 0.5) / 2 - 0.5 for d in range(dim)]
            for m in range(dim + 2):
                exp = 2.0 + m + sum(g[d] * xc[d] for d in range(dim))
                worst = max(worst, abs(fab[(m,) + idx] - exp))
                count += 1
    assert count > 0 and worst <= 1e-14, worst


def test_cf_interpolation_conserves_and_is_bounded(oracle_bxpow):
    """Random coarse data: the children of a coarse cell average to the
    coarse value (the face ghost slab, two cells deep, holds whole coarse
    cells) and every child lies within the min/max of the coarse 3x3
    neighbourhood."""
    rng = np.random.default_rng(5)
    table = rng.uniform(0.5, 2.0, size=(4, 40, 40))

    def rnd(c, m):
        return table[m][c[1] + 8, c[0] + 8]

    flay, Uf = fill_case(oracle_bxpow, rnd)
    groups = {}
    for b in range(len(flay.boxes)):
        fab = flay.fab(Uf, b)
        for idx, x in face_ghosts(flay, b):
            key = (x[0] // 2, x[1] // 2)
            groups.setdefault(key, []).append(fab[(slice(None),) + idx])
    checked = 0
    for (cx, cy), vals in groups.items():
        nb = table[:, cy + 8 - 1:cy + 8 + 2, cx + 8 - 1:cx + 8 + 2]
        for v in vals:
            assert np.all(v >= nb.min(axis=(1, 2)) - 1e-15) and np.all(v <= nb.max(axis=(1, 2)) + 1e-15)
        if len(vals) == 4:
            mean = (vals[0] + vals[1] + vals[2] + vals[3]) / 4
            assert np.allclose(mean, table[:, cy + 8, cx + 8], rtol=4e-16, atol=0)
            checked += 1
    assert checked > 10


def test_cf_time_interpolation_weights(oracle_bxpow):
    """New coarse state = old + 3 (a constant shift keeps the slopes): at
    t = t_old the fill uses the old state, at 1/4 of the interval it adds
    3/4 of the shift, at t_new all of it."""
    def base(c, m):
        return 1.0 + 0.1 * c[0] - 0.05 * c[1] + m

    def shifted(c, m):
        return base(c, m) + 3.0

    res = {}
    for t in (0.0, 0.25, 1.0):
        flay, Uf = fill_case(oracle_bxpow, base, t=(0.0, 1.0, t), new_fn=shifted)
        res[t] = Uf
    cells = [(b, idx) for b in range(len(flay.boxes)) for idx, _ in face_ghosts(flay, b)]
    for b, idx in cells:
        for m in range(4):
            a0 = flay.fab(res[0.0], b)[(m,) + idx]
            assert abs(flay.fab(res[0.25], b)[(m,) + idx] - a0 - 0.75) <= 1e-14
            assert abs(flay.fab(res[1.0], b)[(m,) + idx] - a0 - 3.0) <= 1e-14
    assert len(cells) > 20


THREE_LEVELS = {
    # level-1 and level-2 regions in their own index spaces (nested by >= 2
    # parent cells), level 0 = 32^2
    "2d": dict(make=vortex2d, ncell=32, mgs=16,
               regions=[((8, 12), (55, 51)), ((32, 40), (79, 71))], mgss=[16, 16]),
    # level 1 spans the periodic x direction; level 2 touches the periodic
    # edge, so its coarse-fine cells there wrap around
    "2d_periodic_edge": dict(make=vortex2d, ncell=32, mgs=16,
                             regions=[((0, 8), (63, 55)), ((0, 24), (39, 87))], mgss=[16, 16]),
}


@pytest.mark.parametrize("case", list(THREE_LEVELS))
def test_three_levels_conserve_and_need_reflux(oracle_bxpow, case):
    """Three levels (ratio 2): the composite totals stay constant to rounding
    with refluxing at both coarse-fine interfaces, and drift without it."""
    c = THREE_LEVELS[case]
    s = amr.refine_levels(c["make"], c["ncell"], c["mgs"], c["regions"], c["mgss"])
    op = oparams(s)
    drift = {}
    for rf in (True, False):
        Us = [U.copy() for U in s["Us"]]
        oracle_bxpow.amr_run_levels(op, s["layouts"], Us, s["domain_lo"], s["domain_hi"], s["bc"],
                                    s["bc"], s["dx"], 0.9, max_step=0)
        tot0 = amr.composite_totals_levels(s["layouts"], Us, s["dx"])
        t, n = oracle_bxpow.amr_run_levels(op, s["layouts"], Us, s["domain_lo"], s["domain_hi"],
                                           s["bc"], s["bc"], s["dx"], 0.9, max_step=6,
                                           do_reflux=rf)
        assert n == 6
        tot1 = amr.composite_totals_levels(s["layouts"], Us, s["dx"])
        drift[rf] = np.abs(tot1 - tot0) / np.abs(tot0).max()
    assert np.all(drift[True] <= 2e-14), drift
    assert np.max(drift[False]) > 1e-9, drift


def test_two_level_entry_equals_levels_entry(oracle_bxpow):
    """or_amr_run (two levels) is or_amr_run_levels with nlev = 2."""
    s = amr.refine_setup(vortex2d, 32, 16, (6, 10), (21, 25), fine_max_grid_size=8)
    op = oparams(s)
    Uc, Uf = s["U"].copy(), s["Uf"].copy()
    r1 = oracle_bxpow.amr_run(op, s["layout"], Uc, s["fine_layout"], Uf, s["domain_lo"],
                              s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9, max_step=4)
    Us = [s["U"].copy(), s["Uf"].copy()]
    r2 = oracle_bxpow.amr_run_levels(op, [s["layout"], s["fine_layout"]], Us, s["domain_lo"],
                                     s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9, max_step=4)
    assert r1 == r2 and np.array_equal(Uc, Us[0]) and np.array_equal(Uf, Us[1])


def test_oracle_refuses_improper_nesting(oracle_bxpow):
    c = THREE_LEVELS["2d"]
    s = amr.refine_levels(c["make"], c["ncell"], c["mgs"], [c["regions"][0], ((18, 40), (79, 71))],
                          c["mgss"])
    with pytest.raises(ValueError):
        oracle_bxpow.amr_run_levels(oparams(s), s["layouts"], [U.copy() for U in s["Us"]],
                                    s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], s["dx"], 0.9,
                                    max_step=1)
