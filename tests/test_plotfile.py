"""CPU tests of the plotfile writer (bx_write_plotfile, host code of the C
ABI): a plotfile of one or two levels read back with an independent reader
of the format gives every valid cell bit for bit, the derived pressure and
velocities of the reference's derive list, the box geometry and the
per-box minima/maxima."""
import numpy as np
import pytest

from bamrex_amd import amr, problems
from bamrex_amd.level import Layout, chop_domain
from bamrex_amd.plotfile import write_plotfile

from plotfile_reader import read_plotfile


def random_state(lay, seed):
    rng = np.random.default_rng(seed)
    U = rng.uniform(-1.0, 1.0, lay.total)
    for b in range(len(lay.boxes)):
        v = lay.valid(U, b)
        v[0] = rng.uniform(0.5, 2.0, v[0].shape)
        v[-1] = rng.uniform(5.0, 6.0, v[-1].shape)
    return U


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_single_level_round_trip(tmp_path, dim):
    n = {1: 40, 2: 24, 3: 10}[dim]
    lay = Layout(dim, chop_domain((0,) * dim, (n - 1,) * dim, 8), 2)
    U = random_state(lay, dim)
    dx = tuple([0.5 / n] * dim)
    prob_lo = tuple([-0.25] * dim)
    prob_hi = tuple(prob_lo[d] + n * dx[d] for d in range(dim))
    path = tmp_path / "plt00007"
    write_plotfile(path, [dict(layout=lay, U=U, domain_lo=(0,) * dim, domain_hi=(n - 1,) * dim,
                               dx=dx, steps=7)], prob_lo, prob_hi, 0.125, 1.4, 0.5, derived=True)
    pf = read_plotfile(str(path))
    names = ["density"] + ["mom_x", "mom_y", "mom_z"][:dim] + ["energy", "pressure"] + \
        ["u_x", "u_y", "u_z"][:dim]
    assert pf["version"] == "HyperCLaw-V1.1" and pf["names"] == names
    assert pf["dim"] == dim and pf["time"] == 0.125 and pf["steps"] == [7]
    assert pf["prob_lo"] == list(prob_lo) and pf["dx"][0] == list(dx)
    assert pf["domains"][0][:2] == ((0,) * dim, (n - 1,) * dim)
    lev = pf["levels"][0]
    assert lev["nbox"] == len(lay.boxes)
    for b, fab in enumerate(lev["fabs"]):
        lo, hi = lay.boxes[b]
        assert fab["lo"] == tuple(lo) and fab["hi"] == tuple(hi)
        for d in range(dim):
            assert lev["phys"][b][d] == (prob_lo[d] + dx[d] * lo[d], prob_lo[d] + dx[d] * (hi[d] + 1))
        v = lay.valid(U, b)
        nc = dim + 2
        assert np.array_equal(fab["data"][:nc], v)
        ke = 0.5 * (v[1:1 + dim] ** 2).sum(axis=0) / (2 * v[0])  # eps |m|^2 / (2 rho), eps 0.5
        assert np.allclose(fab["data"][nc], 0.4 * (v[-1] - ke), rtol=1e-14, atol=1e-14)
        for d in range(dim):
            assert np.array_equal(fab["data"][nc + 1 + d], v[1 + d] / v[0])
        assert fab["min"] == [float(x.min()) for x in fab["data"]]
        assert fab["max"] == [float(x.max()) for x in fab["data"]]


def test_two_level_round_trip(tmp_path):
    """A coarse level and a ratio-2 fine patch in one plotfile."""
    s = amr.refine_setup(lambda ncell, max_grid_size: problems.quadrant_setup(ncell, max_grid_size),
                         32, 16, (8, 4), (19, 27), fine_max_grid_size=8)
    lay, flay = s["layout"], s["fine_layout"]
    levels = [dict(layout=lay, U=s["U"], domain_lo=(0, 0), domain_hi=(31, 31), dx=s["dx"], steps=4),
              dict(layout=flay, U=s["Uf"], domain_lo=(0, 0), domain_hi=(63, 63),
                   dx=tuple(x / 2 for x in s["dx"]), steps=8)]
    path = tmp_path / "plt00004"
    write_plotfile(path, levels, (0.0, 0.0), (1.0, 1.0), 0.5, 1.4)
    pf = read_plotfile(str(path))
    assert pf["ref_ratio"] == [2] and pf["steps"] == [4, 8] and len(pf["levels"]) == 2
    assert pf["domains"][1][:2] == ((0, 0), (63, 63))
    assert pf["dx"][1] == [1 / 64, 1 / 64]
    for L, lv in zip((lay, flay), pf["levels"]):
        U = s["U"] if L is lay else s["Uf"]
        assert lv["nbox"] == len(L.boxes)
        for b, fab in enumerate(lv["fabs"]):
            assert np.array_equal(fab["data"], L.valid(U, b))
