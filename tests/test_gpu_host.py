"""GPU tests of the C++ host driver (bamrex_amd/bamrex_run): the reference's
main / AmrLevelAdv time loop reading the reference's input keys, run on the
GPU through the C ABI, checked bit for bit against the CPU oracle (both the
whole-level path and the per-box flux-function loop, adv.box_loop = 1)."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from bamrex_amd import problems
from bamrex_amd.level import Layout, make_params

from conftest import ROOT, announce
from plotfile_reader import read_plotfile

pytestmark = pytest.mark.gpu

HOST = os.path.join(ROOT, "bamrex_amd", "bamrex_run")
INPUTS = os.path.join(ROOT, "bamrex_amd", "inputs")


def read_final_plotfile(tmp_path):
    """The last plotfile the host wrote (plt<step>), as the packed level."""
    names = sorted(p.name for p in tmp_path.iterdir() if p.name.startswith("plt"))
    pf = read_plotfile(str(tmp_path / names[-1]))
    dim = pf["dim"]
    fabs = pf["levels"][0]["fabs"]
    boxes = [(f["lo"], f["hi"]) for f in fabs]
    lay = Layout(dim, boxes, 2)
    U = lay.new_array()
    for b, f in enumerate(fabs):
        lay.valid(U, b)[...] = f["data"][:dim + 2]
    return dict(dim=dim, time=pf["time"], step=pf["steps"][0], boxes=boxes, layout=lay, U=U,
                names=names)


def run_host(tmp_path, inputs, *overrides, plot_int=1000):
    cmd = [HOST, os.path.join(INPUTS, inputs)] + list(overrides) + [
        "amr.plot_file=" + str(tmp_path / "plt"), "amr.plot_int=%d" % plot_int,
        "amr.check_file=" + str(tmp_path / "chk")]
    announce("host " + " ".join(cmd[1:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return read_final_plotfile(tmp_path)


@pytest.mark.skipif(not os.path.exists(HOST), reason="C++ host not built")
@pytest.mark.parametrize("box_loop", [0, 1])
def test_host_quadrant_matches_oracle(tmp_path, oracle_bxpow, box_loop):
    d = run_host(tmp_path, "quadrant-4096", "amr.n_cell=128 128", "amr.max_grid_size=32",
                 "max_step=8", "adv.box_loop=%d" % box_loop, plot_int=4)
    assert d["names"] == ["plt00000", "plt00004", "plt00008"]
    lay = d["layout"]
    s = problems.quadrant_setup(ncell=128, max_grid_size=32)
    U = lay.new_array()
    problems.quadrant_2d(lay, U, (0.0, 0.0), s["dx"], problems.QUADRANT_CONFIG, 1.4)
    op = pyoracle.params(2, 1.4, 1.0, limiter="minbee")
    t, n = oracle_bxpow.run(op, lay, U, (0, 0), (127, 127), ["transmissive"] * 2,
                            ["transmissive"] * 2, s["dx"], 0.8, max_step=8)
    assert (n, t) == (d["step"], d["time"])
    for b in range(len(lay.boxes)):
        assert np.array_equal(lay.valid(d["U"], b), lay.valid(U, b)), b


@pytest.mark.skipif(not os.path.exists(HOST), reason="C++ host not built")
def test_host_sod_matches_oracle(tmp_path, oracle_bxpow):
    d = run_host(tmp_path, "sod-1024", "stop_time=0.05", "max_step=-1")
    lay = d["layout"]
    s = problems.sod_setup(ncell=1024)
    U = lay.new_array()
    problems.riemann_1d(lay, U, (0.0,), (1.0,), s["dx"], (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    op = pyoracle.params(1, 1.4, 1.0, limiter="van_leer")
    t, n = oracle_bxpow.run(op, lay, U, (0,), (1023,), ["transmissive"], ["transmissive"], s["dx"],
                            0.9, stop_time=0.05)
    assert n == d["step"] and t == d["time"]
    for b in range(len(lay.boxes)):
        assert np.array_equal(lay.valid(d["U"], b), lay.valid(U, b)), b
