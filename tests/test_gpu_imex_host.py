"""GPU test of the C++ host's IMEX path (adv.num_method = imex;
AmrLevelAdv_advance.cpp:131-157, IMEX_RK.cpp:22-96, the imex.* keys of
IMEXSettings.cpp:11-78): the reference's PEEK-MUSCL-SSP222 Gresho input
(periodic variant) run by bamrex_run, every stage's explicit work and
pressure solve on the GPU, against the oracle's IMEX Runge-Kutta steps with
the oracle's plain multigrid solves (oracle_helm), from the host's own
initial state.  The bar is bit-exact: the GPU multigrid iterate equals the
oracle's, and every stage kernel is an operation-order restatement."""
import os
import subprocess
import sys

import numpy as np
import pytest

from bamrex_amd.imex import gresho_state
from conftest import ROOT, announce
from plotfile_reader import read_plotfile

sys.path.insert(0, os.path.join(ROOT, "oracle"))
import helm_oracle as ho  # noqa: E402
import imex_oracle as io  # noqa: E402

pytestmark = pytest.mark.gpu

HOST = os.path.join(ROOT, "bamrex_amd", "bamrex_run")
INPUT = os.path.join(ROOT, "bamrex_amd", "inputs", "gresho-peek-muscl-ssp222")


def state_of(path):
    pf = read_plotfile(path)
    (fab,) = pf["levels"][0]["fabs"]
    return np.array(fab["data"][:4]), pf["time"], pf["steps"][0]


def run(tmp_path, *keys):
    cmd = [HOST, INPUT, "amr.plot_file=" + str(tmp_path / "plt"), "amr.plot_int=1"] + list(keys)
    announce("imex host " + " ".join(cmd[2:]))
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.skipif(not os.path.exists(HOST), reason="C++ host not built")
@pytest.mark.parametrize("tableau,n,steps,solver", [("SSP222", 128, 3, "mg_fused"),
                                                    ("SSP222", 128, 3, "mg"),
                                                    ("SP111", 64, 3, "mg_fused"),
                                                    ("SSP433", 64, 2, "mg_fused")])
def test_host_imex_matches_oracle(tmp_path, tableau, n, steps, solver):
    """The host's default pressure solver is the blocked-smoother multigrid
    (mg_fused); the per-colour one (mg) gives the same bits."""
    out = run(tmp_path, "amr.n_cell=%d %d" % (n, n), "max_step=%d" % steps,
              "imex.butcher_tableau=" + tableau, "imex.pressure_rtol=1e-10",
              "imex.pressure_maxiter=100", "imex.pressure_solver=" + solver)
    assert "Steps = %d" % steps in out
    U, t, st = state_of(str(tmp_path / "plt00000"))
    assert (t, st) == (0.0, 0)
    # the host's initial data is the Gresho vortex (Exec/GreshoVortex/Prob.cpp)
    U_py, dx = gresho_state(n)
    np.testing.assert_allclose(U, U_py, rtol=1e-13, atol=1e-13)
    p = io.pressure_from_state(U, 1.4, 1.0)
    time = 0.0

    def solve(sy, p0):
        x, its, _ = ho.mg_solve(sy["rhs"], sy["a"], sy["bX"], sy["bY"], sy["alpha"], sy["beta"],
                                dx, dx, x0=p0, rtol=1e-10, maxiter=100)
        assert its < 100
        return x

    for k in range(1, steps + 1):
        dt = io.advective_dt(U, dx, dx, 0.45)
        U, p = io.advance_imex_rk(U, p, dt, dx, dx, tableau, solve, 1.4, 1.0, "muscl_rusanov")
        time = time + dt
        G, tg, sg = state_of(str(tmp_path / ("plt%05d" % k)))
        assert sg == k and tg == time, (k, tg, time)
        assert np.array_equal(G, U), (k, np.argwhere(G != U)[:3])


@pytest.mark.skipif(not os.path.exists(HOST), reason="C++ host not built")
def test_host_imex_rejects_unsupported_settings(tmp_path):
    for key in ("imex.linearise=false", "imex.stabilize=true", "geometry.is_periodic=0 0",
                "imex.butcher_tableau=SSP999", "imex.ke_method=bogus"):
        cmd = [HOST, INPUT, "amr.n_cell=32 32", "max_step=1", "amr.plot_files_output=0", key]
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=60)
        assert r.returncode != 0, key
        assert "bamrex_run:" in r.stderr, (key, r.stderr)
