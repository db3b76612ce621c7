"""GPU tests of checkpoint / restart (bx_write_checkpoint,
bx_read_checkpoint_*; AmrLevelAdv::checkPoint / restart,
Source/AmrLevelAdv.cpp:95-113, and main.cpp's final checkpoint, :97-100):
a run that stops, writes a checkpoint, and is restarted from it in a fresh
context continues bit for bit as the run that never stopped -- on one level
through the Python binding, on two AMR levels, and through the C++ host's
amr.check_int / amr.restart keys."""
import os
import subprocess

import numpy as np
import pytest

from bamrex_amd import amr, checkpoint, problems
from bamrex_amd.level import Level, make_params

from conftest import ROOT, announce
from test_gpu_amr import assert_equal_valid, gparams, quadrant

pytestmark = pytest.mark.gpu

HOST = os.path.join(ROOT, "bamrex_amd", "bamrex_run")
INPUTS = os.path.join(ROOT, "bamrex_amd", "inputs")


def test_level_restart_continues_bit_for_bit(tmp_path):
    s = problems.quadrant_setup(ncell=256, max_grid_size=64)
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    args = (gp, s["layout"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    announce("checkpoint: quadrant 256^2, 12 steps straight through")
    a = Level(*args)
    try:
        a.set_state(s["U"])
        t_ref, n_ref = a.run(s["dx"], s["cfl"], max_step=12)
        U_ref = a.get_state()
    finally:
        a.close()
    announce("checkpoint: 6 steps, checkpoint, restart in a new context, 6 steps")
    b = Level(*args)
    try:
        b.set_state(s["U"])
        t6, n6 = b.run(s["dx"], s["cfl"], max_step=6)
        dt_next = b.estTimeStep(s["dx"], s["cfl"], time=t6)
        checkpoint.write(str(tmp_path / "chk00006"), [b], t6, s["dx"], steps=[n6], dts=[dt_next])
    finally:
        b.close()
    run, levels = checkpoint.read_info(str(tmp_path / "chk00006"))
    assert run["time"] == t6 and run["nlev"] == 1 and run["dim"] == 2
    assert levels[0]["level_steps"] == 6 and levels[0]["dt"] == dt_next
    lay = checkpoint.read_layout(str(tmp_path / "chk00006"), 0)
    assert lay.boxes == s["layout"].boxes and lay.total == s["layout"].total
    c = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        checkpoint.read_state(str(tmp_path / "chk00006"), 0, c)
        t, n = c.run(s["dx"], s["cfl"], max_step=6, start_time=run["time"])
        U = c.get_state()
    finally:
        c.close()
    assert (t, n6 + n) == (t_ref, n_ref)
    assert_equal_valid(lay, U, U_ref, "restart")


def test_checkpoint_reader_rejects_bad_input(tmp_path):
    s = problems.quadrant_setup(ncell=64, max_grid_size=32)
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    lev = Level(gp, s["layout"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        lev.set_state(s["U"])
        path = str(tmp_path / "chk")
        checkpoint.write(path, [lev], 0.0, s["dx"], steps=[0], dts=[0.0])
        small = problems.quadrant_setup(ncell=32, max_grid_size=32)
        other = Level(gp, small["layout"], small["domain_lo"], small["domain_hi"], s["bc"],
                      s["bc"])
        try:
            with pytest.raises(Exception):  # the buffer is not the level's size
                checkpoint.read_state(path, 0, other)
        finally:
            other.close()
        with pytest.raises(Exception):  # no level 1
            checkpoint.read_layout(path, 1)
        state = os.path.join(path, "Level_0", "State")
        raw = bytearray(open(state, "rb").read())
        raw[len(raw) // 2] ^= 1
        open(state, "wb").write(bytes(raw))
        with pytest.raises(Exception):  # checksum
            checkpoint.read_state(path, 0, lev)
        with pytest.raises(Exception):  # not a checkpoint
            checkpoint.read_info(str(tmp_path))
    finally:
        lev.close()


def test_two_level_restart_continues_bit_for_bit(tmp_path):
    s = amr.refine_setup(quadrant, 48, 16, (12, 8), (35, 31), fine_max_grid_size=16)
    gp = gparams(s)
    args = (gp, s["layout"], s["fine_layout"], s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    announce("checkpoint: two levels, 6 coarse steps straight through")
    h = amr.AmrHierarchy(*args)
    try:
        h.set_state(s["U"], s["Uf"])
        t_ref, n_ref = h.run(s["dx"], 0.8, max_step=6)
        Uc_ref, Uf_ref = h.get_state()
    finally:
        h.close()
    announce("checkpoint: two levels, 3 + checkpoint + restart + 3")
    h = amr.AmrHierarchy(*args)
    try:
        h.set_state(s["U"], s["Uf"])
        t3, n3 = h.run(s["dx"], 0.8, max_step=3)
        checkpoint.write(str(tmp_path / "chk"), [h.coarse, h.fine], t3, s["dx"],
                         steps=[n3, 2 * n3], dts=[0.0, 0.0])
    finally:
        h.close()
    lc = checkpoint.read_layout(str(tmp_path / "chk"), 0)
    lf = checkpoint.read_layout(str(tmp_path / "chk"), 1)
    run, levels = checkpoint.read_info(str(tmp_path / "chk"))
    assert [l["level_steps"] for l in levels] == [3, 6]
    h = amr.AmrHierarchy(gp, lc, lf, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        checkpoint.read_state(str(tmp_path / "chk"), 0, h.coarse)
        checkpoint.read_state(str(tmp_path / "chk"), 1, h.fine)
        t, n = h.run(s["dx"], 0.8, max_step=3, start_time=run["time"])
        Uc, Uf = h.get_state()
    finally:
        h.close()
    assert (t, n3 + n) == (t_ref, n_ref)
    assert_equal_valid(lc, Uc, Uc_ref, "restart coarse")
    assert_equal_valid(lf, Uf, Uf_ref, "restart fine")


@pytest.mark.skipif(not os.path.exists(HOST), reason="C++ host not built")
def test_host_checkpoints_and_restart(tmp_path):
    """amr.check_int writes chk00000, chk00004, chk00008; a run restarted
    from chk00004 writes a chk00008 whose state is byte-identical."""
    inputs = os.path.join(INPUTS, "quadrant-4096")
    common = ["amr.n_cell=128 128", "amr.max_grid_size=32", "max_step=8", "amr.plot_int=-1"]
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    announce("checkpoint: host run with amr.check_int=4")
    r = subprocess.run([HOST, inputs] + common + ["amr.check_int=4", "amr.check_file=" + str(a / "chk")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert sorted(p.name for p in a.iterdir()) == ["chk00000", "chk00004", "chk00008"]
    announce("checkpoint: host restart from chk00004")
    r = subprocess.run([HOST, inputs] + common + ["amr.restart=" + str(a / "chk00004"),
                                                  "amr.check_file=" + str(b / "chk")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Restarted from" in r.stdout
    assert sorted(p.name for p in b.iterdir()) == ["chk00008"]
    sa = open(a / "chk00008" / "Level_0" / "State", "rb").read()
    sb = open(b / "chk00008" / "Level_0" / "State", "rb").read()
    assert sa == sb
    ra, _ = checkpoint.read_info(str(a / "chk00008"))
    rb, _ = checkpoint.read_info(str(b / "chk00008"))
    assert ra["time"] == rb["time"]
