"""Checkpoint / restart of device-resident levels (bx_write_checkpoint,
bx_read_checkpoint_*).

Host-side mirror of AmrLevelAdv::checkPoint / restart (Source/AmrLevelAdv.cpp:
95-113) and the final checkpoint of Source/main.cpp:97-100: the time, each
level's step count, dt, box list and state (ghost cells included), so a run
restarted from it continues bit for bit.
"""
import ctypes

from . import _lib as L
from .level import Layout, check


def write(path, levels, time, dx0, prob_lo=(0.0, 0.0, 0.0), prob_hi=(1.0, 1.0, 1.0),
          steps=None, dts=None, ref_ratio=2):
    """Write Levels (finest last) to directory `path`.  Level l's cell size
    is dx0 / ref_ratio^l; steps / dts: per-level step counts and dt."""
    nlev = len(levels)
    arr = (L.bx_chk_level * nlev)()
    for l, lev in enumerate(levels):
        c = arr[l]
        c.boxes = ctypes.cast(lev.layout.c_boxes, ctypes.c_void_p)
        c.nbox = len(lev.layout.boxes)
        c.ng = lev.layout.ng
        c.U = lev.U.ptr
        for d in range(3):
            c.domain_lo[d] = lev.geom.domain_lo[d]
            c.domain_hi[d] = lev.geom.domain_hi[d]
            c.dx[d] = (dx0[d] if d < len(dx0) else 1.0) / ref_ratio ** l
        c.level_steps = int(steps[l]) if steps is not None else 0
        c.dt = float(dts[l]) if dts is not None else 0.0
    check(L.lib().bx_write_checkpoint(path.encode(), levels[0].dim, nlev, arr, ref_ratio,
                                      L.dvec3(prob_lo), L.dvec3(prob_hi), float(time),
                                      levels[0].stream), "bx_write_checkpoint")


def read_info(path):
    """(run info dict, per-level info dicts) from the checkpoint header."""
    info = L.bx_chk_info()
    lv = (L.bx_chk_level_info * 16)()
    check(L.lib().bx_read_checkpoint_info(path.encode(), ctypes.byref(info), lv, 16),
          "bx_read_checkpoint_info")
    run = {"dim": info.dim, "nlev": info.nlev, "ref_ratio": info.ref_ratio, "time": info.time,
           "prob_lo": list(info.prob_lo), "prob_hi": list(info.prob_hi)}
    levels = []
    for l in range(info.nlev):
        x = lv[l]
        levels.append({"nbox": x.nbox, "ng": x.ng, "level_steps": x.level_steps, "dt": x.dt,
                       "domain_lo": list(x.domain_lo), "domain_hi": list(x.domain_hi),
                       "dx": list(x.dx), "ndoubles": x.ndoubles, "checksum": x.checksum})
    return run, levels


def read_layout(path, level):
    """The packed Layout of level `level` as the checkpoint stored it."""
    run, levels = read_info(path)
    li = levels[level]
    boxes = (L.bx_box * li["nbox"])()
    check(L.lib().bx_read_checkpoint_boxes(path.encode(), level, boxes, li["nbox"]),
          "bx_read_checkpoint_boxes")
    dim = run["dim"]
    return Layout(dim, [(tuple(b.lo[:dim]), tuple(b.hi[:dim])) for b in boxes], li["ng"])


def read_state(path, level, lev):
    """Level `level`'s state into Level lev's U (same packed layout)."""
    check(L.lib().bx_read_checkpoint_level(path.encode(), level, lev.U.ptr, lev.layout.total,
                                           lev.stream), "bx_read_checkpoint_level")
