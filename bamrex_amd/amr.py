"""Two AMR levels of the explicit hot path on the GPU (bx_amr_*).

Host-side mirror of what the reference's AmrLevelAdv gets from AMReX for a
coarse level and one fine level: subcycled advance (Amr::timeStep), the
coarse-fine FillPatch, flux registers, reflux and avgDown (post_timestep,
Source/AmrLevelAdv.cpp:628-652, 918-987), computeNewDt over the levels
(:589-618).  Refinement ratio 2; the fine boxes are fixed (no regridding).
"""
import ctypes

import numpy as np

from . import _lib as L
from .level import DeviceBuffer, Layout, Level, chop_domain, check


def refine_setup(make, ncell, max_grid_size, region_lo, region_hi, fine_max_grid_size=None,
                 **kw):
    """Two-level initial data: make(ncell=..., max_grid_size=..., **kw) on
    the coarse level, and the same problem at twice the resolution on fine
    boxes covering coarse cells region_lo..region_hi (inclusive)."""
    s = make(ncell=ncell, max_grid_size=max_grid_size, **kw)
    sf = make(ncell=2 * ncell, max_grid_size=2 * ncell, **kw)
    dim = s["layout"].dim
    flo = tuple(2 * x for x in region_lo)
    fhi = tuple(2 * x + 1 for x in region_hi)
    fboxes = chop_domain(flo, fhi, fine_max_grid_size or max_grid_size)
    flay = Layout(dim, fboxes, s["layout"].ng)
    Uf = flay.new_array()
    full = sf["layout"].valid(sf["U"], 0)
    for b, (lo, hi) in enumerate(fboxes):
        sl = tuple(slice(lo[d], hi[d] + 1) for d in reversed(range(dim)))
        flay.valid(Uf, b)[...] = full[(slice(None),) + sl]
    s.update(fine_layout=flay, Uf=Uf, fine_lo=flo, fine_hi=fhi)
    return s


class AmrHierarchy:
    """A coarse Level and a fine Level (ratio 2) coupled by bx_amr."""

    def __init__(self, params, coarse_layout, fine_layout, domain_lo, domain_hi, bc_lo=None,
                 bc_hi=None):
        self.dim = coarse_layout.dim
        self.coarse = Level(params, coarse_layout, domain_lo, domain_hi, bc_lo, bc_hi)
        self.fine = Level(params, fine_layout, [2 * x for x in domain_lo],
                          [2 * x + 1 for x in domain_hi], bc_lo, bc_hi)
        self.Uc_old = DeviceBuffer(coarse_layout.total * 8)
        self.a = ctypes.c_void_p()
        self.stream = self.coarse.stream
        try:
            check(L.lib().bx_amr_create(self.coarse.ctx, self.fine.ctx, ctypes.byref(self.a)),
                  "bx_amr_create")
        except Exception:
            self.close()
            raise

    def info(self):
        n = [ctypes.c_int() for _ in range(3)]
        check(L.lib().bx_amr_info(self.a, *[ctypes.byref(x) for x in n]), "bx_amr_info")
        return dict(cf_segments=n[0].value, register_faces=n[1].value, reflux_cells=n[2].value)

    def set_state(self, Uc, Uf):
        self.coarse.set_state(Uc)
        self.fine.set_state(Uf)

    def get_state(self):
        return self.coarse.get_state(), self.fine.get_state()

    def avg_down(self):
        check(L.lib().bx_amr_avg_down(self.a, self.fine.U.ptr, self.coarse.U.ptr, self.stream),
              "bx_amr_avg_down")
        L.lib().bx_stream_sync(self.stream)

    def fill_fine(self, t_old, t_new, time):
        """Fine ghost cells at `time` from the coarse states held in Uc_old
        (at t_old) and in the coarse level's U (at t_new)."""
        check(L.lib().bx_amr_fill_fine(self.a, self.fine.U.ptr, self.Uc_old.ptr,
                                       self.coarse.U.ptr, t_old, t_new, time, self.stream),
              "bx_amr_fill_fine")
        L.lib().bx_stream_sync(self.stream)

    def set_coarse_old(self, U_host):
        self.Uc_old.upload(U_host)

    def est_dt(self, dx_coarse, cfl, time=0.0, stop_time=-1.0):
        out = ctypes.c_double()
        check(L.lib().bx_amr_est_dt(self.a, L.dvec3(dx_coarse), cfl, time, stop_time,
                                    self.coarse.U.ptr, self.fine.U.ptr, ctypes.byref(out),
                                    self.stream), "bx_amr_est_dt")
        return out.value

    def advance(self, dt, dx_coarse, time):
        check(L.lib().bx_amr_advance(self.a, dt, L.dvec3(dx_coarse), time, self.coarse.U.ptr,
                                     self.coarse.U_tmp.ptr, self.Uc_old.ptr, self.fine.U.ptr,
                                     self.fine.U_tmp.ptr, self.stream), "bx_amr_advance")
        L.lib().bx_stream_sync(self.stream)

    def set_graph(self, use_graph):
        """Runs of a fixed number of steps replay one captured coarse step
        (default) or launch every kernel from the host (same bits)."""
        check(L.lib().bx_amr_set_graph(self.a, int(bool(use_graph))), "bx_amr_set_graph")

    def run(self, dx_coarse, cfl, stop_time=-1.0, max_step=-1, start_time=0.0):
        t = ctypes.c_double(start_time)
        n = ctypes.c_int()
        check(L.lib().bx_amr_run(self.a, L.dvec3(dx_coarse), cfl, stop_time, max_step,
                                 self.coarse.U.ptr, self.coarse.U_tmp.ptr, self.Uc_old.ptr,
                                 self.fine.U.ptr, self.fine.U_tmp.ptr, ctypes.byref(t),
                                 ctypes.byref(n), self.stream), "bx_amr_run")
        return t.value, n.value

    def write_plotfile(self, path, time, dx_coarse, prob_lo, prob_hi, steps=0, derived=False):
        """Both levels as one AMReX plotfile (AmrLevel::writePlotFile)."""
        from .plotfile import write_plotfile
        Uc, Uf = self.get_state()
        g, gf = self.coarse.geom, self.fine.geom
        dim = self.dim
        levels = [dict(layout=self.coarse.layout, U=Uc, domain_lo=list(g.domain_lo)[:dim],
                       domain_hi=list(g.domain_hi)[:dim], dx=tuple(dx_coarse), steps=steps),
                  dict(layout=self.fine.layout, U=Uf, domain_lo=list(gf.domain_lo)[:dim],
                       domain_hi=list(gf.domain_hi)[:dim], dx=tuple(x / 2 for x in dx_coarse),
                       steps=2 * steps)]
        p = self.coarse.params
        write_plotfile(path, levels, prob_lo, prob_hi, time, p.adiabatic, p.epsilon, derived=derived)

    def close(self):
        if self.a:
            L.lib().bx_amr_destroy(self.a)
            self.a = None
        self.Uc_old.free()
        self.coarse.close()
        self.fine.close()


class AmrLevels:
    """Any number of levels (ratio 2 between neighbours) coupled by bx_amr:
    layouts[l] in level-l index space; domain and dx0 are level 0's."""

    def __init__(self, params, layouts, domain_lo, domain_hi, bc_lo=None, bc_hi=None):
        self.dim = layouts[0].dim
        self.levels = []
        self.olds = []
        self.a = ctypes.c_void_p()
        try:
            for l, lay in enumerate(layouts):
                self.levels.append(Level(params, lay, [x << l for x in domain_lo],
                                         [((x + 1) << l) - 1 for x in domain_hi], bc_lo, bc_hi))
                self.olds.append(DeviceBuffer(lay.total * 8))
            self.stream = self.levels[0].stream
            ctxs = (ctypes.c_void_p * len(self.levels))(*[lv.ctx.value for lv in self.levels])
            check(L.lib().bx_amr_create_levels(ctypes.cast(ctxs, ctypes.c_void_p), len(ctxs),
                                               ctypes.byref(self.a)), "bx_amr_create_levels")
        except Exception:
            self.close()
            raise

    def _ptrs(self, which):
        bufs = {"U": [lv.U for lv in self.levels], "T": [lv.U_tmp for lv in self.levels],
                "O": self.olds}[which]
        arr = (ctypes.c_void_p * len(bufs))(*[b.ptr.value for b in bufs])
        return ctypes.cast(arr, ctypes.c_void_p), arr

    def set_state(self, Us):
        for lv, U in zip(self.levels, Us):
            lv.set_state(U)

    def get_state(self):
        return [lv.get_state() for lv in self.levels]

    def set_graph(self, use_graph):
        """Runs of a fixed number of steps replay one captured coarse step
        (default) or launch every kernel from the host (same bits)."""
        check(L.lib().bx_amr_set_graph(self.a, int(bool(use_graph))), "bx_amr_set_graph")

    def run(self, dx0, cfl, stop_time=-1.0, max_step=-1, start_time=0.0):
        t = ctypes.c_double(start_time)
        n = ctypes.c_int()
        (pu, ku), (pt, kt), (po, ko) = self._ptrs("U"), self._ptrs("T"), self._ptrs("O")
        check(L.lib().bx_amr_run_levels(self.a, L.dvec3(dx0), cfl, stop_time, max_step, pu, pt, po,
                                        ctypes.byref(t), ctypes.byref(n), self.stream),
              "bx_amr_run_levels")
        return t.value, n.value

    def write_plotfile(self, path, time, dx0, prob_lo, prob_hi, steps=0, derived=False):
        """Every level in one AMReX plotfile (AmrLevel::writePlotFile)."""
        from .plotfile import write_plotfile
        dim = self.dim
        levels = []
        for l, (lv, U) in enumerate(zip(self.levels, self.get_state())):
            g = lv.geom
            levels.append(dict(layout=lv.layout, U=U, domain_lo=list(g.domain_lo)[:dim],
                               domain_hi=list(g.domain_hi)[:dim],
                               dx=tuple(x / 2 ** l for x in dx0), steps=steps << l))
        p = self.levels[0].params
        write_plotfile(path, levels, prob_lo, prob_hi, time, p.adiabatic, p.epsilon, derived=derived)

    def close(self):
        if self.a:
            L.lib().bx_amr_destroy(self.a)
            self.a = None
        for b in self.olds:
            b.free()
        for lv in self.levels:
            lv.close()
        self.olds, self.levels = [], []


def refine_levels(make, ncell, max_grid_size, regions, max_grid_sizes=None, **kw):
    """Initial data of a hierarchy: make(...) on level 0 and, for each
    region (lo, hi in level-l index space, inclusive) of levels 1.., the
    same problem at 2^l times the resolution on boxes covering it."""
    s = make(ncell=ncell, max_grid_size=max_grid_size, **kw)
    dim = s["layout"].dim
    layouts, Us = [s["layout"]], [s["U"]]
    for l, (lo, hi) in enumerate(regions, start=1):
        sf = make(ncell=ncell << l, max_grid_size=ncell << l, **kw)
        mgs = (max_grid_sizes or [max_grid_size] * len(regions))[l - 1]
        boxes = chop_domain(tuple(lo), tuple(hi), mgs)
        lay = Layout(dim, boxes, s["layout"].ng)
        U = lay.new_array()
        full = sf["layout"].valid(sf["U"], 0)
        for b, (blo, bhi) in enumerate(boxes):
            sl = tuple(slice(blo[d], bhi[d] + 1) for d in reversed(range(dim)))
            lay.valid(U, b)[...] = full[(slice(None),) + sl]
        layouts.append(lay)
        Us.append(U)
    s.update(layouts=layouts, Us=Us)
    return s


def composite_totals(coarse_layout, Uc, fine_layout, Uf, dx_coarse):
    """Sum over the composite grid of U * cell volume (coarse cells not
    covered by the fine level, and fine cells), per component."""
    dim = coarse_layout.dim
    volc = float(np.prod(dx_coarse))
    volf = volc / 2 ** dim
    tot = np.zeros(dim + 2)
    fine_boxes = fine_layout.boxes
    for b, (lo, hi) in enumerate(coarse_layout.boxes):
        v = coarse_layout.valid(Uc, b).copy()
        mask = np.ones(v.shape[1:], dtype=bool)
        for flo, fhi in fine_boxes:
            clo = [max(x // 2, l) for x, l in zip(flo, lo)]
            chi = [min(x // 2, h) for x, h in zip(fhi, hi)]
            if all(a <= c for a, c in zip(clo, chi)):
                sl = tuple(slice(clo[d] - lo[d], chi[d] - lo[d] + 1) for d in reversed(range(dim)))
                mask[sl] = False
        tot += volc * np.array([v[n][mask].sum() for n in range(dim + 2)])
    for b in range(len(fine_boxes)):
        v = fine_layout.valid(Uf, b)
        tot += volf * np.array([v[n].sum() for n in range(dim + 2)])
    return tot


def composite_totals_levels(layouts, Us, dx0):
    """Sum over the composite grid of U * cell volume: each level's cells
    not covered by the next finer level, per component."""
    dim = layouts[0].dim
    tot = np.zeros(dim + 2)
    for l, (lay, U) in enumerate(zip(layouts, Us)):
        vol = float(np.prod(dx0)) / 2 ** (dim * l)
        finer = layouts[l + 1].boxes if l + 1 < len(layouts) else []
        for b, (lo, hi) in enumerate(lay.boxes):
            v = lay.valid(U, b)
            mask = np.ones(v.shape[1:], dtype=bool)
            for flo, fhi in finer:
                clo = [max(x // 2, q) for x, q in zip(flo, lo)]
                chi = [min(x // 2, q) for x, q in zip(fhi, hi)]
                if all(a <= c for a, c in zip(clo, chi)):
                    sl = tuple(slice(clo[d] - lo[d], chi[d] - lo[d] + 1) for d in reversed(range(dim)))
                    mask[sl] = False
            tot += vol * np.array([v[n][mask].sum() for n in range(dim + 2)])
    return tot
