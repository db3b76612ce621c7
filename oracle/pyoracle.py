"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline.  The product
(bamrex_amd) never imports it.

Build with `make -C oracle` (done by __graft_entry__.build()).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")

BC = {"transmissive": 0, "periodic": 1, "reflective": 2}
LIMITER = {"minbee": 0, "van_leer": 1, "van_albada": 2, "superbee": 3}
METHOD = {"muscl_hancock": 0, "hllc": 1, "slic": 2, "force": 3, "lax_friedrichs": 4}


class or_params(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("num_method", ctypes.c_int),
        ("adiabatic", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("limiter", ctypes.c_int),
        ("limiter_indices", ctypes.c_int * 5),
        ("limit_on_primv", ctypes.c_int * 5),
    ]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}

D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)
I64 = ctypes.POINTER(ctypes.c_int64)
PP = ctypes.POINTER(or_params)
P_ = ctypes.c_void_p
_SIGS = {
    "or_run": [PP, I, ctypes.c_int, ctypes.c_int, I, I, I, I, D, ctypes.c_double, ctypes.c_double,
               ctypes.c_double, ctypes.c_int, D, I64, D, I],
    "or_sweep": [PP, I, ctypes.c_int, ctypes.c_int, I, I, I, I, ctypes.c_int, ctypes.c_double, D, D,
                 D, I64],
    "or_fill_boundary": [PP, I, ctypes.c_int, ctypes.c_int, I, I, I, I, D, I64],
    "or_max_wave_speed": [PP, I, ctypes.c_int, ctypes.c_int, D, I64, D],
    "or_est_dt": [PP, I, ctypes.c_int, ctypes.c_int, D, I64, D, ctypes.c_double, ctypes.c_double,
                  ctypes.c_double, D],
    "or_flux_box": [PP, I, ctypes.c_int, ctypes.c_int, ctypes.c_double, D, D, D],
    "or_wave_speed_toro": [ctypes.c_int, D, D, ctypes.c_double, ctypes.c_double, D],
    "or_consv_from_primv": [ctypes.c_int, D, ctypes.c_double, ctypes.c_double, D],
    "or_primv_from_consv": [ctypes.c_int, D, ctypes.c_double, ctypes.c_double, D],
    "or_hllc_face": [ctypes.c_int, ctypes.c_int, D, D, ctypes.c_double, ctypes.c_double, D],
    "or_amr_run": [PP, I, ctypes.c_int, I, ctypes.c_int, ctypes.c_int, I, I, I, I, D, ctypes.c_double,
                   ctypes.c_double, ctypes.c_double, ctypes.c_int, D, I64, D, I64, ctypes.c_int, D, I],
    "or_amr_run_levels": [PP, ctypes.c_int, P_, I, ctypes.c_int, I, I, I, I, D, ctypes.c_double,
                          ctypes.c_double, ctypes.c_double, ctypes.c_int, P_, P_, ctypes.c_int, D, I],
    "or_amr_fill_fine": [PP, I, ctypes.c_int, I, ctypes.c_int, ctypes.c_int, I, I, I, I, D, D, I64, D,
                         I64, ctypes.c_double, ctypes.c_double, ctypes.c_double],
}


def _load(bxpow):
    key = bool(bxpow)
    if key in _libs:
        return _libs[key]
    path = os.path.join(BUILD, "liboracle_bxpow.so" if bxpow else "liboracle.so")
    if not os.path.exists(path):
        build()
    l = ctypes.CDLL(path)
    for n, a in _SIGS.items():
        getattr(l, n).argtypes = a
        getattr(l, n).restype = ctypes.c_int
    l.or_last_run_seconds.argtypes = []
    l.or_last_run_seconds.restype = ctypes.c_double
    l.or_threads.argtypes = [ctypes.c_int]
    l.or_threads.restype = ctypes.c_int
    for n in ("or_limiter",):
        getattr(l, n).argtypes = [ctypes.c_int, ctypes.c_double]
        getattr(l, n).restype = ctypes.c_double
    for n in ("or_minmod", "or_pow_fn", "or_pow_negz_fn", "or_pow_invz_fn"):
        getattr(l, n).argtypes = [ctypes.c_double, ctypes.c_double]
        getattr(l, n).restype = ctypes.c_double
    _libs[key] = l
    return l


def _dp(a):
    return a.ctypes.data_as(D)


def _iarr(v, n=3):
    v = list(v) + [0] * (n - len(v))
    return (ctypes.c_int * n)(*v)


def params(dim, adiabatic, epsilon=1.0, num_method="muscl_hancock", limiter="van_leer",
           limiter_indices=None, limit_on_primv=None):
    p = or_params()
    p.dim = dim
    p.num_method = METHOD[num_method] if isinstance(num_method, str) else int(num_method)
    p.adiabatic = adiabatic
    p.epsilon = epsilon
    p.limiter = LIMITER[limiter] if isinstance(limiter, str) else int(limiter)
    li = limiter_indices if limiter_indices is not None else [dim + 1] * (dim + 2)
    lp = limit_on_primv if limit_on_primv is not None else [0] * (dim + 2)
    for n in range(5):
        p.limiter_indices[n] = li[n] if n < len(li) else 0
        p.limit_on_primv[n] = int(lp[n]) if n < len(lp) else 0
    return p


class Oracle:
    """CPU restatement of the reference hot path on a packed level."""

    def __init__(self, bxpow=False):
        self.lib = _load(bxpow)

    def last_run_seconds(self):
        """Wall time of the step loop of the last run() (no level set-up or
        state copies)."""
        return self.lib.or_last_run_seconds()

    def threads(self, n=0):
        """Set the OpenMP threads of the box loop (n > 0); returns the count."""
        return self.lib.or_threads(n)

    @staticmethod
    def _geom(layout, domain_lo, domain_hi, bc_lo, bc_hi):
        dim = layout.dim
        boxes = []
        for lo, hi in layout.boxes:
            boxes += list(lo) + [0] * (3 - dim) + list(hi) + [0] * (3 - dim)
        barr = (ctypes.c_int * len(boxes))(*boxes)
        bl = [BC[b] if isinstance(b, str) else b for b in bc_lo]
        bh = [BC[b] if isinstance(b, str) else b for b in bc_hi]
        offs = np.ascontiguousarray(layout.offsets, dtype=np.int64)
        return barr, _iarr(domain_lo), _iarr(domain_hi), _iarr(bl), _iarr(bh), offs

    def run(self, p, layout, U, domain_lo, domain_hi, bc_lo, bc_hi, dx, cfl, stop_time=-1.0,
            max_step=-1, start_time=0.0):
        barr, dlo, dhi, bl, bh, offs = self._geom(layout, domain_lo, domain_hi, bc_lo, bc_hi)
        t = ctypes.c_double()
        n = ctypes.c_int()
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        self.lib.or_run(ctypes.byref(p), barr, len(layout.boxes), layout.ng, dlo, dhi, bl, bh,
                        _dp(dxa), cfl, start_time, stop_time, max_step, _dp(U),
                        offs.ctypes.data_as(I64), ctypes.byref(t), ctypes.byref(n))
        return t.value, n.value

    def amr_run(self, p, clayout, Uc, flayout, Uf, domain_lo, domain_hi, bc_lo, bc_hi, dx, cfl,
                stop_time=-1.0, max_step=-1, start_time=0.0, do_reflux=True):
        """Two levels, ratio 2 (or_amr_run); Uc, Uf updated in place."""
        cb, dlo, dhi, bl, bh, coffs = self._geom(clayout, domain_lo, domain_hi, bc_lo, bc_hi)
        fb, *_, foffs = self._geom(flayout, domain_lo, domain_hi, bc_lo, bc_hi)
        t = ctypes.c_double()
        n = ctypes.c_int()
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        self.lib.or_amr_run(ctypes.byref(p), cb, len(clayout.boxes), fb, len(flayout.boxes),
                            clayout.ng, dlo, dhi, bl, bh, _dp(dxa), cfl, start_time, stop_time,
                            max_step, _dp(Uc), coffs.ctypes.data_as(I64), _dp(Uf),
                            foffs.ctypes.data_as(I64), int(bool(do_reflux)), ctypes.byref(t),
                            ctypes.byref(n))
        return t.value, n.value

    def amr_run_levels(self, p, layouts, Us, domain_lo, domain_hi, bc_lo, bc_hi, dx, cfl,
                       stop_time=-1.0, max_step=-1, start_time=0.0, do_reflux=True):
        """Any number of levels, ratio 2 (or_amr_run_levels); each U updated
        in place.  domain_lo/hi and dx are level 0's."""
        n = len(layouts)
        geo = [self._geom(lay, domain_lo, domain_hi, bc_lo, bc_hi) for lay in layouts]
        barrs = (ctypes.c_void_p * n)(*[ctypes.cast(g[0], ctypes.c_void_p) for g in geo])
        nbox = (ctypes.c_int * n)(*[len(lay.boxes) for lay in layouts])
        Ups = (ctypes.c_void_p * n)(*[U.ctypes.data for U in Us])
        offs = [g[5] for g in geo]
        Ops = (ctypes.c_void_p * n)(*[o.ctypes.data for o in offs])
        _, dlo, dhi, bl, bh, _ = geo[0]
        t = ctypes.c_double()
        k = ctypes.c_int()
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        rc = self.lib.or_amr_run_levels(ctypes.byref(p), n, ctypes.cast(barrs, ctypes.c_void_p),
                                        nbox, layouts[0].ng, dlo, dhi, bl, bh, _dp(dxa), cfl,
                                        start_time, stop_time, max_step,
                                        ctypes.cast(Ups, ctypes.c_void_p),
                                        ctypes.cast(Ops, ctypes.c_void_p), int(bool(do_reflux)),
                                        ctypes.byref(t), ctypes.byref(k))
        if rc != 0:
            raise ValueError("or_amr_run_levels: levels not properly nested")
        return t.value, k.value

    def amr_fill_fine(self, p, clayout, Uc_old, Uc_new, flayout, Uf, domain_lo, domain_hi, bc_lo,
                      bc_hi, t_old, t_new, time):
        cb, dlo, dhi, bl, bh, coffs = self._geom(clayout, domain_lo, domain_hi, bc_lo, bc_hi)
        fb, *_, foffs = self._geom(flayout, domain_lo, domain_hi, bc_lo, bc_hi)
        self.lib.or_amr_fill_fine(ctypes.byref(p), cb, len(clayout.boxes), fb, len(flayout.boxes),
                                  clayout.ng, dlo, dhi, bl, bh, _dp(Uc_old), _dp(Uc_new),
                                  coffs.ctypes.data_as(I64), _dp(Uf), foffs.ctypes.data_as(I64),
                                  t_old, t_new, time)

    def sweep(self, p, layout, Uin, Uout, dir, dt, dx, domain_lo, domain_hi, bc_lo, bc_hi):
        barr, dlo, dhi, bl, bh, offs = self._geom(layout, domain_lo, domain_hi, bc_lo, bc_hi)
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        self.lib.or_sweep(ctypes.byref(p), barr, len(layout.boxes), layout.ng, dlo, dhi, bl, bh,
                          dir, dt, _dp(dxa), _dp(Uin), _dp(Uout), offs.ctypes.data_as(I64))

    def fill_boundary(self, p, layout, U, domain_lo, domain_hi, bc_lo, bc_hi):
        barr, dlo, dhi, bl, bh, offs = self._geom(layout, domain_lo, domain_hi, bc_lo, bc_hi)
        self.lib.or_fill_boundary(ctypes.byref(p), barr, len(layout.boxes), layout.ng, dlo, dhi,
                                  bl, bh, _dp(U), offs.ctypes.data_as(I64))

    def max_wave_speed(self, p, layout, U):
        barr, *_, offs = self._geom(layout, (0,), (0,), (0,), (0,))
        out = ctypes.c_double()
        self.lib.or_max_wave_speed(ctypes.byref(p), barr, len(layout.boxes), layout.ng, _dp(U),
                                   offs.ctypes.data_as(I64), ctypes.byref(out))
        return out.value

    def est_dt(self, p, layout, U, dx, cfl, time=0.0, stop_time=-1.0):
        barr, *_, offs = self._geom(layout, (0,), (0,), (0,), (0,))
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        out = ctypes.c_double()
        self.lib.or_est_dt(ctypes.byref(p), barr, len(layout.boxes), layout.ng, _dp(U),
                           offs.ctypes.data_as(I64), _dp(dxa), cfl, time, stop_time,
                           ctypes.byref(out))
        return out.value

    def flux_box(self, p, layout, U, b, dir, dt, dx):
        lo, hi = layout.boxes[b]
        dim = layout.dim
        box = (ctypes.c_int * 6)(*(list(lo) + [0] * (3 - dim) + list(hi) + [0] * (3 - dim)))
        fab = np.ascontiguousarray(layout.fab(U, b)).ravel()
        ext = [hi[d] - lo[d] + 1 + (1 if d == dir else 0) for d in range(dim)]
        flux = np.zeros((dim + 2,) + tuple(reversed(ext)), dtype=np.float64)
        dxa = np.array(list(dx) + [0.0] * (3 - len(dx)), dtype=np.float64)
        self.lib.or_flux_box(ctypes.byref(p), box, layout.ng, dir, dt, _dp(dxa), _dp(fab),
                             _dp(flux))
        return flux

    # scalar helpers
    def wave_speed_toro(self, dim, qL, qR, adiabatic, epsilon):
        a = np.array(qL, dtype=np.float64)
        b = np.array(qR, dtype=np.float64)
        out = np.zeros(3)
        self.lib.or_wave_speed_toro(dim, _dp(a), _dp(b), adiabatic, epsilon, _dp(out))
        return out  # S_L, S_R, S_interm

    def consv_from_primv(self, dim, q, adiabatic, epsilon):
        a = np.array(q, dtype=np.float64)
        out = np.zeros(dim + 2)
        self.lib.or_consv_from_primv(dim, _dp(a), adiabatic, epsilon, _dp(out))
        return out

    def primv_from_consv(self, dim, u, adiabatic, epsilon):
        a = np.array(u, dtype=np.float64)
        out = np.zeros(dim + 2)
        self.lib.or_primv_from_consv(dim, _dp(a), adiabatic, epsilon, _dp(out))
        return out

    def hllc_face(self, dim, dir, UL, UR, adiabatic, epsilon):
        a = np.array(UL, dtype=np.float64)
        b = np.array(UR, dtype=np.float64)
        out = np.zeros(dim + 2)
        self.lib.or_hllc_face(dim, dir, _dp(a), _dp(b), adiabatic, epsilon, _dp(out))
        return out

    def limiter(self, kind, r):
        return self.lib.or_limiter(LIMITER[kind] if isinstance(kind, str) else kind, r)

    def minmod(self, a, b):
        return self.lib.or_minmod(a, b)

    def pow(self, x, y):
        return self.lib.or_pow_fn(x, y)

    def pow_negz(self, x, gamma):
        """pow(x, -z), z = (gamma-1)/(2 gamma), as the HLLC estimate calls it."""
        return self.lib.or_pow_negz_fn(x, gamma)

    def pow_invz(self, x, gamma):
        """pow(x, 1/z) as the HLLC estimate calls it."""
        return self.lib.or_pow_invz_fn(x, gamma)
