"""Host-side mirror of the reference's level integrator over the C ABI.

`Level` plays the part of AmrLevelAdv + its MultiFab on one level: it owns the
box list (BoxArray), the packed FAB layout (Array4 order, component-major, x
fastest, ng ghost cells), two device state buffers (Sborder / Sborder2 of
AmrLevelAdv_advance.cpp:103-104) and a bx_ctx.  Method names follow the
reference: compute_MUSCL_hancock_flux, advance, estTimeStep, ...
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import BC, LIMITER, METHOD, check


def chop_domain(domain_lo, domain_hi, max_grid_size):
    """BoxArray(domain).maxSize(max_grid_size): chop the domain box into
    boxes of at most max_grid_size cells per direction (x fastest)."""
    dim = len(domain_lo)
    ranges = []
    for d in range(dim):
        r = []
        a = domain_lo[d]
        while a <= domain_hi[d]:
            b = min(a + max_grid_size - 1, domain_hi[d])
            r.append((a, b))
            a = b + 1
        ranges.append(r)
    boxes = []
    if dim == 1:
        for (a, b) in ranges[0]:
            boxes.append(((a,), (b,)))
    elif dim == 2:
        for (ya, yb) in ranges[1]:
            for (xa, xb) in ranges[0]:
                boxes.append(((xa, ya), (xb, yb)))
    else:
        for (za, zb) in ranges[2]:
            for (ya, yb) in ranges[1]:
                for (xa, xb) in ranges[0]:
                    boxes.append(((xa, ya, za), (xb, yb, zb)))
    return boxes


def default_limiting(dim):
    """FluidSystem::set_parm_defaults (FluidSystem.cpp:52-71): every
    component limited on the conserved energy (index dim+1)."""
    return [dim + 1] * (dim + 2), [0] * (dim + 2)


def make_params(dim, adiabatic, epsilon=1.0, num_method="muscl_hancock",
                limiter="van_leer", limiter_indices=None, limit_on_primv=None):
    p = L.bx_params()
    p.dim = dim
    p.num_method = METHOD[num_method] if isinstance(num_method, str) else int(num_method)
    p.adiabatic = float(adiabatic)
    p.epsilon = float(epsilon)
    p.limiter = LIMITER[limiter] if isinstance(limiter, str) else int(limiter)
    li, lp = default_limiting(dim)
    if limiter_indices is not None:
        li = list(limiter_indices)
    if limit_on_primv is not None:
        lp = [int(bool(x)) for x in limit_on_primv]
    for n in range(5):
        p.limiter_indices[n] = li[n] if n < len(li) else 0
        p.limit_on_primv[n] = lp[n] if n < len(lp) else 0
    return p


class Layout:
    """Packed FAB layout of a box list (shared by the device library, the
    oracle and the tests)."""

    def __init__(self, dim, boxes, ng):
        self.dim = dim
        self.ng = ng
        self.nc = dim + 2
        self.boxes = [(tuple(lo), tuple(hi)) for lo, hi in boxes]
        arr = (L.bx_box * len(boxes))()
        for b, (lo, hi) in enumerate(self.boxes):
            for d in range(3):
                arr[b].lo[d] = lo[d] if d < dim else 0
                arr[b].hi[d] = hi[d] if d < dim else 0
            arr[b].offset = 0
        offs = (ctypes.c_int64 * len(boxes))()
        total = ctypes.c_int64()
        check(L.lib().bx_pack_offsets(dim, ng, arr, len(boxes), offs, ctypes.byref(total)),
              "bx_pack_offsets")
        for b in range(len(boxes)):
            arr[b].offset = offs[b]
        self.c_boxes = arr
        self.offsets = np.array(list(offs), dtype=np.int64)
        self.total = int(total.value)

    def fab_shape(self, b):
        lo, hi = self.boxes[b]
        ext = [hi[d] - lo[d] + 1 + 2 * self.ng for d in range(self.dim)]
        return (self.nc,) + tuple(reversed(ext))  # (n, [z,] [y,] x)

    def fab(self, U, b):
        """View of box b's FAB inside a packed array, shaped (n, [z, y,] x)."""
        shp = self.fab_shape(b)
        n = int(np.prod(shp))
        o = int(self.offsets[b])
        return U[o:o + n].reshape(shp)

    def valid(self, U, b):
        g = self.ng
        f = self.fab(U, b)
        sl = (slice(None),) + tuple(slice(g, s - g) for s in f.shape[1:])
        return f[sl]

    def cell_centers(self, b, prob_lo, dx):
        """Cell-centre coordinates (x, y, z arrays) of the VALID cells of box b,
        prob_lo[d] + (i + 0.5) * dx[d] (as every Prob.cpp computes them)."""
        lo, hi = self.boxes[b]
        axes = []
        for d in range(self.dim):
            idx = np.arange(lo[d], hi[d] + 1, dtype=np.float64)
            axes.append(prob_lo[d] + (idx + 0.5) * dx[d])
        grids = np.meshgrid(*reversed(axes), indexing="ij")
        return list(reversed(grids))

    def ncells(self):
        tot = 0
        for lo, hi in self.boxes:
            v = 1
            for d in range(self.dim):
                v *= hi[d] - lo[d] + 1
            tot += v
        return tot

    def new_array(self):
        return np.zeros(self.total, dtype=np.float64)


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        check(L.lib().bx_malloc(ctypes.byref(self.ptr), nbytes), "bx_malloc")

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        check(L.lib().bx_memcpy_htod(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes), "htod")

    def download(self, arr):
        assert arr.flags["C_CONTIGUOUS"] and arr.nbytes <= self.nbytes
        check(L.lib().bx_memcpy_dtoh(arr.ctypes.data_as(ctypes.c_void_p), self.ptr, arr.nbytes), "dtoh")

    def free(self):
        if self.ptr:
            L.lib().bx_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Level:
    """One AMR level of the explicit hot path on the GPU."""

    def __init__(self, params, layout, domain_lo, domain_hi, bc_lo=None, bc_hi=None,
                 stream=None):
        self.layout = layout
        self.dim = layout.dim
        self.params = params
        lib = L.lib()
        self.ctx = ctypes.c_void_p()
        check(lib.bx_create(ctypes.byref(params), ctypes.byref(self.ctx)), "bx_create")
        g = L.bx_geometry()
        bc_lo = bc_lo or ["transmissive"] * self.dim
        bc_hi = bc_hi or ["transmissive"] * self.dim
        for d in range(3):
            g.domain_lo[d] = domain_lo[d] if d < self.dim else 0
            g.domain_hi[d] = domain_hi[d] if d < self.dim else 0
            g.bc_lo[d] = (BC[bc_lo[d]] if isinstance(bc_lo[d], str) else bc_lo[d]) if d < self.dim else 0
            g.bc_hi[d] = (BC[bc_hi[d]] if isinstance(bc_hi[d], str) else bc_hi[d]) if d < self.dim else 0
        g.ng = layout.ng
        self.geom = g
        check(lib.bx_set_level(self.ctx, ctypes.byref(g), layout.c_boxes, len(layout.boxes)),
              "bx_set_level")
        nbytes = layout.total * 8
        self.U = DeviceBuffer(nbytes)
        self.U_tmp = DeviceBuffer(nbytes)
        self.stream = stream if stream is not None else ctypes.c_void_p()

    # ---------------------------------------------------------- state I/O
    def set_state(self, U_host):
        self.U.upload(U_host)

    def get_state(self, which="U"):
        out = self.layout.new_array()
        (self.U if which == "U" else self.U_tmp).download(out)
        return out

    def set_tmp(self, U_host):
        self.U_tmp.upload(U_host)

    # ---------------------------------------------------- hot-path calls
    def fill_boundary(self, which="U"):
        buf = self.U if which == "U" else self.U_tmp
        check(L.lib().bx_fill_boundary(self.ctx, buf.ptr, self.stream), "bx_fill_boundary")

    def sweep(self, dir, dt, dx):
        """U_tmp <- one MUSCL-Hancock/HLLC sweep of U (valid cells)."""
        check(L.lib().bx_sweep(self.ctx, dir, dt, L.dvec3(dx), self.U.ptr, self.U_tmp.ptr,
                               self.stream), "bx_sweep")

    def max_wave_speed(self):
        out = ctypes.c_double()
        check(L.lib().bx_max_wave_speed(self.ctx, self.U.ptr, ctypes.byref(out), self.stream),
              "bx_max_wave_speed")
        return out.value

    def estTimeStep(self, dx, cfl, time=0.0, stop_time=-1.0):
        out = ctypes.c_double()
        check(L.lib().bx_est_dt(self.ctx, self.U.ptr, L.dvec3(dx), cfl, time, stop_time,
                                ctypes.byref(out), self.stream), "bx_est_dt")
        return out.value

    def advance(self, dt, dx):
        check(L.lib().bx_advance(self.ctx, dt, L.dvec3(dx), self.U.ptr, self.U_tmp.ptr,
                                 self.stream), "bx_advance")

    def run(self, dx, cfl, stop_time=-1.0, max_step=-1, start_time=0.0, use_graph=True):
        t = ctypes.c_double()
        n = ctypes.c_int()
        in_tmp = ctypes.c_int()
        check(L.lib().bx_run(self.ctx, L.dvec3(dx), cfl, start_time, stop_time, max_step,
                             self.U.ptr, self.U_tmp.ptr, 1 if use_graph else 0, self.stream,
                             ctypes.byref(t), ctypes.byref(n), ctypes.byref(in_tmp)), "bx_run")
        return t.value, n.value

    def flux_box(self, ibox, dir, dt, dx, flux_buf, flux_offset=0, which="U"):
        buf = self.U if which == "U" else self.U_tmp
        check(L.lib().bx_flux_box(self.ctx, ibox, dir, dt, L.dvec3(dx), buf.ptr, flux_buf.ptr,
                                  flux_offset, self.stream), "bx_flux_box")

    def update_box(self, ibox, dir, dt, dx, flux_buf, flux_offset=0):
        check(L.lib().bx_update_box(self.ctx, ibox, dir, dt, L.dvec3(dx), self.U.ptr,
                                    flux_buf.ptr, flux_offset, self.U_tmp.ptr, self.stream),
              "bx_update_box")

    def get_tuning(self):
        t = L.bx_tuning()
        check(L.lib().bx_get_tuning(self.ctx, ctypes.byref(t)), "bx_get_tuning")
        return {"block_x": t.block_x, "block_t": t.block_t, "seg_k": t.seg_k, "grid_cap": t.grid_cap}

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        t = L.bx_tuning(cur["block_x"], cur["block_t"], cur["seg_k"], cur["grid_cap"])
        check(L.lib().bx_set_tuning(self.ctx, ctypes.byref(t)), "bx_set_tuning")

    def autotune(self, dx, dt):
        ms = (ctypes.c_double * 3)()
        check(L.lib().bx_autotune(self.ctx, L.dvec3(dx), dt, self.U.ptr, self.U_tmp.ptr,
                                  self.stream, ms), "bx_autotune")
        return [ms[d] for d in range(self.dim)]

    def close(self):
        if self.ctx:
            L.lib().bx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()
        self.U.free()
        self.U_tmp.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
