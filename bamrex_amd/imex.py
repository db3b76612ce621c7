"""IMEX PEEK stage on a 2D periodic level and the pressure Helmholtz solve.

Host side of the learned-preconditioner path.  The stage assembly is the
caller of the pressure solve (in the reference, AMReX host code around
MLMG); the solve itself runs on the GPU through the C ABI
(`bx_helm_*`, `bx_cnn_*`).

PEEK-SP111 stage (one implicit stage, adt = dt), following
  compute_flux_imex_o1            Source/IMEX/IMEX_O1.cpp:18-175
  advance_rusanov_adv             Source/IMEX/IMEX_Fluxes.cpp:10-118
  adv_flux / kinetic_energy       Source/IMEX_K/euler_K.H:10-65
  compute_enthalpy (reuse_pressure), compute_ke (ex)
                                  Source/IMEX/IMEX_O1.cpp:177-262
  pressure_source_vector          Source/IMEX/IMEX_O1.cpp:416-452
  setup_pressure_op               Source/IMEX/IMEX_O1.cpp:570-614
  update_momentum_flux / update_energy_flux
                                  Source/IMEX/IMEX_O1.cpp:480-566
  conservative_update + sum_fluxes (SP111: b = 1)
                                  Source/IMEX/IMEX_RK.cpp:22-96
Arrays are numpy (ny, nx) in x-fastest order, the same order as the device
buffers; U has 4 components (rho, m_x, m_y, E).
"""
import ctypes
import math

import numpy as np

from . import _lib
from ._lib import lib


def _chk(rc, what, err="bx_helm_last_error"):
    if rc != 0:
        msg = getattr(lib(), err)()
        raise _lib.BamrexError("%s failed (code %d): %s" % (what, rc, msg.decode() if msg else ""))


# ------------------------------------------------------------------ problem
def gresho_state(n, mach=0.01, gamma=1.4, eps=1.0):
    """Gresho vortex on [-0.5,0.5]^2 (Exec/GreshoVortex/Prob.cpp:17-106),
    rho = 1, p0 = rho / (gamma M^2); periodic variant of the problem."""
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx - 0.5
    x, y = np.meshgrid(c, c)          # (ny, nx), x fastest
    r = np.sqrt(x * x + y * y)
    th = np.arctan2(y, x)
    ut = np.where(r < 0.2, 5 * r, np.where(r < 0.4, 2 - 5 * r, 0.0))
    p0 = 1.0 / (gamma * mach * mach)
    with np.errstate(divide="ignore", invalid="ignore"):
        p = np.where(r < 0.2, p0 + 12.5 * r * r,
                     np.where(r < 0.4, p0 + 12.5 * r * r + 4 * (1 - 5 * r + np.log(5 * r)),
                              p0 - 2 + 4 * math.log(2)))
    ux = -ut * np.sin(th)
    uy = ut * np.cos(th)
    rho = np.ones_like(p)
    U = np.empty((4, n, n))
    U[0] = rho
    U[1] = rho * ux
    U[2] = rho * uy
    U[3] = p / (gamma - 1) + 0.5 * eps * rho * (ux * ux + uy * uy)
    return U, dx


def kinetic_energy(U, eps):
    return eps * (U[1] * U[1] + U[2] * U[2]) / (2 * U[0])


def pressure_from_state(U, gamma, eps):
    return (gamma - 1) * (U[3] - kinetic_energy(U, eps))


def advective_dt(U, dx, cfl=0.45):
    """Advective time step of the IMEX schemes (AmrLevelAdv.cpp:435):
    cfl / max(|u|/dx + |v|/dy)."""
    s = np.max(np.abs(U[1] / U[0]) / dx + np.abs(U[2] / U[0]) / dx)
    return cfl / s


def _sh(a, d, k):
    """a at index i + k along direction d (periodic); arrays are (.., ny, nx)."""
    return np.roll(a, -k, axis=-1 - d)


def _rusanov_faces(U, eps):
    """Rusanov advective face fluxes F_d at faces i-1/2 (index i),
    IMEX_Fluxes.cpp:73-99."""
    ke = kinetic_energy(U, eps)
    out = []
    for d in range(2):
        u = U[1 + d] / U[0]
        F = np.empty_like(U)
        F[0] = U[1 + d]
        F[1] = U[1] * u
        F[2] = U[2] * u
        F[3] = u * ke
        Fm = _sh(F, d, -1)
        Um = _sh(U, d, -1)
        umax = np.maximum(np.abs(u), np.abs(_sh(u, d, -1)))
        out.append(0.5 * ((F + Fm) - umax * (U - Um)))
    return out


class PeekStage:
    """One PEEK-SP111 step: explicit advection, pressure system, solve,
    flux correction and conservative update."""

    def __init__(self, gamma=1.4, eps=1.0):
        self.gamma = gamma
        self.eps = eps

    def assemble(self, U, p, dt, dx):
        """Returns (U_ex, enthalpy, rhs, a, bX, bY, beta) for the stage with
        adt = dt.  bX is (ny, nx+1), bY is (ny+1, nx)."""
        g, eps = self.gamma, self.eps
        Fx, Fy = _rusanov_faces(U, eps)
        dtdx = dt / dx
        Uex = U - (dtdx * (_sh(Fx, 0, 1) - Fx) + dtdx * (_sh(Fy, 1, 1) - Fy))
        enth = p * g / ((g - 1) * Uex[0])
        ke = kinetic_energy(Uex, eps)
        hdt = 0.5 * dt / dx
        rhs = eps * (Uex[3] - ke - (hdt * (_sh(enth, 0, 1) * _sh(Uex[1], 0, 1)
                                           - _sh(enth, 0, -1) * _sh(Uex[1], 0, -1))
                                    + hdt * (_sh(enth, 1, 1) * _sh(Uex[2], 1, 1)
                                             - _sh(enth, 1, -1) * _sh(Uex[2], 1, -1))))
        # face enthalpy: arithmetic average of the two neighbours
        hx = 0.5 * (_sh(enth, 0, -1) + enth)       # face i-1/2 at index i
        hy = 0.5 * (_sh(enth, 1, -1) + enth)
        ny, nx = p.shape
        bX = np.empty((ny, nx + 1))
        bX[:, :nx] = hx
        bX[:, nx] = hx[:, 0]
        bY = np.empty((ny + 1, nx))
        bY[:ny] = hy
        bY[ny] = hy[0]
        a = np.full_like(p, eps / (g - 1))
        return dict(Uex=Uex, enth=enth, rhs=rhs, a=a, bX=bX, bY=bY,
                    alpha=1.0, beta=dt * dt, Fx=Fx, Fy=Fy)

    def finish(self, U, st, p_new, dt, dx):
        """Pressure and enthalpy flux corrections, then U_new = U - dt/dx div F
        (IMEX_O1.cpp:480-566, IMEX_RK.cpp:93-95)."""
        eps = self.eps
        heps = 0.5 / eps
        hdtdxe = 0.5 * dt / dx / eps
        Fx = st["Fx"].copy()
        Fy = st["Fy"].copy()
        h, ex = st["enth"], st["Uex"]
        pm = _sh(p_new, 0, -1)
        Fx[1] += heps * (p_new + pm)
        Fx[3] += (0.5 * (h * ex[1] + _sh(h, 0, -1) * _sh(ex[1], 0, -1))
                  - hdtdxe * ((h + _sh(h, 0, -1)) * (p_new - pm)))
        pm = _sh(p_new, 1, -1)
        Fy[2] += heps * (p_new + pm)
        Fy[3] += (0.5 * (h * ex[2] + _sh(h, 1, -1) * _sh(ex[2], 1, -1))
                  - hdtdxe * ((h + _sh(h, 1, -1)) * (p_new - pm)))
        dtdx = dt / dx
        return U - (dtdx * (_sh(Fx, 0, 1) - Fx) + dtdx * (_sh(Fy, 1, 1) - Fy))


# ------------------------------------------------------------- GPU solver
class DeviceArray:
    """A device allocation holding one float64/float32 numpy-shaped array."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _lib.check(lib().bx_malloc(ctypes.byref(p), max(self.nbytes, 8)), "bx_malloc")
        self.ptr = p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _lib.check(lib().bx_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes), "htod")
        return self

    def download(self):
        out = np.empty(self.shape, self.dtype)
        _lib.check(lib().bx_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().bx_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def dev(a, dtype=np.float64):
    return DeviceArray(np.shape(a), dtype).upload(a)


class HelmholtzSolver:
    """Owns a bx_helm solver for one nx x ny periodic level."""

    def __init__(self, nx, ny, dx, dy=None):
        self.nx, self.ny = nx, ny
        self.dx = dx
        self.dy = dx if dy is None else dy
        h = ctypes.c_void_p()
        _chk(lib().bx_helm_create(nx, ny, dx, self.dy, 1, ctypes.byref(h)), "bx_helm_create")
        self.h = h
        self._keep = []
        self.cnn = None

    def set_coeffs(self, alpha, beta, a, bX, bY, gamma=0.0, cX=None, cY=None):
        da, dbx, dby = dev(a), dev(bX), dev(bY)
        dcx = dev(cX) if cX is not None else None
        dcy = dev(cY) if cY is not None else None
        _chk(lib().bx_helm_set_coeffs(self.h, alpha, beta, gamma, da.ptr, dbx.ptr, dby.ptr,
                                      dcx.ptr if dcx else None, dcy.ptr if dcy else None, None),
             "bx_helm_set_coeffs")
        # the network reads the coefficients at training time
        self._keep = [da, dbx, dby, dcx, dcy]
        self.coeffs = dict(alpha=alpha, beta=beta, a=da, bX=dbx, bY=dby)
        if self.cnn is not None:
            self.cnn.set_operator(self)

    def set_options(self, nu1=2, nu2=2, nu_bottom=16, precond_fp32=1, restart=20):
        _chk(lib().bx_helm_set_options(self.h, nu1, nu2, nu_bottom, precond_fp32, restart),
             "bx_helm_set_options")

    def apply(self, x_dev, y_dev):
        _chk(lib().bx_helm_apply(self.h, x_dev.ptr, y_dev.ptr, None), "bx_helm_apply")
        lib().bx_stream_sync(None)

    def residual(self, b_dev, x_dev, r_dev):
        _chk(lib().bx_helm_residual(self.h, b_dev.ptr, x_dev.ptr, r_dev.ptr, None), "bx_helm_residual")
        lib().bx_stream_sync(None)

    def smooth(self, phi_dev, rhs_dev, redblack):
        _chk(lib().bx_helm_smooth(self.h, phi_dev.ptr, rhs_dev.ptr, redblack, None), "bx_helm_smooth")
        lib().bx_stream_sync(None)

    def precond(self, r_dev, e_dev):
        _chk(lib().bx_helm_precond(self.h, r_dev.ptr, e_dev.ptr, None), "bx_helm_precond")
        lib().bx_stream_sync(None)

    def attach(self, cnn, omega=1.0, mode="corrector"):
        self.cnn = cnn
        if cnn is not None:
            cnn.set_operator(self)
        _chk(lib().bx_helm_attach_cnn(self.h, cnn.c if cnn else None, omega, _lib.BX_CNN_MODE[mode]),
             "bx_helm_attach_cnn")

    def set_graph(self, use_graph):
        _chk(lib().bx_helm_set_graph(self.h, int(bool(use_graph))), "bx_helm_set_graph")

    def set_fused(self, fused):
        _chk(lib().bx_helm_set_fused(self.h, int(bool(fused))), "bx_helm_set_fused")

    def solve(self, b_dev, x_dev, method="mg", rtol=1e-10, maxiter=200):
        st = _lib.bx_solve_stats()
        _chk(lib().bx_helm_solve(self.h, b_dev.ptr, x_dev.ptr, _lib.BX_SOLVE[method], rtol, maxiter,
                                 ctypes.byref(st), None), "bx_helm_solve")
        return dict(iters=st.iters, converged=bool(st.converged), fell_back=bool(st.fell_back),
                    precond_applies=st.precond_applies, rel_resid=st.rel_resid, seconds=st.seconds)

    def close(self):
        if self.h:
            lib().bx_helm_destroy(self.h)
            self.h = None


class LearnedSmoother:
    """Owns a bx_cnn network."""

    NPARAMS = _lib.BX_CNN_NPARAMS

    def __init__(self, seed=1):
        c = ctypes.c_void_p()
        _chk(lib().bx_cnn_create(seed, ctypes.byref(c)), "bx_cnn_create", "bx_cnn_last_error")
        self.c = c
        self.nx = self.ny = None

    def set_operator(self, solver):
        k = solver.coeffs
        _chk(lib().bx_cnn_set_operator(self.c, solver.nx, solver.ny, k["alpha"], k["beta"], solver.dx,
                                       solver.dy, k["a"].ptr, k["bX"].ptr, k["bY"].ptr, None),
             "bx_cnn_set_operator", "bx_cnn_last_error")
        lib().bx_stream_sync(None)
        self.nx, self.ny = solver.nx, solver.ny

    def get_params(self):
        w = np.empty(self.NPARAMS, np.float32)
        _chk(lib().bx_cnn_get_params(self.c, w.ctypes.data, self.NPARAMS), "bx_cnn_get_params",
             "bx_cnn_last_error")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _chk(lib().bx_cnn_set_params(self.c, w.ctypes.data, self.NPARAMS), "bx_cnn_set_params",
             "bx_cnn_last_error")

    def net_forward(self, x, periodic=False, layered=False):
        """y = N(x) for x of shape (nimg, h, w) float32 (rounded to bf16):
        the fused inference kernel, or the layer-by-layer training kernels."""
        x = np.ascontiguousarray(x, np.float32)
        nimg, hh, ww = x.shape
        dx_, dy_ = dev(x, np.float32), DeviceArray(x.shape, np.float32)
        flags = (1 if periodic else 0) | (2 if layered else 0)
        _chk(lib().bx_cnn_net_forward(self.c, dx_.ptr, dy_.ptr, ww, hh, nimg, flags, None),
             "bx_cnn_net_forward", "bx_cnn_last_error")
        return dy_.download()

    def infer(self, r_dev, e_dev, omega=1.0):
        _chk(lib().bx_cnn_infer(self.c, r_dev.ptr, e_dev.ptr, self.nx, self.ny, omega, None),
             "bx_cnn_infer", "bx_cnn_last_error")
        lib().bx_stream_sync(None)

    def loss_grad(self, R_dev, nfields, crops, crop):
        crops = np.ascontiguousarray(crops, np.int32)
        loss = ctypes.c_double()
        g = np.empty(self.NPARAMS, np.float32)
        _chk(lib().bx_cnn_loss_grad(self.c, R_dev.ptr, nfields, self.nx, self.ny, crops.ctypes.data,
                                    len(crops), crop, ctypes.byref(loss), g.ctypes.data, None),
             "bx_cnn_loss_grad", "bx_cnn_last_error")
        return loss.value, g

    def train(self, R_dev, nfields, steps, crop=64, batch=32, lr=1e-3, seed=7):
        hist = np.zeros(steps)
        _chk(lib().bx_cnn_train(self.c, R_dev.ptr, nfields, self.nx, self.ny, crop, batch, steps,
                                ctypes.c_float(lr), seed, hist.ctypes.data, None),
             "bx_cnn_train", "bx_cnn_last_error")
        return hist

    def close(self):
        if self.c:
            lib().bx_cnn_destroy(self.c)
            self.c = None


# ------------------------------------------------------- workload drivers
def _offset(darr, nbytes):
    return ctypes.c_void_p(darr.ptr.value + nbytes)


def harvest_residuals(solver, b_dev, x_dev, R_dev, first, count, rtol=1e-10):
    """Plain-multigrid solve of A x = b one V-cycle at a time; the residual
    after each cycle (the training inputs: A e = r with e the remaining
    error) is copied into fields first, first+1, ... of R_dev (at most
    `count`).  Returns (fields written, solve stats of the last cycle)."""
    n = solver.nx * solver.ny
    r = DeviceArray((solver.ny, solver.nx))
    got = 0
    st = None
    for _ in range(50):
        solver.residual(b_dev, x_dev, r)
        if got < count:
            _lib.check(lib().bx_memcpy_dtod(_offset(R_dev, (first + got) * n * 8), r.ptr, n * 8), "dtod")
            got += 1
        st = solver.solve(b_dev, x_dev, "mg", rtol=rtol, maxiter=1)
        if st["converged"] and st["iters"] == 0:
            break
    r.free()
    return got, st


class GreshoWorkload:
    """Successive PEEK-SP111 pressure systems of the low-Mach Gresho vortex
    on an n x n periodic level: the state advances with the plain multigrid
    pressure so every method solves the same sequence of systems."""

    def __init__(self, n=1024, mach=0.01, gamma=1.4, eps=1.0, cfl=0.45):
        self.U, self.dx = gresho_state(n, mach, gamma, eps)
        self.n = n
        self.stage = PeekStage(gamma, eps)
        self.p = pressure_from_state(self.U, gamma, eps)
        self.cfl = cfl

    def next_system(self):
        dt = advective_dt(self.U, self.dx, self.cfl)
        sy = self.stage.assemble(self.U, self.p, dt, self.dx)
        sy["dt"] = dt
        sy["p_prev"] = self.p.copy()
        return sy

    def advance(self, sy, p_new):
        self.U = self.stage.finish(self.U, sy, p_new, sy["dt"], self.dx)
        self.p = p_new
