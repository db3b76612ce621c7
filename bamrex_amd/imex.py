"""IMEX PEEK stage on a 2D periodic level and the pressure Helmholtz solve.

Host side of the learned-preconditioner path.  Every array of a step lives
on the GPU: the stage's explicit work runs in the bx_imex_* kernels, the
pressure solve in bx_helm_* / bx_cnn_*; the host only sequences the calls
and holds the scalars (dt, alpha, beta).

PEEK-SP111 stage (one implicit stage, adt = dt), linearised, enthalpy from
the previous pressure, kinetic energy of the explicit state:
  compute_flux_imex_o1            Source/IMEX/IMEX_O1.cpp:18-175
  advance_rusanov_adv             Source/IMEX/IMEX_Fluxes.cpp:10-118
  compute_enthalpy / pressure_source_vector / setup_pressure_op
                                  Source/IMEX/IMEX_O1.cpp:177-228, 416-452, 570-614
  update_momentum_flux / update_energy_flux
                                  Source/IMEX/IMEX_O1.cpp:480-566
  conservative_update + sum_fluxes (SP111: b = 1)
                                  Source/IMEX/IMEX_RK.cpp:22-96
Device arrays are (ny, nx) in x-fastest order; U has 4 components
(rho, m_x, m_y, E).
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
    """Initial data of the Gresho vortex on [-0.5,0.5]^2
    (Exec/GreshoVortex/Prob.cpp:17-106), rho = 1, p0 = rho / (gamma M^2);
    periodic variant of the problem.  Host arrays (problem set-up)."""
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

    def copy_from(self, other):
        """Device-to-device copy of a DeviceArray of the same size."""
        assert other.nbytes == self.nbytes, (other.shape, self.shape)
        _lib.check(lib().bx_memcpy_dtod(self.ptr, other.ptr, self.nbytes), "dtod")
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
    """A new device array holding a (host array) or a copy of a (DeviceArray)."""
    if isinstance(a, DeviceArray):
        return _copy_of(a)
    return DeviceArray(np.shape(a), dtype).upload(a)


def _copy_of(d):
    """A new device array holding a device-to-device copy of d."""
    out = DeviceArray(d.shape, d.dtype)
    _lib.check(lib().bx_memcpy_dtod(out.ptr, d.ptr, d.nbytes), "dtod")
    return out


class PeekStage:
    """Explicit work of one PEEK-SP111 step on an nx x ny periodic level
    (bx_imex_*): owns the stage's device arrays Uex, enth, rhs, a, bX, bY."""

    def __init__(self, nx, ny, dx, dy=None, gamma=1.4, eps=1.0, advection="rusanov"):
        self.nx, self.ny = nx, ny
        self.dx = dx
        self.dy = dx if dy is None else dy
        self.gamma, self.eps = gamma, eps
        # imex.advection_flux (IMEXSettings.cpp:19-29)
        self.adv = _lib.BX_IMEX_ADVECTION[advection]
        self.Uex = DeviceArray((4, ny, nx))
        self.enth = DeviceArray((ny, nx))
        self.rhs = DeviceArray((ny, nx))
        self.a = DeviceArray((ny, nx))
        self.bX = DeviceArray((ny, nx + 1))
        self.bY = DeviceArray((ny + 1, nx))

    def pressure(self, U, p):
        """p = (gamma - 1)(E - ke) on the device."""
        _chk(lib().bx_imex_pressure(U.ptr, self.nx, self.ny, self.gamma, self.eps, p.ptr, None),
             "bx_imex_pressure")

    def advective_dt(self, U, cfl=0.45):
        out = ctypes.c_double()
        _chk(lib().bx_imex_advective_dt(U.ptr, self.nx, self.ny, self.dx, self.dy, cfl,
                                        ctypes.byref(out), None), "bx_imex_advective_dt")
        return out.value

    def acoustic_dt(self, U, cfl=0.45):
        """The acoustic time step (before imex.acoustic_timestep_end_time)."""
        out = ctypes.c_double()
        _chk(lib().bx_imex_acoustic_dt(U.ptr, self.nx, self.ny, self.dx, self.dy, self.gamma, self.eps,
                                       cfl, ctypes.byref(out), None), "bx_imex_acoustic_dt")
        return out.value

    def assemble(self, U, p, dt):
        """Explicit update and pressure system from the state U and the
        previous pressure p; returns the system (device arrays and the
        operator scalars alpha = 1, beta = dt^2)."""
        _chk(lib().bx_imex_peek_assemble(U.ptr, p.ptr, self.nx, self.ny, self.dx, self.dy, dt,
                                         self.gamma, self.eps, self.adv, self.Uex.ptr, self.enth.ptr,
                                         self.rhs.ptr, self.a.ptr, self.bX.ptr, self.bY.ptr, None),
             "bx_imex_peek_assemble")
        return dict(Uex=self.Uex, enth=self.enth, rhs=self.rhs, a=self.a, bX=self.bX, bY=self.bY,
                    alpha=1.0, beta=dt * dt, dt=dt)

    def finish(self, U, p_new, dt, U_out):
        """Flux corrections with the solved pressure and the conservative
        update into U_out (not U)."""
        _chk(lib().bx_imex_peek_finish(U.ptr, self.Uex.ptr, self.enth.ptr, p_new.ptr, self.nx,
                                       self.ny, self.dx, self.dy, dt, self.eps, self.adv, U_out.ptr,
                                       None),
             "bx_imex_peek_finish")

    def free(self):
        for d in (self.Uex, self.enth, self.rhs, self.a, self.bX, self.bY):
            d.free()


class ButcherTableau:
    """IMEXButcherTableau (Source/IMEX/ButcherTableau.cpp:5-101): the
    implicit and explicit matrices and weights of the reference's IMEX
    Runge-Kutta schemes, by their imex.butcher_tableau names."""
    NAMES = ("SP111", "SSP222", "SASSP322", "SASSP332", "SSP433")

    def __init__(self, name):
        third = 1.0 / 3
        if name == "SP111":
            self.n_imp, self.n_exp, self.order = 1, 1, 1
            self.A_imp, self.A_exp, self.b = [[1.0]], [[0.0]], [1.0]
        elif name == "SSP222":
            g = 1 - 1 / math.sqrt(2)
            self.n_imp, self.n_exp, self.order = 2, 2, 2
            self.A_imp = [[g, 0.0], [1 - 2 * g, g]]
            self.A_exp = [[0.0, 0.0], [1.0, 0.0]]
            self.b = [0.5, 0.5]
        elif name == "SASSP322":
            self.n_imp, self.n_exp, self.order = 3, 2, 2
            self.A_imp = [[0.5, 0.0, 0.0], [-0.5, 0.5, 0.0], [0.0, 0.5, 0.5]]
            self.A_exp = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
            self.b = [0.0, 0.5, 0.5]
        elif name == "SASSP332":
            self.n_imp, self.n_exp, self.order = 3, 3, 2
            self.A_imp = [[0.25, 0.0, 0.0], [0.0, 0.25, 0.0], [third, third, third]]
            self.A_exp = [[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.5, 0.5, 0.0]]
            self.b = [third, third, third]
        elif name == "SSP433":
            a, be, et = 0.24169426078821, 0.06042356519705, 0.12915286960590
            self.n_imp, self.n_exp, self.order = 4, 3, 3
            self.A_imp = [[a, 0.0, 0.0, 0.0], [-a, a, 0.0, 0.0], [0.0, 1 - a, a, 0.0],
                          [be, et, 0.5 - be - et - a, a]]
            self.A_exp = [[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0],
                          [0.0, 0.25, 0.25, 0.0]]
            self.b = [0.0, 1.0 / 6, 1.0 / 6, 2.0 / 3]
        else:
            raise ValueError("imex.butcher_tableau: unknown tableau %r (one of %s)"
                             % (name, ", ".join(self.NAMES)))
        self.name = name

    def imp_row(self, r, mult):
        """get_A_imp_row (ButcherTableau.H:53-59)."""
        return [x * mult for x in self.A_imp[r]]

    def exp_row(self, r, mult):
        """get_A_exp_row (ButcherTableau.H:61-67)."""
        return [self.A_exp[r][i] * mult for i in range(self.n_imp)]


def _ptrs(arrays):
    return (ctypes.c_void_p * _lib.BX_IMEX_MAX_STAGES)(*([a.ptr.value for a in arrays]
                                                          + [None] * (_lib.BX_IMEX_MAX_STAGES
                                                                      - len(arrays))))


def _dbls(vals):
    return (ctypes.c_double * _lib.BX_IMEX_MAX_STAGES)(*(list(vals) + [0.0] * (_lib.BX_IMEX_MAX_STAGES
                                                                              - len(vals))))


class ImexRK:
    """advance_imex_rk (Source/IMEX/IMEX_RK.cpp:22-96) on an nx x ny periodic
    level, every stage's explicit work on the GPU (bx_imex_rk_*,
    bx_imex_stage_*); the pressure solve of each stage is the caller's.

    Per step: begin(U, p, dt); then per stage sy = system(), a solve of
    sy's pressure equation from the initial guess sy["p_prev"], and
    complete(p_new, U_out), which returns True after the last stage has
    written U_out.  p (a device array) holds the stage pressures as the
    reference's pressure MultiFab does: the EK pressure of a later stage,
    then each solution; after the step, the last stage's.  SP111 ends with
    the fused finish (bx_imex_peek_finish), later tableaux store the stage
    fluxes and sum them (bx_imex_rk_update).

    ke_method (imex.ke_method): "ex" (PEEK) or "split" (PESK); with split the
    system carries the operator's velocity term (gamma_c and the device
    arrays cX, cY) for HelmholtzSolver.set_coeffs.  "fi" is the
    non-linearised form (imex.linearise = false): each stage solves
    picard_max_iter times (imex.picard_max_iter), calling picard(p) after
    every solve, which updates the system in place.  "conservative_kinetic"
    takes the kinetic energy of the explicit stage state.  enthalpy_method
    (imex.enthalpy_method): reuse_pressure, conservative or conservative_ex
    (the last two replace the stage's pressure by that of the advected
    state before the solve)."""

    def __init__(self, nx, ny, dx, dy=None, gamma=1.4, eps=1.0, advection="rusanov",
                 tableau="SP111", ke_method="ex", picard_max_iter=2,
                 enthalpy_method="reuse_pressure"):
        self.tab = ButcherTableau(tableau)
        if ke_method not in _lib.BX_IMEX_KE:
            raise ValueError("imex.ke_method: ex, split, conservative_kinetic or fi, not %r"
                             % (ke_method,))
        if enthalpy_method not in _lib.BX_IMEX_ENTHALPY:
            raise ValueError("imex.enthalpy_method: reuse_pressure, conservative or "
                             "conservative_ex, not %r" % (enthalpy_method,))
        self.enth_method = _lib.BX_IMEX_ENTHALPY[enthalpy_method]
        self.ke = _lib.BX_IMEX_KE[ke_method]
        self.split = ke_method == "split"
        self.fi = ke_method == "fi"
        if self.fi and not picard_max_iter >= 1:
            raise ValueError("imex.picard_max_iter must be at least 1")
        self.picard_max_iter = picard_max_iter
        self.stage = PeekStage(nx, ny, dx, dy, gamma, eps, advection)
        self.nx, self.ny, self.dx, self.dy = nx, ny, self.stage.dx, self.stage.dy
        self.gamma, self.eps, self.adv = gamma, eps, self.stage.adv
        ns = self.tab.n_imp
        # SP111 ends with the fused finish: it recomputes the stage fluxes per
        # cell (their pressure corrections do not depend on ke_method)
        self.fused = self.tab.name == "SP111"
        self.cX = DeviceArray((ny, nx)) if self.split else None
        self.cY = DeviceArray((ny, nx)) if self.split else None
        self.Fx = [] if self.fused else [DeviceArray((4, ny, nx)) for _ in range(ns)]
        self.Fy = [] if self.fused else [DeviceArray((4, ny, nx)) for _ in range(ns)]
        self.U_imp = DeviceArray((4, ny, nx)) if ns > 1 else None
        self.U_exp = [DeviceArray((4, ny, nx)), DeviceArray((4, ny, nx))] if ns > 1 else []
        self.s = None

    def begin(self, U, p, dt):
        self.U, self.p, self.dt, self.s = U, p, dt, 0
        self._exp = [None] * self.tab.n_imp   # each stage's explicit state
        self._built = False

    def system(self):
        """The current stage's pressure system (device arrays of the stage
        and the operator scalars alpha = 1, beta = adt^2)."""
        s, tab = self.s, self.tab
        if s is None or s >= tab.n_imp:
            raise RuntimeError("ImexRK.system: no step in progress (call begin)")
        dt = self.dt
        if self._built:
            raise RuntimeError("ImexRK.system: stage %d already assembled (call complete)" % s)
        adt = tab.A_imp[s][s] * dt
        nx, ny, dx, dy = self.nx, self.ny, self.dx, self.dy
        if s == 0:
            Ui = Ue = self.U
        else:
            Ui, Ue = self.U_imp, self.U_exp[s % 2]
            _chk(lib().bx_imex_rk_combine(self.U.ptr, _ptrs(self.Fx[:s]), _ptrs(self.Fy[:s]), s,
                                          _dbls(tab.imp_row(s, dt)[:s]), _dbls(tab.exp_row(s, dt)[:s]),
                                          nx, ny, dx, dy, Ui.ptr, Ue.ptr, None), "bx_imex_rk_combine")
            _chk(lib().bx_imex_rk_pressure(Ue.ptr, self._exp[s - 1].ptr, nx, ny, dx, dy, adt,
                                           self.gamma, self.eps, self.adv, self.ke, self.p.ptr, None),
                 "bx_imex_rk_pressure")
        self._exp[s] = Ue
        self._Ue, self._adt, self._built = Ue, adt, True
        st = self.stage
        _chk(lib().bx_imex_stage_assemble(Ui.ptr, Ue.ptr, self.p.ptr, nx, ny, dx, dy, adt, self.gamma,
                                          self.eps, self.adv, int(tab.name == "SP111"), self.ke,
                                          self.enth_method, st.Uex.ptr, st.enth.ptr, st.rhs.ptr, st.a.ptr, st.bX.ptr,
                                          st.bY.ptr, self.cX.ptr if self.split else None,
                                          self.cY.ptr if self.split else None, None),
             "bx_imex_stage_assemble")
        return dict(Uex=st.Uex, enth=st.enth, rhs=st.rhs, a=st.a, bX=st.bX, bY=st.bY, alpha=1.0,
                    beta=adt * adt, gamma_c=(-0.5 * self.eps * adt) if self.split else 0.0,
                    cX=self.cX, cY=self.cY, dt=dt, adt=adt, stage=s, p_prev=self.p)

    def picard(self, p_new):
        """FI: the end of a Picard iteration after a solve gave p_new; the
        stage's system (enth, rhs, bX, bY) is updated for the next solve,
        or for the fluxes after the last one."""
        if not (self.fi and getattr(self, "_built", False)):
            raise RuntimeError("ImexRK.picard: only inside an FI stage")
        st = self.stage
        _chk(lib().bx_imex_picard_update(st.Uex.ptr, p_new.ptr, self.nx, self.ny, self.dx, self.dy,
                                         self._adt, self.gamma, self.eps, st.enth.ptr, st.rhs.ptr,
                                         st.bX.ptr, st.bY.ptr, None), "bx_imex_picard_update")

    def complete(self, p_new, U_out):
        """The stage's fluxes with the solved pressure p_new; after the last
        stage the step's update into U_out (not the step's U).  Returns
        whether the step is done."""
        s, st = self.s, self.stage
        if not getattr(self, "_built", False):
            raise RuntimeError("ImexRK.complete: the stage's system was not assembled")
        nx, ny, dx, dy = self.nx, self.ny, self.dx, self.dy
        last = s + 1 == self.tab.n_imp
        if self.fused:
            _chk(lib().bx_imex_peek_finish(self.U.ptr, st.Uex.ptr, st.enth.ptr, p_new.ptr, nx, ny, dx,
                                           dy, self.dt, self.eps, self.adv, U_out.ptr, None),
                 "bx_imex_peek_finish")
        else:
            _chk(lib().bx_imex_stage_fluxes(self._Ue.ptr, st.Uex.ptr, st.enth.ptr, p_new.ptr, nx, ny,
                                            dx, dy, self._adt, self.eps, self.adv, 0, self.Fx[s].ptr,
                                            self.Fy[s].ptr, None), "bx_imex_stage_fluxes")
            if last:
                n = self.tab.n_imp
                _chk(lib().bx_imex_rk_update(self.U.ptr, _ptrs(self.Fx[:n]), _ptrs(self.Fy[:n]), n,
                                             _dbls(self.tab.b), nx, ny, dx, dy, self.dt, U_out.ptr,
                                             None), "bx_imex_rk_update")
        if p_new.ptr.value != self.p.ptr.value:
            _lib.check(lib().bx_memcpy_dtod(self.p.ptr, p_new.ptr, self.p.nbytes), "dtod")
        self.s = s + 1
        self._built = False
        return last

    def advance(self, U, p, dt, U_out, solve, residual=None, res_tol=1e-12):
        """One whole step; solve(sy) returns the stage's solved pressure (a
        device array).  FI with residual (imex.compute_residual):
        residual(sy) returns ||b - A p|| / ||b|| of the current pressure
        sy["p_prev"] for the current system, and a stage's Picard loop ends
        before a solve once it is below res_tol (imex.res_tol;
        picard_iterate, IMEX_O1.cpp:736-751)."""
        self.begin(U, p, dt)
        while True:
            sy = self.system()
            if self.fi:
                p_new = self.p
                for _ in range(self.picard_max_iter):
                    if residual is not None and abs(residual(sy)) < res_tol:
                        break
                    p_new = solve(sy)
                    if p_new.ptr.value != self.p.ptr.value:
                        self.p.copy_from(p_new)   # the next solve starts from it
                    self.picard(self.p)
            else:
                p_new = solve(sy)
            if self.complete(p_new, U_out):
                return

    def free(self):
        for d in (self.Fx + self.Fy + self.U_exp + ([self.U_imp] if self.U_imp else [])
                  + ([self.cX, self.cY] if self.split else [])):
            d.free()
        self.stage.free()


class ImexRKStab:
    """advance_imex_rk_stab (Source/IMEX/IMEX_RK.cpp:98-194; imex.stabilize):
    the tableau's step (an ImexRK); for a tableau of order above 1, where
    the shock indicator of the new state is positive (imex.k1), the convex
    combination with the first-order step (SP111, Rusanov fluxes, from the
    incoming pressure) of the state and the pressure."""

    def __init__(self, nx, ny, dx, dy=None, gamma=1.4, eps=1.0, advection="rusanov",
                 tableau="SSP222", ke_method="ex", k1=1e-3):
        self.hi = ImexRK(nx, ny, dx, dy, gamma, eps, advection, tableau, ke_method)
        self.k1 = k1
        self.order = self.hi.tab.order
        self.lo = (ImexRK(nx, ny, dx, dy, gamma, eps, "rusanov", "SP111", ke_method)
                   if self.order > 1 else None)
        self.nx, self.ny = nx, ny
        self.p_o1 = DeviceArray((ny, nx))
        self.U_o1 = DeviceArray((4, ny, nx))
        self.ind = DeviceArray((ny, nx))
        self.work = DeviceArray((ny, nx))
        self.ind_max = None
        self.blended = False

    def indicator(self, U):
        """The shock indicator of U into self.ind; returns its maximum."""
        m = ctypes.c_double()
        _chk(lib().bx_imex_shock_indicator(U.ptr, self.nx, self.ny, self.hi.dx, self.hi.dy,
                                           self.hi.gamma, self.hi.eps, self.k1, self.work.ptr,
                                           self.ind.ptr, ctypes.byref(m), None),
             "bx_imex_shock_indicator")
        return m.value

    def advance(self, U, p, dt, U_out, solve):
        """One stabilised step: U_out and p (in place) as advance_imex_rk_stab
        leaves them; solve(sy) as for ImexRK.advance."""
        self.blended = False
        self.ind_max = None
        self.p_o1.copy_from(p)
        self.hi.advance(U, p, dt, U_out, solve)
        if self.order <= 1:
            return
        self.ind_max = self.indicator(U_out)
        if self.ind_max <= 0:
            return
        self.lo.advance(U, self.p_o1, dt, self.U_o1, solve)
        _chk(lib().bx_imex_stab_blend(self.ind.ptr, self.U_o1.ptr, self.p_o1.ptr, self.nx, self.ny,
                                      U_out.ptr, p.ptr, None), "bx_imex_stab_blend")
        self.blended = True

    def free(self):
        for d in (self.p_o1, self.U_o1, self.ind, self.work):
            d.free()
        self.hi.free()
        if self.lo is not None:
            self.lo.free()


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
        self._own = {}
        self.cnn = None

    def set_coeffs(self, alpha, beta, a, bX, bY, gamma=0.0, cX=None, cY=None, copy=True):
        """Coefficients as host arrays or DeviceArrays.  Device arrays are
        copied unless copy=False: then they are used in place and must keep
        their values until the next set_coeffs (the attached network keeps
        pointers to them)."""
        def own(name, x):
            cur = self._own.get(name)
            if isinstance(x, DeviceArray) and not copy:
                return x
            if isinstance(x, DeviceArray):
                if cur is None or cur.nbytes != x.nbytes:
                    cur = self._own[name] = DeviceArray(x.shape)
                return cur.copy_from(x)
            x = np.ascontiguousarray(x, dtype=np.float64)
            if cur is None or cur.nbytes != x.nbytes:
                cur = self._own[name] = DeviceArray(x.shape)
            return cur.upload(x)
        da, dbx, dby = own("a", a), own("bX", bX), own("bY", bY)
        dcx = own("cX", cX) if cX is not None else None
        dcy = own("cY", cY) if cY is not None else None
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
        for d in self._own.values():
            d.free()
        self._own = {}


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

    def save(self, path):
        _chk(lib().bx_cnn_save(self.c, str(path).encode()), "bx_cnn_save", "bx_cnn_last_error")

    def load(self, path):
        _chk(lib().bx_cnn_load(self.c, str(path).encode()), "bx_cnn_load", "bx_cnn_last_error")

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
    """Successive pressure systems of IMEX PEEK steps of the low-Mach Gresho
    vortex on an n x n periodic level, all on the GPU: one system per stage
    of the tableau (SP111: one per step, SSP222: two).  The state advances
    with the pressure passed to advance() (the benchmark passes the plain
    multigrid solution, so every method solves the same sequence)."""

    def __init__(self, n=1024, mach=0.01, gamma=1.4, eps=1.0, cfl=0.45, advection="rusanov",
                 tableau="SP111", ke_method="ex", acoustic_timestep_end_time=-1.0):
        U, self.dx = gresho_state(n, mach, gamma, eps)
        self.n = n
        self.U = dev(U)
        self.U_next = DeviceArray(U.shape)
        self.p = DeviceArray((n, n))
        self.rk = ImexRK(n, n, self.dx, gamma=gamma, eps=eps, advection=advection, tableau=tableau,
                         ke_method=ke_method)
        self.stage = self.rk.stage
        self.stage.pressure(self.U, self.p)
        self.cfl = cfl
        self.in_step = False
        self.time = 0.0
        self.acoustic_end = acoustic_timestep_end_time

    def step_dt(self):
        """The dt of the next step (estTimeStep, AmrLevelAdv.cpp:424-437):
        acoustic before imex.acoustic_timestep_end_time, then advective."""
        if self.time < self.acoustic_end:
            return self.stage.acoustic_dt(self.U, self.cfl)
        return self.stage.advective_dt(self.U, self.cfl)

    def next_system(self, dt=None):
        """Assembles the next system to solve (starting a step with the
        estimated dt, or dt, when the last one is done): a stage's, or for
        FI the stage's system after the previous Picard update; "p_prev" is
        the solve's initial guess (device array)."""
        if getattr(self, "_picard_sy", None) is not None:
            return self._picard_sy
        if not self.in_step:
            self.rk.begin(self.U, self.p, self.step_dt() if dt is None else dt)
            self.in_step = True
        self._picard_it = 0
        return self.rk.system()

    def advance(self, sy, p_new):
        """Completes the system's solve with the solved pressure p_new
        (DeviceArray): for FI with Picard iterations left, the Picard update
        (the next next_system() returns the updated system); otherwise the
        stage, and after a step's last stage U <- the new state.  Returns
        whether a step ended."""
        if self.rk.fi:
            if p_new.ptr.value != self.p.ptr.value:
                self.p.copy_from(p_new)
            self.rk.picard(self.p)
            self._picard_it += 1
            if self._picard_it < self.rk.picard_max_iter:
                self._picard_sy = sy
                lib().bx_stream_sync(None)
                return False
            self._picard_sy = None
        done = self.rk.complete(p_new, self.U_next)
        if done:
            self.U, self.U_next = self.U_next, self.U
            self.in_step = False
            self.time += self.rk.dt
        lib().bx_stream_sync(None)
        return done

    def free(self):
        for d in (self.U, self.U_next, self.p):
            d.free()
        self.rk.free()
