"""CPU restatement (numpy) of the explicit work of the IMEX PEEK-SP111 stage.

TEST INFRASTRUCTURE ONLY: the checker for bx_imex_* (bamrex_amd/csrc/
bx_imex.hip).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import it; the product path never does.

Settings restated: linearised PEEK, enthalpy from the previous pressure
(IMEXSettings::reuse_pressure), kinetic energy of the explicit state
(IMEXSettings::ex), one implicit stage with adt = dt (SP111), 2D periodic.

  kinetic_energy / pressure / adv_flux   Source/IMEX_K/euler_K.H:8-65
  max_dx_scaled_speed                    Source/System/Euler/Euler.cpp:107-129
  estTimeStep (advective, cfl)           Source/AmrLevelAdv.cpp:431-436, 466
  advance_rusanov_adv                    Source/IMEX/IMEX_Fluxes.cpp:10-118
  advance_MUSCL_rusanov_adv (SP111)      Source/IMEX/IMEX_Fluxes.cpp:120-324
  hdtdx = 0.5 dt idx, hdtdxe = hdtdx/eps Source/IMEX/IMEX_O1.cpp:66-70
  compute_enthalpy (reuse_pressure)      Source/IMEX/IMEX_O1.cpp:177-228
  pressure_source_vector                 Source/IMEX/IMEX_O1.cpp:416-452
  setup_pressure_op                      Source/IMEX/IMEX_O1.cpp:570-614
  update_momentum_flux / update_energy_flux
                                         Source/IMEX/IMEX_O1.cpp:480-566
  conservative_update (b = 1)            Source/IMEX/IMEX_RK.cpp:22-96

Arrays are (ny, nx), x fastest; U has 4 components (rho, m_x, m_y, E).
numpy evaluates each binary operation in IEEE double with no contraction,
in the order written, which is the order of the reference's expressions.

Pinned by tests/test_imex_stage.py: the reference's Rusanov known answers
(Tests/RusanovAdv01.cpp), the energy/pressure identity of
Tests/IMEX_O1_01.cpp (ExplicitKEEnergyMatch), free-stream preservation and
conservation, brute-force per-cell evaluation of the advective dt, and for
the MUSCL form: equality with the first-order fluxes on piecewise-constant
data (minmod vanishes at a single jump) and second-order convergence on a
smooth advected density wave.
"""
import numpy as np


def kinetic_energy(U, eps):
    """euler_K.H:8-20: eps (m_x^2 + m_y^2) / (2 rho)."""
    return eps * (U[1] * U[1] + U[2] * U[2]) / (2 * U[0])


def pressure_from_state(U, gamma, eps):
    """euler_K.H:23-35: (gamma - 1)(E - ke)."""
    return (gamma - 1) * (U[3] - kinetic_energy(U, eps))


def advective_dt(U, dx, dy=None, cfl=0.45):
    """cfl * (1 / max((|m_x|/dx + |m_y|/dy) / rho))."""
    dy = dx if dy is None else dy
    s = np.max((np.abs(U[1]) / dx + np.abs(U[2]) / dy) / U[0])
    dt = 1 / float(s)
    dt *= cfl
    return dt


def _sh(a, d, k):
    """a at index i + k along direction d (periodic); arrays are (.., ny, nx)."""
    return np.roll(a, -k, axis=-1 - d)


def rusanov_faces(U, eps):
    """Rusanov advective face fluxes F_d at faces i-1/2 (index i),
    IMEX_Fluxes.cpp:56-90."""
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
        au, aum = np.abs(u), np.abs(_sh(u, d, -1))
        umax = np.where(au < aum, aum, au)   # amrex::max(a, b) = a < b ? b : a
        out.append(0.5 * ((F + Fm) - umax * (U - Um)))
    return out


def minmod(dl, dr):
    """Source/Fluxes/Limiters.H:75-81."""
    return np.where((dl > 0) & (dr > 0), np.fmin(dl, dr),
                    np.where((dl < 0) & (dr < 0), np.fmax(dl, dr), 0.0))


def adv_flux(S, d, eps):
    """adv_flux<d> (euler_K.H:53-65) of the state S."""
    u = S[1 + d] / S[0]
    F = np.empty_like(S)
    F[0] = S[1 + d]
    F[1] = S[1] * u
    F[2] = S[2] * u
    F[3] = u * kinetic_energy(S, eps)
    return F


def muscl_rusanov_faces(U, eps, dt, dx, dy):
    """advance_MUSCL_rusanov_adv (IMEX_Fluxes.cpp:120-324) with the SP111
    local step: per direction, minmod half-slopes, a half step of the
    states with their own fluxes (hdtdx = 0.5 * (dt/dx)), then Rusanov fluxes
    between R of cell i-1 and L of cell i (face index i)."""
    out = []
    for d, h in ((0, dx), (1, dy)):
        hd = 0.5 * (dt / h)
        hdelta = 0.5 * minmod(U - _sh(U, d, -1), _sh(U, d, 1) - U)
        L = U - hdelta
        R = U + hdelta
        lu = hd * (adv_flux(R, d, eps) - adv_flux(L, d, eps))
        L = L - lu
        R = R - lu
        FL, FR = adv_flux(L, d, eps), adv_flux(R, d, eps)
        Rm, FRm = _sh(R, d, -1), _sh(FR, d, -1)
        a, b = np.abs(L[1 + d] / L[0]), np.abs(Rm[1 + d] / Rm[0])
        umax = np.where(a < b, b, a)
        out.append(0.5 * ((FL + FRm) - umax * (L - Rm)))
    return out


class PeekStage:
    """One PEEK-SP111 step: explicit advection, pressure system, and after
    the solve the flux correction and conservative update.  advection:
    "rusanov" (advance_rusanov_adv) or "muscl_rusanov"."""

    def __init__(self, gamma=1.4, eps=1.0, advection="rusanov"):
        self.gamma = gamma
        self.eps = eps
        self.muscl = advection == "muscl_rusanov"

    def _faces(self, U, dt, dx, dy):
        if self.muscl:
            return muscl_rusanov_faces(U, self.eps, dt, dx, dy)
        return rusanov_faces(U, self.eps)

    def assemble(self, U, p, dt, dx, dy=None):
        """Returns Uex, enth, rhs, a, bX (ny, nx+1), bY (ny+1, nx), alpha,
        beta and the Rusanov fluxes Fx, Fy."""
        dy = dx if dy is None else dy
        g, eps = self.gamma, self.eps
        Fx, Fy = self._faces(U, dt, dx, dy)
        dtdx, dtdy = dt / dx, dt / dy
        Uex = U - (dtdx * (_sh(Fx, 0, 1) - Fx) + dtdy * (_sh(Fy, 1, 1) - Fy))
        enth = p * g / ((g - 1) * Uex[0])
        ke = kinetic_energy(Uex, eps)
        hdtdx = 0.5 * dt * (1.0 / dx)
        hdtdy = 0.5 * dt * (1.0 / dy)
        rhs = eps * (Uex[3] - ke - (hdtdx * (_sh(enth, 0, 1) * _sh(Uex[1], 0, 1)
                                             - _sh(enth, 0, -1) * _sh(Uex[1], 0, -1))
                                    + hdtdy * (_sh(enth, 1, 1) * _sh(Uex[2], 1, 1)
                                               - _sh(enth, 1, -1) * _sh(Uex[2], 1, -1))))
        # face enthalpy (average_cellcenter_to_face): 0.5 (h(i-1) + h(i))
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

    def finish(self, U, st, p_new, dt, dx, dy=None):
        """Pressure and enthalpy flux corrections, then U_new = U - dt div F."""
        dy = dx if dy is None else dy
        eps = self.eps
        heps = 0.5 / eps
        hdtdxe = 0.5 * dt * (1.0 / dx) / eps
        hdtdye = 0.5 * dt * (1.0 / dy) / eps
        Fx = st["Fx"].copy()
        Fy = st["Fy"].copy()
        h, ex = st["enth"], st["Uex"]
        for d, F, hd in ((0, Fx, hdtdxe), (1, Fy, hdtdye)):
            pm = _sh(p_new, d, -1)
            hm = _sh(h, d, -1)
            F[1 + d] += heps * (p_new + pm)
            F[3] += (0.5 * (h * ex[1 + d] + hm * _sh(ex[1 + d], d, -1))
                     - hd * ((h + hm) * (p_new - pm)))
        dtdx, dtdy = dt / dx, dt / dy
        return U - (dtdx * (_sh(Fx, 0, 1) - Fx) + dtdy * (_sh(Fy, 1, 1) - Fy))
