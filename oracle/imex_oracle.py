"""CPU restatement (numpy) of the explicit work of the IMEX PEEK stages.

TEST INFRASTRUCTURE ONLY: the checker for bx_imex_* (bamrex_amd/csrc/
bx_imex.hip).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import it; the product path never does.

Settings restated: linearised, enthalpy from the previous pressure
(IMEXSettings::reuse_pressure), kinetic energy of the explicit state
(ke_method ex: PEEK) or split (ke_method split: PESK, with the velocity
term of the pressure operator), 2D periodic; one implicit stage with
adt = dt (SP111, PeekStage) or the stage loop of any of the reference's
IMEX Runge-Kutta tableaux (advance_imex_rk).

  kinetic_energy / pressure / adv_flux   Source/IMEX_K/euler_K.H:8-65
  max_dx_scaled_speed                    Source/System/Euler/Euler.cpp:107-129
  estTimeStep (advective, cfl)           Source/AmrLevelAdv.cpp:431-436, 466
  ExplicitTimeStepDimSplit (acoustic)    Source/System/FluidSystem.cpp:195-206
  advance_rusanov_adv                    Source/IMEX/IMEX_Fluxes.cpp:10-118
  advance_MUSCL_rusanov_adv (SP111)      Source/IMEX/IMEX_Fluxes.cpp:120-324
  hdtdx = 0.5 dt idx, hdtdxe = hdtdx/eps Source/IMEX/IMEX_O1.cpp:66-70
  compute_enthalpy (reuse_pressure)      Source/IMEX/IMEX_O1.cpp:177-228
  pressure_source_vector                 Source/IMEX/IMEX_O1.cpp:416-452
  setup_pressure_op                      Source/IMEX/IMEX_O1.cpp:570-614
  update_momentum_flux / update_energy_flux
                                         Source/IMEX/IMEX_O1.cpp:480-566
  conservative_update (one flux set)     Source/Fluxes/Update.H:74-130
  conservative_update (stage states)     Source/Fluxes/Update.H:218-302
  sum_fluxes                             Source/IMEX/IMEX_RK.cpp:196-241
  advance_imex_rk (stage loop)           Source/IMEX/IMEX_RK.cpp:22-96
  compute_pressure (EK and SK branches)  Source/IMEX/IMEX_O1.cpp:300-391
  compute_ke (ex, split), compute_velocity
                                         Source/IMEX/IMEX_O1.cpp:232-298
  picard_iterate, update_momentum (FI)   Source/IMEX/IMEX_O1.cpp:455-482, 702-789
  IMEXButcherTableau                     Source/IMEX/ButcherTableau.cpp:5-101
  advance_imex_rk_stab, compute_shock_indicator
                                         Source/IMEX/IMEX_RK.cpp:98-194, 243-345

Arrays are (ny, nx), x fastest; U has 4 components (rho, m_x, m_y, E).
numpy evaluates each binary operation in IEEE double with no contraction,
in the order written, which is the order of the reference's expressions.

Pinned by tests/test_imex_stage.py: the reference's Rusanov known answers
(Tests/RusanovAdv01.cpp), the energy/pressure identity of
Tests/IMEX_O1_01.cpp (ExplicitKEEnergyMatch), free-stream preservation and
conservation, brute-force per-cell evaluation of the advective dt, and for
the MUSCL form: equality with the first-order fluxes on piecewise-constant
data (minmod vanishes at a single jump) and second-order convergence on a
smooth advected density wave.  The split form: the identity of
Tests/IMEX_O1_01.cpp (SplitKEEnergyMatch) with a real solve of the
operator with its velocity term.  The stage loop: the tableaux against the
reference's Tests/IMEXButcherTableaux01.cpp values and their order
conditions, SP111 through the loop equal to PeekStage, second order in
time of SSP222 with MUSCL fluxes without the local step (first order with
SP111), and conservation with a real pressure solve.
"""
import math

import numpy as np


def kinetic_energy(U, eps):
    """euler_K.H:8-20: eps (m_x^2 + m_y^2) / (2 rho)."""
    return eps * (U[1] * U[1] + U[2] * U[2]) / (2 * U[0])


def pressure_from_state(U, gamma, eps):
    """euler_K.H:23-35: (gamma - 1)(E - ke)."""
    return (gamma - 1) * (U[3] - kinetic_energy(U, eps))


def _reduce_max(s):
    """ParReduce with ReduceOpMax from the lowest value: d = max(d, s) =
    (d < s) ? s : d skips NaN; a positive maximum is what remains (0 if no
    value is positive, which then gives an infinite dt as in C)."""
    pos = s[s > 0]
    return float(pos.max()) if pos.size else 0.0


def advective_dt(U, dx, dy=None, cfl=0.45):
    """cfl * (1 / max((|m_x|/dx + |m_y|/dy) / rho))."""
    dy = dx if dy is None else dy
    s = _reduce_max((np.abs(U[1]) / dx + np.abs(U[2]) / dy) / U[0])
    dt = 1 / s if s > 0 else math.inf
    dt *= cfl
    return dt


def acoustic_dt(U, dx, dy, gamma, eps, cfl=0.45):
    """ExplicitTimeStepDimSplit (FluidSystem.cpp:195-206) with max_wave_speed
    (Euler.cpp:47-74): s = max over cells of sqrt(gamma p / (rho eps)) +
    sqrt(u^2 + v^2); dt = min(1e20, dx/s, dy/s) (std::min in that order),
    then dt *= cfl (AmrLevelAdv.cpp:466)."""
    u = U[1] / U[0]
    v = U[2] / U[0]
    usq = u * u + v * v
    p = (gamma - 1) * (U[3] - 0.5 * eps * U[0] * usq)
    with np.errstate(invalid="ignore"):
        s = np.sqrt(gamma * p / (U[0] * eps)) + np.sqrt(u * u + v * v)
    smax = _reduce_max(s)
    dt = 1.0e20
    for h in (dx, dy):
        q = h / smax if smax > 0 else math.inf
        dt = q if q < dt else dt
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


def muscl_rusanov_faces(U, eps, dt, dx, dy, local=True):
    """advance_MUSCL_rusanov_adv (IMEX_Fluxes.cpp:120-324): per direction,
    minmod half-slopes; with local (the SP111 local time step) a half step
    of the states with their own fluxes (hdtdx = 0.5 * (dt/dx)); then
    Rusanov fluxes between R of cell i-1 and L of cell i (face index i)."""
    out = []
    for d, h in ((0, dx), (1, dy)):
        hd = 0.5 * (dt / h)
        hdelta = 0.5 * minmod(U - _sh(U, d, -1), _sh(U, d, 1) - U)
        L = U - hdelta
        R = U + hdelta
        if local:
            lu = hd * (adv_flux(R, d, eps) - adv_flux(L, d, eps))
            L = L - lu
            R = R - lu
        FL, FR = adv_flux(L, d, eps), adv_flux(R, d, eps)
        Rm, FRm = _sh(R, d, -1), _sh(FR, d, -1)
        a, b = np.abs(L[1 + d] / L[0]), np.abs(Rm[1 + d] / Rm[0])
        umax = np.where(a < b, b, a)
        out.append(0.5 * ((FL + FRm) - umax * (L - Rm)))
    return out


def advect(Uexp, Uimp, dt, dx, dy, eps, muscl, local):
    """advance_rusanov_adv / advance_MUSCL_rusanov_adv: fluxes of the
    explicit state Uexp applied to the state Uimp (IMEX_Fluxes.cpp:92-111,
    290-310): Uimp - (dt/dx (Fx(i+1) - Fx(i)) + dt/dy (Fy(j+1) - Fy(j))).
    Returns the updated state and the face fluxes."""
    if muscl:
        Fx, Fy = muscl_rusanov_faces(Uexp, eps, dt, dx, dy, local)
    else:
        Fx, Fy = rusanov_faces(Uexp, eps)
    dtdx, dtdy = dt / dx, dt / dy
    return Uimp - (dtdx * (_sh(Fx, 0, 1) - Fx) + dtdy * (_sh(Fy, 1, 1) - Fy)), Fx, Fy


class PeekStage:
    """One PEEK stage: explicit advection, pressure system, and after the
    solve the flux correction; for SP111 also the conservative update.
    advection: "rusanov" (advance_rusanov_adv) or "muscl_rusanov"."""

    def __init__(self, gamma=1.4, eps=1.0, advection="rusanov", ke_method="ex",
                 enthalpy_method="reuse_pressure"):
        if ke_method not in ("ex", "split", "conservative_kinetic", "fi"):
            raise ValueError("ke_method: ex, split, conservative_kinetic or fi (not linearised)")
        if enthalpy_method not in ("reuse_pressure", "conservative", "conservative_ex"):
            raise ValueError("enthalpy_method: reuse_pressure, conservative or conservative_ex")
        self.gamma = gamma
        self.eps = eps
        self.muscl = advection == "muscl_rusanov"
        self.split = ke_method == "split"
        self.fi = ke_method == "fi"
        self.ke_method = ke_method
        # compute_enthalpy (IMEX_O1.cpp:177-229): FI always from the pressure
        self.enthalpy = "reuse_pressure" if self.fi else enthalpy_method

    def assemble(self, U, p, dt, dx, dy=None, U_exp=None, local=True):
        """compute_flux_imex_o1 up to the solve (IMEX_O1.cpp:64-125): the
        advective fluxes of U_exp (default U) update U (the implicit stage
        state); local: the MUSCL local half step (SP111 only).  Returns
        Uex, enth, rhs, a, bX (ny, nx+1), bY (ny+1, nx), alpha, beta, the
        velocity term gamma_c, cX, cY (split: -eps dt / 2 and the velocity of
        U_exp, compute_velocity and setup_pressure_op; ex: 0, None) and the
        advective fluxes Fx, Fy."""
        dy = dx if dy is None else dy
        g, eps = self.gamma, self.eps
        Uexp = U if U_exp is None else U_exp
        Uex, Fx, Fy = advect(Uexp, U, dt, dx, dy, eps, self.muscl, local)
        if self.enthalpy == "reuse_pressure":
            enth = p * g / ((g - 1) * Uex[0])
        else:
            # conservative: vol_enthalpy(U_exp) / rho_ex; conservative_ex:
            # vol_enthalpy(U_ex) / rho_ex (euler_K.H:41-50: gamma (E - ke));
            # both replace the pressure by that of U_ex (the solve's guess)
            src = Uexp if self.enthalpy == "conservative" else Uex
            enth = g * (src[3] - kinetic_energy(src, eps)) / Uex[0]
            p = pressure_from_state(Uex, g, eps)
        if self.split:
            # compute_ke, split: 0.5 eps (m_in . m_ex) / rho_in, in = U_exp
            ke = 0.5 * eps * (Uexp[1] * Uex[1] + Uexp[2] * Uex[2]) / Uexp[0]
            cX, cY = Uexp[1] / Uexp[0], Uexp[2] / Uexp[0]
            gamma_c = -0.5 * eps * dt
        elif self.fi:
            # compute_ke, not linearised: the kinetic energy of the implicit
            # stage state (statenew = statein_imp, IMEX_O1.cpp:108-109)
            ke = kinetic_energy(U, eps)
            cX = cY = None
            gamma_c = 0.0
        elif self.ke_method == "conservative_kinetic":
            ke = kinetic_energy(Uexp, eps)   # compute_ke: of statein = U_exp
            cX = cY = None
            gamma_c = 0.0
        else:
            ke = kinetic_energy(Uex, eps)
            cX = cY = None
            gamma_c = 0.0
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
        return dict(Uex=Uex, enth=enth, rhs=rhs, a=a, bX=bX, bY=bY, ke=ke, p=p,
                    alpha=1.0, beta=dt * dt, gamma_c=gamma_c, cX=cX, cY=cY, Fx=Fx, Fy=Fy)

    def picard_update(self, st, p, dt, dx, dy=None):
        """The end of one Picard iteration (picard_iterate, IMEX_O1.cpp:
        755-781) after the solve gave p: the momentum update_momentum
        (m_ex - hdtde (p(+) - p(-))), the enthalpy from p, the kinetic energy
        of the updated state, the source and the face coefficients.  Returns
        the updated system (a new dict)."""
        dy = dx if dy is None else dy
        g, eps = self.gamma, self.eps
        ex = st["Uex"]
        hdtdx = 0.5 * dt * (1.0 / dx)
        hdtdy = 0.5 * dt * (1.0 / dy)
        hdtdxe, hdtdye = hdtdx / eps, hdtdy / eps
        mx = ex[1] - hdtdxe * (_sh(p, 0, 1) - _sh(p, 0, -1))
        my = ex[2] - hdtdye * (_sh(p, 1, 1) - _sh(p, 1, -1))
        enth = p * g / ((g - 1) * ex[0])
        ke = eps * (mx * mx + my * my) / (2 * ex[0])
        rhs = eps * (ex[3] - ke - (hdtdx * (_sh(enth, 0, 1) * _sh(ex[1], 0, 1)
                                            - _sh(enth, 0, -1) * _sh(ex[1], 0, -1))
                                   + hdtdy * (_sh(enth, 1, 1) * _sh(ex[2], 1, 1)
                                              - _sh(enth, 1, -1) * _sh(ex[2], 1, -1))))
        hx = 0.5 * (_sh(enth, 0, -1) + enth)
        hy = 0.5 * (_sh(enth, 1, -1) + enth)
        ny, nx = p.shape
        bX = np.empty((ny, nx + 1))
        bX[:, :nx] = hx
        bX[:, nx] = hx[:, 0]
        bY = np.empty((ny + 1, nx))
        bY[:ny] = hy
        bY[ny] = hy[0]
        out = dict(st)
        out.update(enth=enth, rhs=rhs, bX=bX, bY=bY, m_new=(mx, my))
        return out

    def fluxes(self, st, p_new, dt, dx, dy=None):
        """update_momentum_flux / update_energy_flux (IMEX_O1.cpp:485-566):
        the stage's face fluxes after the solve (face i-1/2 at index i)."""
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
        return Fx, Fy

    def finish(self, U, st, p_new, dt, dx, dy=None):
        """SP111: the flux corrections, then the conservative update with the
        weight b = 1 (sum_fluxes multiplies by 1 exactly)."""
        dy = dx if dy is None else dy
        Fx, Fy = self.fluxes(st, p_new, dt, dx, dy)
        return conservative_update(U, [Fx], [Fy], [1.0], dt, dx, dy)


def conservative_update(U, Fxs, Fys, b, dt, dx, dy):
    """sum_fluxes (IMEX_RK.cpp:196-241: F = F_0 b_0, then F += F_s b_s) and
    the one-flux conservative_update (Update.H:100-130):
    U - dtdx (Fx(i+1) - Fx(i)) - dtdy (Fy(j+1) - Fy(j)), dtdx = dt (1/dx)."""
    Fx = Fxs[0] * b[0]
    Fy = Fys[0] * b[0]
    for s in range(1, len(b)):
        Fx = Fx + Fxs[s] * b[s]
        Fy = Fy + Fys[s] * b[s]
    dtdx = dt * (1.0 / dx)
    dtdy = dt * (1.0 / dy)
    return U - dtdx * (_sh(Fx, 0, 1) - Fx) - dtdy * (_sh(Fy, 1, 1) - Fy)


def stage_states(U, Fxs, Fys, dt_imp, dt_exp, dx, dy):
    """conservative_update into the two stage states (Update.H:218-302):
    out = U, then for each earlier stage s
    out -= dt_s (1/dx) (Fx_s(i+1) - Fx_s(i)) + dt_s (1/dy) (Fy_s(j+1) - Fy_s(j))
    with dt_s = A[stage][s] dt for the implicit and the explicit row."""
    idx, idy = 1.0 / dx, 1.0 / dy
    out1, out2 = U.copy(), U.copy()
    for s in range(len(Fxs)):
        dFx = _sh(Fxs[s], 0, 1) - Fxs[s]
        dFy = _sh(Fys[s], 1, 1) - Fys[s]
        out1 = out1 - ((dt_imp[s] * idx) * dFx + (dt_imp[s] * idy) * dFy)
        out2 = out2 - ((dt_exp[s] * idx) * dFx + (dt_exp[s] * idy) * dFy)
    return out1, out2


def pressure_ek(U_new, U_old, adt, dx, dy, gamma, eps, muscl):
    """compute_pressure, EK branch (IMEX_O1.cpp:345-384): the previous
    stage's explicit state advanced by its own advective fluxes over adt
    (MUSCL without the local step), then
    p = (gamma - 1)(E_new - 0.5 eps (m_ex . m_ex) / rho_new)."""
    ex, _, _ = advect(U_old, U_old, adt, dx, dy, eps, muscl, False)
    return (gamma - 1) * (U_new[3] - 0.5 * eps * (ex[1] * ex[1] + ex[2] * ex[2]) / U_new[0])


def pressure_sk(U_new, U_old, gamma, eps):
    """compute_pressure, SK branch (IMEX_O1.cpp:326-344):
    p = (gamma - 1)(E_new - 0.5 eps (m_old . m_new) / rho_old)."""
    return (gamma - 1) * (U_new[3] - 0.5 * eps * (U_old[1] * U_new[1] + U_old[2] * U_new[2])
                          / U_old[0])


class ButcherTableau:
    """IMEXButcherTableau (ButcherTableau.cpp:5-101)."""
    NAMES = ("SP111", "SSP222", "SASSP322", "SASSP332", "SSP433")

    def __init__(self, name):
        if name == "SP111":
            self.n_imp, self.n_exp, self.order = 1, 1, 1
            self.A_imp, self.A_exp = [[1.0]], [[0.0]]
            self.b_imp = self.b_exp = [1.0]
        elif name == "SSP222":
            self.n_imp, self.n_exp, self.order = 2, 2, 2
            al = 1 - 1 / math.sqrt(2)
            self.A_imp = [[al, 0.0], [1 - 2 * al, al]]
            self.A_exp = [[0.0, 0.0], [1.0, 0.0]]
            self.b_imp = self.b_exp = [0.5, 0.5]
        elif name == "SASSP322":
            self.n_imp, self.n_exp, self.order = 3, 2, 2
            self.A_imp = [[0.5, 0.0, 0.0], [-0.5, 0.5, 0.0], [0.0, 0.5, 0.5]]
            self.A_exp = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0]]
            self.b_imp = self.b_exp = [0.0, 0.5, 0.5]
        elif name == "SASSP332":
            self.n_imp, self.n_exp, self.order = 3, 3, 2
            t = 1.0 / 3
            self.A_imp = [[0.25, 0.0, 0.0], [0.0, 0.25, 0.0], [t, t, t]]
            self.A_exp = [[0.0, 0.0, 0.0], [0.5, 0.0, 0.0], [0.5, 0.5, 0.0]]
            self.b_imp = self.b_exp = [t, t, t]
        elif name == "SSP433":
            self.n_imp, self.n_exp, self.order = 4, 3, 3
            al, be, et = 0.24169426078821, 0.06042356519705, 0.12915286960590
            self.A_imp = [[al, 0.0, 0.0, 0.0], [-al, al, 0.0, 0.0], [0.0, 1 - al, al, 0.0],
                          [be, et, 0.5 - be - et - al, al]]
            self.A_exp = [[0.0, 0.0, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0], [0.0, 1.0, 0.0, 0.0],
                          [0.0, 0.25, 0.25, 0.0]]
            self.b_imp = self.b_exp = [0.0, 1.0 / 6, 1.0 / 6, 2.0 / 3]
        else:
            raise ValueError("unknown IMEX Butcher tableau %r" % (name,))
        self.name = name

    def imp_row(self, r, mult):
        """get_A_imp_row (ButcherTableau.H:53-59)."""
        return [a * mult for a in self.A_imp[r]]

    def exp_row(self, r, mult):
        """get_A_exp_row (ButcherTableau.H:61-67): n_imp entries."""
        return [self.A_exp[r][i] * mult for i in range(self.n_imp)]


def advance_imex_rk(U, p, dt, dx, dy, tableau, solve, gamma=1.4, eps=1.0, advection="rusanov",
                    record=None, ke_method="ex", picard_max_iter=2,
                    enthalpy_method="reuse_pressure", residual=None, res_tol=1e-12):
    """advance_imex_rk (IMEX_RK.cpp:22-96) on a 2D periodic level.  solve(sy,
    p0) returns the stage's pressure (sy: the PeekStage.assemble system,
    p0: the initial guess).  Returns (U_new, p) with p the last stage's
    pressure.  ke_method: "ex" (PEEK), "split" (PESK: the system carries
    the velocity term gamma_c, cX, cY) or "fi" (not linearised: each stage
    solves picard_max_iter times, updating the system after every solve;
    later stages start from the state's own pressure).  record, a list, receives each stage's system, its states, the
    initial and solved pressure and its corrected fluxes (Fx, Fy; the
    advective ones as Fx_adv, Fy_adv)."""
    tab = ButcherTableau(tableau) if isinstance(tableau, str) else tableau
    st = PeekStage(gamma, eps, advection, ke_method, enthalpy_method)
    Fxs, Fys, exps = [], [], []
    for s in range(tab.n_imp):
        Uimp, Uexp = stage_states(U, Fxs, Fys, tab.imp_row(s, dt)[:s], tab.exp_row(s, dt)[:s],
                                  dx, dy)
        adt = tab.A_imp[s][s] * dt
        if s > 0 and st.fi:
            p = pressure_from_state(Uexp, gamma, eps)   # compute_pressure, FI
        elif s > 0 and st.split:
            p = pressure_sk(Uexp, exps[-1], gamma, eps)
        elif s > 0 and ke_method == "ex":
            p = pressure_ek(Uexp, exps[-1], adt, dx, dy, gamma, eps, st.muscl)
        elif s > 0:
            p = pressure_from_state(Uexp, gamma, eps)    # any other method
        sy = st.assemble(Uimp, p, adt, dx, dy, U_exp=Uexp, local=tab.name == "SP111")
        p = sy["p"]      # the conservative enthalpy methods replace it
        p_in = p
        if st.fi:
            for _ in range(picard_max_iter):
                # imex.compute_residual (IMEX_O1.cpp:736-751): residual(sy, p)
                # is ||b - A p|| / ||b|| of the current system
                if residual is not None and abs(residual(sy, p)) < res_tol:
                    break
                p = solve(sy, p)
                sy = st.picard_update(sy, p, adt, dx, dy)
        else:
            p = solve(sy, p)
        Fx, Fy = st.fluxes(sy, p, adt, dx, dy)
        Fxs.append(Fx)
        Fys.append(Fy)
        exps.append(Uexp)
        if record is not None:
            rec = dict(sy)
            rec.update(U_imp=Uimp, U_exp=Uexp, p_in=p_in, p=p, adt=adt, Fx=Fx, Fy=Fy,
                       Fx_adv=sy["Fx"], Fy_adv=sy["Fy"])
            record.append(rec)
    return conservative_update(U, Fxs, Fys, tab.b_imp, dt, dx, dy), p


def shock_indicator(U, dx, dy, gamma, eps, k1):
    """compute_shock_indicator (IMEX_RK.cpp:243-345) on a periodic level:
    per cell, from the primitive states of the cell and its four face
    neighbours (primv_from_consv, sound_speed: Euler.H:167-198, 245-254),
    c_min = the smallest sound speed, div_u = 0.5 (idx (u_E - u_W) +
    idy (v_N - v_S)), premax = fmin(1, fmax(0, -(div_u + k1 c_min) /
    (k1 c_min))); the indicator is the maximum of premax over the 3 x 3
    neighbourhood.  Returns (indicator, its maximum)."""
    rho = U[0]
    u = U[1] / U[0]
    v = U[2] / U[0]
    usq = u * u + v * v
    p = (gamma - 1) * (U[3] - 0.5 * eps * U[0] * usq)
    c = np.sqrt(gamma * p / (rho * eps))
    c_min = c
    for d, k in ((0, 1), (0, -1), (1, 1), (1, -1)):
        cn = _sh(c, d, k)
        c_min = np.where(cn < c_min, cn, c_min)
    idx, idy = 1.0 / dx, 1.0 / dy
    div_u = 0.5 * (idx * (_sh(u, 0, 1) - _sh(u, 0, -1)) + idy * (_sh(v, 1, 1) - _sh(v, 1, -1)))
    premax = np.fmin(1.0, np.fmax(0.0, -(div_u + k1 * c_min) / (k1 * c_min)))
    ind = premax.copy()
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            nb = _sh(_sh(premax, 0, di), 1, dj)
            ind = np.where(nb > ind, nb, ind)
    return ind, float(ind.max())


def advance_imex_rk_stab(U, p, dt, dx, dy, tableau, solve, gamma=1.4, eps=1.0,
                         advection="rusanov", ke_method="ex", k1=1e-3, record=None):
    """advance_imex_rk_stab (IMEX_RK.cpp:98-194): the tableau's step; for a
    tableau of order above 1, where the shock indicator of the new state is
    positive, the convex combination ind U_o1 + (1 - ind) U (and likewise
    the pressure) with the first-order step U_o1 (SP111, Rusanov fluxes,
    from the incoming pressure).  Returns (U_new, p)."""
    tab = ButcherTableau(tableau) if isinstance(tableau, str) else tableau
    U_hi, p_hi = advance_imex_rk(U, p, dt, dx, dy, tab, solve, gamma, eps, advection,
                                 ke_method=ke_method)
    if tab.order <= 1:
        return U_hi, p_hi
    ind, ind_max = shock_indicator(U_hi, dx, dy, gamma, eps, k1)
    if record is not None:
        record.update(ind=ind, ind_max=ind_max, U_hi=U_hi, p_hi=p_hi)
    if ind_max <= 0:
        return U_hi, p_hi
    U_lo, p_lo = advance_imex_rk(U, p, dt, dx, dy, "SP111", solve, gamma, eps, "rusanov",
                                 ke_method=ke_method)
    if record is not None:
        record.update(U_lo=U_lo, p_lo=p_lo)
    on = ind > 0
    U_new = np.where(on, ind * U_lo + (1 - ind) * U_hi, U_hi)
    p_new = np.where(on, ind * p_lo + (1 - ind) * p_hi, p_hi)
    return U_new, p_new
