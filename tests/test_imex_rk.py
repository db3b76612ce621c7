"""CPU tests of the oracle's IMEX Runge-Kutta step (advance_imex_rk in
oracle/imex_oracle.py), the checker of the bx_imex_rk_* kernels: the
reference's tableau tests, the tableaux' order conditions, SP111 through
the stage loop, the time order of SSP222, conservation and uniform-flow
preservation with a real pressure solve."""
import math
import os
import sys

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))

import helm_oracle as ho  # noqa: E402
import imex_oracle as io  # noqa: E402
from bamrex_amd import imex  # noqa: E402


def state_from_primv(rho, u, v, p, gamma, eps):
    U = np.empty((4,) + rho.shape)
    U[0] = rho
    U[1] = rho * u
    U[2] = rho * v
    U[3] = p / (gamma - 1) + 0.5 * eps * rho * (u * u + v * v)
    return U


def direct_solve(dx, dy):
    """A solve callback for advance_imex_rk: sparse LU of the periodic
    5-point operator alpha a p - beta div(b grad p) + gamma_c c . grad p
    (centred differences), assembled here and checked against the oracle's
    operator (helm_oracle.apply) on a random vector before use."""
    def solve(sy, p0):
        a, bX, bY, alpha, beta = sy["a"], sy["bX"], sy["bY"], sy["alpha"], sy["beta"]
        ny, nx = a.shape
        idx = np.arange(nx * ny).reshape(ny, nx)
        rows, cols, vals = [], [], []
        bw, be = bX[:, :nx], bX[:, 1:]
        bs, bn = bY[:ny], bY[1:]
        diag = alpha * a + beta * ((bw + be) / (dx * dx) + (bs + bn) / (dy * dy))
        gc = sy.get("gamma_c", 0.0)
        cX = sy.get("cX") if gc != 0.0 else None
        cx = np.zeros((ny, nx)) if cX is None else gc * cX * (0.5 / dx)
        cy = np.zeros((ny, nx)) if cX is None else gc * sy["cY"] * (0.5 / dy)
        for nb, w in ((np.roll(idx, 1, 1), -beta * bw / (dx * dx) - cx),
                      (np.roll(idx, -1, 1), -beta * be / (dx * dx) + cx),
                      (np.roll(idx, 1, 0), -beta * bs / (dy * dy) - cy),
                      (np.roll(idx, -1, 0), -beta * bn / (dy * dy) + cy)):
            rows.append(idx.ravel())
            cols.append(nb.ravel())
            vals.append(w.ravel())
        rows.append(idx.ravel())
        cols.append(idx.ravel())
        vals.append(diag.ravel())
        A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(nx * ny, nx * ny))
        x = np.random.default_rng(0).random((ny, nx))
        if cX is None:
            ref = ho.apply(x, a, bX, bY, alpha, beta, dx, dy)
        else:
            ref = ho.apply(x, a, bX, bY, alpha, beta, dx, dy, gc, cX, sy["cY"])
        assert np.allclose((A @ x.ravel()).reshape(ny, nx), ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
        return spla.spsolve(A.tocsc(), sy["rhs"].ravel()).reshape(ny, nx)
    return solve


def test_tableaux_match_the_reference_tests():
    """Tests/IMEXButcherTableaux01.cpp: sizes, orders and entries."""
    t = io.ButcherTableau("SP111")
    assert (t.n_imp, t.n_exp, t.order) == (1, 1, 1)
    assert t.A_imp == [[1]] and t.A_exp == [[0]] and t.b_imp == [1] and t.b_exp == [1]
    t = io.ButcherTableau("SSP222")
    assert (t.n_imp, t.n_exp, t.order) == (2, 2, 2)
    g = 1 - 1.0 / math.sqrt(2.0)
    assert t.A_exp == [[0, 0], [1, 0]]
    assert np.allclose(t.A_imp, [[g, 0], [1 - 2 * g, g]], atol=1e-12, rtol=0)
    assert t.b_imp == [0.5, 0.5]
    t = io.ButcherTableau("SASSP322")
    assert (t.n_imp, t.n_exp, t.order) == (3, 2, 2)
    assert t.A_exp == [[0, 0, 0], [0, 0, 0], [0, 1, 0]]
    assert t.A_imp == [[0.5, 0, 0], [-0.5, 0.5, 0], [0, 0.5, 0.5]]
    assert t.b_imp == [0, 0.5, 0.5]
    t = io.ButcherTableau("SASSP332")
    assert (t.n_imp, t.n_exp, t.order) == (3, 3, 2)
    assert t.A_exp == [[0, 0, 0], [0.5, 0, 0], [0.5, 0.5, 0]]
    assert t.A_imp[0] == [0.25, 0, 0] and t.A_imp[1] == [0, 0.25, 0]
    assert np.allclose(t.A_imp[2], [1 / 3.0] * 3, atol=1e-12, rtol=0)
    assert np.allclose(t.b_imp, [1 / 3.0] * 3, atol=1e-12, rtol=0)
    t = io.ButcherTableau("SSP433")
    assert (t.n_imp, t.n_exp, t.order) == (4, 3, 3)
    assert t.A_exp == [[0, 0, 0, 0], [0, 0, 0, 0], [0, 1, 0, 0], [0, 0.25, 0.25, 0]]
    al, be, et = 0.24169426078821, 0.06042356519705, 0.12915286960590
    assert np.allclose(t.A_imp, [[al, 0, 0, 0], [-al, al, 0, 0], [0, 1 - al, al, 0],
                                 [be, et, 0.5 - be - et - al, al]], atol=1e-12, rtol=0)
    for name in io.ButcherTableau.NAMES:
        t = io.ButcherTableau(name)
        assert t.b_imp == t.b_exp
        for r in range(t.n_imp):
            assert t.imp_row(r, 0.37) == [a * 0.37 for a in t.A_imp[r]]
            assert len(t.exp_row(r, 0.37)) == t.n_imp
    with pytest.raises(ValueError):
        io.ButcherTableau("SSP999")


@pytest.mark.parametrize("name", io.ButcherTableau.NAMES)
def test_tableaux_satisfy_their_order_conditions(name):
    """Additive Runge-Kutta order conditions (Pareschi & Russo 2005) up to
    each tableau's stated order, for both parts and their coupling: a
    mistyped entry fails one of them."""
    t = io.ButcherTableau(name)
    Ai, Ae = np.array(t.A_imp), np.array(t.A_exp)
    b = np.array(t.b_imp)
    ci, ce = Ai.sum(1), Ae.sum(1)
    tol = 1e-10
    assert abs(b.sum() - 1) < tol
    # explicit part strictly lower triangular, implicit lower triangular
    assert np.all(np.triu(Ae) == 0) and np.all(np.triu(Ai, 1) == 0)
    if t.order >= 2:
        for c in (ci, ce):
            assert abs(b @ c - 0.5) < tol
    if t.order >= 3:
        for c1 in (ci, ce):
            for c2 in (ci, ce):
                assert abs(b @ (c1 * c2) - 1 / 3.0) < tol
        for A in (Ai, Ae):
            for c in (ci, ce):
                assert abs(b @ (A @ c) - 1 / 6.0) < tol


def test_sp111_stage_loop_equals_the_single_stage():
    """SP111 through the stage loop (stage states equal to U, the MUSCL
    local step, fluxes summed with weight 1) gives PeekStage's step bit for
    bit, for both advection fluxes."""
    gamma, eps = 1.4, 0.6
    rng = np.random.default_rng(5)
    ny, nx = 12, 20
    U = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    p = 2 + rng.random((ny, nx))
    dx, dy, dt = 1.0 / nx, 0.8 / ny, 0.011
    for adv in ("rusanov", "muscl_rusanov"):
        def stand_in(sy, p0):
            return io.pressure_from_state(sy["Uex"], gamma, eps)
        U1, p1 = io.advance_imex_rk(U, p, dt, dx, dy, "SP111", stand_in, gamma, eps, adv)
        st = io.PeekStage(gamma, eps, adv)
        sy = st.assemble(U, p, dt, dx, dy)
        pn = stand_in(sy, p)
        assert np.array_equal(p1, pn)
        assert np.array_equal(U1, st.finish(U, sy, pn, dt, dx, dy))


def test_stage_states_follow_the_tableau_rows():
    """The SSP222 second stage: U_imp = U - (1 - 2 alpha) dt div F_0 and
    U_exp = U - dt div F_0, and the EK pressure is formed from the first
    explicit state (U) advanced over adt = alpha dt -- checked against a
    per-cell evaluation in plain Python on a small grid."""
    gamma, eps = 1.4, 1.0
    rng = np.random.default_rng(8)
    ny, nx = 5, 7
    U = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    p = io.pressure_from_state(U, gamma, eps)
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.02
    rec = []
    io.advance_imex_rk(U, p, dt, dx, dy, "SSP222", lambda sy, p0: p0 * 1.001, gamma, eps,
                       "muscl_rusanov", record=rec)
    al = 1 - 1 / math.sqrt(2)
    assert rec[0]["adt"] == al * dt and rec[1]["adt"] == al * dt
    assert np.array_equal(rec[0]["U_imp"], U) and np.array_equal(rec[0]["U_exp"], U)
    F0x, F0y = rec[0]["Fx"], rec[0]["Fy"]
    for k in range(4):
        for j in range(ny):
            for i in range(nx):
                dFx = F0x[k, j, (i + 1) % nx] - F0x[k, j, i]
                dFy = F0y[k, (j + 1) % ny, i] - F0y[k, j, i]
                wi = (1 - 2 * al) * dt
                imp = U[k, j, i] - ((wi * (1 / dx)) * dFx + (wi * (1 / dy)) * dFy)
                exp_ = U[k, j, i] - ((dt * (1 / dx)) * dFx + (dt * (1 / dy)) * dFy)
                assert rec[1]["U_imp"][k, j, i] == imp
                assert rec[1]["U_exp"][k, j, i] == exp_
    # EK pressure of stage 1: momentum of U advanced by its own fluxes
    # (MUSCL without the local step) over adt, density and energy of U_exp
    ex, _, _ = io.advect(U, U, al * dt, dx, dy, eps, True, False)
    Ue = rec[1]["U_exp"]
    for j in range(ny):
        for i in range(nx):
            m2 = ex[1, j, i] * ex[1, j, i] + ex[2, j, i] * ex[2, j, i]
            pe = (gamma - 1) * (Ue[3, j, i] - 0.5 * eps * m2 / Ue[0, j, i])
            assert rec[1]["p_in"][j, i] == pe


def _wave_errors(n, tableau, advection, solve_kind):
    """L1 errors of rho and p after advecting rho = 1 + 0.2 sin(2 pi x) with
    u = 1, p = 1 to t = 0.25 on an n x 4 periodic grid at CFL 0.4."""
    gamma, eps = 1.4, 1.0
    x = (np.arange(n) + 0.5) / n
    rho0 = 1.0 + 0.2 * np.sin(2 * np.pi * x)[None, :].repeat(4, 0)
    U = state_from_primv(rho0, np.ones((4, n)), np.zeros((4, n)), np.ones((4, n)), gamma, eps)
    dx, dy = 1.0 / n, 0.25
    solve = direct_solve(dx, dy) if solve_kind == "direct" else (lambda sy, p0: np.ones_like(p0))
    p = np.ones((4, n))
    t, t_end = 0.0, 0.25
    while t < t_end - 1e-14:
        dt = min(0.4 * dx, t_end - t)
        U, p = io.advance_imex_rk(U, p, dt, dx, dy, tableau, solve, gamma, eps, advection)
        t += dt
    exact = 1.0 + 0.2 * np.sin(2 * np.pi * (x - t_end))
    e_rho = np.mean(np.abs(U[0, 0] - exact))
    e_p = np.mean(np.abs(io.pressure_from_state(U, gamma, eps) - 1.0))
    return e_rho, e_p


def test_ssp222_is_second_order_in_time():
    """MUSCL fluxes without the local step are second order in space only:
    the SSP222 stages make the step second order (error ratio ~4 per
    halving of dx and dt), where a first-order time combination -- the
    planted mistake of an SSP222 whose explicit row takes half the first
    stage's flux -- drops to ratio ~2."""
    e = [_wave_errors(n, "SSP222", "muscl_rusanov", "stand_in")[0] for n in (32, 64, 128)]
    assert e[1] / e[2] > 3.0, e
    bad = io.ButcherTableau("SSP222")
    bad.A_exp = [[0.0, 0.0], [0.5, 0.0]]
    eb = [_wave_errors(n, bad, "muscl_rusanov", "stand_in")[0] for n in (32, 64, 128)]
    assert eb[1] / eb[2] < 2.5, eb
    assert eb[2] > 3 * e[2], (e, eb)


def test_solved_pressure_of_an_advected_wave():
    """With the pressure solved at every stage: SP111 keeps the uniform
    pressure of the advected wave to rounding (its energy update and
    pressure equation are one identity), and the multi-stage tableaux keep
    it to first order in dx.  The first-order part comes from the EK
    pressure of the later stages (compute_pressure, IMEX_O1.cpp:345-384),
    which combines the momentum of the previous explicit state advanced
    over A_ss dt with the density and energy of the current explicit state
    (advanced over its whole row of A_exp); rho still converges faster than
    first order."""
    for adv in ("rusanov", "muscl_rusanov"):
        r = [_wave_errors(n, "SP111", adv, "direct") for n in (32, 64)]
        assert max(x[1] for x in r) < 1e-14, r
    r = [_wave_errors(n, "SSP222", "muscl_rusanov", "direct") for n in (32, 64, 128)]
    assert r[1][0] / r[2][0] > 2.4, r
    assert 1.8 < r[1][1] / r[2][1] < 2.3 and r[2][1] < 1e-3, r


@pytest.mark.parametrize("tableau", ["SSP222", "SASSP322", "SASSP332", "SSP433"])
def test_multistage_conserves_and_keeps_uniform_flow(tableau):
    """Every tableau: uniform flow unchanged, and the Gresho vortex's mass,
    momentum and energy totals conserved with a real pressure solve."""
    gamma, eps = 1.4, 1.0
    n = 16
    U = state_from_primv(np.full((n, n), 1.2), np.full((n, n), 0.3), np.full((n, n), -0.2),
                         np.full((n, n), 2.0), gamma, eps)
    dx = 1.0 / n
    U1, p1 = io.advance_imex_rk(U, io.pressure_from_state(U, gamma, eps), 0.02, dx, dx, tableau,
                                direct_solve(dx, dx), gamma, eps, "muscl_rusanov")
    assert np.allclose(U1, U, rtol=1e-12, atol=0)
    assert np.allclose(p1, 2.0, rtol=1e-12, atol=0)
    Ug, dxg = imex.gresho_state(32)
    pg = io.pressure_from_state(Ug, gamma, eps)
    dtg = io.advective_dt(Ug, dxg)
    Un, pn = io.advance_imex_rk(Ug, pg, dtg, dxg, dxg, tableau, direct_solve(dxg, dxg), gamma, eps,
                                "muscl_rusanov")
    scale = np.abs(Ug).sum(axis=(1, 2))
    assert np.all(np.abs(Un.sum(axis=(1, 2)) - Ug.sum(axis=(1, 2))) <= 1e-13 * scale)
    # the vortex is steady: one step at 32^2 moves rho by the discretisation
    # error (about 1e-2, as SP111 does) and the pressure by ~1e-5 relative
    assert np.max(np.abs(Un[0] - Ug[0])) < 2e-2
    assert np.max(np.abs(pn - pg)) < 1e-4 * pg.max()


def test_product_tableaux_equal_the_oracle_tableaux():
    """The host's tableaux (bamrex_amd.imex.ButcherTableau, which drives the
    GPU stages) are the oracle's, entry for entry and row for row."""
    for name in io.ButcherTableau.NAMES:
        a, b = imex.ButcherTableau(name), io.ButcherTableau(name)
        assert (a.n_imp, a.n_exp, a.order) == (b.n_imp, b.n_exp, b.order)
        assert a.A_imp == b.A_imp and a.A_exp == b.A_exp and a.b == b.b_imp
        for r in range(a.n_imp):
            assert a.imp_row(r, 0.013) == b.imp_row(r, 0.013)
            assert a.exp_row(r, 0.013) == b.exp_row(r, 0.013)
    assert imex.ButcherTableau.NAMES == io.ButcherTableau.NAMES


def test_split_ke_energy_matches_pressure():
    """Tests/IMEX_O1_01.cpp:56-117 (SplitKEEnergyMatch, eps 0.5, gamma 1.6,
    dt 0.01, Rusanov): with the split kinetic energy and the operator's
    velocity term, the new energy equals
    p/(gamma-1) + 0.5 eps (m_in . m_new) / rho_in -- a wrong sign of
    gamma_c or a wrong split product breaks it."""
    n = 32
    gamma, eps, dt = 1.6, 0.5, 0.01
    rng = np.random.default_rng(4)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.05 * rng.random((n, n))
    u = 0.4 * np.sin(2 * np.pi * Y)
    v = -0.3 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y))
    U = state_from_primv(rho, u, v, p, gamma, eps)
    dx = 1.0 / n
    for adv in ("rusanov", "muscl_rusanov"):
        U1, pn = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps, adv,
                                    ke_method="split")
        expected = pn / (gamma - 1) + 0.5 * eps / U[0] * (U[1] * U1[1] + U[2] * U1[2])
        assert np.max(np.abs(U1[3] - expected) / np.abs(expected)) <= 1e-12, adv
    # the identity fails with the explicit-state kinetic energy (PEEK)
    U1, pn = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps,
                                "rusanov")
    expected = pn / (gamma - 1) + 0.5 * eps / U[0] * (U[1] * U1[1] + U[2] * U1[2])
    assert np.max(np.abs(U1[3] - expected) / np.abs(expected)) > 1e-8


def test_split_system_and_sk_pressure_per_cell():
    """The split system's kinetic energy, velocity term and the SK pressure
    of SSP222's second stage against per-cell plain-Python evaluations."""
    gamma, eps = 1.4, 0.8
    rng = np.random.default_rng(9)
    ny, nx = 4, 6
    U = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    p = io.pressure_from_state(U, gamma, eps)
    dx, dy, dt = 1.0 / nx, 1.0 / ny, 0.02
    rec = []
    io.advance_imex_rk(U, p, dt, dx, dy, "SSP222", lambda sy, p0: p0 * 0.999, gamma, eps,
                       "muscl_rusanov", record=rec, ke_method="split")
    al = 1 - 1 / math.sqrt(2)
    r0, r1 = rec
    assert r0["gamma_c"] == -0.5 * eps * (al * dt)
    hx = 0.5 * (al * dt) * (1.0 / dx)
    hy = 0.5 * (al * dt) * (1.0 / dy)
    ex, h = r0["Uex"], r0["enth"]
    for j in range(ny):
        for i in range(nx):
            assert r0["cX"][j, i] == U[1, j, i] / U[0, j, i]
            assert r0["cY"][j, i] == U[2, j, i] / U[0, j, i]
            ke = 0.5 * eps * (U[1, j, i] * ex[1, j, i] + U[2, j, i] * ex[2, j, i]) / U[0, j, i]
            ip, im, jp, jm = (i + 1) % nx, (i - 1) % nx, (j + 1) % ny, (j - 1) % ny
            rhs = eps * (ex[3, j, i] - ke
                         - (hx * (h[j, ip] * ex[1, j, ip] - h[j, im] * ex[1, j, im])
                            + hy * (h[jp, i] * ex[2, jp, i] - h[jm, i] * ex[2, jm, i])))
            assert r0["rhs"][j, i] == rhs
            Un, Uo = r1["U_exp"], U
            m2 = Uo[1, j, i] * Un[1, j, i] + Uo[2, j, i] * Un[2, j, i]
            assert r1["p_in"][j, i] == (gamma - 1) * (Un[3, j, i] - 0.5 * eps * m2 / Uo[0, j, i])


@pytest.mark.parametrize("tableau", ["SP111", "SSP222", "SSP433"])
def test_split_conserves_and_keeps_uniform_flow(tableau):
    """PESK steps: uniform flow unchanged, and the Gresho vortex's totals
    conserved with a real solve of the operator with its velocity term."""
    gamma, eps = 1.4, 1.0
    n = 16
    U = state_from_primv(np.full((n, n), 1.2), np.full((n, n), 0.3), np.full((n, n), -0.2),
                         np.full((n, n), 2.0), gamma, eps)
    dx = 1.0 / n
    U1, p1 = io.advance_imex_rk(U, io.pressure_from_state(U, gamma, eps), 0.02, dx, dx, tableau,
                                direct_solve(dx, dx), gamma, eps, "muscl_rusanov", ke_method="split")
    assert np.allclose(U1, U, rtol=1e-12, atol=0)
    assert np.allclose(p1, 2.0, rtol=1e-12, atol=0)
    Ug, dxg = imex.gresho_state(32)
    pg = io.pressure_from_state(Ug, gamma, eps)
    dtg = io.advective_dt(Ug, dxg)
    Un, pn = io.advance_imex_rk(Ug, pg, dtg, dxg, dxg, tableau, direct_solve(dxg, dxg), gamma, eps,
                                "muscl_rusanov", ke_method="split")
    scale = np.abs(Ug).sum(axis=(1, 2))
    assert np.all(np.abs(Un.sum(axis=(1, 2)) - Ug.sum(axis=(1, 2))) <= 1e-13 * scale)
    assert np.max(np.abs(pn - pg)) < 1e-4 * pg.max()


def _compressive_state(ny, nx, gamma, eps, seed):
    """A state with converging flow in the middle (negative div u)."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.1 * rng.random((ny, nx))
    u = -0.8 * np.sin(2 * np.pi * X) + 0.05 * rng.standard_normal((ny, nx))
    v = -0.5 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal((ny, nx))
    p = 1.0 + 0.1 * rng.random((ny, nx))
    return state_from_primv(rho, u, v, p, gamma, eps)


def test_shock_indicator_per_cell():
    """compute_shock_indicator (IMEX_RK.cpp:243-345) against a per-cell
    plain-Python evaluation on a small periodic grid (dx != dy, eps != 1),
    with values across 0 < ind < 1 present; uniform flow gives 0."""
    gamma, eps, k1 = 1.4, 0.7, 0.05
    ny, nx = 7, 9
    dx, dy = 1.0 / nx, 0.8 / ny
    U = _compressive_state(ny, nx, gamma, eps, 3)
    ind, ind_max = io.shock_indicator(U, dx, dy, gamma, eps, k1)

    def prim(j, i):
        j, i = j % ny, i % nx
        r = U[0, j, i]
        u, v = U[1, j, i] / r, U[2, j, i] / r
        pp = (gamma - 1) * (U[3, j, i] - 0.5 * eps * r * (u * u + v * v))
        return u, v, math.sqrt(gamma * pp / (r * eps))

    pre = np.empty((ny, nx))
    for j in range(ny):
        for i in range(nx):
            cs = [prim(j, i)[2], prim(j, i + 1)[2], prim(j, i - 1)[2], prim(j + 1, i)[2],
                  prim(j - 1, i)[2]]
            cm = min(cs)
            div = 0.5 * ((1.0 / dx) * (prim(j, i + 1)[0] - prim(j, i - 1)[0])
                         + (1.0 / dy) * (prim(j + 1, i)[1] - prim(j - 1, i)[1]))
            pre[j, i] = min(1.0, max(0.0, -(div + k1 * cm) / (k1 * cm)))
    for j in range(ny):
        for i in range(nx):
            m = max(pre[(j + dj) % ny, (i + di) % nx] for dj in (-1, 0, 1) for di in (-1, 0, 1))
            assert ind[j, i] == m, (j, i)
    assert ind_max == ind.max() and 0 < ind_max
    assert np.any((ind > 0) & (ind < 1)) and np.any(ind == 0)
    Uu = state_from_primv(np.full((4, 5), 1.1), np.full((4, 5), 0.3), np.full((4, 5), -0.2),
                          np.full((4, 5), 2.0), gamma, eps)
    assert io.shock_indicator(Uu, 0.2, 0.25, gamma, eps, k1)[1] == 0.0


def test_stabilised_step_blends_only_where_indicated():
    """advance_imex_rk_stab: on a smooth uniform flow it is the plain step;
    on a compressive flow the new state is ind U_o1 + (1 - ind) U_hi where
    ind > 0 and U_hi elsewhere, U_o1 being the SP111 Rusanov step; SP111
    itself is never blended."""
    gamma, eps, k1 = 1.4, 1.0, 0.05
    n = 16
    dx = 1.0 / n
    solve = direct_solve(dx, dx)
    Uu = state_from_primv(np.full((n, n), 1.2), np.full((n, n), 0.3), np.full((n, n), -0.2),
                          np.full((n, n), 2.0), gamma, eps)
    pu = io.pressure_from_state(Uu, gamma, eps)
    a = io.advance_imex_rk_stab(Uu, pu, 0.01, dx, dx, "SSP222", solve, gamma, eps, "muscl_rusanov",
                                k1=k1)
    b = io.advance_imex_rk(Uu, pu, 0.01, dx, dx, "SSP222", solve, gamma, eps, "muscl_rusanov")
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    U = _compressive_state(n, n, gamma, eps, 5)
    p = io.pressure_from_state(U, gamma, eps)
    rec = {}
    Us, ps = io.advance_imex_rk_stab(U, p, 0.004, dx, dx, "SSP222", solve, gamma, eps,
                                     "muscl_rusanov", k1=k1, record=rec)
    ind = rec["ind"]
    assert rec["ind_max"] > 0 and np.any(ind == 1) and np.any(ind == 0)
    lo, hi = rec["U_lo"], rec["U_hi"]
    assert np.array_equal(Us[:, ind == 0], hi[:, ind == 0])
    assert np.array_equal(Us[:, ind == 1], lo[:, ind == 1])
    mid = (ind > 0) & (ind < 1)
    assert np.array_equal(Us[:, mid], (ind * lo + (1 - ind) * hi)[:, mid])
    assert np.array_equal(ps[ind == 1], rec["p_lo"][ind == 1])
    U1, _ = io.advance_imex_rk(U, p, 0.004, dx, dx, "SP111", solve, gamma, eps, "muscl_rusanov")
    U2, _ = io.advance_imex_rk_stab(U, p, 0.004, dx, dx, "SP111", solve, gamma, eps, "muscl_rusanov",
                                    k1=k1)
    assert np.array_equal(U1, U2)


def test_fi_picard_fixed_point_satisfies_the_equation_of_state():
    """Not linearised (FI): at a converged Picard fixed point the stage's
    new energy equals p/(gamma-1) + eps |m_new|^2 / (2 rho_new) -- the
    nonlinear system the iteration solves (picard_iterate, IMEX_O1.cpp:
    702-789).  One iteration is not enough for it (negative control); the
    iterates converge."""
    n = 32
    gamma, eps, dt = 1.6, 0.5, 0.01
    rng = np.random.default_rng(4)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.05 * rng.random((n, n))
    u = 0.4 * np.sin(2 * np.pi * Y)
    v = -0.3 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y))
    U = state_from_primv(rho, u, v, p, gamma, eps)
    dx = 1.0 / n

    def eos_gap(U1, p1):
        expected = p1 / (gamma - 1) + eps * (U1[1] * U1[1] + U1[2] * U1[2]) / (2 * U1[0])
        return np.max(np.abs(U1[3] - expected) / np.abs(expected))

    for adv in ("rusanov", "muscl_rusanov"):
        U1, p1 = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps, adv,
                                    ke_method="fi", picard_max_iter=40)
        assert eos_gap(U1, p1) <= 1e-11, adv
    U1, p1 = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps,
                                "rusanov", ke_method="fi", picard_max_iter=1)
    assert eos_gap(U1, p1) > 1e-9
    gaps = []
    for it in (2, 4, 8):
        U1, p1 = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps,
                                    "rusanov", ke_method="fi", picard_max_iter=it)
        gaps.append(eos_gap(U1, p1))
    assert gaps[0] > gaps[1] > gaps[2], gaps


def test_fi_picard_update_per_cell():
    """One Picard update (update_momentum, compute_enthalpy, compute_ke of
    the updated state, pressure_source_vector, face enthalpy) against a
    per-cell plain-Python evaluation, dx != dy."""
    gamma, eps = 1.4, 0.8
    rng = np.random.default_rng(12)
    ny, nx = 4, 5
    U = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    dx, dy, dt = 1.0 / nx, 0.7 / ny, 0.02
    st = io.PeekStage(gamma, eps, "rusanov", "fi")
    sy = st.assemble(U, io.pressure_from_state(U, gamma, eps), dt, dx, dy)
    assert np.array_equal(sy["rhs"], eps * (sy["Uex"][3] - io.kinetic_energy(U, eps) - (
        (0.5 * dt * (1.0 / dx)) * (np.roll(sy["enth"] * sy["Uex"][1], -1, 1)
                                   - np.roll(sy["enth"] * sy["Uex"][1], 1, 1))
        + (0.5 * dt * (1.0 / dy)) * (np.roll(sy["enth"] * sy["Uex"][2], -1, 0)
                                     - np.roll(sy["enth"] * sy["Uex"][2], 1, 0)))))
    pn = 2 + rng.random((ny, nx))
    up = st.picard_update(sy, pn, dt, dx, dy)
    ex = sy["Uex"]
    hx, hy = 0.5 * dt * (1.0 / dx), 0.5 * dt * (1.0 / dy)
    for j in range(ny):
        for i in range(nx):
            ip, im, jp, jm = (i + 1) % nx, (i - 1) % nx, (j + 1) % ny, (j - 1) % ny
            mx = ex[1, j, i] - (hx / eps) * (pn[j, ip] - pn[j, im])
            my = ex[2, j, i] - (hy / eps) * (pn[jp, i] - pn[jm, i])
            h = lambda jj, ii: pn[jj, ii] * gamma / ((gamma - 1) * ex[0, jj, ii])  # noqa: E731
            ke = eps * (mx * mx + my * my) / (2 * ex[0, j, i])
            rhs = eps * (ex[3, j, i] - ke - (hx * (h(j, ip) * ex[1, j, ip] - h(j, im) * ex[1, j, im])
                                             + hy * (h(jp, i) * ex[2, jp, i] - h(jm, i) * ex[2, jm, i])))
            assert up["enth"][j, i] == h(j, i)
            assert up["rhs"][j, i] == rhs
            assert up["bX"][j, i] == 0.5 * (h(j, im) + h(j, i))
            assert up["bY"][j, i] == 0.5 * (h(jm, i) + h(j, i))


@pytest.mark.parametrize("tableau", ["SP111", "SSP222"])
def test_fi_conserves_and_keeps_uniform_flow(tableau):
    gamma, eps = 1.4, 1.0
    n = 16
    U = state_from_primv(np.full((n, n), 1.2), np.full((n, n), 0.3), np.full((n, n), -0.2),
                         np.full((n, n), 2.0), gamma, eps)
    dx = 1.0 / n
    U1, p1 = io.advance_imex_rk(U, io.pressure_from_state(U, gamma, eps), 0.02, dx, dx, tableau,
                                direct_solve(dx, dx), gamma, eps, "muscl_rusanov", ke_method="fi")
    assert np.allclose(U1, U, rtol=1e-12, atol=0)
    assert np.allclose(p1, 2.0, rtol=1e-12, atol=0)
    Ug, dxg = imex.gresho_state(32)
    pg = io.pressure_from_state(Ug, gamma, eps)
    dtg = io.advective_dt(Ug, dxg)
    Un, pn = io.advance_imex_rk(Ug, pg, dtg, dxg, dxg, tableau, direct_solve(dxg, dxg), gamma, eps,
                                "muscl_rusanov", ke_method="fi")
    scale = np.abs(Ug).sum(axis=(1, 2))
    assert np.all(np.abs(Un.sum(axis=(1, 2)) - Ug.sum(axis=(1, 2))) <= 1e-13 * scale)
    assert np.max(np.abs(pn - pg)) < 1e-4 * pg.max()


@pytest.mark.parametrize("enthalpy_method", ["reuse_pressure", "conservative", "conservative_ex"])
@pytest.mark.parametrize("ke_method", ["ex", "split", "conservative_kinetic"])
def test_linearised_energy_identity_for_every_method(ke_method, enthalpy_method):
    """The identity behind Tests/IMEX_O1_01.cpp's EnergyMatch tests, for
    every linearised kinetic-energy and enthalpy method: the new energy
    equals p/(gamma-1) plus the kinetic energy the source used (the energy
    flux correction and the pressure equation share the enthalpy), to the
    solve's precision."""
    n = 24
    gamma, eps, dt = 1.6, 0.5, 0.01
    rng = np.random.default_rng(6)
    x = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * Y) + 0.05 * rng.random((n, n))
    u = 0.4 * np.sin(2 * np.pi * Y)
    v = -0.3 * np.cos(2 * np.pi * X)
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (X + Y))
    U = state_from_primv(rho, u, v, p, gamma, eps)
    dx = 1.0 / n
    rec = []
    U1, pn = io.advance_imex_rk(U, p, dt, dx, dx, "SP111", direct_solve(dx, dx), gamma, eps,
                                "rusanov", record=rec, ke_method=ke_method,
                                enthalpy_method=enthalpy_method)
    r = rec[0]
    if ke_method == "split":
        ke_new = 0.5 * eps * (U[1] * U1[1] + U[2] * U1[2]) / U[0]
    else:
        ke_new = r["ke"]
    expected = pn / (gamma - 1) + ke_new
    assert np.max(np.abs(U1[3] - expected) / np.abs(expected)) <= 1e-12


def test_conservative_enthalpy_per_cell():
    """compute_enthalpy, conservative and conservative_ex (IMEX_O1.cpp:
    211-228): gamma (E - ke)/rho_ex of the explicit stage state or of the
    advected state, and the pressure replaced by that of the advected
    state; conservative_kinetic takes the kinetic energy of the explicit
    stage state -- per cell in plain Python."""
    gamma, eps = 1.4, 0.8
    rng = np.random.default_rng(14)
    ny, nx = 3, 4
    U = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    p = 2 + rng.random((ny, nx))
    for meth in ("conservative", "conservative_ex"):
        st = io.PeekStage(gamma, eps, "rusanov", "conservative_kinetic", meth)
        sy = st.assemble(U, p, 0.01, 1.0 / nx)
        ex = sy["Uex"]
        for j in range(ny):
            for i in range(nx):
                src = U if meth == "conservative" else ex
                ke_s = eps * (src[1, j, i] ** 2 + src[2, j, i] ** 2) / (2 * src[0, j, i])
                assert sy["enth"][j, i] == gamma * (src[3, j, i] - ke_s) / ex[0, j, i]
                ke_x = eps * (ex[1, j, i] ** 2 + ex[2, j, i] ** 2) / (2 * ex[0, j, i])
                assert sy["p"][j, i] == (gamma - 1) * (ex[3, j, i] - ke_x)
                ke_u = eps * (U[1, j, i] ** 2 + U[2, j, i] ** 2) / (2 * U[0, j, i])
                assert sy["ke"][j, i] == ke_u


def test_time_step_estimates_brute_force_and_nan_skipping():
    """estTimeStep of the IMEX schemes: the acoustic estimate
    (ExplicitTimeStepDimSplit with max_wave_speed) and the advective one,
    cell by cell in plain Python with dx != dy; a NaN speed (negative
    pressure) or a negative one (negative density) is passed over, as
    ReduceOpMax's max(d, s) = (d < s) ? s : d passes over it."""
    gamma, eps, cfl = 1.4, 0.6, 0.45
    rng = np.random.default_rng(21)
    ny, nx = 5, 6
    U = state_from_primv(1 + 0.3 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                         rng.uniform(-1, 1, (ny, nx)), 1 + rng.random((ny, nx)), gamma, eps)
    dx, dy = 1.0 / nx, 0.7 / ny
    sa, sv = [], []
    for j in range(ny):
        for i in range(nx):
            r, mx, my, E = U[:, j, i]
            u, v = mx / r, my / r
            p = (gamma - 1) * (E - 0.5 * eps * r * (u * u + v * v))
            sa.append(math.sqrt(gamma * p / (r * eps)) + math.sqrt(u * u + v * v))
            sv.append((abs(mx) / dx + abs(my) / dy) / r)
    dt_a = min(1e20, dx / max(sa), dy / max(sa)) * cfl
    assert io.acoustic_dt(U, dx, dy, gamma, eps, cfl) == dt_a
    assert io.advective_dt(U, dx, dy, cfl) == (1 / max(sv)) * cfl
    bad = U.copy()
    bad[3, 2, 3] = 0.0          # negative pressure: NaN sound speed
    bad[0, 4, 1] = -bad[0, 4, 1]  # negative density: negative advective speed
    sa2 = [x for k, x in enumerate(sa) if k != 2 * nx + 3]
    assert io.acoustic_dt(bad, dx, dy, gamma, eps, cfl) >= min(1e20, dx / max(sa2), dy / max(sa2)) * cfl
    sv2 = [x for k, x in enumerate(sv) if k != 4 * nx + 1]
    assert io.advective_dt(bad, dx, dy, cfl) == (1 / max(sv2)) * cfl
