"""CPU tests of the oracle's PEEK-SP111 stage (oracle/imex_oracle.py), the
checker of the bx_imex_* kernels, restating the reference's own IMEX tests."""
import os
import sys

import numpy as np
import pytest

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


def test_rusanov_flux_known_answers():
    """Tests/RusanovAdv01.cpp:116-210 (ToroTest2: rho 1, u -2 | +2, p 0.4,
    gamma 1.4): face fluxes (-2, 4, -4) inside the left state, 0 at the
    interface, (2, 4, 4) inside the right state."""
    nx, ny = 8, 2
    u = np.where(np.arange(nx) < 4, -2.0, 2.0)[None, :].repeat(ny, 0)
    U = state_from_primv(np.ones((ny, nx)), u, np.zeros((ny, nx)), np.full((ny, nx), 0.4), 1.4, 1.0)
    Fx, _ = io.rusanov_faces(U, 1.0)
    # face index i is the face between cells i-1 and i
    for i, exp in ((1, (-2, 4, -4)), (2, (-2, 4, -4)), (4, (0, 0, 0)), (5, (2, 4, 4)), (6, (2, 4, 4))):
        got = (Fx[0][0, i], Fx[1][0, i], Fx[3][0, i])
        assert np.allclose(got, exp, rtol=0, atol=1e-12), (i, got)
    assert np.all(Fx[2] == 0)


def test_peek_energy_matches_pressure_and_explicit_ke():
    """Tests/IMEX_O1_01.cpp:201-268 (ExplicitKEEnergyMatch, eps 0.5,
    gamma 1.6, dt 0.01): after the stage, the new energy equals
    p/(gamma-1) + eps |m_ex|^2 / (2 rho_ex) with p the solved pressure and
    (rho, m)_ex the Rusanov-advected state -- the energy update and the
    pressure equation are one identity, so a wrong sign or a dropped term in
    the rhs, the enthalpy fluxes or the operator breaks it."""
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
    stage = io.PeekStage(gamma, eps)
    sy = stage.assemble(U, p, dt, dx)
    pn, its, _ = ho.mg_solve(sy["rhs"], sy["a"], sy["bX"], sy["bY"], sy["alpha"], sy["beta"], dx, dx,
                             x0=p, rtol=1e-15, maxiter=60)
    Unew = stage.finish(U, sy, pn, dt, dx)
    ex = sy["Uex"]
    expected = pn / (gamma - 1) + 0.5 * eps / ex[0] * (ex[1] * ex[1] + ex[2] * ex[2])
    assert np.max(np.abs(Unew[3] - expected) / np.abs(expected)) <= 1e-12
    # density is advected only: the stage leaves the explicit density (to
    # rounding: the final update forms dt (1/dx) and subtracts the two
    # directions one after the other, the advection dt/dx and their sum)
    assert np.max(np.abs(Unew[0] - ex[0]) / ex[0]) <= 4e-16


def test_peek_step_conserves_and_keeps_uniform_flow():
    n = 16
    gamma, eps = 1.4, 1.0
    U = state_from_primv(np.full((n, n), 1.2), np.full((n, n), 0.3), np.full((n, n), -0.2),
                         np.full((n, n), 2.0), gamma, eps)
    p = io.pressure_from_state(U, gamma, eps)
    stage = io.PeekStage(gamma, eps)
    dx = 1.0 / n
    sy = stage.assemble(U, p, 0.02, dx)
    pn, its, _ = ho.mg_solve(sy["rhs"], sy["a"], sy["bX"], sy["bY"], sy["alpha"], sy["beta"], dx, dx,
                             x0=p, rtol=1e-15, maxiter=60)
    Unew = stage.finish(U, sy, pn, 0.02, dx)
    assert np.allclose(Unew, U, rtol=1e-13, atol=0)
    # Gresho vortex: mass, momentum and energy totals are conserved
    Ug, dxg = imex.gresho_state(64)
    pg = io.pressure_from_state(Ug, gamma, eps)
    dtg = io.advective_dt(Ug, dxg)
    sy = stage.assemble(Ug, pg, dtg, dxg)
    pn, its, _ = ho.mg_solve(sy["rhs"], sy["a"], sy["bX"], sy["bY"], sy["alpha"], sy["beta"], dxg, dxg,
                             x0=pg, rtol=1e-12, maxiter=60)
    Un = stage.finish(Ug, sy, pn, dtg, dxg)
    tot0, tot1 = Ug.sum(axis=(1, 2)), Un.sum(axis=(1, 2))
    scale = np.abs(Ug).sum(axis=(1, 2))
    assert np.all(np.abs(tot1 - tot0) <= 1e-13 * scale)


def test_advective_dt_brute_force_and_anisotropic_cells():
    """cfl / max (|m_x|/dx + |m_y|/dy)/rho, cell by cell in plain Python;
    with dy != dx a transposed operand or swapped spacing changes it."""
    rng = np.random.default_rng(11)
    U = np.empty((4, 5, 7))
    U[0] = rng.uniform(0.5, 2.0, (5, 7))
    U[1] = rng.uniform(-1, 1, (5, 7))
    U[2] = rng.uniform(-3, 3, (5, 7))
    U[3] = 10.0
    dx, dy = 0.1, 0.025
    best = 0.0
    for j in range(5):
        for i in range(7):
            best = max(best, (abs(U[1, j, i]) / dx + abs(U[2, j, i]) / dy) / U[0, j, i])
    assert io.advective_dt(U, dx, dy, cfl=0.45) == (1 / best) * 0.45
    assert io.advective_dt(U, dx, dy, cfl=0.45) != io.advective_dt(U, dy, dx, cfl=0.45)


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
def test_stage_anisotropic_free_stream_and_conservation(advection):
    """Uniform flow stays uniform with dx != dy, and a non-uniform state's
    totals are conserved (periodic fluxes cancel) with dx != dy."""
    ny, nx = 12, 20
    gamma, eps = 1.4, 0.3
    U = state_from_primv(np.full((ny, nx), 0.8), np.full((ny, nx), 0.5),
                         np.full((ny, nx), -0.25), np.full((ny, nx), 3.0), gamma, eps)
    st = io.PeekStage(gamma, eps, advection)
    dx, dy, dt = 1.0 / nx, 0.5 / ny, 0.01
    p = io.pressure_from_state(U, gamma, eps)
    sy = st.assemble(U, p, dt, dx, dy)
    Un = st.finish(U, sy, p, dt, dx, dy)
    assert np.allclose(Un, U, rtol=1e-14, atol=0)
    rng = np.random.default_rng(2)
    U2 = state_from_primv(1 + 0.2 * rng.random((ny, nx)), rng.uniform(-1, 1, (ny, nx)),
                          rng.uniform(-1, 1, (ny, nx)), 2 + rng.random((ny, nx)), gamma, eps)
    p2 = io.pressure_from_state(U2, gamma, eps)
    sy2 = st.assemble(U2, p2, dt, dx, dy)
    Un2 = st.finish(U2, sy2, 1.01 * p2, dt, dx, dy)
    scale = np.abs(U2).sum(axis=(1, 2))
    assert np.all(np.abs(Un2.sum(axis=(1, 2)) - U2.sum(axis=(1, 2))) <= 1e-13 * scale)


def test_muscl_equals_first_order_on_piecewise_constant_data():
    """ToroTest2 data (constant states, two jumps): every minmod slope is
    zero (one of the two differences vanishes at each jump), so the MUSCL
    states are the cell values, the local step changes nothing, and the
    MUSCL fluxes equal the first-order ones bit for bit."""
    nx, ny = 8, 2
    u = np.where(np.arange(nx) < 4, -2.0, 2.0)[None, :].repeat(ny, 0)
    U = state_from_primv(np.ones((ny, nx)), u, np.zeros((ny, nx)), np.full((ny, nx), 0.4), 1.4, 1.0)
    f1 = io.rusanov_faces(U, 1.0)
    f2 = io.muscl_rusanov_faces(U, 1.0, 0.01, 1.0 / nx, 1.0 / ny)
    for a, b in zip(f1, f2):
        assert np.array_equal(a, b)


def _density_wave_error(n, advection):
    """L1 error of rho after advecting rho = 1 + 0.2 sin(2 pi x) with u = 1,
    p = 1 (exact pressure, no solve) to t = 0.25 on an n x 4 periodic grid."""
    gamma, eps = 1.4, 1.0
    x = (np.arange(n) + 0.5) / n
    rho0 = 1.0 + 0.2 * np.sin(2 * np.pi * x)[None, :].repeat(4, 0)
    U = state_from_primv(rho0, np.ones((4, n)), np.zeros((4, n)), np.ones((4, n)), gamma, eps)
    st = io.PeekStage(gamma, eps, advection)
    dx, dy = 1.0 / n, 0.25
    t, t_end = 0.0, 0.25
    while t < t_end - 1e-14:
        dt = min(0.4 * dx, t_end - t)
        p = io.pressure_from_state(U, gamma, eps)
        sy = st.assemble(U, p, dt, dx, dy)
        U = st.finish(U, sy, np.ones_like(p), dt, dx, dy)
        t += dt
    exact = 1.0 + 0.2 * np.sin(2 * np.pi * (x - t_end))
    return np.mean(np.abs(U[0, 0] - exact))


def test_muscl_is_second_order_on_a_smooth_wave():
    e = {a: [_density_wave_error(n, a) for n in (32, 64, 128)] for a in ("rusanov", "muscl_rusanov")}
    r1 = e["rusanov"][1] / e["rusanov"][2]
    r2 = e["muscl_rusanov"][1] / e["muscl_rusanov"][2]
    assert 1.6 < r1 < 2.4, e          # first order
    assert r2 > 3.0, e                 # second order (minmod clips the extrema)
    assert e["muscl_rusanov"][2] < 0.2 * e["rusanov"][2], e
