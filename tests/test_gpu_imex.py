"""GPU parity of the IMEX PEEK stage kernels and of the IMEX Runge-Kutta
step built on them (bx_imex_*, through the C ABI) against the oracle's numpy
restatement (oracle/imex_oracle.py).

Bar: bit-exact -- every output of the kernels is evaluated in the oracle's
operation order with no contraction."""
import ctypes

import numpy as np
import pytest

import imex_oracle as io
from bamrex_amd import imex
from bamrex_amd.imex import DeviceArray, PeekStage, dev

from conftest import announce

pytestmark = pytest.mark.gpu


def random_state(ny, nx, gamma, eps, seed):
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.3 * rng.random((ny, nx))
    u = rng.uniform(-1, 1, (ny, nx))
    v = rng.uniform(-1, 1, (ny, nx))
    p = 2.0 + rng.random((ny, nx))
    U = np.empty((4, ny, nx))
    U[0] = rho
    U[1] = rho * u
    U[2] = rho * v
    U[3] = p / (gamma - 1) + 0.5 * eps * rho * (u * u + v * v)
    return U


def check_stage(U, p, p_new, dt, dx, dy, gamma, eps, tag, advection="rusanov"):
    ny, nx = p.shape
    announce("peek stage %s %s %dx%d" % (tag, advection, nx, ny))
    st = PeekStage(nx, ny, dx, dy, gamma, eps, advection)
    ost = io.PeekStage(gamma, eps, advection)
    dU, dp, dpn, dout = dev(U), dev(p), dev(p_new), DeviceArray(U.shape)
    try:
        sy = st.assemble(dU, dp, dt)
        ref = ost.assemble(U, p, dt, dx, dy)
        for key in ("Uex", "enth", "rhs", "a", "bX", "bY"):
            got = sy[key].download()
            assert np.array_equal(got, ref[key]), (tag, key, np.argwhere(got != ref[key])[:3])
        assert sy["beta"] == ref["beta"] and sy["alpha"] == ref["alpha"]
        st.finish(dU, dpn, dt, dout)
        exp = ost.finish(U, ref, p_new, dt, dx, dy)
        got = dout.download()
        assert np.array_equal(got, exp), (tag, "U_new", np.argwhere(got != exp)[:3])
        dpp = DeviceArray(p.shape)
        st.pressure(dU, dpp)
        assert np.array_equal(dpp.download(), io.pressure_from_state(U, gamma, eps))
        assert st.advective_dt(dU, 0.45) == io.advective_dt(U, dx, dy, cfl=0.45)
        dpp.free()
    finally:
        st.free()
        for d in (dU, dp, dpn, dout):
            d.free()


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
@pytest.mark.parametrize("shape", [(20, 48), (1, 1), (5, 1), (1, 7), (2, 3), (3, 65), (130, 256)])
def test_peek_stage_random_states(shape, advection):
    """Random states, anisotropic cells, eps != 1, sizes across the
    64 x 4 thread-block edges and degenerate one-cell directions (where the
    periodic neighbour is the cell itself)."""
    ny, nx = shape
    gamma, eps = 1.4, 0.3
    U = random_state(ny, nx, gamma, eps, seed=ny * 1000 + nx)
    rng = np.random.default_rng(7)
    p = 2.0 + rng.random((ny, nx))
    p_new = p * (1 + 0.01 * rng.standard_normal((ny, nx)))
    check_stage(U, p, p_new, 0.013, 1.0 / nx, 0.7 / ny, gamma, eps, "random", advection)


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
def test_peek_stage_gresho_steps_and_conservation(advection):
    """Five steps of the Gresho vortex at 256^2 (eps = 1) driven through the
    GPU stage and the oracle side by side with the same pressure: identical
    dt and state after every step, and the GPU totals conserved."""
    n = 256
    gamma, eps = 1.4, 1.0
    U, dx = imex.gresho_state(n)
    st = PeekStage(n, n, dx, gamma=gamma, eps=eps, advection=advection)
    ost = io.PeekStage(gamma, eps, advection)
    dU, dUn, dp = dev(U), DeviceArray(U.shape), DeviceArray((n, n))
    st.pressure(dU, dp)
    p = io.pressure_from_state(U, gamma, eps)
    assert np.array_equal(dp.download(), p)
    try:
        for k in range(5):
            announce("peek gresho step %d" % k)
            dt = st.advective_dt(dU)
            assert dt == io.advective_dt(U, dx)
            st.assemble(dU, dp, dt)
            ref = ost.assemble(U, p, dt, dx)
            # a stand-in for the solved pressure: deterministic, not the input
            p_new = io.pressure_from_state(ref["Uex"], gamma, eps)
            dpn = dev(p_new)
            st.finish(dU, dpn, dt, dUn)
            U_next = ost.finish(U, ref, p_new, dt, dx)
            got = dUn.download()
            assert np.array_equal(got, U_next), (k, np.argwhere(got != U_next)[:3])
            scale = np.abs(U).sum(axis=(1, 2))
            assert np.all(np.abs(got.sum(axis=(1, 2)) - U.sum(axis=(1, 2))) <= 1e-12 * scale)
            U, p = U_next, p_new
            dU, dUn = dUn, dU
            dp.copy_from(dpn)
            dpn.free()
    finally:
        st.free()
        for d in (dU, dUn, dp):
            d.free()


def test_workload_full_size_step_matches_oracle():
    """The benchmark workload's first system at its full size (Gresho 1024^2):
    the GPU system equals the oracle's, and the finish with the oracle's own
    stand-in pressure gives the oracle's state."""
    announce("peek workload 1024")
    n = 1024
    wl = imex.GreshoWorkload(n)
    U0, dx = imex.gresho_state(n)
    try:
        sy = wl.next_system()
        ost = io.PeekStage(1.4, 1.0)
        p0 = io.pressure_from_state(U0, 1.4, 1.0)
        dt = io.advective_dt(U0, dx)
        assert sy["dt"] == dt
        ref = ost.assemble(U0, p0, dt, dx)
        for key in ("rhs", "bX", "bY", "enth"):
            assert np.array_equal(sy[key].download(), ref[key]), key
        p_new = 1.0001 * p0
        wl.advance(sy, dev(p_new))
        assert np.array_equal(wl.U.download(), ost.finish(U0, ref, p_new, dt, dx))
    finally:
        wl.free()


def test_peek_assemble_rejects_aliased_outputs():
    """An output that aliases an input is refused before any launch."""
    from bamrex_amd import _lib
    n = 8
    st = PeekStage(n, n, 1.0 / n)
    dU, dp = DeviceArray((4, n, n)), DeviceArray((n, n))
    try:
        rc = _lib.lib().bx_imex_peek_assemble(dU.ptr, dp.ptr, n, n, st.dx, st.dy, 0.01, 1.4, 1.0, 0,
                                              st.Uex.ptr, st.enth.ptr, dp.ptr, st.a.ptr, st.bX.ptr,
                                              st.bY.ptr, None)
        assert rc != 0 and b"alias" in _lib.lib().bx_helm_last_error()
    finally:
        st.free()
        dU.free()
        dp.free()


def _rk_parity(U, p, dt, dx, dy, gamma, eps, tableau, advection, tag, ke_method="ex",
               enthalpy_method="reuse_pressure"):
    """One IMEX RK step on the GPU (ImexRK, through the C ABI) and in the
    oracle (advance_imex_rk), with the stand-in solve p_new = pressure of
    the stage's explicit state on both sides (bx_imex_pressure is
    bit-exact with the oracle's pressure): every stage's system and the
    step's state and pressure must be identical."""
    ny, nx = p.shape
    announce("imex rk %s %s %s %s %dx%d" % (tableau, advection, ke_method, tag, nx, ny))
    rec = []
    U_ref, p_ref = io.advance_imex_rk(U, p, dt, dx, dy, tableau,
                                      lambda sy, p0: io.pressure_from_state(sy["Uex"], gamma, eps),
                                      gamma, eps, advection, record=rec, ke_method=ke_method,
                                      enthalpy_method=enthalpy_method)
    rk = imex.ImexRK(nx, ny, dx, dy, gamma, eps, advection, tableau, ke_method,
                     enthalpy_method=enthalpy_method)
    dU, dp, dout, dpn = dev(U), dev(p), DeviceArray(U.shape), DeviceArray(p.shape)
    try:
        seen = []

        def solve(sy):
            r = rec[sy["stage"]]
            assert sy["adt"] == r["adt"] and sy["beta"] == r["beta"]
            assert sy["gamma_c"] == r["gamma_c"]
            keys = ("Uex", "enth", "rhs", "bX", "bY") + (("cX", "cY") if ke_method == "split" else ())
            for key in keys:
                got = sy[key].download()
                assert np.array_equal(got, r[key]), (tableau, sy["stage"], key,
                                                      np.argwhere(got != r[key])[:3])
            assert np.array_equal(sy["p_prev"].download(), r["p_in"]), (tableau, sy["stage"], "p_in")
            rk.stage.pressure(sy["Uex"], dpn)
            seen.append(sy["stage"])
            return dpn

        rk.advance(dU, dp, dt, dout, solve)
        assert seen == list(range(rk.tab.n_imp))
        got = dout.download()
        assert np.array_equal(got, U_ref), (tableau, "U_new", np.argwhere(got != U_ref)[:3])
        assert np.array_equal(dp.download(), p_ref)
    finally:
        rk.free()
        for d in (dU, dp, dout, dpn):
            d.free()


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
@pytest.mark.parametrize("tableau", ["SP111", "SSP222", "SASSP322", "SASSP332", "SSP433"])
def test_imex_rk_step_matches_oracle(tableau, advection):
    """Every IMEX Butcher tableau of the reference, both advection fluxes,
    random states on anisotropic cells, eps != 1, across the 64 x 4
    thread-block edges and with a one-cell direction."""
    gamma, eps = 1.4, 0.7
    for ny, nx in ((20, 48), (3, 65), (1, 9)):
        U = random_state(ny, nx, gamma, eps, seed=7 * nx + ny)
        p = io.pressure_from_state(U, gamma, eps) * 1.01
        _rk_parity(U, p, 0.009, 1.0 / nx, 0.6 / ny, gamma, eps, tableau, advection,
                   "random")


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
@pytest.mark.parametrize("tableau", ["SP111", "SSP222", "SASSP322", "SSP433"])
def test_imex_rk_split_ke_step_matches_oracle(tableau, advection):
    """PESK (ke_method split): the split kinetic energy, the velocity
    coefficients of the operator and the SK pressure of later stages,
    bit-exact stage by stage."""
    gamma, eps = 1.4, 0.7
    for ny, nx in ((20, 48), (1, 9)):
        U = random_state(ny, nx, gamma, eps, seed=5 * nx + ny)
        p = io.pressure_from_state(U, gamma, eps) * 0.99
        _rk_parity(U, p, 0.009, 1.0 / nx, 0.6 / ny, gamma, eps, tableau, advection, "random",
                   ke_method="split")


def test_imex_rk_pesk_ssp222_gresho_with_multigrid_solves():
    """PESK-MUSCL-SSP222 steps of the Gresho vortex at 128^2, every stage's
    non-symmetric operator (with the velocity term) solved by GPU
    multigrid to 1e-10: the oracle, given the same solved pressures,
    reaches the identical state; totals are conserved."""
    n = 128
    announce("imex rk pesk ssp222 gresho 128")
    U0, dx = imex.gresho_state(n)
    rk = imex.ImexRK(n, n, dx, advection="muscl_rusanov", tableau="SSP222", ke_method="split")
    solver = imex.HelmholtzSolver(n, n, dx)
    dU, dUn, dp = dev(U0), DeviceArray(U0.shape), DeviceArray((n, n))
    rk.stage.pressure(dU, dp)
    U, p = U0, dp.download()
    try:
        for k in range(2):
            dt = rk.stage.advective_dt(dU)
            solved = []

            def solve(sy):
                solver.set_coeffs(sy["alpha"], sy["beta"], sy["a"], sy["bX"], sy["bY"],
                                  gamma=sy["gamma_c"], cX=sy["cX"], cY=sy["cY"])
                x = imex.dev(sy["p_prev"].download())
                st = solver.solve(imex.dev(sy["rhs"].download()), x, "mg", rtol=1e-10, maxiter=100)
                assert st["converged"] and st["iters"] < 100
                solved.append(x.download())
                return x

            rk.advance(dU, dp, dt, dUn, solve)
            it = iter(solved)
            U_ref, p_ref = io.advance_imex_rk(U, p, dt, dx, dx, "SSP222", lambda sy, p0: next(it), 1.4,
                                              1.0, "muscl_rusanov", ke_method="split")
            got = dUn.download()
            assert np.array_equal(got, U_ref), (k, np.argwhere(got != U_ref)[:3])
            assert np.array_equal(dp.download(), p_ref)
            scale = np.abs(U).sum(axis=(1, 2))
            assert np.all(np.abs(got.sum(axis=(1, 2)) - U.sum(axis=(1, 2))) <= 1e-12 * scale)
            U, p = U_ref, p_ref
            dU, dUn = dUn, dU
    finally:
        solver.close()
        rk.free()
        for d in (dU, dUn, dp):
            d.free()


def test_imex_rk_ssp222_gresho_with_multigrid_solves():
    """PEEK-MUSCL-SSP222 steps of the Gresho vortex at 256^2 with the
    pressure of every stage solved on the GPU (plain multigrid, 1e-10): the
    oracle, given the same solved pressures, reaches the identical state;
    the GPU totals are conserved."""
    n = 256
    announce("imex rk ssp222 gresho 256")
    wl = imex.GreshoWorkload(n, advection="muscl_rusanov", tableau="SSP222")
    solver = imex.HelmholtzSolver(n, n, wl.dx)
    ost_U = wl.U.download()
    ost_p = wl.p.download()
    try:
        for k in range(3):
            solved = []
            U0 = wl.U.download()
            dt = wl.step_dt()
            assert dt == io.advective_dt(U0, wl.dx)
            while True:
                sy = wl.next_system(dt)
                solver.set_coeffs(sy["alpha"], sy["beta"], sy["a"], sy["bX"], sy["bY"])
                x = imex.dev(sy["p_prev"].download())
                st = solver.solve(imex.dev(sy["rhs"].download()), x, "mg", rtol=1e-10, maxiter=100)
                assert st["iters"] < 100
                solved.append(x.download())
                done = wl.advance(sy, x)
                x.free()
                if done:
                    break
            assert len(solved) == 2
            it = iter(solved)
            ost_U, ost_p = io.advance_imex_rk(ost_U, ost_p, dt, wl.dx, wl.dx, "SSP222",
                                              lambda sy, p0: next(it), 1.4, 1.0, "muscl_rusanov")
            got = wl.U.download()
            assert np.array_equal(got, ost_U), (k, np.argwhere(got != ost_U)[:3])
            assert np.array_equal(wl.p.download(), ost_p)
            scale = np.abs(U0).sum(axis=(1, 2))
            assert np.all(np.abs(got.sum(axis=(1, 2)) - U0.sum(axis=(1, 2))) <= 1e-12 * scale)
    finally:
        solver.close()
        wl.free()


def test_imex_rk_rejects_bad_arguments():
    """Stage counts beyond the maximum, aliased outputs and unknown
    tableaux are refused before any launch."""
    from bamrex_amd import _lib
    n = 8
    L = _lib.lib()
    dU, a1, a2 = DeviceArray((4, n, n)), DeviceArray((4, n, n)), DeviceArray((4, n, n))
    try:
        ptrs = (ctypes.c_void_p * 5)(*([a1.ptr.value] * 5))
        w = (ctypes.c_double * 5)(*([0.1] * 5))
        rc = L.bx_imex_rk_combine(dU.ptr, ptrs, ptrs, 5, w, w, n, n, 0.1, 0.1, a2.ptr, dU.ptr, None)
        assert rc != 0 and b"bad argument" in L.bx_helm_last_error()
        one = (ctypes.c_void_p * 4)(a1.ptr.value, None, None, None)
        rc = L.bx_imex_rk_combine(dU.ptr, one, one, 1, w, w, n, n, 0.1, 0.1, a1.ptr, a2.ptr, None)
        assert rc != 0    # a flux array is also an output
        rc = L.bx_imex_rk_update(dU.ptr, one, one, 0, w, n, n, 0.1, 0.1, 0.01, a2.ptr, None)
        assert rc != 0    # no stage
        rc = L.bx_imex_rk_pressure(dU.ptr, a1.ptr, n, n, 0.1, 0.1, 0.01, 1.4, 1.0, 0, 0, dU.ptr, None)
        assert rc != 0    # p aliases the state
        rc = L.bx_imex_rk_pressure(dU.ptr, a1.ptr, n, n, 0.1, 0.1, 0.01, 1.4, 1.0, 0, 7, a2.ptr, None)
        assert rc != 0    # unknown ke_method
        sc = [DeviceArray((n, n)) for _ in range(6)]
        try:
            rc = L.bx_imex_stage_assemble(dU.ptr, dU.ptr, a1.ptr, n, n, 0.1, 0.1, 0.01, 1.4, 1.0, 0, 0,
                                          1, 0, a2.ptr, *[d.ptr for d in sc[:5]], None, None, None)
            assert rc != 0    # split without velocity outputs
            rc = L.bx_imex_stage_assemble(dU.ptr, dU.ptr, a1.ptr, n, n, 0.1, 0.1, 0.01, 1.4, 1.0, 0, 0,
                                          0, 5, a2.ptr, *[d.ptr for d in sc[:5]], None, None, None)
            assert rc != 0    # unknown enthalpy_method
        finally:
            for d in sc:
                d.free()
    finally:
        for d in (dU, a1, a2):
            d.free()
    with pytest.raises(ValueError):
        imex.ImexRK(n, n, 0.1, tableau="RK4")
    with pytest.raises(ValueError):
        imex.ImexRK(n, n, 0.1, ke_method="conservative")
    # stages out of order are refused on the host, before any launch
    rk = imex.ImexRK(n, n, 0.1, tableau="SSP222")
    try:
        with pytest.raises(RuntimeError):
            rk.system()
        with pytest.raises(RuntimeError):
            rk.complete(rk.stage.rhs, rk.stage.Uex)
    finally:
        rk.free()


def _compressive_state(ny, nx, gamma, eps, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    X, Y = np.meshgrid(x, y)
    rho = 1.0 + 0.1 * rng.random((ny, nx))
    u = -0.8 * np.sin(2 * np.pi * X) + 0.05 * rng.standard_normal((ny, nx))
    v = -0.5 * np.sin(2 * np.pi * Y) + 0.05 * rng.standard_normal((ny, nx))
    p = 1.0 + 0.1 * rng.random((ny, nx))
    U = np.empty((4, ny, nx))
    U[0], U[1], U[2] = rho, rho * u, rho * v
    U[3] = p / (gamma - 1) + 0.5 * eps * rho * (u * u + v * v)
    return U


@pytest.mark.parametrize("shape", [(20, 48), (1, 9), (67, 130)])
def test_shock_indicator_matches_oracle(shape):
    """compute_shock_indicator on the GPU: the indicator field and its
    maximum bit-exact against the oracle (dx != dy, eps != 1)."""
    ny, nx = shape
    announce("shock indicator %dx%d" % (nx, ny))
    gamma, eps, k1 = 1.4, 0.7, 0.05
    U = _compressive_state(ny, nx, gamma, eps, nx + ny)
    dx, dy = 1.0 / nx, 0.8 / ny
    ref, ref_max = io.shock_indicator(U, dx, dy, gamma, eps, k1)
    st = imex.ImexRKStab(nx, ny, dx, dy, gamma, eps, k1=k1)
    dU = dev(U)
    try:
        m = st.indicator(dU)
        assert np.array_equal(st.ind.download(), ref)
        assert m == ref_max
    finally:
        st.free()
        dU.free()


@pytest.mark.parametrize("ke_method", ["ex", "split"])
@pytest.mark.parametrize("tableau", ["SSP222", "SSP433", "SP111"])
def test_stabilised_step_matches_oracle(tableau, ke_method):
    """advance_imex_rk_stab on the GPU (ImexRKStab) against the oracle with
    the stand-in solve on both sides: a compressive state (blended, for the
    second-order tableaux) and a uniform flow (not blended)."""
    gamma, eps, k1 = 1.4, 1.0, 0.05
    n = 32
    dx = 1.0 / n
    uni = np.empty((4, n, n))
    uni[0], uni[1], uni[2], uni[3] = 1.2, 0.36, -0.24, 2.0 / 0.4 + 0.5 * 1.2 * (0.09 + 0.04)
    for tag, U in (("compressive", _compressive_state(n, n, gamma, eps, 11)), ("uniform", uni)):
        announce("stabilised step %s %s %s" % (tableau, ke_method, tag))
        p = io.pressure_from_state(U, gamma, eps)
        rec = {}
        U_ref, p_ref = io.advance_imex_rk_stab(
            U, p, 0.004, dx, dx, tableau,
            lambda sy, p0: io.pressure_from_state(sy["Uex"], gamma, eps), gamma, eps,
            "muscl_rusanov", ke_method=ke_method, k1=k1, record=rec)
        st = imex.ImexRKStab(n, n, dx, dx, gamma, eps, "muscl_rusanov", tableau, ke_method, k1)
        dU, dp, dout, dpn = dev(U), dev(p), DeviceArray(U.shape), DeviceArray((n, n))
        try:
            def stand_in(sy):
                st.hi.stage.pressure(sy["Uex"], dpn)
                return dpn
            st.advance(dU, dp, 0.004, dout, stand_in)
            blended = "U_lo" in rec
            assert st.blended == blended and (blended or tag == "uniform" or tableau == "SP111")
            if tableau != "SP111":
                assert st.ind_max == rec["ind_max"]
            got = dout.download()
            assert np.array_equal(got, U_ref), (tag, np.argwhere(got != U_ref)[:3])
            assert np.array_equal(dp.download(), p_ref)
        finally:
            st.free()
            for d in (dU, dp, dout, dpn):
                d.free()


@pytest.mark.parametrize("advection", ["rusanov", "muscl_rusanov"])
@pytest.mark.parametrize("tableau", ["SP111", "SSP222", "SSP433"])
def test_imex_rk_fi_step_matches_oracle(tableau, advection):
    """FI (not linearised, two Picard iterations per stage): every system
    handed to the solve -- the stage's first and each Picard update -- and
    the step's state and pressure bit-exact against the oracle, with the
    stand-in solve p = 1.01 x the pressure of the stage's explicit state."""
    gamma, eps = 1.4, 0.7
    for ny, nx in ((20, 48), (1, 9)):
        announce("imex rk fi %s %s %dx%d" % (tableau, advection, nx, ny))
        U = random_state(ny, nx, gamma, eps, seed=3 * nx + ny)
        p = io.pressure_from_state(U, gamma, eps) * 1.02
        dx, dy, dt = 1.0 / nx, 0.6 / ny, 0.009
        calls = []

        def oracle_solve(sy, p0):
            calls.append({k: sy[k].copy() for k in ("rhs", "bX", "bY", "enth")})
            calls[-1]["p0"] = p0.copy()
            return 1.01 * io.pressure_from_state(sy["Uex"], gamma, eps)

        U_ref, p_ref = io.advance_imex_rk(U, p, dt, dx, dy, tableau, oracle_solve, gamma, eps,
                                          advection, ke_method="fi", picard_max_iter=2)
        rk = imex.ImexRK(nx, ny, dx, dy, gamma, eps, advection, tableau, "fi", picard_max_iter=2)
        dU, dp, dout, dpn = dev(U), dev(p), DeviceArray(U.shape), DeviceArray(p.shape)
        k = [0]
        try:
            def solve(sy):
                ref = calls[k[0]]
                for key in ("rhs", "bX", "bY", "enth"):
                    got = sy[key].download()
                    assert np.array_equal(got, ref[key]), (tableau, k[0], key)
                assert np.array_equal(sy["p_prev"].download(), ref["p0"]), (tableau, k[0], "p0")
                k[0] += 1
                rk.stage.pressure(sy["Uex"], dpn)
                dpn.upload(1.01 * dpn.download())
                return dpn

            rk.advance(dU, dp, dt, dout, solve)
            assert k[0] == len(calls) == 2 * rk.tab.n_imp
            got = dout.download()
            assert np.array_equal(got, U_ref), (tableau, np.argwhere(got != U_ref)[:3])
            assert np.array_equal(dp.download(), p_ref)
        finally:
            rk.free()
            for d in (dU, dp, dout, dpn):
                d.free()


@pytest.mark.parametrize("enthalpy_method", ["conservative", "conservative_ex"])
@pytest.mark.parametrize("ke_method", ["ex", "split", "conservative_kinetic"])
def test_imex_rk_enthalpy_methods_match_oracle(ke_method, enthalpy_method):
    """The conservative enthalpy methods (EEEK / EESK inputs) and the
    conservative_kinetic energy, SSP222 and SASSP322 with MUSCL advection:
    every stage's system, initial pressure and the step bit-exact."""
    gamma, eps = 1.4, 0.7
    for tableau in ("SSP222", "SASSP322"):
        U = random_state(20, 48, gamma, eps, seed=77)
        p = io.pressure_from_state(U, gamma, eps) * 1.01
        _rk_parity(U, p, 0.009, 1.0 / 48, 0.6 / 20, gamma, eps, tableau, "muscl_rusanov",
                   "%s/%s" % (ke_method, enthalpy_method), ke_method=ke_method,
                   enthalpy_method=enthalpy_method)


def test_time_step_estimates_match_oracle():
    """The acoustic and advective dt on the GPU equal the oracle's, also
    with a NaN speed (negative pressure) and a negative one (negative
    density) that the reduction passes over."""
    gamma, eps = 1.4, 0.6
    for ny, nx in ((37, 70), (1, 5)):
        announce("imex dt estimates %dx%d" % (nx, ny))
        U = random_state(ny, nx, gamma, eps, seed=nx)
        st = PeekStage(nx, ny, 1.0 / nx, 0.7 / ny, gamma, eps)
        try:
            for case in ("plain", "bad"):
                V = U.copy()
                if case == "bad":
                    V[3, 0, 2] = 0.0
                    V[0, ny - 1, 3] = -V[0, ny - 1, 3]
                dV = dev(V)
                try:
                    assert st.acoustic_dt(dV, 0.45) == io.acoustic_dt(V, st.dx, st.dy, gamma, eps, 0.45)
                    assert st.advective_dt(dV, 0.45) == io.advective_dt(V, st.dx, st.dy, 0.45)
                finally:
                    dV.free()
        finally:
            st.free()


def test_imex_rk_fi_residual_check_ends_picard_loops():
    """imex.compute_residual: a residual below res_tol before a solve ends
    the stage's Picard loop (here after the first solve of each stage, and
    before any solve in the second stage of a second run); the GPU step
    follows the oracle's bit for bit."""
    gamma, eps = 1.4, 0.7
    ny, nx = 20, 48
    U = random_state(ny, nx, gamma, eps, seed=91)
    p = io.pressure_from_state(U, gamma, eps)
    dx, dy, dt = 1.0 / nx, 0.6 / ny, 0.009
    for pattern in ((1.0, 0.0, 1.0, 0.0), (1.0, 0.0, 0.0)):
        announce("imex rk fi residual pattern %s" % (pattern,))
        seq = iter(pattern)
        U_ref, p_ref = io.advance_imex_rk(
            U, p, dt, dx, dy, "SSP222", lambda sy, p0: 1.01 * io.pressure_from_state(sy["Uex"], gamma, eps),
            gamma, eps, "muscl_rusanov", ke_method="fi", picard_max_iter=3,
            residual=lambda sy, p0: next(seq))
        rk = imex.ImexRK(nx, ny, dx, dy, gamma, eps, "muscl_rusanov", "SSP222", "fi", picard_max_iter=3)
        dU, dp, dout, dpn = dev(U), dev(p), DeviceArray(U.shape), DeviceArray(p.shape)
        seq2 = iter(pattern)
        try:
            def solve(sy):
                rk.stage.pressure(sy["Uex"], dpn)
                dpn.upload(1.01 * dpn.download())
                return dpn
            rk.advance(dU, dp, dt, dout, solve, residual=lambda sy: next(seq2))
            assert np.array_equal(dout.download(), U_ref)
            assert np.array_equal(dp.download(), p_ref)
        finally:
            rk.free()
            for d in (dU, dp, dout, dpn):
                d.free()


def test_workload_fi_step_with_multigrid_solves():
    """The benchmark's path for FI-MUSCL-SSP222 (GreshoWorkload stepping
    through the Picard iterations, 128^2, plain multigrid to 1e-10): the
    oracle, given the same four solved pressures, reaches the identical
    state and pressure."""
    n = 128
    announce("workload fi ssp222 128")
    wl = imex.GreshoWorkload(n, advection="muscl_rusanov", tableau="SSP222", ke_method="fi")
    solver = imex.HelmholtzSolver(n, n, wl.dx)
    U0, p0 = wl.U.download(), wl.p.download()
    solved = []
    try:
        dt = wl.step_dt()
        while True:
            sy = wl.next_system(dt)
            solver.set_coeffs(sy["alpha"], sy["beta"], sy["a"], sy["bX"], sy["bY"])
            x = imex.dev(sy["p_prev"].download())
            st = solver.solve(imex.dev(sy["rhs"].download()), x, "mg", rtol=1e-10, maxiter=100)
            assert st["converged"]
            solved.append(x.download())
            done = wl.advance(sy, x)
            x.free()
            if done:
                break
        assert len(solved) == 4
        it = iter(solved)
        U_ref, p_ref = io.advance_imex_rk(U0, p0, dt, wl.dx, wl.dx, "SSP222", lambda sy, p: next(it),
                                          1.4, 1.0, "muscl_rusanov", ke_method="fi", picard_max_iter=2)
        assert np.array_equal(wl.U.download(), U_ref)
        assert np.array_equal(wl.p.download(), p_ref)
        assert wl.time == dt
    finally:
        solver.close()
        wl.free()
