"""GPU parity of the IMEX PEEK-SP111 stage kernels (bx_imex_*, through the
C ABI) against the oracle's numpy restatement (oracle/imex_oracle.py).

Bar: bit-exact -- every output of the kernels is evaluated in the oracle's
operation order with no contraction."""
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
