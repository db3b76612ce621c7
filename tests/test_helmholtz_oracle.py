"""CPU tests that pin the pressure-Helmholtz and network oracles.

The oracle (oracle/oracle_helm.c, oracle/cnn_oracle.py) is the checker of the
HIP solver and network; these tests check the oracle itself against closed
forms, the reference's own test formula and finite differences.
"""
import ctypes
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import cnn_oracle as co  # noqa: E402
import helm_oracle as ho  # noqa: E402


def rand_coeffs(rng, nx, ny):
    a = 0.5 + rng.random((ny, nx))
    hc = 1.0 + rng.random((ny, nx))
    # face values: arithmetic average of the two neighbours, periodic
    bX = np.empty((ny, nx + 1))
    bX[:, :nx] = 0.5 * (np.roll(hc, 1, axis=1) + hc)
    bX[:, nx] = bX[:, 0]
    bY = np.empty((ny + 1, nx))
    bY[:ny] = 0.5 * (np.roll(hc, 1, axis=0) + hc)
    bY[ny] = bY[0]
    return a, bX, bY


def test_apply_fourier_eigenvalue():
    """Constant coefficients: a Fourier mode is an eigenvector with the
    eigenvalue alpha a + beta b (4/dx^2 sin^2(pi kx/nx) + 4/dy^2 sin^2(pi ky/ny))."""
    nx, ny, dx, dy = 32, 16, 1 / 32, 1 / 16
    alpha, beta, a0, b0 = 1.3, 0.7e-3, 2.5, 3.0
    a = np.full((ny, nx), a0)
    bX = np.full((ny, nx + 1), b0)
    bY = np.full((ny + 1, nx), b0)
    j, i = np.mgrid[0:ny, 0:nx]
    for kx, ky in [(0, 0), (1, 0), (3, 2), (16, 8), (5, 7)]:
        x = np.cos(2 * math.pi * (kx * i / nx + ky * j / ny))
        y = ho.apply(x, a, bX, bY, alpha, beta, dx, dy)
        lam = alpha * a0 + beta * b0 * (4 / dx ** 2 * math.sin(math.pi * kx / nx) ** 2
                                        + 4 / dy ** 2 * math.sin(math.pi * ky / ny) ** 2)
        assert np.max(np.abs(y - lam * x)) <= 1e-12 * lam


def test_apply_matches_reference_test_residual():
    """Tests/MLABecLapGrad01.cpp:176-213 writes the operator independently
    (a alpha x - b sum ddx(...) + 0.5 c sum idx gamma (...)); restated here on
    the test's coefficient fields at interior cells of an 8x8 unit domain."""
    n = 8
    dx = 1.0 / n
    c = (np.arange(n) + 0.5) * dx
    x, y = np.meshgrid(c, c)
    A_, B_, C_ = 1.0, 2.0, -4.0
    alpha = np.sin(2 * np.pi * x) ** 2 * np.cos(4 * np.pi * y) ** 2
    gam0 = np.cos(np.pi * x) ** 4 * np.cos(2 * np.pi * y) ** 2
    gam1 = np.cos(3 * np.pi * x) ** 2 * np.sin(2 * np.pi * y) ** 4
    beta_cc = np.sin(2 * np.pi * x) + 2 + 0.5 * np.cos(4 * np.pi * y)
    bX = np.empty((n, n + 1))
    bX[:, :n] = 0.5 * (np.roll(beta_cc, 1, axis=1) + beta_cc)
    bX[:, n] = bX[:, 0]
    bY = np.empty((n + 1, n))
    bY[:n] = 0.5 * (np.roll(beta_cc, 1, axis=0) + beta_cc)
    bY[n] = bY[0]
    rng = np.random.default_rng(5)
    s = rng.normal(size=(n, n))
    out = ho.apply(s, alpha, bX, bY, A_, B_, dx, dx, gamma=C_, cX=gam0, cY=gam1)
    ddx = 1 / (dx * dx)
    idx = 1 / dx
    for j in range(1, n - 1):
        for i in range(1, n - 1):
            aterm = A_ * alpha[j, i] * s[j, i]
            bterm = -B_ * (ddx * (bX[j, i + 1] * (s[j, i + 1] - s[j, i]) - bX[j, i] * (s[j, i] - s[j, i - 1]))
                           + ddx * (bY[j + 1, i] * (s[j + 1, i] - s[j, i]) - bY[j, i] * (s[j, i] - s[j - 1, i])))
            cterm = 0.5 * C_ * (idx * gam0[j, i] * (s[j, i + 1] - s[j, i - 1])
                                + idx * gam1[j, i] * (s[j + 1, i] - s[j - 1, i]))
            ref = aterm + bterm + cterm
            assert abs(out[j, i] - ref) <= 1e-12 * (abs(ref) + 1.0)


def test_apply_symmetric_positive():
    rng = np.random.default_rng(1)
    nx, ny = 24, 20
    a, bX, bY = rand_coeffs(rng, nx, ny)
    u, v = rng.normal(size=(2, ny, nx))
    Au = ho.apply(u, a, bX, bY, 1.0, 1e-3, 1 / nx, 1 / ny)
    Av = ho.apply(v, a, bX, bY, 1.0, 1e-3, 1 / nx, 1 / ny)
    assert abs(np.sum(v * Au) - np.sum(u * Av)) <= 1e-12 * np.sum(np.abs(v * Au))
    assert np.sum(u * Au) > 0


def test_gsrb_zeroes_residual_on_updated_colour():
    """After a colour sweep the residual at the cells of that colour is
    zero up to rounding (the cell equation is solved exactly)."""
    rng = np.random.default_rng(2)
    nx, ny = 16, 12
    a, bX, bY = rand_coeffs(rng, nx, ny)
    rhs = rng.normal(size=(ny, nx))
    phi = rng.normal(size=(ny, nx))
    j, i = np.mgrid[0:ny, 0:nx]
    for rb in (0, 1):
        phi = ho.gsrb(phi, rhs, a, bX, bY, 1.0, 2e-3, 1 / nx, 1 / ny, rb)
        r = ho.apply(phi, a, bX, bY, 1.0, 2e-3, 1 / nx, 1 / ny, rhs=rhs)
        colour = (i + j + rb) % 2 == 0
        assert np.max(np.abs(r[colour])) <= 1e-12 * np.max(np.abs(rhs))
        assert np.max(np.abs(r[~colour])) > 1e-6


def test_restrict_and_face_average():
    rng = np.random.default_rng(3)
    f = rng.normal(size=(8, 12))
    c = ho.restrict(f)
    assert c.shape == (4, 6)
    assert abs(4 * c.sum() - f.sum()) <= 1e-12 * np.abs(f).sum()
    assert np.allclose(ho.restrict(np.full((8, 8), 3.25)), 3.25, rtol=0, atol=0)
    fX = np.full((8, 9), 1.5)
    fY = np.full((9, 8), 2.5)
    cX, cY = ho.avg_faces(fX, fY)
    assert cX.shape == (4, 5) and cY.shape == (5, 4)
    assert np.all(cX == 1.5) and np.all(cY == 2.5)


def test_mg_solve_matches_direct_solve():
    scipy_sparse = pytest.importorskip("scipy.sparse")
    from scipy.sparse.linalg import spsolve
    rng = np.random.default_rng(4)
    nx = ny = 32
    a, bX, bY = rand_coeffs(rng, nx, ny)
    alpha, beta, dx = 1.0, 5e-3, 1 / 32
    b = rng.normal(size=(ny, nx))
    x, its, hist = ho.mg_solve(b, a, bX, bY, alpha, beta, dx, dx, rtol=1e-10, maxiter=50)
    assert its > 0
    # assemble A column by column
    n = nx * ny
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        cols.append(ho.apply(e.reshape(ny, nx), a, bX, bY, alpha, beta, dx, dx).ravel())
    A = scipy_sparse.csr_matrix(np.array(cols).T)
    xd = spsolve(A, b.ravel()).reshape(ny, nx)
    assert np.max(np.abs(x - xd)) <= 1e-8 * np.max(np.abs(xd))
    # geometric convergence of the V-cycle
    rates = hist[1:its + 1] / hist[:its]
    assert np.all(rates < 0.3)


def test_mg_solve_at_extent_halving_to_odd_matches_direct_solve():
    """n = 100 halves to 50 and then to 25, an odd periodic extent on which
    red/black colours meet across the seam; the hierarchy stops at 50 and
    the solve still reaches the direct solution."""
    scipy_sparse = pytest.importorskip("scipy.sparse")
    from scipy.sparse.linalg import spsolve
    rng = np.random.default_rng(14)
    nx = ny = 100
    a, bX, bY = rand_coeffs(rng, nx, ny)
    alpha, beta, dx = 1.0, 2e-4, 1 / 100
    b = rng.normal(size=(ny, nx))
    x, its, hist = ho.mg_solve(b, a, bX, bY, alpha, beta, dx, dx, rtol=1e-10, maxiter=200)
    assert 0 < its < 200
    n = nx * ny
    cols = []
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1
        cols.append(ho.apply(e.reshape(ny, nx), a, bX, bY, alpha, beta, dx, dx).ravel())
    A = scipy_sparse.csr_matrix(np.array(cols).T)
    xd = spsolve(A, b.ravel()).reshape(ny, nx)
    assert np.max(np.abs(x - xd)) <= 1e-8 * np.max(np.abs(xd))


def test_vcycle_is_linear_and_symmetric():
    """One V-cycle from a zero guess with symmetric smoothing (red, black
    pre; red, black post) is a linear operator M; check linearity."""
    rng = np.random.default_rng(6)
    nx = ny = 32
    a, bX, bY = rand_coeffs(rng, nx, ny)
    r1, r2 = rng.normal(size=(2, ny, nx))
    e1 = ho.vcycle(r1, a, bX, bY, 1.0, 5e-3, 1 / 32, 1 / 32)
    e2 = ho.vcycle(r2, a, bX, bY, 1.0, 5e-3, 1 / 32, 1 / 32)
    e12 = ho.vcycle(2 * r1 - 3 * r2, a, bX, bY, 1.0, 5e-3, 1 / 32, 1 / 32)
    assert np.max(np.abs(e12 - (2 * e1 - 3 * e2))) <= 1e-12 * np.max(np.abs(e12))


# ------------------------------------------------------------ network oracle
def test_cnn_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(size=co.NPARAMS) * 0.2
    x = rng.normal(size=(2, 16, 16))
    y, cache = co.forward(p, x, quantize=False)
    gy = rng.normal(size=y.shape)
    g = co.backward(cache, gy)
    idx = list(rng.choice(co.NPARAMS, 12, replace=False)) + [0, 143, 144, 2447, 2448, 4751, 4752, 4895]
    for i in idx:
        h = 1e-6
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (np.sum(co.forward(pp, x, quantize=False)[0] * gy)
              - np.sum(co.forward(pm, x, quantize=False)[0] * gy)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * (abs(fd) + 1.0), (i, g[i], fd)


def test_cnn_forward_exact_integer_case():
    """Small-integer inputs and sparse {-1,0,1} weights keep every BF16
    rounding and FP32 sum exact, so the emulated and float64 networks agree
    exactly; a transposed weight layout would change the answer."""
    p = exact_params(np.random.default_rng(11))
    x = np.random.default_rng(12).integers(-2, 3, size=(2, 16, 32)).astype(np.float64)
    yq, _ = co.forward(p, x, quantize=True)
    yd, _ = co.forward(p, x, quantize=False)
    assert np.array_equal(yq.astype(np.float64), yd)
    assert np.any(yd != 0)
    # transposing W2 changes the result
    W1, W2, W3, W4 = co.unpack(p)
    pt = np.concatenate([W1.ravel(), np.transpose(W2, (0, 2, 1)).ravel(), W3.ravel(), W4.ravel()])
    assert not np.array_equal(co.forward(pt, x, quantize=False)[0], yd)


def exact_params(rng):
    """Sparse small-integer weights for exact-arithmetic network tests."""
    p = np.zeros(co.NPARAMS)
    for off, n in ((0, co.NW1), (co.NW1, co.NW2), (co.NW1 + co.NW2, co.NW3),
                   (co.NW1 + co.NW2 + co.NW3, co.NW4)):
        k = max(4, n // 24)
        sel = rng.choice(n, k, replace=False)
        p[off + sel] = rng.choice([-1.0, 1.0], k)
    return p


def test_cnn_homogeneous():
    rng = np.random.default_rng(3)
    p = rng.normal(size=co.NPARAMS) * 0.3
    x = rng.normal(size=(1, 16, 16))
    y1, _ = co.forward(p, x, quantize=True)
    y4, _ = co.forward(p, 4 * x, quantize=True)
    assert np.array_equal(4 * y1, y4)   # scaling by 2^k commutes with rounding


def test_crop_loss_gradient_matches_finite_differences():
    """The training objective's gradient (oracle, float64 network) against
    central differences of the loss."""
    rng = np.random.default_rng(8)
    nx = ny = 48
    a, bX, bY = rand_coeffs(rng, nx, ny)
    R = rng.normal(size=(2, ny, nx))
    crops = [(0, 3, 40), (1, 20, 7)]
    p = rng.normal(size=co.NPARAMS) * 0.2
    kw = dict(a=a, bX=bX, bY=bY, alpha=1.0, beta=2e-3, dx=1 / nx, dy=1 / ny)
    L, g = ho.crop_loss_grad64(p, R, crops, 16, **kw)
    for i in [5, 200, 3000, 4800]:
        h = 1e-6
        pp = p.copy(); pp[i] += h
        pm = p.copy(); pm[i] -= h
        fd = (ho.crop_loss_grad64(pp, R, crops, 16, **kw)[0]
              - ho.crop_loss_grad64(pm, R, crops, 16, **kw)[0]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-5 * (abs(fd) + 1e-3), (i, g[i], fd)
