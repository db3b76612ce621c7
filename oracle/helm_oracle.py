"""ctypes binding of the pressure-Helmholtz oracle (oracle/oracle_helm.c).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
Helmholtz solver; the product never imports it.  Arrays are numpy (ny, nx),
x fastest; bX is (ny, nx+1), bY is (ny+1, nx).
"""
import ctypes
import os
import subprocess

import numpy as np

import cnn_oracle as co

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_build", "liboracle_helm.so")
_lib = None
D = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        _lib = ctypes.CDLL(PATH)
        _lib.or_helm_apply.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, D, D, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, D]
        _lib.or_helm_apply.restype = None
        _lib.or_helm_gsrb.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D, D, D, D, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int]
        _lib.or_helm_gsrb.restype = None
        _lib.or_helm_restrict.argtypes = [ctypes.c_int, ctypes.c_int, D, D]
        _lib.or_helm_restrict.restype = None
        _lib.or_helm_avg_faces.argtypes = [ctypes.c_int, ctypes.c_int, D, D, D, D]
        _lib.or_helm_avg_faces.restype = None
        _lib.or_helm_vcycle.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_double, D, D, D, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, D, D]
        _lib.or_helm_vcycle.restype = ctypes.c_int
        _lib.or_helm_mg_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double, D, D, D, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, D, D, ctypes.c_double, ctypes.c_int, D]
        _lib.or_helm_mg_solve.restype = ctypes.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(D)


def _c(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def apply(x, a, bX, bY, alpha, beta, dx, dy, gamma=0.0, cX=None, cY=None, rhs=None):
    x, a, bX, bY, cX, cY, rhs = map(_c, (x, a, bX, bY, cX, cY, rhs))
    ny, nx = x.shape
    y = np.empty_like(x)
    lib().or_helm_apply(nx, ny, _p(x), _p(y), _p(a), _p(bX), _p(bY), _p(cX), _p(cY), alpha, beta, gamma,
                        1.0 / dx, 1.0 / dy, _p(rhs))
    return y


def gsrb(phi, rhs, a, bX, bY, alpha, beta, dx, dy, redblack, gamma=0.0, cX=None, cY=None):
    phi = np.array(phi, np.float64, copy=True)
    rhs, a, bX, bY, cX, cY = map(_c, (rhs, a, bX, bY, cX, cY))
    ny, nx = phi.shape
    lib().or_helm_gsrb(nx, ny, _p(phi), _p(rhs), _p(a), _p(bX), _p(bY), _p(cX), _p(cY), alpha, beta, gamma,
                       dx, dy, redblack)
    return phi


def restrict(f):
    f = _c(f)
    c = np.empty((f.shape[0] // 2, f.shape[1] // 2))
    lib().or_helm_restrict(c.shape[1], c.shape[0], _p(f), _p(c))
    return c


def avg_faces(fX, fY):
    fX, fY = _c(fX), _c(fY)
    ncy, ncx = fX.shape[0] // 2, (fX.shape[1] - 1) // 2
    cX = np.empty((ncy, ncx + 1))
    cY = np.empty((ncy + 1, ncx))
    lib().or_helm_avg_faces(ncx, ncy, _p(fX), _p(fY), _p(cX), _p(cY))
    return cX, cY


def vcycle(r, a, bX, bY, alpha, beta, dx, dy, nu1=2, nu2=2, nu_bottom=16):
    r, a, bX, bY = map(_c, (r, a, bX, bY))
    ny, nx = r.shape
    e = np.empty_like(r)
    lib().or_helm_vcycle(nx, ny, dx, dy, alpha, beta, _p(a), _p(bX), _p(bY), nu1, nu2, nu_bottom, _p(r),
                         _p(e))
    return e


def mg_solve(b, a, bX, bY, alpha, beta, dx, dy, x0=None, rtol=1e-10, maxiter=100, nu1=2, nu2=2,
             nu_bottom=16):
    b, a, bX, bY = map(_c, (b, a, bX, bY))
    ny, nx = b.shape
    x = np.zeros_like(b) if x0 is None else np.array(x0, np.float64, copy=True)
    hist = np.zeros(maxiter + 1)
    its = lib().or_helm_mg_solve(nx, ny, dx, dy, alpha, beta, _p(a), _p(bX), _p(bY), nu1, nu2, nu_bottom,
                                 _p(b), _p(x), rtol, maxiter, _p(hist))
    return x, its, hist


def crop_loss_grad64(params, R, crops, cs, a, bX, bY, alpha, beta, dx, dy, border=5):
    """The training objective of the device with a float64 network (no BF16
    rounding): a smooth function of the parameters, for finite-difference
    checks of the gradient formula."""
    ny, nx = a.shape
    Dg = co.op_diag(a, bX, bY, alpha, beta, dx, dy)
    dhx, dhy = beta / (dx * dx), beta / (dy * dy)
    coef = (bX[:, :nx], bX[:, 1:], bY[:ny], bY[1:])
    B = len(crops)
    xs = np.empty((B, cs, cs))
    meta = []
    for b, (f, x0, y0) in enumerate(crops):
        r = co._gather(R, f, x0, y0, cs)
        sig = np.sqrt(np.mean(r * r))
        xs[b] = r / sig
        meta.append((r, sig))
    y, cache = co.forward(params, xs, periodic=False, quantize=False)
    mask = np.zeros((cs, cs), bool)
    mask[border:cs - border, border:cs - border] = True
    loss = 0.0
    gy = np.zeros((B, cs, cs))
    for b, (f, x0, y0) in enumerate(crops):
        r, sig = meta[b]
        Dl = co._gather(Dg[None], 0, x0, y0, cs)
        al = co._gather(a[None], 0, x0, y0, cs)
        c = [co._gather(z[None], 0, x0, y0, cs) for z in coef]
        e = sig * y[b] / Dl
        res = np.where(mask, r - co._apply_local(e, al, *c, alpha, dhx, dhy), 0.0)
        rn2 = max(np.sum(np.where(mask, r * r, 0.0)), np.sum(r * r) * (mask.sum() / float(cs * cs)))
        loss += np.sum(res * res) / rn2
        gy[b] = (-2.0 * co._apply_local(res, al, *c, alpha, dhx, dhy) / rn2) * sig / Dl / B
    return loss / B, co.backward(cache, gy, periodic=False)
