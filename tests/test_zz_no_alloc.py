"""GPU check of bx_imex_acoustic_dt_noalloc, the entry point that allocates
nothing: the time step it writes to device memory equals, bit for bit, the
one bx_imex_acoustic_dt returns for the same state, on two calls in a row.
The second call is on a smaller level whose largest speed is lower, so a
partial maximum left in the shared scratch by the first call would change
its result.  Both states hold cells of negative pressure (NaN sound speed),
which both entry points skip."""
import ctypes

import numpy as np
import pytest

from bamrex_amd import _lib
from bamrex_amd.imex import DeviceArray, _chk, dev

pytestmark = pytest.mark.gpu

GAMMA, EPS, CFL = 1.4, 1.0, 0.45


def state(nx, ny, seed, speed):
    rng = np.random.default_rng(seed)
    rho = 0.5 + rng.random((ny, nx))
    u = speed * (rng.random((ny, nx)) - 0.5)
    v = speed * (rng.random((ny, nx)) - 0.5)
    p = 0.5 + rng.random((ny, nx))
    p[rng.random((ny, nx)) < 0.01] = -0.3  # NaN sound speed
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def test_noalloc_acoustic_dt_matches_the_allocating_entry_point():
    l = _lib.lib()
    dt_dev = DeviceArray((1,))
    cases = [(700, 500, 3, 4.0, 1.0 / 700, 1.0 / 500),  # 1024 partial maxima
             (96, 64, 5, 0.5, 1.0 / 96, 1.0 / 64)]      # 24, and a lower maximum
    for nx, ny, seed, speed, dx, dy in cases:
        print("[no_alloc] acoustic dt %dx%d seed %d" % (nx, ny, seed), flush=True)
        dU = dev(state(nx, ny, seed, speed))
        try:
            ref = ctypes.c_double()
            _chk(l.bx_imex_acoustic_dt(dU.ptr, nx, ny, dx, dy, GAMMA, EPS, CFL, ctypes.byref(ref),
                                       None), "bx_imex_acoustic_dt")
            dt_dev.upload(np.array([-1.0]))
            _chk(l.bx_imex_acoustic_dt_noalloc(dU.ptr, nx, ny, dx, dy, GAMMA, EPS, CFL, dt_dev.ptr,
                                               None), "bx_imex_acoustic_dt_noalloc")
            out = dt_dev.download()[0]  # blocking copy on the same (null) stream
        finally:
            dU.free()
        assert np.float64(out).tobytes() == np.float64(ref.value).tobytes(), (nx, ny, out, ref.value)
    dt_dev.free()


def test_noalloc_acoustic_dt_rejects_bad_arguments():
    l = _lib.lib()
    dt_dev = DeviceArray((1,))
    dU = DeviceArray((4, 8, 8))
    try:
        assert l.bx_imex_acoustic_dt_noalloc(None, 8, 8, 0.1, 0.1, GAMMA, EPS, CFL, dt_dev.ptr,
                                             None) != 0
        assert l.bx_imex_acoustic_dt_noalloc(dU.ptr, 8, 8, 0.1, 0.1, GAMMA, EPS, CFL, None,
                                             None) != 0
        assert l.bx_imex_acoustic_dt_noalloc(dU.ptr, 0, 8, 0.1, 0.1, GAMMA, EPS, CFL, dt_dev.ptr,
                                             None) != 0
        assert l.bx_imex_acoustic_dt_noalloc(dU.ptr, 8, 8, -0.1, 0.1, GAMMA, EPS, CFL, dt_dev.ptr,
                                             None) != 0
    finally:
        dU.free()
        dt_dev.free()
