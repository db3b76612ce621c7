"""Initial conditions of the workloads (the reference's Exec/*/Prob.cpp
initdata functions), written into the packed FAB layout of a `Layout`.

All states are set through consv_from_primv with the reference's operand
order (Source/System/Euler/Euler.H:138-165).
"""
import math

import numpy as np

from .level import Layout, chop_domain


def consv_from_primv(q, adiabatic, epsilon, dim):
    """q: sequence of dim+2 arrays (rho, u.., p) -> list of conserved arrays."""
    rho = q[0]
    out = [rho]
    for d in range(dim):
        out.append(rho * q[1 + d])
    p = q[1 + dim]
    vol_internal_energy = p / (adiabatic - 1)
    u_sq = q[1] * q[1]
    for d in range(1, dim):
        u_sq = u_sq + q[1 + d] * q[1 + d]
    out.append(vol_internal_energy + 0.5 * epsilon * rho * u_sq)
    return out


def primv_from_consv(U, adiabatic, epsilon, dim):
    """U: array (dim+2, ...) -> array (dim+2, ...) of (rho, u.., p)."""
    q = np.empty_like(U)
    q[0] = U[0]
    for d in range(dim):
        q[1 + d] = U[1 + d] / U[0]
    u_sq = q[1] * q[1]
    for d in range(1, dim):
        u_sq = u_sq + q[1 + d] * q[1 + d]
    q[1 + dim] = (adiabatic - 1) * (U[1 + dim] - 0.5 * epsilon * U[0] * u_sq)
    return q


def _set_valid(layout, U, b, comps):
    v = layout.valid(U, b)
    for n, c in enumerate(comps):
        v[n][...] = c


# --------------------------------------------------------------- Sod (1D)
def riemann_1d(layout, U, prob_lo, prob_hi, dx, left, right, adiabatic, epsilon=1.0):
    """Exec/RiemannProblem/Prob.cpp:33-121 (1D branch): primitive state
    `left` for x < x_mid, `right` otherwise, x_mid = (prob_lo + prob_hi)/2."""
    x_mid = (prob_lo[0] + prob_hi[0]) * 0.5
    for b in range(len(layout.boxes)):
        (x,) = layout.cell_centers(b, prob_lo, dx)
        sel = x < x_mid
        q = [np.where(sel, left[n], right[n]) for n in range(3)]
        _set_valid(layout, U, b, consv_from_primv(q, adiabatic, epsilon, 1))


def sod_setup(ncell=1024, max_grid_size=None, ng=2):
    """Config: Sod shock tube, 1024 cells on [0,1], gamma 1.4, (1,0,1) |
    (0.125,0,0.1), van Leer, CFL 0.9, transmissive, t_end = 0.25."""
    mgs = max_grid_size or ncell
    boxes = chop_domain((0,), (ncell - 1,), mgs)
    layout = Layout(1, boxes, ng)
    U = layout.new_array()
    dx = (1.0 / ncell,)
    riemann_1d(layout, U, (0.0,), (1.0,), dx, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    return dict(layout=layout, U=U, dx=dx, domain_lo=(0,), domain_hi=(ncell - 1,),
                bc=["transmissive"], adiabatic=1.4, epsilon=1.0, limiter="van_leer",
                cfl=0.9, stop_time=0.25)


# ------------------------------------------------------ quadrant (2D)
QUADRANT_CONFIG = {
    # (rho, u, v, p) per quadrant
    "TR": (1.5, 0.0, 0.0, 1.5),
    "TL": (0.5323, 1.206, 0.0, 0.3),
    "BL": (0.138, 1.206, 1.206, 0.029),
    "BR": (0.5323, 0.0, 1.206, 0.3),
}


def quadrant_2d(layout, U, prob_lo, dx, states, adiabatic, epsilon=1.0):
    """Exec/QuadrantProblem/Prob.cpp:20-97."""
    cs = {k: consv_from_primv([np.float64(x) for x in v], adiabatic, epsilon, 2)
          for k, v in states.items()}
    for b in range(len(layout.boxes)):
        x, y = layout.cell_centers(b, prob_lo, dx)
        tr = (x > 0.5) & (y > 0.5)
        tl = (x < 0.5) & (y > 0.5)
        bl = (x < 0.5) & (y < 0.5)
        comps = []
        for n in range(4):
            c = np.where(tr, cs["TR"][n],
                         np.where(tl, cs["TL"][n], np.where(bl, cs["BL"][n], cs["BR"][n])))
            comps.append(c)
        _set_valid(layout, U, b, comps)


def quadrant_setup(ncell=4096, max_grid_size=128, ng=2):
    """Config: 2D four-quadrant Riemann problem, single uniform level
    ncell^2 on [0,1]^2, max_grid_size 128, gamma 1.4, minbee, CFL 0.8."""
    boxes = chop_domain((0, 0), (ncell - 1, ncell - 1), max_grid_size)
    layout = Layout(2, boxes, ng)
    U = layout.new_array()
    dx = (1.0 / ncell, 1.0 / ncell)
    quadrant_2d(layout, U, (0.0, 0.0), dx, QUADRANT_CONFIG, 1.4)
    return dict(layout=layout, U=U, dx=dx, domain_lo=(0, 0), domain_hi=(ncell - 1, ncell - 1),
                bc=["transmissive"] * 2, adiabatic=1.4, epsilon=1.0, limiter="minbee",
                cfl=0.8, stop_time=-1.0)


# -------------------------------------------------- isentropic vortex (3D)
def _domain_wrap(c, lo, hi):
    """details::domain_wrap, Exec/IsentropicVortex/Prob.cpp:145-156."""
    size = hi - lo
    out = np.where(c < lo, c + np.ceil((lo - c) / size) * size, c)
    out = np.where(c > hi, c - np.ceil((c - hi) / size) * size, out)
    return out


def isentropic_vortex_fields(x, y, time, adiabatic, lam, density_bg, u_bg, v_bg,
                             pressure_bg, x_c, y_c, prob_lo, prob_hi):
    """init_exact_solution, Exec/IsentropicVortex/Prob.cpp:67-129 (primitive
    fields at cell centres x, y; column vortex, independent of z)."""
    pi = math.pi
    dT_prefactor = -(adiabatic - 1) * lam * lam / (8 * adiabatic * pi * pi)
    vel_prefactor = lam / (2 * pi)
    xw = _domain_wrap(x - u_bg * time, prob_lo[0], prob_hi[0])
    yw = _domain_wrap(y - v_bg * time, prob_lo[1], prob_hi[1])
    xx = xw - x_c
    yy = yw - y_c
    r_sq = xx * xx + yy * yy
    dT = dT_prefactor * np.exp(1 - r_sq)
    vel_mult = vel_prefactor * np.exp((1 - r_sq) * 0.5)
    density = density_bg + np.power(1 + dT, 1 / (adiabatic - 1)) - 1
    pressure = pressure_bg + np.power(1 + dT, adiabatic / (adiabatic - 1)) - 1
    u = u_bg - vel_mult * yy
    v = v_bg + vel_mult * xx
    return density, u, v, pressure


def isentropic_vortex(layout, U, prob_lo, prob_hi, dx, time=0.0, adiabatic=1.4, epsilon=1.0,
                      lam=5.0, density_bg=1.0, u_bg=1.0, v_bg=1.0, w_bg=0.0, pressure_bg=1.0,
                      x_c=0.0, y_c=0.0):
    dim = layout.dim
    for b in range(len(layout.boxes)):
        cc = layout.cell_centers(b, prob_lo, dx)
        x, y = cc[0], cc[1]
        if dim == 3:
            # column vortex: the fields do not depend on z; evaluate one slice
            x, y = x[0], y[0]
        rho, u, v, p = isentropic_vortex_fields(x, y, time, adiabatic, lam, density_bg,
                                                u_bg, v_bg, pressure_bg, x_c, y_c, prob_lo,
                                                prob_hi)
        q = [rho, u, v] + ([np.full_like(u, w_bg)] if dim == 3 else []) + [p]
        Q = consv_from_primv(q, adiabatic, epsilon, dim)
        if dim == 3:
            nz = cc[0].shape[0]
            Q = [np.broadcast_to(c, (nz,) + c.shape) for c in Q]
        _set_valid(layout, U, b, Q)


def vortex_setup(ncell=512, dim=3, max_grid_size=64, ng=2):
    """Config: 3D isentropic vortex, periodic ncell^3 on [-5,5]^3, gamma 1.4,
    lambda 5 advected at (1,1,0); superbee limiting density/mom/energy on
    themselves (problem_inputs_moving), CFL 0.9."""
    lo = tuple([0] * dim)
    hi = tuple([ncell - 1] * dim)
    boxes = chop_domain(lo, hi, max_grid_size)
    layout = Layout(dim, boxes, ng)
    U = layout.new_array()
    prob_lo = tuple([-5.0] * dim)
    prob_hi = tuple([5.0] * dim)
    dx = tuple([10.0 / ncell] * dim)
    isentropic_vortex(layout, U, prob_lo, prob_hi, dx)
    return dict(layout=layout, U=U, dx=dx, domain_lo=lo, domain_hi=hi, bc=["periodic"] * dim,
                adiabatic=1.4, epsilon=1.0, limiter="superbee",
                limiter_indices=list(range(dim + 2)), limit_on_primv=[0] * (dim + 2),
                cfl=0.9, stop_time=-1.0, prob_lo=prob_lo, prob_hi=prob_hi)


# ----------------------------------------------- batched uneven boxes (2D)
def post_shock(M, rho_amb, p_amb, adiabatic):
    """Post-shock primitive state (rho, u, p) behind a shock of Mach M moving
    into gas at rest (Source/System/Euler/RH.H:17-29)."""
    g = adiabatic
    c_amb = np.sqrt(p_amb * g / rho_amb)
    rho = rho_amb * ((g + 1) * M * M) / (2 + (g - 1) * M * M)
    u = c_amb * 2 * (M * M - 1) / ((g + 1) * M)
    p = p_amb * (1 + 2 * (g / (g + 1)) * (M * M - 1))
    return rho, u, p


def uneven_setup(nboxes=3000, smin=8, smax=96, ng=2, seed=20261202, adiabatic=1.4):
    """Config: batched uneven AMR level -- nboxes 2D boxes of sizes drawn
    uniformly from smin..smax per side (seeded), ng ghost cells, seeded smooth
    background rho, p in [0.5, 2], |u| <= 1, plus 1-3 random planar shock
    fronts of Mach 1.5-3 per box.  Boxes sit on a grid of slots separated by
    gaps, so they share no ghost data; the state is defined on ghost cells too.
    """
    rng = np.random.default_rng(seed)
    sizes = rng.integers(smin, smax + 1, size=(nboxes, 2))
    ncol = int(math.ceil(math.sqrt(nboxes)))
    slot = smax + 4 * ng
    boxes = []
    for b in range(nboxes):
        cx, cy = b % ncol, b // ncol
        lo = (cx * slot + 2 * ng, cy * slot + 2 * ng)
        hi = (lo[0] + int(sizes[b, 0]) - 1, lo[1] + int(sizes[b, 1]) - 1)
        boxes.append((lo, hi))
    nrow = (nboxes + ncol - 1) // ncol
    domain_lo = (0, 0)
    domain_hi = (ncol * slot - 1, nrow * slot - 1)
    layout = Layout(2, boxes, ng)
    U = layout.new_array()
    h = 1.0 / 1024.0
    dx = (h, h)
    for b in range(nboxes):
        lo, hi = boxes[b]
        f = layout.fab(U, b)
        _, nyd, nxd = f.shape
        xi = (np.arange(nxd) + lo[0] - ng + 0.5) * h
        yi = (np.arange(nyd) + lo[1] - ng + 0.5) * h
        Y, X = np.meshgrid(yi, xi, indexing="ij")
        k = rng.uniform(2.0, 12.0, size=(4, 2)) * 2 * np.pi
        ph = rng.uniform(0, 2 * np.pi, size=(4,))
        s_rho = 0.5 * (np.sin(k[0, 0] * X + ph[0]) * np.cos(k[0, 1] * Y) + np.sin(k[1, 0] * X + k[1, 1] * Y + ph[1]))
        s_p = 0.5 * (np.cos(k[2, 0] * X + ph[2]) * np.sin(k[2, 1] * Y + ph[3]) + np.cos(k[3, 0] * X - k[3, 1] * Y))
        rho = 1.25 + 0.75 * s_rho
        p = 1.25 + 0.75 * s_p
        u = 0.7 * np.sin(k[1, 1] * Y + ph[2])
        v = 0.7 * np.cos(k[0, 0] * X + ph[1])
        nshock = int(rng.integers(1, 4))
        for _ in range(nshock):
            ang = rng.uniform(0, 2 * np.pi)
            nx_, ny_ = math.cos(ang), math.sin(ang)
            x0 = (lo[0] + rng.uniform(0.2, 0.8) * (hi[0] - lo[0] + 1)) * h
            y0 = (lo[1] + rng.uniform(0.2, 0.8) * (hi[1] - lo[1] + 1)) * h
            M = rng.uniform(1.5, 3.0)
            behind = (X - x0) * nx_ + (Y - y0) * ny_ < 0
            r2, du, p2 = post_shock(M, rho, p, adiabatic)
            rho = np.where(behind, r2, rho)
            p = np.where(behind, p2, p)
            u = np.where(behind, u + du * nx_, u)
            v = np.where(behind, v + du * ny_, v)
        # keep |u| <= 1 in the background sense: shocked regions may exceed it
        comps = consv_from_primv([rho, u, v, p], adiabatic, 1.0, 2)
        for n in range(4):
            f[n][...] = comps[n]
    return dict(layout=layout, U=U, dx=dx, domain_lo=domain_lo, domain_hi=domain_hi,
                bc=["transmissive"] * 2, adiabatic=adiabatic, epsilon=1.0, limiter="minbee",
                cfl=0.8, stop_time=-1.0)
