// CPU ORACLE -- test infrastructure only.
//
// A plain C++ restatement of the reference's explicit hot path, used by the
// tests (as the checker), by __graft_entry__.smoke() and by bench.py's
// cpu_baseline leg.  It is never part of the product path.
//
// It follows the reference operation by operation, per box, with the same
// temporaries the reference allocates:
//   Euler.H:138-254        consv_from_primv / primv_from_consv / euler_flux
//   Euler.cpp:10-74        compute_flux_function, max_wave_speed
//   Limiters.H:29-81       slope limiters, minmod
//   CalcLimitingVal.H:9-22 limiting variable
//   Flux_Helpers.cpp:52-149 reconstruct_and_half_time_step (temporaries on
//                          grow(bx,1), Q on grow(gbx,dir,1))
//   HLLC.H:213-337         estimate_wave_speed_HLLC_Toro, HLLC_interm_state_replace
//   HLLC.cpp:64-173        compute_HLLC_flux_LR_euler
//   MUSCL_Hancock.cpp:15-52 compute_MUSCL_hancock_flux (swapped L/R arrays)
//   Fluxes.H:192-227       compute_HLLC_flux (first order)
//   Force.cpp:24-146       compute_force_flux_LR, compute_LF_flux
//   SLIC.cpp:12-50         compute_SLIC_flux
//   AmrLevelAdv_advance.cpp:261-360 advance_explicit (x,y,z order, full dt,
//                          update, fill_consv_boundary after every sweep)
//   BCs.cpp:215-442        bc types -> FillBoundary + PhysBCFunct
//   AmrLevelAdv.cpp:395-619 estTimeStep, computeInitialDt / computeNewDt
//   main.cpp:76-94         time loop
// Compiled with -ffp-contract=off and no fast-math.  ORACLE_BXPOW selects the
// portable pow of the product (bamrex_amd/csrc/bx_pow.h) instead of
// std::pow, to separate the one libm dependency from everything else.
//
// Parity status of each part (which check pins it) is listed in the header of
// tests/test_oracle.py.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <vector>

#include <omp.h>

#ifdef ORACLE_BXPOW
#include "../bamrex_amd/csrc/bx_pow.h"
static inline double or_pow(double x, double y) { return bx_pow(x, y); }
// the specialised exponents of the HLLC estimate (same operations as the kernels)
static inline double or_pow_negz(double x, double y, double gamma)
{
    return bx_pow_negz(x, y, bx_pow_hllc_setup(gamma));
}
static inline double or_pow_invz(double x, double y, double gamma)
{
    return bx_pow_invz(x, y, bx_pow_hllc_setup(gamma));
}
#else
static inline double or_pow(double x, double y) { return std::pow(x, y); }
static inline double or_pow_negz(double x, double y, double) { return std::pow(x, y); }
static inline double or_pow_invz(double x, double y, double) { return std::pow(x, y); }
#endif

namespace oracle {

typedef std::array<double, 5> Vec;

struct Parm
{
    int    D;
    double adiabatic;
    double epsilon;
    int    limiter; // 0 minbee, 1 van_leer, 2 van_albada, 3 superbee
    int    limiter_indices[5];
    bool   limit_on_primv[5];
    bool   need_primitive_for_limiting;
};

struct Box
{
    int lo[3];
    int hi[3];
};

static Box grow(const Box &b, int D, int n)
{
    Box r = b;
    for (int d = 0; d < D; ++d)
    {
        r.lo[d] -= n;
        r.hi[d] += n;
    }
    return r;
}

static Box grow_dir(const Box &b, int dir, int n)
{
    Box r = b;
    r.lo[dir] -= n;
    r.hi[dir] += n;
    return r;
}

static Box surrounding_nodes(const Box &b, int dir)
{
    Box r = b;
    r.hi[dir] += 1;
    return r;
}

// amrex::FArrayBox / Array4 with (i, j, k, n) indexing over box bx
struct Fab
{
    Box                 bx;
    int                 nc = 0;
    long long           jst = 0, kst = 0, nst = 0;
    std::vector<double> a;
    void define(const Box &b, int ncomp)
    {
        bx  = b;
        nc  = ncomp;
        jst = bx.hi[0] - bx.lo[0] + 1;
        kst = jst * (bx.hi[1] - bx.lo[1] + 1);
        nst = kst * (bx.hi[2] - bx.lo[2] + 1);
        a.assign(nst * nc, 0.0);
    }
    double &operator()(int i, int j, int k, int n)
    {
        return a[(i - bx.lo[0]) + (j - bx.lo[1]) * jst + (k - bx.lo[2]) * kst
                 + n * nst];
    }
    double operator()(int i, int j, int k, int n) const
    {
        return a[(i - bx.lo[0]) + (j - bx.lo[1]) * jst + (k - bx.lo[2]) * kst
                 + n * nst];
    }
};

template <class F>
static void for_box(const Box &b, F f)
{
    for (int k = b.lo[2]; k <= b.hi[2]; ++k)
        for (int j = b.lo[1]; j <= b.hi[1]; ++j)
            for (int i = b.lo[0]; i <= b.hi[0]; ++i)
                f(i, j, k);
}

// ------------------------------------------------------------ Euler.H
// consv_from_primv  Euler.H:138-165
static Vec consv_from_primv(const Vec &q, double adiabatic, double epsilon,
                            int D)
{
    Vec c{};
    c[0] = q[0];
    for (int d = 0; d < D; ++d)
        c[1 + d] = q[0] * q[1 + d];
    const double pressure            = q[1 + D];
    const double vol_internal_energy = pressure / (adiabatic - 1);
    double       u_sq                = q[1] * q[1];
    for (int d = 1; d < D; ++d)
        u_sq = u_sq + q[1 + d] * q[1 + d];
    c[1 + D] = vol_internal_energy + 0.5 * epsilon * q[0] * u_sq;
    return c;
}

// primv_from_consv  Euler.H:167-198
static Vec primv_from_consv(const Fab &U, double adiabatic, double epsilon,
                            int i, int j, int k, int D)
{
    Vec q{};
    q[0] = U(i, j, k, 0);
    for (int d = 0; d < D; ++d)
        q[1 + d] = U(i, j, k, 1 + d) / U(i, j, k, 0);
    double u_sq = q[1] * q[1];
    for (int d = 1; d < D; ++d)
        u_sq = u_sq + q[1 + d] * q[1 + d];
    q[1 + D] = (adiabatic - 1)
               * (U(i, j, k, 1 + D) - 0.5 * epsilon * U(i, j, k, 0) * u_sq);
    return q;
}

// euler_flux<dir>  Euler.H:211-244
static Vec euler_flux(const Fab &U, double adiabatic, double epsilon, int i,
                      int j, int k, int dir, int D)
{
    const Vec    q        = primv_from_consv(U, adiabatic, epsilon, i, j, k, D);
    const double pressure = q[1 + D];
    const double energy   = U(i, j, k, 1 + D);
    Vec          f{};
    f[0] = U(i, j, k, 1 + dir);
    for (int d = 0; d < D; ++d)
        f[1 + d] = U(i, j, k, 1 + dir) * q[1 + d];
    f[1 + dir] += pressure / epsilon;
    f[1 + D] = (energy + pressure) * q[1 + dir];
    return f;
}

// compute_flux_function<dir>  Euler.cpp:10-44
static void compute_flux_function(const Box &bx, Fab &flux_func, const Fab &U,
                                  double adiabatic, double epsilon, int dir,
                                  int D)
{
    for_box(bx, [&](int i, int j, int k) {
        const Vec    q        = primv_from_consv(U, adiabatic, epsilon, i, j, k, D);
        const double pressure = q[1 + D];
        const double energy   = U(i, j, k, 1 + D);
        flux_func(i, j, k, 0) = U(i, j, k, 1 + dir);
        for (int d = 0; d < D; ++d)
            flux_func(i, j, k, 1 + d) = U(i, j, k, 1 + dir) * q[1 + d];
        flux_func(i, j, k, 1 + dir) += pressure / epsilon;
        flux_func(i, j, k, 1 + D) = (energy + pressure) * q[1 + dir];
    });
}

// ---------------------------------------------------------- Limiters.H
static double minbee(double r)
{
    return r <= 0 ? 0 : (r > 1 ? std::fmin(2. / (1 + r), 1.) : r);
}
static double van_leer(double r)
{
    return r <= 0 ? 0 : std::fmin(2. * r / (1 + r), 2. / (1 + r));
}
static double van_albada(double r)
{
    return r <= 0 ? 0 : std::fmin(r * (1 + r) / (1 + r * r), 2. / (1 + r));
}
static double superbee(double r)
{
    return (r <= 0) ? 0
                    : ((0 < r && r <= 0.5)
                           ? 2 * r
                           : ((0.5 < r && r <= 1)
                                  ? 1
                                  : std::fmin(r, std::fmin(2. / (1 + r), 2.))));
}
static double slope_limit(int limiter, double r)
{
    switch (limiter)
    {
    case 0: return minbee(r);
    case 1: return van_leer(r);
    case 2: return van_albada(r);
    default: return superbee(r);
    }
}
static double minmod(double delta_L, double delta_R)
{
    return (delta_L > 0 && delta_R > 0)
               ? std::fmin(delta_L, delta_R)
               : ((delta_L < 0 && delta_R < 0) ? std::fmax(delta_L, delta_R)
                                               : 0);
}

// calc_limiting_var  CalcLimitingVal.H:9-22
static double calc_limiting_var(const Fab &U, const Fab &Q, const Parm &parm,
                                int i, int j, int k, int n)
{
    if (parm.limit_on_primv[n])
        return Q(i, j, k, parm.limiter_indices[n]);
    return U(i, j, k, parm.limiter_indices[n]);
}

// UtoQ  FluidSystem.cpp:161-178
static void UtoQ(const Box &bx, Fab &Q, const Fab &U, const Parm &parm)
{
    for_box(bx, [&](int i, int j, int k) {
        const Vec q = primv_from_consv(U, parm.adiabatic, parm.epsilon, i, j, k,
                                       parm.D);
        for (int n = 0; n < parm.D + 2; ++n)
            Q(i, j, k, n) = q[n];
    });
}

// reconstruct_and_half_time_step<limiter,dir>  Flux_Helpers.cpp:52-149
static void reconstruct_and_half_time_step(const Box &gbx, Fab &half_L,
                                           Fab &half_R, const Fab &consv,
                                           const double dx_arr[3], double dt,
                                           const Parm &parm, int dir)
{
    const int    D      = parm.D;
    const int    NSTATE = D + 2;
    const double dx     = dx_arr[dir];
    const Box    gbx2   = grow_dir(gbx, dir, 1);
    Fab          flux_func_L, flux_func_R, Q;
    flux_func_L.define(gbx, NSTATE);
    flux_func_R.define(gbx, NSTATE);
    if (parm.need_primitive_for_limiting)
    {
        Q.define(gbx2, NSTATE);
        UtoQ(gbx2, Q, consv, parm);
    }
    const int i_off = (dir == 0) ? 1 : 0;
    const int j_off = (dir == 1) ? 1 : 0;
    const int k_off = (dir == 2) ? 1 : 0;

    for (int n = 0; n < NSTATE; ++n)
        for_box(gbx, [&](int i, int j, int k) {
            const double val_C = calc_limiting_var(consv, Q, parm, i, j, k, n);
            const double val_L = calc_limiting_var(consv, Q, parm, i - i_off,
                                                   j - j_off, k - k_off, n);
            const double val_R = calc_limiting_var(consv, Q, parm, i + i_off,
                                                   j + j_off, k + k_off, n);
            const double numerator = val_C - val_L;
            const double denom     = val_R - val_C;
            double       limiter_val;
            if (numerator == 0. && denom == 0.)
                limiter_val = slope_limit(parm.limiter, 1);
            else if (denom == 0)
                limiter_val = 0;
            else
                limiter_val = slope_limit(parm.limiter, numerator / denom);

            const double omega      = 0;
            const double delta_left = consv(i, j, k, n)
                                      - consv(i - i_off, j - j_off, k - k_off, n);
            const double delta_right
                = consv(i + i_off, j + j_off, k + k_off, n) - consv(i, j, k, n);
            const double delta = 0.5 * (1 + omega) * delta_left
                                 + 0.5 * (1 - omega) * delta_right;
            half_L(i, j, k, n) = consv(i, j, k, n) - 0.5 * limiter_val * delta;
            half_R(i, j, k, n) = consv(i, j, k, n) + 0.5 * limiter_val * delta;
        });

    compute_flux_function(gbx, flux_func_L, half_L, parm.adiabatic,
                          parm.epsilon, dir, D);
    compute_flux_function(gbx, flux_func_R, half_R, parm.adiabatic,
                          parm.epsilon, dir, D);

    const double hdtdx = 0.5 * dt / dx;
    for (int n = 0; n < NSTATE; ++n)
        for_box(gbx, [&](int i, int j, int k) {
            half_L(i, j, k, n)
                -= hdtdx * (flux_func_R(i, j, k, n) - flux_func_L(i, j, k, n));
            half_R(i, j, k, n)
                -= hdtdx * (flux_func_R(i, j, k, n) - flux_func_L(i, j, k, n));
        });
}

// estimate_wave_speed_HLLC_Toro<dir>  HLLC.H:213-309
static void estimate_wave_speed_HLLC_Toro(const Fab &cL, const Fab &cR, int i,
                                          int j, int k, double adiabatic_L,
                                          double adiabatic_R, double epsilon,
                                          int dir, int D, double &speed_L,
                                          double &speed_R, double &speed_interm)
{
    const int i_off = (dir == 0) ? 1 : 0;
    const int j_off = (dir == 1) ? 1 : 0;
    const int k_off = (dir == 2) ? 1 : 0;
    Vec primv_L = primv_from_consv(cL, adiabatic_L, epsilon, i - i_off,
                                   j - j_off, k - k_off, D);
    Vec primv_R = primv_from_consv(cR, adiabatic_R, epsilon, i, j, k, D);

    double       density_L = primv_L[0];
    double       density_R = primv_R[0];
    double       p_L       = primv_L[1 + D];
    double       p_R       = primv_R[1 + D];
    const double u_L       = primv_L[1 + dir];
    const double u_R       = primv_R[1 + dir];

    constexpr double min_rho = 1e-14;
    constexpr double min_p   = 1e-14;
    density_L                = std::max(density_L, min_rho);
    density_R                = std::max(density_R, min_rho);
    p_L                      = std::max(p_L, min_p);
    p_R                      = std::max(p_R, min_p);

    const double c_L = std::sqrt(adiabatic_L * p_L / (density_L * epsilon));
    const double c_R = std::sqrt(adiabatic_R * p_R / (density_R * epsilon));
    const double adiabatic = adiabatic_L;

    const double z = (adiabatic - 1) / (2 * adiabatic);
    const double p_rr = or_pow_invz(
        (c_L + c_R - (adiabatic - 1) / 2 * (u_R - u_L))
            / (c_L * or_pow_negz(p_L, -z, adiabatic) + c_R * or_pow_negz(p_R, -z, adiabatic)),
        1 / z, adiabatic);
    const double q_L
        = (p_rr > p_L) ? std::sqrt(1 + (adiabatic + 1) / (2 * adiabatic)
                                           * (p_rr / p_L - 1))
                       : 1;
    const double q_R
        = (p_rr > p_R) ? std::sqrt(1 + (adiabatic + 1) / (2 * adiabatic)
                                           * (p_rr / p_R - 1))
                       : 1;
    speed_L = u_L - c_L * q_L;
    speed_R = u_R + c_R * q_R;
    const double denom
        = (density_L * (speed_L - u_L) - density_R * (speed_R - u_R));
    speed_interm = (denom != 0) ? ((p_R - p_L) / epsilon
                                   + density_L * u_L * (speed_L - u_L)
                                   - density_R * u_R * (speed_R - u_R))
                                      / denom
                                : 0;
}

// HLLC_interm_state_replace<dir>  HLLC.H:311-337
static void HLLC_interm_state_replace(Vec &primv, double energy, double speed_K,
                                      double speed_interm, double epsilon,
                                      int dir, int D)
{
    const double density  = primv[0];
    const double u_K      = primv[1 + dir];
    const double pressure = primv[1 + D];
    primv[0]              = 1;
    primv[1 + dir]        = speed_interm;
    const double scale_factor
        = density * (speed_K - u_K) / (speed_K - speed_interm);
    primv[1 + D] = energy / density
                   + (speed_interm - u_K)
                         * (speed_interm * epsilon
                            + pressure / (density * (speed_K - u_K)));
    for (int i = 0; i < D + 2; ++i)
        primv[i] *= scale_factor;
}

// compute_HLLC_flux_LR_euler<dir>  HLLC.cpp:64-173
static void compute_HLLC_flux_LR(const Box &bx, Fab &flux, const Fab &cL,
                                 const Fab &cR, const Parm &parm, int dir)
{
    const int    D       = parm.D;
    const int    NSTATE  = D + 2;
    const Box    flux_bx = surrounding_nodes(bx, dir);
    const int    i_off   = (dir == 0) ? 1 : 0;
    const int    j_off   = (dir == 1) ? 1 : 0;
    const int    k_off   = (dir == 2) ? 1 : 0;
    const double aL = parm.adiabatic, aR = parm.adiabatic, eps = parm.epsilon;
    for_box(flux_bx, [&](int i, int j, int k) {
        double speed_L, speed_R, speed_interm;
        estimate_wave_speed_HLLC_Toro(cL, cR, i, j, k, aL, aR, eps, dir, D,
                                      speed_L, speed_R, speed_interm);
        if (0 <= speed_L)
        {
            const Vec f = euler_flux(cL, aL, eps, i - i_off, j - j_off,
                                     k - k_off, dir, D);
            for (int n = 0; n < NSTATE; ++n)
                flux(i, j, k, n) = f[n];
        }
        else if (speed_L <= 0 && 0 <= speed_interm)
        {
            const Vec f = euler_flux(cL, aL, eps, i - i_off, j - j_off,
                                     k - k_off, dir, D);
            Vec       pv = primv_from_consv(cL, aL, eps, i - i_off, j - j_off,
                                            k - k_off, D);
            HLLC_interm_state_replace(
                pv, cL(i - i_off, j - j_off, k - k_off, 1 + D), speed_L,
                speed_interm, eps, dir, D);
            for (int n = 0; n < NSTATE; ++n)
                flux(i, j, k, n)
                    = f[n]
                      + speed_L * (pv[n] - cL(i - i_off, j - j_off, k - k_off, n));
        }
        else if (speed_interm <= 0 && 0 <= speed_R)
        {
            const Vec f  = euler_flux(cR, aR, eps, i, j, k, dir, D);
            Vec       pv = primv_from_consv(cR, aR, eps, i, j, k, D);
            HLLC_interm_state_replace(pv, cR(i, j, k, 1 + D), speed_R,
                                      speed_interm, eps, dir, D);
            for (int n = 0; n < NSTATE; ++n)
                flux(i, j, k, n) = f[n] + speed_R * (pv[n] - cR(i, j, k, n));
        }
        else if (0 >= speed_R)
        {
            const Vec f = euler_flux(cR, aR, eps, i, j, k, dir, D);
            for (int n = 0; n < NSTATE; ++n)
                flux(i, j, k, n) = f[n];
        }
        else
        {
            for (int n = 0; n < NSTATE; ++n)
                flux(i, j, k, n) = std::nan("");
        }
    });
}

// compute_MUSCL_hancock_flux<dir>  MUSCL_Hancock.cpp:15-52
static void compute_MUSCL_hancock_flux(const Box &bx, Fab &flux,
                                       const Fab &consv, const double dx[3],
                                       double dt, const Parm &parm, int dir)
{
    const Box bx_g1 = grow(bx, parm.D, 1);
    Fab       half_L, half_R;
    half_L.define(bx_g1, parm.D + 2);
    half_R.define(bx_g1, parm.D + 2);
    reconstruct_and_half_time_step(bx_g1, half_L, half_R, consv, dx, dt, parm,
                                   dir);
    compute_HLLC_flux_LR(bx, flux, half_R, half_L, parm, dir);
}

// compute_force_flux_LR<dir>  Force.cpp:24-104.  Face i takes cL at i-1
// and cR at i; the flux array covers surroundingNodes(bx, dir) (the
// reference's flux_bx = growHi(bx, dir, 1), the same index range).
static void compute_force_flux_LR(const Box &bx, Fab &flux, const Fab &cL,
                                  const Fab &cR, const double dx[3], double dt,
                                  const Parm &parm, int dir)
{
    const int    D      = parm.D;
    const int    NSTATE = D + 2;
    const double dxd    = dx[dir];
    const double a = parm.adiabatic, e = parm.epsilon;
    const Box    ghost_bx = grow_dir(bx, dir, 1);
    const Box    flux_bx  = surrounding_nodes(bx, dir);
    Fab          flux_func_values_L, flux_func_values_R, half_updated_values;
    flux_func_values_L.define(ghost_bx, NSTATE);
    flux_func_values_R.define(ghost_bx, NSTATE);
    half_updated_values.define(ghost_bx, NSTATE);
    compute_flux_function(ghost_bx, flux_func_values_L, cL, a, e, dir, D);
    compute_flux_function(ghost_bx, flux_func_values_R, cR, a, e, dir, D);
    const int io = (dir == 0) ? 1 : 0;
    const int jo = (dir == 1) ? 1 : 0;
    const int ko = (dir == 2) ? 1 : 0;
    // half of the Lax-Friedrichs flux, and the Richtmyer half-step state
    for (int n = 0; n < NSTATE; ++n)
        for_box(flux_bx, [&](int i, int j, int k) {
            flux(i, j, k, n)
                = 0.5
                  * ((0.5 * dxd / dt) * (cL(i - io, j - jo, k - ko, n) - cR(i, j, k, n))
                     + 0.5
                           * (flux_func_values_L(i - io, j - jo, k - ko, n)
                              + flux_func_values_R(i, j, k, n)));
            half_updated_values(i, j, k, n)
                = 0.5 * (cL(i - io, j - jo, k - ko, n) + cR(i, j, k, n))
                  + (dt / (2 * dxd))
                        * (flux_func_values_L(i - io, j - jo, k - ko, n)
                           - flux_func_values_R(i, j, k, n));
        });
    // plus half of the Richtmyer flux (the L buffer is reused, as there)
    compute_flux_function(flux_bx, flux_func_values_L, half_updated_values, a, e,
                          dir, D);
    for (int n = 0; n < NSTATE; ++n)
        for_box(flux_bx, [&](int i, int j, int k) {
            flux(i, j, k, n) += 0.5 * flux_func_values_L(i, j, k, n);
        });
}

// compute_LF_flux<dir>  Force.cpp:106-146
static void compute_LF_flux(const Box &bx, Fab &flux, const Fab &consv,
                            const double dx[3], double dt, const Parm &parm,
                            int dir)
{
    const int    D        = parm.D;
    const int    NSTATE   = D + 2;
    const double dxd      = dx[dir];
    const Box    ghost_bx = grow(bx, D, 1);
    const Box    flux_bx  = surrounding_nodes(bx, dir);
    Fab          flux_func_values;
    flux_func_values.define(ghost_bx, NSTATE);
    compute_flux_function(ghost_bx, flux_func_values, consv, parm.adiabatic,
                          parm.epsilon, dir, D);
    const int io = (dir == 0) ? 1 : 0;
    const int jo = (dir == 1) ? 1 : 0;
    const int ko = (dir == 2) ? 1 : 0;
    for (int n = 0; n < NSTATE; ++n)
        for_box(flux_bx, [&](int i, int j, int k) {
            flux(i, j, k, n)
                = 0.5 * dxd / dt * (consv(i - io, j - jo, k - ko, n) - consv(i, j, k, n))
                  + 0.5 * (flux_func_values(i - io, j - jo, k - ko, n)
                           + flux_func_values(i, j, k, n));
        });
}

// compute_SLIC_flux<dir>  SLIC.cpp:12-50: reconstruction + half step on
// grow(bx, dir, 1), then the FORCE flux of (half_R at i-1, half_L at i)
static void compute_SLIC_flux(const Box &bx, Fab &flux, const Fab &consv,
                              const double dx[3], double dt, const Parm &parm,
                              int dir)
{
    const Box bx_g1 = grow_dir(bx, dir, 1);
    Fab       half_stepped_L, half_stepped_R;
    half_stepped_L.define(bx_g1, parm.D + 2);
    half_stepped_R.define(bx_g1, parm.D + 2);
    reconstruct_and_half_time_step(bx_g1, half_stepped_L, half_stepped_R, consv,
                                   dx, dt, parm, dir);
    compute_force_flux_LR(bx, flux, half_stepped_R, half_stepped_L, dx, dt, parm,
                          dir);
}

// advance_explicit's flux dispatch (AmrLevelAdv_advance.cpp:311-326):
// 0 muscl_hancock, 1 hllc, 2 slic, 3 force, 4 lax_friedrichs
static void compute_flux(int num_method, const Box &bx, Fab &flux, const Fab &consv,
                         const double dx[3], double dt, const Parm &parm, int dir)
{
    switch (num_method)
    {
    case 0: compute_MUSCL_hancock_flux(bx, flux, consv, dx, dt, parm, dir); break;
    case 1: compute_HLLC_flux_LR(bx, flux, consv, consv, parm, dir); break;
    case 2: compute_SLIC_flux(bx, flux, consv, dx, dt, parm, dir); break;
    case 3: compute_force_flux_LR(bx, flux, consv, consv, dx, dt, parm, dir); break;
    case 4: compute_LF_flux(bx, flux, consv, dx, dt, parm, dir); break;
    default: throw std::logic_error("Internal error");
    }
}

// ----------------------------------------------------------------- level
struct Level
{
    Parm             parm;
    int              D, ng;
    int              dlo[3], dhi[3];
    int              bclo[3], bchi[3]; // 0 transmissive, 1 periodic, 2 reflective
    std::vector<Box> valid;
    std::vector<Fab> fabs;
};

static bool inside(const Box &b, int i, int j, int k)
{
    return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1]
           && k >= b.lo[2] && k <= b.hi[2];
}

// FillBoundary(periodicity) + PhysBCFunct, BCs.cpp:429-442.
// FillBoundary copies valid data of every box (and of its periodic images)
// into the ghost cells of the others; PhysBCFunct then fills, direction by
// direction, the ghost cells outside a non-periodic domain face:
// foextrap = nearest interior value, reflect_even/odd = mirror (BCs.cpp:215-397).
static void fill_consv_boundary(Level &L)
{
    const int D  = L.D;
    const int NC = D + 2;
    int       len[3];
    for (int d = 0; d < 3; ++d)
        len[d] = L.dhi[d] - L.dlo[d] + 1;
    // boxes fill their own ghost cells from valid cells only, so the boxes
    // are independent; the source box of the previous ghost cell is tried
    // first (the boxes do not overlap, so the result does not depend on it)
#pragma omp parallel for schedule(dynamic)
    for (long bl = 0; bl < (long)L.fabs.size(); ++bl)
    {
        const size_t b    = (size_t)bl;
        Fab         &F    = L.fabs[b];
        const Box   &vb   = L.valid[b];
        size_t       last = 0;
        for_box(F.bx, [&](int i, int j, int k) {
            if (inside(vb, i, j, k))
                return;
            int x[3] = { i, j, k };
            for (int d = 0; d < D; ++d)
            {
                if (L.bclo[d] == 1)
                {
                    if (x[d] < L.dlo[d])
                        x[d] += len[d];
                    else if (x[d] > L.dhi[d])
                        x[d] -= len[d];
                }
                if (x[d] < L.dlo[d] || x[d] > L.dhi[d])
                    return; // physical boundary: PhysBCFunct below
            }
            for (size_t t = 0; t <= L.valid.size(); ++t)
            {
                const size_t s = t == 0 ? last : t - 1;
                if (inside(L.valid[s], x[0], x[1], x[2]))
                {
                    for (int n = 0; n < NC; ++n)
                        F(i, j, k, n) = L.fabs[s](x[0], x[1], x[2], n);
                    last = s;
                    return;
                }
            }
        });
        for (int d = 0; d < D; ++d)
        {
            if (L.bclo[d] == 1)
                continue;
            for_box(F.bx, [&](int i, int j, int k) {
                int x[3] = { i, j, k };
                int side;
                if (x[d] < L.dlo[d])
                    side = 0;
                else if (x[d] > L.dhi[d])
                    side = 1;
                else
                    return;
                const int type = side ? L.bchi[d] : L.bclo[d];
                int       y[3] = { i, j, k };
                bool      odd  = false;
                if (type == 2)
                {
                    y[d] = side ? 2 * L.dhi[d] + 1 - x[d] : 2 * L.dlo[d] - 1 - x[d];
                    odd  = true;
                }
                else
                    y[d] = side ? L.dhi[d] : L.dlo[d];
                for (int n = 0; n < NC; ++n)
                {
                    double v = F(y[0], y[1], y[2], n);
                    if (odd && n == 1 + d)
                        v = -v;
                    F(i, j, k, n) = v;
                }
            });
        }
    }
}

// one sweep of advance_explicit (AmrLevelAdv_advance.cpp:275-357) over all
// boxes, writing the valid cells of out
// Sees the face fluxes of one box and direction (the flux-register hook of
// the two-level driver; called concurrently for different boxes).
typedef std::function<void(int box, int dir, const Fab &flux)> FluxHook;

static void sweep(Level &L, std::vector<Fab> &in, std::vector<Fab> &out,
                  int d, double dt, const double dx[3], int num_method,
                  const FluxHook *hook = nullptr)
{
    const int D     = L.D;
    const int ncomp = D + 2;
    const int iOff = (d == 0), jOff = (d == 1), kOff = (d == 2);
#pragma omp parallel for schedule(dynamic)
    for (long b = 0; b < (long)L.valid.size(); ++b)
    {
        const Box &bx = L.valid[b];
        Fab        flux;
        flux.define(surrounding_nodes(bx, d), ncomp);
        compute_flux(num_method, bx, flux, in[b], dx, dt, L.parm, d);
        if (hook)
            (*hook)((int)b, d, flux);
        const Fab &Uin  = in[b];
        Fab       &Uout = out[b];
        for (int n = 0; n < ncomp; ++n)
            for_box(bx, [&](int i, int j, int k) {
                Uout(i, j, k, n)
                    = Uin(i, j, k, n)
                      - (dt / dx[d])
                            * (flux(i + iOff, j + jOff, k + kOff, n)
                               - flux(i, j, k, n));
            });
    }
}

// max_wave_speed  Euler.cpp:46-74
static double max_wave_speed(const Level &L)
{
    // identity of ReduceOpMax over non-negative values; a maximum does not
    // depend on the order it is taken in, so boxes reduce independently
    const double        a = L.parm.adiabatic, e = L.parm.epsilon;
    const int           D = L.D;
    std::vector<double> wb(L.valid.size(), 0.0);
#pragma omp parallel for schedule(dynamic)
    for (long bl = 0; bl < (long)L.valid.size(); ++bl)
    {
        const size_t b  = (size_t)bl;
        double       ws = 0;
        for_box(L.valid[b], [&](int i, int j, int k) {
            const Vec    q   = primv_from_consv(L.fabs[b], a, e, i, j, k, D);
            double       vsq = q[1] * q[1];
            for (int d = 1; d < D; ++d)
                vsq = vsq + q[1 + d] * q[1 + d];
            const double velocity = std::sqrt(vsq);
            const double s = std::sqrt(a * q[1 + D] / (q[0] * e)) + velocity;
            ws             = std::max(ws, s);
        });
        wb[b] = ws;
    }
    double ws = 0;
    for (double w : wb)
        ws = std::max(ws, w);
    return ws;
}

// estTimeStep (AmrLevelAdv.cpp:395-477) + computeNewDt clamp (:589-618)
static double compute_dt(const Level &L, const double dx[3], double cfl,
                         double cur_time, double stop_time)
{
    const double max_speed = max_wave_speed(L);
    double       dt_est    = 1.0e20;
    for (int d = 0; d < L.D; ++d)
        dt_est = std::min(dt_est, dx[d] / max_speed);
    dt_est *= cfl;
    double       dt_0 = std::min(1.0e+100, 1 * dt_est);
    const double eps  = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((cur_time + dt_0) > (stop_time - eps))
            dt_0 = stop_time - cur_time;
    return dt_0 / 1;
}

static void advance_explicit(Level &L, double dt, const double dx[3],
                             int num_method, const FluxHook *hook = nullptr)
{
    std::vector<Fab> tmp = L.fabs; // Sborder2 = copy of Sborder (:121)
    std::vector<Fab> *in = &L.fabs, *out = &tmp;
    for (int d = 0; d < L.D; ++d)
    {
        sweep(L, *in, *out, d, dt, dx, num_method, hook);
        std::swap(L.fabs, *out); // fill boundary needs the level view
        fill_consv_boundary(L);
        std::swap(L.fabs, *out);
        std::swap(in, out);
    }
    if (in != &L.fabs)
        L.fabs = *in;
}


// ============================================================ two levels
// Refinement ratio 2, fine level subcycled (Amr::timeStep), restated from
// AMReX's documented semantics (AMReX is not available here):
//   FillPatchTwoLevels: fine ghost cells not covered by fine boxes and inside
//     the (periodic) domain get coarse data interpolated linearly in time
//     (copy at t_old; else alpha*old + beta*new) and conservatively in space
//     (cell_cons_interp with MC slopes limited to the neighbourhood range);
//     face ghost cells only (the sweeps read no others);
//   FluxRegister: fluxes scaled by dt * face area (_advance.cpp:192-207);
//     CrseInit(-1) sets -F_c, FineAdd(+1) adds the fine faces of each coarse
//     face (transverse children x fastest), one FineAdd per fine step;
//   Reflux (scale 1): coarse cell outside the fine region += sign*reg/vol,
//     sign -1 on the fine boxes' low faces, +1 on their high faces, in the
//     orientation order x lo, y lo, z lo, x hi, y hi, z hi;
//   avgDown: covered coarse cell = volfrac * sum of children (x fastest);
//   computeNewDt over both levels (AmrLevelAdv.cpp:589-618).
// Parity against AMReX of the interpolation formula and of these summation
// orders is unpinned; the tests pin conservation, free streams, linear
// exactness and bounds of the interpolation, and the sign of the reflux.
static int box_of(const Level &L, const int x[3])
{
    for (size_t s = 0; s < L.valid.size(); ++s)
        if (inside(L.valid[s], x[0], x[1], x[2]))
            return (int)s;
    return -1;
}

static bool wrap_level(const Level &L, int x[3])
{
    for (int d = 0; d < L.D; ++d)
    {
        const int len = L.dhi[d] - L.dlo[d] + 1;
        if (x[d] < L.dlo[d] || x[d] > L.dhi[d])
        {
            if (L.bclo[d] != 1)
                return false;
            x[d] += x[d] < L.dlo[d] ? len : -len;
        }
    }
    return true;
}

struct AmrSlot
{
    int d, orient, sign;
    int cbox, cface[3];     // coarse face: low face of cell cface in cbox
    int fbox, fface[4][3];  // fine faces (low faces of these cells) in fbox
    int tbox, tcell[3];     // coarse cell receiving the correction
};

// Register faces: for every fine box, direction and side, each coarse face
// of that side whose outside neighbour (periodically wrapped) is in the
// domain and not refined.
static std::vector<AmrSlot> amr_slots(const Level &C, const Level &F)
{
    const int            D = C.D;
    std::vector<AmrSlot> out;
    for (int side = 0; side < 2; ++side)
        for (int d = 0; d < D; ++d)
            for (size_t fb = 0; fb < F.valid.size(); ++fb)
            {
                const Box &v = F.valid[fb];
                Box        cv;
                for (int e = 0; e < 3; ++e)
                {
                    cv.lo[e] = e < D ? v.lo[e] / 2 : 0;
                    cv.hi[e] = e < D ? v.hi[e] / 2 : 0;
                }
                const int fine_face = side == 0 ? v.lo[d] : v.hi[d] + 1;
                const int crse_face = fine_face / 2;
                cv.lo[d] = cv.hi[d] = 0;
                for_box(cv, [&](int i, int j, int k) {
                    AmrSlot s;
                    int     t[3] = { i, j, k };
                    int     o[3] = { i, j, k };
                    o[d]         = side == 0 ? crse_face - 1 : crse_face;
                    int child[3] = { 2 * o[0], D > 1 ? 2 * o[1] : 0, D > 2 ? 2 * o[2] : 0 };
                    if (!wrap_level(F, child) || box_of(F, child) >= 0)
                        return;
                    int ow[3] = { o[0], o[1], o[2] };
                    wrap_level(C, ow);
                    int in[3] = { t[0], t[1], t[2] };
                    in[d]     = side == 0 ? crse_face : crse_face - 1;
                    s.d       = d;
                    s.orient  = side * 3 + d;
                    s.sign    = side == 0 ? -1 : 1;
                    s.cbox    = box_of(C, in);
                    s.tbox    = box_of(C, ow);
                    if (s.cbox < 0 || s.tbox < 0)
                        throw std::runtime_error("coarse level does not cover the domain");
                    for (int e = 0; e < 3; ++e)
                    {
                        s.cface[e] = t[e];
                        s.tcell[e] = ow[e];
                    }
                    s.cface[d] = crse_face;
                    s.fbox     = (int)fb;
                    int q      = 0;
                    // transverse children, x fastest
                    for (int c2 = 0; c2 < (D > 2 && d != 2 ? 2 : 1); ++c2)
                        for (int c1 = 0; c1 < (D > 1 && d != 1 ? 2 : 1); ++c1)
                            for (int c0 = 0; c0 < (d != 0 ? 2 : 1); ++c0)
                            {
                                s.fface[q][0] = d == 0 ? fine_face : 2 * t[0] + c0;
                                s.fface[q][1] = D < 2 ? 0 : (d == 1 ? fine_face : 2 * t[1] + c1);
                                s.fface[q][2] = D < 3 ? 0 : (d == 2 ? fine_face : 2 * t[2] + c2);
                                ++q;
                            }
                    out.push_back(s);
                });
            }
    return out;
}

static double mc_slope(double um, double uc, double up)
{
    const double dc = 0.5 * (up - um);
    const double dl = 2 * (uc - um);
    const double dr = 2 * (up - uc);
    if (!(dl * dr > 0))
        return 0.0;
    const double m = std::fmin(std::fabs(dc), std::fmin(std::fabs(dl), std::fabs(dr)));
    return dc < 0 ? -m : m;
}

// Coarse-fine face ghost cells of the fine level at a time between the two
// coarse states (mode 0: Cold; mode 1: alpha*Cold + beta*Cnew).
static void cf_fill(Level &F, const Level &Cold, const Level &Cnew, int mode,
                    double alpha, double beta)
{
    const int D  = F.D;
    const int NC = D + 2;
    for (size_t b = 0; b < F.fabs.size(); ++b)
    {
        const Box &v = F.valid[b];
        for_box(F.fabs[b].bx, [&](int i, int j, int k) {
            int x[3] = { i, j, k }, nout = 0;
            for (int d = 0; d < D; ++d)
                nout += (x[d] < v.lo[d] || x[d] > v.hi[d]);
            if (nout != 1)
                return;
            int img[3] = { i, j, k };
            if (!wrap_level(F, img) || box_of(F, img) >= 0)
                return;
            int    ci[3] = { 0, 0, 0 };
            double off[3] = { 0, 0, 0 };
            for (int d = 0; d < D; ++d)
            {
                ci[d]  = img[d] >> 1;
                off[d] = (img[d] & 1) ? 0.25 : -0.25;
            }
            const int cb = box_of(Cold, ci);
            for (int m = 0; m < NC; ++m)
            {
                auto val = [&](int a0, int a1, int a2) {
                    const double o = Cold.fabs[cb](ci[0] + a0, ci[1] + a1, ci[2] + a2, m);
                    if (mode == 0)
                        return o;
                    const double nw = Cnew.fabs[cb](ci[0] + a0, ci[1] + a1, ci[2] + a2, m);
                    return alpha * o + beta * nw;
                };
                const double uc   = val(0, 0, 0);
                double       s[3] = { 0, 0, 0 };
                s[0]              = mc_slope(val(-1, 0, 0), uc, val(1, 0, 0));
                if (D > 1)
                    s[1] = mc_slope(val(0, -1, 0), uc, val(0, 1, 0));
                if (D > 2)
                    s[2] = mc_slope(val(0, 0, -1), uc, val(0, 0, 1));
                double umin = uc, umax = uc;
                for (int kk = (D > 2 ? -1 : 0); kk <= (D > 2 ? 1 : 0); ++kk)
                    for (int jj = (D > 1 ? -1 : 0); jj <= (D > 1 ? 1 : 0); ++jj)
                        for (int ii = -1; ii <= 1; ++ii)
                        {
                            const double w = val(ii, jj, kk);
                            umin           = std::fmin(umin, w);
                            umax           = std::fmax(umax, w);
                        }
                double ext = 0.0;
                for (int d = 0; d < D; ++d)
                    ext = ext + 0.25 * std::fabs(s[d]);
                double lim = 1.0;
                if (ext > 0)
                    lim = std::fmin(1.0, std::fmin((umax - uc) / ext, (uc - umin) / ext));
                double r = uc;
                for (int d = 0; d < D; ++d)
                    r = r + off[d] * (lim * s[d]);
                F.fabs[b](i, j, k, m) = r;
            }
        });
    }
}

static void avg_down(Level &C, const Level &F)
{
    const int    D       = C.D;
    const double volfrac = 1.0 / (double)(1 << D);
    for (size_t fb = 0; fb < F.valid.size(); ++fb)
    {
        const Box &v = F.valid[fb];
        Box        cv;
        for (int e = 0; e < 3; ++e)
        {
            cv.lo[e] = e < D ? v.lo[e] / 2 : 0;
            cv.hi[e] = e < D ? v.hi[e] / 2 : 0;
        }
        for_box(cv, [&](int i, int j, int k) {
            int       c[3] = { i, j, k };
            const int cb   = box_of(C, c);
            for (int m = 0; m < D + 2; ++m)
            {
                double sum = 0;
                for (int kr = 0; kr < (D > 2 ? 2 : 1); ++kr)
                    for (int jr = 0; jr < (D > 1 ? 2 : 1); ++jr)
                        for (int ir = 0; ir < 2; ++ir)
                            sum = sum + F.fabs[fb](2 * i + ir, D > 1 ? 2 * j + jr : 0,
                                                   D > 2 ? 2 * k + kr : 0, m);
                C.fabs[cb](i, j, k, m) = volfrac * sum;
            }
        });
    }
}

static double level_dt_est(const Level &L, const double dx[3], double cfl)
{
    const double max_speed = max_wave_speed(L);
    double       dt_est    = 1.0e20;
    for (int d = 0; d < L.D; ++d)
        dt_est = std::min(dt_est, dx[d] / max_speed);
    return dt_est * cfl;
}

// computeNewDt over levels: min of 2^l dt_est(l), then the stop clamp.
static double amr_dt(const std::vector<Level> &L, const double dx0[3], double cfl,
                     double cur_time, double stop_time)
{
    double dt_0     = 1.0e+100;
    int    n_factor = 1;
    for (size_t l = 0; l < L.size(); ++l)
    {
        double dxl[3];
        for (int d = 0; d < 3; ++d)
            dxl[d] = dx0[d] / (double)(1 << l);
        n_factor *= (l == 0 ? 1 : 2);
        dt_0 = std::min(dt_0, n_factor * level_dt_est(L[l], dxl, cfl));
    }
    const double eps = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((cur_time + dt_0) > (stop_time - eps))
            dt_0 = stop_time - cur_time;
    return dt_0;
}

struct AmrH
{
    std::vector<Level>                L, old;
    std::vector<std::vector<AmrSlot>> slots; // slots[l]: levels l and l+1
    std::vector<std::vector<double>>  reg;
    int                               num_method = 0;
    bool                              do_reflux  = true;
};

static void reflux_level(Level &C, const std::vector<AmrSlot> &slots,
                         const std::vector<double> &reg, const double dxc[3])
{
    const int NC  = C.D + 2;
    double    vol = dxc[0];
    for (int d = 1; d < C.D; ++d)
        vol *= dxc[d];
    for (int orient = 0; orient < 6; ++orient)
        for (size_t q = 0; q < slots.size(); ++q)
        {
            const AmrSlot &s = slots[q];
            if (s.orient != orient)
                continue;
            for (int m = 0; m < NC; ++m)
            {
                double &u = C.fabs[s.tbox](s.tcell[0], s.tcell[1], s.tcell[2], m);
                u         = u + (s.sign * reg[q * NC + m]) / vol;
            }
        }
}

// Amr::timeStep(l): advance level l, subcycle the finer levels twice, then
// reflux and avgDown (post_timestep).  t_par, dt_par: the parent's step.
static void amr_time_step(AmrH &h, int l, double time, double dt, double t_par,
                          double dt_par, const double dx0[3])
{
    const int nlev = (int)h.L.size();
    const int D    = h.L[0].D;
    const int NC   = D + 2;
    double    dxl[3];
    for (int d = 0; d < 3; ++d)
        dxl[d] = dx0[d] / (double)(1 << l);
    if (l > 0)
    {
        fill_consv_boundary(h.L[l]);
        const double t_new = t_par + dt_par;
        if (time == t_par)
            cf_fill(h.L[l], h.old[l - 1], h.L[l - 1], 0, 1.0, 0.0);
        else
            cf_fill(h.L[l], h.old[l - 1], h.L[l - 1], 1, (t_new - time) / (t_new - t_par),
                    (time - t_par) / (t_new - t_par));
    }
    if (l + 1 < nlev)
        h.old[l] = h.L[l];
    auto scale_of = [&](int d, double hstep, const double *dxx) {
        double sc = hstep;
        for (int e = 0; e < D; ++e)
            if (e != d)
                sc *= dxx[e];
        return sc;
    };
    FluxHook hook = [&](int box, int dir, const Fab &flux) {
        const double sc = scale_of(dir, dt, dxl);
        if (l > 0) // fine side of the pair below: FineAdd(+1)
        {
            const std::vector<AmrSlot> &sl   = h.slots[l - 1];
            std::vector<double>        &reg  = h.reg[l - 1];
            const int                   nsub = 1 << (D - 1);
            for (size_t q = 0; q < sl.size(); ++q)
                if (sl[q].d == dir && sl[q].fbox == box)
                    for (int c = 0; c < nsub; ++c)
                        for (int m = 0; m < NC; ++m)
                            reg[q * NC + m]
                                = reg[q * NC + m]
                                  + flux(sl[q].fface[c][0], sl[q].fface[c][1],
                                         sl[q].fface[c][2], m)
                                        * sc;
        }
        if (l + 1 < nlev) // coarse side of the pair above: CrseInit(-1)
        {
            const std::vector<AmrSlot> &sl  = h.slots[l];
            std::vector<double>        &reg = h.reg[l];
            for (size_t q = 0; q < sl.size(); ++q)
                if (sl[q].d == dir && sl[q].cbox == box)
                    for (int m = 0; m < NC; ++m)
                        reg[q * NC + m]
                            = -1.0
                              * (flux(sl[q].cface[0], sl[q].cface[1], sl[q].cface[2], m) * sc);
        }
    };
    advance_explicit(h.L[l], dt, dxl, h.num_method, &hook);
    if (l + 1 < nlev)
    {
        const double dtf = dt / 2;
        for (int sub = 0; sub < 2; ++sub)
            amr_time_step(h, l + 1, sub == 0 ? time : time + dtf, dtf, time, dt, dx0);
        if (h.do_reflux)
            reflux_level(h.L[l], h.slots[l], h.reg[l], dxl);
        avg_down(h.L[l], h.L[l + 1]);
        fill_consv_boundary(h.L[l]);
    }
}
} // namespace oracle

// ===================================================================== C ABI
using namespace oracle;

extern "C" {

typedef struct
{
    int    dim;
    int    num_method;
    double adiabatic;
    double epsilon;
    int    limiter;
    int    limiter_indices[5];
    int    limit_on_primv[5];
} or_params;

static Parm make_parm(const or_params *p)
{
    Parm m;
    m.D                           = p->dim;
    m.adiabatic                   = p->adiabatic;
    m.epsilon                     = p->epsilon;
    m.limiter                     = p->limiter;
    m.need_primitive_for_limiting = false;
    for (int n = 0; n < 5; ++n)
    {
        m.limiter_indices[n] = p->limiter_indices[n];
        m.limit_on_primv[n]  = p->limit_on_primv[n] != 0;
        if (n < p->dim + 2)
            m.need_primitive_for_limiting
                = m.need_primitive_for_limiting || m.limit_on_primv[n];
    }
    return m;
}

// Level from the packed layout shared with the GPU library:
// boxes: nbox x 6 ints (lo0 lo1 lo2 hi0 hi1 hi2), offsets: FAB starts.
static void make_level(Level &L, const or_params *p, const int *boxes,
                       int nbox, int ng, const int *dlo, const int *dhi,
                       const int *bclo, const int *bchi)
{
    L.parm = make_parm(p);
    L.D    = p->dim;
    L.ng   = ng;
    for (int d = 0; d < 3; ++d)
    {
        L.dlo[d]  = d < L.D ? dlo[d] : 0;
        L.dhi[d]  = d < L.D ? dhi[d] : 0;
        L.bclo[d] = bclo[d];
        L.bchi[d] = bchi[d];
    }
    L.valid.resize(nbox);
    L.fabs.resize(nbox);
    for (int b = 0; b < nbox; ++b)
    {
        Box v;
        for (int d = 0; d < 3; ++d)
        {
            v.lo[d] = d < L.D ? boxes[6 * b + d] : 0;
            v.hi[d] = d < L.D ? boxes[6 * b + 3 + d] : 0;
        }
        L.valid[b] = v;
        L.fabs[b].define(grow(v, L.D, ng), L.D + 2);
    }
}

static void level_in(Level &L, const double *U, const int64_t *offsets)
{
    for (size_t b = 0; b < L.fabs.size(); ++b)
        std::memcpy(L.fabs[b].a.data(), U + offsets[b],
                    sizeof(double) * L.fabs[b].a.size());
}

static void level_out(const Level &L, double *U, const int64_t *offsets)
{
    for (size_t b = 0; b < L.fabs.size(); ++b)
        std::memcpy(U + offsets[b], L.fabs[b].a.data(),
                    sizeof(double) * L.fabs[b].a.size());
}

// The main.cpp loop on a single level.  U: packed level (in/out).
// Wall time of the step loop of the last or_run on this thread (the time
// steps, without building the level and copying the state in and out).
static thread_local double g_run_loop_seconds = 0.0;
double or_last_run_seconds(void) { return g_run_loop_seconds; }

int or_run(const or_params *p, const int *boxes, int nbox, int ng,
           const int *dlo, const int *dhi, const int *bclo, const int *bchi,
           const double *dx, double cfl, double start_time, double stop_time,
           int max_step, double *U, const int64_t *offsets, double *time_out,
           int *steps_out)
{
    Level L;
    make_level(L, p, boxes, nbox, ng, dlo, dhi, bclo, bchi);
    level_in(L, U, offsets);
    fill_consv_boundary(L);
    double     time  = start_time;
    int        steps = 0;
    const auto t0    = std::chrono::steady_clock::now();
    while ((steps < max_step || max_step < 0)
           && (time < stop_time || stop_time < 0.0))
    {
        const double dt = compute_dt(L, dx, cfl, time, stop_time);
        advance_explicit(L, dt, dx, p->num_method);
        time = time + dt;
        ++steps;
    }
    g_run_loop_seconds
        = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    level_out(L, U, offsets);
    if (time_out)
        *time_out = time;
    if (steps_out)
        *steps_out = steps;
    return 0;
}

// One sweep of direction dir: Uout valid cells <- sweep(Uin).  Ghosts of Uin
// must be filled; Uout ghost cells are left unchanged.
int or_sweep(const or_params *p, const int *boxes, int nbox, int ng,
             const int *dlo, const int *dhi, const int *bclo, const int *bchi,
             int dir, double dt, const double *dx, const double *Uin,
             double *Uout, const int64_t *offsets)
{
    Level L;
    make_level(L, p, boxes, nbox, ng, dlo, dhi, bclo, bchi);
    level_in(L, Uin, offsets);
    std::vector<Fab> out(L.fabs.size());
    for (size_t b = 0; b < out.size(); ++b)
    {
        out[b] = L.fabs[b];
        std::memcpy(out[b].a.data(), Uout + offsets[b],
                    sizeof(double) * out[b].a.size());
    }
    sweep(L, L.fabs, out, dir, dt, dx, p->num_method);
    L.fabs = out;
    level_out(L, Uout, offsets);
    return 0;
}

int or_fill_boundary(const or_params *p, const int *boxes, int nbox, int ng,
                     const int *dlo, const int *dhi, const int *bclo,
                     const int *bchi, double *U, const int64_t *offsets)
{
    Level L;
    make_level(L, p, boxes, nbox, ng, dlo, dhi, bclo, bchi);
    level_in(L, U, offsets);
    fill_consv_boundary(L);
    level_out(L, U, offsets);
    return 0;
}

int or_max_wave_speed(const or_params *p, const int *boxes, int nbox, int ng,
                      const double *U, const int64_t *offsets, double *out)
{
    int   z[3] = { 0, 0, 0 };
    Level L;
    make_level(L, p, boxes, nbox, ng, z, z, z, z);
    level_in(L, U, offsets);
    *out = max_wave_speed(L);
    return 0;
}

int or_est_dt(const or_params *p, const int *boxes, int nbox, int ng,
              const double *U, const int64_t *offsets, const double *dx,
              double cfl, double time, double stop_time, double *out)
{
    int   z[3] = { 0, 0, 0 };
    Level L;
    make_level(L, p, boxes, nbox, ng, z, z, z, z);
    level_in(L, U, offsets);
    *out = compute_dt(L, dx, cfl, time, stop_time);
    return 0;
}

// Face fluxes of one box on surroundingNodes(bx, dir) (Array4 without ghosts,
// NC components), as compute_MUSCL_hancock_flux / compute_HLLC_flux write them.
int or_flux_box(const or_params *p, const int *box, int ng, int dir, double dt,
                const double *dx, const double *Ufab, double *flux)
{
    const int D = p->dim;
    Box       v;
    for (int d = 0; d < 3; ++d)
    {
        v.lo[d] = d < D ? box[d] : 0;
        v.hi[d] = d < D ? box[3 + d] : 0;
    }
    Parm parm = make_parm(p);
    Fab  U;
    U.define(grow(v, D, ng), D + 2);
    std::memcpy(U.a.data(), Ufab, sizeof(double) * U.a.size());
    Fab F;
    F.define(surrounding_nodes(v, dir), D + 2);
    compute_flux(p->num_method, v, F, U, dx, dt, parm, dir);
    std::memcpy(flux, F.a.data(), sizeof(double) * F.a.size());
    return 0;
}

// ---- scalar entry points for known-answer tests (2-cell FABs as in
// Tests/hllc_test.H; dir 0, face at i = 1)
int or_wave_speed_toro(int dim, const double *primv_L, const double *primv_R,
                       double adiabatic, double epsilon, double *out3)
{
    Box b{};
    b.hi[0] = 1;
    Fab L, R;
    L.define(b, dim + 2);
    R.define(b, dim + 2);
    Vec qL{}, qR{};
    for (int n = 0; n < dim + 2; ++n)
    {
        qL[n] = primv_L[n];
        qR[n] = primv_R[n];
    }
    const Vec cL = consv_from_primv(qL, adiabatic, epsilon, dim);
    const Vec cR = consv_from_primv(qR, adiabatic, epsilon, dim);
    for (int i = 0; i < 2; ++i)
        for (int n = 0; n < dim + 2; ++n)
        {
            L(i, 0, 0, n) = cL[n];
            R(i, 0, 0, n) = cR[n];
        }
    estimate_wave_speed_HLLC_Toro(L, R, 1, 0, 0, adiabatic, adiabatic, epsilon,
                                  0, dim, out3[0], out3[1], out3[2]);
    return 0;
}

int or_consv_from_primv(int dim, const double *q, double adiabatic,
                        double epsilon, double *u)
{
    Vec qq{};
    for (int n = 0; n < dim + 2; ++n)
        qq[n] = q[n];
    const Vec c = consv_from_primv(qq, adiabatic, epsilon, dim);
    for (int n = 0; n < dim + 2; ++n)
        u[n] = c[n];
    return 0;
}

int or_primv_from_consv(int dim, const double *u, double adiabatic,
                        double epsilon, double *q)
{
    Box b{};
    Fab U;
    U.define(b, dim + 2);
    for (int n = 0; n < dim + 2; ++n)
        U(0, 0, 0, n) = u[n];
    const Vec r = primv_from_consv(U, adiabatic, epsilon, 0, 0, 0, dim);
    for (int n = 0; n < dim + 2; ++n)
        q[n] = r[n];
    return 0;
}

// HLLC face flux between conserved states UL (cell i-1) and UR (cell i) in
// direction dir, through compute_HLLC_flux_LR on two-cell FABs
int or_hllc_face(int dim, int dir, const double *UL, const double *UR,
                 double adiabatic, double epsilon, double *flux)
{
    Box b{};
    b.hi[dir] = 1;
    Fab L, R;
    L.define(b, dim + 2);
    R.define(b, dim + 2);
    const int io = dir == 0, jo = dir == 1, ko = dir == 2;
    for (int n = 0; n < dim + 2; ++n)
    {
        L(0, 0, 0, n)    = UL[n];
        L(io, jo, ko, n) = UL[n];
        R(0, 0, 0, n)    = UR[n];
        R(io, jo, ko, n) = UR[n];
    }
    Parm pm{};
    pm.D         = dim;
    pm.adiabatic = adiabatic;
    pm.epsilon   = epsilon;
    Box cellbox{}; // zero-width cell box whose surroundingNodes is face 1
    cellbox.lo[dir] = 1;
    cellbox.hi[dir] = 0;
    Fab F;
    F.define(surrounding_nodes(cellbox, dir), dim + 2);
    compute_HLLC_flux_LR(cellbox, F, L, R, pm, dir);
    for (int n = 0; n < dim + 2; ++n)
        flux[n] = F(io, jo, ko, n);
    return 0;
}

double or_limiter(int kind, double r) { return slope_limit(kind, r); }
double or_minmod(double a, double b) { return minmod(a, b); }
double or_pow_fn(double x, double y) { return or_pow(x, y); }
double or_pow_negz_fn(double x, double gamma)
{
    return or_pow_negz(x, -((gamma - 1) / (2 * gamma)), gamma);
}
double or_pow_invz_fn(double x, double gamma)
{
    return or_pow_invz(x, 1 / ((gamma - 1) / (2 * gamma)), gamma);
}

// nlev levels (ratio 2): the main.cpp loop with Amr::timeStep subcycling,
// post_init avgDown first.  Level l: boxes[l] (nbox[l] x 6 ints, level-l
// index space), packed state U[l] with FAB offsets offsets[l]; level 0's
// domain dlo..dhi, level l's refined by 2^l, same BCs.
int or_amr_run_levels(const or_params *p, int nlev, const int *const *boxes,
                      const int *nbox, int ng, const int *dlo, const int *dhi,
                      const int *bclo, const int *bchi, const double *dx, double cfl,
                      double start_time, double stop_time, int max_step, double *const *U,
                      const int64_t *const *offsets, int do_reflux, double *time_out,
                      int *steps_out)
{
    AmrH h;
    h.L.resize(nlev);
    h.old.resize(nlev);
    h.num_method = p->num_method;
    h.do_reflux  = do_reflux != 0;
    for (int l = 0; l < nlev; ++l)
    {
        int ldlo[3], ldhi[3];
        for (int d = 0; d < 3; ++d)
        {
            ldlo[d] = dlo[d] << l;
            ldhi[d] = ((dhi[d] + 1) << l) - 1;
        }
        make_level(h.L[l], p, boxes[l], nbox[l], ng, ldlo, ldhi, bclo, bchi);
        level_in(h.L[l], U[l], offsets[l]);
    }
    // proper nesting (every coarse cell within 2 of a coarsened fine box is
    // a valid cell of the coarser level, periodic images included)
    for (int l = 0; l + 1 < nlev; ++l)
        for (const Box &v : h.L[l + 1].valid)
        {
            Box g;
            for (int d = 0; d < 3; ++d)
            {
                g.lo[d] = d < p->dim ? (v.lo[d] >> 1) - 2 : 0;
                g.hi[d] = d < p->dim ? (v.hi[d] >> 1) + 2 : 0;
            }
            bool ok = true;
            for_box(g, [&](int i, int j, int k) {
                int c[3] = { i, j, k };
                if (wrap_level(h.L[l], c) && box_of(h.L[l], c) < 0)
                    ok = false;
            });
            if (!ok)
                return 1;
        }
    for (int l = 0; l + 1 < nlev; ++l)
    {
        h.slots.push_back(amr_slots(h.L[l], h.L[l + 1]));
        h.reg.push_back(std::vector<double>(h.slots.back().size() * (p->dim + 2), 0.0));
    }
    for (int l = nlev - 2; l >= 0; --l)
        avg_down(h.L[l], h.L[l + 1]);
    for (int l = 0; l < nlev; ++l)
        fill_consv_boundary(h.L[l]);
    double time  = start_time;
    int    steps = 0;
    while ((steps < max_step || max_step < 0)
           && (time < stop_time || stop_time < 0.0))
    {
        const double dt = amr_dt(h.L, dx, cfl, time, stop_time);
        amr_time_step(h, 0, time, dt, time, dt, dx);
        time = time + dt;
        ++steps;
    }
    for (int l = 0; l < nlev; ++l)
        level_out(h.L[l], U[l], offsets[l]);
    if (time_out)
        *time_out = time;
    if (steps_out)
        *steps_out = steps;
    return 0;
}

int or_amr_run(const or_params *p, const int *cboxes, int ncbox, const int *fboxes,
               int nfbox, int ng, const int *dlo, const int *dhi, const int *bclo,
               const int *bchi, const double *dx, double cfl, double start_time,
               double stop_time, int max_step, double *Uc, const int64_t *coffsets,
               double *Uf, const int64_t *foffsets, int do_reflux, double *time_out,
               int *steps_out)
{
    const int     *b[2] = { cboxes, fboxes };
    const int      n[2] = { ncbox, nfbox };
    double        *U[2] = { Uc, Uf };
    const int64_t *o[2] = { coffsets, foffsets };
    return or_amr_run_levels(p, 2, b, n, ng, dlo, dhi, bclo, bchi, dx, cfl, start_time,
                             stop_time, max_step, U, o, do_reflux, time_out, steps_out);
}

// The fine level's ghost cells at `time` between coarse states at t_old and
// t_new: fine-fine and physical (fill_consv_boundary), then coarse-fine.
int or_amr_fill_fine(const or_params *p, const int *cboxes, int ncbox, const int *fboxes,
                     int nfbox, int ng, const int *dlo, const int *dhi, const int *bclo,
                     const int *bchi, const double *Uc_old, const double *Uc_new,
                     const int64_t *coffsets, double *Uf, const int64_t *foffsets,
                     double t_old, double t_new, double time)
{
    Level Co, Cn, F;
    int   fdlo[3], fdhi[3];
    for (int d = 0; d < 3; ++d)
    {
        fdlo[d] = 2 * dlo[d];
        fdhi[d] = 2 * dhi[d] + 1;
    }
    make_level(Co, p, cboxes, ncbox, ng, dlo, dhi, bclo, bchi);
    make_level(Cn, p, cboxes, ncbox, ng, dlo, dhi, bclo, bchi);
    make_level(F, p, fboxes, nfbox, ng, fdlo, fdhi, bclo, bchi);
    level_in(Co, Uc_old, coffsets);
    level_in(Cn, Uc_new, coffsets);
    level_in(F, Uf, foffsets);
    fill_consv_boundary(F);
    if (time == t_old || !(t_new != t_old))
        cf_fill(F, Co, Cn, 0, 1.0, 0.0);
    else
        cf_fill(F, Co, Cn, 1, (t_new - time) / (t_new - t_old), (time - t_old) / (t_new - t_old));
    level_out(F, Uf, foffsets);
    return 0;
}

// Threads of the box loop (omp_set_num_threads); n <= 0 leaves the count
// unchanged.  Returns the count the next parallel region uses.
int or_threads(int n)
{
    if (n > 0)
        omp_set_num_threads(n);
    return omp_get_max_threads();
}

} // extern "C"
