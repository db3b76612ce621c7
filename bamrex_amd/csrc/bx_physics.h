// Per-cell Euler physics for the fused MUSCL-Hancock / HLLC sweep kernels.
//
// Every expression keeps the operand order of the reference so that, with
// floating-point contraction disabled (-ffp-contract=off), each result is
// bit-identical to the CPU evaluation:
//   primv_from_consv   Source/System/Euler/Euler.H:167-198
//   euler_flux         Source/System/Euler/Euler.H:211-244 (= Euler.cpp:18-43)
//   limiters           Source/Fluxes/Limiters.H:42-73
//   reconstruction     Source/Fluxes/Flux_Helpers.cpp:92-148
//   Toro wave speeds   Source/Fluxes/HLLC.H:213-309
//   HLLC star state    Source/Fluxes/HLLC.H:311-337
//   HLLC flux          Source/Fluxes/HLLC.cpp:90-172
//   max wave speed     Source/System/Euler/Euler.cpp:53-68
//   FORCE / LF flux    Source/Fluxes/Force.cpp:24-104, 106-146
// State vectors are conserved variables U = (rho, m_0..m_{D-1}, E), NC = D+2.
#pragma once
#include "bx_pow.h"

#define BX_DEV __device__ __forceinline__

namespace bx {

enum LimiterKind
{
    LIM_MINBEE     = 0,
    LIM_VAN_LEER   = 1,
    LIM_VAN_ALBADA = 2,
    LIM_SUPERBEE   = 3
};

struct EulerParm
{
    double    gamma;
    double    eps;
    BxPowHllc ph; // exponent data of the HLLC pressure estimate
};

// Q = (rho, u_0..u_{D-1}, p)
// Multiplication / division by epsilon.  With E1 (epsilon == 1, the only
// value the explicit configs use) they are dropped: x*1 and x/1 are exact,
// so the result is unchanged bit for bit.
template <bool E1> BX_DEV double mul_eps(double x, double eps)
{
    return E1 ? x : x * eps;
}
template <bool E1> BX_DEV double div_eps(double x, double eps)
{
    return E1 ? x : x / eps;
}

// a / b for several numerators sharing b: with y = RN(1/b) (one IEEE
// division) and q0 = RN(a y), the Markstein correction q0 + (a - b q0) y,
// rounded once, is exactly RN(a / b) whenever nothing under- or overflows
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990) -- the same bits as a / b.
BX_DEV double div_by_recip(double a, double b, double y)
{
    const double q0 = a * y;
    const double e  = __builtin_fma(-q0, b, a);
    return __builtin_fma(e, y, q0);
}

// sqrt(x), correctly rounded: the bits of IEEE sqrt.  For x in
// [2^-767, 2^1024) -- every positive finite x except the tiny ones -- the
// coupled Newton iteration from the hardware reciprocal square root
// (g ~ sqrt(x), h ~ 1/(2 sqrt(x)), two final corrections g += (x - g^2) h;
// Markstein, "Software division and square root using Goldschmidt's
// algorithms", 2004) needs neither the rescaling of tiny arguments nor the
// fix-up of zero and infinity that the library sqrt adds around the same
// iteration; every other x takes the library sqrt.  Z0: x == 0 is common
// (a fluid at rest), so it joins the fast path and returns x.
template <bool Z0 = false>
BX_DEV double sqrt_rn(double x)
{
    const unsigned hi = (unsigned)((unsigned long long)__double_as_longlong(x) >> 32);
    const bool     in = hi - 0x10000000u < 0x6ff00000u;
    if (__builtin_expect(Z0 ? (in || x == 0.0) : in, 1))
    {
        const double y = __builtin_amdgcn_rsq(x);
        double       g = x * y;
        double       h = y * 0.5;
        const double r = __builtin_fma(-h, g, 0.5);
        g              = __builtin_fma(g, r, g);
        h              = __builtin_fma(h, r, h);
        double d       = __builtin_fma(-g, g, x);
        g              = __builtin_fma(d, h, g);
        d              = __builtin_fma(-g, g, x);
        g              = __builtin_fma(d, h, g);
        return (Z0 && !in) ? x : g;
    }
    return sqrt(x);
}

// Also returns y = RN(1 / rho) for further quotients by rho.
template <int D, bool E1 = false>
BX_DEV void primv_from_consv(const double (&U)[D + 2], double (&Q)[D + 2],
                             const EulerParm &pm, double &y)
{
    Q[0] = U[0];
    y    = 1.0 / U[0];
#pragma unroll
    for (int d = 0; d < D; ++d)
        Q[1 + d] = div_by_recip(U[1 + d], U[0], y);
    double u_sq = Q[1] * Q[1];
#pragma unroll
    for (int d = 1; d < D; ++d)
        u_sq = u_sq + Q[1 + d] * Q[1 + d];
    // 0.5 * eps * rho: (0.5*eps)*rho == 0.5*rho exactly when eps == 1
    Q[1 + D] = (pm.gamma - 1) * (U[1 + D] - mul_eps<E1>(0.5, pm.eps) * U[0] * u_sq);
}

template <int D, bool E1 = false>
BX_DEV void primv_from_consv(const double (&U)[D + 2], double (&Q)[D + 2],
                             const EulerParm &pm)
{
    double y;
    primv_from_consv<D, E1>(U, Q, pm, y);
}

template <int D, int dir, bool E1 = false>
BX_DEV void euler_flux(const double (&U)[D + 2], const double (&Q)[D + 2],
                       double (&F)[D + 2], const EulerParm &pm)
{
    F[0] = U[1 + dir];
#pragma unroll
    for (int d = 0; d < D; ++d)
        F[1 + d] = U[1 + dir] * Q[1 + d];
    F[1 + dir] = F[1 + dir] + div_eps<E1>(Q[1 + D], pm.eps);
    F[1 + D]   = (U[1 + D] + Q[1 + D]) * Q[1 + dir];
}

template <int LIM>
BX_DEV double slope_limit(double r)
{
    if (LIM == LIM_MINBEE)
        return r <= 0 ? 0. : (r > 1 ? fmin(2. / (1 + r), 1.) : r);
    else if (LIM == LIM_VAN_LEER)
        return r <= 0 ? 0. : fmin(2. * r / (1 + r), 2. / (1 + r));
    else if (LIM == LIM_VAN_ALBADA)
        return r <= 0 ? 0. : fmin(r * (1 + r) / (1 + r * r), 2. / (1 + r));
    else
        return (r <= 0) ? 0.
                        : ((0 < r && r <= 0.5)
                               ? 2 * r
                               : ((0.5 < r && r <= 1)
                                      ? 1.
                                      : fmin(r, fmin(2. / (1 + r), 2.))));
}

template <int LIM>
BX_DEV double limiter_value(double vl, double vc, double vr)
{
    const double numerator = vc - vl;
    const double denom     = vr - vc;
    if (numerator == 0. && denom == 0.)
        return slope_limit<LIM>(1.);
    else if (denom == 0)
        return 0.;
    return slope_limit<LIM>(numerator / denom);
}

// Limiting-variable selection (CalcLimitingVal.H:9-22).  The host folds the
// runtime (limiter_indices, limit_on_primv) table into one of three modes:
//   LV_SHARED: every component limited on the conserved energy (the default
//              table, FluidSystem.cpp:67), fixed at compile time
//   LV_OWN:    component n limited on conserved variable n
//   LV_TABLE:  any other table (evaluated with selects)
enum LimVarMode
{
    LV_SHARED = 0,
    LV_OWN    = 1,
    LV_TABLE  = 2
};

struct LimTable
{
    int idx[5];
    int prim[5];
};

template <int D>
BX_DEV double pick_var(const double (&U)[D + 2], const double (&Q)[D + 2],
                       int idx, int prim)
{
    double v = 0.;
#pragma unroll
    for (int m = 0; m < D + 2; ++m)
        if (idx == m)
            v = prim ? Q[m] : U[m];
    return v;
}

// Reconstruct + half time step at one cell (Flux_Helpers.cpp:92-148).
// Um/U0/Up: conserved states at c-1, c, c+1 along dir.  Outputs the
// half-stepped low-side (hL) and high-side (hR) states of cell c.
template <int D, int dir, int LIM, int LV, bool E1 = false>
BX_DEV void reconstruct_half_step(const double (&Um)[D + 2],
                                  const double (&U0)[D + 2],
                                  const double (&Up)[D + 2],
                                  double (&hL)[D + 2], double (&hR)[D + 2],
                                  double hdtdx, const EulerParm &pm,
                                  const LimTable &lt)
{
    constexpr int NC = D + 2;
    double        phi[NC];
    if (LV == LV_SHARED)
    {
        const double p = limiter_value<LIM>(Um[NC - 1], U0[NC - 1], Up[NC - 1]);
#pragma unroll
        for (int n = 0; n < NC; ++n)
            phi[n] = p;
    }
    else if (LV == LV_OWN)
    {
#pragma unroll
        for (int n = 0; n < NC; ++n)
            phi[n] = limiter_value<LIM>(Um[n], U0[n], Up[n]);
    }
    else
    {
        double Qm[NC], Q0[NC], Qp[NC];
        primv_from_consv<D>(Um, Qm, pm);
        primv_from_consv<D>(U0, Q0, pm);
        primv_from_consv<D>(Up, Qp, pm);
#pragma unroll
        for (int n = 0; n < NC; ++n)
            phi[n] = limiter_value<LIM>(pick_var<D>(Um, Qm, lt.idx[n], lt.prim[n]),
                                        pick_var<D>(U0, Q0, lt.idx[n], lt.prim[n]),
                                        pick_var<D>(Up, Qp, lt.idx[n], lt.prim[n]));
    }

#pragma unroll
    for (int n = 0; n < NC; ++n)
    {
        const double delta_left  = U0[n] - Um[n];
        const double delta_right = Up[n] - U0[n];
        const double delta       = 0.5 * delta_left + 0.5 * delta_right;
        hL[n]                    = U0[n] - 0.5 * phi[n] * delta;
        hR[n]                    = U0[n] + 0.5 * phi[n] * delta;
    }
    double QL[NC], QR[NC], FL[NC], FR[NC];
    primv_from_consv<D, E1>(hL, QL, pm);
    primv_from_consv<D, E1>(hR, QR, pm);
    euler_flux<D, dir, E1>(hL, QL, FL, pm);
    euler_flux<D, dir, E1>(hR, QR, FR, pm);
#pragma unroll
    for (int n = 0; n < NC; ++n)
    {
        hL[n] = hL[n] - hdtdx * (FR[n] - FL[n]);
        hR[n] = hR[n] - hdtdx * (FR[n] - FL[n]);
    }
}

BX_DEV double amrex_max(double a, double b) { return (a < b) ? b : a; }

// HLLC flux at the face between left state UL and right state UR
// (HLLC.cpp:90-172 with estimate_wave_speed_HLLC_Toro, HLLC.H:213-309).
// NI: pm.ph.n == 7 (gamma = 1.4), so the two exponents take the near-integer
// forms for n = 7 fixed at compile time, with no general pow inlined (same
// results as the dispatching forms: the floored pressures are >= 1e-14 or
// NaN, and a compile-time n performs the operations of a run-time one).
template <int D, int dir, bool E1 = false, bool NI = false>
BX_DEV void hllc_flux(const double (&UL)[D + 2], const double (&UR)[D + 2],
                      double (&F)[D + 2], const EulerParm &pm)
{
    constexpr int NC = D + 2;
    double        QL[NC], QR[NC], yL, yR;
    primv_from_consv<D, E1>(UL, QL, pm, yL);
    primv_from_consv<D, E1>(UR, QR, pm, yR);

    const double adiabatic = pm.gamma;
    const double epsilon   = pm.eps;
    constexpr double min_rho = 1e-14;
    constexpr double min_p   = 1e-14;
    const double density_L = amrex_max(QL[0], min_rho);
    const double density_R = amrex_max(QR[0], min_rho);
    const double p_L       = amrex_max(QL[1 + D], min_p);
    const double p_R       = amrex_max(QR[1 + D], min_p);
    const double u_L       = QL[1 + dir];
    const double u_R       = QR[1 + dir];

    double c_L, c_R;
    if (E1)
    {
        // gamma p / rho with the reciprocal of the (floored) density
        constexpr double inv_min_rho = 1.0 / min_rho;
        const double     yLd = QL[0] < min_rho ? inv_min_rho : yL;
        const double     yRd = QR[0] < min_rho ? inv_min_rho : yR;
        c_L                  = sqrt_rn(div_by_recip(adiabatic * p_L, density_L, yLd));
        c_R                  = sqrt_rn(div_by_recip(adiabatic * p_R, density_R, yRd));
    }
    else
    {
        c_L = sqrt_rn(adiabatic * p_L / mul_eps<E1>(density_L, epsilon));
        c_R = sqrt_rn(adiabatic * p_R / mul_eps<E1>(density_R, epsilon));
    }

    const double z = (adiabatic - 1) / (2 * adiabatic);
    double       pzL, pzR, p_rr;
    if (NI)
    {
        pzL = bx_pow_negz_ni<7>(p_L, pm.ph);
        pzR = bx_pow_negz_ni<7>(p_R, pm.ph);
    }
    else
    {
        pzL = bx_pow_negz(p_L, -z, pm.ph);
        pzR = bx_pow_negz(p_R, -z, pm.ph);
    }
    const double ratio = (c_L + c_R - (adiabatic - 1) / 2 * (u_R - u_L))
                         / (c_L * pzL + c_R * pzR);
    if (NI)
        p_rr = bx_pow_invz_ni<7>(ratio, pm.ph);
    else
        p_rr = bx_pow_invz(ratio, 1 / z, pm.ph);
    const double q_L
        = (p_rr > p_L)
              ? sqrt_rn(1 + (adiabatic + 1) / (2 * adiabatic) * (p_rr / p_L - 1))
              : 1.;
    const double q_R
        = (p_rr > p_R)
              ? sqrt_rn(1 + (adiabatic + 1) / (2 * adiabatic) * (p_rr / p_R - 1))
              : 1.;
    const double speed_L = u_L - c_L * q_L;
    const double speed_R = u_R + c_R * q_R;
    const double denom
        = (density_L * (speed_L - u_L) - density_R * (speed_R - u_R));
    const double speed_interm
        = (denom != 0) ? (div_eps<E1>(p_R - p_L, epsilon)
                          + density_L * u_L * (speed_L - u_L)
                          - density_R * u_R * (speed_R - u_R))
                             / denom
                       : 0.;

    // Branch selection as in HLLC.cpp:101-167; sides share the star-state
    // arithmetic of HLLC_interm_state_replace on un-floored primitives.
    const bool use_L     = (0 <= speed_L);
    const bool use_starL = !use_L && (speed_L <= 0 && 0 <= speed_interm);
    const bool use_starR
        = !use_L && !use_starL && (speed_interm <= 0 && 0 <= speed_R);
    const bool use_R = !use_L && !use_starL && !use_starR && (0 >= speed_R);
    const bool left  = use_L || use_starL;

    double Us[NC], Qs[NC];
#pragma unroll
    for (int n = 0; n < NC; ++n)
    {
        Us[n] = left ? UL[n] : UR[n];
        Qs[n] = left ? QL[n] : QR[n];
    }
    double Fs[NC];
    euler_flux<D, dir, E1>(Us, Qs, Fs, pm);

    if (use_L || use_R)
    {
#pragma unroll
        for (int n = 0; n < NC; ++n)
            F[n] = Fs[n];
    }
    else if (use_starL || use_starR)
    {
        const double speed_K = use_starL ? speed_L : speed_R;
        // HLLC_interm_state_replace (HLLC.H:311-337)
        const double density  = Qs[0];
        const double u_K      = Qs[1 + dir];
        const double pressure = Qs[1 + D];
        const double energy   = Us[1 + D];
        const double y_K      = left ? yL : yR; // RN(1 / density)
        double       st[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            st[n] = Qs[n];
        st[0]       = 1;
        st[1 + dir] = speed_interm;
        const double scale_factor
            = density * (speed_K - u_K) / (speed_K - speed_interm);
        st[1 + D] = div_by_recip(energy, density, y_K)
                    + (speed_interm - u_K)
                          * (mul_eps<E1>(speed_interm, epsilon)
                             + pressure / (density * (speed_K - u_K)));
#pragma unroll
        for (int n = 0; n < NC; ++n)
            st[n] = st[n] * scale_factor;
#pragma unroll
        for (int n = 0; n < NC; ++n)
            F[n] = Fs[n] + speed_K * (st[n] - Us[n]);
    }
    else
    {
        // NaN wave speeds: the reference leaves the face unwritten
#pragma unroll
        for (int n = 0; n < NC; ++n)
            F[n] = __builtin_nan("");
    }
}

// Lax-Friedrichs flux at the face between UL and UR (compute_LF_flux,
// Force.cpp:106-146): 0.5 dx/dt (UL - UR) + 0.5 (F(UL) + F(UR)).
// lfc = 0.5 * dx / dt, formed once per launch as the reference's expression
// evaluates it (left to right).
template <int D, int dir, bool E1 = false>
BX_DEV void lf_flux(const double (&UL)[D + 2], const double (&UR)[D + 2],
                    double (&F)[D + 2], const EulerParm &pm, double lfc)
{
    constexpr int NC = D + 2;
    double        QL[NC], QR[NC], FL[NC], FR[NC];
    primv_from_consv<D, E1>(UL, QL, pm);
    primv_from_consv<D, E1>(UR, QR, pm);
    euler_flux<D, dir, E1>(UL, QL, FL, pm);
    euler_flux<D, dir, E1>(UR, QR, FR, pm);
#pragma unroll
    for (int n = 0; n < NC; ++n)
        F[n] = lfc * (UL[n] - UR[n]) + 0.5 * (FL[n] + FR[n]);
}

// FORCE flux (compute_force_flux_LR, Force.cpp:24-104): half the
// Lax-Friedrichs flux plus half the flux of the Richtmyer half-step state,
//   F = 0.5 (fc (UL - UR) + 0.5 (FL + FR)) + 0.5 F(0.5 (UL + UR) + rc (FL - FR))
// with fc = (0.5 dx) / dt and rc = dt / (2 dx), in the reference's order.
template <int D, int dir, bool E1 = false>
BX_DEV void force_flux(const double (&UL)[D + 2], const double (&UR)[D + 2],
                       double (&F)[D + 2], const EulerParm &pm, double fc,
                       double rc)
{
    constexpr int NC = D + 2;
    double        Q[NC], FL[NC], FR[NC], H[NC];
    primv_from_consv<D, E1>(UL, Q, pm);
    euler_flux<D, dir, E1>(UL, Q, FL, pm);
    primv_from_consv<D, E1>(UR, Q, pm);
    euler_flux<D, dir, E1>(UR, Q, FR, pm);
#pragma unroll
    for (int n = 0; n < NC; ++n)
    {
        F[n] = 0.5 * (fc * (UL[n] - UR[n]) + 0.5 * (FL[n] + FR[n]));
        H[n] = 0.5 * (UL[n] + UR[n]) + rc * (FL[n] - FR[n]);
    }
    primv_from_consv<D, E1>(H, Q, pm);
    euler_flux<D, dir, E1>(H, Q, FL, pm);
#pragma unroll
    for (int n = 0; n < NC; ++n)
        F[n] += 0.5 * FL[n];
}

// Explicit schemes (adv.num_method, NumericalMethods.H:10-20), numbered as
// the BX_METHOD_* codes of the C ABI.
enum Scheme
{
    SCH_MHM   = 0, // MUSCL-Hancock: reconstruct + half step, HLLC flux
    SCH_HLLC  = 1, // first-order HLLC on the cell states
    SCH_SLIC  = 2, // reconstruct + half step, FORCE flux
    SCH_FORCE = 3, // first-order FORCE
    SCH_LF    = 4  // first-order Lax-Friedrichs
};
constexpr bool scheme_reconstructs(int s) { return s == SCH_MHM || s == SCH_SLIC; }
constexpr bool scheme_uses_hllc(int s) { return s == SCH_MHM || s == SCH_HLLC; }

// Face flux of scheme SCH between the left and right states of a face.
// fc = (0.5 dx) / dt, rc = dt / (2 dx) (FORCE / LF only).
template <int D, int dir, bool E1, bool NI, int SCH>
BX_DEV void face_flux(const double (&UL)[D + 2], const double (&UR)[D + 2],
                      double (&F)[D + 2], const EulerParm &pm, double fc, double rc)
{
    if (scheme_uses_hllc(SCH))
        hllc_flux<D, dir, E1, NI>(UL, UR, F, pm);
    else if (SCH == SCH_LF)
        lf_flux<D, dir, E1>(UL, UR, F, pm, fc);
    else
        force_flux<D, dir, E1>(UL, UR, F, pm, fc, rc);
}

// First-order HLLC (num_method = hllc) uses the same face flux on the cell
// states themselves (Fluxes.H:192-204).

// max over a cell of sqrt(gamma p / (rho eps)) + |u| (Euler.cpp:59-67)
template <int D, bool E1 = false>
BX_DEV double cell_wave_speed(const double (&U)[D + 2], const EulerParm &pm)
{
    double Q[D + 2], y;
    primv_from_consv<D, E1>(U, Q, pm, y);
    double vsq = Q[1] * Q[1];
#pragma unroll
    for (int d = 1; d < D; ++d)
        vsq = vsq + Q[1 + d] * Q[1 + d];
    const double velocity = sqrt_rn<true>(vsq);
    if (E1)
        return sqrt_rn(div_by_recip(pm.gamma * Q[1 + D], Q[0], y)) + velocity;
    return sqrt_rn(pm.gamma * Q[1 + D] / mul_eps<E1>(Q[0], pm.eps)) + velocity;
}

} // namespace bx
