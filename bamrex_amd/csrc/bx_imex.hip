// Explicit work of the IMEX PEEK stages on one 2D periodic level, and the
// IMEX Runge-Kutta step around them: the Rusanov / MUSCL-Rusanov advective
// update, the enthalpy, the pressure-equation source and coefficients,
// after the pressure solve the momentum/energy flux corrections, and for
// the multi-stage tableaux the stage states from stored stage fluxes, the
// EK pressure of later stages and the weighted flux sum with the update.
// In the reference this is host code around MLMG (the multi-flux updates
// are CPU-only there):
//   compute_flux_imex_o1            Source/IMEX/IMEX_O1.cpp:18-175
//   advance_rusanov_adv             Source/IMEX/IMEX_Fluxes.cpp:10-118
//   advance_MUSCL_rusanov_adv       Source/IMEX/IMEX_Fluxes.cpp:120-324
//   adv_flux / kinetic_energy / pressure
//                                   Source/IMEX_K/euler_K.H:8-65
//   compute_enthalpy (reuse_pressure)
//                                   Source/IMEX/IMEX_O1.cpp:177-228
//   compute_pressure (EK branch)    Source/IMEX/IMEX_O1.cpp:300-391
//   pressure_source_vector          Source/IMEX/IMEX_O1.cpp:416-452
//   setup_pressure_op (a, face b)   Source/IMEX/IMEX_O1.cpp:570-614
//   update_momentum_flux / update_energy_flux
//                                   Source/IMEX/IMEX_O1.cpp:485-566
//   advance_imex_rk, sum_fluxes     Source/IMEX/IMEX_RK.cpp:22-96, 196-241
//   conservative_update             Source/Fluxes/Update.H:74-130, 218-302
//   max_dx_scaled_speed             Source/System/Euler/Euler.cpp:107-129
//
// Layout: component-major (rho, m_x, m_y, E), x fastest, nx*ny cells per
// component, no ghost cells (periodic neighbours are wrapped in the
// kernels).  Face arrays: bX is ny rows of nx+1 faces, bY ny+1 rows of nx;
// stage fluxes hold the lower face of every cell.  Every expression keeps
// the reference's operation order, so results are bit-identical to a CPU
// evaluation without contraction.
//
// Kernels and HBM bytes per cell (reads + writes, neighbours from cache):
//   k_peek_advect    U_exp, U_imp -> Uex             64 + 32
//   k_peek_system    Uex, p -> enth, rhs, a, bX, bY  40 + 40
//                    (split: + U_exp 24, + cX, cY 16)
//   k_rk_pressure_sk U_exp new, old -> p (split)     48 + 8
//   k_peek_finish    U, Uex(m), enth, p -> U_new     64 + 32   (SP111)
//   k_stage_fluxes   U_exp, Uex(m), enth, p -> F     64 + 64
//   k_rk_combine     U, 2 s stage fluxes -> 2 states 32 + 64 s + 64
//   k_rk_pressure_ek U_exp new (rho, E), old -> p    48 + 8
//   k_rk_update      U, 2 n stage fluxes -> U_new    32 + 64 n + 32
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/bamrex_hip.h"
#include "bx_helmholtz.h"

namespace bx {

struct PeekGeom
{
    int       nx, ny;
    long long n; // cells per component
};

__device__ __forceinline__ int wrap_lo(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrap_hi(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// kinetic_energy (euler_K.H:8-20): eps (m_x^2 + m_y^2) / (2 rho)
__device__ __forceinline__ double peek_ke(double r, double mx, double my, double eps)
{
    return eps * (mx * mx + my * my) / (2 * r);
}

// Cell state and its advective flux in direction DIR (adv_flux, euler_K.H:
// 53-65), with |u_DIR| for the Rusanov speed.
template <int DIR>
struct AdvCell
{
    double U[4], F[4], au;
    __device__ __forceinline__ void flux(double eps)
    {
        const double u = U[1 + DIR] / U[0];
        F[0]           = U[1 + DIR];
        F[1]           = U[1] * u;
        F[2]           = U[2] * u;
        F[3]           = u * peek_ke(U[0], U[1], U[2], eps);
        au             = fabs(u);
    }
    __device__ __forceinline__ void load(const double *__restrict__ S, long long c,
                                         long long n, double eps)
    {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            U[k] = S[c + k * n];
        flux(eps);
    }
};

// minmod (Source/Fluxes/Limiters.H:75-81)
__device__ __forceinline__ double peek_minmod(double dl, double dr)
{
    return (dl > 0 && dr > 0) ? fmin(dl, dr) : ((dl < 0 && dr < 0) ? fmax(dl, dr) : 0.0);
}

// MUSCL states of cell c in direction DIR (advance_MUSCL_rusanov_adv,
// IMEX_Fluxes.cpp:183-250): minmod half-slopes, then with LOCAL (the SP111
// local time step) a half step with the states' own advective fluxes; L
// and R hold the states with their fluxes.
template <int DIR, bool LOCAL>
__device__ __forceinline__ void muscl_states(const AdvCell<DIR> &m, const AdvCell<DIR> &c,
                                             const AdvCell<DIR> &p, double hdtdx, double eps,
                                             AdvCell<DIR> &L, AdvCell<DIR> &R)
{
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const double h = 0.5 * peek_minmod(c.U[k] - m.U[k], p.U[k] - c.U[k]);
        L.U[k]         = c.U[k] - h;
        R.U[k]         = c.U[k] + h;
    }
    if (LOCAL)
    {
        L.flux(eps);
        R.flux(eps);
#pragma unroll
        for (int k = 0; k < 4; ++k)
        {
            const double lu = hdtdx * (R.F[k] - L.F[k]);
            L.U[k]          = L.U[k] - lu;
            R.U[k]          = R.U[k] - lu;
        }
    }
    L.flux(eps);
    R.flux(eps);
}

// Rusanov face flux between cell m (lower) and c (upper)
// (IMEX_Fluxes.cpp:70-90): 0.5 ((F_c + F_m) - max(|u_c|,|u_m|) (U_c - U_m))
template <int DIR>
__device__ __forceinline__ void rusanov(const AdvCell<DIR> &m, const AdvCell<DIR> &c,
                                        double (&F)[4])
{
    // amrex::max(a, b) = (a < b) ? b : a
    const double umax = (c.au < m.au) ? m.au : c.au;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        F[k] = 0.5 * ((c.F[k] + m.F[k]) - umax * (c.U[k] - m.U[k]));
}

// Rusanov fluxes at the two faces of cell x along direction DIR, from the
// cells at x-2 .. x+2 (offsets in elements of the periodic neighbours).
// First order (advance_rusanov_adv) or MUSCL (advance_MUSCL_rusanov_adv):
// the face between cells m and c takes R of m and L of c.
template <int DIR, bool MUSCL, bool LOCAL>
__device__ __forceinline__ void dir_faces(const double *__restrict__ U, long long n,
                                          long long cm2, long long cm1, long long c0,
                                          long long cp1, long long cp2, double eps,
                                          double hdtdx, double (&Fl)[4], double (&Fr)[4])
{
    AdvCell<DIR> a1, a2, a3;
    a1.load(U, cm1, n, eps);
    a2.load(U, c0, n, eps);
    a3.load(U, cp1, n, eps);
    if (!MUSCL)
    {
        rusanov(a1, a2, Fl);
        rusanov(a2, a3, Fr);
        return;
    }
    AdvCell<DIR> a0, a4;
    a0.load(U, cm2, n, eps);
    a4.load(U, cp2, n, eps);
    AdvCell<DIR> L1, R1, L2, R2, L3, R3;
    muscl_states<DIR, LOCAL>(a0, a1, a2, hdtdx, eps, L1, R1);
    muscl_states<DIR, LOCAL>(a1, a2, a3, hdtdx, eps, L2, R2);
    muscl_states<DIR, LOCAL>(a2, a3, a4, hdtdx, eps, L3, R3);
    rusanov(R1, L2, Fl);
    rusanov(R2, L3, Fr);
}

// The Rusanov flux at the lower face of cell c0 along DIR (between cm1 and
// c0), from the cells cm2 .. cp1; the same arithmetic as dir_faces' Fl.
template <int DIR, bool MUSCL, bool LOCAL>
__device__ __forceinline__ void low_face(const double *__restrict__ U, long long n,
                                         long long cm2, long long cm1, long long c0,
                                         long long cp1, double eps, double hdtdx,
                                         double (&F)[4])
{
    AdvCell<DIR> a1, a2;
    a1.load(U, cm1, n, eps);
    a2.load(U, c0, n, eps);
    if (!MUSCL)
    {
        rusanov(a1, a2, F);
        return;
    }
    AdvCell<DIR> a0, a3;
    a0.load(U, cm2, n, eps);
    a3.load(U, cp1, n, eps);
    AdvCell<DIR> L1, R1, L2, R2;
    muscl_states<DIR, LOCAL>(a0, a1, a2, hdtdx, eps, L1, R1);
    muscl_states<DIR, LOCAL>(a1, a2, a3, hdtdx, eps, L2, R2);
    rusanov(R1, L2, F);
}

// Rusanov fluxes at the four faces of cell (i, j); hdtdx/hdtdy: half of
// dt/dx, dt/dy (the MUSCL local step).
template <bool MUSCL, bool LOCAL>
__device__ __forceinline__ void cell_faces(const double *__restrict__ U, const PeekGeom &g,
                                           int i, int j, double eps, double hdtdx,
                                           double hdtdy, double (&Fxl)[4], double (&Fxr)[4],
                                           double (&Fyl)[4], double (&Fyr)[4])
{
    const long long row = (long long)j * g.nx;
    const int       im1 = wrap_lo(i, g.nx), ip1 = wrap_hi(i, g.nx);
    const int       jm1 = wrap_lo(j, g.ny), jp1 = wrap_hi(j, g.ny);
    const int       im2 = MUSCL ? wrap_lo(im1, g.nx) : im1;
    const int       ip2 = MUSCL ? wrap_hi(ip1, g.nx) : ip1;
    const int       jm2 = MUSCL ? wrap_lo(jm1, g.ny) : jm1;
    const int       jp2 = MUSCL ? wrap_hi(jp1, g.ny) : jp1;
    dir_faces<0, MUSCL, LOCAL>(U, g.n, row + im2, row + im1, row + i, row + ip1, row + ip2, eps,
                               hdtdx, Fxl, Fxr);
    const long long nx = g.nx;
    dir_faces<1, MUSCL, LOCAL>(U, g.n, jm2 * nx + i, jm1 * nx + i, row + i, jp1 * nx + i,
                               jp2 * nx + i, eps, hdtdy, Fyl, Fyr);
}

// Explicit advective update (IMEX_Fluxes.cpp:92-111): the fluxes of the
// explicit stage state Ue applied to the implicit stage state Ui,
//   Uex = Ui - (dtdx (Fx(i+1) - Fx(i)) + dtdy (Fy(j+1) - Fy(j)))
template <bool MUSCL, bool LOCAL>
__global__ __launch_bounds__(256) void k_peek_advect(const double *__restrict__ Ue,
                                                     const double *__restrict__ Ui,
                                                     double *__restrict__ Uex, PeekGeom g,
                                                     double eps, double dtdx, double dtdy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    double Fxl[4], Fxr[4], Fyl[4], Fyr[4];
    cell_faces<MUSCL, LOCAL>(Ue, g, i, j, eps, 0.5 * dtdx, 0.5 * dtdy, Fxl, Fxr, Fyl, Fyr);
    const long long c = (long long)j * g.nx + i;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        Uex[c + k * g.n] = Ui[c + k * g.n] - (dtdx * (Fxr[k] - Fxl[k]) + dtdy * (Fyr[k] - Fyl[k]));
}

// enth = p gamma / ((gamma - 1) rho_ex)   (compute_enthalpy, reuse_pressure)
__device__ __forceinline__ double peek_enth(const double *__restrict__ p,
                                            const double *__restrict__ Uex, long long c,
                                            double adia)
{
    return p[c] * adia / ((adia - 1) * Uex[c]);
}

// Pressure system of the stage: enthalpy, the source b
// (pressure_source_vector) and the operator coefficients a = eps/(gamma-1),
// b = face average of the enthalpy (average_cellcenter_to_face:
// 0.5 (h(i-1) + h(i))).  The periodic image faces (bX column nx, bY row ny)
// repeat faces 0.  SPLIT (ke_method split, compute_ke / compute_velocity,
// IMEX_O1.cpp:232-298): ke = 0.5 eps (m_e . m_ex) / rho_e of the stage's
// explicit state Ue, and the operator's velocity c = m_e / rho_e.  FI (not
// linearised): ke of the implicit stage state Ui (statenew = statein_imp);
// conservative_kinetic: ke of Ue.  ENTH (compute_enthalpy, IMEX_O1.cpp:
// 177-229): reuse_pressure, h = p gamma / ((gamma - 1) rho_ex); conservative
// (_ex), h = gamma (E - ke) / rho_ex of Ue (of Uex), and p is replaced by
// the pressure of Uex (p is then only written).
template <int KE, int ENTH>
__global__ __launch_bounds__(256) void k_peek_system(
    const double *__restrict__ Ue, const double *__restrict__ Ui,
    const double *__restrict__ Uex, double *__restrict__ p,
    double *__restrict__ enth, double *__restrict__ rhs, double *__restrict__ acoef,
    double *__restrict__ bX, double *__restrict__ bY, double *__restrict__ cX,
    double *__restrict__ cY, PeekGeom g, double adia, double eps, double hdtdx, double hdtdy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    const long long cw  = row + wrap_lo(i, g.nx);
    const long long ce  = row + wrap_hi(i, g.nx);
    const long long cs  = (long long)wrap_lo(j, g.ny) * g.nx + i;
    const long long cn  = (long long)wrap_hi(j, g.ny) * g.nx + i;
    const long long n     = g.n;
    auto            enthf = [&](long long x) {
        if (ENTH == BX_IMEX_ENTHALPY_REUSE_PRESSURE)
            return peek_enth(p, Uex, x, adia);
        const double *S = (ENTH == BX_IMEX_ENTHALPY_CONSERVATIVE) ? Ue : Uex;
        // vol_enthalpy (euler_K.H:41-50) / rho_ex
        return adia * (S[x + 3 * n] - peek_ke(S[x], S[x + n], S[x + 2 * n], eps)) / Uex[x];
    };
    const double h  = enthf(c);
    const double hw = enthf(cw);
    const double he = enthf(ce);
    const double hs = enthf(cs);
    const double hn = enthf(cn);
    double       ke;
    if (KE == BX_IMEX_KE_SPLIT)
    {
        const double r0 = Ue[c], m1 = Ue[c + n], m2 = Ue[c + 2 * n];
        ke    = 0.5 * eps * (m1 * Uex[c + n] + m2 * Uex[c + 2 * n]) / r0;
        cX[c] = m1 / r0;
        cY[c] = m2 / r0;
    }
    else if (KE == BX_IMEX_KE_FI)
        ke = peek_ke(Ui[c], Ui[c + n], Ui[c + 2 * n], eps);
    else if (KE == BX_IMEX_KE_CONSERVATIVE_KINETIC)
        ke = peek_ke(Ue[c], Ue[c + n], Ue[c + 2 * n], eps);
    else
        ke = peek_ke(Uex[c], Uex[c + n], Uex[c + 2 * n], eps);
    if (ENTH != BX_IMEX_ENTHALPY_REUSE_PRESSURE)
        p[c] = (adia - 1) * (Uex[c + 3 * n] - peek_ke(Uex[c], Uex[c + n], Uex[c + 2 * n], eps));
    rhs[c] = eps
             * (Uex[c + 3 * n] - ke
                - (hdtdx * (he * Uex[ce + n] - hw * Uex[cw + n])
                   + hdtdy * (hn * Uex[cn + 2 * n] - hs * Uex[cs + 2 * n])));
    enth[c]           = h;
    acoef[c]          = eps / (adia - 1);
    const double bx_f = 0.5 * (hw + h);
    const double by_f = 0.5 * (hs + h);
    const int    bxw  = g.nx + 1;
    bX[(long long)j * bxw + i] = bx_f;
    if (i == 0)
        bX[(long long)j * bxw + g.nx] = bx_f;
    bY[c] = by_f;
    if (j == 0)
        bY[(long long)g.ny * g.nx + i] = by_f;
}

// Face fluxes after the pressure solve (update_momentum_flux,
// update_energy_flux) for the face between cells m (lower) and c (upper) in
// direction DIR; F holds the Rusanov flux of that face on entry.
template <int DIR>
__device__ __forceinline__ void peek_correct(double (&F)[4], const double *__restrict__ Uex,
                                             const double *__restrict__ enth,
                                             const double *__restrict__ p, long long m,
                                             long long c, long long n, double heps,
                                             double hdtde)
{
    F[1 + DIR] += heps * (p[c] + p[m]);
    const double hc = enth[c], hm = enth[m];
    F[3] += 0.5 * (hc * Uex[c + (1 + DIR) * n] + hm * Uex[m + (1 + DIR) * n])
            - hdtde * ((hc + hm) * (p[c] - p[m]));
}

// SP111 step end: the corrected fluxes and, with the weight b = 1 of
// sum_fluxes, the conservative update (Update.H:100-130)
//   U_new = U - dtdx (Fx(i+1) - Fx(i)) - dtdy (Fy(j+1) - Fy(j)),
// dtdx = dt (1/dx); the fluxes are recomputed per cell, not stored.
template <bool MUSCL>
__global__ __launch_bounds__(256) void k_peek_finish(
    const double *__restrict__ U, const double *__restrict__ Uex,
    const double *__restrict__ enth, const double *__restrict__ p, double *__restrict__ Unew,
    PeekGeom g, double eps, double hdtdx_adv, double hdtdy_adv, double dtdx, double dtdy,
    double heps, double hdtdxe, double hdtdye)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    double Fxl[4], Fxr[4], Fyl[4], Fyr[4];
    cell_faces<MUSCL, MUSCL>(U, g, i, j, eps, hdtdx_adv, hdtdy_adv, Fxl, Fxr, Fyl, Fyr);
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    const long long cw  = row + wrap_lo(i, g.nx);
    const long long ce  = row + wrap_hi(i, g.nx);
    const long long cs  = (long long)wrap_lo(j, g.ny) * g.nx + i;
    const long long cn  = (long long)wrap_hi(j, g.ny) * g.nx + i;
    peek_correct<0>(Fxl, Uex, enth, p, cw, c, g.n, heps, hdtdxe);
    peek_correct<0>(Fxr, Uex, enth, p, c, ce, g.n, heps, hdtdxe);
    peek_correct<1>(Fyl, Uex, enth, p, cs, c, g.n, heps, hdtdye);
    peek_correct<1>(Fyr, Uex, enth, p, c, cn, g.n, heps, hdtdye);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        Unew[c + k * g.n] = U[c + k * g.n] - dtdx * (Fxr[k] - Fxl[k]) - dtdy * (Fyr[k] - Fyl[k]);
}

// Stage fluxes after the solve (compute_flux_imex_o1's flux output,
// IMEX_O1.cpp:150-174): each cell writes its lower x face and lower y face,
// Fx[k n + c] the face between (i-1, j) and (i, j), Fy[k n + c] the face
// between (i, j-1) and (i, j).
template <bool MUSCL, bool LOCAL>
__global__ __launch_bounds__(256) void k_stage_fluxes(
    const double *__restrict__ Ue, const double *__restrict__ Uex,
    const double *__restrict__ enth, const double *__restrict__ p, double *__restrict__ Fx,
    double *__restrict__ Fy, PeekGeom g, double eps, double hdtdx_adv, double hdtdy_adv,
    double heps, double hdtdxe, double hdtdye)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long nx  = g.nx;
    const long long row = (long long)j * nx;
    const long long c   = row + i;
    const int       im1 = wrap_lo(i, g.nx), jm1 = wrap_lo(j, g.ny);
    const int       im2 = MUSCL ? wrap_lo(im1, g.nx) : im1;
    const int       jm2 = MUSCL ? wrap_lo(jm1, g.ny) : jm1;
    const int       ip1 = MUSCL ? wrap_hi(i, g.nx) : i;
    const int       jp1 = MUSCL ? wrap_hi(j, g.ny) : j;
    double          F[4];
    low_face<0, MUSCL, LOCAL>(Ue, g.n, row + im2, row + im1, c, row + ip1, eps, hdtdx_adv, F);
    peek_correct<0>(F, Uex, enth, p, row + im1, c, g.n, heps, hdtdxe);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        Fx[c + k * g.n] = F[k];
    low_face<1, MUSCL, LOCAL>(Ue, g.n, jm2 * nx + i, jm1 * nx + i, c, jp1 * nx + i, eps,
                              hdtdy_adv, F);
    peek_correct<1>(F, Uex, enth, p, jm1 * nx + i, c, g.n, heps, hdtdye);
#pragma unroll
    for (int k = 0; k < 4; ++k)
        Fy[c + k * g.n] = F[k];
}

// Stored fluxes of the earlier stages (face layout of k_stage_fluxes) and
// per-stage weights.
struct RKFluxes
{
    const double *fx[BX_IMEX_MAX_STAGES];
    const double *fy[BX_IMEX_MAX_STAGES];
};

struct RKWeights
{
    double x1[BX_IMEX_MAX_STAGES], y1[BX_IMEX_MAX_STAGES];
    double x2[BX_IMEX_MAX_STAGES], y2[BX_IMEX_MAX_STAGES];
};

// The two stage states from the stored fluxes (conservative_update,
// Update.H:218-302): out = U, then for each earlier stage s
//   out -= dtdx_s (Fx_s(i+1) - Fx_s(i)) + dtdy_s (Fy_s(j+1) - Fy_s(j))
// with dtdx_s = (A[stage][s] dt)(1/dx) of the implicit (1) and explicit (2)
// rows.
__global__ __launch_bounds__(256) void k_rk_combine(const double *__restrict__ U, RKFluxes f,
                                                    int nst, RKWeights w,
                                                    double *__restrict__ out1,
                                                    double *__restrict__ out2, PeekGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    const long long ce  = row + wrap_hi(i, g.nx);
    const long long cn  = (long long)wrap_hi(j, g.ny) * g.nx + i;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const long long o  = k * g.n;
        double          o1 = U[c + o], o2 = o1;
        for (int s = 0; s < nst; ++s)
        {
            const double dfx = f.fx[s][ce + o] - f.fx[s][c + o];
            const double dfy = f.fy[s][cn + o] - f.fy[s][c + o];
            o1 -= w.x1[s] * dfx + w.y1[s] * dfy;
            o2 -= w.x2[s] * dfx + w.y2[s] * dfy;
        }
        out1[c + o] = o1;
        out2[c + o] = o2;
    }
}

// EK pressure of a later stage (compute_pressure, IMEX_O1.cpp:345-384): the
// previous explicit state Uo advanced by its own advective fluxes over adt
// (MUSCL without the local step), then
//   p = (gamma - 1)(E_new - 0.5 eps (m_ex . m_ex) / rho_new).
template <bool MUSCL>
__global__ __launch_bounds__(256) void k_rk_pressure_ek(const double *__restrict__ Un,
                                                        const double *__restrict__ Uo,
                                                        double *__restrict__ p, PeekGeom g,
                                                        double adia, double eps, double dtdx,
                                                        double dtdy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    double Fxl[4], Fxr[4], Fyl[4], Fyr[4];
    cell_faces<MUSCL, false>(Uo, g, i, j, eps, 0.5 * dtdx, 0.5 * dtdy, Fxl, Fxr, Fyl, Fyr);
    const long long c  = (long long)j * g.nx + i;
    const double    e1 = Uo[c + g.n] - (dtdx * (Fxr[1] - Fxl[1]) + dtdy * (Fyr[1] - Fyl[1]));
    const double    e2 = Uo[c + 2 * g.n] - (dtdx * (Fxr[2] - Fxl[2]) + dtdy * (Fyr[2] - Fyl[2]));
    p[c] = (adia - 1) * (Un[c + 3 * g.n] - 0.5 * eps * (e1 * e1 + e2 * e2) / Un[c]);
}

// The end of one Picard iteration of the FI stage (picard_iterate,
// IMEX_O1.cpp:755-781) after the solve gave p: momentum m_ex - hdtde
// (p(+) - p(-)) (update_momentum, :455-482), enthalpy from p, kinetic
// energy of the updated state, the source and the face enthalpy.
__global__ __launch_bounds__(256) void k_picard_update(
    const double *__restrict__ Uex, const double *__restrict__ p, double *__restrict__ enth,
    double *__restrict__ rhs, double *__restrict__ bX, double *__restrict__ bY, PeekGeom g,
    double adia, double eps, double hdtdx, double hdtdy, double hdtdxe, double hdtdye)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    const long long cw  = row + wrap_lo(i, g.nx);
    const long long ce  = row + wrap_hi(i, g.nx);
    const long long cs  = (long long)wrap_lo(j, g.ny) * g.nx + i;
    const long long cn  = (long long)wrap_hi(j, g.ny) * g.nx + i;
    const long long n   = g.n;
    const double    mx  = Uex[c + n] - hdtdxe * (p[ce] - p[cw]);
    const double    my  = Uex[c + 2 * n] - hdtdye * (p[cn] - p[cs]);
    const double    h   = peek_enth(p, Uex, c, adia);
    const double    hw  = peek_enth(p, Uex, cw, adia);
    const double    he  = peek_enth(p, Uex, ce, adia);
    const double    hs  = peek_enth(p, Uex, cs, adia);
    const double    hn  = peek_enth(p, Uex, cn, adia);
    const double    ke  = peek_ke(Uex[c], mx, my, eps);
    rhs[c] = eps
             * (Uex[c + 3 * n] - ke
                - (hdtdx * (he * Uex[ce + n] - hw * Uex[cw + n])
                   + hdtdy * (hn * Uex[cn + 2 * n] - hs * Uex[cs + 2 * n])));
    enth[c]           = h;
    const double bx_f = 0.5 * (hw + h);
    const double by_f = 0.5 * (hs + h);
    const int    bxw  = g.nx + 1;
    bX[(long long)j * bxw + i] = bx_f;
    if (i == 0)
        bX[(long long)j * bxw + g.nx] = bx_f;
    bY[c] = by_f;
    if (j == 0)
        bY[(long long)g.ny * g.nx + i] = by_f;
}

// SK pressure of a later stage (compute_pressure, IMEX_O1.cpp:326-344):
//   p = (gamma - 1)(E_new - 0.5 eps (m_old . m_new) / rho_old).
__global__ __launch_bounds__(256) void k_rk_pressure_sk(const double *__restrict__ Un,
                                                        const double *__restrict__ Uo,
                                                        double *__restrict__ p, PeekGeom g,
                                                        double adia, double eps)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long c = (long long)j * g.nx + i, n = g.n;
    p[c] = (adia - 1)
           * (Un[c + 3 * n] - 0.5 * eps * (Uo[c + n] * Un[c + n] + Uo[c + 2 * n] * Un[c + 2 * n]) / Uo[c]);
}

// Step end of a multi-stage tableau: sum_fluxes (IMEX_RK.cpp:196-241,
// F = F_0 b_0, then F += F_s b_s) at the two faces of each direction and
// the conservative update (Update.H:100-130)
//   U_new = U - dtdx (F(i+1) - F(i)) - dtdy (F(j+1) - F(j)), dtdx = dt (1/dx).
__global__ __launch_bounds__(256) void k_rk_update(const double *__restrict__ U, RKFluxes f,
                                                   int nst, RKWeights b,
                                                   double *__restrict__ Unew, PeekGeom g,
                                                   double dtdx, double dtdy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    const long long ce  = row + wrap_hi(i, g.nx);
    const long long cn  = (long long)wrap_hi(j, g.ny) * g.nx + i;
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        const long long o   = k * g.n;
        double          fxl = f.fx[0][c + o] * b.x1[0], fxr = f.fx[0][ce + o] * b.x1[0];
        double          fyl = f.fy[0][c + o] * b.x1[0], fyr = f.fy[0][cn + o] * b.x1[0];
        for (int s = 1; s < nst; ++s)
        {
            fxl += f.fx[s][c + o] * b.x1[s];
            fxr += f.fx[s][ce + o] * b.x1[s];
            fyl += f.fy[s][c + o] * b.x1[s];
            fyr += f.fy[s][cn + o] * b.x1[s];
        }
        Unew[c + o] = U[c + o] - dtdx * (fxr - fxl) - dtdy * (fyr - fyl);
    }
}

// p = (gamma - 1) (E - ke)   (pressure, euler_K.H:23-35)
__global__ __launch_bounds__(256) void k_peek_pressure(const double *__restrict__ U,
                                                       double *__restrict__ p, long long n,
                                                       double adia, double eps)
{
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n)
        return;
    p[c] = (adia - 1) * (U[c + 3 * n] - peek_ke(U[c], U[c + n], U[c + 2 * n], eps));
}

// Maximum over cells of a speed, as ParReduce with ReduceOpMax leaves it
// (max(d, s) = (d < s) ? s : d from the lowest value skips NaN): only
// positive speeds enter, whose bit patterns order like their values.
//   advective (max_dx_scaled_speed, Euler.cpp:107-129):
//     (|m_x|/dx + |m_y|/dy) / rho
//   ACOUSTIC (max_wave_speed, Euler.cpp:47-74):
//     sqrt(gamma p / (rho eps)) + sqrt(u^2 + v^2)
// Block maximum of the speeds as ordered bits (valid in thread 0); speeds
// that are NaN or not positive are skipped.
template <bool ACOUSTIC>
__device__ __forceinline__ unsigned long long block_speed_max(const double *__restrict__ U,
                                                              long long n, double dx, double dy,
                                                              double adia, double eps)
{
    // four cells per thread per pass, spaced by the grid's width, with all
    // their loads issued before any arithmetic; a cell past the end repeats
    // the pass's first cell (the maximum does not change)
    constexpr int      K = 4;
    const long long    T = (long long)gridDim.x * blockDim.x;
    unsigned long long m = 0ULL;
    for (long long c0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; c0 < n; c0 += K * T)
    {
        double q[K][4];
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            const long long c = c0 + k * T < n ? c0 + k * T : c0;
            q[k][0]           = U[c];
            q[k][1]           = U[c + n];
            q[k][2]           = U[c + 2 * n];
            q[k][3]           = ACOUSTIC ? U[c + 3 * n] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
        {
            double s;
            if (ACOUSTIC)
            {
                const double r   = q[k][0];
                const double u   = q[k][1] / r;
                const double v   = q[k][2] / r;
                const double usq = u * u + v * v;
                const double p   = (adia - 1) * (q[k][3] - 0.5 * eps * r * usq);
                s                = sqrt(adia * p / (r * eps)) + sqrt(u * u + v * v);
            }
            else
                s = (fabs(q[k][1]) / dx + fabs(q[k][2]) / dy) / q[k][0];
            const unsigned long long bits
                = s > 0 ? (unsigned long long)__double_as_longlong(s) : 0ULL;
            m = bits > m ? bits : m;
        }
    }
    // wave maxima, then the block's in thread 0
    __shared__ unsigned long long red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const unsigned long long o = __shfl_xor(m, off);
        m                          = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            m = red[i] > m ? red[i] : m;
    return m;
}

// Block maxima, one partial per block with no atomics: every block writes
// its slot (0 when it saw no positive speed).
template <bool ACOUSTIC>
__global__ __launch_bounds__(256) void k_peek_speed_part(const double *__restrict__ U,
                                                         long long n, double dx, double dy,
                                                         double adia, double eps,
                                                         unsigned long long *__restrict__ part)
{
    const unsigned long long m = block_speed_max<ACOUSTIC>(U, n, dx, dy, adia, eps);
    if (threadIdx.x == 0)
        part[blockIdx.x] = m;
}

// One block: the maximum of the first nparts partials, then the acoustic
// time step from it with the host path's arithmetic (ExplicitTimeStepDimSplit,
// FluidSystem.cpp:200-205, then dt *= cfl), written to device memory.
__global__ __launch_bounds__(256) void k_acoustic_dt_from_parts(
    const unsigned long long *__restrict__ part, int nparts, double dx, double dy, double cfl,
    double *__restrict__ dt_out)
{
    unsigned long long m = 0ULL;
    for (int i = threadIdx.x; i < nparts; i += blockDim.x)
        m = part[i] > m ? part[i] : m;
    __shared__ unsigned long long red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const unsigned long long o = __shfl_xor(m, off);
        m                          = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int i = 1; i < (int)(blockDim.x >> 6); ++i)
            m = red[i] > m ? red[i] : m;
        const double smax   = __longlong_as_double((long long)m);
        double       dt_est = 1.0e20;
        const double q0 = dx / smax, q1 = dy / smax;
        dt_est = (q0 < dt_est) ? q0 : dt_est;
        dt_est = (q1 < dt_est) ? q1 : dt_est;
        dt_est *= cfl;
        *dt_out = dt_est;
    }
}

// One block: out[0] = the maximum of the first nparts partials.
__global__ __launch_bounds__(256) void k_max_of_parts(const unsigned long long *__restrict__ part,
                                                      int nparts,
                                                      unsigned long long *__restrict__ out)
{
    unsigned long long m = 0ULL;
    for (int i = threadIdx.x; i < nparts; i += 256)
        m = part[i] > m ? part[i] : m;
    __shared__ unsigned long long red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const unsigned long long o = __shfl_xor(m, off);
        m                          = o > m ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        for (int i = 1; i < 4; ++i)
            m = red[i] > m ? red[i] : m;
        out[0] = m;
    }
}

__global__ void k_peek_zero_u64(unsigned long long *p) { *p = 0ULL; }

// primitive x/y velocity and sound speed of cell c (primv_from_consv and
// sound_speed, Source/System/Euler/Euler.H:167-198, 245-254)
__device__ __forceinline__ void shock_prim(const double *__restrict__ U, long long c, long long n,
                                           double adia, double eps, double &u, double &v,
                                           double &cs)
{
    const double r = U[c];
    u              = U[c + n] / r;
    v              = U[c + 2 * n] / r;
    const double usq = u * u + v * v;
    const double p   = (adia - 1) * (U[c + 3 * n] - 0.5 * eps * r * usq);
    cs               = sqrt(adia * p / (r * eps));
}

// compute_shock_indicator, first pass (IMEX_RK.cpp:277-311): per cell,
// c_min over the cell and its four face neighbours, the centred velocity
// divergence, premax = fmin(1, fmax(0, -(div_u + k1 c_min)/(k1 c_min))).
__global__ __launch_bounds__(256) void k_shock_premax(const double *__restrict__ U,
                                                      double *__restrict__ premax, PeekGeom g,
                                                      double adia, double eps, double idx,
                                                      double idy, double k1)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long row = (long long)j * g.nx;
    const long long c   = row + i;
    // neighbours in the reference's order: i+1, i-1, j+1, j-1
    const long long nb[4] = { row + wrap_hi(i, g.nx), row + wrap_lo(i, g.nx),
                              (long long)wrap_hi(j, g.ny) * g.nx + i,
                              (long long)wrap_lo(j, g.ny) * g.nx + i };
    double          u, v, c_min;
    shock_prim(U, c, g.n, adia, eps, u, v, c_min);
    double un[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
    {
        double uk, vk, ck;
        shock_prim(U, nb[k], g.n, adia, eps, uk, vk, ck);
        c_min = (ck < c_min) ? ck : c_min; // std::min(c_min, ck)
        un[k] = k < 2 ? uk : vk;
    }
    const double div_u = 0.5 * (idx * (un[0] - un[1]) + idy * (un[2] - un[3]));
    premax[c]          = fmin(1.0, fmax(0.0, -(div_u + k1 * c_min) / (k1 * c_min)));
}

// second pass (IMEX_RK.cpp:313-339): the maximum of premax over the 3 x 3
// neighbourhood, and the level maximum (values in [0, 1]: positive ones
// order like their bit patterns; zeros of either sign contribute nothing).
__global__ __launch_bounds__(256) void k_shock_max(const double *__restrict__ premax,
                                                   double *__restrict__ ind, PeekGeom g,
                                                   unsigned long long *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    unsigned long long m = 0ULL;
    if (i < g.nx && j < g.ny)
    {
        const int xs[3] = { wrap_lo(i, g.nx), i, wrap_hi(i, g.nx) };
        const int ys[3] = { wrap_lo(j, g.ny), j, wrap_hi(j, g.ny) };
        double    mx    = premax[(long long)j * g.nx + i];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b)
            {
                const double q = premax[(long long)ys[b] * g.nx + xs[a]];
                mx             = (mx < q) ? q : mx; // std::max(mx, q)
            }
        ind[(long long)j * g.nx + i] = mx;
        if (mx > 0.0)
            m = (unsigned long long)__double_as_longlong(mx);
    }
    __shared__ unsigned long long red[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        const unsigned long long o = __shfl_xor(m, off);
        m                          = o > m ? o : m;
    }
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0)
        red[t >> 6] = m;
    __syncthreads();
    if (t == 0)
    {
        for (int k = 1; k < (int)((blockDim.x * blockDim.y) >> 6); ++k)
            m = red[k] > m ? red[k] : m;
        if (m != 0ULL)
            atomicMax(out, m);
    }
}

// the convex combination of advance_imex_rk_stab (IMEX_RK.cpp:157-193):
// where ind > 0, U = ind U_o1 + (1 - ind) U and p likewise
__global__ __launch_bounds__(256) void k_stab_blend(const double *__restrict__ ind,
                                                   const double *__restrict__ Uo1,
                                                   const double *__restrict__ po1,
                                                   double *__restrict__ U, double *__restrict__ p,
                                                   PeekGeom g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= g.nx || j >= g.ny)
        return;
    const long long c = (long long)j * g.nx + i;
    const double    w = ind[c];
    if (!(w > 0))
        return;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        U[c + k * g.n] = w * Uo1[c + k * g.n] + (1 - w) * U[c + k * g.n];
    p[c] = w * po1[c] + (1 - w) * p[c];
}

} // namespace bx

using namespace bx;

// the 2D launches put 4 rows in a block and the rows on gridDim.y (<= 65535)
static bool peek_sizes_ok(int nx, int ny)
{
    return nx >= 1 && ny >= 1 && ny <= 4 * 65535 && (long long)nx * ny <= (1LL << 31);
}

static dim3 peek_grid(int nx, int ny) { return dim3((nx + 63) / 64, (ny + 3) / 4); }

#define PEEK_CHECK(expr)                                                         \
    do                                                                           \
    {                                                                            \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess)                                                    \
            return bx_helm_fail(BX_ERR_HIP, std::string(#expr) + ": "            \
                                                + hipGetErrorString(e_));        \
    } while (0)

// One 8-byte accumulator per device and purpose, kept until the process
// ends (calls of one purpose on one device are stream-ordered by the caller).
enum
{
    ACC_SPEED = 0,
    ACC_SHOCK = 1
};
// Accumulators of the speed and shock reductions, one set per device: word
// 0 the result, words 1 .. SPEED_PARTS the per-block partial maxima (the
// partials are reduced by one block instead of by atomics on one address,
// which serialise when a thousand blocks finish together).
constexpr int SPEED_PARTS = 1024;

static int peek_acc(const char *who, int slot, unsigned long long **out)
{
    static std::mutex          mu;
    static unsigned long long *acc[64][2] = {};
    int                        dev        = 0;
    PEEK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return bx_helm_fail(BX_ERR_UNSUPPORTED, std::string(who) + ": device id");
    std::lock_guard<std::mutex> lk(mu);
    if (!acc[dev][slot])
        PEEK_CHECK(hipMalloc((void **)&acc[dev][slot],
                             sizeof(unsigned long long) * (1 + SPEED_PARTS)));
    *out = acc[dev][slot];
    return BX_OK;
}

extern "C" int bx_imex_pressure(const double *U, int nx, int ny, double gamma, double eps,
                                double *p, void *stream)
{
    if (!U || !p || !peek_sizes_ok(nx, ny))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_pressure: bad argument");
    const long long n = (long long)nx * ny;
    hipLaunchKernelGGL(k_peek_pressure, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, U, p, n, gamma, eps);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_advective_dt(const double *U, int nx, int ny, double dx, double dy,
                                    double cfl, double *dt, void *stream)
{
    if (!U || !dt || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_advective_dt: bad argument");
    hipStream_t         s = (hipStream_t)stream;
    unsigned long long *d = nullptr;
    int                 rc;
    if ((rc = peek_acc("bx_imex_advective_dt", ACC_SPEED, &d)) != BX_OK)
        return rc;
    const long long n    = (long long)nx * ny;
    const int       grid = (int)std::min<long long>((n + 255) / 256, SPEED_PARTS);
    hipLaunchKernelGGL(k_peek_speed_part<false>, dim3(grid), dim3(256), 0, s, U, n, dx, dy, 0.0, 0.0,
                       d + 1);
    hipLaunchKernelGGL(k_max_of_parts, dim3(1), dim3(256), 0, s, d + 1, grid, d);
    PEEK_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    PEEK_CHECK(hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, s));
    PEEK_CHECK(hipStreamSynchronize(s));
    // dt_est = 1 / max_dx_scaled_speed; dt_est *= cfl (AmrLevelAdv.cpp:435, 466)
    double smax;
    static_assert(sizeof(smax) == sizeof(bits), "");
    memcpy(&smax, &bits, sizeof(smax));
    double dt_est = 1 / smax;
    dt_est *= cfl;
    *dt = dt_est;
    return BX_OK;
}

static bool adv_ok(int advection_flux)
{
    return advection_flux == BX_IMEX_RUSANOV || advection_flux == BX_IMEX_MUSCL_RUSANOV;
}

// fluxes of Ue applied to Ui (Uex), then the pressure system
static int stage_assemble(const char *who, const double *Ui, const double *Ue, double *p, int nx,
                          int ny, double dx, double dy, double dt, double gamma, double eps,
                          int advection_flux, int local, int ke_method, int enthalpy_method,
                          double *Uex, double *enth, double *rhs, double *a, double *bX,
                          double *bY, double *cX, double *cY, void *stream)
{
    const bool split = ke_method == BX_IMEX_KE_SPLIT;
    if (!Ui || !Ue || !p || !Uex || !enth || !rhs || !a || !bX || !bY
        || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0) || !adv_ok(advection_flux)
        || ke_method < BX_IMEX_KE_EX || ke_method > BX_IMEX_KE_CONSERVATIVE_KINETIC
        || enthalpy_method < BX_IMEX_ENTHALPY_REUSE_PRESSURE
        || enthalpy_method > BX_IMEX_ENTHALPY_CONSERVATIVE_EX || (split && (!cX || !cY)))
        return bx_helm_fail(BX_ERR_INVALID, std::string(who) + ": bad argument");
    // outputs are written while inputs and other outputs are read
    const int   nout    = split ? 8 : 6;
    const void *outs[8] = { Uex, enth, rhs, a, bX, bY, cX, cY };
    for (int i = 0; i < nout; ++i)
    {
        if (outs[i] == (const void *)Ui || outs[i] == (const void *)Ue
            || outs[i] == (const void *)p)
            return bx_helm_fail(BX_ERR_INVALID, std::string(who) + ": output aliases an input");
        for (int j = i + 1; j < nout; ++j)
            if (outs[i] == outs[j])
                return bx_helm_fail(BX_ERR_INVALID, std::string(who) + ": outputs alias");
    }
    hipStream_t s = (hipStream_t)stream;
    PeekGeom    g{ nx, ny, (long long)nx * ny };
    // dtdx = dt / dx (IMEX_Fluxes.cpp:93); hdtdx = 0.5 dt idx (IMEX_O1.cpp:66)
    const double dtdx  = dt / dx;
    const double dtdy  = dt / dy;
    const double hdtdx = 0.5 * dt * (1.0 / dx);
    const double hdtdy = 0.5 * dt * (1.0 / dy);
    if (advection_flux == BX_IMEX_MUSCL_RUSANOV && local)
        hipLaunchKernelGGL((k_peek_advect<true, true>), peek_grid(nx, ny), dim3(64, 4), 0, s, Ue,
                           Ui, Uex, g, eps, dtdx, dtdy);
    else if (advection_flux == BX_IMEX_MUSCL_RUSANOV)
        hipLaunchKernelGGL((k_peek_advect<true, false>), peek_grid(nx, ny), dim3(64, 4), 0, s, Ue,
                           Ui, Uex, g, eps, dtdx, dtdy);
    else
        hipLaunchKernelGGL((k_peek_advect<false, false>), peek_grid(nx, ny), dim3(64, 4), 0, s,
                           Ue, Ui, Uex, g, eps, dtdx, dtdy);
    PEEK_CHECK(hipGetLastError());
    // the FI stage (not linearised) always takes the enthalpy from p
    const int em = ke_method == BX_IMEX_KE_FI ? BX_IMEX_ENTHALPY_REUSE_PRESSURE : enthalpy_method;
    double   *vx = split ? cX : nullptr, *vy = split ? cY : nullptr;
#define BX_SYSTEM(KE, EN)                                                                      \
    hipLaunchKernelGGL((k_peek_system<KE, EN>), peek_grid(nx, ny), dim3(64, 4), 0, s, Ue, Ui, Uex, \
                       p, enth, rhs, a, bX, bY, vx, vy, g, gamma, eps, hdtdx, hdtdy)
#define BX_SYSTEM_KE(KE)                                                                       \
    do                                                                                         \
    {                                                                                          \
        if (em == BX_IMEX_ENTHALPY_CONSERVATIVE)                                               \
            BX_SYSTEM(KE, BX_IMEX_ENTHALPY_CONSERVATIVE);                                      \
        else if (em == BX_IMEX_ENTHALPY_CONSERVATIVE_EX)                                       \
            BX_SYSTEM(KE, BX_IMEX_ENTHALPY_CONSERVATIVE_EX);                                   \
        else                                                                                   \
            BX_SYSTEM(KE, BX_IMEX_ENTHALPY_REUSE_PRESSURE);                                    \
    } while (0)
    switch (ke_method)
    {
    case BX_IMEX_KE_SPLIT: BX_SYSTEM_KE(BX_IMEX_KE_SPLIT); break;
    case BX_IMEX_KE_FI: BX_SYSTEM(BX_IMEX_KE_FI, BX_IMEX_ENTHALPY_REUSE_PRESSURE); break;
    case BX_IMEX_KE_CONSERVATIVE_KINETIC: BX_SYSTEM_KE(BX_IMEX_KE_CONSERVATIVE_KINETIC); break;
    default: BX_SYSTEM_KE(BX_IMEX_KE_EX); break;
    }
#undef BX_SYSTEM_KE
#undef BX_SYSTEM
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_peek_assemble(const double *U, const double *p, int nx, int ny,
                                     double dx, double dy, double dt, double gamma,
                                     double eps, int advection_flux, double *Uex, double *enth,
                                     double *rhs, double *a, double *bX, double *bY,
                                     void *stream)
{
    // reuse_pressure only reads p
    return stage_assemble("bx_imex_peek_assemble", U, U, const_cast<double *>(p), nx, ny, dx, dy,
                          dt, gamma, eps, advection_flux, 1, BX_IMEX_KE_EX,
                          BX_IMEX_ENTHALPY_REUSE_PRESSURE, Uex, enth, rhs, a, bX, bY, nullptr,
                          nullptr, stream);
}

extern "C" int bx_imex_stage_assemble(const double *U_imp, const double *U_exp, double *p,
                                      int nx, int ny, double dx, double dy, double adt,
                                      double gamma, double eps, int advection_flux,
                                      int local_time_step, int ke_method, int enthalpy_method,
                                      double *Uex, double *enth, double *rhs, double *a,
                                      double *bX, double *bY, double *cX, double *cY,
                                      void *stream)
{
    return stage_assemble("bx_imex_stage_assemble", U_imp, U_exp, p, nx, ny, dx, dy, adt, gamma,
                          eps, advection_flux, local_time_step, ke_method, enthalpy_method, Uex,
                          enth, rhs, a, bX, bY, cX, cY, stream);
}

extern "C" int bx_imex_peek_finish(const double *U, const double *Uex, const double *enth,
                                   const double *p_new, int nx, int ny, double dx, double dy,
                                   double dt, double eps, int advection_flux, double *U_new,
                                   void *stream)
{
    if (!adv_ok(advection_flux))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_peek_finish: bad advection_flux");
    if (!U || !Uex || !enth || !p_new || !U_new || !peek_sizes_ok(nx, ny) || !(dx > 0)
        || !(dy > 0) || U == U_new || Uex == U_new || (const double *)U_new == enth
        || (const double *)U_new == p_new)
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_peek_finish: bad argument");
    PeekGeom g{ nx, ny, (long long)nx * ny };
    // MUSCL local step: 0.5 (dt/dx) (IMEX_Fluxes.cpp:177-180); update:
    // dt (1/dx) (Update.H:113)
    const double hdtdx_adv = 0.5 * (dt / dx);
    const double hdtdy_adv = 0.5 * (dt / dy);
    const double dtdx      = dt * (1.0 / dx);
    const double dtdy      = dt * (1.0 / dy);
    const double heps      = 0.5 / eps;                   // IMEX_O1.cpp:499
    const double hdtdxe    = 0.5 * dt * (1.0 / dx) / eps; // IMEX_O1.cpp:66-70
    const double hdtdye    = 0.5 * dt * (1.0 / dy) / eps;
    if (advection_flux == BX_IMEX_MUSCL_RUSANOV)
        hipLaunchKernelGGL(k_peek_finish<true>, peek_grid(nx, ny), dim3(64, 4), 0,
                           (hipStream_t)stream, U, Uex, enth, p_new, U_new, g, eps, hdtdx_adv,
                           hdtdy_adv, dtdx, dtdy, heps, hdtdxe, hdtdye);
    else
        hipLaunchKernelGGL(k_peek_finish<false>, peek_grid(nx, ny), dim3(64, 4), 0,
                           (hipStream_t)stream, U, Uex, enth, p_new, U_new, g, eps, hdtdx_adv,
                           hdtdy_adv, dtdx, dtdy, heps, hdtdxe, hdtdye);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_stage_fluxes(const double *U_exp, const double *Uex, const double *enth,
                                    const double *p_new, int nx, int ny, double dx, double dy,
                                    double adt, double eps, int advection_flux,
                                    int local_time_step, double *Fx, double *Fy, void *stream)
{
    if (!U_exp || !Uex || !enth || !p_new || !Fx || !Fy || !peek_sizes_ok(nx, ny) || !(dx > 0)
        || !(dy > 0) || !adv_ok(advection_flux) || Fx == Fy)
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_stage_fluxes: bad argument");
    const void *ins[4] = { U_exp, Uex, enth, p_new };
    for (int i = 0; i < 4; ++i)
        if (ins[i] == (const void *)Fx || ins[i] == (const void *)Fy)
            return bx_helm_fail(BX_ERR_INVALID, "bx_imex_stage_fluxes: output aliases an input");
    PeekGeom     g{ nx, ny, (long long)nx * ny };
    hipStream_t  s         = (hipStream_t)stream;
    const double hdtdx_adv = 0.5 * (adt / dx);
    const double hdtdy_adv = 0.5 * (adt / dy);
    const double heps      = 0.5 / eps;
    const double hdtdxe    = 0.5 * adt * (1.0 / dx) / eps;
    const double hdtdye    = 0.5 * adt * (1.0 / dy) / eps;
    if (advection_flux == BX_IMEX_MUSCL_RUSANOV && local_time_step)
        hipLaunchKernelGGL((k_stage_fluxes<true, true>), peek_grid(nx, ny), dim3(64, 4), 0, s,
                           U_exp, Uex, enth, p_new, Fx, Fy, g, eps, hdtdx_adv, hdtdy_adv, heps,
                           hdtdxe, hdtdye);
    else if (advection_flux == BX_IMEX_MUSCL_RUSANOV)
        hipLaunchKernelGGL((k_stage_fluxes<true, false>), peek_grid(nx, ny), dim3(64, 4), 0, s,
                           U_exp, Uex, enth, p_new, Fx, Fy, g, eps, hdtdx_adv, hdtdy_adv, heps,
                           hdtdxe, hdtdye);
    else
        hipLaunchKernelGGL((k_stage_fluxes<false, false>), peek_grid(nx, ny), dim3(64, 4), 0, s,
                           U_exp, Uex, enth, p_new, Fx, Fy, g, eps, hdtdx_adv, hdtdy_adv, heps,
                           hdtdxe, hdtdye);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

// the stored fluxes of nst stages: non-null, none an output
static bool rk_fluxes(const double *const *Fx, const double *const *Fy, int nst,
                      const void *out1, const void *out2, RKFluxes &f)
{
    if (!Fx || !Fy || nst < 0 || nst > BX_IMEX_MAX_STAGES)
        return false;
    for (int k = 0; k < BX_IMEX_MAX_STAGES; ++k)
    {
        f.fx[k] = k < nst ? Fx[k] : nullptr;
        f.fy[k] = k < nst ? Fy[k] : nullptr;
        if (k < nst
            && (!f.fx[k] || !f.fy[k] || f.fx[k] == out1 || f.fy[k] == out1 || f.fx[k] == out2
                || f.fy[k] == out2))
            return false;
    }
    return true;
}

extern "C" int bx_imex_rk_combine(const double *U, const double *const *Fx,
                                  const double *const *Fy, int nstage, const double *dt_imp,
                                  const double *dt_exp, int nx, int ny, double dx, double dy,
                                  double *U_imp, double *U_exp, void *stream)
{
    RKFluxes f;
    if (!U || !U_imp || !U_exp || U_imp == U_exp || U == U_imp || U == U_exp
        || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0)
        || !rk_fluxes(Fx, Fy, nstage, U_imp, U_exp, f) || (nstage > 0 && (!dt_imp || !dt_exp)))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_rk_combine: bad argument");
    RKWeights w{};
    // dtdx[iN] = dt[iN] * idx[0] (Update.H:250-255)
    for (int k = 0; k < nstage; ++k)
    {
        w.x1[k] = dt_imp[k] * (1.0 / dx);
        w.y1[k] = dt_imp[k] * (1.0 / dy);
        w.x2[k] = dt_exp[k] * (1.0 / dx);
        w.y2[k] = dt_exp[k] * (1.0 / dy);
    }
    PeekGeom g{ nx, ny, (long long)nx * ny };
    hipLaunchKernelGGL(k_rk_combine, peek_grid(nx, ny), dim3(64, 4), 0, (hipStream_t)stream, U,
                       f, nstage, w, U_imp, U_exp, g);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_rk_pressure(const double *U_exp_new, const double *U_exp_old, int nx,
                                   int ny, double dx, double dy, double adt, double gamma,
                                   double eps, int advection_flux, int ke_method, double *p,
                                   void *stream)
{
    if (!U_exp_new || !U_exp_old || !p || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0)
        || !adv_ok(advection_flux) || (const double *)p == U_exp_new
        || (const double *)p == U_exp_old
        || ke_method < BX_IMEX_KE_EX || ke_method > BX_IMEX_KE_CONSERVATIVE_KINETIC)
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_rk_pressure: bad argument");
    PeekGeom g{ nx, ny, (long long)nx * ny };
    if (ke_method == BX_IMEX_KE_FI || ke_method == BX_IMEX_KE_CONSERVATIVE_KINETIC)
    {
        // FI and the other methods: the new explicit state's own pressure
        // (IMEX_O1.cpp:322-325, 385-387)
        hipLaunchKernelGGL(k_peek_pressure, dim3((unsigned)((g.n + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, U_exp_new, p, g.n, gamma, eps);
        PEEK_CHECK(hipGetLastError());
        return BX_OK;
    }
    if (ke_method == BX_IMEX_KE_SPLIT)
    {
        hipLaunchKernelGGL(k_rk_pressure_sk, peek_grid(nx, ny), dim3(64, 4), 0,
                           (hipStream_t)stream, U_exp_new, U_exp_old, p, g, gamma, eps);
        PEEK_CHECK(hipGetLastError());
        return BX_OK;
    }
    const double dtdx = adt / dx;
    const double dtdy = adt / dy;
    if (advection_flux == BX_IMEX_MUSCL_RUSANOV)
        hipLaunchKernelGGL(k_rk_pressure_ek<true>, peek_grid(nx, ny), dim3(64, 4), 0,
                           (hipStream_t)stream, U_exp_new, U_exp_old, p, g, gamma, eps, dtdx,
                           dtdy);
    else
        hipLaunchKernelGGL(k_rk_pressure_ek<false>, peek_grid(nx, ny), dim3(64, 4), 0,
                           (hipStream_t)stream, U_exp_new, U_exp_old, p, g, gamma, eps, dtdx,
                           dtdy);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_rk_update(const double *U, const double *const *Fx,
                                 const double *const *Fy, int nstage, const double *b, int nx,
                                 int ny, double dx, double dy, double dt, double *U_new,
                                 void *stream)
{
    RKFluxes f;
    if (!U || !U_new || U == U_new || !b || nstage < 1 || !peek_sizes_ok(nx, ny) || !(dx > 0)
        || !(dy > 0) || !rk_fluxes(Fx, Fy, nstage, U_new, U_new, f))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_rk_update: bad argument");
    RKWeights w{};
    for (int k = 0; k < nstage; ++k)
        w.x1[k] = b[k];
    PeekGeom g{ nx, ny, (long long)nx * ny };
    hipLaunchKernelGGL(k_rk_update, peek_grid(nx, ny), dim3(64, 4), 0, (hipStream_t)stream, U, f,
                       nstage, w, U_new, g, dt * (1.0 / dx), dt * (1.0 / dy));
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

// The indicator launches and the read-back of its maximum; work holds the
// per-cell indicator before the 3x3 maximum, acc the level maximum's bits.
static int shock_indicator_run(const double *U, int nx, int ny, double dx, double dy,
                               double gamma, double eps, double k1, double *work, double *ind,
                               unsigned long long *acc, double *ind_max, hipStream_t s)
{
    PeekGeom g{ nx, ny, (long long)nx * ny };
    hipLaunchKernelGGL(k_peek_zero_u64, dim3(1), dim3(1), 0, s, acc);
    // idx = geom.InvCellSizeArray() (IMEX_RK.cpp:272)
    hipLaunchKernelGGL(k_shock_premax, peek_grid(nx, ny), dim3(64, 4), 0, s, U, work, g, gamma,
                       eps, 1.0 / dx, 1.0 / dy, k1);
    PEEK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_shock_max, peek_grid(nx, ny), dim3(64, 4), 0, s, work, ind, g, acc);
    PEEK_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    PEEK_CHECK(hipMemcpyAsync(&bits, acc, sizeof(bits), hipMemcpyDeviceToHost, s));
    PEEK_CHECK(hipStreamSynchronize(s));
    double m;
    memcpy(&m, &bits, sizeof(m));
    *ind_max = m;
    return BX_OK;
}

extern "C" int bx_imex_shock_indicator(const double *U, int nx, int ny, double dx, double dy,
                                       double gamma, double eps, double k1, double *work,
                                       double *ind, double *ind_max, void *stream)
{
    if (!U || !work || !ind || !ind_max || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0)
        || !(k1 > 0) || work == ind || (const double *)work == U || (const double *)ind == U)
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_shock_indicator: bad argument");
    unsigned long long *d = nullptr;
    int                 rc;
    if ((rc = peek_acc("bx_imex_shock_indicator", ACC_SHOCK, &d)) != BX_OK)
        return rc;
    return shock_indicator_run(U, nx, ny, dx, dy, gamma, eps, k1, work, ind, d, ind_max,
                               (hipStream_t)stream);
}

extern "C" int bx_imex_stab_blend(const double *ind, const double *U_o1, const double *p_o1,
                                  int nx, int ny, double *U, double *p, void *stream)
{
    if (!ind || !U_o1 || !p_o1 || !U || !p || !peek_sizes_ok(nx, ny) || U == p
        || (const double *)U == U_o1 || (const double *)p == p_o1 || (const double *)U == ind
        || (const double *)p == ind)
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_stab_blend: bad argument");
    PeekGeom g{ nx, ny, (long long)nx * ny };
    hipLaunchKernelGGL(k_stab_blend, peek_grid(nx, ny), dim3(64, 4), 0, (hipStream_t)stream, ind,
                       U_o1, p_o1, U, p, g);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_picard_update(const double *Uex, const double *p, int nx, int ny,
                                     double dx, double dy, double adt, double gamma, double eps,
                                     double *enth, double *rhs, double *bX, double *bY,
                                     void *stream)
{
    if (!Uex || !p || !enth || !rhs || !bX || !bY || !peek_sizes_ok(nx, ny) || !(dx > 0)
        || !(dy > 0))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_picard_update: bad argument");
    const void *outs[4] = { enth, rhs, bX, bY };
    for (int i = 0; i < 4; ++i)
    {
        if (outs[i] == (const void *)Uex || outs[i] == (const void *)p)
            return bx_helm_fail(BX_ERR_INVALID, "bx_imex_picard_update: output aliases an input");
        for (int j = i + 1; j < 4; ++j)
            if (outs[i] == outs[j])
                return bx_helm_fail(BX_ERR_INVALID, "bx_imex_picard_update: outputs alias");
    }
    PeekGeom     g{ nx, ny, (long long)nx * ny };
    const double hdtdx = 0.5 * adt * (1.0 / dx); // IMEX_O1.cpp:66-70
    const double hdtdy = 0.5 * adt * (1.0 / dy);
    hipLaunchKernelGGL(k_picard_update, peek_grid(nx, ny), dim3(64, 4), 0, (hipStream_t)stream,
                       Uex, p, enth, rhs, bX, bY, g, gamma, eps, hdtdx, hdtdy, hdtdx / eps,
                       hdtdy / eps);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

// ---------------------------------------------------------------------------
// Scratch of bx_imex_acoustic_dt_noalloc: one partial maximum per block of
// its speed kernel, at most NOALLOC_PARTS blocks (8 KiB), in a file-scope
// device array that each GPU holds until the process ends.  Host code uses
// only the address hipGetSymbolAddress returns for the current device
// (noalloc_scratch), never the array's name.  Every call on one GPU shares
// that GPU's array: calls that overlap on one GPU (different streams, or
// host threads) race.
constexpr int NOALLOC_PARTS = 1024;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_PARTS];

static int noalloc_scratch(const char *who, unsigned long long **out)
{
    static std::mutex          mu;
    static unsigned long long *ptr[64] = {};
    int                        dev     = 0;
    PEEK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64)
        return bx_helm_fail(BX_ERR_UNSUPPORTED, std::string(who) + ": device id");
    std::lock_guard<std::mutex> lk(mu);
    if (!ptr[dev])
    {
        void            *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return bx_helm_fail(BX_ERR_HIP, std::string(who) + ": hipGetSymbolAddress(scratch): "
                                                + hipGetErrorString(e));
        ptr[dev] = (unsigned long long *)p;
    }
    *out = ptr[dev];
    return BX_OK;
}

extern "C" int bx_imex_acoustic_dt_noalloc(const double *U, int nx, int ny, double dx,
                                           double dy, double gamma, double eps, double cfl,
                                           double *dt_dev, void *stream)
{
    const char *who = "bx_imex_acoustic_dt_noalloc";
    if (!U || !dt_dev || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0))
        return bx_helm_fail(BX_ERR_INVALID, std::string(who) + ": bad argument");
    const long long n    = (long long)nx * ny;
    const long long grid = std::min<long long>((n + 255) / 256, NOALLOC_PARTS);
    if (grid < 1 || grid > NOALLOC_PARTS) // the grid-stride loop covers any n
        return bx_helm_fail(BX_ERR_INVALID,
                            std::string(who) + ": more partial maxima than the scratch holds");
    unsigned long long *part = nullptr;
    int                 rc;
    if ((rc = noalloc_scratch(who, &part)) != BX_OK)
        return rc;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_peek_speed_part<true>, dim3((unsigned)grid), dim3(256), 0, s, U, n, dx,
                       dy, gamma, eps, part);
    PEEK_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_acoustic_dt_from_parts, dim3(1), dim3(256), 0, s, part, (int)grid, dx,
                       dy, cfl, dt_dev);
    PEEK_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_imex_acoustic_dt(const double *U, int nx, int ny, double dx, double dy,
                                   double gamma, double eps, double cfl, double *dt,
                                   void *stream)
{
    if (!U || !dt || !peek_sizes_ok(nx, ny) || !(dx > 0) || !(dy > 0))
        return bx_helm_fail(BX_ERR_INVALID, "bx_imex_acoustic_dt: bad argument");
    hipStream_t         s = (hipStream_t)stream;
    unsigned long long *d = nullptr;
    int                 rc;
    if ((rc = peek_acc("bx_imex_acoustic_dt", ACC_SPEED, &d)) != BX_OK)
        return rc;
    const long long n    = (long long)nx * ny;
    const int       grid = (int)std::min<long long>((n + 255) / 256, SPEED_PARTS);
    hipLaunchKernelGGL(k_peek_speed_part<true>, dim3(grid), dim3(256), 0, s, U, n, dx, dy, gamma, eps,
                       d + 1);
    hipLaunchKernelGGL(k_max_of_parts, dim3(1), dim3(256), 0, s, d + 1, grid, d);
    PEEK_CHECK(hipGetLastError());
    unsigned long long bits = 0;
    PEEK_CHECK(hipMemcpyAsync(&bits, d, sizeof(bits), hipMemcpyDeviceToHost, s));
    PEEK_CHECK(hipStreamSynchronize(s));
    double smax;
    memcpy(&smax, &bits, sizeof(smax));
    // ExplicitTimeStepDimSplit (FluidSystem.cpp:200-205), then dt *= cfl
    double dt_est = 1.0e20;
    const double q0 = dx / smax, q1 = dy / smax;
    dt_est = (q0 < dt_est) ? q0 : dt_est;
    dt_est = (q1 < dt_est) ? q1 : dt_est;
    dt_est *= cfl;
    *dt = dt_est;
    return BX_OK;
}
