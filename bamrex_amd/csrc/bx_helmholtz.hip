// Variable-coefficient pressure Helmholtz solve of the IMEX (PEEK/PESK)
// schemes on one 2D periodic level: the operator, a geometric multigrid with
// the semantics of AMReX MLMG (the reference's "plain MLMG" solver) and an
// FP64 flexible GMRES whose preconditioner is a multigrid V-cycle or a
// V-cycle whose fine-level smoothing is a learned convolutional correction.
//
// Reference path (Source/IMEX/IMEX_O1.cpp:392-414, 570-640):
//   A p = alpha a p - beta div(b grad p) + gamma c . grad p
//   alpha = 1, a = eps/(gamma-1), beta = adt^2, b = face-averaged enthalpy,
//   gamma = -eps adt/2 (split KE) or 0 (PEEK)      setup_pressure_op
//   apply:   mlalapgrad_adotx   Source/LinearSolver/MLALapGrad_2D_K.H:9-33
//   smooth:  alapgrad_gsrb      Source/LinearSolver/MLALapGrad_2D_K.H:179-229
//            with dhx = beta/(h*h), cdhx = gamma/(2h) (MLALaplacianGrad.H:1243-1249)
//   coarse coefficients: average_down of a and c, average_down_faces of b
//            (MLALaplacianGrad.H:950-1027)
// MLMG defaults restated (AMReX is not available here): V-cycle, 2 pre- and
// 2 post-smoothing red-black sweeps, cell-average restriction, piecewise
// constant interpolation, coarsening by 2 while both extents stay even and
// >= 8; the bottom level gets nu_bottom (default 16) red-black sweep pairs
// (AMReX's GPU default bottom solver is BiCGStab; on these mass-term
// operators the coarsest level is diagonally dominant and the sweeps
// converge it).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/bamrex_hip.h"
#include "bx_helmholtz.h"

namespace bx {

// ------------------------------------------------------------- kernels
// Cell operations shared by the grid kernels and the single-workgroup tail
// of the V-cycle, so both give the same bits.

// (A x) of one cell from its value, its four neighbours' and its
// coefficients (mlalapgrad_adotx); dhx = beta dxinv dxinv, dhy likewise
template <class T>
__device__ __forceinline__ T apply_vals(T xc, T xl, T xr, T xd, T xu, T a_c, T bxl, T bxr,
                                        T byl, T byr, T cx, T cy, bool with_c, T alpha,
                                        T dhx, T dhy, T gamma, T dxinv, T dyinv)
{
    T v = alpha * a_c * xc - dhx * (bxr * (xr - xc) - bxl * (xc - xl))
          - dhy * (byr * (xu - xc) - byl * (xc - xd));
    if (with_c)
    {
        const T chdx = gamma * cx * T(0.5) * dxinv;
        const T chdy = gamma * cy * T(0.5) * dyinv;
        v = v + chdx * (xr - xl) + chdy * (xu - xd);
    }
    return v;
}

// (A x)_ij on a periodic level (mlalapgrad_adotx)
template <class T>
__device__ __forceinline__ T apply_at(const T *__restrict__ x, const T *__restrict__ a,
                                      const T *__restrict__ bX, const T *__restrict__ bY,
                                      const T *__restrict__ cX, const T *__restrict__ cY,
                                      int nx, int ny, T alpha, T beta, T gamma, T dxinv,
                                      T dyinv, int i, int j)
{
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const long long c   = (long long)j * nx + i;
    const T         xc  = x[c];
    const T         dhx = beta * dxinv * dxinv;
    const T         dhy = beta * dyinv * dyinv;
    // faces: bX(i,j) at j*(nx+1)+i, bY(i,j) at j*nx+i
    const T bxl = bX[(long long)j * (nx + 1) + i];
    const T bxr = bX[(long long)j * (nx + 1) + i + 1];
    const T byl = bY[(long long)j * nx + i];
    const T byr = bY[(long long)(j + 1) * nx + i];
    const T xl = x[(long long)j * nx + im], xr = x[(long long)j * nx + ip];
    const T xd = x[(long long)jm * nx + i], xu = x[(long long)jp * nx + i];
    return apply_vals(xc, xl, xr, xd, xu, a[c], bxl, bxr, byl, byr, cX ? cX[c] : T(0),
                      cY ? cY[c] : T(0), cX != nullptr, alpha, dhx, dhy, gamma, dxinv, dyinv);
}

// red/black Gauss-Seidel update of cell (i, j) (alapgrad_gsrb; periodic
// level: no boundary correction, delta = 0)
// the cell update of red/black Gauss-Seidel from its neighbours' values
template <class T>
__device__ __forceinline__ T gsrb_value(T rhs_c, T a_c, T bxl, T bxr, T byl, T byr, T cx,
                                        T cy, bool with_c, T pl, T pr, T pd, T pu, T phi_c,
                                        T alpha, T dhx, T dhy, T cdhx, T cdhy)
{
    const T gamma = alpha * a_c + dhx * (bxl + bxr) + dhy * (byl + byr);
    T rho = dhx * (bxl * pl + bxr * pr) + dhy * (byl * pd + byr * pu);
    if (with_c)
        rho = rho - cdhx * cx * (pr - pl) - cdhy * cy * (pu - pd);
    return (rhs_c + rho - phi_c * T(0)) / (gamma - T(0));
}

template <class T>
__device__ __forceinline__ void gsrb_at(T *__restrict__ phi, const T *__restrict__ rhs,
                                        const T *__restrict__ a, const T *__restrict__ bX,
                                        const T *__restrict__ bY, const T *__restrict__ cX,
                                        const T *__restrict__ cY, int nx, int ny, T alpha,
                                        T dhx, T dhy, T cdhx, T cdhy, int i, int j)
{
    const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
    const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
    const long long c   = (long long)j * nx + i;
    const T bxl = bX[(long long)j * (nx + 1) + i];
    const T bxr = bX[(long long)j * (nx + 1) + i + 1];
    const T byl = bY[(long long)j * nx + i];
    const T byr = bY[(long long)(j + 1) * nx + i];
    const T pl = phi[(long long)j * nx + im], pr = phi[(long long)j * nx + ip];
    const T pd = phi[(long long)jm * nx + i], pu = phi[(long long)jp * nx + i];
    phi[c] = gsrb_value(rhs[c], a[c], bxl, bxr, byl, byr, cX ? cX[c] : T(0),
                        cY ? cY[c] : T(0), cX != nullptr, pl, pr, pd, pu, phi[c], alpha, dhx,
                        dhy, cdhx, cdhy);
}

template <class T>
__device__ __forceinline__ T restrict_at(const T *__restrict__ f, int ncx, int I, int J)
{
    const int       nfx = 2 * ncx;
    const long long f0  = (long long)(2 * J) * nfx + 2 * I;
    return T(0.25) * (f[f0] + f[f0 + 1] + f[f0 + nfx] + f[f0 + nfx + 1]);
}

template <class T>
__global__ __launch_bounds__(256) void k_apply(
    const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ a,
    const T *__restrict__ bX, const T *__restrict__ bY,
    const T *__restrict__ cX, const T *__restrict__ cY, int nx, int ny,
    const T *__restrict__ sc, T dxinv, T dyinv, const T *__restrict__ rhs)
{
    // sc = (alpha, beta, gamma) in device memory, so that a captured graph
    // stays valid when the scalars change from one solve to the next
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nx)
        return;
    const long long c = (long long)j * nx + i;
    const T v = apply_at(x, a, bX, bY, cX, cY, nx, ny, sc[0], sc[1], sc[2], dxinv, dyinv, i, j);
    // residual form when rhs is given: y = rhs - A x
    y[c] = rhs ? rhs[c] - v : v;
}

template <class T>
__global__ __launch_bounds__(256) void k_gsrb(
    T *__restrict__ phi, const T *__restrict__ rhs, const T *__restrict__ a,
    const T *__restrict__ bX, const T *__restrict__ bY,
    const T *__restrict__ cX, const T *__restrict__ cY, int nx, int ny,
    const T *__restrict__ sc, T hx, T hy, int redblack)
{
    const T alpha = sc[0];
    const T dhx = sc[1] / (hx * hx), dhy = sc[1] / (hy * hy);
    const T cdhx = sc[2] / (2 * hx), cdhy = sc[2] / (2 * hy);
    // thread covers cells of one colour: i = 2*t + ((j + redblack) & 1)
    const int j = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = 2 * t + ((j + redblack) & 1);
    if (i >= nx)
        return;
    gsrb_at(phi, rhs, a, bX, bY, cX, cY, nx, ny, alpha, dhx, dhy, cdhx, cdhy, i, j);
}

template <class T>
__global__ __launch_bounds__(256) void k_restrict(const T *__restrict__ f,
                                                  T *__restrict__ c, int ncx,
                                                  int ncy)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    if (I >= ncx)
        return;
    c[(long long)J * ncx + I] = restrict_at(f, ncx, I, J);
}

template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(const T *__restrict__ c,
                                                     T *__restrict__ f, int nfx,
                                                     int nfy)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (i >= nfx)
        return;
    f[(long long)j * nfx + i] += c[(long long)(j >> 1) * (nfx >> 1) + (i >> 1)];
}

// Red/black smoothing of a whole level in one launch (temporal blocking):
// a workgroup loads phi on its TX x TY tile plus a halo of NH = 2 * pairs
// cells into LDS and runs the NH half-sweeps there, updating at half-sweep k
// the cells at distance > k from the loaded region's edge, whose neighbours
// are still exact; the tile interior is then written back.  Each cell value
// is computed by the same gsrb_value as the one-colour launches, from the
// same neighbour values, so the result is identical to 2*pairs launches of
// k_gsrb.  Input and output are different arrays: other workgroups of the
// same launch still read the halo of a tile after it has been written.  With `coarse` the prolongation phi += P e_c is applied at load
// (the post-smoothing of a V-cycle); with `accumulate` the result is added
// to phi_out instead of stored (a multigrid iteration's x += e); with
// `zero_in` phi is taken as zero without being read (a V-cycle's initial
// guess); with rhs_in the right-hand side is read from that FP64 array and
// with out64 the result is written there in FP64 instead of to phi_out (a
// preconditioner's input residual and output correction, the conversions
// of k_convert / k_scale).  Coefficients are read from global
// memory (read-only, cache-resident across the half-sweeps).
// Tiles: 64 x 16 on large levels; 32 x 8 where 64 x 16 tiles would leave
// most CUs idle (a small level is latency-bound: more, shorter workgroups).
constexpr int SM_TX = 64, SM_TY = 16, SM_MAXH = 8;
constexpr int SM_TX_SMALL = 32, SM_TY_SMALL = 8;

template <class T, int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_gsrb_tiled(
    const T *__restrict__ phi, T *__restrict__ phi_out, const T *__restrict__ rhs,
    const T *__restrict__ a,
    const T *__restrict__ bX, const T *__restrict__ bY, const T *__restrict__ cX,
    const T *__restrict__ cY, int nx, int ny, const T *__restrict__ sc, T hx, T hy,
    int pairs, const T *__restrict__ coarse, int accumulate, int zero_in,
    T *__restrict__ crhs, T *__restrict__ cxz, T dxinv, T dyinv,
    const double *__restrict__ rhs_in, double *__restrict__ out64)
{
    // each thread owns the region cells tid + NT m; their coefficients stay
    // in registers across the half-sweeps, phi lives in LDS.  Every global
    // load of the region is issued before the first LDS write (phi and the
    // coarse correction land in registers first), so a thread waits for
    // memory once instead of once per region cell.  With crhs the region
    // has one more halo cell, so that the smoothed values are exact on the
    // tile's neighbours too, and the block also writes the coarse
    // right-hand side R (rhs - A x) of its tile and zeroes the coarse
    // iterate (the residuals and averaging order of k_resid_restrict).
    constexpr int MAXC = ((TY + 2 * SM_MAXH) * (TX + 2 * SM_MAXH) + NT - 1) / NT;
    __shared__ T ph[(TY + 2 * SM_MAXH) * (TX + 2 * SM_MAXH)];
    __shared__ T rt[TX * TY];
    const int NH = 2 * pairs;
    const int HR = crhs ? NH + 1 : NH; // region halo
    const int RW = TX + 2 * HR, RH = TY + 2 * HR;
    const int x0 = blockIdx.x * TX - HR, y0 = blockIdx.y * TY - HR;
    const T alpha = sc[0];
    const T dhx = sc[1] / (hx * hx), dhy = sc[1] / (hy * hy);
    const T cdhx = sc[2] / (2 * hx), cdhy = sc[2] / (2 * hy);
    auto wrapx = [&](int g) { g %= nx; return g < 0 ? g + nx : g; };
    auto wrapy = [&](int g) { g %= ny; return g < 0 ? g + ny : g; };
    // the two x faces of a cell are adjacent in memory: one paired load
    // straight into a register pair
    typedef T T2 __attribute__((ext_vector_type(2)));
    T   r_rhs[MAXC], r_a[MAXC], r_byl[MAXC], r_byr[MAXC];
    T   r_cx[MAXC], r_cy[MAXC], r_ph[MAXC], r_co[MAXC];
    T2  r_bx[MAXC];
    int r_dist[MAXC], r_par[MAXC];
#pragma unroll
    for (int m = 0; m < MAXC; ++m)
    {
        const int c = threadIdx.x + NT * m;
        r_dist[m]   = -1;
        if (c >= RW * RH)
            continue;
        const int lj = c / RW, li = c - lj * RW;
        const int gi = wrapx(x0 + li), gj = wrapy(y0 + lj);
        const long long g = (long long)gj * nx + gi;
        r_ph[m] = zero_in ? T(0) : phi[g];
        if (coarse)
            r_co[m] = coarse[(long long)(gj >> 1) * (nx >> 1) + (gi >> 1)];
        // distance from the region edge; only interior cells are updated
        const int d = min(min(li, RW - 1 - li), min(lj, RH - 1 - lj));
        r_dist[m]   = d;
        r_par[m]    = (gi + gj) & 1;
        if (d == 0)
            continue;
        r_rhs[m] = rhs_in ? (T)rhs_in[g] : rhs[g];
        r_a[m]   = a[g];
        __builtin_memcpy(&r_bx[m], bX + (long long)gj * (nx + 1) + gi, sizeof(T2));
        r_byl[m] = bY[(long long)gj * nx + gi];
        r_byr[m] = bY[(long long)(gj + 1) * nx + gi];
        r_cx[m]  = cX ? cX[g] : T(0);
        r_cy[m]  = cY ? cY[g] : T(0);
    }
#pragma unroll
    for (int m = 0; m < MAXC; ++m)
        if (r_dist[m] >= 0)
            ph[threadIdx.x + NT * m] = coarse ? r_ph[m] + r_co[m] : r_ph[m];
    __syncthreads();
    for (int k = 0; k < NH; ++k)
    {
        const int rb = k & 1;
#pragma unroll
        for (int m = 0; m < MAXC; ++m)
        {
            // updated at half-sweep k: distance > k and the colour of k
            if (r_dist[m] <= k || ((r_par[m] + rb) & 1))
                continue;
            const int c = threadIdx.x + NT * m;
            ph[c] = gsrb_value(r_rhs[m], r_a[m], r_bx[m].x, r_bx[m].y, r_byl[m], r_byr[m],
                               r_cx[m], r_cy[m], cX != nullptr, ph[c - 1], ph[c + 1],
                               ph[c - RW], ph[c + RW], ph[c], alpha, dhx, dhy, cdhx, cdhy);
        }
        __syncthreads();
    }
    if (crhs)
    {
        // residuals of the tile's cells; the coefficients are read again
        // (from cache) rather than kept live past the half-sweeps
        const T beta = sc[1], gamma = sc[2];
        const T ahx = beta * dxinv * dxinv, ahy = beta * dyinv * dyinv;
        for (int c = threadIdx.x; c < TX * TY; c += NT)
        {
            const int tj = c / TX, ti = c - tj * TX;
            const int gi = wrapx(blockIdx.x * TX + ti), gj = wrapy(blockIdx.y * TY + tj);
            const long long g  = (long long)gj * nx + gi;
            const int       lc = (tj + HR) * RW + ti + HR;
            const long long fx = (long long)gj * (nx + 1) + gi;
            rt[c] = (rhs_in ? (T)rhs_in[g] : rhs[g])
                    - apply_vals(ph[lc], ph[lc - 1], ph[lc + 1], ph[lc - RW], ph[lc + RW], a[g],
                                 bX[fx], bX[fx + 1], bY[g], bY[g + nx], cX ? cX[g] : T(0),
                                 cY ? cY[g] : T(0), cX != nullptr, alpha, ahx, ahy, gamma,
                                 dxinv, dyinv);
        }
        __syncthreads();
        const int ncx = nx >> 1;
        for (int c = threadIdx.x; c < (TX / 2) * (TY / 2); c += NT)
        {
            const int cj = c / (TX / 2), ci = c - cj * (TX / 2);
            const int I = blockIdx.x * (TX / 2) + ci, J = blockIdx.y * (TY / 2) + cj;
            if (2 * I >= nx || 2 * J >= ny)
                continue;
            const T *q = rt + (2 * cj) * TX + 2 * ci;
            crhs[(long long)J * ncx + I] = T(0.25) * (q[0] + q[1] + q[TX] + q[TX + 1]);
            cxz[(long long)J * ncx + I]  = T(0);
        }
    }
    for (int c = threadIdx.x; c < TX * TY; c += NT)
    {
        const int tj = c / TX, ti = c - tj * TX;
        const int gi = blockIdx.x * TX + ti, gj = blockIdx.y * TY + tj;
        if (gi < nx && gj < ny)
        {
            const long long g = (long long)gj * nx + gi;
            const T         v = ph[(tj + HR) * RW + ti + HR];
            // accumulate: phi_out += 1 * v, the bits of k_axpy adding the
            // correction to an iterate (each cell is read and written by
            // its own thread only)
            if (out64)
                out64[g] = (double)v;
            else
                phi_out[g] = accumulate ? phi_out[g] + T(1) * v : v;
        }
    }
}

// coarse rhs = R (rhs - A x) on the fine level, and coarse x = 0, in one
// launch (the residual values and the averaging order of k_apply +
// k_restrict)
template <class T>
__global__ __launch_bounds__(256) void k_resid_restrict(
    const T *__restrict__ x, const T *__restrict__ rhs, const T *__restrict__ a,
    const T *__restrict__ bX, const T *__restrict__ bY, const T *__restrict__ cX,
    const T *__restrict__ cY, int nx, int ny, const T *__restrict__ sc, T dxinv, T dyinv,
    T *__restrict__ crhs, T *__restrict__ cx)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y;
    const int ncx = nx >> 1;
    if (I >= ncx)
        return;
    const T alpha = sc[0], beta = sc[1], gamma = sc[2];
    T r[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
    {
        const int i = 2 * I + (q & 1), j = 2 * J + (q >> 1);
        r[q] = rhs[(long long)j * nx + i]
               - apply_at(x, a, bX, bY, cX, cY, nx, ny, alpha, beta, gamma, dxinv, dyinv, i, j);
    }
    crhs[(long long)J * ncx + I] = T(0.25) * (r[0] + r[1] + r[2] + r[3]);
    cx[(long long)J * ncx + I]   = T(0);
}

// The coarse end of a V-cycle (the coarsest levels, as many as fit in
// the device's LDS per workgroup, at most TAIL_LDS_BYTES) in one workgroup of
// TAIL_THREADS threads: the coefficients and the work arrays of those
// levels are copied into LDS, the same phases as the per-level launches run
// there separated by workgroup barriers, and the correction of the first
// tail level is written back.  (gfx950: 160 KB of LDS, so an FP32 V-cycle's
// tail starts at 64 x 64, an FP64 one at 32 x 32.)
constexpr int TAIL_MAX_LEVELS = 8;
constexpr int TAIL_LDS_BYTES  = 160 * 1024;
constexpr int TAIL_THREADS    = 1024;
// cells of one colour per thread on a tail level (FP64: fewer, so that the
// unrolled half-sweep stays within 128 VGPRs)
template <class T> constexpr int tail_maxk() { return sizeof(T) <= 4 ? 4 : 2; }

template <class T> struct TailLevel
{
    const T *a, *bX, *bY, *cX, *cY; // global coefficients
    T       *x, *rhs;                // global work arrays (read/written for level 0 only)
    int      nx, ny;
    T        hx, hy, dxinv, dyinv;
    int      o_a, o_bX, o_bY, o_cX, o_cY, o_x, o_rhs; // LDS offsets (elements)
};

template <class T> struct TailArgs
{
    TailLevel<T> lev[TAIL_MAX_LEVELS];
};

// LDS view of one tail level
template <class T> struct LdsLevel
{
    T  *a, *bX, *bY, *cX, *cY, *x, *rhs;
    int nx, ny;
    T   hx, hy, dxinv, dyinv;
};

template <class T>
__device__ LdsLevel<T> lds_level(const TailLevel<T> &G, T *sm)
{
    LdsLevel<T> L;
    L.a   = sm + G.o_a;
    L.bX  = sm + G.o_bX;
    L.bY  = sm + G.o_bY;
    L.cX  = G.cX ? sm + G.o_cX : nullptr;
    L.cY  = G.cY ? sm + G.o_cY : nullptr;
    L.x   = sm + G.o_x;
    L.rhs = sm + G.o_rhs;
    L.nx = G.nx; L.ny = G.ny; L.hx = G.hx; L.hy = G.hy; L.dxinv = G.dxinv; L.dyinv = G.dyinv;
    return L;
}

template <class T>
__device__ void tail_smooth(const LdsLevel<T> &L, T alpha, T beta, T gam, int pairs)
{
    const T dhx = beta / (L.hx * L.hx), dhy = beta / (L.hy * L.hy);
    const T cdhx = gam / (2 * L.hx), cdhy = gam / (2 * L.hy);
    const int n = L.nx * L.ny;
    if (n <= 64)
    {
        // one wave owns the level: colours are separated by wave-level
        // ordering of the LDS accesses instead of workgroup barriers
        if (threadIdx.x < 64)
        {
            const int c = threadIdx.x;
            const int j = c / L.nx, i = c - j * L.nx;
            for (int it = 0; it < pairs; ++it)
                for (int rb = 0; rb < 2; ++rb)
                {
                    if (c < n && !((i + j + rb) & 1))
                        gsrb_at(L.x, L.rhs, L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny, alpha,
                                dhx, dhy, cdhx, cdhy, i, j);
                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                }
        }
        __syncthreads();
        return;
    }
    // the n / 2 cells of colour rb only (nx is even): cell k of the colour
    // is row j = k / (nx / 2), column 2 (k mod nx / 2) + ((j + rb) & 1), so
    // every lane of a wave has a cell to update.  A thread's cells
    // k = tid + TAIL_THREADS m (at most tail_maxk, checked when the tail is
    // planned) are located once per level; the half-sweeps only pick the
    // column of their colour.  The tail runs on one CU, so its time is the
    // instruction count per cell: the update takes 32-bit LDS indices.
    constexpr int MK  = tail_maxk<T>();
    const int     hx2 = L.nx >> 1, h = n >> 1;
    int           jr[MK], ie[MK];
#pragma unroll
    for (int m = 0; m < MK; ++m)
    {
        const int k = threadIdx.x + TAIL_THREADS * m;
        jr[m]       = k < h ? k / hx2 : -1;
        ie[m]       = 2 * (k - jr[m] * hx2);
    }
    const int nx = L.nx, ny = L.ny;
    for (int it = 0; it < pairs; ++it)
        for (int rb = 0; rb < 2; ++rb)
        {
#pragma unroll
            for (int m = 0; m < MK; ++m)
            {
                const int j = jr[m];
                if (j < 0)
                    break;
                const int i  = ie[m] + ((j + rb) & 1);
                const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
                const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
                const int c  = j * nx + i;
                const T   bxl = L.bX[j * (nx + 1) + i], bxr = L.bX[j * (nx + 1) + i + 1];
                const T   byl = L.bY[c], byr = L.bY[c + nx];
                L.x[c] = gsrb_value(L.rhs[c], L.a[c], bxl, bxr, byl, byr,
                                    L.cX ? L.cX[c] : T(0), L.cY ? L.cY[c] : T(0),
                                    L.cX != nullptr, L.x[j * nx + im], L.x[j * nx + ip],
                                    L.x[jm * nx + i], L.x[jp * nx + i], L.x[c], alpha, dhx,
                                    dhy, cdhx, cdhy);
            }
            __syncthreads();
        }
}

// Workgroup copy of n elements from global memory into LDS, one dword per
// lane per iteration, each going from memory straight into LDS
// (global_load_lds_dword: lane i of a wave writes the dword at M0 + 4 i, so
// the wave passes the LDS address of its lane 0, which is the same for every
// lane because TAIL_THREADS is a multiple of 64).  No thread waits for a load
// before issuing the next one; lds_copy_wait() completes them all.
typedef __attribute__((address_space(1))) void global_void;
typedef __attribute__((address_space(3))) void lds_void;

template <class T>
__device__ void copy_in(T *dst, const T *__restrict__ src, int n)
{
    static_assert(sizeof(T) % 4 == 0, "dword copies");
    const int       nw   = n * (int)(sizeof(T) / 4);
    const unsigned *s32  = reinterpret_cast<const unsigned *>(src);
    unsigned       *d32  = reinterpret_cast<unsigned *>(dst);
    const int       lane = threadIdx.x & 63;
    for (int c = threadIdx.x; c < nw; c += TAIL_THREADS)
        __builtin_amdgcn_global_load_lds((global_void *)(s32 + c), (lds_void *)(d32 + (c - lane)),
                                         4, 0, 0);
}

// This wave's copy_in loads have reached LDS (vmcnt 0); a workgroup barrier
// after it makes every wave's copies visible.
__device__ __forceinline__ void lds_copy_wait() { __builtin_amdgcn_s_waitcnt(0); }

// A: level table in device memory (the levels from the first tail level)
template <class T>
__global__ __launch_bounds__(TAIL_THREADS) void k_vcycle_tail(const TailArgs<T> *__restrict__ A,
                                                      int nlev, int nu1, int nu2,
                                                      int nu_bottom,
                                                      const T *__restrict__ sc)
{
    extern __shared__ __align__(16) unsigned char tail_smem[];
    T *sm = reinterpret_cast<T *>(tail_smem);
    const T alpha = sc[0], beta = sc[1], gam = sc[2];
    for (int l = 0; l < nlev; ++l)
    {
        const TailLevel<T> &G = A->lev[l];
        const int n = G.nx * G.ny;
        copy_in(sm + G.o_a, G.a, n);
        copy_in(sm + G.o_bX, G.bX, (G.nx + 1) * G.ny);
        copy_in(sm + G.o_bY, G.bY, G.nx * (G.ny + 1));
        if (G.cX)
        {
            copy_in(sm + G.o_cX, G.cX, n);
            copy_in(sm + G.o_cY, G.cY, n);
        }
        if (l == 0)
        {
            copy_in(sm + G.o_x, G.x, n);
            copy_in(sm + G.o_rhs, G.rhs, n);
        }
    }
    lds_copy_wait();
    __syncthreads();
    const int last = nlev - 1;
    for (int l = 0; l < last; ++l)
    {
        const LdsLevel<T> L = lds_level(A->lev[l], sm);
        const LdsLevel<T> C = lds_level(A->lev[l + 1], sm);
        tail_smooth(L, alpha, beta, gam, nu1);
        // coarse rhs = R (rhs - A x): the four fine residuals of a coarse
        // cell computed where they are averaged (the values and the
        // averaging order of restrict_at over a stored residual)
        const int nc = C.nx * C.ny;
        for (int c = threadIdx.x; c < nc; c += TAIL_THREADS)
        {
            const int J = c / C.nx, I = c - J * C.nx;
            T         r[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
            {
                const int i = 2 * I + (q & 1), j = 2 * J + (q >> 1);
                r[q] = L.rhs[j * L.nx + i]
                       - apply_at(L.x, L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny, alpha, beta,
                                  gam, L.dxinv, L.dyinv, i, j);
            }
            C.rhs[c] = T(0.25) * (r[0] + r[1] + r[2] + r[3]);
            C.x[c]   = T(0);
        }
        __syncthreads();
    }
    tail_smooth(lds_level(A->lev[last], sm), alpha, beta, gam, nu_bottom);
    for (int l = last - 1; l >= 0; --l)
    {
        const LdsLevel<T> L = lds_level(A->lev[l], sm);
        const LdsLevel<T> C = lds_level(A->lev[l + 1], sm);
        const int n = L.nx * L.ny;
        for (int c = threadIdx.x; c < n; c += TAIL_THREADS)
        {
            const int j = c / L.nx, i = c - j * L.nx;
            L.x[c] += C.x[(j >> 1) * C.nx + (i >> 1)];
        }
        __syncthreads();
        tail_smooth(L, alpha, beta, gam, nu2);
    }
    const TailLevel<T> &G0 = A->lev[0];
    for (int c = threadIdx.x; c < G0.nx * G0.ny; c += TAIL_THREADS)
        G0.x[c] = sm[G0.o_x + c];
}

// coarse face coefficients: average of the two fine faces on the coarse face
template <class T>
__global__ void k_avg_faces_x(const T *__restrict__ f, T *__restrict__ c,
                              int ncx, int ncy)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x; // 0..ncx
    const int J = blockIdx.y;
    if (I > ncx)
        return;
    const int nfx = 2 * ncx;
    c[(long long)J * (ncx + 1) + I]
        = T(0.5) * (f[(long long)(2 * J) * (nfx + 1) + 2 * I]
                    + f[(long long)(2 * J + 1) * (nfx + 1) + 2 * I]);
}

// One launch per coarse level for all its coefficients: a (and c) by cell
// average (k_restrict), bX / bY by face average (k_avg_faces_x / _y); the
// same expressions, so the same bits.  Grid (ncx + 1) x (ncy + 1).
template <class T>
__device__ __forceinline__ void coarsen_at(
    const T *__restrict__ fa, const T *__restrict__ fbX, const T *__restrict__ fbY,
    const T *__restrict__ fcX, const T *__restrict__ fcY, T *__restrict__ ca,
    T *__restrict__ cbX, T *__restrict__ cbY, T *__restrict__ ccX, T *__restrict__ ccY,
    int ncx, int ncy, int I, int J)
{
    const int nfx = 2 * ncx;
    if (J < ncy)
    {
        cbX[(long long)J * (ncx + 1) + I]
            = T(0.5) * (fbX[(long long)(2 * J) * (nfx + 1) + 2 * I]
                        + fbX[(long long)(2 * J + 1) * (nfx + 1) + 2 * I]);
        if (I < ncx)
        {
            ca[(long long)J * ncx + I] = restrict_at(fa, ncx, I, J);
            if (fcX)
            {
                ccX[(long long)J * ncx + I] = restrict_at(fcX, ncx, I, J);
                ccY[(long long)J * ncx + I] = restrict_at(fcY, ncx, I, J);
            }
        }
    }
    if (I < ncx)
        cbY[(long long)J * ncx + I] = T(0.5)
                                      * (fbY[(long long)(2 * J) * nfx + 2 * I]
                                         + fbY[(long long)(2 * J) * nfx + 2 * I + 1]);
}

template <class T>
__global__ __launch_bounds__(256) void k_coarsen_coeffs(
    const T *__restrict__ fa, const T *__restrict__ fbX, const T *__restrict__ fbY,
    const T *__restrict__ fcX, const T *__restrict__ fcY, T *__restrict__ ca,
    T *__restrict__ cbX, T *__restrict__ cbY, T *__restrict__ ccX, T *__restrict__ ccY,
    int ncx, int ncy)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x; // 0..ncx
    const int J = blockIdx.y;                            // 0..ncy
    if (I <= ncx)
        coarsen_at(fa, fbX, fbY, fcX, fcY, ca, cbX, cbY, ccX, ccY, ncx, ncy, I, J);
}

// The small coarse levels of a hierarchy, finest first, in one workgroup
// (a barrier between levels): the same expressions as k_coarsen_coeffs.
constexpr int COARSEN_MAX_LEVELS = 12;
constexpr int COARSEN_SMALL      = 64 * 64; // coarse levels up to this many cells
template <class T> struct CoarsenLevel
{
    const T *fa, *fbX, *fbY, *fcX, *fcY;
    T       *ca, *cbX, *cbY, *ccX, *ccY;
    int      ncx, ncy;
};
template <class T> struct CoarsenList
{
    int             n;
    CoarsenLevel<T> lev[COARSEN_MAX_LEVELS];
};

template <class T>
__global__ __launch_bounds__(1024) void k_coarsen_small(CoarsenList<T> cl)
{
    for (int l = 0; l < cl.n; ++l)
    {
        const CoarsenLevel<T> &c  = cl.lev[l];
        const int              w  = c.ncx + 1;
        const int              nq = w * (c.ncy + 1);
        for (int q = threadIdx.x; q < nq; q += 1024)
        {
            const int J = q / w, I = q - J * w;
            coarsen_at(c.fa, c.fbX, c.fbY, c.fcX, c.fcY, c.ca, c.cbX, c.cbY, c.ccX, c.ccY,
                       c.ncx, c.ncy, I, J);
        }
        __syncthreads();
    }
}

template <class T>
__global__ void k_avg_faces_y(const T *__restrict__ f, T *__restrict__ c,
                              int ncx, int ncy)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int J = blockIdx.y; // 0..ncy
    if (I >= ncx)
        return;
    const int nfx = 2 * ncx;
    c[(long long)J * ncx + I] = T(0.5)
                                * (f[(long long)(2 * J) * nfx + 2 * I]
                                   + f[(long long)(2 * J) * nfx + 2 * I + 1]);
}

// Bottom solve of the V-cycle in one launch: `iters` red/black sweeps of a
// level of at most BOTTOM_MAX cells, held in LDS by a single workgroup.  The
// cell update is the one of k_gsrb, so the result is the same bits as
// 2 * iters separate colour launches.  Launched with 256 threads (its loops
// step by that constant, so no phase reads the block size from memory).
constexpr int BOTTOM_MAX = 1024;

template <class T>
__global__ __launch_bounds__(256) void k_bottom_gsrb(
    T *__restrict__ phi_g, const T *__restrict__ rhs_g, const T *__restrict__ a_g,
    const T *__restrict__ bX_g, const T *__restrict__ bY_g,
    const T *__restrict__ cX_g, const T *__restrict__ cY_g, int nx, int ny,
    const T *__restrict__ sc, T hx, T hy, int iters)
{
    const T alpha = sc[0];
    const T dhx = sc[1] / (hx * hx), dhy = sc[1] / (hy * hy);
    const T cdhx = sc[2] / (2 * hx), cdhy = sc[2] / (2 * hy);
    __shared__ T phi[BOTTOM_MAX], rhs[BOTTOM_MAX], a[BOTTOM_MAX];
    __shared__ T bX[BOTTOM_MAX + 64], bY[BOTTOM_MAX + 64];
    __shared__ T cX[BOTTOM_MAX], cY[BOTTOM_MAX];
    const int n   = nx * ny;
    const int tid = threadIdx.x;
    for (int c = tid; c < n; c += 256)
    {
        phi[c] = phi_g[c];
        rhs[c] = rhs_g[c];
        a[c]   = a_g[c];
        if (cX_g)
        {
            cX[c] = cX_g[c];
            cY[c] = cY_g[c];
        }
    }
    for (int c = tid; c < (nx + 1) * ny; c += 256)
        bX[c] = bX_g[c];
    for (int c = tid; c < nx * (ny + 1); c += 256)
        bY[c] = bY_g[c];
    __syncthreads();
    for (int it = 0; it < iters; ++it)
        for (int redblack = 0; redblack < 2; ++redblack)
        {
            for (int c = tid; c < n; c += 256)
            {
                const int j = c / nx, i = c - j * nx;
                if ((i + j + redblack) & 1)
                    continue;
                const int im = i == 0 ? nx - 1 : i - 1, ip = i == nx - 1 ? 0 : i + 1;
                const int jm = j == 0 ? ny - 1 : j - 1, jp = j == ny - 1 ? 0 : j + 1;
                const T bxl = bX[j * (nx + 1) + i];
                const T bxr = bX[j * (nx + 1) + i + 1];
                const T byl = bY[j * nx + i];
                const T byr = bY[(j + 1) * nx + i];
                const T pl = phi[j * nx + im], pr = phi[j * nx + ip];
                const T pd = phi[jm * nx + i], pu = phi[jp * nx + i];
                const T gamma = alpha * a[c] + dhx * (bxl + bxr) + dhy * (byl + byr);
                T rho = dhx * (bxl * pl + bxr * pr) + dhy * (byl * pd + byr * pu);
                if (cX_g)
                    rho = rho - cdhx * cX[c] * (pr - pl) - cdhy * cY[c] * (pu - pd);
                phi[c] = (rhs[c] + rho - phi[c] * T(0)) / (gamma - T(0));
            }
            __syncthreads();
        }
    for (int c = tid; c < n; c += 256)
        phi_g[c] = phi[c];
}

template <class TO, class TI>
__global__ __launch_bounds__(256) void k_convert(const TI *__restrict__ in,
                                                 TO *__restrict__ out,
                                                 long long n, TI scale)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (TO)(in[i] * scale);
}

// Level-0 coefficients of a hierarchy from FP64 arrays, and its scalars
// (alpha, beta, gamma) into device memory, one launch:
// a, c over nx*ny cells, bX over (nx+1)*ny faces, bY over nx*(ny+1).
template <class T>
__global__ __launch_bounds__(256) void k_convert_coeffs(
    const double *__restrict__ a, const double *__restrict__ bX,
    const double *__restrict__ bY, const double *__restrict__ cX,
    const double *__restrict__ cY, T *__restrict__ oa, T *__restrict__ obX,
    T *__restrict__ obY, T *__restrict__ ocX, T *__restrict__ ocY, int nx, int ny,
    T *__restrict__ sc, T alpha, T beta, T gamma)
{
    if (blockIdx.x == 0 && threadIdx.x == 0)
    {
        sc[0] = alpha;
        sc[1] = beta;
        sc[2] = gamma;
    }
    const long long n   = (long long)nx * ny;
    const long long nbx = (long long)(nx + 1) * ny;
    const long long nby = (long long)nx * (ny + 1);
    const long long nm  = nbx > nby ? nbx : nby;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nm;
         i += (long long)gridDim.x * blockDim.x)
    {
        if (i < n)
        {
            oa[i] = (T)a[i];
            if (cX)
            {
                ocX[i] = (T)cX[i];
                ocY[i] = (T)cY[i];
            }
        }
        if (i < nbx)
            obX[i] = (T)bX[i];
        if (i < nby)
            obY[i] = (T)bY[i];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_fill(T *__restrict__ p, long long n,
                                              T v)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// y += alpha * x  (alpha from a device scalar times a host factor)
__global__ __launch_bounds__(256) void k_axpy(double *__restrict__ y,
                                              const double *__restrict__ x,
                                              long long n, double alpha)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = y[i] + alpha * x[i];
}

__global__ __launch_bounds__(256) void k_scale(double *__restrict__ y,
                                               const double *__restrict__ x,
                                               long long n, double s)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i] * s;
}

// partial dot products of w with k vectors V[0..k-1] (stride n)
__global__ __launch_bounds__(256) void k_multidot(
    const double *__restrict__ V, const double *__restrict__ w, long long n,
    int k, double *__restrict__ partial)
{
    __shared__ double sh[8][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int v0 = 0; v0 < k; v0 += 16)
    {
        const int nv = min(16, k - v0);
        double    acc[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            acc[q] = 0;
        for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
             i < n; i += (long long)gridDim.x * blockDim.x)
        {
            const double wi = w[i];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q < nv)
                    acc[q] += V[(long long)(v0 + q) * n + i] * wi;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q)
        {
            double s = acc[q];
            for (int off = 32; off > 0; off >>= 1)
                s += __shfl_xor(s, off);
            if (lane == 0)
                sh[wv][q] = s;
        }
        __syncthreads();
        if (threadIdx.x < nv)
        {
            double s = 0;
            for (int q = 0; q < (int)(blockDim.x >> 6); ++q)
                s += sh[q][threadIdx.x];
            partial[(long long)(v0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
        }
        __syncthreads();
    }
}

// r = rhs - A x (k_apply's residual) and the partial sums of r^2 with the
// grid, the per-thread order and the reduction tree of k_multidot(r, r, k =
// 1), so k_reduce_partials gives the bits of norm2's sum: the residual and
// its norm in one pass.  Launched with nb_dot blocks of 256 threads; the
// level has at most 2^31 - 1 cells.
__global__ __launch_bounds__(256) void k_residual_sumsq(
    const double *__restrict__ x, double *__restrict__ r, const double *__restrict__ rhs,
    const double *__restrict__ a, const double *__restrict__ bX,
    const double *__restrict__ bY, const double *__restrict__ cX,
    const double *__restrict__ cY, int nx, int ny, const double *__restrict__ sc,
    double dxinv, double dyinv, double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int    lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int    n    = nx * ny;
    const double al = sc[0], be = sc[1], ga = sc[2];
    double       acc = 0;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256)
    {
        const int    j  = c / nx, i = c - j * nx;
        const double ri = rhs[c] - apply_at(x, a, bX, bY, cX, cY, nx, ny, al, be, ga, dxinv,
                                            dyinv, i, j);
        r[c] = ri;
        acc += ri * ri;
    }
    double s = acc;
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if (lane == 0)
        sh[wv] = s;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double t = 0;
        for (int q = 0; q < 4; ++q)
            t += sh[q];
        partial[blockIdx.x] = t;
    }
}

// y = x / sqrt(d[0]) when d[0] > 0 (the new Arnoldi vector)
__global__ __launch_bounds__(256) void k_scale_inv_norm(double *__restrict__ y,
                                                        const double *__restrict__ x,
                                                        long long n,
                                                        const double *__restrict__ d)
{
    const double hn = sqrt(d[0]);
    if (!(hn > 0))
        return;
    const double s = 1.0 / hn;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i] * s;
}

__global__ void k_reduce_partials(const double *__restrict__ partial, int nb,
                                  int k, double *__restrict__ out)
{
    // one block per vector
    __shared__ double sh[256];
    const int v = blockIdx.x;
    double    s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        s += partial[(long long)v * nb + b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[v] = sh[0];
}

// w -= sum_q h[q] V[q]
__global__ __launch_bounds__(256) void k_subtract_combo(
    double *__restrict__ w, const double *__restrict__ V,
    const double *__restrict__ h, int k, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        double s = w[i];
        for (int q = 0; q < k; ++q)
            s -= h[q] * V[(long long)q * n + i];
        w[i] = s;
    }
}

// x += sum_q y[q] Z[q]
__global__ __launch_bounds__(256) void k_add_combo(double *__restrict__ x,
                                                   const double *__restrict__ Z,
                                                   const double *__restrict__ y,
                                                   int k, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        double s = x[i];
        for (int q = 0; q < k; ++q)
            s += y[q] * Z[(long long)q * n + i];
        x[i] = s;
    }
}

// ---- one FGMRES iteration's vector work in fewer passes ----------------
// Block partial sums of k dot products (k <= 16), one row of gridDim.x
// partials per vector (layout as k_multidot's), from per-thread sums.
__device__ inline void block_partials16(const double (&acc)[16], int k,
                                        double *__restrict__ partial)
{
    __shared__ double sh[4][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
        if (q >= k)
            break;
        double v = acc[q];
        for (int off = 32; off > 0; off >>= 1)
            v += __shfl_xor(v, off);
        if (lane == 0)
            sh[wv][q] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < k)
        partial[(long long)threadIdx.x * gridDim.x + blockIdx.x]
            = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// w = A z (FP64 operator, the k_apply expression) and the partial dot
// products of w with V[0..k) (the first Gram-Schmidt projection); 256
// threads per block, grid-stride over cells
__global__ __launch_bounds__(256) void k_apply_multidot(
    const double *__restrict__ z, double *__restrict__ w, const double *__restrict__ a,
    const double *__restrict__ bX, const double *__restrict__ bY,
    const double *__restrict__ cX, const double *__restrict__ cY, int nx, int ny,
    const double *__restrict__ sc, double dxinv, double dyinv,
    const double *__restrict__ V, int k, double *__restrict__ partial)
{
    const long long n = (long long)nx * ny;
    double          acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        acc[q] = 0;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x)
    {
        const int    j = (int)(c / nx), i = (int)(c - (long long)j * nx);
        const double v = apply_at(z, a, bX, bY, cX, cY, nx, ny, sc[0], sc[1], sc[2], dxinv,
                                  dyinv, i, j);
        w[c] = v;
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (q < k)
                acc[q] += V[(long long)q * n + c] * v;
    }
    block_partials16(acc, k, partial);
}

// w -= sum_q h[q] V[q] (q < k, the previous projection's coefficients), then
// the partial dot products of the new w with V[0..k) (second = true) or
// with itself (second = false: its squared norm, one row)
__global__ __launch_bounds__(256) void k_subtract_dots(double *__restrict__ w,
                                                       const double *__restrict__ V,
                                                       const double *__restrict__ h, int k,
                                                       long long n, int second,
                                                       double *__restrict__ partial)
{
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        acc[q] = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
    {
        double vq[16];
        double s = w[i];
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if (q < k)
            {
                vq[q] = V[(long long)q * n + i];
                s -= h[q] * vq[q];
            }
        w[i] = s;
        if (second)
        {
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q < k)
                    acc[q] += vq[q] * s;
        }
        else
            acc[0] += s * s;
    }
    block_partials16(acc, second ? k : 1, partial);
}

// FGMRES least-squares state on the device: the Hessenberg matrix H
// ((m+1) x m, row stride m), the Givens rotations (cs, sn), the rotated
// right-hand side g, the solution y of the triangular system, and a status
// word the host reads (est = |g[j+1]|, the new column's norm, and whether y
// came out non-finite).
struct Arnoldi
{
    double *H, *cs, *sn, *g, *y, *status;
    int     m;
};

// g = beta e_0
__global__ void k_arnoldi_init(Arnoldi a, double beta)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        for (int i = 0; i <= a.m; ++i)
            a.g[i] = i == 0 ? beta : 0.0;
}

// Column j of H from the two Gram-Schmidt passes (d1 + d2) and the new
// vector's squared norm d3, its previous rotations, a new rotation that
// zeroes H[j+1][j], and g rotated with it: one thread, in the order of a
// textbook Givens update (Saad, Iterative Methods, Alg. 6.9).
__device__ void arnoldi_step(Arnoldi a, const double *__restrict__ d1,
                             const double *__restrict__ d2, const double *__restrict__ d3,
                             int j)
{
    const int m  = a.m;
    double   *H  = a.H;
    for (int i = 0; i <= j; ++i)
        H[i * m + j] = d1[i] + d2[i];
    const double hn    = sqrt(d3[0]);
    H[(j + 1) * m + j] = hn;
    for (int i = 0; i < j; ++i)
    {
        const double t0    = a.cs[i] * H[i * m + j] + a.sn[i] * H[(i + 1) * m + j];
        const double t1    = -a.sn[i] * H[i * m + j] + a.cs[i] * H[(i + 1) * m + j];
        H[i * m + j]       = t0;
        H[(i + 1) * m + j] = t1;
    }
    const double den   = hypot(H[j * m + j], H[(j + 1) * m + j]);
    a.cs[j]            = den > 0 ? H[j * m + j] / den : 1.0;
    a.sn[j]            = den > 0 ? H[(j + 1) * m + j] / den : 0.0;
    H[j * m + j]       = den;
    H[(j + 1) * m + j] = 0;
    a.g[j + 1]         = -a.sn[j] * a.g[j];
    a.g[j]             = a.cs[j] * a.g[j];
    a.status[0]        = fabs(a.g[j + 1]);
    a.status[1]        = hn;
}

__global__ void k_arnoldi_step(Arnoldi a, const double *__restrict__ d1,
                               const double *__restrict__ d2,
                               const double *__restrict__ d3, int j)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        arnoldi_step(a, d1, d2, d3, j);
}

// the norm's partials summed by one block, then (thread 0) the Arnoldi step
__global__ __launch_bounds__(256) void k_norm_arnoldi(const double *__restrict__ partial, int nb,
                                                      double *__restrict__ d3, Arnoldi ar,
                                                      const double *__restrict__ d1,
                                                      const double *__restrict__ d2, int j)
{
    __shared__ double sh[256];
    double            s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
        s += partial[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1)
    {
        if ((int)threadIdx.x < off)
            sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        d3[0] = sh[0];
        arnoldi_step(ar, d1, d2, d3, j);
    }
}

// y = R^-1 g over the first j columns (back substitution); a non-finite y
// is replaced by zeros (the iterate is kept) and flagged in status[2]
__global__ void k_arnoldi_solve(Arnoldi a, int j)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const int m      = a.m;
    bool      finite = true;
    for (int i = j - 1; i >= 0; --i)
    {
        double t = a.g[i];
        for (int q = i + 1; q < j; ++q)
            t -= a.H[i * m + q] * a.y[q];
        a.y[i] = t / a.H[i * m + i];
        finite = finite && isfinite(a.y[i]);
    }
    if (!finite)
        for (int i = 0; i < j; ++i)
            a.y[i] = 0.0;
    a.status[2] = finite ? 0.0 : 1.0;
}

// doubles of the least-squares state for restart length up to m_max
static long long arnoldi_doubles(int m_max)
{
    return (long long)(m_max + 1) * m_max + 3LL * m_max + (m_max + 1) + 4;
}
static Arnoldi arnoldi_view(double *buf, int m, int m_max)
{
    Arnoldi a;
    a.m      = m;
    a.H      = buf;
    a.cs     = a.H + (long long)(m_max + 1) * m_max;
    a.sn     = a.cs + m_max;
    a.g      = a.sn + m_max;
    a.y      = a.g + (m_max + 1);
    a.status = a.y + m_max;
    return a;
}

// ------------------------------------------------------------ host class
static dim3 grid2(int nx, int ny, int bs = 256)
{
    return dim3((unsigned)((nx + bs - 1) / bs), (unsigned)ny, 1);
}

template <class T> struct MGLevel
{
    int nx = 0, ny = 0;
    T   hx = 0, hy = 0;       // cell sizes (smoother: beta / h^2)
    T   dxinv = 0, dyinv = 0; // 1 / h (operator apply)
    T  *a = nullptr, *bX = nullptr, *bY = nullptr, *cX = nullptr, *cY = nullptr;
    T  *x = nullptr, *rhs = nullptr, *res = nullptr;
};

template <class T> struct MGHierarchy
{
    std::vector<MGLevel<T>> lev;
    T                       alpha = 0, beta = 0, gamma = 0;
    T                       host_sc[3] = { 0, 0, 0 };
    T                      *d_sc       = nullptr; // (alpha, beta, gamma) on the device
    bool                    with_c     = false;
    int                     tail_l0    = -1;      // first level run by k_vcycle_tail
    int                     tail_lds   = 0;       // its LDS bytes
    TailArgs<T>            *d_tail     = nullptr;
    int                     nu1 = 2, nu2 = 2, nu_bottom = 16;
    // the finest level's coefficient arrays are the solver's own FP64
    // copies (FP64 hierarchy): not freed here
    bool                    alias0 = false;

    void free_all()
    {
        for (size_t l = 0; l < lev.size(); ++l)
        {
            const MGLevel<T> &L = lev[l];
            if (!(l == 0 && alias0))
                for (T *p : { L.a, L.bX, L.bY, L.cX, L.cY })
                    if (p)
                        (void)hipFree(p);
            for (T *p : { L.x, L.rhs, L.res })
                if (p)
                    (void)hipFree(p);
        }
        alias0 = false;
        lev.clear();
        if (d_sc)
            (void)hipFree(d_sc);
        d_sc = nullptr;
        if (d_tail)
            (void)hipFree(d_tail);
        d_tail  = nullptr;
        tail_l0 = -1;
    }
};

} // namespace bx

using namespace bx;

struct bx_helm
{
    int    nx = 0, ny = 0;
    double dx = 0, dy = 0;
    long long n = 0;
    bool   has_c = false;
    double alpha = 0, beta = 0, gamma = 0;
    // fine-level FP64 operator (owned copies)
    double *a = nullptr, *bX = nullptr, *bY = nullptr, *cX = nullptr,
           *cY = nullptr;
    MGHierarchy<double> mg64;
    // the FP32 hierarchy of the FGMRES preconditioner, built from the
    // current coefficients when a preconditioner first needs it after
    // bx_helm_set_coeffs (multigrid solves never do)
    MGHierarchy<float>  mg32;
    bool                mg32_stale   = true;
    int                 precond_fp32 = 1;
    // FGMRES workspace (restart length m <= m_max, the allocated basis)
    int     m = 20, m_max = 20;
    double *V = nullptr, *Z = nullptr, *w = nullptr, *r = nullptr,
           *partial = nullptr, *dots = nullptr, *hdev = nullptr;
    int     nb_dot = 512;
    // the least-squares state (Arnoldi) lives in arn; the host reads the
    // status words through hstat (pinned)
    double *arn = nullptr, *hstat = nullptr;
    // learned smoother
    bx_cnn *cnn       = nullptr;
    double  cnn_omega = 1.0;
    int     cnn_mode  = BX_CNN_SMOOTHER;
    double *tmp1 = nullptr, *tmp2 = nullptr;
    float  *fr = nullptr, *fe = nullptr;
    bool    coeffs_set = false;
    // preconditioner replayed as captured graphs (FGMRES), one per
    // (network attached or not, input vector, output vector): FGMRES applies
    // it to its basis vectors V_j into Z_j, a bounded set, so the graphs
    // read and write those vectors directly; an entry is recaptured when
    // anything it bakes in changes (gen, the network's weights epoch)
    struct PGraph
    {
        int            slot;
        const double  *in;
        double        *out;
        hipGraphExec_t exec;
        long long      gen, epoch;
    };
    int                 use_graph  = 1;
    int                 fused      = 1; // blocked-smoother V-cycle in the FGMRES preconditioner
    hipStream_t         cap_stream = nullptr;
    std::vector<PGraph> pgraphs;
    long long           gen = 0;
};

static thread_local std::string g_helm_err;
static int herr(int code, const std::string &m)
{
    g_helm_err = m;
    return code;
}
#define H_CHECK(expr)                                                           \
    do                                                                          \
    {                                                                           \
        hipError_t e__ = (expr);                                                \
        if (e__ != hipSuccess)                                                  \
            return herr(BX_ERR_HIP, std::string(#expr) + ": "                   \
                                        + hipGetErrorString(e__));              \
    } while (0)

extern "C" const char *bx_helm_last_error(void) { return g_helm_err.c_str(); }

int bx_helm_fail(int code, const std::string &msg) { return herr(code, msg); }

template <class T> static int halloc(T **p, long long n)
{
    H_CHECK(hipMalloc((void **)p, sizeof(T) * std::max<long long>(n, 1)));
    return BX_OK;
}

// src*: the FP64 finest-level coefficients the hierarchy is built from
// (copied into, or converted to, its level 0)
template <class T>
static int build_hierarchy(bx_helm *h, MGHierarchy<T> &mg, hipStream_t s,
                           bool *reallocated, const double *src_a, const double *src_bX,
                           const double *src_bY, const double *src_cX,
                           const double *src_cY)
{
    mg.alpha      = (T)h->alpha;
    mg.beta       = (T)h->beta;
    mg.gamma      = (T)h->gamma;
    mg.host_sc[0] = mg.alpha;
    mg.host_sc[1] = mg.beta;
    mg.host_sc[2] = mg.gamma;
    int rc;
    // the level arrays depend only on the sizes and on whether the gradient
    // term is present: keep them (and every pointer a captured graph holds)
    // when those are unchanged, and refresh the values only
    const bool keep = !mg.lev.empty() && mg.with_c == h->has_c;
    if (!keep)
    {
        *reallocated = true;
        mg.free_all();
        if ((rc = halloc(&mg.d_sc, 3)))
            return rc;
        mg.with_c = h->has_c;
    }
    // the FP64 hierarchy's finest level is the solver's own FP64 copy of the
    // coefficients
    constexpr bool alias = std::is_same<T, double>::value;
    int nx = h->nx, ny = h->ny;
    double dx = h->dx, dy = h->dy;
    while (!keep)
    {
        MGLevel<T> L;
        L.nx    = nx;
        L.ny    = ny;
        L.hx    = (T)dx;
        L.hy    = (T)dy;
        L.dxinv = (T)(1.0 / dx);
        L.dyinv = (T)(1.0 / dy);
        const long long nc = (long long)nx * ny;
        if (alias && mg.lev.empty())
        {
            L.a  = (T *)h->a;
            L.bX = (T *)h->bX;
            L.bY = (T *)h->bY;
            L.cX = h->has_c ? (T *)h->cX : nullptr;
            L.cY = h->has_c ? (T *)h->cY : nullptr;
            mg.alias0 = true;
        }
        else
        {
            if ((rc = halloc(&L.a, nc)) || (rc = halloc(&L.bX, (long long)(nx + 1) * ny))
                || (rc = halloc(&L.bY, (long long)nx * (ny + 1))))
                return rc;
            if (h->has_c && ((rc = halloc(&L.cX, nc)) || (rc = halloc(&L.cY, nc))))
                return rc;
        }
        if ((rc = halloc(&L.x, nc)) || (rc = halloc(&L.rhs, nc)) || (rc = halloc(&L.res, nc)))
            return rc;
        mg.lev.push_back(L);
        // coarsen only to even extents: on a periodic level of odd extent
        // the cells on either side of the seam have the same red/black
        // colour, so one colour pass would update neighbours concurrently
        if (nx % 4 || ny % 4 || nx / 2 < 8 || ny / 2 < 8)
            break;
        nx /= 2;
        ny /= 2;
        dx *= 2;
        dy *= 2;
    }
    if (!keep)
    {
        // the coarsest levels whose arrays fit in LDS run in one launch
        const int nl   = (int)mg.lev.size();
        int       l0   = nl;
        long long used = 0;
        int       dev = 0, lds_max = 0;
        H_CHECK(hipGetDevice(&dev));
        H_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const long long tail_bytes = std::min<long long>(lds_max, TAIL_LDS_BYTES);
        auto      need = [&](const MGLevel<T> &L) {
            const long long n = (long long)L.nx * L.ny;
            return (n + (L.nx + 1) * (long long)L.ny + L.nx * (long long)(L.ny + 1)
                    + (L.cX ? 2 * n : 0) + 2 * n)
                   * (long long)sizeof(T);
        };
        while (l0 > 0 && nl - (l0 - 1) <= TAIL_MAX_LEVELS
               && used + need(mg.lev[l0 - 1]) <= tail_bytes
               && (long long)mg.lev[l0 - 1].nx * mg.lev[l0 - 1].ny
                      <= 2LL * tail_maxk<T>() * TAIL_THREADS)
            used += need(mg.lev[--l0]);
        if (l0 < nl)
        {
            TailArgs<T> ta{};
            int         off = 0;
            for (int l = l0; l < nl; ++l)
            {
                const MGLevel<T> &L  = mg.lev[l];
                const int         n  = L.nx * L.ny;
                TailLevel<T>     &tl = ta.lev[l - l0];
                tl.a = L.a; tl.bX = L.bX; tl.bY = L.bY; tl.cX = L.cX; tl.cY = L.cY;
                tl.x = L.x; tl.rhs = L.rhs;
                tl.nx = L.nx; tl.ny = L.ny; tl.hx = L.hx; tl.hy = L.hy;
                tl.dxinv = L.dxinv; tl.dyinv = L.dyinv;
                tl.o_a   = off; off += n;
                tl.o_bX  = off; off += (L.nx + 1) * L.ny;
                tl.o_bY  = off; off += L.nx * (L.ny + 1);
                tl.o_cX  = off; off += L.cX ? n : 0;
                tl.o_cY  = off; off += L.cY ? n : 0;
                tl.o_x   = off; off += n;
                tl.o_rhs = off; off += n;
            }
            H_CHECK(hipMalloc((void **)&mg.d_tail, sizeof(TailArgs<T>)));
            H_CHECK(hipMemcpy(mg.d_tail, &ta, sizeof(ta), hipMemcpyHostToDevice));
            mg.tail_l0  = l0;
            mg.tail_lds = (int)used;
        }
    }
    // level 0 and the scalars in one launch, then one averaging launch per
    // large coarse level and one for all the small ones (the set-up is
    // bound by launches, not bytes)
    MGLevel<T> &L0 = mg.lev[0];
    const bool  c  = h->has_c;
    hipLaunchKernelGGL((k_convert_coeffs<T>), dim3(1024), dim3(256), 0, s, src_a, src_bX,
                       src_bY, c ? src_cX : nullptr, c ? src_cY : nullptr, L0.a, L0.bX, L0.bY,
                       L0.cX, L0.cY, h->nx, h->ny, mg.d_sc, mg.alpha, mg.beta, mg.gamma);
    CoarsenList<T> small{};
    for (size_t l = 1; l < mg.lev.size(); ++l)
    {
        MGLevel<T> &F = mg.lev[l - 1], &C = mg.lev[l];
        if ((long long)C.nx * C.ny <= COARSEN_SMALL && small.n < COARSEN_MAX_LEVELS)
        {
            small.lev[small.n++] = { F.a,  F.bX, F.bY, c ? F.cX : nullptr, c ? F.cY : nullptr,
                                     C.a,  C.bX, C.bY, C.cX, C.cY, C.nx, C.ny };
            continue;
        }
        if (small.n) // levels coarsen in order: flush before a large one
        {
            hipLaunchKernelGGL((k_coarsen_small<T>), dim3(1), dim3(1024), 0, s, small);
            small.n = 0;
        }
        hipLaunchKernelGGL((k_coarsen_coeffs<T>), grid2(C.nx + 1, C.ny + 1), dim3(256), 0, s,
                           F.a, F.bX, F.bY, c ? F.cX : nullptr, c ? F.cY : nullptr, C.a, C.bX,
                           C.bY, C.cX, C.cY, C.nx, C.ny);
    }
    if (small.n)
        hipLaunchKernelGGL((k_coarsen_small<T>), dim3(1), dim3(1024), 0, s, small);
    H_CHECK(hipGetLastError());
    return BX_OK;
}

template <class T>
static void launch_apply(const MGLevel<T> &L, const MGHierarchy<T> &mg,
                         const T *x, T *y, const T *rhs, hipStream_t s)
{
    hipLaunchKernelGGL((k_apply<T>), grid2(L.nx, L.ny), dim3(256), 0, s, x, y,
                       L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny, mg.d_sc, L.dxinv,
                       L.dyinv, rhs);
}

template <class T>
static void launch_smooth(const MGLevel<T> &L, const MGHierarchy<T> &mg,
                          T *phi, const T *rhs, int redblack, hipStream_t s)
{
    hipLaunchKernelGGL((k_gsrb<T>), grid2((L.nx + 1) / 2, L.ny), dim3(256), 0,
                       s, phi, rhs, L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny,
                       mg.d_sc, L.hx, L.hy, redblack);
}

// `pairs` blocked red/black pairs on level L, in -> out (+ P e_c with
// coarse), in 64 x 16 tiles of 512 threads, or 32 x 8 tiles of 256 threads
// when the level has fewer than one 64 x 16 tile per CU of an MI355X (256).
template <class T>
static void launch_tiled(const MGLevel<T> &L, const MGHierarchy<T> &mg, const T *in, T *out,
                         int pairs, const T *coarse, hipStream_t s, int accumulate = 0,
                         int zero_in = 0, T *crhs = nullptr, T *cxz = nullptr,
                         const double *rhs_in = nullptr, double *out64 = nullptr)
{
    const long long big = (long long)((L.nx + SM_TX - 1) / SM_TX) * ((L.ny + SM_TY - 1) / SM_TY);
    if (big >= 256)
        hipLaunchKernelGGL((k_gsrb_tiled<T, SM_TX, SM_TY, 512>),
                           dim3((unsigned)((L.nx + SM_TX - 1) / SM_TX),
                                (unsigned)((L.ny + SM_TY - 1) / SM_TY)),
                           dim3(512), 0, s, in, out, L.rhs, L.a, L.bX, L.bY, L.cX, L.cY, L.nx,
                           L.ny, mg.d_sc, L.hx, L.hy, pairs, coarse, accumulate, zero_in, crhs, cxz, L.dxinv,
                           L.dyinv, rhs_in, out64);
    else
        hipLaunchKernelGGL((k_gsrb_tiled<T, SM_TX_SMALL, SM_TY_SMALL, 256>),
                           dim3((unsigned)((L.nx + SM_TX_SMALL - 1) / SM_TX_SMALL),
                                (unsigned)((L.ny + SM_TY_SMALL - 1) / SM_TY_SMALL)),
                           dim3(256), 0, s, in, out, L.rhs, L.a, L.bX, L.bY, L.cX, L.cY, L.nx,
                           L.ny, mg.d_sc, L.hx, L.hy, pairs, coarse, accumulate, zero_in, crhs, cxz, L.dxinv,
                           L.dyinv, rhs_in, out64);
}

// One V-cycle from level l.  acc (level 0 only): when the blocked path
// runs, its last smoothing adds the correction to acc (acc += e, as k_axpy)
// instead of storing it, and the call returns true; otherwise the
// correction is in mg.lev[l].x and the call returns false.  x_zero: the
// cycle starts from a zero guess on level l (the blocked pre-smoothing
// takes it as zero without reading it; any other path zeroes the array
// first).
// Level l takes the blocked path of vcycle().
template <class T>
static bool blocked_level(const MGHierarchy<T> &mg, int l, bool fused)
{
    const MGLevel<T> &L      = mg.lev[l];
    const bool        tail   = mg.tail_l0 >= 0 && l == mg.tail_l0;
    const bool        bottom = l + 1 == (int)mg.lev.size();
    return !tail && !bottom && fused && mg.nu2 > 0 && 2 * mg.nu1 <= SM_MAXH
           && 2 * mg.nu2 <= SM_MAXH && L.nx % 2 == 0 && L.ny % 2 == 0;
}

// Level l's blocked path also restricts the residual in its pre-smoothing
// launch, so no other launch reads the level's right-hand side array.
template <class T>
static bool direct_io_level(const MGHierarchy<T> &mg, int l, bool fused)
{
    return blocked_level(mg, l, fused) && mg.nu1 > 0 && 2 * mg.nu1 + 1 <= SM_MAXH;
}

// rhs_in / out64 (only where direct_io_level holds): level l's right-hand
// side is read from rhs_in (FP64) and the correction written to out64
// (FP64) by the smoothing launches themselves.
template <class T>
static bool vcycle(MGHierarchy<T> &mg, int l, hipStream_t s, bool fused = false,
                   T *acc = nullptr, bool x_zero = false, const double *rhs_in = nullptr,
                   double *out64 = nullptr)
{
    MGLevel<T> &L       = mg.lev[l];
    const bool  tail    = mg.tail_l0 >= 0 && l == mg.tail_l0;
    const bool  bottom  = l + 1 == (int)mg.lev.size();
    const bool  blocked = blocked_level(mg, l, fused);
    if (x_zero && !(blocked && mg.nu1 > 0))
        hipLaunchKernelGGL((k_fill<T>), dim3(2048), dim3(256), 0, s, L.x,
                           (long long)L.nx * L.ny, (T)0);
    if (tail)
    {
        hipLaunchKernelGGL((k_vcycle_tail<T>), dim3(1), dim3(TAIL_THREADS), mg.tail_lds, s,
                           (const TailArgs<T> *)mg.d_tail, (int)mg.lev.size() - l, mg.nu1,
                           mg.nu2, mg.nu_bottom, mg.d_sc);
        return false;
    }
    if (bottom)
    {
        if (L.nx * L.ny <= BOTTOM_MAX && (L.nx + 1) * L.ny <= BOTTOM_MAX + 64
            && L.nx * (L.ny + 1) <= BOTTOM_MAX + 64)
        {
            hipLaunchKernelGGL((k_bottom_gsrb<T>), dim3(1), dim3(256), 0, s, L.x, L.rhs,
                               L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny, mg.d_sc, L.hx,
                               L.hy, mg.nu_bottom);
            return false;
        }
        for (int it = 0; it < mg.nu_bottom; ++it)
        {
            launch_smooth(L, mg, L.x, L.rhs, 0, s);
            launch_smooth(L, mg, L.x, L.rhs, 1, s);
        }
        return false;
    }
    MGLevel<T> &C = mg.lev[l + 1];
    if (blocked)
    {
        // two launches per level: blocked pre-smoothing with the residual's
        // restriction and the coarse zeroing, prolongation + blocked
        // post-smoothing
        // pre-smoothing x -> res, post-smoothing res + P e_c -> x (the res
        // array is free in this path: the residual goes straight to C.rhs)
        const T *xs = L.x;
        if (mg.nu1 > 0 && 2 * mg.nu1 + 1 <= SM_MAXH)
        {
            // pre-smoothing with the residual's restriction in the same launch
            launch_tiled(L, mg, L.x, L.res, mg.nu1, (const T *)nullptr, s, 0, x_zero ? 1 : 0,
                         C.rhs, C.x, rhs_in);
            xs = L.res;
        }
        else
        {
            if (mg.nu1 > 0)
            {
                launch_tiled(L, mg, L.x, L.res, mg.nu1, (const T *)nullptr, s, 0,
                             x_zero ? 1 : 0);
                xs = L.res;
            }
            hipLaunchKernelGGL((k_resid_restrict<T>), grid2(C.nx, C.ny), dim3(256), 0, s, xs,
                               L.rhs, L.a, L.bX, L.bY, L.cX, L.cY, L.nx, L.ny, mg.d_sc,
                               L.dxinv, L.dyinv, C.rhs, C.x);
        }
        vcycle(mg, l + 1, s, fused);
        if (acc)
        {
            launch_tiled(L, mg, xs, acc, mg.nu2, (const T *)C.x, s, 1, 0, (T *)nullptr,
                         (T *)nullptr, rhs_in);
            return true;
        }
        launch_tiled(L, mg, xs, L.x, mg.nu2, (const T *)C.x, s, 0, 0, (T *)nullptr,
                     (T *)nullptr, rhs_in, out64);
        return false;
    }
    for (int it = 0; it < mg.nu1; ++it)
    {
        launch_smooth(L, mg, L.x, L.rhs, 0, s);
        launch_smooth(L, mg, L.x, L.rhs, 1, s);
    }
    launch_apply(L, mg, L.x, L.res, L.rhs, s);
    hipLaunchKernelGGL((k_restrict<T>), grid2(C.nx, C.ny), dim3(256), 0, s,
                       L.res, C.rhs, C.nx, C.ny);
    hipLaunchKernelGGL((k_fill<T>), dim3(256), dim3(256), 0, s, C.x,
                       (long long)C.nx * C.ny, (T)0);
    vcycle(mg, l + 1, s, fused);
    hipLaunchKernelGGL((k_prolong_add<T>), grid2(L.nx, L.ny), dim3(256), 0, s,
                       C.x, L.x, L.nx, L.ny);
    for (int it = 0; it < mg.nu2; ++it)
    {
        launch_smooth(L, mg, L.x, L.rhs, 0, s);
        launch_smooth(L, mg, L.x, L.rhs, 1, s);
    }
    return false;
}

// coarse-grid part of the V-cycle only (levels >= 1), from L.res on level 0
template <class T>
static void coarse_correction(MGHierarchy<T> &mg, hipStream_t s, bool fused)
{
    MGLevel<T> &L = mg.lev[0];
    MGLevel<T> &C = mg.lev[1];
    hipLaunchKernelGGL((k_restrict<T>), grid2(C.nx, C.ny), dim3(256), 0, s,
                       L.res, C.rhs, C.nx, C.ny);
    hipLaunchKernelGGL((k_fill<T>), dim3(256), dim3(256), 0, s, C.x,
                       (long long)C.nx * C.ny, (T)0);
    vcycle(mg, 1, s, fused);
    hipLaunchKernelGGL((k_prolong_add<T>), grid2(L.nx, L.ny), dim3(256), 0, s,
                       C.x, L.x, L.nx, L.ny);
}

static int dot_reduce(bx_helm *h, const double *V, const double *w, int k,
                      hipStream_t s, double *host_out)
{
    hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, V, w, h->n,
                       k, h->partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(256), 0, s, h->partial,
                       h->nb_dot, k, h->dots);
    H_CHECK(hipGetLastError());
    H_CHECK(hipMemcpyAsync(host_out, h->dots, sizeof(double) * k,
                           hipMemcpyDeviceToHost, s));
    H_CHECK(hipStreamSynchronize(s));
    return BX_OK;
}

static int norm2(bx_helm *h, const double *v, hipStream_t s, double *out)
{
    double d = 0;
    int    rc = dot_reduce(h, v, v, 1, s, &d);
    *out      = std::sqrt(d);
    return rc;
}

// ||u||_2 and ||v||_2 with one host synchronisation
static int norm2_pair(bx_helm *h, const double *u, const double *v, hipStream_t s,
                      double *nu, double *nv)
{
    hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, u, u, h->n, 1,
                       h->partial);
    hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, v, v, h->n, 1,
                       h->partial + h->nb_dot);
    hipLaunchKernelGGL(k_reduce_partials, dim3(2), dim3(256), 0, s, h->partial, h->nb_dot,
                       2, h->dots);
    H_CHECK(hipGetLastError());
    H_CHECK(hipMemcpyAsync(h->hstat, h->dots, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    H_CHECK(hipStreamSynchronize(s));
    *nu = std::sqrt(h->hstat[0]);
    *nv = std::sqrt(h->hstat[1]);
    return BX_OK;
}

// r = b - A x at FP64
static void residual64(bx_helm *h, const double *b, const double *x, double *r,
                       hipStream_t s)
{
    launch_apply(h->mg64.lev[0], h->mg64, x, r, b, s);
}

// ------------------------------------------------------------- C ABI
extern "C" int bx_helm_create(int nx, int ny, double dx, double dy,
                              int periodic, bx_helm **out)
{
    if (!out || nx < 8 || ny < 8 || !(dx > 0) || !(dy > 0))
        return herr(BX_ERR_INVALID, "bx_helm_create: bad argument");
    if (!periodic)
        return herr(BX_ERR_UNSUPPORTED,
                    "bx_helm_create: only periodic levels are implemented");
    // red/black Gauss-Seidel on a periodic level needs even extents (an odd
    // extent makes the two cells across the seam the same colour)
    if (nx % 2 || ny % 2)
        return herr(BX_ERR_UNSUPPORTED, "bx_helm_create: nx and ny must be even");
    if ((long long)nx * ny > 2147483647LL)
        return herr(BX_ERR_UNSUPPORTED, "bx_helm_create: more than 2^31 - 1 cells");
    bx_helm *h = new bx_helm();
    h->nx = nx;
    h->ny = ny;
    h->dx = dx;
    h->dy = dy;
    h->n  = (long long)nx * ny;
    int rc;
    if ((rc = halloc(&h->a, h->n)) || (rc = halloc(&h->bX, (long long)(nx + 1) * ny))
        || (rc = halloc(&h->bY, (long long)nx * (ny + 1)))
        || (rc = halloc(&h->V, h->n * (h->m_max + 1))) || (rc = halloc(&h->Z, h->n * h->m_max))
        || (rc = halloc(&h->w, h->n)) || (rc = halloc(&h->r, h->n))
        || (rc = halloc(&h->tmp1, h->n)) || (rc = halloc(&h->tmp2, h->n))
        || (rc = halloc(&h->fr, h->n)) || (rc = halloc(&h->fe, h->n))
        || (rc = halloc(&h->partial, (long long)h->nb_dot * (h->m_max + 1)))
        || (rc = halloc(&h->dots, 2 * (h->m_max + 1) + 2)) || (rc = halloc(&h->hdev, h->m_max + 1))
        || (rc = halloc(&h->arn, arnoldi_doubles(h->m_max))))
    {
        bx_helm_destroy(h);
        return rc;
    }
    if (hipHostMalloc((void **)&h->hstat, 8 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    {
        h->hstat = nullptr;
        bx_helm_destroy(h);
        return herr(BX_ERR_HIP, "bx_helm_create: hipHostMalloc failed");
    }
    *out = h;
    return BX_OK;
}

extern "C" void bx_helm_destroy(bx_helm *h)
{
    if (!h)
        return;
    h->mg64.free_all();
    h->mg32.free_all();
    for (const auto &pg : h->pgraphs)
        if (pg.exec)
            (void)hipGraphExecDestroy(pg.exec);
    if (h->cap_stream)
        (void)hipStreamDestroy(h->cap_stream);
    for (void *p : { (void *)h->a, (void *)h->bX, (void *)h->bY, (void *)h->cX,
                     (void *)h->cY, (void *)h->V, (void *)h->Z, (void *)h->w,
                     (void *)h->r, (void *)h->partial, (void *)h->dots,
                     (void *)h->hdev, (void *)h->tmp1, (void *)h->tmp2,
                     (void *)h->fr, (void *)h->fe,
                     (void *)h->arn })
        if (p)
            (void)hipFree(p);
    if (h->hstat)
        (void)hipHostFree(h->hstat);
    delete h;
}

extern "C" int bx_helm_set_coeffs(bx_helm *h, double alpha, double beta,
                                  double gamma, const double *a,
                                  const double *bXf, const double *bYf,
                                  const double *cX, const double *cY,
                                  void *stream)
{
    if (!h || !a || !bXf || !bYf || ((cX == nullptr) != (cY == nullptr)))
        return herr(BX_ERR_INVALID, "bx_helm_set_coeffs: bad argument");
    hipStream_t s = (hipStream_t)stream;
    h->alpha = alpha;
    h->beta  = beta;
    h->gamma = gamma;
    h->has_c = cX != nullptr && gamma != 0.0;
    int rc;
    if (h->has_c && !h->cX && ((rc = halloc(&h->cX, h->n)) || (rc = halloc(&h->cY, h->n))))
        return rc;
    // the FP64 hierarchy's level 0 is the solver's copy (h->a, ...): its
    // level-0 launch copies the caller's arrays there
    bool realloc = false;
    if ((rc = build_hierarchy(h, h->mg64, s, &realloc, a, bXf, bYf, cX, cY)))
        return rc;
    if (realloc)
        ++h->gen;
    h->mg32_stale = true;
    H_CHECK(hipStreamSynchronize(s));
    h->coeffs_set = true;
    return BX_OK;
}

// Builds the FP32 hierarchy if a preconditioner will read it and it is
// older than the coefficients.  Never called inside a graph capture.
static int ensure_mg32(bx_helm *h, hipStream_t s)
{
    if (!h->mg32_stale || !h->precond_fp32)
        return BX_OK;
    bool realloc = false;
    int  rc;
    if ((rc = build_hierarchy(h, h->mg32, s, &realloc, h->a, h->bX, h->bY, h->cX, h->cY)))
        return rc;
    if (realloc)
        ++h->gen;
    h->mg32_stale = false;
    return BX_OK;
}

extern "C" int bx_helm_set_options(bx_helm *h, int nu1, int nu2, int nu_bottom,
                                   int precond_fp32, int restart)
{
    if (!h || nu1 < 0 || nu2 < 0 || nu_bottom < 1 || restart < 1 || restart > h->m_max)
        return herr(BX_ERR_INVALID, "bx_helm_set_options: bad argument");
    for (auto *mg : { &h->mg64 })
    {
        mg->nu1       = nu1;
        mg->nu2       = nu2;
        mg->nu_bottom = nu_bottom;
    }
    h->mg32.nu1       = nu1;
    h->mg32.nu2       = nu2;
    h->mg32.nu_bottom = nu_bottom;
    h->precond_fp32   = precond_fp32;
    h->m              = restart;
    ++h->gen;
    return BX_OK;
}

extern "C" int bx_helm_set_fused(bx_helm *h, int fused)
{
    if (!h)
        return herr(BX_ERR_INVALID, "bx_helm_set_fused: null solver");
    h->fused = fused ? 1 : 0;
    ++h->gen;
    return BX_OK;
}

extern "C" int bx_helm_set_graph(bx_helm *h, int use_graph)
{
    if (!h)
        return herr(BX_ERR_INVALID, "bx_helm_set_graph: null solver");
    h->use_graph = use_graph ? 1 : 0;
    return BX_OK;
}

extern "C" int bx_helm_apply(bx_helm *h, const double *x, double *y,
                             void *stream)
{
    if (!h || !h->coeffs_set || !x || !y)
        return herr(BX_ERR_STATE, "bx_helm_apply: coefficients not set");
    launch_apply(h->mg64.lev[0], h->mg64, x, y, (const double *)nullptr,
                 (hipStream_t)stream);
    H_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_helm_residual(bx_helm *h, const double *b, const double *x,
                                double *r, void *stream)
{
    if (!h || !h->coeffs_set || !b || !x || !r)
        return herr(BX_ERR_STATE, "bx_helm_residual: coefficients not set");
    residual64(h, b, x, r, (hipStream_t)stream);
    H_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_helm_smooth(bx_helm *h, double *phi, const double *rhs,
                              int redblack, void *stream)
{
    if (!h || !h->coeffs_set || !phi || !rhs)
        return herr(BX_ERR_STATE, "bx_helm_smooth: coefficients not set");
    launch_smooth(h->mg64.lev[0], h->mg64, phi, rhs, redblack & 1,
                  (hipStream_t)stream);
    H_CHECK(hipGetLastError());
    return BX_OK;
}

// e = M r : one V-cycle from a zero guess (FP32 or FP64), optionally with the
// learned correction as fine-level smoother:
//   e  = omega * CNN(r)                 (learned pre-smoothing)
//   e += P Vcycle_coarse(R (r - A e))
//   e += omega * CNN(r - A e)           (learned post-smoothing)
static int precond_vcycle(bx_helm *h, const double *rv, double *e, hipStream_t s);

static int precond(bx_helm *h, const double *rv, double *e, hipStream_t s)
{
    const long long n = h->n;
    if (h->cnn && h->cnn_mode == BX_CNN_CORRECTOR)
    {
        // e = V(r); e += omega * CNN(r - A e)
        int rc;
        if ((rc = precond_vcycle(h, rv, e, s)))
            return rc;
        residual64(h, rv, e, h->tmp1, s);
        if ((rc = bx_cnn_forward(h->cnn, h->tmp1, h->tmp2, h->nx, h->ny,
                                 h->cnn_omega, s)))
            return rc;
        hipLaunchKernelGGL(k_axpy, dim3(2048), dim3(256), 0, s, e, h->tmp2, n, 1.0);
        H_CHECK(hipGetLastError());
        return BX_OK;
    }
    if (h->cnn)
    {
        int rc;
        // pre: e = omega*CNN(r)
        if ((rc = bx_cnn_forward(h->cnn, rv, e, h->nx, h->ny, h->cnn_omega, s)))
            return rc;
        // residual r1 = r - A e (FP64), coarse correction in FP32 or FP64
        residual64(h, rv, e, h->tmp1, s);
        if (h->mg64.lev.size() < 2)
        {
            // single-level hierarchy: no coarse correction
        }
        else if (h->precond_fp32)
        {
            MGHierarchy<float> &mg = h->mg32;
            hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0,
                               s, h->tmp1, mg.lev[0].res, n, 1.0);
            hipLaunchKernelGGL((k_fill<float>), dim3(2048), dim3(256), 0, s,
                               mg.lev[0].x, n, 0.f);
            coarse_correction(mg, s, h->fused);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(2048), dim3(256), 0,
                               s, mg.lev[0].x, h->tmp2, n, 1.0f);
        }
        else
        {
            MGHierarchy<double> &mg = h->mg64;
            hipLaunchKernelGGL((k_scale), dim3(2048), dim3(256), 0, s,
                               mg.lev[0].res, h->tmp1, n, 1.0);
            hipLaunchKernelGGL((k_fill<double>), dim3(2048), dim3(256), 0, s,
                               mg.lev[0].x, n, 0.0);
            coarse_correction(mg, s, h->fused);
            hipLaunchKernelGGL((k_scale), dim3(2048), dim3(256), 0, s, h->tmp2,
                               mg.lev[0].x, n, 1.0);
        }
        if (h->mg64.lev.size() >= 2)
            hipLaunchKernelGGL(k_axpy, dim3(2048), dim3(256), 0, s, e, h->tmp2, n,
                               1.0);
        // post: e += omega*CNN(r - A e)
        residual64(h, rv, e, h->tmp1, s);
        if ((rc = bx_cnn_forward(h->cnn, h->tmp1, h->tmp2, h->nx, h->ny,
                                 h->cnn_omega, s)))
            return rc;
        hipLaunchKernelGGL(k_axpy, dim3(2048), dim3(256), 0, s, e, h->tmp2, n, 1.0);
        H_CHECK(hipGetLastError());
        return BX_OK;
    }
    return precond_vcycle(h, rv, e, s);
}

// e = one V-cycle on r from a zero guess, FP32 or FP64 hierarchy
static int precond_vcycle(bx_helm *h, const double *rv, double *e, hipStream_t s)
{
    const long long n = h->n;
    // where the finest level's smoothing launches can read r and write e
    // themselves, no conversion launches
    if (h->precond_fp32)
    {
        MGHierarchy<float> &mg = h->mg32;
        if (direct_io_level(mg, 0, h->fused != 0))
            vcycle(mg, 0, s, true, (float *)nullptr, true, rv, e);
        else
        {
            hipLaunchKernelGGL((k_convert<float, double>), dim3(2048), dim3(256), 0, s,
                               rv, mg.lev[0].rhs, n, 1.0);
            vcycle(mg, 0, s, h->fused != 0, (float *)nullptr, true);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(2048), dim3(256), 0, s,
                               mg.lev[0].x, e, n, 1.0f);
        }
    }
    else
    {
        MGHierarchy<double> &mg = h->mg64;
        if (direct_io_level(mg, 0, h->fused != 0))
            vcycle(mg, 0, s, true, (double *)nullptr, true, rv, e);
        else
        {
            hipLaunchKernelGGL((k_scale), dim3(2048), dim3(256), 0, s, mg.lev[0].rhs,
                               rv, n, 1.0);
            vcycle(mg, 0, s, h->fused != 0, (double *)nullptr, true);
            hipLaunchKernelGGL((k_scale), dim3(2048), dim3(256), 0, s, e,
                               mg.lev[0].x, n, 1.0);
        }
    }
    H_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_helm_precond(bx_helm *h, const double *r, double *e,
                               void *stream)
{
    if (!h || !h->coeffs_set || !r || !e)
        return herr(BX_ERR_STATE, "bx_helm_precond: coefficients not set");
    int rc;
    if ((rc = ensure_mg32(h, (hipStream_t)stream)))
        return rc;
    return precond(h, r, e, (hipStream_t)stream);
}

extern "C" int bx_helm_attach_cnn(bx_helm *h, bx_cnn *cnn, double omega, int mode)
{
    if (!h || (mode != BX_CNN_SMOOTHER && mode != BX_CNN_CORRECTOR))
        return herr(BX_ERR_INVALID, "bx_helm_attach_cnn: bad argument");
    h->cnn       = cnn;
    h->cnn_omega = omega;
    h->cnn_mode  = mode;
    ++h->gen;
    return BX_OK;
}

// e = M r through the captured graph for (r, e)
static int precond_replay(bx_helm *h, const double *rv, double *e, hipStream_t s)
{
    if (!h->use_graph)
        return precond(h, rv, e, s);
    const int       slot  = h->cnn ? 1 : 0;
    const long long epoch = h->cnn ? bx_cnn_epoch(h->cnn) : 0;
    bx_helm::PGraph *pg   = nullptr;
    for (auto &q : h->pgraphs)
        if (q.slot == slot && q.in == rv && q.out == e)
            pg = &q;
    if (!pg)
    {
        h->pgraphs.push_back({ slot, rv, e, nullptr, -1, -1 });
        pg = &h->pgraphs.back();
    }
    if (!pg->exec || pg->gen != h->gen || pg->epoch != epoch)
    {
        if (pg->exec)
        {
            H_CHECK(hipGraphExecDestroy(pg->exec));
            pg->exec = nullptr;
        }
        if (!h->cap_stream)
            H_CHECK(hipStreamCreateWithFlags(&h->cap_stream, hipStreamNonBlocking));
        // everything the capture touches must be allocated beforehand
        if (h->cnn)
        {
            int rc = bx_cnn_reserve(h->cnn, h->n);
            if (rc)
                return rc;
        }
        H_CHECK(hipStreamSynchronize(s));
        hipGraph_t g = nullptr;
        H_CHECK(hipStreamBeginCapture(h->cap_stream, hipStreamCaptureModeThreadLocal));
        int rc = precond(h, rv, e, h->cap_stream);
        hipError_t ec = hipStreamEndCapture(h->cap_stream, &g);
        if (rc)
        {
            if (g)
                (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess)
            return herr(BX_ERR_HIP, std::string("precond capture: ") + hipGetErrorString(ec));
        ec = hipGraphInstantiate(&pg->exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ec != hipSuccess)
        {
            pg->exec = nullptr;
            return herr(BX_ERR_HIP, std::string("precond instantiate: ") + hipGetErrorString(ec));
        }
        pg->gen   = h->gen;
        pg->epoch = epoch;
    }
    H_CHECK(hipGraphLaunch(pg->exec, s));
    return BX_OK;
}

// b - A x into the finest level's right-hand side (the V-cycle only reads
// it), with the partial sums of its squares reduced into dots[0] (the bits
// of residual64 + norm2's kernels).
static void mg_residual(bx_helm *h, const double *b, const double *x, hipStream_t s)
{
    MGHierarchy<double> &mg = h->mg64;
    MGLevel<double>     &L0 = mg.lev[0];
    hipLaunchKernelGGL(k_residual_sumsq, dim3(h->nb_dot), dim3(256), 0, s, x, L0.rhs, b, L0.a,
                       L0.bX, L0.bY, L0.cX, L0.cY, L0.nx, L0.ny, mg.d_sc, L0.dxinv, L0.dyinv,
                       h->partial);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, h->partial, h->nb_dot, 1,
                       h->dots);
}

// One multigrid iteration on the device: x += V(r) with r in the finest
// level's right-hand side, then the new residual there and its squared
// norm in dots[0].
static void mg_iteration(bx_helm *h, const double *b, double *x, bool fused, hipStream_t s)
{
    MGHierarchy<double> &mg = h->mg64;
    MGLevel<double>     &L0 = mg.lev[0];
    const long long      n  = h->n;
    if (!vcycle(mg, 0, s, fused, x, true))
        hipLaunchKernelGGL(k_axpy, dim3(2048), dim3(256), 0, s, x, L0.x, n, 1.0);
    mg_residual(h, b, x, s);
}

static int read_rnorm(bx_helm *h, hipStream_t s, double *rn)
{
    H_CHECK(hipGetLastError());
    H_CHECK(hipMemcpyAsync(h->hstat, h->dots, sizeof(double), hipMemcpyDeviceToHost, s));
    H_CHECK(hipStreamSynchronize(s));
    *rn = std::sqrt(h->hstat[0]);
    return BX_OK;
}

// Plain multigrid solve (MLMG::solve semantics, FP64): x <- x + V(b - A x)
// until ||b - A x||_2 <= rtol ||b||_2.  The iterations are launched kernel by
// kernel: at 1024^2 the GPU is the bottleneck, and a captured iteration
// (which needs copies of b and x into fixed buffers) measured slower.  The
// residual lives in the finest level's right-hand side.
// rn0 >= 0: h->r already holds b - A x, of norm rn0.
static int solve_mg(bx_helm *h, const double *b, double *x, double rtol,
                    int maxiter, double bnorm, bx_solve_stats *st,
                    hipStream_t s, bool fused = false, double rn0 = -1)
{
    MGLevel<double> &L0 = h->mg64.lev[0];
    double           rn = rn0;
    int              rc;
    if (rn0 < 0)
    {
        mg_residual(h, b, x, s);
        if ((rc = read_rnorm(h, s, &rn)))
            return rc;
    }
    else if (rn > rtol * bnorm && maxiter > 0)
        hipLaunchKernelGGL((k_scale), dim3(2048), dim3(256), 0, s, L0.rhs, h->r, h->n, 1.0);
    int it = 0;
    while (rn > rtol * bnorm && it < maxiter)
    {
        mg_iteration(h, b, x, fused, s);
        if ((rc = read_rnorm(h, s, &rn)))
            return rc;
        ++it;
        ++st->precond_applies;
    }
    st->iters += it;
    st->rel_resid = rn / bnorm;
    st->converged = rn <= rtol * bnorm;
    return BX_OK;
}

// Flexible GMRES(m), right preconditioned, FP64 vectors and dot products,
// classical Gram-Schmidt applied twice.  The Hessenberg matrix, the Givens
// rotations and the triangular solve stay on the device; the host reads two
// status words per iteration (the residual estimate and the new column's
// norm), the one synchronisation an iteration needs to decide whether to
// go on.  Stall detection: if the residual has not dropped by 2x over 8
// iterations, or the iteration cap is reached, the caller falls back to
// plain multigrid.  rn0 >= 0: h->r already holds b - A x, of norm rn0.
static int solve_fgmres(bx_helm *h, const double *b, double *x, double rtol,
                        int maxiter, double bnorm, bx_solve_stats *st,
                        hipStream_t s, bool *stalled, double rn0 = -1)
{
    const long long n = h->n;
    const int       m = h->m;
    const Arnoldi   A = arnoldi_view(h->arn, m, h->m_max);
    *stalled          = false;
    int    total      = 0;
    double best       = 1e300;
    int    best_at    = 0;
    int    rc;
    bool   stop       = false; // a stall found inside the last cycle
    int    pending_j  = 0;     // directions added by the last cycle
    double beta       = rn0;
    while (true)
    {
        if (beta < 0)
        {
            residual64(h, b, x, h->r, s);
            if ((rc = norm2(h, h->r, s, &beta))) // also completes the status copy
                return rc;
        }
        if (pending_j > 0)
        {
            if (h->hstat[2] != 0) // non-finite y: the iterate was kept
                stop = true;
            else
                st->iters += pending_j;
            pending_j = 0;
        }
        st->rel_resid = beta / bnorm;
        if (beta <= rtol * bnorm)
        {
            st->converged = 1;
            break;
        }
        if (stop || total >= maxiter)
        {
            *stalled = true;
            break;
        }
        hipLaunchKernelGGL(k_scale, dim3(2048), dim3(256), 0, s, h->V, h->r, n,
                           1.0 / beta);
        hipLaunchKernelGGL(k_arnoldi_init, dim3(1), dim3(64), 0, s, A, beta);
        beta  = -1;
        int j = 0;
        for (; j < m && total < maxiter; ++j)
        {
            double *vj = h->V + (long long)j * n;
            double *zj = h->Z + (long long)j * n;
            if ((rc = precond_replay(h, vj, zj, s)))
                return rc;
            ++st->precond_applies;
            // CGS2 chained on the device (the projections are read from
            // device memory)
            double *d1 = h->dots, *d2 = h->dots + (m + 1), *d3 = h->dots + 2 * (m + 1);
            if (j + 1 <= 16)
            {
                // w = A z with the first projection's partial sums, then
                // each subtraction with the next pass's partial sums, the
                // norm's reduction with the Arnoldi step: 7 launches, not 11
                const MGLevel<double> &L0 = h->mg64.lev[0];
                hipLaunchKernelGGL(k_apply_multidot, dim3(h->nb_dot), dim3(256), 0, s, zj, h->w,
                                   L0.a, L0.bX, L0.bY, L0.cX, L0.cY, L0.nx, L0.ny, h->mg64.d_sc,
                                   L0.dxinv, L0.dyinv, h->V, j + 1, h->partial);
                hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, j + 1, d1);
                hipLaunchKernelGGL(k_subtract_dots, dim3(h->nb_dot), dim3(256), 0, s, h->w, h->V,
                                   d1, j + 1, n, 1, h->partial);
                hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, j + 1, d2);
                hipLaunchKernelGGL(k_subtract_dots, dim3(h->nb_dot), dim3(256), 0, s, h->w, h->V,
                                   d2, j + 1, n, 0, h->partial);
                hipLaunchKernelGGL(k_norm_arnoldi, dim3(1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, d3, A, d1, d2, j);
                hipLaunchKernelGGL(k_scale_inv_norm, dim3(2048), dim3(256), 0, s,
                                   h->V + (long long)(j + 1) * n, h->w, n, d3);
            }
            else
            {
                launch_apply(h->mg64.lev[0], h->mg64, zj, h->w,
                             (const double *)nullptr, s);
                hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, h->V, h->w, n,
                                   j + 1, h->partial);
                hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, j + 1, d1);
                hipLaunchKernelGGL(k_subtract_combo, dim3(2048), dim3(256), 0, s, h->w,
                                   h->V, d1, j + 1, n);
                hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, h->V, h->w, n,
                                   j + 1, h->partial);
                hipLaunchKernelGGL(k_reduce_partials, dim3(j + 1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, j + 1, d2);
                hipLaunchKernelGGL(k_subtract_combo, dim3(2048), dim3(256), 0, s, h->w,
                                   h->V, d2, j + 1, n);
                hipLaunchKernelGGL(k_multidot, dim3(h->nb_dot), dim3(256), 0, s, h->w, h->w, n,
                                   1, h->partial);
                hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, s, h->partial,
                                   h->nb_dot, 1, d3);
                hipLaunchKernelGGL(k_scale_inv_norm, dim3(2048), dim3(256), 0, s,
                                   h->V + (long long)(j + 1) * n, h->w, n, d3);
                hipLaunchKernelGGL(k_arnoldi_step, dim3(1), dim3(64), 0, s, A, d1, d2, d3, j);
            }
            H_CHECK(hipGetLastError());
            H_CHECK(hipMemcpyAsync(h->hstat, A.status, 2 * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
            H_CHECK(hipStreamSynchronize(s));
            const double est = h->hstat[0], hn = h->hstat[1];
            ++total;
            if (!std::isfinite(est) || !std::isfinite(hn))
            {
                // a non-finite preconditioner output: drop this direction
                stop = true;
                break;
            }
            if (est < best * 0.5)
            {
                best    = est;
                best_at = total;
            }
            if (est <= rtol * bnorm * 0.5 || hn == 0)
            {
                ++j;
                break;
            }
            if (total - best_at >= 8)
            {
                stop = true;
                ++j;
                break;
            }
        }
        // x += Z y, y from the triangular system of the first j columns
        if (j > 0)
        {
            hipLaunchKernelGGL(k_arnoldi_solve, dim3(1), dim3(64), 0, s, A, j);
            H_CHECK(hipMemcpyAsync(h->hstat + 2, A.status + 2, sizeof(double),
                                   hipMemcpyDeviceToHost, s));
            hipLaunchKernelGGL(k_add_combo, dim3(2048), dim3(256), 0, s, x, h->Z, A.y, j, n);
            H_CHECK(hipGetLastError());
            pending_j = j;
        }
    }
    return BX_OK;
}

extern "C" int bx_helm_solve(bx_helm *h, const double *b, double *x, int method,
                             double rtol, int maxiter, bx_solve_stats *st,
                             void *stream)
{
    if (!h || !h->coeffs_set || !b || !x || !st || !(rtol > 0) || maxiter < 1
        || method < BX_SOLVE_MG || method > BX_SOLVE_MG_FUSED)
        return herr(BX_ERR_INVALID, "bx_helm_solve: bad argument");
    if (method == BX_SOLVE_FGMRES_CNN && !h->cnn)
        return herr(BX_ERR_STATE, "bx_helm_solve: no network attached");
    hipStream_t s = (hipStream_t)stream;
    std::memset(st, 0, sizeof(*st));
    const auto t0 = std::chrono::steady_clock::now();
    double     bnorm = 0, rn0 = 0;
    int        rc;
    // ||b|| and the initial residual's norm with one synchronisation
    residual64(h, b, x, h->r, s);
    if ((rc = norm2_pair(h, b, h->r, s, &bnorm, &rn0)))
        return rc;
    if (bnorm == 0)
    {
        hipLaunchKernelGGL((k_fill<double>), dim3(2048), dim3(256), 0, s, x, h->n,
                           0.0);
        st->converged = 1;
        H_CHECK(hipStreamSynchronize(s));
        return BX_OK;
    }
    if (method == BX_SOLVE_MG || method == BX_SOLVE_MG_FUSED)
    {
        if ((rc = solve_mg(h, b, x, rtol, maxiter, bnorm, st, s,
                           method == BX_SOLVE_MG_FUSED, rn0)))
            return rc;
    }
    else
    {
        if ((rc = ensure_mg32(h, s)))
            return rc;
        bx_cnn *saved = h->cnn;
        if (method == BX_SOLVE_FGMRES_MG)
            h->cnn = nullptr;
        bool stalled = false;
        rc           = solve_fgmres(h, b, x, rtol, maxiter, bnorm, st, s, &stalled, rn0);
        h->cnn       = saved;
        if (rc)
            return rc;
        if (!st->converged)
        {
            // fall back to plain multigrid from the current iterate
            st->fell_back = 1;
            if ((rc = solve_mg(h, b, x, rtol, 1000, bnorm, st, s)))
                return rc;
        }
    }
    H_CHECK(hipStreamSynchronize(s));
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                      .count();
    return BX_OK;
}
