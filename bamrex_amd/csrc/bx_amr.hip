// Two AMR levels of the explicit path, refinement ratio 2, fine level
// subcycled: the coupling the reference gets from AMReX's Amr/AmrLevel/
// FluxRegister machinery around AmrLevelAdv.
//
//   coarse advance, then two fine advances of dt/2   Amr::timeStep (AMReX)
//   fine ghost cells at the coarse-fine boundary      FillPatcherFill
//     (AmrLevelAdv_advance.cpp:121-129): coarse data interpolated in time
//     (linear between the coarse old and new states) and space (conservative
//     linear, cell_cons_interp, AmrLevelAdv.cpp:171-175)
//   flux registers                                    _advance.cpp:61-81, 192-254
//     (fluxes scaled by dt and the face area; CrseInit with -1, FineAdd with +1)
//   reflux, avgDown after the fine advances           post_timestep
//     (AmrLevelAdv.cpp:628-652, 918-987)
//   dt over both levels                               computeNewDt (:589-618)
//
// Fine boxes are in fine index space, aligned to the ratio (even lo, odd
// hi), inside the fine domain; each coarse cell is either covered by fine
// cells or not.  The split sweeps run exactly as on a single level; the
// coarse-fine ghost cells are set once per fine step (on both ping-pong
// buffers) and, as in the reference, are not refreshed between the
// directions (fill_consv_boundary fills fine-fine and physical ghosts only).
// Edge and corner ghost cells at the coarse-fine boundary are not filled:
// the sweeps read face ghost cells only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "bx_level.h"

namespace bx {

// Fine ghost cells [lo, hi] of fine box dst (global fine indices) take the
// interpolated value of the coarse data around their image x + shift
// (shift: periodic wrap into the domain), coarse cells of box src.
struct CFSeg
{
    int dst, src;
    int lo[3], hi[3];
    int shift[3];
    int pad;
};

// The face on the low side of cell i (global index) of box `box`.
struct RegFace
{
    int box;
    int i[3];
};

// A coarse cell receiving reflux corrections terms[begin, end).
struct RefluxTarget
{
    int box;
    int i[3];
    int begin, end;
};

struct RefluxTerm
{
    int slot;
    int sign;
};

// Coarse cells [lo, hi] of box cbox covered by fine box fbox.
struct AvgSeg
{
    int fbox, cbox;
    int lo[3], hi[3];
};

__device__ __forceinline__ long long cell_off(const DevBox &b, int ng, int D, int i,
                                              int j, int k)
{
    const int ngy = (D >= 2) ? ng : 0;
    const int ngz = (D >= 3) ? ng : 0;
    return b.base + (i - b.lo[0] + ng) + (long long)(j - b.lo[1] + ngy) * b.jstride
           + (long long)(k - b.lo[2] + ngz) * b.kstride;
}

// Monotonised-central slope of the coarse data: min(|c|, |l|, |r|) with the
// sign of the central difference c, zero at an extremum.
__device__ __forceinline__ double cf_mc_slope(double um, double uc, double up)
{
    const double dc = 0.5 * (up - um);
    const double dl = 2 * (uc - um);
    const double dr = 2 * (up - uc);
    if (!(dl * dr > 0))
        return 0.0;
    const double m = fmin(fabs(dc), fmin(fabs(dl), fabs(dr)));
    return dc < 0 ? -m : m;
}

// Coarse-fine ghost fill.  One block per segment; threads over its cells.
// Time weights: mode 0 = the coarse old state, 1 = alpha old + beta new
// (from w[0..2] when w is given).
// Per component: MC slopes in each direction, scaled by the largest
// alpha <= 1 that keeps the four (2D) / eight (3D) children of the coarse
// cell within the minimum and maximum of its 3^D neighbourhood; the child
// value is u_c + sum_d off_d (alpha s_d), off_d = -1/4 or +1/4.
template <int D>
__global__ __launch_bounds__(256) void k_cf_fill(
    double *__restrict__ Uf, const DevBox *__restrict__ fboxes,
    const double *__restrict__ Uold, const double *__restrict__ Unew,
    const DevBox *__restrict__ cboxes, const CFSeg *__restrict__ segs, int nseg, int ngf,
    int ngc, int mode, double alpha, double beta, const double *__restrict__ w)
{
    constexpr int NC = D + 2;
    if (w) // mode, alpha, beta from device memory (a replayed step)
    {
        mode  = (int)w[0];
        alpha = w[1];
        beta  = w[2];
    }
    for (int si = blockIdx.x; si < nseg; si += gridDim.x)
    {
        const CFSeg  &g  = segs[si];
        const DevBox &fb = fboxes[g.dst];
        const DevBox &cb = cboxes[g.src];
        const int     ex = g.hi[0] - g.lo[0] + 1;
        const int     ey = g.hi[1] - g.lo[1] + 1;
        const int     ez = g.hi[2] - g.lo[2] + 1;
        const int     n  = ex * ey * ez;
        for (int c = threadIdx.x; c < n; c += blockDim.x)
        {
            int x[3];
            x[0] = g.lo[0] + c % ex;
            x[1] = g.lo[1] + (c / ex) % ey;
            x[2] = g.lo[2] + c / (ex * ey);
            int    ci[3] = { 0, 0, 0 };
            double off[3] = { 0, 0, 0 };
            for (int d = 0; d < D; ++d)
            {
                const int img = x[d] + g.shift[d];
                ci[d]         = img >> 1; // floor(img / 2)
                off[d]        = (img & 1) ? 0.25 : -0.25;
            }
            const long long doff = cell_off(fb, ngf, D, x[0], x[1], x[2]);
            const long long cc   = cell_off(cb, ngc, D, ci[0], ci[1], ci[2]);
            const long long sd[3] = { 1, cb.jstride, cb.kstride };
            for (int m = 0; m < NC; ++m)
            {
                auto val = [&](long long o) {
                    const long long a = o + m * cb.nstride;
                    return mode == 0 ? Uold[a] : alpha * Uold[a] + beta * Unew[a];
                };
                const double uc = val(cc);
                double       s[3] = { 0, 0, 0 };
                for (int d = 0; d < D; ++d)
                    s[d] = cf_mc_slope(val(cc - sd[d]), uc, val(cc + sd[d]));
                double umin = uc, umax = uc;
                for (int kk = (D >= 3 ? -1 : 0); kk <= (D >= 3 ? 1 : 0); ++kk)
                    for (int jj = (D >= 2 ? -1 : 0); jj <= (D >= 2 ? 1 : 0); ++jj)
                        for (int ii = -1; ii <= 1; ++ii)
                        {
                            const double v = val(cc + ii + jj * sd[1] + kk * sd[2]);
                            umin           = fmin(umin, v);
                            umax           = fmax(umax, v);
                        }
                double ext = 0.0;
                for (int d = 0; d < D; ++d)
                    ext = ext + 0.25 * fabs(s[d]);
                double lim = 1.0;
                if (ext > 0)
                    lim = fmin(1.0, fmin((umax - uc) / ext, (uc - umin) / ext));
                double r = uc;
                for (int d = 0; d < D; ++d)
                    r = r + off[d] * (lim * s[d]);
                Uf[doff + m * fb.nstride] = r;
            }
        }
    }
}

// Flux-register contributions of direction DIR: slot s gets the fluxes of
// faces[s*nsub .. s*nsub+nsub-1] scaled by `scale` (dt times the face
// area), each evaluated by the same face computation the sweeps use.
// accumulate = 0: reg = -F (CrseInit with -1, nsub = 1);
// accumulate = 1: reg += F for each face in order (FineAdd with +1).
// coef: when given, the coefficients come from device memory.
template <int D, int DIR, int LIM, int LV, int SCH>
__global__ __launch_bounds__(256) void k_reg_flux(
    const double *__restrict__ U, const DevBox *__restrict__ boxes,
    const RegFace *__restrict__ faces, int nslot, int nsub, double *__restrict__ reg,
    int accumulate, int ng, double hdtdx, double fc, double rc, double scale,
    const double *__restrict__ coef, EulerParm pm, LimTable lt)
{
    constexpr int NC = D + 2;
    const int     sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nslot)
        return;
    if (coef) // hdtdx, fc, rc, scale from device memory (a replayed step)
    {
        hdtdx = coef[0];
        fc    = coef[1];
        rc    = coef[2];
        scale = coef[3];
    }
    for (int q = 0; q < nsub; ++q)
    {
        const RegFace  &f  = faces[(long long)sl * nsub + q];
        const DevBox   &bx = boxes[f.box];
        const long long sd = (DIR == 0) ? 1 : (DIR == 1 ? bx.jstride : bx.kstride);
        const long long c  = cell_off(bx, ng, D, f.i[0], f.i[1], f.i[2]);
        double          W[4][NC];
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int n = 0; n < NC; ++n)
                W[s][n] = U[c + (s - 2) * sd + n * bx.nstride];
        double F[NC];
        if (!scheme_reconstructs(SCH))
        {
            face_flux<D, DIR, false, false, SCH>(W[1], W[2], F, pm, fc, rc);
        }
        else
        {
            double hL0[NC], hR0[NC], hL1[NC], hR1[NC];
            reconstruct_half_step<D, DIR, LIM, LV>(W[0], W[1], W[2], hL0, hR0, hdtdx, pm, lt);
            reconstruct_half_step<D, DIR, LIM, LV>(W[1], W[2], W[3], hL1, hR1, hdtdx, pm, lt);
            face_flux<D, DIR, false, false, SCH>(hR0, hL1, F, pm, fc, rc);
        }
#pragma unroll
        for (int n = 0; n < NC; ++n)
        {
            const double v = F[n] * scale;
            double      &r = reg[(long long)sl * NC + n];
            if (accumulate)
                r = r + v;
            else
                r = -1.0 * v;
        }
    }
}

// Reflux (FluxRegister::Reflux with scale 1): each coarse cell next to the
// fine region gets sign * reg / vol per register face, faces in the order
// (x lo, y lo, z lo, x hi, y hi, z hi); sign = -1 on the low faces of the
// fine boxes, +1 on the high faces.
template <int D>
__global__ __launch_bounds__(256) void k_reflux(double *__restrict__ U,
                                                const DevBox *__restrict__ boxes,
                                                const RefluxTarget *__restrict__ tg, int nt,
                                                const RefluxTerm *__restrict__ terms,
                                                const double *__restrict__ reg, int ng,
                                                double vol)
{
    constexpr int NC = D + 2;
    const int     t  = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt)
        return;
    const RefluxTarget &g  = tg[t];
    const DevBox       &bx = boxes[g.box];
    const long long     o  = cell_off(bx, ng, D, g.i[0], g.i[1], g.i[2]);
    for (int n = 0; n < NC; ++n)
    {
        double u = U[o + n * bx.nstride];
        for (int q = g.begin; q < g.end; ++q)
            u = u + (terms[q].sign * reg[(long long)terms[q].slot * NC + n]) / vol;
        U[o + n * bx.nstride] = u;
    }
}

// avgDown (average_down, AmrLevelAdv.cpp:961-987): a covered coarse cell
// becomes volfrac * (sum of its 2^D children), children summed with x
// fastest.  One block per segment.
template <int D>
__global__ __launch_bounds__(256) void k_avg_down(double *__restrict__ Uc,
                                                  const DevBox *__restrict__ cboxes,
                                                  const double *__restrict__ Uf,
                                                  const DevBox *__restrict__ fboxes,
                                                  const AvgSeg *__restrict__ segs, int nseg,
                                                  int ngc, int ngf, double volfrac)
{
    constexpr int NC = D + 2;
    for (int si = blockIdx.x; si < nseg; si += gridDim.x)
    {
        const AvgSeg &g  = segs[si];
        const DevBox &cb = cboxes[g.cbox];
        const DevBox &fb = fboxes[g.fbox];
        const int     ex = g.hi[0] - g.lo[0] + 1;
        const int     ey = g.hi[1] - g.lo[1] + 1;
        const int     ez = g.hi[2] - g.lo[2] + 1;
        const int     n  = ex * ey * ez;
        for (int c = threadIdx.x; c < n; c += blockDim.x)
        {
            const int       i  = g.lo[0] + c % ex;
            const int       j  = g.lo[1] + (c / ex) % ey;
            const int       k  = g.lo[2] + c / (ex * ey);
            const long long co = cell_off(cb, ngc, D, i, j, k);
            for (int m = 0; m < NC; ++m)
            {
                double s = 0;
                for (int kr = 0; kr < (D >= 3 ? 2 : 1); ++kr)
                    for (int jr = 0; jr < (D >= 2 ? 2 : 1); ++jr)
                        for (int ir = 0; ir < 2; ++ir)
                        {
                            const long long fo
                                = cell_off(fb, ngf, D, 2 * i + ir, D >= 2 ? 2 * j + jr : 0,
                                           D >= 3 ? 2 * k + kr : 0);
                            s = s + Uf[fo + m * fb.nstride];
                        }
                Uc[co + m * cb.nstride] = volfrac * s;
            }
        }
    }
}

// y = x, n doubles (the level copies of a step: kernels, so a captured step
// holds no copy nodes)
__global__ __launch_bounds__(256) void k_copy_f64(double *__restrict__ y,
                                                  const double *__restrict__ x, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i];
}

// ---- device-side step schedule (a captured coarse step, replayed) ------
// What the host computes before and during a coarse step from the time t
// and the levels' maximum wave speeds, computed on the device with the same
// expressions in the same order (so the same bits): dt of every level, the
// flux-register coefficients of every level and direction, and the time
// weights of every coarse-fine fill in call order.
constexpr int AMR_MAXLEV = 8;
constexpr int AMR_MAXCF  = 256;
struct AmrSched
{
    double t, dt0;
    double coef[AMR_MAXLEV][3][4]; // hdtdx, fc, rc, scale
    double cfw[AMR_MAXCF][3];      // mode, alpha, beta
};
struct SchedArgs
{
    const unsigned long long *maxbits[AMR_MAXLEV];
    double                   *dt[AMR_MAXLEV];
    double                    dx0[3];
    double                    cfl;
    int                       nlev, D;
};

// Amr::timeStep's recursion, in the same order (depth first, substep 0
// before substep 1) with an explicit stack: the fine levels' ghost-fill
// weights (two fills per call: the level's state and its ping-pong
// buffer), each level's dt and register coefficients
__device__ void sched_walk(AmrSched *S, const SchedArgs &a, double t, double dt_0)
{
    struct Call
    {
        int    l;
        double time, dt, t_par, dt_par;
    };
    Call st[2 * AMR_MAXLEV];
    int  sp = 0, slot = 0;
    st[sp++] = { 0, t, dt_0, t, dt_0 };
    while (sp > 0)
    {
        const Call c = st[--sp];
        if (c.l > 0)
        {
            const double t_old = c.t_par, t_new = c.t_par + c.dt_par;
            double       w[3];
            if (c.time == t_old || !(t_new != t_old))
            {
                w[0] = 0;
                w[1] = 1.0;
                w[2] = 0.0;
            }
            else
            {
                w[0] = 1;
                w[1] = (t_new - c.time) / (t_new - t_old);
                w[2] = (c.time - t_old) / (t_new - t_old);
            }
            for (int k = 0; k < 2; ++k, ++slot)
                for (int q = 0; q < 3; ++q)
                    S->cfw[slot][q] = w[q];
        }
        *a.dt[c.l] = c.dt;
        double dx[3] = { 0, 0, 0 };
        for (int d = 0; d < a.D; ++d)
            dx[d] = a.dx0[d] / (double)(1 << c.l);
        for (int d = 0; d < a.D; ++d)
        {
            double scale = c.dt;
            for (int e = 0; e < a.D; ++e)
                if (e != d)
                    scale *= dx[e];
            S->coef[c.l][d][0] = 0.5 * c.dt / dx[d];
            S->coef[c.l][d][1] = 0.5 * dx[d] / c.dt;
            S->coef[c.l][d][2] = c.dt / (2 * dx[d]);
            S->coef[c.l][d][3] = scale;
        }
        if (c.l + 1 < a.nlev)
        {
            const double dtf = c.dt / 2;
            st[sp++]         = { c.l + 1, c.time + dtf, dtf, c.time, c.dt }; // substep 1
            st[sp++]         = { c.l + 1, c.time, dtf, c.time, c.dt };       // substep 0
        }
    }
}

// computeNewDt without a stop time (dt_0 = min over levels of 2^l times the
// level's CFL estimate from its maximum wave speed), the step's schedule
// from S->t, then S->t += dt_0
__global__ void k_amr_schedule(AmrSched *S, SchedArgs a)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double dt_0     = 1.0e+100;
    int    n_factor = 1;
    for (int l = 0; l < a.nlev; ++l)
    {
        double dxl[3] = { 0, 0, 0 };
        for (int d = 0; d < a.D; ++d)
            dxl[d] = a.dx0[d] / (double)(1 << l);
        const double max_speed = __longlong_as_double((long long)*a.maxbits[l]);
        double       dt_est    = 1.0e20;
        for (int d = 0; d < a.D; ++d)
        {
            const double c = dxl[d] / max_speed;
            dt_est         = (c < dt_est) ? c : dt_est;
        }
        const double dmin = dt_est * a.cfl;
        n_factor *= (l == 0 ? 1 : 2);
        const double cand = n_factor * dmin;
        dt_0              = (cand < dt_0) ? cand : dt_0;
    }
    const double t = S->t;
    sched_walk(S, a, t, dt_0);
    S->dt0 = dt_0;
    S->t   = t + dt_0;
}

// ------------------------------------------------------ launch dispatch
template <int D, int DIR, int LIM, int LV, int SCH>
static hipError_t launch_reg(const double *U, const DevBox *boxes, const RegFace *faces,
                             int nslot, int nsub, double *reg, int acc, int ng,
                             double hdtdx, double fc, double rc, double scale,
                             const double *coef, EulerParm pm, LimTable lt, hipStream_t s)
{
    hipLaunchKernelGGL((k_reg_flux<D, DIR, LIM, LV, SCH>), dim3((nslot + 127) / 128),
                       dim3(128), 0, s, U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx,
                       fc, rc, scale, coef, pm, lt);
    return hipGetLastError();
}

template <int D, int DIR, int SCH>
static hipError_t reg_lim(int lim, int lv, const double *U, const DevBox *boxes,
                          const RegFace *faces, int nslot, int nsub, double *reg, int acc,
                          int ng, double hdtdx, double fc, double rc, double scale,
                          const double *coef, EulerParm pm, LimTable lt, hipStream_t s)
{
#define BX_RL(LIMV)                                                                 \
    do                                                                              \
    {                                                                               \
        if (lv == LV_SHARED)                                                        \
            return launch_reg<D, DIR, LIMV, LV_SHARED, SCH>(                        \
                U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx, fc, rc, scale, coef, pm, lt, s); \
        if (lv == LV_OWN)                                                           \
            return launch_reg<D, DIR, LIMV, LV_OWN, SCH>(                           \
                U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx, fc, rc, scale, coef, pm, lt, s); \
        return launch_reg<D, DIR, LIMV, LV_TABLE, SCH>(                             \
            U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx, fc, rc, scale, coef, pm, lt, s); \
    } while (0)
    if constexpr (!scheme_reconstructs(SCH))
        return launch_reg<D, DIR, LIM_MINBEE, LV_SHARED, SCH>(
            U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx, fc, rc, scale, coef, pm, lt, s);
    else
        switch (lim)
        {
        case LIM_MINBEE: BX_RL(LIM_MINBEE);
        case LIM_VAN_LEER: BX_RL(LIM_VAN_LEER);
        case LIM_VAN_ALBADA: BX_RL(LIM_VAN_ALBADA);
        default: BX_RL(LIM_SUPERBEE);
        }
#undef BX_RL
}

template <int D, int DIR>
static hipError_t reg_scheme(int sch, int lim, int lv, const double *U, const DevBox *boxes,
                             const RegFace *faces, int nslot, int nsub, double *reg, int acc,
                             int ng, double hdtdx, double fc, double rc, double scale,
                             const double *coef, EulerParm pm, LimTable lt, hipStream_t s)
{
#define BX_RS(S)                                                                        \
    return reg_lim<D, DIR, S>(lim, lv, U, boxes, faces, nslot, nsub, reg, acc, ng, hdtdx, \
                              fc, rc, scale, coef, pm, lt, s)
    switch (sch)
    {
    case SCH_MHM: BX_RS(SCH_MHM);
    case SCH_HLLC: BX_RS(SCH_HLLC);
    case SCH_SLIC: BX_RS(SCH_SLIC);
    case SCH_FORCE: BX_RS(SCH_FORCE);
    default: BX_RS(SCH_LF);
    }
#undef BX_RS
}

} // namespace bx

// ================================================================ host side
// The coupling of two adjacent levels: its plan and flux registers.
struct AmrPair
{
    bx_ctx *crse = nullptr, *fine = nullptr;
    int     D = 0, NC = 0;
    // coarse-fine ghost fill
    std::vector<CFSeg> cf;
    CFSeg             *d_cf = nullptr;
    // flux registers: slots of direction d are [slot0[d], slot0[d+1])
    int                  nslot = 0, slot0[4] = { 0, 0, 0, 0 };
    std::vector<RegFace> cface, fface; // 1 and 2^(D-1) faces per slot
    RegFace             *d_cface = nullptr, *d_fface = nullptr;
    double              *d_reg   = nullptr;
    std::vector<RefluxTarget> tg;
    std::vector<RefluxTerm>   terms;
    RefluxTarget             *d_tg    = nullptr;
    RefluxTerm               *d_terms = nullptr;
    std::vector<AvgSeg>       avg;
    AvgSeg                   *d_avg = nullptr;
};

template <class T>
static int amr_upload(T **dptr, const std::vector<T> &v)
{
    if (*dptr)
        (void)hipFree(*dptr);
    *dptr = nullptr;
    if (v.empty())
        return BX_OK;
    BX_HIP_CHECK(hipMalloc((void **)dptr, sizeof(T) * v.size()));
    BX_HIP_CHECK(hipMemcpy(*dptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return BX_OK;
}

static bool in_box(const bx_box &b, const int x[3], int D)
{
    for (int d = 0; d < D; ++d)
        if (x[d] < b.lo[d] || x[d] > b.hi[d])
            return false;
    return true;
}

static int find_box(const std::vector<bx_box> &B, const int x[3], int D)
{
    for (size_t i = 0; i < B.size(); ++i)
        if (in_box(B[i], x, D))
            return (int)i;
    return -1;
}

// Wraps x into the domain along periodic directions; false if x lies
// outside the domain along a non-periodic direction.
static bool wrap_into(const bx_geometry &g, int D, int x[3])
{
    for (int d = 0; d < D; ++d)
    {
        const int L = g.domain_hi[d] - g.domain_lo[d] + 1;
        if (x[d] < g.domain_lo[d] || x[d] > g.domain_hi[d])
        {
            if (g.bc_lo[d] != BX_BC_PERIODIC)
                return false;
            x[d] = x[d] < g.domain_lo[d] ? x[d] + L : x[d] - L;
        }
    }
    return true;
}

static int amr_plan(AmrPair *a)
{
    const bx_ctx      *C = a->crse, *F = a->fine;
    const int          D = a->D;
    const bx_geometry &gc = C->geom, &gf = F->geom;
    const auto        &CB = C->hboxes, &FB = F->hboxes;
    auto               odd_cell = [](int v) { return (v & 1) != 0; };
    // ---- validation
    for (int d = 0; d < D; ++d)
    {
        if (gf.domain_lo[d] != 2 * gc.domain_lo[d] || gf.domain_hi[d] != 2 * gc.domain_hi[d] + 1
            || gf.bc_lo[d] != gc.bc_lo[d] || gf.bc_hi[d] != gc.bc_hi[d])
            return set_error(BX_ERR_INVALID,
                             "bx_amr_create: the fine geometry must be the coarse one refined by 2");
    }
    for (const bx_box &b : FB)
        for (int d = 0; d < D; ++d)
            if (odd_cell(b.lo[d]) || !odd_cell(b.hi[d]) || b.lo[d] < gf.domain_lo[d]
                || b.hi[d] > gf.domain_hi[d])
                return set_error(BX_ERR_INVALID,
                                 "bx_amr_create: fine boxes must be ratio-2 aligned and inside the domain");
    if (gc.ng < 2 || gf.ng < 2)
        return set_error(BX_ERR_INVALID, "bx_amr_create: both levels need 2 ghost cells");
    // proper nesting: the coarse-fine interpolation reads the coarse cells
    // within 2 of each fine box (coarsened), which must be coarse valid
    // cells or coarse ghosts filled from other coarse boxes
    for (const bx_box &b : FB)
    {
        int lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
        for (int d = 0; d < D; ++d)
        {
            lo[d] = (b.lo[d] >> 1) - 2;
            hi[d] = (b.hi[d] >> 1) + 2;
        }
        for (int z = lo[2]; z <= hi[2]; ++z)
            for (int y = lo[1]; y <= hi[1]; ++y)
                for (int x = lo[0]; x <= hi[0]; ++x)
                {
                    int c[3] = { x, y, z };
                    if (wrap_into(gc, D, c) && find_box(CB, c, D) < 0)
                        return set_error(BX_ERR_INVALID,
                                         "bx_amr_create: the fine level is not properly nested "
                                         "(2 coarse cells around every fine box)");
                }
    }
    for (size_t i = 0; i < FB.size(); ++i)
        for (size_t j = i + 1; j < FB.size(); ++j)
        {
            bool overlap = true;
            for (int d = 0; d < D; ++d)
                overlap = overlap && FB[i].lo[d] <= FB[j].hi[d] && FB[j].lo[d] <= FB[i].hi[d];
            if (overlap)
                return set_error(BX_ERR_INVALID, "bx_amr_create: fine boxes overlap");
        }

    // ---- coarse-fine ghost cells: face ghost slabs of each fine box, cell
    // by cell, those whose (periodically wrapped) image is inside the domain
    // and not covered by a fine box; runs along x become segments
    a->cf.clear();
    const int ngf = gf.ng;
    for (int ib = 0; ib < (int)FB.size(); ++ib)
    {
        const bx_box &b = FB[ib];
        for (int d = 0; d < D; ++d)
            for (int side = 0; side < 2; ++side)
            {
                int lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
                for (int e = 0; e < D; ++e)
                {
                    lo[e] = b.lo[e];
                    hi[e] = b.hi[e];
                }
                if (side == 0)
                {
                    lo[d] = b.lo[d] - ngf;
                    hi[d] = b.lo[d] - 1;
                }
                else
                {
                    lo[d] = b.hi[d] + 1;
                    hi[d] = b.hi[d] + ngf;
                }
                for (int z = lo[2]; z <= hi[2]; ++z)
                    for (int y = lo[1]; y <= hi[1]; ++y)
                    {
                        CFSeg cur{};
                        bool  open = false;
                        for (int x = lo[0]; x <= hi[0]; ++x)
                        {
                            int p[3] = { x, y, z }, img[3] = { x, y, z };
                            bool cf = wrap_into(gf, D, img) && find_box(FB, img, D) < 0;
                            int  src = -1, sh[3] = { img[0] - p[0], img[1] - p[1], img[2] - p[2] };
                            if (cf)
                            {
                                int cc[3] = { img[0] >> 1, img[1] >> 1, img[2] >> 1 };
                                src       = find_box(CB, cc, D);
                                if (src < 0)
                                    return set_error(BX_ERR_INVALID,
                                                     "bx_amr_create: coarse level does not cover the domain");
                            }
                            const bool extend = open && cf && src == cur.src
                                                && sh[0] == cur.shift[0] && sh[1] == cur.shift[1]
                                                && sh[2] == cur.shift[2] && x == cur.hi[0] + 1;
                            if (extend)
                            {
                                cur.hi[0] = x;
                                continue;
                            }
                            if (open)
                                a->cf.push_back(cur);
                            open = cf;
                            if (cf)
                            {
                                cur       = CFSeg{};
                                cur.dst   = ib;
                                cur.src   = src;
                                cur.lo[0] = cur.hi[0] = x;
                                cur.lo[1] = cur.hi[1] = y;
                                cur.lo[2] = cur.hi[2] = z;
                                for (int e = 0; e < 3; ++e)
                                    cur.shift[e] = sh[e];
                            }
                        }
                        if (open)
                            a->cf.push_back(cur);
                    }
            }
    }

    // merge runs of consecutive rows (then planes) with the same x range,
    // boxes and shift into rectangles: fewer, larger segments
    for (int dim_merge = 1; dim_merge < 3; ++dim_merge)
    {
        std::vector<CFSeg> merged;
        for (const CFSeg &g : a->cf)
        {
            if (!merged.empty())
            {
                CFSeg &m    = merged.back();
                bool   same = m.dst == g.dst && m.src == g.src;
                for (int e = 0; e < 3 && same; ++e)
                {
                    same = same && m.shift[e] == g.shift[e];
                    if (e == dim_merge)
                        same = same && g.lo[e] == m.hi[e] + 1 && g.hi[e] == g.lo[e];
                    else if (e < dim_merge)
                        same = same && m.lo[e] == g.lo[e] && m.hi[e] == g.hi[e];
                    else
                        same = same && m.lo[e] == g.lo[e] && m.hi[e] == g.hi[e];
                }
                if (same)
                {
                    m.hi[dim_merge] = g.hi[dim_merge];
                    continue;
                }
            }
            merged.push_back(g);
        }
        a->cf.swap(merged);
    }

    // split large rectangles along their longest direction so that a
    // block (one segment) has at most CF_MAX_CELLS cells
    {
        const long long    CF_MAX_CELLS = 512;
        std::vector<CFSeg> split;
        for (const CFSeg &g : a->cf)
        {
            long long n = 1;
            int       e_long = 0, len_long = 0;
            for (int e = 0; e < 3; ++e)
            {
                const int len = g.hi[e] - g.lo[e] + 1;
                n *= len;
                if (len > len_long)
                {
                    len_long = len;
                    e_long   = e;
                }
            }
            const long long per_unit = n / len_long; // cells per unit of the long side
            const int       step     = (int)std::max<long long>(1, CF_MAX_CELLS / per_unit);
            for (int a0 = g.lo[e_long]; a0 <= g.hi[e_long]; a0 += step)
            {
                CFSeg h       = g;
                h.lo[e_long]  = a0;
                h.hi[e_long]  = std::min(g.hi[e_long], a0 + step - 1);
                split.push_back(h);
            }
        }
        a->cf.swap(split);
    }

    // ---- flux registers: per direction, per fine box side, per coarse
    // face of that side whose outside neighbour is not fine
    a->cface.clear();
    a->fface.clear();
    struct Tgt
    {
        int box, i[3], orient;
        RefluxTerm t;
    };
    std::vector<Tgt> tgts;
    const int nsub = 1 << (D - 1);
    a->nslot       = 0;
    for (int d = 0; d < D; ++d)
    {
        a->slot0[d] = a->nslot;
        for (int side = 0; side < 2; ++side)
            for (int ib = 0; ib < (int)FB.size(); ++ib)
            {
                const bx_box &b = FB[ib];
                int           clo[3] = { 0, 0, 0 }, chi[3] = { 0, 0, 0 };
                for (int e = 0; e < D; ++e)
                {
                    clo[e] = b.lo[e] >> 1;
                    chi[e] = b.hi[e] >> 1;
                }
                const int Ff = side == 0 ? b.lo[d] : b.hi[d] + 1; // fine face
                const int Fc = Ff >> 1;                           // coarse face
                clo[d] = chi[d] = 0;
                for (int z = clo[2]; z <= chi[2]; ++z)
                    for (int y = clo[1]; y <= chi[1]; ++y)
                        for (int x = clo[0]; x <= chi[0]; ++x)
                        {
                            int t[3] = { x, y, z };
                            // outside coarse cell and one of its fine children
                            int oc[3] = { t[0], t[1], t[2] };
                            oc[d]     = side == 0 ? Fc - 1 : Fc;
                            int ofine[3] = { 2 * oc[0], 2 * oc[1], 2 * oc[2] };
                            if (!wrap_into(gf, D, ofine) || find_box(FB, ofine, D) >= 0)
                                continue;
                            int ocw[3] = { oc[0], oc[1], oc[2] };
                            wrap_into(gc, D, ocw);
                            // coarse face: high-side cell index Fc, evaluated in the
                            // coarse box holding the inside cell
                            int inside[3] = { t[0], t[1], t[2] };
                            inside[d]     = side == 0 ? Fc : Fc - 1;
                            const int cbox = find_box(CB, inside, D);
                            const int tbox = find_box(CB, ocw, D);
                            if (cbox < 0 || tbox < 0)
                                return set_error(BX_ERR_INVALID,
                                                 "bx_amr_create: coarse level does not cover the domain");
                            RegFace cf{ cbox, { t[0], t[1], t[2] } };
                            cf.i[d] = Fc;
                            a->cface.push_back(cf);
                            // fine faces: children in the transverse directions, x then y then z
                            for (int q = 0; q < nsub; ++q)
                            {
                                RegFace ff{ ib, { 0, 0, 0 } };
                                int     bit = 0;
                                for (int e = 0; e < D; ++e)
                                {
                                    if (e == d)
                                    {
                                        ff.i[e] = Ff;
                                        continue;
                                    }
                                    ff.i[e] = 2 * t[e] + ((q >> bit) & 1);
                                    ++bit;
                                }
                                a->fface.push_back(ff);
                            }
                            Tgt tt;
                            tt.box = tbox;
                            for (int e = 0; e < 3; ++e)
                                tt.i[e] = ocw[e];
                            tt.orient = side * 3 + d; // x lo, y lo, z lo, x hi, ...
                            tt.t      = RefluxTerm{ a->nslot, side == 0 ? -1 : 1 };
                            tgts.push_back(tt);
                            ++a->nslot;
                        }
            }
    }
    a->slot0[D] = a->nslot;
    // group the reflux terms by target cell, orientations in order
    std::stable_sort(tgts.begin(), tgts.end(), [](const Tgt &p, const Tgt &q) {
        if (p.box != q.box)
            return p.box < q.box;
        for (int e = 2; e >= 0; --e)
            if (p.i[e] != q.i[e])
                return p.i[e] < q.i[e];
        return p.orient < q.orient;
    });
    a->tg.clear();
    a->terms.clear();
    for (size_t k = 0; k < tgts.size(); ++k)
    {
        const bool same = !a->tg.empty() && a->tg.back().box == tgts[k].box
                          && a->tg.back().i[0] == tgts[k].i[0] && a->tg.back().i[1] == tgts[k].i[1]
                          && a->tg.back().i[2] == tgts[k].i[2];
        if (!same)
            a->tg.push_back(RefluxTarget{ tgts[k].box, { tgts[k].i[0], tgts[k].i[1], tgts[k].i[2] },
                                          (int)a->terms.size(), (int)a->terms.size() });
        a->terms.push_back(tgts[k].t);
        a->tg.back().end = (int)a->terms.size();
    }

    // ---- avgDown: coarsened fine boxes intersected with the coarse boxes
    a->avg.clear();
    for (int ib = 0; ib < (int)FB.size(); ++ib)
        for (int jb = 0; jb < (int)CB.size(); ++jb)
        {
            AvgSeg s{ ib, jb, { 0, 0, 0 }, { 0, 0, 0 } };
            bool   ok = true;
            for (int e = 0; e < D && ok; ++e)
            {
                s.lo[e] = std::max(FB[ib].lo[e] >> 1, CB[jb].lo[e]);
                s.hi[e] = std::min(FB[ib].hi[e] >> 1, CB[jb].hi[e]);
                ok      = s.lo[e] <= s.hi[e];
            }
            if (ok)
                a->avg.push_back(s);
        }

    int rc;
    if ((rc = amr_upload(&a->d_cf, a->cf)) || (rc = amr_upload(&a->d_cface, a->cface))
        || (rc = amr_upload(&a->d_fface, a->fface)) || (rc = amr_upload(&a->d_tg, a->tg))
        || (rc = amr_upload(&a->d_terms, a->terms)) || (rc = amr_upload(&a->d_avg, a->avg)))
        return rc;
    if (a->d_reg)
        (void)hipFree(a->d_reg);
    a->d_reg = nullptr;
    if (a->nslot)
        BX_HIP_CHECK(hipMalloc((void **)&a->d_reg, sizeof(double) * a->nslot * a->NC));
    return BX_OK;
}

static void pair_destroy(AmrPair *a)
{
    if (!a)
        return;
    for (void *p : { (void *)a->d_cf, (void *)a->d_cface, (void *)a->d_fface, (void *)a->d_reg,
                     (void *)a->d_tg, (void *)a->d_terms, (void *)a->d_avg })
        if (p)
            (void)hipFree(p);
    delete a;
}

// A hierarchy of levels 0..nlev-1 (ratio 2 between neighbours) and the
// couplings of adjacent levels.
struct bx_amr
{
    int                    D = 0, NC = 0;
    std::vector<bx_ctx *>  lev;
    std::vector<AmrPair *> pairs; // pairs[l] couples lev[l] and lev[l+1]
    // the coarse step captured as a hipGraph (runs of a fixed number of
    // steps): its device schedule, capture stream, executable and what it
    // bakes in (buffers, cell sizes, CFL, each level's boxes and tuning)
    int                 use_graph = 1;
    AmrSched           *d_sched   = nullptr;
    hipStream_t         cap       = nullptr;
    hipGraphExec_t      gx[2]     = { nullptr, nullptr }; // even / odd steps
    std::vector<uint64_t> gkey;
};

extern "C" void bx_amr_destroy(bx_amr *a)
{
    if (!a)
        return;
    for (AmrPair *p : a->pairs)
        pair_destroy(p);
    for (hipGraphExec_t g : a->gx)
        if (g)
            (void)hipGraphExecDestroy(g);
    if (a->cap)
        (void)hipStreamDestroy(a->cap);
    if (a->d_sched)
        (void)hipFree(a->d_sched);
    delete a;
}

extern "C" int bx_amr_set_graph(bx_amr *a, int use_graph)
{
    if (!a)
        return set_error(BX_ERR_INVALID, "bx_amr_set_graph: null hierarchy");
    a->use_graph = use_graph ? 1 : 0;
    return BX_OK;
}

static bool same_params(const bx_params &pc, const bx_params &pf)
{
    bool same = pc.dim == pf.dim && pc.num_method == pf.num_method
                && pc.adiabatic == pf.adiabatic && pc.epsilon == pf.epsilon
                && pc.limiter == pf.limiter;
    for (int n = 0; n < 5; ++n)
        same = same && pc.limiter_indices[n] == pf.limiter_indices[n]
               && pc.limit_on_primv[n] == pf.limit_on_primv[n];
    return same;
}

extern "C" int bx_amr_create_levels(bx_ctx *const *levels, int nlev, bx_amr **out)
{
    if (!levels || nlev < 2 || nlev > 16 || !out)
        return set_error(BX_ERR_INVALID, "bx_amr_create: bad argument");
    for (int l = 0; l < nlev; ++l)
    {
        if (!levels[l])
            return set_error(BX_ERR_INVALID, "bx_amr_create: null level");
        if (!levels[l]->level_set)
            return set_error(BX_ERR_INVALID, "bx_amr_create: every level needs bx_set_level first");
        for (int m = 0; m < l; ++m)
            if (levels[m] == levels[l])
                return set_error(BX_ERR_INVALID, "bx_amr_create: a level appears twice");
        if (l && (levels[l]->D != levels[0]->D || !same_params(levels[0]->p, levels[l]->p)))
            return set_error(BX_ERR_INVALID, "bx_amr_create: the levels must share bx_params");
    }
    bx_amr *a = new bx_amr;
    a->D      = levels[0]->D;
    a->NC     = levels[0]->NC;
    a->lev.assign(levels, levels + nlev);
    for (int l = 0; l + 1 < nlev; ++l)
    {
        AmrPair *p = new AmrPair;
        p->crse    = levels[l];
        p->fine    = levels[l + 1];
        p->D       = a->D;
        p->NC      = a->NC;
        a->pairs.push_back(p);
        const int rc = amr_plan(p);
        if (rc != BX_OK)
        {
            bx_amr_destroy(a);
            return rc;
        }
    }
    *out = a;
    return BX_OK;
}

extern "C" int bx_amr_create(bx_ctx *coarse, bx_ctx *fine, bx_amr **out)
{
    bx_ctx *l[2] = { coarse, fine };
    if (!coarse || !fine)
        return set_error(BX_ERR_INVALID, "bx_amr_create: bad argument");
    return bx_amr_create_levels(l, 2, out);
}

extern "C" int bx_amr_info(const bx_amr *a, int *n_cf_segs, int *n_register_faces,
                           int *n_reflux_cells)
{
    if (!a)
        return set_error(BX_ERR_INVALID, "bx_amr_info: null");
    int nc = 0, nr = 0, nt = 0;
    for (const AmrPair *p : a->pairs)
    {
        nc += (int)p->cf.size();
        nr += p->nslot;
        nt += (int)p->tg.size();
    }
    if (n_cf_segs)
        *n_cf_segs = nc;
    if (n_register_faces)
        *n_register_faces = nr;
    if (n_reflux_cells)
        *n_reflux_cells = nt;
    return BX_OK;
}

// Flux-register contribution of one level's sweep input in direction d.
// coef: the coefficients in device memory (a captured step), else from dt.
static int reg_fluxes(AmrPair *a, bool fine_level, int d, const double *U, double dt,
                      const double dx[3], hipStream_t s, const double *coef)
{
    const int nslot = a->slot0[d + 1] - a->slot0[d];
    if (nslot == 0)
        return BX_OK;
    bx_ctx *L = fine_level ? a->fine : a->crse;
    // the sweeps' coefficients (same expressions, so the same bits)
    const double hdtdx = 0.5 * dt / dx[d];
    const double fc    = 0.5 * dx[d] / dt;
    const double rc    = dt / (2 * dx[d]);
    // fluxes[d].mult(dt * prod of the other cell sizes) (_advance.cpp:192-207)
    double scale = dt;
    for (int e = 0; e < a->D; ++e)
        if (e != d)
            scale *= dx[e];
    const int      nsub  = fine_level ? (1 << (a->D - 1)) : 1;
    const RegFace *faces = (fine_level ? a->d_fface : a->d_cface) + (long long)a->slot0[d] * nsub;
    double        *reg   = a->d_reg + (long long)a->slot0[d] * a->NC;
    const int      sch = L->p.num_method, lim = L->p.limiter, lv = L->lv_mode, ng = L->geom.ng;
    hipError_t     e   = hipErrorInvalidValue;
#define BX_R(DD, DIR)                                                                    \
    reg_scheme<DD, DIR>(sch, lim, lv, U, L->d_boxes, faces, nslot, nsub, reg, fine_level, \
                        ng, hdtdx, fc, rc, scale, coef, L->pm, L->lt, s)
    switch (a->D * 4 + d)
    {
    case 4: e = BX_R(1, 0); break;
    case 8: e = BX_R(2, 0); break;
    case 9: e = BX_R(2, 1); break;
    case 12: e = BX_R(3, 0); break;
    case 13: e = BX_R(3, 1); break;
    case 14: e = BX_R(3, 2); break;
    default: break;
    }
#undef BX_R
    BX_HIP_CHECK(e);
    return BX_OK;
}

// A captured coarse step's device schedule and the slot of the next
// coarse-fine fill in it; nullptr for a host-driven step.
struct CapturedStep
{
    AmrSched *S;
    int       slot;
    // the finest level's steps still to come in this coarse step: its last
    // one computes the level's maximum wave speed for the next step's dt
    int fine_left;
    // each level's current state and scratch buffers: a coarser level's
    // step leaves its result in the scratch buffer and the two exchange
    // roles (level l > 0 does so 2^l times per coarse step, so only level
    // 0's roles differ between consecutive coarse steps)
    std::vector<double *> U, O;
};

static int copy_level(double *dst, const double *src, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_f64, dim3(2048), dim3(256), 0, s, dst, src, n);
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

// One split step of level l from A (valid + ghosts filled), result in A, or
// in Y when Y is given (A then keeps the old state); B is the ping-pong
// buffer.  Each sweep's input feeds the flux registers of the pair below (as
// its fine side) and of the pair above (coarse side).  In a captured step
// the schedule kernel has written the level's dt and register coefficients.
static int level_step(bx_amr *h, int l, double dt, const double dx[3], double *A, double *B,
                      hipStream_t s, CapturedStep *g, double *Y = nullptr)
{
    bx_ctx       *L = h->lev[l];
    const int     D = h->D;
    int           rc;
    const double *in = A;
    if (!g && (rc = set_scalar(L->d_dt, dt, s)) != BX_OK)
        return rc;
    for (int d = 0; d < D; ++d)
    {
        // in place: A -> B -> A ...; into Y: the chain alternates Y and B so
        // that the last sweep writes Y
        double *out = !Y ? ((d % 2 == 0) ? B : A) : (((D - 1 - d) % 2 == 0) ? Y : B);
        const double *coef = g ? g->S->coef[l][d] : nullptr;
        // the finest level's last sweep of the coarse step reduces the wave
        // speed of its output, the level's final state (avgDown does not
        // change the finest level)
        unsigned long long *mb
            = (g && l + 1 == (int)h->lev.size() && g->fine_left == 1 && d == D - 1)
                  ? L->d_maxbits
                  : nullptr;
        if ((rc = do_sweep(L, d, dx, in, out, mb, s)) != BX_OK)
            return rc;
        if (l > 0 && (rc = reg_fluxes(h->pairs[l - 1], true, d, in, dt, dx, s, coef)) != BX_OK)
            return rc;
        if (l + 1 < (int)h->lev.size()
            && (rc = reg_fluxes(h->pairs[l], false, d, in, dt, dx, s, coef)) != BX_OK)
            return rc;
        // the next sweep's ghosts; after the last sweep nothing reads the
        // level's ghosts before a full fill (time_step's fill after a coarser
        // level's step, fill_level before a finer level's, the run's last)
        if (d + 1 < D && (rc = do_fill(L, out, d + 1, s)) != BX_OK)
            return rc;
        in = out;
    }
    if (g && l + 1 == (int)h->lev.size())
        --g->fine_left;
    if (!Y && D % 2 == 1)
        return copy_level(A, B, L->level_extent, s);
    return BX_OK;
}

static int cf_fill(AmrPair *a, double *Uf, const double *Uold, const double *Unew, int mode,
                   double alpha, double beta, hipStream_t s, const double *w = nullptr)
{
    const int nseg = (int)a->cf.size();
    if (nseg == 0)
        return BX_OK;
    const int grid = std::min(nseg, 8192);
    bx_ctx   *C = a->crse, *F = a->fine;
#define BX_CF(DD)                                                                          \
    hipLaunchKernelGGL((k_cf_fill<DD>), dim3(grid), dim3(256), 0, s, Uf, F->d_boxes, Uold, \
                       Unew, C->d_boxes, a->d_cf, nseg, F->geom.ng, C->geom.ng, mode, alpha, \
                       beta, w)
    if (a->D == 1)
        BX_CF(1);
    else if (a->D == 2)
        BX_CF(2);
    else
        BX_CF(3);
#undef BX_CF
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

static int avg_down(AmrPair *a, const double *Uf, double *Uc, hipStream_t s)
{
    const int nseg = (int)a->avg.size();
    if (nseg == 0)
        return BX_OK;
    bx_ctx      *C = a->crse, *F = a->fine;
    const double volfrac = 1.0 / (double)(1 << a->D);
#define BX_AV(DD)                                                                           \
    hipLaunchKernelGGL((k_avg_down<DD>), dim3(std::min(nseg, 4096)), dim3(256), 0, s, Uc,    \
                       C->d_boxes, Uf, F->d_boxes, a->d_avg, nseg, C->geom.ng, F->geom.ng,  \
                       volfrac)
    if (a->D == 1)
        BX_AV(1);
    else if (a->D == 2)
        BX_AV(2);
    else
        BX_AV(3);
#undef BX_AV
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

static bool amr_dx_ok(const bx_amr *a, const double dxc[3])
{
    for (int d = 0; d < a->D; ++d)
        if (!(dxc[d] > 0))
            return false;
    return true;
}

static int reflux(AmrPair *p, double *Uc, const double dx_crse[3], hipStream_t s)
{
    if (p->tg.empty())
        return BX_OK;
    double vol = dx_crse[0];
    for (int d = 1; d < p->D; ++d)
        vol *= dx_crse[d];
    const int nt = (int)p->tg.size();
    bx_ctx   *C  = p->crse;
#define BX_RF(DD)                                                                          \
    hipLaunchKernelGGL((k_reflux<DD>), dim3((nt + 127) / 128), dim3(128), 0, s, Uc,          \
                       C->d_boxes, p->d_tg, nt, p->d_terms, p->d_reg, C->geom.ng, vol)
    if (p->D == 1)
        BX_RF(1);
    else if (p->D == 2)
        BX_RF(2);
    else
        BX_RF(3);
#undef BX_RF
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

// Ghost cells of level l (> 0) at `time`: fine-fine and physical, then the
// coarse-fine ones from level l-1's states at t_old (Uold) and t_new (Unew).
// Applied to both of the level's buffers when B is given (B's coarse-fine
// ghosts only: the sweeps refill the rest).
// In a captured step the weights come from the schedule's next slot.
static int fill_level(bx_amr *h, int l, double *A, double *B, const double *Uold,
                      const double *Unew, double t_old, double t_new, double time, hipStream_t s,
                      CapturedStep *g = nullptr)
{
    int rc;
    if ((rc = do_fill_all(h->lev[l], A, s)) != BX_OK)
        return rc;
    AmrPair *p = h->pairs[l - 1];
    for (double *X : { A, B })
    {
        if (!X)
            continue;
        if (g)
            rc = cf_fill(p, X, Uold, Unew, 0, 1.0, 0.0, s, g->S->cfw[g->slot++]);
        else if (time == t_old || !(t_new != t_old))
            rc = cf_fill(p, X, Uold, Unew, 0, 1.0, 0.0, s);
        else // FillPatchTwoLevels time interpolation: alpha old + beta new
            rc = cf_fill(p, X, Uold, Unew, 1, (t_new - time) / (t_new - t_old),
                         (time - t_old) / (t_new - t_old), s);
        if (rc != BX_OK)
            return rc;
    }
    return BX_OK;
}

extern "C" int bx_amr_avg_down(bx_amr *a, const double *Uf, double *Uc, void *stream)
{
    if (!a || !Uf || !Uc)
        return set_error(BX_ERR_INVALID, "bx_amr_avg_down: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = avg_down(a->pairs[0], Uf, Uc, s)) != BX_OK)
        return rc;
    return do_fill_all(a->lev[0], Uc, s);
}

extern "C" int bx_amr_fill_fine(bx_amr *a, double *Uf, const double *Uc_old,
                                const double *Uc_new, double t_old, double t_new, double time,
                                void *stream)
{
    if (!a || !Uf || !Uc_old || !Uc_new)
        return set_error(BX_ERR_INVALID, "bx_amr_fill_fine: bad argument");
    return fill_level(a, 1, Uf, nullptr, Uc_old, Uc_new, t_old, t_new, time,
                      (hipStream_t)stream);
}

struct LevelBufs
{
    double *const *U, *const *T, *const *O;
};

// Amr::timeStep(l) (AMReX): advance level l by dt from `time`; the finer
// levels subcycle twice, then post_timestep(l) refluxes and averages down.
// t_par / dt_par: the parent level's step (for the ghost interpolation).
// g: a step being captured (time, dt, t_par, dt_par then only shape the
// recursion; the values come from the device schedule).
static int time_step(bx_amr *h, int l, double time, double dt, double t_par, double dt_par,
                     const double dx0[3], const LevelBufs &b, hipStream_t s, bool fill_after,
                     CapturedStep *g = nullptr)
{
    const int nlev = (int)h->lev.size();
    double    dx[3] = { 0, 0, 0 };
    for (int d = 0; d < h->D; ++d)
        dx[d] = dx0[d] / (double)(1 << l);
    int rc;
    // the buffers in their current roles (captured: views of g->U, g->O, so
    // they follow the exchanges below)
    const LevelBufs bb = g ? LevelBufs{ g->U.data(), b.T, g->O.data() } : b;
    if (l > 0 && (rc = fill_level(h, l, bb.U[l], bb.T[l], bb.O[l - 1], bb.U[l - 1], t_par,
                                  t_par + dt_par, time, s, g))
                     != BX_OK)
        return rc;
    // The finer levels' fills need this level's old state (ghosts filled).
    // Host-driven: copied to O[l] before the step.  Captured: the step writes
    // its result into O[l] and U[l] keeps the old state, then the two
    // exchange roles (the run alternates two graphs for level 0 and copies
    // back once after an odd number of steps).
    if (l + 1 < nlev && g)
    {
        if ((rc = level_step(h, l, dt, dx, bb.U[l], bb.T[l], s, g, bb.O[l])) != BX_OK)
            return rc;
        std::swap(g->U[l], g->O[l]);
    }
    else
    {
        if (l + 1 < nlev
            && (rc = copy_level(b.O[l], b.U[l], h->lev[l]->level_extent, s)) != BX_OK)
            return rc;
        if ((rc = level_step(h, l, dt, dx, b.U[l], b.T[l], s, g)) != BX_OK)
            return rc;
    }
    if (l + 1 < nlev)
    {
        if ((rc = do_fill_all(h->lev[l], bb.U[l], s)) != BX_OK)
            return rc;
        const double dtf = dt / 2; // dt_level[l+1] = dt_level[l] / 2
        for (int sub = 0; sub < 2; ++sub)
            if ((rc = time_step(h, l + 1, sub == 0 ? time : time + dtf, dtf, time, dt, dx0, bb,
                                s, false, g))
                != BX_OK)
                return rc;
        AmrPair *p = h->pairs[l];
        if ((rc = reflux(p, bb.U[l], dx, s)) != BX_OK
            || (rc = avg_down(p, bb.U[l + 1], bb.U[l], s))
            || (rc = do_fill_all(h->lev[l], bb.U[l], s)) != BX_OK)
            return rc;
    }
    else if (fill_after && (rc = do_fill_all(h->lev[l], b.U[l], s)) != BX_OK)
        return rc;
    return BX_OK;
}

// estTimeStep of one level (AmrLevelAdv.cpp:395-477, explicit methods) from
// the bits of its maximum wave speed
static void level_dt_est(const bx_ctx *L, unsigned long long bits, const double dx[3],
                         double cfl, double *out)
{
    double max_speed;
    std::memcpy(&max_speed, &bits, sizeof(max_speed));
    double dt_est = 1.0e20;
    for (int d = 0; d < L->D; ++d)
    {
        const double c = dx[d] / max_speed;
        dt_est         = (c < dt_est) ? c : dt_est;
    }
    *out = dt_est * cfl;
}

// computeNewDt (AmrLevelAdv.cpp:589-618): dt_0 = min over levels of
// n_factor * dt_est(l), n_factor = 2^l, then the stop-time clamp.
static int est_dt_levels(bx_amr *a, const double dx0[3], double cfl, double time,
                         double stop_time, const double *const *U, double *dt, hipStream_t s)
{
    const int nlev = (int)a->lev.size();
    int       rc;
    for (int l = 0; l < nlev; ++l)
        if ((rc = do_maxspeed(a->lev[l], U[l], s)) != BX_OK)
            return rc;
    std::vector<unsigned long long> bits(nlev, 0ULL);
    for (int l = 0; l < nlev; ++l)
        BX_HIP_CHECK(hipMemcpyAsync(&bits[l], a->lev[l]->d_maxbits, sizeof(bits[l]),
                                    hipMemcpyDeviceToHost, s));
    BX_HIP_CHECK(hipStreamSynchronize(s));
    double dt_0     = 1.0e+100;
    int    n_factor = 1;
    for (int l = 0; l < nlev; ++l)
    {
        double dxl[3] = { 0, 0, 0 };
        for (int d = 0; d < a->D; ++d)
            dxl[d] = dx0[d] / (double)(1 << l);
        double dmin;
        level_dt_est(a->lev[l], bits[l], dxl, cfl, &dmin);
        n_factor *= (l == 0 ? 1 : 2);
        dt_0 = std::min(dt_0, n_factor * dmin);
    }
    const double eps = 0.001 * dt_0;
    if (stop_time >= 0.0)
        if ((time + dt_0) > (stop_time - eps))
            dt_0 = stop_time - time;
    *dt = dt_0;
    return BX_OK;
}

static bool bufs_ok(const bx_amr *a, const LevelBufs &b)
{
    if (!b.U || !b.T || !b.O)
        return false;
    for (size_t l = 0; l < a->lev.size(); ++l)
        if (!b.U[l] || !b.T[l] || (l + 1 < a->lev.size() && !b.O[l]))
            return false;
    return true;
}

extern "C" int bx_amr_advance_levels(bx_amr *a, double dt, const double dx0[3], double time,
                                     double *const *U, double *const *U_tmp,
                                     double *const *U_old, void *stream)
{
    LevelBufs b{ U, U_tmp, U_old };
    if (!a || !amr_dx_ok(a, dx0) || !(dt > 0) || !bufs_ok(a, b))
        return set_error(BX_ERR_INVALID, "bx_amr_advance: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = time_step(a, 0, time, dt, time, dt, dx0, b, s, true)) != BX_OK)
        return rc;
    for (size_t l = 1; l < a->lev.size(); ++l)
        if ((rc = do_fill_all(a->lev[l], U[l], s)) != BX_OK)
            return rc;
    return BX_OK;
}

extern "C" int bx_amr_est_dt_levels(bx_amr *a, const double dx0[3], double cfl, double time,
                                    double stop_time, const double *const *U, double *dt,
                                    void *stream)
{
    if (!a || !U || !dt || !amr_dx_ok(a, dx0))
        return set_error(BX_ERR_INVALID, "bx_amr_est_dt: bad argument");
    for (size_t l = 0; l < a->lev.size(); ++l)
        if (!U[l])
            return set_error(BX_ERR_INVALID, "bx_amr_est_dt: bad argument");
    return est_dt_levels(a, dx0, cfl, time, stop_time, U, dt, (hipStream_t)stream);
}

// the coarse step fits the device schedule's slots
static bool graph_fits(const bx_amr *a)
{
    const int nlev = (int)a->lev.size();
    if (nlev > AMR_MAXLEV)
        return false;
    long long ncf = 0; // two fills per fine-level step, 2^l steps of level l
    for (int l = 1; l < nlev; ++l)
        ncf += 2LL << l;
    return ncf <= AMR_MAXCF;
}

static int capture_one(bx_amr *a, const SchedArgs &sa, const LevelBufs &b, const double dx0[3],
                       hipGraphExec_t *gx);

// Captures one coarse step (the levels' maximum wave speeds, the schedule,
// Amr::timeStep(0)) into a->gx, for even and for odd steps, unless the
// graphs already hold this configuration.
static int capture_step(bx_amr *a, const double dx0[3], double cfl, const LevelBufs &b)
{
    const int             nlev = (int)a->lev.size();
    std::vector<uint64_t> key;
    for (double v : { dx0[0], dx0[1], dx0[2], cfl })
    {
        uint64_t u;
        std::memcpy(&u, &v, sizeof(u));
        key.push_back(u);
    }
    for (int l = 0; l < nlev; ++l)
    {
        const bx_ctx *L = a->lev[l];
        // every device table a captured launch reads (a re-upload that moves
        // one forces a new capture)
        for (const void *p : { (const void *)b.U[l], (const void *)b.T[l],
                               (const void *)(l + 1 < nlev ? b.O[l] : nullptr),
                               (const void *)L->d_boxes, (const void *)L->d_ib_x,
                               (const void *)L->d_ib_t[0], (const void *)L->d_ib_t[1],
                               (const void *)L->d_ib_t[2], (const void *)L->d_cbeg,
                               (const void *)L->d_segs[0], (const void *)L->d_segs[1],
                               (const void *)L->d_segs[2], (const void *)L->d_segs[3],
                               (const void *)L->d_segs[4], (const void *)L->d_maxbits,
                               (const void *)L->d_dt })
            key.push_back((uint64_t)(uintptr_t)p);
        for (int v : { L->tune.block_x, L->tune.block_t, L->tune.seg_k, L->tune.grid_cap,
                       L->occ_x, L->occ_t })
            key.push_back((uint64_t)(int64_t)v);
    }
    if (a->gx[0] && a->gx[1] && key == a->gkey)
        return BX_OK;
    for (hipGraphExec_t &gx : a->gx)
        if (gx)
        {
            BX_HIP_CHECK(hipGraphExecDestroy(gx));
            gx = nullptr;
        }
    if (!a->d_sched)
        BX_HIP_CHECK(hipMalloc((void **)&a->d_sched, sizeof(AmrSched)));
    if (!a->cap)
        BX_HIP_CHECK(hipStreamCreateWithFlags(&a->cap, hipStreamNonBlocking));
    SchedArgs sa{};
    for (int l = 0; l < nlev; ++l)
    {
        sa.maxbits[l] = a->lev[l]->d_maxbits;
        sa.dt[l]      = a->lev[l]->d_dt;
    }
    for (int d = 0; d < 3; ++d)
        sa.dx0[d] = dx0[d];
    sa.cfl  = cfl;
    sa.nlev = nlev;
    sa.D    = a->D;
    // odd steps start with level 0's state and scratch buffers exchanged
    for (int q = 0; q < 2; ++q)
    {
        std::vector<double *> U0(b.U, b.U + nlev), O0(b.O, b.O + nlev);
        if (q == 1)
            std::swap(U0[0], O0[0]);
        const int rc = capture_one(a, sa, LevelBufs{ U0.data(), b.T, O0.data() }, dx0,
                                   &a->gx[q]);
        if (rc != BX_OK)
            return rc;
    }
    a->gkey = key;
    return BX_OK;
}

// One coarse step of the buffers b captured into *gx.
static int capture_one(bx_amr *a, const SchedArgs &sa, const LevelBufs &b, const double dx0[3],
                       hipGraphExec_t *gx)
{
    const int   nlev = (int)a->lev.size();
    hipStream_t s    = a->cap;
    BX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = BX_OK;
    // the finest level's maximum comes from its last sweep of the previous
    // step (or from the run's start)
    for (int l = 0; l + 1 < nlev && rc == BX_OK; ++l)
        rc = do_maxspeed(a->lev[l], b.U[l], s);
    if (rc == BX_OK)
    {
        hipLaunchKernelGGL(k_amr_schedule, dim3(1), dim3(64), 0, s, a->d_sched, sa);
        rc = launch_set_u64(a->lev[nlev - 1]->d_maxbits, 0ULL, s) == hipSuccess ? BX_OK
                                                                                 : BX_ERR_HIP;
    }
    if (rc == BX_OK)
    {
        CapturedStep g{ a->d_sched, 0, 1 << (nlev - 1),
                        std::vector<double *>(b.U, b.U + nlev),
                        std::vector<double *>(b.O, b.O + nlev) };
        // the host's time values only shape the recursion here
        rc = time_step(a, 0, 0.0, 1.0, 0.0, 1.0, dx0, b, s, false, &g);
    }
    hipGraph_t      graph = nullptr;
    const hipError_t ec   = hipStreamEndCapture(s, &graph);
    if (rc != BX_OK || ec != hipSuccess)
    {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc != BX_OK ? rc
                           : set_error(BX_ERR_HIP, std::string("amr step capture: ")
                                                       + hipGetErrorString(ec));
    }
    const hipError_t ei = hipGraphInstantiate(gx, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess)
    {
        *gx = nullptr;
        return set_error(BX_ERR_HIP, std::string("amr step instantiate: ")
                                         + hipGetErrorString(ei));
    }
    return BX_OK;
}

extern "C" int bx_amr_run_levels(bx_amr *a, const double dx0[3], double cfl, double stop_time,
                                 int max_step, double *const *U, double *const *U_tmp,
                                 double *const *U_old, double *time, int *steps, void *stream)
{
    LevelBufs b{ U, U_tmp, U_old };
    if (!a || !time || !amr_dx_ok(a, dx0) || !bufs_ok(a, b))
        return set_error(BX_ERR_INVALID, "bx_amr_run: bad argument");
    hipStream_t s    = (hipStream_t)stream;
    const int   nlev = (int)a->lev.size();
    int         rc;
    // post_init: average down from the finest level, then every ghost cell
    for (int l = nlev - 2; l >= 0; --l)
        if ((rc = avg_down(a->pairs[l], U[l + 1], U[l], s)) != BX_OK)
            return rc;
    for (int l = 0; l < nlev; ++l)
        if ((rc = do_fill_all(a->lev[l], U[l], s)) != BX_OK)
            return rc;
    double t = *time;
    int    n = 0;
    if (a->use_graph && stop_time < 0.0 && max_step > 0 && graph_fits(a))
    {
        // a fixed number of steps: the step is captured once and replayed,
        // its dt, time and interpolation weights computed on the device
        if ((rc = capture_step(a, dx0, cfl, b)) != BX_OK)
            return rc;
        if ((rc = set_scalar(&a->d_sched->t, t, s)) != BX_OK
            || (rc = do_maxspeed(a->lev[nlev - 1], U[nlev - 1], s)) != BX_OK)
            return rc;
        for (; n < max_step; ++n)
            BX_HIP_CHECK(hipGraphLaunch(a->gx[n & 1], s));
        if ((n & 1) // level 0's state is in its scratch buffer
            && (rc = copy_level(U[0], U_old[0], a->lev[0]->level_extent, s)) != BX_OK)
            return rc;
        BX_HIP_CHECK(hipMemcpyAsync(&t, &a->d_sched->t, sizeof(double), hipMemcpyDeviceToHost,
                                    s));
        BX_HIP_CHECK(hipStreamSynchronize(s));
    }
    while ((n < max_step || max_step < 0) && (t < stop_time || stop_time < 0.0))
    {
        double dt;
        if ((rc = est_dt_levels(a, dx0, cfl, t, stop_time, U, &dt, s)) != BX_OK)
            return rc;
        if ((rc = time_step(a, 0, t, dt, t, dt, dx0, b, s, false)) != BX_OK)
            return rc;
        t = t + dt;
        ++n;
    }
    for (int l = 1; l < nlev; ++l)
        if ((rc = do_fill_all(a->lev[l], U[l], s)) != BX_OK)
            return rc;
    BX_HIP_CHECK(hipStreamSynchronize(s));
    *time = t;
    if (steps)
        *steps = n;
    return BX_OK;
}

// ------------------------------------------------- two-level entry points
static bool two_levels(const bx_amr *a) { return a && a->lev.size() == 2; }

extern "C" int bx_amr_advance(bx_amr *a, double dt, const double dx_coarse[3], double time,
                              double *Uc, double *Uc_tmp, double *Uc_old, double *Uf,
                              double *Uf_tmp, void *stream)
{
    if (!two_levels(a))
        return set_error(BX_ERR_INVALID, "bx_amr_advance: needs a two-level hierarchy");
    double *U[2] = { Uc, Uf }, *T[2] = { Uc_tmp, Uf_tmp }, *O[2] = { Uc_old, nullptr };
    return bx_amr_advance_levels(a, dt, dx_coarse, time, U, T, O, stream);
}

extern "C" int bx_amr_est_dt(bx_amr *a, const double dx_coarse[3], double cfl, double time,
                             double stop_time, const double *Uc, const double *Uf, double *dt,
                             void *stream)
{
    if (!two_levels(a))
        return set_error(BX_ERR_INVALID, "bx_amr_est_dt: needs a two-level hierarchy");
    const double *U[2] = { Uc, Uf };
    return bx_amr_est_dt_levels(a, dx_coarse, cfl, time, stop_time, U, dt, stream);
}

extern "C" int bx_amr_run(bx_amr *a, const double dx_coarse[3], double cfl, double stop_time,
                          int max_step, double *Uc, double *Uc_tmp, double *Uc_old, double *Uf,
                          double *Uf_tmp, double *time, int *steps, void *stream)
{
    if (!two_levels(a))
        return set_error(BX_ERR_INVALID, "bx_amr_run: needs a two-level hierarchy");
    double *U[2] = { Uc, Uf }, *T[2] = { Uc_tmp, Uf_tmp }, *O[2] = { Uc_old, nullptr };
    return bx_amr_run_levels(a, dx_coarse, cfl, stop_time, max_step, U, T, O, time, steps,
                             stream);
}
