// Fused MUSCL-Hancock + HLLC + conservative-update sweep kernels, ghost
// fill and CFL reduction for all boxes of a level.
//
// Reference path replaced (per box, per direction d):
//   compute_MUSCL_hancock_flux<d>  Source/Fluxes/MUSCL_Hancock.cpp:15-52
//   conservative update            Source/AmrLevelAdv_advance.cpp:330-341
//   fill_consv_boundary            Source/BoundaryConditions/BCs.cpp:429-442
//   max_wave_speed / estTimeStep   Source/System/Euler/Euler.cpp:46-74,
//                                  Source/AmrLevelAdv.cpp:395-477, 537-619
//
// Each sweep is ONE persistent launch over a device-resident box list.  The
// level's work is cut into items.
//
//  * x sweep (unit stride): an item is B consecutive positions of one box
//    plane (B = the block size), flattened over (valid rows) x (row incl.
//    ghosts); all waves of a block work on one item, and the blocks take
//    chunks of X_CHUNK consecutive items in turn.  Thread t holds one
//    position; the neighbour values the stencil needs (U at t-1 and t+1, the
//    high-side half-stepped state at t-1, the face flux at t+1) pass through
//    LDS, one buffer each, with a block barrier after each write.  B - 4 of
//    B threads produce an update.
//  * y/z sweeps (strided): an item is 64 (transverse cell, segment) pairs,
//    and every wave owns a contiguous range of items (consecutive ranges go
//    to the waves of one XCD, so a range stays in one L2); lane t walks its column along the sweep direction through K cells,
//    keeping the 3-cell window, the previous half-stepped state and the
//    previous face flux in registers (K+2 reconstructions per K updates).
// No temporaries leave registers: per cell the sweep reads U once and
// writes U once (plus halo re-reads that hit L2).
#include <hip/hip_runtime.h>

#include "bx_internal.h"
#include "bx_physics.h"
#include "bx_kernels.h"

// Occupancy targets of the sweep kernels (waves per SIMD; 0 = compiler's
// choice).  Raising them trades register spills for latency hiding.
#ifndef BX_SWEEP_X_WAVES
#define BX_SWEEP_X_WAVES 0
#endif
#ifndef BX_SWEEP_T_WAVES
#define BX_SWEEP_T_WAVES 0
#endif
#if BX_SWEEP_X_WAVES > 0
#define BX_X_OCC __attribute__((amdgpu_waves_per_eu(BX_SWEEP_X_WAVES)))
#else
#define BX_X_OCC
#endif
#if BX_SWEEP_T_WAVES > 0
#define BX_T_OCC __attribute__((amdgpu_waves_per_eu(BX_SWEEP_T_WAVES)))
#else
#define BX_T_OCC
#endif

namespace bx {

__device__ __forceinline__ int find_box(const long long *__restrict__ begin,
                                        int nbox, long long item)
{
    int lo = 0, hi = nbox; // invariant: begin[lo] <= item < begin[hi]
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (begin[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__device__ __forceinline__ unsigned long long umax64(unsigned long long a,
                                                     unsigned long long b)
{
    return a > b ? a : b;
}

// Running maximum of non-negative wave speeds as ordered bits.  A NaN speed
// (negative pressure) is skipped, as the reference's reduction skips it:
// amrex::max(acc, s) = (acc < s) ? s : acc keeps acc when s is NaN
// (max_wave_speed, Euler.cpp:53-68).
__device__ __forceinline__ unsigned long long speed_max(unsigned long long acc,
                                                        double s)
{
    return s == s ? umax64(acc, (unsigned long long)__double_as_longlong(s)) : acc;
}

// Contiguous item range of this wave, XCD-aware: blocks are dealt to the 8
// XCDs round-robin, so block b runs on XCD b%8; renumber so that XCD x owns
// a contiguous slice of the item space.
__device__ __forceinline__ void wave_item_range(long long nitems,
                                                long long &first,
                                                long long &last)
{
    const int  wpb    = blockDim.x >> 6;
    const int  nblk   = gridDim.x;
    int        blk    = blockIdx.x;
    if ((nblk & 7) == 0)
        blk = (blk & 7) * (nblk >> 3) + (blk >> 3);
    // the wave index is uniform across the wave: readfirstlane lets the
    // compiler keep all per-item bookkeeping in scalar registers
    const int       wib  = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long nw   = (long long)nblk * wpb;
    const long long w    = (long long)blk * wpb + wib;
    const long long per  = (nitems + nw - 1) / nw;
    first                = w * per;
    last                 = first + per;
    if (last > nitems)
        last = nitems;
}

// Block maximum, then one atomicMax per block (many waves hitting one
// address serialise in L2).  Every thread of the block must call it.
__device__ __forceinline__ void block_max_to_global(unsigned long long v,
                                                    unsigned long long *dst)
{
    __shared__ unsigned long long red[16];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = umax64(v, __shfl_xor(v, off));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        red[w] = v;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long m = 0ULL;
        for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i)
            m = umax64(m, red[i]);
        if (m != 0ULL)
            atomicMax(dst, m);
    }
}

// ---------------------------------------------------------------- x sweep
// Geometry of one x-sweep item (B consecutive positions of one box plane,
// B - 4 = step of them updated).
struct XItem
{
    long long pbase; // element offset of position 0 of the plane
    long long p;     // this lane's position
    long long P;     // positions in the plane (valid rows x padded row)
    long long nst;   // component stride
    int       nx, nxd;
};

// Walk over a block's consecutive x-sweep items.  The box geometry and the
// item's (plane, chunk) position are carried from one item to the next, so
// the divisions that locate an item run only when the walk enters a box
// (the first item of a chunk of items, with its box from a per-chunk
// table, or the next box at a box boundary); within a box the next item is
// the next chunk of the plane, or the first chunk of the next plane.  Every
// value is uniform over the block.
struct XCursor
{
    int       b;                      // current box
    long long bend;                   // first item past the current box
    long long base, kstride, jstride; // box geometry
    long long nst, P;
    int       nx, nxd, cpp, kpl, chunk;
};

// b: the box that holds item
template <int D>
__device__ __forceinline__ void x_enter(const DevBox *__restrict__ boxes,
                                        const long long *__restrict__ ibeg, int b,
                                        long long item, int ng, int step, XCursor &q)
{
    const DevBox &bx = boxes[b];
    q.b              = b;
    q.bend           = ibeg[b + 1];
    q.nx             = bx.hi[0] - bx.lo[0] + 1;
    const int ny     = (D >= 2) ? bx.hi[1] - bx.lo[1] + 1 : 1;
    q.nxd            = q.nx + 2 * ng;
    q.P              = (long long)ny * q.nxd;
    q.cpp            = (int)((q.P + step - 1) / step);
    q.base           = bx.base;
    q.kstride        = bx.kstride;
    q.jstride        = bx.jstride;
    q.nst            = bx.nstride;
    const unsigned local = (unsigned)(item - ibeg[b]);
    q.kpl                = (int)(local / (unsigned)q.cpp);
    q.chunk              = (int)(local - (unsigned)q.kpl * (unsigned)q.cpp);
}

// item: the item the cursor moves to (the one after its current item)
template <int D>
__device__ __forceinline__ void x_advance(const DevBox *__restrict__ boxes,
                                          const long long *__restrict__ ibeg,
                                          long long item, int ng, int step, XCursor &q)
{
    if (item >= q.bend) // every box has at least one item
    {
        x_enter<D>(boxes, ibeg, q.b + 1, item, ng, step, q);
        return;
    }
    if (++q.chunk == q.cpp)
    {
        q.chunk = 0;
        ++q.kpl;
    }
}

template <int D>
__device__ __forceinline__ void x_item(const XCursor &q, int ng, int lane, int step, XItem &it)
{
    const int ngy = (D >= 2) ? ng : 0;
    const int ngz = (D >= 3) ? ng : 0;
    it.pbase      = q.base + (long long)(q.kpl + ngz) * q.kstride + (long long)ngy * q.jstride;
    it.p          = (long long)q.chunk * step - 2 + lane;
    it.P          = q.P;
    it.nst        = q.nst;
    it.nx         = q.nx;
    it.nxd        = q.nxd;
}

// PM (physics mode) bits: PM_NI -- the HLLC exponents take the near-integer
// pow forms with n = 7 (gamma = 1.4) fixed at compile time; PM_E1 --
// epsilon == 1 (set only together with PM_NI).
constexpr int PM_NI = 1;
constexpr int PM_E1 = 2;

template <int D, int LIM, int LV, int SCH, int PM>
__global__ __launch_bounds__(256) BX_X_OCC void k_sweep_x(
    const double *__restrict__ Uin, double *__restrict__ Uout,
    const DevBox *__restrict__ boxes, const long long *__restrict__ ibeg,
    const int *__restrict__ xbox, long long nitems, int ng,
    const double *__restrict__ d_dt, double dx, EulerParm pm, LimTable lt,
    unsigned long long *d_maxbits)
{
    constexpr int  NC    = D + 2;
    constexpr bool E1    = (PM & PM_E1) != 0;
    constexpr bool NI    = (PM & PM_NI) != 0;
    const int      t     = threadIdx.x;
    const int      B     = blockDim.x;
    const int      step  = x_item_step(B);
    const int      tm    = t > 0 ? t - 1 : 0;     // neighbours inside the item
    const int      tp    = t < B - 1 ? t + 1 : t; // (clamped at its ends)
    const double   dt    = *d_dt;
    const double   hdtdx = 0.5 * dt / dx;
    const double   dtdx  = dt / dx;
    const double   fc    = 0.5 * dx / dt;   // FORCE / LF
    const double   rc    = dt / (2 * dx);   // FORCE
    constexpr bool RECON = scheme_reconstructs(SCH);
    constexpr int  XS    = X_LDS_STRIDE;
    extern __shared__ double x_lds[];
    double *const  sU = x_lds;
    double *const  sH = x_lds + NC * XS;
    double *const  sF = x_lds + 2 * NC * XS;

    // Chunks of X_CHUNK consecutive items are dealt to the blocks in turn,
    // so every block (and every XCD) samples the whole level: items whose
    // HLLC branches diverge (along shocks) or take the costlier star-state
    // branch are spread over all blocks instead of lengthening the blocks
    // that own one region.  Within a chunk the cursor advances without
    // divisions.
    const long long nblk = gridDim.x;
    long long       item = (long long)blockIdx.x * X_CHUNK;
    const long long end  = nitems;
    auto next_of = [&](long long it) -> long long {
        return ((it + 1) % X_CHUNK) ? it + 1 : it + 1 + (nblk - 1) * X_CHUNK;
    };
    long long          nxt_item = next_of(item);
    unsigned long long vmax = 0ULL;
    XCursor            q;
    long long          qitem = -1; // the item the cursor is on
    XItem              cur, nxt;
    double             Un_[NC]; // prefetched state of the next item
    auto issue = [&](const XItem &it, double (&V)[NC]) {
        const long long pc = it.p < 0 ? 0 : (it.p >= it.P ? it.P - 1 : it.p);
#pragma unroll
        for (int n = 0; n < NC; ++n)
            V[n] = Uin[it.pbase + pc + n * it.nst];
    };
    if (item < end)
    {
        x_enter<D>(boxes, ibeg, xbox[item / X_CHUNK], item, ng, step, q);
        qitem = item;
        x_item<D>(q, ng, t, step, nxt);
        issue(nxt, Un_);
    }
    // Every thread of the block runs the same items, so the barriers below
    // are reached uniformly.  Each buffer is written before one barrier and
    // read after it; its next write comes after the following barrier.
    while (item < end)
    {
        cur = nxt;
        double U[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            U[n] = Un_[n];
        if (nxt_item < end)
        {
            if (nxt_item == qitem + 1)
                x_advance<D>(boxes, ibeg, nxt_item, ng, step, q);
            else // the first item of a chunk
                x_enter<D>(boxes, ibeg, xbox[nxt_item / X_CHUNK], nxt_item, ng, step, q);
            qitem = nxt_item;
            x_item<D>(q, ng, t, step, nxt);
            issue(nxt, Un_);
        }
#pragma unroll
        for (int n = 0; n < NC; ++n)
            sU[n * XS + t] = U[n];
        __syncthreads();
        double Um[NC], F[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            Um[n] = sU[n * XS + tm];
        if (!RECON)
        {
            face_flux<D, 0, E1, NI, SCH>(Um, U, F, pm, fc, rc);
        }
        else
        {
            double Up[NC], hL[NC], hR[NC], hRm[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n)
                Up[n] = sU[n * XS + tp];
            reconstruct_half_step<D, 0, LIM, LV, E1>(Um, U, Up, hL, hR, hdtdx,
                                                     pm, lt);
#pragma unroll
            for (int n = 0; n < NC; ++n)
                sH[n * XS + t] = hR[n];
            __syncthreads();
#pragma unroll
            for (int n = 0; n < NC; ++n)
                hRm[n] = sH[n * XS + tm];
            face_flux<D, 0, E1, NI, SCH>(hRm, hL, F, pm, fc, rc);
        }
#pragma unroll
        for (int n = 0; n < NC; ++n)
            sF[n * XS + t] = F[n];
        __syncthreads();
        double Fp[NC];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            Fp[n] = sF[n * XS + tp];

        const long long pc  = cur.p < 0 ? 0 : (cur.p >= cur.P ? cur.P - 1 : cur.p);
        const int       col = (int)((unsigned)pc % (unsigned)cur.nxd);
        const bool upd = t >= 2 && t < B - 2 && cur.p < cur.P && col >= ng
                         && col < ng + cur.nx;
        if (upd)
        {
            double Unew[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n)
            {
                Unew[n] = U[n] - dtdx * (Fp[n] - F[n]);
                Uout[cur.pbase + cur.p + n * cur.nst] = Unew[n];
            }
            if (d_maxbits)
            {
                const double sp = cell_wave_speed<D, E1>(Unew, pm);
                vmax = speed_max(vmax, sp);
            }
        }
        item     = nxt_item;
        nxt_item = next_of(nxt_item);
    }
    if (d_maxbits) // uniform over the block; every thread gets here
        block_max_to_global(vmax, d_maxbits);
}

// -------------------------------------------------------------- y/z sweep
template <int D, int DIR, int LIM, int LV, int SCH, int PM>
__global__ __launch_bounds__(256) BX_T_OCC void k_sweep_t(
    const double *__restrict__ Uin, double *__restrict__ Uout,
    const DevBox *__restrict__ boxes, const long long *__restrict__ ibeg,
    int nbox, long long nitems, int ng, int K, const double *__restrict__ d_dt,
    double dx, EulerParm pm, LimTable lt, unsigned long long *d_maxbits)
{
    constexpr int  NC    = D + 2;
    constexpr bool E1    = (PM & PM_E1) != 0;
    constexpr bool NI    = (PM & PM_NI) != 0;
    const int      lane  = threadIdx.x & 63;
    const double   dt    = *d_dt;
    const double   hdtdx = 0.5 * dt / dx;
    const double   dtdx  = dt / dx;
    const double   fc    = 0.5 * dx / dt;   // FORCE / LF
    const double   rc    = dt / (2 * dx);   // FORCE
    constexpr bool RECON = scheme_reconstructs(SCH);
    const int      ngy   = (D >= 2) ? ng : 0;
    const int      ngz   = (D >= 3) ? ng : 0;

    long long first, last;
    wave_item_range(nitems, first, last);
    unsigned long long vmax = 0ULL;
    int                b    = -1;
    long long          bend = -1;
    DevBox             bx;
    int na = 0, nb = 0, L = 0, nseg = 0;
    long long Q = 0, sd = 0;
    for (long long item = first; item < last; ++item)
    {
        if (item >= bend)
        {
            b    = find_box(ibeg, nbox, item);
            bend = ibeg[b + 1];
            bx   = boxes[b];
            na   = bx.hi[0] - bx.lo[0] + 1;
            if (DIR == 1)
            {
                nb = (D >= 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
                L  = bx.hi[1] - bx.lo[1] + 1;
                sd = bx.jstride;
            }
            else
            {
                nb = bx.hi[1] - bx.lo[1] + 1;
                L  = bx.hi[2] - bx.lo[2] + 1;
                sd = bx.kstride;
            }
            nseg = (L + K - 1) / K;
            Q    = (long long)na * nb * nseg;
        }
        const long long q = (item - ibeg[b]) * 64 + lane;
        if (q >= Q)
            continue;
        const unsigned  qu  = (unsigned)q;
        const int       a   = (int)(qu % (unsigned)na);
        const unsigned  q2  = qu / (unsigned)na;
        const int       bb  = (int)(q2 % (unsigned)nb);
        const int       seg = (int)(q2 / (unsigned)nb);
        long long       col;
        if (DIR == 1)
            col = bx.base + (a + ng) + (long long)(bb + ngz) * bx.kstride
                  + (long long)ngy * bx.jstride;
        else
            col = bx.base + (a + ng) + (long long)(bb + ngy) * bx.jstride
                  + (long long)ngz * bx.kstride;
        const int s0 = seg * K;
        const int s1 = (s0 + K < L) ? s0 + K : L; // cells [s0, s1)

        auto load = [&](int s, double (&V)[NC]) {
            const long long o = col + (long long)s * sd;
#pragma unroll
            for (int n = 0; n < NC; ++n)
                V[n] = Uin[o + n * bx.nstride];
        };
        // window U(s-1), U(s), and U(s+1) in flight one iteration ahead
        double Um1[NC], U0[NC], Up[NC], hRprev[NC], Fprev[NC];
        load(s0 - 1, Um1);
        load(s0, U0);
        load(s0 + 1, Up);
        if (RECON)
        {
            double Um2[NC], hLd[NC];
            load(s0 - 2, Um2);
            reconstruct_half_step<D, DIR, LIM, LV, E1>(Um2, Um1, U0, hLd,
                                                       hRprev, hdtdx, pm, lt);
        }
        for (int s = s0; s <= s1; ++s)
        {
            double Unext[NC];
            // U(s+2) is only needed by iteration s+1 (s+1 <= s1)
            if (s < s1)
                load(s + 2, Unext);
            double F[NC];
            if (!RECON)
            {
                face_flux<D, DIR, E1, NI, SCH>(Um1, U0, F, pm, fc, rc);
            }
            else
            {
                double hL0[NC], hR0[NC];
                reconstruct_half_step<D, DIR, LIM, LV, E1>(Um1, U0, Up, hL0,
                                                           hR0, hdtdx, pm, lt);
                face_flux<D, DIR, E1, NI, SCH>(hRprev, hL0, F, pm, fc, rc);
#pragma unroll
                for (int n = 0; n < NC; ++n)
                    hRprev[n] = hR0[n];
            }
            if (s > s0)
            {
                // update cell s-1 from faces s-1 (Fprev) and s (F)
                const long long o = col + (long long)(s - 1) * sd;
                double          Un[NC];
#pragma unroll
                for (int n = 0; n < NC; ++n)
                {
                    Un[n] = Um1[n] - dtdx * (F[n] - Fprev[n]);
                    Uout[o + n * bx.nstride] = Un[n];
                }
                if (d_maxbits)
                {
                    const double sp = cell_wave_speed<D, E1>(Un, pm);
                    vmax = speed_max(vmax, sp);
                }
            }
#pragma unroll
            for (int n = 0; n < NC; ++n)
            {
                Fprev[n] = F[n];
                Um1[n]   = U0[n];
                U0[n]    = Up[n];
            }
            if (s < s1)
#pragma unroll
                for (int n = 0; n < NC; ++n)
                    Up[n] = Unext[n];
        }
    }
    if (d_maxbits) // uniform over the block; every thread gets here
        block_max_to_global(vmax, d_maxbits);
}

#ifndef BX_SWEEP_PART // kernels outside the per-dimension sweep instantiations
__device__ void step_tail(const StepTail &t);

// ------------------------------------------------------------- ghost fill
// One block per ghost segment (block-stride over segments), so the segment
// and its two boxes are wave-uniform loads; lanes run over the segment's
// cells, x fastest.
__global__ __launch_bounds__(256) void k_fill_ghosts(
    double *__restrict__ U, const DevBox *__restrict__ boxes,
    const GhostSeg *__restrict__ segs, int nseg, int ng, int D, int NC,
    StepTail tail)
{
    if (tail.d_time && blockIdx.x == 0 && threadIdx.x == 0)
        step_tail(tail);
    const int ngy = (D >= 2) ? ng : 0;
    const int ngz = (D >= 3) ? ng : 0;
    for (int si = blockIdx.x; si < nseg; si += gridDim.x)
    {
        const GhostSeg &g  = segs[si];
        const int       ex = g.hi[0] - g.lo[0] + 1;
        const int       ey = g.hi[1] - g.lo[1] + 1;
        const int       ez = g.hi[2] - g.lo[2] + 1;
        const int       n  = ex * ey * ez;
        const DevBox   &db = boxes[g.dst];
        const DevBox   &sb = boxes[g.src];
        for (int c = threadIdx.x; c < n; c += blockDim.x)
        {
            const int x  = g.lo[0] + c % ex;
            const int t  = c / ex;
            const int y  = g.lo[1] + t % ey;
            const int z  = g.lo[2] + t / ey;
            const int sx = g.a[0] * x + g.b[0];
            const int sy = g.a[1] * y + g.b[1];
            const int sz = g.a[2] * z + g.b[2];
            const long long doff = db.base + (x - db.lo[0] + ng)
                                   + (long long)(y - db.lo[1] + ngy) * db.jstride
                                   + (long long)(z - db.lo[2] + ngz) * db.kstride;
            const long long soff = sb.base + (sx - sb.lo[0] + ng)
                                   + (long long)(sy - sb.lo[1] + ngy) * sb.jstride
                                   + (long long)(sz - sb.lo[2] + ngz) * sb.kstride;
            for (int m = 0; m < NC; ++m)
            {
                double v = U[soff + m * sb.nstride];
                if ((g.flip_mask >> m) & 1)
                    v = -v;
                U[doff + m * db.nstride] = v;
            }
        }
    }
}

// ------------------------------------------------------ max wave speed
// Blocks take chunks of MWS_CHUNK consecutive cells of the flattened level
// (boxes in order, x fastest); a thread locates its box once per chunk and
// walks forward from there, so the per-cell index work is 32-bit.
constexpr int MWS_CHUNK = 4096;

template <int D>
__global__ __launch_bounds__(256) void k_max_wave_speed(
    const double *__restrict__ U, const DevBox *__restrict__ boxes,
    const long long *__restrict__ cbeg, int nbox, long long ncells, int ng,
    EulerParm pm, unsigned long long *d_maxbits)
{
    constexpr int      NC   = D + 2;
    const int          ngy  = (D >= 2) ? ng : 0;
    const int          ngz  = (D >= 3) ? ng : 0;
    unsigned long long vmax = 0ULL;
    for (long long c0 = (long long)blockIdx.x * MWS_CHUNK; c0 < ncells;
         c0 += (long long)gridDim.x * MWS_CHUNK)
    {
        const long long cend = c0 + MWS_CHUNK < ncells ? c0 + MWS_CHUNK : ncells;
        const long long cfst = c0 + threadIdx.x;
        if (cfst >= cend)
            continue;
        // box geometry held in registers, reloaded only when the walk
        // crosses into the next box
        int       b    = find_box(cbeg, nbox, cfst);
        long long bbeg = cbeg[b], bend = cbeg[b + 1];
        DevBox    bx   = boxes[b];
        for (long long c = cfst; c < cend; c += blockDim.x)
        {
            if (c >= bend)
            {
                do
                {
                    ++b;
                    bbeg = bend;
                    bend = cbeg[b + 1];
                } while (c >= bend);
                bx = boxes[b];
            }
            const unsigned local = (unsigned)(c - bbeg);
            const unsigned nx    = (unsigned)(bx.hi[0] - bx.lo[0] + 1);
            const unsigned ny    = (D >= 2) ? (unsigned)(bx.hi[1] - bx.lo[1] + 1) : 1u;
            const unsigned i     = local % nx;
            const unsigned r     = local / nx;
            const unsigned j     = r % ny;
            const unsigned k     = r / ny;
            const long long o = bx.base + (i + ng) + (long long)(j + ngy) * bx.jstride
                                + (long long)(k + ngz) * bx.kstride;
            double V[NC];
#pragma unroll
            for (int n = 0; n < NC; ++n)
                V[n] = U[o + n * bx.nstride];
            vmax = speed_max(vmax, cell_wave_speed<D>(V, pm));
        }
    }
    block_max_to_global(vmax, d_maxbits);
}

// ----------------------------------------------------------- dt control
// estTimeStep (AmrLevelAdv.cpp:395-477) for the explicit methods, followed by
// the stop-time clamp of computeNewDt (AmrLevelAdv.cpp:589-618).  The
// accumulator is cleared for the next fused reduction.
__device__ void compute_dt(unsigned long long *d_maxbits, double *d_dt,
                           const double *d_time, int D, double dx0, double dx1,
                           double dx2, double cfl, double stop_time,
                           double fixed_dt)
{
    double dt_0;
    if (fixed_dt > 0)
    {
        dt_0 = fixed_dt;
    }
    else
    {
        const double max_speed = __longlong_as_double(*d_maxbits);
        const double dxs[3]    = { dx0, dx1, dx2 };
        double       dt_est    = 1.0e20;
        for (int d = 0; d < D; ++d)
        {
            const double c = dxs[d] / max_speed;
            dt_est         = (c < dt_est) ? c : dt_est;
        }
        dt_est *= cfl;
        dt_0 = 1.0e+100;
        dt_0 = (dt_est < dt_0) ? dt_est : dt_0;
    }
    const double eps      = 0.001 * dt_0;
    const double cur_time = *d_time;
    if (stop_time >= 0.0)
    {
        if ((cur_time + dt_0) > (stop_time - eps))
            dt_0 = stop_time - cur_time;
    }
    *d_dt      = dt_0;
    *d_maxbits = 0ULL;
}

__global__ void k_compute_dt(unsigned long long *d_maxbits, double *d_dt,
                             const double *d_time, int D, double dx0,
                             double dx1, double dx2, double cfl,
                             double stop_time, double fixed_dt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    compute_dt(d_maxbits, d_dt, d_time, D, dx0, dx1, dx2, cfl, stop_time,
               fixed_dt);
}

// End of a step, run by one thread of the last ghost fill: advance the time
// by the step's dt, then the dt of the next step from the maximum the last
// sweep accumulated (the stop-time clamp uses the advanced time).
__device__ void step_tail(const StepTail &t)
{
    *t.d_time = *t.d_time + *t.d_dt;
    *t.d_step = *t.d_step + 1;
    compute_dt(t.d_maxbits, t.d_dt, t.d_time, t.D, t.dx[0], t.dx[1], t.dx[2],
               t.cfl, t.stop_time, -1.0);
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

#endif // BX_SWEEP_PART

// ------------------------------------------------------------ launchers
template <int D, int LIM, int LV, int SCH, int PM>
static hipError_t launch_x(const SweepLaunch &L)
{
    hipLaunchKernelGGL((k_sweep_x<D, LIM, LV, SCH, PM>), dim3(L.grid),
                       dim3(L.block), x_lds_bytes(D + 2), L.stream,
                       L.Uin, L.Uout, L.boxes,
                       L.item_begin, L.xbox, L.nitems, L.ng, L.d_dt, L.dx,
                       L.pm, L.lt, L.d_maxbits);
    return hipGetLastError();
}

template <int D, int DIR, int LIM, int LV, int SCH, int PM>
static hipError_t launch_t(const SweepLaunch &L)
{
    hipLaunchKernelGGL((k_sweep_t<D, DIR, LIM, LV, SCH, PM>), dim3(L.grid),
                       dim3(L.block), 0, L.stream, L.Uin, L.Uout, L.boxes,
                       L.item_begin, L.nbox, L.nitems, L.ng, L.K, L.d_dt,
                       L.dx, L.pm, L.lt, L.d_maxbits);
    return hipGetLastError();
}

template <int D, int LIM, int LV, int SCH, int PM>
static hipError_t dispatch_dir_e(int dir, const SweepLaunch &L)
{
    if (dir == 0)
        return launch_x<D, LIM, LV, SCH, PM>(L);
    if constexpr (D >= 2)
        if (dir == 1)
            return launch_t<D, 1, LIM, LV, SCH, PM>(L);
    if constexpr (D >= 3)
        if (dir == 2)
            return launch_t<D, 2, LIM, LV, SCH, PM>(L);
    return hipErrorInvalidValue;
}

// Physics mode.  The HLLC-based schemes take the compile-time n = 7
// near-integer pow forms when gamma gives n = 7 (PM_NI) and drop epsilon
// when it is 1 (PM_E1); any other gamma runs the dispatching forms (0),
// which choose the general or the run-time-n near-integer routine.  FORCE,
// LF and SLIC use no pow, so only epsilon == 1 (PM_NI | PM_E1) or not (0)
// matters.
template <int D, int LIM, int LV, int SCH>
static hipError_t dispatch_pm(int dir, const SweepLaunch &L)
{
    if constexpr (scheme_uses_hllc(SCH))
    {
        if (L.pm.ph.n != 7) // dispatching pow (and any epsilon)
            return dispatch_dir_e<D, LIM, LV, SCH, 0>(dir, L);
        if (L.pm.eps == 1.0)
            return dispatch_dir_e<D, LIM, LV, SCH, PM_NI | PM_E1>(dir, L);
        return dispatch_dir_e<D, LIM, LV, SCH, PM_NI>(dir, L);
    }
    else
    {
        if (L.pm.eps == 1.0)
            return dispatch_dir_e<D, LIM, LV, SCH, PM_NI | PM_E1>(dir, L);
        return dispatch_dir_e<D, LIM, LV, SCH, 0>(dir, L);
    }
}

template <int D, int LIM, int SCH>
static hipError_t dispatch_lv(int dir, int lv, const SweepLaunch &L)
{
    switch (lv)
    {
    case LV_SHARED: return dispatch_pm<D, LIM, LV_SHARED, SCH>(dir, L);
    case LV_OWN: return dispatch_pm<D, LIM, LV_OWN, SCH>(dir, L);
    default: return dispatch_pm<D, LIM, LV_TABLE, SCH>(dir, L);
    }
}

template <int D, int SCH>
static hipError_t dispatch_lim(int dir, int lim, int lv, const SweepLaunch &L)
{
    if constexpr (!scheme_reconstructs(SCH)) // no reconstruction: no limiter
        return dispatch_pm<D, LIM_MINBEE, LV_SHARED, SCH>(dir, L);
    else
        switch (lim)
        {
        case LIM_MINBEE: return dispatch_lv<D, LIM_MINBEE, SCH>(dir, lv, L);
        case LIM_VAN_LEER: return dispatch_lv<D, LIM_VAN_LEER, SCH>(dir, lv, L);
        case LIM_VAN_ALBADA: return dispatch_lv<D, LIM_VAN_ALBADA, SCH>(dir, lv, L);
        case LIM_SUPERBEE: return dispatch_lv<D, LIM_SUPERBEE, SCH>(dir, lv, L);
        default: return hipErrorInvalidValue;
        }
}

template <int D>
static hipError_t dispatch_scheme(int dir, int sch, int lim, int lv,
                                  const SweepLaunch &L)
{
    switch (sch)
    {
    case SCH_MHM: return dispatch_lim<D, SCH_MHM>(dir, lim, lv, L);
    case SCH_HLLC: return dispatch_lim<D, SCH_HLLC>(dir, lim, lv, L);
    case SCH_SLIC: return dispatch_lim<D, SCH_SLIC>(dir, lim, lv, L);
    case SCH_FORCE: return dispatch_lim<D, SCH_FORCE>(dir, lim, lv, L);
    case SCH_LF: return dispatch_lim<D, SCH_LF>(dir, lim, lv, L);
    default: return hipErrorInvalidValue;
    }
}

// One dimension's instantiations per translation unit: bx_explicit.hip
// holds 1D and everything else, bx_sweep2d.hip / bx_sweep3d.hip include this
// file with BX_SWEEP_PART 2 / 3, so the three build in parallel.
#if !defined(BX_SWEEP_PART)
hipError_t launch_sweep_d1(int dir, int sch, int lim, int lv, const SweepLaunch &L)
{
    return dispatch_scheme<1>(dir, sch, lim, lv, L);
}
const void *sweep_kernel_ptr_d1(int)
{
    return (const void *)&k_sweep_x<1, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
}
#elif BX_SWEEP_PART == 2
hipError_t launch_sweep_d2(int dir, int sch, int lim, int lv, const SweepLaunch &L)
{
    return dispatch_scheme<2>(dir, sch, lim, lv, L);
}
const void *sweep_kernel_ptr_d2(int dir)
{
    if (dir == 0)
        return (const void *)&k_sweep_x<2, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
    return (const void *)&k_sweep_t<2, 1, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
}
#elif BX_SWEEP_PART == 3
hipError_t launch_sweep_d3(int dir, int sch, int lim, int lv, const SweepLaunch &L)
{
    return dispatch_scheme<3>(dir, sch, lim, lv, L);
}
const void *sweep_kernel_ptr_d3(int dir)
{
    if (dir == 0)
        return (const void *)&k_sweep_x<3, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
    if (dir == 1)
        return (const void *)&k_sweep_t<3, 1, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
    return (const void *)&k_sweep_t<3, 2, LIM_MINBEE, LV_SHARED, SCH_MHM, PM_NI | PM_E1>;
}
#endif

#ifndef BX_SWEEP_PART
hipError_t launch_sweep(int D, int dir, int scheme, int lim, int lv,
                        const SweepLaunch &L)
{
    switch (D)
    {
    case 1: return launch_sweep_d1(dir, scheme, lim, lv, L);
    case 2: return launch_sweep_d2(dir, scheme, lim, lv, L);
    case 3: return launch_sweep_d3(dir, scheme, lim, lv, L);
    default: return hipErrorInvalidValue;
    }
}

const void *sweep_kernel_ptr(int D, int dir)
{
    // representative instantiation for occupancy queries
    switch (D)
    {
    case 1: return sweep_kernel_ptr_d1(dir);
    case 2: return sweep_kernel_ptr_d2(dir);
    default: return sweep_kernel_ptr_d3(dir);
    }
}

hipError_t launch_fill_ghosts(double *U, const DevBox *boxes,
                              const GhostSeg *segs, int nseg, int ng, int D,
                              int NC, int grid, const StepTail &tail,
                              hipStream_t stream)
{
    if (nseg == 0 && !tail.d_time)
        return hipSuccess;
    hipLaunchKernelGGL(k_fill_ghosts, dim3(grid), dim3(256), 0, stream, U,
                       boxes, segs, nseg, ng, D, NC, tail);
    return hipGetLastError();
}

hipError_t launch_max_wave_speed(int D, const double *U, const DevBox *boxes,
                                 const long long *cbeg, int nbox,
                                 long long ncells, int ng, EulerParm pm,
                                 unsigned long long *d_maxbits, int grid,
                                 hipStream_t stream)
{
    switch (D)
    {
    case 1:
        hipLaunchKernelGGL(k_max_wave_speed<1>, dim3(grid), dim3(256), 0,
                           stream, U, boxes, cbeg, nbox, ncells, ng, pm,
                           d_maxbits);
        break;
    case 2:
        hipLaunchKernelGGL(k_max_wave_speed<2>, dim3(grid), dim3(256), 0,
                           stream, U, boxes, cbeg, nbox, ncells, ng, pm,
                           d_maxbits);
        break;
    case 3:
        hipLaunchKernelGGL(k_max_wave_speed<3>, dim3(grid), dim3(256), 0,
                           stream, U, boxes, cbeg, nbox, ncells, ng, pm,
                           d_maxbits);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_compute_dt(unsigned long long *d_maxbits, double *d_dt,
                             const double *d_time, int D, const double dx[3],
                             double cfl, double stop_time, double fixed_dt,
                             hipStream_t stream)
{
    hipLaunchKernelGGL(k_compute_dt, dim3(1), dim3(64), 0, stream, d_maxbits,
                       d_dt, d_time, D, dx[0], dx[1], dx[2], cfl, stop_time,
                       fixed_dt);
    return hipGetLastError();
}

hipError_t launch_set_u64(unsigned long long *p, unsigned long long v,
                          hipStream_t stream)
{
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(64), 0, stream, p, v);
    return hipGetLastError();
}

#endif // BX_SWEEP_PART

} // namespace bx
