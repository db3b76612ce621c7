// Host-side launch interface of the explicit-path kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "bx_internal.h"
#include "bx_physics.h"

namespace bx {

// Optional end-of-step work of a ghost fill (d_time == nullptr: none): one
// thread advances the time by dt, then computes the next dt from the
// maximum the last sweep accumulated.
struct StepTail
{
    unsigned long long *d_maxbits = nullptr;
    double             *d_dt      = nullptr;
    double             *d_time    = nullptr;
    long long          *d_step    = nullptr;
    int                 D         = 0;
    double              dx[3]     = { 0, 0, 0 };
    double              cfl       = 0;
    double              stop_time = -1;
};

// x sweep: an item is B consecutive positions of a box plane (B = block
// size, at most X_LDS_STRIDE), of which B - 4 are updated; the block
// exchanges neighbour values through three LDS buffers (U, high-side half
// state, face flux) of NC rows of X_LDS_STRIDE doubles, allocated
// dynamically.  The compile-time row stride keeps every LDS address an
// immediate offset from the thread's own.
constexpr int X_LDS_STRIDE = 256;
constexpr int x_item_step(int B) { return B - 4; }
constexpr int x_lds_bytes(int NC) { return 3 * NC * X_LDS_STRIDE * 8; }
// consecutive x-sweep items a block takes before the next block's turn
constexpr int X_CHUNK = 8;

struct SweepLaunch
{
    const double       *Uin;
    double             *Uout;
    const DevBox       *boxes;
    const long long    *item_begin;
    const int          *xbox = nullptr; // x sweep: box of each chunk's first item
    int                 nbox;
    long long           nitems;
    int                 ng;
    int                 K; // cells per lane segment (y/z sweeps)
    const double       *d_dt;
    double              dx;
    EulerParm           pm;
    LimTable            lt;
    unsigned long long *d_maxbits; // nullable: fused CFL reduction target
    int                 grid;
    int                 block;
    hipStream_t         stream;
};

// scheme: SCH_* (bx_physics.h), the BX_METHOD_* code
hipError_t  launch_sweep(int D, int dir, int scheme, int lim, int lv,
                         const SweepLaunch &L);
const void *sweep_kernel_ptr(int D, int dir);
// per-dimension instantiations (bx_explicit.hip, bx_sweep2d.hip, bx_sweep3d.hip)
hipError_t  launch_sweep_d1(int dir, int scheme, int lim, int lv, const SweepLaunch &L);
hipError_t  launch_sweep_d2(int dir, int scheme, int lim, int lv, const SweepLaunch &L);
hipError_t  launch_sweep_d3(int dir, int scheme, int lim, int lv, const SweepLaunch &L);
const void *sweep_kernel_ptr_d1(int dir);
const void *sweep_kernel_ptr_d2(int dir);
const void *sweep_kernel_ptr_d3(int dir);

hipError_t launch_fill_ghosts(double *U, const DevBox *boxes,
                              const GhostSeg *segs, int nseg, int ng, int D,
                              int NC, int grid, const StepTail &tail,
                              hipStream_t stream);

hipError_t launch_max_wave_speed(int D, const double *U, const DevBox *boxes,
                                 const long long *cbeg, int nbox,
                                 long long ncells, int ng, EulerParm pm,
                                 unsigned long long *d_maxbits, int grid,
                                 hipStream_t stream);

hipError_t launch_compute_dt(unsigned long long *d_maxbits, double *d_dt,
                             const double *d_time, int D, const double dx[3],
                             double cfl, double stop_time, double fixed_dt,
                             hipStream_t stream);

hipError_t launch_set_u64(unsigned long long *p, unsigned long long v,
                          hipStream_t stream);

} // namespace bx
