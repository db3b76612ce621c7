// Structures shared by the host planner, the HIP kernels and the C ABI.
#pragma once
#include <stdint.h>

namespace bx {

// Longest segment (cells per lane along y/z) a launch shape may use.
constexpr int BX_MAX_SEG_K = 4096;

// One box of a level as seen by the kernels.  The FAB of the box covers
// grow(valid, ng) (ng only in the first D directions), component-major with
// x fastest -- the amrex::Array4 layout:
//   elem(i,j,k,n) = base + (i-lo0+ng) + (j-lo1+ngy)*jstride
//                        + (k-lo2+ngz)*kstride + n*nstride
struct DevBox
{
    int       lo[3];
    int       hi[3];
    int       jstride;
    int       pad0;
    long long kstride;
    long long nstride;
    long long base; // element offset of (lo-ng, comp 0) from the level pointer
};

// A rectangular ghost-cell copy: every cell x of [lo,hi] (global indices) of
// box dst takes the value of cell (a*x + b) (per direction) of box src.
// flip_mask bit n set => component n changes sign (reflect_odd).
struct GhostSeg
{
    int dst;
    int src;
    int lo[3];
    int hi[3];
    int a[3];
    int b[3];
    int flip_mask;
    int pad0;
    long long cell_begin; // prefix of cell counts over segments
};

// Work decomposition of one sweep kernel: item_begin[b] = first work item
// of box b (exclusive prefix sum, nbox+1 entries).
struct SweepPlan
{
    long long *d_item_begin = nullptr;
    long long  nitems       = 0;
    int        chunk        = 0; // positions (x) or rows per segment (y/z)
};

} // namespace bx
