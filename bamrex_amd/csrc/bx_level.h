// Level context of the explicit path, shared by the C ABI (bx_api.hip) and
// the two-level driver (bx_amr.hip).  Internal: not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/bamrex_hip.h"
#include "bx_internal.h"
#include "bx_kernels.h"
#include "bx_physics.h"

namespace bx {
// Records msg as the text bx_last_error returns and gives back code.
int set_error(int code, const std::string &msg);
} // namespace bx

#define BX_HIP_CHECK(expr)                                                      \
    do                                                                          \
    {                                                                           \
        hipError_t e__ = (expr);                                                \
        if (e__ != hipSuccess)                                                  \
            return bx::set_error(BX_ERR_HIP, std::string(#expr) + ": "          \
                                                 + hipGetErrorString(e__));     \
    } while (0)

using namespace bx;

// ------------------------------------------------------------- context
struct StepGraph
{
    hipGraphExec_t exec = nullptr;
    hipGraph_t     graph = nullptr;
    const double  *a = nullptr, *b = nullptr;
    double         dx[3] = { 0, 0, 0 };
    double         cfl = 0, stop = 0;
    void          *stream = nullptr;
};

struct bx_ctx
{
    bx_params p;
    int       D = 0, NC = 0;
    EulerParm pm;
    LimTable  lt;
    int       lv_mode = LV_SHARED;

    bool                 level_set = false;
    bx_geometry          geom;
    std::vector<bx_box>  hboxes;
    std::vector<DevBox>  hdev;
    DevBox              *d_boxes = nullptr;
    int64_t              level_extent = 0; // doubles spanned by all FABs

    std::vector<long long> ib_x, ib_t[3], cbeg;
    std::vector<int>       xbox; // box of the first item of each x-sweep chunk
    int                   *d_xbox = nullptr;
    long long *d_ib_x = nullptr, *d_ib_t[3] = { nullptr, nullptr, nullptr },
              *d_cbeg = nullptr;
    long long  ncells = 0;

    // [0..D-1]: ghosts in dir d only; [3]: rest; [4]: all of them (each
    // segment copies valid cells only, so one launch fills every ghost)
    std::vector<GhostSeg> segs[5];
    GhostSeg             *d_segs[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
    long long             gcells[5] = { 0, 0, 0, 0, 0 };

    unsigned long long *d_maxbits = nullptr;
    double             *d_dt = nullptr, *d_time = nullptr;
    long long          *d_step = nullptr;

    bx_tuning tune;
    int       ncu   = 0;
    int       occ_x = 1, occ_t = 1;

    StepGraph graphs[2];
};


// Step pieces (bx_api.hip).  dx: the level's cell sizes; all launches are
// asynchronous on s.
// One sweep in direction dir from Uin into the valid cells of Uout (dt from
// the level's device dt); maxbits != nullptr also reduces the wave speed.
int do_sweep(bx_ctx *c, int dir, const double dx[3], const double *Uin,
             double *Uout, unsigned long long *maxbits, hipStream_t s);
// Ghost cells of one segment list: 0..D-1 = ghosts in that direction only,
// 3 = edges and corners.
int do_fill(bx_ctx *c, double *U, int list, hipStream_t s,
            const bx::StepTail &tail = bx::StepTail());
int do_fill_all(bx_ctx *c, double *U, hipStream_t s);
// Maximum wave speed of the valid cells into c->d_maxbits.
int do_maxspeed(bx_ctx *c, const double *U, hipStream_t s);
bool dx_ok(const bx_ctx *c, const double dx[3]);
int  set_scalar(double *p, double v, hipStream_t s);
