// C ABI of the explicit hot path (declared in include/bamrex_hip.h).
//
// Host side: level planning (device box list, persistent-kernel work tables,
// ghost-copy segments), the time loop with device-side dt control, hipGraph
// capture of whole steps, and the launch-shape autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bamrex_hip.h"
#include "bx_internal.h"
#include "bx_kernels.h"
#include "bx_level.h"
#include "bx_physics.h"

using namespace bx;

// ------------------------------------------------------------ error state
static thread_local std::string g_last_error;

int bx::set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

extern "C" const char *bx_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------ per-box kernels
// Drop-in of compute_MUSCL_hancock_flux / compute_HLLC_flux /
// compute_SLIC_flux / compute_force_flux / compute_LF_flux for one box: one
// thread per face of surroundingNodes(bx, dir).
template <int D, int DIR, int LIM, int LV, int SCH>
__global__ __launch_bounds__(256) void k_flux_box(
    const double *__restrict__ U, DevBox bx, int ng, double *__restrict__ flux,
    double hdtdx, double fc, double rc, EulerParm pm, LimTable lt)
{
    constexpr int NC = D + 2;
    const int     fx = bx.hi[0] - bx.lo[0] + 1 + (DIR == 0);
    const int     fy = (D >= 2) ? bx.hi[1] - bx.lo[1] + 1 + (DIR == 1) : 1;
    const int     fz = (D >= 3) ? bx.hi[2] - bx.lo[2] + 1 + (DIR == 2) : 1;
    const long long nf = (long long)fx * fy * fz;
    const int       ngy = (D >= 2) ? ng : 0;
    const int       ngz = (D >= 3) ? ng : 0;
    for (long long f = (long long)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
         f += (long long)gridDim.x * blockDim.x)
    {
        const int i = (int)(f % fx);
        const int j = (int)((f / fx) % fy);
        const int k = (int)(f / ((long long)fx * fy));
        // cell (i,j,k) is the high-side cell of the face (local indices)
        const long long sd = (DIR == 0) ? 1 : (DIR == 1 ? bx.jstride : bx.kstride);
        const long long c  = bx.base + (i + ng) + (long long)(j + ngy) * bx.jstride
                            + (long long)(k + ngz) * bx.kstride;
        double W[4][NC]; // U at c-2, c-1, c, c+1
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
            reconstruct_half_step<D, DIR, LIM, LV>(W[0], W[1], W[2], hL0, hR0,
                                                   hdtdx, pm, lt);
            reconstruct_half_step<D, DIR, LIM, LV>(W[1], W[2], W[3], hL1, hR1,
                                                   hdtdx, pm, lt);
            // swapped arrays: left state = half_R(c-1), right = half_L(c)
            face_flux<D, DIR, false, false, SCH>(hR0, hL1, F, pm, fc, rc);
        }
        const long long fvol = nf;
#pragma unroll
        for (int n = 0; n < NC; ++n)
            flux[f + n * fvol] = F[n];
    }
}

template <int D, int DIR>
__global__ __launch_bounds__(256) void k_update_box(
    const double *__restrict__ Uin, double *__restrict__ Uout, DevBox bx,
    int ng, const double *__restrict__ flux, double dt, double dxd)
{
    constexpr int NC = D + 2;
    const int     nx = bx.hi[0] - bx.lo[0] + 1;
    const int     ny = (D >= 2) ? bx.hi[1] - bx.lo[1] + 1 : 1;
    const int     nz = (D >= 3) ? bx.hi[2] - bx.lo[2] + 1 : 1;
    const int     fx = nx + (DIR == 0), fy = ny + (DIR == 1), fz = nz + (DIR == 2);
    const long long fvol = (long long)fx * fy * fz;
    const long long nc   = (long long)nx * ny * nz;
    const int       ngy  = (D >= 2) ? ng : 0;
    const int       ngz  = (D >= 3) ? ng : 0;
    const int       off  = (DIR == 0) ? 1 : (DIR == 1 ? fx : fx * fy);
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
         c += (long long)gridDim.x * blockDim.x)
    {
        const int       i  = (int)(c % nx);
        const int       j  = (int)((c / nx) % ny);
        const int       k  = (int)(c / ((long long)nx * ny));
        const long long u  = bx.base + (i + ng) + (long long)(j + ngy) * bx.jstride
                            + (long long)(k + ngz) * bx.kstride;
        const long long fo = i + (long long)j * fx + (long long)k * fx * fy;
#pragma unroll
        for (int n = 0; n < NC; ++n)
            Uout[u + n * bx.nstride]
                = Uin[u + n * bx.nstride]
                  - (dt / dxd) * (flux[fo + off + n * fvol] - flux[fo + n * fvol]);
    }
}

// Plain streaming copy (calibrates the HBM byte counters against a kernel
// of known traffic: reads and writes `n` doubles once).
__global__ __launch_bounds__(256) void k_copy(const double *__restrict__ src,
                                              double *__restrict__ dst,
                                              long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_set_f64(double *p, double v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0)
        *p = v;
}

static void free_dev(void *p)
{
    if (p)
        (void)hipFree(p);
}

template <class T>
static int upload(T **dptr, const std::vector<T> &v)
{
    free_dev(*dptr);
    *dptr = nullptr;
    if (v.empty())
        return BX_OK;
    BX_HIP_CHECK(hipMalloc((void **)dptr, sizeof(T) * v.size()));
    BX_HIP_CHECK(hipMemcpy(*dptr, v.data(), sizeof(T) * v.size(),
                           hipMemcpyHostToDevice));
    return BX_OK;
}

static void destroy_graphs(bx_ctx *c)
{
    for (auto &g : c->graphs)
    {
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
        if (g.graph)
            (void)hipGraphDestroy(g.graph);
        g = StepGraph();
    }
}

extern "C" int bx_create(const bx_params *params, bx_ctx **out)
{
    if (!params || !out)
        return set_error(BX_ERR_INVALID, "bx_create: null argument");
    const bx_params &p = *params;
    if (p.dim < 1 || p.dim > 3)
        return set_error(BX_ERR_INVALID, "bx_create: dim must be 1, 2 or 3");
    if (p.num_method < BX_METHOD_MUSCL_HANCOCK || p.num_method > BX_METHOD_LAX_FRIEDRICHS)
        return set_error(BX_ERR_INVALID,
                         "Invalid numerical method: valid explicit options are "
                         "muscl_hancock, hllc, slic, force and lax_friedrichs");
    if (p.limiter < BX_LIMITER_MINBEE || p.limiter > BX_LIMITER_SUPERBEE)
        return set_error(BX_ERR_INVALID, "bx_create: invalid limiter");
    if (!(p.adiabatic > 1.0))
        return set_error(BX_ERR_INVALID,
                         "bx_create: prob.adiabatic must be set (> 1)");
    if (!(p.epsilon > 0.0))
        return set_error(BX_ERR_INVALID, "bx_create: epsilon must be > 0");
    const int NC = p.dim + 2;
    for (int n = 0; n < NC; ++n)
    {
        if (p.limiter_indices[n] < 0 || p.limiter_indices[n] >= NC)
            return set_error(BX_ERR_UNSUPPORTED,
                             "bx_create: limiting variable index out of range "
                             "(only density/mom/energy/u/p are supported)");
    }
    bx_ctx *c = new bx_ctx();
    c->p      = p;
    c->D      = p.dim;
    c->NC     = NC;
    c->pm     = EulerParm{ p.adiabatic, p.epsilon, bx_pow_hllc_setup(p.adiabatic) };
    // LV_SHARED: every component limited on the conserved energy
    bool shared = true, own = true;
    for (int n = 0; n < NC; ++n)
    {
        c->lt.idx[n]  = p.limiter_indices[n];
        c->lt.prim[n] = p.limit_on_primv[n] ? 1 : 0;
        if (p.limiter_indices[n] != NC - 1 || p.limit_on_primv[n])
            shared = false;
        if (p.limiter_indices[n] != n || p.limit_on_primv[n])
            own = false;
    }
    for (int n = NC; n < 5; ++n)
    {
        c->lt.idx[n]  = 0;
        c->lt.prim[n] = 0;
    }
    c->lv_mode = shared ? LV_SHARED : (own ? LV_OWN : LV_TABLE);

    hipDeviceProp_t prop;
    int             dev = 0;
    if (hipGetDevice(&dev) != hipSuccess
        || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    {
        delete c;
        return set_error(BX_ERR_HIP, "bx_create: no HIP device");
    }
    c->ncu = prop.multiProcessorCount;
    c->tune.block_x  = 256;
    c->tune.block_t  = 256;
    c->tune.seg_k    = 16;
    c->tune.grid_cap = 0; // 0 = occupancy limit
    if (hipMalloc((void **)&c->d_maxbits, 64) != hipSuccess
        || hipMalloc((void **)&c->d_dt, 64) != hipSuccess
        || hipMalloc((void **)&c->d_time, 64) != hipSuccess
        || hipMalloc((void **)&c->d_step, 64) != hipSuccess)
    {
        bx_destroy(c);
        return set_error(BX_ERR_HIP, "bx_create: hipMalloc failed");
    }
    (void)hipMemset(c->d_maxbits, 0, 64);
    (void)hipMemset(c->d_dt, 0, 64);
    (void)hipMemset(c->d_time, 0, 64);
    (void)hipMemset(c->d_step, 0, 64);
    *out = c;
    return BX_OK;
}

extern "C" void bx_destroy(bx_ctx *c)
{
    if (!c)
        return;
    destroy_graphs(c);
    free_dev(c->d_boxes);
    free_dev(c->d_ib_x);
    free_dev(c->d_xbox);
    for (auto &p : c->d_ib_t)
        free_dev(p);
    free_dev(c->d_cbeg);
    for (auto &p : c->d_segs)
        free_dev(p);
    free_dev(c->d_maxbits);
    free_dev(c->d_dt);
    free_dev(c->d_time);
    free_dev(c->d_step);
    delete c;
}

// ------------------------------------------------------------ planning
static int box_n(const bx_box &b, int d) { return b.hi[d] - b.lo[d] + 1; }

extern "C" int bx_pack_offsets(int dim, int ng, const bx_box *boxes, int nboxes,
                               int64_t *offsets_out, int64_t *total_doubles)
{
    if (dim < 1 || dim > 3 || ng < 0 || nboxes < 0 || (nboxes && !boxes)
        || !total_doubles)
        return set_error(BX_ERR_INVALID, "bx_pack_offsets: bad argument");
    int64_t off = 0;
    for (int b = 0; b < nboxes; ++b)
    {
        int64_t vol = 1;
        for (int d = 0; d < dim; ++d)
            vol *= (int64_t)(box_n(boxes[b], d) + 2 * ng);
        if (offsets_out)
            offsets_out[b] = off;
        off += vol * (dim + 2);
        off = (off + 31) / 32 * 32;
    }
    *total_doubles = off;
    return BX_OK;
}

static void build_sweep_tables(bx_ctx *c)
{
    const int D = c->D, ng = c->geom.ng;
    const int nb = (int)c->hboxes.size();
    c->ib_x.assign(nb + 1, 0);
    for (int d = 0; d < 3; ++d)
        c->ib_t[d].assign(nb + 1, 0);
    c->cbeg.assign(nb + 1, 0);
    const int K    = c->tune.seg_k;
    const int xstp = x_item_step(c->tune.block_x);
    for (int b = 0; b < nb; ++b)
    {
        const bx_box &bx  = c->hboxes[b];
        const long long nx = box_n(bx, 0);
        const long long ny = D >= 2 ? box_n(bx, 1) : 1;
        const long long nz = D >= 3 ? box_n(bx, 2) : 1;
        const long long P  = ny * (nx + 2 * ng);
        c->ib_x[b + 1]     = c->ib_x[b] + nz * ((P + xstp - 1) / xstp);
        if (D >= 2)
        {
            const long long q = nx * nz * ((ny + K - 1) / K);
            c->ib_t[1][b + 1] = c->ib_t[1][b] + (q + 63) / 64;
        }
        if (D >= 3)
        {
            const long long q = nx * ny * ((nz + K - 1) / K);
            c->ib_t[2][b + 1] = c->ib_t[2][b] + (q + 63) / 64;
        }
        c->cbeg[b + 1] = c->cbeg[b] + nx * ny * nz;
    }
    c->ncells = c->cbeg[nb];
    const long long nch = (c->ib_x[nb] + X_CHUNK - 1) / X_CHUNK;
    c->xbox.assign(nch, 0);
    for (long long ch = 0, b = 0; ch < nch; ++ch)
    {
        while (c->ib_x[b + 1] <= ch * X_CHUNK)
            ++b;
        c->xbox[ch] = (int)b;
    }
}

static int upload_sweep_tables(bx_ctx *c)
{
    int rc;
    if ((rc = upload(&c->d_ib_x, c->ib_x)) != BX_OK)
        return rc;
    if ((rc = upload(&c->d_xbox, c->xbox)) != BX_OK)
        return rc;
    for (int d = 1; d < c->D; ++d)
        if ((rc = upload(&c->d_ib_t[d], c->ib_t[d])) != BX_OK)
            return rc;
    return upload(&c->d_cbeg, c->cbeg);
}

struct Piece
{
    int lo, hi, a, b, flip;
};

static void build_ghost_segments(bx_ctx *c)
{
    const int            D  = c->D;
    const int            ng = c->geom.ng;
    const bx_geometry   &g  = c->geom;
    const auto          &B  = c->hboxes;
    const int            nb = (int)B.size();
    for (auto &s : c->segs)
        s.clear();

    for (int ib = 0; ib < nb; ++ib)
    {
        const bx_box &bx = B[ib];
        int           part[3];
        const int     ncomb = (D == 1) ? 3 : (D == 2 ? 9 : 27);
        for (int comb = 0; comb < ncomb; ++comb)
        {
            int t = comb, nz = 0, cls = -1;
            for (int d = 0; d < 3; ++d)
            {
                if (d < D)
                {
                    part[d] = (t % 3) - 1;
                    t /= 3;
                }
                else
                    part[d] = 0;
                if (part[d] != 0)
                {
                    ++nz;
                    cls = d;
                }
            }
            if (nz == 0)
                continue;
            const int list = (nz == 1) ? cls : 3;
            // pieces per direction
            std::vector<Piece> pcs[3];
            for (int d = 0; d < 3; ++d)
            {
                int r0, r1;
                if (d >= D)
                {
                    r0 = r1 = 0;
                    pcs[d].push_back({ 0, 0, 1, 0, 0 });
                    continue;
                }
                if (part[d] < 0)
                {
                    r0 = bx.lo[d] - ng;
                    r1 = bx.lo[d] - 1;
                }
                else if (part[d] > 0)
                {
                    r0 = bx.hi[d] + 1;
                    r1 = bx.hi[d] + ng;
                }
                else
                {
                    r0 = bx.lo[d];
                    r1 = bx.hi[d];
                }
                const int dlo = g.domain_lo[d], dhi = g.domain_hi[d];
                const int L   = dhi - dlo + 1;
                // below the domain
                if (r0 < dlo)
                {
                    Piece p{ r0, std::min(r1, dlo - 1), 1, 0, 0 };
                    switch (g.bc_lo[d])
                    {
                    case BX_BC_PERIODIC: p.a = 1; p.b = L; break;
                    case BX_BC_REFLECTIVE:
                        p.a    = -1;
                        p.b    = 2 * dlo - 1;
                        p.flip = 1 << (1 + d);
                        break;
                    default: p.a = 0; p.b = dlo; break;
                    }
                    pcs[d].push_back(p);
                }
                const int i0 = std::max(r0, dlo), i1 = std::min(r1, dhi);
                if (i0 <= i1)
                    pcs[d].push_back({ i0, i1, 1, 0, 0 });
                if (r1 > dhi)
                {
                    Piece p{ std::max(r0, dhi + 1), r1, 1, 0, 0 };
                    switch (g.bc_hi[d])
                    {
                    case BX_BC_PERIODIC: p.a = 1; p.b = -L; break;
                    case BX_BC_REFLECTIVE:
                        p.a    = -1;
                        p.b    = 2 * dhi + 1;
                        p.flip = 1 << (1 + d);
                        break;
                    default: p.a = 0; p.b = dhi; break;
                    }
                    pcs[d].push_back(p);
                }
            }
            for (const Piece &px : pcs[0])
                for (const Piece &py : pcs[1])
                    for (const Piece &pz : pcs[2])
                    {
                        const Piece *pp[3] = { &px, &py, &pz };
                        int          ilo[3], ihi[3];
                        for (int d = 0; d < 3; ++d)
                        {
                            const int v0 = pp[d]->a * pp[d]->lo + pp[d]->b;
                            const int v1 = pp[d]->a * pp[d]->hi + pp[d]->b;
                            ilo[d]       = std::min(v0, v1);
                            ihi[d]       = std::max(v0, v1);
                        }
                        for (int is = 0; is < nb; ++is)
                        {
                            const bx_box &sb = B[is];
                            int           qlo[3], qhi[3];
                            bool          ok = true;
                            for (int d = 0; d < 3 && ok; ++d)
                            {
                                const int slo = d < D ? sb.lo[d] : 0;
                                const int shi = d < D ? sb.hi[d] : 0;
                                const int lo  = std::max(ilo[d], slo);
                                const int hi  = std::min(ihi[d], shi);
                                if (lo > hi)
                                {
                                    ok = false;
                                    break;
                                }
                                // preimage of [lo, hi]
                                if (pp[d]->a == 0)
                                {
                                    qlo[d] = pp[d]->lo;
                                    qhi[d] = pp[d]->hi;
                                }
                                else if (pp[d]->a == 1)
                                {
                                    qlo[d] = lo - pp[d]->b;
                                    qhi[d] = hi - pp[d]->b;
                                }
                                else
                                {
                                    qlo[d] = pp[d]->b - hi;
                                    qhi[d] = pp[d]->b - lo;
                                }
                            }
                            if (!ok)
                                continue;
                            GhostSeg s;
                            s.dst = ib;
                            s.src = is;
                            for (int d = 0; d < 3; ++d)
                            {
                                s.lo[d] = qlo[d];
                                s.hi[d] = qhi[d];
                                s.a[d]  = pp[d]->a;
                                s.b[d]  = pp[d]->b;
                            }
                            s.flip_mask  = px.flip | py.flip | pz.flip;
                            s.pad0       = 0;
                            s.cell_begin = 0;
                            c->segs[list].push_back(s);
                        }
                    }
        }
    }
    for (int l = 0; l < 4; ++l)
        c->segs[4].insert(c->segs[4].end(), c->segs[l].begin(), c->segs[l].end());
    for (int l = 0; l < 5; ++l)
    {
        long long acc = 0;
        for (auto &s : c->segs[l])
        {
            s.cell_begin = acc;
            acc += (long long)(s.hi[0] - s.lo[0] + 1) * (s.hi[1] - s.lo[1] + 1)
                   * (s.hi[2] - s.lo[2] + 1);
        }
        c->gcells[l] = acc;
    }
}

static int occupancy(bx_ctx *c, int dir, int block, int *occ)
{
    int n = 0;
    BX_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &n, sweep_kernel_ptr(c->D, dir), block,
        dir == 0 ? x_lds_bytes(c->NC) : 0));
    *occ = std::max(1, n);
    return BX_OK;
}

extern "C" int bx_set_level(bx_ctx *c, const bx_geometry *geom,
                            const bx_box *boxes, int nboxes)
{
    if (!c || !geom || (nboxes > 0 && !boxes) || nboxes <= 0)
        return set_error(BX_ERR_INVALID, "bx_set_level: bad argument");
    const int D = c->D;
    if (geom->ng < 2)
        return set_error(BX_ERR_INVALID,
                         "bx_set_level: MUSCL-Hancock needs ng >= 2 "
                         "(AMREX_ASSERT(NUM_GROW >= 2), MUSCL_Hancock.cpp:27)");
    for (int d = 0; d < D; ++d)
    {
        if (geom->domain_hi[d] < geom->domain_lo[d])
            return set_error(BX_ERR_INVALID, "bx_set_level: empty domain");
        const bool pl = geom->bc_lo[d] == BX_BC_PERIODIC;
        const bool ph = geom->bc_hi[d] == BX_BC_PERIODIC;
        if (pl != ph)
            return set_error(BX_ERR_INVALID,
                             "bx_set_level: periodic must be set on both sides");
        for (int s = 0; s < 2; ++s)
        {
            const int t = s ? geom->bc_hi[d] : geom->bc_lo[d];
            if (t < BX_BC_TRANSMISSIVE || t > BX_BC_REFLECTIVE)
                return set_error(BX_ERR_INVALID,
                                 "Boundary type not implemented!");
        }
    }
    destroy_graphs(c);
    c->geom = *geom;
    c->hboxes.assign(boxes, boxes + nboxes);
    c->hdev.resize(nboxes);
    const int ng = geom->ng;
    c->level_extent = 0;
    for (int b = 0; b < nboxes; ++b)
    {
        bx_box &bx = c->hboxes[b];
        for (int d = D; d < 3; ++d)
        {
            bx.lo[d] = 0;
            bx.hi[d] = 0;
        }
        for (int d = 0; d < D; ++d)
        {
            if (bx.hi[d] < bx.lo[d])
                return set_error(BX_ERR_INVALID, "bx_set_level: empty box");
            if (bx.lo[d] < geom->domain_lo[d] || bx.hi[d] > geom->domain_hi[d])
                return set_error(BX_ERR_INVALID,
                                 "bx_set_level: box outside the domain");
        }
        DevBox &db       = c->hdev[b];
        const int nxd    = box_n(bx, 0) + 2 * ng;
        const int nyd    = D >= 2 ? box_n(bx, 1) + 2 * ng : 1;
        const int nzd    = D >= 3 ? box_n(bx, 2) + 2 * ng : 1;
        for (int d = 0; d < 3; ++d)
        {
            db.lo[d] = bx.lo[d];
            db.hi[d] = bx.hi[d];
        }
        db.jstride = nxd;
        db.pad0    = 0;
        db.kstride = (long long)nxd * nyd;
        db.nstride = db.kstride * nzd;
        db.base    = bx.offset;
        if (bx.offset < 0)
            return set_error(BX_ERR_INVALID, "bx_set_level: negative offset");
        c->level_extent
            = std::max<int64_t>(c->level_extent, bx.offset + db.nstride * c->NC);
    }
    int rc;
    if ((rc = upload(&c->d_boxes, c->hdev)) != BX_OK)
        return rc;
    build_sweep_tables(c);
    if ((rc = upload_sweep_tables(c)) != BX_OK)
        return rc;
    build_ghost_segments(c);
    for (int l = 0; l < 5; ++l)
        if ((rc = upload(&c->d_segs[l], c->segs[l])) != BX_OK)
            return rc;
    if ((rc = occupancy(c, 0, c->tune.block_x, &c->occ_x)) != BX_OK)
        return rc;
    if (D >= 2 && (rc = occupancy(c, 1, c->tune.block_t, &c->occ_t)) != BX_OK)
        return rc;
    c->level_set = true;
    return BX_OK;
}

// ------------------------------------------------------------- launches
// Occupancy-limited grid.  Items per block: one at a time per wave
// (wave_items, the y/z sweeps) or per block (the x sweep).
static int persistent_grid(const bx_ctx *c, long long nitems, int block, int occ,
                           bool wave_items)
{
    const int       wpb   = wave_items ? block / 64 : 1;
    const long long need  = (nitems + wpb - 1) / wpb;
    const int       cap   = c->tune.grid_cap > 0 ? std::min(c->tune.grid_cap, occ) : occ;
    long long       limit = (long long)c->ncu * cap;
    long long       g     = std::min(need, limit);
    if (g >= 8)
        g = g / 8 * 8; // lets the kernel map contiguous item ranges per XCD
    return (int)std::max<long long>(g, 1);
}

int do_sweep(bx_ctx *c, int dir, const double dx[3], const double *Uin,
                    double *Uout, unsigned long long *maxbits, hipStream_t s)
{
    SweepLaunch L;
    L.Uin   = Uin;
    L.Uout  = Uout;
    L.boxes = c->d_boxes;
    L.nbox  = (int)c->hboxes.size();
    L.ng    = c->geom.ng;
    L.K     = c->tune.seg_k;
    L.d_dt  = c->d_dt;
    L.dx    = dx[dir];
    L.pm    = c->pm;
    L.lt    = c->lt;
    L.d_maxbits = maxbits;
    L.stream    = s;
    if (dir == 0)
    {
        L.item_begin = c->d_ib_x;
        L.xbox       = c->d_xbox;
        L.nitems     = c->ib_x.back();
        L.block      = c->tune.block_x;
        L.grid       = persistent_grid(c, L.nitems, L.block, c->occ_x, false);
    }
    else
    {
        L.item_begin = c->d_ib_t[dir];
        L.nitems     = c->ib_t[dir].back();
        L.block      = c->tune.block_t;
        L.grid       = persistent_grid(c, L.nitems, L.block, c->occ_t, true);
    }
    if (L.nitems == 0)
        return BX_OK;
    BX_HIP_CHECK(launch_sweep(c->D, dir, c->p.num_method,
                              c->p.limiter, c->lv_mode, L));
    return BX_OK;
}

int do_fill(bx_ctx *c, double *U, int list, hipStream_t s, const StepTail &tail)
{
    const int nseg = (int)c->segs[list].size();
    if (nseg == 0 && !tail.d_time)
        return BX_OK;
    const int grid = std::max(1, std::min(nseg, c->ncu * 16));
    BX_HIP_CHECK(launch_fill_ghosts(U, c->d_boxes, c->d_segs[list], nseg,
                                    c->geom.ng, c->D, c->NC, grid, tail, s));
    return BX_OK;
}

int do_fill_all(bx_ctx *c, double *U, hipStream_t s)
{
    return do_fill(c, U, 4, s);
}

int do_maxspeed(bx_ctx *c, const double *U, hipStream_t s)
{
    BX_HIP_CHECK(launch_set_u64(c->d_maxbits, 0ULL, s));
    // chunks of 4096 cells per block (k_max_wave_speed)
    const int grid = (int)std::min<long long>((c->ncells + 4095) / 4096,
                                              (long long)c->ncu * 8);
    BX_HIP_CHECK(launch_max_wave_speed(c->D, U, c->d_boxes, c->d_cbeg,
                                       (int)c->hboxes.size(), c->ncells,
                                       c->geom.ng, c->pm, c->d_maxbits,
                                       std::max(grid, 1), s));
    return BX_OK;
}

#define BX_REQUIRE_LEVEL(c)                                                     \
    do                                                                          \
    {                                                                           \
        if (!(c))                                                               \
            return set_error(BX_ERR_INVALID, "null context");                   \
        if (!(c)->level_set)                                                    \
            return set_error(BX_ERR_STATE, "bx_set_level has not been called"); \
    } while (0)

bool dx_ok(const bx_ctx *c, const double dx[3])
{
    if (!dx)
        return false;
    for (int d = 0; d < c->D; ++d)
        if (!(dx[d] > 0))
            return false;
    return true;
}

int set_scalar(double *p, double v, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_f64, dim3(1), dim3(64), 0, s, p, v);
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_sweep(bx_ctx *c, int dir, double dt, const double dx[3],
                        const double *U_in, double *U_out, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (dir < 0 || dir >= c->D || !dx_ok(c, dx) || !U_in || !U_out
        || U_in == U_out)
        return set_error(BX_ERR_INVALID, "bx_sweep: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = set_scalar(c->d_dt, dt, s)) != BX_OK)
        return rc;
    return do_sweep(c, dir, dx, U_in, U_out, nullptr, s);
}

template <int D, int DIR, int LIM, int LV, int SCH>
static hipError_t launch_flux_box(const double *U, const DevBox &bx, int ng,
                                  double *flux, double hdtdx, double fc, double rc,
                                  EulerParm pm, LimTable lt, hipStream_t s)
{
    hipLaunchKernelGGL((k_flux_box<D, DIR, LIM, LV, SCH>), dim3(256), dim3(256),
                       0, s, U, bx, ng, flux, hdtdx, fc, rc, pm, lt);
    return hipGetLastError();
}

template <int D, int DIR, int SCH>
static hipError_t flux_box_lim(int lim, int lv, const double *U, const DevBox &bx,
                               int ng, double *flux, double hdtdx, double fc,
                               double rc, EulerParm pm, LimTable lt, hipStream_t s)
{
#define BX_FB(LIMV)                                                            \
    do                                                                         \
    {                                                                          \
        if (lv == LV_SHARED)                                                   \
            return launch_flux_box<D, DIR, LIMV, LV_SHARED, SCH>(              \
                U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);                    \
        if (lv == LV_OWN)                                                      \
            return launch_flux_box<D, DIR, LIMV, LV_OWN, SCH>(                 \
                U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);                    \
        return launch_flux_box<D, DIR, LIMV, LV_TABLE, SCH>(                   \
            U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);                        \
    } while (0)
    if constexpr (!scheme_reconstructs(SCH))
        return launch_flux_box<D, DIR, LIM_MINBEE, LV_SHARED, SCH>(
            U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    else
        switch (lim)
        {
        case LIM_MINBEE: BX_FB(LIM_MINBEE);
        case LIM_VAN_LEER: BX_FB(LIM_VAN_LEER);
        case LIM_VAN_ALBADA: BX_FB(LIM_VAN_ALBADA);
        default: BX_FB(LIM_SUPERBEE);
        }
#undef BX_FB
}

template <int D, int DIR>
static hipError_t flux_box_scheme(int sch, int lim, int lv, const double *U,
                                  const DevBox &bx, int ng, double *flux, double hdtdx,
                                  double fc, double rc, EulerParm pm, LimTable lt,
                                  hipStream_t s)
{
    switch (sch)
    {
    case SCH_MHM:
        return flux_box_lim<D, DIR, SCH_MHM>(lim, lv, U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    case SCH_HLLC:
        return flux_box_lim<D, DIR, SCH_HLLC>(lim, lv, U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    case SCH_SLIC:
        return flux_box_lim<D, DIR, SCH_SLIC>(lim, lv, U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    case SCH_FORCE:
        return flux_box_lim<D, DIR, SCH_FORCE>(lim, lv, U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    default:
        return flux_box_lim<D, DIR, SCH_LF>(lim, lv, U, bx, ng, flux, hdtdx, fc, rc, pm, lt, s);
    }
}

extern "C" int bx_flux_box(bx_ctx *c, int ibox, int dir, double dt,
                           const double dx[3], const double *U,
                           double *flux_base, int64_t flux_offset, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (ibox < 0 || ibox >= (int)c->hboxes.size() || dir < 0 || dir >= c->D
        || !dx_ok(c, dx) || !U || !flux_base || flux_offset < 0)
        return set_error(BX_ERR_INVALID, "bx_flux_box: bad argument");
    const DevBox &bx    = c->hdev[ibox];
    const double  hdtdx = 0.5 * dt / dx[dir];
    const double  fc    = 0.5 * dx[dir] / dt; // FORCE / LF
    const double  rc    = dt / (2 * dx[dir]); // FORCE
    double       *flux  = flux_base + flux_offset;
    const int     sch   = c->p.num_method;
    hipStream_t   s     = (hipStream_t)stream;
    hipError_t    e     = hipErrorInvalidValue;
    const int     lim = c->p.limiter, lv = c->lv_mode, ng = c->geom.ng;
#define BX_FBS(DD, DIR)                                                         \
    flux_box_scheme<DD, DIR>(sch, lim, lv, U, bx, ng, flux, hdtdx, fc, rc, c->pm, \
                             c->lt, s)
    switch (c->D * 4 + dir)
    {
    case 4: e = BX_FBS(1, 0); break;
    case 8: e = BX_FBS(2, 0); break;
    case 9: e = BX_FBS(2, 1); break;
    case 12: e = BX_FBS(3, 0); break;
    case 13: e = BX_FBS(3, 1); break;
    case 14: e = BX_FBS(3, 2); break;
    default: break;
    }
#undef BX_FBS
    BX_HIP_CHECK(e);
    return BX_OK;
}

extern "C" int bx_update_box(bx_ctx *c, int ibox, int dir, double dt,
                             const double dx[3], const double *U_in,
                             const double *flux_base, int64_t flux_offset,
                             double *U_out, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (ibox < 0 || ibox >= (int)c->hboxes.size() || dir < 0 || dir >= c->D
        || !dx_ok(c, dx) || !U_in || !U_out || !flux_base || flux_offset < 0)
        return set_error(BX_ERR_INVALID, "bx_update_box: bad argument");
    const DevBox &bx   = c->hdev[ibox];
    const double *flux = flux_base + flux_offset;
    hipStream_t   s    = (hipStream_t)stream;
    const int     ng   = c->geom.ng;
#define BX_UB(DD, DIR)                                                           \
    hipLaunchKernelGGL((k_update_box<DD, DIR>), dim3(256), dim3(256), 0, s, U_in, \
                       U_out, bx, ng, flux, dt, dx[dir])
    switch (c->D * 4 + dir)
    {
    case 4: BX_UB(1, 0); break;
    case 8: BX_UB(2, 0); break;
    case 9: BX_UB(2, 1); break;
    case 12: BX_UB(3, 0); break;
    case 13: BX_UB(3, 1); break;
    case 14: BX_UB(3, 2); break;
    default: return set_error(BX_ERR_INVALID, "bx_update_box: bad dir");
    }
#undef BX_UB
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_fill_boundary(bx_ctx *c, double *U, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (!U)
        return set_error(BX_ERR_INVALID, "bx_fill_boundary: null U");
    return do_fill_all(c, U, (hipStream_t)stream);
}

extern "C" int bx_max_wave_speed(bx_ctx *c, const double *U, double *out,
                                 void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (!U || !out)
        return set_error(BX_ERR_INVALID, "bx_max_wave_speed: null argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = do_maxspeed(c, U, s)) != BX_OK)
        return rc;
    unsigned long long bits = 0;
    BX_HIP_CHECK(hipMemcpyAsync(&bits, c->d_maxbits, 8, hipMemcpyDeviceToHost, s));
    BX_HIP_CHECK(hipStreamSynchronize(s));
    std::memcpy(out, &bits, 8);
    return BX_OK;
}

extern "C" int bx_est_dt(bx_ctx *c, const double *U, const double dx[3],
                         double cfl, double time, double stop_time,
                         double *dt_out, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (!U || !dt_out || !dx_ok(c, dx) || !(cfl > 0))
        return set_error(BX_ERR_INVALID, "bx_est_dt: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = do_maxspeed(c, U, s)) != BX_OK)
        return rc;
    if ((rc = set_scalar(c->d_time, time, s)) != BX_OK)
        return rc;
    BX_HIP_CHECK(launch_compute_dt(c->d_maxbits, c->d_dt, c->d_time, c->D, dx,
                                   cfl, stop_time, -1.0, s));
    BX_HIP_CHECK(hipMemcpyAsync(dt_out, c->d_dt, 8, hipMemcpyDeviceToHost, s));
    BX_HIP_CHECK(hipStreamSynchronize(s));
    return BX_OK;
}

// One split step on device buffers.  dt is read from c->d_dt.  When
// fuse_max is set the last sweep accumulates the CFL maximum of the new
// state into c->d_maxbits.  A non-empty tail runs with the last ghost fill.
static int step_body(bx_ctx *c, const double dx[3], double *A, double *B,
                     bool fuse_max, hipStream_t s,
                     const StepTail &tail = StepTail())
{
    const int D = c->D;
    int       rc;
    for (int d = 0; d < D; ++d)
    {
        const double *in  = (d % 2 == 0) ? A : B;
        double       *out = (d % 2 == 0) ? B : A;
        if ((rc = do_sweep(c, d, dx, in, out,
                           (fuse_max && d == D - 1) ? c->d_maxbits : nullptr, s))
            != BX_OK)
            return rc;
        // ghost cells the next sweep reads: direction d+1, or x for the
        // next step after the last sweep
        if ((rc = do_fill(c, out, (d + 1) % D, s,
                          d == D - 1 ? tail : StepTail()))
            != BX_OK)
            return rc;
    }
    return BX_OK;
}

extern "C" int bx_advance(bx_ctx *c, double dt, const double dx[3], double *U,
                          double *U_tmp, void *stream)
{
    BX_REQUIRE_LEVEL(c);
    if (!U || !U_tmp || U == U_tmp || !dx_ok(c, dx))
        return set_error(BX_ERR_INVALID, "bx_advance: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = set_scalar(c->d_dt, dt, s)) != BX_OK)
        return rc;
    if ((rc = step_body(c, dx, U, U_tmp, false, s)) != BX_OK)
        return rc;
    if (c->D % 2 == 1)
        BX_HIP_CHECK(hipMemcpyAsync(U, U_tmp, sizeof(double) * c->level_extent,
                                    hipMemcpyDeviceToDevice, s));
    // the reference leaves every ghost cell filled (fill_consv_boundary)
    return do_fill_all(c, U, s);
}

// One step of the device-controlled time loop.  The step starts with the
// dt already in c->d_dt; the last ghost fill advances the time by it and
// computes the next step's dt from the maximum the last sweep accumulated
// (the loop's first dt comes from launch_compute_dt before the first step).
static int step_device(bx_ctx *c, const double dx[3], double cfl,
                       double stop_time, double *A, double *B, hipStream_t s)
{
    StepTail t;
    t.d_maxbits = c->d_maxbits;
    t.d_dt      = c->d_dt;
    t.d_time    = c->d_time;
    t.d_step    = c->d_step;
    t.D         = c->D;
    for (int d = 0; d < 3; ++d)
        t.dx[d] = dx[d];
    t.cfl       = cfl;
    t.stop_time = stop_time;
    return step_body(c, dx, A, B, true, s, t);
}

static int get_graph(bx_ctx *c, int which, const double dx[3], double cfl,
                     double stop_time, double *A, double *B, hipStream_t s,
                     hipGraphExec_t *exec)
{
    StepGraph &g = c->graphs[which];
    if (g.exec && g.a == A && g.b == B && g.dx[0] == dx[0] && g.dx[1] == dx[1]
        && g.dx[2] == dx[2] && g.cfl == cfl && g.stop == stop_time
        && g.stream == (void *)s)
    {
        *exec = g.exec;
        return BX_OK;
    }
    if (g.exec)
        (void)hipGraphExecDestroy(g.exec);
    if (g.graph)
        (void)hipGraphDestroy(g.graph);
    g = StepGraph();
    BX_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = step_device(c, dx, cfl, stop_time, A, B, s);
    hipGraph_t graph = nullptr;
    hipError_t e     = hipStreamEndCapture(s, &graph);
    if (rc != BX_OK)
    {
        if (graph)
            (void)hipGraphDestroy(graph);
        return rc;
    }
    BX_HIP_CHECK(e);
    BX_HIP_CHECK(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0));
    g.graph = graph;
    g.a     = A;
    g.b     = B;
    for (int d = 0; d < 3; ++d)
        g.dx[d] = dx[d];
    g.cfl    = cfl;
    g.stop   = stop_time;
    g.stream = (void *)s;
    *exec    = g.exec;
    return BX_OK;
}

extern "C" int bx_run(bx_ctx *c, const double dx[3], double cfl,
                      double start_time, double stop_time, int max_step,
                      double *U, double *U_tmp, int use_graph, void *stream,
                      double *time_out, int *steps_out, int *result_in_tmp)
{
    BX_REQUIRE_LEVEL(c);
    if (!U || !U_tmp || U == U_tmp || !dx_ok(c, dx) || !(cfl > 0))
        return set_error(BX_ERR_INVALID, "bx_run: bad argument");
    if (max_step < 0 && stop_time < 0)
        return set_error(BX_ERR_INVALID,
                         "Exiting because neither max_step nor stop_time is "
                         "non-negative.");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = do_fill_all(c, U, s)) != BX_OK)
        return rc;
    if ((rc = do_maxspeed(c, U, s)) != BX_OK)
        return rc;
    if ((rc = set_scalar(c->d_time, start_time, s)) != BX_OK)
        return rc;
    BX_HIP_CHECK(launch_set_u64((unsigned long long *)c->d_step, 0ULL, s));
    BX_HIP_CHECK(launch_compute_dt(c->d_maxbits, c->d_dt, c->d_time, c->D, dx,
                                   cfl, stop_time, -1.0, s));

    double *A = U, *B = U_tmp;
    const bool odd = (c->D % 2) == 1;
    int        steps = 0;
    double     time  = start_time;
    // stream capture needs a created stream; the null stream runs the host loop
    if (stop_time < 0 && use_graph && s != nullptr)
    {
        hipGraphExec_t g0 = nullptr, g1 = nullptr;
        if ((rc = get_graph(c, 0, dx, cfl, stop_time, U, U_tmp, s, &g0)) != BX_OK)
            return rc;
        if (odd && (rc = get_graph(c, 1, dx, cfl, stop_time, U_tmp, U, s, &g1))
                       != BX_OK)
            return rc;
        for (int n = 0; n < max_step; ++n)
        {
            BX_HIP_CHECK(hipGraphLaunch((odd && (n & 1)) ? g1 : g0, s));
            if (odd)
                std::swap(A, B);
        }
        steps = max_step;
        BX_HIP_CHECK(hipMemcpyAsync(&time, c->d_time, 8, hipMemcpyDeviceToHost, s));
        BX_HIP_CHECK(hipStreamSynchronize(s));
    }
    else
    {
        while ((steps < max_step || max_step < 0)
               && (time < stop_time || stop_time < 0.0))
        {
            if ((rc = step_device(c, dx, cfl, stop_time, A, B, s)) != BX_OK)
                return rc;
            if (odd)
                std::swap(A, B);
            ++steps;
            BX_HIP_CHECK(hipMemcpyAsync(&time, c->d_time, 8,
                                        hipMemcpyDeviceToHost, s));
            BX_HIP_CHECK(hipStreamSynchronize(s));
        }
    }
    int in_tmp = (A != U);
    if (in_tmp)
    {
        BX_HIP_CHECK(hipMemcpyAsync(U, U_tmp, sizeof(double) * c->level_extent,
                                    hipMemcpyDeviceToDevice, s));
        in_tmp = 0;
    }
    if ((rc = do_fill_all(c, U, s)) != BX_OK)
        return rc;
    BX_HIP_CHECK(hipStreamSynchronize(s));
    if (time_out)
        *time_out = time;
    if (steps_out)
        *steps_out = steps;
    if (result_in_tmp)
        *result_in_tmp = in_tmp;
    return BX_OK;
}

// --------------------------------------------------------------- tuning
extern "C" int bx_get_tuning(const bx_ctx *c, bx_tuning *out)
{
    if (!c || !out)
        return set_error(BX_ERR_INVALID, "bx_get_tuning: null argument");
    *out = c->tune;
    return BX_OK;
}

extern "C" int bx_set_tuning(bx_ctx *c, const bx_tuning *t)
{
    if (!c || !t)
        return set_error(BX_ERR_INVALID, "bx_set_tuning: null argument");
    auto blk_ok = [](int b) { return b >= 64 && b <= 256 && b % 64 == 0; };
    if (!blk_ok(t->block_x) || !blk_ok(t->block_t) || t->seg_k < 1
        || t->seg_k > BX_MAX_SEG_K || t->grid_cap < 0)
        return set_error(BX_ERR_INVALID, "bx_set_tuning: out of range");
    destroy_graphs(c);
    c->tune = *t;
    if (c->level_set)
    {
        build_sweep_tables(c);
        int rc;
        if ((rc = upload_sweep_tables(c)) != BX_OK)
            return rc;
        if ((rc = occupancy(c, 0, c->tune.block_x, &c->occ_x)) != BX_OK)
            return rc;
        if (c->D >= 2
            && (rc = occupancy(c, 1, c->tune.block_t, &c->occ_t)) != BX_OK)
            return rc;
    }
    return BX_OK;
}

static int time_sweep(bx_ctx *c, int dir, const double dx[3], double *U,
                      double *U_tmp, hipStream_t s, int reps, double *ms)
{
    hipEvent_t e0, e1;
    BX_HIP_CHECK(hipEventCreate(&e0));
    BX_HIP_CHECK(hipEventCreate(&e1));
    // untimed launches first: a new shape follows a host synchronisation
    // (table upload), and the GPU clock dips at the start of a busy period
    int rc = BX_OK;
    for (int r = 0; r < 8 && rc == BX_OK; ++r)
        rc = do_sweep(c, dir, dx, U, U_tmp, nullptr, s);
    if (rc == BX_OK)
    {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && rc == BX_OK; ++r)
            rc = do_sweep(c, dir, dx, U, U_tmp, nullptr, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float f = 0;
        (void)hipEventElapsedTime(&f, e0, e1);
        *ms = f / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// The search itself; on an error the caller restores the entry shape.
static int autotune_search(bx_ctx *c, const double dx[3], double *U, double *U_tmp,
                           hipStream_t s, double *best_ms)
{
    int       rc;
    bx_tuning best = c->tune;
    // untimed sweeps at the entry shape first: after an idle period the GPU
    // clock dips over the first ~10 ms of back-to-back work, which would
    // bias the first candidates
    for (int r = 0; r < 16; ++r)
        for (int d = 0; d < c->D; ++d)
            if ((rc = do_sweep(c, d, dx, U, U_tmp, nullptr, s)) != BX_OK)
                return rc;
    const int blocks[3] = { 64, 128, 256 };
    const int caps[4]   = { 0, 1, 2, 4 };
    // segment lengths: powers of two, and the median box length along the
    // transverse sweeps split into m equal parts (m = 1..8), which lets the
    // item count divide evenly over 2, 3, ... resident waves per SIMD;
    // lengths above what bx_set_tuning accepts are skipped
    std::vector<int> ks = { 8, 16, 32, 64, 128 };
    if (c->D >= 2)
    {
        std::vector<int> len;
        for (const bx_box &bx : c->hboxes)
            len.push_back(bx.hi[1] - bx.lo[1] + 1);
        std::nth_element(len.begin(), len.begin() + len.size() / 2, len.end());
        const int L = len[len.size() / 2];
        for (int m = 1; m <= 8; ++m)
        {
            const int k = (L + m - 1) / m;
            if (k >= 4 && k <= BX_MAX_SEG_K)
                ks.push_back(k);
        }
        std::sort(ks.begin(), ks.end());
        ks.erase(std::unique(ks.begin(), ks.end()), ks.end());
    }
    // x sweep: block size x grid cap
    double bx_ms = 1e30;
    for (int b : blocks)
        for (int cap : caps)
        {
            bx_tuning t = best;
            t.block_x   = b;
            t.grid_cap  = cap;
            if ((rc = bx_set_tuning(c, &t)) != BX_OK)
                return rc;
            double ms = 0;
            if ((rc = time_sweep(c, 0, dx, U, U_tmp, s, 5, &ms)) != BX_OK)
                return rc;
            if (ms < bx_ms)
            {
                bx_ms         = ms;
                best.block_x  = b;
                best.grid_cap = cap;
            }
        }
    if (c->D >= 2)
    {
        double t_ms = 1e30;
        for (int b : blocks)
            for (int k : ks)
            {
                bx_tuning t = best;
                t.block_t   = b;
                t.seg_k     = k;
                if ((rc = bx_set_tuning(c, &t)) != BX_OK)
                    return rc;
                double ms = 0;
                if ((rc = time_sweep(c, 1, dx, U, U_tmp, s, 5, &ms)) != BX_OK)
                    return rc;
                if (ms < t_ms)
                {
                    t_ms         = ms;
                    best.block_t = b;
                    best.seg_k   = k;
                }
            }
    }
    if ((rc = bx_set_tuning(c, &best)) != BX_OK)
        return rc;
    if (best_ms)
    {
        best_ms[0] = bx_ms;
        for (int d = 1; d < c->D; ++d)
        {
            double ms = 0;
            if ((rc = time_sweep(c, d, dx, U, U_tmp, s, 5, &ms)) != BX_OK)
                return rc;
            best_ms[d] = ms;
        }
    }
    return BX_OK;
}

extern "C" int bx_autotune(bx_ctx *c, const double dx[3], double dt, double *U,
                           double *U_tmp, void *stream, double *best_ms)
{
    BX_REQUIRE_LEVEL(c);
    if (!U || !U_tmp || U == U_tmp || !dx_ok(c, dx))
        return set_error(BX_ERR_INVALID, "bx_autotune: bad argument");
    hipStream_t s = (hipStream_t)stream;
    int         rc;
    if ((rc = do_fill_all(c, U, s)) != BX_OK)
        return rc;
    if ((rc = set_scalar(c->d_dt, dt, s)) != BX_OK)
        return rc;
    const bx_tuning entry = c->tune;
    if ((rc = autotune_search(c, dx, U, U_tmp, s, best_ms)) != BX_OK)
    {
        // keep the error of the failed step, not that of the restore
        const std::string msg = bx_last_error();
        (void)bx_set_tuning(c, &entry);
        return set_error(rc, msg);
    }
    return BX_OK;
}

// --------------------------------------------------------- HIP helpers
extern "C" int bx_device_count(int *n)
{
    if (!n)
        return set_error(BX_ERR_INVALID, "null");
    BX_HIP_CHECK(hipGetDeviceCount(n));
    return BX_OK;
}
extern "C" int bx_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return set_error(BX_ERR_INVALID, "null");
    BX_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
    return BX_OK;
}
extern "C" int bx_free(void *ptr)
{
    BX_HIP_CHECK(hipFree(ptr));
    return BX_OK;
}
extern "C" int bx_memcpy_htod(void *dst, const void *src, size_t bytes)
{
    BX_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return BX_OK;
}
extern "C" int bx_memcpy_dtoh(void *dst, const void *src, size_t bytes)
{
    BX_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return BX_OK;
}
extern "C" int bx_memcpy_dtod(void *dst, const void *src, size_t bytes)
{
    BX_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return BX_OK;
}
extern "C" int bx_memset(void *dst, int value, size_t bytes)
{
    BX_HIP_CHECK(hipMemset(dst, value, bytes));
    BX_HIP_CHECK(hipDeviceSynchronize());
    return BX_OK;
}
extern "C" int bx_stream_create(void **stream)
{
    if (!stream)
        return set_error(BX_ERR_INVALID, "null");
    hipStream_t s;
    BX_HIP_CHECK(hipStreamCreate(&s));
    *stream = (void *)s;
    return BX_OK;
}
extern "C" int bx_stream_destroy(void *stream)
{
    BX_HIP_CHECK(hipStreamDestroy((hipStream_t)stream));
    return BX_OK;
}
extern "C" int bx_stream_sync(void *stream)
{
    BX_HIP_CHECK(hipStreamSynchronize((hipStream_t)stream));
    return BX_OK;
}
extern "C" int bx_event_create(void **ev)
{
    if (!ev)
        return set_error(BX_ERR_INVALID, "null");
    hipEvent_t e;
    BX_HIP_CHECK(hipEventCreate(&e));
    *ev = (void *)e;
    return BX_OK;
}
extern "C" int bx_event_destroy(void *ev)
{
    BX_HIP_CHECK(hipEventDestroy((hipEvent_t)ev));
    return BX_OK;
}
extern "C" int bx_event_record(void *ev, void *stream)
{
    BX_HIP_CHECK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return BX_OK;
}
extern "C" int bx_event_elapsed_ms(void *start, void *stop, float *ms)
{
    if (!ms)
        return set_error(BX_ERR_INVALID, "null");
    BX_HIP_CHECK(hipEventSynchronize((hipEvent_t)stop));
    BX_HIP_CHECK(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return BX_OK;
}
extern "C" int bx_copy_kernel(void *dst, const void *src, size_t bytes,
                              void *stream)
{
    if (!dst || !src || bytes % 8)
        return set_error(BX_ERR_INVALID, "bx_copy_kernel: bad argument");
    const long long n = (long long)(bytes / 8);
    hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, (hipStream_t)stream,
                       (const double *)src, (double *)dst, n);
    BX_HIP_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" int bx_device_name(char *buf, int len)
{
    if (!buf || len <= 0)
        return set_error(BX_ERR_INVALID, "null");
    hipDeviceProp_t prop;
    int             dev = 0;
    BX_HIP_CHECK(hipGetDevice(&dev));
    BX_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name,
                  prop.gcnArchName, prop.multiProcessorCount);
    return BX_OK;
}
