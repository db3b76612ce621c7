/*
 * CPU ORACLE -- TEST INFRASTRUCTURE ONLY.  Used by tests/ as the checker of the
 * pressure Helmholtz solver; never linked into or called by the product.
 *
 * Plain-C restatement of the variable-coefficient Helmholtz operator of the
 * IMEX pressure equation and of the multigrid pieces built on it, on one 2D
 * periodic level, in the operation order of the reference kernels:
 *   or_helm_apply    mlalapgrad_adotx   Source/LinearSolver/MLALapGrad_2D_K.H:9-33
 *   or_helm_gsrb     alapgrad_gsrb      Source/LinearSolver/MLALapGrad_2D_K.H:179-229
 *                    (periodic level: every boundary mask is "covered", so the
 *                    delta terms vanish), with dhx = beta/(h*h), cdhx =
 *                    gamma/(2h) as set in Fsmooth (MLALaplacianGrad.H:1190-1260)
 *   or_helm_restrict / or_helm_avg_faces   averageDownCoeffsSameAmrLevel
 *                    (MLALaplacianGrad.H:950-1027: amrex::average_down of cell
 *                    data, average_down_faces of face data)
 *   or_helm_vcycle / or_helm_mg_solve       AMReX MLMG V-cycle, restated from
 *                    the AMReX documentation (AMReX is not part of the
 *                    reference tree): nu1 pre- and nu2 post- red/black sweeps,
 *                    residual restriction by cell averaging, piecewise-constant
 *                    prolongation, coarsening while both halved extents are
 *                    even and >= 8 (an odd periodic extent would give the two
 *                    cells across the seam one colour), bottom = nu_bottom red/black sweep
 *                    pairs (16 by default).
 *                    Parity of these MLMG choices with AMReX itself is
 *                    unpinned (no AMReX build here); the pieces they use are
 *                    pinned by the tests (closed forms, symmetry, Fourier
 *                    eigenvalues).
 * Layout: cell arrays ny x nx (x fastest); bX is ny x (nx+1) (face i of row j at
 * j*(nx+1)+i), bY is (ny+1) x nx.  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* y = A x, or y = rhs - A x when rhs != NULL.  cX == NULL drops the c term. */
void or_helm_apply(int nx, int ny, const double *x, double *y, const double *a,
                   const double *bX, const double *bY, const double *cX,
                   const double *cY, double alpha, double beta, double gamma,
                   double dxinv, double dyinv, const double *rhs)
{
    const double dhx = beta * dxinv * dxinv;
    const double dhy = beta * dyinv * dyinv;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            const long   c   = (long)j * nx + i;
            const double xc  = x[c];
            const double xl  = x[(long)j * nx + wrap(i - 1, nx)];
            const double xr  = x[(long)j * nx + wrap(i + 1, nx)];
            const double xd  = x[(long)wrap(j - 1, ny) * nx + i];
            const double xu  = x[(long)wrap(j + 1, ny) * nx + i];
            const double bxl = bX[(long)j * (nx + 1) + i];
            const double bxr = bX[(long)j * (nx + 1) + i + 1];
            const double byl = bY[(long)j * nx + i];
            const double byr = bY[(long)(j + 1) * nx + i];
            double v = alpha * a[c] * xc - dhx * (bxr * (xr - xc) - bxl * (xc - xl))
                       - dhy * (byr * (xu - xc) - byl * (xc - xd));
            if (cX)
            {
                const double chdx = gamma * cX[c] * 0.5 * dxinv;
                const double chdy = gamma * cY[c] * 0.5 * dyinv;
                v = v + chdx * (xr - xl) + chdy * (xu - xd);
            }
            y[c] = rhs ? rhs[c] - v : v;
        }
}

/* One colour of red/black Gauss-Seidel: cells with (i + j + redblack) even. */
void or_helm_gsrb(int nx, int ny, double *phi, const double *rhs, const double *a,
                  const double *bX, const double *bY, const double *cX,
                  const double *cY, double alpha, double beta, double gamma,
                  double hx, double hy, int redblack)
{
    const double dhx = beta / (hx * hx), dhy = beta / (hy * hy);
    const double cdhx = gamma / (2 * hx), cdhy = gamma / (2 * hy);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
        {
            if ((i + j + redblack) % 2 != 0)
                continue;
            const long   c   = (long)j * nx + i;
            const double bxl = bX[(long)j * (nx + 1) + i];
            const double bxr = bX[(long)j * (nx + 1) + i + 1];
            const double byl = bY[(long)j * nx + i];
            const double byr = bY[(long)(j + 1) * nx + i];
            const double pl  = phi[(long)j * nx + wrap(i - 1, nx)];
            const double pr  = phi[(long)j * nx + wrap(i + 1, nx)];
            const double pd  = phi[(long)wrap(j - 1, ny) * nx + i];
            const double pu  = phi[(long)wrap(j + 1, ny) * nx + i];
            const double gdiag = alpha * a[c] + dhx * (bxl + bxr) + dhy * (byl + byr);
            double rho = dhx * (bxl * pl + bxr * pr) + dhy * (byl * pd + byr * pu);
            if (cX)
                rho = rho - cdhx * cX[c] * (pr - pl) - cdhy * cY[c] * (pu - pd);
            const double delta = 0.0;
            phi[c] = (rhs[c] + rho - phi[c] * delta) / (gdiag - delta);
        }
}

void or_helm_restrict(int ncx, int ncy, const double *f, double *c)
{
    const int nfx = 2 * ncx;
    for (int J = 0; J < ncy; ++J)
        for (int I = 0; I < ncx; ++I)
        {
            const long f0 = (long)(2 * J) * nfx + 2 * I;
            c[(long)J * ncx + I] = 0.25 * (f[f0] + f[f0 + 1] + f[f0 + nfx] + f[f0 + nfx + 1]);
        }
}

void or_helm_avg_faces(int ncx, int ncy, const double *fX, const double *fY,
                       double *cXo, double *cYo)
{
    const int nfx = 2 * ncx;
    for (int J = 0; J < ncy; ++J)
        for (int I = 0; I <= ncx; ++I)
            cXo[(long)J * (ncx + 1) + I]
                = 0.5 * (fX[(long)(2 * J) * (nfx + 1) + 2 * I]
                         + fX[(long)(2 * J + 1) * (nfx + 1) + 2 * I]);
    for (int J = 0; J <= ncy; ++J)
        for (int I = 0; I < ncx; ++I)
            cYo[(long)J * ncx + I] = 0.5 * (fY[(long)(2 * J) * nfx + 2 * I]
                                            + fY[(long)(2 * J) * nfx + 2 * I + 1]);
}

typedef struct
{
    int     nx, ny;
    double  hx, hy;
    double *a, *bX, *bY, *x, *rhs, *res;
} Lev;

typedef struct
{
    int    nlev;
    Lev    L[24];
    double alpha, beta;
    int    nu1, nu2, nu_bottom;
} Hier;

static void hier_build(Hier *H, int nx, int ny, double dx, double dy,
                       const double *a, const double *bX, const double *bY)
{
    H->nlev = 0;
    for (;;)
    {
        Lev *L = &H->L[H->nlev++];
        L->nx  = nx;
        L->ny  = ny;
        L->hx  = dx;
        L->hy  = dy;
        const long n = (long)nx * ny;
        L->a   = (double *)calloc(n, 8);
        L->bX  = (double *)calloc((long)(nx + 1) * ny, 8);
        L->bY  = (double *)calloc((long)nx * (ny + 1), 8);
        L->x   = (double *)calloc(n, 8);
        L->rhs = (double *)calloc(n, 8);
        L->res = (double *)calloc(n, 8);
        if (nx % 4 || ny % 4 || nx / 2 < 8 || ny / 2 < 8 || H->nlev == 24)
            break;
        nx /= 2;
        ny /= 2;
        dx *= 2;
        dy *= 2;
    }
    Lev *L0 = &H->L[0];
    memcpy(L0->a, a, sizeof(double) * L0->nx * L0->ny);
    memcpy(L0->bX, bX, sizeof(double) * (L0->nx + 1) * L0->ny);
    memcpy(L0->bY, bY, sizeof(double) * L0->nx * (L0->ny + 1));
    for (int l = 1; l < H->nlev; ++l)
    {
        Lev *F = &H->L[l - 1], *C = &H->L[l];
        or_helm_restrict(C->nx, C->ny, F->a, C->a);
        or_helm_avg_faces(C->nx, C->ny, F->bX, F->bY, C->bX, C->bY);
    }
}

static void hier_free(Hier *H)
{
    for (int l = 0; l < H->nlev; ++l)
    {
        Lev *L = &H->L[l];
        free(L->a); free(L->bX); free(L->bY); free(L->x); free(L->rhs); free(L->res);
    }
}

static void smooth2(Hier *H, Lev *L)
{
    or_helm_gsrb(L->nx, L->ny, L->x, L->rhs, L->a, L->bX, L->bY, NULL, NULL, H->alpha,
                 H->beta, 0.0, L->hx, L->hy, 0);
    or_helm_gsrb(L->nx, L->ny, L->x, L->rhs, L->a, L->bX, L->bY, NULL, NULL, H->alpha,
                 H->beta, 0.0, L->hx, L->hy, 1);
}

static void vcycle(Hier *H, int l)
{
    Lev *L = &H->L[l];
    if (l + 1 == H->nlev)
    {
        for (int it = 0; it < H->nu_bottom; ++it)
            smooth2(H, L);
        return;
    }
    for (int it = 0; it < H->nu1; ++it)
        smooth2(H, L);
    or_helm_apply(L->nx, L->ny, L->x, L->res, L->a, L->bX, L->bY, NULL, NULL, H->alpha,
                  H->beta, 0.0, 1.0 / L->hx, 1.0 / L->hy, L->rhs);
    Lev *C = &H->L[l + 1];
    or_helm_restrict(C->nx, C->ny, L->res, C->rhs);
    memset(C->x, 0, sizeof(double) * C->nx * C->ny);
    vcycle(H, l + 1);
    for (int j = 0; j < L->ny; ++j)
        for (int i = 0; i < L->nx; ++i)
            L->x[(long)j * L->nx + i] += C->x[(long)(j >> 1) * C->nx + (i >> 1)];
    for (int it = 0; it < H->nu2; ++it)
        smooth2(H, L);
}

/* e = V-cycle(r) from a zero guess (gamma = 0 operators). */
int or_helm_vcycle(int nx, int ny, double dx, double dy, double alpha, double beta,
                   const double *a, const double *bX, const double *bY, int nu1,
                   int nu2, int nu_bottom, const double *r, double *e)
{
    Hier H;
    H.alpha = alpha; H.beta = beta; H.nu1 = nu1; H.nu2 = nu2; H.nu_bottom = nu_bottom;
    hier_build(&H, nx, ny, dx, dy, a, bX, bY);
    memcpy(H.L[0].rhs, r, sizeof(double) * nx * ny);
    memset(H.L[0].x, 0, sizeof(double) * nx * ny);
    vcycle(&H, 0);
    memcpy(e, H.L[0].x, sizeof(double) * nx * ny);
    int nl = H.nlev;
    hier_free(&H);
    return nl;
}

static double norm2(const double *v, long n)
{
    double s = 0;
    for (long i = 0; i < n; ++i)
        s += v[i] * v[i];
    return sqrt(s);
}

/* Plain multigrid: x += V(b - A x) until ||b - A x||_2 <= rtol ||b||_2.
 * Returns the number of V-cycles (or -1 if maxiter is reached); the
 * history of relative residuals goes to hist (maxiter+1 entries) if given. */
int or_helm_mg_solve(int nx, int ny, double dx, double dy, double alpha, double beta,
                     const double *a, const double *bX, const double *bY, int nu1,
                     int nu2, int nu_bottom, const double *b, double *x, double rtol,
                     int maxiter, double *hist)
{
    const long n = (long)nx * ny;
    Hier       H;
    H.alpha = alpha; H.beta = beta; H.nu1 = nu1; H.nu2 = nu2; H.nu_bottom = nu_bottom;
    hier_build(&H, nx, ny, dx, dy, a, bX, bY);
    double *r = (double *)malloc(sizeof(double) * n);
    const double bn = norm2(b, n);
    int it = 0, ok = 0;
    for (;;)
    {
        or_helm_apply(nx, ny, x, r, a, bX, bY, NULL, NULL, alpha, beta, 0.0, 1.0 / dx,
                      1.0 / dy, b);
        const double rn = norm2(r, n);
        if (hist)
            hist[it] = rn / bn;
        if (rn <= rtol * bn)
        {
            ok = 1;
            break;
        }
        if (it >= maxiter)
            break;
        memcpy(H.L[0].rhs, r, sizeof(double) * n);
        memset(H.L[0].x, 0, sizeof(double) * n);
        vcycle(&H, 0);
        for (long i = 0; i < n; ++i)
            x[i] = x[i] + 1.0 * H.L[0].x[i];
        ++it;
    }
    free(r);
    hier_free(&H);
    return ok ? it : -1;
}
