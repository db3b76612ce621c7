// initdata of the reference problems the hot path is exercised on.  The
// problem is chosen from the init.* keys of the input file, as each Exec/
// directory of the reference links its own Prob.cpp:
//   init.density_TR ...   Exec/QuadrantProblem/Prob.cpp:20-97
//   init.lambda ...       Exec/IsentropicVortex/Prob.cpp:23-58, 67-129
//   init.density_L ...    Exec/RiemannProblem/Prob.cpp:33-121
#include <cmath>
#include <stdexcept>

#include "AmrLevelAdv.H"

namespace bamrex {

// consv_from_primv, Source/System/Euler/Euler.H:138-165
static void consv_from_primv(const double *q, double *u, double adiabatic,
                             double epsilon, int dim)
{
    u[0] = q[0];
    for (int d = 0; d < dim; ++d)
        u[1 + d] = q[0] * q[1 + d];
    const double pressure            = q[1 + dim];
    const double vol_internal_energy = pressure / (adiabatic - 1);
    double       u_sq                = q[1] * q[1];
    for (int d = 1; d < dim; ++d)
        u_sq = u_sq + q[1 + d] * q[1 + d];
    u[1 + dim] = vol_internal_energy + 0.5 * epsilon * q[0] * u_sq;
}

static double domain_wrap(double c, double lo, double hi)
{
    const double size = hi - lo;
    if (c < lo)
        return c + std::ceil((lo - c) / size) * size;
    if (c > hi)
        return c - std::ceil((c - hi) / size) * size;
    return c;
}

template <class F>
static void for_valid(const BoxLayout &lay, F f)
{
    for (int b = 0; b < (int)lay.boxes.size(); ++b)
    {
        const bx_box &bx = lay.boxes[b];
        for (int k = bx.lo[2]; k <= bx.hi[2]; ++k)
            for (int j = bx.lo[1]; j <= bx.hi[1]; ++j)
                for (int i = bx.lo[0]; i <= bx.hi[0]; ++i)
                    f(b, i, j, k);
    }
}

void initdata(std::vector<double> &U, const BoxLayout &lay, const Geometry &g,
              const bx_params &parm, const ParmParse &)
{
    ParmParse    pp("init");
    const int    dim = g.dim;
    const double a = parm.adiabatic, e = parm.epsilon;
    double       q[5], u[5];
    if (pp.contains("density_TR"))
    {
        const char *Q[4] = { "TR", "TL", "BL", "BR" };
        double      cs[4][5];
        for (int s = 0; s < 4; ++s)
        {
            const std::string t = Q[s];
            pp.get("density_" + t, q[0]);
            pp.get("velocity_x_" + t, q[1]);
            pp.get("velocity_y_" + t, q[2]);
            pp.get("pressure_" + t, q[dim + 1]);
            if (dim == 3)
                q[3] = 0;
            consv_from_primv(q, cs[s], a, e, dim);
        }
        for_valid(lay, [&](int b, int i, int j, int k) {
            const double x = g.prob_lo[0] + (i + 0.5) * g.dx[0];
            const double y = g.prob_lo[1] + (j + 0.5) * g.dx[1];
            int          s;
            if (x > 0.5 && y > 0.5)
                s = 0;
            else if (x < 0.5 && y > 0.5)
                s = 1;
            else if (x < 0.5 && y < 0.5)
                s = 2;
            else
                s = 3;
            for (int n = 0; n < dim + 2; ++n)
                const_cast<BoxLayout &>(lay).at(U, b, i, j, k, n) = cs[s][n];
        });
        return;
    }
    if (pp.contains("lambda"))
    {
        double lambda, rho_bg, u_bg, v_bg, w_bg, p_bg, x_c, y_c;
        pp.get("lambda", lambda);
        pp.get("density_bg", rho_bg);
        pp.get("u_bg", u_bg);
        pp.get("v_bg", v_bg);
        pp.get("w_bg", w_bg);
        pp.get("pressure_bg", p_bg);
        pp.get("x_c", x_c);
        pp.get("y_c", y_c);
        const double pi   = 3.141592653589793;
        const double dT_p = -(a - 1) * lambda * lambda / (8 * a * pi * pi);
        const double v_p  = lambda / (2 * pi);
        for_valid(lay, [&](int b, int i, int j, int k) {
            const double xw = domain_wrap(g.prob_lo[0] + (i + 0.5) * g.dx[0],
                                          g.prob_lo[0], g.prob_hi[0]);
            const double yw = domain_wrap(g.prob_lo[1] + (j + 0.5) * g.dx[1],
                                          g.prob_lo[1], g.prob_hi[1]);
            const double x = xw - x_c, y = yw - y_c;
            const double r_sq = x * x + y * y;
            const double dT   = dT_p * std::exp(1 - r_sq);
            const double vm   = v_p * std::exp((1 - r_sq) * 0.5);
            double       qq[5];
            qq[0] = rho_bg + std::pow(1 + dT, 1 / (a - 1)) - 1;
            qq[1] = u_bg - vm * y;
            qq[2] = v_bg + vm * x;
            if (dim == 3)
                qq[3] = w_bg;
            qq[dim + 1] = p_bg + std::pow(1 + dT, a / (a - 1)) - 1;
            double uu[5];
            consv_from_primv(qq, uu, a, e, dim);
            for (int n = 0; n < dim + 2; ++n)
                const_cast<BoxLayout &>(lay).at(U, b, i, j, k, n) = uu[n];
        });
        return;
    }
    if (pp.contains("density_L"))
    {
        double ql[5] = { 0, 0, 0, 0, 0 }, qr[5] = { 0, 0, 0, 0, 0 };
        double vl, vr;
        pp.get("density_L", ql[0]);
        pp.get("velocity_L", vl);
        pp.get("pressure_L", ql[dim + 1]);
        pp.get("density_R", qr[0]);
        pp.get("velocity_R", vr);
        pp.get("pressure_R", qr[dim + 1]);
        // 1D: along x; 2D/3D: along y (Prob.cpp:62-89)
        const int vdir = dim == 1 ? 1 : 2;
        ql[vdir]       = vl;
        qr[vdir]       = vr;
        double cl[5], cr[5];
        consv_from_primv(ql, cl, a, e, dim);
        consv_from_primv(qr, cr, a, e, dim);
        (void)u;
        for_valid(lay, [&](int b, int i, int j, int k) {
            bool left;
            if (dim == 1)
            {
                const double x_mid = (g.prob_lo[0] + g.prob_hi[0]) * 0.5;
                left               = g.prob_lo[0] + (i + 0.5) * g.dx[0] < x_mid;
            }
            else
                left = g.prob_lo[1] + (j + 0.5) * g.dx[1] < 0.5;
            for (int n = 0; n < dim + 2; ++n)
                const_cast<BoxLayout &>(lay).at(U, b, i, j, k, n)
                    = left ? cl[n] : cr[n];
        });
        return;
    }
    throw std::runtime_error("initdata: no known problem in the init.* keys");
}

} // namespace bamrex
