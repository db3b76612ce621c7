// AMReX plotfile writer (host code): the on-disk format the reference
// produces through AmrLevel::writePlotFile (Source/AmrLevelAdv.cpp:116-136),
// so outputs open in VisIt / yt and the reference's post-processing.
//
// Layout (HyperCLaw-V1.1, one file of FAB data per level):
//   <dir>/Header                 variables, geometry, boxes of every level
//   <dir>/Level_<l>/Cell_H       VisMF header: box list, data offsets, per-box
//                                component minima and maxima
//   <dir>/Level_<l>/Cell_D_00000 per box: a FAB header line, then the valid
//                                cells as IEEE little-endian doubles,
//                                component after component, x fastest
// Variables: the conserved state (FluidSystem::VariableNames,
// Source/System/FluidSystem.cpp:103-109) and optionally the derived
// pressure and velocities (SetupDeriveList, :111-137; the derive kernels of
// Source/System/Euler/EulerDerive.cpp:9-33).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bx_level.h"

namespace {

std::string fmt(double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

std::string ivect(const int *v, int dim)
{
    std::string s = "(";
    for (int d = 0; d < dim; ++d)
        s += (d ? "," : "") + std::to_string(v[d]);
    return s + ")";
}

std::string box_str(const int *lo, const int *hi, int dim)
{
    const int zero[3] = { 0, 0, 0 };
    return "(" + ivect(lo, dim) + " " + ivect(hi, dim) + " " + ivect(zero, dim) + ")";
}

bool make_dir(const std::string &p)
{
    return mkdir(p.c_str(), 0755) == 0 || errno == EEXIST;
}

// one plotted variable of cell (box data pointer, cell offset, component stride)
double plot_value(int var, int dim, const double *U, long long o, long long ns,
                  double gamma, double eps)
{
    const int nc = dim + 2;
    if (var < nc)
        return U[o + var * ns];
    const double rho = U[o];
    if (var == nc) // pressure: (gamma - 1)(E - ke), ke = eps |m|^2 / (2 rho)
    {
        double m2 = U[o + ns] * U[o + ns];
        for (int d = 1; d < dim; ++d)
            m2 = m2 + U[o + (1 + d) * ns] * U[o + (1 + d) * ns];
        const double ke = eps * m2 / (2 * rho);
        return (gamma - 1) * (U[o + (nc - 1) * ns] - ke);
    }
    return U[o + (var - nc) * ns] / rho; // u_d = m_d / rho
}

} // namespace

extern "C" int bx_write_plotfile(const char *dir, int dim, int nlev,
                                 const bx_plot_level *levels, int ref_ratio,
                                 const double prob_lo[3], const double prob_hi[3],
                                 double time, double gamma, double eps, int derived)
{
    if (!dir || dim < 1 || dim > 3 || nlev < 1 || !levels || !prob_lo || !prob_hi
        || (nlev > 1 && ref_ratio < 2))
        return set_error(BX_ERR_INVALID, "bx_write_plotfile: bad argument");
    for (int l = 0; l < nlev; ++l)
        if (!levels[l].boxes || levels[l].nbox < 1 || !levels[l].U || levels[l].ng < 0)
            return set_error(BX_ERR_INVALID, "bx_write_plotfile: bad level");
    const int                nc = dim + 2;
    std::vector<std::string> names = { "density" };
    const char              *mom[3] = { "mom_x", "mom_y", "mom_z" };
    const char              *vel[3] = { "u_x", "u_y", "u_z" };
    for (int d = 0; d < dim; ++d)
        names.push_back(mom[d]);
    names.push_back("energy");
    if (derived)
    {
        names.push_back("pressure");
        for (int d = 0; d < dim; ++d)
            names.push_back(vel[d]);
    }
    const int   nvar = (int)names.size();
    std::string root(dir);
    if (!make_dir(root))
        return set_error(BX_ERR_INVALID, "bx_write_plotfile: cannot create " + root);

    std::ostringstream H;
    H << "HyperCLaw-V1.1\n" << nvar << '\n';
    for (auto &n : names)
        H << n << '\n';
    H << dim << '\n' << fmt(time) << '\n' << nlev - 1 << '\n';
    for (int d = 0; d < dim; ++d)
        H << fmt(prob_lo[d]) << ' ';
    H << '\n';
    for (int d = 0; d < dim; ++d)
        H << fmt(prob_hi[d]) << ' ';
    H << '\n';
    for (int l = 0; l + 1 < nlev; ++l)
        H << ref_ratio << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l)
        H << box_str(levels[l].domain_lo, levels[l].domain_hi, dim) << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l)
        H << levels[l].level_steps << ' ';
    H << '\n';
    for (int l = 0; l < nlev; ++l)
    {
        for (int d = 0; d < dim; ++d)
            H << fmt(levels[l].dx[d]) << ' ';
        H << '\n';
    }
    H << "0\n0\n"; // cartesian coordinates, boundary width 0

    for (int l = 0; l < nlev; ++l)
    {
        const bx_plot_level &L   = levels[l];
        const std::string    ldir = root + "/Level_" + std::to_string(l);
        if (!make_dir(ldir))
            return set_error(BX_ERR_INVALID, "bx_write_plotfile: cannot create " + ldir);
        H << l << ' ' << L.nbox << ' ' << fmt(time) << '\n' << L.level_steps << '\n';
        for (int b = 0; b < L.nbox; ++b)
            for (int d = 0; d < dim; ++d)
                H << fmt(prob_lo[d] + L.dx[d] * L.boxes[b].lo[d]) << ' '
                  << fmt(prob_lo[d] + L.dx[d] * (L.boxes[b].hi[d] + 1)) << '\n';
        H << "Level_" << l << "/Cell\n";

        std::ofstream D(ldir + "/Cell_D_00000", std::ios::binary);
        if (!D)
            return set_error(BX_ERR_INVALID, "bx_write_plotfile: cannot write " + ldir);
        std::vector<long long>           offsets;
        std::vector<std::vector<double>> mins, maxs;
        std::vector<double>              buf;
        for (int b = 0; b < L.nbox; ++b)
        {
            const bx_box &bx = L.boxes[b];
            int           n[3] = { 1, 1, 1 }, g[3] = { 0, 0, 0 };
            for (int d = 0; d < dim; ++d)
            {
                n[d] = bx.hi[d] - bx.lo[d] + 1;
                g[d] = L.ng;
            }
            const long long sj = n[0] + 2 * g[0];
            const long long sk = sj * (n[1] + 2 * g[1]);
            const long long ns = sk * (n[2] + 2 * g[2]);
            const double   *U  = L.U + bx.offset;
            const long long nv = (long long)n[0] * n[1] * n[2];
            buf.resize((size_t)(nv * nvar));
            std::vector<double> mn(nvar), mx(nvar);
            for (int v = 0; v < nvar; ++v)
            {
                double   *out = buf.data() + (size_t)v * nv;
                long long q   = 0;
                for (int k = 0; k < n[2]; ++k)
                    for (int j = 0; j < n[1]; ++j)
                        for (int i = 0; i < n[0]; ++i)
                        {
                            const long long o
                                = (i + g[0]) + (j + g[1]) * sj + (k + g[2]) * sk;
                            out[q++] = plot_value(v, dim, U, o, ns, gamma, eps);
                        }
                mn[v] = *std::min_element(out, out + nv);
                mx[v] = *std::max_element(out, out + nv);
            }
            offsets.push_back((long long)D.tellp());
            const int lo[3] = { bx.lo[0], bx.lo[1], bx.lo[2] };
            const int hi[3] = { bx.hi[0], bx.hi[1], bx.hi[2] };
            D << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))"
              << box_str(lo, hi, dim) << ' ' << nvar << '\n';
            D.write(reinterpret_cast<const char *>(buf.data()),
                    (std::streamsize)(sizeof(double) * buf.size()));
            mins.push_back(mn);
            maxs.push_back(mx);
        }
        if (!D)
            return set_error(BX_ERR_INVALID, "bx_write_plotfile: write failed in " + ldir);

        std::ofstream C(ldir + "/Cell_H");
        C << "1\n1\n" << nvar << "\n0\n(" << L.nbox << " 0\n";
        for (int b = 0; b < L.nbox; ++b)
            C << box_str(L.boxes[b].lo, L.boxes[b].hi, dim) << '\n';
        C << ")\n" << L.nbox << '\n';
        for (int b = 0; b < L.nbox; ++b)
            C << "FabOnDisk: Cell_D_00000 " << offsets[b] << '\n';
        C << '\n' << L.nbox << ',' << nvar << '\n';
        for (auto &m : mins)
        {
            for (double v : m)
                C << fmt(v) << ',';
            C << '\n';
        }
        C << '\n' << L.nbox << ',' << nvar << '\n';
        for (auto &m : maxs)
        {
            for (double v : m)
                C << fmt(v) << ',';
            C << '\n';
        }
        if (!C)
            return set_error(BX_ERR_INVALID, "bx_write_plotfile: cannot write Cell_H");
    }
    std::ofstream hf(root + "/Header");
    hf << H.str();
    if (!hf)
        return set_error(BX_ERR_INVALID, "bx_write_plotfile: cannot write Header");
    return BX_OK;
}
