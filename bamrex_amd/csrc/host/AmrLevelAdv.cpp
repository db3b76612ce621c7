#include "AmrLevelAdv.H"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <stdexcept>

namespace bamrex {

int         AmrLevelAdv::verbose    = 0;
double      AmrLevelAdv::cfl        = 0.9;
std::string AmrLevelAdv::num_method = "muscl_hancock";
int         AmrLevelAdv::box_loop   = 0;

static void check(int rc, const char *what)
{
    if (rc != BX_OK)
        throw std::runtime_error(std::string(what) + ": " + bx_last_error());
}

double &BoxLayout::at(std::vector<double> &U, int b, int i, int j, int k,
                      int n) const
{
    const bx_box &bx  = boxes[b];
    const int     nxd = (int)nx(b, 0) + 2 * ng;
    const int     nyd = dim >= 2 ? (int)nx(b, 1) + 2 * ng : 1;
    const int     nzd = dim >= 3 ? (int)nx(b, 2) + 2 * ng : 1;
    const int     gy = dim >= 2 ? ng : 0, gz = dim >= 3 ? ng : 0;
    const long long o = bx.offset + (i - bx.lo[0] + ng)
                        + (long long)(j - bx.lo[1] + gy) * nxd
                        + (long long)(k - bx.lo[2] + gz) * nxd * nyd
                        + (long long)n * nxd * nyd * nzd;
    return U[o];
}

AmrLevelAdv::~AmrLevelAdv()
{
    if (m_ctx)
        bx_destroy(m_ctx);
    if (m_S1)
        bx_free(m_S1);
    if (m_S2)
        bx_free(m_S2);
    if (m_flux)
        bx_free(m_flux);
    if (m_stream)
        bx_stream_destroy(m_stream);
}

// AmrLevelAdv::read_params (AmrLevelAdv.cpp:837-909) and
// FluidSystem::ReadParameters (FluidSystem.cpp:10-71) for the hot path.
void AmrLevelAdv::read_params(const ParmParse &root)
{
    ParmParse pp("adv");
    pp.query("v", verbose);
    pp.query("cfl", cfl);
    pp.get("num_method", num_method);
    pp.query("box_loop", box_loop);

    std::vector<int>    n_cell;
    std::vector<double> lo, hi;
    std::vector<int>    per;
    ParmParse           pa("amr"), pg("geometry");
    pa.getarr("n_cell", n_cell);
    pg.getarr("prob_lo", lo);
    pg.getarr("prob_hi", hi);
    pg.queryarr("is_periodic", per);
    int dim = 0;
    root.query("dim", dim);
    if (dim == 0)
        dim = (int)std::min<size_t>(3, n_cell.size());
    m_geom.dim = dim;
    for (int d = 0; d < dim; ++d)
    {
        m_geom.n_cell[d]      = n_cell.at(d);
        m_geom.prob_lo[d]     = lo.at(d);
        m_geom.prob_hi[d]     = hi.at(d);
        m_geom.is_periodic[d] = d < (int)per.size() ? per[d] : 0;
        m_geom.dx[d] = (m_geom.prob_hi[d] - m_geom.prob_lo[d]) / m_geom.n_cell[d];
    }

    m_parm.dim = dim;
    // NumericalMethods::enum_from_string (Source/NumericalMethods.H:22-44),
    // restricted to the explicit methods
    if (num_method == "muscl_hancock")
        m_parm.num_method = BX_METHOD_MUSCL_HANCOCK;
    else if (num_method == "hllc")
        m_parm.num_method = BX_METHOD_HLLC;
    else if (num_method == "slic")
        m_parm.num_method = BX_METHOD_SLIC;
    else if (num_method == "force")
        m_parm.num_method = BX_METHOD_FORCE;
    else if (num_method == "lax_friedrichs")
        m_parm.num_method = BX_METHOD_LAX_FRIEDRICHS;
    else
        throw std::runtime_error("Invalid numerical method \"" + num_method
                                 + "\" for the explicit hot path");
    ParmParse pprob("prob"), plim("limiting");
    pprob.get("adiabatic", m_parm.adiabatic);
    m_parm.epsilon = 1;
    pprob.query("epsilon", m_parm.epsilon);
    std::string limiter = "van_leer";
    plim.query("limiter", limiter);
    if (limiter == "minbee")
        m_parm.limiter = BX_LIMITER_MINBEE;
    else if (limiter == "van_leer")
        m_parm.limiter = BX_LIMITER_VAN_LEER;
    else if (limiter == "van_albada")
        m_parm.limiter = BX_LIMITER_VAN_ALBADA;
    else if (limiter == "superbee")
        m_parm.limiter = BX_LIMITER_SUPERBEE;
    else
        throw std::runtime_error("Invalid limiter " + limiter);
    const int E = dim + 1;
    for (int n = 0; n < 5; ++n)
    {
        m_parm.limiter_indices[n] = E;
        m_parm.limit_on_primv[n]  = 0;
    }
    auto var = [&](const std::string &name, int &idx, int &prim) {
        if (name == "density") { idx = 0; prim = 0; }
        else if (name == "mom_x") { idx = 1; prim = 0; }
        else if (name == "mom_y" && dim >= 2) { idx = 2; prim = 0; }
        else if (name == "mom_z" && dim >= 3) { idx = 3; prim = 0; }
        else if (name == "energy") { idx = E; prim = 0; }
        else if (name == "u_x") { idx = 1; prim = 1; }
        else if (name == "u_y" && dim >= 2) { idx = 2; prim = 1; }
        else if (name == "u_z" && dim >= 3) { idx = 3; prim = 1; }
        else if (name == "p") { idx = E; prim = 1; }
        else throw std::runtime_error("Invalid variable name " + name);
    };
    std::string                s;
    std::vector<std::string>   moms;
    if (plim.query("density", s))
        var(s, m_parm.limiter_indices[0], m_parm.limit_on_primv[0]);
    if (plim.query("energy", s))
        var(s, m_parm.limiter_indices[E], m_parm.limit_on_primv[E]);
    if (plim.queryarr("mom", moms))
        for (int d = 0; d < dim && d < (int)moms.size(); ++d)
            var(moms[d], m_parm.limiter_indices[1 + d], m_parm.limit_on_primv[1 + d]);

    // bc.{x,y,z} (BCs.cpp:35-74), periodic geometry overrides (:220-224)
    ParmParse   pbc("bc");
    const char *ax[3] = { "x", "y", "z" };
    for (int d = 0; d < dim; ++d)
    {
        std::vector<std::string> b;
        m_bc_lo[d] = m_bc_hi[d] = BX_BC_TRANSMISSIVE;
        if (pbc.queryarr(ax[d], b))
        {
            auto conv = [](const std::string &t) {
                if (t == "transmissive") return BX_BC_TRANSMISSIVE;
                if (t == "periodic") return BX_BC_PERIODIC;
                if (t == "reflective") return BX_BC_REFLECTIVE;
                throw std::runtime_error("Boundary type not implemented!");
            };
            m_bc_lo[d] = conv(b.at(0));
            m_bc_hi[d] = conv(b.at(1));
        }
        if (m_geom.is_periodic[d])
            m_bc_lo[d] = m_bc_hi[d] = BX_BC_PERIODIC;
    }
}

// BoxArray(domain).maxSize(max_grid_size), packed FABs, device buffers
void AmrLevelAdv::define(int max_grid_size, int ng)
{
    const int dim = m_geom.dim;
    std::vector<bx_box> boxes;
    std::vector<std::vector<std::pair<int, int>>> r(3);
    for (int d = 0; d < 3; ++d)
    {
        if (d >= dim)
        {
            r[d].push_back({ 0, 0 });
            continue;
        }
        for (int a = 0; a < m_geom.n_cell[d]; a += max_grid_size)
            r[d].push_back({ a, std::min(a + max_grid_size, m_geom.n_cell[d]) - 1 });
    }
    for (auto &z : r[2])
        for (auto &y : r[1])
            for (auto &x : r[0])
            {
                bx_box b{};
                b.lo[0] = x.first;  b.hi[0] = x.second;
                b.lo[1] = y.first;  b.hi[1] = y.second;
                b.lo[2] = z.first;  b.hi[2] = z.second;
                boxes.push_back(b);
            }
    define(boxes, ng);
}

// The level on a given box list (a restart takes the checkpoint's)
void AmrLevelAdv::define(const std::vector<bx_box> &boxes, int ng)
{
    const int dim = m_geom.dim;
    m_layout.dim  = dim;
    m_layout.ng   = ng;
    m_layout.nc   = dim + 2;
    m_layout.boxes = boxes;
    std::vector<int64_t> offs(m_layout.boxes.size());
    check(bx_pack_offsets(dim, ng, m_layout.boxes.data(), (int)offs.size(),
                          offs.data(), &m_layout.total),
          "bx_pack_offsets");
    for (size_t b = 0; b < offs.size(); ++b)
        m_layout.boxes[b].offset = offs[b];

    check(bx_create(&m_parm, &m_ctx), "bx_create");
    bx_geometry g{};
    for (int d = 0; d < 3; ++d)
    {
        g.domain_lo[d] = 0;
        g.domain_hi[d] = d < dim ? m_geom.n_cell[d] - 1 : 0;
        g.bc_lo[d]     = m_bc_lo[d];
        g.bc_hi[d]     = m_bc_hi[d];
    }
    g.ng = ng;
    check(bx_set_level(m_ctx, &g, m_layout.boxes.data(),
                       (int)m_layout.boxes.size()),
          "bx_set_level");
    check(bx_stream_create(&m_stream), "bx_stream_create");
    check(bx_malloc((void **)&m_S1, sizeof(double) * m_layout.total), "bx_malloc");
    check(bx_malloc((void **)&m_S2, sizeof(double) * m_layout.total), "bx_malloc");
    check(bx_memset(m_S1, 0, sizeof(double) * m_layout.total), "bx_memset");
    check(bx_memset(m_S2, 0, sizeof(double) * m_layout.total), "bx_memset");
    if (box_loop)
    {
        long long mx = 0;
        for (size_t b = 0; b < m_layout.boxes.size(); ++b)
        {
            long long v = 1;
            for (int d = 0; d < dim; ++d)
                v *= m_layout.nx((int)b, d) + 1;
            mx = std::max(mx, v);
        }
        check(bx_malloc((void **)&m_flux, sizeof(double) * mx * (dim + 2)),
              "bx_malloc");
    }
}

void AmrLevelAdv::initData(const ParmParse &root)
{
    std::vector<double> U(m_layout.total, 0.0);
    initdata(U, m_layout, m_geom, m_parm, root);
    upload(U);
    check(bx_fill_boundary(m_ctx, m_S1, m_stream), "bx_fill_boundary");
}

void AmrLevelAdv::upload(const std::vector<double> &U)
{
    check(bx_memcpy_htod(m_S1, U.data(), sizeof(double) * U.size()), "htod");
}

void AmrLevelAdv::download(std::vector<double> &U) const
{
    U.resize(m_layout.total);
    check(bx_stream_sync(m_stream), "sync");
    check(bx_memcpy_dtoh(U.data(), m_S1, sizeof(double) * U.size()), "dtoh");
}

// estTimeStep (AmrLevelAdv.cpp:395-477): cfl * min_d dx_d / max wave speed
double AmrLevelAdv::estTimeStep(double cur_time)
{
    double dt = 0;
    check(bx_est_dt(m_ctx, m_S1, m_geom.dx, cfl, cur_time, -1.0, &dt, m_stream),
          "bx_est_dt");
    if (verbose)
        std::printf("AmrLevelAdv::estTimeStep at level 0:  dt_est = %.17g\n", dt);
    return dt;
}

// computeNewDt (AmrLevelAdv.cpp:537-619) on a single level
double AmrLevelAdv::computeNewDt(double cur_time, double stop_time)
{
    double dt = 0;
    check(bx_est_dt(m_ctx, m_S1, m_geom.dx, cfl, cur_time, stop_time, &dt,
                    m_stream),
          "bx_est_dt");
    return dt;
}

// advance (AmrLevelAdv_advance.cpp:21-259), explicit methods, one level
double AmrLevelAdv::advance(double time, double dt, int /*iteration*/,
                            int /*ncycle*/)
{
    if (verbose)
        std::printf("Starting advance:\n");
    advance_explicit(time, dt);
    if (verbose)
        std::printf("Advance complete!\n");
    return dt;
}

// advance_explicit (AmrLevelAdv_advance.cpp:261-439): x, y, z sweeps with the
// full dt, ghost refill after each sweep.
void AmrLevelAdv::advance_explicit(double /*time*/, double dt)
{
    if (!box_loop)
    {
        check(bx_advance(m_ctx, dt, m_geom.dx, m_S1, m_S2, m_stream), "bx_advance");
        return;
    }
    // Reference-shaped loop: for each direction, for each box (MFIter):
    // compute_MUSCL_hancock_flux / compute_HLLC_flux, conservative update;
    // then fill_consv_boundary; swap(MFin, MFout).
    double *in = m_S1, *out = m_S2;
    for (int d = 0; d < m_geom.dim; ++d)
    {
        for (int b = 0; b < (int)m_layout.boxes.size(); ++b)
        {
            check(bx_flux_box(m_ctx, b, d, dt, m_geom.dx, in, m_flux, 0, m_stream),
                  "bx_flux_box");
            check(bx_update_box(m_ctx, b, d, dt, m_geom.dx, in, m_flux, 0, out,
                                m_stream),
                  "bx_update_box");
        }
        check(bx_fill_boundary(m_ctx, out, m_stream), "bx_fill_boundary");
        std::swap(in, out);
    }
    if (in != m_S1) // MultiFab::Copy(S_new, *MFin) (:438)
        check(bx_memcpy_dtod(m_S1, in, sizeof(double) * m_layout.total), "dtod");
}

int AmrLevelAdv::run(double stop_time, int max_step, double &time)
{
    if (box_loop)
    {
        int steps = 0;
        while ((steps < max_step || max_step < 0)
               && (time < stop_time || stop_time < 0.0))
        {
            const double dt = computeNewDt(time, stop_time);
            advance(time, dt, 1, 1);
            time = time + dt;
            ++steps;
        }
        return steps;
    }
    double t_out = 0;
    int    steps = 0, in_tmp = 0;
    check(bx_run(m_ctx, m_geom.dx, cfl, time, stop_time, max_step, m_S1, m_S2,
                 1, m_stream, &t_out, &steps, &in_tmp),
          "bx_run");
    if (in_tmp) // an odd number of sweeps in graph mode ends in the second buffer
        std::swap(m_S1, m_S2);
    time = t_out;
    return steps;
}

// AmrLevel::checkPoint (AmrLevelAdv.cpp:108-113): time, step, the dt of the
// next step, boxes and the state with its ghost cells, written by the library
void AmrLevelAdv::checkPoint(const std::string &path, double time, int step,
                             double dt_next) const
{
    check(bx_stream_sync(m_stream), "sync");
    bx_chk_level lv{};
    lv.boxes = m_layout.boxes.data();
    lv.nbox  = (int)m_layout.boxes.size();
    lv.ng    = m_layout.ng;
    lv.U     = m_S1;
    for (int d = 0; d < 3; ++d)
    {
        lv.domain_lo[d] = 0;
        lv.domain_hi[d] = d < m_geom.dim ? m_geom.n_cell[d] - 1 : 0;
        lv.dx[d]        = d < m_geom.dim ? m_geom.dx[d] : 0.0;
    }
    lv.level_steps = step;
    lv.dt          = dt_next;
    check(bx_write_checkpoint(path.c_str(), m_geom.dim, 1, &lv, 2, m_geom.prob_lo,
                              m_geom.prob_hi, time, m_stream),
          "bx_write_checkpoint");
}

// AmrLevelAdv::restart (AmrLevelAdv.cpp:95-106): the level on the
// checkpoint's boxes with its state; returns the checkpoint's time and step
void AmrLevelAdv::restart(const std::string &path, double &time, int &step)
{
    bx_chk_info       info{};
    bx_chk_level_info li{};
    check(bx_read_checkpoint_info(path.c_str(), &info, &li, 1), "bx_read_checkpoint_info");
    if (info.dim != m_geom.dim)
        throw std::runtime_error("restart: the checkpoint is " + std::to_string(info.dim)
                                 + "D, the inputs " + std::to_string(m_geom.dim) + "D");
    for (int d = 0; d < m_geom.dim; ++d)
        if (li.domain_hi[d] - li.domain_lo[d] + 1 != m_geom.n_cell[d])
            throw std::runtime_error("restart: the checkpoint's domain differs from amr.n_cell");
    std::vector<bx_box> boxes(li.nbox);
    check(bx_read_checkpoint_boxes(path.c_str(), 0, boxes.data(), li.nbox),
          "bx_read_checkpoint_boxes");
    define(boxes, li.ng);
    check(bx_read_checkpoint_level(path.c_str(), 0, m_S1, m_layout.total, m_stream),
          "bx_read_checkpoint_level");
    time = info.time;
    step = li.level_steps;
}

// AmrLevel::writePlotFile (AmrLevelAdv.cpp:116-136): an AMReX plotfile
// directory of this level's state, written by the library.
void AmrLevelAdv::writePlotFile(const std::string &path, double time,
                                int step) const
{
    std::vector<double> U;
    download(U);
    bx_plot_level lv{};
    lv.boxes = m_layout.boxes.data();
    lv.nbox  = (int)m_layout.boxes.size();
    lv.ng    = m_layout.ng;
    lv.U     = U.data();
    for (int d = 0; d < 3; ++d)
    {
        lv.domain_lo[d] = 0;
        lv.domain_hi[d] = d < m_geom.dim ? m_geom.n_cell[d] - 1 : 0;
        lv.dx[d]        = d < m_geom.dim ? m_geom.dx[d] : 0.0;
    }
    lv.level_steps = step;
    check(bx_write_plotfile(path.c_str(), m_geom.dim, 1, &lv, 2, m_geom.prob_lo,
                            m_geom.prob_hi, time, m_parm.adiabatic, m_parm.epsilon, 0),
          "bx_write_plotfile");
}

} // namespace bamrex
