#include "ImexLevel.H"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace bamrex {

static void hcheck(int rc, const char *what)
{
    if (rc != BX_OK)
        throw std::runtime_error(std::string(what) + ": " + bx_helm_last_error());
}

static void bcheck(int rc, const char *what)
{
    if (rc != BX_OK)
        throw std::runtime_error(std::string(what) + ": " + bx_last_error());
}

// get_imex_settings (IMEXSettings.cpp:11-78) and acoustic_timestep_end_time
// (AmrLevelAdv.cpp:866-869)
IMEXSettings get_imex_settings(double stop_time)
{
    IMEXSettings s;
    ParmParse    pp("imex");
    std::string  adv_flux;
    pp.query("advection_flux", adv_flux);
    if (adv_flux.empty() || adv_flux == "rusanov")
        s.advection_flux = BX_IMEX_RUSANOV;
    else if (adv_flux == "muscl_rusanov")
        s.advection_flux = BX_IMEX_MUSCL_RUSANOV;
    else
        throw std::runtime_error("Invalid advection flux specified! Valid options are "
                                 "rusanov and muscl_rusanov.");
    bool linearise = true;
    pp.get("linearise", linearise);
    if (!linearise)
        throw std::runtime_error("imex.linearise = false (Picard iteration) is not run by "
                                 "this host; the library's bx_imex_picard_update provides it");
    std::string enth;
    pp.get("enthalpy_method", enth);
    if (enth == "reuse_pressure")
        s.enthalpy_method = BX_IMEX_ENTHALPY_REUSE_PRESSURE;
    else if (enth == "conservative")
        s.enthalpy_method = BX_IMEX_ENTHALPY_CONSERVATIVE;
    else if (enth == "conservative_ex")
        s.enthalpy_method = BX_IMEX_ENTHALPY_CONSERVATIVE_EX;
    else
        throw std::runtime_error("Invalid enthalpy method specified! Valid options are "
                                 "reuse_pressure, conservative, and conservative_ex");
    std::string ke;
    pp.get("ke_method", ke);
    if (ke == "conservative")
        s.ke_method = BX_IMEX_KE_CONSERVATIVE_KINETIC;
    else if (ke == "ex")
        s.ke_method = BX_IMEX_KE_EX;
    else if (ke == "split")
        s.ke_method = BX_IMEX_KE_SPLIT;
    else
        throw std::runtime_error("Invalid KE method specified! Valid options are ex and split");
    pp.get("butcher_tableau", s.butcher_tableau);
    bool stabilize = false;
    pp.query("stabilize", stabilize);
    if (stabilize)
        throw std::runtime_error("imex.stabilize = true is not run by this host; the library's "
                                 "bx_imex_shock_indicator / bx_imex_stab_blend provide it");
    s.acoustic_timestep_end_time = 0.1 * stop_time;
    pp.query("acoustic_timestep_end_time", s.acoustic_timestep_end_time);
    // default: the blocked-smoother V-cycle, whose iterates equal the
    // per-colour one's bit for bit
    std::string solver = "mg_fused";
    pp.query("pressure_solver", solver);
    if (solver == "mg")
        s.solver = BX_SOLVE_MG;
    else if (solver == "mg_fused")
        s.solver = BX_SOLVE_MG_FUSED;
    else if (solver == "fgmres_mg")
        s.solver = BX_SOLVE_FGMRES_MG;
    else
        throw std::runtime_error("imex.pressure_solver: mg, mg_fused or fgmres_mg, not " + solver);
    pp.query("pressure_rtol", s.rtol);
    pp.query("pressure_maxiter", s.maxiter);
    return s;
}

// IMEXButcherTableau::{get_A_imp, get_A_exp, get_b} (ButcherTableau.cpp:5-101)
ButcherTableau ButcherTableau::from_string(const std::string &name)
{
    ButcherTableau t;
    t.name = name;
    if (name == "SP111")
    {
        t.A_imp = { { 1.0 } };
        t.A_exp = { { 0.0 } };
        t.b     = { 1.0 };
    }
    else if (name == "SSP222")
    {
        const double g = 1 - 1 / std::sqrt(2.0);
        t.A_imp = { { g, 0.0 }, { 1 - 2 * g, g } };
        t.A_exp = { { 0.0, 0.0 }, { 1.0, 0.0 } };
        t.b     = { 0.5, 0.5 };
    }
    else if (name == "SASSP322")
    {
        t.A_imp = { { 0.5, 0.0, 0.0 }, { -0.5, 0.5, 0.0 }, { 0.0, 0.5, 0.5 } };
        t.A_exp = { { 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0 }, { 0.0, 1.0, 0.0 } };
        t.b     = { 0.0, 0.5, 0.5 };
    }
    else if (name == "SASSP332")
    {
        const double third = 1.0 / 3;
        t.A_imp = { { 0.25, 0.0, 0.0 }, { 0.0, 0.25, 0.0 }, { third, third, third } };
        t.A_exp = { { 0.0, 0.0, 0.0 }, { 0.5, 0.0, 0.0 }, { 0.5, 0.5, 0.0 } };
        t.b     = { third, third, third };
    }
    else if (name == "SSP433")
    {
        const double a = 0.24169426078821, be = 0.06042356519705, et = 0.12915286960590;
        t.A_imp = { { a, 0.0, 0.0, 0.0 },
                    { -a, a, 0.0, 0.0 },
                    { 0.0, 1 - a, a, 0.0 },
                    { be, et, 0.5 - be - et - a, a } };
        t.A_exp = { { 0.0, 0.0, 0.0, 0.0 },
                    { 0.0, 0.0, 0.0, 0.0 },
                    { 0.0, 1.0, 0.0, 0.0 },
                    { 0.0, 0.25, 0.25, 0.0 } };
        t.b     = { 0.0, 1.0 / 6, 1.0 / 6, 2.0 / 3 };
    }
    else
        throw std::runtime_error("Invalid Butcher tableau " + name
                                 + " (SP111, SSP222, SASSP322, SASSP332, SSP433)");
    t.n_imp = (int)t.A_imp.size();
    return t;
}

ImexLevel::~ImexLevel()
{
    if (m_helm)
        bx_helm_destroy(m_helm);
    for (double *p : m_bufs)
        bx_free(p);
    if (m_stream)
        bx_stream_destroy(m_stream);
}

double *ImexLevel::alloc(size_t n)
{
    double *p = nullptr;
    bcheck(bx_malloc((void **)&p, sizeof(double) * n), "bx_malloc");
    m_bufs.push_back(p);
    return p;
}

void ImexLevel::read_params(double stop_time)
{
    ParmParse           pa("amr"), pg("geometry"), pprob("prob"), padv("adv"), pbc("bc");
    std::vector<int>    n_cell, per;
    std::vector<double> lo, hi;
    pa.getarr("n_cell", n_cell);
    pg.getarr("prob_lo", lo);
    pg.getarr("prob_hi", hi);
    pg.queryarr("is_periodic", per);
    if (n_cell.size() < 2 || (n_cell.size() > 2 && n_cell[2] > 1))
        throw std::runtime_error("adv.num_method = imex runs on 2D levels");
    m_nx = n_cell[0];
    m_ny = n_cell[1];
    for (int d = 0; d < 2; ++d)
    {
        m_prob_lo[d] = lo.at(d);
        m_prob_hi[d] = hi.at(d);
        std::vector<std::string> b;
        const bool periodic = (d < (int)per.size() && per[d])
                              || (pbc.queryarr(d == 0 ? "x" : "y", b) && b.size() == 2
                                  && b[0] == "periodic" && b[1] == "periodic");
        if (!periodic)
            throw std::runtime_error("adv.num_method = imex runs on periodic levels here "
                                     "(geometry.is_periodic = 1 1)");
    }
    m_dx = (m_prob_hi[0] - m_prob_lo[0]) / m_nx;
    m_dy = (m_prob_hi[1] - m_prob_lo[1]) / m_ny;
    pprob.get("adiabatic", m_gamma);
    pprob.query("epsilon", m_eps);
    padv.query("cfl", m_cfl);
    m_set = get_imex_settings(stop_time);
    m_tab = ButcherTableau::from_string(m_set.butcher_tableau);
}

void ImexLevel::define()
{
    const size_t n = (size_t)m_nx * m_ny;
    bcheck(bx_stream_create(&m_stream), "bx_stream_create");
    m_U    = alloc(4 * n);
    m_Unew = alloc(4 * n);
    m_p    = alloc(n);
    m_Uex  = alloc(4 * n);
    m_enth = alloc(n);
    m_rhs  = alloc(n);
    m_a    = alloc(n);
    m_bX   = alloc((size_t)(m_nx + 1) * m_ny);
    m_bY   = alloc((size_t)m_nx * (m_ny + 1));
    if (m_set.ke_method == BX_IMEX_KE_SPLIT)
    {
        m_cX = alloc(n);
        m_cY = alloc(n);
    }
    if (m_tab.n_imp > 1)
    {
        m_Uimp    = alloc(4 * n);
        m_Uexp[0] = alloc(4 * n);
        m_Uexp[1] = alloc(4 * n);
        for (int s = 0; s < m_tab.n_imp; ++s)
        {
            m_Fx.push_back(alloc(4 * n));
            m_Fy.push_back(alloc(4 * n));
        }
    }
    hcheck(bx_helm_create(m_nx, m_ny, m_dx, m_dy, 1, &m_helm), "bx_helm_create");
}

// Exec/GreshoVortex/Prob.cpp:17-106 with consv_from_primv (Euler.H:138-165)
void ImexLevel::initData()
{
    ParmParse pinit("init");
    double    M = 0;
    pinit.get("M", M);
    if (m_eps != 1)
        throw std::runtime_error("the Gresho vortex needs prob.epsilon = 1");
    const double dens = 1;
    const double p_0  = dens / (m_gamma * M * M);
    const size_t n    = (size_t)m_nx * m_ny;
    std::vector<double> U(4 * n);
    for (int j = 0; j < m_ny; ++j)
        for (int i = 0; i < m_nx; ++i)
        {
            const double x     = (i + 0.5) * m_dx + m_prob_lo[0];
            const double y     = (j + 0.5) * m_dy + m_prob_lo[1];
            const double r     = std::sqrt(x * x + y * y);
            const double theta = std::atan2(y, x);
            const double u_theta = (0 <= r && r < 0.2) ? 5 * r : ((r < 0.4) ? 2 - 5 * r : 0);
            const double u_x     = -u_theta * std::sin(theta);
            const double u_y     = u_theta * std::cos(theta);
            const double p
                = (0 <= r && r < 0.2)
                      ? p_0 + (double)25 / 2 * r * r
                      : ((r < 0.4) ? p_0 + (double)25 / 2 * r * r + 4 * (1 - 5 * r + std::log(5 * r))
                                   : p_0 - 2 + 4 * std::log(2));
            const size_t c = (size_t)j * m_nx + i;
            U[c]         = dens;
            U[n + c]     = dens * u_x;
            U[2 * n + c] = dens * u_y;
            double u_sq  = u_x * u_x;
            u_sq         = u_sq + u_y * u_y;
            U[3 * n + c] = p / (m_gamma - 1) + 0.5 * m_eps * dens * u_sq;
        }
    bcheck(bx_memcpy_htod(m_U, U.data(), sizeof(double) * U.size()), "htod");
    hcheck(bx_imex_pressure(m_U, m_nx, m_ny, m_gamma, m_eps, m_p, m_stream), "bx_imex_pressure");
    bcheck(bx_stream_sync(m_stream), "sync");
}

double ImexLevel::estTimeStep(double cur_time)
{
    double dt = 0;
    if (cur_time < m_set.acoustic_timestep_end_time)
        hcheck(bx_imex_acoustic_dt(m_U, m_nx, m_ny, m_dx, m_dy, m_gamma, m_eps, m_cfl, &dt,
                                   m_stream),
               "bx_imex_acoustic_dt");
    else
        hcheck(bx_imex_advective_dt(m_U, m_nx, m_ny, m_dx, m_dy, m_cfl, &dt, m_stream),
               "bx_imex_advective_dt");
    return dt;
}

double ImexLevel::computeNewDt(double cur_time, double stop_time)
{
    double       dt_0 = 1.0e+100;
    const double est  = estTimeStep(cur_time);
    dt_0              = (est < dt_0) ? est : dt_0;
    const double eps  = 0.001 * dt_0;
    if (stop_time >= 0.0 && (cur_time + dt_0) > (stop_time - eps))
        dt_0 = stop_time - cur_time;
    return dt_0;
}

// advance_imex_rk (IMEX_RK.cpp:22-96): per stage the stage states from the
// stored fluxes (stages after the first), the EK/SK pressure of the stage,
// the explicit update and the pressure system, the solve from the stage's
// pressure, and the corrected stage fluxes; then the b-weighted update.
// SP111 ends with the fused finish instead (the same arithmetic).
void ImexLevel::advance(double /*time*/, double dt)
{
    const int    ns    = m_tab.n_imp;
    const bool   sp111 = m_tab.name == "SP111";
    const bool   split = m_set.ke_method == BX_IMEX_KE_SPLIT;
    const double *exp_prev = nullptr;
    for (int s = 0; s < ns; ++s)
    {
        const double adt = m_tab.A_imp[s][s] * dt;
        const double *Ui = m_U, *Ue = m_U;
        if (s > 0)
        {
            double dt_imp[BX_IMEX_MAX_STAGES] = { 0 }, dt_exp[BX_IMEX_MAX_STAGES] = { 0 };
            for (int j = 0; j < s; ++j)
            {
                dt_imp[j] = m_tab.A_imp[s][j] * dt;
                dt_exp[j] = m_tab.A_exp[s][j] * dt;
            }
            double *ue = m_Uexp[s % 2];
            hcheck(bx_imex_rk_combine(m_U, m_Fx.data(), m_Fy.data(), s, dt_imp, dt_exp, m_nx,
                                      m_ny, m_dx, m_dy, m_Uimp, ue, m_stream),
                   "bx_imex_rk_combine");
            hcheck(bx_imex_rk_pressure(ue, exp_prev, m_nx, m_ny, m_dx, m_dy, adt, m_gamma, m_eps,
                                       m_set.advection_flux, m_set.ke_method, m_p, m_stream),
                   "bx_imex_rk_pressure");
            Ui = m_Uimp;
            Ue = ue;
        }
        exp_prev = Ue;
        hcheck(bx_imex_stage_assemble(Ui, Ue, m_p, m_nx, m_ny, m_dx, m_dy, adt, m_gamma, m_eps,
                                      m_set.advection_flux, sp111 ? 1 : 0, m_set.ke_method,
                                      m_set.enthalpy_method, m_Uex, m_enth, m_rhs, m_a, m_bX,
                                      m_bY, m_cX, m_cY, m_stream),
               "bx_imex_stage_assemble");
        // details::setup_pressure_op and solve_pressure (IMEX_O1.cpp:570-640):
        // alpha = 1, beta = adt^2, gamma_c = -eps adt / 2 with the split KE
        hcheck(bx_helm_set_coeffs(m_helm, 1.0, adt * adt, split ? -0.5 * m_eps * adt : 0.0, m_a,
                                  m_bX, m_bY, split ? m_cX : nullptr, split ? m_cY : nullptr,
                                  m_stream),
               "bx_helm_set_coeffs");
        bx_solve_stats st{};
        hcheck(bx_helm_solve(m_helm, m_rhs, m_p, m_set.solver, m_set.rtol, m_set.maxiter, &st,
                             m_stream),
               "bx_helm_solve");
        if (!st.converged)
            throw std::runtime_error("pressure solve did not converge (stage "
                                     + std::to_string(s) + ")");
        ++m_solves;
        m_solve_iters += st.iters;
        if (sp111)
            hcheck(bx_imex_peek_finish(m_U, m_Uex, m_enth, m_p, m_nx, m_ny, m_dx, m_dy, dt, m_eps,
                                       m_set.advection_flux, m_Unew, m_stream),
                   "bx_imex_peek_finish");
        else
        {
            hcheck(bx_imex_stage_fluxes(Ue, m_Uex, m_enth, m_p, m_nx, m_ny, m_dx, m_dy, adt, m_eps,
                                        m_set.advection_flux, 0, m_Fx[s], m_Fy[s], m_stream),
                   "bx_imex_stage_fluxes");
            if (s + 1 == ns)
                hcheck(bx_imex_rk_update(m_U, m_Fx.data(), m_Fy.data(), ns, m_tab.b.data(), m_nx,
                                         m_ny, m_dx, m_dy, dt, m_Unew, m_stream),
                       "bx_imex_rk_update");
        }
    }
    bcheck(bx_stream_sync(m_stream), "sync");
    std::swap(m_U, m_Unew);
}

int ImexLevel::run(double stop_time, int max_step, double &time)
{
    int steps = 0;
    while ((steps < max_step || max_step < 0) && (time < stop_time || stop_time < 0.0))
    {
        const double dt = computeNewDt(time, stop_time);
        advance(time, dt);
        time = time + dt;
        ++steps;
    }
    return steps;
}

// AmrLevel::writePlotFile: the level as one box without ghost cells (the
// IMEX arrays' layout is that FAB's)
void ImexLevel::writePlotFile(const std::string &path, double time, int step) const
{
    std::vector<double> U((size_t)4 * m_nx * m_ny);
    bcheck(bx_stream_sync(m_stream), "sync");
    bcheck(bx_memcpy_dtoh(U.data(), m_U, sizeof(double) * U.size()), "dtoh");
    bx_box box{};
    box.hi[0] = m_nx - 1;
    box.hi[1] = m_ny - 1;
    bx_plot_level lv{};
    lv.boxes = &box;
    lv.nbox  = 1;
    lv.ng    = 0;
    lv.U     = U.data();
    lv.domain_hi[0] = m_nx - 1;
    lv.domain_hi[1] = m_ny - 1;
    lv.dx[0] = m_dx;
    lv.dx[1] = m_dy;
    lv.level_steps = step;
    bcheck(bx_write_plotfile(path.c_str(), 2, 1, &lv, 2, m_prob_lo, m_prob_hi, time, m_gamma,
                             m_eps, 0),
           "bx_write_plotfile");
}

} // namespace bamrex
