// IMEX Runge-Kutta level of the C++ host (adv.num_method = imex): the
// reference's AmrLevelAdv::advance dispatch to advance_imex_rk_stab
// (Source/AmrLevelAdv_advance.cpp:131-157) with the IMEX settings of
// Source/IMEX/IMEXSettings.cpp:11-78, on one periodic 2D level, every
// stage's explicit work and every pressure solve on the GPU through the C
// ABI (bx_imex_*, bx_helm_*).
#pragma once
#include <string>
#include <vector>

#include "../../../include/bamrex_hip.h"
#include "ParmParse.H"

namespace bamrex {

// IMEXSettings (Source/IMEX/IMEXSettings.H, read by get_imex_settings,
// IMEXSettings.cpp:11-78) for the linearised stages this host runs, plus the
// pressure solve's method and tolerance.  The reference fixes the solve in
// details::solve_pressure (IMEX_O1.cpp:616-640): MLMG to rtol 1e-12 (atol
// 1e-15), at most 1000 iterations; imex.pressure_solver / pressure_rtol /
// pressure_maxiter (this build's keys) choose another method or tolerance.
struct IMEXSettings
{
    int         advection_flux  = BX_IMEX_RUSANOV;
    int         enthalpy_method = BX_IMEX_ENTHALPY_REUSE_PRESSURE;
    int         ke_method       = BX_IMEX_KE_EX;
    std::string butcher_tableau;
    double      acoustic_timestep_end_time = 0;
    int         solver   = BX_SOLVE_MG_FUSED;
    double      rtol     = 1e-12;
    int         maxiter  = 1000;
};

IMEXSettings get_imex_settings(double stop_time);

// IMEXButcherTableau (Source/IMEX/ButcherTableau.cpp:5-101)
struct ButcherTableau
{
    std::string                      name;
    int                              n_imp = 0;
    std::vector<std::vector<double>> A_imp, A_exp;
    std::vector<double>              b;
    static ButcherTableau            from_string(const std::string &name);
};

class ImexLevel
{
  public:
    ~ImexLevel();

    // geometry, prob.*, adv.cfl and imex.* keys; the level must be 2D and
    // periodic (geometry.is_periodic = 1 1 or bc.x / bc.y = periodic)
    void read_params(double stop_time);
    void define();
    // Exec/GreshoVortex/Prob.cpp:51-106 (init.M), and the initial pressure
    // of the state (the pressure MultiFab's initial values)
    void initData();

    // estTimeStep for imex (AmrLevelAdv.cpp:417-437): acoustic before
    // imex.acoustic_timestep_end_time, advective after
    double estTimeStep(double cur_time);
    // computeNewDt (AmrLevelAdv.cpp:537-619): the stop-time clamp
    double computeNewDt(double cur_time, double stop_time);
    // advance_imex_rk (Source/IMEX/IMEX_RK.cpp:22-96): one step of dt
    void advance(double time, double dt);
    // the main.cpp loop: steps until max_step or stop_time
    int run(double stop_time, int max_step, double &time);

    void writePlotFile(const std::string &path, double time, int step) const;

    int    nx() const { return m_nx; }
    int    ny() const { return m_ny; }
    // pressure solves so far and their total iterations
    int    solves() const { return m_solves; }
    long   solve_iters() const { return m_solve_iters; }

  private:
    double *alloc(size_t n);

    IMEXSettings   m_set;
    ButcherTableau m_tab;
    int            m_nx = 0, m_ny = 0;
    double         m_dx = 0, m_dy = 0, m_prob_lo[3] = { 0, 0, 0 }, m_prob_hi[3] = { 1, 1, 1 };
    double         m_gamma = 1.4, m_eps = 1, m_cfl = 0.45;
    std::vector<double *> m_bufs;
    double        *m_U = nullptr, *m_Unew = nullptr, *m_p = nullptr;
    double        *m_Uex = nullptr, *m_enth = nullptr, *m_rhs = nullptr, *m_a = nullptr;
    double        *m_bX = nullptr, *m_bY = nullptr, *m_cX = nullptr, *m_cY = nullptr;
    double        *m_Uimp = nullptr, *m_Uexp[2] = { nullptr, nullptr };
    std::vector<double *> m_Fx, m_Fy;
    bx_helm       *m_helm   = nullptr;
    void          *m_stream = nullptr;
    int            m_solves = 0;
    long           m_solve_iters = 0;
};

} // namespace bamrex
