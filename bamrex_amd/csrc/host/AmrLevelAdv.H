// Single-level host integrator with the shape of the reference's AmrLevelAdv
// (Source/AmrLevelAdv.H:37-283): read_params, initData, estTimeStep,
// computeNewDt, advance -> advance_explicit, driving the HIP hot path through
// the C ABI of include/bamrex_hip.h.
#pragma once
#include <string>
#include <vector>

#include "../../../include/bamrex_hip.h"
#include "ParmParse.H"

namespace bamrex {

struct Geometry
{
    int    dim = 2;
    int    n_cell[3]      = { 1, 1, 1 };
    double prob_lo[3]     = { 0, 0, 0 };
    double prob_hi[3]     = { 1, 1, 1 };
    int    is_periodic[3] = { 0, 0, 0 };
    double dx[3]          = { 1, 1, 1 };
};

struct BoxLayout
{
    int                    dim = 2, ng = 2, nc = 4;
    std::vector<bx_box>    boxes;
    int64_t                total = 0;
    // host view helpers
    long long nx(int b, int d) const { return boxes[b].hi[d] - boxes[b].lo[d] + 1; }
    double   &at(std::vector<double> &U, int b, int i, int j, int k, int n) const;
};

class AmrLevelAdv
{
  public:
    // AmrLevelAdv static parameters (AmrLevelAdv.cpp:31-48)
    static int         verbose;
    static double      cfl;
    static std::string num_method;
    static int         box_loop; // 1: per-box flux loop shaped like the reference

    ~AmrLevelAdv();

    void read_params(const ParmParse &pp_root);
    void define(int max_grid_size, int ng = 2);
    void define(const std::vector<bx_box> &boxes, int ng);
    void initData(const ParmParse &pp_root);

    double estTimeStep(double cur_time);
    double computeNewDt(double cur_time, double stop_time);
    double advance(double time, double dt, int iteration, int ncycle);
    void   advance_explicit(double time, double dt);

    // main.cpp loop (graph replay when no stop time is set)
    int run(double stop_time, int max_step, double &time);

    void download(std::vector<double> &U) const;
    void upload(const std::vector<double> &U);
    void writePlotFile(const std::string &path, double time, int step) const;
    void checkPoint(const std::string &path, double time, int step, double dt_next) const;
    void restart(const std::string &path, double &time, int &step);

    const BoxLayout &layout() const { return m_layout; }
    const Geometry  &geom() const { return m_geom; }
    bx_ctx          *ctx() const { return m_ctx; }

  private:
    Geometry        m_geom;
    bx_params       m_parm{};
    int             m_bc_lo[3] = { 0, 0, 0 }, m_bc_hi[3] = { 0, 0, 0 };
    BoxLayout       m_layout;
    bx_ctx         *m_ctx    = nullptr;
    double         *m_S1     = nullptr; // Sborder
    double         *m_S2     = nullptr; // Sborder2
    double         *m_flux   = nullptr; // per-box flux FAB (box_loop mode)
    void           *m_stream = nullptr;
};

// Exec/*/Prob.cpp initdata: fills the valid cells of U from init.* keys
void initdata(std::vector<double> &U, const BoxLayout &lay, const Geometry &geom,
              const bx_params &parm, const ParmParse &pp_root);

} // namespace bamrex
