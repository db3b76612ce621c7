// bamrex_run: the reference's main (Source/main.cpp:18-118) for the explicit
// hot path on one level.
//   bamrex_run <inputs> [key=value ...]
// Reads max_step / stop_time / amr.* / geometry.* / adv.* / prob.* /
// limiting.* / bc.* / init.*, runs the time loop and prints "Run time = ".
// amr.plot_file, when present, receives the final state (raw level dump).
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "AmrLevelAdv.H"

using namespace bamrex;

int main(int argc, char **argv)
{
    if (argc < 2)
    {
        std::fprintf(stderr, "usage: %s <inputs> [key=value ...]\n", argv[0]);
        return 1;
    }
    try
    {
        const auto t0 = std::chrono::steady_clock::now();
        ParmParse::Initialize(argc, argv);
        ParmParse pp;
        int       max_step  = -1;
        double    strt_time = 0.0, stop_time = -1.0;
        pp.query("max_step", max_step);
        pp.query("strt_time", strt_time);
        pp.query("stop_time", stop_time);
        if (strt_time < 0.0)
            throw std::runtime_error("MUST SPECIFY a non-negative strt_time");
        if (max_step < 0 && stop_time < 0.0)
            throw std::runtime_error(
                "Exiting because neither max_step nor stop_time is non-negative.");

        AmrLevelAdv level;
        level.read_params(pp);
        ParmParse pa("amr");
        int       mgs = level.geom().dim == 3 ? 32 : 128;
        pa.query("max_grid_size", mgs);
        level.define(mgs);
        level.initData(pp);

        double    time  = strt_time;
        const int steps = level.run(stop_time, max_step, time);

        std::string plot;
        if (pa.query("plot_file", plot))
            level.writePlotFile(plot, time, steps);
        const double run_time
            = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                  .count();
        std::printf("Steps = %d  time = %.17g\n", steps, time);
        std::printf("Run time = %g\n", run_time);
    }
    catch (const std::exception &e)
    {
        std::fprintf(stderr, "bamrex_run: %s\n", e.what());
        return 2;
    }
    return 0;
}
