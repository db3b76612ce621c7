// bamrex_run: the reference's main (Source/main.cpp:18-118) for the explicit
// hot path on one level.
//   bamrex_run <inputs> [key=value ...]
// Reads max_step / stop_time / amr.* / geometry.* / adv.* / prob.* /
// limiting.* / bc.* / init.*, runs the time loop and prints "Run time = ".
// Plotfiles as the reference's Amr writes them: amr.plot_file (root name,
// default "plt") + the step in 5 digits, at the start, every amr.plot_int
// steps and at the end; amr.plot_files_output = 0 turns them off.
#include <algorithm>
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

        std::string plot_root = "plt";
        int         plot_int = -1, plot_on = 1;
        pa.query("plot_file", plot_root);
        pa.query("plot_int", plot_int);
        pa.query("plot_files_output", plot_on);
        const bool plots = plot_on != 0 && plot_int > 0;
        auto       plot_name = [&](int step) {
            char b[16];
            std::snprintf(b, sizeof b, "%05d", step);
            return plot_root + b;
        };

        double time  = strt_time;
        int    steps = 0, last_plot = -1;
        if (plots)
        {
            level.writePlotFile(plot_name(0), time, 0);
            last_plot = 0;
        }
        while ((steps < max_step || max_step < 0) && (time < stop_time || stop_time < 0.0))
        {
            // run up to the next plot step (or to the end)
            int chunk = max_step < 0 ? -1 : max_step - steps;
            if (plots)
            {
                const int to_plot = plot_int - steps % plot_int;
                chunk             = chunk < 0 ? to_plot : std::min(chunk, to_plot);
            }
            const int done = level.run(stop_time, chunk, time);
            steps += done;
            if (plots && steps % plot_int == 0 && steps != last_plot)
            {
                level.writePlotFile(plot_name(steps), time, steps);
                last_plot = steps;
            }
            if (done == 0)
                break;
        }
        if (plots && last_plot != steps)
            level.writePlotFile(plot_name(steps), time, steps);
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
