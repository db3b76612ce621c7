// bamrex_run: the reference's main (Source/main.cpp:18-118) for the explicit
// hot path on one level.
//   bamrex_run <inputs> [key=value ...]
// Reads max_step / stop_time / amr.* / geometry.* / adv.* / prob.* /
// limiting.* / bc.* / init.*, runs the time loop and prints "Run time = ".
// Plotfiles as the reference's Amr writes them: amr.plot_file (root name,
// default "plt") + the step in 5 digits, at the start, every amr.plot_int
// steps and at the end; amr.plot_files_output = 0 turns them off.
// Checkpoints likewise: amr.check_file (default "chk") + the step, every
// amr.check_int steps (and at the start of a new run when check_int > 0), and
// the final one of Source/main.cpp:97-100 unless
// amr.checkpoint_files_output = 0.  amr.restart = <checkpoint directory>
// continues a run from a checkpoint (AmrLevelAdv::restart,
// Source/AmrLevelAdv.cpp:95-106) instead of from the initial data.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "AmrLevelAdv.H"
#include "ImexLevel.H"

using namespace bamrex;

// adv.num_method = imex (AmrLevelAdv_advance.cpp:131-157): the IMEX
// Runge-Kutta level on the GPU, with plotfiles as for the explicit methods.
// Checkpoints are written for the explicit methods only (the IMEX state
// also needs its pressure field); the reference's IMEX inputs turn them off.
static int run_imex(int max_step, double strt_time, double stop_time)
{
    ParmParse   pa("amr");
    std::string plot_root = "plt";
    int         plot_int = -1, plot_on = 1;
    pa.query("plot_file", plot_root);
    pa.query("plot_int", plot_int);
    pa.query("plot_files_output", plot_on);
    const bool plots = plot_on != 0 && plot_int > 0;
    auto       name  = [](const std::string &root, int step) {
        char b[16];
        std::snprintf(b, sizeof b, "%05d", step);
        return root + b;
    };
    ImexLevel level;
    level.read_params(stop_time);
    level.define();
    level.initData();
    double time  = strt_time;
    int    steps = 0, last_plot = -1;
    if (plots)
    {
        level.writePlotFile(name(plot_root, 0), time, 0);
        last_plot = 0;
    }
    while ((steps < max_step || max_step < 0) && (time < stop_time || stop_time < 0.0))
    {
        int chunk = max_step < 0 ? -1 : max_step - steps;
        if (plots)
        {
            const int to_next = plot_int - steps % plot_int;
            chunk             = chunk < 0 ? to_next : std::min(chunk, to_next);
        }
        const int done = level.run(stop_time, chunk, time);
        steps += done;
        if (plots && steps % plot_int == 0 && steps != last_plot)
        {
            level.writePlotFile(name(plot_root, steps), time, steps);
            last_plot = steps;
        }
        if (done == 0)
            break;
    }
    if (plots && last_plot != steps)
        level.writePlotFile(name(plot_root, steps), time, steps);
    std::printf("Pressure solves = %d  iterations = %ld\n", level.solves(), level.solve_iters());
    std::printf("Steps = %d  time = %.17g\n", steps, time);
    return steps;
}

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

        std::string method;
        ParmParse("adv").get("num_method", method);
        if (method == "imex")
        {
            run_imex(max_step, strt_time, stop_time);
            const double run_time
                = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            std::printf("Run time = %g\n", run_time);
            return 0;
        }

        AmrLevelAdv level;
        level.read_params(pp);
        ParmParse pa("amr");
        int       mgs = level.geom().dim == 3 ? 32 : 128;
        pa.query("max_grid_size", mgs);

        std::string plot_root = "plt", check_root = "chk", restart_dir;
        int         plot_int = -1, plot_on = 1, check_int = -1, check_on = 1;
        pa.query("plot_file", plot_root);
        pa.query("plot_int", plot_int);
        pa.query("plot_files_output", plot_on);
        pa.query("check_file", check_root);
        pa.query("check_int", check_int);
        pa.query("checkpoint_files_output", check_on);
        pa.query("restart", restart_dir);
        const bool plots  = plot_on != 0 && plot_int > 0;
        const bool checks = check_on != 0 && check_int > 0;
        auto       name   = [](const std::string &root, int step) {
            char b[16];
            std::snprintf(b, sizeof b, "%05d", step);
            return root + b;
        };

        double time  = strt_time;
        int    steps = 0, last_plot = -1, last_check = -1;
        if (!restart_dir.empty())
        {
            level.restart(restart_dir, time, steps);
            last_plot = last_check = steps;
            std::printf("Restarted from %s at step %d, time %.17g\n", restart_dir.c_str(),
                        steps, time);
        }
        else
        {
            level.define(mgs);
            level.initData(pp);
            if (plots)
            {
                level.writePlotFile(name(plot_root, 0), time, 0);
                last_plot = 0;
            }
            if (checks)
            {
                level.checkPoint(name(check_root, 0), time, 0,
                                 level.computeNewDt(time, stop_time));
                last_check = 0;
            }
        }
        while ((steps < max_step || max_step < 0) && (time < stop_time || stop_time < 0.0))
        {
            // run up to the next plot or checkpoint step (or to the end)
            int chunk = max_step < 0 ? -1 : max_step - steps;
            for (int every : { plots ? plot_int : 0, checks ? check_int : 0 })
                if (every > 0)
                {
                    const int to_next = every - steps % every;
                    chunk             = chunk < 0 ? to_next : std::min(chunk, to_next);
                }
            const int done = level.run(stop_time, chunk, time);
            steps += done;
            if (plots && steps % plot_int == 0 && steps != last_plot)
            {
                level.writePlotFile(name(plot_root, steps), time, steps);
                last_plot = steps;
            }
            if (checks && steps % check_int == 0 && steps != last_check)
            {
                level.checkPoint(name(check_root, steps), time, steps,
                                 level.computeNewDt(time, stop_time));
                last_check = steps;
            }
            if (done == 0)
                break;
        }
        // the final checkpoint and plotfile (Source/main.cpp:97-105)
        if (check_on != 0 && last_check != steps)
            level.checkPoint(name(check_root, steps), time, steps,
                             level.computeNewDt(time, stop_time));
        if (plots && last_plot != steps)
            level.writePlotFile(name(plot_root, steps), time, steps);
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
