#!/usr/bin/env python3
"""Benchmark of the explicit hot path (one JSON line on stdout).

Workload (the configuration the headline metric is quoted on): 2D
four-quadrant Riemann problem, one uniform level of 4096^2 cells,
max_grid_size 128 (1024 boxes), gamma 1.4, MUSCL-Hancock + HLLC with the
minbee limiter, CFL 0.8, FP64.  A step is one full split step of the level:
device-side CFL dt, x sweep, ghost fill, y sweep (with the fused CFL
reduction), ghost fill -- replayed as a captured hipGraph.

value      = cell-updates/s over K timed steps (state resident in HBM).
roofline   = the slowest sweep kernel: algorithmic bytes 2*NC*8 = 64 B per
             valid cell per sweep / its average duration (HIP events around
             the kernel on its stream), against 8 TB/s.
cpu_baseline = the CPU oracle (same arithmetic) on a bounded sample of the
             same workload, timed on this host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MHM-HLLC FP64 cell-updates/s; IMEX solve s/step learned-precond vs plain MLMG"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ncell", type=int, default=4096)
    ap.add_argument("--mgs", type=int, default=128)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-repeats", type=int, default=3)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--tuning", default="", help="block_x,block_t,seg_k,grid_cap (skips autotune)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain copy kernel of this size (counter calibration)")
    ap.add_argument("--no-imex", action="store_true", help="skip the IMEX pressure-solve leg")
    ap.add_argument("--imex-only", action="store_true", help="run only the IMEX pressure-solve leg")
    ap.add_argument("--imex-profile-only", action="store_true",
                    help="short IMEX workload for kernel profiling: --imex-steps steps, each "
                    "stage solved by every method (untrained network), no harvest or training")
    ap.add_argument("--imex-n", type=int, default=1024)
    ap.add_argument("--imex-advection", default="muscl_rusanov", choices=["rusanov", "muscl_rusanov"])
    ap.add_argument("--imex-tableau", default="SSP222",
                    choices=["SP111", "SSP222", "SASSP322", "SASSP332", "SSP433"])
    ap.add_argument("--imex-ke", default="ex", choices=["ex", "split", "conservative_kinetic", "fi"],
                    help="imex.ke_method: ex (PEEK), split (PESK), conservative_kinetic, or fi "
                         "(not linearised, 2 Picard iterations per stage)")
    ap.add_argument("--imex-harvest-systems", type=int, default=104,
                    help="pressure systems harvested for training (4 more are held out)")
    ap.add_argument("--imex-steps", type=int, default=4, help="time steps (one solve per stage)")
    ap.add_argument("--imex-train-steps", type=int, default=2000)
    ap.add_argument("--imex-batch", type=int, default=32)
    ap.add_argument("--imex-crop", type=int, default=64)
    ap.add_argument("--no-vortex", action="store_true", help="skip the 3D isentropic vortex leg")
    ap.add_argument("--vortex-n", type=int, default=512)
    ap.add_argument("--vortex-mgs", type=int, default=128)
    ap.add_argument("--vortex-steps", type=int, default=50)
    ap.add_argument("--no-uneven", action="store_true", help="skip the batched uneven-box leg")
    ap.add_argument("--no-amr", action="store_true", help="skip the two-level AMR leg")
    ap.add_argument("--amr-only", action="store_true", help="run only the two-level AMR leg")
    ap.add_argument("--amr-n", type=int, default=2048)
    ap.add_argument("--amr-steps", type=int, default=20)
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (each GPU box has one "
              "MI355X); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from bamrex_amd import _lib, problems
    from bamrex_amd.level import Level, make_params

    lib = _lib.lib()
    name = ctypes.create_string_buffer(256)
    _lib.check(lib.bx_device_name(name, 256), "bx_device_name")
    log("device: " + name.value.decode())
    if args.imex_only:
        print(json.dumps({"imex": imex_leg(args)}), flush=True)
        return
    if args.imex_profile_only:
        print(json.dumps({"imex_profile": imex_profile_leg(args)}), flush=True)
        return
    if args.amr_only:
        print(json.dumps({"amr": amr_leg(args)}), flush=True)
        return

    log("building the %d^2 quadrant level (mgs %d)" % (args.ncell, args.mgs))
    s = problems.quadrant_setup(ncell=args.ncell, max_grid_size=args.mgs)
    lay = s["layout"]
    ncells = lay.ncells()
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    stream = ctypes.c_void_p()
    _lib.check(lib.bx_stream_create(ctypes.byref(stream)), "stream")
    lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    lev.set_state(s["U"])
    lev.fill_boundary()
    dt0 = lev.estTimeStep(s["dx"], s["cfl"])
    tuning = None
    if args.tuning:
        bx_, bt_, k_, cap_ = [int(v) for v in args.tuning.split(",")]
        lev.set_tuning(block_x=bx_, block_t=bt_, seg_k=k_, grid_cap=cap_)
        tuning = lev.get_tuning()
    elif not args.no_autotune:
        log("autotuning launch shapes")
        lev.autotune(s["dx"], dt0)
        tuning = lev.get_tuning()
        log("tuning: %s" % tuning)
        lev.set_state(s["U"])

    if args.copy_mib:
        from bamrex_amd.level import DeviceBuffer
        nb = args.copy_mib << 20
        a, b = DeviceBuffer(nb), DeviceBuffer(nb)
        for _ in range(3):
            _lib.check(lib.bx_copy_kernel(b.ptr, a.ptr, nb, stream), "copy")
        _lib.check(lib.bx_stream_sync(stream), "sync")
        log("copy kernel: %d MiB read + %d MiB written per launch, 3 launches" % (
            args.copy_mib, args.copy_mib))
        a.free()
        b.free()

    # ------------------------------------------------- per-kernel timing
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(lib.bx_event_create(ctypes.byref(ev0)), "event")
    _lib.check(lib.bx_event_create(ctypes.byref(ev1)), "event")
    lev.fill_boundary()
    lev.set_tmp(s["U"])
    kern_ms = []
    for d in range(2):
        # untimed launches first: the GPU's clock dips and recovers over the
        # first ~10 ms of back-to-back sweeps, and the timed steps run after it
        for _ in range(args.kernel_reps):
            lev.sweep(d, dt0, s["dx"])
        _lib.check(lib.bx_stream_sync(stream), "sync")
        _lib.check(lib.bx_event_record(ev0, stream), "record")
        for _ in range(args.kernel_reps):
            lev.sweep(d, dt0, s["dx"])
        _lib.check(lib.bx_event_record(ev1, stream), "record")
        ms = ctypes.c_float()
        _lib.check(lib.bx_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)), "elapsed")
        kern_ms.append(ms.value / args.kernel_reps)
    log("sweep kernel avg ms: x %.3f  y %.3f" % tuple(kern_ms))

    # ----------------------------------------------------- timed steps
    lev.set_state(s["U"])
    if args.warmup > 0:
        lev.run(s["dx"], s["cfl"], max_step=args.warmup, use_graph=True)
    lev.set_state(s["U"])
    _lib.check(lib.bx_stream_sync(stream), "sync")
    _lib.check(lib.bx_event_record(ev0, stream), "record")
    t0 = time.perf_counter()
    t_end, nsteps = lev.run(s["dx"], s["cfl"], max_step=args.steps, use_graph=True)
    _lib.check(lib.bx_event_record(ev1, stream), "record")
    _lib.check(lib.bx_stream_sync(stream), "sync")
    wall = time.perf_counter() - t0
    ms = ctypes.c_float()
    _lib.check(lib.bx_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)), "elapsed")
    elapsed = max(wall, ms.value / 1000.0)
    value = nsteps * ncells / elapsed
    log("%d steps in %.3f ms (t=%.6g): %.3e cell-updates/s" % (nsteps, elapsed * 1e3, t_end, value))

    # the timed run (tuned launch shapes, graph replay) against the same
    # steps under the library's default shapes, launched from the host loop:
    # the same bits in every valid cell, or the benchmark fails
    check = state_check(s, gp, lev, nsteps, stream)
    log("state check: %s" % check)
    if not check["equal"]:
        log("FAILED: the timed run's final state differs from the default-shape run")
        sys.exit(3)

    dom = max(range(2), key=lambda d: kern_ms[d])
    algo_bytes = ncells * 2 * 4 * 8  # read + write NC doubles per valid cell
    achieved = algo_bytes / (kern_ms[dom] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic("k_sweep_x" if dom == 0 else "k_sweep_t")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": traffic_src,
                "kernel": "k_sweep_%s" % ("x" if dom == 0 else "t (y)"),
                "kernel_ms": round(kern_ms[dom], 4),
                "sweep_ms": [round(x, 4) for x in kern_ms],
                "algo_bytes_per_launch": algo_bytes}

    lev.close()
    uneven = None
    if not args.no_uneven:
        uneven = uneven_leg(args)
    vortex = None
    if not args.no_vortex:
        vortex = vortex_leg(args)
    amr_out = None
    if not args.no_amr:
        amr_out = amr_leg(args)
    imex = None
    if not args.no_imex:
        imex = imex_leg(args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_repeats)

    out = {
        "metric": METRIC, "value": value, "unit": "cell-updates/s", "n_gpus": 1,
        "steps": nsteps, "warmup": args.warmup, "ms_per_step": elapsed * 1e3 / nsteps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "quadrant-2d-%dx%d-mgs%d-mhm-hllc-minbee-cfl0.8" % (
            args.ncell, args.ncell, args.mgs), "cells": ncells, "boxes": len(lay.boxes),
            "ghost_cells": lay.ng, "graph": True, "tuning": tuning},
        "state_check": check,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "uneven": uneven,
        "vortex3d": vortex,
        "amr": amr_out,
        "imex": imex,
    }
    print(json.dumps(out), flush=True)


def state_check(s, gp, lev, nsteps, stream):
    """sha256 of the valid cells of the timed level's final state and of a
    second level run for the same steps with the default launch shape and
    no graph (host-driven steps)."""
    import hashlib
    import numpy as np
    from bamrex_amd.level import Level
    lay = s["layout"]
    mask = np.zeros(lay.total, dtype=bool)
    for b in range(len(lay.boxes)):
        lay.valid(mask, b)[...] = True
    got = lev.get_state()[mask]
    ref_lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    try:
        ref_lev.set_state(s["U"])
        _, n_ref = ref_lev.run(s["dx"], s["cfl"], max_step=nsteps, use_graph=False)
        ref = ref_lev.get_state()[mask]
        default_tuning = ref_lev.get_tuning()
    finally:
        ref_lev.close()
    h_got = hashlib.sha256(got.tobytes()).hexdigest()
    h_ref = hashlib.sha256(ref.tobytes()).hexdigest()
    return {"steps": nsteps, "sha256_timed": h_got, "sha256_default_shape_host_loop": h_ref,
            "default_tuning": default_tuning, "equal": bool(h_got == h_ref and n_ref == nsteps)}


def measured_traffic(kernel):
    """HBM bytes per launch of a sweep kernel on this workload, as measured
    by rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (calibrated against a copy
    kernel) and recorded in the profile summary; None when not recorded."""
    path = os.path.join(ROOT, "profiles", "sweep_quadrant4096.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                fields = dict(f.split("=") for f in line.split(":", 1)[1].split())
                if kernel in fields:
                    return int(fields[kernel]), "profiles/sweep_quadrant4096.md"
    except OSError:
        pass
    return None, None


def _time_sweeps(lib, lev, dt, dx, stream, reps):
    """Average ms per launch of each directional sweep (HIP events on the
    sweeps' stream)."""
    from bamrex_amd import _lib
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(lib.bx_event_create(ctypes.byref(ev0)), "event")
    _lib.check(lib.bx_event_create(ctypes.byref(ev1)), "event")
    out = []
    for d in range(lev.layout.dim):
        for _ in range(reps):  # untimed: past the clock transient of a busy start
            lev.sweep(d, dt, dx)
        _lib.check(lib.bx_stream_sync(stream), "sync")
        _lib.check(lib.bx_event_record(ev0, stream), "record")
        for _ in range(reps):
            lev.sweep(d, dt, dx)
        _lib.check(lib.bx_event_record(ev1, stream), "record")
        ms = ctypes.c_float()
        _lib.check(lib.bx_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)), "elapsed")
        out.append(ms.value / reps)
    lib.bx_event_destroy(ev0)
    lib.bx_event_destroy(ev1)
    return out


def uneven_leg(args):
    """Batched uneven AMR level: 3000 2D boxes of 8..96 cells per side
    (seed 20261202), 2 ghost cells, smooth background plus Mach 1.5-3 shock
    fronts; one MHM-HLLC sweep per direction over all boxes in one persistent
    launch, autotuned, and graph-replayed steps."""
    from bamrex_amd import _lib, problems
    from bamrex_amd.level import Level, make_params
    lib = _lib.lib()
    s = problems.uneven_setup()
    lay = s["layout"]
    ncells = lay.ncells()
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    stream = ctypes.c_void_p()
    _lib.check(lib.bx_stream_create(ctypes.byref(stream)), "stream")
    lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    lev.set_state(s["U"])
    lev.set_tmp(s["U"])
    dt = lev.estTimeStep(s["dx"], s["cfl"])
    lev.autotune(s["dx"], dt)
    tuning = lev.get_tuning()
    lev.set_state(s["U"])
    lev.set_tmp(s["U"])
    ms = _time_sweeps(lib, lev, dt, s["dx"], stream, args.kernel_reps)
    lev.set_state(s["U"])
    lev.run(s["dx"], s["cfl"], max_step=3, use_graph=True)
    lev.set_state(s["U"])
    t0 = time.perf_counter()
    _, nsteps = lev.run(s["dx"], s["cfl"], max_step=20, use_graph=True)
    _lib.check(lib.bx_stream_sync(stream), "sync")
    el = time.perf_counter() - t0
    lev.close()
    lib.bx_stream_destroy(stream)
    bytes_per_sweep = ncells * 2 * 4 * 8
    out = {"workload": "uneven-2d-3000-boxes-8..96-seed20261202-mhm-hllc", "boxes": len(lay.boxes),
           "cells": ncells, "sweep_ms": [round(x, 4) for x in ms],
           "sweep_cell_updates_per_s": [ncells / (x * 1e-3) for x in ms],
           "sweep_hbm_frac": [round(bytes_per_sweep / (x * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for x in ms],
           "graph_steps": nsteps, "step_cell_updates_per_s": nsteps * ncells / el, "tuning": tuning}
    log("uneven: %d boxes, %d cells: sweeps %s ms, %.3e cell-updates/s per step" % (
        len(lay.boxes), ncells, out["sweep_ms"], out["step_cell_updates_per_s"]))
    return out


def vortex_leg(args):
    """3D isentropic vortex, periodic n^3 on [-5,5]^3, lambda 5 advected at
    (1,1,0), MHM-HLLC with superbee limiting each conserved variable on
    itself, CFL 0.9, graph-replayed steps; L1 error of the final state against
    the analytic solution at the final time (Exec/IsentropicVortex/Prob.cpp)."""
    import numpy as np
    from bamrex_amd import _lib, problems
    from bamrex_amd.level import Level, make_params
    lib = _lib.lib()
    n = args.vortex_n
    log("vortex: building %d^3 level (mgs %d)" % (n, args.vortex_mgs))
    s = problems.vortex_setup(ncell=n, dim=3, max_grid_size=args.vortex_mgs)
    lay = s["layout"]
    ncells = lay.ncells()
    gp = make_params(3, s["adiabatic"], s["epsilon"], limiter=s["limiter"],
                     limiter_indices=s["limiter_indices"], limit_on_primv=s["limit_on_primv"])
    stream = ctypes.c_void_p()
    _lib.check(lib.bx_stream_create(ctypes.byref(stream)), "stream")
    lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], stream=stream)
    lev.set_state(s["U"])
    lev.fill_boundary()
    dt0 = lev.estTimeStep(s["dx"], s["cfl"])
    lev.set_tmp(s["U"])
    ms = _time_sweeps(lib, lev, dt0, s["dx"], stream, 3)
    lev.set_state(s["U"])
    _lib.check(lib.bx_stream_sync(stream), "sync")
    t0 = time.perf_counter()
    t_end, nsteps = lev.run(s["dx"], s["cfl"], max_step=args.vortex_steps, use_graph=True)
    _lib.check(lib.bx_stream_sync(stream), "sync")
    el = time.perf_counter() - t0
    U = lev.get_state()
    lev.close()
    lib.bx_stream_destroy(stream)
    ex = lay.new_array()
    problems.isentropic_vortex(lay, ex, s["prob_lo"], s["prob_hi"], s["dx"], time=t_end)
    l1 = np.zeros(lay.nc)
    for b in range(len(lay.boxes)):
        d = np.abs(lay.valid(U, b) - lay.valid(ex, b))
        l1 += d.reshape(lay.nc, -1).sum(axis=1)
    l1 = l1 / ncells          # mean absolute error per cell (volume-weighted L1 / volume)
    out = {"workload": "isentropic-vortex-3d-%d^3-mgs%d-mhm-hllc-superbee-cfl0.9" % (n, args.vortex_mgs),
           "cells": ncells, "steps": nsteps, "t_end": t_end,
           "cell_updates_per_s": nsteps * ncells / el, "ms_per_step": el * 1e3 / nsteps,
           "sweep_ms": [round(x, 4) for x in ms],
           "sweep_hbm_frac": [round(ncells * 2 * 5 * 8 / (x * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for x in ms],
           "l1_error": [float(x) for x in l1]}
    log("vortex: %d steps to t=%.4g, %.3e cell-updates/s, L1 rho %.3e" % (
        nsteps, t_end, out["cell_updates_per_s"], l1[0]))
    if not args.no_cpu:
        out["l1_cpu_vs_gpu"] = vortex_l1_cpu_gpu(t_end)
    return out


def vortex_l1_cpu_gpu(t_end, n=128, mgs=32):
    """The isentropic vortex at n^3 on the GPU and in the CPU oracle (glibc
    pow, as the reference), both to the stop time t_end: their L1 errors
    against the analytic solution side by side (Exec/IsentropicVortex/
    Prob.cpp:49-57, 67-135)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from bamrex_amd import problems
    from bamrex_amd.level import Level, make_params
    s = problems.vortex_setup(ncell=n, dim=3, max_grid_size=mgs)
    lay = s["layout"]
    kw = dict(limiter=s["limiter"], limiter_indices=s["limiter_indices"],
              limit_on_primv=s["limit_on_primv"])
    gp = make_params(3, s["adiabatic"], s["epsilon"], **kw)
    op = pyoracle.params(3, s["adiabatic"], s["epsilon"], **kw)
    lev = Level(gp, lay, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"])
    try:
        lev.set_state(s["U"])
        t_gpu, n_gpu = lev.run(s["dx"], s["cfl"], stop_time=t_end, use_graph=False)
        U_gpu = lev.get_state()
    finally:
        lev.close()
    U_cpu = s["U"].copy()
    t_cpu, n_cpu = pyoracle.Oracle(bxpow=False).run(op, lay, U_cpu, s["domain_lo"], s["domain_hi"],
                                                     s["bc"], s["bc"], s["dx"], s["cfl"],
                                                     stop_time=t_end)
    ex = lay.new_array()
    problems.isentropic_vortex(lay, ex, s["prob_lo"], s["prob_hi"], s["dx"], time=t_gpu)
    l1g = np.zeros(lay.nc)
    l1c = np.zeros(lay.nc)
    dmax = 0.0
    for b in range(len(lay.boxes)):
        vg, vc, ve = lay.valid(U_gpu, b), lay.valid(U_cpu, b), lay.valid(ex, b)
        l1g += np.abs(vg - ve).reshape(lay.nc, -1).sum(axis=1)
        l1c += np.abs(vc - ve).reshape(lay.nc, -1).sum(axis=1)
        dmax = max(dmax, float(np.max(np.abs(vg - vc))))
    ncell = lay.ncells()
    out = {"cells": ncell, "max_grid_size": mgs, "t_end": t_gpu, "steps_gpu": n_gpu,
           "steps_cpu": n_cpu, "l1_gpu": [float(x) for x in l1g / ncell],
           "l1_cpu": [float(x) for x in l1c / ncell], "max_abs_diff_gpu_cpu": dmax}
    log("vortex %d^3 to t=%.4g: L1 rho GPU %.6e CPU %.6e (%d / %d steps)" % (
        n, t_gpu, out["l1_gpu"][0], out["l1_cpu"][0], n_gpu, n_cpu))
    return out


def amr_leg(args):
    """Two levels, ratio 2: the quadrant problem on an n^2 coarse level (mgs
    128) with a fixed fine level over the central half of the domain in
    each direction (an n^2 fine region, boxes of 128), MHM-HLLC, minbee,
    CFL 0.8.  One coarse step = coarse advance + two fine advances of dt/2
    with coarse-fine interpolation, flux registers, reflux, avgDown and the
    two-level dt.  The run replays one captured coarse step (hipGraph): dt,
    the time and the coarse-fine time weights are computed on the device, so
    a step needs no host synchronisation."""
    from bamrex_amd import _lib, amr, problems
    from bamrex_amd.level import make_params
    lib = _lib.lib()
    n = args.amr_n
    log("amr: building %d^2 coarse + %d^2 fine levels" % (n, n))

    def quadrant(ncell, max_grid_size):
        return problems.quadrant_setup(ncell=ncell, max_grid_size=max_grid_size)

    s = amr.refine_setup(quadrant, n, 128, (n // 4, n // 4), (3 * n // 4 - 1, 3 * n // 4 - 1),
                         fine_max_grid_size=128)
    gp = make_params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])
    h = amr.AmrHierarchy(gp, s["layout"], s["fine_layout"], s["domain_lo"], s["domain_hi"],
                         s["bc"], s["bc"])
    # launch shapes tuned per level on the initial state, which is then reloaded
    h.set_state(s["U"], s["Uf"])
    h.avg_down()
    dt = h.est_dt(s["dx"], 0.8)
    h.coarse.autotune(s["dx"], dt)
    h.fine.autotune(tuple(x / 2 for x in s["dx"]), dt / 2)
    h.set_state(s["U"], s["Uf"])
    del s["U"], s["Uf"]
    h.run(s["dx"], 0.8, max_step=2)
    _lib.check(lib.bx_stream_sync(h.stream), "sync")
    t0 = time.perf_counter()
    t_end, nsteps = h.run(s["dx"], 0.8, max_step=args.amr_steps, start_time=0.0)
    el = time.perf_counter() - t0
    info = h.info()
    h.close()
    nc, nf = s["layout"].ncells(), s["fine_layout"].ncells()
    out = {"workload": "quadrant-2d-%d^2-mgs128+fine-%d^2-ratio2-mhm-hllc-minbee-cfl0.8" % (n, n),
           "coarse_cells": nc, "fine_cells": nf, "coarse_steps": nsteps,
           "ms_per_coarse_step": round(el * 1e3 / nsteps, 4),
           "cell_updates_per_s": (nc + 2 * nf) * nsteps / el,
           "cf_segments": info["cf_segments"], "register_faces": info["register_faces"],
           "graph_replayed": True}
    log("amr: %d coarse steps, %.3f ms/step, %.3e cell-updates/s" % (
        nsteps, out["ms_per_coarse_step"], out["cell_updates_per_s"]))
    return out


def imex_leg(args):
    """IMEX PEEK low-Mach Gresho vortex (M = 0.01, gamma 1.4, n^2 periodic):
    the pressure Helmholtz solve of successive steps to 1e-10 relative
    residual, plain multigrid (FP64 V-cycles, the reference's MLMG::solve
    semantics, kernels launched one by one) against FP64 flexible GMRES
    whose preconditioner (FP32 V-cycle, replayed as a captured graph) is
    followed by a learned correction of the V-cycle's residual ("corrector")
    or has the learned network as its fine-level smoother ("smoother"), and
    against the same FGMRES with the plain V-cycle.  The network is trained on the GPU
    first, on crops of residual fields harvested from plain multigrid solves
    of the preceding steps.  Every solve starts from the previous pressure
    and is checked by an FP64 residual computed independently."""
    import numpy as np
    from bamrex_amd import _lib, imex
    n = args.imex_n
    log("imex: Gresho %d^2, harvesting %d systems (+4 held out)" % (n, args.imex_harvest_systems))
    wl = imex.GreshoWorkload(n, advection=args.imex_advection, tableau=args.imex_tableau,
                             ke_method=args.imex_ke)

    def set_coeffs(sy):
        # the stage's arrays in place: they hold until the next system
        solver.set_coeffs(sy["alpha"], sy["beta"], sy["a"], sy["bX"], sy["bY"], gamma=sy["gamma_c"],
                          cX=sy["cX"], cY=sy["cY"], copy=False)
    nst = wl.rk.tab.n_imp * (wl.rk.picard_max_iter if wl.rk.fi else 1)   # solves per step
    solver = imex.HelmholtzSolver(n, n, wl.dx)
    nsys, n_hold = args.imex_harvest_systems, 4
    nf_max = 4 * nsys
    R = imex.DeviceArray((nf_max, n, n))
    Rh = imex.DeviceArray((4 * n_hold, n, n))
    nf = nh = 0
    held = []   # per held-out system: host coefficients and its field range in Rh
    # one pressure system per stage; each plain-multigrid solve, one V-cycle
    # at a time, gives its residual fields (A e = r, e the remaining error)
    for k in range(nsys + n_hold):
        sy = wl.next_system()
        set_coeffs(sy)
        db, dx_ = imex.dev(sy["rhs"]), imex.dev(sy["p_prev"])
        if k < nsys:
            got, st = imex.harvest_residuals(solver, db, dx_, R, nf, nf_max - nf)
            nf += got
        else:
            got, st = imex.harvest_residuals(solver, db, dx_, Rh, nh, 4 * n_hold - nh)
            held.append(dict(alpha=sy["alpha"], beta=sy["beta"], gamma=sy["gamma_c"],
                             a=sy["a"].download(), bX=sy["bX"].download(),
                             bY=sy["bY"].download(),
                             cX=sy["cX"].download() if sy["cX"] is not None else None,
                             cY=sy["cY"].download() if sy["cY"] is not None else None,
                             fields=(nh, nh + got)))
            nh += got
        wl.advance(sy, dx_)
        db.free()
        dx_.free()
    while wl.in_step:        # finish a step left open by the harvest
        sy = wl.next_system()
        set_coeffs(sy)
        db, dx_ = imex.dev(sy["rhs"]), imex.dev(sy["p_prev"])
        solver.solve(db, dx_, "mg", rtol=1e-10, maxiter=200)
        wl.advance(sy, dx_)
    log("imex: %d residual fields harvested from %d systems, %d held out from %d" % (
        nf, nsys, nh, n_hold))

    def held_out_reduction():
        """mean ||r - A M r|| / ||r|| of one preconditioner application on
        the held-out systems' fields, each with its own system's operator"""
        r, e, q = (imex.DeviceArray((n, n)) for _ in range(3))
        out = []
        rh = Rh.download()
        for h in held:
            solver.set_coeffs(h["alpha"], h["beta"], h["a"], h["bX"], h["bY"], gamma=h["gamma"],
                              cX=h["cX"], cY=h["cY"])
            for i in range(*h["fields"]):
                r.upload(rh[i])
                solver.precond(r, e)
                solver.residual(r, e, q)
                out.append(float(np.linalg.norm(q.download()) / np.linalg.norm(rh[i])))
        for d in (r, e, q):
            d.free()
        return float(np.mean(out))
    rho = {"V": held_out_reduction()}
    net = imex.LearnedSmoother(seed=1)
    solver.attach(net, omega=1.0, mode="corrector")
    rho["V+untrained"] = held_out_reduction()
    t0 = time.perf_counter()
    hist = net.train(R, nf, steps=args.imex_train_steps, crop=args.imex_crop, batch=args.imex_batch,
                     lr=1e-3, seed=11)
    train_s = time.perf_counter() - t0
    rho["V+trained"] = held_out_reduction()
    log("imex: trained %d steps x %d crops in %.2f s, loss %.4g -> %.4g; held-out one-application "
        "residual reduction V %.4g, V+untrained %.4g, V+trained %.4g" % (
            args.imex_train_steps, args.imex_batch, train_s, float(np.mean(hist[:10])),
            float(np.mean(hist[-10:])), rho["V"], rho["V+untrained"], rho["V+trained"]))
    # one network application (bx_cnn_infer: sigma partial sums + the fused
    # kernel) on a held-out field, 30 back-to-back calls between HIP events
    # on the default stream the calls use
    r_in, e_out = imex.DeviceArray((n, n)), imex.DeviceArray((n, n))
    r_in.upload(Rh.download()[0])
    L = imex.lib()
    for _ in range(5):
        _lib.check(L.bx_cnn_infer(net.c, r_in.ptr, e_out.ptr, n, n, 1.0, None), "bx_cnn_infer")
    ev0, ev1 = ctypes.c_void_p(), ctypes.c_void_p()
    _lib.check(L.bx_event_create(ctypes.byref(ev0)), "event")
    _lib.check(L.bx_event_create(ctypes.byref(ev1)), "event")
    _lib.check(L.bx_event_record(ev0, None), "record")
    for _ in range(30):
        _lib.check(L.bx_cnn_infer(net.c, r_in.ptr, e_out.ptr, n, n, 1.0, None), "bx_cnn_infer")
    _lib.check(L.bx_event_record(ev1, None), "record")
    ms = ctypes.c_float()
    _lib.check(L.bx_event_elapsed_ms(ev0, ev1, ctypes.byref(ms)), "elapsed")
    L.bx_event_destroy(ev0)
    L.bx_event_destroy(ev1)
    infer_us = ms.value * 1e3 / 30
    log("imex: network application %.1f us at %d^2" % (infer_us, n))
    r_in.free()
    e_out.free()
    R.free()
    Rh.free()
    # (label, solve method, network mode, graph-replayed preconditioner)
    methods = [("mg", "mg", None, False),
               ("mg_fused", "mg_fused", None, False),
               ("fgmres_mg", "fgmres_mg", None, True),
               ("fgmres_cnn", "fgmres_cnn", "corrector", True),
               ("fgmres_cnn_smoother", "fgmres_cnn", "smoother", True)]
    res = {m[0]: {"seconds": [], "iters": [], "rel_resid": [], "fell_back": 0} for m in methods}
    r = imex.DeviceArray((n, n))
    # explicit work per step (host clock around synchronised calls):
    # advective dt; per stage the stage states, EK pressure, advection and
    # pressure system ("assemble"), the coefficient upload and multigrid
    # hierarchy build ("set_coeffs"), the stage fluxes and the step's update
    # ("finish"), summed over the step's stages
    stage_ms = {"advective_dt": [], "assemble": [], "set_coeffs": [], "finish": []}
    cold = {}
    for k in range(args.imex_steps):
        t0 = time.perf_counter()
        dt = wl.step_dt()
        stage_ms["advective_dt"].append((time.perf_counter() - t0) * 1e3)
        acc = {"assemble": 0.0, "set_coeffs": 0.0, "finish": 0.0}
        done = False
        while not done:
            t1 = time.perf_counter()
            sy = wl.next_system(dt)
            imex.lib().bx_stream_sync(None)
            t2 = time.perf_counter()
            set_coeffs(sy)
            t3 = time.perf_counter()
            acc["assemble"] += t2 - t1
            acc["set_coeffs"] += t3 - t2
            db = imex.dev(sy["rhs"])
            bn = float(np.linalg.norm(sy["rhs"].download()))
            p_mg = None
            for label, meth, mode, graph in methods:
                if mode is not None:
                    solver.attach(net, omega=1.0, mode=mode)
                solver.set_graph(graph)
                dx_ = imex.dev(sy["p_prev"])
                # untimed solve first: graph capture after a configuration change
                solver.solve(db, dx_, meth, rtol=1e-10, maxiter=200)
                dx_.copy_from(sy["p_prev"])
                st = solver.solve(db, dx_, meth, rtol=1e-10, maxiter=200)
                solver.residual(db, dx_, r)
                rel = float(np.linalg.norm(r.download())) / bn
                res[label]["seconds"].append(st["seconds"])
                res[label]["iters"].append(st["iters"])
                res[label]["rel_resid"].append(rel)
                res[label]["fell_back"] += int(st["fell_back"])
                if label == "mg":
                    p_mg = dx_
            if k == args.imex_steps - 1 and sy["stage"] == wl.rk.tab.n_imp - 1:
                # cold start (zero initial guess) on the last system: where a
                # better preconditioner saves iterations
                for label, meth, mode, graph in methods:
                    if mode is not None:
                        solver.attach(net, omega=1.0, mode=mode)
                    solver.set_graph(graph)
                    dx_ = imex.dev(np.zeros((n, n)))
                    solver.solve(db, dx_, meth, rtol=1e-10, maxiter=200)
                    dx_.upload(np.zeros((n, n)))
                    st = solver.solve(db, dx_, meth, rtol=1e-10, maxiter=200)
                    cold[label] = {"s_per_solve": st["seconds"], "iters": st["iters"],
                                   "fell_back": int(st["fell_back"])}
            t0 = time.perf_counter()
            done = wl.advance(sy, p_mg)
            acc["finish"] += time.perf_counter() - t0
            log("imex step %d stage %d: " % (k, sy["stage"]) + "  ".join(
                "%s %.3f ms/%d it" % (m[0], res[m[0]]["seconds"][-1] * 1e3, res[m[0]]["iters"][-1])
                for m in methods))
        for key, v in acc.items():
            stage_ms[key].append(v * 1e3)
    out = {"workload": "gresho-%s-%s-%s-%dx%d-M0.01-periodic" % (
        {"ex": "peek", "split": "pesk", "conservative_kinetic": "pe-ck", "fi": "fi"}[args.imex_ke],
        args.imex_tableau.lower(), args.imex_advection, n, n),
           "tableau": args.imex_tableau, "solves_per_step": nst,
           "rtol": 1e-10,
           "harvest_systems": nsys, "harvest_fields": nf, "held_out_systems": n_hold,
           "held_out_fields": nh, "held_out_residual_reduction": rho,
           "train_steps": args.imex_train_steps, "train_batch": args.imex_batch,
           "train_crop": args.imex_crop, "train_crops_total": args.imex_train_steps * args.imex_batch,
           "train_seconds": round(train_s, 3), "network_application_us": round(infer_us, 2),
           "train_loss_first_last": [float(np.mean(hist[:10])), float(np.mean(hist[-10:]))]}
    for label, meth, mode, graph in methods:
        out[label] = {"s_per_solve": float(np.mean(res[label]["seconds"])),
                      "iters_mean": float(np.mean(res[label]["iters"])),
                      "max_rel_resid": float(np.max(res[label]["rel_resid"])),
                      "fell_back": res[label]["fell_back"], "network_mode": mode,
                      "graph_preconditioner": graph}
    out["cold_start"] = cold
    out["stage_ms_median"] = {k: round(float(np.median(v)), 4) for k, v in stage_ms.items()}
    # a whole PEEK step per solve method: the explicit work of its stages
    # plus one warm-started solve per stage
    stage_total = sum(out["stage_ms_median"].values())
    out["peek_step_ms"] = {m[0]: round(stage_total + nst * out[m[0]]["s_per_solve"] * 1e3, 4)
                           for m in methods}
    # against plain multigrid (kernels launched one by one, as MLMG), against
    # the faster of the two plain multigrid variants, and against the same
    # FGMRES without the network; warm starts (the previous pressure) and
    # the cold start (zero guess) on the last system
    best_mg = min(out["mg"]["s_per_solve"], out["mg_fused"]["s_per_solve"])
    out["speedup_cnn_vs_mg"] = out["mg"]["s_per_solve"] / out["fgmres_cnn"]["s_per_solve"]
    out["speedup_cnn_vs_best_mg"] = best_mg / out["fgmres_cnn"]["s_per_solve"]
    out["speedup_cnn_vs_fgmres_mg"] = (out["fgmres_mg"]["s_per_solve"]
                                       / out["fgmres_cnn"]["s_per_solve"])
    out["speedup_fgmres_mg_vs_mg"] = out["mg"]["s_per_solve"] / out["fgmres_mg"]["s_per_solve"]
    out["speedup_mg_fused_vs_mg"] = out["mg"]["s_per_solve"] / out["mg_fused"]["s_per_solve"]
    if cold:
        cbest = min(cold["mg"]["s_per_solve"], cold["mg_fused"]["s_per_solve"])
        out["cold_speedup_cnn_vs_best_mg"] = cbest / cold["fgmres_cnn"]["s_per_solve"]
        out["cold_speedup_cnn_vs_fgmres_mg"] = (cold["fgmres_mg"]["s_per_solve"]
                                                / cold["fgmres_cnn"]["s_per_solve"])
    solver.close()
    net.close()
    wl.free()
    return out


def imex_profile_leg(args):
    """A short IMEX workload whose kernels a profiler can attribute: the
    Gresho vortex at --imex-n, --imex-steps steps of --imex-tableau with the
    explicit stage work on the GPU, and every stage's pressure system solved
    to 1e-10 by each method (mg, mg_fused, fgmres_mg, fgmres_cnn with an
    untrained network; FGMRES preconditioners graph-replayed), warm-started
    from the previous pressure; the state advances with the mg solution."""
    from bamrex_amd import imex
    n = args.imex_n
    wl = imex.GreshoWorkload(n, advection=args.imex_advection, tableau=args.imex_tableau,
                             ke_method=args.imex_ke)
    solver = imex.HelmholtzSolver(n, n, wl.dx)
    net = imex.LearnedSmoother(seed=1)
    methods = ["mg", "mg_fused", "fgmres_mg", "fgmres_cnn"]
    iters = {m: [] for m in methods}
    for k in range(args.imex_steps):
        dt = wl.step_dt()
        done = False
        while not done:
            sy = wl.next_system(dt)
            solver.set_coeffs(sy["alpha"], sy["beta"], sy["a"], sy["bX"], sy["bY"],
                              gamma=sy["gamma_c"], cX=sy["cX"], cY=sy["cY"], copy=False)
            db = imex.dev(sy["rhs"])
            p_mg = None
            for meth in methods:
                if meth == "fgmres_cnn":
                    solver.attach(net, omega=1.0, mode="corrector")
                dx_ = imex.dev(sy["p_prev"])
                st = solver.solve(db, dx_, meth, rtol=1e-10, maxiter=200)
                iters[meth].append(st["iters"])
                if meth == "mg":
                    p_mg = dx_
            done = wl.advance(sy, p_mg)
        log("imex profile: step %d done" % k)
    solver.close()
    net.close()
    wl.free()
    return {"workload": "gresho %d^2 %s, %d steps" % (n, args.imex_tableau, args.imex_steps),
            "iters": iters}


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, avail


def _median_rate(fn, cells, repeats):
    """Median over `repeats` runs of cells / seconds, fn() returning the
    seconds of one run."""
    rates = sorted(cells / fn() for _ in range(repeats))
    return rates[len(rates) // 2]


def cpu_baseline(repeats=3):
    """The oracle (the same arithmetic as the kernels, glibc pow as the
    reference calls it) on bounded samples of the C2, C3 and C4 workloads,
    with all host threads and with one, median of `repeats` runs each (as
    Exec/GreshoVortex/profiling/profile.sh repeats its runs).  The samples:
    C2 one step of the 4096^2 quadrant level (one thread: of the 1024^2
    level); C3 two steps of the isentropic vortex at 128^3 in boxes of 32^3
    (one thread: 64^3);
    C4 one step of the 3000-box uneven level (one thread: 600 of its boxes)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from bamrex_amd import problems
    model, avail = host_cpu()
    threads_all = int(os.environ.get("OMP_NUM_THREADS", avail))
    o = pyoracle.Oracle(bxpow=False)

    def run_rate(s, p, steps, threads):
        o.threads(threads)
        lay = s["layout"]

        def go():
            # the oracle times its own step loop (dt, sweeps, ghost fills),
            # without building the level or copying the state in and out
            U = s["U"].copy()
            o.run(p, lay, U, s["domain_lo"], s["domain_hi"], s["bc"], s["bc"], s["dx"], s["cfl"],
                  max_step=steps)
            return o.last_run_seconds()
        return _median_rate(go, lay.ncells() * steps, repeats)

    def quad(n):
        s = problems.quadrant_setup(ncell=n, max_grid_size=128)
        return s, pyoracle.params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])

    def vortex(n):
        # boxes of 32^3, so the oracle's box loop has work for every thread
        s = problems.vortex_setup(ncell=n, dim=3, max_grid_size=32)
        return s, pyoracle.params(3, s["adiabatic"], s["epsilon"], limiter=s["limiter"],
                                  limiter_indices=s["limiter_indices"],
                                  limit_on_primv=s["limit_on_primv"])

    def uneven(nb):
        s = problems.uneven_setup(nboxes=nb)
        return s, pyoracle.params(2, s["adiabatic"], s["epsilon"], limiter=s["limiter"])

    out = {}
    t0 = time.perf_counter()
    for key, (big, small, steps) in {"C2": (lambda: quad(4096), lambda: quad(1024), 1),
                                      "C3": (lambda: vortex(128), lambda: vortex(64), 2),
                                      "C4": (lambda: uneven(3000), lambda: uneven(600), 1)}.items():
        s, p = big()
        all_rate = run_rate(s, p, steps, threads_all)
        cells_big = s["layout"].ncells()
        del s
        s, p = small()
        one_rate = run_rate(s, p, steps, 1)
        cells_small = s["layout"].ncells()
        del s
        out[key] = {"all_threads": all_rate, "one_thread": one_rate,
                    "cells_all": cells_big, "cells_one": cells_small, "steps": steps}
        log("cpu oracle %s: %.3e cell-updates/s on %d threads, %.3e on 1" % (
            key, all_rate, threads_all, one_rate))
    o.threads(threads_all)
    log("cpu baseline: %.1f s" % (time.perf_counter() - t0))
    return {"value": out["C2"]["all_threads"], "unit": "cell-updates/s", "cores": threads_all,
            "kind": "port",
            "sample": "1 step of the 4096^2 quadrant level (1024 boxes of 128^2), median of %d "
                      "runs on %d threads" % (repeats, threads_all),
            "one_thread": {"value": out["C2"]["one_thread"],
                           "sample": "1 step of the 1024^2 quadrant level, median of %d runs" % repeats},
            "cpu_model": model, "cpus_available": avail, "repeats": repeats,
            "configs": out}


if __name__ == "__main__":
    main()
