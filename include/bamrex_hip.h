/*
 * bamrex_hip.h -- C ABI of the MI355X-native BAMReX explicit hot path.
 *
 * Plain C types only (pointers, sizes, ints, doubles).  Device buffers are
 * owned by the caller; the library owns only its context and the device
 * copies of the box list / copy plans it builds.  Every call that takes a
 * `stream` (a hipStream_t, NULL = default stream) is stream-ordered and
 * asynchronous unless its comment says otherwise.  A context is not
 * thread-safe: use one context per stream.  Status: 0 = BX_OK, otherwise a
 * BX_ERR_* code and bx_last_error() describes it.  No exception crosses the ABI.
 *
 * Data layout (the amrex::Array4 / FArrayBox layout): the FAB of a box with
 * valid region [lo, hi] and ng ghost cells covers grow([lo,hi], ng) in the
 * first `dim` directions; element (i,j,k,n) sits at
 *   offset + (i-lo0+ng) + (j-lo1+ng)*jstride + (k-lo2+ng)*kstride + n*nstride
 * with jstride = nx+2ng, kstride = jstride*(ny+2ng), nstride = FAB volume,
 * counted in doubles from the level base pointer passed to each call.
 * Components are conserved variables (rho, mom_x..mom_{dim-1}, E), dim+2 of them.
 */
#ifndef BAMREX_HIP_H
#define BAMREX_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BX_OK 0
#define BX_ERR_INVALID 1     /* bad argument (amrex::Abort / AMREX_ASSERT in the reference) */
#define BX_ERR_HIP 2         /* HIP runtime error */
#define BX_ERR_UNSUPPORTED 3 /* valid input the reference also rejects (RuntimeError) */
#define BX_ERR_STATE 4       /* call order violated (e.g. no level set) */

/* adv.num_method (Source/NumericalMethods.H:10-44); the explicit methods
 * AmrLevelAdv::advance_explicit dispatches (AmrLevelAdv_advance.cpp:311-326):
 * compute_MUSCL_hancock_flux, compute_HLLC_flux, compute_SLIC_flux,
 * compute_force_flux, compute_LF_flux (Fluxes.H, Force.cpp, SLIC.cpp). */
#define BX_METHOD_MUSCL_HANCOCK 0
#define BX_METHOD_HLLC 1
#define BX_METHOD_SLIC 2
#define BX_METHOD_FORCE 3
#define BX_METHOD_LAX_FRIEDRICHS 4

/* limiting.limiter (Source/Fluxes/Limiters.H:13) */
#define BX_LIMITER_MINBEE 0
#define BX_LIMITER_VAN_LEER 1
#define BX_LIMITER_VAN_ALBADA 2
#define BX_LIMITER_SUPERBEE 3

/* bc.{x,y,z} (Source/BoundaryConditions/BCTypes.H; mapping BCs.cpp:215-397) */
#define BX_BC_TRANSMISSIVE 0 /* amrex foextrap */
#define BX_BC_PERIODIC 1     /* amrex int_dir */
#define BX_BC_REFLECTIVE 2   /* reflect_even, normal momentum reflect_odd */

typedef struct bx_ctx bx_ctx;

/* struct Parm (Source/System/Parm.H:14-61) restricted to the hot path.
 * limiter_indices[n] / limit_on_primv[n]: variable that limits component n
 * (FluidSystem.cpp:27-71; default: conserved energy, index dim+1, for all n). */
typedef struct
{
    int    dim;
    int    num_method;
    double adiabatic;
    double epsilon;
    int    limiter;
    int    limiter_indices[5];
    int    limit_on_primv[5];
} bx_params;

typedef struct
{
    int     lo[3]; /* valid region, inclusive, global cell index (unused dims 0) */
    int     hi[3];
    int64_t offset; /* doubles from the level base pointer to the FAB start */
} bx_box;

typedef struct
{
    int domain_lo[3];
    int domain_hi[3];
    int bc_lo[3]; /* BX_BC_* per direction */
    int bc_hi[3];
    int ng;       /* ghost cells of every FAB, >= 2 */
} bx_geometry;

/* ------------------------------------------------------------------ setup */

/* Parm set-up: AmrLevelAdv::read_params + FluidSystem::ReadParameters
 * (Source/AmrLevelAdv.cpp:837-909, Source/System/FluidSystem.cpp:10-71). */
int  bx_create(const bx_params *params, bx_ctx **out);
void bx_destroy(bx_ctx *ctx);

/* Define the level: box list (BoxArray + FAB offsets), domain and boundary
 * conditions.  Builds the device box list, the persistent-kernel work tables
 * and the ghost-copy plan (replaces MFIter set-up and the FillBoundary /
 * PhysBCFunct copy schedule, Source/BoundaryConditions/BCs.cpp:429-442). */
int bx_set_level(bx_ctx *ctx, const bx_geometry *geom, const bx_box *boxes,
                 int nboxes);

/* Packed FAB offsets for a box list (each FAB 256-byte aligned);
 * total_doubles receives the level buffer size.  Host only. */
int bx_pack_offsets(int dim, int ng, const bx_box *boxes, int nboxes,
                    int64_t *offsets_out, int64_t *total_doubles);

/* ------------------------------------------------------- hot-path calls */

/* One directional sweep over every box of the level: MUSCL-Hancock
 * reconstruction + half step + HLLC flux + conservative update for
 * num_method muscl_hancock (compute_MUSCL_hancock_flux, Fluxes.H:237-257 /
 * MUSCL_Hancock.cpp:15-52, then AmrLevelAdv_advance.cpp:330-341), or the
 * first-order HLLC flux + update for num_method hllc (Fluxes.H:206-227).
 * Reads U_in (valid + 2 ghost layers along dir), writes the valid cells of
 * U_out. */
int bx_sweep(bx_ctx *ctx, int dir, double dt, const double dx[3],
             const double *U_in, double *U_out, void *stream);

/* Per-box drop-in of the reference flux function:
 * compute_MUSCL_hancock_flux(dir, time, bx, flux, consv_values, dx, dt, ...)
 * (Fluxes.H:237-257) / compute_HLLC_flux (Fluxes.H:206-227) for box ibox of
 * the level.  flux_base + flux_offset addresses an Array4 on
 * surroundingNodes(bx, dir) (no ghosts, nodal in dir) with dim+2 components. */
int bx_flux_box(bx_ctx *ctx, int ibox, int dir, double dt, const double dx[3],
                const double *U, double *flux_base, int64_t flux_offset,
                void *stream);

/* Split conservative update of box ibox from the fluxes bx_flux_box wrote
 * (AmrLevelAdv_advance.cpp:330-341). */
int bx_update_box(bx_ctx *ctx, int ibox, int dir, double dt, const double dx[3],
                  const double *U_in, const double *flux_base,
                  int64_t flux_offset, double *U_out, void *stream);

/* bamrexBCData::fill_consv_boundary (BCs.cpp:429-442): all ghost cells of U
 * from the valid cells of the level (periodic images included) and the
 * physical boundary conditions. */
int bx_fill_boundary(bx_ctx *ctx, double *U, void *stream);

/* max_wave_speed (Source/System/Euler/Euler.cpp:46-74): max over valid cells
 * of sqrt(gamma p / (rho eps)) + |u|.  Synchronises the stream. */
int bx_max_wave_speed(bx_ctx *ctx, const double *U, double *out,
                      void *stream);

/* estTimeStep for the explicit methods (AmrLevelAdv.cpp:395-477, 468) plus the
 * stop-time clamp of computeNewDt (AmrLevelAdv.cpp:600-611); stop_time < 0
 * disables the clamp.  Synchronises the stream. */
int bx_est_dt(bx_ctx *ctx, const double *U, const double dx[3], double cfl,
              double time, double stop_time, double *dt_out, void *stream);

/* AmrLevelAdv::advance / advance_explicit for the explicit methods
 * (AmrLevelAdv_advance.cpp:21-259, 261-439) on a single level: sweep x, y, z
 * in that order with the full dt, refilling ghosts after each sweep.  U holds
 * the state with filled ghosts on entry and the new state (ghosts filled) on
 * return; U_tmp is a second level buffer of the same layout. */
int bx_advance(bx_ctx *ctx, double dt, const double dx[3], double *U,
               double *U_tmp, void *stream);

/* The main.cpp time loop (Source/main.cpp:76-94) for a single level:
 *   while (step < max_step && time < stop_time) { dt = clamp(cfl * est);
 *   advance; time += dt }  (max_step < 0 or stop_time < 0 disables a bound).
 * U holds the initial state (ghosts need not be filled).  The dt reduction is
 * fused into the last sweep of each step and the dt control runs on the
 * device.  With stop_time < 0 and use_graph != 0 each step is captured once as
 * a hipGraph and replayed, with no host synchronisation inside the loop.
 * On return *result_in_tmp tells whether the final state is in U_tmp (odd
 * number of sweeps in graph mode) and time / steps are filled.  Synchronous. */
int bx_run(bx_ctx *ctx, const double dx[3], double cfl, double start_time,
           double stop_time, int max_step, double *U, double *U_tmp,
           int use_graph, void *stream, double *time_out, int *steps_out,
           int *result_in_tmp);

/* ------------------------------------------------------------- tuning */

typedef struct
{
    int block_x;      /* threads per block of the x sweep (64..256, multiple of 64) */
    int block_t;      /* threads per block of the y/z sweeps */
    int seg_k;        /* cells per lane along y/z */
    int grid_cap;     /* max resident blocks per CU used for persistent grids */
} bx_tuning;

int bx_get_tuning(const bx_ctx *ctx, bx_tuning *out);
int bx_set_tuning(bx_ctx *ctx, const bx_tuning *t);

/* Times every candidate launch shape of every sweep direction on (U, U_tmp)
 * and keeps the fastest (block size x segment length x grid cap).  The state
 * buffers are overwritten.  Synchronous; writes the timings (ms per sweep)
 * of the winning shapes to best_ms[dir] when best_ms is not NULL. */
int bx_autotune(bx_ctx *ctx, const double dx[3], double dt, double *U,
                double *U_tmp, void *stream, double *best_ms);

/* -------------------------------------------- device memory / timing helpers
 * Thin wrappers so that hosts without a HIP toolchain (ctypes, other FFIs)
 * can drive the library. */
int         bx_device_count(int *n);
int         bx_malloc(void **ptr, size_t bytes);
int         bx_free(void *ptr);
int         bx_memcpy_htod(void *dst, const void *src, size_t bytes);
int         bx_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int         bx_memcpy_dtod(void *dst, const void *src, size_t bytes);
int         bx_memset(void *dst, int value, size_t bytes);
int         bx_stream_create(void **stream);
int         bx_stream_destroy(void *stream);
int         bx_stream_sync(void *stream);
int         bx_event_create(void **ev);
int         bx_event_destroy(void *ev);
int         bx_event_record(void *ev, void *stream);
int         bx_event_elapsed_ms(void *start, void *stop, float *ms);
int         bx_device_name(char *buf, int len);
/* Streaming device copy kernel (reads and writes `bytes` once); used to
 * calibrate HBM byte counters against a kernel of known traffic. */
int         bx_copy_kernel(void *dst, const void *src, size_t bytes, void *stream);
const char *bx_last_error(void);

/* ==========================================================================
 * Plotfiles
 * ========================================================================== */
/* One level of a plotfile: a host copy of a packed level (bx_pack_offsets
 * layout, ng ghost cells per FAB; only valid cells are written). */
typedef struct
{
    const bx_box *boxes;
    int           nbox;
    int           ng;
    const double *U;
    int           domain_lo[3];
    int           domain_hi[3];
    double        dx[3];
    int           level_steps;
} bx_plot_level;
/* AmrLevel::writePlotFile (Source/AmrLevelAdv.cpp:116-136): an AMReX
 * plotfile directory (HyperCLaw-V1.1: Header, Level_<l>/Cell_H,
 * Level_<l>/Cell_D_00000) of nlev levels (refinement ratio ref_ratio).
 * Variables: density, mom_x[, mom_y[, mom_z]], energy
 * (FluidSystem::VariableNames); derived != 0 adds pressure and u_x[, u_y[,
 * u_z]] (FluidSystem::SetupDeriveList).  Host code, synchronous. */
int bx_write_plotfile(const char *dir, int dim, int nlev,
                      const bx_plot_level *levels, int ref_ratio,
                      const double prob_lo[3], const double prob_hi[3],
                      double time, double gamma, double eps, int derived);

/* ==========================================================================
 * Checkpoint / restart
 * ==========================================================================
 * AmrLevelAdv::checkPoint / restart (Source/AmrLevelAdv.cpp:95-113, through
 * AMReX's AmrLevel::checkPoint / restart) and the final checkpoint of
 * Source/main.cpp:97-100: what a restart needs to continue bit for bit --
 * the time, each level's step count, dt, box list and state.  A directory
 * holds a text Header and per level Level_<l>/State, the level's packed
 * state (bx_pack_offsets layout, ghost cells included) as little-endian
 * doubles with an FNV-1a checksum in the header.  Host code; the state
 * pointers are device pointers; synchronous on `stream`. */
typedef struct
{
    const bx_box *boxes;
    int           nbox;
    int           ng;
    const double *U;            /* device: the packed level state */
    int           domain_lo[3];
    int           domain_hi[3];
    double        dx[3];
    int           level_steps;
    double        dt;
} bx_chk_level;
typedef struct
{
    int    dim;
    int    nlev;
    int    ref_ratio;
    double time;
    double prob_lo[3];
    double prob_hi[3];
} bx_chk_info;
typedef struct
{
    int      nbox;
    int      ng;
    int      level_steps;
    double   dt;
    int      domain_lo[3];
    int      domain_hi[3];
    double   dx[3];
    int64_t  ndoubles;
    uint64_t checksum;
} bx_chk_level_info;
int bx_write_checkpoint(const char *dir, int dim, int nlev,
                        const bx_chk_level *levels, int ref_ratio,
                        const double prob_lo[3], const double prob_hi[3],
                        double time, void *stream);
/* The header: run data and up to max_levels levels' data. */
int bx_read_checkpoint_info(const char *dir, bx_chk_info *info,
                            bx_chk_level_info *levels, int max_levels);
/* Level `level`'s box list (nbox must be the level's box count). */
int bx_read_checkpoint_boxes(const char *dir, int level, bx_box *boxes,
                             int nbox);
/* Level `level`'s state into the device buffer U of ndoubles doubles (the
 * level's packed size), checked against the header's checksum. */
int bx_read_checkpoint_level(const char *dir, int level, double *U,
                             int64_t ndoubles, void *stream);

/* ==========================================================================
 * Two AMR levels (refinement ratio 2, fine level subcycled)
 * ==========================================================================
 * What AMReX's Amr / FluxRegister / FillPatch machinery does around
 * AmrLevelAdv for a coarse level and one fine level: the coarse advance, two
 * fine advances of dt/2 whose coarse-fine ghost cells are interpolated from
 * the coarse old and new states (FillPatcherFill, AmrLevelAdv_advance.cpp:
 * 121-129; cell_cons_interp, AmrLevelAdv.cpp:171-175), flux registers
 * (_advance.cpp:61-81, 192-254), then reflux and avgDown (post_timestep,
 * AmrLevelAdv.cpp:628-652, 918-987).  Both levels are bx_ctx with the same
 * bx_params; the fine geometry is the coarse one refined by 2 and its boxes
 * are ratio-aligned (even lo, odd hi) in fine index space.  Device state
 * buffers belong to the caller, each level's packed layout as for bx_run. */
typedef struct bx_amr bx_amr;
/* Plans the coarse-fine ghost cells, register faces and covered cells. */
int  bx_amr_create(bx_ctx *coarse, bx_ctx *fine, bx_amr **out);
void bx_amr_destroy(bx_amr *a);
/* Totals over the level pairs. */
int  bx_amr_info(const bx_amr *a, int *n_cf_segs, int *n_register_faces,
                 int *n_reflux_cells);
/* AmrLevelAdv::avgDown (AmrLevelAdv.cpp:947-987): covered coarse cells
 * become the average of their fine cells; coarse ghosts refilled. */
int bx_amr_avg_down(bx_amr *a, const double *Uf, double *Uc, void *stream);
/* FillPatch of the fine level at `time` (t_old <= time <= t_new): fine-fine
 * and physical ghosts, then the coarse-fine face ghosts from the coarse
 * states at t_old and t_new (ghosts of both filled). */
int bx_amr_fill_fine(bx_amr *a, double *Uf, const double *Uc_old,
                     const double *Uc_new, double t_old, double t_new,
                     double time, void *stream);
/* One coarse step of dt at `time` (Amr::timeStep for levels 0 and 1 plus
 * post_timestep).  Uc, Uf: states with ghosts filled (in/out); Uc_tmp,
 * Uf_tmp: ping-pong buffers; Uc_old receives the coarse state at `time`. */
int bx_amr_advance(bx_amr *a, double dt, const double dx_coarse[3], double time,
                   double *Uc, double *Uc_tmp, double *Uc_old, double *Uf,
                   double *Uf_tmp, void *stream);
/* computeNewDt over both levels (AmrLevelAdv.cpp:589-618): min(dt_est(0),
 * 2 dt_est(1)) with the stop-time clamp; *dt is the coarse dt.  Synchronous. */
int bx_amr_est_dt(bx_amr *a, const double dx_coarse[3], double cfl, double time,
                  double stop_time, const double *Uc, const double *Uf,
                  double *dt, void *stream);
/* The main.cpp:76-94 loop on two levels: post_init avgDown, then steps until
 * max_step or stop_time.  *time in: start time, out: final time.
 * Synchronous. */
int bx_amr_run(bx_amr *a, const double dx_coarse[3], double cfl,
               double stop_time, int max_step, double *Uc, double *Uc_tmp,
               double *Uc_old, double *Uf, double *Uf_tmp, double *time,
               int *steps, void *stream);
/* Any number of levels (2..16), refinement ratio 2 between neighbours
 * (amr.max_level = nlev - 1, amr.ref_ratio = 2 ...): levels[l] is level l,
 * its geometry the level-0 one refined by 2^l.  Level l advances by
 * dt / 2^l, subcycled recursively as Amr::timeStep does; each level's
 * post_timestep refluxes and averages down from the level above.  Per
 * level l: U[l] (state, ghosts filled, in/out), U_tmp[l] (ping-pong),
 * U_old[l] (scratch for the old state; not used for the finest level).
 * dx0: level 0's cell sizes; dt: level 0's step.  The two-level calls above
 * are these with nlev = 2. */
int bx_amr_create_levels(bx_ctx *const *levels, int nlev, bx_amr **out);
int bx_amr_advance_levels(bx_amr *a, double dt, const double dx0[3],
                          double time, double *const *U, double *const *U_tmp,
                          double *const *U_old, void *stream);
int bx_amr_est_dt_levels(bx_amr *a, const double dx0[3], double cfl,
                         double time, double stop_time, const double *const *U,
                         double *dt, void *stream);
int bx_amr_run_levels(bx_amr *a, const double dx0[3], double cfl,
                      double stop_time, int max_step, double *const *U,
                      double *const *U_tmp, double *const *U_old, double *time,
                      int *steps, void *stream);
/* How bx_amr_run / bx_amr_run_levels launch a run of a fixed number of steps
 * (stop_time < 0, max_step > 0, at most 7 levels).  use_graph = 1 (the
 * default): one coarse step (the levels' maximum wave speeds, computeNewDt,
 * Amr::timeStep) is captured once as a hipGraph and replayed max_step times;
 * dt, the time, the flux-register coefficients and the coarse-fine time
 * weights are computed on the device with the host's expressions, so the
 * states and the final time are the same bits as with use_graph = 0, which
 * reads dt back and launches every kernel from the host each step.  Runs
 * with a stop time are always host-driven.  No counterpart in the reference
 * (its steps are host-driven); this only chooses the launch mode. */
int bx_amr_set_graph(bx_amr *a, int use_graph);

/* ============================================================================
 * Pressure Helmholtz solve of the IMEX schemes (PEEK / PESK) on one 2D
 * periodic level, and the learned preconditioner.
 *
 * Operator (MLALaplacianGrad, Source/LinearSolver/MLALapGrad_2D_K.H:9-33, set
 * up by setup_pressure_op, Source/IMEX/IMEX_O1.cpp:392-414):
 *   (A p)_ij = alpha a_ij p_ij
 *            - beta/dx^2 (bX_{i+1,j}(p_{i+1,j}-p_ij) - bX_ij(p_ij-p_{i-1,j}))
 *            - beta/dy^2 (bY_{i,j+1}(p_{i,j+1}-p_ij) - bY_ij(p_ij-p_{i,j-1}))
 *            + gamma (cX_ij (p_{i+1,j}-p_{i-1,j})/(2dx) + cY_ij (p_{i,j+1}-p_{i,j-1})/(2dy))
 * Cell arrays are nx*ny doubles, x fastest.  Face arrays: bX is (nx+1) x ny
 * (face i of row j at j*(nx+1)+i), bY is nx x (ny+1) (face j of column i at
 * j*nx+i); on a periodic level face nx equals face 0.  All arrays are device
 * pointers.  A solver object is not thread-safe.
 * ========================================================================== */
typedef struct bx_helm bx_helm;
typedef struct bx_cnn  bx_cnn;

/* bx_helm_solve methods */
#define BX_SOLVE_MG 0         /* MLMG::solve: V-cycles until converged (FP64) */
#define BX_SOLVE_FGMRES_MG 1  /* FGMRES, V-cycle preconditioner */
#define BX_SOLVE_FGMRES_CNN 2 /* FGMRES, V-cycle with the attached network */
#define BX_SOLVE_MG_FUSED 3   /* as BX_SOLVE_MG with the blocked-smoother V-cycle */

typedef struct
{
    int    iters;           /* outer iterations (V-cycles or Krylov steps) */
    int    converged;       /* final ||b - A x||_2 <= rtol ||b||_2 */
    int    fell_back;       /* FGMRES stalled and plain multigrid finished */
    int    precond_applies; /* V-cycles / preconditioner applications */
    double rel_resid;       /* final ||b - A x||_2 / ||b||_2 */
    double seconds;         /* wall time of the solve (host clock) */
} bx_solve_stats;

/* MLABecLaplacianGradT::define for one periodic level (MLALaplacianGrad.H:312-375).
 * nx, ny >= 8 and even (an odd periodic extent puts two cells of one red/black
 * colour side by side across the seam); non-periodic levels and odd extents
 * return BX_ERR_UNSUPPORTED.  Coarser multigrid levels halve both extents
 * while the halves stay even and >= 8. */
int  bx_helm_create(int nx, int ny, double dx, double dy, int periodic,
                    bx_helm **out);
void bx_helm_destroy(bx_helm *h);
/* setScalars / setACoeffs / setBCoeffs / setCCoeffs (MLALaplacianGrad.H:
 * 380-560); builds the coarse-level coefficients (averageDownCoeffs,
 * MLALaplacianGrad.H:930-1027).  cX = cY = NULL (or gamma = 0) drops the
 * gradient term.  Synchronous. */
int bx_helm_set_coeffs(bx_helm *h, double alpha, double beta, double gamma,
                       const double *a, const double *bX, const double *bY,
                       const double *cX, const double *cY, void *stream);
/* MLMG::setPreSmooth / setPostSmooth / setFinalSmooth (nu1, nu2, bottom red-
 * black sweeps), preconditioner precision and FGMRES restart length (<= 20). */
int bx_helm_set_options(bx_helm *h, int nu1, int nu2, int nu_bottom,
                        int precond_fp32, int restart);
/* y = A x (Fapply, MLALaplacianGrad.H:1095-1188). */
int bx_helm_apply(bx_helm *h, const double *x, double *y, void *stream);
/* r = b - A x (MLLinOp residual). */
int bx_helm_residual(bx_helm *h, const double *b, const double *x, double *r,
                     void *stream);
/* One red (redblack = 0) or black (1) Gauss-Seidel half sweep on the fine
 * level (Fsmooth, MLALaplacianGrad.H:1190-1300). */
int bx_helm_smooth(bx_helm *h, double *phi, const double *rhs, int redblack,
                   void *stream);
/* e = M r: one preconditioner application from a zero guess. */
int bx_helm_precond(bx_helm *h, const double *r, double *e, void *stream);
/* Learned preconditioner variants (bx_helm_attach_cnn mode):
 *   BX_CNN_SMOOTHER   e = w N(r); e += P V_coarse(R(r - A e)); e += w N(r - A e)
 *                     (the network replaces the fine-level smoothing)
 *   BX_CNN_CORRECTOR  e = V(r); e += w N(r - A e)
 *                     (the network corrects what the V-cycle leaves) */
#define BX_CNN_SMOOTHER 0
#define BX_CNN_CORRECTOR 1
/* Use `cnn` (NULL detaches) in the preconditioner, correction scaled by
 * omega, in the given mode. */
int bx_helm_attach_cnn(bx_helm *h, bx_cnn *cnn, double omega, int mode);
/* FGMRES replays its preconditioner as a captured hipGraph (default 1);
 * 0 launches every kernel of it directly. */
int bx_helm_set_graph(bx_helm *h, int use_graph);
/* The preconditioner's V-cycle smooths each level in one launch (red/black
 * half-sweeps blocked in LDS, residual and restriction fused; default 1) or
 * with one launch per colour (0), the structure of the plain solve.  Same
 * results bit for bit.  No reference counterpart. */
int bx_helm_set_fused(bx_helm *h, int fused);
/* MLMG::solve(phi, rhs, rtol, 0) (IMEX_O1.cpp:602-614) with the chosen method;
 * x holds the initial guess.  FGMRES that stalls (no 2x residual drop in 8
 * iterations) or hits maxiter hands its iterate to plain multigrid
 * (st->fell_back = 1).  Synchronous. */
int bx_helm_solve(bx_helm *h, const double *b, double *x, int method,
                  double rtol, int maxiter, bx_solve_stats *st, void *stream);
/* Describes the last failure of a bx_helm_* or bx_imex_* call. */
const char *bx_helm_last_error(void);

/* Explicit work of the IMEX PEEK-SP111 stage (linearised, enthalpy from the
 * previous pressure, kinetic energy of the explicit state) on one nx x ny
 * periodic 2D level.  Device arrays: U, Uex, U_new hold 4 components
 * (rho, m_x, m_y, E) of nx*ny doubles each, x fastest, no ghost cells;
 * p, enth, rhs, a hold nx*ny; bX is ny rows of nx+1 faces, bY ny+1 rows of
 * nx (the layout bx_helm_set_coeffs takes).  Asynchronous on `stream`
 * unless stated.
 *
 * pressure (Source/IMEX_K/euler_K.H:23-35): p = (gamma-1)(E - ke). */
int bx_imex_pressure(const double *U, int nx, int ny, double gamma, double eps,
                     double *p, void *stream);
/* Advective time step of the IMEX schemes (estTimeStep, AmrLevelAdv.cpp:
 * 431-436 and 466): cfl / max over cells of (|m_x|/dx + |m_y|/dy)/rho
 * (max_dx_scaled_speed, Source/System/Euler/Euler.cpp:107-129).
 * Synchronous; writes the host double *dt. */
int bx_imex_advective_dt(const double *U, int nx, int ny, double dx, double dy,
                         double cfl, double *dt, void *stream);
/* Acoustic time step of the IMEX schemes before imex.acoustic_timestep_end_time
 * (estTimeStep, AmrLevelAdv.cpp:424-430 and 466): ExplicitTimeStepDimSplit
 * (Source/System/FluidSystem.cpp:195-206), min(dx, dy) / max over cells of
 * sqrt(gamma p/(rho eps)) + |u| (max_wave_speed, Euler.cpp:47-74), times
 * cfl.  Both reductions skip NaN speeds as AMReX's ReduceOpMax does.
 * Synchronous; writes the host double *dt. */
int bx_imex_acoustic_dt(const double *U, int nx, int ny, double dx, double dy,
                        double gamma, double eps, double cfl, double *dt,
                        void *stream);
/* No counterpart in the reference: bx_imex_acoustic_dt for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  Same time step, bit for bit, but written to the device
 * double *dt_dev and asynchronous on `stream`; it calls no allocator.  Its
 * scratch (one partial maximum per block, at most 1024 blocks: 8 KiB) is a
 * file-scope device array that each GPU holds until the process ends, shared
 * by every call on that GPU: calls that overlap on one GPU race. */
int bx_imex_acoustic_dt_noalloc(const double *U, int nx, int ny, double dx,
                                double dy, double gamma, double eps, double cfl,
                                double *dt_dev, void *stream);
/* First half of compute_flux_imex_o1 (Source/IMEX/IMEX_O1.cpp:18-131):
 * Rusanov advective update U -> Uex (advance_rusanov_adv,
 * Source/IMEX/IMEX_Fluxes.cpp:10-118, or its MUSCL form), enthalpy p gamma/((gamma-1) rho_ex)
 * from the previous pressure p (compute_enthalpy, IMEX_O1.cpp:177-228), the
 * pressure equation's source rhs (pressure_source_vector, :416-452) and its
 * coefficients a = eps/(gamma-1), b = face-averaged enthalpy
 * (setup_pressure_op, :570-614).  The operator's scalars are alpha = 1,
 * beta = dt*dt. */
#define BX_IMEX_RUSANOV 0       /* imex.advection_flux = rusanov (1st order) */
#define BX_IMEX_MUSCL_RUSANOV 1 /* = muscl_rusanov (advance_MUSCL_rusanov_adv,
                                   Source/IMEX/IMEX_Fluxes.cpp:120-324; here
                                   with the SP111 local half step) */
int bx_imex_peek_assemble(const double *U, const double *p, int nx, int ny,
                          double dx, double dy, double dt, double gamma,
                          double eps, int advection_flux, double *Uex,
                          double *enth, double *rhs, double *a, double *bX,
                          double *bY, void *stream);
/* Second half of an SP111 step (IMEX_O1.cpp:132-175 with
 * update_momentum_flux / update_energy_flux, :485-566, then sum_fluxes with
 * b = 1 and conservative_update, Source/IMEX/IMEX_RK.cpp:93-95 and
 * Source/Fluxes/Update.H:100-130): the advective fluxes of U corrected with
 * the solved pressure p_new, then
 * U_new = U - dt(1/dx) dFx - dt(1/dy) dFy.  U_new must not alias U or Uex. */
int bx_imex_peek_finish(const double *U, const double *Uex, const double *enth,
                        const double *p_new, int nx, int ny, double dx, double dy,
                        double dt, double eps, int advection_flux, double *U_new,
                        void *stream);

/* Stages of an IMEX Runge-Kutta step (advance_imex_rk, Source/IMEX/
 * IMEX_RK.cpp:22-96, any IMEXButcherTableau of ButcherTableau.cpp:5-101).
 * Stage fluxes Fx, Fy hold 4 components of nx*ny faces each: Fx[k*n + c]
 * is the face between cells (i-1, j) and (i, j), Fy[k*n + c] between
 * (i, j-1) and (i, j), periodic.  The host loop per stage s:
 *   bx_imex_rk_combine     (s > 0; stage 0 uses U for both states)
 *   bx_imex_rk_pressure    (s > 0)
 *   bx_imex_stage_assemble, the pressure solve, bx_imex_stage_fluxes
 * and after the last stage bx_imex_rk_update. */
#define BX_IMEX_MAX_STAGES 4
/* imex.ke_method (IMEXSettings, Source/IMEX/IMEX_O1.cpp:232-298): the
 * kinetic energy of the explicit state (ex: the PEEK schemes) or split,
 * 0.5 eps (m_e . m_ex)/rho_e with the explicit stage state e (the PESK
 * schemes, whose operator has the velocity term gamma_c c . grad p with
 * gamma_c = -eps adt / 2 and c = m_e / rho_e). */
#define BX_IMEX_KE_EX 0
#define BX_IMEX_KE_SPLIT 1
/* imex.linearise = false (the FI schemes): the stage's first system takes
 * the kinetic energy of the implicit stage state, each solve is followed
 * by bx_imex_picard_update (imex.picard_max_iter solves per stage), and a
 * later stage starts from the explicit state's own pressure. */
#define BX_IMEX_KE_FI 2
/* imex.ke_method = conservative_kinetic: the kinetic energy of the explicit
 * stage state; a later stage starts from that state's own pressure. */
#define BX_IMEX_KE_CONSERVATIVE_KINETIC 3
/* imex.enthalpy_method (compute_enthalpy, Source/IMEX/IMEX_O1.cpp:177-229)
 * of the linearised stages: reuse_pressure, h = p gamma/((gamma-1) rho_ex);
 * conservative, h = gamma (E - ke)/rho_ex of the explicit stage state;
 * conservative_ex, the same of the advected state U_ex.  The conservative
 * methods replace p by the pressure of U_ex (the solve's initial guess). */
#define BX_IMEX_ENTHALPY_REUSE_PRESSURE 0
#define BX_IMEX_ENTHALPY_CONSERVATIVE 1
#define BX_IMEX_ENTHALPY_CONSERVATIVE_EX 2
/* The stage states (conservative_update into two outputs, Source/Fluxes/
 * Update.H:218-302): U_imp = U - sum_{s<nstage} dt_imp[s] div F_s and
 * U_exp likewise with dt_exp, dt_x[s] = A_x[stage][s] dt (host arrays of
 * nstage values); Fx, Fy: host arrays of nstage device pointers. */
int bx_imex_rk_combine(const double *U, const double *const *Fx,
                       const double *const *Fy, int nstage, const double *dt_imp,
                       const double *dt_exp, int nx, int ny, double dx, double dy,
                       double *U_imp, double *U_exp, void *stream);
/* The pressure of a later stage (compute_pressure, Source/IMEX/
 * IMEX_O1.cpp:300-391, linearised).  ke_method ex (EK): the previous
 * explicit state advanced over adt by its own advective fluxes (MUSCL
 * without the local step) supplies the momentum, U_exp_new the density
 * and energy: p = (gamma-1)(E - 0.5 eps |m_ex|^2 / rho).  ke_method split
 * (SK): p = (gamma-1)(E_new - 0.5 eps (m_old . m_new) / rho_old).
 * BX_IMEX_KE_FI and BX_IMEX_KE_CONSERVATIVE_KINETIC: the pressure of
 * U_exp_new, (gamma-1)(E - ke). */
int bx_imex_rk_pressure(const double *U_exp_new, const double *U_exp_old, int nx,
                        int ny, double dx, double dy, double adt, double gamma,
                        double eps, int advection_flux, int ke_method, double *p,
                        void *stream);
/* The end of one Picard iteration of an FI stage (picard_iterate,
 * Source/IMEX/IMEX_O1.cpp:702-789) after a solve gave p: the momentum
 * m_ex - (0.5 adt/dx/eps)(p(+) - p(-)) (update_momentum, :455-482), the
 * enthalpy from p, the kinetic energy of that state, and from them the
 * new source rhs and face coefficients bX, bY of the next solve (a is
 * unchanged). */
int bx_imex_picard_update(const double *Uex, const double *p, int nx, int ny,
                          double dx, double dy, double adt, double gamma, double eps,
                          double *enth, double *rhs, double *bX, double *bY,
                          void *stream);
/* compute_flux_imex_o1 up to the solve with separate stage states: the
 * advective fluxes of U_exp update U_imp; local_time_step selects the
 * MUSCL local half step (the reference takes it for SP111 only,
 * IMEX_O1.cpp:95-102).  Outputs as bx_imex_peek_assemble, whose call is
 * this one with U_imp = U_exp = U, local_time_step = 1, ke_method ex;
 * operator scalars alpha = 1, beta = adt*adt, gamma_c = 0 (ex) or
 * -0.5 eps adt (split).  cX, cY (nx*ny each, written for split, may be
 * null otherwise): the velocity coefficients of bx_helm_set_coeffs.  p is
 * written by the conservative enthalpy methods (ignored for FI). */
int bx_imex_stage_assemble(const double *U_imp, const double *U_exp, double *p,
                           int nx, int ny, double dx, double dy, double adt,
                           double gamma, double eps, int advection_flux,
                           int local_time_step, int ke_method, int enthalpy_method,
                           double *Uex, double *enth, double *rhs, double *a,
                           double *bX, double *bY, double *cX, double *cY,
                           void *stream);
/* The stage's fluxes after the solve (update_momentum_flux /
 * update_energy_flux, IMEX_O1.cpp:485-566): the advective fluxes of U_exp
 * corrected with the solved pressure p_new, into Fx, Fy. */
int bx_imex_stage_fluxes(const double *U_exp, const double *Uex, const double *enth,
                         const double *p_new, int nx, int ny, double dx, double dy,
                         double adt, double eps, int advection_flux,
                         int local_time_step, double *Fx, double *Fy, void *stream);
/* Step end (sum_fluxes, IMEX_RK.cpp:196-241, and conservative_update,
 * Update.H:100-130): F = sum_s b[s] F_s, U_new = U - dt(1/dx) dFx -
 * dt(1/dy) dFy.  b: host array of nstage weights. */
int bx_imex_rk_update(const double *U, const double *const *Fx,
                      const double *const *Fy, int nstage, const double *b, int nx,
                      int ny, double dx, double dy, double dt, double *U_new,
                      void *stream);

/* High-Mach stabilisation of a step of order above 1
 * (advance_imex_rk_stab, Source/IMEX/IMEX_RK.cpp:98-194; imex.stabilize,
 * imex.k1).  The shock indicator of the step's new state U
 * (compute_shock_indicator, :243-345): per cell c_min over the cell and
 * its face neighbours and the centred velocity divergence,
 * fmin(1, fmax(0, -(div u + k1 c_min)/(k1 c_min))), then its maximum over
 * the 3 x 3 neighbourhood into ind; work: nx*ny doubles of scratch.
 * Synchronous; writes the level maximum to the host double *ind_max. */
int bx_imex_shock_indicator(const double *U, int nx, int ny, double dx, double dy,
                            double gamma, double eps, double k1, double *work,
                            double *ind, double *ind_max, void *stream);
/* Where ind > 0: U = ind U_o1 + (1 - ind) U and p = ind p_o1 + (1 - ind) p,
 * with the first-order step (SP111, Rusanov) U_o1, p_o1. */
int bx_imex_stab_blend(const double *ind, const double *U_o1, const double *p_o1,
                       int nx, int ny, double *U, double *p, void *stream);

/* Learned smoother: bias-free 3x3 convolutions 1->16->16->16->1 with ReLU,
 * BF16 weights/activations, FP32 accumulation (MFMA in every layer).
 * Correction e = omega (sigma / D) N(r / sigma), sigma = rms(r), D = diag(A).
 * Parameters: W1[9][16], W2[9][16][16], W3[9][16][16], W4[9][16] in that
 * order ([tap][out][in]; tap = (dy+1)*3 + dx+1). */
#define BX_CNN_NPARAMS 4896
int  bx_cnn_create(unsigned seed, bx_cnn **out);
void bx_cnn_destroy(bx_cnn *c);
/* FP32 host arrays of BX_CNN_NPARAMS values.  set resets the Adam state. */
int bx_cnn_set_params(bx_cnn *c, const float *params, int n);
int bx_cnn_get_params(bx_cnn *c, float *params, int n);
/* The trained network to and from a file (bx_cnn_load/save of the proposed
 * boundary): magic "BXCNN001", the parameter count, an FNV-1a checksum, then
 * the FP32 parameters in the order above.  load checks all three
 * (BX_ERR_INVALID otherwise) and resets the Adam state. */
int bx_cnn_save(bx_cnn *c, const char *path);
int bx_cnn_load(bx_cnn *c, const char *path);
/* Registers the operator the network is trained for and applied with (its
 * diagonal and, for training, its coefficients; device pointers must stay
 * valid).  nx must be a multiple of 16. */
int bx_cnn_set_operator(bx_cnn *c, int nx, int ny, double alpha, double beta,
                        double dx, double dy, const double *a, const double *bX,
                        const double *bY, void *stream);
/* e = omega (sigma/D) N(r/sigma) on the registered periodic level. */
int bx_cnn_infer(bx_cnn *c, const double *r, double *e, int nx, int ny,
                 double omega, void *stream);
/* y = N(x) on nimg images of w x h, x rounded to BF16, zero padding or
 * wrap-around (BX_CNN_PERIODIC).  Device FP32 arrays.  By default the fused
 * inference kernel runs (all four layers per output tile in LDS, the kernel
 * bx_cnn_infer uses); BX_CNN_LAYERED runs the layer-by-layer kernels of the
 * training forward instead (w multiple of 16).  Both give identical results.
 * Synchronous. */
#define BX_CNN_PERIODIC 1
#define BX_CNN_LAYERED 2
int bx_cnn_net_forward(bx_cnn *c, const float *x, float *y, int w, int h,
                       int nimg, int flags, void *stream);
/* Loss and gradient on crops of the residual fields R (nfields x ny x nx,
 * device): crop b is the crop x crop window at (x0, y0) of field f, with
 * crops_host[3b..3b+2] = {f, x0, y0} (periodic wrap).  Loss per crop:
 * ||m (r - A e)||^2 / ||m r||^2 with e = (sigma/D) N(r/sigma) (omega = 1),
 * m = cells at distance >= 5 from the crop edge; averaged over the batch.
 * grad_host (BX_CNN_NPARAMS floats) may be NULL.  Synchronous. */
int bx_cnn_loss_grad(bx_cnn *c, const double *R, int nfields, int nx, int ny,
                     const int *crops_host, int batch, int crop, double *loss,
                     float *grad_host, void *stream);
/* `steps` Adam steps (lr, beta1 0.9, beta2 0.999) on random crops drawn with
 * `seed`; loss_history (steps doubles) may be NULL.  Synchronous. */
int bx_cnn_train(bx_cnn *c, const double *R, int nfields, int nx, int ny,
                 int crop, int batch, int steps, float lr, unsigned seed,
                 double *loss_history, void *stream);
const char *bx_cnn_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BAMREX_HIP_H */
