/*
 * libteuk -- MI355X-native FP64 2+1D Teukolsky evolution (spin weight s=-2,
 * one azimuthal mode m) behind a plain C ABI.  No HIP, CUDA or torch types
 * appear in these signatures: host buffers are plain double arrays, sizes are
 * ints, and every entry point returns 0 or a negative TEUK_E* code.
 *
 * Grid layout is the reference's (TD_GPU_code/HighOrder/parm.h:13):
 *   idx(b, a) = N*b + a,  b = theta row (0..M-1), a = rho column (0..N-1).
 * A field state is 4 planes of N*M doubles: q_re, q_im, p_re, p_im
 * (Psi and Pi of kern.cu:10-13).
 *
 * Threading: one solver handle is driven by one host thread at a time.
 * Different handles are independent (each owns its device memory and a HIP
 * stream), which is how batched and served solves are built.
 */
#ifndef TEUK_H
#define TEUK_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define TEUK_OK          0
#define TEUK_EINVAL     -1   /* bad argument or unsupported size            */
#define TEUK_EHIP       -2   /* a HIP runtime call failed                    */
#define TEUK_ENOMEM     -3   /* device or host allocation failed             */
#define TEUK_EDIVERGED  -4   /* the solve failed: a non-finite value, or one
                                above 1e60 in magnitude, in its fields       */
#define TEUK_EIO        -5   /* trajectory or waveform file error            */
#define TEUK_ERANGE     -6   /* trajectory exhausted / step beyond schedule  */
#define TEUK_ETOOBIG    -7   /* more memory than teuk_solve_noalloc's fixed
                                scratch (TEUK_NOALLOC_BYTES)                 */

/* ---- run parameters ---------------------------------------------------- */
/* Field names follow parm.h:1-11 and main.c:55-61 of the reference. */
typedef struct teuk_params {
  int    N, M;          /* radial (rho) and angular (theta) points           */
  double xmin, xmax;    /* rho domain, parm.h:3-4 (Xmin, Xmax)               */
  double mass, a;       /* main.c:55-56 (mass, aa)                           */
  double s;             /* spin weight; only -2 is supported (kern.cu:142)   */
  double m;             /* azimuthal mode number (body() argument mm)        */
  double dx, dtheta, dt;/* main.c:59-61; dtheta is 3.14159265358979/M        */
  double T;             /* stop time, parm.h:8                               */
  int    Fr;            /* output cadence in steps, parm.h:7                 */
  int    ptsx, ptsy;    /* source patch size, parm.h:10-11                   */
  int    integrator;    /* 0 = SSP-RK3 (reference), 1 = classic RK4,
                           2 = the staggered second-order scheme of
                           TD_GPU_code/Old (kern.cu, body.cu there)         */
  int    j_extract;     /* extraction column; -1 = N-1 (scri, kern.cu:750)   */
} teuk_params;

/* Per-step particle state seen by the source deposit: the values body.cu
 * hands to `sourced` at body.cu:122-124 after reading two trajectory lines
 * (body.cu:112-120).  `timer` is the solver time at the start of the step. */
typedef struct teuk_source_step {
  double timer, start;
  double rp, tp, phip, E, lz, Q;
  double drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2;
} teuk_source_step;

/* One output record, written every Fr steps (body.cu:204-212).  `t` is the
 * solver time before the increment (timer), `h` the 7 (h+, hx) pairs for
 * l = |m| .. |m|+6 (body.cu:189-202), `dedt` the energy flux
 * (body.cu:174-183), `write` 1 when body.cu:207 would write the row. */
typedef struct teuk_output_row {
  int    k;
  int    write;
  double t;
  double dedt;
  double h[14];
} teuk_output_row;

typedef struct teuk_solver teuk_solver;

/* ---- host set-up --------------------------------------------------------- */
/* main.c:79-198 grid(): x_c[N], r_c[N] (Boyer-Lindquist r by Newton
 * inversion of r*(r) = x/Omega(x)), theta[M].  Replaces grid(). */
int teuk_grid(const teuk_params *p, double *x_c, double *r_c, double *theta);

/* Half points of the staggered scheme (integrator 2): x_h[l] = xmin +
 * (l - 1/2) dx, where TD_GPU_code/Old/kern.cu:93 places its half-point
 * coefficients, and r_h[N] by the same Newton inversion as teuk_grid.  Old/
 * has no main.c; this supplies the r_h its body() receives. */
int teuk_grid_half(const teuk_params *p, double *x_h, double *r_h);

/* Convenience: fill a teuk_params with the reference's derived quantities
 * (dx = (xmax-xmin)/N, dtheta = Pi/M, main.c:59-60), dt = dt_over_dx*dx,
 * s = -2, Fr = 1, ptsx = ptsy = 128, j_extract = -1. */
int teuk_default_params(teuk_params *p, int N, int M, double xmin, double xmax,
                        double a, double m, double dt_over_dx, double T);

/* ---- solver lifecycle ---------------------------------------------------- */
/* Allocates device state, uploads r_c/theta, computes the Kerr coefficient
 * fields and zeroes the fields (kernel_init, kern.cu:84-169; body.cu:32-54).
 * N must be a multiple of 64 and >= 128; M a multiple of 8, >= 16 and
 * <= 1024 (<= 256 for the staggered scheme); N * M <= 2^26. */
int teuk_create(const teuk_params *p, const double *r_c, const double *theta,
                teuk_solver **out);

/* The staggered scheme (integrator 2) with caller-supplied half-point radii,
 * as TD_GPU_code/Old/body.cu:28-30 receives r_h beside r_c and theta (and
 * kernel_init computes both coefficient sets, Old/kern.cu:61-193).
 * teuk_create with integrator 2 computes r_h by teuk_grid_half. */
int teuk_create_staggered(const teuk_params *p, const double *r_h, const double *r_c,
                          const double *theta, teuk_solver **out);
void teuk_destroy(teuk_solver *s);

/* Host <-> device field transfer, layout 4 x N x M (see top). */
int teuk_set_fields(teuk_solver *s, const double *u);
int teuk_get_fields(teuk_solver *s, double *u);

/* Source schedule: one entry per step (NULL / 0 = no source).  Replaces the
 * per-step reads of body.cu:112-129.  For the staggered scheme two entries
 * per step: the half step's (evaluated at the half points), then the full
 * step's (Old/body.cu:127-180). */
int teuk_set_source(teuk_solver *s, const teuk_source_step *steps, int nsteps);

/* Builds the source schedule from raw .traj rows (10 columns, the format of
 * GenOT/WL_BL_RK.cc and body.cu:97) with the reference's consumption pattern:
 * prime with one line then read until t >= start (body.cu:97-103), then two
 * lines per step, r += 1e-6, accelerations by backward difference over dt
 * with zero initial velocities (body.cu:112-129). */
int teuk_set_trajectory(teuk_solver *s, const double *rows, int nrows, int nsteps);

/* Host-only: the per-step schedule teuk_set_trajectory builds, written to
 * out[nsteps] (start = solver start time, 0 for the reference's p = 1). */
int teuk_trajectory_schedule(const double *rows, int nrows, int nsteps, double start, double dt,
                             teuk_source_step *out);

/* Host-only: the staggered scheme's schedule, out[2 * nsteps]: after the
 * same priming, one line per half step, each used as read, accelerations by
 * backward difference over dt/2 ((v - v_prev) / dt * 2), the step's timer
 * in both entries (Old/body.cu:113-180). */
int teuk_trajectory_schedule_staggered(const double *rows, int nrows, int nsteps, double start, double dt,
                                       teuk_source_step *out);

/* Advance nsteps full time steps (SSP-RK3: three fused stage kernels per
 * step; RK4: four; staggered: a half-step and a full-step kernel, each
 * followed by its edge fill).  Asynchronous on the solver's stream; use
 * teuk_sync. */
int teuk_step(teuk_solver *s, int nsteps);
/* Waits for the solver's queued steps; TEUK_EDIVERGED once the solve has
 * failed (see teuk_status). */
int teuk_sync(teuk_solver *s);

/* TEUK_OK, or TEUK_EDIVERGED once this solve has failed: after each chunk
 * of steps (at most 32, or one teuk_step call) a device sentinel checks
 * every value of the fields for non-finite values and magnitudes above
 * 1e60, and each step checks Psi at the extraction column.  Works alone or
 * in a batch (waits for the stream that steps the solver); a failed solve
 * never touches another solve's memory, so its batch neighbours go on
 * unchanged and the caller retires it with teuk_batch_remove.  No
 * counterpart in the reference, which checks nothing (body.cu:105-216). */
int teuk_status(teuk_solver *s);

/* Number of completed steps and current solver time. */
int teuk_progress(teuk_solver *s, int *k, double *t);

/* Output records produced so far (every Fr steps).  Copies up to max rows
 * starting at record `first`; *n receives the count copied. */
int teuk_read_output(teuk_solver *s, int first, int max, teuk_output_row *rows, int *n);

/* Psi at the extraction column (M complex values, re plane then im plane)
 * recorded with each output row: psi[(row - first)][2][M]. */
int teuk_read_psi(teuk_solver *s, int first, int max, double *psi, int *n);

/* Kernel timing: when enabled, teuk_step brackets each step's back-to-back
 * stage-kernel launches with a pair of HIP events on the solver stream.
 * teuk_timing returns the number of timed stage launches and their summed
 * duration in milliseconds. */
int teuk_enable_timing(teuk_solver *s, int on);
int teuk_timing(teuk_solver *s, long long *launches, double *ms);

/* Device bytes held by one solver of these parameters for `nsteps` of
 * schedule (admission control for batching/serving). */
long long teuk_footprint(const teuk_params *p, int nsteps);

/* Launch shape of the stage kernels: wavefronts per block (1, 2, 4; each
 * wavefront takes one chunk of theta rows, see teuk_set_theta_rows, and
 * when the chunks of a rho strip are not a multiple of it the surplus
 * wavefronts of the strip's last block stay idle) and the occupancy the
 * kernels were compiled for (2, 3 or 4 waves per SIMD).  Results do not depend
 * on it. */
int teuk_set_tiling(teuk_solver *s, int waves_per_block, int waves_per_eu);

/* Theta rows per stage-kernel wavefront: 2 (the default), 4 or 8.  The
 * first and last chunk of a rho strip compute 2, 3 or 5 rows next to the
 * theta boundary rows they fill, the chunks between `rows` rows each; a
 * 64-column strip of an M-row grid then takes 2 + (M - 10) / 2,
 * 2 + (M - 12) / 4 or M / 8 wavefronts.  Short chunks give more wavefronts
 * and fewer registers per wavefront, long ones re-read fewer theta halo
 * rows.  Results do not depend on it.  No counterpart in the reference (its
 * kernels take one thread per point, body.cu:131-165). */
int teuk_set_theta_rows(teuk_solver *s, int rows);

/* The 12 per-step source coefficients (the canonical T_4 coefficients times
 * e^{-i m phi_p}/pi, (0,0) (1,0) (0,1) (2,0) (1,1) (0,2) as re, im pairs)
 * the device computed for steps first .. first+n-1, into c[n][12].  No
 * counterpart in the reference (its sourced() keeps them in registers,
 * kern.cu:1229-2679); for checking the source stage by stage. */
int teuk_source_coeffs(teuk_solver *s, int first, int n, double *c);

/* The source patch of step k (2 x ptsy x ptsx: real plane, imaginary
 * plane; cell (j, i) at grid point (origin[0] + j, origin[1] + i)), as the
 * device computes it for that step (kern.cu:770-2683).  For checking; call
 * it while the solver is idle. */
int teuk_source_patch(teuk_solver *s, int k, double *patch, int *origin);

/* Steps left in the solver's schedule ((int)(T/dt), or the source schedule
 * if shorter). */
int teuk_remaining(teuk_solver *s, int *steps);

/* ---- batches ------------------------------------------------------------------ */
/* No counterpart in the reference (body() solves one mode per call; the
 * reference runs modes and trajectories as separate jobs, main.c:50-55,
 * run.sh:3-16).  A batch steps many solvers together: each RK stage of every
 * member is one launch (blocks index a table of {solver, tile}); chunks of
 * steps are captured in a hipGraph.  Members may differ in N, M, a, m, dt,
 * source and length; each stops at its own schedule end.  A member's results
 * are bitwise those of stepping it alone.  Members must share the
 * integrator.  While in a batch a solver is stepped only by its batch;
 * teuk_read_output / teuk_read_psi / teuk_get_fields / teuk_progress still
 * work on it. */
typedef struct teuk_batch teuk_batch;
/* hbm_budget: bytes of member footprint (teuk_footprint) the batch admits;
 * 0 = no limit.  Integrators 0 and 1 only. */
int teuk_batch_create(int integrator, long long hbm_budget, teuk_batch **out);
void teuk_batch_destroy(teuk_batch *b);      /* members stay valid */
/* TEUK_ENOMEM when the member would exceed the budget. */
int teuk_batch_add(teuk_batch *b, teuk_solver *s);
int teuk_batch_remove(teuk_batch *b, teuk_solver *s);
/* order 0: each XCD gets a contiguous run of tiles; 1: tiles dealt in list
 * order.  chunk: steps per captured graph. */
int teuk_batch_set_tiling(teuk_batch *b, int waves_per_block, int waves_per_eu, int order, int chunk);
/* Theta rows per stage-kernel wavefront for every member (as
 * teuk_set_theta_rows). */
int teuk_batch_set_theta_rows(teuk_batch *b, int rows);
/* Cache groups: the unfinished members, in list order, are split into
 * groups whose field state (RK state planes and the dbdx plane, 136 B per
 * point for RK4) stays within group_bytes, so that a group's fields stay in
 * the last-level cache while it runs a chunk of steps; each chunk then runs
 * group after group.  0 (the default) = one group.  Results do not depend
 * on it. */
int teuk_batch_set_cache_groups(teuk_batch *b, long long group_bytes);
/* Advance every unfinished member by up to nsteps steps (asynchronous). */
int teuk_batch_step(teuk_batch *b, int nsteps);
/* Waits for the batch; TEUK_EDIVERGED if any member has failed (find which
 * with teuk_status on each). */
int teuk_batch_sync(teuk_batch *b);
int teuk_batch_info(teuk_batch *b, int *members, int *active, long long *bytes_used);

/* ---- drop-in for body() ---------------------------------------------------- */
/* body.cu:25-225: one full m-mode solve.  Reads the trajectory file
 * (traj_path, may be NULL for no source), evolves from zero fields to T,
 * writes the waveform file in the reference format "%lf %g x14" to out_path
 * (NULL: no file) and, if full_path is non-NULL, the same rows at %.17g.
 * Prints the reference's stdout lines when verbose.  With integrator 2 it
 * is Old/body.cu:28-258 instead: r_h from teuk_grid_half, one trajectory
 * line per half step, rows "%lf %g x10" (l = |m| .. |m|+4) written only
 * while the particle radius is <= 12 (Old/body.cu:233-237). */
int teuk_body(const teuk_params *p, const double *r_c, const double *theta,
              const char *traj_path, const char *out_path, const char *full_path,
              int verbose);

/* ---- trajectory source ----------------------------------------------------- */
/* Seeded inspiral-transition-plunge worldline (mass ratio eta, inclination
 * iota in radians, 0 <= iota <= 85 degrees, start at r0 = r0_over_isco x the
 * inclined ISCO radius, or where the transition begins if that is further
 * out): the quadrupole-flux inspiral, the generalised Ori-Thorne transition
 * and the geodesic plunge of GenOT/ParamWLMino.cc, sampled
 * every dt_out of Boyer-Lindquist time up to t_final, as .traj rows
 * (10 doubles each: t r theta dr/dt dtheta/dt dphi/dt phi E Lz Q, the format
 * GenOT/WL_BL_RK.cc:214-221 writes and body.cu:97 reads).  Use
 * dt_out = dt/2 for a solver step dt (body.cu:112-113 reads two rows per
 * step).  Replaces running GenOT, whose adiabatic input files are not
 * part of the reference tree. */
int teuk_traj_plunge(double a, double eta, double iota, double r0_over_isco, double dt_out,
                     double t_final, unsigned long long seed, double *rows, int cap, int *nrows);

/* ---- a solve that allocates nothing --------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation): one whole solve -- what teuk_create, teuk_set_fields,
 * teuk_set_source, teuk_step(nsteps), teuk_sync, teuk_get_fields,
 * teuk_read_output and teuk_destroy do together, with the same kernels and
 * bit for bit the same results -- that calls no allocating routine, host or
 * device.  All of its device memory (fields, tables, schedule, records) is a
 * fixed scratch array of TEUK_NOALLOC_BYTES that each GPU the library runs
 * on holds until the process ends; a solve needing more returns
 * TEUK_ETOOBIG (e.g. RK4 at 4096 x 32 fits with a 6000-step schedule and
 * Fr >= 2; 16384 x 64 does not).  Integrators 0 and 1.
 *   u0:    initial fields (layout of teuk_set_fields), NULL = zero fields
 *   steps: nsteps source entries (teuk_set_source), NULL = no source
 *   u_out: final fields (teuk_get_fields layout), may be NULL
 *   rows:  up to max_rows output records (teuk_read_output); *nrows = count
 * Calls are serialised by a process-wide lock (they share the scratch, which
 * keeps the last call's contents; each call clears what must start at
 * zero).  Runs on the current device's null stream and returns when done;
 * TEUK_EDIVERGED if the sentinel fired. */
#define TEUK_NOALLOC_BYTES (64LL << 20)
int teuk_solve_noalloc(const teuk_params *p, const double *r_c, const double *theta,
                       const double *u0, const teuk_source_step *steps, int nsteps,
                       double *u_out, teuk_output_row *rows, int max_rows, int *nrows);

/* The universal Ori-Thorne transition function X(L) (X'' = -X^2 - L,
 * X -> sqrt(-L) as L -> -inf) and X'(L), for -40 <= L <= 3.4, as
 * teuk_traj_plunge's transition uses it.  Replaces the table GenOT reads,
 * GenOT/OT_X_vs_L.dat (ParamWLMino.cc:190-205). */
int teuk_ot_transition(double L, double *X, double *dX);

/* Human-readable text for an error code. */
const char *teuk_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif
