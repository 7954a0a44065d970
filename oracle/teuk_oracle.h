/*
 * CPU oracle for the 2+1D Teukolsky hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain FP64 host restatement of the
 * reference update loop in TD_GPU_code/HighOrder (main.c, kern.cu, body.cu,
 * proj.c).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it, and only as the checker or the timed CPU baseline.  The
 * product (bh_ringdown_amd, libteuk) never links or calls anything here.
 *
 * Storage layout follows parm.h:13, idx(b,a) = N*b + a: theta rows b (slow),
 * rho columns a (fast).  A state is 4 planes of N*M doubles in the order
 * q_re, q_im, p_re, p_im (kern.cu:10-13).  The coefficient block is 9 planes
 * in the order bb, cc, ee, a_re, a_im, d_re, d_im, v_re, v_im (kern.cu:23-31).
 *
 * Pinning status of each function is listed in the header of teuk_oracle.cpp.
 */
#ifndef TEUK_ORACLE_H
#define TEUK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  int N, M;            /* parm.h:1-2 */
  double xmin, xmax;   /* parm.h:3-4 (rho domain) */
  double mass, a, s, m;/* main.c:55-57, body() mm */
  double dx, dtheta, dt;
  int integrator;      /* 0 = SSP-RK3 (reference), 1 = classic RK4,
                          2 = staggered 2nd order (TD_GPU_code/Old, orc_*_so2) */
  int j_extract;       /* extraction column, -1 = N-1 (kern.cu:750-751) */
  int ptsx, ptsy;      /* parm.h:10-11 source patch size */
  int nthreads;        /* OpenMP threads in the timing build (<= 1: serial) */
} orc_params;

/* Per-step particle state handed to the source (body.cu:112-124). */
typedef struct orc_src_step {
  double timer, start;
  double rp, tp, phip, E, lz, Q;
  double drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2;
} orc_src_step;

/* main.c:79-198 grid(): rho/theta grids and the Newton inversion r(r*). */
void orc_grid(int N, int M, double xmin, double xmax, double mass, double a,
              double *x_c, double *r_c, double *theta);

/* kern.cu:84-169 kernel_init: 9 coefficient planes (zero outside 3<=k<M-3). */
void orc_coeffs(const orc_params *p, const double *theta, const double *r_c,
                double *coef);

/* kern.cu:171-299 kernel_rhs1/2.  src may be NULL (no particle source). */
void orc_rhs(const orc_params *p, const double *coef, const double *theta,
             const double *u, const double *src, double *rhs);

/* kern.cu:320-409 (theta edge rows) and kern.cu:412-449 (rho edge columns). */
void orc_bc_theta(const orc_params *p, double *u);
void orc_bc_rho(const orc_params *p, double *u);

/* One full time step (body.cu:131-165), in place on u.  work must hold
 * 16*N*M doubles.  src (2*N*M: re plane then im plane) may be NULL. */
void orc_step(const orc_params *p, const double *coef, const double *theta,
              double *u, const double *src, double *work);

/* kern.cu:770-2683 sourced: particle source on the PTSX x PTSY patch.
 * Writes the 2*N*M planes src (re, im) after zeroing them (kern.cu:755-768). */
void orc_source(const orc_params *p, const double *theta, const double *r_c,
                const orc_src_step *st, double *src);

/* proj.c:8-75 */
double orc_fact(double n);
double orc_swsh(double ell, double ml, double the);
/* qq must hold M+1 doubles; qq[M] is overwritten with qq[M-1] (proj.c:45). */
double orc_project(double *qq, int M, double l, double m);

/* A full solve driver: nsteps steps from u, per-step extraction and
 * integration (body.cu:168-202).  Outputs (any may be NULL):
 *   psi_hist  [nsteps][2][M]  extraction column after each step
 *   h_hist    [nsteps][14]    (h+, hx) for l=|m|..|m|+6 after each step
 *   dedt_hist [nsteps]
 * src_steps may be NULL (no source) or hold nsteps entries. */
int orc_run(const orc_params *p, const double *r_c, const double *theta,
            double *u, int nsteps, const orc_src_step *src_steps,
            double *psi_hist, double *h_hist, double *dedt_hist);

/* Staggered second-order scheme of TD_GPU_code/Old (teuk_oracle_so2.inc):
 * half-point radii, the two coefficient sets (9 planes each), one step
 * (work: 12 planes, zeroed before the first step and kept between steps),
 * and a run with two source entries per step (half point, then grid point). */
void orc_grid_half(int N, double xmin, double xmax, double mass, double a, double *x_h, double *r_h);
void orc_coeffs_so2(const orc_params *p, const double *theta, const double *r_h, const double *r_c,
                    double *coef_h, double *coef_c);
void orc_step_so2(const orc_params *p, const double *coef_h, const double *coef_c, const double *theta,
                  double *u, const double *src_h, const double *src_c, double *work);
int orc_run_so2(const orc_params *p, const double *r_h, const double *r_c, const double *theta, double *u,
                int nsteps, const orc_src_step *src_steps, double *psi_hist, double *h_hist, double *dedt_hist);

#ifdef __cplusplus
}
#endif
#endif
