// Internal declarations shared by the libteuk host and device sources.
// Not part of the public C ABI (include/teuk.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace teuk {

// parm.h:12 -- the 15-digit pi the reference uses for dtheta, cot and the
// spherical harmonics.
constexpr double kPi15 = 3.14159265358979;
// kern.cu:46
constexpr double kThird = 0.33333333333333333;
// kern.cu:196 -- first-derivative weight literal
constexpr double kD1 = 0.016666666666667;

// Stage-kernel tile: one wavefront owns 64 consecutive rho columns and a
// chunk of kRows consecutive theta rows; up to 4 wavefronts per block stack
// along theta.
constexpr int kWave = 64;
constexpr int kRows = 8;
constexpr int kHalo = 4;            // KO reach in rho (kern.cu:247-256)
constexpr int kLdsCols = kWave + 2 * kHalo;

// Scalar run description passed by value to every kernel.
struct Grid {
  int N, M;
  int even_m;          // ((int)m) % 2 == 0  (kern.cu:330)
  int ptsx, ptsy;
  int je;              // extraction column
  int Fr;
  int nrec_cap;        // output records the device buffers hold
  double xmin, xmax, dx, dtheta, dt;
  double mass, a, m;
  double start;
};

// Device control block, one per solver.
struct Ctl {
  int k;               // completed steps
  int nrec;            // output records written
  int status;          // 0 ok, 1 non-finite field seen
  int pad;
  double timer;        // solver time, advanced by repeated addition (body.cu:214)
};

// Per-step particle scalars on the device (mirrors teuk_source_step).
struct SrcStep {
  double timer, start;
  double rp, tp, phip, E, lz, Q;
  double drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2;
};

// Device pointers of one solver.
struct DevState {
  double *U, *C, *D, *A;   // field states (4 planes each); A only for RK4
  double *coef;            // 9 planes: bb cc ee a_re a_im dre d_im v_re v_im
  double *ctan;            // [M] tan(Pi/2 - theta), host libm
  double *col4;            // [N][4] x, Omega, Omega', Omega'' (host libm)
  double *row4;            // [M][4] sin, cos, cot, csc of theta (host libm)
  double *theta, *r_c;     // [M], [N]
  double *patch;           // [2][ptsy][ptsx] source patch
  int *patch_org;          // [2] {row0, col0}; row0 = INT_MIN when empty
  SrcStep *steps;          // [nsteps] source schedule, or null
  int nsteps_src;
  double *ext;             // [4][M] int_re, int_im, dint_re, dint_im
  double *extw;            // [M] dvol, [M] rho2^4, [1] rr^6
  double *projw;           // [7][5*(M/4)][3] Boole/swsh factors per (l, term)
  double *rec;             // [nrec_cap][2+14] t, dedt, h[14]
  int *rec_k;              // [nrec_cap][2] k, write flag
  double *psi;             // [nrec_cap][2][M]
  Ctl *ctl;
};

// Launchers (teuk_kernels.hip).
hipError_t launch_init_coeffs(const Grid &g, const DevState &d, hipStream_t s);
hipError_t launch_source(const Grid &g, const DevState &d, hipStream_t s);
hipError_t launch_stage(const Grid &g, const DevState &d, int integrator, int stage,
                        bool has_src, hipStream_t s);
hipError_t launch_extract(const Grid &g, const DevState &d, hipStream_t s);

// Host helpers (teuk_host.cpp).
void host_grid(int N, int M, double xmin, double xmax, double mass, double a,
               double *x_c, double *r_c, double *theta);
double host_swsh(double ell, double ml, double the);
void host_projection_weights(int M, double m, double *w /*[7][5*(M/4)][3]*/);
// Transcendental inputs of the coefficient fields (kern.cu:110-127, 204).
void host_coeff_tables(int N, int M, double xmin, double xmax, double dx, const double *theta,
                       double *col4, double *row4, double *ctan);
bool source_available();

}  // namespace teuk
