/*
 * CPU oracle for the 2+1D Teukolsky hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates, in plain FP64 C++, the update loop of the reference solver
 * TD_GPU_code/HighOrder.  Every function cites the reference lines it
 * follows.  The product library never links this file; tests, the smoke
 * check and bench.py's cpu_baseline leg load it as the checker.
 *
 * Parity build:  g++ -O2 -ffp-contract=off  (no implicit FMA contraction: the
 *                point arithmetic of teuk_point.h, shared with the device,
 *                places its fused multiply-adds explicitly, everything else
 *                is evaluated in the order written)
 * Timing build:  g++ -O3 -fopenmp           (cpu_baseline leg of bench.py)
 *
 * Pinning (how each function is checked, in tests/):
 *   orc_grid      bitwise against the reference main.c grid() compiled by
 *                 oracle/Makefile into oracle/_ref, plus the Newton residual
 *                 r*(r) = x/Omega to 1e-12 at any size.
 *   orc_fact/orc_swsh/orc_project
 *                 bitwise against the reference proj.c compiled into
 *                 oracle/_ref, plus closed forms of -2Y_lm and orthonormality.
 *   orc_coeffs    a restatement of kern.cu:132-158 in factored form, pinned
 *                 against Teukolsky's equation: the second-order operator the
 *                 nine fields define equals the one derived symbolically from
 *                 the s = -2 master equation in the compactified coordinates
 *                 (Kerr azimuth, r^3 rescaling, hyperboloidal layer), to
 *                 1e-11 inside and outside the layer; bb is the negative root.
 *   orc_rhs       stencil exactness on polynomials, KO annihilation of degree
 *                 <= 7, assembly against a direct evaluation on smooth fields.
 *   orc_bc_*      exact extrapolation formulas and ordering invariants.
 *   orc_step      SSP-RK3 / RK4 amplification polynomials on linear modes.
 *   orc_source    parity unpinned against the reference (kern.cu needs the
 *                 CUDA compiler, absent here); pinned by the distributional
 *                 canonical form and physical checks (see teuk_source.inc).
 *   orc_*_so2     the staggered second-order scheme of TD_GPU_code/Old
 *                 (teuk_oracle_so2.inc lists its checks).
 */
#include "teuk_oracle.h"
#include "teuk_point.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* parm.h:12 -- the 15-digit pi used for dtheta (main.c:60), cot (kern.cu:204)
 * and the spherical harmonics (proj.c:33,43). */
const double kPi15 = 3.14159265358979;
/* kern.cu:46 */
const double kThird = 0.33333333333333333;
/* kern.cu:196 -- first-derivative weight literal (not exactly 1/60) */
const double kD1 = 0.016666666666667;

inline size_t plane(const orc_params *p) { return (size_t)p->N * (size_t)p->M; }

inline int extract_col(const orc_params *p) {
  return (p->j_extract < 0) ? p->N - 1 : p->j_extract;
}

}  // namespace

/* ---------------------------------------------------------------- grid --- */
namespace {

/* main.c:79-198 from the rho grid on: the hyperboloidal layer Omega,
 * r* = rho/Omega, and a Newton solve for the Boyer-Lindquist r(r*) marching
 * from the outer edge inward, each point warm-started from its neighbour. */
void newton_radii(int N, const double *x, double xmin, double xmax, double mass, double a, double *r_out) {
  const double rin = xmin + (xmax - xmin) * 0.64;
  const double outer = xmax + 1e-6;
  std::vector<double> rstar(N);
  for (int l = 0; l < N; l++) {
    double om = 1.0;
    if (!(x[l] < rin)) om = 1.0 - std::pow((x[l] - rin) / (outer - rin), 4);
    rstar[l] = x[l] / om;
  }

  const double root = std::sqrt(mass * mass - a * a);
  const double rplus = mass + root, rminus = mass - root;
  if (mass == 0.0) {
    for (int l = 0; l < N; l++) r_out[l] = rstar[l];
    return;
  }
  const double tol = 1e-12;
  double r = rstar[N - 1];
  for (int l = N - 1; l >= 0; l--) {
    double rel;
    double rnext;
    do {
      /* dr/dr* = Delta/(r^2+a^2) */
      const double slope = (r * r + a * a - 2.0 * mass * r) / (r * r + a * a);
      double tort;
      if (mass == a) {
        tort = r + 2.0 * a * std::log(r / a - 1.0) - 2.0 * a * a / (r - a);
      } else if (a == 0.0) {
        tort = r + 2.0 * mass * std::log(r / (2.0 * mass) - 1.0);
      } else {
        tort = r + (rplus * rplus + a * a) / (rplus - rminus) * std::log(r / rplus - 1.0)
                 - (rminus * rminus + a * a) / (rplus - rminus) * std::log(r / rminus - 1.0);
      }
      rnext = r + (rstar[l] - tort) * slope;
      /* Deviation from main.c: where the iterate reaches the horizon (or
       * NaN) the reference goes on to log(0) and fills the rest of the grid
       * with NaN; here r is clamped to r+ (the r* -> -inf limit) instead.
       * Wherever the reference result is finite the two agree bitwise. */
      if (!(rnext > rplus)) { rnext = rplus; r = rnext; break; }
      rel = std::fabs(rnext - r) / r;
      r = rnext;
    } while (rel > tol);
    r_out[l] = rnext;
  }
}

}  // namespace

extern "C" {

/* main.c:79-198 grid(): uniform rho and theta grids and the radii. */
void orc_grid(int N, int M, double xmin, double xmax, double mass, double a,
              double *x_c, double *r_c, double *theta) {
  const double dx = (xmax - xmin) / N;
  const double dth = kPi15 / M;
  for (int l = 0; l < N; l++) x_c[l] = xmin + (double)l * dx;
  for (int k = 0; k < M; k++) theta[k] = (double)k * dth;
  newton_radii(N, x_c, xmin, xmax, mass, a, r_c);
}

/* ---------------------------------------------------------- coefficients --- */
}  // extern "C"

namespace {

/* The theta dependence of kernel_init's formulas (kern.cu:132-145) enters
 * only through sin^2(theta) in bir, cos(theta) in B^tau_im and the angular
 * potential Y = 4 cot^2 - 4 m cot csc + m^2 csc^2 in C_re.  Each field is
 * rewritten as (column constant [+ column constant x row value]) / bir, with
 * bir = K2 sin^2 - K1; the algebra is that of kern.cu:132-145 term by term,
 * only regrouped.  Column j sits at rho = Xmin + (j + shift) dx with radius
 * r[j] (shift 0: the grid points; -1/2: the half points of the staggered
 * scheme, Old/kern.cu:93). */
enum { K1, K2, KA, KR, KD, KBt, KBr, KC0, KI0, KI1, KBi, KCi, NK };   /* teuk_point.h order */

void column_constants(const orc_params *p, double shift, const double *r_col, std::vector<double> &K) {
  const int N = p->N;
  const double mass = p->mass, spin = p->a, em = p->m;
  const double rin = p->xmin + (p->xmax - p->xmin) * 0.64;
  const double outer = p->xmax + 1e-6;
  K.assign((size_t)NK * N, 0.0);
  for (int j = 0; j < N; j++) {
    const double x = p->xmin + ((double)j + shift) * p->dx;
    double om, om1, om2;
    if (x < rin) {
      om = 1.0; om1 = 0.0; om2 = 0.0;
    } else {
      om = 1.0 - std::pow((x - rin) / (outer - rin), 4);
      om1 = -4.0 * std::pow(x - rin, 3) / std::pow(outer - rin, 4);
      om2 = -12.0 * std::pow(x - rin, 2) / std::pow(outer - rin, 4);
    }
    const double L = om - x * om1;      /* Omega^2 times d(r*)/d(rho) */
    const double R = om * r_col[j];     /* compactified radius */
    const double a2 = spin * spin, a4 = a2 * a2;
    const double om_2 = om * om, om3 = om_2 * om, om4 = om_2 * om_2, om5 = om4 * om;
    const double S = R * R + om_2 * a2;                 /* (r^2+a^2) Omega^2 */
    const double D = R * R - 2.0 * R * mass * om + om_2 * a2; /* Delta Omega^2 */
    const double L2 = L * L, L3 = L2 * L;
    const double R2 = R * R, R3 = R2 * R, R5 = R3 * R2;
    double *c = &K[0];
    c[K1 * N + j] = ((2.0 * L - om_2) * (S * S)) / L2;
    c[K2 * N + j] = a2 * D;
    c[KA * N + j] = (2.0 * (L - om_2) * (S * S)) / L2;
    c[KR * N + j] = -((om_2 * (S * S)) / L2);
    c[KD * N + j] = D;
    c[KBt * N + j] = -((om * (R3 * (4.0 * L2 * (R - 3.0 * mass * om) + 2.0 * L * R2 * om1 + R2 * om * om2 * x)
                              + 2.0 * R * (L2 * om_2 * (5.0 * R - 4.0 * mass * om)
                                           - 3.0 * L3 * (R - 2.0 * mass * om)
                                           + 2.0 * L * R2 * om_2 * om1 + R2 * om3 * om2 * x) * a2
                              + (-6.0 * L3 * om_2 + 6.0 * L2 * om4 + 2.0 * L * R * om4 * om1
                                 + R * om5 * om2 * x) * a4))
                       / (L3 * R));
    c[KBr * N + j] = -((om * (R5 * (2.0 * om * om1 - 2.0 * om1 * om1 * x + om * om2 * x)
                              + 2.0 * R3 * om_2 * (2.0 * om * om1 - 2.0 * om1 * om1 * x + om * om2 * x) * a2
                              + om4 * (6.0 * om_2 + 2.0 * om1 * om1 * x * (-R + 3.0 * x)
                                       + om * (2.0 * R * om1 - 12.0 * om1 * x + R * om2 * x)) * a4
                              + 2.0 * L2 * R * (2.0 * R3 - 6.0 * R2 * mass * om + 5.0 * R * om_2 * a2
                                                - 4.0 * mass * om3 * a2)))
                       / (L3 * R));
    c[KC0 * N + j] = (D * (2.0 * R2 + 6.0 * R * mass * om - 6.0 * om_2 * a2)) / R2;
    c[KI0 * N + j] = (2.0 * spin * (em * (-(om_2 * S) + L * (R2 + 2.0 * R * mass * om + om_2 * a2)))) / L;
    c[KI1 * N + j] = 4.0 * spin * D;
    c[KBi * N + j] = (-2.0 * em * om_2 * spin * S) / L;
    c[KCi * N + j] = (2.0 * em * om * spin * (R2 + 2.0 * R * mass * om - 3.0 * om_2 * a2)) / R;
  }
}

/* The nine fields (kern.cu:147-158) on rows lo <= k < hi from the column
 * constants; zero elsewhere (kern.cu:113). */
void fill_coeffs(const orc_params *p, const double *theta, const std::vector<double> &K, int lo, int hi,
                 double *coef) {
  const int N = p->N;
  const size_t NM = plane(p);
  const double em = p->m;
  std::memset(coef, 0, 9 * NM * sizeof(double));
  for (int k = lo; k < hi; k++) {
    const double y = theta[k];
    const double sn = std::sin(y), cs = std::cos(y);
    const double cot = 1.0 / std::tan(y), csc = 1.0 / std::sin(y);
    const double sn2 = sn * sn;
    const double Y = 4.0 * (cot * cot) - 4.0 * em * cot * csc + (em * em) * (csc * csc);
    for (int j = 0; j < N; j++) {
      double Kj[NK];
      for (int i = 0; i < NK; i++) Kj[i] = K[(size_t)i * N + j];
      const teuk_pt::Coef cf = teuk_pt::coef_point(Kj, sn2, cs, Y, [](double v) { return std::sqrt(v); });
      const size_t id = (size_t)k * N + j;
      coef[0 * NM + id] = cf.bb;
      coef[1 * NM + id] = cf.cc;            /* A^yy = -Delta/bir */
      coef[2 * NM + id] = cf.ee;
      coef[3 * NM + id] = cf.are;
      coef[4 * NM + id] = cf.aim;
      coef[5 * NM + id] = cf.dre;
      coef[6 * NM + id] = cf.dim;
      coef[7 * NM + id] = cf.vre;
      coef[8 * NM + id] = cf.vim;
    }
  }
}

}  // namespace

extern "C" {

/* kern.cu:84-169.  The compactified Teukolsky operator is written as
 *   d_tau^2 Psi = A^tr d_tau d_rho Psi + A^rr d_rho^2 Psi
 *               + A^yy (d_th^2 + cot d_th) Psi + B^tau d_tau Psi
 *               + B^rho d_rho Psi + C Psi
 * and reduced to first order with Pi = d_tau Psi + bb d_rho Psi (kern.cu:147-158);
 * the fields are set on rows 3 <= k < M-3 (kern.cu:113). */
void orc_coeffs(const orc_params *p, const double *theta, const double *r_c,
                double *coef) {
  std::vector<double> K;
  column_constants(p, 0.0, r_c, K);
  fill_coeffs(p, theta, K, 3, p->M - 3, coef);
}

/* ------------------------------------------------------------------- rhs --- */
/* kern.cu:171-299 (kernel_rhs1; kernel_rhs2 at kern.cu:452-580 is the same
 * arithmetic on the stage arrays). */
void orc_rhs(const orc_params *p, const double *coef, const double *theta,
             const double *u, const double *src, double *rhs) {
  const int N = p->N, M = p->M;
  const size_t NM = plane(p);
  const double *qr = u, *qi = u + NM, *pr = u + 2 * NM, *pi = u + 3 * NM;
  const double *bb = coef, *cc = coef + NM, *ee = coef + 2 * NM;
  const double *are = coef + 3 * NM, *aim = coef + 4 * NM;
  const double *dre_c = coef + 5 * NM, *dim = coef + 6 * NM;
  const double *vre = coef + 7 * NM, *vim = coef + 8 * NM;
  const double over_dx = 1.0 / p->dx;
  const double over_dth = 1.0 / p->dtheta;
  const double kof = teuk_pt::ko_factor(p->dt);
  const double wy = 0.05 * kThird * over_dth;
  const double wyy = 0.05 * kThird * kThird * over_dth * over_dth;
  const double wx = kD1 * over_dx;

  /* Point arithmetic (stencils with explicit fused multiply-adds, RHS
   * assembly) from teuk_point.h, the same code the device compiles. */
  auto dx1 = [&](const double *f, size_t c) {
    return teuk_pt::d1(f[c + 1], f[c - 1], f[c + 2], f[c - 2], f[c + 3], f[c - 3], wx);
  };
  auto ko8 = [&](const double *f, size_t c) {
    return teuk_pt::ko8(f[c], f[c + 1], f[c - 1], f[c + 2], f[c - 2], f[c + 3], f[c - 3], f[c + 4], f[c - 4]);
  };
  const size_t sN = (size_t)N;
  auto dy1 = [&](const double *f, size_t c) {
    return teuk_pt::d1(f[c + sN], f[c - sN], f[c + 2 * sN], f[c - 2 * sN], f[c + 3 * sN], f[c - 3 * sN], wy);
  };
  auto dy2 = [&](const double *f, size_t c) {
    return teuk_pt::d2(f[c], f[c + sN], f[c - sN], f[c + 2 * sN], f[c - 2 * sN], f[c + 3 * sN], f[c - 3 * sN],
                       wyy);
  };

  const int nth = p->nthreads > 1 ? p->nthreads : 1;
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
  for (int b = 3; b < M - 3; b++) {
    const double ctan = std::tan(0.5 * kPi15 - theta[b]);
    for (int a = 3; a < N - 3; a++) {
      const size_t c = (size_t)b * N + a;
      const double dbdx = dx1(bb, c);
      teuk_pt::Coef cf;
      cf.bb = bb[c]; cf.cc = cc[c]; cf.ee = ee[c]; cf.are = are[c]; cf.aim = aim[c];
      cf.dre = dre_c[c]; cf.dim = dim[c]; cf.vre = vre[c]; cf.vim = vim[c];
      const double dre = teuk_pt::fma_(-dbdx, cf.ee, cf.dre);
      teuk_pt::Pt x;
      x.qr = qr[c]; x.qi = qi[c]; x.pr = pr[c]; x.pi = pi[c];
      x.llr = teuk_pt::fma_(ctan, dy1(qr, c), dy2(qr, c));
      x.lli = teuk_pt::fma_(ctan, dy1(qi, c), dy2(qi, c));
      x.qrx = dx1(qr, c); x.qix = dx1(qi, c);
      x.prx = dx1(pr, c); x.pix = dx1(pi, c);
      x.kqr = x.kqi = x.kpr = x.kpi = 0.0;
      if (a >= 4 && a < N - 4) {
        x.kqr = ko8(qr, c) * kof; x.kqi = ko8(qi, c) * kof;
        x.kpr = ko8(pr, c) * kof; x.kpi = ko8(pi, c) * kof;
      }
      x.sr = src ? src[c] : 0.0;
      x.si = src ? src[NM + c] : 0.0;
      double r[4];
      teuk_pt::rhs(cf, dre, x, r);
      rhs[c] = r[0];
      rhs[NM + c] = r[1];
      rhs[2 * NM + c] = r[2];
      rhs[3 * NM + c] = r[3];
    }
  }
}

/* ------------------------------------------------------------ boundaries --- */
/* kern.cu:320-409: theta edge rows for columns 3 <= a < N-3.  Even m:
 * quadratic extrapolation chained inward-out; odd m: the six rows vanish. */
void orc_bc_theta(const orc_params *p, double *u) {
  const int N = p->N, M = p->M;
  const size_t NM = plane(p);
  const bool even = (((int)p->m) % 2) == 0;
  for (int f = 0; f < 4; f++) {
    double *g = u + f * NM;
    auto at = [&](int b, int a) -> double & { return g[(size_t)b * N + a]; };
    for (int a = 3; a < N - 3; a++) {
      if (even) {
        at(2, a) = (4.0 * at(3, a) - at(4, a)) * kThird;
        at(1, a) = (4.0 * at(2, a) - at(3, a)) * kThird;
        at(M - 3, a) = (4.0 * at(M - 4, a) - at(M - 5, a)) * kThird;
        at(M - 2, a) = (4.0 * at(M - 3, a) - at(M - 4, a)) * kThird;
        at(0, a) = (4.0 * at(1, a) - at(2, a)) * kThird;
        at(M - 1, a) = (4.0 * at(M - 2, a) - at(M - 3, a)) * kThird;
      } else {
        at(0, a) = 0.0; at(M - 1, a) = 0.0; at(1, a) = 0.0;
        at(2, a) = 0.0; at(M - 2, a) = 0.0; at(M - 3, a) = 0.0;
      }
    }
  }
}

/* kern.cu:412-449: rho edge columns on every row, after the theta rows. */
void orc_bc_rho(const orc_params *p, double *u) {
  const int N = p->N, M = p->M;
  const size_t NM = plane(p);
  for (int f = 0; f < 4; f++) {
    double *g = u + f * NM;
    for (int b = 0; b < M; b++) {
      double *r = g + (size_t)b * N;
      r[2] = (4.0 * r[3] - r[4]) * kThird;
      r[N - 3] = (4.0 * r[N - 4] - r[N - 5]) * kThird;
      r[1] = (4.0 * r[2] - r[3]) * kThird;
      r[N - 2] = (4.0 * r[N - 3] - r[N - 4]) * kThird;
      r[0] = (4.0 * r[1] - r[2]) * kThird;
      r[N - 1] = (4.0 * r[N - 2] - r[N - 3]) * kThird;
    }
  }
}

/* ------------------------------------------------------------------ step --- */
/* body.cu:131-165 step order with kern.cu:301-318 (u1), :582-600 (u2),
 * :602-620 (u3, in place).  Updates touch 3<=b<M-3, 3<=a<N-3 only; the
 * boundary fills then define every other point.
 * RK4 (classic, same RHS, source frozen over the step like the reference)
 * is the alternative integrator, written with its stage values (no
 * accumulator).  work: 16 planes. */
void orc_step(const orc_params *p, const double *coef, const double *theta,
              double *u, const double *src, double *work) {
  const int N = p->N, M = p->M;
  const size_t NM = plane(p);
  const double dt = p->dt;
  double *uc = work;             /* 4 planes (RK4: u1, then u3) */
  double *rhs = work + 4 * NM;   /* 4 planes */
  double *u2 = work + 8 * NM;    /* 4 planes (RK4) */
  double *w = work + 12 * NM;    /* 4 planes (RK4) */
  const int nth = p->nthreads > 1 ? p->nthreads : 1;
  auto interior = [&](auto &&fn) {
#pragma omp parallel for collapse(2) schedule(static) num_threads(nth) if (nth > 1)
    for (int f = 0; f < 4; f++)
      for (int b = 3; b < M - 3; b++)
        for (int a = 3; a < N - 3; a++) fn(f * NM + (size_t)b * N + a);
  };
  if (p->integrator == 0) {
    orc_rhs(p, coef, theta, u, src, rhs);
    interior([&](size_t c) { uc[c] = teuk_pt::rk3_s1(u[c], dt, rhs[c]); });
    orc_bc_theta(p, uc); orc_bc_rho(p, uc);
    orc_rhs(p, coef, theta, uc, src, rhs);
    interior([&](size_t c) { uc[c] = teuk_pt::rk3_s2(u[c], uc[c], dt, rhs[c]); });
    orc_bc_theta(p, uc); orc_bc_rho(p, uc);
    orc_rhs(p, coef, theta, uc, src, rhs);
    interior([&](size_t c) { u[c] = teuk_pt::rk3_s3(u[c], uc[c], dt, rhs[c], kThird); });
    orc_bc_theta(p, u); orc_bc_rho(p, u);
  } else {
    /* from the stage values: u1, u2 with w = (u1 + 2 u2 - u)/3, u3, then
     * u + dt/6 (L(u) + 2 L(u1) + 2 L(u2) + L(u3)) = w + u3/3 + dt/6 L(u3) */
    double *u3 = uc;               /* u1 is dead once w holds it */
    const double h2 = 0.5 * dt, h6 = dt / 6.0;
    orc_rhs(p, coef, theta, u, src, rhs);
    interior([&](size_t c) { uc[c] = teuk_pt::axpy(h2, rhs[c], u[c]); });
    orc_bc_theta(p, uc); orc_bc_rho(p, uc);
    orc_rhs(p, coef, theta, uc, src, rhs);
    interior([&](size_t c) {
      u2[c] = teuk_pt::axpy(h2, rhs[c], u[c]);
      w[c] = teuk_pt::rk4_w(u[c], uc[c], u2[c]);
    });
    orc_bc_theta(p, u2); orc_bc_rho(p, u2);
    orc_rhs(p, coef, theta, u2, src, rhs);
    interior([&](size_t c) { u3[c] = teuk_pt::axpy(dt, rhs[c], u[c]); });
    orc_bc_theta(p, u3); orc_bc_rho(p, u3);
    orc_rhs(p, coef, theta, u3, src, rhs);
    interior([&](size_t c) { u[c] = teuk_pt::rk4_last(w[c], u3[c], h6, rhs[c]); });
    orc_bc_theta(p, u); orc_bc_rho(p, u);
  }
}

/* ---------------------------------------------------------- projection --- */
/* proj.c:8-14 */
double orc_fact(double n) {
  double f = 1.0;
  for (double c = 1.0; c <= n; c++) f = f * c;
  return f;
}

/* proj.c:16-35: spin-weight -2 spherical harmonic by the Wigner sum. */
double orc_swsh(double ell, double ml, double the) {
  double lo = 0.0;
  if ((ml - 2.0) > 0) lo = ml - 2.0;
  double hi = ell - 2.0;
  if ((ell - 2.0) > (ell + ml)) hi = ml + ell;
  double sum = 0.0;
  for (double t = lo; t <= hi; t++) {
    double w = std::pow(-1.0, t) * std::sqrt(orc_fact(ell + ml) * orc_fact(ell - ml)
                                             * orc_fact(ell + 2.0) * orc_fact(ell - 2.0));
    w = w / (orc_fact(ell + ml - t) * orc_fact(ell - 2.0 - t) * orc_fact(t) * orc_fact(t + 2.0 - ml));
    sum = sum + w * std::pow(std::cos(the / 2.0), 2.0 * ell + ml - 2.0 - 2.0 * t)
                  * std::pow(std::sin(the / 2.0), 2.0 * t + 2.0 - ml);
  }
  return std::sqrt((2.0 * ell + 1.0) / 4.0 / kPi15) * sum;
}

/* proj.c:37-75: composite Boole rule over M/4 panels of the theta grid. */
double orc_project(double *qq, int M, double l, double m) {
  const double dth = kPi15 / M;
  static const double w[5] = {7.0, 32.0, 12.0, 32.0, 7.0};
  double h = 0.0;
  qq[M] = qq[M - 1];
  for (int i = 0; i < M / 4; i++) {
    double th = 4.0 * dth * (double)i;
    for (int j = 0; j < 5; j++) {
      if (j > 0) th += dth;
      h += w[j] * 4.0 * kPi15 * qq[4 * i + j] * std::sin(th) * dth * orc_swsh(l, m, th);
    }
  }
  return 2.0 * h / 45.0;
}

/* ------------------------------------------------------------------- run --- */
/* body.cu:105-215 without file I/O: source, step, extraction of the column,
 * rectangle-rule double time integral, energy flux and the 14 projections. */
int orc_run(const orc_params *p, const double *r_c, const double *theta,
            double *u, int nsteps, const orc_src_step *src_steps,
            double *psi_hist, double *h_hist, double *dedt_hist) {
  const int N = p->N, M = p->M;
  const size_t NM = plane(p);
  std::vector<double> coef(9 * NM), work(16 * NM);
  std::vector<double> src;
  if (src_steps) src.assign(2 * NM, 0.0);
  orc_coeffs(p, theta, r_c, coef.data());
  const int je = extract_col(p);
  std::vector<double> ire(M, 0.0), iim(M, 0.0), dre(M + 1, 0.0), dim(M + 1, 0.0);
  const double dt = p->dt;
  const int am = std::abs((int)p->m);
  for (int k = 0; k < nsteps; k++) {
    if (src_steps) orc_source(p, theta, r_c, &src_steps[k], src.data());
    orc_step(p, coef.data(), theta, u, src_steps ? src.data() : nullptr, work.data());
    const double rr = r_c[je];
    double dedt = 0.0;
    for (int q = 0; q < M; q++) {
      const double pr = u[(size_t)q * N + je], pi = u[NM + (size_t)q * N + je];
      if (psi_hist) {
        psi_hist[(size_t)k * 2 * M + q] = pr;
        psi_hist[(size_t)k * 2 * M + M + q] = pi;
      }
      ire[q] = ire[q] + pr * dt;
      iim[q] = iim[q] + pi * dt;
      const double psi2 = ire[q] * ire[q] + iim[q] * iim[q];
      const double rho2 = rr * rr + std::pow(p->a * std::cos(theta[q]), 2);
      const double dvol = 0.5 * p->dtheta * std::sin(theta[q]) * rr * rr;
      dedt = dedt + psi2 * dvol * std::pow(rr, 6) / std::pow(rho2, 4);
      dre[q] = dre[q] + ire[q] * dt;
      dim[q] = dim[q] + iim[q] * dt;
    }
    if (dedt_hist) dedt_hist[k] = dedt;
    if (h_hist) {
      for (int l = 0; l < 7; l++) {
        h_hist[(size_t)k * 14 + 2 * l] = orc_project(dre.data(), M, (double)(l + am), p->m);
        h_hist[(size_t)k * 14 + 2 * l + 1] = -orc_project(dim.data(), M, (double)(l + am), p->m);
      }
    }
    if (!std::isfinite(u[(size_t)(M / 2) * N + N / 2])) return -4;
  }
  return 0;
}

}  // extern "C"

#include "teuk_source.inc"
#include "teuk_oracle_so2.inc"
