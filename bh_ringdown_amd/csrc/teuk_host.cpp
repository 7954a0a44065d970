// Host-side set-up for libteuk: the compactified grid with its Newton
// inversion r(r*), the -2Y_lm projection weights and the trajectory
// consumption pattern of the reference driver.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "teuk_internal.h"

namespace teuk {

// TD_GPU_code/HighOrder/main.c:79-198 from the rho grid on: Omega(rho) = 1
// inside rin = Xmin + 0.64 (Xmax - Xmin) and 1 - ((rho - rin)/(Xmax + 1e-6 -
// rin))^4 beyond, r* = rho/Omega, then Newton on r*(r) - r* = 0 with
// dr*/dr = (r^2+a^2)/Delta, marching inward from the outer edge with each
// point warm-started from its neighbour.
static void host_radii(int N, const double *x, double xmin, double xmax, double mass, double a, double *r_out) {
  const double rin = xmin + (xmax - xmin) * 0.64;
  const double outer = xmax + 1e-6;
  std::vector<double> target(N);
  for (int l = 0; l < N; l++) {
    const double om = (x[l] < rin) ? 1.0 : 1.0 - std::pow((x[l] - rin) / (outer - rin), 4);
    target[l] = x[l] / om;
  }
  if (mass == 0.0) {
    for (int l = 0; l < N; l++) r_out[l] = target[l];
    return;
  }
  const double disc = std::sqrt(mass * mass - a * a);
  const double rp = mass + disc, rm = mass - disc;
  const double tol = 1e-12;
  double r = target[N - 1];
  for (int l = N - 1; l >= 0; l--) {
    double step_rel, r_new;
    do {
      const double jac = (r * r + a * a - 2.0 * mass * r) / (r * r + a * a);
      double rstar;
      if (mass == a)
        rstar = r + 2.0 * a * std::log(r / a - 1.0) - 2.0 * a * a / (r - a);
      else if (a == 0.0)
        rstar = r + 2.0 * mass * std::log(r / (2.0 * mass) - 1.0);
      else
        rstar = r + (rp * rp + a * a) / (rp - rm) * std::log(r / rp - 1.0) -
                (rm * rm + a * a) / (rp - rm) * std::log(r / rm - 1.0);
      r_new = r + (target[l] - rstar) * jac;
      // Past double resolution of r - r+ the reference iteration hits
      // log(0) and returns NaN for every inner point; clamp to r+ (the
      // r* -> -inf limit) there.  Finite reference results are unchanged.
      if (!(r_new > rp)) {
        r_new = rp;
        r = r_new;
        break;
      }
      step_rel = std::fabs(r_new - r) / r;
      r = r_new;
    } while (step_rel > tol);
    r_out[l] = r_new;
  }
}

// main.c:79-198: rho_l = Xmin + l dx, theta_k = k dtheta and the radii.
void host_grid(int N, int M, double xmin, double xmax, double mass, double a, double *x_c, double *r_c,
               double *theta) {
  const double dx = (xmax - xmin) / N;
  const double dth = kPi15 / M;
  for (int l = 0; l < N; l++) x_c[l] = xmin + (double)l * dx;
  for (int k = 0; k < M; k++) theta[k] = (double)k * dth;
  host_radii(N, x_c, xmin, xmax, mass, a, r_c);
}

// The staggered scheme's half points (TD_GPU_code/Old/kern.cu:93 places
// its half-point coefficients at Xmin + (j - 1/2) dx with radius r_h[j]; Old/
// has no main.c, so r_h is the same inversion at those points).
void host_grid_half(int N, double xmin, double xmax, double mass, double a, double *x_h, double *r_h) {
  const double dx = (xmax - xmin) / N;
  for (int l = 0; l < N; l++) x_h[l] = xmin + ((double)l - 0.5) * dx;
  host_radii(N, x_h, xmin, xmax, mass, a, r_h);
}

static double factorial(double n) {
  double f = 1.0;
  for (double c = 1.0; c <= n; c++) f = f * c;
  return f;
}

// proj.c:16-35: -2Y_lm(theta) (phi part dropped) from the Wigner d-matrix sum.
double host_swsh(double ell, double ml, double the) {
  const double tlo = (ml - 2.0) > 0 ? ml - 2.0 : 0.0;
  const double thi = (ell - 2.0) > (ell + ml) ? ml + ell : ell - 2.0;
  double acc = 0.0;
  for (double t = tlo; t <= thi; t++) {
    double c = std::pow(-1.0, t) *
               std::sqrt(factorial(ell + ml) * factorial(ell - ml) * factorial(ell + 2.0) * factorial(ell - 2.0));
    c = c / (factorial(ell + ml - t) * factorial(ell - 2.0 - t) * factorial(t) * factorial(t + 2.0 - ml));
    acc = acc + c * std::pow(std::cos(the / 2.0), 2.0 * ell + ml - 2.0 - 2.0 * t) *
                    std::pow(std::sin(the / 2.0), 2.0 * t + 2.0 - ml);
  }
  return std::sqrt((2.0 * ell + 1.0) / 4.0 / kPi15) * acc;
}

// Factors of the composite Boole rule of proj.c:47-72 for l = |m| .. |m|+6.
// The reference sum has 5 terms per panel (panel joints appear twice, at
// slightly different accumulated theta), so the table holds one entry per
// term t = 5 i + j, evaluated at grid point 4 i + j (point M repeats M-1,
// proj.c:45): (w_j*4*Pi, sin(theta_t), -2Y_lm(theta_t)), with theta_t
// accumulated exactly as the reference does (4 dtheta i, then += dtheta).
void host_projection_weights(int M, double m, double *w) {
  const double dth = kPi15 / M;
  static const double boole[5] = {7.0, 32.0, 12.0, 32.0, 7.0};
  const int am = std::abs((int)m);
  const int nterm = 5 * (M / 4);
  for (int l = 0; l < 7; l++) {
    double *wl = w + (size_t)l * nterm * 3;
    for (int i = 0; i < M / 4; i++) {
      double th = 4.0 * dth * (double)i;
      for (int j = 0; j < 5; j++) {
        if (j > 0) th += dth;
        double *e = wl + 3 * (5 * i + j);
        e[0] = boole[j] * 4.0 * kPi15;
        e[1] = std::sin(th);
        e[2] = host_swsh((double)(l + am), m, th);
      }
    }
  }
}

// Coefficient fields of kernel_init (kern.cu:110-158) split into what depends
// on the rho column only and what depends on the theta row only.  The layer
// function Omega(rho) and its derivatives (kern.cu:119-127), the compactified
// radius eR = Omega r and the Kerr factors are folded into 12 constants per
// column; the device then needs one division and one square root per point.
// Evaluated with the host libm so that the device assembly reproduces the
// host computation bit for bit.
void host_coeff_tables(int N, int M, double xmin, double xmax, double dx, double mass, double a,
                       double m, const double *r, const double *theta, double *ktab, double *rtab, double shift,
                       bool staggered) {
  const double rin = xmin + (xmax - xmin) * 0.64;
  const double outer = xmax + 1e-6;
  const double spin = a, em = m;
  for (int j = 0; j < N; j++) {
    const double x = xmin + ((double)j + shift) * dx;
    double om = 1.0, om1 = 0.0, om2 = 0.0;
    if (!(x < rin)) {
      om = 1.0 - std::pow((x - rin) / (outer - rin), 4);
      om1 = -4.0 * std::pow(x - rin, 3) / std::pow(outer - rin, 4);
      om2 = -12.0 * std::pow(x - rin, 2) / std::pow(outer - rin, 4);
    }
    const double L = om - x * om1;  // Omega^2 d(r*)/d(rho)
    const double R = om * r[j];     // eR of kern.cu:130
    const double a2 = spin * spin, a4 = a2 * a2;
    const double om_2 = om * om, om3 = om_2 * om, om4 = om_2 * om_2, om5 = om4 * om;
    const double S = R * R + om_2 * a2;
    const double D = R * R - 2.0 * R * mass * om + om_2 * a2;
    const double L2 = L * L, L3 = L2 * L;
    const double R2 = R * R, R3 = R2 * R, R5 = R3 * R2;
    double K[kNK];
    K[kK1] = ((2.0 * L - om_2) * (S * S)) / L2;
    K[kK2] = a2 * D;
    K[kKA] = (2.0 * (L - om_2) * (S * S)) / L2;
    K[kKR] = -((om_2 * (S * S)) / L2);
    K[kKD] = D;
    K[kKBt] = -((om * (R3 * (4.0 * L2 * (R - 3.0 * mass * om) + 2.0 * L * R2 * om1 + R2 * om * om2 * x) +
                       2.0 * R *
                           (L2 * om_2 * (5.0 * R - 4.0 * mass * om) - 3.0 * L3 * (R - 2.0 * mass * om) +
                            2.0 * L * R2 * om_2 * om1 + R2 * om3 * om2 * x) *
                           a2 +
                       (-6.0 * L3 * om_2 + 6.0 * L2 * om4 + 2.0 * L * R * om4 * om1 + R * om5 * om2 * x) * a4)) /
                (L3 * R));
    K[kKBr] = -((om * (R5 * (2.0 * om * om1 - 2.0 * om1 * om1 * x + om * om2 * x) +
                       2.0 * R3 * om_2 * (2.0 * om * om1 - 2.0 * om1 * om1 * x + om * om2 * x) * a2 +
                       om4 * (6.0 * om_2 + 2.0 * om1 * om1 * x * (-R + 3.0 * x) +
                              om * (2.0 * R * om1 - 12.0 * om1 * x + R * om2 * x)) *
                           a4 +
                       2.0 * L2 * R * (2.0 * R3 - 6.0 * R2 * mass * om + 5.0 * R * om_2 * a2 - 4.0 * mass * om3 * a2))) /
                (L3 * R));
    K[kKC0] = (D * (2.0 * R2 + 6.0 * R * mass * om - 6.0 * om_2 * a2)) / R2;
    K[kKI0] = (2.0 * spin * (em * (-(om_2 * S) + L * (R2 + 2.0 * R * mass * om + om_2 * a2)))) / L;
    K[kKI1] = 4.0 * spin * D;
    K[kKBi] = (-2.0 * em * om_2 * spin * S) / L;
    K[kKCi] = (2.0 * em * om * spin * (R2 + 2.0 * R * mass * om - 3.0 * om_2 * a2)) / R;
    for (int i = 0; i < kNK; i++) ktab[(size_t)i * N + j] = K[i];
  }
  for (int k = 0; k < M; k++) {
    const double y = theta[k];
    const double sn = std::sin(y), cs = std::cos(y);
    const double cot = 1.0 / std::tan(y), csc = 1.0 / std::sin(y);
    rtab[k] = sn * sn;
    rtab[M + k] = cs;
    rtab[2 * M + k] = 4.0 * (cot * cot) - 4.0 * em * cot * csc + (em * em) * (csc * csc);
    rtab[3 * M + k] = staggered ? cs / sn : std::tan(0.5 * kPi15 - y);  // Old/kern.cu:241-243, kern.cu:204
  }
}

// kern.cu:2671-2672: the source picks up e^{-i m a/(r+ - r-) ln((r - r+)/(r - r-))},
// evaluated at the grid radius of each column.
void host_source_phase(int N, double mass, double a, double m, const double *r_c, double *phase) {
  const double rpl = mass + std::sqrt(mass * mass - a * a);
  const double rmi = mass - std::sqrt(mass * mass - a * a);
  for (int j = 0; j < N; j++) {
    const double r = r_c[j];
    double arg = 0.0;
    if (a != 0.0) arg = -m * a / (rpl - rmi) * std::log((r - rpl) / (r - rmi));
    phase[j] = std::cos(arg);
    phase[N + j] = std::sin(arg);
  }
}

// kern.cu:1186-1206 and 2673 for one step: the patch origin from the
// particle's tortoise radius and theta, the radial Gaussian width
// wr = 2 dx (1 - r+/r_p), the turn-on (1 - e^{-(t - t0)/25})^4, and the
// transcendental inputs of the T_4 coefficients.  For a = 0 the tortoise
// radius uses its Schwarzschild form (the reference's Kerr form is 0 * inf
// there).
void host_source_prepare(const Grid &g, const SrcStep &st, SrcPrep &out) {
  const double a = g.a, mass = g.mass;
  const double rpl = mass + std::sqrt(mass * mass - a * a);
  const double rmi = mass - std::sqrt(mass * mass - a * a);
  double xp;
  if (a != 0.0)
    xp = st.rp + (rpl * rpl + a * a) / (rpl - rmi) * std::log(st.rp / rpl - 1.0) -
         (rmi * rmi + a * a) / (rpl - rmi) * std::log(st.rp / rmi - 1.0);
  else
    xp = st.rp + 2.0 * mass * std::log(st.rp / (2.0 * mass) - 1.0);
  const int ip = (int)((xp - g.xmin) / g.dx);
  const int it = (int)(st.tp / g.dtheta);
  SrcDev &d = out.dev;
  for (int i = 0; i < 12; i++) d.c[i] = 0.0;
  d.rp = st.rp;
  d.tp = st.tp;
  d.wr = (2.0 * g.dx) * (1.0 - rpl / st.rp);
  d.turnon = std::pow(1.0 - std::exp(-(st.timer - st.start) / 25.0), 4);
  d.ip0 = ip - g.ptsx / 2;
  d.it0 = it - g.ptsy / 2;
  teuk_t4::StepIn &in = out.in;
  in.mass = mass;
  in.a = a;
  in.E = st.E;
  in.lz = st.lz;
  in.nmu = 1.0;  // kern.cu:1190
  in.mm = g.m;
  in.rp = st.rp;
  in.drdt = st.drdt;
  in.d2rdt2 = st.d2rdt2;
  in.dthdt = st.dthdt;
  in.d2thdt2 = st.d2thdt2;
  in.dphidt = st.dphidt;
  in.d2phidt2 = st.d2phidt2;
  in.stp = std::sin(st.tp);
  in.ctp = std::cos(st.tp);
  in.cmphi = std::cos(g.m * st.phip);
  in.smphi = std::sin(g.m * st.phip);
}

}  // namespace teuk
