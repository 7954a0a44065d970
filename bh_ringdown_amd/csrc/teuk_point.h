// Arithmetic of one Runge-Kutta stage at one grid point: the nine Kerr
// coefficients from the factored column/row constants, the 6th-order
// stencils, Kreiss-Oliger dissipation, the right-hand side and the stage
// combinations (TD_GPU_code/HighOrder/kern.cu:132-158, 192-299, 301-318,
// 582-620).  Host and device compile this same code, so both give the same
// bits: every multiply-add the code wants fused is an explicit fma
// (correctly rounded on both sides), everything else is built without
// contraction.  TEUK_HD sets the function qualifiers.
#pragma once

#ifndef TEUK_HD
#define TEUK_HD inline
#endif

namespace teuk_pt {

TEUK_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Column constants of the coefficient fields (kern.cu:132-145 with the theta
// dependence factored out).  With bir = K2 sin^2 - K1 and ib = 1/bir:
//   A^tr = KA ib, A^rr = KR ib, A^yy = -KD ib, B^tau_re = KBt ib,
//   B^rho_re = KBr ib, C_re = (KC0 + KD Y) ib, B^tau_im = (KI0 - KI1 cos) ib,
//   B^rho_im = KBi ib, C_im = KCi ib,  Y = 4 cot^2 - 4 m cot csc + m^2 csc^2.
enum { kK1, kK2, kKA, kKR, kKD, kKBt, kKBr, kKC0, kKI0, kKI1, kKBi, kKCi, kNK };

// The nine fields at one point (kern.cu:147-158); d_re without the
// -dbdx ee term (dbdx needs neighbouring columns).
struct Coef {
  double bb, cc, ee, are, aim, dre, dim, vre, vim;
};

template <class SQRT>
TEUK_HD Coef coef_point(const double *K, double sn2, double cs, double Y, SQRT sqrt_rn) {
  const double bir = fma_(K[kK2], sn2, -K[kK1]);
  const double ib = 1.0 / bir;
  const double Atr = K[kKA] * ib, Arr = K[kKR] * ib;
  const double BtRe = K[kKBt] * ib, BrRe = K[kKBr] * ib;
  const double CRe = fma_(K[kKD], Y, K[kKC0]) * ib;
  const double BtIm = fma_(-K[kKI1], cs, K[kKI0]) * ib;
  Coef c;
  c.bb = -0.5 * (Atr + sqrt_rn(fma_(Atr, Atr, 4.0 * Arr)));
  c.cc = -(K[kKD] * ib);
  c.ee = Atr + c.bb;
  c.are = -BtRe;
  c.aim = -BtIm;
  c.dre = fma_(-c.bb, BtRe, BrRe);
  c.dim = fma_(-c.bb, BtIm, K[kKBi] * ib);
  c.vre = -CRe;
  c.vim = -(K[kKCi] * ib);
  return c;
}

// Centered 6th-order first derivative from the differences at distance 1, 2,
// 3 (kern.cu:235-244; w = 0.016666666666667/dx, or 0.05/3/dtheta in theta).
TEUK_HD double d1(double fp1, double fm1, double fp2, double fm2, double fp3, double fm3, double w) {
  return fma_(45.0, fp1 - fm1, fma_(-9.0, fp2 - fm2, fp3 - fm3)) * w;
}

// Centered 6th-order second derivative (kern.cu:213; w = 0.05/9/dtheta^2).
TEUK_HD double d2(double f0, double fp1, double fm1, double fp2, double fm2, double fp3, double fm3, double w) {
  double t = 2.0 * (fp3 + fm3);
  t = fma_(-27.0, fp2 + fm2, t);
  t = fma_(270.0, fp1 + fm1, t);
  t = fma_(-490.0, f0, t);
  return t * w;
}

// 8th difference of the Kreiss-Oliger term (kern.cu:247-256), before its
// factor ko_factor(dt).
TEUK_HD double ko8(double f0, double fp1, double fm1, double fp2, double fm2, double fp3, double fm3, double fp4,
                   double fm4) {
  double t = fma_(-8.0, fp3 + fm3, fp4 + fm4);
  t = fma_(70.0, f0, t);
  t = fma_(-56.0, fp1 + fm1, t);
  return fma_(28.0, fp2 + fm2, t);
}

// 0.1 x 1/16 / dt (kern.cu:258-261).
TEUK_HD double ko_factor(double dt) { return 0.1 * 0.0625 * (1.0 / dt); }

// Right-hand side at one point (kern.cu:262-296): q' = -bb q_x + p - KO,
// p' = src + ee p_x - KO + cc L q + d q_x - a p - v q with complex a, d, v;
// dre includes -dbdx ee.
struct Pt {
  double qr, qi, pr, pi;      // fields at the point
  double qrx, qix, prx, pix;  // rho derivatives
  double llr, lli;            // theta operator on q
  double kqr, kqi, kpr, kpi;  // dissipation terms (already scaled)
  double sr, si;              // source
};

TEUK_HD void rhs(const Coef &cf, double dre, const Pt &x, double r[4]) {
  r[0] = fma_(-cf.bb, x.qrx, x.pr) - x.kqr;
  r[1] = fma_(-cf.bb, x.qix, x.pi) - x.kqi;
  double t = fma_(cf.ee, x.prx, x.sr) - x.kpr;
  t = fma_(cf.cc, x.llr, t);
  t = fma_(dre, x.qrx, t);
  t = fma_(-cf.dim, x.qix, t);
  t = fma_(-cf.are, x.pr, t);
  t = fma_(cf.aim, x.pi, t);
  t = fma_(-cf.vre, x.qr, t);
  r[2] = fma_(cf.vim, x.qi, t);
  t = fma_(cf.ee, x.pix, x.si) - x.kpi;
  t = fma_(cf.cc, x.lli, t);
  t = fma_(dre, x.qix, t);
  t = fma_(cf.dim, x.qrx, t);
  t = fma_(-cf.are, x.pi, t);
  t = fma_(-cf.aim, x.pr, t);
  t = fma_(-cf.vre, x.qi, t);
  r[3] = fma_(-cf.vim, x.qr, t);
}

// SSP-RK3 stage combinations (kern.cu:301-318, 582-600, 602-620):
//   u1 = u + dt L(u);  u2 = (3 u + u1 + dt L(u1)) / 4;  u3 = (u + 2 u2 + 2 dt L(u2)) / 3.
TEUK_HD double rk3_s1(double u, double dt, double r) { return fma_(dt, r, u); }
TEUK_HD double rk3_s2(double base, double u, double dt, double r) {
  return 0.25 * fma_(dt, r, fma_(3.0, base, u));
}
TEUK_HD double rk3_s3(double base, double u, double dt, double r, double third) {
  return third * fma_(2.0 * dt, r, fma_(2.0, u, base));
}
// Classic RK4 from its stage values: u1 = u + dt/2 L(u), u2 = u + dt/2 L(u1),
// u3 = u + dt L(u2) (each an axpy), w = (u1 + 2 u2 - u)/3 formed with u2, and
//   u(t + dt) = w + u3/3 + dt/6 L(u3),
// which is u + dt/6 L(u) + dt/3 L(u1) + dt/3 L(u2) + dt/6 L(u3) without an
// accumulator array (and without u in the last stage).
TEUK_HD double axpy(double h, double r, double y) { return fma_(h, r, y); }
TEUK_HD double rk4_w(double u, double u1, double u2) { return (fma_(2.0, u2, u1) - u) * (1.0 / 3.0); }
TEUK_HD double rk4_last(double w, double u3, double h6, double r) {
  return fma_(h6, r, fma_(u3, 1.0 / 3.0, w));
}

// Staggered second-order scheme of TD_GPU_code/Old/kern.cu: fields averaged
// to the half points rho_{a-1/2} (kernel_average1, :196-212) and back
// (kernel_average2, :346-362), centred theta differences (:245-254,
// :395-404), one-cell rho differences (:260-263, :410-413), the 2nd-order
// dbdx (:231-233, :381-383), half and full updates (:297-300, :448-451) and
// the theta/rho edge extrapolation (:305-340, :456-522), which divides by 3.
TEUK_HD double so2_avg(double f0, double f1) { return 0.5 * (f0 + f1); }
TEUK_HD double so2_dth(double fp, double fm, double dth) { return (fp - fm) / (2.0 * dth); }
TEUK_HD double so2_dth2(double f0, double fp, double fm, double dth) {
  return (fp + fm - 2.0 * f0) / (dth * dth);
}
TEUK_HD double so2_dx(double f1, double f0, double dx) { return (f1 - f0) / dx; }
TEUK_HD double so2_dbdx(double bp, double bm, double dx) { return (bp - bm) / dx / 2.0; }
TEUK_HD double so2_edge(double f1, double f2) { return (4.0 * f1 - f2) / 3.0; }

}  // namespace teuk_pt
