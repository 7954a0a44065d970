// Point-particle source of the s = -2 Teukolsky equation, evaluated with
// truncated Taylor jets.  Host and device compile the same code; it uses
// only +, -, *, / (no libm), so both give the same bits when built without
// FMA contraction.  TEUK_HD sets the function qualifiers.
//
// What is computed (the physics of TD_GPU_code/HighOrder/kern.cu:770-2679,
// derived here from the published equations):
//
//   T_4 = (D + 3g - g* + 4mu + mu*)[(db - 2tau* + 2al) T_nmb - (D + 2g - 2g* + mu*) T_mbmb]
//       + (db - tau* + beta* + 3al + 4pi)[(D + 2g + 2mu*) T_nmb - (db - tau* + 2beta* + 2al) T_nn]
//
// (Teukolsky 1973, ApJ 185, 635, eq. 2.15; Kinnersley tetrad and spin
// coefficients of its eqs. 4.4-4.5, signature +---, D = n^a d_a,
// db = mbar^a d_a) for the stress tensor of a point mass
// T^ab = mu u^a u^b delta^3(x - z(t)) / (Sigma sin(theta) u^t), with the
// 4-velocity in the reference's representation (kern.cu:1258-1300):
// u_t = E, u_phi = -L_z, u_r = -(Sigma/Delta) r' u^t, u_theta = -Sigma theta' u^t,
// Sigma u^t = E[(r^2+a^2)^2/Delta - a^2 sin^2] + a L_z[1 - (r^2+a^2)/Delta],
// third time derivatives of r and theta zero (body.cu:124), and the azimuthal
// mode delta(phi - phi_p) -> e^{-i m phi_p}/pi, d_phi -> i m (kern.cu:1270).
//
// Normalisation: the reference's generated T_4 is HALF of that expression.
// Its delta''(r) delta(theta) coefficient for r' = theta' = 0 (kern.cu:2630
// t7 * t3198 / 2, with t3198 = -t2717 = t26 t2479 / 2 (kern.cu:2552, 2363,
// 2281), t2479 = t7 t2444 / 4 (:2280), t2444 = -t590 t147 t900 (:2262),
// t900 = t224 e^{-i m phi_p} / (pi sin), t224 = t51^2 / t38 (:1688, :1368))
// is -Delta^2 mu (mbar.u)^2 e^{-i m phi_p} / (8 pi Sigma^2 Sigma u^t sin),
// where -D^2 T_mbmb above gives twice that; and the reference's own waveform
// for its plunge example (QNM_Extraction/Examples/hm2_a0.7_thi000_thf090.0_n
// .dat) is half of what the undivided source produces on the same grid and
// trajectory, for every l and every time (ratio 1.98 at N = 16000, M = 256,
// 2.00 in the continuum limit).  So the coefficients carry a factor 1/2 with
// the 1/pi: a circular orbit then radiates half the quadrupole-formula
// amplitude, |h_22| -> 4 sqrt(pi/5) mu v^2.
//
// The result is reduced to the canonical form
//   T_4 = sum_{i+j<=2} c_ij delta^(i)(r - r_p) delta^(j)(theta - theta_p)
// with coefficients on the worldline, the form the reference evaluates with
// DelR, DelR1, DelR2 x DelTH, DelTH1, DelTH2 (kern.cu:1211-1227, 2597-2662).
//
// Method: every quantity is a complex polynomial in x = r - r_p,
// y = theta - theta_p and s = t - t_k truncated at total degree 2; T_4
// applies at most two derivatives to T_ab, and the canonical reduction of a
// delta^(i) delta^(j) term needs i + j further derivatives of a coefficient
// that has already lost the same number, so degree 2 is exact throughout.
// A time derivative of delta(r - r_p(t)) is -r_p'(t) delta'(r - r_p(t)).
#pragma once

#ifndef TEUK_HD
#define TEUK_HD inline
#endif

namespace teuk_t4 {

// Monomials: 1, x, y, s, xx, xy, xs, yy, ys, ss.
enum { k1, kx, ky, ks, kxx, kxy, kxs, kyy, kys, kss, kN };

struct J {
  double re[kN], im[kN];
};

TEUK_HD J jzero() {
  J a;
  for (int i = 0; i < kN; i++) a.re[i] = a.im[i] = 0.0;
  return a;
}
TEUK_HD J jc(double re, double im = 0.0) {
  J a = jzero();
  a.re[k1] = re;
  a.im[k1] = im;
  return a;
}
TEUK_HD J jadd(const J &a, const J &b) {
  J c;
  for (int i = 0; i < kN; i++) {
    c.re[i] = a.re[i] + b.re[i];
    c.im[i] = a.im[i] + b.im[i];
  }
  return c;
}
TEUK_HD J jsub(const J &a, const J &b) {
  J c;
  for (int i = 0; i < kN; i++) {
    c.re[i] = a.re[i] - b.re[i];
    c.im[i] = a.im[i] - b.im[i];
  }
  return c;
}
// complex scalar times jet
TEUK_HD J jscale(const J &a, double sr, double si = 0.0) {
  J c;
  for (int i = 0; i < kN; i++) {
    c.re[i] = sr * a.re[i] - si * a.im[i];
    c.im[i] = sr * a.im[i] + si * a.re[i];
  }
  return c;
}

// Product table: (i, j) -> k for deg(i) + deg(j) <= 2, in a fixed order.
struct MulTerm {
  unsigned char i, j, k;
};
constexpr MulTerm kMul[] = {
    {k1, k1, k1},
    {k1, kx, kx},   {kx, k1, kx},   {k1, ky, ky},   {ky, k1, ky},   {k1, ks, ks},   {ks, k1, ks},
    {k1, kxx, kxx}, {kx, kx, kxx},  {kxx, k1, kxx},
    {k1, kxy, kxy}, {kx, ky, kxy},  {ky, kx, kxy},  {kxy, k1, kxy},
    {k1, kxs, kxs}, {kx, ks, kxs},  {ks, kx, kxs},  {kxs, k1, kxs},
    {k1, kyy, kyy}, {ky, ky, kyy},  {kyy, k1, kyy},
    {k1, kys, kys}, {ky, ks, kys},  {ks, ky, kys},  {kys, k1, kys},
    {k1, kss, kss}, {ks, ks, kss},  {kss, k1, kss},
};

TEUK_HD J jmul(const J &a, const J &b) {
  J c = jzero();
  for (unsigned t = 0; t < sizeof(kMul) / sizeof(kMul[0]); t++) {
    const int i = kMul[t].i, j = kMul[t].j, k = kMul[t].k;
    c.re[k] = c.re[k] + (a.re[i] * b.re[j] - a.im[i] * b.im[j]);
    c.im[k] = c.im[k] + (a.re[i] * b.im[j] + a.im[i] * b.re[j]);
  }
  return c;
}

// 1/a = (1/a0) (1 - g + g^2) with g = (a - a0)/a0.
TEUK_HD J jinv(const J &a) {
  const double d = a.re[k1] * a.re[k1] + a.im[k1] * a.im[k1];
  const double ir = a.re[k1] / d, ii = -a.im[k1] / d;
  J g = jscale(a, ir, ii);
  g.re[k1] = 0.0;
  g.im[k1] = 0.0;
  J r = jsub(jadd(jc(1.0), jmul(g, g)), g);
  return jscale(r, ir, ii);
}

// Partial derivatives; the result is exact to degree 1.
TEUK_HD J jdx(const J &a) {
  J c = jzero();
  for (int p = 0; p < 2; p++) {
    const double *s = p ? a.im : a.re;
    double *d = p ? c.im : c.re;
    d[k1] = s[kx];
    d[kx] = 2.0 * s[kxx];
    d[ky] = s[kxy];
    d[ks] = s[kxs];
  }
  return c;
}
TEUK_HD J jdy(const J &a) {
  J c = jzero();
  for (int p = 0; p < 2; p++) {
    const double *s = p ? a.im : a.re;
    double *d = p ? c.im : c.re;
    d[k1] = s[ky];
    d[kx] = s[kxy];
    d[ky] = 2.0 * s[kyy];
    d[ks] = s[kys];
  }
  return c;
}
TEUK_HD J jds(const J &a) {
  J c = jzero();
  for (int p = 0; p < 2; p++) {
    const double *s = p ? a.im : a.re;
    double *d = p ? c.im : c.re;
    d[k1] = s[ks];
    d[kx] = s[kxs];
    d[ky] = s[kys];
    d[ks] = 2.0 * s[kss];
  }
  return c;
}

// Distribution sum_{i+j<=2} G_ij delta^(i)(x - X(s)) delta^(j)(y - Y(s)).
enum { d00, d10, d01, d20, d11, d02, dN };
TEUK_HD int didx(int i, int j) {
  return (i == 0) ? (j == 0 ? d00 : (j == 1 ? d01 : d02)) : (i == 1 ? (j == 0 ? d10 : d11) : d20);
}
struct Dist {
  J g[dN];
};
constexpr unsigned char kDI[dN] = {0, 1, 0, 2, 1, 0};
constexpr unsigned char kDJ[dN] = {0, 0, 1, 0, 1, 2};

TEUK_HD Dist dzero() {
  Dist d;
  for (int t = 0; t < dN; t++) d.g[t] = jzero();
  return d;
}
TEUK_HD Dist dadd(const Dist &a, const Dist &b) {
  Dist d;
  for (int t = 0; t < dN; t++) d.g[t] = jadd(a.g[t], b.g[t]);
  return d;
}
TEUK_HD Dist dsub(const Dist &a, const Dist &b) {
  Dist d;
  for (int t = 0; t < dN; t++) d.g[t] = jsub(a.g[t], b.g[t]);
  return d;
}
TEUK_HD Dist dmul(const Dist &a, const J &f) {
  Dist d;
  for (int t = 0; t < dN; t++) d.g[t] = jmul(a.g[t], f);
  return d;
}
// d/dr: derivative of the coefficient plus one more derivative on delta(x).
TEUK_HD Dist ddr(const Dist &a) {
  Dist d = dzero();
  for (int t = 0; t < dN; t++) {
    d.g[t] = jadd(d.g[t], jdx(a.g[t]));
    if (kDI[t] + kDJ[t] < 2) {
      const int u = didx(kDI[t] + 1, kDJ[t]);
      d.g[u] = jadd(d.g[u], a.g[t]);
    }
  }
  return d;
}
TEUK_HD Dist ddth(const Dist &a) {
  Dist d = dzero();
  for (int t = 0; t < dN; t++) {
    d.g[t] = jadd(d.g[t], jdy(a.g[t]));
    if (kDI[t] + kDJ[t] < 2) {
      const int u = didx(kDI[t], kDJ[t] + 1);
      d.g[u] = jadd(d.g[u], a.g[t]);
    }
  }
  return d;
}
// d/dt with the deltas riding on the worldline: Xd = dX/ds, Yd = dY/ds.
TEUK_HD Dist ddt(const Dist &a, const J &Xd, const J &Yd) {
  Dist d = dzero();
  for (int t = 0; t < dN; t++) {
    d.g[t] = jadd(d.g[t], jds(a.g[t]));
    if (kDI[t] + kDJ[t] < 2) {
      const int ux = didx(kDI[t] + 1, kDJ[t]), uy = didx(kDI[t], kDJ[t] + 1);
      d.g[ux] = jsub(d.g[ux], jmul(Xd, a.g[t]));
      d.g[uy] = jsub(d.g[uy], jmul(Yd, a.g[t]));
    }
  }
  return d;
}

// Inputs of one time step.
struct StepIn {
  double mass, a, E, lz, nmu, mm;
  double rp, drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2;
  double stp, ctp;        // sin, cos of theta_p
  double cmphi, smphi;    // cos, sin of m phi_p
};

// Canonical coefficients times e^{-i m phi_p}/(2 pi) (the reference's
// normalisation, see the top): out[2*t], out[2*t+1] for
// (i, j) = (0,0) (1,0) (0,1) (2,0) (1,1) (0,2).
TEUK_HD void t4_coeffs(const StepIn &p, double *out) {
  const double M = p.mass, a = p.a, E = p.E, Lz = p.lz, em = p.mm;
  const double rs2 = 1.4142135623730951;  // sqrt(2)
  // field point r = r_p + x, theta = theta_p + y
  J r = jc(p.rp);
  r.re[kx] = 1.0;
  J sn = jc(p.stp), cs = jc(p.ctp);
  sn.re[ky] = p.ctp;
  sn.re[kyy] = -0.5 * p.stp;
  cs.re[ky] = -p.stp;
  cs.re[kyy] = -0.5 * p.ctp;
  const J r2 = jmul(r, r);
  const J a2 = jc(a * a);
  const J Sig = jadd(r2, jmul(a2, jmul(cs, cs)));
  const J Dlt = jadd(jsub(r2, jscale(r, 2.0 * M)), a2);
  const J acs = jscale(cs, a);
  const J rmi = jsub(r, jscale(acs, 0.0, 1.0));   // r - i a cos
  const J rpi = jadd(r, jscale(acs, 0.0, 1.0));   // r + i a cos
  const J rho = jscale(jinv(rmi), -1.0);
  const J rhob = jscale(jinv(rpi), -1.0);
  const J isig2 = jscale(jinv(Sig), 0.5);
  const J izb = jscale(jinv(rmi), 1.0 / rs2);     // 1/(sqrt2 (r - i a cos))
  const J isn = jinv(sn);
  const J cot = jmul(cs, isn);
  // tetrad
  const J nT = jmul(jadd(r2, a2), isig2);
  const J nR = jscale(jmul(Dlt, isig2), -1.0);
  const J nP = jscale(isig2, a);
  const J bT = jscale(jmul(sn, izb), 0.0, -a);
  const J bH = izb;
  const J bP = jscale(jmul(isn, izb), 0.0, -1.0);
  const J w = jinv(jmul(Sig, sn));
  // spin coefficients
  const J rr_ = jmul(rho, rhob);
  const J betab = jscale(jmul(rho, cot), -1.0 / (2.0 * rs2));
  const J pi_ = jscale(jmul(jmul(rho, rho), sn), 0.0, a / rs2);
  const J taub = jscale(jmul(rr_, sn), 0.0, a / rs2);
  const J mu_ = jscale(jmul(jmul(jmul(rho, rho), rhob), Dlt), 0.5);
  const J mub = jscale(jmul(jmul(jmul(rhob, rhob), rho), Dlt), 0.5);
  const J rm = jmul(rr_, jsub(r, jc(M)));
  const J gam = jadd(mu_, jscale(rm, 0.5));
  const J gamb = jadd(mub, jscale(rm, 0.5));
  const J alp = jsub(pi_, betab);
  const J s1 = jadd(jscale(taub, -2.0), jscale(alp, 2.0));
  const J s2 = jadd(jsub(jscale(gam, 2.0), jscale(gamb, 2.0)), mub);
  const J s3 = jadd(jscale(gam, 2.0), jscale(mub, 2.0));
  const J s4 = jadd(jadd(jscale(taub, -1.0), jscale(betab, 2.0)), jscale(alp, 2.0));
  const J s5 = jadd(jadd(jsub(jscale(gam, 3.0), gamb), jscale(mu_, 4.0)), mub);
  const J s6 = jadd(jadd(jadd(jscale(taub, -1.0), betab), jscale(alp, 3.0)), jscale(pi_, 4.0));
  // worldline in s = t - t_k
  J R = jc(p.rp), Rd = jc(p.drdt), Yd = jc(p.dthdt), dth = jzero();
  R.re[ks] = p.drdt;
  R.re[kss] = 0.5 * p.d2rdt2;
  Rd.re[ks] = p.d2rdt2;
  Yd.re[ks] = p.d2thdt2;
  dth.re[ks] = p.dthdt;
  dth.re[kss] = 0.5 * p.d2thdt2;
  const J dth2h = jscale(jmul(dth, dth), 0.5);
  const J snP = jadd(jscale(jsub(jc(1.0), dth2h), p.stp), jscale(dth, p.ctp));
  const J csP = jsub(jscale(jsub(jc(1.0), dth2h), p.ctp), jscale(dth, p.stp));
  const J R2 = jmul(R, R);
  const J Sp = jadd(R2, jmul(a2, jmul(csP, csP)));
  const J iDp = jinv(jadd(jsub(R2, jscale(R, 2.0 * M)), a2));
  const J R2a2 = jadd(R2, a2);
  const J U = jadd(jscale(jsub(jmul(jmul(R2a2, R2a2), iDp), jmul(a2, jmul(snP, snP))), E),
                   jscale(jsub(jc(1.0), jmul(R2a2, iDp)), a * Lz));
  const J u_r = jscale(jmul(jmul(Rd, U), iDp), -1.0);
  const J u_h = jscale(jmul(Yd, U), -1.0);
  // e^{-i m (phi(s) - phi_p)} to second order
  J z = jzero();
  z.im[ks] = -em * p.dphidt;
  z.im[kss] = -em * 0.5 * p.d2phidt2;
  const J mode = jadd(jadd(jc(1.0), z), jscale(jmul(z, z), 0.5));
  const J norm = jmul(jmul(jscale(Sp, p.nmu), jinv(U)), jmul(w, mode));
  const J nu = jadd(jadd(jscale(nT, E), jmul(nR, u_r)), jscale(nP, -Lz));
  const J mbu = jadd(jadd(jscale(bT, E), jmul(bH, u_h)), jscale(bP, -Lz));

  Dist Tnn = dzero(), Tnm = dzero(), Tmm = dzero();
  Tnn.g[d00] = jmul(norm, jmul(nu, nu));
  Tnm.g[d00] = jmul(norm, jmul(nu, mbu));
  Tmm.g[d00] = jmul(norm, jmul(mbu, mbu));
  auto Dop = [&](const Dist &X) {
    return dadd(dadd(dmul(ddt(X, Rd, Yd), nT), dmul(ddr(X), nR)), dmul(X, jscale(nP, 0.0, em)));
  };
  auto Dbop = [&](const Dist &X) {
    return dadd(dadd(dmul(ddt(X, Rd, Yd), bT), dmul(ddth(X), bH)), dmul(X, jscale(bP, 0.0, em)));
  };
  const Dist A = dsub(dadd(Dbop(Tnm), dmul(Tnm, s1)), dadd(Dop(Tmm), dmul(Tmm, s2)));
  const Dist B = dsub(dadd(Dop(Tnm), dmul(Tnm, s3)), dadd(Dbop(Tnn), dmul(Tnn, s4)));
  const Dist T4 = dadd(dadd(Dop(A), dmul(A, s5)), dadd(Dbop(B), dmul(B, s6)));

  // canonical reduction at x = y = 0, s = 0
  double cr[dN], ci[dN];
  for (int t = 0; t < dN; t++) cr[t] = ci[t] = 0.0;
  for (int t = 0; t < dN; t++) {
    const int i = kDI[t], j = kDJ[t];
    for (int kr = 0; kr <= i; kr++) {
      for (int kt = 0; kt <= j; kt++) {
        // d^kr/dx^kr d^kt/dy^kt of the coefficient at the origin
        int mono = k1;
        double fac = 1.0;
        if (kr == 1 && kt == 0) mono = kx;
        if (kr == 0 && kt == 1) mono = ky;
        if (kr == 2) { mono = kxx; fac = 2.0; }
        if (kt == 2) { mono = kyy; fac = 2.0; }
        if (kr == 1 && kt == 1) mono = kxy;
        const double bin = (i == 2 && kr == 1 ? 2.0 : 1.0) * (j == 2 && kt == 1 ? 2.0 : 1.0);
        const double sgn = ((kr + kt) & 1) ? -1.0 : 1.0;
        const int u = didx(i - kr, j - kt);
        cr[u] = cr[u] + sgn * bin * fac * T4.g[t].re[mono];
        ci[u] = ci[u] + sgn * bin * fac * T4.g[t].im[mono];
      }
    }
  }
  // times e^{-i m phi_p}/pi and the reference's factor 1/2 (exact: a power of 2)
  const double ipi = 0.5 * 0.3183098861837907;
  const double mr = p.cmphi * ipi, mi = -p.smphi * ipi;
  for (int t = 0; t < dN; t++) {
    out[2 * t] = cr[t] * mr - ci[t] * mi;
    out[2 * t + 1] = cr[t] * mi + ci[t] * mr;
  }
}

// exp(x) for x <= 0 by Cody-Waite reduction x = k ln2 + f, |f| <= ln2/2, and
// the Taylor series of e^f to degree 14 (truncation below 1e-17); scaled by
// 2^k with exact multiplications.  Same bits on host and device.
TEUK_HD double exp_neg(double x) {
  if (!(x > -708.0)) return 0.0;
  const double ln2_hi = 0.6931471803691238;      // 33 significant bits
  const double ln2_lo = 1.9082149292705877e-10;
  const double inv_ln2 = 1.4426950408889634;
  double kf = x * inv_ln2;
  kf = (kf < 0.0) ? (double)(long long)(kf - 0.5) : (double)(long long)(kf + 0.5);
  const double f = (x - kf * ln2_hi) - kf * ln2_lo;
  double s = 1.0 / 87178291200.0;  // 1/14!
  const double inv_fact[14] = {1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0,
                               1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,      1.0 / 720.0,
                               1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,         0.5,
                               1.0,                1.0};
  for (int i = 0; i < 14; i++) s = s * f + inv_fact[i];
  // 2^k for k in [-1022, 0] by repeated halving of an exact power of two
  int k = (int)kf;
  double scale = 1.0;
  double step = 0.5;
  int n = -k;
  while (n > 0) {
    if (n & 1) scale = scale * step;
    step = step * step;
    n >>= 1;
  }
  return s * scale;
}

}  // namespace teuk_t4
