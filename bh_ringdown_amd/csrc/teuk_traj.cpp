// Seeded inspiral-transition-plunge worldlines in Boyer-Lindquist
// coordinates, written as the 10-column .traj rows the solver consumes
// (t r theta dr/dt dtheta/dt dphi/dt phi E Lz Q; GenOT/WL_BL_RK.cc:214-221,
// read at body.cu:97-129).
//
// Three phases, the ones GenOT joins (GenOT/ParamWLMino.cc; Apte & Hughes
// 2019), with a quadrupole-formula flux in place of GenOT's tables of
// adiabatic Teukolsky fluxes (GREMLIN output, not part of the reference tree):
//
// 1. Inspiral: from the circular orbit of fixed inclination iota at
//    r0 = f r_isco(iota), energy and angular momentum are lost at the
//    quadrupole rates of a circular orbit of the particle's angular speed
//    Omega, dE/dt = -(32/5) eta Omega^(10/3), dLz/dt = cos(iota) dE/dt / Omega,
//    with Q = Lz^2 tan^2(iota) (fixed inclination), while the positions follow
//    Kerr's second-order geodesic equations in Mino time lambda:
//      (dr/dl)^2 = R(r) = [E(r^2+a^2) - a Lz]^2 - Delta [r^2 + (Lz - a E)^2 + Q],
//      d^2r/dl^2 = R'(r)/2 (r tracks the potential minimum as E, Lz shrink);
//      with z = cos(theta): (dz/dl)^2 = Z(z) = Q - z^2 [Q + Lz^2 + a^2(1-E^2)(1-z^2)]
//      = b (z-^2 - z^2)(z+^2 - z^2), b = a^2(1-E^2), integrated through the
//      polar angle chi of z = z- cos(chi) (Drasco & Hughes 2004),
//      dchi/dl = sqrt(b z+^2 - b z-^2 cos^2 chi), so that |z| <= z- < 1 holds
//      at every step while the constants, and with them z-, change;
//      dt/dl = -a(a E sin^2 - Lz) + (r^2+a^2)/Delta [E(r^2+a^2) - a Lz],
//      dphi/dl = -(a E - Lz/sin^2) + a/Delta [E(r^2+a^2) - a Lz].
// 2. Transition (the generalised Ori-Thorne transition of
//    ParamWLMino.cc:180-300): from where r = r_isco + R0 X(L_i),
//      r(lambda) = r_isco + R0 X(lambda / l0),
//    X the universal solution of X'' = -X^2 - L with X -> sqrt(-L) (the
//    function GenOT/OT_X_vs_L.dat tabulates), R0 = (eta B)^(2/5) A^(-3/5),
//    l0 = (eta A B)^(-1/5), where at the ISCO A = -R'''/4 and
//    B = -(R_Er fE + R_Lr fL + R_Qr fQ)/2 with the eta-free fluxes fE, fL, fQ
//    per unit lambda; E, Lz and Q follow the quadratic-plus-cubic model of
//    ParamWLMino.cc:229-241 (the inspiral's values and rates at L_i joined to
//    the ISCO values and fluxes at L = 0), until L_f.  L_i = -3 and
//    L_f = 2.35 are the values Apte & Hughes recommend (GenOT/README).
// 3. Plunge: the geodesic with the constants frozen at L_f, from the
//    transition's r and dr/dl there (ParamWLMino.cc:300-354).
// Boyer-Lindquist time is the independent variable throughout, as in
// GenOT/WL_BL_RK.cc (RK4 with Mino time, the polar and azimuthal motion and,
// outside the transition, r and dr/dl as the integrated variables).  The
// seed picks the initial azimuth and the initial polar phase.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

#include "teuk.h"

namespace {

// E, Lz and Carter's Q; Lt is the total angular momentum of the orbit's
// inclination convention, Lz = Lt cos(iota), Q = Lt^2 sin^2(iota) (so
// Q = Lz^2 tan^2(iota), defined up to polar orbits).
struct Consts {
  double E, L, Q, Lt;
};

void set_lt(Consts &c, double Lt, double ci, double si) {
  c.Lt = Lt;
  c.L = Lt * ci;
  c.Q = Lt * Lt * si * si;
}

double Rfun(double r, double a, const Consts &c) {
  const double P = c.E * (r * r + a * a) - a * c.L;
  const double D = r * r - 2.0 * r + a * a;
  return P * P - D * (r * r + (c.L - a * c.E) * (c.L - a * c.E) + c.Q);
}
double dRdr(double r, double a, const Consts &c) {
  const double P = c.E * (r * r + a * a) - a * c.L;
  const double D = r * r - 2.0 * r + a * a;
  return 4.0 * c.E * r * P - (2.0 * r - 2.0) * (r * r + (c.L - a * c.E) * (c.L - a * c.E) + c.Q) - D * 2.0 * r;
}
double d2Rdr2(double r, double a, const Consts &c) {
  const double h = 1e-5 * (1.0 + r);
  return (dRdr(r + h, a, c) - dRdr(r - h, a, c)) / (2.0 * h);
}

// Circular orbit of inclination iota at radius r: R = R' = 0, by Newton in
// (E, Lt) from the equatorial prograde constants.
bool circular(double r, double a, double iota, Consts &c) {
  const double v = 1.0 / std::sqrt(r);
  const double den = std::sqrt(1.0 - 3.0 * v * v + 2.0 * a * v * v * v);
  const double ci = std::cos(iota), si = std::sin(iota);
  c.E = (1.0 - 2.0 * v * v + a * v * v * v) / den;
  set_lt(c, std::sqrt(r) * (1.0 - 2.0 * a * v * v * v + a * a * v * v * v * v) / den, ci, si);
  for (int it = 0; it < 100; it++) {
    const double f1 = Rfun(r, a, c), f2 = dRdr(r, a, c);
    const double hE = 1e-7, hL = 1e-7;
    Consts cE = c, cL = c;
    cE.E += hE;
    set_lt(cL, c.Lt + hL, ci, si);
    const double j11 = (Rfun(r, a, cE) - f1) / hE, j12 = (Rfun(r, a, cL) - f1) / hL;
    const double j21 = (dRdr(r, a, cE) - f2) / hE, j22 = (dRdr(r, a, cL) - f2) / hL;
    const double det = j11 * j22 - j12 * j21;
    if (det == 0.0) return false;
    const double dE = (f1 * j22 - f2 * j12) / det, dL = (j11 * f2 - j21 * f1) / det;
    c.E -= dE;
    set_lt(c, c.Lt - dL, ci, si);
    if (std::fabs(dE) < 1e-14 && std::fabs(dL) < 1e-13) return true;
  }
  return false;
}

// ISCO of the inclined family: the smallest r with R'' >= 0 on circular orbits.
double isco(double a, double iota) {
  double lo = 1.0 + std::sqrt(1.0 - a * a) + 0.05, hi = 10.0;
  Consts c;
  for (int it = 0; it < 80; it++) {
    const double mid = 0.5 * (lo + hi);
    if (circular(mid, a, iota, c) && d2Rdr2(mid, a, c) < 0.0)
      hi = mid;
    else
      lo = mid;
  }
  return hi;
}

// ---- the Ori-Thorne transition function ---------------------------------
// X(L) on [kOtL0, kOtL1]: the solution of X'' = -X^2 - L that follows
// sqrt(-L) + 1/(8 L^2) as L -> -inf (that asymptotic series gives the data at
// kOtL0), by RK4 with step kOtH, tabulated with X' once per process.
constexpr double kOtL0 = -40.0, kOtL1 = 3.4, kOtH = 1e-3;

struct OtTable {
  std::vector<double> X, dX;
};

const OtTable &ot_table() {
  static OtTable tab;
  static std::once_flag once;
  std::call_once(once, [] {
    const int n = (int)std::lround((kOtL1 - kOtL0) / kOtH) + 1;
    tab.X.resize(n);
    tab.dX.resize(n);
    const double L0 = kOtL0;
    double x = std::sqrt(-L0) + 1.0 / (8.0 * L0 * L0), v = -0.5 / std::sqrt(-L0) - 1.0 / (4.0 * L0 * L0 * L0);
    auto acc = [](double l, double xx) { return -xx * xx - l; };
    for (int i = 0; i < n; i++) {
      tab.X[i] = x;
      tab.dX[i] = v;
      const double L = kOtL0 + i * kOtH, h = kOtH;
      const double k1x = v, k1v = acc(L, x);
      const double k2x = v + 0.5 * h * k1v, k2v = acc(L + 0.5 * h, x + 0.5 * h * k1x);
      const double k3x = v + 0.5 * h * k2v, k3v = acc(L + 0.5 * h, x + 0.5 * h * k2x);
      const double k4x = v + h * k3v, k4v = acc(L + h, x + h * k3x);
      x += h / 6.0 * (k1x + 2.0 * k2x + 2.0 * k3x + k4x);
      v += h / 6.0 * (k1v + 2.0 * k2v + 2.0 * k3v + k4v);
    }
  });
  return tab;
}

// X(L) and X'(L), cubic Hermite interpolation of the table.
double ot_x(double L, double *dX) {
  const OtTable &t = ot_table();
  const int n = (int)t.X.size();
  const double u = (L - kOtL0) / kOtH;
  int i = (int)std::floor(u);
  if (i < 0) i = 0;
  if (i > n - 2) i = n - 2;
  const double s = u - i, h = kOtH;
  const double x0 = t.X[i], x1 = t.X[i + 1], d0 = t.dX[i] * h, d1 = t.dX[i + 1] * h;
  const double s2 = s * s, s3 = s2 * s;
  const double x = (2 * s3 - 3 * s2 + 1) * x0 + (s3 - 2 * s2 + s) * d0 + (-2 * s3 + 3 * s2) * x1 + (s3 - s2) * d1;
  if (dX)
    *dX = ((6 * s2 - 6 * s) * x0 + (3 * s2 - 4 * s + 1) * d0 + (-6 * s2 + 6 * s) * x1 + (3 * s2 - 2 * s) * d1) / h;
  return x;
}

// The transition's limits in the universal time L (Apte & Hughes 2019).
constexpr double kLTi = -3.0, kLTf = 2.35;

struct State {
  double t, r, chi, phi, ur;    // cos(theta) = z- cos(chi); ur = dr/dlambda
  double lam;                   // Mino time from the transition's origin (transition only)
};

struct Deriv {
  double dt, dr, dchi, dphi, dur;  // per unit lambda
  double z, dz;                    // cos(theta) and its rate per unit lambda
};

// The polar turning point z- and b z+^2 of Z(z) = b (z-^2 - z^2)(z+^2 - z^2),
// the roots of b y^2 - (b + Lz^2 + Q) y + Q = 0 in y = z^2, in forms that
// stay finite as b -> 0 (z-^2 -> Q / (Lz^2 + Q), b z+^2 -> Lz^2 + Q).
struct Polar {
  double zm, b, bzp2;
};
Polar polar(double a, const Consts &c) {
  Polar p;
  p.b = a * a * (1.0 - c.E * c.E);
  const double S = p.b + c.L * c.L + c.Q;
  const double sum = S + std::sqrt(std::max(0.0, S * S - 4.0 * p.b * c.Q));
  p.bzp2 = 0.5 * sum;
  p.zm = sum > 0.0 ? std::sqrt(2.0 * c.Q / sum) : 0.0;
  return p;
}

Deriv rhs(const State &s, double a, const Consts &c) {
  const Polar pz = polar(a, c);
  const double cx = std::cos(s.chi), sx = std::sin(s.chi);
  const double r = s.r, z = pz.zm * cx;
  const double sn2 = 1.0 - z * z;
  const double D = r * r - 2.0 * r + a * a;
  const double P = c.E * (r * r + a * a) - a * c.L;
  Deriv d;
  d.dt = -a * (a * c.E * sn2 - c.L) + (r * r + a * a) / D * P;
  d.dphi = -(a * c.E - c.L / sn2) + a / D * P;
  d.dr = s.ur;
  d.dur = 0.5 * dRdr(r, a, c);
  d.dchi = std::sqrt(std::max(0.0, pz.bzp2 - pz.b * pz.zm * pz.zm * cx * cx));
  d.z = z;
  d.dz = pz.zm > 0.0 ? -pz.zm * sx * d.dchi : 0.0;
  return d;
}

// Quadrupole energy and angular-momentum loss per unit time and eta at the
// particle's angular speed on the sphere, sqrt(phidot^2 sin^2 + thetadot^2).
void quad_rates(const State &s, double a, const Consts &c, double *dEdt, double *dLtdt) {
  const Deriv d = rhs(s, a, c);
  const double sn2 = std::max(1e-300, 1.0 - d.z * d.z);
  const double phd = d.dphi / d.dt, zd = d.dz / d.dt;
  const double Om = std::sqrt(phd * phd * sn2 + zd * zd / sn2);
  *dEdt = -(32.0 / 5.0) * std::pow(Om, 10.0 / 3.0);
  *dLtdt = *dEdt / Om;
}

// The transition: ISCO constants, eta-free fluxes there, the Ori-Thorne
// scales and the quadratic-plus-cubic coefficients of E, Lt (Lz and Q follow
// Lt at fixed inclination).
struct Ot {
  double rl, El, Ltl;          // ISCO radius and constants
  double fE, fLt;              // eta-free fluxes per unit lambda at the ISCO
  double A, B, R0, l0;
  double e2, e3, t2, t3;       // model coefficients (set when the transition starts)
};

double r_trans(const Ot &o, double lam, double *drdl) {
  double dX = 0.0;
  const double X = ot_x(lam / o.l0, &dX);
  if (drdl) *drdl = dX * o.R0 / o.l0;
  return o.rl + o.R0 * X;
}

// ParamWLMino.cc:229-241 (the quadratic-plus-cubic model): Y(l) = Y_l + l (eta f
// + l (c2 / 2 + l c3 / 6)) with c2, c3 fixed by Y(li) = Yi, Y'(li) = dYi.
void model2(double Yl, double f, double Yi, double dYi, double li, double eta, double &c2, double &c3) {
  c2 = -2.0 * (3.0 * Yl - 3.0 * Yi + (2.0 * eta * f + dYi) * li) / (li * li);
  c3 = 6.0 * (2.0 * Yl - 2.0 * Yi + (eta * f + dYi) * li) / (li * li * li);
}
double model2_at(double Yl, double f, double c2, double c3, double eta, double l) {
  return Yl + l * (eta * f + l * (0.5 * c2 + l / 6.0 * c3));
}

// Polar passes (theta -> 0) are excluded: the (theta, phi) description of
// the worldline and of the source is singular on the axis.
constexpr double kMaxIota = 85.0 * M_PI / 180.0;

}  // namespace

extern "C" int teuk_traj_plunge(double a, double eta, double iota, double r0_over_isco, double dt_out,
                                double t_final, unsigned long long seed, double *rows, int cap, int *nrows) {
  if (!rows || !nrows || cap <= 0 || !(a >= 0.0 && a < 1.0) || !(eta > 0.0) || !(dt_out > 0.0) ||
      !(t_final > 0.0) || !(r0_over_isco >= 1.0) || !(iota >= 0.0 && iota <= kMaxIota))
    return TEUK_EINVAL;
  const double rh = 1.0 + std::sqrt(1.0 - a * a);
  const double ci = std::cos(iota), si = std::sin(iota);
  const double risco = isco(a, iota);

  // Ori-Thorne parameters at the ISCO.  Its circular orbit's Mino-time rate
  // dt/dl and angular speed are taken where it crosses the equator.
  Ot o{};
  {
    Consts cl;
    if (!circular(risco, a, iota, cl)) return TEUK_EINVAL;
    o.rl = risco;
    o.El = cl.E;
    o.Ltl = cl.Lt;
    State sl{0.0, risco, 0.5 * M_PI, 0.0, 0.0, 0.0};
    double dEdt, dLtdt;
    quad_rates(sl, a, cl, &dEdt, &dLtdt);
    const double G = rhs(sl, a, cl).dt;
    o.fE = dEdt * G;
    o.fLt = dLtdt * G;
    const double fL = o.fLt * ci, fQ = 2.0 * cl.Lt * si * si * o.fLt;
    const double r = risco, E = cl.E, L = cl.L;
    o.A = -0.25 * (12.0 + 24.0 * (E * E - 1.0) * r);
    const double REr = 8.0 * E * r * r * r + 4.0 * a * a * E * (1.0 + r) - 4.0 * a * L;
    const double RLr = -4.0 * (a * E + L * (r - 1.0));
    const double RQr = 2.0 * (1.0 - r);
    o.B = -0.5 * (REr * o.fE + RLr * fL + RQr * fQ);
    if (!(o.A > 0.0 && o.B > 0.0)) return TEUK_EINVAL;
    o.R0 = std::pow(eta * o.B, 0.4) * std::pow(o.A, -0.6);
    o.l0 = std::pow(eta * o.A * o.B, -0.2);
  }
  const double r_ti = r_trans(o, kLTi * o.l0, nullptr);
  const double lam_f = kLTf * o.l0;
  // The worldline enters the transition from the inspiral, at L_i, as GenOT's
  // does (ParamWLMino.cc:396-416): a start inside the transition region
  // (r0 below r_isco + R0 X(L_i), which for eta = 1e-3 lies 0.1 - 0.2 M
  // outside the ISCO) begins at that radius instead.
  const double r0 = std::max(r0_over_isco * risco, r_ti * (1.0 + 1e-9));
  Consts c;
  if (!circular(r0, a, iota, c)) return TEUK_EINVAL;

  // seeded initial azimuth and polar phase (splitmix64)
  uint64_t x = seed + 0x9E3779B97F4A7C15ull;
  auto next = [&]() {
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
  };
  const double phi0 = 2.0 * M_PI * next();
  const double psi0 = 2.0 * M_PI * next();
  // polar motion: z = z- cos(chi) oscillates between +-z-; chi starts at psi0
  State s{0.0, r0, psi0, phi0, 0.0, 0.0};

  // 0 inspiral, 1 transition, 2 plunge
  int phase = 0;
  const double E_l = o.El, Lt_l = o.Ltl;
  // Constants of the transition at Mino time lam (model 2).
  auto trans_consts = [&](double lam) {
    Consts k;
    k.E = model2_at(E_l, o.fE, o.e2, o.e3, eta, lam);
    set_lt(k, model2_at(Lt_l, o.fLt, o.t2, o.t3, eta, lam), ci, si);
    return k;
  };
  // Start the transition at Mino time li from the current constants and rates.
  auto start_transition = [&](double li) {
    double dEdt, dLtdt;
    quad_rates(s, a, c, &dEdt, &dLtdt);
    const double G = rhs(s, a, c).dt;
    model2(E_l, o.fE, c.E, eta * dEdt * G, li, eta, o.e2, o.e3);
    model2(Lt_l, o.fLt, c.Lt, eta * dLtdt * G, li, eta, o.t2, o.t3);
    s.lam = li;
    s.r = r_trans(o, li, &s.ur);
    c = trans_consts(li);
    phase = 1;
  };

  const int sub = 8;
  const double h = dt_out / sub;
  int n = 0;
  auto emit = [&](const State &st) {
    const Deriv d = rhs(st, a, c);
    double *row = rows + 10 * (size_t)n;
    const double sn = std::sqrt(1.0 - d.z * d.z);
    row[0] = st.t;
    row[1] = st.r;
    row[2] = std::acos(d.z);
    row[3] = d.dr / d.dt;
    row[4] = -(d.dz / d.dt) / (sn > 1e-300 ? sn : 1e-300);
    row[5] = d.dphi / d.dt;
    row[6] = st.phi;
    row[7] = c.E;
    row[8] = c.L;
    row[9] = c.Q;
    n++;
  };
  emit(s);
  while (n < cap && n * dt_out < t_final + 0.5 * dt_out) {
    for (int k = 0; k < sub; k++) {
      if (s.r - rh < 1e-7) {  // BL time never reaches the horizon: co-rotate with it
        s.t += h;
        s.phi += h * a / (rh * rh + a * a);
        continue;
      }
      // RK4 in BL time: dy/dt = (dy/dl) / (dt/dl).  In the transition r and
      // the constants are functions of lambda, which is integrated instead.
      auto f = [&](const State &y) {
        State yy = y;
        Consts cc = c;
        if (phase == 1) {
          cc = trans_consts(y.lam);
          yy.r = r_trans(o, y.lam, &yy.ur);
        }
        Deriv d = rhs(yy, a, cc);
        const double it = 1.0 / d.dt;
        return std::pair<Deriv, double>(Deriv{1.0, d.dr * it, d.dchi * it, d.dphi * it, d.dur * it, 0.0, 0.0}, it);
      };
      auto add = [](const State &y, const std::pair<Deriv, double> &dd, double w) {
        const Deriv &d = dd.first;
        return State{y.t + w * d.dt, y.r + w * d.dr, y.chi + w * d.dchi, y.phi + w * d.dphi, y.ur + w * d.dur,
                     y.lam + w * dd.second};
      };
      const auto k1 = f(s), k2 = f(add(s, k1, 0.5 * h)), k3 = f(add(s, k2, 0.5 * h)), k4 = f(add(s, k3, h));
      auto comb = [&](double State::*m, double Deriv::*dm) {
        return s.*m + h / 6.0 * (k1.first.*dm + 2.0 * k2.first.*dm + 2.0 * k3.first.*dm + k4.first.*dm);
      };
      State ns = s;
      ns.t += h;
      ns.chi = comb(&State::chi, &Deriv::dchi);
      ns.phi = comb(&State::phi, &Deriv::dphi);
      if (phase == 1) {
        ns.lam = s.lam + h / 6.0 * (k1.second + 2.0 * k2.second + 2.0 * k3.second + k4.second);
        ns.r = r_trans(o, ns.lam, &ns.ur);
      } else {
        ns.r = comb(&State::r, &Deriv::dr);
        ns.ur = comb(&State::ur, &Deriv::dur);
      }
      if (!(ns.r > rh)) ns.r = rh + 1e-7;
      s = ns;
      if (phase == 0) {
        // radiation reaction of the inspiral
        double dEdt, dLtdt;
        quad_rates(s, a, c, &dEdt, &dLtdt);
        c.E += eta * dEdt * h;
        set_lt(c, c.Lt + eta * dLtdt * h, ci, si);
        if (s.r <= r_ti) start_transition(kLTi * o.l0);
      } else if (phase == 1) {
        c = trans_consts(s.lam);
        if (s.lam >= lam_f) phase = 2;   // the plunge: constants frozen from here
      }
    }
    s.t = n * dt_out;  // rows sit on the exact output grid
    emit(s);
  }
  for (size_t i = 0; i < 10 * (size_t)n; i++)
    if (!std::isfinite(rows[i])) return TEUK_EINVAL;   // never hand out a broken worldline
  *nrows = n;
  return TEUK_OK;
}

extern "C" int teuk_ot_transition(double L, double *X, double *dX) {
  if (!X || !(L >= kOtL0 && L <= kOtL1)) return TEUK_EINVAL;
  *X = ot_x(L, dX);
  return TEUK_OK;
}
