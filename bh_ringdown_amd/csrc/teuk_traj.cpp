// Seeded transition-plus-plunge worldlines in Boyer-Lindquist coordinates,
// written as the 10-column .traj rows the solver consumes
// (t r theta dr/dt dtheta/dt dphi/dt phi E Lz Q; GenOT/WL_BL_RK.cc:214-221,
// read at body.cu:97-129).
//
// Model (a compact Ori-Thorne-style transition, not a port of GenOT): the
// particle starts on the circular orbit of fixed inclination iota at
// r0 = f r_isco(iota), loses energy and angular momentum at the quadrupole
// rates of a circular orbit with the same angular speed Omega,
// dE/dt = -(32/5) eta Omega^(10/3), dLz/dt = cos(iota) dE/dt / Omega, with Q = Lz^2 tan^2(iota)
// (fixed inclination); once r falls below r_isco - 0.5 M the constants are
// frozen and the motion relaxes onto the geodesic plunge.  Positions follow Kerr's
// second-order geodesic equations in Mino time, integrated by RK4 in BL time:
//   (dr/dl)^2 = R(r) = [E(r^2+a^2) - a Lz]^2 - Delta [r^2 + (Lz - a E)^2 + Q],
//   d^2r/dl^2 = R'(r)/2, so r tracks the moving potential minimum while E and
//   Lz shrink and falls away from it once the minimum disappears;
//   with z = cos(theta): (dz/dl)^2 = Z(z) = Q - z^2 [Q + Lz^2 + a^2(1-E^2)(1-z^2)],
//   d^2z/dl^2 = Z'(z)/2;
//   dt/dl = -a(a E sin^2 - Lz) + (r^2+a^2)/Delta [E(r^2+a^2) - a Lz],
//   dphi/dl = -(a E - Lz/sin^2) + a/Delta [E(r^2+a^2) - a Lz].
// The seed picks the initial azimuth and the initial polar phase.
#include <cmath>
#include <cstdint>
#include <cstring>
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

struct State {
  double t, r, z, phi, ur, uz;  // z = cos(theta); u* = d/dlambda
};

struct Deriv {
  double dt, dr, dz, dphi, dur, duz;  // per unit lambda
};

Deriv rhs(const State &s, double a, const Consts &c) {
  const double r = s.r, z = s.z;
  const double sn2 = 1.0 - z * z;
  const double D = r * r - 2.0 * r + a * a;
  const double P = c.E * (r * r + a * a) - a * c.L;
  Deriv d;
  d.dt = -a * (a * c.E * sn2 - c.L) + (r * r + a * a) / D * P;
  d.dphi = -(a * c.E - c.L / sn2) + a / D * P;
  d.dr = s.ur;
  d.dz = s.uz;
  d.dur = 0.5 * dRdr(r, a, c);
  // (dz/dl)^2 = Zf(z) = Q - z^2 [a^2 (1-E^2)(1-z^2) + L^2 + Q];  d2z/dl2 = Zf'(z)/2
  const double b = a * a * (1.0 - c.E * c.E);
  d.duz = 0.5 * (-2.0 * z * (b * (1.0 - z * z) + c.L * c.L + c.Q) + 2.0 * b * z * z * z);
  return d;
}

// Polar passes (theta -> 0) are excluded: the (theta, phi) description of
// the worldline and of the source is singular on the axis.
constexpr double kMaxIota = 85.0 * M_PI / 180.0;
// Time constant (M) of the velocity relaxation after the constants freeze.
constexpr double kRelax = 2.0;

}  // namespace

extern "C" int teuk_traj_plunge(double a, double eta, double iota, double r0_over_isco, double dt_out,
                                double t_final, unsigned long long seed, double *rows, int cap, int *nrows) {
  if (!rows || !nrows || cap <= 0 || !(a >= 0.0 && a < 1.0) || !(eta > 0.0) || !(dt_out > 0.0) ||
      !(t_final > 0.0) || !(r0_over_isco >= 1.0) || !(iota >= 0.0 && iota <= kMaxIota))
    return TEUK_EINVAL;
  const double rh = 1.0 + std::sqrt(1.0 - a * a);
  const double risco = isco(a, iota);
  const double r0 = r0_over_isco * risco;
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
  // polar motion: z oscillates between +-z_max; start at z_max cos(psi0)
  const double b0 = a * a * (1.0 - c.E * c.E);
  const double sq = c.L * c.L + c.Q + b0;
  const double zmax2 = (sq - std::sqrt(sq * sq - 4.0 * b0 * c.Q)) / (2.0 * b0 > 0.0 ? 2.0 * b0 : 1.0);
  double zmax = b0 > 1e-14 ? std::sqrt(std::max(0.0, zmax2)) : std::sqrt(c.Q / (c.L * c.L + c.Q));
  State s{0.0, r0, zmax * std::cos(psi0), phi0, 0.0, 0.0};
  {
    const double Z = c.Q - s.z * s.z * (b0 * (1.0 - s.z * s.z) + c.L * c.L + c.Q);
    s.uz = -std::sqrt(std::max(0.0, Z)) * (std::sin(psi0) >= 0 ? 1.0 : -1.0);
  }
  const int sub = 8;
  const double h = dt_out / sub;
  bool plunging = false;
  int n = 0;
  auto emit = [&](const State &st) {
    const Deriv d = rhs(st, a, c);
    double *row = rows + 10 * (size_t)n;
    const double sn = std::sqrt(1.0 - st.z * st.z);
    row[0] = st.t;
    row[1] = st.r;
    row[2] = std::acos(st.z);
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
      // RK4 in BL time: dy/dt = (dy/dl) / (dt/dl)
      auto f = [&](const State &y) {
        Deriv d = rhs(y, a, c);
        const double it = 1.0 / d.dt;
        return Deriv{1.0, d.dr * it, d.dz * it, d.dphi * it, d.dur * it, d.duz * it};
      };
      auto add = [](const State &y, const Deriv &d, double w) {
        return State{y.t + w * d.dt, y.r + w * d.dr, y.z + w * d.dz, y.phi + w * d.dphi, y.ur + w * d.dur,
                     y.uz + w * d.duz};
      };
      if (s.r - rh < 1e-7) {  // BL time never reaches the horizon: co-rotate with it
        s.t += h;
        s.phi += h * a / (rh * rh + a * a);
        continue;
      }
      const Deriv k1 = f(s), k2 = f(add(s, k1, 0.5 * h)), k3 = f(add(s, k2, 0.5 * h)), k4 = f(add(s, k3, h));
      State ns = s;
      ns.t += h;
      ns.r += h / 6.0 * (k1.dr + 2.0 * k2.dr + 2.0 * k3.dr + k4.dr);
      ns.z += h / 6.0 * (k1.dz + 2.0 * k2.dz + 2.0 * k3.dz + k4.dz);
      ns.phi += h / 6.0 * (k1.dphi + 2.0 * k2.dphi + 2.0 * k3.dphi + k4.dphi);
      ns.ur += h / 6.0 * (k1.dur + 2.0 * k2.dur + 2.0 * k3.dur + k4.dur);
      ns.uz += h / 6.0 * (k1.duz + 2.0 * k2.duz + 2.0 * k3.duz + k4.duz);
      if (!(ns.r > rh)) ns.r = rh + 1e-7;
      s = ns;
      // radiation reaction until the plunge
      if (!plunging) {
        // angular speed on the sphere, sqrt(phidot^2 sin^2 + thetadot^2)
        const Deriv d = rhs(s, a, c);
        const double sn2 = std::max(1e-300, 1.0 - s.z * s.z);
        const double phd = d.dphi / d.dt, zd = d.dz / d.dt;
        const double Om = std::sqrt(phd * phd * sn2 + zd * zd / sn2);
        const double dE = -(32.0 / 5.0) * eta * std::pow(Om, 10.0 / 3.0) * h;
        c.E += dE;
        set_lt(c, c.Lt + dE / Om, std::cos(iota), std::sin(iota));
        if (s.r < risco - 0.5) plunging = true;
      } else if (s.ur < 0.0) {
        // With E and Lz changing, (dr/dl)^2 - R(r) drifts from zero; after the
        // constants freeze, relax dr/dl onto the plunge geodesic's
        // -sqrt(R(r)) over a few M of time so the velocity stays continuous.
        const double Rv = Rfun(s.r, a, c);
        if (Rv > 0.0) s.ur += (-std::sqrt(Rv) - s.ur) * std::min(1.0, h / kRelax);
      }
    }
    s.t = n * dt_out;  // rows sit on the exact output grid
    emit(s);
  }
  *nrows = n;
  return TEUK_OK;
}
