// Host-side code of libteuk run under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitizers.py builds it with
// g++ -fsanitize=address,undefined and runs it): the grid inversions, the
// coefficient and projection tables, the source preparation and T4
// coefficients, the trajectory generator over steep, fast, near-polar and
// high-spin worldlines, and the stage kernels' theta-chunk layout and block
// order.  Every output must be finite; any out-of-bounds
// access or undefined operation aborts the run.
#include <cmath>
#include <cstdio>
#include <vector>

#include "teuk.h"
#include "teuk_internal.h"

namespace {

int failures = 0;

void check_finite(const char *what, const double *v, size_t n) {
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(v[i])) {
      std::printf("non-finite %s[%zu]\n", what, i);
      failures++;
      return;
    }
}

void grids_and_tables(int N, int M, double xmin, double xmax, double a, double m) {
  using namespace teuk;
  const double dx = (xmax - xmin) / N;
  std::vector<double> x(N), r(N), th(M), xh(N), rh(N);
  host_grid(N, M, xmin, xmax, 1.0, a, x.data(), r.data(), th.data());
  host_grid_half(N, xmin, xmax, 1.0, a, xh.data(), rh.data());
  check_finite("r_c", r.data(), N);
  check_finite("r_h", rh.data(), N);
  std::vector<double> ktab((size_t)kNK * N), rtab((size_t)kNR * M);
  host_coeff_tables(N, M, xmin, xmax, dx, 1.0, a, m, r.data(), th.data(), ktab.data(), rtab.data(), 0.0, false);
  host_coeff_tables(N, M, xmin, xmax, dx, 1.0, a, m, rh.data(), th.data(), ktab.data(), rtab.data(), -0.5, true);
  std::vector<double> w((size_t)7 * 5 * (M / 4) * 3), ph(2 * (size_t)N);
  host_projection_weights(M, m, w.data());
  check_finite("projection weights", w.data(), w.size());
  host_source_phase(N, 1.0, a, m, r.data(), ph.data());
}

void trajectory_and_source(double a, double eta, double iota_deg, double dt_out, double t_final) {
  using namespace teuk;
  const int cap = (int)(t_final / dt_out) + 8;
  std::vector<double> rows((size_t)cap * 10);
  int n = 0;
  const int rc = teuk_traj_plunge(a, eta, iota_deg * M_PI / 180.0, 1.02, dt_out, t_final, 7, rows.data(), cap, &n);
  if (rc != TEUK_OK || n < 2) {
    std::printf("trajectory a=%g eta=%g iota=%g failed (%d)\n", a, eta, iota_deg, rc);
    failures++;
    return;
  }
  check_finite("trajectory", rows.data(), (size_t)n * 10);
  // source preparation and T4 coefficients along the worldline (one row in
  // every 64), on a grid that holds the particle
  Grid g{};
  g.N = 4096;
  g.M = 64;
  g.ptsx = g.ptsy = 128;
  g.xmin = -400.0;
  g.xmax = 1200.0;
  g.dx = (g.xmax - g.xmin) / g.N;
  g.dtheta = 3.14159265358979 / g.M;
  g.dt = 0.5 * g.dx;
  g.mass = 1.0;
  g.a = a;
  g.m = 2.0;
  for (int k = 1; k < n; k += 64) {
    const double *q = &rows[10 * (size_t)k];
    if (!(q[1] > 1.0 + std::sqrt(1.0 - a * a) + 1e-3)) break;   // at the horizon
    SrcStep st{};
    st.timer = q[0];
    st.start = 0.0;
    st.rp = q[1];
    st.tp = q[2];
    st.drdt = q[3];
    st.dthdt = q[4];
    st.dphidt = q[5];
    st.phip = q[6];
    st.E = q[7];
    st.lz = q[8];
    st.Q = q[9];
    SrcPrep prep;
    host_source_prepare(g, st, prep);
    double c[12];
    teuk_t4::t4_coeffs(prep.in, c);
    check_finite("T4 coefficients", c, 12);
  }
}

// The stage kernels' theta chunks (theta_chunks, theta_chunk_b0) cover the
// rows 0 .. M-1 once and compute the rows 3 .. M-4 once, each edge chunk
// holding the two rows it extrapolates from, for every M the solver accepts;
// the single-solve block order (stage_block_tile) visits every (strip,
// block-row) once, in both orders.
void chunk_layout_and_block_order() {
  using namespace teuk;
  for (int M = 16; M <= 1024; M += 8)
    for (int R : {2, 4, 8}) {
      const int nc = theta_chunks(M, R), E = edge_rows(R);
      std::vector<int> spanned(M, 0), computed(M, 0);
      for (int c = 0; c < nc; c++) {
        const bool first = c == 0, last = c == nc - 1;
        const int b0 = theta_chunk_b0(M, R, c), span = (first || last) ? E + 3 : R;
        const int c0 = b0 + (first ? 3 : 0), c1 = b0 + (last ? span - 3 : span);   // computed rows [c0, c1)
        if (b0 < 0 || b0 + span > M) {
          std::printf("chunk %d of M=%d R=%d outside the grid\n", c, M, R);
          failures++;
          continue;
        }
        for (int r = b0; r < b0 + span; r++) spanned[r]++;
        for (int r = c0; r < c1; r++) computed[r]++;
        if ((first && !(c0 <= 3 && 4 < c1)) || (last && !(c0 <= M - 5 && M - 4 < c1))) {
          std::printf("edge chunk %d of M=%d R=%d lacks its extrapolation rows\n", c, M, R);
          failures++;
        }
      }
      for (int r = 0; r < M; r++)
        if (spanned[r] != 1 || computed[r] != (r >= 3 && r < M - 3 ? 1 : 0)) {
          std::printf("row %d of M=%d R=%d: spanned %d, computed %d times\n", r, M, R, spanned[r], computed[r]);
          failures++;
          break;
        }
    }
  for (int S : {1, 2, 3, 5, 8, 12, 16, 24, 40, 64, 256})
    for (int ng = 1; ng <= 40; ng++)
      for (bool edge_first : {false, true}) {
        std::vector<int> seen((size_t)S * ng, 0);
        for (int b = 0; b < S * ng; b++) {
          int strip = -1, grp = -1;
          stage_block_tile(b, S * ng, ng, edge_first, strip, grp);
          if (strip < 0 || strip >= S || grp < 0 || grp >= ng) {
            std::printf("block %d of %dx%d maps outside\n", b, S, ng);
            failures++;
            return;
          }
          seen[(size_t)strip * ng + grp]++;
        }
        for (int v : seen)
          if (v != 1) {
            std::printf("block order %dx%d (edge first %d) is not a permutation\n", S, ng, (int)edge_first);
            failures++;
            break;
          }
      }
}

}  // namespace

int main() {
  chunk_layout_and_block_order();
  grids_and_tables(1024, 32, -100.0, 300.0, 0.5, 2.0);
  grids_and_tables(512, 16, -40.0, 60.0, 0.99, -3.0);
  grids_and_tables(256, 64, -50.0, 50.0, 1e-8, 0.0);
  grids_and_tables(256, 16, -100.0, 300.0, 0.0, 1.0);
  const double cases[][3] = {{0.9, 1e-3, 30.0}, {0.0926, 8.4e-3, 85.0}, {0.9366, 7.4e-3, 84.0},
                             {0.99, 1e-2, 60.0}, {0.0, 1e-4, 0.0}, {0.7, 1.3e-3, 0.0}, {0.6, 4.1e-3, 85.0}};
  for (const auto &c : cases) trajectory_and_source(c[0], c[1], c[2], 0.05, 600.0);
  double X, dX;
  for (double L = -40.0; L <= 3.4; L += 0.37)
    if (teuk_ot_transition(L, &X, &dX) != TEUK_OK || !std::isfinite(X) || !std::isfinite(dX)) failures++;
  std::printf("host checks: %d failures\n", failures);
  return failures == 0 ? 0 : 1;
}
