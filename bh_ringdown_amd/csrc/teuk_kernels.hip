// HIP kernels for the FP64 2+1D Teukolsky evolution on gfx950 (MI355X).
//
// Hot kernel: k_stage<INTEG, STAGE> -- one Runge-Kutta stage fused end to end:
// 6th-order rho/theta stencils, Kreiss-Oliger dissipation, the RHS assembly,
// the stage combination and both boundary fills, in a single pass over the
// grid.  It replaces the reference's kernel_rhs1/2 + kernel_update1/2/3 +
// kernel_boundary1a/1b/2a/2b sequence (TD_GPU_code/HighOrder/kern.cu:171-753,
// launched 12 times per step at body.cu:131-165).
//
// Tile: a wavefront owns 64 consecutive rho columns (coalesced, rho is the
// contiguous index of parm.h:13) and a chunk of 8 theta rows.  Psi's theta
// stencil (+-3 rows) lives in a 14-deep per-lane register window, so each
// theta row is fetched once per chunk.  The rho stencil (+-4 columns, for
// KO dissipation) is read from an LDS copy of the current row with a 4-column
// halo on each side.  Boundary rows/columns are produced by the tile that
// owns them: the first and last theta chunk extrapolate from their own rows
// 3,4 / M-5,M-4 (kern.cu:320-409) and the first and last rho wavefront fill
// columns 0..2 / N-3..N-1 from lanes 3,4 / N-5,N-4 by cross-lane shuffles
// (kern.cu:412-449), keeping the reference's theta-then-rho order.
#include "teuk_internal.h"

#include <climits>

namespace teuk {
namespace {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Coefficient fields (kernel_init, kern.cu:84-169).  One thread per point.
// The transcendental inputs come from host tables -- per rho column
// (x, Omega, Omega', Omega'') and per theta row (sin, cos, cot, csc) -- and
// the per-point assembly uses only correctly rounded IEEE operations without
// contraction, so every coefficient equals the host computation bit for bit.
// Plane 5 holds d_re until k_fix_dre folds in -dbdx*ee (kern.cu:192-197),
// which is time independent and therefore computed once here.
__global__ __launch_bounds__(256) void k_init_coeffs(Grid g, const double *__restrict__ col4,
                                                     const double *__restrict__ row4,
                                                     const double *__restrict__ r_c,
                                                     double *__restrict__ coef) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int k = (int)(id / g.N), j = (int)(id - (size_t)k * g.N);
  double out[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (k >= 3 && k < g.M - 3) {
    const double mass = g.mass, spin = g.a, em = g.m;
    const double sn = row4[4 * k], cs = row4[4 * k + 1], cot = row4[4 * k + 2], csc = row4[4 * k + 3];
    const double x = col4[4 * j], om = col4[4 * j + 1], om1 = col4[4 * j + 2], om2 = col4[4 * j + 3];
    const double L = om - x * om1;
    const double R = om * r_c[j];
    const double a2 = spin * spin;
    const double om_2 = om * om;
    const double S = R * R + om_2 * a2;
    const double D = R * R - 2.0 * R * mass * om + om_2 * a2;
    const double L2 = L * L, L3 = L2 * L;
    const double bir = -(((2.0 * L - om_2) * (S * S)) / L2) + a2 * D * (sn * sn);
    const double Atr = (2.0 * (L - om_2) * (S * S)) / (bir * L2);
    const double Arr = -((om_2 * (S * S)) / (bir * L2));
    const double Ayy = -(D / bir);
    const double R2 = R * R, R3 = R2 * R, R5 = R3 * R2;
    const double a4 = a2 * a2;
    const double om3 = om_2 * om, om4 = om_2 * om_2, om5 = om4 * om;
    const double BtRe = -((om * (R3 * (4.0 * L2 * (R - 3.0 * mass * om) + 2.0 * L * R2 * om1
                                        + R2 * om * om2 * x)
                                 + 2.0 * R * (L2 * om_2 * (5.0 * R - 4.0 * mass * om)
                                              - 3.0 * L3 * (R - 2.0 * mass * om)
                                              + 2.0 * L * R2 * om_2 * om1
                                              + R2 * om3 * om2 * x) * a2
                                 + (-6.0 * L3 * om_2 + 6.0 * L2 * om4 + 2.0 * L * R * om4 * om1
                                    + R * om5 * om2 * x) * a4))
                          / (bir * L3 * R));
    const double BrRe = -((om * (R5 * (2.0 * om * om1 - 2.0 * om1 * om1 * x + om * om2 * x)
                                 + 2.0 * R3 * om_2 * (2.0 * om * om1 - 2.0 * om1 * om1 * x
                                                      + om * om2 * x) * a2
                                 + om4 * (6.0 * om_2 + 2.0 * om1 * om1 * x * (-R + 3.0 * x)
                                          + om * (2.0 * R * om1 - 12.0 * om1 * x + R * om2 * x)) * a4
                                 + 2.0 * L2 * R * (2 * R3 - 6.0 * R2 * mass * om
                                                   + 5.0 * R * om_2 * a2
                                                   - 4.0 * mass * om3 * a2)))
                          / (bir * L3 * R));
    const double CRe = (D * (2.0 * R2 + 6.0 * R * mass * om - 6.0 * om_2 * a2
                             + 4.0 * R2 * (cot * cot) - 4.0 * em * R2 * cot * csc
                             + (em * em) * R2 * (csc * csc)))
                       / (bir * R2);
    const double BtIm = (2.0 * spin * (em * (-(om_2 * S) + L * (R2 + 2.0 * R * mass * om + om_2 * a2))
                                       - 2.0 * L * D * cs))
                        / (bir * L);
    const double BrIm = (-2.0 * em * om_2 * spin * S) / (bir * L);
    const double CIm = (2.0 * em * om * spin * (R2 + 2.0 * R * mass * om - 3.0 * om_2 * a2)) / (bir * R);
    const double bb = -(Atr + __dsqrt_rn(Atr * Atr + 4.0 * Arr)) / 2.0;
    out[0] = bb;
    out[1] = Ayy;
    out[2] = Atr + bb;
    out[3] = -BtRe;
    out[4] = -BtIm;
    out[5] = BrRe - bb * BtRe;
    out[6] = BrIm - bb * BtIm;
    out[7] = -CRe;
    out[8] = -CIm;
  }
#pragma unroll
  for (int c = 0; c < 9; c++) coef[c * NM + id] = out[c];
}

// dre = d_re - dbdx*ee with dbdx the 6th-order rho derivative of bb
// (kern.cu:192-197); zero dbdx outside 3 <= a < N-3.
__global__ __launch_bounds__(256) void k_fix_dre(Grid g, double *__restrict__ coef) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int k = (int)(id / g.N), a = (int)(id - (size_t)k * g.N);
  const double *bb = coef;
  double dbdx = 0.0;
  if (a >= 3 && a < g.N - 3)
    dbdx = (45.0 * bb[id + 1] - 45.0 * bb[id - 1] - 9.0 * bb[id + 2] + 9.0 * bb[id - 2] + bb[id + 3] -
            bb[id - 3]) *
           (kD1 * (1.0 / g.dx));
  coef[5 * NM + id] = coef[5 * NM + id] - dbdx * coef[2 * NM + id];
}

// ---------------------------------------------------------------------------
// Fused Runge-Kutta stage.
//   INTEG 0 (SSP-RK3, kern.cu:301-318, 582-620):
//     STAGE 1: out = in + dt L(in)
//     STAGE 2: out = 1/4 (3 base + in + dt L(in))
//     STAGE 3: out = 1/3 (base + 2 in + 2 dt L(in))        (out may == base)
//   INTEG 1 (classic RK4, accumulator form):
//     STAGE 1: out = in + dt/2 L(in);  acc = in + dt/6 L(in)
//     STAGE 2: out = base + dt/2 L(in); acc += dt/3 L(in)
//     STAGE 3: out = base + dt L(in);   acc += dt/3 L(in)
//     STAGE 4: out = acc + dt/6 L(in)
template <int INTEG, int STAGE>
__global__ __launch_bounds__(256) void k_stage(Grid g, const double *__restrict__ in, const double *base,
                                               double *out, double *acc, const double *__restrict__ coef,
                                               const double *__restrict__ ctan,
                                               const double *__restrict__ patch,
                                               const int *__restrict__ porg, int has_src, int ngroups) {
  __shared__ double lds[4][4][kLdsCols];
  const int N = g.N, M = g.M;
  const size_t NM = (size_t)N * M;
  const int lane = threadIdx.x, w = threadIdx.y;

  // XCD-aware tile order: the hardware deals consecutive blocks round-robin
  // over the 8 XCDs; remap so each XCD walks a contiguous run of tiles and
  // the theta/rho halos it re-reads stay in its own L2.
  const int nblocks = gridDim.x;
  int tile = blockIdx.x;
  if ((nblocks & 7) == 0) tile = (blockIdx.x & 7) * (nblocks >> 3) + (blockIdx.x >> 3);
  const int strip = tile / ngroups, grp = tile - strip * ngroups;
  const int a0 = strip * kWave;
  const int a = a0 + lane;
  const int b0 = (grp * blockDim.y + w) * kRows;
  double *L = &lds[w][0][0];

  const bool lo_edge = (a0 == 0), hi_edge = (a0 + kWave == N);
  const bool inner_col = (a >= 3 && a < N - 3);
  const bool ko_col = (a >= 4 && a < N - 4);

  const double dt = g.dt;
  const double over_dx = 1.0 / g.dx, over_dth = 1.0 / g.dtheta, over_dt = 1.0 / dt;
  const double wx = kD1 * over_dx;
  const double wy = 0.05 * kThird * over_dth;
  const double wyy = 0.05 * kThird * kThird * over_dth * over_dth;

  // Psi (q) theta window: rows b0-3 .. b0+10 of this column.
  double wqr[kRows + 6], wqi[kRows + 6];
#pragma unroll
  for (int i = 0; i < kRows + 6; i++) {
    const int r = b0 - 3 + i;
    if (r >= 0 && r < M) {
      wqr[i] = in[(size_t)r * N + a];
      wqi[i] = in[NM + (size_t)r * N + a];
    } else {
      wqr[i] = 0.0;
      wqi[i] = 0.0;
    }
  }

  int it0 = INT_MIN / 2, ip0 = INT_MIN / 2;
  if (has_src) {
    it0 = porg[0];
    ip0 = porg[1];
  }

  double lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0};

  // Fill the rho-edge columns of one row in the reference order (kern.cu:412-449).
  auto rho_fix = [&](double (&v)[4]) {
    if (lo_edge) {
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const double v3 = __shfl(v[f], 3), v4 = __shfl(v[f], 4);
        const double v2 = (4.0 * v3 - v4) * kThird;
        const double v1 = (4.0 * v2 - v3) * kThird;
        const double v0 = (4.0 * v1 - v2) * kThird;
        if (lane == 2) v[f] = v2;
        if (lane == 1) v[f] = v1;
        if (lane == 0) v[f] = v0;
      }
    }
    if (hi_edge) {
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const double v4 = __shfl(v[f], kWave - 4), v5 = __shfl(v[f], kWave - 5);
        const double v3 = (4.0 * v4 - v5) * kThird;
        const double v2 = (4.0 * v3 - v4) * kThird;
        const double v1 = (4.0 * v2 - v3) * kThird;
        if (lane == kWave - 3) v[f] = v3;
        if (lane == kWave - 2) v[f] = v2;
        if (lane == kWave - 1) v[f] = v1;
      }
    }
  };
  auto store_row = [&](int b, const double (&v)[4]) {
    const size_t c = (size_t)b * N + a;
#pragma unroll
    for (int f = 0; f < 4; f++) out[f * NM + c] = v[f];
  };

#pragma unroll
  for (int i = 0; i < kRows; i++) {
    const int b = b0 + i;
    if (b < 3 || b >= M - 3) continue;  // uniform over the wavefront
    const size_t c0 = (size_t)b * N;
    const double pr0 = in[2 * NM + c0 + a], pi0 = in[3 * NM + c0 + a];
    const double qr0 = wqr[i + 3], qi0 = wqi[i + 3];

    // Stage the row (4 fields, 64 columns + 4-column halos) in LDS.
    L[0 * kLdsCols + lane + kHalo] = qr0;
    L[1 * kLdsCols + lane + kHalo] = qi0;
    L[2 * kLdsCols + lane + kHalo] = pr0;
    L[3 * kLdsCols + lane + kHalo] = pi0;
    if (lane < 2 * kHalo) {
      const int slot = lane < kHalo ? lane : kWave + lane;
      const int col = a0 - kHalo + slot;
      if (col >= 0 && col < N) {
#pragma unroll
        for (int f = 0; f < 4; f++) L[f * kLdsCols + slot] = in[f * NM + c0 + col];
      }
    }
    wave_sync();

    double v[4] = {0, 0, 0, 0};
    if (inner_col) {
      const int c = lane + kHalo;
      auto dx1 = [&](int f) {
        const double *r = L + f * kLdsCols;
        return (45.0 * r[c + 1] - 45.0 * r[c - 1] - 9.0 * r[c + 2] + 9.0 * r[c - 2] + r[c + 3] - r[c - 3]) * wx;
      };
      auto ko8 = [&](int f) {
        const double *r = L + f * kLdsCols;
        return r[c + 4] + r[c - 4] - 8.0 * (r[c + 3] + r[c - 3]) + 70.0 * r[c] - 56.0 * (r[c + 1] + r[c - 1]) +
               28.0 * (r[c + 2] + r[c - 2]);
      };
      const double ct = ctan[b];
      const double qry = ((wqr[i + 6] - wqr[i]) - 9.0 * (wqr[i + 5] - wqr[i + 1]) + 45.0 * (wqr[i + 4] - wqr[i + 2])) * wy;
      const double qryy = (2.0 * (wqr[i + 6] + wqr[i]) - 27.0 * (wqr[i + 5] + wqr[i + 1]) +
                           270.0 * (wqr[i + 4] + wqr[i + 2]) - 490.0 * wqr[i + 3]) * wyy;
      const double qiy = ((wqi[i + 6] - wqi[i]) - 9.0 * (wqi[i + 5] - wqi[i + 1]) + 45.0 * (wqi[i + 4] - wqi[i + 2])) * wy;
      const double qiyy = (2.0 * (wqi[i + 6] + wqi[i]) - 27.0 * (wqi[i + 5] + wqi[i + 1]) +
                           270.0 * (wqi[i + 4] + wqi[i + 2]) - 490.0 * wqi[i + 3]) * wyy;
      const double llr = qryy + ct * qry;
      const double lli = qiyy + ct * qiy;
      const double qrx = dx1(0), qix = dx1(1), prx = dx1(2), pix = dx1(3);
      double dqr = 0.0, dqi = 0.0, dpr = 0.0, dpi = 0.0;
      if (ko_col) {
        dqr = ko8(0);
        dqi = ko8(1);
        dpr = ko8(2);
        dpi = ko8(3);
      }
      dqr = 0.1 * dqr * 0.0625 * over_dt;
      dqi = 0.1 * dqi * 0.0625 * over_dt;
      dpr = 0.1 * dpr * 0.0625 * over_dt;
      dpi = 0.1 * dpi * 0.0625 * over_dt;

      const size_t id = c0 + a;
      const double bb = coef[id], cc = coef[NM + id], ee = coef[2 * NM + id];
      const double are = coef[3 * NM + id], aim = coef[4 * NM + id];
      const double dre = coef[5 * NM + id], dim = coef[6 * NM + id];
      const double vre = coef[7 * NM + id], vim = coef[8 * NM + id];

      double sr = 0.0, si = 0.0;
      const int sb = b - it0, sa = a - ip0;
      if ((unsigned)sb < (unsigned)g.ptsy && (unsigned)sa < (unsigned)g.ptsx) {
        sr = patch[sb * g.ptsx + sa];
        si = patch[g.ptsx * g.ptsy + sb * g.ptsx + sa];
      }

      double r[4];
      r[0] = -bb * qrx + pr0 - dqr;
      r[1] = -bb * qix + pi0 - dqi;
      r[2] = sr + ee * prx - dpr + cc * llr + dre * qrx - dim * qix - are * pr0 + aim * pi0 - vre * qr0 + vim * qi0;
      r[3] = si + ee * pix - dpi + cc * lli + dre * qix + dim * qrx - are * pi0 - aim * pr0 - vre * qi0 - vim * qr0;

      const double u[4] = {qr0, qi0, pr0, pi0};
      if (INTEG == 0) {
        if (STAGE == 1) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = u[f] + dt * r[f];
        } else if (STAGE == 2) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = 0.25 * (3.0 * base[f * NM + id] + u[f] + dt * r[f]);
        } else {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = kThird * (base[f * NM + id] + 2.0 * u[f] + 2.0 * dt * r[f]);
        }
      } else {
        const double h2 = 0.5 * dt, h6 = dt / 6.0, h3 = dt / 3.0;
        if (STAGE == 1) {
#pragma unroll
          for (int f = 0; f < 4; f++) {
            v[f] = u[f] + h2 * r[f];
            acc[f * NM + id] = u[f] + h6 * r[f];
          }
        } else if (STAGE == 2 || STAGE == 3) {
          const double h = (STAGE == 2) ? h2 : dt;
#pragma unroll
          for (int f = 0; f < 4; f++) {
            v[f] = base[f * NM + id] + h * r[f];
            acc[f * NM + id] = acc[f * NM + id] + h3 * r[f];
          }
        } else {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = acc[f * NM + id] + h6 * r[f];
        }
      }
    }
    wave_sync();  // LDS row is free for the next iteration
    rho_fix(v);
    store_row(b, v);
    if (b == 3 || b == M - 5) {
#pragma unroll
      for (int f = 0; f < 4; f++) lo[f] = v[f];
    }
    if (b == 4 || b == M - 4) {
#pragma unroll
      for (int f = 0; f < 4; f++) hi[f] = v[f];
    }
  }

  // Theta edge rows (kern.cu:320-409), then their rho edges.
  if (b0 == 0) {
    double r2[4], r1[4], r0[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
      if (g.even_m) {
        r2[f] = (4.0 * lo[f] - hi[f]) * kThird;
        r1[f] = (4.0 * r2[f] - lo[f]) * kThird;
        r0[f] = (4.0 * r1[f] - r2[f]) * kThird;
      } else {
        r2[f] = r1[f] = r0[f] = 0.0;
      }
    }
    rho_fix(r2);
    rho_fix(r1);
    rho_fix(r0);
    store_row(2, r2);
    store_row(1, r1);
    store_row(0, r0);
  }
  if (b0 + kRows == M) {
    double s3[4], s2[4], s1[4];
#pragma unroll
    for (int f = 0; f < 4; f++) {
      if (g.even_m) {
        s3[f] = (4.0 * hi[f] - lo[f]) * kThird;
        s2[f] = (4.0 * s3[f] - hi[f]) * kThird;
        s1[f] = (4.0 * s2[f] - s3[f]) * kThird;
      } else {
        s3[f] = s2[f] = s1[f] = 0.0;
      }
    }
    rho_fix(s3);
    rho_fix(s2);
    rho_fix(s1);
    store_row(M - 3, s3);
    store_row(M - 2, s2);
    store_row(M - 1, s1);
  }
}

// ---------------------------------------------------------------------------
// Extraction (kernel_boundary2b's column copy, kern.cu:750-751) fused with
// the host-side bookkeeping of body.cu:174-202: the rectangle-rule double
// time integral, the energy flux and, every Fr steps, the 14 projections
// onto -2Y_lm (proj.c:37-75).  Arithmetic uses explicit round-to-nearest
// operations in the reference's evaluation order, so the records equal the
// host computation bit for bit given equal field values.  Also advances the
// step counter and the solver time.
__global__ __launch_bounds__(256) void k_extract(Grid g, const double *__restrict__ U, double *__restrict__ ext,
                                                 const double *__restrict__ extw,
                                                 const double *__restrict__ projw, double *__restrict__ rec,
                                                 int *__restrict__ rec_k, double *__restrict__ psi, Ctl *ctl) {
  __shared__ double term[256];
  __shared__ double dint[2][256];
  __shared__ int bad;
  const int q = threadIdx.x, M = g.M, N = g.N;
  const size_t NM = (size_t)N * M;
  const int k = ctl->k;
  const double timer = ctl->timer;
  const int r = ctl->nrec;
  const bool out_step = (k % g.Fr) == 0 && r < g.nrec_cap;
  const double dt = g.dt;
  if (q == 0) bad = 0;
  __syncthreads();
  if (q < M) {
    const double pr = U[(size_t)q * N + g.je], pi = U[NM + (size_t)q * N + g.je];
    const double ir = __dadd_rn(ext[q], __dmul_rn(pr, dt));
    const double ii = __dadd_rn(ext[M + q], __dmul_rn(pi, dt));
    const double psi2 = __dadd_rn(__dmul_rn(ir, ir), __dmul_rn(ii, ii));
    term[q] = __ddiv_rn(__dmul_rn(__dmul_rn(psi2, extw[q]), extw[2 * M]), extw[M + q]);
    const double dr = __dadd_rn(ext[2 * M + q], __dmul_rn(ir, dt));
    const double di = __dadd_rn(ext[3 * M + q], __dmul_rn(ii, dt));
    ext[q] = ir;
    ext[M + q] = ii;
    ext[2 * M + q] = dr;
    ext[3 * M + q] = di;
    dint[0][q] = dr;
    dint[1][q] = di;
    if (!isfinite(pr) || !isfinite(pi)) bad = 1;
    if (out_step) {
      psi[((size_t)r * 2) * M + q] = pr;
      psi[((size_t)r * 2 + 1) * M + q] = pi;
    }
  }
  __syncthreads();
  if (out_step) {
    if (q == 0) {
      double dedt = 0.0;
      for (int i = 0; i < M; i++) dedt = __dadd_rn(dedt, term[i]);
      rec[(size_t)r * 16] = timer;
      rec[(size_t)r * 16 + 1] = dedt;
      rec_k[2 * r] = k;
      rec_k[2 * r + 1] = ((double)k * dt) >= (g.xmax + 200.0) ? 1 : 0;
    }
    if (q >= 64 && q < 64 + 14) {
      const int l = (q - 64) >> 1, part = (q - 64) & 1;
      const int nterm = 5 * (M / 4);
      const double *wl = projw + (size_t)l * nterm * 3;
      const double *dq = dint[part];
      double h = 0.0;
      for (int t = 0; t < nterm; t++) {
        const int pt = 4 * (t / 5) + (t % 5);
        const double v = dq[pt < M ? pt : M - 1];
        const double term = __dmul_rn(__dmul_rn(__dmul_rn(__dmul_rn(wl[3 * t], v), wl[3 * t + 1]), g.dtheta), wl[3 * t + 2]);
        h = __dadd_rn(h, term);
      }
      h = __ddiv_rn(__dmul_rn(2.0, h), 45.0);
      rec[(size_t)r * 16 + 2 + (q - 64)] = part ? -h : h;
    }
  }
  __syncthreads();
  if (q == 0) {
    ctl->k = k + 1;
    ctl->timer = timer + dt;
    if (out_step) ctl->nrec = r + 1;
    if (bad) ctl->status = 1;
  }
}

// ---------------------------------------------------------------------------
// Particle source on the compact ptsx x ptsy patch (kern.cu:770-2683).
#include "teuk_source_dev.inc"

}  // namespace

bool source_available() { return kSourceAvailable; }

// ---------------------------------------------------------------------------
// Launchers.
hipError_t launch_init_coeffs(const Grid &g, const DevState &d, hipStream_t s) {
  const size_t NM = (size_t)g.N * g.M;
  const unsigned nb = (unsigned)((NM + 255) / 256);
  hipLaunchKernelGGL(k_init_coeffs, dim3(nb), dim3(256), 0, s, g, d.col4, d.row4, d.r_c, d.coef);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_fix_dre, dim3(nb), dim3(256), 0, s, g, d.coef);
  return hipGetLastError();
}

static inline int waves_per_block(int M) {
  const int chunks = M / kRows;
  return chunks >= 4 ? 4 : (chunks >= 2 ? 2 : 1);
}

hipError_t launch_stage(const Grid &g, const DevState &d, int integrator, int stage, bool has_src,
                        hipStream_t s) {
  const int wpb = waves_per_block(g.M);
  const int ngroups = g.M / (kRows * wpb);
  const int nb = (g.N / kWave) * ngroups;
  const dim3 grid(nb), block(kWave, wpb);
  const int hs = has_src ? 1 : 0;
  if (integrator == 0) {
    switch (stage) {
      case 1:
        hipLaunchKernelGGL((k_stage<0, 1>), grid, block, 0, s, g, d.U, d.U, d.C, nullptr, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      case 2:
        hipLaunchKernelGGL((k_stage<0, 2>), grid, block, 0, s, g, d.C, d.U, d.D, nullptr, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      case 3:
        hipLaunchKernelGGL((k_stage<0, 3>), grid, block, 0, s, g, d.D, d.U, d.U, nullptr, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      default:
        return hipErrorInvalidValue;
    }
  } else {
    switch (stage) {
      case 1:
        hipLaunchKernelGGL((k_stage<1, 1>), grid, block, 0, s, g, d.U, d.U, d.C, d.A, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      case 2:
        hipLaunchKernelGGL((k_stage<1, 2>), grid, block, 0, s, g, d.C, d.U, d.D, d.A, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      case 3:
        hipLaunchKernelGGL((k_stage<1, 3>), grid, block, 0, s, g, d.D, d.U, d.C, d.A, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      case 4:
        hipLaunchKernelGGL((k_stage<1, 4>), grid, block, 0, s, g, d.C, d.U, d.U, d.A, d.coef, d.ctan,
                           d.patch, d.patch_org, hs, ngroups);
        break;
      default:
        return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}

hipError_t launch_extract(const Grid &g, const DevState &d, hipStream_t s) {
  hipLaunchKernelGGL(k_extract, dim3(1), dim3(256), 0, s, g, d.U, d.ext, d.extw, d.projw, d.rec, d.rec_k, d.psi,
                     d.ctl);
  return hipGetLastError();
}

hipError_t launch_source(const Grid &g, const DevState &d, hipStream_t s) {
  const int cells = g.ptsx * g.ptsy;
  hipLaunchKernelGGL(k_source, dim3((cells + 255) / 256), dim3(256), 0, s, g, d.theta, d.r_c, d.steps,
                     d.nsteps_src, d.patch, d.patch_org, d.ctl);
  return hipGetLastError();
}

}  // namespace teuk
