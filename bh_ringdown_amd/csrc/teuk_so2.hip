// HIP kernels of the staggered second-order scheme of TD_GPU_code/Old
// (integrator 2) on gfx950.
//
// The reference runs nine kernels and a host read-back per step
// (Old/body.cu:121-243): average to the half points, source, rhs, half
// update, theta edges; average back, source, rhs, full update, theta edges,
// rho edges.  Here one step is two launches:
//   k_so2_half     averages, right-hand side and half update at the half
//                  points rho_{a-1/2}, straight from the grid-point fields
//                  (Old/kern.cu:196-303);
//   k_so2_full     re-averages (with the half-point theta edge rows of
//                  :305-340 formed on the fly), right-hand side and full
//                  update at the grid points in place (:346-453), then theta
//                  edge rows, rho edge columns (:456-522) and the time
//                  integrals of Psi at the extraction column
//                  (Old/body.cu:199-216), the edges through LDS.
// Each thread owns one point; a wavefront covers 64 consecutive rho columns
// of one theta row, so every access is coalesced; the neighbour reads (one
// column to the side, one row up and down) are cache hits.  A complex value
// (re, im) is one 16-byte access, addressed by a 32-bit byte offset from the
// plane base.  The coefficient fields are rebuilt per point from the column
// constants of their point set, staged once per block in LDS, and the row
// values, as in the high-order stage kernel.
//
// All arithmetic goes through teuk_point.h, built without contraction, so the
// results equal the oracle's bit for bit.
#include "teuk_internal.h"

namespace teuk {
namespace {

using teuk_pt::Coef;

constexpr int kSo2Rows = 4;   // theta rows per block (64 x 4 threads)

// Field f (q_re, q_im, p_re, p_im) of point c: a state is two complex planes
// (Psi, Pi), each [M][N] of (re, im).
__device__ __forceinline__ size_t fidx(int f, size_t c, size_t NM) {
  return (size_t)(f >> 1) * 2 * NM + 2 * c + (size_t)(f & 1);
}

__device__ __forceinline__ Coef so2_coef(const double *__restrict__ ktab, int N, int a, double sn2, double cs,
                                         double Y) {
  double K[kNK];
#pragma unroll
  for (int i = 0; i < kNK; i++) K[i] = ktab[(size_t)i * N + a];
  return teuk_pt::coef_point(K, sn2, cs, Y, [](double x) { return __dsqrt_rn(x); });
}

// Complex value of plane h (0: Psi, 1: Pi) at point c: byte h cb + 16 c.
__device__ __forceinline__ double2 cld(const double *base, unsigned off) {
  return *reinterpret_cast<const double2 *>(reinterpret_cast<const char *>(base) + off);
}
__device__ __forceinline__ void cst(double *base, unsigned off, double re, double im) {
  *reinterpret_cast<double2 *>(reinterpret_cast<char *>(base) + off) = make_double2(re, im);
}

// The column constants of the block's 64 columns into LDS (Kl[i][tx]); every
// thread of the block takes part, then the block waits.
__device__ __forceinline__ void so2_stage_K(const double *__restrict__ ktab, int N, int a0,
                                            double (&Kl)[kNK][kWave]) {
  const int tx = threadIdx.x, ty = threadIdx.y;
  for (int i = ty; i < kNK; i += kSo2Rows) Kl[i][tx] = ktab[(size_t)i * N + a0 + tx];
  __syncthreads();
}

// bb on rows 1 <= k < M-1 (Old/kern.cu:90 and :143, bb at :123 and :176) of
// one point set into a scratch plane, then its second-order rho derivative
// on 1 <= a < N-1 (Old/kern.cu:231-233, :381-383), zero elsewhere.
__global__ __launch_bounds__(256) void k_so2_init_bb(Grid g, const double *__restrict__ ktab,
                                                     const double *__restrict__ rtab, double *__restrict__ bbp) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int k = (int)(id / g.N), j = (int)(id - (size_t)k * g.N);
  double bb = 0.0;
  if (k >= 1 && k < g.M - 1) bb = so2_coef(ktab, g.N, j, rtab[k], rtab[g.M + k], rtab[2 * g.M + k]).bb;
  bbp[id] = bb;
}

__global__ __launch_bounds__(256) void k_so2_init_dbdx(Grid g, const double *__restrict__ bbp,
                                                       double *__restrict__ dbdx) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int a = (int)(id % g.N);
  double v = 0.0;
  if (a >= 1 && a < g.N - 1) v = teuk_pt::so2_dbdx(bbp[id + 1], bbp[id - 1], g.dx);
  dbdx[id] = v;
}

// The source value of point (b, a) from the patch of half step kh, or zero.
__device__ __forceinline__ void so2_source(const Grid &g, const DevState &d, int kh, int b, int a, double &sr,
                                           double &si) {
  const int slot = kh % kRing;
  const int it0 = d.patch_org[2 * slot], ip0 = d.patch_org[2 * slot + 1];
  const int sb = b - it0, sa = a - ip0;
  sr = 0.0;
  si = 0.0;
  if ((unsigned)sb < (unsigned)g.ptsy && (unsigned)sa < (unsigned)g.ptsx) {
    const double *patch = d.patch + (size_t)slot * 2 * g.ptsx * g.ptsy;
    const int pc = sb * g.ptsx + sa;
    sr = patch[pc];
    si = patch[g.ptsx * g.ptsy + pc];
  }
}

// Right-hand side at one point of either set (Old/kern.cu:214-284,
// :365-434): pv = the averaged fields at rows b-1, b, b+1 (q) and b (p),
// dxv = the four rho differences.
__device__ __forceinline__ void so2_rhs_point(const Grid &g, const DevState &d, const double (&Kl)[kNK][kWave],
                                              const double *__restrict__ dbdx, int b, int a, const double qr[3],
                                              const double qi[3], double pr, double pi, const double dxv[4],
                                              double sr, double si, double r[4]) {
  const double *__restrict__ rtab = d.rtab;
  const int M = g.M;
  double K[kNK];
#pragma unroll
  for (int i = 0; i < kNK; i++) K[i] = Kl[i][threadIdx.x];
  const Coef cf = teuk_pt::coef_point(K, rtab[b], rtab[M + b], rtab[2 * M + b], [](double x) { return __dsqrt_rn(x); });
  const double dre = teuk_pt::fma_(-dbdx[(size_t)b * g.N + a], cf.ee, cf.dre);
  const double ct = rtab[3 * M + b];   // cos/sin (Old/kern.cu:241-243)
  teuk_pt::Pt x;
  x.qr = qr[1];
  x.qi = qi[1];
  x.pr = pr;
  x.pi = pi;
  x.llr = teuk_pt::fma_(ct, teuk_pt::so2_dth(qr[2], qr[0], g.dtheta), teuk_pt::so2_dth2(qr[1], qr[2], qr[0], g.dtheta));
  x.lli = teuk_pt::fma_(ct, teuk_pt::so2_dth(qi[2], qi[0], g.dtheta), teuk_pt::so2_dth2(qi[1], qi[2], qi[0], g.dtheta));
  x.qrx = dxv[0];
  x.qix = dxv[1];
  x.prx = dxv[2];
  x.pix = dxv[3];
  x.kqr = x.kqi = x.kpr = x.kpi = 0.0;
  x.sr = sr;
  x.si = si;
  teuk_pt::rhs(cf, dre, x, r);
}

// Half step: X_h = avg(X[a], X[a-1]) + dt/2 L_h, rows 1..M-2, 1 <= a < N
// (Old/kern.cu:196-303).  Reads X (d.U) only, writes X_h (d.C).
__global__ __launch_bounds__(256) void k_so2_half(Grid g, DevState d, int has_src, int k_rel) {
  __shared__ double Kl[kNK][kWave];
  const int a0 = blockIdx.x * kWave;
  const int a = a0 + threadIdx.x;
  // the row is uniform over the wavefront: row tables become scalar loads
  const int b = blockIdx.y * kSo2Rows + __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int N = g.N, M = g.M;
  so2_stage_K(d.ktab_h, N, a0, Kl);
  if (a < 1 || a >= N || b < 1 || b >= M - 1) return;
  const unsigned cb = 16u * (unsigned)N * (unsigned)M;
  const double *__restrict__ X = d.U;
  double *__restrict__ Xh = d.C;
  const unsigned c = 16u * (unsigned)(b * N + a);
  const unsigned rowb = 16u * (unsigned)N;
  double qr[3], qi[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const unsigned ci = c + rowb * i - rowb;
    const double2 x1 = cld(X, ci), x0 = cld(X, ci - 16u);
    qr[i] = teuk_pt::so2_avg(x1.x, x0.x);
    qi[i] = teuk_pt::so2_avg(x1.y, x0.y);
  }
  const double2 p1 = cld(X, cb + c), p0 = cld(X, cb + c - 16u);
  const double2 q1 = cld(X, c), q0 = cld(X, c - 16u);
  const double pr = teuk_pt::so2_avg(p1.x, p0.x);
  const double pi = teuk_pt::so2_avg(p1.y, p0.y);
  double dxv[4];
  dxv[0] = teuk_pt::so2_dx(q1.x, q0.x, g.dx);
  dxv[1] = teuk_pt::so2_dx(q1.y, q0.y, g.dx);
  dxv[2] = teuk_pt::so2_dx(p1.x, p0.x, g.dx);
  dxv[3] = teuk_pt::so2_dx(p1.y, p0.y, g.dx);
  double sr = 0.0, si = 0.0;
  if (has_src) so2_source(g, d, 2 * (d.ctl->k + k_rel), b, a, sr, si);
  double r[4];
  so2_rhs_point(g, d, Kl, d.dbdx_h, b, a, qr, qi, pr, pi, dxv, sr, si, r);
  const double h = 0.5 * g.dt;
  cst(Xh, c, teuk_pt::axpy(h, r[0], qr[1]), teuk_pt::axpy(h, r[1], qi[1]));
  cst(Xh, cb + c, teuk_pt::axpy(h, r[2], pr), teuk_pt::axpy(h, r[3], pi));
}

// Full step, edges and extraction in one pass (Old/kern.cu:305-522 and
// Old/body.cu:199-216).  A block holds 64 columns x 4 rows.
//  1. Interior points (rows 1..M-2, 1 <= a < N-1): X += dt L_c with
//     X_c = avg(X_h[a], X_h[a+1]); the half-point fields' theta edge rows
//     (kernel_boundary1, :305-340), needed only here, are formed on the fly
//     from rows 1, 2 / M-2, M-3 -- the same operations in the same order.
//     Each thread reads only its own point of X, so X is updated in place.
//  2. Theta edge rows of X, 1 <= a <= N-2 (:456-497), from the block's new
//     rows 1, 2 / M-2, M-3 through LDS (the first and last block rows hold
//     them; column N-1 is the rho fill's).
//  3. Rho edge columns of every row (:499-522), from the block's columns
//     1, 2 / N-3, N-2 (the first and last block columns), after 2.
//  4. The thread at the extraction column advances its row's time integrals
//     from the final value and snapshots them every Fr steps.
__global__ __launch_bounds__(256) void k_so2_full(Grid g, DevState d, int has_src, int k_rel) {
  __shared__ double tile[4][kSo2Rows][kWave];
  __shared__ double Kl[kNK][kWave];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int a0 = blockIdx.x * kWave;
  const int a = a0 + tx;
  // the row is uniform over the wavefront: row tables become scalar loads
  const int b = blockIdx.y * kSo2Rows + __builtin_amdgcn_readfirstlane(ty);
  const int N = g.N, M = g.M;
  const size_t NM = (size_t)N * M;
  const unsigned cb = 16u * (unsigned)N * (unsigned)M;
  const double *__restrict__ Xh = d.C;
  double *__restrict__ X = d.U;
  const size_t c = (size_t)b * N + a;
  const unsigned cB = 16u * (unsigned)c;
  so2_stage_K(d.ktab, N, a0, Kl);
  const bool inner = a >= 1 && a < N - 1 && b >= 1 && b < M - 1;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (inner) {
    // Psi_h (re, im) at row rr, columns a and a+1, with the theta edge rows
    // of kernel_boundary1 formed from the next two rows
    auto xh = [&](int rr, double2 &c0, double2 &c1) {
      auto at = [&](int row, double2 &v0, double2 &v1) {
        const unsigned o = 16u * (unsigned)(row * N + a);
        v0 = cld(Xh, o);
        v1 = cld(Xh, o + 16u);
      };
      if (rr == 0 || rr == M - 1) {
        double2 e1a, e1b, e2a, e2b;
        at(rr == 0 ? 1 : M - 2, e1a, e1b);
        at(rr == 0 ? 2 : M - 3, e2a, e2b);
        if (g.even_m) {
          c0 = make_double2(teuk_pt::so2_edge(e1a.x, e2a.x), teuk_pt::so2_edge(e1a.y, e2a.y));
          c1 = make_double2(teuk_pt::so2_edge(e1b.x, e2b.x), teuk_pt::so2_edge(e1b.y, e2b.y));
        } else {
          c0 = c1 = make_double2(0.0, 0.0);
        }
        return;
      }
      at(rr, c0, c1);
    };
    double qr[3], qi[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double2 x0, x1;
      xh(b + i - 1, x0, x1);
      qr[i] = teuk_pt::so2_avg(x0.x, x1.x);
      qi[i] = teuk_pt::so2_avg(x0.y, x1.y);
    }
    const double2 q0 = cld(Xh, cB), q1 = cld(Xh, cB + 16u);
    const double2 p0 = cld(Xh, cb + cB), p1 = cld(Xh, cb + cB + 16u);
    const double pr = teuk_pt::so2_avg(p0.x, p1.x);
    const double pi = teuk_pt::so2_avg(p0.y, p1.y);
    double dxv[4];
    dxv[0] = teuk_pt::so2_dx(q1.x, q0.x, g.dx);
    dxv[1] = teuk_pt::so2_dx(q1.y, q0.y, g.dx);
    dxv[2] = teuk_pt::so2_dx(p1.x, p0.x, g.dx);
    dxv[3] = teuk_pt::so2_dx(p1.y, p0.y, g.dx);
    double sr = 0.0, si = 0.0;
    if (has_src) so2_source(g, d, 2 * (d.ctl->k + k_rel) + 1, b, a, sr, si);
    double r[4];
    so2_rhs_point(g, d, Kl, d.dbdx, b, a, qr, qi, pr, pi, dxv, sr, si, r);
    const double2 u0 = cld(X, cB), u1 = cld(X, cb + cB);
    v[0] = teuk_pt::axpy(g.dt, r[0], u0.x);
    v[1] = teuk_pt::axpy(g.dt, r[1], u0.y);
    v[2] = teuk_pt::axpy(g.dt, r[2], u1.x);
    v[3] = teuk_pt::axpy(g.dt, r[3], u1.y);
    cst(X, cB, v[0], v[1]);
    cst(X, cb + cB, v[2], v[3]);
  }
#pragma unroll
  for (int f = 0; f < 4; f++) tile[f][ty][tx] = v[f];
  __syncthreads();
  // 2. theta edge rows
  const bool theta_edge = (b == 0 || b == M - 1) && a >= 1 && a <= N - 2;
  if (theta_edge) {
    const int t1 = b == 0 ? 1 : ty - 1, t2 = b == 0 ? 2 : ty - 2;
#pragma unroll
    for (int f = 0; f < 4; f++) {
      v[f] = g.even_m ? teuk_pt::so2_edge(tile[f][t1][tx], tile[f][t2][tx]) : 0.0;
      X[fidx(f, c, NM)] = v[f];
    }
  }
  __syncthreads();
  if (theta_edge) {
#pragma unroll
    for (int f = 0; f < 4; f++) tile[f][ty][tx] = v[f];
  }
  __syncthreads();
  // 3. rho edge columns
  if (a == 0 || a == N - 1) {
    const int t1 = a == 0 ? 1 : tx - 1, t2 = a == 0 ? 2 : tx - 2;
#pragma unroll
    for (int f = 0; f < 4; f++) {
      v[f] = teuk_pt::so2_edge(tile[f][ty][t1], tile[f][ty][t2]);
      X[fidx(f, c, NM)] = v[f];
    }
  }
  // 4. extraction (Old/body.cu:209-216), every step
  if (a == g.je) {
    const Ext &ext = d.ext;
    Ctl *ctl = d.ctl;
    const int kstep = ctl->k + k_rel;
    const double dt = g.dt;
    const double q0 = v[0], q1 = v[1];
    const double ir = ext.integ[b] + q0 * dt;
    const double ii = ext.integ[M + b] + q1 * dt;
    const double dr = ext.integ[2 * M + b] + ir * dt;
    const double di = ext.integ[3 * M + b] + ii * dt;
    ext.integ[b] = ir;
    ext.integ[M + b] = ii;
    ext.integ[2 * M + b] = dr;
    ext.integ[3 * M + b] = di;
    if (!isfinite(q0) || !isfinite(q1)) ctl->status = 1;
    if ((kstep % g.Fr) == 0 && kstep / g.Fr < g.nrec_cap) {
      double *sn = ext.snap + (size_t)(kstep / g.Fr) * 6 * M;
      sn[b] = q0;
      sn[M + b] = q1;
      sn[2 * M + b] = ir;
      sn[3 * M + b] = ii;
      sn[4 * M + b] = dr;
      sn[5 * M + b] = di;
    }
  }
}

}  // namespace

hipError_t launch_init_so2(const Grid &g, const DevState &d, hipStream_t s) {
  const size_t NM = (size_t)g.N * g.M;
  const dim3 nb((unsigned)((NM + 255) / 256));
  hipLaunchKernelGGL(k_so2_init_bb, nb, dim3(256), 0, s, g, d.ktab_h, d.rtab, d.scratch);
  hipLaunchKernelGGL(k_so2_init_dbdx, nb, dim3(256), 0, s, g, d.scratch, d.dbdx_h);
  hipLaunchKernelGGL(k_so2_init_bb, nb, dim3(256), 0, s, g, d.ktab, d.rtab, d.scratch);
  hipLaunchKernelGGL(k_so2_init_dbdx, nb, dim3(256), 0, s, g, d.scratch, d.dbdx);
  return hipGetLastError();
}

hipError_t launch_so2_step(const Grid &g, const DevState &d, bool has_src, int k_rel, hipStream_t s) {
  if (g.N % kWave != 0 || g.M % kSo2Rows != 0 || g.M > 256) return hipErrorInvalidValue;
  const dim3 grid(g.N / kWave, g.M / kSo2Rows), block(kWave, kSo2Rows);
  const int hs = has_src ? 1 : 0;
  hipLaunchKernelGGL(k_so2_half, grid, block, 0, s, g, d, hs, k_rel);
  hipLaunchKernelGGL(k_so2_full, grid, block, 0, s, g, d, hs, k_rel);
  return hipGetLastError();
}

}  // namespace teuk
