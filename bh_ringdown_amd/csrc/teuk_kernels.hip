// HIP kernels of libteuk other than the Runge-Kutta stage (teuk_stage.inc):
// the time-independent coefficient planes, the particle-source patches, the
// output records and the whole-field sentinel, and the launchers of all of
// them, the stage included (launch_stage picks the stage build for the
// launch shape's theta rows).
#include "teuk_dev.h"

namespace teuk {
namespace {


// ---------------------------------------------------------------------------
// bb on interior rows (kern.cu:147) into a scratch plane, then its 6th-order
// rho derivative (kern.cu:192-196), which is time independent.
__global__ __launch_bounds__(256) void k_init_bb(Grid g, const double *__restrict__ ktab,
                                                 const double *__restrict__ rtab, double *__restrict__ bbp) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int k = (int)(id / g.N), j = (int)(id - (size_t)k * g.N);
  double bb = 0.0;
  if (k >= 3 && k < g.M - 3) {
    double K[kNK];
#pragma unroll
    for (int i = 0; i < kNK; i++) K[i] = ktab[(size_t)i * g.N + j];
    bb = coef_point(K, rtab[k], rtab[g.M + k], rtab[2 * g.M + k]).bb;
  }
  bbp[id] = bb;
}

__global__ __launch_bounds__(256) void k_init_dbdx(Grid g, const double *__restrict__ bbp,
                                                   double *__restrict__ dbdx) {
  const size_t NM = (size_t)g.N * g.M;
  const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= NM) return;
  const int a = (int)(id % g.N);
  double d = 0.0;
  if (a >= 3 && a < g.N - 3)
    d = teuk_pt::d1(bbp[id + 1], bbp[id - 1], bbp[id + 2], bbp[id - 2], bbp[id + 3], bbp[id - 3],
                    kD1 * (1.0 / g.dx));
  dbdx[id] = d;
}


// ---------------------------------------------------------------------------
// Output records for a batch of n completed steps (body.cu:174-212): for each
// step k with k % Fr == 0, the energy flux from the snapshot of int Psi dt
// and the 14 projections of int int Psi dt onto -2Y_lm (proj.c:37-75); the
// solver time advanced by repeated addition (body.cu:214).  Explicit
// round-to-nearest operations in the reference's evaluation order make the
// records equal the host computation bit for bit.  Then advances the step
// counter by n.
__device__ __forceinline__ void finalize_solve(const Grid &g, const Ext &ext_in, Ctl *ctl_in, int n) {
  const int M = g.M, t = threadIdx.x;
  // global-memory views of the solve's buffers (see global_ptr)
  struct {
    const TEUK_GLOBAL double *snap, *extw, *projw;
    TEUK_GLOBAL double *rec;
    TEUK_GLOBAL int *rec_k;
  } ext = {global_ptr(const_cast<const double *>(ext_in.snap)), global_ptr(ext_in.extw), global_ptr(ext_in.projw),
           global_ptr(ext_in.rec), global_ptr(ext_in.rec_k)};
  TEUK_GLOBAL Ctl *ctl = global_ptr(ctl_in);
  const int k0 = ctl->k;
  const double timer0 = ctl->timer;
  const int r_lo = (k0 + g.Fr - 1) / g.Fr;                 // first record in [k0, k0+n)
  int r_hi = (k0 + n - 1) / g.Fr + 1;                       // one past the last
  if (r_hi > g.nrec_cap) r_hi = g.nrec_cap;
  const int nrec = r_hi > r_lo ? r_hi - r_lo : 0;
  const double dt = g.dt;
  const int nterm = 5 * (M / 4);
  for (int task = t; task < nrec * 15; task += blockDim.x) {
    const int r = r_lo + task / 15, part = task % 15;
    const TEUK_GLOBAL double *sn = ext.snap + (size_t)r * 6 * M;
    TEUK_GLOBAL double *rec = ext.rec + (size_t)r * 16;
    // The sums run in the reference's order; their inputs are loaded a group
    // at a time (M is a multiple of 8, so groups of 8 rows and of 5 Boole
    // terms divide the loops) so the loads overlap instead of each waiting
    // out the memory latency in turn.
    if (part == 14) {
      double dedt = 0.0;
      const double r6 = ext.extw[2 * M];
      for (int q0 = 0; q0 < M; q0 += 8) {
        double ir[8], ii[8], w0[8], w1[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          ir[j] = sn[2 * M + q0 + j];
          ii[j] = sn[3 * M + q0 + j];
          w0[j] = ext.extw[q0 + j];
          w1[j] = ext.extw[M + q0 + j];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const double psi2 = __dadd_rn(__dmul_rn(ir[j], ir[j]), __dmul_rn(ii[j], ii[j]));
          dedt = __dadd_rn(dedt, __ddiv_rn(__dmul_rn(__dmul_rn(psi2, w0[j]), r6), w1[j]));
        }
      }
      rec[1] = dedt;
    } else {
      const int l = part >> 1, im = part & 1;
      const TEUK_GLOBAL double *wl = ext.projw + (size_t)l * nterm * 3;
      const TEUK_GLOBAL double *dq = sn + (4 + im) * M;
      double h = 0.0;
      for (int g5 = 0; g5 < nterm / 5; g5++) {   // Boole group g5: rows 4 g5 .. 4 g5 + 4
        double v[5], w[15];
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const int pt = 4 * g5 + j;
          v[j] = dq[pt < M ? pt : M - 1];
        }
#pragma unroll
        for (int j = 0; j < 15; j++) w[j] = wl[15 * g5 + j];
#pragma unroll
        for (int j = 0; j < 5; j++) {
          const double term =
              __dmul_rn(__dmul_rn(__dmul_rn(__dmul_rn(w[3 * j], v[j]), w[3 * j + 1]), g.dtheta), w[3 * j + 2]);
          h = __dadd_rn(h, term);
        }
      }
      h = __ddiv_rn(__dmul_rn(2.0, h), 45.0);
      rec[2 + part] = im ? -h : h;
    }
  }
  __syncthreads();
  if (t == 0) {
    double timer = timer0;
    for (int k = k0; k < k0 + n; k++) {
      if (k % g.Fr == 0 && k / g.Fr < g.nrec_cap) {
        const int r = k / g.Fr;
        ext.rec[(size_t)r * 16] = timer;
        ext.rec_k[2 * r] = k;
        ext.rec_k[2 * r + 1] = ((double)k * dt) >= (g.xmax + 200.0) ? 1 : 0;
      }
      timer = __dadd_rn(timer, dt);
    }
    ctl->k = k0 + n;
    ctl->timer = timer;
    ctl->nrec = r_hi > ctl->nrec ? r_hi : ctl->nrec;
  }
}

__global__ __launch_bounds__(256) void k_finalize(Grid g, Ext ext, Ctl *ctl, int n) { finalize_solve(g, ext, ctl, n); }

// One block per solve of the batch; a solve closes at its scheduled step count.
__global__ __launch_bounds__(256) void k_finalize_batch(const SolveDesc *__restrict__ descs, int n) {
  const SolveDesc &sd = descs[blockIdx.x];
  const int left = sd.nt - sd.d.ctl->k;
  const int n_eff = n < left ? n : left;
  if (n_eff <= 0) return;
  finalize_solve(sd.g, sd.d.ext, sd.d.ctl, n_eff);
}

// Whole-field sentinel (no counterpart in the reference, which checks
// nothing): every value of the current fields, as (re, im) pairs, against
// non-finite values and magnitudes above kGrowthLimit; a hit sets status
// bit 1 of the solve.  `!(|x| <= L)` is also true for NaN.  Grid-stride over
// the 2 N M pairs of the state.
__device__ __forceinline__ void sentinel_scan(const double2 *__restrict__ u, size_t n2, Ctl *ctl, size_t i0,
                                              size_t stride) {
  const TEUK_GLOBAL dvec2 *__restrict__ ug = global_ptr(reinterpret_cast<const dvec2 *>(u));
  bool bad = false;
  for (size_t i = i0; i < n2; i += stride) {
    const dvec2 v = ug[i];
    bad = bad || !(fabs(v.x) <= kGrowthLimit) || !(fabs(v.y) <= kGrowthLimit);
  }
  if (bad) atomicOr(&ctl->status, 2);
}

__global__ __launch_bounds__(256) void k_sentinel(const double2 *__restrict__ u, size_t n2, Ctl *ctl) {
  sentinel_scan(u, n2, ctl, (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// blockIdx.y = member; a member that had finished before this chunk (its
// step counter, which the chunk's finalize advances afterwards, is at its
// scheduled count) is skipped.
__global__ __launch_bounds__(256) void k_sentinel_batch(const SolveDesc *__restrict__ descs) {
  const SolveDesc &sd = descs[blockIdx.y];
  if (sd.d.ctl->k >= sd.nt) return;
  const size_t n2 = 2 * (size_t)sd.g.N * sd.g.M;
  sentinel_scan(reinterpret_cast<const double2 *>(sd.d.U), n2, sd.d.ctl,
                (size_t)blockIdx.x * blockDim.x + threadIdx.x, (size_t)gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------------------
// Particle source on the compact ptsx x ptsy patch (kern.cu:770-2683).
#include "teuk_source_dev.inc"

}  // namespace

bool source_available() { return true; }

// ---------------------------------------------------------------------------
// Launchers.
hipError_t launch_init_coeffs(const Grid &g, const DevState &d, hipStream_t s) {
  const size_t NM = (size_t)g.N * g.M;
  const unsigned nb = (unsigned)((NM + 255) / 256);
  hipLaunchKernelGGL(k_init_bb, dim3(nb), dim3(256), 0, s, g, d.ktab, d.rtab, d.scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_init_dbdx, dim3(nb), dim3(256), 0, s, g, d.scratch, d.dbdx);
  return hipGetLastError();
}

// Launch shape for an N x M grid, from measurements on an MI355X: 2-row
// wavefronts (2 + (M - 10) / 2 per strip, an odd count, so one wavefront per
// block), compiled for 3 waves per SIMD from M = 64 up; 4-row wavefronts at
// 8 M points and more, where there are wavefronts enough and fewer theta
// halo re-reads pay.
Tiling default_tiling(int N, int M) {
  Tiling t;
  const bool large = (long long)N * M >= (1LL << 23);
  t.rows = large ? 4 : kDefaultThetaRows;
  t.wpe = (!large && M >= 64) ? 3 : 2;
  const int chunks = theta_chunks(M, t.rows);
  t.wpb = chunks % 4 == 0 ? 4 : (chunks % 2 == 0 ? 2 : 1);
  return t;
}

// Blocks of wpb waves per rho strip; when wpb does not divide the chunks,
// the surplus waves of the last block find no chunk and return at once (the
// stage kernel has no block-level barrier).
int stage_groups(int M, int wpb, int rows) { return (theta_chunks(M, rows) + wpb - 1) / wpb; }

hipError_t launch_stage(const Grid &g, const DevState &d, int integrator, int stage, bool has_src, int k_rel,
                        Tiling t, hipStream_t s) {
  if (t.wpb < 1 || t.wpb > 4 || !theta_rows_ok(g.M, t.rows)) return hipErrorInvalidValue;
  const int ngroups = stage_groups(g.M, t.wpb, t.rows);
  const int nb = (g.N / kWave) * ngroups;
  const dim3 grid(nb), block(kWave, t.wpb);
  switch (t.rows) {
    case 2: return launch_stage_rows<2>(g, d, integrator, stage, has_src, k_rel, ngroups, grid, block, t.wpe, s);
    case 4: return launch_stage_rows<4>(g, d, integrator, stage, has_src, k_rel, ngroups, grid, block, t.wpe, s);
    case 8: return launch_stage_rows<8>(g, d, integrator, stage, has_src, k_rel, ngroups, grid, block, t.wpe, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_batch(const BatchTile *tiles, int ntiles, int integrator, int stage, int k_rel, Tiling t,
                              hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (t.wpb < 1 || t.wpb > 4) return hipErrorInvalidValue;
  const dim3 block(kWave, t.wpb);
  switch (t.rows) {
    case 2: return launch_stage_batch_rows<2>(tiles, ntiles, integrator, stage, k_rel, block, t.wpe, s);
    case 4: return launch_stage_batch_rows<4>(tiles, ntiles, integrator, stage, k_rel, block, t.wpe, s);
    case 8: return launch_stage_batch_rows<8>(tiles, ntiles, integrator, stage, k_rel, block, t.wpe, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_finalize_batch(const SolveDesc *descs, int nsolves, int n, hipStream_t s) {
  if (nsolves <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize_batch, dim3(nsolves), dim3(256), 0, s, descs, n);
  return hipGetLastError();
}

hipError_t launch_source_batch(const SolveDesc *descs, int nsolves, int max_cells, int k_rel0, int n,
                               hipStream_t s) {
  if (nsolves <= 0 || max_cells <= 0 || n <= 0) return hipSuccess;
  if (n > kRing) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_source_batch, dim3((max_cells + 255) / 256, nsolves, n), dim3(256), 0, s, descs, k_rel0);
  return hipGetLastError();
}

hipError_t launch_finalize(const Grid &g, const DevState &d, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, g, d.ext, d.ctl, n);
  return hipGetLastError();
}

// Blocks of 256 threads, about four pairs per thread, at most 1024 blocks.
static unsigned sentinel_blocks(size_t n2) {
  const size_t nb = (n2 + 1023) / 1024;
  return (unsigned)(nb < 1 ? 1 : (nb > 1024 ? 1024 : nb));
}

hipError_t launch_sentinel(const Grid &g, const DevState &d, hipStream_t s) {
  const size_t n2 = 2 * (size_t)g.N * g.M;
  hipLaunchKernelGGL(k_sentinel, dim3(sentinel_blocks(n2)), dim3(256), 0, s,
                     reinterpret_cast<const double2 *>(d.U), n2, d.ctl);
  return hipGetLastError();
}

hipError_t launch_sentinel_batch(const SolveDesc *descs, int nsolves, hipStream_t s) {
  if (nsolves <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sentinel_batch, dim3(64, nsolves), dim3(256), 0, s, descs);
  return hipGetLastError();
}

hipError_t launch_source(const Grid &g, const DevState &d, int k_rel0, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (n > kRing) return hipErrorInvalidValue;
  const int cells = g.ptsx * g.ptsy;
  hipLaunchKernelGGL(k_source, dim3((cells + 255) / 256, n), dim3(256), 0, s, g, d, k_rel0);
  return hipGetLastError();
}

hipError_t launch_source_at(const Grid &g, const DevState &d, int kstep, hipStream_t s) {
  const int cells = g.ptsx * g.ptsy;
  hipLaunchKernelGGL(k_source_at, dim3((cells + 255) / 256), dim3(256), 0, s, g, d, kstep);
  return hipGetLastError();
}

hipError_t launch_source_coeffs(const teuk_t4::StepIn *in, SrcDev *sched, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_source_coeffs, dim3((n + 63) / 64), dim3(64), 0, s, in, sched, n);
  return hipGetLastError();
}

}  // namespace teuk
