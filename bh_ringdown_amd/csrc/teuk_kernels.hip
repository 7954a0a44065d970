// HIP kernels for the FP64 2+1D Teukolsky evolution on gfx950 (MI355X).
//
// Hot kernel: k_stage<INTEG, STAGE> -- one Runge-Kutta stage fused end to end:
// Kerr coefficients evaluated in registers, 6th-order rho/theta stencils,
// Kreiss-Oliger dissipation, the RHS assembly, the stage combination, both
// boundary fills and (last stage) the extraction-column bookkeeping, in one
// pass over the grid.  It replaces the reference's kernel_rhs1/2 +
// kernel_update1/2/3 + kernel_boundary1a/1b/2a/2b sequence and the host loop
// of body.cu:168-187 (TD_GPU_code/HighOrder/kern.cu:171-753, launched 12
// times per step at body.cu:131-165).
//
// Data per point and stage: the stage input (4 doubles, read once), the step
// base for stages 2-3 (4 doubles), the output (4 doubles) and the
// time-independent dbdx plane (1 double).  The nine coefficient fields of
// kernel_init are not stored: each is rebuilt per point from 13 per-column and
// 3 per-row values with one division and one square root.
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
//
// All arithmetic is built without FMA contraction and with the host's libm
// values fed in through tables, so results equal the host evaluation of the
// same expressions bit for bit.
#include "teuk_internal.h"

#include <climits>
#include <type_traits>

namespace teuk {
namespace {

// Every buffer a kernel touches is device global memory.  Pointers that
// arrive in kernel arguments are known to be global; pointers read from a
// table in memory (the batched kernels' descriptors) are not, and accesses
// through them would be flat instructions, which count against the scalar
// memory counter too, so that waiting for any scalar load also waits for
// every load in flight.  global_ptr states the address space.
#define TEUK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ TEUK_GLOBAL T *global_ptr(T *p) {
  return (TEUK_GLOBAL T *)p;
}

// Global accesses by a 32-bit byte offset from a uniform base: the address is
// an SGPR pair plus one VGPR (no 64-bit address arithmetic per access).  A
// solve's planes stay below 4 GiB (teuk_create checks N*M).
__device__ __forceinline__ double ld(const double *base, unsigned off) {
  return *global_ptr(reinterpret_cast<const double *>(reinterpret_cast<const char *>(base) + off));
}
__device__ __forceinline__ void st(double *base, unsigned off, double v) {
  *global_ptr(reinterpret_cast<double *>(reinterpret_cast<char *>(base) + off)) = v;
}
// A complex value (re, im) of a state: 16 bytes, one load or store (as a
// native two-double vector: HIP's double2 class has no address-space
// qualified copies).
typedef double dvec2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ld2(const double *base, unsigned off) {
  const dvec2 v = *global_ptr(reinterpret_cast<const dvec2 *>(reinterpret_cast<const char *>(base) + off));
  return make_double2(v.x, v.y);
}
__device__ __forceinline__ void st2(double *base, unsigned off, double re, double im) {
  const dvec2 v = {re, im};
  *global_ptr(reinterpret_cast<dvec2 *>(reinterpret_cast<char *>(base) + off)) = v;
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

using teuk_pt::Coef;

// The nine coefficient fields at one point (kern.cu:132-158) from the
// column constants and the row values sin^2, cos and Y (teuk_point.h).
__device__ __forceinline__ Coef coef_point(const double (&K)[kNK], double sn2, double cs, double Y) {
  return teuk_pt::coef_point(K, sn2, cs, Y, [](double x) { return __dsqrt_rn(x); });
}

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
// Fused Runge-Kutta stage.
//   INTEG 0 (SSP-RK3, kern.cu:301-318, 582-620):
//     STAGE 1: out = in + dt L(in)
//     STAGE 2: out = 1/4 (3 base + in + dt L(in))
//     STAGE 3: out = 1/3 (base + 2 in + 2 dt L(in))        (out may == base)
//   INTEG 1 (classic RK4 from its stage values u1, u2, u3):
//     STAGE 1: u1 = in + dt/2 L(in)                                (U -> C)
//     STAGE 2: u2 = base + dt/2 L(in);  w = (in + 2 u2 - base)/3   (C -> D, A)
//     STAGE 3: u3 = base + dt L(in)                                (D -> C)
//     STAGE 4: out = w + in/3 + dt/6 L(in)                         (C, A -> U)
//   State traffic 64 + 128 + 96 + 96 = 384 B per point, against
//   96 + 160 + 160 + 96 B with an accumulator array.
// The last stage also integrates Psi at the extraction column twice in time
// (body.cu:177-187) and snapshots it every Fr steps.
#ifndef TEUK_STAGE_WAVES_PER_EU
#define TEUK_STAGE_WAVES_PER_EU 2
#endif

// Buffers of one stage (the ping-pong between the state planes); w is RK4's
// partial combination, written by its second stage and read by its last.
struct StageIO {
  const double *in;
  const double *base;
  double *out;
  double *w;
};

template <int INTEG, int STAGE>
__device__ __forceinline__ StageIO stage_io(const DevState &d) {
  if (INTEG == 0) {
    if (STAGE == 1) return {d.U, d.U, d.C, nullptr};
    if (STAGE == 2) return {d.C, d.U, d.D, nullptr};
    return {d.D, d.U, d.U, nullptr};
  }
  if (STAGE == 1) return {d.U, d.U, d.C, nullptr};
  if (STAGE == 2) return {d.C, d.U, d.D, d.A};
  if (STAGE == 3) return {d.D, d.U, d.C, nullptr};
  return {d.C, d.U, d.U, d.A};
}

// One tile of one stage: the wavefront threadIdx.y of block-row `grp` of
// rho strip `strip` (64 columns), wpb wavefronts per block.  kGuard: the
// tile belongs to a batch member that may have reached its last scheduled
// step nt before this step; checked once the step counter is in (after the
// first row's loads are issued), before anything is stored.
template <int INTEG, int STAGE, bool kGuard>
__device__ __forceinline__ void stage_tile(const Grid &g, const DevState &d, int has_src, int k_rel, int strip,
                                           int grp, int wpb, int nt) {
  constexpr bool kLast = (INTEG == 0 && STAGE == 3) || (INTEG == 1 && STAGE == 4);
  __shared__ double lds[4][4][kLdsCols];
  const int N = g.N, M = g.M;
  const size_t NM = (size_t)N * M;
  // A state is two complex planes, Psi then Pi, each [M][N] of (re, im):
  // field f of point c sits at byte (f / 2) cb + 16 c + 8 (f % 2).
  const unsigned cb = 16u * (unsigned)NM;
  // The wavefront index is uniform: make that visible to the compiler so the
  // row index b is a scalar and the per-row tables are scalar loads.
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const StageIO io = stage_io<INTEG, STAGE>(d);
  const double *__restrict__ in = io.in;
  const double *base = io.base;
  double *out = io.out;
  double *wp = io.w;
  const TEUK_GLOBAL double *__restrict__ ktab = global_ptr(d.ktab);
  const TEUK_GLOBAL double *__restrict__ rtab = global_ptr(d.rtab);
  const double *__restrict__ dbdx = d.dbdx;
  const TEUK_GLOBAL double *__restrict__ patch = global_ptr(d.patch);   // advanced to this step's ring slot below
  const TEUK_GLOBAL int *__restrict__ porg = global_ptr(d.patch_org);
  TEUK_GLOBAL double *integ = global_ptr(d.ext.integ);
  TEUK_GLOBAL double *snap_buf = global_ptr(d.ext.snap);
  TEUK_GLOBAL Ctl *ctl = global_ptr(d.ctl);

  const int a0 = strip * kWave;
  const int a = a0 + lane;
  const int b0 = (grp * wpb + w) * kRows;
  if (b0 >= M) return;  // uniform over the wavefront; no block-level barrier follows
  double *L = &lds[w][0][0];

  // The step counter and the source-patch origins of all kRing ring slots
  // (one int per lane) are requested together, first: this step's origin is
  // picked out of the lanes once the counter is in, so the patch loads wait
  // for one memory round trip instead of two, and the first row's other
  // loads are issued before anything waits for either.
  // Both are vector loads: a scalar load would join the kernel-argument
  // loads, which can only be waited for all together.  (lane >> 6 is 0 for
  // every lane of a 64-wide block row; as the compiler cannot see that, the
  // counter's address is per lane and the load a vector load.  Any index
  // 0..3 stays inside Ctl's four leading ints.)
  static_assert(2 * kRing == kWave, "one patch-origin entry per lane");
  const int org_lane = porg[lane];
  const int kctl_lane = (&ctl->k)[lane >> 6];

  const bool lo_edge = (a0 == 0), hi_edge = (a0 + kWave == N);
  const bool inner_col = (a >= 3 && a < N - 3);
  const bool ko_col = (a >= 4 && a < N - 4);

  const double dt = g.dt;
  const double over_dx = 1.0 / g.dx, over_dth = 1.0 / g.dtheta;
  const double kof = teuk_pt::ko_factor(dt);
  // RK4 weights, once per chunk (not per row inside the lane-divergent body)
  const double h2 = 0.5 * dt, h6 = dt / 6.0;
  const double wx = kD1 * over_dx;
  const double wy = 0.05 * kThird * over_dth;
  const double wyy = 0.05 * kThird * kThird * over_dth * over_dth;

  // Column constants of the coefficient fields, loaded once per chunk.
  double K[kNK];
#pragma unroll
  for (int i = 0; i < kNK; i++) K[i] = ktab[(size_t)i * N + a];

  // Psi (q) theta window: rows b0-3 .. b0+10 of this column.
  double wqr[kRows + 6], wqi[kRows + 6];
#pragma unroll
  for (int i = 0; i < kRows + 6; i++) {
    const int r = b0 - 3 + i;
    if (r >= 0 && r < M) {
      const double2 q = ld2(in, 16u * (unsigned)(r * N + a));
      wqr[i] = q.x;
      wqi[i] = q.y;
    } else {
      wqr[i] = 0.0;
      wqi[i] = 0.0;
    }
  }

  int kstep = 0;
  int it0 = INT_MIN / 2, ip0 = INT_MIN / 2;
  // The step, and this step's patch slot and origin: the first use of the
  // loads above, after the first row's loads are issued.
  auto find_patch = [&]() {
    kstep = __builtin_amdgcn_readfirstlane(kctl_lane) + k_rel;
    if (has_src) {
      const int slot = kstep % kRing;
      patch += (size_t)slot * 2 * g.ptsx * g.ptsy;
      it0 = __builtin_amdgcn_readlane(org_lane, 2 * slot);
      ip0 = __builtin_amdgcn_readlane(org_lane, 2 * slot + 1);
    }
  };

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
  // Store one finished row (no conditions: the compiler's memory wait counts
  // stay exact across rows).
  auto store_row = [&](int b, const double (&v)[4]) {
    const unsigned c = 16u * (unsigned)(b * N + a);
    st2(out, c, v[0], v[1]);
    st2(out, cb + c, v[2], v[3]);
  };

  // Global inputs of a row.  The stencil inputs (p and the halo columns) are
  // fetched one row ahead; the inputs used only after the LDS stencils
  // (stage base, RK4 partial combination, dbdx, source patch, row values) are issued at the top
  // of their own row, before the next row's prefetch, so that waiting for
  // them never waits for the prefetch (loads retire in issue order).
  struct RowIn {
    double pr, pi;
    double2 halo;   // this lane's entry of the row's 2 x 8 complex halo values
  };
  struct RowLate {
    double base[4], w[4], dbdx, sr, si, rt[kNR];  // rt: the row's sin^2, cos, Y, cot
  };
  // The halo columns (4 left, 4 right of the strip) of a row, both complex
  // fields, come from one load: lane l serves plane (l / 8) % 2 (Psi, Pi),
  // halo slot l % 8 (lanes 16..63 repeat lanes 0..15).  A halo column
  // outside the grid (first and last strip) reads the strip's own edge column
  // instead; no update reads those slots.  Every load below is issued by all
  // lanes from an in-bounds address (lanes outside the patch re-read their
  // own column): no branches around loads, so the compiler's wait counts
  // stay exact and never cover the prefetch.
  const int hplane = (lane >> 3) & 1, hs = lane & 7;
  const int hslot = hs < kHalo ? hs : kWave + hs;          // LDS slot of the halo value
  const int hcol = a0 - kHalo + hslot;
  const int hsafe = (hcol >= 0 && hcol < N) ? hcol : (hs < kHalo ? a0 : a0 + kWave - 1);
  const unsigned hoff = (unsigned)hplane * cb + 16u * (unsigned)hsafe;
  constexpr bool kBase = INTEG == 0 ? STAGE >= 2 : (STAGE == 2 || STAGE == 3);
  constexpr bool kWIn = INTEG == 1 && STAGE == 4;
  auto load_row = [&](int b) {
    RowIn x;
    const unsigned c0 = 16u * (unsigned)(b * N);
    const double2 p = ld2(in, cb + c0 + 16u * a);
    x.pr = p.x;
    x.pi = p.y;
    x.halo = ld2(in, hoff + c0);
    return x;
  };
  // The row inputs other than the source patch ...
  auto load_late_fixed = [&](int b) {
    RowLate x;
    const unsigned c = 8u * (unsigned)(b * N + a);
    const unsigned c2 = 16u * (unsigned)(b * N + a);
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const double2 bv = kBase ? ld2(base, h * cb + c2) : make_double2(0.0, 0.0);
      x.base[2 * h] = bv.x;
      x.base[2 * h + 1] = bv.y;
      const double2 wv = kWIn ? ld2(wp, h * cb + c2) : make_double2(0.0, 0.0);
      x.w[2 * h] = wv.x;
      x.w[2 * h + 1] = wv.y;
    }
    x.dbdx = ld(dbdx, c);
#pragma unroll
    for (int j = 0; j < kNR; j++) x.rt[j] = rtab[j * M + b];
    return x;
  };
  // ... and the patch's source values at the row (0 outside the patch).
  auto load_patch = [&](RowLate &x, int b) {
    const int sb = b - it0, sa = a - ip0;
    const bool in_patch = (unsigned)sb < (unsigned)g.ptsy && (unsigned)sa < (unsigned)g.ptsx;
    const int pc = in_patch ? sb * g.ptsx + sa : 0;
    const double pr_ = patch[pc], pi_ = patch[g.ptsx * g.ptsy + pc];
    x.sr = in_patch ? pr_ : 0.0;
    x.si = in_patch ? pi_ : 0.0;
  };
  // Rows of the chunk, i0 <= i < i1.  The first chunk (kPart 1) skips rows
  // 0-2 and the last (kPart 2) rows M-3 .. M-1, which the theta boundary
  // fill produces; the row range is known at compile time, so the body has
  // no row conditions at all (exact wait counts, prefetch always in flight).
  auto run_rows = [&](auto part_tag) {
    constexpr int kPart = decltype(part_tag)::value;
    constexpr int i0 = kPart == 1 ? 3 : 0;
    constexpr int i1 = kPart == 2 ? kRows - 3 : kRows;
    RowIn cur = load_row(b0 + i0);
    RowLate late0 = load_late_fixed(b0 + i0);
    find_patch();
    if (kGuard && kstep >= nt) return;  // uniform: the member has finished
    load_patch(late0, b0 + i0);
#pragma unroll
    for (int i = i0; i < i1; i++) {
    const int b = b0 + i;
    RowLate late = late0;
    if (i > i0) {
      late = load_late_fixed(b);
      load_patch(late, b);
    }
    RowIn nxt = cur;
    if (i + 1 < i1) nxt = load_row(b + 1);
    const double pr0 = cur.pr, pi0 = cur.pi;
    const double qr0 = wqr[i + 3], qi0 = wqi[i + 3];

    // Stage the row (4 fields, 64 columns + 4-column halos) in LDS.
    L[0 * kLdsCols + lane + kHalo] = qr0;
    L[1 * kLdsCols + lane + kHalo] = qi0;
    L[2 * kLdsCols + lane + kHalo] = pr0;
    L[3 * kLdsCols + lane + kHalo] = pi0;
    L[(2 * hplane) * kLdsCols + hslot] = cur.halo.x;
    L[(2 * hplane + 1) * kLdsCols + hslot] = cur.halo.y;
    wave_sync();

    double v[4] = {0, 0, 0, 0};
    if (inner_col) {
      const int c = lane + kHalo;
      auto dx1 = [&](int f) {
        const double *r = L + f * kLdsCols;
        return teuk_pt::d1(r[c + 1], r[c - 1], r[c + 2], r[c - 2], r[c + 3], r[c - 3], wx);
      };
      auto ko8 = [&](int f) {
        const double *r = L + f * kLdsCols;
        return teuk_pt::ko8(r[c], r[c + 1], r[c - 1], r[c + 2], r[c - 2], r[c + 3], r[c - 3], r[c + 4], r[c - 4]);
      };
      const double ct = late.rt[3];
      teuk_pt::Pt x;
      x.qr = qr0;
      x.qi = qi0;
      x.pr = pr0;
      x.pi = pi0;
      const double qry = teuk_pt::d1(wqr[i + 4], wqr[i + 2], wqr[i + 5], wqr[i + 1], wqr[i + 6], wqr[i], wy);
      const double qryy =
          teuk_pt::d2(wqr[i + 3], wqr[i + 4], wqr[i + 2], wqr[i + 5], wqr[i + 1], wqr[i + 6], wqr[i], wyy);
      const double qiy = teuk_pt::d1(wqi[i + 4], wqi[i + 2], wqi[i + 5], wqi[i + 1], wqi[i + 6], wqi[i], wy);
      const double qiyy =
          teuk_pt::d2(wqi[i + 3], wqi[i + 4], wqi[i + 2], wqi[i + 5], wqi[i + 1], wqi[i + 6], wqi[i], wyy);
      x.llr = teuk_pt::fma_(ct, qry, qryy);
      x.lli = teuk_pt::fma_(ct, qiy, qiyy);
      x.qrx = dx1(0);
      x.qix = dx1(1);
      x.prx = dx1(2);
      x.pix = dx1(3);
      x.kqr = x.kqi = x.kpr = x.kpi = 0.0;
      if (ko_col) {
        x.kqr = ko8(0) * kof;
        x.kqi = ko8(1) * kof;
        x.kpr = ko8(2) * kof;
        x.kpi = ko8(3) * kof;
      }
      x.sr = late.sr;
      x.si = late.si;

      const Coef cf = coef_point(K, late.rt[0], late.rt[1], late.rt[2]);
      const double dre = teuk_pt::fma_(-late.dbdx, cf.ee, cf.dre);
      double r[4];
      teuk_pt::rhs(cf, dre, x, r);

      const double u[4] = {qr0, qi0, pr0, pi0};
      if (INTEG == 0) {
        if (STAGE == 1) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::rk3_s1(u[f], dt, r[f]);
        } else if (STAGE == 2) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::rk3_s2(late.base[f], u[f], dt, r[f]);
        } else {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::rk3_s3(late.base[f], u[f], dt, r[f], kThird);
        }
      } else {
        if (STAGE == 1) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::axpy(h2, r[f], u[f]);
        } else if (STAGE == 2) {
          const unsigned c = 16u * (unsigned)(b * N + a);
          double wv[4];
#pragma unroll
          for (int f = 0; f < 4; f++) {
            v[f] = teuk_pt::axpy(h2, r[f], late.base[f]);
            wv[f] = teuk_pt::rk4_w(late.base[f], u[f], v[f]);
          }
          st2(wp, c, wv[0], wv[1]);
          st2(wp, cb + c, wv[2], wv[3]);
        } else if (STAGE == 3) {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::axpy(dt, r[f], late.base[f]);
        } else {
#pragma unroll
          for (int f = 0; f < 4; f++) v[f] = teuk_pt::rk4_last(late.w[f], u[f], h6, r[f]);
        }
      }
    }
    wave_sync();  // LDS row is free for the next iteration
    rho_fix(v);
    store_row(b, v);
    cur = nxt;
    if (kPart != 0) {
      if (b == 3 || b == M - 5) {
#pragma unroll
        for (int f = 0; f < 4; f++) lo[f] = v[f];
      }
      if (b == 4 || b == M - 4) {
#pragma unroll
        for (int f = 0; f < 4; f++) hi[f] = v[f];
      }
    }
    }
  };
  if (b0 == 0)
    run_rows(std::integral_constant<int, 1>{});
  else if (b0 + kRows == M)
    run_rows(std::integral_constant<int, 2>{});
  else
    run_rows(std::integral_constant<int, 0>{});
  if (kGuard && kstep >= nt) return;
  const bool snap = kLast && (kstep % g.Fr) == 0 && kstep / g.Fr < g.nrec_cap;

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

  // Last stage: the extraction-column lane advances the time integrals of
  // its chunk's rows (body.cu:177-187) from the values it just stored, and
  // snapshots them every Fr steps.
  if (kLast && g.je >= a0 && g.je < a0 + kWave && a == g.je) {
    for (int b = b0; b < b0 + kRows; b++) {
      const double2 q = ld2(out, 16u * (unsigned)(b * N + a));
      const double q0 = q.x, q1 = q.y;
      const double ir = integ[b] + q0 * dt;
      const double ii = integ[M + b] + q1 * dt;
      const double dr = integ[2 * M + b] + ir * dt;
      const double di = integ[3 * M + b] + ii * dt;
      integ[b] = ir;
      integ[M + b] = ii;
      integ[2 * M + b] = dr;
      integ[3 * M + b] = di;
      if (!isfinite(q0) || !isfinite(q1)) ctl->status = 1;
      if (snap) {
        TEUK_GLOBAL double *sn = snap_buf + (size_t)(kstep / g.Fr) * 6 * M;
        sn[b] = q0;
        sn[M + b] = q1;
        sn[2 * M + b] = ir;
        sn[3 * M + b] = ii;
        sn[4 * M + b] = dr;
        sn[5 * M + b] = di;
      }
    }
  }
}

// Single-solve stage launch.  XCD-aware tile order: the hardware deals
// consecutive blocks round-robin over the 8 XCDs; remap so each XCD walks a
// contiguous run of tiles and the theta/rho halos it re-reads stay in its
// own L2.
template <int INTEG, int STAGE, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
void k_stage(Grid g, DevState d, int has_src, int ngroups, int k_rel) {
  const int nblocks = gridDim.x;
  int tile = blockIdx.x;
  if ((nblocks & 7) == 0) tile = (blockIdx.x & 7) * (nblocks >> 3) + (blockIdx.x >> 3);
  const int strip = tile / ngroups, grp = tile - strip * ngroups;
  stage_tile<INTEG, STAGE, false>(g, d, has_src, k_rel, strip, grp, blockDim.y, 0);
}

// Batched stage launch: block i works on tiles[i], a member's description
// and one of its tiles (the host orders the table for the XCD dealing).  A
// member that has completed its scheduled steps stores nothing.
template <int INTEG, int STAGE, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W)))
void k_stage_batch(const BatchTile *__restrict__ tiles, int k_rel) {
  const BatchTile &bt = tiles[blockIdx.x];
  if (bt.tile < 0) return;  // padding of the tile table
  const SolveDesc &sd = bt.sd;
  const int strip = bt.tile / sd.ngroups, grp = bt.tile - strip * sd.ngroups;
  stage_tile<INTEG, STAGE, true>(sd.g, sd.d, sd.has_src, k_rel, strip, grp, blockDim.y, sd.nt);
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

// Waves per block: the largest of 4, 2, 1 that divides the theta chunks, so
// no block carries idle waves.
Tiling default_tiling(int M) {
  const int chunks = M / kRows;
  Tiling t;
  t.wpb = chunks % 4 == 0 ? 4 : (chunks % 2 == 0 ? 2 : 1);
  t.wpe = TEUK_STAGE_WAVES_PER_EU;
  return t;
}

// Blocks of wpb waves per rho strip; when wpb does not divide the chunks,
// the surplus waves of the last block find b0 >= M and return at once (the
// stage kernel has no block-level barrier).
int stage_groups(int M, int wpb) { return (M / kRows + wpb - 1) / wpb; }

// k_stage<I, S, W> for the runtime (stage, wpe) pair.
template <int I, int S>
static hipError_t stage_one(const Grid &g, const DevState &d, int hs, int ngroups, int k_rel, dim3 grid,
                            dim3 block, int wpe, hipStream_t s) {
  if (wpe == 4)
    hipLaunchKernelGGL((k_stage<I, S, 4>), grid, block, 0, s, g, d, hs, ngroups, k_rel);
  else if (wpe == 2)
    hipLaunchKernelGGL((k_stage<I, S, 2>), grid, block, 0, s, g, d, hs, ngroups, k_rel);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <int I, int S>
static hipError_t stage_batch_one(const BatchTile *tiles, int ntiles, int k_rel, dim3 block, int wpe,
                                  hipStream_t s) {
  if (wpe == 4)
    hipLaunchKernelGGL((k_stage_batch<I, S, 4>), dim3(ntiles), block, 0, s, tiles, k_rel);
  else if (wpe == 2)
    hipLaunchKernelGGL((k_stage_batch<I, S, 2>), dim3(ntiles), block, 0, s, tiles, k_rel);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_stage(const Grid &g, const DevState &d, int integrator, int stage, bool has_src, int k_rel,
                        Tiling t, hipStream_t s) {
  if (t.wpb < 1 || t.wpb > 4 || g.M % kRows != 0) return hipErrorInvalidValue;
  const int ngroups = stage_groups(g.M, t.wpb);
  const int nb = (g.N / kWave) * ngroups;
  const dim3 grid(nb), block(kWave, t.wpb);
  const int hs = has_src ? 1 : 0;
  if (integrator == 0) {
    switch (stage) {
      case 1: return stage_one<0, 1>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
      case 2: return stage_one<0, 2>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
      case 3: return stage_one<0, 3>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (stage) {
    case 1: return stage_one<1, 1>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
    case 2: return stage_one<1, 2>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
    case 3: return stage_one<1, 3>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
    case 4: return stage_one<1, 4>(g, d, hs, ngroups, k_rel, grid, block, t.wpe, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_stage_batch(const BatchTile *tiles, int ntiles, int integrator, int stage, int k_rel, Tiling t,
                              hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  if (t.wpb < 1 || t.wpb > 4) return hipErrorInvalidValue;
  const dim3 block(kWave, t.wpb);
  if (integrator == 0) {
    switch (stage) {
      case 1: return stage_batch_one<0, 1>(tiles, ntiles, k_rel, block, t.wpe, s);
      case 2: return stage_batch_one<0, 2>(tiles, ntiles, k_rel, block, t.wpe, s);
      case 3: return stage_batch_one<0, 3>(tiles, ntiles, k_rel, block, t.wpe, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (stage) {
    case 1: return stage_batch_one<1, 1>(tiles, ntiles, k_rel, block, t.wpe, s);
    case 2: return stage_batch_one<1, 2>(tiles, ntiles, k_rel, block, t.wpe, s);
    case 3: return stage_batch_one<1, 3>(tiles, ntiles, k_rel, block, t.wpe, s);
    case 4: return stage_batch_one<1, 4>(tiles, ntiles, k_rel, block, t.wpe, s);
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
