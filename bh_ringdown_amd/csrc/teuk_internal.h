// Internal declarations shared by the libteuk host and device sources.
// Not part of the public C ABI (include/teuk.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#define TEUK_HD __host__ __device__ inline
#include "teuk_point.h"
#include "teuk_t4.h"

namespace teuk {

// parm.h:12 -- the 15-digit pi the reference uses for dtheta, cot and the
// spherical harmonics.
constexpr double kPi15 = 3.14159265358979;
// kern.cu:46
constexpr double kThird = 0.33333333333333333;
// kern.cu:196 -- first-derivative weight literal
constexpr double kD1 = 0.016666666666667;

// Stage-kernel tile: one wavefront owns 64 consecutive rho columns and a
// chunk of consecutive theta rows (theta_chunks); up to 4 wavefronts per
// block stack along theta.
constexpr int kWave = 64;
constexpr int kHalo = 4;            // KO reach in rho (kern.cu:247-256)
// Source patches are prepared for up to kRing steps per launch and kept in a
// ring of that many slots (step k uses slot k % kRing).
constexpr int kRing = 32;
constexpr int kLdsCols = kWave + 2 * kHalo;
// Whole-field sentinel: once per chunk of steps every field value is checked;
// a non-finite value, or a magnitude above this bound, marks the solve failed.
constexpr double kGrowthLimit = 1e60;

// Per-column constants of the coefficient fields (teuk_point.h).
using teuk_pt::kK1;
using teuk_pt::kK2;
using teuk_pt::kKA;
using teuk_pt::kKR;
using teuk_pt::kKD;
using teuk_pt::kKBt;
using teuk_pt::kKBr;
using teuk_pt::kKC0;
using teuk_pt::kKI0;
using teuk_pt::kKI1;
using teuk_pt::kKBi;
using teuk_pt::kKCi;
using teuk_pt::kNK;
// Per-row values: sin^2, cos, Y, tan(Pi/2 - theta); stored [4][M].
constexpr int kNR = 4;

// Scalar run description passed by value to every kernel.
struct Grid {
  int N, M;
  int even_m;          // ((int)m) % 2 == 0  (kern.cu:330)
  int ptsx, ptsy;
  int je;              // extraction column
  int Fr;
  int nrec_cap;        // output records the device buffers hold
  double xmin, xmax, dx, dtheta, dt;
  double mass, a, m;
  double start;
  int stag;            // 1: staggered second-order scheme (TD_GPU_code/Old);
                       // its source schedule holds two entries per step
};

// Device control block, one per solver.
struct Ctl {
  int k;               // completed steps
  int nrec;            // output records written
  int status;          // 0 ok; bit 0: non-finite Psi at the extraction column,
                       // bit 1: the whole-field sentinel fired (see kGrowthLimit)
  int pad;
  double timer;        // solver time, advanced by repeated addition (body.cu:214)
};

// Extraction state and output records (device pointers).
struct Ext {
  double *integ;       // [4][M] int Psi dt (re, im), int int Psi dt (re, im)
  double *snap;        // [nrec_cap][6][M] Psi, int, int int at output steps
  double *rec;         // [nrec_cap][16] t, dE/dt, h[14]
  int *rec_k;          // [nrec_cap][2] k, write flag
  const double *extw;  // [2M+1] dvol, rho2^4, rr^6 (body.cu:181-183)
  const double *projw; // [7][5*(M/4)][3] Boole/swsh factors per (l, term)
};

// Per-step particle scalars on the device (mirrors teuk_source_step).
struct SrcStep {
  double timer, start;
  double rp, tp, phip, E, lz, Q;
  double drdt, d2rdt2, dthdt, d2thdt2, dphidt, d2phidt2;
};

// One step of the source schedule on the device: canonical T_4 coefficients
// (filled by k_source_coeffs), particle position, radial Gaussian width,
// turn-on factor and the patch origin (kern.cu:1191-1206, 2673).
struct SrcDev {
  double c[12];
  double rp, tp, wr, turnon;
  int ip0, it0;
};

// Host-side preparation of one step (teuk_host.cpp).
struct SrcPrep {
  teuk_t4::StepIn in;
  SrcDev dev;
};

// Device pointers of one solver.
struct DevState {
  double *U, *C, *D, *A;   // field states (4 planes each); A only for RK4
  double *ktab;            // [kNK][N] column constants
  double *rtab;            // [kNR][M] row values
  double *dbdx;            // [M][N] rho derivative of bb (kern.cu:192-196)
  double *scratch;         // [M][N] set-up scratch (bb plane)
  double *theta, *r_c;     // [M], [N]
  double *phase;           // [2][N] cos, sin of the source phase (kern.cu:2671-2672)
  // Staggered scheme only: the half-point set (rho_a - dx/2, r_h).
  double *ktab_h;          // [kNK][N] column constants at the half points
  double *dbdx_h;          // [M][N] rho derivative of bb at the half points
  double *r_h;             // [N]
  double *phase_h;         // [2][N]
  double *patch;           // [kRing][2][ptsy][ptsx] source patches
  int *patch_org;          // [kRing][2] {row0, col0}
  SrcDev *sched;           // [nsched] source schedule, or null
  int nsched;
  Ext ext;
  Ctl *ctl;
};

// One member of a batch as the batched kernels see it: the run description,
// the solver's device pointers, its scheduled step count (steps at or past
// nt are skipped) and its tile count for the batch's block shape.
struct SolveDesc {
  Grid g;
  DevState d;
  int has_src;
  int ngroups;         // theta block-rows per rho strip
  int nt;
  int tiles;           // (N / 64) * ngroups
};

// One block's work in a batched stage launch: the member's description
// inline (one table read per block, not a tile entry and then a descriptor)
// and the tile within the member; tile < 0 pads the table.
struct BatchTile {
  SolveDesc sd;
  int tile;
  int pad[3];
};

// Theta chunks of one rho strip for wavefronts of R rows (R = 2, 4 or 8):
// the first chunk computes rows 3 .. 2+E and fills rows 0-2 from rows 3 and
// 4, the last computes rows M-3-E .. M-4 and fills M-3 .. M-1 from M-5 and
// M-4, and the rows between go R at a time.  E = edge_rows(R) >= 2 makes R
// divide what is left for every M that is a multiple of 8 and at least 16
// (the grids teuk_create accepts).
TEUK_HD constexpr int edge_rows(int R) { return R == 8 ? 5 : (R == 4 ? 3 : 2); }
TEUK_HD constexpr int theta_chunks(int M, int R) { return 2 + (M - 2 * edge_rows(R) - 6) / R; }
TEUK_HD constexpr bool theta_rows_ok(int M, int R) {
  return (R == 2 || R == 4 || R == 8) && M >= 16 && M % 8 == 0;
}
static_assert(theta_chunks(16, 8) == 2 && theta_chunks(64, 8) == 8 && theta_chunks(64, 4) == 15 &&
                  theta_chunks(64, 2) == 29 && theta_chunks(16, 2) == 5,
              "theta chunk layout");
// First row of theta chunk c, 0 <= c < theta_chunks(M, R): the first chunk
// starts at row 0, the last at theta_last_chunk_b0, chunk c in between at
// theta_inner_chunk_b0.  The edge chunks span edge_rows(R) + 3 rows (their
// computed rows and the three boundary rows), the others R.  (The stage
// kernel calls the per-kind forms inside its per-kind branches: a first row
// the compiler cannot see as 0 in the first chunk cost 15 % of the stage.)
TEUK_HD constexpr int theta_last_chunk_b0(int M, int R) { return M - edge_rows(R) - 3; }
TEUK_HD constexpr int theta_inner_chunk_b0(int R, int c) { return edge_rows(R) + 3 + (c - 1) * R; }
TEUK_HD constexpr int theta_chunk_b0(int M, int R, int c) {
  return c == 0 ? 0 : (c == theta_chunks(M, R) - 1 ? theta_last_chunk_b0(M, R) : theta_inner_chunk_b0(R, c));
}

// Block b of a single-solve stage launch of nblocks = S * ngroups blocks (S
// rho strips, ngroups block-rows per strip): its strip and block-row.  The
// hardware deals blocks round-robin over the 8 XCDs in dispatch order; each
// XCD gets a contiguous run of tiles, so the halos its waves re-read stay in
// its own L2.  edge_first: XCD x owns strips [x S/8, (x+1) S/8) when the
// strips split evenly, and the block-rows holding a strip's first and last
// theta chunk are dispatched before the others.
TEUK_HD void stage_block_tile(int b, int nblocks, int ngroups, bool edge_first, int &strip, int &grp) {
  if (!edge_first) {
    int tile = b;
    if ((nblocks & 7) == 0) tile = (b & 7) * (nblocks >> 3) + (b >> 3);
    strip = tile / ngroups;
    grp = tile - strip * ngroups;
    return;
  }
  const int S = nblocks / ngroups;
  int j = b, s0 = 0, ns = S;
  if ((S & 7) == 0) {
    ns = S >> 3;
    s0 = (b & 7) * ns;
    j = b >> 3;
  }
  const int ne = ngroups >= 2 ? 2 : 1;   // edge block-rows per strip
  if (j < ne * ns) {
    strip = s0 + j / ne;
    grp = (j % ne) ? ngroups - 1 : 0;
  } else {
    const int jj = j - ne * ns;
    strip = s0 + jj / (ngroups - 2);
    grp = 1 + jj % (ngroups - 2);
  }
}
constexpr int kDefaultThetaRows = 2;
constexpr int kDefaultWavesPerEU = 2;   // a batch's until set (teuk_batch_set_tiling)

// Stage-kernel launch shape: wavefronts per block (1, 2 or 4), the
// occupancy hint the kernel was compiled for (waves per SIMD: 2 or 4) and
// the theta rows per wavefront (2, 4 or 8).
struct Tiling {
  int wpb;
  int wpe;
  int rows;
};

// Launchers (teuk_kernels.hip).
hipError_t launch_init_coeffs(const Grid &g, const DevState &d, hipStream_t s);
hipError_t launch_source_coeffs(const teuk_t4::StepIn *in, SrcDev *sched, int n, hipStream_t s);
// Patches of n <= kRing consecutive steps from (step counter + k_rel0).
hipError_t launch_source(const Grid &g, const DevState &d, int k_rel0, int n, hipStream_t s);
// The patch of absolute step kstep (into its ring slot).
hipError_t launch_source_at(const Grid &g, const DevState &d, int kstep, hipStream_t s);
hipError_t launch_stage(const Grid &g, const DevState &d, int integrator, int stage,
                        bool has_src, int k_rel, Tiling t, hipStream_t s);
hipError_t launch_finalize(const Grid &g, const DevState &d, int n, hipStream_t s);
// The whole-field sentinel over the current fields (state U); enqueued once
// per chunk of steps, before the chunk's launch_finalize.
hipError_t launch_sentinel(const Grid &g, const DevState &d, hipStream_t s);
// Staggered second-order scheme (teuk_so2.hip): bb and dbdx of both point
// sets, and one step (half step, its theta edges, full step, edges and
// extraction).
hipError_t launch_init_so2(const Grid &g, const DevState &d, hipStream_t s);
hipError_t launch_so2_step(const Grid &g, const DevState &d, bool has_src, int k_rel, hipStream_t s);
// Batched forms: one launch covers every member of the batch.
hipError_t launch_source_batch(const SolveDesc *descs, int nsolves, int max_cells, int k_rel0, int n,
                               hipStream_t s);
hipError_t launch_stage_batch(const BatchTile *tiles, int ntiles, int integrator, int stage, int k_rel, Tiling t,
                              hipStream_t s);
// The stage builds for R theta rows per wavefront (teuk_stage_r<R>.hip),
// behind launch_stage / launch_stage_batch.
template <int R>
hipError_t launch_stage_rows(const Grid &g, const DevState &d, int integrator, int stage, bool has_src, int k_rel,
                             int ngroups, dim3 grid, dim3 block, int wpe, hipStream_t s);
template <int R>
hipError_t launch_stage_batch_rows(const BatchTile *tiles, int ntiles, int integrator, int stage, int k_rel,
                                   dim3 block, int wpe, hipStream_t s);
hipError_t launch_finalize_batch(const SolveDesc *descs, int nsolves, int n, hipStream_t s);
// The sentinel of every member still running before the chunk's finalize.
hipError_t launch_sentinel_batch(const SolveDesc *descs, int nsolves, hipStream_t s);
// Default launch shape for an N x M grid.
Tiling default_tiling(int N, int M);
// Blocks per rho strip for wpb waves per block of `rows` theta rows each
// (the last may hold surplus waves, which return at once).
int stage_groups(int M, int wpb, int rows);
bool source_available();

// Host helpers (teuk_host.cpp).
void host_grid(int N, int M, double xmin, double xmax, double mass, double a,
               double *x_c, double *r_c, double *theta);
// Half points of the staggered scheme: rho_h = xmin + (l - 1/2) dx, r_h(rho_h).
void host_grid_half(int N, double xmin, double xmax, double mass, double a, double *x_h, double *r_h);
double host_swsh(double ell, double ml, double the);
void host_projection_weights(int M, double m, double *w /*[7][5*(M/4)][3]*/);
// Column constants [kNK][N] at rho = xmin + (j + shift) dx with radius r[j]
// and row values [kNR][M] of the coefficient fields; `staggered` selects the
// staggered scheme's cot (cos/sin, Old/kern.cu:241-243) for the last row value.
void host_coeff_tables(int N, int M, double xmin, double xmax, double dx, double mass, double a,
                       double m, const double *r, const double *theta, double *ktab,
                       double *rtab, double shift = 0.0, bool staggered = false);
// Source phase e^{-i m a ln((r-r+)/(r-r-))/(r+-r-)} per column, [2][N].
void host_source_phase(int N, double mass, double a, double m, const double *r_c, double *phase);
// Per-step transcendentals and patch placement of the source (kern.cu:1186-1206, 2673).
void host_source_prepare(const Grid &g, const SrcStep &st, SrcPrep &out);

}  // namespace teuk
