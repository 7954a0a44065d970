// libteuk C ABI (include/teuk.h): solver lifecycle, stepping with
// hipGraph-captured steps, trajectory consumption and the body() drop-in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "teuk.h"
#include "teuk_internal.h"

using namespace teuk;

struct teuk_batch;

struct teuk_solver {
  teuk_params P;
  Grid g;
  DevState d;
  Tiling tiling;
  teuk_batch *batch = nullptr;   // set while a batch drives this solver
  hipStream_t stream = nullptr;
  int nt = 0;              // scheduled steps: (int)(T/dt), body.cu:59
  int k_host = 0;          // steps enqueued
  bool has_src = false;
  // Graph cache: one captured chunk of `graph_steps` steps.
  hipGraphExec_t gexec = nullptr;
  int graph_steps = 0;
  // Kernel timing.
  bool timing = false;
  std::vector<hipEvent_t> ev;   // pairs
  size_t ev_used = 0;
  long long timed_launches = 0;
  double timed_ms = 0.0;
  std::vector<void *> allocs;
  void *sched_blk = nullptr;   // the source schedule's allocation (replaced by each teuk_set_source)
};

// A set of solvers stepped together by the batched kernels (one launch per
// stage for every member), with graph-captured chunks of steps.
struct teuk_batch {
  int integrator = 0;
  Tiling tiling{4, kDefaultWavesPerEU, kDefaultThetaRows};
  int chunk = 32;          // steps per captured graph
  int order = 0;           // tile table: 0 = contiguous run per XCD, 1 = dealt in list order
  long long budget = 0;    // bytes of member footprint allowed (0 = no limit)
  long long used = 0;
  std::vector<teuk_solver *> members;
  std::vector<long long> member_bytes;
  hipStream_t stream = nullptr;
  // Cache groups: the members still stepping, in groups whose field state
  // (group_bytes, 0 = all in one group) stays in the last-level cache while
  // the group runs a chunk of steps; each group is a contiguous run of the
  // descriptor table, has its own tile table and its own captured graph.
  long long group_bytes = 0;
  struct Group {
    int first = 0, count = 0;      // descriptor range
    size_t tile_off = 0;           // into d_tiles
    int ntiles = 0;
    int max_cells = 0;             // largest source patch of the group (0: no sources)
    hipGraphExec_t gexec = nullptr;
  };
  std::vector<Group> groups;
  SolveDesc *d_desc = nullptr;
  int desc_cap = 0;
  BatchTile *d_tiles = nullptr;
  size_t tiles_cap = 0;
  bool dirty = true;
  int graph_steps = 0;
};

namespace {

// The stream that runs a solver's steps: its own, or its batch's.
hipStream_t run_stream(const teuk_solver *s) { return s->batch ? s->batch->stream : s->stream; }

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) return TEUK_EHIP;    \
  } while (0)

template <typename T>
int dalloc(teuk_solver *s, T **p, size_t count) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return TEUK_ENOMEM;
  s->allocs.push_back(q);
  *p = static_cast<T *>(q);
  return TEUK_OK;
}

int validate(const teuk_params *p) {
  if (!p) return TEUK_EINVAL;
  if (p->N < 128 || p->N % kWave != 0) return TEUK_EINVAL;
  if (p->M < 16 || p->M % 8 != 0) return TEUK_EINVAL;  // the stage kernels' theta chunks (theta_chunks)
  if (p->M > (p->integrator == 2 ? 256 : 1024)) return TEUK_EINVAL;
  if ((long long)p->N * p->M > (1LL << 26)) return TEUK_EINVAL;  // kernels address planes with 32-bit offsets
  if (p->s != -2.0) return TEUK_EINVAL;  // kern.cu:142-145 bake in s = -2
  if (!(p->dt > 0.0) || !(p->dx > 0.0) || !(p->dtheta > 0.0)) return TEUK_EINVAL;
  if (!(p->xmax > p->xmin)) return TEUK_EINVAL;
  if (p->Fr < 1 || p->ptsx < 1 || p->ptsy < 1) return TEUK_EINVAL;
  if (p->integrator < 0 || p->integrator > 2) return TEUK_EINVAL;
  if (p->j_extract >= p->N || p->j_extract < -1) return TEUK_EINVAL;
  if (!(std::fabs(p->a) < p->mass)) return TEUK_EINVAL;  // horizon formulas need |a| < M
  if (p->m != std::floor(p->m)) return TEUK_EINVAL;
  if (!(p->T >= 0.0)) return TEUK_EINVAL;
  return TEUK_OK;
}

int scheduled_steps(const teuk_params *p) { return (int)(p->T / p->dt); }

// Kernel launches per step that the timing events cover as "stages": the
// RK stages, or the staggered scheme's half and full step.
int stages_per_step(int integrator) { return integrator == 0 ? 3 : (integrator == 1 ? 4 : 2); }

// Steps the source schedule covers (two entries per step for the staggered
// scheme).
int sched_steps(const teuk_solver *s) { return s->g.stag ? s->d.nsched / 2 : s->d.nsched; }

void flush_timing(teuk_solver *s) {
  if (s->ev_used == 0) return;
  (void)hipStreamSynchronize(s->stream);
  for (size_t i = 0; i < s->ev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, s->ev[i], s->ev[i + 1]) == hipSuccess) {
      s->timed_ms += ms;
      s->timed_launches += stages_per_step(s->P.integrator);
    }
  }
  s->ev_used = 0;
}

// One time step whose index is (device step counter + k_rel): source patches
// (every kRing steps, for the next kRing steps), then the fused stages.
// Output records are completed by launch_finalize once per batch of steps.
int enqueue_step(teuk_solver *s, int k_rel) {
  const int nstage = stages_per_step(s->P.integrator);
  // the patches of this step and the next ones (up to the ring size) come
  // from one launch at the first step of each group of kRing steps (kRing
  // half steps for the staggered scheme)
  const int group = s->g.stag ? kRing / 2 : kRing;
  if (s->has_src && k_rel % group == 0) HIPCHK(launch_source(s->g, s->d, k_rel, kRing, s->stream));
  // Timing: one event pair around the step's back-to-back stage launches
  // (an event between two launches would add its own gap to both).
  if (s->timing) {
    if (s->ev_used + 2 > s->ev.size()) {
      if (s->ev.size() < 4096) {
        for (int i = 0; i < 64; i++) {
          hipEvent_t e;
          HIPCHK(hipEventCreate(&e));
          s->ev.push_back(e);
        }
      } else {
        flush_timing(s);
      }
    }
    HIPCHK(hipEventRecord(s->ev[s->ev_used], s->stream));
  }
  if (s->g.stag) {
    HIPCHK(launch_so2_step(s->g, s->d, s->has_src, k_rel, s->stream));
  } else {
    for (int st = 1; st <= nstage; st++)
      HIPCHK(launch_stage(s->g, s->d, s->P.integrator, st, s->has_src, k_rel, s->tiling, s->stream));
  }
  if (s->timing) {
    HIPCHK(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
    s->ev_used += 2;
  }
  return TEUK_OK;
}

int build_graph(teuk_solver *s, int steps) {
  if (s->gexec) {
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = TEUK_OK;
  for (int i = 0; i < steps && rc == TEUK_OK; i++) rc = enqueue_step(s, i);
  if (rc == TEUK_OK && launch_sentinel(s->g, s->d, s->stream) != hipSuccess) rc = TEUK_EHIP;
  if (rc == TEUK_OK && launch_finalize(s->g, s->d, steps, s->stream) != hipSuccess) rc = TEUK_EHIP;
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  if (rc != TEUK_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return TEUK_EHIP;
  e = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return TEUK_EHIP;
  s->graph_steps = steps;
  return TEUK_OK;
}

// body.cu:97-129 consumption of the .traj stream (10 columns per row):
// read one row, then rows until t >= start; per step read two rows and use
// the second; r += 1e-6; accelerations are backward differences of the
// velocity columns over dt with zero velocities before the first step.
int consume_trajectory(const double *rows, int nrows, int nsteps, double start, double dt,
                       std::vector<SrcStep> &out) {
  out.clear();
  out.reserve(nsteps);
  int cur = 0;
  auto read = [&](const double *&row) {
    if (cur >= nrows) return false;
    row = rows + (size_t)10 * cur++;
    return true;
  };
  const double *row = nullptr;
  if (!read(row)) return TEUK_ERANGE;
  do {
    if (!read(row)) return TEUK_ERANGE;
  } while (row[0] < start);
  double v_r = 0.0, v_th = 0.0, v_ph = 0.0;
  double timer = start;
  for (int k = 0; k < nsteps; k++) {
    if (!read(row) || !read(row)) return TEUK_ERANGE;
    SrcStep st;
    st.timer = timer;
    st.start = start;
    st.rp = row[1] + 1e-6;
    st.tp = row[2];
    st.drdt = row[3];
    st.dthdt = row[4];
    st.dphidt = row[5];
    st.phip = row[6];
    st.E = row[7];
    st.lz = row[8];
    st.Q = row[9];
    st.d2rdt2 = (row[3] - v_r) / dt;
    st.d2thdt2 = (row[4] - v_th) / dt;
    st.d2phidt2 = (row[5] - v_ph) / dt;
    v_r = row[3];
    v_th = row[4];
    v_ph = row[5];
    out.push_back(st);
    timer += dt;
  }
  return TEUK_OK;
}

// Old/body.cu:113-180, the staggered scheme's reading of the same stream:
// the same priming, then one row per half step, each used as read, with
// r += 1e-6 and accelerations (v - v_prev) / dt * 2 over the half step; both
// entries of a step carry the step's timer.
int consume_trajectory_half(const double *rows, int nrows, int nsteps, double start, double dt,
                            std::vector<SrcStep> &out) {
  out.clear();
  out.reserve(2 * (size_t)nsteps);
  int cur = 0;
  auto read = [&](const double *&row) {
    if (cur >= nrows) return false;
    row = rows + (size_t)10 * cur++;
    return true;
  };
  const double *row = nullptr;
  if (!read(row)) return TEUK_ERANGE;
  do {
    if (!read(row)) return TEUK_ERANGE;
  } while (row[0] < start);
  double v_r = 0.0, v_th = 0.0, v_ph = 0.0;
  double timer = start;
  for (int k = 0; k < nsteps; k++) {
    for (int h = 0; h < 2; h++) {
      if (!read(row)) return TEUK_ERANGE;
      SrcStep st;
      st.timer = timer;
      st.start = start;
      st.rp = row[1] + 1e-6;
      st.tp = row[2];
      st.drdt = row[3];
      st.dthdt = row[4];
      st.dphidt = row[5];
      st.phip = row[6];
      st.E = row[7];
      st.lz = row[8];
      st.Q = row[9];
      st.d2rdt2 = (row[3] - v_r) / dt * 2.0;
      st.d2thdt2 = (row[4] - v_th) / dt * 2.0;
      st.d2phidt2 = (row[5] - v_ph) / dt * 2.0;
      v_r = row[3];
      v_th = row[4];
      v_ph = row[5];
      out.push_back(st);
    }
    timer += dt;
  }
  return TEUK_OK;
}

int create_solver(const teuk_params *p, const double *r_h, const double *r_c, const double *theta,
                  teuk_solver **out);

int read_traj_file(const char *path, std::vector<double> &rows) {
  FILE *f = std::fopen(path, "r");
  if (!f) return TEUK_EIO;
  double v[10];
  while (std::fscanf(f, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                     &v[6], &v[7], &v[8], &v[9]) == 10)
    rows.insert(rows.end(), v, v + 10);
  std::fclose(f);
  return rows.empty() ? TEUK_EIO : TEUK_OK;
}

}  // namespace

extern "C" {

const char *teuk_strerror(int code) {
  switch (code) {
    case TEUK_OK: return "ok";
    case TEUK_EINVAL: return "invalid argument or unsupported size";
    case TEUK_EHIP: return "HIP runtime error";
    case TEUK_ENOMEM: return "out of memory";
    case TEUK_EDIVERGED: return "non-finite field (solve diverged)";
    case TEUK_EIO: return "trajectory or waveform file error";
    case TEUK_ERANGE: return "step beyond the scheduled run or trajectory";
    case TEUK_ETOOBIG: return "the solve needs more than teuk_solve_noalloc's fixed scratch memory";
    default: return "unknown error";
  }
}

int teuk_default_params(teuk_params *p, int N, int M, double xmin, double xmax, double a, double m,
                        double dt_over_dx, double T) {
  if (!p || N <= 0 || M <= 0) return TEUK_EINVAL;
  std::memset(p, 0, sizeof(*p));
  p->N = N;
  p->M = M;
  p->xmin = xmin;
  p->xmax = xmax;
  p->mass = 1.0;
  p->a = a;
  p->s = -2.0;
  p->m = m;
  p->dx = (xmax - xmin) / N;
  p->dtheta = kPi15 / M;
  p->dt = dt_over_dx * p->dx;
  p->T = T;
  p->Fr = 1;
  p->ptsx = 128;
  p->ptsy = 128;
  p->integrator = 0;
  p->j_extract = -1;
  return TEUK_OK;
}

int teuk_grid(const teuk_params *p, double *x_c, double *r_c, double *theta) {
  if (!p || !x_c || !r_c || !theta || p->N <= 0 || p->M <= 0) return TEUK_EINVAL;
  host_grid(p->N, p->M, p->xmin, p->xmax, p->mass, p->a, x_c, r_c, theta);
  return TEUK_OK;
}

int teuk_grid_half(const teuk_params *p, double *x_h, double *r_h) {
  if (!p || !x_h || !r_h || p->N <= 0) return TEUK_EINVAL;
  host_grid_half(p->N, p->xmin, p->xmax, p->mass, p->a, x_h, r_h);
  return TEUK_OK;
}

long long teuk_footprint(const teuk_params *p, int nsteps) {
  if (!p) return -1;
  const long long NM = (long long)p->N * p->M;
  const long long states = (p->integrator == 0 ? 3 : (p->integrator == 1 ? 4 : 2)) * 4 * NM;
  const long long entries = (long long)std::max(nsteps, 0) * (p->integrator == 2 ? 2 : 1);
  const long long nrec = (long long)(std::max(nsteps, 0) / std::max(p->Fr, 1)) + 1;
  long long bytes = 8 * (states + 2 * NM + (kNK + 1) * (long long)p->N + (kNR + 1) * (long long)p->M);
  bytes += 8 * ((long long)kRing * 2 * p->ptsx * p->ptsy) + 4LL * 2 * kRing;
  bytes += 8 * (4LL * p->M + 2 * p->M + 1 + 7LL * 5 * (p->M / 4) * 3);
  bytes += nrec * (16 * 8 + 2 * 4 + 6LL * p->M * 8);
  bytes += entries * (long long)(sizeof(SrcDev) + sizeof(teuk_t4::StepIn));
  if (p->integrator == 2) bytes += 8 * (NM + (kNK + 3) * (long long)p->N);  // dbdx_h, ktab_h, r_h, phase_h
  return bytes;
}

int teuk_create(const teuk_params *p, const double *r_c, const double *theta, teuk_solver **out) {
  if (!out || !r_c || !theta) return TEUK_EINVAL;
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  if (p->integrator == 2) {
    std::vector<double> x_h(p->N), r_h(p->N);
    host_grid_half(p->N, p->xmin, p->xmax, p->mass, p->a, x_h.data(), r_h.data());
    return teuk_create_staggered(p, r_h.data(), r_c, theta, out);
  }
  return create_solver(p, nullptr, r_c, theta, out);
}

int teuk_create_staggered(const teuk_params *p, const double *r_h, const double *r_c, const double *theta,
                          teuk_solver **out) {
  if (!out || !r_h || !r_c || !theta) return TEUK_EINVAL;
  *out = nullptr;
  int rc = validate(p);
  if (rc) return rc;
  if (p->integrator != 2) return TEUK_EINVAL;
  return create_solver(p, r_h, r_c, theta, out);
}

}  // extern "C"

namespace {

// Byte offsets of one solver's pieces in its device memory, each piece
// 256-byte aligned: the tables copied from the host first (one copy), then
// the memory that starts at zero (one fill), then the rest.
struct Layout {
  size_t phase = 0, theta = 0, r_c = 0, ktab = 0, rtab = 0, extw = 0, projw = 0, org = 0, ctl = 0;
  size_t phase_h = 0, r_h = 0, ktab_h = 0;   // staggered scheme only
  size_t init_end = 0;
  size_t states = 0, integ = 0, patch = 0;
  size_t zero_end = 0;
  size_t dbdx = 0, scratch = 0, snap = 0, rec = 0, rec_k = 0, dbdx_h = 0;
  size_t end = 0;
};

size_t piece(size_t &at, size_t bytes) {
  const size_t off = at;
  at += (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  return off;
}

int state_count(int integrator) { return integrator == 0 ? 3 : (integrator == 1 ? 4 : 2); }

Layout plan_layout(const teuk_params *p, int nrec_cap) {
  Layout L;
  const bool stag = p->integrator == 2;
  const size_t N = (size_t)p->N, M = (size_t)p->M, NM = N * M, D = sizeof(double);
  size_t at = 0;
  L.phase = piece(at, 2 * N * D);
  L.theta = piece(at, M * D);
  L.r_c = piece(at, N * D);
  L.ktab = piece(at, kNK * N * D);
  L.rtab = piece(at, kNR * M * D);
  L.extw = piece(at, (2 * M + 1) * D);
  L.projw = piece(at, 7 * (5 * (M / 4)) * 3 * D);
  L.org = piece(at, 2 * kRing * sizeof(int));
  L.ctl = piece(at, sizeof(Ctl));
  if (stag) {
    L.phase_h = piece(at, 2 * N * D);
    L.r_h = piece(at, N * D);
    L.ktab_h = piece(at, kNK * N * D);
  }
  L.init_end = at;
  // Zero fields and integrals (kern.cu:164-167, body.cu:86-91), empty patches.
  L.states = piece(at, (size_t)state_count(p->integrator) * 4 * NM * D);
  L.integ = piece(at, 4 * M * D);
  L.patch = piece(at, (size_t)kRing * 2 * p->ptsx * p->ptsy * D);
  L.zero_end = at;
  L.dbdx = piece(at, NM * D);
  L.scratch = piece(at, NM * D);
  L.snap = piece(at, (size_t)nrec_cap * 6 * M * D);
  L.rec = piece(at, (size_t)nrec_cap * 16 * D);
  L.rec_k = piece(at, (size_t)nrec_cap * 2 * sizeof(int));
  if (stag) L.dbdx_h = piece(at, NM * D);
  L.end = at;
  return L;
}

// Host contents of the copied pieces, written at their offsets into `stage`
// (L.init_end bytes): source phases, coefficient column and row tables,
// extraction weights (body.cu:175-183, rr the extraction radius), projection
// weights, patch origins, the control block; for the staggered scheme the
// same at the half points (rtab_h: kNR * M doubles of workspace).
void fill_init(const teuk_params *p, const double *r_c, const double *theta, const double *r_h,
               double *rtab_h, const Layout &L, char *stage) {
  auto at = [&](size_t off) { return reinterpret_cast<double *>(stage + off); };
  const int N = p->N, M = p->M;
  const bool stag = p->integrator == 2;
  host_source_phase(N, p->mass, p->a, p->m, r_c, at(L.phase));
  std::memcpy(at(L.theta), theta, (size_t)M * sizeof(double));
  std::memcpy(at(L.r_c), r_c, (size_t)N * sizeof(double));
  host_coeff_tables(N, M, p->xmin, p->xmax, p->dx, p->mass, p->a, p->m, r_c, theta, at(L.ktab), at(L.rtab), 0.0,
                    stag);
  double *w = at(L.extw);
  const double rr = r_c[p->j_extract < 0 ? N - 1 : p->j_extract];
  for (int q = 0; q < M; q++) {
    const double rho2 = rr * rr + std::pow(p->a * std::cos(theta[q]), 2);
    w[q] = 0.5 * p->dtheta * std::sin(theta[q]) * rr * rr;
    w[M + q] = std::pow(rho2, 4);
  }
  w[2 * M] = std::pow(rr, 6);
  host_projection_weights(M, p->m, at(L.projw));
  int *org = reinterpret_cast<int *>(stage + L.org);
  for (int i = 0; i < 2 * kRing; i++) org[i] = -(1 << 29);
  Ctl c0{};
  c0.k = 0;
  c0.nrec = 0;
  c0.status = 0;
  c0.timer = 0.0;
  std::memcpy(stage + L.ctl, &c0, sizeof(Ctl));
  if (stag) {  // Old/kern.cu:90-135: the same fields at rho - dx/2 with r_h
    host_source_phase(N, p->mass, p->a, p->m, r_h, at(L.phase_h));
    std::memcpy(at(L.r_h), r_h, (size_t)N * sizeof(double));
    host_coeff_tables(N, M, p->xmin, p->xmax, p->dx, p->mass, p->a, p->m, r_h, theta, at(L.ktab_h), rtab_h, -0.5,
                      true);
  }
}

// Point a solver's device pointers into memory laid out by L at `base`.
void bind_layout(const teuk_params *p, const Layout &L, char *base, DevState &d) {
  auto at = [&](size_t off) { return reinterpret_cast<double *>(base + off); };
  const size_t NM = (size_t)p->N * p->M;
  const int nstates = state_count(p->integrator);
  d.phase = at(L.phase);
  d.theta = at(L.theta);
  d.r_c = at(L.r_c);
  d.ktab = at(L.ktab);
  d.rtab = at(L.rtab);
  d.ext.extw = at(L.extw);
  d.ext.projw = at(L.projw);
  d.patch_org = reinterpret_cast<int *>(base + L.org);
  d.ctl = reinterpret_cast<Ctl *>(base + L.ctl);
  double *states = at(L.states);
  d.U = states;
  d.C = states + 4 * NM;   // staggered scheme: the half-point fields
  d.D = nstates >= 3 ? states + 8 * NM : nullptr;
  d.A = nstates == 4 ? states + 12 * NM : nullptr;
  d.ext.integ = at(L.integ);
  d.patch = at(L.patch);
  d.dbdx = at(L.dbdx);
  d.scratch = at(L.scratch);
  d.ext.snap = at(L.snap);
  d.ext.rec = at(L.rec);
  d.ext.rec_k = reinterpret_cast<int *>(base + L.rec_k);
  if (p->integrator == 2) {
    d.phase_h = at(L.phase_h);
    d.r_h = at(L.r_h);
    d.ktab_h = at(L.ktab_h);
    d.dbdx_h = at(L.dbdx_h);
  }
  d.sched = nullptr;
  d.nsched = 0;
}

// The run description of a solve.
Grid make_grid(const teuk_params *p, int nrec_cap) {
  Grid g{};
  g.N = p->N;
  g.M = p->M;
  g.even_m = (((int)p->m) % 2) == 0;
  g.ptsx = p->ptsx;
  g.ptsy = p->ptsy;
  g.je = p->j_extract < 0 ? p->N - 1 : p->j_extract;
  g.Fr = p->Fr;
  g.nrec_cap = nrec_cap;
  g.xmin = p->xmin;
  g.xmax = p->xmax;
  g.dx = p->dx;
  g.dtheta = p->dtheta;
  g.dt = p->dt;
  g.mass = p->mass;
  g.a = p->a;
  g.m = p->m;
  g.start = 0.0;
  g.stag = p->integrator == 2 ? 1 : 0;
  return g;
}

// Device state, tables and coefficient set-up of one solver; r_h only for
// the staggered scheme.  All of the solver's device memory is one
// allocation (plan_layout); its tables go up in one copy from one staging
// buffer and the memory that starts at zero is cleared by one fill.
int create_solver(const teuk_params *p, const double *r_h, const double *r_c, const double *theta,
                  teuk_solver **out) {
  int rc = TEUK_OK;
  teuk_solver *s = new teuk_solver();
  s->P = *p;
  s->nt = scheduled_steps(p);
  s->tiling = default_tiling(p->N, p->M);
  s->g = make_grid(p, s->nt / p->Fr + 1);
  Grid &g = s->g;
  auto fail = [&](int code) {
    teuk_destroy(s);
    return code;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) return fail(TEUK_EHIP);
  const Layout L = plan_layout(p, g.nrec_cap);
  char *base = nullptr;
  if ((rc = dalloc(s, &base, L.end))) return fail(rc);
  bind_layout(p, L, base, s->d);
  {
    std::vector<char> staging(L.init_end, 0);
    std::vector<double> rtab_h(g.stag ? kNR * (size_t)p->M : 0);
    fill_init(p, r_c, theta, r_h, rtab_h.data(), L, staging.data());
    if (hipMemcpyAsync(base, staging.data(), L.init_end, hipMemcpyHostToDevice, s->stream) != hipSuccess)
      return fail(TEUK_EHIP);
    if (hipMemsetAsync(base + L.init_end, 0, L.zero_end - L.init_end, s->stream) != hipSuccess)
      return fail(TEUK_EHIP);
    // the staging buffer is pageable: the copy has completed when the call returns
    if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(TEUK_EHIP);
  }
  if ((g.stag ? launch_init_so2(g, s->d, s->stream) : launch_init_coeffs(g, s->d, s->stream)) != hipSuccess)
    return fail(TEUK_EHIP);
  if (hipStreamSynchronize(s->stream) != hipSuccess) return fail(TEUK_EHIP);
  *out = s;
  return TEUK_OK;
}

}  // namespace

extern "C" {

void teuk_destroy(teuk_solver *s) {
  if (!s) return;
  if (s->batch) (void)teuk_batch_remove(s->batch, s);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  for (hipEvent_t e : s->ev) (void)hipEventDestroy(e);
  for (void *q : s->allocs) (void)hipFree(q);
  if (s->sched_blk) (void)hipFree(s->sched_blk);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// On the device a state is two complex planes, Psi and Pi, each [M][N] of
// (re, im) pairs (one 16-byte access per complex value); the ABI's layout is
// four real planes q_re, q_im, p_re, p_im (kern.cu:10-13).
int teuk_set_fields(teuk_solver *s, const double *u) {
  if (!s || !u) return TEUK_EINVAL;
  const size_t NM = (size_t)s->P.N * s->P.M;
  std::vector<double> dev(4 * NM);
  for (size_t h = 0; h < 2; h++)
    for (size_t c = 0; c < NM; c++) {
      dev[h * 2 * NM + 2 * c] = u[(2 * h) * NM + c];
      dev[h * 2 * NM + 2 * c + 1] = u[(2 * h + 1) * NM + c];
    }
  HIPCHK(hipMemcpyAsync(s->d.U, dev.data(), dev.size() * sizeof(double), hipMemcpyHostToDevice, run_stream(s)));
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  return TEUK_OK;
}

int teuk_get_fields(teuk_solver *s, double *u) {
  if (!s || !u) return TEUK_EINVAL;
  const size_t NM = (size_t)s->P.N * s->P.M;
  std::vector<double> dev(4 * NM);
  HIPCHK(hipMemcpyAsync(dev.data(), s->d.U, dev.size() * sizeof(double), hipMemcpyDeviceToHost, run_stream(s)));
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  for (size_t h = 0; h < 2; h++)
    for (size_t c = 0; c < NM; c++) {
      u[(2 * h) * NM + c] = dev[h * 2 * NM + 2 * c];
      u[(2 * h + 1) * NM + c] = dev[h * 2 * NM + 2 * c + 1];
    }
  return TEUK_OK;
}

int teuk_set_source(teuk_solver *s, const teuk_source_step *steps, int nsteps) {
  if (!s || nsteps < 0) return TEUK_EINVAL;
  if (s->batch) return TEUK_EINVAL;  // the batch's descriptors hold the schedule
  if (!steps || nsteps == 0) {
    // A captured graph holds the source launches and the source-on stage
    // kernels: drop it with the schedule.
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->gexec) {
      (void)hipGraphExecDestroy(s->gexec);
      s->gexec = nullptr;
    }
    if (s->sched_blk) (void)hipFree(s->sched_blk);
    s->sched_blk = nullptr;
    s->d.sched = nullptr;
    s->d.nsched = 0;
    s->has_src = false;
    return TEUK_OK;
  }
  if (!source_available()) return TEUK_EINVAL;
  static_assert(sizeof(teuk_source_step) == sizeof(SrcStep), "source step layout");
  const SrcStep *raw = reinterpret_cast<const SrcStep *>(steps);
  for (int k = 0; k < nsteps; k++)
    if (!(raw[k].rp > s->g.mass + std::sqrt(s->g.mass * s->g.mass - s->g.a * s->g.a))) return TEUK_EINVAL;
  // Host: per-step transcendentals and patch placement; device: the T_4
  // coefficients of every step, one thread each.
  std::vector<SrcPrep> prep((size_t)nsteps);
  for (int k = 0; k < nsteps; k++) host_source_prepare(s->g, raw[k], prep[k]);
  std::vector<teuk_t4::StepIn> hin((size_t)nsteps);
  std::vector<SrcDev> hdev((size_t)nsteps);
  for (int k = 0; k < nsteps; k++) {
    hin[k] = prep[k].in;
    hdev[k] = prep[k].dev;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  // The schedule and the particle states it is computed from, in one
  // allocation and one copy.
  const size_t b_dev = ((size_t)nsteps * sizeof(SrcDev) + 255) & ~(size_t)255;
  const size_t b_in = (size_t)nsteps * sizeof(teuk_t4::StepIn);
  char *blk = nullptr;
  if (hipMalloc(reinterpret_cast<void **>(&blk), b_dev + b_in) != hipSuccess) return TEUK_ENOMEM;
  SrcDev *dst = reinterpret_cast<SrcDev *>(blk);
  teuk_t4::StepIn *din = reinterpret_cast<teuk_t4::StepIn *>(blk + b_dev);
  {
    std::vector<char> staging(b_dev + b_in, 0);
    std::memcpy(staging.data(), hdev.data(), (size_t)nsteps * sizeof(SrcDev));
    std::memcpy(staging.data() + b_dev, hin.data(), b_in);
    if (hipMemcpy(blk, staging.data(), b_dev + b_in, hipMemcpyHostToDevice) != hipSuccess ||
        launch_source_coeffs(din, dst, nsteps, s->stream) != hipSuccess ||
        hipStreamSynchronize(s->stream) != hipSuccess) {
      (void)hipFree(blk);
      return TEUK_EHIP;
    }
  }
  // the stream is idle: the previous schedule is no longer read
  if (s->sched_blk) (void)hipFree(s->sched_blk);
  s->sched_blk = blk;
  s->d.sched = dst;
  s->d.nsched = nsteps;
  s->has_src = true;
  if (s->gexec) {  // captured graphs hold the old pointers
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  return TEUK_OK;
}

int teuk_set_trajectory(teuk_solver *s, const double *rows, int nrows, int nsteps) {
  if (!s || !rows || nrows <= 0 || nsteps <= 0) return TEUK_EINVAL;
  std::vector<SrcStep> sched;
  int rc = s->g.stag ? consume_trajectory_half(rows, nrows, nsteps, s->g.start, s->P.dt, sched)
                     : consume_trajectory(rows, nrows, nsteps, s->g.start, s->P.dt, sched);
  if (rc) return rc;
  return teuk_set_source(s, reinterpret_cast<const teuk_source_step *>(sched.data()), (int)sched.size());
}

int teuk_trajectory_schedule(const double *rows, int nrows, int nsteps, double start, double dt,
                             teuk_source_step *out) {
  if (!rows || !out || nrows <= 0 || nsteps <= 0 || !(dt > 0.0)) return TEUK_EINVAL;
  std::vector<SrcStep> sched;
  int rc = consume_trajectory(rows, nrows, nsteps, start, dt, sched);
  if (rc) return rc;
  std::memcpy(out, sched.data(), sched.size() * sizeof(SrcStep));
  return TEUK_OK;
}

int teuk_trajectory_schedule_staggered(const double *rows, int nrows, int nsteps, double start, double dt,
                                       teuk_source_step *out) {
  if (!rows || !out || nrows <= 0 || nsteps <= 0 || !(dt > 0.0)) return TEUK_EINVAL;
  std::vector<SrcStep> sched;
  int rc = consume_trajectory_half(rows, nrows, nsteps, start, dt, sched);
  if (rc) return rc;
  std::memcpy(out, sched.data(), sched.size() * sizeof(SrcStep));
  return TEUK_OK;
}

int teuk_step(teuk_solver *s, int nsteps) {
  if (!s || nsteps < 0) return TEUK_EINVAL;
  if (s->batch) return TEUK_EINVAL;  // a batch drives this solver
  if (s->k_host + nsteps > s->nt) return TEUK_ERANGE;
  if (s->has_src && s->k_host + nsteps > sched_steps(s)) return TEUK_ERANGE;
  int rc = TEUK_OK;
  // Direct launches of n steps, closed by one finalize.
  auto direct = [&](int n) {
    int r = TEUK_OK;
    for (int i = 0; i < n && r == TEUK_OK; i++) r = enqueue_step(s, i);
    if (r == TEUK_OK && n > 0 && launch_sentinel(s->g, s->d, s->stream) != hipSuccess) r = TEUK_EHIP;
    if (r == TEUK_OK && n > 0 && launch_finalize(s->g, s->d, n, s->stream) != hipSuccess) r = TEUK_EHIP;
    return r;
  };
  if (s->timing) {
    rc = direct(nsteps);
  } else {
    const int G = 32;
    int left = nsteps;
    if (left >= G) {
      if (!s->gexec || s->graph_steps != G) rc = build_graph(s, G);
      while (rc == TEUK_OK && left >= G) {
        if (hipGraphLaunch(s->gexec, s->stream) != hipSuccess) rc = TEUK_EHIP;
        left -= G;
      }
    }
    if (rc == TEUK_OK) rc = direct(left);
  }
  if (rc == TEUK_OK) s->k_host += nsteps;
  return rc;
}

int teuk_sync(teuk_solver *s) {
  if (!s) return TEUK_EINVAL;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  if (s->timing) flush_timing(s);
  Ctl c;
  HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  return c.status ? TEUK_EDIVERGED : TEUK_OK;
}

int teuk_status(teuk_solver *s) {
  if (!s) return TEUK_EINVAL;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  Ctl c;
  HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  return c.status ? TEUK_EDIVERGED : TEUK_OK;
}

int teuk_progress(teuk_solver *s, int *k, double *t) {
  if (!s) return TEUK_EINVAL;
  Ctl c;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  if (k) *k = c.k;
  if (t) *t = c.timer;
  return TEUK_OK;
}

int teuk_read_output(teuk_solver *s, int first, int max, teuk_output_row *rows, int *n) {
  if (!s || !rows || first < 0 || max < 0) return TEUK_EINVAL;
  Ctl c;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  const int cnt = std::max(0, std::min(max, c.nrec - first));
  if (cnt > 0) {
    std::vector<double> rec((size_t)cnt * 16);
    std::vector<int> rk((size_t)cnt * 2);
    HIPCHK(hipMemcpy(rec.data(), s->d.ext.rec + (size_t)first * 16, rec.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(rk.data(), s->d.ext.rec_k + (size_t)first * 2, rk.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < cnt; i++) {
      rows[i].k = rk[2 * i];
      rows[i].write = rk[2 * i + 1];
      rows[i].t = rec[16 * i];
      rows[i].dedt = rec[16 * i + 1];
      for (int j = 0; j < 14; j++) rows[i].h[j] = rec[16 * i + 2 + j];
    }
  }
  if (n) *n = cnt;
  return TEUK_OK;
}

int teuk_read_psi(teuk_solver *s, int first, int max, double *psi, int *n) {
  if (!s || !psi || first < 0 || max < 0) return TEUK_EINVAL;
  Ctl c;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  const int cnt = std::max(0, std::min(max, c.nrec - first));
  // Each snapshot holds Psi (re, im) followed by the two time integrals.
  // One contiguous copy and a host-side gather: a strided 2D copy into
  // pageable memory costs milliseconds.
  const size_t M = (size_t)s->P.M;
  if (cnt > 0) {
    std::vector<double> snap((size_t)cnt * 6 * M);
    HIPCHK(hipMemcpy(snap.data(), s->d.ext.snap + (size_t)first * 6 * M, snap.size() * sizeof(double),
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < cnt; i++)
      std::memcpy(psi + (size_t)i * 2 * M, snap.data() + (size_t)i * 6 * M, 2 * M * sizeof(double));
  }
  if (n) *n = cnt;
  return TEUK_OK;
}

int teuk_enable_timing(teuk_solver *s, int on) {
  if (!s) return TEUK_EINVAL;
  if (s->timing && !on) flush_timing(s);
  s->timing = on != 0;
  s->timed_ms = 0.0;
  s->timed_launches = 0;
  return TEUK_OK;
}

int teuk_timing(teuk_solver *s, long long *launches, double *ms) {
  if (!s) return TEUK_EINVAL;
  flush_timing(s);
  if (launches) *launches = s->timed_launches;
  if (ms) *ms = s->timed_ms;
  return TEUK_OK;
}

int teuk_body(const teuk_params *p, const double *r_c, const double *theta, const char *traj_path,
              const char *out_path, const char *full_path, int verbose) {
  int rc = validate(p);
  if (rc) return rc;
  teuk_solver *s = nullptr;
  if ((rc = teuk_create(p, r_c, theta, &s))) return rc;
  const int nt = s->nt;
  const bool stag = p->integrator == 2;
  // Staggered scheme: the particle radius after each step's second read,
  // which gates the output rows (Old/body.cu:166-169, :236).
  std::vector<double> rp_step;
  if (traj_path) {
    std::vector<double> rows;
    if ((rc = read_traj_file(traj_path, rows))) {
      teuk_destroy(s);
      return rc;
    }
    if ((rc = teuk_set_trajectory(s, rows.data(), (int)(rows.size() / 10), nt))) {
      teuk_destroy(s);
      return rc;
    }
    if (stag) {
      std::vector<SrcStep> sch;
      if ((rc = consume_trajectory_half(rows.data(), (int)(rows.size() / 10), nt, s->g.start, p->dt, sch))) {
        teuk_destroy(s);
        return rc;
      }
      rp_step.resize(nt);
      for (int k = 0; k < nt; k++) rp_step[k] = sch[2 * (size_t)k + 1].rp;
    }
  }
  // body.cu:208 writes 7 (h+, hx) pairs, Old/body.cu:237 five.
  const int nh = stag ? 10 : 14;
  FILE *fo = out_path ? std::fopen(out_path, "w") : nullptr;
  FILE *ff = full_path ? std::fopen(full_path, "w") : nullptr;
  if ((out_path && !fo) || (full_path && !ff)) {
    if (fo) std::fclose(fo);
    if (ff) std::fclose(ff);
    teuk_destroy(s);
    return TEUK_EIO;
  }
  if (verbose) {
    std::printf(" procs %d my_rank %d \n", 1, 0);
    std::printf(" start %lf endtime %lf \n", 0.0, p->T);
  }
  // Advance in slices so output rows stream out while the GPU runs.
  const int slice = std::max(p->Fr, 1) * 64;
  int done = 0, nread = 0;
  std::vector<teuk_output_row> rows(1024);
  while (rc == TEUK_OK && done < nt) {
    const int n = std::min(slice, nt - done);
    rc = teuk_step(s, n);
    if (rc) break;
    done += n;
    rc = teuk_sync(s);
    int got = 0;
    do {
      int r2 = teuk_read_output(s, nread, (int)rows.size(), rows.data(), &got);
      if (r2) {
        rc = r2;
        break;
      }
      for (int i = 0; i < got; i++) {
        const teuk_output_row &o = rows[i];
        const bool keep = o.write && (rp_step.empty() || rp_step[o.k] <= 12.0);
        if (keep) {
          // body.cu:208 (time column is timer - Xmax)
          if (fo) {
            std::fprintf(fo, "%lf", o.t - p->xmax);
            for (int j = 0; j < nh; j++) std::fprintf(fo, " %g", o.h[j]);
            std::fprintf(fo, "\n");
          }
          if (ff) {
            std::fprintf(ff, "%.17g", o.t - p->xmax);
            for (int j = 0; j < nh; j++) std::fprintf(ff, " %.17g", o.h[j]);
            std::fprintf(ff, "\n");
          }
        }
        if (verbose) std::printf("%lf %g \n", o.t, o.dedt);  // body.cu:210
      }
      nread += got;
    } while (got == (int)rows.size());
  }
  if (fo) std::fclose(fo);
  if (ff) std::fclose(ff);
  teuk_destroy(s);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Launch shape of one solver's stage kernels.
int teuk_set_tiling(teuk_solver *s, int waves_per_block, int waves_per_eu) {
  if (!s || s->batch || s->g.stag) return TEUK_EINVAL;
  if (!(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4)) return TEUK_EINVAL;
  if (!(waves_per_eu >= 2 && waves_per_eu <= 4)) return TEUK_EINVAL;
  s->tiling.wpb = waves_per_block;
  s->tiling.wpe = waves_per_eu;
  if (s->gexec) {
    HIPCHK(hipStreamSynchronize(s->stream));
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  return TEUK_OK;
}

int teuk_set_theta_rows(teuk_solver *s, int rows) {
  if (!s || s->batch || s->g.stag) return TEUK_EINVAL;
  if (!theta_rows_ok(s->P.M, rows)) return TEUK_EINVAL;
  s->tiling.rows = rows;
  if (s->gexec) {
    HIPCHK(hipStreamSynchronize(s->stream));
    (void)hipGraphExecDestroy(s->gexec);
    s->gexec = nullptr;
  }
  return TEUK_OK;
}

// ---------------------------------------------------------------------------
// Batches.
namespace {

int member_steps(const teuk_solver *s) { return s->has_src ? std::min(s->nt, sched_steps(s)) : s->nt; }

void batch_drop_graph(teuk_batch *b) {
  for (auto &g : b->groups)
    if (g.gexec) {
      (void)hipGraphExecDestroy(g.gexec);
      g.gexec = nullptr;
    }
}

// Field-state bytes of a member, what its group keeps in cache: the RK
// state planes and the dbdx plane.
long long member_state_bytes(const teuk_solver *s) {
  const long long NM = (long long)s->P.N * s->P.M;
  return (long long)(state_count(s->P.integrator) * 32 + 8) * NM;
}

// Descriptors and tile tables of the members that still have steps, in
// cache groups (members in list order, a new group whenever the next member
// would take the group past group_bytes).
int batch_rebuild(teuk_batch *b) {
  HIPCHK(hipStreamSynchronize(b->stream));  // kernels may still read the old tables
  batch_drop_graph(b);
  b->groups.clear();
  std::vector<SolveDesc> hd;
  std::vector<BatchTile> table;
  long long in_group = 0;
  for (teuk_solver *s : b->members) {
    SolveDesc sd;
    std::memset(&sd, 0, sizeof sd);
    sd.g = s->g;
    sd.d = s->d;
    sd.has_src = s->has_src ? 1 : 0;
    sd.ngroups = stage_groups(s->P.M, b->tiling.wpb, b->tiling.rows);
    sd.nt = member_steps(s);
    sd.tiles = (s->P.N / kWave) * sd.ngroups;
    if (s->k_host >= sd.nt) continue;
    const long long sb = member_state_bytes(s);
    if (b->groups.empty() || (b->group_bytes > 0 && in_group > 0 && in_group + sb > b->group_bytes)) {
      teuk_batch::Group g;
      g.first = (int)hd.size();
      b->groups.push_back(g);
      in_group = 0;
    }
    teuk_batch::Group &g = b->groups.back();
    g.count++;
    in_group += sb;
    if (s->has_src) g.max_cells = std::max(g.max_cells, s->P.ptsx * s->P.ptsy);
    hd.push_back(sd);
  }
  // Tile tables, one per group (member index relative to the group's first
  // descriptor).  Blocks are dealt round-robin over the 8 XCDs: block i lands
  // on XCD i % 8.  Order 0 gives XCD x the contiguous run x*per ..
  // (x+1)*per of the group's list, so neighbouring tiles of one solve share
  // an L2; no-op tiles pad the table.  With 2-row theta chunks, within each
  // XCD's run (order 0) or the whole list (order 1) the block-rows holding a
  // strip's first or last chunk come first: their waves also fill the theta
  // boundary rows, the longest of the launch, and dispatched last they would
  // set its end (4- and 8-row edge chunks compute fewer rows than the others,
  // and there the plain order measured faster).
  BatchTile pad;
  std::memset(&pad, 0, sizeof pad);
  pad.tile = -1;
  const bool edge_rows_first = b->tiling.rows == 2;
  auto edge_first = [edge_rows_first](std::vector<BatchTile>::iterator first, std::vector<BatchTile>::iterator last) {
    if (!edge_rows_first) return;
    std::stable_partition(first, last, [](const BatchTile &bt) {
      const int grp = bt.tile % bt.sd.ngroups;
      return grp == 0 || grp == bt.sd.ngroups - 1;
    });
  };
  for (auto &g : b->groups) {
    std::vector<BatchTile> list;
    for (int i = 0; i < g.count; i++)
      for (int t = 0; t < hd[g.first + i].tiles; t++) {
        BatchTile bt = pad;
        bt.sd = hd[g.first + i];
        bt.tile = t;
        list.push_back(bt);
      }
    g.tile_off = table.size();
    if (b->order == 0) {
      const size_t T = list.size(), per = (T + 7) / 8;
      for (size_t x = 0; x < 8; x++) {
        const size_t lo = std::min(T, x * per), hi = std::min(T, (x + 1) * per);
        edge_first(list.begin() + lo, list.begin() + hi);
      }
      for (size_t blk = 0; blk < per * 8; blk++) {
        const size_t src = (blk & 7) * per + (blk >> 3);
        table.push_back(src < T ? list[src] : pad);
      }
    } else {
      edge_first(list.begin(), list.end());
      table.insert(table.end(), list.begin(), list.end());
    }
    g.ntiles = (int)(table.size() - g.tile_off);
  }
  const int n = (int)hd.size();
  if (n > b->desc_cap) {
    if (b->d_desc) (void)hipFree(b->d_desc);
    b->d_desc = nullptr;
    b->desc_cap = 0;
    if (hipMalloc(&b->d_desc, (size_t)std::max(n, 1) * sizeof(SolveDesc)) != hipSuccess) return TEUK_ENOMEM;
    b->desc_cap = std::max(n, 1);
  }
  if (table.size() > b->tiles_cap) {
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    b->d_tiles = nullptr;
    b->tiles_cap = 0;
    if (hipMalloc(&b->d_tiles, std::max<size_t>(table.size(), 1) * sizeof(BatchTile)) != hipSuccess)
      return TEUK_ENOMEM;
    b->tiles_cap = std::max<size_t>(table.size(), 1);
  }
  if (n > 0) HIPCHK(hipMemcpy(b->d_desc, hd.data(), (size_t)n * sizeof(SolveDesc), hipMemcpyHostToDevice));
  if (!table.empty())
    HIPCHK(hipMemcpy(b->d_tiles, table.data(), table.size() * sizeof(BatchTile), hipMemcpyHostToDevice));
  b->dirty = false;
  return TEUK_OK;
}

// nsteps steps of one group: per step the source patches (every kRing
// steps) and one launch per RK stage over the group's tiles; then the
// sentinel and the output records of the group's members.
int batch_enqueue(teuk_batch *b, const teuk_batch::Group &g, int nsteps) {
  const SolveDesc *descs = b->d_desc + g.first;
  const BatchTile *tiles = b->d_tiles + g.tile_off;
  const int nstage = b->integrator == 0 ? 3 : 4;
  for (int k = 0; k < nsteps; k++) {
    if (g.max_cells > 0 && k % kRing == 0)
      HIPCHK(launch_source_batch(descs, g.count, g.max_cells, k, kRing, b->stream));
    for (int st = 1; st <= nstage; st++)
      HIPCHK(launch_stage_batch(tiles, g.ntiles, b->integrator, st, k, b->tiling, b->stream));
  }
  HIPCHK(launch_sentinel_batch(descs, g.count, b->stream));
  HIPCHK(launch_finalize_batch(descs, g.count, nsteps, b->stream));
  return TEUK_OK;
}

int batch_build_graph(teuk_batch *b, teuk_batch::Group &g) {
  if (g.gexec) {
    (void)hipGraphExecDestroy(g.gexec);
    g.gexec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = batch_enqueue(b, g, b->chunk);
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc != TEUK_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return TEUK_EHIP;
  e = hipGraphInstantiate(&g.gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return TEUK_EHIP;
  b->graph_steps = b->chunk;
  return TEUK_OK;
}

}  // namespace

int teuk_batch_create(int integrator, long long hbm_budget, teuk_batch **out) {
  if (!out || (integrator != 0 && integrator != 1) || hbm_budget < 0) return TEUK_EINVAL;  // no staggered batches
  *out = nullptr;
  teuk_batch *b = new teuk_batch();
  b->integrator = integrator;
  b->budget = hbm_budget;
  b->tiling = Tiling{1, kDefaultWavesPerEU, kDefaultThetaRows};
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return TEUK_EHIP;
  }
  *out = b;
  return TEUK_OK;
}

void teuk_batch_destroy(teuk_batch *b) {
  if (!b) return;
  (void)hipStreamSynchronize(b->stream);
  for (teuk_solver *s : b->members) s->batch = nullptr;
  batch_drop_graph(b);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  (void)hipStreamDestroy(b->stream);
  delete b;
}

int teuk_batch_set_tiling(teuk_batch *b, int waves_per_block, int waves_per_eu, int order, int chunk) {
  if (!b) return TEUK_EINVAL;
  if (!(waves_per_block == 1 || waves_per_block == 2 || waves_per_block == 4)) return TEUK_EINVAL;
  if (!(waves_per_eu >= 2 && waves_per_eu <= 4) || (order != 0 && order != 1)) return TEUK_EINVAL;
  if (chunk < 1 || chunk > 256) return TEUK_EINVAL;
  b->tiling.wpb = waves_per_block;
  b->tiling.wpe = waves_per_eu;
  b->order = order;
  b->chunk = chunk;
  b->dirty = true;
  return TEUK_OK;
}

int teuk_batch_set_theta_rows(teuk_batch *b, int rows) {
  if (!b || !(rows == 2 || rows == 4 || rows == 8)) return TEUK_EINVAL;
  b->tiling.rows = rows;
  b->dirty = true;
  return TEUK_OK;
}

int teuk_batch_add(teuk_batch *b, teuk_solver *s) {
  if (!b || !s || s->batch) return TEUK_EINVAL;
  if (s->P.integrator != b->integrator) return TEUK_EINVAL;
  const long long bytes = teuk_footprint(&s->P, member_steps(s));
  if (b->budget > 0 && b->used + bytes > b->budget) return TEUK_ENOMEM;
  HIPCHK(hipStreamSynchronize(s->stream));  // set-up work of the solver's own stream
  s->batch = b;
  b->members.push_back(s);
  b->member_bytes.push_back(bytes);
  b->used += bytes;
  b->dirty = true;
  return TEUK_OK;
}

int teuk_batch_remove(teuk_batch *b, teuk_solver *s) {
  if (!b || !s || s->batch != b) return TEUK_EINVAL;
  HIPCHK(hipStreamSynchronize(b->stream));
  for (size_t i = 0; i < b->members.size(); i++) {
    if (b->members[i] == s) {
      b->used -= b->member_bytes[i];
      b->members.erase(b->members.begin() + i);
      b->member_bytes.erase(b->member_bytes.begin() + i);
      break;
    }
  }
  s->batch = nullptr;
  b->dirty = true;
  return TEUK_OK;
}

int teuk_batch_step(teuk_batch *b, int nsteps) {
  if (!b || nsteps < 0) return TEUK_EINVAL;
  int left = nsteps;
  while (left > 0) {
    bool any = false;
    for (teuk_solver *s : b->members) any = any || s->k_host < member_steps(s);
    if (!any) break;
    if (b->dirty) {
      int rc = batch_rebuild(b);
      if (rc) return rc;
    }
    // one chunk of steps, group after group (each group's state stays in
    // cache through its chunk)
    const int n = std::min(left, b->chunk);
    for (auto &g : b->groups) {
      if (n == b->chunk) {
        if (!g.gexec || b->graph_steps != b->chunk) {
          int rc = batch_build_graph(b, g);
          if (rc) return rc;
        }
        HIPCHK(hipGraphLaunch(g.gexec, b->stream));
      } else {
        int rc = batch_enqueue(b, g, n);
        if (rc) return rc;
      }
    }
    for (teuk_solver *s : b->members) {
      const int nt = member_steps(s);
      const int before = s->k_host;
      s->k_host = std::min(nt, s->k_host + n);
      if (before < nt && s->k_host >= nt) b->dirty = true;  // drop its tiles next chunk
    }
    left -= n;
  }
  return TEUK_OK;
}

int teuk_batch_sync(teuk_batch *b) {
  if (!b) return TEUK_EINVAL;
  HIPCHK(hipStreamSynchronize(b->stream));
  int rc = TEUK_OK;
  for (teuk_solver *s : b->members) {
    Ctl c;
    HIPCHK(hipMemcpy(&c, s->d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    if (c.status) rc = TEUK_EDIVERGED;
  }
  return rc;
}

int teuk_batch_set_cache_groups(teuk_batch *b, long long group_bytes) {
  if (!b || group_bytes < 0) return TEUK_EINVAL;
  b->group_bytes = group_bytes;
  b->dirty = true;
  return TEUK_OK;
}

int teuk_batch_info(teuk_batch *b, int *members, int *active, long long *bytes_used) {
  if (!b) return TEUK_EINVAL;
  int act = 0;
  for (teuk_solver *s : b->members) act += s->k_host < member_steps(s) ? 1 : 0;
  if (members) *members = (int)b->members.size();
  if (active) *active = act;
  if (bytes_used) *bytes_used = b->used;
  return TEUK_OK;
}

int teuk_remaining(teuk_solver *s, int *steps) {
  if (!s || !steps) return TEUK_EINVAL;
  *steps = member_steps(s) - s->k_host;
  return TEUK_OK;
}

int teuk_source_coeffs(teuk_solver *s, int first, int n, double *c) {
  if (!s || !c || first < 0 || n < 0) return TEUK_EINVAL;
  if (!s->has_src || first + n > s->d.nsched) return TEUK_ERANGE;
  if (n == 0) return TEUK_OK;
  HIPCHK(hipStreamSynchronize(run_stream(s)));
  HIPCHK(hipMemcpy2D(c, 12 * sizeof(double), s->d.sched + first, sizeof(SrcDev), 12 * sizeof(double), (size_t)n,
                     hipMemcpyDeviceToHost));
  return TEUK_OK;
}

int teuk_source_patch(teuk_solver *s, int k, double *patch, int *origin) {
  if (!s || !patch || !origin || k < 0) return TEUK_EINVAL;
  if (!s->has_src || k >= s->d.nsched) return TEUK_ERANGE;
  hipStream_t st = run_stream(s);
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(launch_source_at(s->g, s->d, k, st));
  HIPCHK(hipStreamSynchronize(st));
  const size_t cells = (size_t)s->P.ptsx * s->P.ptsy;
  HIPCHK(hipMemcpy(patch, s->d.patch + (size_t)(k % kRing) * 2 * cells, 2 * cells * sizeof(double),
                   hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(origin, s->d.patch_org + 2 * (k % kRing), 2 * sizeof(int), hipMemcpyDeviceToHost));
  return TEUK_OK;
}

// ---------------------------------------------------------------------------
// A whole solve in fixed, preallocated memory (teuk_solve_noalloc).  All of
// its device memory is `scratch` below, a file-scope device array that each
// GPU holds from the moment the library is loaded on it until the process
// ends; host-side staging uses a file-scope host array of the same size.
// Calls are serialised by one process-wide lock (they share both arrays), so
// two calls never overlap; the arrays keep the last call's contents, and
// every call clears what must start at zero.
namespace {

constexpr size_t kNoAllocBytes = (size_t)TEUK_NOALLOC_BYTES;
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocBytes];
alignas(256) char g_noalloc_stage[kNoAllocBytes];
void *g_noalloc_dev[kNoAllocDevices];   // `scratch` on each device, from hipGetSymbolAddress
std::mutex g_noalloc_lock;

// The device address of `scratch` on the current device: asked of the
// runtime once per device (with that device current), then kept.  Host code
// uses only this address for the array's memory.
int noalloc_base(char **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocDevices) return TEUK_EHIP;
  if (!g_noalloc_dev[dev]) {
    void *q = nullptr;
    if (hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)) != hipSuccess || !q) return TEUK_EHIP;
    g_noalloc_dev[dev] = q;
  }
  *out = static_cast<char *>(g_noalloc_dev[dev]);
  return TEUK_OK;
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

int teuk_solve_noalloc(const teuk_params *p, const double *r_c, const double *theta, const double *u0,
                       const teuk_source_step *steps, int nsteps, double *u_out, teuk_output_row *rows,
                       int max_rows, int *nrows) {
  if (nrows) *nrows = 0;
  if (!p || !r_c || !theta || nsteps < 0 || max_rows < 0 || (max_rows > 0 && !rows)) return TEUK_EINVAL;
  int rc = validate(p);
  if (rc) return rc;
  if (p->integrator == 2) return TEUK_EINVAL;   // the RK integrators only
  if (nsteps > scheduled_steps(p)) return TEUK_ERANGE;
  if (steps && nsteps > 0) {
    if (!source_available()) return TEUK_EINVAL;
    const double rpl = p->mass + std::sqrt(p->mass * p->mass - p->a * p->a);
    for (int k = 0; k < nsteps; k++)
      if (!(steps[k].rp > rpl)) return TEUK_EINVAL;
  }
  const bool has_src = steps && nsteps > 0;
  const size_t NM = (size_t)p->N * p->M;
  const int nrec_cap = nsteps / p->Fr + 1;
  const Layout L = plan_layout(p, nrec_cap);
  // the source schedule after the solver's pieces: SrcDev entries, then the
  // particle states the device turns into T_4 coefficients
  const size_t off_sched = L.end;
  const size_t b_dev = has_src ? round256((size_t)nsteps * sizeof(SrcDev)) : 0;
  const size_t b_in = has_src ? round256((size_t)nsteps * sizeof(teuk_t4::StepIn)) : 0;
  if (off_sched + b_dev + b_in > kNoAllocBytes) return TEUK_ETOOBIG;

  std::lock_guard<std::mutex> hold(g_noalloc_lock);
  char *base = nullptr;
  if ((rc = noalloc_base(&base))) return rc;
  char *stage = g_noalloc_stage;
  const hipStream_t st = nullptr;
  const Grid g = make_grid(p, nrec_cap);
  DevState d{};
  bind_layout(p, L, base, d);

  // tables up, the zero region cleared, coefficient set-up (kernel_init)
  std::memset(stage, 0, L.init_end);
  fill_init(p, r_c, theta, nullptr, nullptr, L, stage);
  HIPCHK(hipMemcpy(base, stage, L.init_end, hipMemcpyHostToDevice));
  HIPCHK(hipMemsetAsync(base + L.init_end, 0, L.zero_end - L.init_end, st));
  HIPCHK(launch_init_coeffs(g, d, st));
  // initial fields as (re, im) pairs of Psi and Pi (teuk_set_fields); zero if none
  if (u0) {
    double *dev = reinterpret_cast<double *>(stage);
    for (size_t h = 0; h < 2; h++)
      for (size_t c = 0; c < NM; c++) {
        dev[h * 2 * NM + 2 * c] = u0[(2 * h) * NM + c];
        dev[h * 2 * NM + 2 * c + 1] = u0[(2 * h + 1) * NM + c];
      }
    HIPCHK(hipMemcpy(d.U, stage, 4 * NM * sizeof(double), hipMemcpyHostToDevice));
  }
  // source schedule (teuk_set_source)
  if (has_src) {
    SrcDev *hdev = reinterpret_cast<SrcDev *>(stage);
    teuk_t4::StepIn *hin = reinterpret_cast<teuk_t4::StepIn *>(stage + b_dev);
    const SrcStep *raw = reinterpret_cast<const SrcStep *>(steps);
    for (int k = 0; k < nsteps; k++) {
      SrcPrep prep;
      host_source_prepare(g, raw[k], prep);
      hdev[k] = prep.dev;
      hin[k] = prep.in;
    }
    HIPCHK(hipMemcpy(base + off_sched, stage, b_dev + b_in, hipMemcpyHostToDevice));
    d.sched = reinterpret_cast<SrcDev *>(base + off_sched);
    d.nsched = nsteps;
    HIPCHK(launch_source_coeffs(reinterpret_cast<const teuk_t4::StepIn *>(base + off_sched + b_dev), d.sched,
                                nsteps, st));
  }
  // the steps, in the launch sequence of teuk_step: chunks of kRing steps
  // (patches for the chunk, the fused stages of each step), each closed by
  // the sentinel and the output records
  const Tiling tiling = default_tiling(p->N, p->M);
  const int nstage = stages_per_step(p->integrator);
  for (int done = 0; done < nsteps;) {
    const int n = std::min(kRing, nsteps - done);
    for (int i = 0; i < n; i++) {
      if (has_src && i == 0) HIPCHK(launch_source(g, d, 0, kRing, st));
      for (int s = 1; s <= nstage; s++) HIPCHK(launch_stage(g, d, p->integrator, s, has_src, i, tiling, st));
    }
    HIPCHK(launch_sentinel(g, d, st));
    HIPCHK(launch_finalize(g, d, n, st));
    done += n;
  }
  HIPCHK(hipStreamSynchronize(st));
  // results back (teuk_get_fields, teuk_read_output)
  Ctl c;
  HIPCHK(hipMemcpy(&c, d.ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  if (u_out) {
    const double *dev = reinterpret_cast<const double *>(stage);
    HIPCHK(hipMemcpy(stage, d.U, 4 * NM * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t h = 0; h < 2; h++)
      for (size_t q = 0; q < NM; q++) {
        u_out[(2 * h) * NM + q] = dev[h * 2 * NM + 2 * q];
        u_out[(2 * h + 1) * NM + q] = dev[h * 2 * NM + 2 * q + 1];
      }
  }
  const int cnt = std::max(0, std::min(max_rows, c.nrec));
  if (cnt > 0) {
    const double *rec = reinterpret_cast<const double *>(stage);
    const int *rk = reinterpret_cast<const int *>(stage + round256((size_t)cnt * 16 * sizeof(double)));
    HIPCHK(hipMemcpy(stage, d.ext.rec, (size_t)cnt * 16 * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(const_cast<int *>(rk), d.ext.rec_k, (size_t)cnt * 2 * sizeof(int), hipMemcpyDeviceToHost));
    for (int i = 0; i < cnt; i++) {
      rows[i].k = rk[2 * i];
      rows[i].write = rk[2 * i + 1];
      rows[i].t = rec[16 * i];
      rows[i].dedt = rec[16 * i + 1];
      for (int j = 0; j < 14; j++) rows[i].h[j] = rec[16 * i + 2 + j];
    }
  }
  if (nrows) *nrows = cnt;
  return c.status ? TEUK_EDIVERGED : TEUK_OK;
}
