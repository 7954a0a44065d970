// body_teuk.cpp -- body() of TD_GPU_code/HighOrder/body.cu:25-27 on top of
// libteuk, for a build that keeps the reference's main.c, proj.c and parm.h
// and replaces body.cu and kern.cu (build.sh:7-22) by this file:
//
//   g++ -O -c main.c proj.c
//   g++ -O -c -I<repo>/include body_teuk.cpp
//   g++ -o teuk_main main.o proj.o body_teuk.o -L<repo>/bh_ringdown_amd/lib -lteuk
//
// build.sh compiles main.c with g++, so body() has C++ linkage here too.
// Per call (one m-mode, main.c:65-70) it reads ../Trajectories/<TRAJ_NAME>.traj
// (body.cu:77; TRAJ_NAME defaults to the name body.cu hard-codes) and writes
// hm0<m>_<TRAJ_NAME>.dat / hmm<|m|>_<TRAJ_NAME>.dat (body.cu:63-70) in the
// "%lf %g x14" format of body.cu:208, printing body.cu's stdout lines.
//
// parm.h's macros (N, M, T, Fr, ...) would rewrite the field names of
// teuk_params, so the parameter block is filled by a function defined before
// parm.h is included.  TRAJ_NAME plays the part of the trajectory name that
// TD_GPU_code/run.sh:12 substitutes into body.cu per run.
#include <cstdio>
#include <cstdlib>

#include "teuk.h"

namespace {

teuk_params body_params(int n, int m, double xmin, double xmax, double t_stop, int fr, int ptsx, int ptsy,
                        double mass, double aa, double ss, double mm, double dx, double dtheta, double dt) {
  teuk_params prm = {};
  prm.N = n;
  prm.M = m;
  prm.xmin = xmin;
  prm.xmax = xmax;
  prm.mass = mass;
  prm.a = aa;
  prm.s = ss;
  prm.m = mm;
  prm.dx = dx;
  prm.dtheta = dtheta;
  prm.dt = dt;
  prm.T = t_stop;
  prm.Fr = fr;
  prm.ptsx = ptsx;
  prm.ptsy = ptsy;
  prm.integrator = 0;  // SSP-RK3, kern.cu's scheme
  prm.j_extract = -1;  // scri, kern.cu:750
  return prm;
}

}  // namespace

#include "parm.h"

#ifndef TRAJ_NAME
#define TRAJ_NAME "a0.0_thi000_thf090.0_n"
#endif

void body(double *r_c, double *theta, double dtheta, double mass, double aa, double ss, double mm, double dx,
          double dt, int p, int my_rank) {
  if (p != 1 || my_rank != 0) {
    // body.cu:56-58 splits the time range over MPI ranks; libteuk runs the
    // whole range on one GPU.
    std::fprintf(stderr, "body: one process only (p = %d)\n", p);
    std::exit(1);
  }
  const teuk_params prm =
      body_params(N, M, Xmin, Xmax, T, Fr, PTSX, PTSY, mass, aa, ss, mm, dx, dtheta, dt);
  char out[256];
  std::snprintf(out, sizeof out, "%s%d_%s.dat", mm < 0 ? "hmm" : "hm0", (int)std::abs((int)mm), TRAJ_NAME);
  const int rc = teuk_body(&prm, r_c, theta, "../Trajectories/" TRAJ_NAME ".traj", out, nullptr, 1);
  if (rc) {
    std::fprintf(stderr, "body: m = %d: %s\n", (int)mm, teuk_strerror(rc));
    std::exit(1);
  }
}
