// teuk_run -- the reference's main.c driver on top of libteuk.
//
// Keeps TD_GPU_code/HighOrder's run interface: the parm.h parameter names,
// the m-loop of main.c:65-70 calling body() per mode, the trajectory file
// ../Trajectories/<name>.traj read two lines per step (body.cu:97-129), and
// the waveform files hm0<m>_<name>.dat / hmm<|m|>_<name>.dat (body.cu:63-70)
// in the "%lf %g x14" format of body.cu:208.  A full-precision copy of each
// waveform file (same rows, %.17g) is written beside it as *.full.dat.
//
// Usage:
//   teuk_run [-p params.txt] [-t traj_dir] [-o out_dir] [-q] [name]
// The parameter file holds "key value" lines; keys (defaults from parm.h and
// main.c):
//   N 16000   M 256   Xmax 50.0   Xmin -50.0   Fr 50   T 1500.0
//   PTSX 128  PTSY 128   mass 1.0   aa 1e-8   dt 0.002
//   mmin -5   mmax 5   integrator 0 (1 = RK4, 2 = the staggered scheme of TD_GPU_code/Old)   j_extract -1
// `name` selects <traj_dir>/<name>.traj (traj_dir defaults to
// ../Trajectories as in body.cu:77); "-" runs without a particle source.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "teuk.h"

namespace {

struct Config {
  std::map<std::string, double> v = {
      {"N", 16000}, {"M", 256},      {"Xmax", 50.0},  {"Xmin", -50.0}, {"Fr", 50},        {"T", 1500.0},
      {"PTSX", 128}, {"PTSY", 128},  {"mass", 1.0},   {"aa", 1e-8},    {"dt", 0.002},     {"mmin", -5},
      {"mmax", 5},   {"integrator", 0}, {"j_extract", -1}};
};

bool read_params(const char *path, Config &c) {
  FILE *f = std::fopen(path, "r");
  if (!f) return false;
  char key[64];
  double val;
  char line[256];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    if (std::sscanf(line, "%63s %lf", key, &val) == 2) {
      if (!c.v.count(key)) {
        std::fprintf(stderr, "teuk_run: unknown parameter '%s'\n", key);
        std::fclose(f);
        return false;
      }
      c.v[key] = val;
    }
  }
  std::fclose(f);
  return true;
}

// body.cu:63-70, without its fixed-size buffers: "hm0" for m >= 0, "hmm"
// for m < 0, then |m|, then "_<name>.dat".
std::string wave_name(int m, const std::string &name) {
  return std::string(m < 0 ? "hmm" : "hm0") + std::to_string(std::abs(m)) + "_" + name + ".dat";
}

}  // namespace

int main(int argc, char **argv) {
  Config cfg;
  std::string traj_dir = "../Trajectories", out_dir = ".", name = "-";
  bool verbose = true;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "-p") && i + 1 < argc) {
      if (!read_params(argv[++i], cfg)) return 2;
    } else if (!std::strcmp(argv[i], "-t") && i + 1 < argc) {
      traj_dir = argv[++i];
    } else if (!std::strcmp(argv[i], "-o") && i + 1 < argc) {
      out_dir = argv[++i];
    } else if (!std::strcmp(argv[i], "-q")) {
      verbose = false;
    } else {
      name = argv[i];
    }
  }
  auto &v = cfg.v;
  teuk_params p;
  // main.c:59-61 derives dx = (Xmax-Xmin)/N and dtheta = Pi/M.
  teuk_default_params(&p, (int)v["N"], (int)v["M"], v["Xmin"], v["Xmax"], v["aa"], 0.0, 1.0, v["T"]);
  p.mass = v["mass"];
  p.dt = v["dt"];
  p.Fr = (int)v["Fr"];
  p.ptsx = (int)v["PTSX"];
  p.ptsy = (int)v["PTSY"];
  p.integrator = (int)v["integrator"];
  p.j_extract = (int)v["j_extract"];

  if (verbose) std::printf("\nWELCOME TO TEUKOLSKY CODE!\n\n");  // main.c:50-52
  std::vector<double> x_c(p.N), r_c(p.N), theta(p.M);
  int rc = teuk_grid(&p, x_c.data(), r_c.data(), theta.data());
  if (rc) {
    std::fprintf(stderr, "teuk_run: grid: %s\n", teuk_strerror(rc));
    return 1;
  }
  const std::string traj = name == "-" ? "" : traj_dir + "/" + name + ".traj";
  const std::string tag = name == "-" ? "nosource" : name;
  for (int m = (int)v["mmin"]; m <= (int)v["mmax"]; m++) {  // main.c:65-70
    p.m = m;
    const std::string out = out_dir + "/" + wave_name(m, tag);
    const std::string full = out.substr(0, out.size() - 4) + ".full.dat";
    rc = teuk_body(&p, r_c.data(), theta.data(), traj.empty() ? nullptr : traj.c_str(), out.c_str(), full.c_str(),
                   verbose ? 1 : 0);
    if (rc) {
      std::fprintf(stderr, "teuk_run: m=%d: %s\n", m, teuk_strerror(rc));
      return 1;
    }
  }
  if (verbose) std::printf("All done.\nBye bye!\n");  // main.c:73-74
  return 0;
}
