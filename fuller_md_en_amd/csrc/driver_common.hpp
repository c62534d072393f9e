// driver_common.hpp -- host pieces shared by the two drivers: the
// --key=value option map, file naming, stop-control polling and the per-step
// phase schedule (COLD / WARM / PROD).
#pragma once

#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "fmd.h"

namespace drv {

using Opts = std::map<std::string, std::string>;

// parse_args (LJ:558-564): "--key=value" (or "--key"), keys lower-cased,
// anything not starting with "--" ignored.
inline Opts parse_args(int argc, char **argv) {
  Opts o;
  for (int i = 1; i < argc; i++) {
    std::string a = argv[i];
    if (a.substr(0, 2) != "--") continue;
    auto eq = a.find('=');
    std::string k = eq != std::string::npos ? a.substr(2, eq - 2) : a.substr(2);
    std::string v = eq != std::string::npos ? a.substr(eq + 1) : "";
    std::transform(k.begin(), k.end(), k.begin(), ::tolower);
    o[k] = v;
  }
  return o;
}

inline std::string get(const Opts &o, const std::string &k, const std::string &d) {
  auto it = o.find(k);
  return it != o.end() ? it->second : d;
}

inline bool file_exists(const std::string &p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}
inline bool dir_exists(const char *p) {
  struct stat st;
  return stat(p, &st) == 0 && S_ISDIR(st.st_mode);
}

// unique_file (LJ:579-582): base.ext, else base_1.ext ... base_9999.ext.
inline std::string unique_file(const std::string &base, const std::string &ext) {
  std::string p = base + ext;
  if (!file_exists(p)) return p;
  for (int n = 1; n <= 9999; n++) {
    std::string c = base + "_" + std::to_string(n) + ext;
    if (!file_exists(c)) return c;
  }
  return base + "_9999" + ext;
}

// build_suffix (LJ:568-577): "_key_value" for each user-set option among the
// run-defining ones, in this fixed order; ':' -> '_', trailing ".0" dropped.
inline std::string build_suffix(const Opts &opts) {
  static const char *keys[] = {"fullerene", "crystal", "cell", "temp", "pres",
                               "step",      "dt",      "init_scale", "seed"};
  std::string sfx;
  for (const char *k : keys) {
    auto it = opts.find(k);
    if (it == opts.end()) continue;
    std::string cv = it->second;
    for (auto &c : cv)
      if (c == ':') c = '_';
    if (cv.size() >= 2 && cv.substr(cv.size() - 2) == ".0") cv = cv.substr(0, cv.size() - 2);
    sfx += "_" + std::string(k) + "_" + cv;
  }
  return sfx;
}

// restart_filename (LJ:583-588): the trajectory name with "ovito_traj" ->
// "restart" and ".rst"; intermediate saves carry "_%0*d" of the step.
inline std::string restart_filename(const std::string &oname, int istep, int nsteps) {
  std::string b = oname;
  auto dt = b.rfind('.');
  if (dt != std::string::npos) b = b.substr(0, dt);
  auto p = b.find("ovito_traj");
  if (p != std::string::npos) b.replace(p, 10, "restart");
  if (istep == nsteps) return b + ".rst";
  int dg = 1;
  for (int x = nsteps; x >= 10; x /= 10) dg++;
  char buf[64];
  snprintf(buf, sizeof buf, "_%0*d", dg, istep);
  return b + buf + ".rst";
}

// Phase schedule of one global step (LJ:848-857, 879-894).
struct Phase {
  int coldstart, warmup;
  double T, T_cold;
  const char *name(int g) const {
    return g <= coldstart ? "COLD" : g <= coldstart + warmup ? "WARM" : "PROD";
  }
  fmd_step_ctrl ctrl(int g) const {
    fmd_step_ctrl c{};
    if (g <= coldstart)
      c.Tt = T_cold;
    else if (g <= coldstart + warmup)
      c.Tt = T_cold + (T - T_cold) * double(g - coldstart) / double(warmup);
    else
      c.Tt = T;
    if (coldstart > 0 && g == coldstart + 1) c.flags |= FMD_CTRL_ZERO_XI_VG;
    if (g <= coldstart) c.flags |= FMD_CTRL_ZERO_VG;
    if (g <= coldstart + warmup) {  // rescale when T > 0.1 (decided on the GPU)
      c.flags |= FMD_CTRL_RESCALE;
      c.rescale_T = g <= coldstart ? T_cold : c.Tt;
      if (g <= coldstart) c.flags |= FMD_CTRL_RESCALE_ZERO_VG;
    }
    return c;
  }
};

inline void die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, fmd_last_error());
  exit(1);
}

}  // namespace drv
