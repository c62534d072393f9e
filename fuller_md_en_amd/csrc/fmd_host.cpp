// fmd_host.cpp -- host-side system construction shared by the C++ drivers and
// the Python package: .cc1 geometry loading, fullerene name resolution,
// crystal builders and the seeded initial state.  Plain C++17 (libstdc++), so
// std::mt19937 + std::normal_distribution draw exactly the stream the
// reference programs draw.
#include <dirent.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "fmd.h"

namespace {
constexpr double kConv = 9.64853321e-3, kB = 8.617333262e-5, kMassC = 12.011;
constexpr int kMaxNatom = 84;  // LJ:109
}  // namespace

extern "C" {

// load_cc1 (LJ:178-207): natom, then "El idx x y z flag nbr..." lines; the
// molecule is centred on its geometric centre.  props = {Rmol, Dmol, I0,
// Mmol}: largest centre distance, largest atom-atom distance, isotropic
// inertia (Ixx+Iyy+Izz)/3 and mass.
int fmd_load_cc1(const char *path, double *coords, int *natom, double *props) {
  std::ifstream f(path);
  if (!f) return FMD_ERR_ARG;
  int n = 0;
  f >> n;
  if (n < 1 || n > kMaxNatom) return FMD_ERR_CAPACITY;
  std::vector<double> t(3 * n);
  for (int i = 0; i < n; i++) {
    std::string line;
    std::getline(f, line);
    while (line.empty() && f.good()) std::getline(f, line);
    std::istringstream ss(line);
    std::string el;
    int idx, flag;
    double x, y, z;
    if (!(ss >> el >> idx >> x >> y >> z >> flag)) return FMD_ERR_ARG;
    t[3 * i] = x;
    t[3 * i + 1] = y;
    t[3 * i + 2] = z;
  }
  double cm[3] = {0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) cm[a] += t[3 * i + a];
  for (int a = 0; a < 3; a++) cm[a] /= n;
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) t[3 * i + a] -= cm[a];
  double rmol = 0, dmol = 0, ixx = 0, iyy = 0, izz = 0;
  for (int i = 0; i < n; i++) {
    const double *p = &t[3 * i];
    double r2 = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
    rmol = std::max(rmol, std::sqrt(r2));
    for (int j = i + 1; j < n; j++) {
      const double *q = &t[3 * j];
      double dx = p[0] - q[0], dy = p[1] - q[1], dz = p[2] - q[2];
      dmol = std::max(dmol, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    ixx += kMassC * (r2 - p[0] * p[0]);
    iyy += kMassC * (r2 - p[1] * p[1]);
    izz += kMassC * (r2 - p[2] * p[2]);
  }
  for (int i = 0; i < 3 * n; i++) coords[i] = t[i];
  *natom = n;
  props[0] = rmol;
  props[1] = dmol;
  props[2] = (ixx + iyy + izz) / 3.0;
  props[3] = n * kMassC;
  return FMD_OK;
}

// resolve_fullerene (LJ:210-226): C60 | C70 | C72 | C74 | C76:<sym> |
// C84:<n>[:<sym>] (case-insensitive) -> library path and display label.
int fmd_resolve_fullerene(const char *spec_c, const char *lib_c, char *path, int pcap,
                          char *label, int lcap) {
  std::string spec = spec_c, lib = lib_c, sl = spec;
  std::transform(sl.begin(), sl.end(), sl.begin(), ::tolower);
  std::string p, l;
  if (sl == "buckyball" || sl == "c60" || sl == "c60:ih") {
    p = lib + "/C60-76/C60-Ih.cc1";
    l = "C60(Ih)";
  } else if (sl == "c70" || sl == "c70:d5h") {
    p = lib + "/C60-76/C70-D5h.cc1";
    l = "C70(D5h)";
  } else if (sl == "c72" || sl == "c72:d6d") {
    p = lib + "/C60-76/C72-D6d.cc1";
    l = "C72(D6d)";
  } else if (sl == "c74" || sl == "c74:d3h") {
    p = lib + "/C60-76/C74-D3h.cc1";
    l = "C74(D3h)";
  } else if (sl.rfind("c76:", 0) == 0 && sl.size() > 4) {
    std::string sym = spec.substr(4);
    p = lib + "/C60-76/C76-" + sym + ".cc1";
    l = "C76(" + sym + ")";
  } else if (sl.rfind("c84:", 0) == 0) {
    std::string rest = spec.substr(4);
    auto c = rest.find(':');
    if (c != std::string::npos) {
      int n = std::atoi(rest.substr(0, c).c_str());
      char buf[64];
      snprintf(buf, sizeof buf, "C84-No.%02d-%s.cc1", n, rest.substr(c + 1).c_str());
      p = lib + "/C84/" + buf;
      l = "C84 No." + std::to_string(n);
    } else {
      int n = std::atoi(rest.c_str());
      char pfx[32];
      snprintf(pfx, sizeof pfx, "C84-No.%02d-", n);
      std::string dpath = lib + "/C84";
      std::vector<std::string> hits;
      if (DIR *dp = opendir(dpath.c_str())) {
        while (dirent *ep = readdir(dp)) {
          std::string fn = ep->d_name;
          if (fn.rfind(pfx, 0) == 0 && fn.size() > 4 && fn.substr(fn.size() - 4) == ".cc1")
            hits.push_back(fn);
        }
        closedir(dp);
      }
      if (!hits.empty()) {
        std::sort(hits.begin(), hits.end());  // deterministic where readdir is not
        p = dpath + "/" + hits[0];
        l = "C84 No." + std::to_string(n);
      }
    }
  }
  if (p.empty()) return FMD_ERR_ARG;
  snprintf(path, pcap, "%s", p.c_str());
  if (label && lcap > 0) snprintf(label, lcap, "%s", l.c_str());
  return FMD_OK;
}

// default_a0 (LJ:251-254): 1.4 x Dmol x {sqrt2 (FCC), 1 (HCP), 2/sqrt3 (BCC)}.
double fmd_default_a0(double dmax, const char *crystal, double scale) {
  std::string st = crystal;
  std::transform(st.begin(), st.end(), st.begin(), ::toupper);
  double m = 1.4, a0;
  if (st == "FCC")
    a0 = dmax * std::sqrt(2.0) * m;
  else if (st == "HCP")
    a0 = dmax * m;
  else
    a0 = dmax * 2.0 / std::sqrt(3.0) * m;
  return a0 * scale;
}

// make_fcc / make_hcp / make_bcc (LJ:229-250).  Returns the molecule count.
int fmd_make_crystal(const char *crystal, double a, int nc, double *pos, double *h) {
  std::string st = crystal;
  std::transform(st.begin(), st.end(), st.begin(), ::toupper);
  int n = 0;
  for (int i = 0; i < 9; i++) h[i] = 0;
  if (st == "FCC") {
    const double bas[4][3] = {{0, 0, 0}, {.5 * a, .5 * a, 0}, {.5 * a, 0, .5 * a},
                              {0, .5 * a, .5 * a}};
    for (int ix = 0; ix < nc; ix++)
      for (int iy = 0; iy < nc; iy++)
        for (int iz = 0; iz < nc; iz++)
          for (int b = 0; b < 4; b++) {
            pos[n * 3] = a * ix + bas[b][0];
            pos[n * 3 + 1] = a * iy + bas[b][1];
            pos[n * 3 + 2] = a * iz + bas[b][2];
            n++;
          }
    h[0] = h[4] = h[8] = nc * a;
  } else if (st == "HCP") {
    const double c = a * std::sqrt(8.0 / 3.0);
    const double a1[3] = {a, 0, 0}, a2[3] = {a / 2, a * std::sqrt(3.0) / 2, 0}, a3[3] = {0, 0, c};
    const double bas[2][3] = {{0, 0, 0}, {1.0 / 3, 2.0 / 3, 0.5}};
    for (int ix = 0; ix < nc; ix++)
      for (int iy = 0; iy < nc; iy++)
        for (int iz = 0; iz < nc; iz++)
          for (int b = 0; b < 2; b++) {
            double fx = ix + bas[b][0], fy = iy + bas[b][1], fz = iz + bas[b][2];
            pos[n * 3] = fx * a1[0] + fy * a2[0] + fz * a3[0];
            pos[n * 3 + 1] = fx * a1[1] + fy * a2[1] + fz * a3[1];
            pos[n * 3 + 2] = fx * a1[2] + fy * a2[2] + fz * a3[2];
            n++;
          }
    h[0] = nc * a1[0];
    h[3] = nc * a1[1];
    h[1] = nc * a2[0];
    h[4] = nc * a2[1];
    h[8] = nc * a3[2];
  } else if (st == "BCC") {
    const double bas[2][3] = {{0, 0, 0}, {.5 * a, .5 * a, .5 * a}};
    for (int ix = 0; ix < nc; ix++)
      for (int iy = 0; iy < nc; iy++)
        for (int iz = 0; iz < nc; iz++)
          for (int b = 0; b < 2; b++) {
            pos[n * 3] = a * ix + bas[b][0];
            pos[n * 3 + 1] = a * iy + bas[b][1];
            pos[n * 3 + 2] = a * iz + bas[b][2];
            n++;
          }
    h[0] = h[4] = h[8] = nc * a;
  } else {
    return FMD_ERR_ARG;
  }
  return n;
}

// Seeded rigid initial state (LJ:790-801): per molecule three (v, w) draw
// pairs interleaved, then four quaternion draws, normalised; COM velocity
// removed.
int fmd_rigid_init_state(int n, unsigned seed, double T, double Mmol, double I0, double *vel,
                         double *omg, double *qv) {
  if (n < 1) return FMD_ERR_ARG;
  std::mt19937 rng(seed);
  std::normal_distribution<double> gauss(0, 1);
  double sv = std::sqrt(kB * T * kConv / Mmol), sw = std::sqrt(kB * T * kConv / I0);
  for (int i = 0; i < n; i++) {
    for (int a = 0; a < 3; a++) {
      vel[i * 3 + a] = sv * gauss(rng);
      omg[i * 3 + a] = sw * gauss(rng);
    }
    for (int a = 0; a < 4; a++) qv[i * 4 + a] = gauss(rng);
    double nn = std::sqrt(qv[i * 4] * qv[i * 4] + qv[i * 4 + 1] * qv[i * 4 + 1] +
                          qv[i * 4 + 2] * qv[i * 4 + 2] + qv[i * 4 + 3] * qv[i * 4 + 3]);
    for (int a = 0; a < 4; a++) qv[i * 4 + a] /= nn;
  }
  double vcm[3] = {0, 0, 0};
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) vcm[a] += vel[i * 3 + a];
  for (int a = 0; a < 3; a++) vcm[a] /= n;
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) vel[i * 3 + a] -= vcm[a];
  return FMD_OK;
}

// Seeded atomic velocities (AB:808-815): per atom three draws scaled by
// sqrt(kB T CONV / m_i); COM velocity removed.
int fmd_atomic_init_velocities(int na, unsigned seed, double T, const double *mass,
                               double *vel) {
  if (na < 1) return FMD_ERR_ARG;
  std::mt19937 rng(seed);
  std::normal_distribution<double> g(0, 1);
  for (int i = 0; i < na; i++) {
    double sv = std::sqrt(kB * T * kConv / mass[i]);
    for (int a = 0; a < 3; a++) vel[i * 3 + a] = sv * g(rng);
  }
  double vcm[3] = {0, 0, 0};
  for (int i = 0; i < na; i++)
    for (int a = 0; a < 3; a++) vcm[a] += vel[i * 3 + a];
  for (int a = 0; a < 3; a++) vcm[a] /= na;
  for (int i = 0; i < na; i++)
    for (int a = 0; a < 3; a++) vel[i * 3 + a] -= vcm[a];
  return FMD_OK;
}

// Normal draws of std::normal_distribution<double>(0,1) over mt19937(seed).
int fmd_normal_stream(unsigned seed, int n, double *out) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> g(0, 1);
  for (int i = 0; i < n; i++) out[i] = g(rng);
  return FMD_OK;
}

}  // extern "C"
