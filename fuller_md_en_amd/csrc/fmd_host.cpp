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

// generate_c60 (C1:228-252): the truncated icosahedron of the core program.
// Its 60 vertices are the cyclic permutations of three sign-varied triples
// in units where the edge is 2, (0, 1, 3 phi), (2, 1 + 2 phi, phi) and
// (1, 2 + phi, 2 phi) with phi the golden ratio, listed family by family, then
// permutation, then signs of the components in order (the first family has
// no sign on its zero component); centred and scaled by 0.72 A (edge 1.44 A,
// radius 3.568 A).  The vertex order is the core program's, so site sums run
// in the same order.  props = {Rmol, Dmol, I0, Mmol} as fmd_load_cc1 gives
// them, with I0 = (2/3) sum m r^2 (equal to the trace form for this shape).
int fmd_generate_c60(double *coords, double *props) {
  if (!coords || !props) return FMD_ERR_ARG;
  const double phi = (1.0 + std::sqrt(5.0)) / 2.0;
  const double fam[3][3] = {{0.0, 1.0, 3.0 * phi},
                            {2.0, 1.0 + 2.0 * phi, phi},
                            {1.0, 2.0 + phi, 2.0 * phi}};
  double t[60][3];
  int n = 0;
  for (int f = 0; f < 3; f++)
    for (int p = 0; p < 3; p++)             // component k of the triple goes to axis (p+k) mod 3
      for (int s0 = (f == 0 ? 1 : -1); s0 <= 1; s0 += 2)
        for (int s1 = -1; s1 <= 1; s1 += 2)
          for (int s2 = -1; s2 <= 1; s2 += 2) {
            const int sg[3] = {s0, s1, s2};
            for (int k = 0; k < 3; k++) t[n][(p + k) % 3] = sg[k] * fam[f][k];
            n++;
          }
  double cm[3] = {0, 0, 0};
  for (int i = 0; i < 60; i++)
    for (int a = 0; a < 3; a++) cm[a] += t[i][a];
  for (int a = 0; a < 3; a++) cm[a] /= 60.0;
  double rmol = 0, isum = 0;
  for (int i = 0; i < 60; i++) {
    for (int a = 0; a < 3; a++) t[i][a] = (t[i][a] - cm[a]) * 0.72;
    const double r2 = t[i][0] * t[i][0] + t[i][1] * t[i][1] + t[i][2] * t[i][2];
    rmol = std::max(rmol, std::sqrt(r2));
    isum += kMassC * r2;
  }
  double dmol = 0;
  for (int i = 0; i < 60; i++)
    for (int j = i + 1; j < 60; j++) {
      const double dx = t[i][0] - t[j][0], dy = t[i][1] - t[j][1], dz = t[i][2] - t[j][2];
      dmol = std::max(dmol, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
  for (int i = 0; i < 60; i++)
    for (int a = 0; a < 3; a++) coords[3 * i + a] = t[i][a];
  props[0] = rmol;
  props[1] = dmol;
  props[2] = isum * 2.0 / 3.0;
  props[3] = 60 * kMassC;
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

// ------------------------------------------------------ molecular mechanics
namespace {
// Adjacency of build_flat_topology (MM:256-257): the bonds in order, each
// appended to both of its atoms' lists.
std::vector<std::vector<int>> mm_adjacency(int natom, const int *bonds, int nbond) {
  std::vector<std::vector<int>> adj(natom);
  for (int b = 0; b < nbond; b++) {
    adj[bonds[2 * b]].push_back(bonds[2 * b + 1]);
    adj[bonds[2 * b + 1]].push_back(bonds[2 * b]);
  }
  return adj;
}

bool mm_bonds_ok(int natom, const int *bonds, int nbond) {
  if (natom < 1 || nbond < 0 || (nbond && !bonds)) return false;
  for (int b = 0; b < 2 * nbond; b++)
    if (bonds[b] < 0 || bonds[b] >= natom) return false;
  return true;
}

// Torsion angle i-j-k-l of the reference geometry (compute_phi0 MM:240-252).
double mm_phi0(const double *c, int i, int j, int k, int l) {
  double b1x = c[j * 3] - c[i * 3], b1y = c[j * 3 + 1] - c[i * 3 + 1],
         b1z = c[j * 3 + 2] - c[i * 3 + 2];
  double b2x = c[k * 3] - c[j * 3], b2y = c[k * 3 + 1] - c[j * 3 + 1],
         b2z = c[k * 3 + 2] - c[j * 3 + 2];
  double b3x = c[l * 3] - c[k * 3], b3y = c[l * 3 + 1] - c[k * 3 + 1],
         b3z = c[l * 3 + 2] - c[k * 3 + 2];
  double mx = b1y * b2z - b1z * b2y, my = b1z * b2x - b1x * b2z, mz = b1x * b2y - b1y * b2x;
  double nx = b2y * b3z - b2z * b3y, ny = b2z * b3x - b2x * b3z, nz = b2x * b3y - b2y * b3x;
  double mm = mx * mx + my * my + mz * mz, nn = nx * nx + ny * ny + nz * nz;
  if (mm < 1e-20 || nn < 1e-20) return 0;
  double b2 = std::sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
  double c_ = (mx * nx + my * ny + mz * nz) / std::sqrt(mm * nn);
  c_ = std::min(1.0, std::max(-1.0, c_));
  double s_ = (mx * b3x + my * b3y + mz * b3z) * b2 / std::sqrt(mm * nn);
  return std::atan2(s_, c_);
}
}  // namespace

extern "C" {

// The bond list of load_cc1 (MM:160-183): every adjacency entry (1-based in
// the file) as the pair (min, max), unique, ascending.
int fmd_load_cc1_bonds(const char *path, int *bonds, int cap, int *nbond) {
  if (!path || !nbond) return FMD_ERR_ARG;
  std::ifstream f(path);
  if (!f) return FMD_ERR_ARG;
  int n = 0;
  f >> n;
  if (n < 1 || n > kMaxNatom) return FMD_ERR_CAPACITY;
  std::vector<std::pair<int, int>> pairs;
  for (int i = 0; i < n; i++) {
    std::string line;
    std::getline(f, line);
    while (line.empty() && f.good()) std::getline(f, line);
    std::istringstream ss(line);
    std::string el;
    int idx, flag, b;
    double x, y, z;
    if (!(ss >> el >> idx >> x >> y >> z >> flag)) return FMD_ERR_ARG;
    while (ss >> b) {
      const int j = b - 1;
      if (j < 0 || j >= n) return FMD_ERR_ARG;
      pairs.push_back({std::min(i, j), std::max(i, j)});
    }
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  *nbond = (int)pairs.size();
  if ((int)pairs.size() > cap) return FMD_ERR_CAPACITY;
  for (size_t b = 0; b < pairs.size(); b++) {
    bonds[2 * b] = pairs[b].first;
    bonds[2 * b + 1] = pairs[b].second;
  }
  return FMD_OK;
}

int fmd_mm_topology_counts(int natom, const int *bonds, int nbond, int *counts4) {
  if (!counts4 || !mm_bonds_ok(natom, bonds, nbond)) return FMD_ERR_ARG;
  auto adj = mm_adjacency(natom, bonds, nbond);
  int nang = 0, ndih = 0, nimp = 0;
  for (int j = 0; j < natom; j++) {
    const int d = (int)adj[j].size();
    nang += d * (d - 1) / 2;
    nimp += d == 3;
  }
  for (int b = 0; b < nbond; b++) {
    const int bj = bonds[2 * b], bk = bonds[2 * b + 1];
    for (int i : adj[bj]) {
      if (i == bk) continue;
      for (int l : adj[bk])
        if (l != bj && l != i) ndih++;
    }
  }
  counts4[0] = nbond;
  counts4[1] = nang;
  counts4[2] = ndih;
  counts4[3] = nimp;
  return FMD_OK;
}

// build_flat_topology (MM:254-334): bonds with their geometric length;
// angles i-j-k over each pair of j's neighbours; dihedrals i-bj-bk-l over
// each bond; impropers (centre, its three neighbours) of three-coordinated
// atoms; gamma = 2 phi0 + pi and multiplicity 2 for the torsions; replicated
// molecule by molecule.
int fmd_mm_build_topology(const double *coords, int natom, const int *bonds, int nbond,
                          int Nmol, double kb, double kth, double v2_dih, double k_imp,
                          fmd_mm_topology *ft) {
  int cnt[4];
  if (!coords || !ft || Nmol < 1 || fmd_mm_topology_counts(natom, bonds, nbond, cnt))
    return FMD_ERR_ARG;
  auto adj = mm_adjacency(natom, bonds, nbond);
  const double pi = 3.14159265358979323846;
  struct Term {
    int a[4];
    double x0;
  };
  std::vector<Term> tb, ta, td, ti;
  for (int b = 0; b < nbond; b++) {
    const int bi = bonds[2 * b], bj = bonds[2 * b + 1];
    double dx = coords[bi * 3] - coords[bj * 3], dy = coords[bi * 3 + 1] - coords[bj * 3 + 1],
           dz = coords[bi * 3 + 2] - coords[bj * 3 + 2];
    tb.push_back({{bi, bj, 0, 0}, std::sqrt(dx * dx + dy * dy + dz * dz)});
  }
  for (int j = 0; j < natom; j++) {
    const auto &nj = adj[j];
    for (size_t ii = 0; ii < nj.size(); ii++)
      for (size_t kk = ii + 1; kk < nj.size(); kk++) {
        const int i = nj[ii], k = nj[kk];
        double rji[3], rjk[3];
        for (int c = 0; c < 3; c++) {
          rji[c] = coords[i * 3 + c] - coords[j * 3 + c];
          rjk[c] = coords[k * 3 + c] - coords[j * 3 + c];
        }
        double dji = std::sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
        double djk = std::sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
        double cs = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (dji * djk);
        cs = std::min(1.0, std::max(-1.0, cs));
        ta.push_back({{i, j, k, 0}, std::acos(cs)});
      }
  }
  for (int b = 0; b < nbond; b++) {
    const int bj = bonds[2 * b], bk = bonds[2 * b + 1];
    for (int i : adj[bj]) {
      if (i == bk) continue;
      for (int l : adj[bk]) {
        if (l == bj || l == i) continue;
        td.push_back({{i, bj, bk, l}, 2 * mm_phi0(coords, i, bj, bk, l) + pi});
      }
    }
  }
  for (int i = 0; i < natom; i++) {
    if (adj[i].size() != 3) continue;
    const int j = adj[i][0], k = adj[i][1], l = adj[i][2];
    ti.push_back({{i, j, k, l}, 2 * mm_phi0(coords, j, i, k, l) + pi});
  }
  ft->Nb = cnt[0] * Nmol;
  ft->Nang = cnt[1] * Nmol;
  ft->Ndih = cnt[2] * Nmol;
  ft->Nimp = cnt[3] * Nmol;
  for (int m = 0; m < Nmol; m++) {
    const int off = m * natom;
    for (int t = 0; t < cnt[0]; t++) {
      const int g = m * cnt[0] + t;
      ft->b_i0[g] = tb[t].a[0] + off;
      ft->b_i1[g] = tb[t].a[1] + off;
      ft->b_kb[g] = kb;
      ft->b_r0[g] = tb[t].x0;
    }
    for (int t = 0; t < cnt[1]; t++) {
      const int g = m * cnt[1] + t;
      ft->ang_i0[g] = ta[t].a[0] + off;
      ft->ang_i1[g] = ta[t].a[1] + off;
      ft->ang_i2[g] = ta[t].a[2] + off;
      ft->ang_kth[g] = kth;
      ft->ang_th0[g] = ta[t].x0;
    }
    for (int t = 0; t < cnt[2]; t++) {
      const int g = m * cnt[2] + t;
      ft->dih_i0[g] = td[t].a[0] + off;
      ft->dih_i1[g] = td[t].a[1] + off;
      ft->dih_i2[g] = td[t].a[2] + off;
      ft->dih_i3[g] = td[t].a[3] + off;
      ft->dih_Vn[g] = v2_dih;
      ft->dih_mult[g] = 2;
      ft->dih_gamma[g] = td[t].x0;
    }
    for (int t = 0; t < cnt[3]; t++) {
      const int g = m * cnt[3] + t;
      ft->imp_i0[g] = ti[t].a[0] + off;
      ft->imp_i1[g] = ti[t].a[1] + off;
      ft->imp_i2[g] = ti[t].a[2] + off;
      ft->imp_i3[g] = ti[t].a[3] + off;
      ft->imp_ki[g] = k_imp;
      ft->imp_gamma[g] = ti[t].x0;
    }
    if (ft->mol_id)
      for (int a = 0; a < natom; a++) ft->mol_id[off + a] = m;
  }
  return FMD_OK;
}

}  // extern "C"
