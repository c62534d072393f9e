// driver_core.cpp -- fuller_LJ_core: the rigid-body LJ "core" program
// (src/fuller_LJ_npt_md_core_serial.cpp, "C1", main at C1:605-745) on the
// GPU engine.  Same command line (the cell count nc, 1..8, as "--cell=nc" or
// a bare number), the same fixed run (analytic C60, FCC a0 = 14.17 A, 300 K,
// 0 GPa, dt = 1 fs, 1000 steps, list every 25 steps, seed 42), the same
// monitor rows every 100 steps and the same "Avg" line over the last quarter.
// The NPT loop runs inside the engine in chunks of 100 steps.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "fmd.h"

namespace {

// The core program's fixed run (C1:615-621) and molecular constants
// (C1:109-112: list radius from the nominal C60 radius 3.55 A).
constexpr int kSteps = 1000, kMon = 100, kNlup = 25;
constexpr double kT = 300.0, kPe = 0.0, kDt = 1.0, kA0 = 14.17, kRC60 = 3.55;
constexpr unsigned kSeed = 42;

void die(const char *what) {
  fprintf(stderr, "Error: %s: %s\n", what, fmd_last_error());
  exit(1);
}

}  // namespace

int main(int argc, char **argv) {
  int nc = 3;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a.rfind("--cell=", 0) == 0)
      nc = std::atoi(a.c_str() + 7);
    else if (a[0] != '-')
      nc = std::atoi(a.c_str());
  }
  if (nc < 1 || nc > 8) {
    printf("Error: nc must be 1-8 (got %d)\n", nc);
    return 1;
  }
  const int avg_from = kSteps - kSteps / 4;

  std::vector<double> body(60 * 3);
  double props[4];
  if (fmd_generate_c60(body.data(), props) != FMD_OK) die("generate_c60");
  const double I0 = props[2], Mmol = props[3];
  const int natom = 60, Nmax = 4 * nc * nc * nc;
  std::vector<double> pos(Nmax * 3), vel(Nmax * 3), omg(Nmax * 3), qv(Nmax * 4);
  double h[9];
  const int N = fmd_make_crystal("FCC", kA0, nc, pos.data(), h);
  if (N < 0) die("make_fcc");

  char dev[256] = "?";
  int cus = 0;
  double hbm = 0;
  fmd_device_info(dev, sizeof dev, &cus, &hbm);
  printf("================================================================\n");
  printf("  C60 LJ NPT-MD Core (MI355X GPU)\n");
  printf("================================================================\n");
  printf("  GPU device      : %s, %d CUs\n", dev, cus);
  printf("  FCC cell        : %dx%dx%d  N=%d molecules\n", nc, nc, nc, N);
  printf("  Atoms/molecule  : %d\n", natom);
  printf("  a0=%.2f A  T=%.0f K  P=%.1f GPa  dt=%.1f fs  steps=%d\n", kA0, kT, kPe, kDt, kSteps);
  printf("================================================================\n\n");

  // seeded state in the core program's draw order (C1:666-680)
  if (fmd_rigid_init_state(N, kSeed, kT, Mmol, I0, vel.data(), omg.data(), qv.data()))
    die("initial state");
  fmd_npt_state npt;
  fmd_make_npt(&npt, kT, kPe, 6 * N - 3);

  fmd_rigid_config cfg;
  fmd_rigid_config_default(&cfg, kRC60);  // sigma 3.431, eps 2.635e-3, rc 3 sigma, skin 3 A
  cfg.n_mol = N;
  cfg.natom = natom;
  cfg.body = body.data();
  cfg.I0 = I0;
  cfg.Mmol = Mmol;
  cfg.dt = kDt;
  cfg.nlup = kNlup;
  fmd_engine *eng = nullptr;
  if (fmd_rigid_create(&cfg, &eng) != FMD_OK) die("engine");
  if (fmd_rigid_set_state(eng, pos.data(), vel.data(), qv.data(), omg.data(), h, &npt) ||
      fmd_engine_prepare(eng))  // list, apply_pbc, forces (C1:686-696)
    die("initial forces");

  double sT = 0, sP = 0, sa = 0, sEp = 0;
  int nav = 0;
  auto t0 = std::chrono::steady_clock::now();
  printf("%8s %7s %9s %8s %10s %7s\n", "step", "T[K]", "P[GPa]", "a[A]", "Ecoh[eV]", "t[s]");
  std::vector<fmd_thermo> th(kMon);
  for (int g = 1; g <= kSteps; g += kMon) {
    const int n = std::min(kMon, kSteps - g + 1);
    if (fmd_engine_run(eng, g - 1, n, nullptr, th.data())) die("run");
    for (int k = 0; k < n; k++) {
      const int gs = g + k;
      const double Ec = th[k].Ep / N, an = th[k].h[0] / nc;
      if (gs >= avg_from) {
        sT += th[k].T;
        sP += th[k].P;
        sa += an;
        sEp += Ec;
        nav++;
      }
      if (gs % kMon == 0 || gs == kSteps) {
        const double el =
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        printf("%8d %7.1f %9.3f %8.3f %10.5f %7.0f\n", gs, th[k].T, th[k].P, an, Ec, el);
      }
    }
    fflush(stdout);
  }
  if (nav > 0)
    printf("Avg(%d): T=%.2f P=%.4f a=%.4f Ecoh=%.5f\n", nav, sT / nav, sP / nav, sa / nav,
           sEp / nav);
  printf("Done %.1fs\n",
         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  long long cnt[3];
  fmd_engine_counters(eng, cnt);
  if (cnt[0]) printf("WARNING: %lld neighbour-list overflows (list capacity)\n", cnt[0]);
  fmd_engine_destroy(eng);
  return 0;
}
