// driver_lj.cpp -- fuller_LJ_npt_md: rigid-body LJ fullerene-crystal NPT MD on
// the GPU, with the command line, output table, OVITO / restart files and
// stop control of the reference program (src/fuller_LJ_npt_md_serial_omp_acc.cpp,
// "LJ", main at LJ:638-953).  The NPT loop runs inside the engine (libfmd.so)
// in chunks between the steps where host work is due.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "driver_common.hpp"

using namespace drv;

namespace {

constexpr double kEv2Kcal = 23.06054783, kT_cold = 4.0;

void help() {
  printf(
      "fuller_LJ_npt_md -- LJ rigid-body fullerene NPT-MD (MI355X)\n\n"
      "Options:\n"
      "  --help                  Show this help\n"
      "  --fullerene=<name>      Fullerene species (default: C60)\n"
      "  --crystal=<fcc|hcp|bcc> Crystal structure (default: fcc)\n"
      "  --cell=<nc>             Unit cell repeats (default: 3)\n"
      "  --temp=<K>              Target temperature [K] (default: 298.0)\n"
      "  --pres=<GPa>            Target pressure [GPa] (default: 0.0)\n"
      "  --step=<N>              Production steps (default: 10000)\n"
      "  --dt=<fs>               Time step [fs] (default: 1.0)\n"
      "  --init_scale=<s>        Lattice scale factor (default: 1.0)\n"
      "  --seed=<n>              Random seed (default: 42)\n"
      "  --coldstart=<N>         Cold-start steps at 4K (default: 0)\n"
      "  --warmup=<N>            Warm-up ramp steps 4K->T (default: 0)\n"
      "  --from=<step>           Averaging start step (default: auto)\n"
      "  --to=<step>             Averaging end step (default: nsteps)\n"
      "  --mon=<N>               Monitor print interval (default: auto)\n"
      "  --warmup_mon=<mode>     Warmup monitor: norm|freq|some (default: norm)\n"
      "  --ovito=<N>             OVITO XYZ output interval, 0=off (default: 0)\n"
      "  --ofile=<filename>      OVITO output filename (default: auto)\n"
      "  --restart=<N>           Restart save interval, 0=off (default: 0)\n"
      "  --resfile=<path>        Resume from restart file\n"
      "  --libdir=<path>         Fullerene library dir (default: FullereneLib)\n"
      "  --sigma=<A> --epsilon=<eV> --rcut=<A>\n"
      "                          LJ parameters (default: 3.431, 2.635e-3, 3 sigma)\n");
}

struct Restart {
  int istep = 0, n = 0;
  double h[9] = {0};
  fmd_npt_state npt{};
  std::vector<double> pos, qv, vel, omg;
  bool ok = false;
};

// read_restart_lj (LJ:612-635): tagged lines; only STEP, NMOL, H, NPT, VG, MOL
// and END are read, everything else (comments, run options) is ignored.
Restart read_restart(const std::string &fname) {
  Restart rd;
  std::ifstream f(fname);
  if (!f) return rd;
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "STEP")
      ss >> rd.istep;
    else if (tag == "NMOL")
      ss >> rd.n;
    else if (tag == "H")
      for (double &x : rd.h) ss >> x;
    else if (tag == "NPT")
      ss >> rd.npt.xi >> rd.npt.Q >> rd.npt.W >> rd.npt.Pe >> rd.npt.Tt >> rd.npt.Nf;
    else if (tag == "VG")
      for (double &x : rd.npt.Vg) ss >> x;
    else if (tag == "MOL") {
      int idx;
      double v[13];
      ss >> idx;
      for (double &x : v) ss >> x;
      rd.pos.insert(rd.pos.end(), v, v + 3);
      rd.qv.insert(rd.qv.end(), v + 3, v + 7);
      rd.vel.insert(rd.vel.end(), v + 7, v + 10);
      rd.omg.insert(rd.omg.end(), v + 10, v + 13);
    } else if (tag == "END")
      break;
  }
  rd.n = (int)rd.pos.size() / 3;
  rd.ok = rd.n > 0;
  if (rd.ok) printf("  Restart loaded: %s (step %d, %d mols)\n", fname.c_str(), rd.istep, rd.n);
  return rd;
}

// write_restart_lj (LJ:590-608): same tags, order and %.15e fields.
void write_restart(const std::string &fname, int istep, const Opts &opts, const std::string &st,
                   int nc, double T, double Pe, int nsteps, double dt, int seed,
                   const std::string &fspec, double init_scale, const double *h,
                   const fmd_npt_state &npt, const std::vector<double> &pos,
                   const std::vector<double> &qv, const std::vector<double> &vel,
                   const std::vector<double> &omg, int n, int natom) {
  FILE *f = fopen(fname.c_str(), "w");
  if (!f) return;
  fprintf(f, "# RESTART fuller_LJ_npt_md (MI355X)\n# OPTIONS:");
  for (auto &kv : opts) fprintf(f, " --%s=%s", kv.first.c_str(), kv.second.c_str());
  fprintf(f, "\nSTEP %d\nNSTEPS %d\nDT %.15e\nTEMP %.15e\nPRES %.15e\n", istep, nsteps, dt, T, Pe);
  fprintf(f, "CRYSTAL %s\nNC %d\nFULLERENE %s\nINIT_SCALE %.15e\nSEED %d\n", st.c_str(), nc,
          fspec.c_str(), init_scale, seed);
  fprintf(f, "NMOL %d\nNATOM_MOL %d\n", n, natom);
  fprintf(f, "H");
  for (int i = 0; i < 9; i++) fprintf(f, " %.15e", h[i]);
  fprintf(f, "\nNPT %.15e %.15e %.15e %.15e %.15e %d\n", npt.xi, npt.Q, npt.W, npt.Pe, npt.Tt,
          npt.Nf);
  fprintf(f, "VG");
  for (int i = 0; i < 9; i++) fprintf(f, " %.15e", npt.Vg[i]);
  fprintf(f, "\n");
  for (int i = 0; i < n; i++)
    fprintf(f,
            "MOL %d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e "
            "%.15e\n",
            i + 1, pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2], qv[i * 4], qv[i * 4 + 1],
            qv[i * 4 + 2], qv[i * 4 + 3], vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2],
            omg[i * 3], omg[i * 3 + 1], omg[i * 3 + 2]);
  fprintf(f, "END\n");
  fclose(f);
}

// write_ovito (LJ:541-555): every site, with its molecule's velocity.
void write_ovito(FILE *fp, int istep, double dt, const std::vector<double> &pos,
                 const std::vector<double> &vel, const std::vector<double> &qv,
                 const std::vector<double> &body, const double *h, int n, int natom) {
  fprintf(fp, "%d\n", n * natom);
  fprintf(fp,
          "Lattice=\"%.8f %.8f %.8f %.8f %.8f %.8f %.8f %.8f %.8f\" "
          "Properties=species:S:1:pos:R:3:c_mol:I:1:vx:R:1:vy:R:1:vz:R:1 Time=%.4f Step=%d "
          "pbc=\"T T T\"\n",
          h[0], h[3], h[6], h[1], h[4], h[7], h[2], h[5], h[8], istep * dt, istep);
  for (int i = 0; i < n; i++) {
    const double *q = &qv[i * 4];
    const double w = q[0], x = q[1], y = q[2], z = q[3];
    const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                         2 * (x * y + w * z),     1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                         2 * (x * z - w * y),     2 * (y * z + w * x), 1 - 2 * (x * x + y * y)};
    for (int a = 0; a < natom; a++) {
      const double bx = body[a * 3], by = body[a * 3 + 1], bz = body[a * 3 + 2];
      fprintf(fp, "C %14.8f %14.8f %14.8f %5d %14.8e %14.8e %14.8e\n",
              pos[i * 3] + R[0] * bx + R[1] * by + R[2] * bz,
              pos[i * 3 + 1] + R[3] * bx + R[4] * by + R[5] * bz,
              pos[i * 3 + 2] + R[6] * bx + R[7] * by + R[8] * bz, i + 1, vel[i * 3],
              vel[i * 3 + 1], vel[i * 3 + 2]);
    }
  }
}

}  // namespace

int main(int argc, char **argv) {
  auto opts = parse_args(argc, argv);
  if (opts.count("help")) {
    help();
    return 0;
  }
  std::string crystal = get(opts, "crystal", "fcc");
  std::string st = crystal;
  std::transform(st.begin(), st.end(), st.begin(), ::toupper);
  const int nc = std::atoi(get(opts, "cell", "3").c_str());
  const double T = std::atof(get(opts, "temp", "298.0").c_str());
  const double Pe = std::atof(get(opts, "pres", "0.0").c_str());
  const int nsteps = std::atoi(get(opts, "step", "10000").c_str());
  const double dt = std::atof(get(opts, "dt", "1.0").c_str());
  const int seed = std::atoi(get(opts, "seed", "42").c_str());
  const double init_scale = std::atof(get(opts, "init_scale", "1.0").c_str());
  const std::string fspec = get(opts, "fullerene", "C60");
  const std::string libdir = get(opts, "libdir", "FullereneLib");
  const int coldstart = std::atoi(get(opts, "coldstart", "0").c_str());
  const int warmup = std::atoi(get(opts, "warmup", "0").c_str());
  int avg_from = std::atoi(get(opts, "from", "0").c_str());
  int avg_to = std::atoi(get(opts, "to", "0").c_str());
  const int nrec_o = std::atoi(get(opts, "ovito", "0").c_str());
  const int nrec_rst = std::atoi(get(opts, "restart", "0").c_str());
  const std::string resfile = get(opts, "resfile", "");
  const int mon_interval = std::atoi(get(opts, "mon", "0").c_str());
  const std::string warmup_mon_mode = get(opts, "warmup_mon", "norm");
  int start_step = 0;

  if (avg_to <= 0) avg_to = nsteps;
  if (avg_from <= 0) avg_from = std::max(1, nsteps - nsteps / 4);
  const int total_steps = coldstart + warmup + nsteps;
  const int gavg_from = coldstart + warmup + avg_from, gavg_to = coldstart + warmup + avg_to;
  if (avg_from < 1 || avg_from >= avg_to || avg_to > nsteps) {
    fprintf(stderr, "Error: invalid --from/--to range\n");
    return 1;
  }
  if (st != "FCC" && st != "HCP" && st != "BCC") {
    fprintf(stderr, "Error: unknown crystal %s\n", crystal.c_str());
    return 1;
  }

  const std::string sfx = build_suffix(opts);
  const std::string ovito_file = opts.count("ofile")
                                     ? get(opts, "ofile", "")
                                     : unique_file("ovito_traj_LJ_gpu" + sfx, ".xyz");

  char path[1024], label[128];
  if (fmd_resolve_fullerene(fspec.c_str(), libdir.c_str(), path, sizeof path, label,
                            sizeof label) != FMD_OK) {
    fprintf(stderr, "Unknown fullerene: %s\n", fspec.c_str());
    return 1;
  }
  std::vector<double> body(84 * 3);
  int natom = 0;
  double props[4];
  if (fmd_load_cc1(path, body.data(), &natom, props) != FMD_OK) {
    fprintf(stderr, "Error: cannot open %s\n", path);
    return 1;
  }
  body.resize(natom * 3);
  const double Rmol = props[0], Dmol = props[1], I0 = props[2], Mmol = props[3];
  fmd_rigid_config cfg;
  fmd_rigid_config_default(&cfg, Rmol);
  if (opts.count("sigma") || opts.count("epsilon") || opts.count("rcut")) {
    const double sg = std::atof(get(opts, "sigma", "3.431").c_str());
    const double ep = std::atof(get(opts, "epsilon", "2.635e-3").c_str());
    const double rc = opts.count("rcut") ? std::atof(get(opts, "rcut", "0").c_str()) : 3.0 * sg;
    const double sr = sg / rc, sr6 = sr * sr * sr * sr * sr * sr;
    cfg.lj.sigma = sg;
    cfg.lj.eps = ep;
    cfg.lj.rcut = rc;
    cfg.lj.vshift = 4.0 * ep * (sr6 * sr6 - sr6);
    cfg.rmcut = rc + 2 * Rmol + 1.0;
  }
  const double a0 = fmd_default_a0(Dmol, st.c_str(), init_scale);
  const int Nmax = (st == "FCC" ? 4 : 2) * nc * nc * nc;
  std::vector<double> pos(Nmax * 3), vel(Nmax * 3), omg(Nmax * 3), qv(Nmax * 4);
  double h[9];
  const int N = fmd_make_crystal(st.c_str(), a0, nc, pos.data(), h);

  char dev[256] = "?";
  int cus = 0;
  double hbm = 0;
  fmd_device_info(dev, sizeof dev, &cus, &hbm);
  printf("========================================================================\n");
  printf("  Fullerene Crystal NPT-MD — LJ rigid-body (MI355X GPU)\n");
  printf("========================================================================\n");
  printf("  GPU device      : %s, %d CUs, %.1f GB HBM\n", dev, cus, hbm / 1073741824.0);
  printf("  Fullerene       : %s (%d atoms/mol)\n", label, natom);
  printf("  Crystal         : %s %dx%dx%d  Nmol=%d\n", st.c_str(), nc, nc, nc, N);
  printf("  a0=%.3f A  T=%.1f K  P=%.4f GPa  dt=%.2f fs\n", a0, T, Pe, dt);
  if (coldstart > 0) printf("  Coldstart       : %d steps at %.1f K\n", coldstart, kT_cold);
  if (warmup > 0) printf("  Warmup          : %d steps (%.1fK->%.1fK)\n", warmup, kT_cold, T);
  printf("  Production      : %d steps  avg=%d-%d\n", nsteps, avg_from, avg_to);
  printf("  Total           : %d steps\n", total_steps);
  printf("========================================================================\n\n");

  const double T_init = (coldstart > 0 || warmup > 0) ? kT_cold : T;
  fmd_rigid_init_state(N, (unsigned)seed, T_init, Mmol, I0, vel.data(), omg.data(), qv.data());
  fmd_npt_state npt;
  fmd_make_npt(&npt, T, Pe, 6 * N - 3);
  npt.Tt = T_init;
  if (!resfile.empty()) {
    Restart rd = read_restart(resfile);
    if (rd.ok) {
      start_step = rd.istep;
      for (int i = 0; i < 9; i++) h[i] = rd.h[i];
      npt = rd.npt;
      const int nn = std::min(N, rd.n);
      for (int i = 0; i < nn * 3; i++) {
        pos[i] = rd.pos[i];
        vel[i] = rd.vel[i];
        omg[i] = rd.omg[i];
      }
      for (int i = 0; i < nn * 4; i++) qv[i] = rd.qv[i];
      printf("  Restarting from global step %d\n", start_step);
    }
  }

  cfg.n_mol = N;
  cfg.natom = natom;
  cfg.body = body.data();
  cfg.I0 = I0;
  cfg.Mmol = Mmol;
  cfg.dt = dt;
  fmd_engine *eng = nullptr;
  if (fmd_rigid_create(&cfg, &eng) != FMD_OK) die("engine");
  if (fmd_rigid_set_state(eng, pos.data(), vel.data(), qv.data(), omg.data(), h, &npt) ||
      fmd_engine_prepare(eng))
    die("initial forces");

  const int prn = mon_interval > 0 ? mon_interval : std::max(1, total_steps / 50);
  int prn_pre = prn;
  if (coldstart + warmup > 0) {
    int div = 100;
    if (warmup_mon_mode == "freq")
      div = 10;
    else if (warmup_mon_mode == "some")
      div = 1000;
    prn_pre = std::max(1, (coldstart + warmup) / div);
  }
  const Phase phase{coldstart, warmup, T, kT_cold};
  double sT = 0, sP = 0, sa = 0, sEp = 0;
  int nav = 0;
  auto t0 = std::chrono::steady_clock::now();
  FILE *io_o = nrec_o > 0 ? fopen(ovito_file.c_str(), "w") : nullptr;
  printf("  %8s %5s %7s %9s %8s %10s %13s %7s\n", "step", "phase", "T[K]", "P[GPa]", "a[A]",
         "Ecoh[eV]", "Ecoh[kcal/m]", "t[s]");

  auto host_work = [&](int g) {
    const int cur = g <= coldstart + warmup ? prn_pre : prn;
    return g % cur == 0 || g == total_steps || (nrec_o > 0 && g % nrec_o == 0) ||
           (nrec_rst > 0 && g % nrec_rst == 0);
  };
  auto fetch = [&]() {
    if (fmd_rigid_get_state(eng, pos.data(), vel.data(), qv.data(), omg.data(), h, &npt))
      die("state");
  };
  std::vector<fmd_step_ctrl> ctrl;
  std::vector<fmd_thermo> th;
  bool stop_requested = false, done = false;
  int g = start_step + 1;
  while (g <= total_steps && !done) {
    int e = g;
    while (e < total_steps && !host_work(e)) e++;
    const int n = e - g + 1;
    ctrl.resize(n);
    th.resize(n);
    for (int k = 0; k < n; k++) ctrl[k] = phase.ctrl(g + k);
    if (fmd_engine_run(eng, g - 1, n, ctrl.data(), th.data())) die("run");
    for (int k = 0; k < n; k++) {
      const int gs = g + k;
      const double Ec = th[k].Ep / N, an = th[k].h[0] / nc;
      if (gs >= gavg_from && gs <= gavg_to) {
        sT += th[k].T;
        sP += th[k].P;
        sa += an;
        sEp += Ec;
        nav++;
      }
    }
    const fmd_thermo &t = th[n - 1];
    const double Ec = t.Ep / N, an = t.h[0] / nc;
    if (io_o && e % nrec_o == 0) {
      fetch();
      write_ovito(io_o, e, dt, pos, vel, qv, body, h, N, natom);
      fflush(io_o);
    }
    if (nrec_rst > 0 && (e % nrec_rst == 0 || e == total_steps)) {
      fetch();
      write_restart(restart_filename(ovito_file, e, total_steps), e, opts, st, nc, T, Pe, nsteps,
                    dt, seed, fspec, init_scale, h, npt, pos, qv, vel, omg, N, natom);
      if (stop_requested) {
        printf("\n  *** Stopped at restart checkpoint (step %d) ***\n", e);
        break;
      }
    }
    const int cur = e <= coldstart + warmup ? prn_pre : prn;
    if (e % cur == 0 || e == total_steps) {
      if (dir_exists("abort.md")) {
        printf("\n  *** abort.md detected at step %d ***\n", e);
        if (nrec_rst > 0) {
          fetch();
          write_restart(restart_filename(ovito_file, e, total_steps), e, opts, st, nc, T, Pe,
                        nsteps, dt, seed, fspec, init_scale, h, npt, pos, qv, vel, omg, N, natom);
        }
        done = true;
        break;
      }
      if (!stop_requested && dir_exists("stop.md")) {
        stop_requested = true;
        printf("\n  *** stop.md detected at step %d — will stop at next checkpoint ***\n", e);
      }
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      printf("  %8d %5s %7.1f %9.3f %8.3f %10.5f %13.4f %7.0f\n", e, phase.name(e), t.T, t.P, an,
             Ec, Ec * kEv2Kcal, el);
      fflush(stdout);
    }
    g = e + 1;
  }
  if (io_o) fclose(io_o);
  if (nav > 0) {
    printf("\n========================================================================\n");
    printf("  Averages (%d samples): T=%.2f K  P=%.4f GPa  a=%.4f A  Ecoh=%.5f eV\n", nav,
           sT / nav, sP / nav, sa / nav, sEp / nav);
    printf("========================================================================\n");
  }
  printf("  Done (%.1f sec)\n",
         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  long long cnt[3];
  fmd_engine_counters(eng, cnt);
  if (cnt[0]) printf("  WARNING: %lld neighbour-list overflows (list capacity)\n", cnt[0]);
  fmd_engine_destroy(eng);
  return 0;
}
