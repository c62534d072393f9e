// driver_airebo.cpp -- fuller_airebo_npt_md: flexible fullerene-crystal NPT MD
// with REBO bonds inside molecules and LJ between them, on the GPU, with the
// command line, output table, OVITO / restart files and stop control of the
// reference program (src/fuller_airebo_npt_md_serial_omp_acc.cpp, "AB", main at
// AB:681-963).
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

constexpr double kT_cold = 4.0, kMassC = 12.011;

void help() {
  printf(
      "fuller_airebo_npt_md -- AIREBO fullerene NPT-MD (MI355X)\n\n"
      "Options:\n"
      "  --help                  Show this help\n"
      "  --fullerene=<name>      Fullerene species (default: C60)\n"
      "  --crystal=<fcc|hcp|bcc> Crystal structure (default: fcc)\n"
      "  --cell=<nc>             Unit cell repeats (default: 3)\n"
      "  --temp=<K>              Target temperature [K] (default: 298.0)\n"
      "  --pres=<GPa>            Target pressure [GPa] (default: 0.0)\n"
      "  --step=<N>              Production steps (default: 10000)\n"
      "  --dt=<fs>               Time step [fs] (default: 0.5)\n"
      "  --init_scale=<s>        Lattice scale factor (default: 1.0)\n"
      "  --seed=<n>              Random seed (default: 42)\n"
      "  --coldstart=<N>         Cold-start steps at 4K (default: 0)\n"
      "  --warmup=<N>            Warm-up ramp steps 4K->T (default: 0)\n"
      "  --from=<step>           Averaging start step (default: auto)\n"
      "  --to=<step>             Averaging end step (default: nsteps)\n"
      "  --mon=<N>               Monitor print interval (default: auto)\n"
      "  --warmup_mon=<mode>     Warmup monitor: norm|freq|some (default: norm)\n"
      "  --ovito=<N>             OVITO XYZ output interval, 0=off (default: 0)\n"
      "  --restart=<N>           Restart save interval, 0=off (default: 0)\n"
      "  --resfile=<path>        Resume from restart file\n"
      "  --libdir=<path>         Fullerene library dir (default: FullereneLib)\n"
      "  --rebo=<compat|physical> REBO forces: the reference's, or the exact\n"
      "                          gradient with the full virial (default: compat)\n"
      "  --ljlist=<compat|exact> LJ pairs: the reference's 400-entry half list,\n"
      "                          or every pair in the cutoff (default: compat)\n");
}

struct Restart {
  int istep = 0, na = 0;
  double h[9] = {0};
  fmd_npt_state npt{};
  std::vector<double> pos, vel, mass;
  std::vector<int> mol_id;
  bool ok = false;
};

// read_restart_airebo (AB:655-677).
Restart read_restart(const std::string &fname) {
  Restart rd;
  std::ifstream f(fname);
  if (!f) {
    fprintf(stderr, "Cannot read: %s\n", fname.c_str());
    return rd;
  }
  std::string line;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ss(line);
    std::string tag;
    ss >> tag;
    if (tag == "STEP")
      ss >> rd.istep;
    else if (tag == "NATOM")
      ss >> rd.na;
    else if (tag == "H")
      for (double &x : rd.h) ss >> x;
    else if (tag == "NPT")
      ss >> rd.npt.xi >> rd.npt.Q >> rd.npt.W >> rd.npt.Pe >> rd.npt.Tt >> rd.npt.Nf;
    else if (tag == "VG")
      for (double &x : rd.npt.Vg) ss >> x;
    else if (tag == "ATOM") {
      int idx, mid;
      double m, p[3], v[3];
      ss >> idx >> mid >> m >> p[0] >> p[1] >> p[2] >> v[0] >> v[1] >> v[2];
      rd.pos.insert(rd.pos.end(), p, p + 3);
      rd.vel.insert(rd.vel.end(), v, v + 3);
      rd.mol_id.push_back(mid);
      rd.mass.push_back(m);
    } else if (tag == "END")
      break;
  }
  rd.na = (int)rd.pos.size() / 3;
  rd.ok = rd.na > 0;
  if (rd.ok) printf("  Restart loaded: %s (step %d, %d atoms)\n", fname.c_str(), rd.istep, rd.na);
  return rd;
}

// write_restart_airebo (AB:629-653).
void write_restart(const std::string &fname, int istep, const Opts &opts, const std::string &st,
                   int nc, double T, double Pe, int nsteps, double dt, int seed,
                   const std::string &fspec, double init_scale, const double *h,
                   const fmd_npt_state &npt, const std::vector<double> &pos,
                   const std::vector<double> &vel, const std::vector<int> &mol_id,
                   const std::vector<double> &mass, int na, int nmol, int natom_mol) {
  FILE *f = fopen(fname.c_str(), "w");
  if (!f) {
    fprintf(stderr, "Cannot write restart: %s\n", fname.c_str());
    return;
  }
  fprintf(f, "# RESTART fuller_airebo_npt_md (MI355X)\n# OPTIONS:");
  for (auto &kv : opts) fprintf(f, " --%s=%s", kv.first.c_str(), kv.second.c_str());
  fprintf(f, "\n");
  fprintf(f, "STEP %d\nNSTEPS %d\nDT %.15e\nTEMP %.15e\nPRES %.15e\n", istep, nsteps, dt, T, Pe);
  fprintf(f, "CRYSTAL %s\nNC %d\nFULLERENE %s\nINIT_SCALE %.15e\nSEED %d\n", st.c_str(), nc,
          fspec.c_str(), init_scale, seed);
  fprintf(f, "NMOL %d\nNATOM_MOL %d\nNATOM %d\n", nmol, natom_mol, na);
  fprintf(f, "H");
  for (int i = 0; i < 9; i++) fprintf(f, " %.15e", h[i]);
  fprintf(f, "\nNPT %.15e %.15e %.15e %.15e %.15e %d\n", npt.xi, npt.Q, npt.W, npt.Pe, npt.Tt,
          npt.Nf);
  fprintf(f, "VG");
  for (int i = 0; i < 9; i++) fprintf(f, " %.15e", npt.Vg[i]);
  fprintf(f, "\n");
  for (int i = 0; i < na; i++) {
    fprintf(f, "ATOM %d %d %.15e", i + 1, mol_id[i], mass[i]);
    fprintf(f, " %.15e %.15e %.15e", pos[i * 3], pos[i * 3 + 1], pos[i * 3 + 2]);
    fprintf(f, " %.15e %.15e %.15e\n", vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2]);
  }
  fprintf(f, "END\n");
  fclose(f);
}

// write_ovito (AB:593-601).
void write_ovito(FILE *fp, int step, double dt, const std::vector<double> &pos,
                 const std::vector<double> &vel, const std::vector<int> &mol_id, const double *h,
                 int na) {
  fprintf(fp, "%d\n", na);
  fprintf(fp,
          "Lattice=\"%.8f %.8f %.8f %.8f %.8f %.8f %.8f %.8f %.8f\" "
          "Properties=species:S:1:pos:R:3:c_mol:I:1:vx:R:1:vy:R:1:vz:R:1 Time=%.4f Step=%d "
          "pbc=\"T T T\"\n",
          h[0], h[3], h[6], h[1], h[4], h[7], h[2], h[5], h[8], step * dt, step);
  for (int i = 0; i < na; i++)
    fprintf(fp, "C %14.8f %14.8f %14.8f %5d %14.8e %14.8e %14.8e\n", pos[i * 3], pos[i * 3 + 1],
            pos[i * 3 + 2], mol_id[i], vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2]);
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
  const double dt = std::atof(get(opts, "dt", "0.5").c_str());
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
  const std::string rebo = get(opts, "rebo", "compat"), ljlist = get(opts, "ljlist", "compat");
  int start_step = 0;

  if (avg_to <= 0) avg_to = nsteps;
  if (avg_from <= 0) avg_from = std::max(1, nsteps - nsteps / 4);
  const int total_steps = coldstart + warmup + nsteps;
  const int gavg_from = coldstart + warmup + avg_from, gavg_to = coldstart + warmup + avg_to;
  if (st != "FCC" && st != "HCP" && st != "BCC") {
    fprintf(stderr, "Error: unknown crystal %s\n", crystal.c_str());
    return 1;
  }

  char path[1024], label[128];
  if (fmd_resolve_fullerene(fspec.c_str(), libdir.c_str(), path, sizeof path, label,
                            sizeof label) != FMD_OK) {
    fprintf(stderr, "Unknown: %s\n", fspec.c_str());
    return 1;
  }
  std::vector<double> coords(84 * 3);
  int natom = 0;
  double props[4];
  if (fmd_load_cc1(path, coords.data(), &natom, props) != FMD_OK) {
    fprintf(stderr, "Cannot open %s\n", path);
    return 1;
  }
  const double a0 = fmd_default_a0(props[1], st.c_str(), init_scale);
  const int Nmol_max = (st == "FCC" ? 4 : 2) * nc * nc * nc;
  std::vector<double> centres(Nmol_max * 3);
  double h[9];
  const int Nmol = fmd_make_crystal(st.c_str(), a0, nc, centres.data(), h);
  const int Na = Nmol * natom;
  std::vector<double> pos(Na * 3), vel(Na * 3), mass(Na);
  std::vector<int> mol_id(Na);
  for (int m = 0; m < Nmol; m++)  // AB:781-786: no rotation, mol_id = m + 1
    for (int a = 0; a < natom; a++) {
      const int idx = m * natom + a;
      for (int k = 0; k < 3; k++) pos[idx * 3 + k] = centres[m * 3 + k] + coords[a * 3 + k];
      mass[idx] = kMassC;
      mol_id[idx] = m + 1;
    }

  char dev[256] = "?";
  int cus = 0;
  double hbm = 0;
  fmd_device_info(dev, sizeof dev, &cus, &hbm);
  printf("========================================================================\n");
  printf("  Fullerene Crystal NPT-MD — AIREBO (MI355X GPU)\n");
  printf("========================================================================\n");
  printf("  GPU device      : %s, %d CUs, %.1f GB HBM\n", dev, cus, hbm / 1073741824.0);
  printf("  Fullerene       : %s (%d atoms/mol)\n", label, natom);
  printf("  Crystal         : %s %dx%dx%d  Nmol=%d  Natom=%d\n", st.c_str(), nc, nc, nc, Nmol, Na);
  printf("  a0=%.3f A  T=%.1f K  P=%.4f GPa  dt=%.3f fs\n", a0, T, Pe, dt);
  printf("  Production      : %d steps  avg=%d-%d  Total=%d\n", nsteps, avg_from, avg_to,
         total_steps);
  printf("  Forces          : REBO %s, LJ list %s\n", rebo.c_str(), ljlist.c_str());
  printf("========================================================================\n\n");

  const double T_init = (coldstart > 0 || warmup > 0) ? kT_cold : T;
  fmd_atomic_init_velocities(Na, (unsigned)seed, T_init, mass.data(), vel.data());
  fmd_npt_state npt;
  fmd_make_npt(&npt, T, Pe, 3 * Na - 3);
  npt.Tt = T_init;
  if (!resfile.empty()) {
    Restart rd = read_restart(resfile);
    if (rd.ok) {
      start_step = rd.istep;
      for (int i = 0; i < 9; i++) h[i] = rd.h[i];
      npt = rd.npt;
      for (int i = 0; i < Na && i < rd.na; i++) {
        for (int k = 0; k < 3; k++) {
          pos[i * 3 + k] = rd.pos[i * 3 + k];
          vel[i * 3 + k] = rd.vel[i * 3 + k];
        }
        mol_id[i] = rd.mol_id[i];
        mass[i] = rd.mass[i];
      }
      printf("  Restarting from global step %d\n", start_step);
    }
  }

  fmd_atomic_config cfg;
  fmd_atomic_config_default(&cfg);
  cfg.n_atoms = Na;
  cfg.mol_id = mol_id.data();
  cfg.mass = mass.data();
  cfg.dt = dt;
  cfg.rebo_mode = rebo == "physical" ? FMD_REBO_PHYSICAL : FMD_REBO_COMPAT;
  if (ljlist == "exact") cfg.lj_max_neigh = 0;
  fmd_engine *eng = nullptr;
  if (fmd_atomic_create(&cfg, &eng) != FMD_OK) die("engine");
  if (fmd_atomic_set_state(eng, pos.data(), vel.data(), h, &npt) || fmd_engine_prepare(eng))
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
  double sT = 0, sP = 0, sa = 0, sR = 0, sL = 0, sE = 0;
  int nav = 0;
  auto t0 = std::chrono::steady_clock::now();
  const std::string ovito_file = unique_file("ovito_traj_airebo_gpu", ".xyz");
  FILE *io_o = nrec_o > 0 ? fopen(ovito_file.c_str(), "w") : nullptr;
  printf("  %8s %5s %7s %9s %8s %11s %11s %11s %7s\n", "step", "phase", "T[K]", "P[GPa]", "a[A]",
         "E_REBO", "E_LJ", "E_total", "t[s]");
  const std::string rst_base = nrec_o > 0 ? ovito_file : std::string("restart_airebo_gpu");

  auto host_work = [&](int g) {
    const int cur = g <= coldstart + warmup ? prn_pre : prn;
    return g % cur == 0 || g == total_steps || (nrec_o > 0 && g % nrec_o == 0) ||
           (nrec_rst > 0 && g % nrec_rst == 0);
  };
  auto fetch = [&]() {
    if (fmd_atomic_get_state(eng, pos.data(), vel.data(), h, &npt)) die("state");
  };
  auto save = [&](int e) {
    fetch();
    write_restart(restart_filename(rst_base, e, total_steps), e, opts, st, nc, T, Pe, nsteps, dt,
                  seed, fspec, init_scale, h, npt, pos, vel, mol_id, mass, Na, Nmol, natom);
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
      if (gs >= gavg_from && gs <= gavg_to) {
        sT += th[k].T;
        sP += th[k].P;
        sa += th[k].h[0] / nc;
        sR += th[k].Ep_rebo / Nmol;
        sL += th[k].Ep_lj / Nmol;
        sE += th[k].Ep / Nmol;
        nav++;
      }
    }
    const fmd_thermo &t = th[n - 1];
    if (io_o && e % nrec_o == 0) {
      fetch();
      write_ovito(io_o, e, dt, pos, vel, mol_id, h, Na);
      fflush(io_o);
    }
    if (nrec_rst > 0 && (e % nrec_rst == 0 || e == total_steps)) {
      save(e);
      if (stop_requested) {
        printf("\n  *** Stopped at restart checkpoint (global step %d) ***\n", e);
        break;
      }
    }
    const int cur = e <= coldstart + warmup ? prn_pre : prn;
    if (e % cur == 0 || e == total_steps) {
      if (dir_exists("abort.md")) {
        printf("\n  *** abort.md detected at step %d ***\n", e);
        if (nrec_rst > 0) save(e);
        done = true;
        break;
      }
      if (!stop_requested && dir_exists("stop.md")) {
        stop_requested = true;
        printf("\n  *** stop.md detected at step %d — will stop at next checkpoint ***\n", e);
        if (nrec_rst == 0) {
          done = true;
          break;
        }
      }
      const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      printf("  %8d %5s %7.1f %9.3f %8.3f %11.4f %11.4f %11.4f %7.0f\n", e, phase.name(e), t.T, t.P,
             t.h[0] / nc, t.Ep_rebo / Nmol, t.Ep_lj / Nmol, t.Ep / Nmol, el);
      fflush(stdout);
    }
    g = e + 1;
  }
  if (io_o) fclose(io_o);
  if (nav > 0)
    printf("\n  Averages (%d): T=%.2f P=%.4f a=%.4f REBO=%.4f LJ=%.4f Total=%.4f\n", nav, sT / nav,
           sP / nav, sa / nav, sR / nav, sL / nav, sE / nav);
  printf("  Done (%.1f sec)\n",
         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  long long cnt[3];
  fmd_engine_counters(eng, cnt);
  if (cnt[0]) printf("  WARNING: %lld neighbour-list overflows (list capacity)\n", cnt[0]);
  fmd_engine_destroy(eng);
  return 0;
}
