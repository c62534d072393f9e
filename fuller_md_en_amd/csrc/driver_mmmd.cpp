// driver_mmmd.cpp -- fuller_LJ_npt_mmmd: flexible fullerene-crystal NPT MD
// with a molecular-mechanics force field (bonds, angles, dihedrals and
// impropers inside molecules, LJ between them), on the GPU, with the command
// line, output table, OVITO / restart files and stop control of the reference
// program (src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp, "MM", main at
// MM:979-1291).
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

constexpr double kT_cold = 4.0, kKcal2eV = 1.0 / 23.06054783;

void help() {
  printf(
      "fuller_LJ_npt_mmmd -- MM+LJ fullerene NPT-MD (MI355X)\n\n"
      "Options:\n"
      "  --help                  Show this help\n"
      "  --fullerene=<name>      Fullerene species (default: C60)\n"
      "  --crystal=<fcc|hcp|bcc> Crystal structure (default: fcc)\n"
      "  --cell=<nc>             Unit cell repeats (default: 3)\n"
      "  --temp=<K>              Target temperature [K] (default: 298.0)\n"
      "  --pres=<GPa>            Target pressure [GPa] (default: 0.0)\n"
      "  --step=<N>              Production steps (default: 10000)\n"
      "  --dt=<fs>               Time step [fs] (default: 0.1)\n"
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
      "  --ff_kb=<kcal/mol>      Bond stretch constant (default: 469.0)\n"
      "  --ff_kth=<kcal/mol>     Angle bend constant (default: 63.0)\n"
      "  --ff_v2=<kcal/mol>      Dihedral constant (default: 14.5)\n"
      "  --ff_kimp=<kcal/mol>    Improper constant (default: 15.0)\n"
      "  --mm=<compat|physical>  Bonded forces: the reference's (not the gradient\n"
      "                          of its energy), or the exact gradient with the\n"
      "                          full virial (default: compat)\n"
      "  --ljlist=<compat|exact> LJ pairs: the reference's 400-entry half list,\n"
      "                          or every pair in the cutoff (default: compat)\n");
}

struct Restart {
  int istep = 0, na = 0;
  double h[9] = {0};
  fmd_npt_state npt{};
  std::vector<double> pos, vel;
  bool ok = false;
};

// read_restart_mmmd (MM:955-977).
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
      int idx;
      double p[3], v[3];
      ss >> idx >> p[0] >> p[1] >> p[2] >> v[0] >> v[1] >> v[2];
      rd.pos.insert(rd.pos.end(), p, p + 3);
      rd.vel.insert(rd.vel.end(), v, v + 3);
    } else if (tag == "END")
      break;
  }
  rd.na = (int)rd.pos.size() / 3;
  rd.ok = rd.na > 0;
  if (rd.ok) printf("  Restart loaded: %s (step %d, %d atoms)\n", fname.c_str(), rd.istep, rd.na);
  return rd;
}

// write_restart_mmmd (MM:930-953).
void write_restart(const std::string &fname, int istep, const Opts &opts, const std::string &st,
                   int nc, double T, double Pe, int nsteps, double dt, int seed,
                   const std::string &fspec, double init_scale, const double *h,
                   const fmd_npt_state &npt, const std::vector<double> &pos,
                   const std::vector<double> &vel, int na, int nmol, int natom_mol) {
  FILE *f = fopen(fname.c_str(), "w");
  if (!f) {
    fprintf(stderr, "Cannot write restart: %s\n", fname.c_str());
    return;
  }
  fprintf(f, "# RESTART fuller_LJ_npt_mmmd (MI355X)\n# OPTIONS:");
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
  for (int i = 0; i < na; i++)
    fprintf(f, "ATOM %d %.15e %.15e %.15e %.15e %.15e %.15e\n", i + 1, pos[i * 3],
            pos[i * 3 + 1], pos[i * 3 + 2], vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2]);
  fprintf(f, "END\n");
  fclose(f);
}

// write_ovito (MM:893-902): molecule ids printed 1-based.
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
            pos[i * 3 + 2], mol_id[i] + 1, vel[i * 3], vel[i * 3 + 1], vel[i * 3 + 2]);
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
  const double dt = std::atof(get(opts, "dt", "0.1").c_str());
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
  const std::string mmf = get(opts, "mm", "compat"), ljlist = get(opts, "ljlist", "compat");
  const double ff_kb = std::atof(get(opts, "ff_kb", "469.0").c_str()) * kKcal2eV;
  const double ff_kth = std::atof(get(opts, "ff_kth", "63.0").c_str()) * kKcal2eV;
  const double ff_v2 = std::atof(get(opts, "ff_v2", "14.5").c_str()) * kKcal2eV;
  const double ff_kimp = std::atof(get(opts, "ff_kimp", "15.0").c_str()) * kKcal2eV;
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
  int natom = 0, nbond = 0;
  double props[4];
  std::vector<int> bonds(2 * 1024);
  if (fmd_load_cc1(path, coords.data(), &natom, props) != FMD_OK ||
      fmd_load_cc1_bonds(path, bonds.data(), 1024, &nbond) != FMD_OK) {
    fprintf(stderr, "Cannot open %s\n", path);
    return 1;
  }
  const double a0 = fmd_default_a0(props[1], st.c_str(), init_scale);
  const int Nmol_max = (st == "FCC" ? 4 : 2) * nc * nc * nc;
  std::vector<double> centres(Nmol_max * 3);
  double h[9];
  const int Nmol = fmd_make_crystal(st.c_str(), a0, nc, centres.data(), h);
  const int Na = Nmol * natom;
  std::vector<double> pos(Na * 3), vel(Na * 3), mass(Na, 12.011);
  for (int m = 0; m < Nmol; m++)  // MM:1100-1106: no rotation
    for (int a = 0; a < natom; a++) {
      const int idx = m * natom + a;
      for (int k = 0; k < 3; k++) pos[idx * 3 + k] = centres[m * 3 + k] + coords[a * 3 + k];
    }
  // build_flat_topology (MM:1114): the caller's arrays, mol_id = molecule
  int cnt[4];
  if (fmd_mm_topology_counts(natom, bonds.data(), nbond, cnt) != FMD_OK) die("topology");
  const size_t nb = (size_t)cnt[0] * Nmol, nang = (size_t)cnt[1] * Nmol,
               ndih = (size_t)cnt[2] * Nmol, nimp = (size_t)cnt[3] * Nmol;
  std::vector<int> ib(2 * nb), ia(3 * nang), id(5 * ndih), ii(4 * nimp), mol_id(Na);
  std::vector<double> pb(2 * nb), pa(2 * nang), pd(2 * ndih), pi(2 * nimp);
  fmd_mm_topology ft{};
  ft.b_i0 = ib.data();
  ft.b_i1 = ib.data() + nb;
  ft.b_kb = pb.data();
  ft.b_r0 = pb.data() + nb;
  ft.ang_i0 = ia.data();
  ft.ang_i1 = ia.data() + nang;
  ft.ang_i2 = ia.data() + 2 * nang;
  ft.ang_kth = pa.data();
  ft.ang_th0 = pa.data() + nang;
  ft.dih_i0 = id.data();
  ft.dih_i1 = id.data() + ndih;
  ft.dih_i2 = id.data() + 2 * ndih;
  ft.dih_i3 = id.data() + 3 * ndih;
  ft.dih_mult = id.data() + 4 * ndih;
  ft.dih_Vn = pd.data();
  ft.dih_gamma = pd.data() + ndih;
  ft.imp_i0 = ii.data();
  ft.imp_i1 = ii.data() + nimp;
  ft.imp_i2 = ii.data() + 2 * nimp;
  ft.imp_i3 = ii.data() + 3 * nimp;
  ft.imp_ki = pi.data();
  ft.imp_gamma = pi.data() + nimp;
  ft.mol_id = mol_id.data();
  if (fmd_mm_build_topology(coords.data(), natom, bonds.data(), nbond, Nmol, ff_kb, ff_kth, ff_v2,
                            ff_kimp, &ft) != FMD_OK)
    die("topology");
  printf("  Topology/mol: %d bonds, %d angles, %d dihedrals, %d impropers\n", cnt[0], cnt[1],
         cnt[2], cnt[3]);
  printf("  Total:        %d bonds, %d angles, %d dihedrals, %d impropers\n", ft.Nb, ft.Nang,
         ft.Ndih, ft.Nimp);

  char dev[256] = "?";
  int cus = 0;
  double hbm = 0;
  fmd_device_info(dev, sizeof dev, &cus, &hbm);
  printf("========================================================================\n");
  printf("  Fullerene Crystal NPT-MD — MM Force Field (MI355X GPU)\n");
  printf("========================================================================\n");
  printf("  GPU device      : %s, %d CUs, %.1f GB HBM\n", dev, cus, hbm / 1073741824.0);
  printf("  Fullerene       : %s (%d atoms/mol)\n", label, natom);
  printf("  Crystal         : %s %dx%dx%d  Nmol=%d  Natom=%d\n", st.c_str(), nc, nc, nc, Nmol, Na);
  printf("  a0=%.3f A  T=%.1f K  P=%.4f GPa  dt=%.3f fs\n", a0, T, Pe, dt);
  printf("  Production      : %d steps  avg=%d-%d  Total=%d\n", nsteps, avg_from, avg_to,
         total_steps);
  printf("  Forces          : MM %s, LJ list %s\n", mmf.c_str(), ljlist.c_str());
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
      }
      printf("  Restarting from global step %d\n", start_step);
    }
  }

  fmd_mm_config cfg;
  fmd_mm_config_default(&cfg);
  cfg.n_atoms = Na;
  cfg.mol_id = mol_id.data();
  cfg.ft = &ft;
  cfg.dt = dt;
  cfg.mm_mode = mmf == "physical" ? FMD_MM_PHYSICAL : FMD_MM_COMPAT;
  if (ljlist == "exact") cfg.lj_max_neigh = 0;
  fmd_engine *eng = nullptr;
  if (fmd_mm_create(&cfg, &eng) != FMD_OK) die("engine");
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
  double sT = 0, sP = 0, sa = 0, sEb = 0, sEa = 0, sEd = 0, sL = 0, sE = 0;
  int nav = 0;
  auto t0 = std::chrono::steady_clock::now();
  const std::string ovito_file = unique_file("ovito_traj_mmmd_gpu", ".xyz");
  FILE *io_o = nrec_o > 0 ? fopen(ovito_file.c_str(), "w") : nullptr;
  printf("  %8s %5s %7s %9s %8s %9s %9s %9s %9s %9s %7s\n", "step", "phase", "T[K]", "P[GPa]",
         "a[A]", "E_bond", "E_angle", "E_dih", "E_LJ", "E_total", "t[s]");
  const std::string rst_base = nrec_o > 0 ? ovito_file : std::string("restart_mmmd_gpu");

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
                  seed, fspec, init_scale, h, npt, pos, vel, Na, Nmol, natom);
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
        sEb += th[k].Eb / Nmol;
        sEa += th[k].Ea / Nmol;
        sEd += (th[k].Ed + th[k].Ei) / Nmol;
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
      printf("  %8d %5s %7.1f %9.3f %8.3f %9.4f %9.4f %9.4f %9.4f %9.4f %7.0f\n", e, phase.name(e),
             t.T, t.P, t.h[0] / nc, t.Eb / Nmol, t.Ea / Nmol, (t.Ed + t.Ei) / Nmol, t.Ep_lj / Nmol,
             t.Ep / Nmol, el);
      fflush(stdout);
    }
    g = e + 1;
  }
  if (io_o) fclose(io_o);
  if (nav > 0)
    printf("\n  Averages (%d): T=%.2f P=%.4f a=%.4f  bond=%.4f ang=%.4f dih=%.4f LJ=%.4f tot=%.4f\n",
           nav, sT / nav, sP / nav, sa / nav, sEb / nav, sEa / nav, sEd / nav, sL / nav, sE / nav);
  printf("  Done (%.1f sec)\n",
         std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  long long ovf[3];
  fmd_engine_counters(eng, ovf);
  if (ovf[0]) printf("  WARNING: %lld neighbour-list overflows (list capacity)\n", ovf[0]);
  fmd_engine_destroy(eng);
  return 0;
}
