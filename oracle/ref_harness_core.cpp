// ref_harness_core.cpp -- exposes the rigid-LJ core reference functions
// (src/fuller_LJ_npt_md_core_serial.cpp, "C1") through a C ABI so that the
// golden-vector generator can call them on chosen inputs.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// source where it lies (never copied), output to oracle/_ref/.  Used only to
// write the numeric fixtures under tests/golden/ and to cross-check the FP64
// oracle in this container.
#define main fmd_ref_core_main
#include "fuller_LJ_npt_md_core_serial.cpp"
#undef main

#include <vector>

static void npt_from(const double *s, NPTState &n) {
  n.xi = s[0];
  n.Q = s[1];
  for (int i = 0; i < 9; i++) n.Vg[i] = s[2 + i];
  n.W = s[11];
  n.Pe = s[12];
  n.Tt = s[13];
  n.Nf = (int)s[14];
}
static void npt_to(const NPTState &n, double *s) {
  s[0] = n.xi;
  s[1] = n.Q;
  for (int i = 0; i < 9; i++) s[2 + i] = n.Vg[i];
  s[11] = n.W;
  s[12] = n.Pe;
  s[13] = n.Tt;
  s[14] = n.Nf;
}

extern "C" {
double refc_rmcut() { return RMCUT; }
int refc_max_neigh() { return MAX_NEIGH; }

void refc_generate_c60(double *coords, double *props) {
  C60Data d = generate_c60();
  for (int i = 0; i < 180; i++) coords[i] = d.coords[i];
  props[0] = d.I0;
  props[1] = d.Mmol;
  props[2] = d.Rmol;
}

int refc_make_fcc(double a, int nc, double *pos, double *h) { return make_fcc(a, nc, pos, h); }

void refc_mat_inv9(const double *h, double *hi) { mat_inv9(h, hi); }

void refc_apply_pbc(double *pos, const double *h, const double *hi, int n) {
  apply_pbc(pos, h, hi, n);
}

void refc_nlist(const double *pos, const double *h, const double *hi, int n, double rmcut,
                int *nl_count, int *nl_list) {
  nlist_build_sym(pos, h, hi, n, rmcut, nl_count, nl_list);
}

double refc_forces(double *Fv, double *Tv, double *Wm9, const double *pos, const double *qv,
                   const double *body, const double *h, const double *hi,
                   const int *nl_count, const int *nl_list, int n, int natom, double rmcut2,
                   double *lab) {
  return forces(Fv, Tv, Wm9, pos, qv, body, h, hi, nl_count, nl_list, n, natom, rmcut2, lab);
}

void refc_make_npt(double T, double Pe, int n, double *state) { npt_to(make_npt(T, Pe, n), state); }

void refc_step_npt(double *pos, double *vel, double *qv, double *omg, double *Fv, double *Tv,
                   double *Wm9, double *h, double *hi, const double *body, double I0,
                   double Mmol, int n, int natom, double rmcut2, double dt, double *state,
                   const int *nl_count, const int *nl_list, double *lab, double *out) {
  NPTState npt;
  npt_from(state, npt);
  auto r = step_npt(pos, vel, qv, omg, Fv, Tv, Wm9, h, hi, body, I0, Mmol, n, natom, rmcut2, dt,
                    npt, nl_count, nl_list, lab);
  npt_to(npt, state);
  out[0] = r.first;
  out[1] = r.second;
}

// Initial state exactly as the core's main (C1:644-682): FCC a0=14.17, then
// per molecule 3 velocity / 3 omega draws interleaved, 4 quaternion draws,
// normalisation, COM velocity removal.
int refc_init(int nc, unsigned seed, double T, double *pos, double *vel, double *omg,
              double *qv, double *h) {
  C60Data c60 = generate_c60();
  int N = make_fcc(14.17, nc, pos, h);
  std::mt19937 rng(seed);
  std::normal_distribution<double> gauss(0, 1);
  double sv = sqrt(kB * T * CONV / c60.Mmol), sw = sqrt(kB * T * CONV / c60.I0);
  for (int i = 0; i < N; i++) {
    for (int a = 0; a < 3; a++) {
      vel[i * 3 + a] = sv * gauss(rng);
      omg[i * 3 + a] = sw * gauss(rng);
    }
    for (int a = 0; a < 4; a++) qv[i * 4 + a] = gauss(rng);
    double nn = sqrt(qv[i * 4] * qv[i * 4] + qv[i * 4 + 1] * qv[i * 4 + 1] +
                     qv[i * 4 + 2] * qv[i * 4 + 2] + qv[i * 4 + 3] * qv[i * 4 + 3]);
    for (int a = 0; a < 4; a++) qv[i * 4 + a] /= nn;
  }
  double vcm[3] = {0, 0, 0};
  for (int i = 0; i < N; i++)
    for (int a = 0; a < 3; a++) vcm[a] += vel[i * 3 + a];
  for (int a = 0; a < 3; a++) vcm[a] /= N;
  for (int i = 0; i < N; i++)
    for (int a = 0; a < 3; a++) vel[i * 3 + a] -= vcm[a];
  return N;
}

// The main loop of the core program (C1:683-733) from the refc_init state of
// `seed` at nc, nsteps steps: list rebuilt every 25 steps, the reference's
// observables per step into series[4 * (g - 1) + {0, 1, 2, 3}] =
// {T, P, a, Ecoh} (C1:716-720).  Returns N.
int refc_main_loop(int nc, unsigned seed, int nsteps, double *series) {
  C60Data c60 = generate_c60();
  const int natom = C60_NATOM, Nmax = 4 * nc * nc * nc;
  std::vector<double> pos(Nmax * 3), vel(Nmax * 3), omg(Nmax * 3), qv(Nmax * 4),
      Fv(Nmax * 3), Tv(Nmax * 3), lab((size_t)Nmax * natom * 3);
  std::vector<int> nl_count(Nmax), nl_list((size_t)Nmax * MAX_NEIGH);
  double h[9] = {}, hi[9] = {}, Wm9[9] = {};
  const int N = refc_init(nc, seed, 300.0, pos.data(), vel.data(), omg.data(), qv.data(), h);
  mat_inv9(h, hi);
  NPTState npt = make_npt(300.0, 0.0, N);
  nlist_build_sym(pos.data(), h, hi, N, RMCUT, nl_count.data(), nl_list.data());
  apply_pbc(pos.data(), h, hi, N);
  forces(Fv.data(), Tv.data(), Wm9, pos.data(), qv.data(), c60.coords, h, hi, nl_count.data(),
         nl_list.data(), N, natom, RMCUT2, lab.data());
  for (int g = 1; g <= nsteps; g++) {
    if (g % 25 == 0) {
      mat_inv9(h, hi);
      nlist_build_sym(pos.data(), h, hi, N, RMCUT, nl_count.data(), nl_list.data());
    }
    auto [Ep, KE] = step_npt(pos.data(), vel.data(), qv.data(), omg.data(), Fv.data(), Tv.data(),
                             Wm9, h, hi, c60.coords, c60.I0, c60.Mmol, N, natom, RMCUT2, 1.0,
                             npt, nl_count.data(), nl_list.data(), lab.data());
    double kt = ke_trans(vel.data(), N, c60.Mmol);
    double V = fabs(mat_det9(h));
    double *o = series + 4 * (size_t)(g - 1);
    o[0] = inst_T(KE, npt.Nf);
    o[1] = inst_P(Wm9, kt, V);
    o[2] = h[0] / nc;
    o[3] = Ep / N;
  }
  return N;
}

// First n outputs of std::normal_distribution<double>(0,1) over mt19937(seed),
// pins the libstdc++ draw sequence the host driver relies on.
void refc_normal_stream(unsigned seed, int n, double *out) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> gauss(0, 1);
  for (int i = 0; i < n; i++) out[i] = gauss(rng);
}
}
