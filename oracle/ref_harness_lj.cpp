// ref_harness_lj.cpp -- exposes the rigid-LJ "full" reference functions
// (src/fuller_LJ_npt_md_serial_omp_acc.cpp, "LJ") through a C ABI.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// source where it lies (never copied), output to oracle/_ref/.
#define main fmd_ref_lj_main
#include "fuller_LJ_npt_md_serial_omp_acc.cpp"
#undef main

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
int reflj_max_neigh() { return MAX_NEIGH; }

// props = {natom, Rmol, Dmol, I0, Mmol}
int reflj_load_cc1(const char *path, double *coords, double *props) {
  MolData m = load_cc1(path);
  for (int i = 0; i < m.natom * 3; i++) coords[i] = m.coords[i];
  props[0] = m.natom;
  props[1] = m.Rmol;
  props[2] = m.Dmol;
  props[3] = m.I0;
  props[4] = m.Mmol;
  return m.natom;
}

double reflj_default_a0(double dmax, const char *st, double s) { return default_a0(dmax, st, s); }

int reflj_make_crystal(const char *st, double a, int nc, double *pos, double *h) {
  std::string s = st;
  if (s == "FCC") return make_fcc(a, nc, pos, h);
  if (s == "HCP") return make_hcp(a, nc, pos, h);
  return make_bcc(a, nc, pos, h);
}

void reflj_apply_pbc(double *pos, const double *h, const double *hi, int n) {
  apply_pbc(pos, h, hi, n);
}

void reflj_nlist(const double *pos, const double *h, const double *hi, int n, double rmcut,
                 int *nl_count, int *nl_list) {
  nlist_build_sym(pos, h, hi, n, rmcut, nl_count, nl_list);
}

double reflj_forces(double *Fv, double *Tv, double *Wm9, const double *pos, const double *qv,
                    const double *body, const double *h, const double *hi,
                    const int *nl_count, const int *nl_list, int n, int natom,
                    double rmcut2, double *lab) {
  return forces(Fv, Tv, Wm9, pos, qv, body, h, hi, nl_count, nl_list, n, natom, rmcut2, lab);
}

void reflj_step_npt(double *pos, double *vel, double *qv, double *omg, double *Fv, double *Tv,
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

void reflj_make_npt(double T, double Pe, int n, double *state) { npt_to(make_npt(T, Pe, n), state); }

void reflj_restart_filename(const char *oname, int istep, int nsteps, char *out, int cap) {
  std::string r = restart_filename(oname, istep, nsteps);
  snprintf(out, cap, "%s", r.c_str());
}
}
