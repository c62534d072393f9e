// ref_harness_airebo.cpp -- exposes the AIREBO reference functions
// (src/fuller_airebo_npt_md_serial_omp_acc.cpp, "AB") through a C ABI.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// source where it lies (never copied), output to oracle/_ref/.
#define main fmd_ref_airebo_main
#include "fuller_airebo_npt_md_serial_omp_acc.cpp"
#undef main

static void npt_from(const double *s, NPTState &n) {
  n.xi = s[0];
  n.Q = s[1];
  for (int i = 0; i < 9; i++) n.Vg[i] = s[2 + i];
  n.W_ = s[11];
  n.Pe = s[12];
  n.Tt = s[13];
  n.Nf = (int)s[14];
}
static void npt_to(const NPTState &n, double *s) {
  s[0] = n.xi;
  s[1] = n.Q;
  for (int i = 0; i < 9; i++) s[2 + i] = n.Vg[i];
  s[11] = n.W_;
  s[12] = n.Pe;
  s[13] = n.Tt;
  s[14] = n.Nf;
}

extern "C" {
int refab_max_rebo() { return MAX_REBO_NEIGH; }
int refab_max_lj() { return MAX_LJ_NEIGH; }

// props = {natom, Rmol, Dmol}
int refab_load_cc1(const char *path, double *coords, double *props) {
  MolData m = load_cc1(path);
  for (int i = 0; i < m.natom; i++)
    for (int a = 0; a < 3; a++) coords[i * 3 + a] = m.coords[i][a];
  props[0] = m.natom;
  props[1] = m.Rmol;
  props[2] = m.Dmol;
  return m.natom;
}

void refab_nlist_rebo(const double *pos, const double *h, const double *hi, int na,
                      const int *mol_id, int *nlc, int *nll) {
  build_nlist_rebo(pos, h, hi, na, mol_id, nlc, nll);
}
void refab_nlist_lj(const double *pos, const double *h, const double *hi, int na,
                    const int *mol_id, int *nlc, int *nll) {
  build_nlist_lj(pos, h, hi, na, mol_id, nlc, nll);
}

double refab_compute_rebo(double *F, double *vir9, const double *pos, const double *h,
                          const double *hi, const int *nlc, const int *nll, int na) {
  return compute_rebo(F, vir9, pos, h, hi, nlc, nll, na);
}
double refab_compute_lj(double *F, double *vir9, const double *pos, const double *h,
                        const double *hi, const int *nlc, const int *nll, int na) {
  return compute_lj(F, vir9, pos, h, hi, nlc, nll, na);
}
void refab_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                          const double *hi, const int *nlc_r, const int *nll_r,
                          const int *nlc_l, const int *nll_l, int na, double *out) {
  ForceResult r = compute_forces(F, vir9, pos, h, hi, nlc_r, nll_r, nlc_l, nll_l, na);
  out[0] = r.Ep;
  out[1] = r.Ep_rebo;
  out[2] = r.Ep_lj;
}
void refab_make_npt(double T, double Pe, int na, double *state) { npt_to(make_npt(T, Pe, na), state); }
void refab_step_npt(double *pos, double *vel, double *F, double *vir9, double *h, double *hi,
                    const double *mass, int na, double dt, double *state, const int *nlc_r,
                    const int *nll_r, const int *nlc_l, const int *nll_l, double *out) {
  NPTState npt;
  npt_from(state, npt);
  StepResult r = step_npt(pos, vel, F, vir9, h, hi, mass, na, dt, npt, nlc_r, nll_r, nlc_l, nll_l);
  npt_to(npt, state);
  out[0] = r.Ep;
  out[1] = r.Ep_rebo;
  out[2] = r.Ep_lj;
  out[3] = r.KE;
}
void refab_apply_pbc(double *pos, const double *h, const double *hi, int n) {
  apply_pbc(pos, h, hi, n);
}
void refab_mat_inv9(const double *h, double *hi) { mat_inv9(h, hi); }
}
