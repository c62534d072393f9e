// ref_harness_mm.cpp -- exposes the molecular-mechanics reference functions
// (src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp, "MM") through a C ABI.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile from the reference
// source where it lies (never copied), output to oracle/_ref/.
//
// The flat topology of the last refmm_build call is kept here; the force and
// step entry points use it, as the reference's main keeps one (MM:1106).
#define main fmd_ref_mm_main
#include "fuller_LJ_npt_mmmd_serial_omp_acc.cpp"
#undef main

static MolData g_mol;
static FlatTopology g_ft;
static bool g_have = false;

static void free_ft() {
  if (!g_have) return;
  delete[] g_ft.b_i0; delete[] g_ft.b_i1; delete[] g_ft.b_kb; delete[] g_ft.b_r0;
  delete[] g_ft.ang_i0; delete[] g_ft.ang_i1; delete[] g_ft.ang_i2;
  delete[] g_ft.ang_kth; delete[] g_ft.ang_th0;
  delete[] g_ft.dih_i0; delete[] g_ft.dih_i1; delete[] g_ft.dih_i2; delete[] g_ft.dih_i3;
  delete[] g_ft.dih_Vn; delete[] g_ft.dih_mult; delete[] g_ft.dih_gamma;
  delete[] g_ft.imp_i0; delete[] g_ft.imp_i1; delete[] g_ft.imp_i2; delete[] g_ft.imp_i3;
  delete[] g_ft.imp_ki; delete[] g_ft.imp_gamma;
  delete[] g_ft.mol_id;
  g_have = false;
}

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
int refmm_max_lj() { return MAX_LJ_NEIGH; }

// coords natom x 3 (centred), bonds nbond x 2 (0-based, i < j, sorted);
// props = {natom, nbond, Rmol, Dmol}.
int refmm_load_cc1(const char *path, double *coords, int *bonds, int cap_bonds,
                   double *props) {
  MolData m = load_cc1(path);
  for (int i = 0; i < m.natom; i++)
    for (int a = 0; a < 3; a++) coords[i * 3 + a] = m.coords[i][a];
  int nb = (int)m.bonds.size();
  for (int b = 0; b < nb && b < cap_bonds; b++) {
    bonds[2 * b] = m.bonds[b].first;
    bonds[2 * b + 1] = m.bonds[b].second;
  }
  props[0] = m.natom;
  props[1] = nb;
  props[2] = m.Rmol;
  props[3] = m.Dmol;
  return m.natom;
}

// Builds (and keeps) the flat topology of Nmol copies of the .cc1 molecule;
// k4 = {kb, kth, v2, kimp} in eV units; counts = {Nb, Nang, Ndih, Nimp}.
void refmm_build(const char *path, int Nmol, const double *k4, int *counts) {
  free_ft();
  g_mol = load_cc1(path);
  g_ft = build_flat_topology(g_mol, Nmol, k4[0], k4[1], k4[2], k4[3]);
  g_have = true;
  counts[0] = g_ft.Nb;
  counts[1] = g_ft.Nang;
  counts[2] = g_ft.Ndih;
  counts[3] = g_ft.Nimp;
}

// The kept topology: b (Nb x 2), bp (Nb x {kb, r0}), a (Nang x 3),
// ap (Nang x {kth, th0}), d (Ndih x 4), dp (Ndih x {Vn, gamma}), dm (Ndih),
// im (Nimp x 4), ip (Nimp x {ki, gamma}), mol_id (Nmol * natom).
void refmm_get(int *b, double *bp, int *a, double *ap, int *d, double *dp, int *dm, int *im,
               double *ip, int *mol_id, int natot) {
  for (int t = 0; t < g_ft.Nb; t++) {
    b[2 * t] = g_ft.b_i0[t];
    b[2 * t + 1] = g_ft.b_i1[t];
    bp[2 * t] = g_ft.b_kb[t];
    bp[2 * t + 1] = g_ft.b_r0[t];
  }
  for (int t = 0; t < g_ft.Nang; t++) {
    a[3 * t] = g_ft.ang_i0[t];
    a[3 * t + 1] = g_ft.ang_i1[t];
    a[3 * t + 2] = g_ft.ang_i2[t];
    ap[2 * t] = g_ft.ang_kth[t];
    ap[2 * t + 1] = g_ft.ang_th0[t];
  }
  for (int t = 0; t < g_ft.Ndih; t++) {
    d[4 * t] = g_ft.dih_i0[t];
    d[4 * t + 1] = g_ft.dih_i1[t];
    d[4 * t + 2] = g_ft.dih_i2[t];
    d[4 * t + 3] = g_ft.dih_i3[t];
    dp[2 * t] = g_ft.dih_Vn[t];
    dp[2 * t + 1] = g_ft.dih_gamma[t];
    dm[t] = g_ft.dih_mult[t];
  }
  for (int t = 0; t < g_ft.Nimp; t++) {
    im[4 * t] = g_ft.imp_i0[t];
    im[4 * t + 1] = g_ft.imp_i1[t];
    im[4 * t + 2] = g_ft.imp_i2[t];
    im[4 * t + 3] = g_ft.imp_i3[t];
    ip[2 * t] = g_ft.imp_ki[t];
    ip[2 * t + 1] = g_ft.imp_gamma[t];
  }
  for (int i = 0; i < natot; i++) mol_id[i] = g_ft.mol_id[i];
}

void refmm_nlist_lj(const double *pos, const double *h, const double *hi, int na, int *nlc,
                    int *nll) {
  build_nlist_lj(pos, h, hi, na, g_ft.mol_id, nlc, nll);
}

// out7 = {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot} over the kept topology.
void refmm_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                          const double *hi, const int *nlc, const int *nll,
                          const double *charge, int has_charge, int na, double *out7) {
  ForceResult r = compute_forces(
      F, vir9, pos, h, hi, g_ft.Nb, g_ft.b_i0, g_ft.b_i1, g_ft.b_kb, g_ft.b_r0, g_ft.Nang,
      g_ft.ang_i0, g_ft.ang_i1, g_ft.ang_i2, g_ft.ang_kth, g_ft.ang_th0, g_ft.Ndih,
      g_ft.dih_i0, g_ft.dih_i1, g_ft.dih_i2, g_ft.dih_i3, g_ft.dih_Vn, g_ft.dih_mult,
      g_ft.dih_gamma, g_ft.Nimp, g_ft.imp_i0, g_ft.imp_i1, g_ft.imp_i2, g_ft.imp_i3,
      g_ft.imp_ki, g_ft.imp_gamma, nlc, nll, charge, g_ft.mol_id, has_charge != 0, na);
  out7[0] = r.Eb;
  out7[1] = r.Ea;
  out7[2] = r.Ed;
  out7[3] = r.Ei;
  out7[4] = r.Elj;
  out7[5] = r.Ecoul;
  out7[6] = r.Etot;
}

void refmm_make_npt(double T, double Pe, int na, double *state) {
  npt_to(make_npt(T, Pe, na), state);
}

// out8 = {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot, KE}.
void refmm_step_npt(double *pos, double *vel, double *F, double *vir9, double *h, double *hi,
                    int na, double dt, double *state, const int *nlc, const int *nll,
                    const double *charge, int has_charge, double *out8) {
  NPTState npt;
  npt_from(state, npt);
  auto r = step_npt(pos, vel, F, vir9, h, hi, na, dt, npt, g_ft, nlc, nll, charge,
                    g_ft.mol_id, has_charge != 0);
  npt_to(npt, state);
  out8[0] = r.first.Eb;
  out8[1] = r.first.Ea;
  out8[2] = r.first.Ed;
  out8[3] = r.first.Ei;
  out8[4] = r.first.Elj;
  out8[5] = r.first.Ecoul;
  out8[6] = r.first.Etot;
  out8[7] = r.second;
}

void refmm_apply_pbc(double *pos, const double *h, const double *hi, int n) {
  apply_pbc(pos, h, hi, n);
}
void refmm_mat_inv9(const double *h, double *hi) { mat_inv9(h, hi); }
}
