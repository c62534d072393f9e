/*
 * fmd_oracle.h -- FP64 CPU restatement of the fuller_md force / integrator
 * path.  TEST INFRASTRUCTURE ONLY (see fmd_oracle.c).
 */
#ifndef FMD_ORACLE_H
#define FMD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double sig2, eps, rc2, vshift;
} orc_lj_params;

/* NPTState (C1:478, AB:511). */
typedef struct {
  double xi, Q, Vg[9], W, Pe, Tt;
  int Nf;
} orc_npt;

enum { ORC_REBO_COMPAT = 0, ORC_REBO_PHYSICAL = 1 };

double orc_mat_det9(const double *h);
void orc_mat_inv9(const double *h, double *hi);
void orc_mimg(double *d, const double *hi, const double *h);
void orc_apply_pbc(double *pos, const double *h, const double *hi, int n);
void orc_q2R(const double *q, double *R);

void orc_lj_params_rigid(orc_lj_params *p);
void orc_lj_params_airebo(orc_lj_params *p);

void orc_rigid_nlist(const double *pos, const double *h, const double *hi, int n,
                     double rmcut, int max_neigh, int *nl_count, int *nl_list);
void orc_rigid_lab(const double *qv, const double *body, int n, int natom, double *lab);
double orc_rigid_forces(double *Fv, double *Tv, double *Wm9, const double *pos,
                        const double *qv, const double *body, const double *h,
                        const double *hi, const int *nl_count, const int *nl_list,
                        int max_neigh, int n, int natom, double rmcut2,
                        const orc_lj_params *p, double *lab);
double orc_ke_trans(const double *vel, int n, double mmol);
double orc_ke_rot(const double *omg, int n, double i0);
double orc_inst_T(double ke, int nf);
double orc_inst_P(const double *W, double ket, double V);
void orc_make_npt(orc_npt *s, double T, double Pe, int nf);
void orc_rigid_step_npt(double *pos, double *vel, double *qv, double *omg, double *Fv,
                        double *Tv, double *Wm9, double *h, double *hi,
                        const double *body, double I0, double Mmol, int n, int natom,
                        double rmcut2, double dt, orc_npt *npt, const int *nl_count,
                        const int *nl_list, int max_neigh, const orc_lj_params *p,
                        double *lab, double *out);

double orc_rebo_fc(double r);
double orc_rebo_dfc(double r);
double orc_rebo_vr(double r);
double orc_rebo_dvr(double r);
double orc_rebo_va(double r);
double orc_rebo_dva(double r);
double orc_rebo_G(double x);
double orc_rebo_dG(double x);

void orc_build_nlist_rebo(const double *pos, const double *h, const double *hi, int na,
                          const int *mol_id, double rcut, int max_neigh, int *nlc,
                          int *nll);
void orc_build_nlist_lj(const double *pos, const double *h, const double *hi, int na,
                        const int *mol_id, double rlist, int max_neigh, int *nlc,
                        int *nll);
double orc_compute_rebo(double *F, double *vir9, const double *pos, const double *h,
                        const double *hi, const int *nlc, const int *nll, int max_neigh,
                        int na, int mode);
double orc_compute_lj(double *F, double *vir9, const double *pos, const double *h,
                      const double *hi, const int *nlc, const int *nll, int max_neigh,
                      int na, const orc_lj_params *p);
/* Exact-gradient REBO with the REBO-II P_ij correction (pcc[5]: P_CC at
 * N = 0..4) and the AIREBO torsion term (well depth eps_t); no reference
 * counterpart.  Adds into F / vir9; returns the total energy, the torsion
 * part in *e_tors. */
double orc_rebo_ext(double *F, double *vir9, const double *pos, const double *h,
                    const double *hi, const int *nlc, const int *nll, int max_neigh, int na,
                    const double *pcc, double eps_t, double *e_tors);
void orc_airebo_forces(double *F, double *vir9, const double *pos, const double *h,
                       const double *hi, const int *nlc_r, const int *nll_r, int max_r,
                       const int *nlc_l, const int *nll_l, int max_l, int na,
                       int rebo_mode, const orc_lj_params *p, double *out);
double orc_ke_total(const double *vel, const double *mass, int na);
void orc_airebo_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                         double *hi, const double *mass, int na, double dt,
                         orc_npt *npt, const int *nlc_r, const int *nll_r, int max_r,
                         const int *nlc_l, const int *nll_l, int max_l, int rebo_mode,
                         const orc_lj_params *p, double *out);

/* Molecular mechanics (MM).  Flat topology arrays of build_flat_topology
 * (MM:228-239) gathered in one struct. */
typedef struct {
  int Nb, Nang, Ndih, Nimp;
  const int *b;     /* Nb x 2 */
  const double *bp; /* Nb x {kb, r0} */
  const int *a;     /* Nang x 3 (i, centre j, k) */
  const double *ap; /* Nang x {kth, th0} */
  const int *d;     /* Ndih x 4 */
  const double *dp; /* Ndih x {Vn, gamma} */
  const int *dm;    /* Ndih multiplicities */
  const int *im;    /* Nimp x 4 (centre first) */
  const double *ip; /* Nimp x {ki, gamma} */
} orc_mm_topo;

enum { ORC_MM_COMPAT = 0, ORC_MM_PHYSICAL = 1 };

void orc_lj_params_mm(orc_lj_params *p);
void orc_mm_topology_counts(int natom, const int *bonds, int nbond, int *counts4);
void orc_mm_topology(const double *coords, int natom, const int *bonds, int nbond, int nmol,
                     const double *k4, int *b, double *bp, int *a, double *ap, int *d,
                     double *dp, int *dm, int *im, double *ip, int *mol_id);
void orc_mm_forces(double *F, double *vir9, const double *pos, const double *h,
                   const double *hi, const orc_mm_topo *t, const int *nlc, const int *nll,
                   int max_l, const double *charge, const int *mol_id, int has_charge, int na,
                   int mode, double *out7);
void orc_mm_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                     double *hi, int na, double dt, orc_npt *npt, const orc_mm_topo *t,
                     const int *nlc, const int *nll, int max_l, const double *charge,
                     const int *mol_id, int has_charge, int mode, double *out8);

#ifdef __cplusplus
}
#endif
#endif
