/*
 * fmd.h -- C ABI of the MI355X fuller_md force engine (libfmd.so).
 *
 * Two layers:
 *
 *  1. Drop-in entry points with the arguments of the reference's force,
 *     list and integrator functions (host pointers, reference layouts).  Each
 *     call copies its inputs to the GPU, runs the gfx950 kernels and copies
 *     the outputs back, so a reference driver can swap one call at a time.
 *     Reference files are cited as
 *       C1 = src/fuller_LJ_npt_md_core_serial.cpp
 *       LJ = src/fuller_LJ_npt_md_serial_omp_acc.cpp
 *       AB = src/fuller_airebo_npt_md_serial_omp_acc.cpp
 *
 *  2. A resident engine (fmd_engine_*) that keeps the whole NPT state in HBM
 *     and runs many steps per call: the fast path used by the C++ drivers and
 *     the benchmark.
 *
 * Layout conventions (all as in the reference): flat AoS x,y,z per particle;
 * quaternions [w,x,y,z]; the cell matrix h is row-major with the cell vectors
 * as its columns (r = h s); virial W[3a+b] = sum d_a f_b; units A, amu, eV,
 * fs, K, GPa.
 *
 * Errors: every function returns FMD_OK (0) or a negative status;
 * fmd_last_error() gives a message (thread-local).  Kernels never fail
 * silently: list overflow, which the reference ignores, is reported through
 * fmd_engine_counters().
 *
 * Threading: one engine per host thread; the drop-in entry points share one
 * per-process device workspace and must be called from one thread at a time.
 */
#ifndef FMD_H
#define FMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define FMD_OK 0
#define FMD_ERR_ARG -1
#define FMD_ERR_HIP -2
#define FMD_ERR_CAPACITY -3
#define FMD_ERR_STATE -4
#define FMD_ERR_NODEVICE -5

/* Message of the last failed call on this thread ("" if none). */
const char *fmd_last_error(void);
/* ABI version (major * 100 + minor). */
int fmd_version(void);
/* Name and compute-unit count of the current HIP device. */
int fmd_device_info(char *name, int cap, int *cu_count, double *hbm_bytes);

/* Shifted 12-6 Lennard-Jones: V = 4 eps ((s/r)^12 - (s/r)^6) - vshift for
 * r < rcut (rigid, C1:403) or r <= rcut (AIREBO, AB:441); r^2 is clamped at
 * rmin2 (0.25 in the reference, C1:404). */
typedef struct {
  double sigma, eps, rcut, vshift, rmin2;
} fmd_lj_params;

/* Reference constants: rigid sigma=3.431 eps=2.635e-3 rc=3 sigma (C1:93-104);
 * AIREBO sigma=3.40 eps=2.84e-3 rc=3 sigma (AB:119-124). */
void fmd_lj_params_rigid(fmd_lj_params *p);
void fmd_lj_params_airebo(fmd_lj_params *p);
/* Molecular mechanics sigma=3.431 eps=2.635e-3 rc=3 sigma (MM:111-117). */
void fmd_lj_params_mm(fmd_lj_params *p);

/* NPTState of the reference (C1:478 / AB:511). */
typedef struct {
  double xi, Q, Vg[9], W, Pe, Tt;
  int Nf;
} fmd_npt_state;

/* make_npt (C1:480-488; AB:512-515): Q = Nf kB T 1e4, W = (Nf+9) kB T 1e6. */
void fmd_make_npt(fmd_npt_state *s, double T, double Pe, int Nf);

/* ------------------------------------------------- system construction */

/* load_cc1 (LJ:178-207, AB:195-210): coords (>= 84 x 3) centred on the
 * geometric centre; props = {Rmol, Dmol, I0, Mmol}. */
int fmd_load_cc1(const char *path, double *coords, int *natom, double *props);
/* generate_c60 (C1:228-252): the core program's analytic C60 (truncated
 * icosahedron, edge 1.44 A, radius 3.568 A), coords 60 x 3 in its vertex
 * order; props = {Rmol, Dmol, I0, Mmol} as fmd_load_cc1 gives them. */
int fmd_generate_c60(double *coords, double *props);
/* resolve_fullerene (LJ:210-226): species spec -> .cc1 path and label. */
int fmd_resolve_fullerene(const char *spec, const char *libdir, char *path, int pcap,
                          char *label, int lcap);
/* default_a0 (LJ:251-254). */
double fmd_default_a0(double dmax, const char *crystal, double scale);
/* make_fcc / make_hcp / make_bcc (LJ:229-250); returns the molecule count
 * (or a negative status).  pos holds 4 nc^3 (FCC) or 2 nc^3 entries. */
int fmd_make_crystal(const char *crystal, double a, int nc, double *pos, double *h);
/* Seeded initial state of LJ:790-801 (std::mt19937 + normal_distribution). */
int fmd_rigid_init_state(int n, unsigned seed, double T, double Mmol, double I0, double *vel,
                         double *omg, double *qv);
/* Seeded atomic velocities of AB:808-815. */
int fmd_atomic_init_velocities(int na, unsigned seed, double T, const double *mass,
                               double *vel);
/* First n draws of normal_distribution<double>(0,1) over mt19937(seed). */
int fmd_normal_stream(unsigned seed, int n, double *out);

/* ----------------------------------------------------- rigid drop-ins */

/* Replaces nlist_build_sym (C1:280-301, LJ:257-268): symmetric list of
 * molecule pairs with minimum-image COM distance < rmcut + 3, ascending,
 * capped at max_neigh (reference MAX_NEIGH = 80, C1:121).  *overflow (may be
 * NULL) receives the number of molecules whose list was truncated. */
int fmd_rigid_nlist(const double *pos, const double *h, const double *hi, int N,
                    double rmcut, int max_neigh, int *nl_count, int *nl_list,
                    int *overflow);

/* Replaces forces (C1:346-351, LJ:271-274): site-site LJ between rigid
 * molecules.  Writes Fv, Tv (N x 3), Wm9 (9) and *Ep.  lab (N*natom*3, may be
 * NULL) receives the lab-frame site offsets as in the reference.  lj may be
 * NULL for the rigid reference constants.  natom <= 128. */
int fmd_rigid_forces(double *Fv, double *Tv, double *Wm9, const double *pos,
                     const double *qv, const double *body, const double *h,
                     const double *hi, const int *nl_count, const int *nl_list,
                     int max_neigh, int N, int natom, double rmcut2,
                     const fmd_lj_params *lj, double *lab, double *Ep);

/* No counterpart in the reference: fmd_rigid_forces for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * fmd_rigid_forces, but every device buffer lives in one fixed 16 MiB
 * per-device scratch array of the library, set up on the first call on each
 * device: no hipMalloc, hipFree or other allocating call.  Work runs on the
 * null stream.  Inputs needing more than the scratch (about
 * N * (40 natom_pad + 4 max_neigh + 190) bytes) return FMD_ERR_CAPACITY.
 * Calls on one device share that scratch, so calls that overlap on one GPU
 * (from several threads) race. */
int fmd_rigid_forces_noalloc(double *Fv, double *Tv, double *Wm9, const double *pos,
                             const double *qv, const double *body, const double *h,
                             const double *hi, const int *nl_count, const int *nl_list,
                             int max_neigh, int N, int natom, double rmcut2,
                             const fmd_lj_params *lj, double *lab, double *Ep);

/* Replaces step_npt (C1:504-510, LJ:434-439): one NH + PR velocity-Verlet
 * step of rigid molecules (clamps xi +-0.1, Vg +-0.01).  Updates pos, vel,
 * qv, omg, Fv, Tv, Wm9, h, hi and *npt in place; *Ep, *KE as the reference
 * returns them. */
int fmd_rigid_step_npt(double *pos, double *vel, double *qv, double *omg, double *Fv,
                       double *Tv, double *Wm9, double *h, double *hi,
                       const double *body, double I0, double Mmol, int N, int natom,
                       double rmcut2, double dt, fmd_npt_state *npt,
                       const int *nl_count, const int *nl_list, int max_neigh,
                       const fmd_lj_params *lj, double *Ep, double *KE);

/* ---------------------------------------------------- AIREBO drop-ins */

#define FMD_REBO_COMPAT 0   /* the reference's REBO forces and pair-only virial */
#define FMD_REBO_PHYSICAL 1 /* exact gradient of the same energy, full virial */

/* Replaces build_nlist_rebo (AB:251-261): symmetric same-molecule list within
 * 2.3 A, ascending, stride 12 (MAX_REBO_NEIGH, AB:126). */
int fmd_airebo_nlist_rebo(const double *pos, const double *h, const double *hi, int Na,
                          const int *mol_id, int *nlc, int *nll);

/* Replaces build_nlist_lj (AB:263-272): half list (j > i) of other-molecule
 * atoms within rc + 2 A, ascending, the first 400 kept (MAX_LJ_NEIGH, AB:127),
 * stride 400.  Atoms of one molecule must be contiguous (as the reference
 * builds them, AB:781-786). */
int fmd_airebo_nlist_lj(const double *pos, const double *h, const double *hi, int Na,
                        const int *mol_id, int *nlc, int *nll);

/* Replaces compute_forces (AB:477-495): REBO + intermolecular LJ forces F
 * (Na x 3), virial vir9 and out3 = {Ep, Ep_rebo, Ep_lj}.  The REBO list is
 * used as given; the LJ half list is used through its truncation points (the
 * last kept j of every full 400-entry list), its pairs being regenerated on
 * the GPU, which is the same pair set when the list was built from these
 * positions (or within its 2 A skin).  mol_id (not an argument of the
 * reference function) gives the molecule of each atom. */
int fmd_airebo_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                              const double *hi, const int *nlc_r, const int *nll_r,
                              const int *nlc_l, const int *nll_l, int Na, const int *mol_id,
                              int rebo_mode, double *out3);

/* Replaces step_npt (AB:518-590): one atomic NH + PR step (clamps xi +-0.05,
 * Vg +-0.005); updates pos, vel, F, vir9, h, hi, *npt in place;
 * out4 = {Ep, Ep_rebo, Ep_lj, KE}. */
int fmd_airebo_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                        double *hi, const double *mass, int Na, double dt,
                        fmd_npt_state *npt, const int *nlc_r, const int *nll_r,
                        const int *nlc_l, const int *nll_l, const int *mol_id,
                        int rebo_mode, double *out4);


/* --------------------------------------------- molecular mechanics (MM)
 * The bonded force field of src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp:
 * harmonic bonds and angles, cosine dihedrals and impropers within each
 * molecule, shifted LJ between molecules over the reference's 400-entry half
 * list, and a cut Coulomb term between charged atoms of different molecules.
 *
 * FMD_MM_COMPAT gives the reference's forces term by term.  They are not the
 * gradient of its energy: bond and angle forces are +dE/dr (MM:391,
 * 468) and the two middle atoms of each torsion take the dot-product
 * coefficients with the opposite sign (MM:528-529, 653, 672), so the
 * reference trajectories heat up without bound.  The virial is the reference's: the
 * diagonal of the bond and LJ parts.  FMD_MM_PHYSICAL: the exact gradient of
 * the same energy and the full virial of every term. */
#define FMD_MM_COMPAT 0
#define FMD_MM_PHYSICAL 1

/* FlatTopology (MM:226-237): the same members in the same order, so the
 * reference's struct converts by pointer.  Terms list global atom indices. */
typedef struct {
  int Nb;
  int *b_i0, *b_i1;
  double *b_kb, *b_r0;
  int Nang;
  int *ang_i0, *ang_i1, *ang_i2;
  double *ang_kth, *ang_th0;
  int Ndih;
  int *dih_i0, *dih_i1, *dih_i2, *dih_i3;
  double *dih_Vn;
  int *dih_mult;
  double *dih_gamma;
  int Nimp;
  int *imp_i0, *imp_i1, *imp_i2, *imp_i3;
  double *imp_ki, *imp_gamma;
  int *mol_id;
} fmd_mm_topology;

/* The bonds of a .cc1 file (load_cc1 MM:160-183): each unique pair of its
 * adjacency lists once, 0-based (i < j), in ascending order; bonds holds
 * cap pairs. */
int fmd_load_cc1_bonds(const char *path, int *bonds, int cap, int *nbond);
/* Terms per molecule {bonds, angles, dihedrals, impropers} that
 * build_flat_topology derives from a bond list. */
int fmd_mm_topology_counts(int natom, const int *bonds, int nbond, int *counts4);
/* Replaces build_flat_topology (MM:254-334): fills the caller's arrays of ft
 * (counts4 x Nmol terms, mol_id of Nmol * natom atoms) for Nmol copies of the
 * molecule; force constants in eV units (the reference's kcal/mol / 23.06).
 * Sets the counts of ft. */
int fmd_mm_build_topology(const double *coords, int natom, const int *bonds, int nbond,
                          int Nmol, double kb, double kth, double v2_dih, double k_imp,
                          fmd_mm_topology *ft);

/* Replaces build_nlist_lj (MM:337-349): half list (j > i) of other-molecule
 * atoms within rc + 2 A, ascending, first 400 kept, stride 400.  Atoms of a
 * molecule contiguous. */
int fmd_mm_nlist_lj(const double *pos, const double *h, const double *hi, int Na,
                    const int *mol_id, int *nlc, int *nll);

/* Replaces compute_forces (MM:351-788): F (Na x 3), vir9 and
 * out7 = {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot}.  Every bonded term must lie
 * within one molecule (as build_flat_topology makes them); molecules with
 * identical terms share one device copy.  The LJ half list is used through
 * its truncation points, as in fmd_airebo_compute_forces; charge is read
 * only when has_charge. */
int fmd_mm_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                          const double *hi, const fmd_mm_topology *ft, const int *nlc,
                          const int *nll, const double *charge, const int *mol_id,
                          int has_charge, int Na, int mm_mode, double *out7);

/* Replaces step_npt (MM:812-890): one NH + PR step with carbon masses
 * (clamps xi +-0.05, Vg +-0.005); updates pos, vel, F, vir9, h, hi, *npt;
 * out8 = {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot, KE}. */
int fmd_mm_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                    double *hi, int Na, double dt, fmd_npt_state *npt,
                    const fmd_mm_topology *ft, const int *nlc, const int *nll,
                    const double *charge, const int *mol_id, int has_charge, int mm_mode,
                    double *out8);

/* ------------------------------------------------------- resident engine */

typedef struct fmd_engine fmd_engine;

#define FMD_KIND_RIGID 1
#define FMD_KIND_ATOMIC 2

typedef struct {
  int n_mol;            /* molecules */
  int natom;            /* sites per molecule (<= 128) */
  const double *body;   /* natom x 3 body-frame offsets from the COM (A) */
  double I0, Mmol;      /* isotropic inertia (amu A^2), molecular mass (amu) */
  fmd_lj_params lj;
  double rmcut;         /* molecule-pair cutoff (reference rc + 2 Rmol + 1) */
  double skin;          /* list radius = rmcut + skin (reference 3.0) */
  int max_neigh;        /* reference 80 */
  int nlup;             /* list rebuild interval in global steps (25) */
  double dt;            /* fs */
  double clamp_xi, clamp_vg; /* 0.1, 0.01 (C1:520, 525) */
  int half_pairs;       /* 1 (default): time steps evaluate each molecule pair once
                           (Newton's third law) and keep only the virial trace, the
                           part the pressure uses; with at most 64 sites one
                           wavefront per pair, 2: the same per molecule (its owned
                           pairs in one workgroup); 0: every pair from both sides
                           and the full tensor, as the reference computes */
  int cell_min_molecules; /* molecule lists from cells of a grid for systems of at
                           least this many molecules (3000; 0 = never), below it a
                           scan of every molecule; same lists */
} fmd_rigid_config;

/* Defaults of the reference rigid programs for the given molecule radius. */
void fmd_rigid_config_default(fmd_rigid_config *c, double rmol);

int fmd_rigid_create(const fmd_rigid_config *c, fmd_engine **out);

/* State in reference layouts.  Any pointer may be NULL (left unchanged). */
int fmd_rigid_set_state(fmd_engine *e, const double *pos, const double *vel,
                        const double *qv, const double *omg, const double *h,
                        const fmd_npt_state *npt);
int fmd_rigid_get_state(fmd_engine *e, double *pos, double *vel, double *qv, double *omg,
                        double *h, fmd_npt_state *npt);

typedef struct {
  int n_atoms;
  const int *mol_id;     /* molecule of each atom; each molecule's atoms contiguous */
  const double *mass;    /* amu */
  fmd_lj_params lj;      /* intermolecular LJ */
  double lj_skin;        /* half-list radius = rcut + lj_skin (2.0, AB:265) */
  int lj_max_neigh;      /* 400 (AB:127): the reference's truncation; 0 = none */
  double rebo_rlist;     /* 2.3 (AB:116) */
  int rebo_max_neigh;    /* 12 (AB:126) */
  int rebo_mode;         /* FMD_REBO_COMPAT / FMD_REBO_PHYSICAL */
  int nlup;              /* list rebuild interval (20, AB:860) */
  double dt;             /* fs (0.5) */
  double clamp_xi, clamp_vg; /* 0.05, 0.005 (AB:524, 530) */
  double al_skin;        /* engines: the per-atom LJ lists hold the pairs within
                            rcut + al_skin (1.0; no counterpart in the reference,
                            whose LJ list is the half list above).  Results do
                            not depend on it beyond summation order: stale lists
                            are rebuilt within the step. */
  int cell_min_molecules; /* engines: molecule lists from cells of a grid for
                            replicas of at least this many molecules (2048),
                            below it a scan of every molecule; same lists */
  int rebo_lds_slots;    /* REBO kernel: bond entries per atom its LDS holds (4,
                            3..8); a molecule with more takes the global-memory
                            path, with the same forces */
  int rebo_threads;      /* REBO kernel: workgroup size, 192 (molecules of up to
                            192 atoms; else 256) or 256 */
  /* REBO-II / AIREBO terms the reference's REBO does not have (no reference
     counterpart; FMD_REBO_PHYSICAL only; all zero, the default, is the
     reference's potential):
       rebo_pcc[n]  P_CC(N, N_H = 0) at N = n = 0..4, the bond-order correction
                    b_ij = (1 + Gs_ij + P(N_ij))^-1/2, N_ij = sum over i's
                    other bonds of fc; a cubic Hermite with zero slope at the
                    knots (FMD_AIREBO_PCC2 at N = 2 for AIREBO carbon);
       torsion_eps  well depth of the AIREBO torsion term
                    w w w eps (256/405 cos^10(omega/2) - 1/10) over bond chains
                    (FMD_AIREBO_TORSION_EPS); its energy is part of Ep_rebo. */
  double rebo_pcc[5];
  double torsion_eps;
} fmd_atomic_config;

/* AIREBO carbon values of the two terms above (Stuart, Tutein & Harrison,
 * J. Chem. Phys. 112, 6472 (2000)). */
#define FMD_AIREBO_PCC2 (-0.0276030)
#define FMD_AIREBO_TORSION_EPS 0.3079

/* Defaults of the reference AIREBO program. */
void fmd_atomic_config_default(fmd_atomic_config *c);
int fmd_atomic_create(const fmd_atomic_config *c, fmd_engine **out);
int fmd_atomic_set_state(fmd_engine *e, const double *pos, const double *vel,
                         const double *h, const fmd_npt_state *npt);
int fmd_atomic_get_state(fmd_engine *e, double *pos, double *vel, double *h,
                         fmd_npt_state *npt);

/* Resident molecular-mechanics engine: an atomic engine (fmd_atomic_set_state,
 * fmd_atomic_get_state, fmd_engine_*) whose intramolecular force is the MM
 * bonded field.  Every atom has the carbon mass (MM:798, 827). */
typedef struct {
  int n_atoms;
  const int *mol_id;          /* each molecule's atoms contiguous */
  const fmd_mm_topology *ft;  /* terms over the n_atoms atoms */
  const double *charge;       /* NULL: no Coulomb term (the reference's case) */
  fmd_lj_params lj;           /* fmd_lj_params_mm */
  double lj_skin;             /* 2.0 (MM:339) */
  int lj_max_neigh;           /* 400 (MM:119); 0 = no truncation */
  int nlup;                   /* 20 (MM:1179) */
  double dt;                  /* 0.1 fs (MM:1023) */
  double clamp_xi, clamp_vg;  /* 0.05, 0.005 (MM:818, 824) */
  int mm_mode;                /* FMD_MM_COMPAT / FMD_MM_PHYSICAL */
} fmd_mm_config;

void fmd_mm_config_default(fmd_mm_config *c);
int fmd_mm_create(const fmd_mm_config *c, fmd_engine **out);
/* Energies of the engine's last force evaluation:
 * {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot}. */
int fmd_mm_energies(fmd_engine *e, double *out7);

/* ---------------------------------------------------- batched replicas
 * R independent AIREBO NPT systems (sizes, species, T and P may differ) in
 * one engine: every kernel launch of a step covers all replicas, and steps
 * are replayed from captured HIP graphs.  No reference counterpart (the
 * reference runs one system per process); replica r alone follows the
 * AIREBO program's step_npt (AB:518-590).  cfg[r] gives replica r's atoms;
 * potential, list, dt, nlup and clamp fields must equal cfg[0]'s.
 * fmd_engine_run then takes ctrl and thermo arrays of nsteps x R entries,
 * ctrl[k * R + r] / thermo[k * R + r] for step k of replica r. */
int fmd_batch_create(const fmd_atomic_config *cfg, int R, fmd_engine **out);
int fmd_batch_size(fmd_engine *e, int *R);
int fmd_batch_set_state(fmd_engine *e, int r, const double *pos, const double *vel,
                        const double *h, const fmd_npt_state *npt);
int fmd_batch_get_state(fmd_engine *e, int r, double *pos, double *vel, double *h,
                        fmd_npt_state *npt);
/* Replica r's forces (its atoms x 3), W9 and {Ep, Ep_rebo, Ep_lj}. */
int fmd_batch_forces(fmd_engine *e, int r, double *F, double *W9, double *energies);

/* ---------------------------------------------- learned REBO surrogate
 * No counterpart in the reference: a per-atom energy model
 * E = sum_i MLP(D_i) (40 invariant same-molecule descriptors within 6 A,
 * 3 tanh layers of width 64, BF16 MFMA) whose forces are exact gradients,
 * standing in for the exact-gradient REBO term (compute_rebo AB:275-421,
 * FMD_REBO_PHYSICAL).  Trained on the GPU on frames the engine generates;
 * while enabled, full REBO re-runs every judge_every steps and the engine
 * falls back to it for good when the force RMSE exceeds the threshold.
 * Atomic engines only; frames and training need a single-system engine.
 * theta: fmd_sg_param_count() floats (W1 [40][64], b1, W2 [64][64], b2,
 * W3 [64][64], b3, w4 [64], b4); mu, sd: fmd_sg_feature_count() floats. */
int fmd_sg_param_count(void);
int fmd_sg_feature_count(void);
int fmd_sg_attach(fmd_engine *e, int seed);
int fmd_sg_set_params(fmd_engine *e, const float *theta, const float *mu, const float *sd);
int fmd_sg_get_params(fmd_engine *e, float *theta, float *mu, float *sd);
/* Store the current positions and their exact-gradient REBO forces as a
 * training frame. */
int fmd_sg_collect(fmd_engine *e);
int fmd_sg_frame_count(fmd_engine *e, int *n);
/* mu, sd from the descriptors of the stored frames. */
int fmd_sg_fit_normalisation(fmd_engine *e);
/* iters Adam steps (one frame each, cycling) on the mean squared force
 * error; loss[it] = that error before step it (eV^2/A^2), or NULL. */
int fmd_sg_train(fmd_engine *e, int iters, double lr, double *loss);
/* Loss and its gradient over theta on one stored frame (no update). */
int fmd_sg_loss_grad(fmd_engine *e, int frame, double *loss, float *grad);
/* Normalised descriptors of a stored frame (Na x 40). */
int fmd_sg_descriptor(fmd_engine *e, int frame, float *feat);
/* Surrogate forces (Na x 3), virial, energy on a stored frame (frame >= 0,
 * rmse against its REBO forces) or on the current state (frame < 0). */
int fmd_sg_evaluate(fmd_engine *e, int frame, double *F, double *W9, double *E,
                    double *rmse);
int fmd_sg_enable(fmd_engine *e, int on, int judge_every, double threshold);
int fmd_sg_stats(fmd_engine *e, double *last_rmse, int *judged, int *fallbacks,
                 int *active);

/* Per-step phase control (the COLD / WARM / PROD logic of LJ:848-894). */
#define FMD_CTRL_ZERO_XI_VG 1  /* xi = 0, Vg = 0 before the step */
#define FMD_CTRL_ZERO_VG 2     /* Vg = 0 before the step */
#define FMD_CTRL_RESCALE 4     /* rescale velocities to rescale_T after it */
#define FMD_CTRL_RESCALE_ZERO_VG 8
#define FMD_CTRL_REBUILD 16    /* force a list rebuild before the step */
typedef struct {
  double Tt;         /* target temperature of the step */
  double rescale_T;
  int flags;
  int pad;
} fmd_step_ctrl;

/* Monitor record of one step (the quantities LJ:874-876 / AB:896 print). */
typedef struct {
  double Ep, KE, KEt, KEr, T, P, V, xi, Ep_rebo, Ep_lj;
  double h[9];
  /* molecular-mechanics engines: the bonded terms and Coulomb (0 otherwise;
   * Ep_rebo = Eb + Ea + Ed + Ei, the MM monitor's E_dih = Ed + Ei) */
  double Eb, Ea, Ed, Ei, Ecoul;
} fmd_thermo;

/* Prepare the current state as the reference does before its loop
 * (LJ:813-828): wrap positions into the cell, build the list, compute forces
 * and the observables the first step needs. */
int fmd_engine_prepare(fmd_engine *e);

/* Run nsteps steps whose global indices are first_step+1 .. first_step+nsteps.
 * ctrl: nsteps entries, or NULL (production: Tt = the state's Tt).  The list
 * is rebuilt before every global step divisible by nlup (LJ:859) or when
 * FMD_CTRL_REBUILD is set.  thermo: nsteps records, or NULL. */
int fmd_engine_run(fmd_engine *e, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
                   fmd_thermo *thermo);

/* Forces of the current state: rigid -> F = Fv (N x 3), T = Tv; atomic ->
 * F (Na x 3), T unused.  W9, energies {Ep, Ep_rebo, Ep_lj}.  Any may be NULL. */
int fmd_engine_forces(fmd_engine *e, double *F, double *T, double *W9, double *energies);

/* Counters: [0] list faults since creation (a list overflow; a rigid
 * molecule list found stale at its rebuild; an atomic engine's cell grown too
 * narrow for per-molecule images), [1] list rebuilds: the scheduled ones, and
 * for atomic engines also each replica's per-atom LJ lists rebuilt within a
 * step because an atom pair may have closed their skin, [2] steps. */
int fmd_engine_counters(fmd_engine *e, long long *out3);

/* Mean duration (ms) of the force kernel over the launches timed since the
 * last call (HIP events on the engine's stream; enabled by
 * fmd_engine_set_timing(e, on): on = 1 times every step, on = k > 1 every
 * k-th step of a run and every list rebuild, 0 none).  *n receives the
 * number of launches timed. */
int fmd_engine_set_timing(fmd_engine *e, int on);
int fmd_engine_kernel_time(fmd_engine *e, double *mean_ms, int *n);
/* Mean durations (ms) and counts of the timed step phases since the last
 * read: [0] the force phase, [1] its intramolecular part (REBO, surrogate or
 * bonded terms; 0 for rigid molecules), [2] its intermolecular LJ part, [3] a
 * list rebuild (rebuild steps only).  Resets the sums, as does
 * fmd_engine_kernel_time. */
int fmd_engine_phase_times(fmd_engine *e, double *mean_ms4, int *n4);

/* Replay step sequences from captured HIP graphs (1) or launch every kernel
 * (0).  Default: on for batches of more than one replica, off otherwise. */
int fmd_engine_set_graph(fmd_engine *e, int on);

/* Wait for all work of the engine. */
int fmd_engine_sync(fmd_engine *e);
void fmd_engine_destroy(fmd_engine *e);

#ifdef __cplusplus
}
#endif
#endif /* FMD_H */
