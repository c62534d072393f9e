/*
 * fmd_oracle.c -- FP64 CPU restatement of the fuller_md force / integrator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP kernels are
 * compared against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it; the product library never links or calls it.
 *
 * Every function follows one reference function; the citation is given as
 * alias:line with
 *   C1 = src/fuller_LJ_npt_md_core_serial.cpp
 *   LJ = src/fuller_LJ_npt_md_serial_omp_acc.cpp
 *   AB = src/fuller_airebo_npt_md_serial_omp_acc.cpp
 *   MM = src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp
 * The arithmetic order follows the reference so that the serial reference and
 * this restatement agree to a few ulps on identical inputs.
 *
 * How each function is pinned (besides the golden vectors written by the
 * reference build, tests/golden/):
 *   orc_rigid_forces     golden vectors (C1 and LJ), Newton-3 on molecule sums,
 *                        finite-difference gradient of the energy.
 *   orc_rigid_nlist      golden lists, set equality with brute force.
 *   orc_rigid_step_npt   golden multi-step trajectories (C1).
 *   orc_compute_rebo     golden vectors (compat); physical mode by FD gradient
 *                        and strain-derivative virial; dimer closed form.
 *   orc_compute_lj       golden vectors; FD gradient.
 *   orc_build_nlist_*    golden lists.
 *   orc_airebo_step_npt  golden multi-step trajectories (AB).
 *   orc_mm_topology      golden per-molecule terms (MM), bitwise.
 *   orc_mm_forces        golden vectors incl. Coulomb (compat); physical mode
 *                        by FD gradient and strain-derivative virial; compat
 *                        bond/angle forces equal +gradient term by term.
 *   orc_mm_step_npt      golden multi-step trajectories (MM).
 *   orc_rebo_ext         no reference counterpart (the REBO-II P_ij correction
 *                        and the AIREBO torsion term): equal to
 *                        orc_compute_rebo (physical) with both terms off; FD
 *                        gradient, strain-derivative virial, closed forms of
 *                        planar and twisted chains and of a three-coordinated
 *                        centre.  Parity with the reference is not defined
 *                        for these terms (the reference has neither).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "fmd_oracle.h"

#define H9(h, i, j) ((h)[3 * (i) + (j)])

/* Units: A, amu, eV, fs (C1:84-87). */
static const double CONV = 9.64853321e-3;
static const double KB = 8.617333262e-5;
static const double EV2GPA = 160.21766208;

/* REBO-II C-C parameters (AB:110-116). */
static const double Q_CC = 0.3134602960833, A_CC = 10953.544162170,
                    ALPHA_CC = 4.7465390606595;
static const double B1_CC = 12388.79197798, BETA1_CC = 4.7204523127;
static const double B2_CC = 17.56740646509, BETA2_CC = 1.4332132499;
static const double B3_CC = 30.71493208065, BETA3_CC = 1.3826912506;
static const double DMIN_CC = 1.7, DMAX_CC = 2.0, BO_DELTA = 0.5;
static const double G_A0 = 0.00020813, G_C0 = 330.0, G_D0 = 3.5;
static const double PI_ = 3.14159265358979323846;

static double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

/* mat_det9 C1:137-141 */
double orc_mat_det9(const double *h) {
  return H9(h, 0, 0) * (H9(h, 1, 1) * H9(h, 2, 2) - H9(h, 1, 2) * H9(h, 2, 1)) -
         H9(h, 0, 1) * (H9(h, 1, 0) * H9(h, 2, 2) - H9(h, 1, 2) * H9(h, 2, 0)) +
         H9(h, 0, 2) * (H9(h, 1, 0) * H9(h, 2, 1) - H9(h, 1, 1) * H9(h, 2, 0));
}

/* mat_inv9 C1:147-158 */
void orc_mat_inv9(const double *h, double *hi) {
  double d = orc_mat_det9(h), id = 1.0 / d;
  hi[0] = id * (h[4] * h[8] - h[5] * h[7]);
  hi[1] = id * (h[2] * h[7] - h[1] * h[8]);
  hi[2] = id * (h[1] * h[5] - h[2] * h[4]);
  hi[3] = id * (h[5] * h[6] - h[3] * h[8]);
  hi[4] = id * (h[0] * h[8] - h[2] * h[6]);
  hi[5] = id * (h[2] * h[3] - h[0] * h[5]);
  hi[6] = id * (h[3] * h[7] - h[4] * h[6]);
  hi[7] = id * (h[1] * h[6] - h[0] * h[7]);
  hi[8] = id * (h[0] * h[4] - h[1] * h[3]);
}

/* mimg_flat C1:166-175: fractional round-to-nearest minimum image. */
void orc_mimg(double *d, const double *hi, const double *h) {
  double dx = d[0], dy = d[1], dz = d[2];
  double s0 = hi[0] * dx + hi[1] * dy + hi[2] * dz;
  double s1 = hi[3] * dx + hi[4] * dy + hi[5] * dz;
  double s2 = hi[6] * dx + hi[7] * dy + hi[8] * dz;
  s0 -= round(s0);
  s1 -= round(s1);
  s2 -= round(s2);
  d[0] = h[0] * s0 + h[1] * s1 + h[2] * s2;
  d[1] = h[3] * s0 + h[4] * s1 + h[5] * s2;
  d[2] = h[6] * s0 + h[7] * s1 + h[8] * s2;
}

/* apply_pbc C1:309-320: wrap into [0,1) fractional. */
void orc_apply_pbc(double *pos, const double *h, const double *hi, int n) {
  for (int i = 0; i < n; i++) {
    double px = pos[i * 3], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
    double s0 = hi[0] * px + hi[1] * py + hi[2] * pz;
    double s1 = hi[3] * px + hi[4] * py + hi[5] * pz;
    double s2 = hi[6] * px + hi[7] * py + hi[8] * pz;
    s0 -= floor(s0);
    s1 -= floor(s1);
    s2 -= floor(s2);
    pos[i * 3] = h[0] * s0 + h[1] * s1 + h[2] * s2;
    pos[i * 3 + 1] = h[3] * s0 + h[4] * s1 + h[5] * s2;
    pos[i * 3 + 2] = h[6] * s0 + h[7] * s1 + h[8] * s2;
  }
}

/* q2R_flat C1:186-191 (q = [w,x,y,z]). */
void orc_q2R(const double *q, double *R) {
  double w = q[0], x = q[1], y = q[2], z = q[3];
  R[0] = 1 - 2 * (y * y + z * z);
  R[1] = 2 * (x * y - w * z);
  R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z);
  R[4] = 1 - 2 * (x * x + z * z);
  R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y);
  R[7] = 2 * (y * z + w * x);
  R[8] = 1 - 2 * (x * x + y * y);
}

/* qmul_flat C1:193-198 */
static void qmul(const double *a, const double *b, double *o) {
  o[0] = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
  o[1] = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
  o[2] = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
  o[3] = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
}

/* qnorm_flat C1:200-204 */
static void qnorm(double *q) {
  double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  double inv = 1.0 / n;
  q[0] *= inv;
  q[1] *= inv;
  q[2] *= inv;
  q[3] *= inv;
}

/* omega2dq_flat C1:206-216: exact rotation by |w| dt about w. */
static void omega2dq(double wx, double wy, double wz, double dt, double *dq) {
  double wm = sqrt(wx * wx + wy * wy + wz * wz);
  double th = wm * dt * 0.5;
  if (th < 1e-14) {
    dq[0] = 1.0;
    dq[1] = 0.5 * dt * wx;
    dq[2] = 0.5 * dt * wy;
    dq[3] = 0.5 * dt * wz;
  } else {
    double s = sin(th) / wm;
    dq[0] = cos(th);
    dq[1] = s * wx;
    dq[2] = s * wy;
    dq[3] = s * wz;
  }
}

/* ---------------------------------------------------------------- LJ params */

/* Rigid constants: sigma=3.431, eps=2.635e-3, rc=3 sigma, VSHFT with
 * sigma/rc = 1/3 (C1:93-104, LJ:103-107). */
void orc_lj_params_rigid(orc_lj_params *p) {
  const double sigma = 3.431, eps = 2.635e-3;
  const double rc = 3.0 * sigma;
  const double sr = 1.0 / 3.0, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
  p->sig2 = sigma * sigma;
  p->eps = eps;
  p->rc2 = rc * rc;
  p->vshift = 4.0 * eps * (sr6 * sr6 - sr6);
}

/* AIREBO intermolecular constants: sigma=3.40, eps=2.84e-3, rc=3 sigma,
 * VSHFT with sigma/rc (AB:119-124). */
void orc_lj_params_airebo(orc_lj_params *p) {
  const double sigma = 3.40, eps = 2.84e-3;
  const double rc = 3.0 * sigma;
  const double sr = sigma / rc, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
  p->sig2 = sigma * sigma;
  p->eps = eps;
  p->rc2 = rc * rc;
  p->vshift = 4 * eps * (sr6 * sr6 - sr6);
}

/* ------------------------------------------------------------- rigid body */

/* nlist_build_sym C1:280-301: symmetric full list of molecule pairs with
 * COM distance < rmcut+3 (minimum image), capped at max_neigh per molecule
 * in discovery order.  Overflowing entries are dropped (the reference only
 * warns). */
void orc_rigid_nlist(const double *pos, const double *h, const double *hi,
                     int n, double rmcut, int max_neigh, int *nl_count,
                     int *nl_list) {
  double rc2 = (rmcut + 3.0) * (rmcut + 3.0);
  for (int i = 0; i < n; i++) nl_count[i] = 0;
  for (int i = 0; i < n; i++) {
    for (int j = i + 1; j < n; j++) {
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 < rc2) {
        int ci = nl_count[i], cj = nl_count[j];
        if (ci < max_neigh) {
          nl_list[i * max_neigh + ci] = j;
          nl_count[i]++;
        }
        if (cj < max_neigh) {
          nl_list[j * max_neigh + cj] = i;
          nl_count[j]++;
        }
      }
    }
  }
}

/* Lab-frame site offsets, C1:354-364. */
void orc_rigid_lab(const double *qv, const double *body, int n, int natom,
                   double *lab) {
  for (int i = 0; i < n; i++) {
    double R[9];
    orc_q2R(&qv[i * 4], R);
    for (int a = 0; a < natom; a++) {
      double bx = body[a * 3], by = body[a * 3 + 1], bz = body[a * 3 + 2];
      int idx = i * natom * 3 + a * 3;
      lab[idx] = R[0] * bx + R[1] * by + R[2] * bz;
      lab[idx + 1] = R[3] * bx + R[4] * by + R[5] * bz;
      lab[idx + 2] = R[6] * bx + R[7] * by + R[8] * bz;
    }
  }
}

/* forces C1:346-444 (= LJ:271-400): site-site shifted LJ between rigid
 * molecules over the symmetric molecule list.  Each molecule accumulates only
 * its own force / torque; energy and virial carry a factor 1/2 for the double
 * counting.  Returns the potential energy. */
double orc_rigid_forces(double *Fv, double *Tv, double *Wm9, const double *pos,
                        const double *qv, const double *body, const double *h,
                        const double *hi, const int *nl_count,
                        const int *nl_list, int max_neigh, int n, int natom,
                        double rmcut2, const orc_lj_params *p, double *lab) {
  orc_rigid_lab(qv, body, n, natom, lab);
  for (int i = 0; i < n * 3; i++) {
    Fv[i] = 0.0;
    Tv[i] = 0.0;
  }
  for (int i = 0; i < 9; i++) Wm9[i] = 0.0;
  const double sig2 = p->sig2, eps = p->eps, rc2 = p->rc2, vsh = p->vshift;
  double Ep = 0.0;
  for (int i = 0; i < n; i++) {
    double fi0 = 0, fi1 = 0, fi2 = 0, ti0 = 0, ti1 = 0, ti2 = 0, my_Ep = 0;
    double w00 = 0, w01 = 0, w02 = 0, w10 = 0, w11 = 0, w12 = 0, w20 = 0,
           w21 = 0, w22 = 0;
    int nni = nl_count[i];
    for (int k = 0; k < nni; k++) {
      int j = nl_list[i * max_neigh + k];
      double dm[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                      pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(dm, hi, h);
      double dmx = dm[0], dmy = dm[1], dmz = dm[2];
      if (dmx * dmx + dmy * dmy + dmz * dmz > rmcut2) continue;
      for (int ai = 0; ai < natom; ai++) {
        int ia = i * natom * 3 + ai * 3;
        double rax = lab[ia], ray = lab[ia + 1], raz = lab[ia + 2];
        for (int bj = 0; bj < natom; bj++) {
          int jb = j * natom * 3 + bj * 3;
          double ddx = dmx + lab[jb] - rax;
          double ddy = dmy + lab[jb + 1] - ray;
          double ddz = dmz + lab[jb + 2] - raz;
          double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
          if (r2 < rc2) {
            if (r2 < 0.25) r2 = 0.25;
            double ri2 = 1.0 / r2;
            double sr2 = sig2 * ri2;
            double sr6 = sr2 * sr2 * sr2;
            double sr12 = sr6 * sr6;
            double fm = 24.0 * eps * (2.0 * sr12 - sr6) * ri2;
            double fx = fm * ddx, fy = fm * ddy, fz = fm * ddz;
            fi0 -= fx;
            fi1 -= fy;
            fi2 -= fz;
            ti0 -= (ray * fz - raz * fy);
            ti1 -= (raz * fx - rax * fz);
            ti2 -= (rax * fy - ray * fx);
            my_Ep += 0.5 * (4.0 * eps * (sr12 - sr6) - vsh);
            w00 += 0.5 * ddx * fx;
            w01 += 0.5 * ddx * fy;
            w02 += 0.5 * ddx * fz;
            w10 += 0.5 * ddy * fx;
            w11 += 0.5 * ddy * fy;
            w12 += 0.5 * ddy * fz;
            w20 += 0.5 * ddz * fx;
            w21 += 0.5 * ddz * fy;
            w22 += 0.5 * ddz * fz;
          }
        }
      }
    }
    Fv[i * 3] = fi0;
    Fv[i * 3 + 1] = fi1;
    Fv[i * 3 + 2] = fi2;
    Tv[i * 3] = ti0;
    Tv[i * 3 + 1] = ti1;
    Tv[i * 3 + 2] = ti2;
    Wm9[0] += w00;
    Wm9[1] += w01;
    Wm9[2] += w02;
    Wm9[3] += w10;
    Wm9[4] += w11;
    Wm9[5] += w12;
    Wm9[6] += w20;
    Wm9[7] += w21;
    Wm9[8] += w22;
    Ep += my_Ep;
  }
  return Ep;
}

/* ke_trans / ke_rot C1:451-463 */
double orc_ke_trans(const double *vel, int n, double mmol) {
  double s = 0;
  for (int i = 0; i < n; i++)
    s += vel[i * 3] * vel[i * 3] + vel[i * 3 + 1] * vel[i * 3 + 1] +
         vel[i * 3 + 2] * vel[i * 3 + 2];
  return 0.5 * mmol * s / CONV;
}

double orc_ke_rot(const double *omg, int n, double i0) {
  double s = 0;
  for (int i = 0; i < n; i++)
    s += omg[i * 3] * omg[i * 3] + omg[i * 3 + 1] * omg[i * 3 + 1] +
         omg[i * 3 + 2] * omg[i * 3 + 2];
  return 0.5 * i0 * s / CONV;
}

/* inst_T / inst_P C1:465-468 */
double orc_inst_T(double ke, int nf) { return 2 * ke / (nf * KB); }
double orc_inst_P(const double *W, double ket, double V) {
  return (2 * ket + W[0] + W[4] + W[8]) / (3 * V) * EV2GPA;
}

/* make_npt C1:480-488 (rigid: Nf = 6N-3) and AB:512-515 (atomic: 3Na-3). */
void orc_make_npt(orc_npt *s, double T, double Pe, int nf) {
  s->xi = 0;
  s->Q = fmax(nf * KB * T * 100.0 * 100.0, 1e-20);
  for (int i = 0; i < 9; i++) s->Vg[i] = 0;
  s->W = fmax((nf + 9) * KB * T * 1000.0 * 1000.0, 1e-20);
  s->Pe = Pe;
  s->Tt = T;
  s->Nf = nf;
}

/* step_npt C1:504-601 (= LJ:434-538): one Nose-Hoover + Parrinello-Rahman
 * velocity-Verlet step of rigid molecules.  Clamps: xi in +-0.1, diagonal
 * Vg in +-0.01.  out[0] = Ep, out[1] = KE after the second half-kick. */
void orc_rigid_step_npt(double *pos, double *vel, double *qv, double *omg,
                        double *Fv, double *Tv, double *Wm9, double *h,
                        double *hi, const double *body, double I0, double Mmol,
                        int n, int natom, double rmcut2, double dt,
                        orc_npt *npt, const int *nl_count, const int *nl_list,
                        int max_neigh, const orc_lj_params *p, double *lab,
                        double *out) {
  double hdt = 0.5 * dt;
  orc_mat_inv9(h, hi);
  double V = fabs(orc_mat_det9(h));
  double kt = orc_ke_trans(vel, n, Mmol), kr = orc_ke_rot(omg, n, I0),
         KE = kt + kr;
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.1, 0.1);
  double dP = orc_inst_P(Wm9, kt, V) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.01, 0.01);
  double eps_tr = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_nh = exp(-hdt * npt->xi);
  double sc_pr = exp(-hdt * eps_tr / 3.0);
  double sc_v = sc_nh * sc_pr;
  double cF = CONV / Mmol, cT = CONV / I0;
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      vel[i * 3 + a] = vel[i * 3 + a] * sc_v + hdt * Fv[i * 3 + a] * cF;
      omg[i * 3 + a] = omg[i * 3 + a] * sc_nh + hdt * Tv[i * 3 + a] * cT;
    }
  for (int i = 0; i < n; i++) {
    double px = pos[i * 3], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
    double vx = vel[i * 3], vy = vel[i * 3 + 1], vz = vel[i * 3 + 2];
    double sx = hi[0] * px + hi[1] * py + hi[2] * pz;
    double sy = hi[3] * px + hi[4] * py + hi[5] * pz;
    double sz = hi[6] * px + hi[7] * py + hi[8] * pz;
    double vsx = hi[0] * vx + hi[1] * vy + hi[2] * vz;
    double vsy = hi[3] * vx + hi[4] * vy + hi[5] * vz;
    double vsz = hi[6] * vx + hi[7] * vy + hi[8] * vz;
    sx += dt * vsx;
    sy += dt * vsy;
    sz += dt * vsz;
    sx -= floor(sx);
    sy -= floor(sy);
    sz -= floor(sz);
    pos[i * 3] = sx;
    pos[i * 3 + 1] = sy;
    pos[i * 3 + 2] = sz;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) h[a * 3 + b] += dt * npt->Vg[a * 3 + b];
  for (int i = 0; i < n; i++) {
    double sx = pos[i * 3], sy = pos[i * 3 + 1], sz = pos[i * 3 + 2];
    pos[i * 3] = h[0] * sx + h[1] * sy + h[2] * sz;
    pos[i * 3 + 1] = h[3] * sx + h[4] * sy + h[5] * sz;
    pos[i * 3 + 2] = h[6] * sx + h[7] * sy + h[8] * sz;
  }
  for (int i = 0; i < n; i++) {
    double dq[4], tmp[4];
    omega2dq(omg[i * 3], omg[i * 3 + 1], omg[i * 3 + 2], dt, dq);
    qmul(&qv[i * 4], dq, tmp);
    qv[i * 4] = tmp[0];
    qv[i * 4 + 1] = tmp[1];
    qv[i * 4 + 2] = tmp[2];
    qv[i * 4 + 3] = tmp[3];
    qnorm(&qv[i * 4]);
  }
  orc_mat_inv9(h, hi);
  double Ep = orc_rigid_forces(Fv, Tv, Wm9, pos, qv, body, h, hi, nl_count,
                               nl_list, max_neigh, n, natom, rmcut2, p, lab);
  double eps_tr2 = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_v2 = sc_nh * exp(-hdt * eps_tr2 / 3.0);
  for (int i = 0; i < n; i++)
    for (int a = 0; a < 3; a++) {
      vel[i * 3 + a] = (vel[i * 3 + a] + hdt * Fv[i * 3 + a] * cF) * sc_v2;
      omg[i * 3 + a] = (omg[i * 3 + a] + hdt * Tv[i * 3 + a] * cT) * sc_nh;
    }
  kt = orc_ke_trans(vel, n, Mmol);
  kr = orc_ke_rot(omg, n, I0);
  KE = kt + kr;
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.1, 0.1);
  double V2 = fabs(orc_mat_det9(h));
  dP = orc_inst_P(Wm9, kt, V2) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V2 * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.01, 0.01);
  out[0] = Ep;
  out[1] = KE;
}

/* ------------------------------------------------------------------ AIREBO */

/* fc_d / dfc_d AB:165-171: cosine switch between DMIN and DMAX. */
double orc_rebo_fc(double r) {
  if (r <= DMIN_CC) return 1.0;
  if (r >= DMAX_CC) return 0.0;
  return 0.5 * (1.0 + cos(PI_ * (r - DMIN_CC) / (DMAX_CC - DMIN_CC)));
}
double orc_rebo_dfc(double r) {
  if (r <= DMIN_CC || r >= DMAX_CC) return 0.0;
  return -0.5 * PI_ / (DMAX_CC - DMIN_CC) * sin(PI_ * (r - DMIN_CC) / (DMAX_CC - DMIN_CC));
}
/* VR_CC_d / dVR_CC_d AB:173-177 */
double orc_rebo_vr(double r) { return (1.0 + Q_CC / r) * A_CC * exp(-ALPHA_CC * r); }
double orc_rebo_dvr(double r) {
  double ex = A_CC * exp(-ALPHA_CC * r);
  return (-Q_CC / (r * r)) * ex + (1.0 + Q_CC / r) * (-ALPHA_CC) * ex;
}
/* VA_CC_d / dVA_CC_d AB:179-183 */
double orc_rebo_va(double r) {
  return B1_CC * exp(-BETA1_CC * r) + B2_CC * exp(-BETA2_CC * r) +
         B3_CC * exp(-BETA3_CC * r);
}
double orc_rebo_dva(double r) {
  return -BETA1_CC * B1_CC * exp(-BETA1_CC * r) - BETA2_CC * B2_CC * exp(-BETA2_CC * r) -
         BETA3_CC * B3_CC * exp(-BETA3_CC * r);
}
/* G_C_d / dG_C_d AB:185-191: Brenner-1990 rational angular function. */
double orc_rebo_G(double x) {
  double c2 = G_C0 * G_C0, d2 = G_D0 * G_D0, hh = 1.0 + x;
  return G_A0 * (1.0 + c2 / d2 - c2 / (d2 + hh * hh));
}
double orc_rebo_dG(double x) {
  double c2 = G_C0 * G_C0, d2 = G_D0 * G_D0, hh = 1.0 + x, dn = d2 + hh * hh;
  return G_A0 * 2.0 * c2 * hh / (dn * dn);
}

/* build_nlist_rebo AB:251-261: symmetric list of same-molecule atoms within
 * rcut (reference 2.3 A), capped in discovery order. */
void orc_build_nlist_rebo(const double *pos, const double *h, const double *hi,
                          int na, const int *mol_id, double rcut, int max_neigh,
                          int *nlc, int *nll) {
  double rc2 = rcut * rcut;
  for (int i = 0; i < na; i++) nlc[i] = 0;
  for (int i = 0; i < na - 1; i++) {
    int mi = mol_id[i];
    for (int j = i + 1; j < na; j++) {
      if (mol_id[j] != mi) continue;
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rc2) {
        if (nlc[i] < max_neigh) {
          nll[i * max_neigh + nlc[i]] = j;
          nlc[i]++;
        }
        if (nlc[j] < max_neigh) {
          nll[j * max_neigh + nlc[j]] = i;
          nlc[j]++;
        }
      }
    }
  }
}

/* build_nlist_lj AB:263-272: half list (j > i) of other-molecule atoms
 * within rlist (reference rc+2 = 12.2 A), keeping the first max_neigh
 * (reference 400) in ascending j; the rest are silently dropped. */
void orc_build_nlist_lj(const double *pos, const double *h, const double *hi,
                        int na, const int *mol_id, double rlist, int max_neigh,
                        int *nlc, int *nll) {
  double rc2 = rlist * rlist;
  for (int i = 0; i < na; i++) nlc[i] = 0;
  for (int i = 0; i < na - 1; i++) {
    int mi = mol_id[i];
    for (int j = i + 1; j < na; j++) {
      if (mol_id[j] == mi) continue;
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < rc2) {
        if (nlc[i] < max_neigh) {
          nll[i * max_neigh + nlc[i]] = j;
          nlc[i]++;
        }
      }
    }
  }
}

/* One angular term of the bond order around centre c for bond c->p with the
 * third atom t (all displacement vectors taken from c).  Returns the two
 * force vectors of the physical gradient: ft on t and fp on p (the centre
 * receives -(ft+fp)).  coeff = 0.5 fc(r_cp) VA(r_cp) db/dGs. */
static void rebo_angular(const double *rcp, double r_cp, const double *rct,
                         double r_ct, double fc_ct, double dfc_ct, double coeff,
                         double *ft, double *fp) {
  double uct[3] = {rct[0] / r_ct, rct[1] / r_ct, rct[2] / r_ct};
  double ucp[3] = {rcp[0] / r_cp, rcp[1] / r_cp, rcp[2] / r_cp};
  double cs = (rcp[0] * rct[0] + rcp[1] * rct[1] + rcp[2] * rct[2]) / (r_cp * r_ct);
  if (cs < -1.0) cs = -1.0;
  if (cs > 1.0) cs = 1.0;
  double g = orc_rebo_G(cs), dg = orc_rebo_dG(cs);
  for (int a = 0; a < 3; a++) {
    ft[a] = coeff * (dfc_ct * g * uct[a] + fc_ct * dg * (ucp[a] - cs * uct[a]) / r_ct);
    fp[a] = coeff * fc_ct * dg * (uct[a] - cs * ucp[a]) / r_cp;
  }
}

/* compute_rebo AB:275-421.
 *   mode ORC_REBO_COMPAT: exactly the reference -- pair force and pair-only
 *     virial per bond i<j; angular forces on k (and l) with the reference's
 *     sign and without the d(cos)/d(r_j) term (AB:360-419).
 *   mode ORC_REBO_PHYSICAL: the exact negative gradient of the same energy
 *     (angular forces on k, j, i) and the full virial sum r_ab (x) F_b.
 * Adds into F and vir9 (the caller zeroes them, as compute_forces does,
 * AB:481-492).  Returns the REBO energy. */
double orc_compute_rebo(double *F, double *vir9, const double *pos,
                        const double *h, const double *hi, const int *nlc,
                        const int *nll, int max_neigh, int na, int mode) {
  double Ep = 0;
  for (int i = 0; i < na; i++) {
    int nni = nlc[i];
    for (int jn = 0; jn < nni; jn++) {
      int j = nll[i * max_neigh + jn];
      if (j <= i) continue;
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      double dx = d[0], dy = d[1], dz = d[2];
      double rij = sqrt(dx * dx + dy * dy + dz * dz);
      if (rij > DMAX_CC) continue;
      double fcut = orc_rebo_fc(rij), dfcut = orc_rebo_dfc(rij);
      if (fcut < 1e-15 && dfcut == 0) continue;
      double vr = orc_rebo_vr(rij), dvr = orc_rebo_dvr(rij);
      double va = orc_rebo_va(rij), dva = orc_rebo_dva(rij);
      double rij_inv = 1.0 / rij;
      double rhat0 = dx * rij_inv, rhat1 = dy * rij_inv, rhat2 = dz * rij_inv;

      double Gs_ij = 0;
      for (int kn = 0; kn < nni; kn++) {
        int k = nll[i * max_neigh + kn];
        if (k == j) continue;
        double dk[3] = {pos[k * 3] - pos[i * 3], pos[k * 3 + 1] - pos[i * 3 + 1],
                        pos[k * 3 + 2] - pos[i * 3 + 2]};
        orc_mimg(dk, hi, h);
        double rik = sqrt(dk[0] * dk[0] + dk[1] * dk[1] + dk[2] * dk[2]);
        if (rik > DMAX_CC) continue;
        double fc_ik = orc_rebo_fc(rik);
        if (fc_ik < 1e-15) continue;
        double costh = (dx * dk[0] + dy * dk[1] + dz * dk[2]) / (rij * rik);
        if (costh < -1.0) costh = -1.0;
        if (costh > 1.0) costh = 1.0;
        Gs_ij += fc_ik * orc_rebo_G(costh);
      }
      double bij = pow(1.0 + Gs_ij, -BO_DELTA);

      double Gs_ji = 0;
      int nnj = nlc[j];
      for (int ln = 0; ln < nnj; ln++) {
        int l = nll[j * max_neigh + ln];
        if (l == i) continue;
        double dl[3] = {pos[l * 3] - pos[j * 3], pos[l * 3 + 1] - pos[j * 3 + 1],
                        pos[l * 3 + 2] - pos[j * 3 + 2]};
        orc_mimg(dl, hi, h);
        double rjl = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
        if (rjl > DMAX_CC) continue;
        double fc_jl = orc_rebo_fc(rjl);
        if (fc_jl < 1e-15) continue;
        double costh = (-dx * dl[0] - dy * dl[1] - dz * dl[2]) / (rij * rjl);
        if (costh < -1.0) costh = -1.0;
        if (costh > 1.0) costh = 1.0;
        Gs_ji += fc_jl * orc_rebo_G(costh);
      }
      double bji = pow(1.0 + Gs_ji, -BO_DELTA);
      double bbar = 0.5 * (bij + bji);

      Ep += fcut * (vr - bbar * va);
      double fpair = (dfcut * (vr - bbar * va) + fcut * (dvr - bbar * dva)) * rij_inv;
      F[i * 3] += fpair * dx;
      F[i * 3 + 1] += fpair * dy;
      F[i * 3 + 2] += fpair * dz;
      F[j * 3] -= fpair * dx;
      F[j * 3 + 1] -= fpair * dy;
      F[j * 3 + 2] -= fpair * dz;
      for (int a = 0; a < 3; a++) {
        double da = d[a];
        for (int b = 0; b < 3; b++) vir9[a * 3 + b] -= da * fpair * d[b];
      }

      /* angular terms: side 0 = centre i (partner j, thirds k),
       *                side 1 = centre j (partner i, thirds l). */
      for (int side = 0; side < 2; side++) {
        int c = side == 0 ? i : j;
        int pp = side == 0 ? j : i;
        double Gs = side == 0 ? Gs_ij : Gs_ji;
        if (!(fabs(Gs) > 1e-20 && va > 1e-20)) continue;
        double dbp = -BO_DELTA * pow(1.0 + Gs, -BO_DELTA - 1.0);
        double vh = 0.5 * fcut * va;
        double rcp[3] = {side == 0 ? dx : -dx, side == 0 ? dy : -dy, side == 0 ? dz : -dz};
        double rhc[3] = {side == 0 ? rhat0 : -rhat0, side == 0 ? rhat1 : -rhat1,
                         side == 0 ? rhat2 : -rhat2};
        int nnc = nlc[c];
        for (int tn = 0; tn < nnc; tn++) {
          int t = nll[c * max_neigh + tn];
          if (t == pp) continue;
          double dt_[3] = {pos[t * 3] - pos[c * 3], pos[t * 3 + 1] - pos[c * 3 + 1],
                           pos[t * 3 + 2] - pos[c * 3 + 2]};
          orc_mimg(dt_, hi, h);
          double rct = sqrt(dt_[0] * dt_[0] + dt_[1] * dt_[1] + dt_[2] * dt_[2]);
          if (rct > DMAX_CC) continue;
          double fc_ct = orc_rebo_fc(rct);
          if (fc_ct < 1e-15) continue;
          double dfc_ct = orc_rebo_dfc(rct);
          if (mode == ORC_REBO_COMPAT) {
            /* AB:370-386 / 398-419, operation order kept. */
            double rct_inv = 1.0 / rct;
            double rk[3] = {dt_[0] * rct_inv, dt_[1] * rct_inv, dt_[2] * rct_inv};
            double costh = (rcp[0] * dt_[0] + rcp[1] * dt_[1] + rcp[2] * dt_[2]) / (rij * rct);
            if (costh < -1.0) costh = -1.0;
            if (costh > 1.0) costh = 1.0;
            double gv = orc_rebo_G(costh), dgv = orc_rebo_dG(costh), coeff = -vh * dbp;
            for (int a = 0; a < 3; a++) {
              double fk1 = coeff * dfc_ct * gv * rk[a];
              double dc = (rhc[a] - costh * rk[a]) * rct_inv;
              double fk = fk1 + coeff * fc_ct * dgv * dc;
              F[t * 3 + a] += fk;
              F[c * 3 + a] -= fk;
            }
          } else {
            double ft[3], fp[3];
            rebo_angular(rcp, rij, dt_, rct, fc_ct, dfc_ct, vh * dbp, ft, fp);
            for (int a = 0; a < 3; a++) {
              F[t * 3 + a] += ft[a];
              F[pp * 3 + a] += fp[a];
              F[c * 3 + a] -= ft[a] + fp[a];
            }
            for (int a = 0; a < 3; a++)
              for (int b = 0; b < 3; b++)
                vir9[a * 3 + b] += dt_[a] * ft[b] + rcp[a] * fp[b];
          }
        }
      }
    }
  }
  return Ep;
}

/* ------------------------------------------ REBO-II / AIREBO extension
 * The reference's REBO (above) has neither the REBO-II P_ij correction nor
 * the AIREBO torsion term.  This is an independent FP64 formulation of the
 * exact-gradient energy with both (Brenner et al., J. Phys.: Condens. Matter
 * 14, 783 (2002); Stuart, Tutein & Harrison, J. Chem. Phys. 112, 6472
 * (2000)), written per bond rather than in the reference's loop order:
 *   E = sum_{bonds i<j} fc_ij (VR_ij - bbar_ij VA_ij)
 *       + sum_{chains i-j-k-l} w_ij w_jk w_kl s(cos th_ijk) s(cos th_jkl) V(cos om)
 *   b_ij = (1 + S_ij)^-1/2,  S_ij = sum_{k != j} fc_ik G(cos th_jik) + P(N_ij),
 *   N_ij = sum_{k != j} fc_ik (carbon neighbours; no hydrogen, N^H = 0),
 *   P: cubic Hermite through pcc[n] at N = n (n = 0..4), zero slope at the
 *      knots, pcc[4] beyond;
 *   V(c) = eps_t (256/405 ((1 + c)/2)^5 - 1/10) = eps_t (256/405 cos^10(om/2)
 *      - 1/10), om the dihedral of the chain (0 = cis), w = fc;
 *   s(c) = 1 for c >= -0.995, t^2 (3 - 2t) with t = (c + 1) / 0.005 below
 *      (the torsion fades out as a bond angle straightens, where the
 *      dihedral is undefined).
 * Each chain (bonds i-j, j-k, k-l, i != l, j-k its central bond) counts
 * once.  Forces are the exact negative gradient; the virial is
 * sum_a r_a (x) F_a with each term's atoms taken through its bond vectors. */
static double pcc_spline(const double *pc, double N, double *dP) {
  if (N <= 0.0) {
    *dP = 0.0;
    return pc[0];
  }
  if (N >= 4.0) {
    *dP = 0.0;
    return pc[4];
  }
  const int n = (int)floor(N);
  const double t = N - n, dp = pc[n + 1] - pc[n];
  *dP = dp * 6.0 * t * (1.0 - t);
  return pc[n] + dp * t * t * (3.0 - 2.0 * t);
}

static double tors_switch(double c, double *ds) {
  const double c0 = -1.0, c1 = -0.995;
  if (c >= c1) {
    *ds = 0.0;
    return 1.0;
  }
  if (c <= c0) {
    *ds = 0.0;
    return 0.0;
  }
  const double t = (c - c0) / (c1 - c0);
  *ds = 6.0 * t * (1.0 - t) / (c1 - c0);
  return t * t * (3.0 - 2.0 * t);
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* One torsion term of the chain with bond vectors b1 = r_j - r_i,
 * b2 = r_k - r_j, b3 = r_l - r_k, switches (w, dw/dr) of the three bonds.
 * Returns the energy; g1, g2, g3 = dE/db1, dE/db2, dE/db3. */
static double tors_term(const double *b1, const double *b2, const double *b3,
                        const double *w, const double *dw, double eps_t,
                        double *g1, double *g2, double *g3) {
  const double r1 = sqrt(dot3(b1, b1)), r2 = sqrt(dot3(b2, b2)), r3 = sqrt(dot3(b3, b3));
  double u1[3], u2[3], u3[3];
  for (int a = 0; a < 3; a++) {
    u1[a] = b1[a] / r1;
    u2[a] = b2[a] / r2;
    u3[a] = b3[a] / r3;
  }
  for (int a = 0; a < 3; a++) g1[a] = g2[a] = g3[a] = 0.0;
  /* bond angles at j (between i and k) and at k (between j and l) */
  const double cj = -dot3(u1, u2), ck = -dot3(u2, u3);
  double dsj, dsk;
  const double sj = tors_switch(cj, &dsj), sk = tors_switch(ck, &dsk);
  if (sj == 0.0 || sk == 0.0) return 0.0;
  double n1[3], n2[3];
  cross3(b1, b2, n1);
  cross3(b2, b3, n2);
  const double m1 = sqrt(dot3(n1, n1)), m2 = sqrt(dot3(n2, n2));
  if (m1 == 0.0 || m2 == 0.0) return 0.0;
  double c = dot3(n1, n2) / (m1 * m2);
  if (c > 1.0) c = 1.0;
  if (c < -1.0) c = -1.0;
  const double hc = 0.5 * (1.0 + c), hc4 = hc * hc * hc * hc;
  const double V = eps_t * (256.0 / 405.0 * hc4 * hc - 0.1);
  const double dV = eps_t * 256.0 / 405.0 * 2.5 * hc4;
  const double W = w[0] * w[1] * w[2], S = sj * sk;
  /* dc/dn1, dc/dn2 */
  double A[3], B[3];
  for (int a = 0; a < 3; a++) {
    A[a] = (n2[a] / m2 - c * n1[a] / m1) / m1;
    B[a] = (n1[a] / m1 - c * n2[a] / m2) / m2;
  }
  /* dc/db1 = b2 x A, dc/db2 = A x b1 + b3 x B, dc/db3 = B x b2 */
  double c1[3], c2a[3], c2b[3], c3[3];
  cross3(b2, A, c1);
  cross3(A, b1, c2a);
  cross3(b3, B, c2b);
  cross3(B, b2, c3);
  const double kc = W * S * dV;
  /* d cj / db1 = -(u2 + cj u1) / r1, d cj / db2 = -(u1 + cj u2) / r2;
     d ck / db2 = -(u3 + ck u2) / r2, d ck / db3 = -(u2 + ck u3) / r3 */
  const double kj = W * V * dsj * sk, kk = W * V * sj * dsk;
  const double kw1 = S * V * dw[0] * w[1] * w[2], kw2 = S * V * w[0] * dw[1] * w[2],
               kw3 = S * V * w[0] * w[1] * dw[2];
  for (int a = 0; a < 3; a++) {
    g1[a] = kc * c1[a] - kj * (u2[a] + cj * u1[a]) / r1 + kw1 * u1[a];
    g2[a] = kc * (c2a[a] + c2b[a]) - kj * (u1[a] + cj * u2[a]) / r2 -
            kk * (u3[a] + ck * u2[a]) / r2 + kw2 * u2[a];
    g3[a] = kc * c3[a] - kk * (u2[a] + ck * u3[a]) / r3 + kw3 * u3[a];
  }
  return W * S * V;
}

double orc_rebo_ext(double *F, double *vir9, const double *pos, const double *h,
                    const double *hi, const int *nlc, const int *nll, int max_neigh, int na,
                    const double *pcc, double eps_t, double *e_tors) {
  /* the bonds of every atom (list entries shorter than DMAX_CC) */
  int *nb = (int *)calloc((size_t)na, sizeof(int));
  int *bj = (int *)malloc((size_t)na * max_neigh * sizeof(int));
  double *bv = (double *)malloc((size_t)na * max_neigh * 6 * sizeof(double));
  for (int i = 0; i < na; i++) {
    for (int jn = 0; jn < nlc[i]; jn++) {
      const int j = nll[i * max_neigh + jn];
      double d[3] = {pos[3 * j] - pos[3 * i], pos[3 * j + 1] - pos[3 * i + 1],
                     pos[3 * j + 2] - pos[3 * i + 2]};
      orc_mimg(d, hi, h);
      const double r = sqrt(dot3(d, d));
      if (r >= DMAX_CC) continue;
      const int q = i * max_neigh + nb[i]++;
      bj[q] = j;
      double *e = bv + 6 * (size_t)q;
      e[0] = d[0];
      e[1] = d[1];
      e[2] = d[2];
      e[3] = r;
      e[4] = orc_rebo_fc(r);
      e[5] = orc_rebo_dfc(r);
    }
  }
  double Ep = 0.0;
  /* bond energies and forces: each bond i<j, both directions' bond orders */
  for (int i = 0; i < na; i++) {
    for (int a = 0; a < nb[i]; a++) {
      const int j = bj[i * max_neigh + a];
      if (j <= i) continue;
      const double *e = bv + 6 * (size_t)(i * max_neigh + a);
      const double r = e[3], fc = e[4], dfc = e[5];
      const double vr = orc_rebo_vr(r), dvr = orc_rebo_dvr(r);
      const double va = orc_rebo_va(r), dva = orc_rebo_dva(r);
      /* bond orders of the two directions: centre c, partner p */
      double bdir[2], Sdir[2], dPdir[2];
      int cq[2] = {i, j}, pq[2] = {j, i};
      for (int side = 0; side < 2; side++) {
        const int c = cq[side], pp = pq[side];
        const double sg = side == 0 ? 1.0 : -1.0;
        const double ucp[3] = {sg * e[0] / r, sg * e[1] / r, sg * e[2] / r};
        double Gs = 0.0, N = 0.0;
        for (int t = 0; t < nb[c]; t++) {
          if (bj[c * max_neigh + t] == pp) continue;
          const double *et = bv + 6 * (size_t)(c * max_neigh + t);
          double cs = (ucp[0] * et[0] + ucp[1] * et[1] + ucp[2] * et[2]) / et[3];
          if (cs > 1.0) cs = 1.0;
          if (cs < -1.0) cs = -1.0;
          Gs += et[4] * orc_rebo_G(cs);
          N += et[4];
        }
        double dP;
        const double P = pcc_spline(pcc, N, &dP);
        Sdir[side] = Gs + P;
        dPdir[side] = dP;
        bdir[side] = 1.0 / sqrt(1.0 + Sdir[side]);
      }
      const double bbar = 0.5 * (bdir[0] + bdir[1]);
      Ep += fc * (vr - bbar * va);
      /* pair force at fixed bond orders, on i along d = r_j - r_i */
      const double fpair = (dfc * (vr - bbar * va) + fc * (dvr - bbar * dva)) / r;
      for (int x = 0; x < 3; x++) {
        F[3 * i + x] += fpair * e[x];
        F[3 * j + x] -= fpair * e[x];
      }
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) vir9[3 * x + y] -= e[x] * fpair * e[y];
      /* through each direction's S: dE/dS = 1/4 fc VA b^3 */
      for (int side = 0; side < 2; side++) {
        const int c = cq[side], pp = pq[side];
        const double sg = side == 0 ? 1.0 : -1.0;
        const double dcp[3] = {sg * e[0], sg * e[1], sg * e[2]};
        const double ucp[3] = {dcp[0] / r, dcp[1] / r, dcp[2] / r};
        const double b = bdir[side];
        const double dEdS = 0.25 * fc * va * b * b * b;
        for (int t = 0; t < nb[c]; t++) {
          if (bj[c * max_neigh + t] == pp) continue;
          const int tt = bj[c * max_neigh + t];
          const double *et = bv + 6 * (size_t)(c * max_neigh + t);
          const double rt = et[3];
          const double uct[3] = {et[0] / rt, et[1] / rt, et[2] / rt};
          double cs = dot3(ucp, uct);
          if (cs > 1.0) cs = 1.0;
          if (cs < -1.0) cs = -1.0;
          const double g = orc_rebo_G(cs), dg = orc_rebo_dG(cs);
          double ft[3], fp[3];
          for (int x = 0; x < 3; x++) {
            /* dS/dd_ct and dS/dd_cp of this third atom */
            const double st = et[5] * (g + dPdir[side]) * uct[x] +
                              et[4] * dg * (ucp[x] - cs * uct[x]) / rt;
            const double sp = et[4] * dg * (uct[x] - cs * ucp[x]) / r;
            ft[x] = -dEdS * st;
            fp[x] = -dEdS * sp;
          }
          for (int x = 0; x < 3; x++) {
            F[3 * tt + x] += ft[x];
            F[3 * pp + x] += fp[x];
            F[3 * c + x] -= ft[x] + fp[x];
          }
          for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) vir9[3 * x + y] += et[x] * ft[y] + dcp[x] * fp[y];
        }
      }
    }
  }
  /* torsion: chains i-j-k-l around each bond j<k */
  double Et = 0.0;
  if (eps_t != 0.0) {
    for (int j = 0; j < na; j++) {
      for (int a = 0; a < nb[j]; a++) {
        const int k = bj[j * max_neigh + a];
        if (k <= j) continue;
        const double *ejk = bv + 6 * (size_t)(j * max_neigh + a);
        for (int bi = 0; bi < nb[j]; bi++) {
          const int i = bj[j * max_neigh + bi];
          if (i == k) continue;
          const double *eji = bv + 6 * (size_t)(j * max_neigh + bi);
          for (int cl = 0; cl < nb[k]; cl++) {
            const int l = bj[k * max_neigh + cl];
            if (l == j || l == i) continue;
            const double *ekl = bv + 6 * (size_t)(k * max_neigh + cl);
            const double b1[3] = {-eji[0], -eji[1], -eji[2]};
            const double b2[3] = {ejk[0], ejk[1], ejk[2]};
            const double b3[3] = {ekl[0], ekl[1], ekl[2]};
            const double w[3] = {eji[4], ejk[4], ekl[4]}, dw[3] = {eji[5], ejk[5], ekl[5]};
            double g1[3], g2[3], g3[3];
            Et += tors_term(b1, b2, b3, w, dw, eps_t, g1, g2, g3);
            /* r_i - r_j = -b1, r_k - r_j = b2, r_l - r_j = b2 + b3 */
            double Fi[3], Fj[3], Fk[3], Fl[3];
            for (int x = 0; x < 3; x++) {
              Fi[x] = g1[x];
              Fj[x] = g2[x] - g1[x];
              Fk[x] = g3[x] - g2[x];
              Fl[x] = -g3[x];
              F[3 * i + x] += Fi[x];
              F[3 * j + x] += Fj[x];
              F[3 * k + x] += Fk[x];
              F[3 * l + x] += Fl[x];
            }
            for (int x = 0; x < 3; x++)
              for (int y = 0; y < 3; y++)
                vir9[3 * x + y] += -b1[x] * Fi[y] + b2[x] * Fk[y] + (b2[x] + b3[x]) * Fl[y];
          }
        }
      }
    }
  }
  free(nb);
  free(bj);
  free(bv);
  if (e_tors) *e_tors = Et;
  return Ep + Et;
}

/* compute_lj AB:424-473: shifted 12-6 over the half list with r^2 <= rc^2
 * (note <=), r^2 clamped at 0.25, Newton-3 pair forces and 9-component
 * virial sum d_a fm d_b.  Adds into F / vir9.  Returns the energy. */
double orc_compute_lj(double *F, double *vir9, const double *pos, const double *h,
                      const double *hi, const int *nlc, const int *nll,
                      int max_neigh, int na, const orc_lj_params *p) {
  const double sig2 = p->sig2, eps = p->eps, rc2 = p->rc2, vsh = p->vshift;
  double Ep = 0;
  for (int i = 0; i < na; i++) {
    int nni = nlc[i];
    for (int jn = 0; jn < nni; jn++) {
      int j = nll[i * max_neigh + jn];
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      double dx = d[0], dy = d[1], dz = d[2];
      double r2 = dx * dx + dy * dy + dz * dz;
      if (r2 > rc2) continue;
      if (r2 < 0.25) r2 = 0.25;
      double ri2 = 1.0 / r2, sr2 = sig2 * ri2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
      double fm = 24 * eps * (2 * sr12 - sr6) * ri2;
      Ep += 4 * eps * (sr12 - sr6) - vsh;
      F[i * 3] -= fm * dx;
      F[i * 3 + 1] -= fm * dy;
      F[i * 3 + 2] -= fm * dz;
      F[j * 3] += fm * dx;
      F[j * 3 + 1] += fm * dy;
      F[j * 3 + 2] += fm * dz;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) vir9[a * 3 + b] += d[a] * fm * d[b];
    }
  }
  return Ep;
}

/* compute_forces AB:477-495: zero, REBO, LJ.  out = {Ep, Ep_rebo, Ep_lj}. */
void orc_airebo_forces(double *F, double *vir9, const double *pos, const double *h,
                       const double *hi, const int *nlc_r, const int *nll_r,
                       int max_r, const int *nlc_l, const int *nll_l, int max_l,
                       int na, int rebo_mode, const orc_lj_params *p, double *out) {
  for (int i = 0; i < na * 3; i++) F[i] = 0;
  for (int i = 0; i < 9; i++) vir9[i] = 0;
  double er = orc_compute_rebo(F, vir9, pos, h, hi, nlc_r, nll_r, max_r, na, rebo_mode);
  double el = orc_compute_lj(F, vir9, pos, h, hi, nlc_l, nll_l, max_l, na, p);
  out[0] = er + el;
  out[1] = er;
  out[2] = el;
}

/* ke_total AB:498-506 */
double orc_ke_total(const double *vel, const double *mass, int na) {
  double s = 0;
  for (int i = 0; i < na; i++)
    s += mass[i] * (vel[i * 3] * vel[i * 3] + vel[i * 3 + 1] * vel[i * 3 + 1] +
                    vel[i * 3 + 2] * vel[i * 3 + 2]);
  return 0.5 * s / CONV;
}

/* step_npt AB:518-590: atomic NH + PR velocity-Verlet step.  Clamps: xi in
 * +-0.05, diagonal Vg in +-0.005.  Note h^-1 is refreshed after the cell
 * update, before the fractional->Cartesian map (AB:555).
 * out = {Ep, Ep_rebo, Ep_lj, KE}. */
void orc_airebo_step_npt(double *pos, double *vel, double *F, double *vir9,
                         double *h, double *hi, const double *mass, int na,
                         double dt, orc_npt *npt, const int *nlc_r,
                         const int *nll_r, int max_r, const int *nlc_l,
                         const int *nll_l, int max_l, int rebo_mode,
                         const orc_lj_params *p, double *out) {
  double hdt = 0.5 * dt, V = fabs(orc_mat_det9(h));
  double KE = orc_ke_total(vel, mass, na);
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.05, 0.05);
  double dP = orc_inst_P(vir9, KE, V) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.005, 0.005);
  double eps = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_nh = exp(-hdt * npt->xi), sc_pr = exp(-hdt * eps / 3.0), sc_v = sc_nh * sc_pr;
  for (int i = 0; i < na; i++) {
    double mi = CONV / mass[i];
    vel[i * 3] = vel[i * 3] * sc_v + hdt * F[i * 3] * mi;
    vel[i * 3 + 1] = vel[i * 3 + 1] * sc_v + hdt * F[i * 3 + 1] * mi;
    vel[i * 3 + 2] = vel[i * 3 + 2] * sc_v + hdt * F[i * 3 + 2] * mi;
  }
  for (int i = 0; i < na; i++) {
    double px = pos[i * 3], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
    double vx = vel[i * 3], vy = vel[i * 3 + 1], vz = vel[i * 3 + 2];
    double sx = hi[0] * px + hi[1] * py + hi[2] * pz, sy = hi[3] * px + hi[4] * py + hi[5] * pz,
           sz = hi[6] * px + hi[7] * py + hi[8] * pz;
    double vsx = hi[0] * vx + hi[1] * vy + hi[2] * vz, vsy = hi[3] * vx + hi[4] * vy + hi[5] * vz,
           vsz = hi[6] * vx + hi[7] * vy + hi[8] * vz;
    sx += dt * vsx;
    sy += dt * vsy;
    sz += dt * vsz;
    sx -= floor(sx);
    sy -= floor(sy);
    sz -= floor(sz);
    pos[i * 3] = sx;
    pos[i * 3 + 1] = sy;
    pos[i * 3 + 2] = sz;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) h[a * 3 + b] += dt * npt->Vg[a * 3 + b];
  orc_mat_inv9(h, hi);
  for (int i = 0; i < na; i++) {
    double sx = pos[i * 3], sy = pos[i * 3 + 1], sz = pos[i * 3 + 2];
    pos[i * 3] = h[0] * sx + h[1] * sy + h[2] * sz;
    pos[i * 3 + 1] = h[3] * sx + h[4] * sy + h[5] * sz;
    pos[i * 3 + 2] = h[6] * sx + h[7] * sy + h[8] * sz;
  }
  double fr[3];
  orc_airebo_forces(F, vir9, pos, h, hi, nlc_r, nll_r, max_r, nlc_l, nll_l, max_l, na,
                    rebo_mode, p, fr);
  double eps2 = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_v2 = sc_nh * exp(-hdt * eps2 / 3.0);
  for (int i = 0; i < na; i++) {
    double mi = CONV / mass[i];
    vel[i * 3] = (vel[i * 3] + hdt * F[i * 3] * mi) * sc_v2;
    vel[i * 3 + 1] = (vel[i * 3 + 1] + hdt * F[i * 3 + 1] * mi) * sc_v2;
    vel[i * 3 + 2] = (vel[i * 3 + 2] + hdt * F[i * 3 + 2] * mi) * sc_v2;
  }
  KE = orc_ke_total(vel, mass, na);
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.05, 0.05);
  double V2 = fabs(orc_mat_det9(h));
  dP = orc_inst_P(vir9, KE, V2) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V2 * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.005, 0.005);
  out[0] = fr[0];
  out[1] = fr[1];
  out[2] = fr[2];
  out[3] = KE;
}

/* --------------------------------------------------- molecular mechanics */
/* MM = src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp.
 *
 * Modes: ORC_MM_COMPAT reproduces the reference forces term by term.  They
 * are not the gradient of the reference energy: the bond and angle forces
 * carry the opposite sign (F = +dE/dr, MM:391 and MM:468), and the
 * middle-atom dihedral/improper forces use the dot-product coefficients with
 * the opposite sign (MM:528-529 and MM:653, 672; b1 = r2 - r1 there, where the
 * analytic formula it follows has r1 - r2).  ORC_MM_PHYSICAL uses the exact
 * gradient F = -dE/dr of the same energy, and the full virial tensor
 * W = sum over terms of (relative position) (x) force, bonded terms included.
 *
 * Pinned: golden vectors (compat), term-by-term finite differences
 * (physical; compat bond/angle equal +gradient), strain derivative of the
 * energy (physical virial), Newton's third law per term. */

/* sigma 3.431 A, eps 2.635 meV, rc = 3 sigma, shifted (MM:111-117). */
void orc_lj_params_mm(orc_lj_params *p) {
  const double sigma = 3.431, eps = 2.635e-3;
  const double rc = 3.0 * sigma;
  const double sr = 1.0 / 3.0, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
  p->sig2 = sigma * sigma;
  p->eps = eps;
  p->rc2 = rc * rc;
  p->vshift = 4.0 * eps * (sr6 * sr6 - sr6);
}

/* Adjacency of build_flat_topology (MM:256-257): bonds in the given order,
 * each appending to both ends.  off[natom+1], adj[2 nbond]. */
static void mm_adjacency(int natom, const int *bonds, int nbond, int *off, int *adj) {
  int i;
  for (i = 0; i <= natom; i++) off[i] = 0;
  for (i = 0; i < nbond; i++) {
    off[bonds[2 * i] + 1]++;
    off[bonds[2 * i + 1] + 1]++;
  }
  for (i = 0; i < natom; i++) off[i + 1] += off[i];
  {
    int fill[4096];
    for (i = 0; i < natom; i++) fill[i] = off[i];
    for (i = 0; i < nbond; i++) {
      int bi = bonds[2 * i], bj = bonds[2 * i + 1];
      adj[fill[bi]++] = bj;
      adj[fill[bj]++] = bi;
    }
  }
}

void orc_mm_topology_counts(int natom, const int *bonds, int nbond, int *counts4) {
  int off[4097], adj[8192];
  mm_adjacency(natom, bonds, nbond, off, adj);
  int nang = 0, ndih = 0, nimp = 0;
  for (int j = 0; j < natom; j++) {
    int dg = off[j + 1] - off[j];
    nang += dg * (dg - 1) / 2;
    if (dg == 3) nimp++;
  }
  for (int b = 0; b < nbond; b++) {
    int bj = bonds[2 * b], bk = bonds[2 * b + 1];
    for (int x = off[bj]; x < off[bj + 1]; x++) {
      int i = adj[x];
      if (i == bk) continue;
      for (int y = off[bk]; y < off[bk + 1]; y++) {
        int l = adj[y];
        if (l == bj || l == i) continue;
        ndih++;
      }
    }
  }
  counts4[0] = nbond;
  counts4[1] = nang;
  counts4[2] = ndih;
  counts4[3] = nimp;
}

/* compute_phi0 MM:240-252 */
static double mm_phi0(const double *c, int i, int j, int k, int l) {
  double b1x = c[j * 3] - c[i * 3], b1y = c[j * 3 + 1] - c[i * 3 + 1],
         b1z = c[j * 3 + 2] - c[i * 3 + 2];
  double b2x = c[k * 3] - c[j * 3], b2y = c[k * 3 + 1] - c[j * 3 + 1],
         b2z = c[k * 3 + 2] - c[j * 3 + 2];
  double b3x = c[l * 3] - c[k * 3], b3y = c[l * 3 + 1] - c[k * 3 + 1],
         b3z = c[l * 3 + 2] - c[k * 3 + 2];
  double mx = b1y * b2z - b1z * b2y, my = b1z * b2x - b1x * b2z, mz = b1x * b2y - b1y * b2x;
  double nx = b2y * b3z - b2z * b3y, ny = b2z * b3x - b2x * b3z, nz = b2x * b3y - b2y * b3x;
  double mm = mx * mx + my * my + mz * mz, nn = nx * nx + ny * ny + nz * nz;
  if (mm < 1e-20 || nn < 1e-20) return 0;
  double b2len = sqrt(b2x * b2x + b2y * b2y + b2z * b2z);
  double cosphi = (mx * nx + my * ny + mz * nz) / sqrt(mm * nn);
  if (cosphi > 1.0) cosphi = 1.0;
  if (cosphi < -1.0) cosphi = -1.0;
  double sinphi = (mx * b3x + my * b3y + mz * b3z) * b2len / sqrt(mm * nn);
  return atan2(sinphi, cosphi);
}

/* build_flat_topology MM:254-334: per-molecule terms (bonds in the given
 * order with r0 from the geometry; angles i-j-k over neighbour pairs of each
 * centre j; dihedrals i-bj-bk-l over each bond; impropers centre-first for
 * three-coordinated atoms, gamma = 2 phi0 + pi, multiplicity 2), replicated
 * over nmol molecules of natom atoms.  k4 = {kb, kth, v2, kimp}. */
void orc_mm_topology(const double *coords, int natom, const int *bonds, int nbond, int nmol,
                     const double *k4, int *b, double *bp, int *a, double *ap, int *d,
                     double *dp, int *dm, int *im, double *ip, int *mol_id) {
  int off[4097], adj[8192], cnt[4];
  mm_adjacency(natom, bonds, nbond, off, adj);
  orc_mm_topology_counts(natom, bonds, nbond, cnt);
  const double PI = 3.14159265358979323846;
  for (int m = 0; m < nmol; m++) {
    const int o = m * natom;
    int nb = 0, na_ = 0, nd = 0, ni = 0;
    for (int t = 0; t < nbond; t++) {
      int bi = bonds[2 * t], bj = bonds[2 * t + 1];
      double dx = coords[bi * 3] - coords[bj * 3], dy = coords[bi * 3 + 1] - coords[bj * 3 + 1],
             dz = coords[bi * 3 + 2] - coords[bj * 3 + 2];
      int g = m * cnt[0] + nb++;
      b[2 * g] = bi + o;
      b[2 * g + 1] = bj + o;
      bp[2 * g] = k4[0];
      bp[2 * g + 1] = sqrt(dx * dx + dy * dy + dz * dz);
    }
    for (int j = 0; j < natom; j++) {
      for (int ii = off[j]; ii < off[j + 1]; ii++)
        for (int kk = ii + 1; kk < off[j + 1]; kk++) {
          int i = adj[ii], k = adj[kk];
          double rji[3], rjk[3];
          for (int c = 0; c < 3; c++) {
            rji[c] = coords[i * 3 + c] - coords[j * 3 + c];
            rjk[c] = coords[k * 3 + c] - coords[j * 3 + c];
          }
          double dji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
          double djk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
          double cs = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (dji * djk);
          cs = clampd(cs, -1.0, 1.0);
          int g = m * cnt[1] + na_++;
          a[3 * g] = i + o;
          a[3 * g + 1] = j + o;
          a[3 * g + 2] = k + o;
          ap[2 * g] = k4[1];
          ap[2 * g + 1] = acos(cs);
        }
    }
    for (int t = 0; t < nbond; t++) {
      int bj = bonds[2 * t], bk = bonds[2 * t + 1];
      for (int x = off[bj]; x < off[bj + 1]; x++) {
        int i = adj[x];
        if (i == bk) continue;
        for (int y = off[bk]; y < off[bk + 1]; y++) {
          int l = adj[y];
          if (l == bj || l == i) continue;
          int g = m * cnt[2] + nd++;
          d[4 * g] = i + o;
          d[4 * g + 1] = bj + o;
          d[4 * g + 2] = bk + o;
          d[4 * g + 3] = l + o;
          dp[2 * g] = k4[2];
          dp[2 * g + 1] = 2 * mm_phi0(coords, i, bj, bk, l) + PI;
          dm[g] = 2;
        }
      }
    }
    for (int i = 0; i < natom; i++) {
      if (off[i + 1] - off[i] != 3) continue;
      int j = adj[off[i]], k = adj[off[i] + 1], l = adj[off[i] + 2];
      int g = m * cnt[3] + ni++;
      im[4 * g] = i + o;
      im[4 * g + 1] = j + o;
      im[4 * g + 2] = k + o;
      im[4 * g + 3] = l + o;
      ip[2 * g] = k4[3];
      ip[2 * g + 1] = 2 * mm_phi0(coords, j, i, k, l) + PI;
    }
    for (int x = 0; x < natom; x++) mol_id[o + x] = m;
  }
}

/* One torsion (dihedral MM:494-603 / improper MM:604-710) over atoms
 * p0-p1-p2-p3 with E = Vn/2 (1 + cos(mult phi - gamma)).  Adds the four
 * forces into F and (physical) the virial; returns the energy.  Skipped
 * (0) when either cross product is degenerate. */
static double mm_torsion(double *F, double *W, const double *pos, const double *h,
                         const double *hi, int p0, int p1, int p2, int p3, double Vn,
                         int mult, double gamma, int mode) {
  double b1[3], b2[3], b3[3];
  for (int c = 0; c < 3; c++) {
    b1[c] = pos[p1 * 3 + c] - pos[p0 * 3 + c];
    b2[c] = pos[p2 * 3 + c] - pos[p1 * 3 + c];
    b3[c] = pos[p3 * 3 + c] - pos[p2 * 3 + c];
  }
  orc_mimg(b1, hi, h);
  orc_mimg(b2, hi, h);
  orc_mimg(b3, hi, h);
  double mx = b1[1] * b2[2] - b1[2] * b2[1], my = b1[2] * b2[0] - b1[0] * b2[2],
         mz = b1[0] * b2[1] - b1[1] * b2[0];
  double nx = b2[1] * b3[2] - b2[2] * b3[1], ny = b2[2] * b3[0] - b2[0] * b3[2],
         nz = b2[0] * b3[1] - b2[1] * b3[0];
  double mm = mx * mx + my * my + mz * mz, nn = nx * nx + ny * ny + nz * nz;
  if (mm < 1e-20 || nn < 1e-20) return 0;
  double imm = 1.0 / mm, inn = 1.0 / nn;
  double b2len = sqrt(b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2]);
  double cosphi = (mx * nx + my * ny + mz * nz) / sqrt(mm * nn);
  if (cosphi > 1.0) cosphi = 1.0;
  if (cosphi < -1.0) cosphi = -1.0;
  double b1b2 = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
  double b2b3 = b2[0] * b3[0] + b2[1] * b3[1] + b2[2] * b3[2];
  double sinphi = (mx * b3[0] + my * b3[1] + mz * b3[2]) * b2len / sqrt(mm * nn);
  double phi = atan2(sinphi, cosphi);
  double E = 0.5 * Vn * (1.0 + cos(mult * phi - gamma));
  double dphi = -0.5 * Vn * mult * sin(mult * phi - gamma);
  double f1[3] = {dphi * b2len * imm * mx, dphi * b2len * imm * my, dphi * b2len * imm * mz};
  double f4[3] = {-dphi * b2len * inn * nx, -dphi * b2len * inn * ny,
                  -dphi * b2len * inn * nz};
  double ci = b1b2 / (b2len * b2len), ck = b2b3 / (b2len * b2len);
  double f2[3], f3[3];
  for (int c = 0; c < 3; c++) {
    if (mode == ORC_MM_PHYSICAL) {
      f2[c] = -f1[c] - ci * f1[c] + ck * f4[c];
      f3[c] = -f4[c] - ck * f4[c] + ci * f1[c];
    } else {
      f2[c] = -f1[c] + ci * f1[c] - ck * f4[c];
      f3[c] = -f4[c] - ci * f1[c] + ck * f4[c];
    }
  }
  for (int c = 0; c < 3; c++) {
    F[p0 * 3 + c] += f1[c];
    F[p1 * 3 + c] += f2[c];
    F[p2 * 3 + c] += f3[c];
    F[p3 * 3 + c] += f4[c];
  }
  if (mode == ORC_MM_PHYSICAL) {
    /* positions relative to p0: b1, b1 + b2, b1 + b2 + b3 */
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++)
        W[3 * x + y] += b1[x] * f2[y] + (b1[x] + b2[x]) * f3[y] +
                        (b1[x] + b2[x] + b3[x]) * f4[y];
  }
  return E;
}

/* compute_forces MM:351-788.  out7 = {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot}.
 * The LJ half list has stride max_l (400 in the reference).  Compat: the
 * virial holds only the diagonal bond and LJ parts, as the reference's. */
void orc_mm_forces(double *F, double *vir9, const double *pos, const double *h,
                   const double *hi, const orc_mm_topo *t, const int *nlc, const int *nll,
                   int max_l, const double *charge, const int *mol_id, int has_charge, int na,
                   int mode, double *out7) {
  const int phys = mode == ORC_MM_PHYSICAL;
  orc_lj_params lj;
  orc_lj_params_mm(&lj);
  for (int i = 0; i < na * 3; i++) F[i] = 0;
  for (int i = 0; i < 9; i++) vir9[i] = 0;
  double Eb = 0, Ea = 0, Ed = 0, Ei = 0, Elj = 0, Ec = 0;
  /* bonds MM:378-448 */
  for (int q = 0; q < t->Nb; q++) {
    int i = t->b[2 * q], j = t->b[2 * q + 1];
    double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                   pos[j * 3 + 2] - pos[i * 3 + 2]};
    orc_mimg(d, hi, h);
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (r < 1e-10) continue;
    double kb = t->bp[2 * q], dr = r - t->bp[2 * q + 1];
    Eb += 0.5 * kb * dr * dr;
    double fm = phys ? kb * dr / r : -kb * dr / r; /* force on i along d */
    double f[3] = {fm * d[0], fm * d[1], fm * d[2]};
    for (int c = 0; c < 3; c++) {
      F[i * 3 + c] += f[c];
      F[j * 3 + c] -= f[c];
    }
    if (phys) {
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) vir9[3 * x + y] -= d[x] * f[y];
    } else {
      vir9[0] += d[0] * f[0];
      vir9[4] += d[1] * f[1];
      vir9[8] += d[2] * f[2];
    }
  }
  /* angles MM:449-493 */
  for (int q = 0; q < t->Nang; q++) {
    int i = t->a[3 * q], j = t->a[3 * q + 1], k = t->a[3 * q + 2];
    double rji[3], rjk[3];
    for (int c = 0; c < 3; c++) {
      rji[c] = pos[i * 3 + c] - pos[j * 3 + c];
      rjk[c] = pos[k * 3 + c] - pos[j * 3 + c];
    }
    orc_mimg(rji, hi, h);
    orc_mimg(rjk, hi, h);
    double dji = sqrt(rji[0] * rji[0] + rji[1] * rji[1] + rji[2] * rji[2]);
    double djk = sqrt(rjk[0] * rjk[0] + rjk[1] * rjk[1] + rjk[2] * rjk[2]);
    if (dji < 1e-10 || djk < 1e-10) continue;
    double cs = (rji[0] * rjk[0] + rji[1] * rjk[1] + rji[2] * rjk[2]) / (dji * djk);
    if (cs > 0.999999) cs = 0.999999;
    if (cs < -0.999999) cs = -0.999999;
    double kth = t->ap[2 * q], dth = acos(cs) - t->ap[2 * q + 1];
    Ea += 0.5 * kth * dth * dth;
    double sn = sqrt(1.0 - cs * cs + 1e-30);
    double dV = phys ? kth * dth / sn : -kth * dth / sn;
    double fi[3], fk[3];
    for (int c = 0; c < 3; c++) {
      fi[c] = dV * (rjk[c] / (dji * djk) - cs * rji[c] / (dji * dji));
      fk[c] = dV * (rji[c] / (dji * djk) - cs * rjk[c] / (djk * djk));
      F[i * 3 + c] += fi[c];
      F[k * 3 + c] += fk[c];
      F[j * 3 + c] -= fi[c] + fk[c];
    }
    if (phys)
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++) vir9[3 * x + y] += rji[x] * fi[y] + rjk[x] * fk[y];
  }
  /* dihedrals MM:494-603 */
  double Wt[9] = {0};
  for (int q = 0; q < t->Ndih; q++)
    Ed += mm_torsion(F, Wt, pos, h, hi, t->d[4 * q], t->d[4 * q + 1], t->d[4 * q + 2],
                     t->d[4 * q + 3], t->dp[2 * q], t->dm[q], t->dp[2 * q + 1], mode);
  /* impropers MM:604-710: atom order c1-c0-c2-c3, multiplicity 2 */
  for (int q = 0; q < t->Nimp; q++)
    Ei += mm_torsion(F, Wt, pos, h, hi, t->im[4 * q + 1], t->im[4 * q], t->im[4 * q + 2],
                     t->im[4 * q + 3], t->ip[2 * q], 2, t->ip[2 * q + 1], mode);
  if (phys)
    for (int x = 0; x < 9; x++) vir9[x] += Wt[x];
  /* LJ MM:711-765 */
  for (int i = 0; i < na; i++) {
    for (int jn = 0; jn < nlc[i]; jn++) {
      int j = nll[i * max_l + jn];
      double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                     pos[j * 3 + 2] - pos[i * 3 + 2]};
      orc_mimg(d, hi, h);
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (r2 > lj.rc2) continue;
      if (r2 < 0.25) r2 = 0.25;
      double ri2 = 1.0 / r2, sr2 = lj.sig2 * ri2, sr6 = sr2 * sr2 * sr2, sr12 = sr6 * sr6;
      double fm = 24 * lj.eps * (2 * sr12 - sr6) * ri2;
      Elj += 4 * lj.eps * (sr12 - sr6) - lj.vshift;
      for (int c = 0; c < 3; c++) {
        F[i * 3 + c] -= fm * d[c];
        F[j * 3 + c] += fm * d[c];
      }
      if (phys) {
        for (int x = 0; x < 3; x++)
          for (int y = 0; y < 3; y++) vir9[3 * x + y] += d[x] * fm * d[y];
      } else {
        vir9[0] += d[0] * fm * d[0];
        vir9[4] += d[1] * fm * d[1];
        vir9[8] += d[2] * fm * d[2];
      }
    }
  }
  /* Coulomb MM:766-787 (all pairs of different molecules within 3 sigma) */
  if (has_charge) {
    const double KC = 14.3996, rc = 3.0 * 3.431;
    for (int i = 0; i < na - 1; i++)
      for (int j = i + 1; j < na; j++) {
        if (mol_id[i] == mol_id[j]) continue;
        double d[3] = {pos[j * 3] - pos[i * 3], pos[j * 3 + 1] - pos[i * 3 + 1],
                       pos[j * 3 + 2] - pos[i * 3 + 2]};
        orc_mimg(d, hi, h);
        double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (r2 > lj.rc2) continue;
        double r = sqrt(r2), qq = KC * charge[i] * charge[j];
        Ec += qq / r - qq / rc;
        double fm = qq / (r2 * r);
        for (int c = 0; c < 3; c++) {
          F[i * 3 + c] -= fm * d[c];
          F[j * 3 + c] += fm * d[c];
        }
        if (phys)
          for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) vir9[3 * x + y] += d[x] * fm * d[y];
      }
  }
  out7[0] = Eb;
  out7[1] = Ea;
  out7[2] = Ed;
  out7[3] = Ei;
  out7[4] = Elj;
  out7[5] = Ec;
  out7[6] = Eb + Ea + Ed + Ei + Elj + Ec;
}

/* step_npt MM:812-890 (carbon mass for every atom).  out8 = out7 + {KE}. */
void orc_mm_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                     double *hi, int na, double dt, orc_npt *npt, const orc_mm_topo *t,
                     const int *nlc, const int *nll, int max_l, const double *charge,
                     const int *mol_id, int has_charge, int mode, double *out8) {
  const double mC = 12.011, mi = CONV / mC;
  double hdt = 0.5 * dt, V = fabs(orc_mat_det9(h));
  double s = 0;
  for (int i = 0; i < na; i++)
    s += mC * (vel[i * 3] * vel[i * 3] + vel[i * 3 + 1] * vel[i * 3 + 1] +
               vel[i * 3 + 2] * vel[i * 3 + 2]);
  double KE = 0.5 * s / CONV;
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.05, 0.05);
  double dP = orc_inst_P(vir9, KE, V) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.005, 0.005);
  double eps = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_nh = exp(-hdt * npt->xi), sc_pr = exp(-hdt * eps / 3.0), sc_v = sc_nh * sc_pr;
  for (int i = 0; i < na * 3; i++) vel[i] = vel[i] * sc_v + hdt * F[i] * mi;
  for (int i = 0; i < na; i++) {
    double px = pos[i * 3], py = pos[i * 3 + 1], pz = pos[i * 3 + 2];
    double vx = vel[i * 3], vy = vel[i * 3 + 1], vz = vel[i * 3 + 2];
    double sx = hi[0] * px + hi[1] * py + hi[2] * pz, sy = hi[3] * px + hi[4] * py + hi[5] * pz,
           sz = hi[6] * px + hi[7] * py + hi[8] * pz;
    double vsx = hi[0] * vx + hi[1] * vy + hi[2] * vz, vsy = hi[3] * vx + hi[4] * vy + hi[5] * vz,
           vsz = hi[6] * vx + hi[7] * vy + hi[8] * vz;
    sx += dt * vsx;
    sy += dt * vsy;
    sz += dt * vsz;
    sx -= floor(sx);
    sy -= floor(sy);
    sz -= floor(sz);
    pos[i * 3] = sx;
    pos[i * 3 + 1] = sy;
    pos[i * 3 + 2] = sz;
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) h[a * 3 + b] += dt * npt->Vg[a * 3 + b];
  orc_mat_inv9(h, hi);
  for (int i = 0; i < na; i++) {
    double sx = pos[i * 3], sy = pos[i * 3 + 1], sz = pos[i * 3 + 2];
    pos[i * 3] = h[0] * sx + h[1] * sy + h[2] * sz;
    pos[i * 3 + 1] = h[3] * sx + h[4] * sy + h[5] * sz;
    pos[i * 3 + 2] = h[6] * sx + h[7] * sy + h[8] * sz;
  }
  orc_mm_forces(F, vir9, pos, h, hi, t, nlc, nll, max_l, charge, mol_id, has_charge, na, mode,
                out8);
  double eps2 = npt->Vg[0] * hi[0] + npt->Vg[4] * hi[4] + npt->Vg[8] * hi[8];
  double sc_v2 = sc_nh * exp(-hdt * eps2 / 3.0);
  for (int i = 0; i < na * 3; i++) vel[i] = (vel[i] + hdt * F[i] * mi) * sc_v2;
  s = 0;
  for (int i = 0; i < na; i++)
    s += mC * (vel[i * 3] * vel[i * 3] + vel[i * 3 + 1] * vel[i * 3 + 1] +
               vel[i * 3 + 2] * vel[i * 3 + 2]);
  KE = 0.5 * s / CONV;
  npt->xi += hdt * (2 * KE - npt->Nf * KB * npt->Tt) / npt->Q;
  npt->xi = clampd(npt->xi, -0.05, 0.05);
  double V2 = fabs(orc_mat_det9(h));
  dP = orc_inst_P(vir9, KE, V2) - npt->Pe;
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] += hdt * V2 * dP / (npt->W * EV2GPA);
  for (int a = 0; a < 3; a++) npt->Vg[a * 4] = clampd(npt->Vg[a * 4], -0.005, 0.005);
  out8[7] = KE;
}
