// fmd_rebo_passes.hpp -- the REBO bond-order force (compute_rebo, AB:275-421)
// as two owner-computes passes with no atomics:
//
//   pass 1, one thread per directed list entry (i -> j):  bond vector and
//           length, switch, pair terms, the bond-order sum Gs_ij over
//           k in N(i)\{j}, b_ij = (1+Gs)^-1/2, the angular coefficient
//           1/2 fc VA db/dGs, validity flags and the reverse entry (j -> i).
//   pass 2, per atom x, one thread per neighbour entry: every force
//           component that lands on x -- the pair terms of its bonds (b-bar
//           from both directions), the angular terms where x is the centre,
//           the third atom (and, in the physical mode, the partner) -- plus
//           x's share of energy and virial, summed over the entries.
//
// Modes: kReboCompat reproduces the reference exactly (angular force on the
// third atom with the reference's sign, no d cos / d r_partner term, pair-only
// virial); kReboPhysical is the negative gradient of the same energy with the
// full virial.
#pragma once

#include "fmd_rebo.hpp"

namespace fmd {

enum : int { kReboCompat = 0, kReboPhysical = 1 };
enum : int { kBondPair = 1, kBondThird = 2, kBondAngular = 4 };

// Per directed-entry storage (slot of entry jn of atom i, see SlotStride /
// SlotPrefix):
//   bd[4*slot]  dx dy dz r      FP32 (r_j - r_i, minimum image): angular terms
//   ba[4*slot]  fc dfc coeff -  FP32 (coeff = 1/2 fc VA db/dGs)
//   bp[4*slot]  vr dvr va dva   FP64 pair terms
//   bq[4*slot]  r fc dfc b      FP64 (b = (1 + Gs)^-1/2)
//   bflag[slot] kBond* bits,  brev[slot] slot of i in j's list
// The pair force is the difference of the repulsive and attractive
// derivatives, each ~77 eV/A at the C60 bond length against a ~1 eV/A
// result; FP32 rounding of r alone moves exp(-alpha r) by ~4e-7 relative, so
// the pair terms and their combination are FP64 (4 exponentials per directed
// bond).  Bond-order sums and angular terms are FP32.
struct ReboBonds {
  float *bd, *ba;
  double *bp, *bq;
  int *bflag, *brev;
};

// Where entry jn of atom i is stored: i * maxr + jn (the reference's list
// layout) or, for a molecule staged in LDS, a prefix-sum offset + jn.
struct SlotStride {
  int maxr;
  FMD_HD int operator()(int i, int jn) const { return i * maxr + jn; }
};
struct SlotPrefix {
  const int *pre;
  FMD_HD int operator()(int i, int jn) const { return pre[i] + jn; }
};

// Entry (i -> j = nll[slot(i, jn)]); vec[kn] = (dx, dy, dz, r) of i's entry
// kn (minimum image r_k - r_i, FP64), so every bond vector of i is taken once.
template <class R, class Slot>
FMD_HD void rebo_pass1(int i, int jn, const double4 *vec, const int *nlc, const int *nll,
                       Slot slot, ReboBonds B) {
  using namespace rebo;
  const int s = slot(i, jn);
  const int j = nll[s];
  const int nni = nlc[i];
  const double dx = vec[jn].x, dy = vec[jn].y, dz = vec[jn].z, rd = vec[jn].w;
  const R r = (R)rd;
  int rev = -1;
  for (int ln = 0; ln < nlc[j]; ln++)
    if (nll[slot(j, ln)] == i) rev = slot(j, ln);
  B.brev[s] = rev;
  B.bd[4 * s] = (float)dx;
  B.bd[4 * s + 1] = (float)dy;
  B.bd[4 * s + 2] = (float)dz;
  B.bd[4 * s + 3] = (float)rd;
  if (rd > kDmax) {
    B.bflag[s] = 0;
    return;
  }
  const double fd = fc(rd), dfd = dfc(rd);
  int flag = 0;
  if (!(fd < 1e-15 && dfd == 0.0)) flag |= kBondPair;
  if (!(fd < 1e-15)) flag |= kBondThird;
  double vr, dvr, va, dva;
  pair_terms(rd, vr, dvr, va, dva);
  const R ux = (R)dx, uy = (R)dy, uz = (R)dz;
  R Gs = 0;
  for (int kn = 0; kn < nni; kn++) {
    if (kn == jn) continue;
    const double ex = vec[kn].x, ey = vec[kn].y, ez = vec[kn].z, rkd = vec[kn].w;
    if (rkd > kDmax) continue;
    const R rk = (R)rkd;
    const R fk = fc(rk);
    if (fk < R(1e-15)) continue;
    R cs = clamp1((ux * (R)ex + uy * (R)ey + uz * (R)ez) / (r * rk));
    Gs += fk * G(cs);
  }
  const double b = rsq(1.0 + (double)Gs);
  const double dbp = -kDelta * b * b * b;  // d b / d Gs = -1/2 (1 + Gs)^-3/2
  const double coeff = 0.5 * fd * va * dbp;
  if ((flag & kBondPair) && Gs > R(1e-20) && va > 1e-20) flag |= kBondAngular;
  B.bflag[s] = flag;
  B.ba[4 * s] = (float)fd;
  B.ba[4 * s + 1] = (float)dfd;
  B.ba[4 * s + 2] = (float)coeff;
  B.ba[4 * s + 3] = 0.f;
  B.bp[4 * s] = vr;
  B.bp[4 * s + 1] = dvr;
  B.bp[4 * s + 2] = va;
  B.bp[4 * s + 3] = dva;
  B.bq[4 * s] = rd;
  B.bq[4 * s + 1] = fd;
  B.bq[4 * s + 2] = dfd;
  B.bq[4 * s + 3] = b;
}

// Contributions to atom x through its neighbour entry yn (neighbour y = c):
// the pair term of bond (x, y) (energy and virial once, from the lower
// index), the angular terms centred on x with partner y, and the angular
// terms centred on y in which x is the third atom (physical mode: also those
// with x as the partner).  Summed over yn they are every force component
// that lands on x.  F += force, E += energy share, W += virial share
// (row-major; TRACE: W[0] += its trace only).
template <class R, int MODE, bool TRACE, class Slot>
FMD_HD void rebo_pass2_entry(int x, int yn, const int *nlc, const int *nll, Slot slot,
                             ReboBonds B, double *F, double &E, double *W) {
  using namespace rebo;
  const int nx = nlc[x];
  const int s = slot(x, yn);
  const int fl = B.bflag[s];
  if (fl & kBondPair) {
    const int y = nll[s];
    const R d[3] = {B.bd[4 * s], B.bd[4 * s + 1], B.bd[4 * s + 2]};
    const R r = B.bd[4 * s + 3];
    const double rd = B.bq[4 * s], f = B.bq[4 * s + 1], df = B.bq[4 * s + 2];
    const int rv = B.brev[s];  // -1 only for an asymmetric (overflowed) list
    const double bbar = 0.5 * (B.bq[4 * s + 3] + B.bq[4 * (rv >= 0 ? rv : s) + 3]);
    const double vr = B.bp[4 * s], dvr = B.bp[4 * s + 1], va = B.bp[4 * s + 2],
                 dva = B.bp[4 * s + 3];
    // pair force (AB:326-330): F_x += fpair d_xy
    const double fpair = (df * (vr - bbar * va) + f * (dvr - bbar * dva)) / rd;
    for (int k = 0; k < 3; k++) F[k] += fpair * (double)d[k];
    if (x < y) {
      E += f * (vr - bbar * va);
      if (TRACE) {
        W[0] -= fpair * ((double)d[0] * d[0] + (double)d[1] * d[1] + (double)d[2] * d[2]);
      } else {
        for (int a = 0; a < 3; a++)
          for (int c = 0; c < 3; c++) W[3 * a + c] -= (double)d[a] * fpair * (double)d[c];
      }
    }
    // angular terms with centre x, partner y
    if (fl & kBondAngular) {
      const R coeff = B.ba[4 * s + 2];
      for (int tn = 0; tn < nx; tn++) {
        if (tn == yn) continue;
        const int t = slot(x, tn);
        if (!(B.bflag[t] & kBondThird)) continue;
        const R dt[3] = {B.bd[4 * t], B.bd[4 * t + 1], B.bd[4 * t + 2]};
        R ft[3], fp[3];
        angular(d, r, dt, (R)B.bd[4 * t + 3], (R)B.ba[4 * t], (R)B.ba[4 * t + 1], coeff, ft,
                fp);
        if (MODE == kReboCompat) {
          for (int k = 0; k < 3; k++) F[k] += (double)ft[k];
        } else {
          for (int k = 0; k < 3; k++) F[k] -= (double)(ft[k] + fp[k]);
          if (TRACE) {
            W[0] += (double)(dt[0] * ft[0] + dt[1] * ft[1] + dt[2] * ft[2] + d[0] * fp[0] +
                             d[1] * fp[1] + d[2] * fp[2]);
          } else {
            for (int a = 0; a < 3; a++)
              for (int c = 0; c < 3; c++)
                W[3 * a + c] += (double)(dt[a] * ft[c] + d[a] * fp[c]);
          }
        }
      }
    }
  }
  // x as the third atom (and, physical, as the partner) of bonds centred on
  // its neighbour c
  const int sx = B.brev[s];  // entry c -> x
  if (sx < 0) return;
  const int c = nll[s];
  const int fl_cx = B.bflag[sx];
  const R dcx[3] = {B.bd[4 * sx], B.bd[4 * sx + 1], B.bd[4 * sx + 2]};
  const R rcx = B.bd[4 * sx + 3];
  const int nc = nlc[c];
  if (fl_cx & kBondThird) {
    const R fcx = B.ba[4 * sx], dfcx = B.ba[4 * sx + 1];
    for (int pn = 0; pn < nc; pn++) {
      const int sp = slot(c, pn);
      if (sp == sx || !(B.bflag[sp] & kBondAngular)) continue;
      const R dcp[3] = {B.bd[4 * sp], B.bd[4 * sp + 1], B.bd[4 * sp + 2]};
      R ft[3], fp[3];
      angular(dcp, (R)B.bd[4 * sp + 3], dcx, rcx, fcx, dfcx, (R)B.ba[4 * sp + 2], ft, fp);
      if (MODE == kReboCompat)
        for (int k = 0; k < 3; k++) F[k] -= (double)ft[k];
      else
        for (int k = 0; k < 3; k++) F[k] += (double)ft[k];
    }
  }
  if (MODE == kReboPhysical && (fl_cx & kBondAngular)) {
    const R coeff = B.ba[4 * sx + 2];
    for (int tn = 0; tn < nc; tn++) {
      const int st = slot(c, tn);
      if (st == sx || !(B.bflag[st] & kBondThird)) continue;
      const R dct[3] = {B.bd[4 * st], B.bd[4 * st + 1], B.bd[4 * st + 2]};
      R ft[3], fp[3];
      angular(dcx, rcx, dct, (R)B.bd[4 * st + 3], (R)B.ba[4 * st], (R)B.ba[4 * st + 1], coeff,
              ft, fp);
      for (int k = 0; k < 3; k++) F[k] += (double)fp[k];
    }
  }
}

}  // namespace fmd
