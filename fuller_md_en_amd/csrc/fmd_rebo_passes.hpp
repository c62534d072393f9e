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
//   ba[4*slot]  fc dfc coeff dP FP32 (coeff = 1/2 fc VA db/dGs, dP = dP_ij/dN)
//   bp[4*slot]  P = bond_terms  FP64 pair quantities (f VR, f VA, C, D)
//   bq[4*slot]  r fc dfc b      FP64 (b = (1 + Gs + P)^-1/2)
//   bflag[slot] kBond* bits,  brev[slot] slot of i in j's list
// The pair force is the difference of the repulsive and attractive
// derivatives, each ~77 eV/A at the C60 bond length against a ~1 eV/A
// result; FP32 rounding of r alone moves exp(-alpha r) by ~4e-7 relative, so
// the pair terms and their combination are FP64 (4 exponentials per bond).
// Bond-order sums and angular terms are FP32.
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

// FP64 terms of a bond of length rd (the same for both directions): switch
// fd = fc(rd) and dfd, and the pair quantities P = (f VR, f VA, C, D) with
// C = (dfc VR + f VR') / r and D = (dfc VA + f VA') / r, so that the pair
// energy is P.x - bbar P.y and the pair force on i along r_j - r_i is
// (P.z - bbar P.w) per unit of that vector (AB:315-330).  Returns the
// kBondPair / kBondThird bits; 0 (and zeros) beyond the switch.
FMD_HD int bond_terms(double rd, double &fd, double &dfd, double4 &P) {
  using namespace rebo;
  if (rd > kDmax) {
    fd = dfd = 0;
    P = make_double4(0, 0, 0, 0);
    return 0;
  }
  fd = fc(rd);
  dfd = dfc(rd);
  int flag = 0;
  if (!(fd < 1e-15 && dfd == 0.0)) flag |= kBondPair;
  if (!(fd < 1e-15)) flag |= kBondThird;
  double vr, dvr, va, dva;
  pair_terms(rd, vr, dvr, va, dva);
  P = make_double4(fd * vr, fd * va, (dfd * vr + fd * dvr) / rd, (dfd * va + fd * dva) / rd);
  return flag;
}

// Bond-order sum Gs_ij = sum over k in N(i)\{j} of fc(r_ik) G(cos theta_ijk)
// (AB:292-305) from FP32 bond vectors: bv(kn) gives (dx, dy, dz, r) of i's
// entry kn.  Also the P_ij term P(N_ij), N_ij = sum of the same fc(r_ik), and
// its slope dP (both 0 when pc is all zero, the reference's bond order).
template <class R, class BV>
FMD_HD R bond_order_sum(int jn, int nni, R ux, R uy, R uz, R r, BV bv, const rebo::ReboPcc &pc,
                        R &P, R &dP) {
#pragma clang fp contract(off)  // the same rounding wherever it is inlined
  using namespace rebo;
  R Gs = 0, N = 0;
  for (int kn = 0; kn < nni; kn++) {
    if (kn == jn) continue;
    const float4 e = bv(kn);
    if (e.w > R(kDmax)) continue;
    const R rk = e.w;
    const R fk = fc(rk);
    if (fk < R(1e-15)) continue;
    R cs = clamp1((ux * (R)e.x + uy * (R)e.y + uz * (R)e.z) / (r * rk));
    Gs += fk * G(cs);
    N += fk;
  }
  if (pc.on) {
    P = rebo::pcc(pc, N, dP);
  } else {
    P = dP = R(0);
  }
  return Gs;
}

// d b / d Gs = -1/2 (1 + Gs)^-3/2 and the angular coefficient
// 1/2 fc VA db/dGs of a direction with bond order b (P.y = fc VA).
FMD_HD float angular_coeff(double Py, double b) {
  const double dbp = -rebo::kDelta * b * b * b;
  return (float)(0.5 * Py * dbp);
}

// Virial share of an angular triple, d_ct (x) f_t + d_cp (x) f_p (TRACE: its
// trace into W[0]).  Products of FP32 numbers are exact in FP64, so the sums
// round the same way wherever this is inlined.
template <bool TRACE>
FMD_HD void triple_virial(const float *dt, const float *ft, const float *dp, const float *fp,
                          double *W) {
  if (TRACE) {
    W[0] += (((double)dt[0] * ft[0] + (double)dt[1] * ft[1]) + (double)dt[2] * ft[2]) +
            (((double)dp[0] * fp[0] + (double)dp[1] * fp[1]) + (double)dp[2] * fp[2]);
  } else {
    for (int a = 0; a < 3; a++)
      for (int c = 0; c < 3; c++) W[3 * a + c] += (double)dt[a] * ft[c] + (double)dp[a] * fp[c];
  }
}

// Entry (i -> j = nll[slot(i, jn)]); vec[kn] = (dx, dy, dz, r) of i's entry
// kn (minimum image r_k - r_i, FP64), so every bond vector of i is taken once.
template <class R, class Slot>
FMD_HD void rebo_pass1(int i, int jn, const double4 *vec, const int *nlc, const int *nll,
                       Slot slot, ReboBonds B, const rebo::ReboPcc &pc) {
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
  double fd, dfd;
  double4 P;
  int flag = bond_terms(rd, fd, dfd, P);
  if (rd > kDmax) {
    B.bflag[s] = 0;
    return;
  }
  R Pn, dPn;
  const R Gs = bond_order_sum<R>(
      jn, nni, (R)dx, (R)dy, (R)dz, r,
      [&](int kn) {
        return make_float4((float)vec[kn].x, (float)vec[kn].y, (float)vec[kn].z,
                           (float)vec[kn].w);
      },
      pc, Pn, dPn);
  const double b = rsq((1.0 + (double)Gs) + (double)Pn);
  if ((flag & kBondPair) && Gs > R(1e-20) && P.y > 0.0) flag |= kBondAngular;
  B.bflag[s] = flag;
  B.ba[4 * s] = (float)fd;
  B.ba[4 * s + 1] = (float)dfd;
  B.ba[4 * s + 2] = angular_coeff(P.y, b);
  B.ba[4 * s + 3] = (float)dPn;
  B.bp[4 * s] = P.x;
  B.bp[4 * s + 1] = P.y;
  B.bp[4 * s + 2] = P.z;
  B.bp[4 * s + 3] = P.w;
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
    const int rv = B.brev[s];  // -1 only for an asymmetric (overflowed) list
    const double bbar = 0.5 * (B.bq[4 * s + 3] + B.bq[4 * (rv >= 0 ? rv : s) + 3]);
    // pair force (AB:326-330): F_x += fpair d_xy
    const double fpair = B.bp[4 * s + 2] - bbar * B.bp[4 * s + 3];
    for (int k = 0; k < 3; k++) F[k] += fpair * (double)d[k];
    if (x < y) {
      E += B.bp[4 * s] - bbar * B.bp[4 * s + 1];
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
        angular(d, r, dt, (R)B.bd[4 * t + 3], (R)B.ba[4 * t], (R)B.ba[4 * t + 1], coeff,
                (R)B.ba[4 * s + 3], ft, fp);
        if (MODE == kReboCompat) {
          for (int k = 0; k < 3; k++) F[k] += (double)ft[k];
        } else {
          for (int k = 0; k < 3; k++) F[k] -= (double)(ft[k] + fp[k]);
          triple_virial<TRACE>(dt, ft, d, fp, W);
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
      angular(dcp, (R)B.bd[4 * sp + 3], dcx, rcx, fcx, dfcx, (R)B.ba[4 * sp + 2],
              (R)B.ba[4 * sp + 3], ft, fp);
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
              (R)B.ba[4 * sx + 3], ft, fp);
      for (int k = 0; k < 3; k++) F[k] += (double)fp[k];
    }
  }
}

}  // namespace fmd
