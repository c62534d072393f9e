// fmd_rebo.hpp -- REBO C-C scalar functions (AB:163-191) templated on the
// working precision, and the per-bond / per-angle terms the REBO kernels are
// built from.  R = float on the GPU; the same code with R = double is the
// reference arithmetic.
#pragma once

#include "fmd_common.hpp"

namespace fmd {
namespace rebo {

// REBO-II C-C pair parameters and the Brenner-1990 angular function (AB:110-116).
constexpr double kQ = 0.3134602960833, kA = 10953.544162170, kAlpha = 4.7465390606595;
constexpr double kB1 = 12388.79197798, kBeta1 = 4.7204523127;
constexpr double kB2 = 17.56740646509, kBeta2 = 1.4332132499;
constexpr double kB3 = 30.71493208065, kBeta3 = 1.3826912506;
constexpr double kDmin = 1.7, kDmax = 2.0, kDelta = 0.5;
constexpr double kGa0 = 0.00020813, kGc0 = 330.0, kGd0 = 3.5;
constexpr double kRlist = kDmax + 0.3;  // REBO_RCUT, AB:116
constexpr double kPi = 3.14159265358979323846;

FMD_HD float ex(float x) { return expf(x); }
FMD_HD double ex(double x) { return exp(x); }
FMD_HD float co(float x) { return cosf(x); }
FMD_HD double co(double x) { return cos(x); }
FMD_HD float si(float x) { return sinf(x); }
FMD_HD double si(double x) { return sin(x); }
FMD_HD float rsq(float x) { return 1.0f / sqrtf(x); }
FMD_HD double rsq(double x) { return 1.0 / sqrt(x); }

// The FP32 terms below are evaluated without FMA contraction, so every
// place they are inlined into rounds them the same way and the REBO kernel's
// two paths agree bit for bit.
//
// Cosine switch fc and its derivative (AB:165-171).
template <class R>
FMD_HD R fc(R r) {
#pragma clang fp contract(off)
  if (r <= R(kDmin)) return R(1);
  if (r >= R(kDmax)) return R(0);
  return R(0.5) * (R(1) + co(R(kPi) * (r - R(kDmin)) / R(kDmax - kDmin)));
}
template <class R>
FMD_HD R dfc(R r) {
#pragma clang fp contract(off)
  if (r <= R(kDmin) || r >= R(kDmax)) return R(0);
  return R(-0.5 * kPi / (kDmax - kDmin)) * si(R(kPi) * (r - R(kDmin)) / R(kDmax - kDmin));
}

// Repulsive / attractive pair terms and derivatives (AB:173-183).
template <class R>
FMD_HD void pair_terms(R r, R &vr, R &dvr, R &va, R &dva) {
  R ex0 = R(kA) * ex(R(-kAlpha) * r);
  R qr = R(kQ) / r;
  vr = (R(1) + qr) * ex0;
  dvr = -(qr / r) * ex0 - (R(1) + qr) * R(kAlpha) * ex0;
  R e1 = R(kB1) * ex(R(-kBeta1) * r);
  R e2 = R(kB2) * ex(R(-kBeta2) * r);
  R e3 = R(kB3) * ex(R(-kBeta3) * r);
  va = e1 + e2 + e3;
  dva = R(-kBeta1) * e1 - R(kBeta2) * e2 - R(kBeta3) * e3;
}

// G(cos) and dG/dcos (AB:185-191).  The reference form
// a0 (1 + c^2/d^2 - c^2/(d^2+h^2)) subtracts two terms near 8.8e3 to leave
// ~1.8e2; the equal form a0 (1 + c^2 h^2 / (d^2 (d^2+h^2))) has no
// cancellation, which keeps FP32 at ~1e-7 relative.
template <class R>
FMD_HD R G(R x) {
#pragma clang fp contract(off)
  const R c2 = R(kGc0 * kGc0), d2 = R(kGd0 * kGd0);
  R hh = R(1) + x, h2 = hh * hh;
  return R(kGa0) * (R(1) + c2 * h2 / (d2 * (d2 + h2)));
}
template <class R>
FMD_HD R dG(R x) {
#pragma clang fp contract(off)
  const R c2 = R(kGc0 * kGc0), d2 = R(kGd0 * kGd0);
  R hh = R(1) + x, dn = d2 + hh * hh;
  return R(kGa0 * 2.0) * c2 * hh / (dn * dn);
}

// The REBO-II P_ij correction of a carbon bond order, P_CC(N, N_H = 0): a
// cubic Hermite through v[n] at N = n (n = 0..4) with zero slope at every
// knot, v[4] beyond (Brenner et al. 2002; the AIREBO carbon values have only
// v[2] nonzero).  No counterpart in the reference, whose REBO has no P term;
// all zero (the default) leaves its bond order unchanged.
struct ReboPcc {
  float v[5];
  int on;  // any v nonzero (0: no P term, the reference's bond order)
};
template <class R>
FMD_HD R pcc(const ReboPcc &pc, R N, R &dP) {
#pragma clang fp contract(off)
  if (!(N > R(0))) {
    dP = R(0);
    return R(pc.v[0]);
  }
  if (N >= R(4)) {
    dP = R(0);
    return R(pc.v[4]);
  }
  R lo = R(pc.v[3]), hi = R(pc.v[4]), n0 = R(3);
  if (N < R(1)) {
    lo = R(pc.v[0]);
    hi = R(pc.v[1]);
    n0 = R(0);
  } else if (N < R(2)) {
    lo = R(pc.v[1]);
    hi = R(pc.v[2]);
    n0 = R(1);
  } else if (N < R(3)) {
    lo = R(pc.v[2]);
    hi = R(pc.v[3]);
    n0 = R(2);
  }
  const R t = N - n0, dv = hi - lo;
  dP = dv * R(6) * t * (R(1) - t);
  return lo + dv * t * t * (R(3) - R(2) * t);
}

template <class R>
FMD_HD R clamp1(R x) {
  return x < R(-1) ? R(-1) : (x > R(1) ? R(1) : x);
}

// Angular term of the bond order of bond c->p at centre c with third atom t
// (vectors from c).  coeff = 1/2 fc(r_cp) VA(r_cp) db/dGs.  Physical forces:
// ft on t, fp on p, the centre gets -(ft + fp).  The reference's force on t
// is -ft, with no fp term (AB:360-419).  dP: dP/dN of the bond's P_ij term
// (N sums fc over c's other bonds, so t's switch carries it too; 0 without
// the term).
template <class R>
FMD_HD void angular(const R *rcp, R r_cp, const R *rct, R r_ct, R fc_ct, R dfc_ct, R coeff,
                    R dP, R *ft, R *fp) {
#pragma clang fp contract(off)
  R inv_ct = R(1) / r_ct, inv_cp = R(1) / r_cp;
  R uct[3] = {rct[0] * inv_ct, rct[1] * inv_ct, rct[2] * inv_ct};
  R ucp[3] = {rcp[0] * inv_cp, rcp[1] * inv_cp, rcp[2] * inv_cp};
  R cs = clamp1((rcp[0] * rct[0] + rcp[1] * rct[1] + rcp[2] * rct[2]) * (inv_cp * inv_ct));
  R g = G(cs), dg = dG(cs);
  R a = coeff * dfc_ct * (g + dP), b = coeff * fc_ct * dg;
  for (int k = 0; k < 3; k++) {
    ft[k] = a * uct[k] + b * (ucp[k] - cs * uct[k]) * inv_ct;
    fp[k] = b * (uct[k] - cs * ucp[k]) * inv_cp;
  }
}

}  // namespace rebo
}  // namespace fmd
