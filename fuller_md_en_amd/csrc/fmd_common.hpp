// fmd_common.hpp -- scalar math shared by the gfx950 kernels and the host
// code: FP64 cell-matrix algebra, the fractional minimum image, quaternion
// kinematics and the FP32 shifted-LJ pair term.
//
// Conventions follow the reference (src/fuller_LJ_npt_md_core_serial.cpp,
// "C1"): h is row-major with the cell vectors as its columns (r = h s),
// quaternions are [w,x,y,z], virial W_ab = sum d_a f_b with d the pair vector
// from the first to the second atom and f the force on the second.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define FMD_HD __host__ __device__ __forceinline__
#else
#define FMD_HD inline
#endif

namespace fmd {

// Units A, amu, eV, fs, K, GPa (C1:84-87).
constexpr double kConv = 9.64853321e-3;  // eV fs^2 / (amu A^2)
constexpr double kB = 8.617333262e-5;    // eV / K
constexpr double kEv2GPa = 160.21766208;
constexpr double kEv2KcalMol = 23.06054783;
constexpr double kMassC = 12.011;

// ------------------------------------------------------------ cell algebra
// The cell algebra has the reference's operation order and no FMA
// contraction: h^-1 feeds the minimum image, where a pair exactly half a
// cell apart is decided by the last bit.
FMD_HD double det9(const double *h) {  // C1:137-141
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return h[0] * (h[4] * h[8] - h[5] * h[7]) - h[1] * (h[3] * h[8] - h[5] * h[6]) +
         h[2] * (h[3] * h[7] - h[4] * h[6]);
}

// Smallest perpendicular width of the cell h (columns are the cell vectors).
FMD_HD double min_width(const double *h) {
  const double a[3] = {h[0], h[3], h[6]}, b[3] = {h[1], h[4], h[7]}, c[3] = {h[2], h[5], h[8]};
  auto cross_norm = [](const double *u, const double *v) {
    const double x = u[1] * v[2] - u[2] * v[1], y = u[2] * v[0] - u[0] * v[2],
                 z = u[0] * v[1] - u[1] * v[0];
    return sqrt(x * x + y * y + z * z);
  };
  const double V = fabs(det9(h));
  const double wa = V / cross_norm(b, c), wb = V / cross_norm(c, a), wc = V / cross_norm(a, b);
  return wa < wb ? (wa < wc ? wa : wc) : (wb < wc ? wb : wc);
}

FMD_HD void inv9(const double *h, double *hi) {  // C1:147-158
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double d = det9(h), id = 1.0 / d;
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

// Fractional round-half-away minimum image (C1:166-175).
FMD_HD void mimg(double &dx, double &dy, double &dz, const double *hi, const double *h) {
  double s0 = hi[0] * dx + hi[1] * dy + hi[2] * dz;
  double s1 = hi[3] * dx + hi[4] * dy + hi[5] * dz;
  double s2 = hi[6] * dx + hi[7] * dy + hi[8] * dz;
  s0 -= round(s0);
  s1 -= round(s1);
  s2 -= round(s2);
  dx = h[0] * s0 + h[1] * s1 + h[2] * s2;
  dy = h[3] * s0 + h[4] * s1 + h[5] * s2;
  dz = h[6] * s0 + h[7] * s1 + h[8] * s2;
}

// The same minimum image with the reference's operation order and no FMA
// contraction: bitwise the reference's FP64 result, for decisions (cutoff,
// list membership) that must agree with it exactly.
FMD_HD void mimg_exact(double &dx, double &dy, double &dz, const double *hi, const double *h) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double s0 = hi[0] * dx + hi[1] * dy + hi[2] * dz;
  double s1 = hi[3] * dx + hi[4] * dy + hi[5] * dz;
  double s2 = hi[6] * dx + hi[7] * dy + hi[8] * dz;
  s0 -= round(s0);
  s1 -= round(s1);
  s2 -= round(s2);
  dx = h[0] * s0 + h[1] * s1 + h[2] * s2;
  dy = h[3] * s0 + h[4] * s1 + h[5] * s2;
  dz = h[6] * s0 + h[7] * s1 + h[8] * s2;
}

// dx^2 + dy^2 + dz^2 without contraction (the reference's r^2).
FMD_HD double norm2_exact(double dx, double dy, double dz) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  return dx * dx + dy * dy + dz * dz;
}

// ------------------------------------------------------------ quaternions
FMD_HD void q2R(const double *q, double *R) {  // C1:186-191
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

// Lab-frame site offset R(q) b with the reference's operation order and no
// contraction (C1:186-191, 357-362): bitwise the reference's lab array.
FMD_HD void lab_exact(const double *q, const double *b, double *o) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double w = q[0], x = q[1], y = q[2], z = q[3];
  double R0 = 1 - 2 * (y * y + z * z), R1 = 2 * (x * y - w * z), R2 = 2 * (x * z + w * y);
  double R3 = 2 * (x * y + w * z), R4 = 1 - 2 * (x * x + z * z), R5 = 2 * (y * z - w * x);
  double R6 = 2 * (x * z - w * y), R7 = 2 * (y * z + w * x), R8 = 1 - 2 * (x * x + y * y);
  o[0] = R0 * b[0] + R1 * b[1] + R2 * b[2];
  o[1] = R3 * b[0] + R4 * b[1] + R5 * b[2];
  o[2] = R6 * b[0] + R7 * b[1] + R8 * b[2];
}

// Squared length of dm + rb - ra with the reference's order (C1:398-401).
FMD_HD double pair_r2_exact(const double *dm, const double *rb, const double *ra) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
  double x = dm[0] + rb[0] - ra[0], y = dm[1] + rb[1] - ra[1], z = dm[2] + rb[2] - ra[2];
  return x * x + y * y + z * z;
}

// q <- normalise(q (x) dq(w dt)), dq the exact rotation by |w| dt about w
// (C1:193-216, 567-574).
FMD_HD void quat_advance(double *q, double wx, double wy, double wz, double dt) {
  double dq[4];
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
  double o0 = q[0] * dq[0] - q[1] * dq[1] - q[2] * dq[2] - q[3] * dq[3];
  double o1 = q[0] * dq[1] + q[1] * dq[0] + q[2] * dq[3] - q[3] * dq[2];
  double o2 = q[0] * dq[2] - q[1] * dq[3] + q[2] * dq[0] + q[3] * dq[1];
  double o3 = q[0] * dq[3] + q[1] * dq[2] - q[2] * dq[1] + q[3] * dq[0];
  double inv = 1.0 / sqrt(o0 * o0 + o1 * o1 + o2 * o2 + o3 * o3);
  q[0] = o0 * inv;
  q[1] = o1 * inv;
  q[2] = o2 * inv;
  q[3] = o3 * inv;
}

// ------------------------------------------------------------ LJ pair term
// Shifted 12-6 in FP32, with the energy kept in units of 4 eps and the shift
// folded in (e = sr6 (sr6 - 1) - vshift / (4 eps)), so one pair costs one
// reciprocal and 14 FMA-class operations.
#if defined(__HIPCC__)
// HIP events of one step on the engine's stream, bracketing its phases; a
// null pointer to this struct means "not timed".  flags goes to
// hipEventRecordWithFlags (hipEventRecordExternal inside a graph capture).
struct StepEvents {
  hipEvent_t f0, fm, f1;  // force phase: start, after the intramolecular part, end
  hipEvent_t l0, l1;      // list rebuild (recorded on rebuild steps only)
  unsigned flags;
};
#endif

struct LJ32 {
  float sig2;   // sigma^2
  float rc2;    // cutoff^2 (pairs with r^2 < rc2, or <= for AIREBO)
  float rc2_lo, rc2_hi;  // rc2 (1 -+ 1e-5): outside this band the FP32 r^2
                         // decides; inside it the pair is re-tested in FP64
  double rc2d;           // the FP64 cutoff^2 of that re-test
  float c48;    // 48 eps
  float c24;    // 24 eps
  float vs4;    // vshift / (4 eps)
  float rmin2;  // r^2 clamp (0.25, C1:404)
};

// Reciprocal: the hardware v_rcp_f32 (1 ulp) on the device, IEEE on the host.
FMD_HD float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// Returns fm / (48 eps): the force on the second atom is 48 eps fm' d (the
// caller applies 48 eps once per partial sum); adds the reduced energy
// sr6 (sr6 - 1) - vshift / (4 eps) into e.  One reciprocal, 10 VALU ops.
FMD_HD float lj_fm(float r2, const LJ32 &p, float &e) {
#if defined(__HIP_DEVICE_COMPILE__)
  // max(r2, rmin2) for finite r2 in one v_med3_f32 (fmaxf adds canonicalizes)
  float ri2 = rcp32(__builtin_amdgcn_fmed3f(r2, p.rmin2, 3.0e38f));
#else
  float ri2 = rcp32(fmaxf(r2, p.rmin2));
#endif
  float sr2 = p.sig2 * ri2;
  float sr6 = sr2 * sr2 * sr2;
  e += sr6 * (sr6 - 1.0f) - p.vs4;
  return sr6 * (sr6 - 0.5f) * ri2;
}

}  // namespace fmd
