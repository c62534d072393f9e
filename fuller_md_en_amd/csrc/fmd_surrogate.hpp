// fmd_surrogate.hpp -- device layout and launchers of the learned REBO
// surrogate: a per-atom energy model E = sum_i MLP(D_i) over same-molecule
// neighbours (no counterpart in the reference; it stands in for compute_rebo,
// AB:275-421, and is judged against it).
//
// Descriptor of atom i, neighbours j of its molecule, d = x_j - x_i (from the
// molecule's FP32 offset tile), r = |d|, u = d / r:
//   b_k(r) = T_k(x(r)) fc(r)   Chebyshev in x = 2 (s - s0) / (1 - s0) - 1,
//            s = exp(-(r - 1) / 1.5), s0 = s(rc); fc = (cos(pi r / rc) + 1) / 2
//   A_k = sum b_k                         k < 16
//   B_k = |sum b_k u|^2                   k < 8
//   C_k = ||sum b_k u (x) u||^2           k < 8
//   I_cl = sum_jj' w_c(r_j) w_c(r_j') cos^l(theta_jj'),  l = 1..4, over the
//          bonded shell (switch 1.8 -> 2.2 A), w_0 = w, w_1 = w (r - 1.42)
// 40 features -> (D - mu) / sd -> 40 x 64 -> 64 x 64 -> 64 x 64 -> 64 x 1,
// tanh, BF16 MFMA with FP32 accumulation.
//
// HBM layout (Na atoms of the engine):
//   feat   bf16 [Na][48]  normalised features (40..47 zero)
//   mom    float [Na][72]  sum b_k u (8 x 3) and sum b_k u (x) u (8 x 6:
//                         xx xy xz yy yz zz) for the force pass
//   nbc/nbl int [Na], int [Na][6]  bonded shell (molecule-local atom index)
//   g      float [Na][40]  dE_i / dD_i (unnormalised)
//   ei, vir  float [Na], float [Na][9]  per-atom energy and virial share
//   u      float [Na][40]  training: residual projected on descriptor space
#pragma once

#include <hip/hip_runtime.h>

#include "fmd_atomic.hpp"  // AtomicDev, SgWeights, SgDev

namespace fmd {

// Descriptor (feat, mom, bonded lists) from the offset tiles.
hipError_t sg_launch_descr(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                           const SgDev &s, hipStream_t st);
// MLP forward + input gradient: ei, g.
hipError_t sg_launch_mlp(const AtomicParams &p, const SgWeights &w, const SgDev &s,
                         hipStream_t st);
// Forces (into out, double [Na][3]) and per-atom virial shares from g.
hipError_t sg_launch_force(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                           const SgDev &s, double *out, hipStream_t st);
// Per-block energy / virial partials (the REBO partial layout, kAtPart per
// integrator block) from ei and vir.
hipError_t sg_launch_parts(const AtomicDev &d, const SgDev &s, double *part, hipStream_t st);
// Training: e = F - Fref (double [Na][3] each) -> u = sum_p c_p (e_j - e_i) / sd.
hipError_t sg_launch_u(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                       const SgDev &s, const double *F, const double *Fref, hipStream_t st);
// Training: parameter-gradient partials of S = sum_i u_i . dE_i/dD_i, one row of
// kSgNParam floats per workgroup (nwg = sg_train_groups(Na)).
int sg_train_groups(int na);
hipError_t sg_launch_train(const AtomicParams &p, const SgWeights &w, const SgDev &s,
                           float *partials, hipStream_t st);
// Per-block sums of (F - Fref)^2 over the integrator blocks (one double each).
hipError_t sg_launch_sqerr(const AtomicDev &d, const double *F, const double *Fref, double *part,
                           hipStream_t st);
// grad = scale * (sum of the nwg partial rows, fixed order); Adam step on
// theta (step t, learning rate lr); BF16 operand copies from theta.
hipError_t sg_launch_reduce_grad(const float *partials, int nwg, float scale, float *grad,
                                 hipStream_t st);
hipError_t sg_launch_adam(const SgWeights &w, const float *grad, float *m, float *v, int t,
                          float lr, hipStream_t st);
hipError_t sg_launch_pack(const SgWeights &w, hipStream_t st);

}  // namespace fmd
