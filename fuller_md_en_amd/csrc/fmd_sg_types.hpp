// fmd_sg_types.hpp -- constants and device buffers of the learned REBO
// surrogate (kernels: surrogate_kernels.hip; layout notes: fmd_surrogate.hpp).
#pragma once

namespace fmd {

constexpr int kSgNB = 16, kSgNV = 8, kSgLB = 4;
constexpr int kSgNF = kSgNB + 2 * kSgNV + 2 * kSgLB;  // 40
constexpr int kSgKF = 48;                             // features padded to 3 MFMA K-steps
constexpr int kSgH = 64;
constexpr int kSgMom = kSgNV * 9;                     // 72
constexpr int kSgNbMax = 6;
constexpr float kSgRc = 6.0f, kSgRb0 = 1.8f, kSgRb1 = 2.2f, kSgR0b = 1.42f;
// FP32 master parameters, in this order: W1 [40][64], b1 [64], W2 [64][64],
// b2, W3 [64][64], b3, w4 [64], b4 [1]
constexpr int kSgOffW1 = 0, kSgOffB1 = kSgOffW1 + kSgNF * kSgH, kSgOffW2 = kSgOffB1 + kSgH,
              kSgOffB2 = kSgOffW2 + kSgH * kSgH, kSgOffW3 = kSgOffB2 + kSgH,
              kSgOffB3 = kSgOffW3 + kSgH * kSgH, kSgOffW4 = kSgOffB3 + kSgH,
              kSgOffB4 = kSgOffW4 + kSgH, kSgNParam = kSgOffB4 + 1;

// BF16 operand copies of the weights: row-major W and W^T per layer (W1
// padded to 48 input rows), as raw bf16 bits.
struct SgWeights {
  unsigned short *w1, *w1t, *w2, *w2t, *w3, *w3t;  // [48][64], [64][48], [64][64] x 4
  float *theta;    // FP32 master parameters (kSgNParam)
  float *mu, *sd;  // feature normalisation (kSgNF each)
};

struct SgDev {
  unsigned short *feat;
  float *mom;
  int *nbc, *nbl;
  float *g, *ei, *vir, *u;
  double *F;        // surrogate forces (double, replaces Frebo)
  int *flags;       // [0] bonded-shell overflow
};

}  // namespace fmd
