// surrogate_kernels.hip -- gfx950 kernels of the learned REBO surrogate.
//
//   k_sg_descr   descriptor of every atom: one wavefront per molecule, lane =
//                atom, the molecule's atoms wavefront-uniform (scalar loads
//                of its FP32 offset tile); bonded shell collected on the way
//   k_sg_mlp     40-64-64-64-1 tanh MLP and its input gradient on BF16 MFMA
//                (v_mfma_f32_16x16x16_bf16, FP32 accumulation): one
//                wavefront per 16 atoms, activations kept transposed
//                (feature x atom) so every layer's accumulator tile is the
//                next layer's B operand without a shuffle; a persistent
//                grid with the weights staged in LDS once per workgroup
//   k_sg_force   forces and per-atom virial shares: one workgroup per
//                molecule, lane = centre atom i, dE_i/dd_ij per neighbour,
//                the neighbours split over four wavefronts (their shares
//                summed in wavefront order); the neighbour side is a
//                fixed-order wavefront sum per neighbour (radial terms) or an
//                LDS exchange over the bonded lists (angular terms)
//   k_sg_parts   per-integrator-block energy / virial partials (REBO layout,
//                component-major)
//   k_sg_u       training: u_i = sum_j dD_i/dd_ij . (e_j - e_i)
//   k_sg_train   training: forward tangent along u and reverse sweep, weight
//                gradients as MFMA products over the atoms of a workgroup
// All sums have a fixed order (no floating-point atomics).
#include "fmd_surrogate.hpp"

namespace fmd {

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));  // four bf16 (raw bits)
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr float kPi = 3.14159265358979f;

__device__ __forceinline__ unsigned short f2bf(float x) {  // round to nearest even
  unsigned int u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
__device__ __forceinline__ float bfround(float x) { return bf2f(f2bf(x)); }
// The network's activation: tanh x = 1 - 2 / (e^(2x) + 1) with the hardware
// base-2 exponential and reciprocal (absolute error ~1e-7, against the 4e-3
// relative rounding of the BF16 activations it feeds); |x| capped at 9, where
// tanh is 1 to FP32 precision.
__device__ __forceinline__ float sg_tanh(float x) {
  const float ax = fminf(fabsf(x), 9.f);
  const float e = __builtin_amdgcn_exp2f(ax * 2.88539008177792681f);  // 2 / ln 2
  return copysignf(1.f - 2.f * __builtin_amdgcn_rcpf(e + 1.f), x);
}
__device__ __forceinline__ bf4 pack4(float a, float b, float c, float d) {
  bf4 r;
  r[0] = (short)f2bf(a);
  r[1] = (short)f2bf(b);
  r[2] = (short)f2bf(c);
  r[3] = (short)f2bf(d);
  return r;
}
__device__ __forceinline__ bf4 ld_bf4(const unsigned short *p) {
  return *reinterpret_cast<const bf4 *>(p);
}
__device__ __forceinline__ f4 mfma(bf4 a, bf4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// b_k(r) and db_k/dr, k < kSgNB (see fmd_surrogate.hpp).
__device__ __forceinline__ void sg_radial(float r, float *b, float *db) {
  const float s0 = 0.035673993f;  // exp(-(6 - 1) / 1.5)
  const float s = __expf(-(r - 1.0f) * (1.0f / 1.5f));
  const float x = 2.0f * (s - s0) / (1.0f - s0) - 1.0f;
  const float dx = (2.0f / (1.0f - s0)) * (-s * (1.0f / 1.5f));
  float sn, cs;
  sincosf(kPi * r / kSgRc, &sn, &cs);
  const float fc = 0.5f * (cs + 1.0f), dfc = -0.5f * kPi / kSgRc * sn;
  float t0 = 1.0f, t1 = x, d0 = 0.0f, d1 = 1.0f;
  b[0] = fc;
  b[1] = x * fc;
  if (db) {
    db[0] = dfc;
    db[1] = dx * fc + x * dfc;
  }
#pragma unroll
  for (int k = 2; k < kSgNB; k++) {
    const float t2 = 2.0f * x * t1 - t0;
    const float d2 = 2.0f * t1 + 2.0f * x * d1 - d0;
    b[k] = t2 * fc;
    if (db) db[k] = d2 * dx * fc + t2 * dfc;
    t0 = t1;
    t1 = t2;
    d0 = d1;
    d1 = d2;
  }
}

// Bonded-shell channels w_0 = w, w_1 = w (r - r0) and their derivatives.
__device__ __forceinline__ void sg_bond(float r, float *w, float *dw) {
  const float t = fminf(fmaxf((r - kSgRb0) / (kSgRb1 - kSgRb0), 0.0f), 1.0f);
  float sn, cs;
  sincosf(kPi * t, &sn, &cs);
  const float w0 = 0.5f * (cs + 1.0f);
  const float dw0 = (r > kSgRb0 && r < kSgRb1) ? -0.5f * kPi / (kSgRb1 - kSgRb0) * sn : 0.0f;
  w[0] = w0;
  dw[0] = dw0;
  w[1] = w0 * (r - kSgR0b);
  dw[1] = dw0 * (r - kSgR0b) + w0;
}

// Fixed-order sum over the 64 lanes (butterfly: every lane gets the same value).
__device__ __forceinline__ float wsum(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ------------------------------------------------------------ descriptor
template <int NS>
__global__ __launch_bounds__(256) void k_sg_descr(AtomicDev d, AtomicParams p, SgWeights w,
                                                  SgDev s) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const int a0 = d.mol_start[m], n = d.mol_start[m + 1] - a0;
  const float4 *tile = d.off + (size_t)m * p.apad;
  const float rc2 = kSgRc * kSgRc;
#pragma unroll 1
  for (int si = 0; si < NS; si++) {
    const int a = lane + 64 * si;
    const bool act = a < n;
    const float4 xa = act ? tile[a] : make_float4(0.f, 0.f, 0.f, 0.f);
    float A[kSgNB], V[kSgNV][3], T[kSgNV][6];
#pragma unroll
    for (int k = 0; k < kSgNB; k++) A[k] = 0.f;
#pragma unroll
    for (int k = 0; k < kSgNV; k++) {
      V[k][0] = V[k][1] = V[k][2] = 0.f;
      for (int c = 0; c < 6; c++) T[k][c] = 0.f;
    }
    int nb = 0;
    int nbl[kSgNbMax];
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) nbl[q] = 0;
    for (int b = 0; b < n; b++) {
      const float4 xb = tile[b];
      const float dx = xb.x - xa.x, dy = xb.y - xa.y, dz = xb.z - xa.z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (act && b != a && r2 < rc2) {
        const float r = sqrtf(r2), ri = 1.0f / r;
        const float ux = dx * ri, uy = dy * ri, uz = dz * ri;
        float bk[kSgNB];
        sg_radial(r, bk, nullptr);
#pragma unroll
        for (int k = 0; k < kSgNB; k++) A[k] += bk[k];
        const float uu[6] = {ux * ux, ux * uy, ux * uz, uy * uy, uy * uz, uz * uz};
#pragma unroll
        for (int k = 0; k < kSgNV; k++) {
          V[k][0] += bk[k] * ux;
          V[k][1] += bk[k] * uy;
          V[k][2] += bk[k] * uz;
#pragma unroll
          for (int c = 0; c < 6; c++) T[k][c] += bk[k] * uu[c];
        }
        if (r < kSgRb1) {
#pragma unroll
          for (int q = 0; q < kSgNbMax; q++)
            if (nb == q) nbl[q] = b;
          nb++;
        }
      }
    }
    if (!act) continue;
    const int i = a0 + a;
    if (nb > kSgNbMax) {
      atomicAdd(&s.flags[0], 1);
      nb = kSgNbMax;
    }
    float D[kSgNF];
#pragma unroll
    for (int k = 0; k < kSgNB; k++) D[k] = A[k];
#pragma unroll
    for (int k = 0; k < kSgNV; k++) {
      D[kSgNB + k] = V[k][0] * V[k][0] + V[k][1] * V[k][1] + V[k][2] * V[k][2];
      D[kSgNB + kSgNV + k] = T[k][0] * T[k][0] + T[k][3] * T[k][3] + T[k][5] * T[k][5] +
                             2.f * (T[k][1] * T[k][1] + T[k][2] * T[k][2] + T[k][4] * T[k][4]);
    }
    // bonded shell: I_cl = sum_qq' w_c(q) w_c(q') cos^l
    float bu[kSgNbMax][3], bw[kSgNbMax][2];
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) {
      bu[q][0] = bu[q][1] = bu[q][2] = 0.f;
      bw[q][0] = bw[q][1] = 0.f;
      if (q < nb) {
        const float4 xq = tile[nbl[q]];
        const float dx = xq.x - xa.x, dy = xq.y - xa.y, dz = xq.z - xa.z;
        const float r = sqrtf(dx * dx + dy * dy + dz * dz), ri = 1.0f / r;
        bu[q][0] = dx * ri;
        bu[q][1] = dy * ri;
        bu[q][2] = dz * ri;
        float dw[2];
        sg_bond(r, bw[q], dw);
      }
    }
    float I[2][kSgLB];
#pragma unroll
    for (int c = 0; c < 2; c++)
      for (int l = 0; l < kSgLB; l++) I[c][l] = 0.f;
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++)
#pragma unroll
      for (int q2 = 0; q2 < kSgNbMax; q2++) {
        const float cs = bu[q][0] * bu[q2][0] + bu[q][1] * bu[q2][1] + bu[q][2] * bu[q2][2];
        float pw = cs;
#pragma unroll
        for (int l = 0; l < kSgLB; l++) {
          I[0][l] += bw[q][0] * bw[q2][0] * pw;
          I[1][l] += bw[q][1] * bw[q2][1] * pw;
          pw *= cs;
        }
      }
#pragma unroll
    for (int c = 0; c < 2; c++)
      for (int l = 0; l < kSgLB; l++) D[kSgNB + 2 * kSgNV + c * kSgLB + l] = I[c][l];
    unsigned short *fo = s.feat + (size_t)i * kSgKF;
#pragma unroll
    for (int f = 0; f < kSgKF; f += 4) {
      float v[4];
      for (int t = 0; t < 4; t++)
        v[t] = f + t < kSgNF ? (D[f + t] - w.mu[f + t]) / w.sd[f + t] : 0.f;
      *reinterpret_cast<bf4 *>(fo + f) = pack4(v[0], v[1], v[2], v[3]);
    }
    float *mo = s.mom + (size_t)i * kSgMom;
#pragma unroll
    for (int k = 0; k < kSgNV; k++) {
      mo[3 * k] = V[k][0];
      mo[3 * k + 1] = V[k][1];
      mo[3 * k + 2] = V[k][2];
#pragma unroll
      for (int c = 0; c < 6; c++) mo[3 * kSgNV + 6 * k + c] = T[k][c];
    }
    s.nbc[i] = nb;
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) s.nbl[(size_t)i * kSgNbMax + q] = nbl[q];
  }
}

// ------------------------------------------------------------------- MLP
// The network's BF16 weight matrices in LDS, rows padded (+4 BF16) so that
// the 16 rows an MFMA operand reads fall in distinct bank pairs.
constexpr int kSgLdK = kSgKF + 4;  // row stride of w1t [64][48]
constexpr int kSgLdH = kSgH + 4;   // row stride of the [.][64] matrices
struct SgLdsW {
  unsigned short w1t[kSgH * kSgLdK], w2t[kSgH * kSgLdH], w3t[kSgH * kSgLdH];
  unsigned short w3[kSgH * kSgLdH], w2[kSgH * kSgLdH], w1[kSgKF * kSgLdH];
};
constexpr int kSgMlpGroupsPerCU = 3;  // workgroups of the persistent grid per CU (48 KB LDS each)

__device__ __forceinline__ void stage_rows(unsigned short *dst, const unsigned short *src, int rows,
                                           int cols, int ld) {
  // four BF16 (8 bytes) per thread and step
  const int q = cols / 4;
  for (int e = threadIdx.x; e < rows * q; e += blockDim.x) {
    const int r = e / q, c = 4 * (e % q);
    *reinterpret_cast<bf4 *>(dst + r * ld + c) = ld_bf4(src + r * cols + c);
  }
}

// One wavefront = 16 atoms.  Transposed activations: tile m of a layer
// holds rows (features) 16m .. 16m+15 x the 16 atoms; lane l keeps rows
// 16m + 4 (l / 16) + v (v < 4) of atom l % 16 -- the MFMA accumulator
// layout, which is also the B-operand layout of the next product.  A
// persistent grid: each workgroup stages the weights in LDS once, then its
// four wavefronts take 16-atom tiles, 64 atoms per workgroup and pass.
__global__ __launch_bounds__(256) void k_sg_mlp(int na, SgWeights w, SgDev s) {
  __shared__ SgLdsW L;
  stage_rows(L.w1t, w.w1t, kSgH, kSgKF, kSgLdK);
  stage_rows(L.w2t, w.w2t, kSgH, kSgH, kSgLdH);
  stage_rows(L.w3t, w.w3t, kSgH, kSgH, kSgLdH);
  stage_rows(L.w3, w.w3, kSgH, kSgH, kSgLdH);
  stage_rows(L.w2, w.w2, kSgH, kSgH, kSgLdH);
  stage_rows(L.w1, w.w1, kSgKF, kSgH, kSgLdH);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int grp = lane >> 4;
  const float *th = w.theta;
  for (int b0 = blockIdx.x * 64; b0 < na; b0 += gridDim.x * 64) {
  const int base = b0 + (threadIdx.x >> 6) * 16;
  if (base >= na) break;
  const int atom = base + (lane & 15);
  const bool ok = atom < na;
  // input tiles (features 16t + 4 grp + v of this lane's atom)
  bf4 x[3];
#pragma unroll
  for (int t = 0; t < 3; t++)
    x[t] = ok ? ld_bf4(s.feat + (size_t)atom * kSgKF + 16 * t + 4 * grp) : bf4{0, 0, 0, 0};
  float h1[4][4], h2[4][4], h3[4][4];
  // layer 1: z1^T = W1^T x^T
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 3; t++)
      acc = mfma(ld_bf4(L.w1t + (16 * mt + (lane & 15)) * kSgLdK + 16 * t + 4 * grp), x[t], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) h1[mt][v] = sg_tanh(acc[v] + th[kSgOffB1 + 16 * mt + 4 * grp + v]);
  }
  bf4 hb[4];
#pragma unroll
  for (int t = 0; t < 4; t++) hb[t] = pack4(h1[t][0], h1[t][1], h1[t][2], h1[t][3]);
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc = mfma(ld_bf4(L.w2t + (16 * mt + (lane & 15)) * kSgLdH + 16 * t + 4 * grp), hb[t], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) h2[mt][v] = sg_tanh(acc[v] + th[kSgOffB2 + 16 * mt + 4 * grp + v]);
  }
#pragma unroll
  for (int t = 0; t < 4; t++) hb[t] = pack4(h2[t][0], h2[t][1], h2[t][2], h2[t][3]);
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc = mfma(ld_bf4(L.w3t + (16 * mt + (lane & 15)) * kSgLdH + 16 * t + 4 * grp), hb[t], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) h3[mt][v] = sg_tanh(acc[v] + th[kSgOffB3 + 16 * mt + 4 * grp + v]);
  }
  // energy: w4 . h3 (w4 in bf16 as the MFMA operands are)
  float e = 0.f;
  bf4 db[4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    float dv[4];
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const float w4 = bfround(th[kSgOffW4 + 16 * mt + 4 * grp + v]);
      e += w4 * h3[mt][v];
      dv[v] = w4 * (1.f - h3[mt][v] * h3[mt][v]);
    }
    db[mt] = pack4(dv[0], dv[1], dv[2], dv[3]);
  }
  e += __shfl_xor(e, 16, 64);
  e += __shfl_xor(e, 32, 64);
  if (ok && grp == 0) s.ei[atom] = e + th[kSgOffB4];
  // backward: dE/dh2^T = W3 delta3^T, delta2 = . (1 - h2^2), ...
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc = mfma(ld_bf4(L.w3 + (16 * mt + (lane & 15)) * kSgLdH + 16 * t + 4 * grp), db[t], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) h2[mt][v] = acc[v] * (1.f - h2[mt][v] * h2[mt][v]);
  }
#pragma unroll
  for (int t = 0; t < 4; t++) db[t] = pack4(h2[t][0], h2[t][1], h2[t][2], h2[t][3]);
#pragma unroll
  for (int mt = 0; mt < 4; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc = mfma(ld_bf4(L.w2 + (16 * mt + (lane & 15)) * kSgLdH + 16 * t + 4 * grp), db[t], acc);
#pragma unroll
    for (int v = 0; v < 4; v++) h1[mt][v] = acc[v] * (1.f - h1[mt][v] * h1[mt][v]);
  }
#pragma unroll
  for (int t = 0; t < 4; t++) db[t] = pack4(h1[t][0], h1[t][1], h1[t][2], h1[t][3]);
#pragma unroll
  for (int mt = 0; mt < 3; mt++) {
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; t++)
      acc = mfma(ld_bf4(L.w1 + (16 * mt + (lane & 15)) * kSgLdH + 16 * t + 4 * grp), db[t], acc);
    if (ok) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int f = 16 * mt + 4 * grp + v;
        if (f < kSgNF) s.g[(size_t)atom * kSgNF + f] = acc[v] / w.sd[f];
      }
    }
  }
  }
}

// --------------------------------------------------------------- forces
// Per-lane gradient coefficients of atom i: gA_k, P_k = gB_k V_k,
// Q_k = gC_k T_k, gI_cl.
struct SgCoef {
  float gA[kSgNB];
  float P[kSgNV][3];
  float Q[kSgNV][6];
  float gI[2][kSgLB];
};

__device__ __forceinline__ void sg_coef(const SgDev &s, int i, SgCoef &c) {
  const float *g = s.g + (size_t)i * kSgNF;
  const float *mo = s.mom + (size_t)i * kSgMom;
#pragma unroll
  for (int k = 0; k < kSgNB; k++) c.gA[k] = g[k];
#pragma unroll
  for (int k = 0; k < kSgNV; k++) {
    const float gb = g[kSgNB + k], gc = g[kSgNB + kSgNV + k];
#pragma unroll
    for (int x = 0; x < 3; x++) c.P[k][x] = gb * mo[3 * k + x];
#pragma unroll
    for (int x = 0; x < 6; x++) c.Q[k][x] = gc * mo[3 * kSgNV + 6 * k + x];
  }
#pragma unroll
  for (int ch = 0; ch < 2; ch++)
    for (int l = 0; l < kSgLB; l++) c.gI[ch][l] = g[kSgNB + 2 * kSgNV + ch * kSgLB + l];
}

// dE_i/dd for one radial pair (i, j): d, r, u.
__device__ __forceinline__ void sg_pair_grad(const SgCoef &c, float r, const float *u,
                                             float *f) {
  float b[kSgNB], db[kSgNB];
  sg_radial(r, b, db);
  float a0 = 0.f;
#pragma unroll
  for (int k = 0; k < kSgNB; k++) a0 += c.gA[k] * db[k];
  float Ph[3] = {0.f, 0.f, 0.f}, Pt[3] = {0.f, 0.f, 0.f};
  float Qh[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, Qt[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < kSgNV; k++) {
#pragma unroll
    for (int x = 0; x < 3; x++) {
      Ph[x] += db[k] * c.P[k][x];
      Pt[x] += b[k] * c.P[k][x];
    }
#pragma unroll
    for (int x = 0; x < 6; x++) {
      Qh[x] += db[k] * c.Q[k][x];
      Qt[x] += b[k] * c.Q[k][x];
    }
  }
  const float ri = 1.0f / r;
  const float Phu = Ph[0] * u[0] + Ph[1] * u[1] + Ph[2] * u[2];
  const float Ptu = Pt[0] * u[0] + Pt[1] * u[1] + Pt[2] * u[2];
  // Q u for symmetric Q = (xx xy xz yy yz zz)
  const float Qhu[3] = {Qh[0] * u[0] + Qh[1] * u[1] + Qh[2] * u[2],
                        Qh[1] * u[0] + Qh[3] * u[1] + Qh[4] * u[2],
                        Qh[2] * u[0] + Qh[4] * u[1] + Qh[5] * u[2]};
  const float Qtu[3] = {Qt[0] * u[0] + Qt[1] * u[1] + Qt[2] * u[2],
                        Qt[1] * u[0] + Qt[3] * u[1] + Qt[4] * u[2],
                        Qt[2] * u[0] + Qt[4] * u[1] + Qt[5] * u[2]};
  const float uQhu = Qhu[0] * u[0] + Qhu[1] * u[1] + Qhu[2] * u[2];
  const float uQtu = Qtu[0] * u[0] + Qtu[1] * u[1] + Qtu[2] * u[2];
  const float alpha = a0 + 2.f * Phu + 2.f * uQhu - ri * (2.f * Ptu + 4.f * uQtu);
#pragma unroll
  for (int x = 0; x < 3; x++) f[x] = alpha * u[x] + ri * (2.f * Pt[x] + 4.f * Qtu[x]);
}

// dE_i/dd_q for the bonded shell of atom i (nb neighbours: unit vectors bu,
// distances br, channels bw / bdw); gradients into fq.
__device__ __forceinline__ void sg_bond_grad(const float gI[2][kSgLB], int nb,
                                             const float bu[kSgNbMax][3],
                                             const float br[kSgNbMax],
                                             const float bw[kSgNbMax][2],
                                             const float bdw[kSgNbMax][2],
                                             float fq[kSgNbMax][3]) {
#pragma unroll
  for (int q = 0; q < kSgNbMax; q++) {
    fq[q][0] = fq[q][1] = fq[q][2] = 0.f;
    if (q >= nb) continue;
    float fu = 0.f, fv[3] = {0.f, 0.f, 0.f};  // coefficient of u_q; vector part
#pragma unroll
    for (int q2 = 0; q2 < kSgNbMax; q2++) {
      if (q2 >= nb) continue;
      const float cs = bu[q][0] * bu[q2][0] + bu[q][1] * bu[q2][1] + bu[q][2] * bu[q2][2];
      float pl = cs, plm = 1.f;  // cos^l, cos^(l-1)
#pragma unroll
      for (int l = 1; l <= kSgLB; l++) {
#pragma unroll
        for (int ch = 0; ch < 2; ch++) {
          const float g = 2.f * gI[ch][l - 1] * bw[q2][ch];
          // 2 [w'_q w_q2 c^l u_q + w_q w_q2 l c^(l-1) (u_q2 - c u_q) / r_q]
          fu += g * (bdw[q][ch] * pl - bw[q][ch] * l * plm * cs / br[q]);
          const float tv = g * bw[q][ch] * l * plm / br[q];
          fv[0] += tv * bu[q2][0];
          fv[1] += tv * bu[q2][1];
          fv[2] += tv * bu[q2][2];
        }
        plm = pl;
        pl *= cs;
      }
    }
#pragma unroll
    for (int x = 0; x < 3; x++) fq[q][x] = fu * bu[q][x] + fv[x];
  }
}

__device__ __forceinline__ void bond_geom(const float4 *tile, const float4 &xa, const int *nbl,
                                          int nb, float bu[kSgNbMax][3], float br[kSgNbMax],
                                          float bd[kSgNbMax][3], float bw[kSgNbMax][2],
                                          float bdw[kSgNbMax][2]) {
#pragma unroll
  for (int q = 0; q < kSgNbMax; q++) {
    bu[q][0] = bu[q][1] = bu[q][2] = 0.f;
    bd[q][0] = bd[q][1] = bd[q][2] = 0.f;
    br[q] = 1.f;
    bw[q][0] = bw[q][1] = bdw[q][0] = bdw[q][1] = 0.f;
    if (q < nb) {
      const float4 xq = tile[nbl[q]];
      bd[q][0] = xq.x - xa.x;
      bd[q][1] = xq.y - xa.y;
      bd[q][2] = xq.z - xa.z;
      br[q] = sqrtf(bd[q][0] * bd[q][0] + bd[q][1] * bd[q][1] + bd[q][2] * bd[q][2]);
      for (int x = 0; x < 3; x++) bu[q][x] = bd[q][x] / br[q];
      sg_bond(br[q], bw[q], bdw[q]);
    }
  }
}

// WPM wavefronts per molecule (1 or 4; NS atoms per lane), 4 / WPM molecules
// per workgroup: wavefront q of a molecule takes the neighbours b of share q
// of its atoms for every atom of the molecule (four shares keep small systems,
// with few molecules per CU, from waiting on one wavefront each).
// F_i += sum_j dE_i/dd_ij, F_j -= dE_i/dd_ij (the latter summed over the
// wavefront, to the lane that owns j); virial share of i:
// -sum_j d_ij (x) dE_i/dd_ij.  A molecule's wavefronts' partial forces and
// virials are summed in LDS in wavefront order; the bonded shell is its
// first wavefront's.
template <int NS, int WPM>
__global__ __launch_bounds__(256) void k_sg_force(AtomicDev d, AtomicParams p, SgWeights w,
                                                  SgDev s, double *out) {
  constexpr int NA = 64 * NS, MPB = 4 / WPM;
  __shared__ float xf[MPB][NA][kSgNbMax][3];
  // each wavefront's F (3) and W (9) per atom (shares of one molecule)
  __shared__ float fw[WPM > 1 ? 4 : 1][WPM > 1 ? NA : 1][12];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ml = wv / WPM, q = wv % WPM;  // molecule of the workgroup, share
  const int m = blockIdx.x * MPB + ml;
  const bool valid = m < p.Nm;  // no early return: the block meets at barriers
  const int a0 = valid ? d.mol_start[m] : 0, n = valid ? d.mol_start[m + 1] - a0 : 0;
  const float4 *tile = d.off + (size_t)(valid ? m : 0) * p.apad;
  const float rc2 = kSgRc * kSgRc;
  const int b_lo = (n * q) / WPM, b_hi = (n * (q + 1)) / WPM;
  float F[NS][3], W[NS][9];
#pragma unroll
  for (int si = 0; si < NS; si++) {
    F[si][0] = F[si][1] = F[si][2] = 0.f;
    for (int k = 0; k < 9; k++) W[si][k] = 0.f;
  }
#pragma unroll 1
  for (int si = 0; si < NS; si++) {
    const int a = lane + 64 * si;
    const bool act = a < n;
    const float4 xa = act ? tile[a] : make_float4(0.f, 0.f, 0.f, 0.f);
    SgCoef c;
    sg_coef(s, act ? a0 + a : a0, c);
    for (int b = b_lo; b < b_hi; b++) {
      const float4 xb = tile[b];
      const float dv[3] = {xb.x - xa.x, xb.y - xa.y, xb.z - xa.z};
      const float r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      float f[3] = {0.f, 0.f, 0.f};
      if (act && b != a && r2 < rc2) {
        const float r = sqrtf(r2), ri = 1.0f / r;
        const float u[3] = {dv[0] * ri, dv[1] * ri, dv[2] * ri};
        sg_pair_grad(c, r, u, f);
#pragma unroll
        for (int x = 0; x < 3; x++) {
          F[si][x] += f[x];
#pragma unroll
          for (int y = 0; y < 3; y++) W[si][3 * x + y] -= dv[x] * f[y];
        }
      }
      // neighbour side: -sum over centres, to the lane that owns b
      const float sx = wsum(f[0]), sy = wsum(f[1]), sz = wsum(f[2]);
      if ((b & 63) == lane) {
#pragma unroll
        for (int sj = 0; sj < NS; sj++)
          if ((b >> 6) == sj) {
            F[sj][0] -= sx;
            F[sj][1] -= sy;
            F[sj][2] -= sz;
          }
      }
    }
  }
  // bonded shell (the molecule's first wavefront): own share now, the
  // neighbour's through LDS
  int nbs[NS], nbls[NS][kSgNbMax];
#pragma unroll
  for (int si = 0; si < NS; si++) {
    const int a = lane + 64 * si;
    const bool act = q == 0 && a < n;
    nbs[si] = act ? s.nbc[a0 + a] : 0;
#pragma unroll
    for (int qb = 0; qb < kSgNbMax; qb++)
      nbls[si][qb] = act ? s.nbl[(size_t)(a0 + a) * kSgNbMax + qb] : 0;
    if (!act) continue;
    const float4 xa = tile[a];
    float bu[kSgNbMax][3], br[kSgNbMax], bd[kSgNbMax][3], bw[kSgNbMax][2], bdw[kSgNbMax][2];
    bond_geom(tile, xa, nbls[si], nbs[si], bu, br, bd, bw, bdw);
    float gI[2][kSgLB];
    const float *g = s.g + (size_t)(a0 + a) * kSgNF;
#pragma unroll
    for (int ch = 0; ch < 2; ch++)
      for (int l = 0; l < kSgLB; l++) gI[ch][l] = g[kSgNB + 2 * kSgNV + ch * kSgLB + l];
    float fq[kSgNbMax][3];
    sg_bond_grad(gI, nbs[si], bu, br, bw, bdw, fq);
#pragma unroll
    for (int qb = 0; qb < kSgNbMax; qb++) {
#pragma unroll
      for (int x = 0; x < 3; x++) {
        F[si][x] += fq[qb][x];
        xf[ml][a][qb][x] = fq[qb][x];
#pragma unroll
        for (int y = 0; y < 3; y++) W[si][3 * x + y] -= bd[qb][x] * fq[qb][y];
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int si = 0; si < NS; si++) {
    const int a = lane + 64 * si;
    if (q == 0 && a < n) {
      for (int qb = 0; qb < nbs[si]; qb++) {
        const int j = nbls[si][qb];  // a is in j's bonded list (same distance test)
        const int cj = s.nbc[a0 + j];
        const int *lj = s.nbl + (size_t)(a0 + j) * kSgNbMax;
        for (int q2 = 0; q2 < cj; q2++)
          if (lj[q2] == a) {
            F[si][0] -= xf[ml][j][q2][0];
            F[si][1] -= xf[ml][j][q2][1];
            F[si][2] -= xf[ml][j][q2][2];
          }
      }
    }
    if constexpr (WPM == 1) {
      if (a < n) {
        const size_t i = a0 + a;
        for (int x = 0; x < 3; x++) out[3 * i + x] = (double)F[si][x];
        for (int k = 0; k < 9; k++) s.vir[9 * i + k] = W[si][k];
      }
    } else {
      for (int x = 0; x < 3; x++) fw[wv][a][x] = F[si][x];
      for (int k = 0; k < 9; k++) fw[wv][a][3 + k] = W[si][k];
    }
  }
  if constexpr (WPM == 1) return;
  __syncthreads();
  // each molecule's shares, in wavefront order: its wavefront q = 0 .. WPM-1
  // takes atoms lane + 64 si
  if (valid) {
#pragma unroll
    for (int si = 0; si < NS; si++) {
      const int a = lane + 64 * si;
      if (q != 0 || a >= n) continue;
      const size_t i = a0 + a;
      float f[12];
      for (int k = 0; k < 12; k++) {
        f[k] = fw[wv][a][k];
        for (int u = 1; u < WPM; u++) f[k] += fw[wv + u][a][k];
      }
      for (int x = 0; x < 3; x++) out[3 * i + x] = (double)f[x];
      for (int k = 0; k < 9; k++) s.vir[9 * i + k] = f[3 + k];
    }
  }
}

// ------------------------------------------------------------- partials
template <int K>
__device__ __forceinline__ void block_sum_d(const double *v, double *out, int stride) {
  __shared__ double red[4][K];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < K; k++) {
    double x = v[k];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
    if (lane == 0) red[wv][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    out[(size_t)k * stride] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

__global__ __launch_bounds__(256) void k_sg_parts(AtomicDev d, SgDev s, double *part) {
  const int i = d.blk_a0[blockIdx.x] + threadIdx.x;
  double v[kAtPart];
  for (int k = 0; k < kAtPart; k++) v[k] = 0;
  if (i < d.blk_a1[blockIdx.x]) {
    for (int k = 0; k < 9; k++) v[k] = s.vir[9 * (size_t)i + k];
    v[9] = s.ei[i];
  }
  block_sum_d<kAtPart>(v, part + blockIdx.x, d.nblk);  // component-major
}

// --------------------------------------------------------------- training
// u_i = sum_j (dD_i/dd_ij) . (e_j - e_i), divided by sd (normalised units).
template <int NS>
__global__ __launch_bounds__(256) void k_sg_u(AtomicDev d, AtomicParams p, SgWeights w,
                                              SgDev s, const double *F, const double *Fref) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const int a0 = d.mol_start[m], n = d.mol_start[m + 1] - a0;
  const float4 *tile = d.off + (size_t)m * p.apad;
  const float rc2 = kSgRc * kSgRc;
#pragma unroll 1
  for (int si = 0; si < NS; si++) {
    const int a = lane + 64 * si;
    const bool act = a < n;
    const int i = act ? a0 + a : a0;
    const float4 xa = act ? tile[a] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float ei[3] = {(float)(F[3 * i] - Fref[3 * i]), (float)(F[3 * i + 1] - Fref[3 * i + 1]),
                         (float)(F[3 * i + 2] - Fref[3 * i + 2])};
    float V[kSgNV][3], T[kSgNV][6];
    const float *mo = s.mom + (size_t)i * kSgMom;
#pragma unroll
    for (int k = 0; k < kSgNV; k++) {
      for (int x = 0; x < 3; x++) V[k][x] = mo[3 * k + x];
      for (int x = 0; x < 6; x++) T[k][x] = mo[3 * kSgNV + 6 * k + x];
    }
    float uf[kSgNF];
#pragma unroll
    for (int f = 0; f < kSgNF; f++) uf[f] = 0.f;
    for (int b = 0; b < n; b++) {
      const float4 xb = tile[b];
      const float dv[3] = {xb.x - xa.x, xb.y - xa.y, xb.z - xa.z};
      const float r2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      const int j = a0 + b;
      const float de[3] = {(float)(F[3 * j] - Fref[3 * j]) - ei[0],
                           (float)(F[3 * j + 1] - Fref[3 * j + 1]) - ei[1],
                           (float)(F[3 * j + 2] - Fref[3 * j + 2]) - ei[2]};
      if (act && b != a && r2 < rc2) {
        const float r = sqrtf(r2), ri = 1.0f / r;
        const float u[3] = {dv[0] * ri, dv[1] * ri, dv[2] * ri};
        float bk[kSgNB], dbk[kSgNB];
        sg_radial(r, bk, dbk);
        const float sd_ = u[0] * de[0] + u[1] * de[1] + u[2] * de[2];
#pragma unroll
        for (int k = 0; k < kSgNB; k++) uf[k] += dbk[k] * sd_;
#pragma unroll
        for (int k = 0; k < kSgNV; k++) {
          const float Vu = V[k][0] * u[0] + V[k][1] * u[1] + V[k][2] * u[2];
          const float Vd = V[k][0] * de[0] + V[k][1] * de[1] + V[k][2] * de[2];
          uf[kSgNB + k] += 2.f * (dbk[k] * Vu * sd_ + bk[k] * ri * (Vd - Vu * sd_));
          const float Tu[3] = {T[k][0] * u[0] + T[k][1] * u[1] + T[k][2] * u[2],
                               T[k][1] * u[0] + T[k][3] * u[1] + T[k][4] * u[2],
                               T[k][2] * u[0] + T[k][4] * u[1] + T[k][5] * u[2]};
          const float uTu = Tu[0] * u[0] + Tu[1] * u[1] + Tu[2] * u[2];
          const float uTd = Tu[0] * de[0] + Tu[1] * de[1] + Tu[2] * de[2];
          uf[kSgNB + kSgNV + k] += 2.f * dbk[k] * uTu * sd_ + 4.f * bk[k] * ri * (uTd - uTu * sd_);
        }
      }
    }
    if (!act) continue;
    // bonded shell: u_cl += (dI_cl/dd_q) . (e_q - e_i), gradient with gI = unit vectors
    const int nb = s.nbc[i];
    int nbl[kSgNbMax];
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) nbl[q] = s.nbl[(size_t)i * kSgNbMax + q];
    float bu[kSgNbMax][3], br[kSgNbMax], bd[kSgNbMax][3], bw[kSgNbMax][2], bdw[kSgNbMax][2];
    bond_geom(tile, xa, nbl, nb, bu, br, bd, bw, bdw);
    float deq[kSgNbMax][3];
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) {
      const int j = a0 + (q < nb ? nbl[q] : a);
      for (int x = 0; x < 3; x++) deq[q][x] = (float)(F[3 * j + x] - Fref[3 * j + x]) - ei[x];
    }
    // (dI_cl/dd_q) . de_q = 2 sum_q2 w_c(q2) [w'_c(q) c^l (u_q.de_q)
    //                        + w_c(q) l c^(l-1) (u_q2.de_q - c u_q.de_q) / r_q]
    float ub[2][kSgLB];
#pragma unroll
    for (int ch = 0; ch < 2; ch++)
      for (int l = 0; l < kSgLB; l++) ub[ch][l] = 0.f;
#pragma unroll
    for (int q = 0; q < kSgNbMax; q++) {
      if (q >= nb) continue;
      const float ue = bu[q][0] * deq[q][0] + bu[q][1] * deq[q][1] + bu[q][2] * deq[q][2];
#pragma unroll
      for (int q2 = 0; q2 < kSgNbMax; q2++) {
        if (q2 >= nb) continue;
        const float cs = bu[q][0] * bu[q2][0] + bu[q][1] * bu[q2][1] + bu[q][2] * bu[q2][2];
        const float u2e = bu[q2][0] * deq[q][0] + bu[q2][1] * deq[q][1] + bu[q2][2] * deq[q][2];
        float pl = cs, plm = 1.f;
#pragma unroll
        for (int l = 1; l <= kSgLB; l++) {
#pragma unroll
          for (int ch = 0; ch < 2; ch++)
            ub[ch][l - 1] += 2.f * bw[q2][ch] *
                             (bdw[q][ch] * pl * ue + bw[q][ch] * l * plm * (u2e - cs * ue) / br[q]);
          plm = pl;
          pl *= cs;
        }
      }
    }
#pragma unroll
    for (int ch = 0; ch < 2; ch++)
      for (int l = 0; l < kSgLB; l++) uf[kSgNB + 2 * kSgNV + ch * kSgLB + l] += ub[ch][l];
#pragma unroll
    for (int f = 0; f < kSgNF; f++) s.u[(size_t)i * kSgNF + f] = uf[f] / w.sd[f];
  }
}

// Gradient of S = sum_i u_i . dE_i/dD~_i over the trained parameters.
// Per workgroup: 4 wavefronts x 16 atoms per pass, looping over the atom
// tiles assigned to it.  Each wavefront runs the forward pass with its
// tangent (h, hdot) and the reverse sweep (zeta, zetadot), stages the
// GEMM operands in LDS as [row][64 atoms] bf16, and after a barrier wave w
// accumulates one weight matrix over the 64 atoms as MFMA products:
// wave 0 W1 (48 x 64), wave 1 W2, wave 2 W3, wave 3 biases and w4.
struct TrainLds {
  unsigned short x0[kSgKF][64], t0[kSgKF][64];  // h0, u
  unsigned short x1[kSgH][64], t1[kSgH][64];    // h1, h1dot
  unsigned short x2[kSgH][64], t2[kSgH][64];    // h2, h2dot
  unsigned short z1[kSgH][64], y1[kSgH][64];    // zeta1, zeta1dot
  unsigned short z2[kSgH][64], y2[kSgH][64];
  unsigned short z3[kSgH][64], y3[kSgH][64];
  unsigned short t3[kSgH][64];                  // h3dot (w4 gradient)
};

int sg_train_groups_impl(int na) {
  const int tiles = (na + 63) / 64;
  return tiles < 512 ? tiles : 512;
}

__global__ __launch_bounds__(256) void k_sg_train(int na, SgWeights w, SgDev s, float *partials) {
  extern __shared__ unsigned char lds_raw[];
  TrainLds &L = *reinterpret_cast<TrainLds *>(lds_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, grp = lane >> 4;
  const float *th = w.theta;
  // gradient accumulators of this wave's matrix: up to 4 x 4 tiles
  f4 G[4][4];
#pragma unroll
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) G[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  float gb[3] = {0.f, 0.f, 0.f};  // wave 3: per-thread bias / w4 sums (row threadIdx.x % 64)
  float gw4 = 0.f;
  const int ntiles = (na + 63) / 64;
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int base = tile * 64 + wv * 16;
    const int atom = base + (lane & 15);
    const bool ok = atom < na;
    const int col0 = wv * 16;
    bf4 x[3], t[3];
    float xs[4][4], ts[4][4];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      x[k] = ok ? ld_bf4(s.feat + (size_t)atom * kSgKF + 16 * k + 4 * grp) : bf4{0, 0, 0, 0};
      float uv[4];
      for (int v = 0; v < 4; v++) {
        const int f = 16 * k + 4 * grp + v;
        uv[v] = (ok && f < kSgNF) ? s.u[(size_t)atom * kSgNF + f] : 0.f;
      }
      t[k] = pack4(uv[0], uv[1], uv[2], uv[3]);
      for (int v = 0; v < 4; v++) {
        xs[k][v] = bf2f((unsigned short)x[k][v]);
        ts[k][v] = uv[v];
      }
    }
    xs[3][0] = xs[3][1] = xs[3][2] = xs[3][3] = 0.f;
    ts[3][0] = ts[3][1] = ts[3][2] = ts[3][3] = 0.f;
    // stage h0 and u (48 rows)
    for (int k = 0; k < 3; k++)
      for (int v = 0; v < 4; v++) {
        L.x0[16 * k + 4 * grp + v][col0 + (lane & 15)] = f2bf(xs[k][v]);
        L.t0[16 * k + 4 * grp + v][col0 + (lane & 15)] = f2bf(ts[k][v]);
      }
    // forward with tangent
    float h[3][4][4], sdot[3][4][4], hd[3][4][4];
    bf4 hin[4], tin[4];
    const unsigned short *wt[3] = {w.w1t, w.w2t, w.w3t};
    const int kin[3] = {kSgKF, kSgH, kSgH};
    const int boff[3] = {kSgOffB1, kSgOffB2, kSgOffB3};
#pragma unroll
    for (int ly = 0; ly < 3; ly++) {
      const int nk = ly == 0 ? 3 : 4;
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f4 az = {0.f, 0.f, 0.f, 0.f}, at = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          if (k >= nk) continue;
          const bf4 A = ld_bf4(wt[ly] + (16 * mt + (lane & 15)) * kin[ly] + 16 * k + 4 * grp);
          az = mfma(A, ly == 0 ? x[k] : hin[k], az);
          at = mfma(A, ly == 0 ? t[k] : tin[k], at);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const float hv = sg_tanh(az[v] + th[boff[ly] + 16 * mt + 4 * grp + v]);
          h[ly][mt][v] = hv;
          sdot[ly][mt][v] = at[v];
          hd[ly][mt][v] = (1.f - hv * hv) * at[v];
        }
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        hin[k] = pack4(h[ly][k][0], h[ly][k][1], h[ly][k][2], h[ly][k][3]);
        tin[k] = pack4(hd[ly][k][0], hd[ly][k][1], hd[ly][k][2], hd[ly][k][3]);
      }
    }
    // stage h1, h1dot, h2, h2dot, h3dot
    for (int mt = 0; mt < 4; mt++)
      for (int v = 0; v < 4; v++) {
        const int r = 16 * mt + 4 * grp + v, cidx = col0 + (lane & 15);
        L.x1[r][cidx] = f2bf(h[0][mt][v]);
        L.t1[r][cidx] = f2bf(hd[0][mt][v]);
        L.x2[r][cidx] = f2bf(h[1][mt][v]);
        L.t2[r][cidx] = f2bf(hd[1][mt][v]);
        L.t3[r][cidx] = f2bf(ok ? hd[2][mt][v] : 0.f);
      }
    // reverse sweep: lt = d S / d hdot, lp = d S / d h
    float lt[4][4], lp[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
      for (int v = 0; v < 4; v++) {
        lt[mt][v] = bfround(th[kSgOffW4 + 16 * mt + 4 * grp + v]);
        lp[mt][v] = 0.f;
      }
    const unsigned short *wr[3] = {w.w1, w.w2, w.w3};
#pragma unroll
    for (int ly = 2; ly >= 0; ly--) {
      float zt[4][4], zp[4][4];
#pragma unroll
      for (int mt = 0; mt < 4; mt++)
        for (int v = 0; v < 4; v++) {
          const float hv = h[ly][mt][v], q = 1.f - hv * hv;
          zt[mt][v] = lt[mt][v] * q;
          const float lpv = lp[mt][v] + lt[mt][v] * (-2.f * hv * sdot[ly][mt][v]);
          zp[mt][v] = ok ? lpv * q : 0.f;
          if (!ok) zt[mt][v] = 0.f;
        }
      unsigned short(*Z)[64] = ly == 0 ? L.z1 : (ly == 1 ? L.z2 : L.z3);
      unsigned short(*Y)[64] = ly == 0 ? L.y1 : (ly == 1 ? L.y2 : L.y3);
      for (int mt = 0; mt < 4; mt++)
        for (int v = 0; v < 4; v++) {
          const int r = 16 * mt + 4 * grp + v, cidx = col0 + (lane & 15);
          Z[r][cidx] = f2bf(zp[mt][v]);
          Y[r][cidx] = f2bf(zt[mt][v]);
        }
      if (ly == 0) break;
      bf4 bt[4], bp[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        bt[k] = pack4(zt[k][0], zt[k][1], zt[k][2], zt[k][3]);
        bp[k] = pack4(zp[k][0], zp[k][1], zp[k][2], zp[k][3]);
      }
#pragma unroll
      for (int mt = 0; mt < 4; mt++) {
        f4 a1 = {0.f, 0.f, 0.f, 0.f}, a2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; k++) {
          const bf4 A = ld_bf4(wr[ly] + (16 * mt + (lane & 15)) * kSgH + 16 * k + 4 * grp);
          a1 = mfma(A, bt[k], a1);
          a2 = mfma(A, bp[k], a2);
        }
        for (int v = 0; v < 4; v++) {
          lt[mt][v] = a1[v];
          lp[mt][v] = a2[v];
        }
      }
    }
    __syncthreads();
    // weight gradients over the 64 staged atoms: G[in][out] += X^T Z + T^T Y
    {
      unsigned short(*X)[64] = wv == 0 ? L.x0 : (wv == 1 ? L.x1 : L.x2);
      unsigned short(*Tg)[64] = wv == 0 ? L.t0 : (wv == 1 ? L.t1 : L.t2);
      unsigned short(*Z)[64] = wv == 0 ? L.z1 : (wv == 1 ? L.z2 : L.z3);
      unsigned short(*Y)[64] = wv == 0 ? L.y1 : (wv == 1 ? L.y2 : L.y3);
      const int nin = wv == 0 ? 3 : 4;
      if (wv < 3) {
#pragma unroll
        for (int mi = 0; mi < 4; mi++) {
          if (mi >= nin) continue;
#pragma unroll
          for (int mo = 0; mo < 4; mo++) {
            f4 acc = G[mi][mo];
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const int r_in = 16 * mi + (lane & 15), r_out = 16 * mo + (lane & 15);
              const int cidx = 16 * k + 4 * grp;
              bf4 A, B, A2, B2;
              for (int v = 0; v < 4; v++) {
                A[v] = (short)X[r_in][cidx + v];
                B[v] = (short)Z[r_out][cidx + v];
                A2[v] = (short)Tg[r_in][cidx + v];
                B2[v] = (short)Y[r_out][cidx + v];
              }
              acc = mfma(A, B, acc);
              acc = mfma(A2, B2, acc);
            }
            G[mi][mo] = acc;
          }
        }
      } else {
        // wave 3: b1, b2, b3 (zeta) and w4 (h3dot): one row per lane
        const int r = lane;
        for (int c = 0; c < 64; c++) {
          gb[0] += bf2f(L.z1[r][c]);
          gb[1] += bf2f(L.z2[r][c]);
          gb[2] += bf2f(L.z3[r][c]);
          gw4 += bf2f(L.t3[r][c]);
        }
      }
    }
    __syncthreads();
  }
  // write this workgroup's partial row
  float *out = partials + (size_t)blockIdx.x * kSgNParam;
  if (wv < 3) {
    const int nin = wv == 0 ? 3 : 4;
    const int off = wv == 0 ? kSgOffW1 : (wv == 1 ? kSgOffW2 : kSgOffW3);
    const int rows = wv == 0 ? kSgNF : kSgH;
    for (int mi = 0; mi < nin; mi++)
      for (int mo = 0; mo < 4; mo++)
        for (int v = 0; v < 4; v++) {
          // accumulator layout: row (input) 16 mi + 4 grp + v, column (output) 16 mo + lane%16
          const int r = 16 * mi + 4 * grp + v, c = 16 * mo + (lane & 15);
          if (r < rows) out[off + r * kSgH + c] = G[mi][mo][v];
        }
  } else {
    out[kSgOffB1 + lane] = gb[0];
    out[kSgOffB2 + lane] = gb[1];
    out[kSgOffB3 + lane] = gb[2];
    out[kSgOffW4 + lane] = gw4;
    if (lane == 0) out[kSgOffB4] = 0.f;
  }
}

__global__ __launch_bounds__(256) void k_sg_sqerr(AtomicDev d, const double *F,
                                                  const double *Fref, double *part) {
  const int i = d.blk_a0[blockIdx.x] + threadIdx.x;
  double v[1] = {0.0};
  if (i < d.blk_a1[blockIdx.x])
    for (int x = 0; x < 3; x++) {
      const double e = F[3 * (size_t)i + x] - Fref[3 * (size_t)i + x];
      v[0] += e * e;
    }
  block_sum_d<1>(v, part + blockIdx.x, 1);
}

// grad[k] = scale * sum_g partials[g][k], g ascending.
__global__ __launch_bounds__(256) void k_sg_reduce_grad(const float *partials, int nwg,
                                                        float scale, float *grad) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= kSgNParam) return;
  float acc = 0.f;
  for (int g = 0; g < nwg; g++) acc += partials[(size_t)g * kSgNParam + k];
  grad[k] = scale * acc;
}

// Adam (beta1 0.9, beta2 0.999, eps 1e-8) on the FP32 master parameters.
__global__ __launch_bounds__(256) void k_sg_adam(float *theta, const float *grad, float *m,
                                                 float *v, int t, float lr) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k >= kSgNParam) return;
  const float b1 = 0.9f, b2 = 0.999f;
  const float g = grad[k];
  const float mk = b1 * m[k] + (1.f - b1) * g, vk = b2 * v[k] + (1.f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  const float mh = mk / (1.f - powf(b1, (float)t)), vh = vk / (1.f - powf(b2, (float)t));
  theta[k] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

// BF16 operand copies: W (row-major [in][out]) and W^T, W1 padded to 48 rows.
__global__ __launch_bounds__(256) void k_sg_pack(SgWeights w) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  if (k < kSgKF * kSgH) {  // W1: [48][64]
    const int r = k / kSgH, c = k % kSgH;
    const unsigned short v = r < kSgNF ? f2bf(w.theta[kSgOffW1 + r * kSgH + c]) : 0;
    w.w1[r * kSgH + c] = v;
    w.w1t[c * kSgKF + r] = v;
  }
  if (k < kSgH * kSgH) {
    const int r = k / kSgH, c = k % kSgH;
    const unsigned short v2 = f2bf(w.theta[kSgOffW2 + k]), v3 = f2bf(w.theta[kSgOffW3 + k]);
    w.w2[k] = v2;
    w.w2t[c * kSgH + r] = v2;
    w.w3[k] = v3;
    w.w3t[c * kSgH + r] = v3;
  }
}

inline int ns_of(const AtomicParams &p) { return p.apad <= 64 ? 1 : 2; }

}  // namespace

hipError_t sg_launch_sqerr(const AtomicDev &d, const double *F, const double *Fref, double *part,
                           hipStream_t st) {
  k_sg_sqerr<<<d.nblk, 256, 0, st>>>(d, F, Fref, part);
  return hipGetLastError();
}

hipError_t sg_launch_reduce_grad(const float *partials, int nwg, float scale, float *grad,
                                 hipStream_t st) {
  k_sg_reduce_grad<<<(kSgNParam + 255) / 256, 256, 0, st>>>(partials, nwg, scale, grad);
  return hipGetLastError();
}

hipError_t sg_launch_adam(const SgWeights &w, const float *grad, float *m, float *v, int t,
                          float lr, hipStream_t st) {
  k_sg_adam<<<(kSgNParam + 255) / 256, 256, 0, st>>>(w.theta, grad, m, v, t, lr);
  return hipGetLastError();
}

hipError_t sg_launch_pack(const SgWeights &w, hipStream_t st) {
  k_sg_pack<<<(kSgH * kSgH + 255) / 256, 256, 0, st>>>(w);
  return hipGetLastError();
}

int sg_train_groups(int na) { return sg_train_groups_impl(na); }

hipError_t sg_launch_descr(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                           const SgDev &s, hipStream_t st) {
  dim3 g((p.Nm + 3) / 4);
  if (ns_of(p) == 1)
    k_sg_descr<1><<<g, 256, 0, st>>>(d, p, w, s);
  else
    k_sg_descr<2><<<g, 256, 0, st>>>(d, p, w, s);
  return hipGetLastError();
}

hipError_t sg_launch_mlp(const AtomicParams &p, const SgWeights &w, const SgDev &s,
                         hipStream_t st) {
  static int ncu = 0;  // CUs of the device (one GPU per process)
  if (!ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      ncu = 256;
  }
  const int groups = std::min((p.Na + 63) / 64, kSgMlpGroupsPerCU * ncu);
  k_sg_mlp<<<groups, 256, 0, st>>>(p.Na, w, s);
  return hipGetLastError();
}

hipError_t sg_launch_force(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                           const SgDev &s, double *out, hipStream_t st) {
  // four wavefronts per molecule for systems too small to give every SIMD
  // several molecules (the training frames), else one
  const bool split = p.Nm < 1024;
  dim3 g(split ? p.Nm : (p.Nm + 3) / 4);
  if (ns_of(p) == 1) {
    if (split)
      k_sg_force<1, 4><<<g, 256, 0, st>>>(d, p, w, s, out);
    else
      k_sg_force<1, 1><<<g, 256, 0, st>>>(d, p, w, s, out);
  } else {
    if (split)
      k_sg_force<2, 4><<<g, 256, 0, st>>>(d, p, w, s, out);
    else
      k_sg_force<2, 1><<<g, 256, 0, st>>>(d, p, w, s, out);
  }
  return hipGetLastError();
}

hipError_t sg_launch_parts(const AtomicDev &d, const SgDev &s, double *part, hipStream_t st) {
  k_sg_parts<<<d.nblk, 256, 0, st>>>(d, s, part);
  return hipGetLastError();
}

hipError_t sg_launch_u(const AtomicDev &d, const AtomicParams &p, const SgWeights &w,
                       const SgDev &s, const double *F, const double *Fref, hipStream_t st) {
  dim3 g((p.Nm + 3) / 4);
  if (ns_of(p) == 1)
    k_sg_u<1><<<g, 256, 0, st>>>(d, p, w, s, F, Fref);
  else
    k_sg_u<2><<<g, 256, 0, st>>>(d, p, w, s, F, Fref);
  return hipGetLastError();
}

hipError_t sg_launch_train(const AtomicParams &p, const SgWeights &w, const SgDev &s,
                           float *partials, hipStream_t st) {
  const size_t lds = sizeof(TrainLds);
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void *)k_sg_train,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    attr = true;
  }
  k_sg_train<<<sg_train_groups_impl(p.Na), 256, lds, st>>>(p.Na, w, s, partials);
  return hipGetLastError();
}

}  // namespace fmd
