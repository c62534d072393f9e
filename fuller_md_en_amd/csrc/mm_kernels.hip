// mm_kernels.hip -- gfx950 kernels of the molecular-mechanics force field
// (compute_forces of src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp, "MM").
//
//   k_mm_bonded   bonds, angles, dihedrals and impropers (MM:378-710): one
//                 256-thread workgroup per molecule.  The molecule's atoms
//                 and its incidence lists sit in LDS; threads evaluate terms
//                 (one each, strided over a chunk) and write each force
//                 contribution into an LDS slot; then four threads per atom
//                 sum the atom's slots.  No atomics: each atom's sum has a
//                 fixed order, and every term is evaluated once.
//   k_mm_coulomb  cut Coulomb between charged atoms of different molecules
//                 (MM:766-787), both sides per atom, position tiles in LDS.
//
// The terms are the reference's formulas with shared reciprocals in place of
// repeated divisions, and the torsions evaluate cos / sin(mult phi - gamma)
// algebraically instead of through atan2, cos and sin; each contribution
// equals the reference's to a few ulps (FP64 throughout).
//
// Intermolecular LJ (MM:711-765) is the atomic engine's LJ with the MM
// parameters: the per-atom lists (k_at_lj_alist, k_at_lj_al) in the engines,
// k_at_lj (every molecule-tile pair) in the drop-ins and in cells narrower
// than twice the list reach.
#include <climits>
#include <cstddef>

#include "fmd_atomic.hpp"
#include "fmd_wave.hpp"

namespace fmd {

namespace {

// Block sum of K doubles in a fixed order over kMmThreads threads (DPP
// wavefront sums, then the wavefronts in order); thread k < K gets sum k.
template <int K>
__device__ __forceinline__ void block_sum(const double *v, double *out) {
  constexpr int NW = kMmThreads / 64;
  __shared__ double red[NW][K];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double t = red[0][k];
    for (int w = 1; w < NW; w++) t += red[w][k];
    out[k] = t;
  }
  __syncthreads();
}

__device__ __forceinline__ const double *hh_of(HHRef hh, int r) { return hh.p + r * hh.stride; }

__device__ __forceinline__ void put3(double *s, double x, double y, double z) {
  s[0] = x;
  s[1] = y;
  s[2] = z;
}

// acc: [0] Eb, [1] Ea, [2] Ed, [3] Ei, then the virial: compat W00, W11, W22;
// physical W9.

// Bond i-j (MM:378-448): E = kb/2 (r - r0)^2.  Slot: the force on i
// (compat: the reference's -kb dr/r d, d = r_j - r_i; physical: +kb dr/r d).
template <int MODE>
__device__ __forceinline__ void mm_bond(const double *xp, int i, int j, double2 pr,
                                        const double *h, const double *hi, double *slot,
                                        double *acc) {
  double dx = xp[3 * j] - xp[3 * i], dy = xp[3 * j + 1] - xp[3 * i + 1],
         dz = xp[3 * j + 2] - xp[3 * i + 2];
  mimg(dx, dy, dz, hi, h);
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  if (r < 1e-10) {
    put3(slot, 0.0, 0.0, 0.0);
    return;
  }
  const double dr = r - pr.y;
  acc[0] += 0.5 * pr.x * dr * dr;
  const double fm = (MODE == kMmPhysical ? pr.x : -pr.x) * dr / r;
  const double fx = fm * dx, fy = fm * dy, fz = fm * dz;
  put3(slot, fx, fy, fz);
  double *W = acc + 4;
  if (MODE == kMmPhysical) {
    const double d[3] = {dx, dy, dz}, f[3] = {fx, fy, fz};
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) W[3 * x + y] -= d[x] * f[y];
  } else {
    W[0] += dx * fx;
    W[1] += dy * fy;
    W[2] += dz * fz;
  }
}

// Angle i-j-k centred on j (MM:449-493): E = kth/2 (theta - th0)^2, cos
// clamped at +-0.999999.  Slots: f_i, f_k, f_i + f_k (j receives minus the
// last).  Compat: the reference's dV = -kth dth / sin; physical: +.
template <int MODE>
__device__ __forceinline__ void mm_angle(const double *xp, int i, int j, int k, double2 pr,
                                         const double *h, const double *hi, double *slot,
                                         double *acc) {
  double a0 = xp[3 * i] - xp[3 * j], a1 = xp[3 * i + 1] - xp[3 * j + 1],
         a2 = xp[3 * i + 2] - xp[3 * j + 2];
  double c0 = xp[3 * k] - xp[3 * j], c1 = xp[3 * k + 1] - xp[3 * j + 1],
         c2 = xp[3 * k + 2] - xp[3 * j + 2];
  mimg(a0, a1, a2, hi, h);
  mimg(c0, c1, c2, hi, h);
  const double dji = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
  const double djk = sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  if (dji < 1e-10 || djk < 1e-10) {
    for (int q = 0; q < 9; q++) slot[q] = 0.0;
    return;
  }
  const double ii = 1.0 / dji, ik = 1.0 / djk, iik = ii * ik;
  double cs = (a0 * c0 + a1 * c1 + a2 * c2) * iik;
  if (cs > 0.999999) cs = 0.999999;
  if (cs < -0.999999) cs = -0.999999;
  const double dth = acos(cs) - pr.y;
  acc[1] += 0.5 * pr.x * dth * dth;
  const double sn = sqrt(1.0 - cs * cs + 1e-30);
  const double dV = (MODE == kMmPhysical ? pr.x : -pr.x) * dth / sn;
  const double gi = cs * ii * ii, gk = cs * ik * ik;
  const double rji[3] = {a0, a1, a2}, rjk[3] = {c0, c1, c2};
  double fi[3], fk[3];
  for (int c = 0; c < 3; c++) {
    fi[c] = dV * (rjk[c] * iik - gi * rji[c]);
    fk[c] = dV * (rji[c] * iik - gk * rjk[c]);
  }
  put3(slot, fi[0], fi[1], fi[2]);
  put3(slot + 3, fk[0], fk[1], fk[2]);
  put3(slot + 6, fi[0] + fk[0], fi[1] + fk[1], fi[2] + fk[2]);
  if (MODE == kMmPhysical) {
    double *W = acc + 4;
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) W[3 * x + y] += rji[x] * fi[y] + rjk[x] * fk[y];
  }
}

// Torsion p0-p1-p2-p3 (dihedral MM:494-603; improper MM:604-710 with its
// atoms stored as c1, c0, c2, c3): E = Vn/2 (1 + cos(mult phi - gamma)).
// Slots: the forces on p0..p3.  Compat: the reference's middle-atom
// coefficients; physical: the exact gradient.  Returns the energy.
// cos / sin of (mult phi - gamma) come from (cos phi, sin phi) -- phi =
// atan2(sin phi, cos phi) in the reference -- by complex powers and the
// angle-difference identity with g = (cos gamma, sin gamma) of the term: no
// transcendental call per torsion.
template <int MODE>
__device__ __forceinline__ double mm_torsion(const double *xp, int4 q, double2 pr, double2 g,
                                             int mult, const double *h, const double *hi,
                                             double *slot, double *W) {
  double b1x = xp[3 * q.y] - xp[3 * q.x], b1y = xp[3 * q.y + 1] - xp[3 * q.x + 1],
         b1z = xp[3 * q.y + 2] - xp[3 * q.x + 2];
  double b2x = xp[3 * q.z] - xp[3 * q.y], b2y = xp[3 * q.z + 1] - xp[3 * q.y + 1],
         b2z = xp[3 * q.z + 2] - xp[3 * q.y + 2];
  double b3x = xp[3 * q.w] - xp[3 * q.z], b3y = xp[3 * q.w + 1] - xp[3 * q.z + 1],
         b3z = xp[3 * q.w + 2] - xp[3 * q.z + 2];
  mimg(b1x, b1y, b1z, hi, h);
  mimg(b2x, b2y, b2z, hi, h);
  mimg(b3x, b3y, b3z, hi, h);
  const double mx = b1y * b2z - b1z * b2y, my = b1z * b2x - b1x * b2z,
               mz = b1x * b2y - b1y * b2x;
  const double nx = b2y * b3z - b2z * b3y, ny = b2z * b3x - b2x * b3z,
               nz = b2x * b3y - b2y * b3x;
  const double mm = mx * mx + my * my + mz * mz, nn = nx * nx + ny * ny + nz * nz;
  if (mm < 1e-20 || nn < 1e-20) {
    for (int k = 0; k < 12; k++) slot[k] = 0.0;
    return 0.0;
  }
  const double imm = 1.0 / mm, inn = 1.0 / nn;
  const double b2sq = b2x * b2x + b2y * b2y + b2z * b2z;
  const double b2len = sqrt(b2sq), ib2 = 1.0 / b2sq;
  const double irmn = 1.0 / sqrt(mm * nn);
  double c1 = (mx * nx + my * ny + mz * nz) * irmn;
  if (c1 > 1.0) c1 = 1.0;
  if (c1 < -1.0) c1 = -1.0;
  const double s1 = (mx * b3x + my * b3y + mz * b3z) * b2len * irmn;
  double ck = 1.0, sk = 0.0;
  for (int k = mult < 0 ? -mult : mult; k > 0; k--) {
    const double c = ck * c1 - sk * s1;
    sk = sk * c1 + ck * s1;
    ck = c;
  }
  if (mult < 0) sk = -sk;
  const double cosv = ck * g.x + sk * g.y;  // cos(mult phi - gamma)
  const double sinv = sk * g.x - ck * g.y;  // sin(mult phi - gamma)
  const double E = 0.5 * pr.x * (1.0 + cosv);
  const double dphi = -0.5 * pr.x * mult * sinv;
  const double p1 = dphi * b2len * imm, p4 = -dphi * b2len * inn;
  const double f1[3] = {p1 * mx, p1 * my, p1 * mz};
  const double f4[3] = {p4 * nx, p4 * ny, p4 * nz};
  const double ci = (b1x * b2x + b1y * b2y + b1z * b2z) * ib2;
  const double cj = (b2x * b3x + b2y * b3y + b2z * b3z) * ib2;
  double f2[3], f3[3];
  for (int c = 0; c < 3; c++) {
    if (MODE == kMmPhysical) {
      f2[c] = -f1[c] - ci * f1[c] + cj * f4[c];
      f3[c] = -f4[c] - cj * f4[c] + ci * f1[c];
    } else {
      f2[c] = -f1[c] + ci * f1[c] - cj * f4[c];
      f3[c] = -f4[c] - ci * f1[c] + cj * f4[c];
    }
  }
  put3(slot, f1[0], f1[1], f1[2]);
  put3(slot + 3, f2[0], f2[1], f2[2]);
  put3(slot + 6, f3[0], f3[1], f3[2]);
  put3(slot + 9, f4[0], f4[1], f4[2]);
  if (MODE == kMmPhysical) {
    const double b1[3] = {b1x, b1y, b1z}, b12[3] = {b1x + b2x, b1y + b2y, b1z + b2z};
    const double b123[3] = {b12[0] + b3x, b12[1] + b3y, b12[2] + b3z};
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) W[3 * x + y] += b1[x] * f2[y] + b12[x] * f3[y] + b123[x] * f4[y];
  }
  return E;
}

// One workgroup per molecule.  Term phase: thread t evaluates terms t,
// t + 256, ... of the current chunk into the LDS slots.  Gather phase: four
// threads per atom (lanes 4a..4a+3 of the atom's quad) add every fourth of
// the atom's slot entries of the chunk; the quad's four partial sums are
// added at the end in a fixed order.  NA: atoms per thread quad (64 NA atoms).
template <int MODE, int NA>
__global__ __launch_bounds__(kMmThreads) void k_mm_bonded(AtomicDev d, AtomicParams p,
                                                          HHRef hh) {
  constexpr int NV = MODE == kMmPhysical ? 13 : 7;  // energies + virial
  extern __shared__ double lds[];
  const MmDev &M = d.mm;
  const int m = blockIdx.x;
  const MmTmpl T = M.tmpl[M.mol_tmpl[m]];
  const int s0 = d.mol_start[m], n = T.natom, tid = threadIdx.x;
  const int qa = tid >> 2, qq = tid & 3;  // atom quad, lane in the quad
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  double *xp = lds;                                     // natom x 3
  double *sl = lds + 3 * M.max_natom;                   // chunk slots x 3
  int *li = reinterpret_cast<int *>(sl + 3 * M.slots);  // incidence entries
  for (int k = tid; k < 3 * n; k += kMmThreads) xp[k] = d.pos[3 * (size_t)s0 + k];
  for (int k = tid; k < T.ninc; k += kMmThreads) li[k] = M.inc[T.inc0 + k];
  int cut_lo[NA];
  double F[NA][3];
#pragma unroll
  for (int s = 0; s < NA; s++) {
    const int a = qa + 64 * s;
    cut_lo[s] = a < n ? M.inc_cut[T.cut0 + a] : 0;
    F[s][0] = F[s][1] = F[s][2] = 0.0;
  }
  double acc[NV];
  for (int k = 0; k < NV; k++) acc[k] = 0.0;
  const int nb = T.nb, nab = T.nb + T.na, nabd = nab + T.nd;
  const int sa = nb, sd = nb + 3 * T.na, si = sd + 4 * T.nd;  // slot bases
  __syncthreads();  // atoms and incidence lists in LDS
  for (int c = 0; c < T.nch; c++) {
    const int u_lo = M.ch_u[T.ch0 + c], u_hi = M.ch_u[T.ch0 + c + 1];
    const int s_lo = M.ch_s[T.ch0 + c];
    int cut_hi[NA];
#pragma unroll
    for (int s = 0; s < NA; s++) {
      const int a = qa + 64 * s;
      cut_hi[s] = a < n ? M.inc_cut[T.cut0 + (c + 1) * n + a] : 0;
    }
    for (int u = u_lo + tid; u < u_hi; u += kMmThreads) {
      const int4 q = M.tq[T.u0 + u];
      const double2 pr = M.tp[T.u0 + u];
      if (u < nb) {
        mm_bond<MODE>(xp, q.x, q.y, pr, h, hi, sl + 3 * (u - s_lo), acc);
      } else if (u < nab) {
        mm_angle<MODE>(xp, q.x, q.y, q.z, pr, h, hi, sl + 3 * (sa + 3 * (u - nb) - s_lo), acc);
      } else if (u < nabd) {
        acc[2] += mm_torsion<MODE>(xp, q, pr, M.tg[T.u0 + u], M.tm[T.u0 + u], h, hi,
                                   sl + 3 * (sd + 4 * (u - nab) - s_lo), acc + 4);
      } else {
        acc[3] += mm_torsion<MODE>(xp, q, pr, M.tg[T.u0 + u], M.tm[T.u0 + u], h, hi,
                                   sl + 3 * (si + 4 * (u - nabd) - s_lo), acc + 4);
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < NA; s++) {
      for (int k = cut_lo[s] + qq; k < cut_hi[s]; k += 4) {
        const int e = li[k];
        const double *v = sl + 3 * ((e >> 1) - s_lo);
        const double sg = (e & 1) ? -1.0 : 1.0;
        F[s][0] += sg * v[0];
        F[s][1] += sg * v[1];
        F[s][2] += sg * v[2];
      }
      cut_lo[s] = cut_hi[s];
    }
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < NA; s++) {
    const int a = qa + 64 * s;
#pragma unroll
    for (int x = 0; x < 3; x++) {
      double f = F[s][x];
      f += dpp64<0xB1>(f);  // quad_perm [1,0,3,2]
      f += dpp64<0x4E>(f);  // quad_perm [2,3,0,1]
      F[s][x] = f;
    }
    if (a < n && qq == 0) {
      double *f = d.Frebo + 3 * ((size_t)s0 + a);
      f[0] = F[s][0];
      f[1] = F[s][1];
      f[2] = F[s][2];
    }
  }
  __shared__ double tot[NV];
  block_sum<NV>(acc, tot);
  const size_t nm = p.Nm;  // part is component-major: W9 then Eb, Ea, Ed, Ei
  if (tid < NV) {
    int k;
    if (tid < 4)
      k = 9 + tid;
    else if (MODE == kMmPhysical)
      k = tid - 4;
    else
      k = 4 * (tid - 4);  // compat: the diagonal W00, W11, W22
    M.part[k * nm + m] = tot[tid];
  }
  if (MODE != kMmPhysical && tid < 6) {
    const int off[6] = {1, 2, 3, 5, 6, 7};
    M.part[off[tid] * nm + m] = 0.0;
  }
}

// One workgroup per integrator block (<= 256 atoms of one replica); thread =
// atom i, looping over every atom j of its replica in LDS tiles of 256.
template <int MODE>
__global__ __launch_bounds__(256) void k_mm_coulomb(AtomicDev d, AtomicParams p, HHRef hh) {
#pragma clang fp contract(off)
  const int b = blockIdx.x, r = d.blk_rep[b];
  const int i = d.blk_a0[b] + threadIdx.x;
  const bool own = i < d.blk_a1[b];
  const double *h = hh_of(hh, r), *hi = h + 9;
  const int j0 = d.rep_atom[r], j1 = d.rep_atom[r + 1];
  __shared__ double xs[256][3], qs[256];
  __shared__ int ms[256];
  double xi[3] = {0, 0, 0}, qi = 0.0;
  int mi = -1;
  if (own) {
    for (int k = 0; k < 3; k++) xi[k] = d.pos[3 * (size_t)i + k];
    qi = d.mm.charge[i];
    mi = d.mm.atom_mol[i];
  }
  const double rc = p.coul_rc, rc2 = p.lj.rc2d;
  double F[3] = {0, 0, 0}, acc[10];
  for (int k = 0; k < 10; k++) acc[k] = 0.0;
  for (int t0 = j0; t0 < j1; t0 += 256) {
    const int j = t0 + threadIdx.x;
    if (j < j1) {
      for (int k = 0; k < 3; k++) xs[threadIdx.x][k] = d.pos[3 * (size_t)j + k];
      qs[threadIdx.x] = d.mm.charge[j];
      ms[threadIdx.x] = d.mm.atom_mol[j];
    }
    __syncthreads();
    const int nt = min(256, j1 - t0);
    if (own) {
      for (int jj = 0; jj < nt; jj++) {
        if (ms[jj] == mi) continue;  // also j == i
        double dx = xs[jj][0] - xi[0], dy = xs[jj][1] - xi[1], dz = xs[jj][2] - xi[2];
        mimg_exact(dx, dy, dz, hi, h);
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > rc2) continue;
        // the pair's charge product in the reference's order (lower index
        // first), so both sides see the same value
        const double rr = sqrt(r2);
        const double qq = t0 + jj < i ? kCoulombK * qs[jj] * qi : kCoulombK * qi * qs[jj];
        acc[9] += 0.5 * (qq / rr - qq / rc);
        const double fm = qq / (r2 * rr);
        F[0] -= fm * dx;
        F[1] -= fm * dy;
        F[2] -= fm * dz;
        if (MODE == kMmPhysical) {
          const double dv[3] = {dx, dy, dz};
          for (int x = 0; x < 3; x++)
            for (int y = 0; y < 3; y++) acc[3 * x + y] += 0.5 * (dv[x] * fm * dv[y]);
        }
      }
    }
    __syncthreads();
  }
  if (own)
    for (int k = 0; k < 3; k++) d.F[3 * (size_t)i + k] += F[k];
  __shared__ double tot[10];
  block_sum<10>(acc, tot);
  if (threadIdx.x < 10)  // component-major
    d.mm.coul_part[(size_t)threadIdx.x * d.nblk + b] = tot[threadIdx.x];
}

}  // namespace

hipError_t mm_launch_bonded(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                            hipStream_t s) {
  const size_t shm = 3 * sizeof(double) * ((size_t)d.mm.max_natom + d.mm.slots) +
                     sizeof(int) * (size_t)d.mm.max_inc;
  const bool one = d.mm.max_natom <= 64;
  if (p.mm_mode == kMmPhysical) {
    if (one)
      k_mm_bonded<kMmPhysical, 1><<<p.Nm, kMmThreads, shm, s>>>(d, p, hh);
    else
      k_mm_bonded<kMmPhysical, 2><<<p.Nm, kMmThreads, shm, s>>>(d, p, hh);
  } else {
    if (one)
      k_mm_bonded<kMmCompat, 1><<<p.Nm, kMmThreads, shm, s>>>(d, p, hh);
    else
      k_mm_bonded<kMmCompat, 2><<<p.Nm, kMmThreads, shm, s>>>(d, p, hh);
  }
  return hipGetLastError();
}

hipError_t mm_launch_coulomb(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                             hipStream_t s) {
  if (p.mm_mode == kMmPhysical)
    k_mm_coulomb<kMmPhysical><<<d.nblk, 256, 0, s>>>(d, p, hh);
  else
    k_mm_coulomb<kMmCompat><<<d.nblk, 256, 0, s>>>(d, p, hh);
  return hipGetLastError();
}

}  // namespace fmd
