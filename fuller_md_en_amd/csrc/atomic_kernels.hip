// atomic_kernels.hip -- gfx950 kernels of the atomic (AIREBO) engine.
//
//   k_at_centers      molecule centres / radii / FP32 offset tiles, one
//                     wavefront per molecule (per step)
//   k_at_mol_nlist    molecule neighbour list, ballot compaction (rebuild)
//   k_at_rebo_nlist   same-molecule REBO list, reference order (build_nlist_rebo
//                     AB:251-261; rebuild)
//   k_at_scan_half    the reference's LJ half list walk (build_nlist_lj
//                     AB:263-272): per atom, ascending j > i of other
//                     molecules within rc + 2, first 400 kept; writes the
//                     list itself (drop-in).  FP64 with the reference's
//                     operation order.
//   k_at_scan_jmax    the same walk's truncation point jmax (engine): counts
//                     per neighbour molecule over a workgroup, then one walk
//   k_at_kick_drift   NPT first half + half-kick + drift + cell update
//                     (step_npt AB:522-568)
//   k_rebo_mol        REBO bond order (compute_rebo AB:275-421), both passes
//                     of fmd_rebo_passes.hpp, one workgroup per molecule
//   k_at_lj           intermolecular LJ (compute_lj AB:424-473): one
//                     256-thread workgroup per molecule, its 4 wavefronts
//                     split the molecule list, lane = atom, neighbour atoms
//                     wavefront-uniform through scalar loads; the half-list
//                     truncation becomes a per-pair predicate on jmax
//   k_at_kick2        second half-kick + KE partials (AB:570-581)
//   k_at_finish       reductions + NPT second half + thermo (AB:581-589, 896)
// No floating-point atomics: all sums have a fixed order.
#include <climits>
#include <cstddef>

#include "fmd_atomic.hpp"
#include "fmd_surrogate.hpp"
#include "fmd_wave.hpp"

namespace fmd {

namespace {

constexpr int kXcd = 8;

__device__ __forceinline__ double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = min(v, __shfl_xor(v, m, 64));
  return v;
}

// |mimg(pos_j - pos_i)|^2 in FP64 with the reference's operation order and
// no contraction (AB:144-147, 257-259), so list decisions at the cutoff agree
// bit for bit with the reference builder.
__device__ __noinline__ double r2_ref(const double *__restrict__ pos, int i, int j,
                                      const double *__restrict__ h,
                                      const double *__restrict__ hi) {
#pragma clang fp contract(off)
  double dx = pos[3 * j] - pos[3 * i], dy = pos[3 * j + 1] - pos[3 * i + 1],
         dz = pos[3 * j + 2] - pos[3 * i + 2];
  double s0 = hi[0] * dx + hi[1] * dy + hi[2] * dz, s1 = hi[3] * dx + hi[4] * dy + hi[5] * dz,
         s2 = hi[6] * dx + hi[7] * dy + hi[8] * dz;
  s0 -= round(s0);
  s1 -= round(s1);
  s2 -= round(s2);
  dx = h[0] * s0 + h[1] * s1 + h[2] * s2;
  dy = h[3] * s0 + h[4] * s1 + h[5] * s2;
  dz = h[6] * s0 + h[7] * s1 + h[8] * s2;
  return dx * dx + dy * dy + dz * dz;
}

__device__ __forceinline__ const double *hh_of(HHRef hh, int r) { return hh.p + r * hh.stride; }

// Sum of NW per-wavefront integer totals in a fixed order.
template <int NW>
__device__ __forceinline__ int wave_tot(const int *w) {
  int s = w[0];
  for (int k = 1; k < NW; k++) s += w[k];
  return s;
}

// Block sum of K doubles in a fixed order (NW wavefronts of 64 threads);
// sum k goes to out[k * stride].
template <int K, int NW = 4>
__device__ __forceinline__ void block_sum(const double *v, double *out, int stride = 1) {
  __shared__ double red[NW][K];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < K; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int k = threadIdx.x;
    double t = red[0][k];
    for (int w = 1; w < NW; w++) t += red[w][k];
    out[(size_t)k * stride] = t;
  }
  __syncthreads();
}

// Per-record partials of a trace-only evaluation: the block sums of the
// virial trace w and the energy e give W = (w/3) 1 and E, written as the
// kAtPart components (W9 row-major, E) at out[k * stride].
template <int NW = 4>
__device__ __forceinline__ void block_sum_trace(double w, double e, double *out, int stride) {
  const double v[2] = {w, e};
  __shared__ double tot[2];
  block_sum<2, NW>(v, tot);
  if (threadIdx.x < kAtPart) {
    const int k = threadIdx.x;
    out[(size_t)k * stride] = k == 9 ? tot[1] : (k == 0 || k == 4 || k == 8) ? tot[0] / 3.0 : 0.0;
  }
}

// ----------------------------------------------------------------- pbc
__global__ __launch_bounds__(256) void k_at_apply_pbc(double *__restrict__ pos,
                                                      const int *__restrict__ atom_rep,
                                                      HHRef hh, int n) {
#pragma clang fp contract(off)  // the reference's apply_pbc bit for bit (see inv9)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double *h = hh_of(hh, atom_rep[i]), *hi = h + 9;
  double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  double s0 = hi[0] * px + hi[1] * py + hi[2] * pz;
  double s1 = hi[3] * px + hi[4] * py + hi[5] * pz;
  double s2 = hi[6] * px + hi[7] * py + hi[8] * pz;
  s0 -= floor(s0);
  s1 -= floor(s1);
  s2 -= floor(s2);
  pos[3 * i] = h[0] * s0 + h[1] * s1 + h[2] * s2;
  pos[3 * i + 1] = h[3] * s0 + h[4] * s1 + h[5] * s2;
  pos[3 * i + 2] = h[6] * s0 + h[7] * s1 + h[8] * s2;
}

// The per-atom lists hold every pair within rc + skin at their build.  A
// pair vector (centre image + offsets) has since changed by at most the two
// molecules' drifts (centre displacement plus largest offset change, the
// replica's largest in al_dmax) and the cell change along the image,
// |dh n| <= sum of the column changes; while that stays below the skin no
// pair within rc is missing from replica r's lists.
// By one whole wavefront: lanes 0-8 carry component k = lane of the cell
// now (hn) and at the build (h0), lanes 0-7 one drift maximum each (dm,
// float bits; the others 0).  Every lane returns the verdict.
__device__ __forceinline__ bool al_stale_wave(const AtomicParams &p, double hn, double h0,
                                              int dm) {
  const double x = hn - h0, y = x * x;
  double dh = 0;
  for (int c = 0; c < 3; c++)  // column c: components c, c + 3, c + 6
    dh += sqrt(__shfl(y, c) + __shfl(y, c + 3) + __shfl(y, c + 6));
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) dm = max(dm, __shfl_xor(dm, m, 64));
  dm = max(dm, __float_as_int(kAlDriftFloor));
  const double skin = sqrt((double)p.al_r2) - sqrt(p.lj.rc2d) - 1e-3;
  return 2.0 * (double)__int_as_float(dm) + dh > skin;
}

// ------------------------------------------------------------- centres
template <int NS>
__global__ __launch_bounds__(256) void k_at_centers(AtomicDev d, AtomicParams p, HHRef hh) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const int rep = d.mol_rep[m];
  const double *h = hh_of(hh, rep), *hi = h + 9;
  // the wavefront of the replica's first molecule also judges its per-atom
  // lists (inputs loaded here, ahead of their use at the end)
  const bool judge = p.lj_alist && !p.small_box && m == d.rep_mol[rep];
  double jn = 0, j0 = 0;
  int jd = 0;
  if (judge && lane < 9) {
    jn = h[lane];
    j0 = d.al_h[9 * rep + lane];
  }
  if (judge && lane < kAlDriftSlots) jd = d.al_dmax[kAlDriftSlots * rep + lane];
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  const double rx = d.pos[3 * s0], ry = d.pos[3 * s0 + 1], rz = d.pos[3 * s0 + 2];
  double da[NS][3];
  double sx = 0, sy = 0, sz = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    da[s][0] = da[s][1] = da[s][2] = 0;
    if (a < n) {
      double dx = d.pos[3 * (s0 + a)] - rx, dy = d.pos[3 * (s0 + a) + 1] - ry,
             dz = d.pos[3 * (s0 + a) + 2] - rz;
      mimg(dx, dy, dz, hi, h);
      da[s][0] = dx;
      da[s][1] = dy;
      da[s][2] = dz;
      sx += dx;
      sy += dy;
      sz += dz;
    }
  }
  const double mx = wave_sum(sx) / n, my = wave_sum(sy) / n, mz = wave_sum(sz) / n;
  float rmax = 0.f;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < p.apad) {
      float4 o;
      if (a < n) {
        o = make_float4((float)(da[s][0] - mx), (float)(da[s][1] - my), (float)(da[s][2] - mz),
                        0.f);
        rmax = fmaxf(rmax, sqrtf(o.x * o.x + o.y * o.y + o.z * o.z));
      } else {
        o = make_float4(1e6f, 1e6f, 1e6f, 0.f);
      }
      d.off[(size_t)m * p.apad + a] = o;
    }
  }
  rmax = wave_max(rmax);
  // groups of four consecutive tile slots: centre and radius of their atoms
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    const bool real = a < n;
    float ox = 0.f, oy = 0.f, oz = 0.f, c = real ? 1.f : 0.f;
    if (real) {
      ox = (float)(da[s][0] - mx);
      oy = (float)(da[s][1] - my);
      oz = (float)(da[s][2] - mz);
    }
    float sx4 = ox, sy4 = oy, sz4 = oz, c4 = c;
    for (int k = 1; k <= 2; k <<= 1) {
      sx4 += __shfl_xor(sx4, k, 64);
      sy4 += __shfl_xor(sy4, k, 64);
      sz4 += __shfl_xor(sz4, k, 64);
      c4 += __shfl_xor(c4, k, 64);
    }
    const float inv = c4 > 0.f ? 1.f / c4 : 0.f;
    const float qx = sx4 * inv, qy = sy4 * inv, qz = sz4 * inv;
    float rq = real ? sqrtf((ox - qx) * (ox - qx) + (oy - qy) * (oy - qy) + (oz - qz) * (oz - qz))
                    : 0.f;
    for (int k = 1; k <= 2; k <<= 1) rq = fmaxf(rq, __shfl_xor(rq, k, 64));
    if ((a & 3) == 0 && a < p.apad)
      d.qcen[(size_t)m * (p.apad >> 2) + (a >> 2)] =
          c4 > 0.f ? make_float4(qx, qy, qz, rq + 1e-4f) : make_float4(1e6f, 1e6f, 1e6f, 0.f);
  }
  // whether the replica's per-atom lists must be rebuilt before this step's
  // LJ walks them (k_at_lj_al; the drifts are final after k_at_kick_drift)
  const bool redo = judge && al_stale_wave(p, jn, j0, jd);
  if (lane == 0) {
    d.center[3 * m] = rx + mx;
    d.center[3 * m + 1] = ry + my;
    d.center[3 * m + 2] = rz + mz;
    d.radius[m] = rmax;
    if (judge) d.al_redo[rep] = redo;
  }
}

// ------------------------------------------------------------ image modes
// At a list rebuild, for replica r (one workgroup each), from its cell and
// the molecule radii of this rebuild: kImgMol when rc + 2 R_max + 1 A is
// below half the narrowest cell width (every atom pair within rc + 1 A is
// then the centres' minimum image plus the two offsets, which the
// per-molecule-image kernels assume), kImgScan when the same holds for the
// half-list radius (the FP32 list walk through the tiles is then exact;
// otherwise the walk takes every pair's own FP64 image).  An engine stepping
// with per-molecule images counts the loss of kImgMol as an image fault.
__global__ __launch_bounds__(256) void k_at_image_check(AtomicDev d, AtomicParams p, HHRef hh) {
  __shared__ float wm[4];
  __shared__ int cdim[1];
  const int r = blockIdx.x;
  float rm = 0.f;
  for (int m = d.rep_mol[r] + (int)threadIdx.x; m < d.rep_mol[r + 1]; m += 256)
    rm = fmaxf(rm, d.radius[m]);
  rm = wave_max(rm);
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = rm;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double rmax = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
    const double half = 0.5 * min_width(hh_of(hh, r));
    int bits = 0;
    if (sqrt(p.lj.rc2d) + 2.0 * rmax + 1.0 < half) bits |= kImgMol;
    if (sqrt(p.lj_rlist2) + 2.0 * rmax + 1.0 < half) bits |= kImgScan;
    d.img[r] = bits;
    for (int k = 0; k < kAlDriftSlots; k++)  // largest drift since this rebuild
      d.al_dmax[kAlDriftSlots * r + k] = 0;
    d.al_redo[r] = 0;
    if (!p.small_box && !(bits & kImgMol)) atomicAdd(&d.flags[3], 1);
    // the molecule-list grid: cells at least one pair reach wide along each
    // cell vector (perpendicular width / cells), three or more per axis so
    // that the 27 neighbouring cells are distinct, at most the replica's
    // allotment; otherwise every molecule is scanned
    const int4 g = cell_grid(hh_of(hh, r), p.mol_rlist + 2.0 * rmax + 1e-3,
                             d.cell_base[r + 1] - d.cell_base[r]);
    const int ok = p.mol_cells && !p.small_box && g.w &&
                   d.rep_mol[r + 1] - d.rep_mol[r] >= p.cell_min;
    d.cell_dim[r] = make_int4(g.x, g.y, g.z, ok);
    cdim[0] = ok ? g.x * g.y * g.z : 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < cdim[0]; c += 256) d.cell_n[d.cell_base[r] + c] = 0;
}

// Molecules into their replica's cells (grids decided by k_at_image_check);
// a full cell sends the replica back to the scan of every molecule.
__global__ __launch_bounds__(256) void k_at_cell_bin(AtomicDev d, AtomicParams p, HHRef hh) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= p.Nm) return;
  const int r = d.mol_rep[m];
  const int4 cd = d.cell_dim[r];
  if (!cd.w) return;
  cell_put(m, d.center + 3 * m, hh_of(hh, r) + 9, cd, d.cell_base[r], d.cell_n, d.cell_m,
           &d.cell_dim[r].w);
}

// --------------------------------------------------------- molecule list
// Row m of the molecule list, by one wavefront (lane = 0..63): the molecules
// q != m of its replica whose centre lies within mol_rlist + R_m + R_q, in
// ascending order.  From the 27 cells around m's when its replica was binned
// (CELLS; the hits then take their places by rank, and a row of more than
// kCellHitCap hits is scanned), else by a scan of every molecule of the
// replica.
template <bool CELLS>
__device__ __forceinline__ void mol_row(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                                        int m, int lane) {
  const int r = d.mol_rep[m];
  const double *h = hh_of(hh, r), *hi = h + 9;
  const double cx = d.center[3 * m], cy = d.center[3 * m + 1], cz = d.center[3 * m + 2];
  const double rm = d.radius[m];
  auto within = [&](int q) {
    if (p.small_box) return true;
    double dx = d.center[3 * q] - cx, dy = d.center[3 * q + 1] - cy,
           dz = d.center[3 * q + 2] - cz;
    mimg(dx, dy, dz, hi, h);
    const double rl = p.mol_rlist + rm + d.radius[q];
    return dx * dx + dy * dy + dz * dz < rl * rl;
  };
  if (CELLS) {
    const int4 cd = d.cell_dim[r];
    if (cd.w) {
      __shared__ int hit[kCellHitCap], seg[64];
      const int nh = cell_hits(cd, cell_of(d.center + 3 * m, hi, cd), d.cell_base[r], d.cell_n,
                               d.cell_m, m, lane, within, hit, seg);
      if (nh <= kCellHitCap) {  // else the scan below
        cell_place(hit, nh, lane, d.mol_nl + (size_t)m * p.mol_cap, p.mol_cap);
        if (lane == 0) {
          d.mol_nl_count[m] = nh < p.mol_cap ? nh : p.mol_cap;
          if (nh > p.mol_cap) atomicAdd(&d.flags[1], 1);
        }
        return;
      }
    }
  }
  const int q0 = d.rep_mol[r], q1 = d.rep_mol[r + 1];  // molecules of this replica
  int count = 0;
  for (int n0 = q0; n0 < q1; n0 += 64) {
    const int q = n0 + lane;
    const bool in = q < q1 && q != m && within(q);
    unsigned long long bits = __ballot(in);
    int rank = __popcll(bits & ((1ull << lane) - 1ull));
    if (in && count + rank < p.mol_cap) d.mol_nl[(size_t)m * p.mol_cap + count + rank] = q;
    count += __popcll(bits);
  }
  if (lane == 0) {
    d.mol_nl_count[m] = count < p.mol_cap ? count : p.mol_cap;
    if (count > p.mol_cap) atomicAdd(&d.flags[1], 1);
  }
}

__global__ __launch_bounds__(64) void k_at_mol_nlist(AtomicDev d, AtomicParams p, HHRef hh) {
  mol_row<true>(d, p, hh, blockIdx.x, threadIdx.x);
}

// -------------------------------------------------------------- REBO list
template <int NS>
__global__ __launch_bounds__(256) void k_at_rebo_nlist(AtomicDev d, AtomicParams p, HHRef hh) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  // FP32 prefilter through the molecule's tile offsets (its atoms' minimum
  // image about the centre): only pairs within 1e-5 relative of the list
  // radius take the reference's FP64 test
  const float rl2 = (float)p.rebo_rlist2, lo = rl2 * (1.0f - 1e-5f), hi2 = rl2 * (1.0f + 1e-5f);
  const float4 *tile = d.off + (size_t)m * p.apad;
  float4 oa[NS];
  int cnt[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    cnt[s] = 0;
    const int a = lane + 64 * s;
    oa[s] = a < n ? tile[a] : make_float4(3e5f, 3e5f, 3e5f, 0.f);
  }
  for (int b = 0; b < n; b++) {
    const int j = s0 + b;
    const float4 ob = tile[b];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int a = lane + 64 * s;
      const int i = s0 + a;
      const float dx = ob.x - oa[s].x, dy = ob.y - oa[s].y, dz = ob.z - oa[s].z;
      const float r2 = dx * dx + dy * dy + dz * dz;
      if (a < n && j != i &&
          (r2 < lo || (r2 < hi2 && r2_ref(d.pos, i, j, h, hi) < p.rebo_rlist2))) {
        if (cnt[s] < p.maxr) d.nll_r[(size_t)i * p.maxr + cnt[s]] = j;
        cnt[s]++;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < n) {
      d.nlc_r[s0 + a] = cnt[s] < p.maxr ? cnt[s] : p.maxr;
      if (cnt[s] > p.maxr) atomicAdd(&d.flags[0], 1);
    }
  }
}

// ------------------------------------------------- LJ half-list walk
// Atom j (tile offset rb) inside atom i's list radius: with the FP32 tiles
// (c = the centres' minimum image minus i's offset) and the reference's FP64
// test within 1e-5 relative of the radius, or the FP64 test alone.
__device__ __forceinline__ bool scan_in(bool fast, float4 rb, float cx, float cy, float cz,
                                        float lo, float hi2, const double *pos, int i, int j,
                                        const double *h, const double *hi, double rl2) {
  if (fast) {
    const float ex = rb.x + cx, ey = rb.y + cy, ez = rb.z + cz;
    const float r2 = ex * ex + ey * ey + ez * ez;
    return r2 < lo || (r2 < hi2 && r2_ref(pos, i, j, h, hi) < rl2);
  }
  return r2_ref(pos, i, j, h, hi) < rl2;
}

// WRITE = 0: truncation point jmax per atom (tile layout) and jmax_min;
// WRITE = 1: the reference-layout half list itself (nlc_l / nll_l).
template <int NS, int WRITE>
__global__ __launch_bounds__(256) void k_at_scan_half(AtomicDev d, AtomicParams p, HHRef hh,
                                                      int *__restrict__ nlc_l,
                                                      int *__restrict__ nll_l, int cap) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  int cnt[NS], jm[NS];
  bool done[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    cnt[s] = 0;
    jm[s] = INT_MAX;
    done[s] = lane + 64 * s >= n;
  }
  // FP32 prefilter through the molecule tiles (when the cell is wide enough
  // that the centres' minimum image is every listed atom pair's, as decided
  // at this rebuild): only pairs within 1e-5 relative of the list radius
  // take the reference's FP64 test.
  const bool fast = p.scan_fast && (d.img[d.mol_rep[m]] & kImgScan);
  const float rl2 = (float)p.lj_rlist2, lo = rl2 * (1.0f - 1e-5f), hi2 = rl2 * (1.0f + 1e-5f);
  const float rl = sqrtf(rl2);
  const double cix = d.center[3 * m], ciy = d.center[3 * m + 1], ciz = d.center[3 * m + 2];
  float oax[NS], oay[NS], oaz[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    const float4 o = a < n ? d.off[(size_t)m * p.apad + a] : make_float4(3e5f, 3e5f, 3e5f, 0.f);
    oax[s] = o.x;
    oay[s] = o.y;
    oaz[s] = o.z;
  }
  const int nq = d.mol_nl_count[m];
  for (int k = 0; k < nq; k++) {
    const int q = uniform(d.mol_nl[(size_t)m * p.mol_cap + k]);
    if (q < m) continue;  // molecule ranges ascend: its atoms are all < i
    const int t0 = d.mol_start[q], tn = d.mol_start[q + 1] - t0;
    float cx[NS], cy[NS], cz[NS];
    if (fast) {
      double dx = d.center[3 * q] - cix, dy = d.center[3 * q + 1] - ciy,
             dz = d.center[3 * q + 2] - ciz;
      mimg(dx, dy, dz, hi, h);
      const float reach = rl + d.radius[q] + 0.05f;
      bool near = false;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        cx[s] = (float)(dx - (double)oax[s]);
        cy[s] = (float)(dy - (double)oay[s]);
        cz[s] = (float)(dz - (double)oaz[s]);
        near |= !done[s] && cx[s] * cx[s] + cy[s] * cy[s] + cz[s] * cz[s] < reach * reach;
      }
      if (!__any(near)) continue;  // no atom of q within the list radius
    }
    const float4 *tq = d.off + (size_t)q * p.apad;
    for (int b = 0; b < tn; b++) {
      const int j = t0 + b;
      const float4 rb = tq[b];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        if (done[s]) continue;
        const int i = s0 + lane + 64 * s;
        if (scan_in(fast, rb, cx[s], cy[s], cz[s], lo, hi2, d.pos, i, j, h, hi, p.lj_rlist2)) {
          if (WRITE) nll_l[(size_t)i * cap + cnt[s]] = j;
          cnt[s]++;
          if (cnt[s] == cap) {
            jm[s] = j;
            done[s] = true;
          }
        }
      }
    }
    bool all = true;
#pragma unroll
    for (int s = 0; s < NS; s++) all = all && done[s];
    if (__all(all)) break;
  }
  int mn = INT_MAX;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < n) {
      if (WRITE)
        nlc_l[s0 + a] = cnt[s];
      else
        d.jmax[(size_t)m * p.apad + a] = jm[s];
      mn = min(mn, jm[s]);
    } else if (!WRITE && a < p.apad) {
      d.jmax[(size_t)m * p.apad + a] = INT_MAX;
    }
  }
  mn = wave_min(mn);
  if (!WRITE && lane == 0) d.jmax_min[m] = mn;
}

// Truncation points jmax (engine path), the walk split in three phases so that
// a molecule's work spreads over a workgroup instead of one wavefront:
//  1. the four wavefronts take the molecule list's entries round robin and
//     count, per atom i of molecule m, the atoms of neighbour molecule q > m
//     inside the list radius (counts in LDS, one byte each);
//  2. per atom, the counts in ascending q give the molecule that holds its
//     cap-th entry;
//  3. that molecule's atoms are walked in ascending order to the cap-th one.
// The radius test is scan_in, the one k_at_scan_half uses, so the decisions
// are the same.
template <int NS>
__global__ __launch_bounds__(256) void k_at_scan_jmax(AtomicDev d, AtomicParams p, HHRef hh) {
  extern __shared__ unsigned char cnt_l[];  // [mol_cap][NS * 64]
  __shared__ int mnw[NS];
  constexpr int NA = NS * 64;
  const int m = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = uniform(threadIdx.x >> 6);
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  const bool fast = p.scan_fast && (d.img[d.mol_rep[m]] & kImgScan);
  const float rl2 = (float)p.lj_rlist2, lo = rl2 * (1.0f - 1e-5f), hi2 = rl2 * (1.0f + 1e-5f);
  const float rl = sqrtf(rl2);
  const float reach_m = sqrtf(hi2) + d.radius[m] + 0.05f;
  const double cix = d.center[3 * m], ciy = d.center[3 * m + 1], ciz = d.center[3 * m + 2];
  float oax[NS], oay[NS], oaz[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    const float4 o = a < n ? d.off[(size_t)m * p.apad + a] : make_float4(3e5f, 3e5f, 3e5f, 0.f);
    oax[s] = o.x;
    oay[s] = o.y;
    oaz[s] = o.z;
  }
  const int nq = d.mol_nl_count[m];
  const int *lq = d.mol_nl + (size_t)m * p.mol_cap;
  for (int k = wv; k < nq; k += 4) {
    const int q = uniform(lq[k]);
    int c[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) c[s] = 0;
    float cx[NS], cy[NS], cz[NS];
    bool near = q > m;  // molecule ranges ascend: q < m holds only atoms < i
    const int nq4 = p.apad >> 2;
    unsigned long long qm = nq4 == 64 ? ~0ull : (1ull << nq4) - 1ull;
    if (near && fast) {
      near = false;
      double dx = d.center[3 * q] - cix, dy = d.center[3 * q + 1] - ciy,
             dz = d.center[3 * q + 2] - ciz;
      mimg(dx, dy, dz, hi, h);
      const float reach = rl + d.radius[q] + 0.05f;
#pragma unroll
      for (int s = 0; s < NS; s++) {
        cx[s] = (float)(dx - (double)oax[s]);
        cy[s] = (float)(dy - (double)oay[s]);
        cz[s] = (float)(dz - (double)oaz[s]);
        near |= lane + 64 * s < n && cx[s] * cx[s] + cy[s] * cy[s] + cz[s] * cz[s] < reach * reach;
      }
      near = __any(near);
      // groups of four atoms of q within the list radius of some atom of m
      // (radius with its FP32 band, plus m's radius and the group's)
      bool qn = false;
      if (lane < nq4) {
        const float4 g = d.qcen[(size_t)q * nq4 + lane];
        const float gx = (float)dx + g.x, gy = (float)dy + g.y, gz = (float)dz + g.z;
        const float rr = reach_m + g.w;
        qn = gx * gx + gy * gy + gz * gz < rr * rr;
      }
      qm = __ballot(qn);
    }
    if (near) {
      const int t0 = d.mol_start[q], tn = d.mol_start[q + 1] - t0;
      const float4 *tq = d.off + (size_t)q * p.apad;
      while (qm) {
        const int b4 = 4 * __builtin_ctzll(qm);
        qm &= qm - 1ull;
        float4 rb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) rb[u] = tq[b4 + u];
        if (fast) {
          // the group's four FP32 tests, then the FP64 re-test of those
          // within the band (one branch per group; tile slots past the atoms
          // hold far-away sentinels, never in the band)
#pragma unroll
          for (int s = 0; s < NS; s++) {
            const bool real = lane + 64 * s < n;
            bool in[4], band = false;
            float r2[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
              const float ex = rb[u].x + cx[s], ey = rb[u].y + cy[s], ez = rb[u].z + cz[s];
              r2[u] = ex * ex + ey * ey + ez * ez;
              in[u] = real && r2[u] < lo;
              band |= real && !(r2[u] < lo) && r2[u] < hi2;
            }
            if (__any(band)) {
#pragma unroll
              for (int u = 0; u < 4; u++)
                if (real && !(r2[u] < lo) && r2[u] < hi2)
                  in[u] = r2_ref(d.pos, s0 + lane + 64 * s, t0 + b4 + u, h, hi) < p.lj_rlist2;
            }
            c[s] += (int)in[0] + (int)in[1] + (int)in[2] + (int)in[3];
          }
        } else {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            if (b4 + u >= tn) break;
#pragma unroll
            for (int s = 0; s < NS; s++)
              if (lane + 64 * s < n)
                c[s] += scan_in(fast, rb[u], cx[s], cy[s], cz[s], lo, hi2, d.pos,
                                s0 + lane + 64 * s, t0 + b4 + u, h, hi, p.lj_rlist2);
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) cnt_l[(size_t)k * NA + lane + 64 * s] = (unsigned char)c[s];
  }
  __syncthreads();
  if (wv < NS) {
    const int a = lane + 64 * wv;
    int jm = INT_MAX;
    if (a < n) {
      const int cap = p.lj_cap;
      int cum = 0, ks = -1;
      // eight counts per round: their LDS reads are issued together
      for (int k0 = 0; k0 < nq && ks < 0; k0 += 8) {
        int ck[8];
#pragma unroll
        for (int u = 0; u < 8; u++) ck[u] = k0 + u < nq ? cnt_l[(size_t)(k0 + u) * NA + a] : 0;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (ks < 0) {
            if (cum + ck[u] >= cap)
              ks = k0 + u;
            else
              cum += ck[u];
          }
        }
      }
      if (ks >= 0) {
        const int q = lq[ks];
        const int t0 = d.mol_start[q], tn = d.mol_start[q + 1] - t0;
        float cx = 0.f, cy = 0.f, cz = 0.f;
        if (fast) {
          double dx = d.center[3 * q] - cix, dy = d.center[3 * q + 1] - ciy,
                 dz = d.center[3 * q + 2] - ciz;
          mimg(dx, dy, dz, hi, h);
          const float ox = wv == 0 ? oax[0] : oax[NS - 1], oy = wv == 0 ? oay[0] : oay[NS - 1],
                      oz = wv == 0 ? oaz[0] : oaz[NS - 1];
          cx = (float)(dx - (double)ox);
          cy = (float)(dy - (double)oy);
          cz = (float)(dz - (double)oz);
        }
        const float4 *tq = d.off + (size_t)q * p.apad;
        // four tile loads per round, then the tests in ascending order
        for (int b0 = 0; b0 < tn && jm == INT_MAX; b0 += 4) {
          float4 rb[4];
#pragma unroll
          for (int u = 0; u < 4; u++) rb[u] = tq[min(b0 + u, tn - 1)];
#pragma unroll
          for (int u = 0; u < 4; u++)
            if (jm == INT_MAX && b0 + u < tn &&
                scan_in(fast, rb[u], cx, cy, cz, lo, hi2, d.pos, s0 + a, t0 + b0 + u, h, hi,
                        p.lj_rlist2) &&
                ++cum == cap)
              jm = t0 + b0 + u;
        }
      }
    }
    if (a < p.apad) d.jmax[(size_t)m * p.apad + a] = jm;
    const int mn = wave_min(jm);
    if (lane == 0) mnw[wv] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int mn = mnw[0];
    for (int s = 1; s < NS; s++) mn = min(mn, mnw[s]);
    d.jmax_min[m] = mn;
  }
}

// jmax from a reference-layout half list (the drop-in path).
template <int NS>
__global__ __launch_bounds__(256) void k_at_jmax_from_list(AtomicDev d, AtomicParams p,
                                                           const int *__restrict__ nlc_l,
                                                           const int *__restrict__ nll_l,
                                                           int cap) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= p.Nm) return;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  int mn = INT_MAX;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < p.apad) {
      int jm = INT_MAX;
      if (a < n) {
        const int i = s0 + a;
        if (nlc_l[i] >= cap) jm = nll_l[(size_t)i * cap + cap - 1];
      }
      d.jmax[(size_t)m * p.apad + a] = jm;
      mn = min(mn, jm);
    }
  }
  mn = wave_min(mn);
  if (lane == 0) d.jmax_min[m] = mn;
}

__global__ __launch_bounds__(256) void k_at_fill_int(int *__restrict__ a, int v, size_t n) {
  size_t i = blockIdx.x * 256ull + threadIdx.x;
  if (i < n) a[i] = v;
}

// ------------------------------------------------------- NPT first half
struct AtHalf1 {
  double hi[9], h1[9], hi1[9], Vg1[9];
  double xi1, sc_nh, sc_v;
};

// step_npt AB:522-555 with the phase controls of AB:875-879.  St: the
// finished state (AtomicState, or AtFinish when the finish of the previous
// step is folded into the kick-drift).
template <class St>
__device__ __forceinline__ void at_half1(const St &S, const StepCtrl &c, const AtomicParams &p,
                                         AtHalf1 &o) {
  const double hdt = 0.5 * p.dt;
  double xi = S.xi, Vg[9];
  for (int k = 0; k < 9; k++) Vg[k] = S.Vg[k];
  if (c.flags & kCtrlZeroXiVg) {
    xi = 0;
    for (int k = 0; k < 9; k++) Vg[k] = 0;
  }
  if (c.flags & kCtrlZeroVg)
    for (int k = 0; k < 9; k++) Vg[k] = 0;
  inv9(S.h, o.hi);
  const double V = fabs(det9(S.h));
  xi += hdt * (2 * S.KE - S.Nf * kB * c.Tt) / S.Q;
  xi = clampd(xi, -p.clamp_xi, p.clamp_xi);
  const double dP = (2 * S.KE + S.W9[0] + S.W9[4] + S.W9[8]) / (3 * V) * kEv2GPa - S.Pe;
  for (int a = 0; a < 3; a++) Vg[a * 4] += hdt * V * dP / (S.W * kEv2GPa);
  for (int a = 0; a < 3; a++) Vg[a * 4] = clampd(Vg[a * 4], -p.clamp_vg, p.clamp_vg);
  const double eps = Vg[0] * o.hi[0] + Vg[4] * o.hi[4] + Vg[8] * o.hi[8];
  o.sc_nh = exp(-hdt * xi);
  o.sc_v = o.sc_nh * exp(-hdt * eps / 3.0);
  for (int k = 0; k < 9; k++) {
    o.h1[k] = S.h[k] + p.dt * Vg[k];
    o.Vg1[k] = Vg[k];
  }
  inv9(o.h1, o.hi1);
  o.xi1 = xi;
}

// First half-kick, drift and wrap of atom i (AB:557-568) with the step's
// first-half NPT scalars H, its velocity first multiplied by vs (the previous
// step's rescale factor when that step's finish is folded in, else 1).
__device__ __forceinline__ void kick_drift_atom(const AtomicDev &d, const AtomicParams &p,
                                                const AtHalf1 &H, int i, int rep, int b,
                                                double vs) {
  const double hdt = 0.5 * p.dt, mi = kConv / d.mass[i];
  double v[3], r[3];
  for (int a = 0; a < 3; a++) {
    v[a] = (vs == 1.0 ? d.vel[3 * i + a] : d.vel[3 * i + a] * vs) * H.sc_v +
           hdt * d.F[3 * i + a] * mi;
    r[a] = d.pos[3 * i + a];
  }
  const double *hi = H.hi;
  double sx = hi[0] * r[0] + hi[1] * r[1] + hi[2] * r[2];
  double sy = hi[3] * r[0] + hi[4] * r[1] + hi[5] * r[2];
  double sz = hi[6] * r[0] + hi[7] * r[1] + hi[8] * r[2];
  sx += p.dt * (hi[0] * v[0] + hi[1] * v[1] + hi[2] * v[2]);
  sy += p.dt * (hi[3] * v[0] + hi[4] * v[1] + hi[5] * v[2]);
  sz += p.dt * (hi[6] * v[0] + hi[7] * v[1] + hi[8] * v[2]);
  sx -= floor(sx);
  sy -= floor(sy);
  sz -= floor(sz);
  const double *h = H.h1;
  d.pos[3 * i] = h[0] * sx + h[1] * sy + h[2] * sz;
  d.pos[3 * i + 1] = h[3] * sx + h[4] * sy + h[5] * sz;
  d.pos[3 * i + 2] = h[6] * sx + h[7] * sy + h[8] * sz;
  for (int a = 0; a < 3; a++) d.vel[3 * i + a] = v[a];
  if (p.lj_alist && !p.small_box) {
    // the atom's drift since the per-atom LJ lists were built, |h (s - s0)|
    // with s0 its fractional position then (minimum image); the pair
    // vectors have changed by at most two drifts plus the cell change
    const float4 q = d.al_pos[i];
    if (q.w != 0.f) {
      float ex = (float)sx - q.x, ey = (float)sy - q.y, ez = (float)sz - q.z;
      ex -= rintf(ex);
      ey -= rintf(ey);
      ez -= rintf(ez);
      const float dx = (float)h[0] * ex + (float)h[1] * ey + (float)h[2] * ez;
      const float dy = (float)h[3] * ex + (float)h[4] * ey + (float)h[5] * ez;
      const float dz = (float)h[6] * ex + (float)h[7] * ey + (float)h[8] * ez;
      const float dr = sqrtf(dx * dx + dy * dy + dz * dz);
      // drifts up to kAlDriftFloor count as that floor and need no atomic;
      // larger ones go to one of the replica's kAlDriftSlots maxima
      if (dr > kAlDriftFloor)
        atomicMax(d.al_dmax + kAlDriftSlots * rep + (b & (kAlDriftSlots - 1)),
                  __float_as_int(dr));
    }
  }
}

__device__ __forceinline__ void write_mid(AtomicMid &M, const AtHalf1 &H) {
  for (int k = 0; k < 9; k++) {
    M.h[k] = H.h1[k];
    M.hi[k] = H.hi1[k];
    M.Vg[k] = H.Vg1[k];
  }
  M.xi = H.xi1;
  M.sc_nh = H.sc_nh;
}

// One block per integrator block of a single replica; every thread derives
// its replica's first-half NPT scalars (redundantly, from the same inputs)
// into midpoint buffer par.
__global__ __launch_bounds__(256) void k_at_kick_drift(AtomicDev d, AtomicParams p,
                                                       int ctrl_slot, int par) {
  const int b = blockIdx.x;
  const int rep = d.blk_rep[b];
  const AtomicState &S = d.state[rep];
  const StepCtrl c = d.ctrl[(size_t)ctrl_slot * p.R + rep];
  AtHalf1 H;
  at_half1(S, c, p, H);
  if (b == d.rep_blk[rep] && threadIdx.x == 0) write_mid(d.mid[(size_t)par * p.R + rep], H);
  const int i = d.blk_a0[b] + threadIdx.x;
  if (i >= d.blk_a1[b]) return;
  kick_drift_atom(d, p, H, i, rep, b, 1.0);
}

// ------------------------------------------------------------------ REBO
// One workgroup per molecule (the REBO list never leaves a molecule): three
// wavefronts when the tile holds at most 192 atoms, else four.  Energy and
// virial: every thread's share, one fixed-order workgroup sum.
//
// Fast path (the molecule has at most rebo_slots bond entries -- list
// entries shorter than the 2.0 A switch -- and kReboTriSlots angular triples
// per atom on average, and the caller sized the LDS for them):
//   B  one thread per list entry: the FP64 minimum-image bond vector and
//      length; the bond entries are then packed atom after atom in LDS, in
//      list order, with local atom indices (entries beyond the switch add
//      nothing to any force, energy or bond order and are dropped);
//   C  one thread per bond entry: the reverse entry, the FP32 bond-order sum
//      Gs and b = (1 + Gs)^-1/2; the bond's owner (lower atom index) computes
//      the FP64 switch and pair quantities once per bond (bond_terms);
//   D  one thread per bond entry c -> p: every angular triple (c, p, t) over
//      c's other bonds t, stored once (force ft on t, fp on p);
//   E  one thread per bond entry x -> y: every force component landing on x
//      through it, in a fixed order -- the pair term, the triples centred on
//      x with partner y, those centred on y with x as the third atom and (in
//      the physical mode) with x as the partner;
// then each atom's force is the sum of its bond entries' in list order.
// Each angular triple is evaluated once (the entry-based form evaluates it
// from all three of its atoms).
//
// Otherwise the entries go through the global ReboBonds arrays in the
// reference list layout and the entry forces through rebo_fe, evaluating
// the triples where they are used (rebo_pass2_entry); both paths compute the
// same FP32 triple terms and FP64 pair terms and add them in the same order,
// and the path depends on the molecule alone.
constexpr int kReboTriSlots = 6;  // angular triples per atom (C60: 6); bond
                                  // entries per atom: AtomicParams::rebo_slots
constexpr int kReboPer = 2;       // list entries per thread in the gather
// fast-path LDS bytes per bond entry: P (double4), FP64 length and b (double
// each), bond vector (float4), fc / dfc (float2), reverse entry and owner
// list (u16 each), local neighbour, owning atom, bond index, bond flags (u8
// each)
constexpr int kReboSlotBytes = 32 + 16 + 16 + 8 + 4 + 4;
constexpr int kReboTriBytes = 24;  // ft, fp (float3 each)
constexpr size_t kReboFastLdsMax = 64 * 1024;
constexpr int kNoRev = 0xffff;  // no reverse entry (an asymmetric, overflowed list)
constexpr int kNoBond = 0x7f;   // entry beyond the switch: not one of the atom's bonds
enum : int { kEntryBond = 1, kBondVaPos = 8 };

// (plus the prefix sums of bonds, triples and list entries per atom, the
// atoms' bond counts, and each list entry's atom and bond flag)
__host__ __device__ inline size_t rebo_fast_lds(int apad, int slots) {
  return (size_t)slots * apad * kReboSlotBytes + (size_t)kReboTriSlots * apad * kReboTriBytes +
         4 * (3 * (size_t)apad + 2) + (size_t)apad + 2 * 256 * kReboPer + 16;
}
// global path: bond vectors of 64 atoms, prefix sums and owning atoms
__host__ __device__ inline size_t rebo_slow_lds(int apad, int maxr) {
  return (size_t)64 * maxr * sizeof(double4) + 4 * ((size_t)apad + 1 + (size_t)apad * maxr);
}

template <int MODE, bool TRACE, int NT>
__global__ __launch_bounds__(NT) void k_rebo_mol(AtomicDev d, AtomicParams p, HHRef hh,
                                                   int fast_ok) {
  extern __shared__ double4 lds[];
  constexpr int NW = NT / 64;
  __shared__ int wtot[NW], wtri[NW], nown;
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int m = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  if (m >= p.Nm) return;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;  // n <= NT (host check)
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  // exclusive prefix sums of the list counts (one atom per thread)
  const int c = t < n ? d.nlc_r[s0 + t] : 0;
  int v = c;
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wtot[wv] = v;
  __syncthreads();
  int off = 0;
  for (int w = 0; w < wv; w++) off += wtot[w];
  const int total = wave_tot<NW>(wtot);
  // the choice depends on the molecule alone (not on the tile width apad)
  bool fast = fast_ok && total <= NT * kReboPer;
  double E = 0, W[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int nbond = 0;  // the molecule's bond entries (list entries shorter than kDmax)
  int yq[kReboPer];         // this thread's list entries: local neighbour
  double dq[kReboPer][4];   // (-1: none), FP64 bond vector and length
#pragma unroll
  for (int q = 0; q < kReboPer; q++) yq[q] = -1;
  if (fast) {
    const int S = p.rebo_slots * p.apad, TS = kReboTriSlots * p.apad;
    char *base = reinterpret_cast<char *>(lds);
    double4 *P = reinterpret_cast<double4 *>(base);
    double *rdv = reinterpret_cast<double *>(P + S);  // FP64 length
    double *bv = rdv + S;                              // b of the direction
    float4 *bd = reinterpret_cast<float4 *>(bv + S);
    float2 *fcd = reinterpret_cast<float2 *>(bd + S);
    float *tri = reinterpret_cast<float *>(fcd + S);  // 6 floats per triple
    int *pre = reinterpret_cast<int *>(tri + 6 * TS);  // each atom's first bond entry
    int *tpre = pre + p.apad + 1;  // triple offsets, exclusive within each wavefront
    int *rpre = tpre + p.apad;     // each atom's first list entry
    unsigned short *rev = reinterpret_cast<unsigned short *>(rpre + p.apad + 1);
    unsigned short *ownl = rev + S;  // the entries that own their bond, packed
    unsigned char *nloc = reinterpret_cast<unsigned char *>(ownl + S);
    unsigned char *eat = nloc + S, *bidx = eat + S, *bflg = bidx + S;
    unsigned char *nb = bflg + S;
    unsigned char *rat = nb + p.apad, *rbd = rat + NT * kReboPer;  // per list entry: atom, bond
    if (t < n) {
      rpre[t] = off + v - c;
      for (int jn = 0; jn < c; jn++) rat[off + v - c + jn] = (unsigned char)t;
    }
    if (t == 0) {
      rpre[n] = total;
      nown = 0;
    }
    __syncthreads();
    // B: list entries, at most kReboPer per thread: local neighbour, FP64
    // minimum-image bond vector and length (kept in registers)
#pragma unroll
    for (int q = 0; q < kReboPer; q++) {
      const int e = t + NT * q;
      if (e < total) {
        const int xl = rat[e], i = s0 + xl;
        const int j = d.nll_r[(size_t)i * p.maxr + (e - rpre[xl])];
        double dx = d.pos[3 * j] - d.pos[3 * i], dy = d.pos[3 * j + 1] - d.pos[3 * i + 1],
               dz = d.pos[3 * j + 2] - d.pos[3 * i + 2];
        mimg(dx, dy, dz, hi, h);
        const double rd = sqrt(dx * dx + dy * dy + dz * dz);
        yq[q] = j - s0;
        dq[q][0] = dx;
        dq[q][1] = dy;
        dq[q][2] = dz;
        dq[q][3] = rd;
        rbd[e] = rd < rebo::kDmax;
      }
    }
    __syncthreads();
    // atoms: bond counts, the bond entries' prefix sums and the triple
    // offsets (within each wavefront); only bonds take part from here on
    // (entries beyond kDmax add nothing to any force, energy or bond order)
    int k = 0;
    if (t < n)
      for (int e = rpre[t]; e < rpre[t + 1]; e++) k += rbd[e];
    int wk = k, w3 = t < n ? k * (k - 1) : 0;
    const int ntri = w3;
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(wk, o, 64), u3 = __shfl_up(w3, o, 64);
      if (lane >= o) {
        wk += u;
        w3 += u3;
      }
    }
    if (lane == 63) {
      wtot[wv] = wk;  // the list totals are no longer needed
      wtri[wv] = w3;
    }
    if (t < n) {
      nb[t] = (unsigned char)k;
      tpre[t] = w3 - ntri;
    }
    __syncthreads();
    int koff = 0;
    for (int w = 0; w < wv; w++) koff += wtot[w];
    nbond = wave_tot<NW>(wtot);
    if (t < n) pre[t] = koff + wk - k;
    if (t == 0) pre[n] = nbond;
    if (nbond > p.rebo_slots * n) fast = false;  // more bonds than the LDS holds
  }
  if (fast) {
    const int S = p.rebo_slots * p.apad, TS = kReboTriSlots * p.apad;
    char *base = reinterpret_cast<char *>(lds);
    double4 *P = reinterpret_cast<double4 *>(base);
    double *rdv = reinterpret_cast<double *>(P + S);
    double *bv = rdv + S;
    float4 *bd = reinterpret_cast<float4 *>(bv + S);
    float2 *fcd = reinterpret_cast<float2 *>(bd + S);
    float *tri = reinterpret_cast<float *>(fcd + S);
    int *pre = reinterpret_cast<int *>(tri + 6 * TS);
    int *tpre = pre + p.apad + 1;
    int *rpre = tpre + p.apad;
    unsigned short *rev = reinterpret_cast<unsigned short *>(rpre + p.apad + 1);
    unsigned short *ownl = rev + S;
    unsigned char *nloc = reinterpret_cast<unsigned char *>(ownl + S);
    unsigned char *eat = nloc + S, *bidx = eat + S, *bflg = bidx + S;
    unsigned char *nb = bflg + S;
    unsigned char *rat = nb + p.apad, *rbd = rat + NT * kReboPer;
    const int total = nbond;
    __syncthreads();
    // the bonds into their slots, atom after atom in list order; those that
    // own their bond (lower atom index first) are listed in any order --
    // which thread takes a bond does not change its result -- so that the
    // FP64 bond terms run on as few wavefronts as possible
#pragma unroll
    for (int q = 0; q < kReboPer; q++) {
      const int e = t + NT * q;
      if (yq[q] < 0 || !rbd[e]) continue;
      const int xl = rat[e];
      int bi = 0;
      for (int f = rpre[xl]; f < e; f++) bi += rbd[f];
      const int ce = pre[xl] + bi;
      nloc[ce] = (unsigned char)yq[q];
      eat[ce] = (unsigned char)xl;
      bd[ce] = make_float4((float)dq[q][0], (float)dq[q][1], (float)dq[q][2], (float)dq[q][3]);
      rdv[ce] = dq[q][3];
      if (xl < yq[q]) ownl[atomicAdd(&nown, 1)] = (unsigned short)ce;
    }
    __syncthreads();
    // C: owners -- the bond's FP64 terms, into its own slot and its reverse's
    auto rev_of = [&](int e) {
      const int xl = eat[e], yl = nloc[e];
      int rv = kNoRev;
      for (int f = pre[yl]; f < pre[yl + 1]; f++)
        if (nloc[f] == xl) rv = f;
      return rv;
    };
    auto owner_terms = [&](int e, int rv) {
      double fd, dfd;
      double4 P4;
      int bf = bond_terms(rdv[e], fd, dfd, P4);
      if (P4.y > 0.0) bf |= kBondVaPos;
      P[e] = P4;
      const float2 fc2 = make_float2((float)fd, (float)dfd);
      fcd[e] = fc2;
      bflg[e] = (unsigned char)bf;
      if (rv != kNoRev) {
        fcd[rv] = fc2;
        bflg[rv] = (unsigned char)bf;
      }
    };
    for (int k = t; k < nown; k += NT) {
      const int e = ownl[k];
      owner_terms(e, rev_of(e));
    }
    //    bonds -- reverse entry, bond index, bond order (the P_ij slope
    //    stays in this thread's registers for its triples in D)
    float dPq[kReboPer];
#pragma unroll
    for (int q = 0; q < kReboPer; q++) {
      dPq[q] = 0.f;
      const int e = t + NT * q;
      if (e >= total) continue;
      const int xl = eat[e], yl = nloc[e];
      const int rv = rev_of(e);
      rev[e] = (unsigned short)rv;
      const int bi = e - pre[xl];
      // an entry above its neighbour whose reverse is missing (an
      // asymmetric, overflowed list) owns its bond too
      if (xl > yl && rv == kNoRev) owner_terms(e, rv);
      const float4 u = bd[e];
      float Pn;
      const float Gs = bond_order_sum<float>(
          bi, pre[xl + 1] - pre[xl], u.x, u.y, u.z, u.w,
          [&](int kn) { return bd[pre[xl] + kn]; }, p.pcc, Pn, dPq[q]);
      bv[e] = rebo::rsq((1.0 + (double)Gs) + (double)Pn);
      bidx[e] = (unsigned char)(bi | (Gs > 1e-20f ? 0x80 : 0));
    }
    __syncthreads();
    const int tri_total = wave_tot<NW>(wtri);
    // triple offset of atom x: its wavefront's scan plus the earlier wavefronts'
    auto tbase = [&](int x) {
      int o = tpre[x];
      for (int w = 0; w < (x >> 6); w++) o += wtri[w];
      return o;
    };
    auto own = [&](int e) {
      const int r = rev[e];
      return (eat[e] < nloc[e] || r == kNoRev) ? e : r;
    };
    if (tri_total > kReboTriSlots * n) {
      fast = false;  // more angular triples than the LDS holds: the global path
    } else {
      // D: angular triples (c, p, t), one thread per bond entry c -> p
#pragma unroll
      for (int q = 0; q < kReboPer; q++) {
        const int e = t + NT * q;
        if (e >= total) continue;
        const int bi = bidx[e] & kNoBond;
        const int xl = eat[e], nbx = nb[xl];
        if (bi == kNoBond || nbx < 2) continue;
        const int bf = bflg[e];
        const bool ang = (bf & kBondPair) && (bidx[e] & 0x80) && (bf & kBondVaPos);
        const float4 u = bd[e];
        const float d3[3] = {u.x, u.y, u.z};
        const float coeff = ang ? angular_coeff(P[own(e)].y, bv[e]) : 0.f;
        float *out = tri + 6 * (tbase(xl) + bi * (nbx - 1));
        for (int f = pre[xl]; f < pre[xl + 1]; f++) {
          if (f == e || (bidx[f] & kNoBond) == kNoBond) continue;
          float ft[3] = {0.f, 0.f, 0.f}, fp[3] = {0.f, 0.f, 0.f};
          if (ang && (bflg[f] & kBondThird)) {
            const float4 w4 = bd[f];
            const float dt[3] = {w4.x, w4.y, w4.z};
            const float2 fc2 = fcd[f];
            rebo::angular(d3, u.w, dt, w4.w, fc2.x, fc2.y, coeff, dPq[q], ft, fp);
            if (MODE == kReboPhysical) triple_virial<TRACE>(dt, ft, d3, fp, W);
          }
          for (int k = 0; k < 3; k++) {
            out[k] = ft[k];
            out[3 + k] = fp[k];
          }
          out += 6;
        }
      }
      __syncthreads();
      // E: every force component landing on the entry's atom, one thread per
      // entry; the forces stay in registers until every thread is done with P
      double Fq[kReboPer][3];
#pragma unroll
      for (int q = 0; q < kReboPer; q++) {
        Fq[q][0] = Fq[q][1] = Fq[q][2] = 0.0;
        const int e = t + NT * q;
        if (e >= total) continue;
        double *F = Fq[q];
        const int xl = eat[e], yl = nloc[e], rv = rev[e], bf = bflg[e];
        const int bi = bidx[e] & kNoBond;
        if (bf & kBondPair) {
          const float4 u = bd[e];
          const double bbar = 0.5 * (bv[e] + bv[rv != kNoRev ? rv : e]);
          const double4 P4 = P[own(e)];
          // pair force (AB:326-330): F_x += fpair d_xy
          const double fpair = P4.z - bbar * P4.w;
          F[0] += fpair * (double)u.x;
          F[1] += fpair * (double)u.y;
          F[2] += fpair * (double)u.z;
          if (xl < yl) {
            E += P4.x - bbar * P4.y;
            const double dd[3] = {u.x, u.y, u.z};
            if (TRACE) {
              W[0] -= fpair * (dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
            } else {
              for (int a = 0; a < 3; a++)
                for (int cc = 0; cc < 3; cc++) W[3 * a + cc] -= dd[a] * fpair * dd[cc];
            }
          }
          // triples centred on x with partner y
          const int nbx = nb[xl];
          if (bi != kNoBond && nbx >= 2) {
            const float *tp = tri + 6 * (tbase(xl) + bi * (nbx - 1));
            for (int q2 = 0; q2 < nbx - 1; q2++, tp += 6) {
              if (MODE == kReboCompat) {
                for (int k = 0; k < 3; k++) F[k] += (double)tp[k];
              } else {
                for (int k = 0; k < 3; k++) F[k] -= (double)(tp[k] + tp[3 + k]);
              }
            }
          }
        }
        // x as the third atom (and, physical, as the partner) of triples
        // centred on its neighbour y
        if (rv == kNoRev || bi == kNoBond) continue;
        const int bix = bidx[rv] & kNoBond;  // x among y's bonds
        const int nby = nb[yl];
        if (nby < 2) continue;
        const int yb = tbase(yl);
        if (bf & kBondThird) {
          for (int pb = 0; pb < nby; pb++) {
            if (pb == bix) continue;
            const float *tp = tri + 6 * (yb + pb * (nby - 1) + (bix < pb ? bix : bix - 1));
            if (MODE == kReboCompat)
              for (int k = 0; k < 3; k++) F[k] -= (double)tp[k];
            else
              for (int k = 0; k < 3; k++) F[k] += (double)tp[k];
          }
        }
        if (MODE == kReboPhysical) {
          const float *tp = tri + 6 * (yb + bix * (nby - 1));
          for (int q2 = 0; q2 < nby - 1; q2++, tp += 6)
            for (int k = 0; k < 3; k++) F[k] += (double)tp[3 + k];
        }
      }
      __syncthreads();
      double *Fe = reinterpret_cast<double *>(P);  // E is done with P
#pragma unroll
      for (int q = 0; q < kReboPer; q++) {
        const int e = t + NT * q;
        if (e < total)
          for (int k = 0; k < 3; k++) Fe[3 * e + k] = Fq[q][k];
      }
      __syncthreads();
      if (t < n) {  // atom t, its c entries
        double F[3] = {0, 0, 0};
        for (int e = pre[t]; e < pre[t + 1]; e++)
          for (int k = 0; k < 3; k++) F[k] += Fe[3 * e + k];
        for (int k = 0; k < 3; k++) d.Frebo[3 * ((size_t)s0 + t) + k] = F[k];
      }
    }
  }
  if (!fast) {
    __syncthreads();  // the fast path may have used the LDS (fast is uniform)
    double4 *vg = lds;  // [64 atoms][maxr]
    int *pre = (int *)(vg + 64 * p.maxr), *eat = pre + p.apad + 1;
    if (t < n) {
      pre[t] = off + v - c;
      for (int jn = 0; jn < c; jn++) eat[off + v - c + jn] = t;
    }
    const SlotStride slot{p.maxr};
    const int q = t & 3;
    for (int base = 0; base < n; base += NT / 4) {
      const int xl = base + (t >> 2), i = s0 + xl;
      const int cx = xl < n ? d.nlc_r[i] : 0;
      double4 *vi = vg + (size_t)(t >> 2) * p.maxr;
      for (int jn = q; jn < cx; jn += 4) {
        const int j = d.nll_r[(size_t)i * p.maxr + jn];
        double dx = d.pos[3 * j] - d.pos[3 * i], dy = d.pos[3 * j + 1] - d.pos[3 * i + 1],
               dz = d.pos[3 * j + 2] - d.pos[3 * i + 2];
        mimg(dx, dy, dz, hi, h);
        vi[jn] = make_double4(dx, dy, dz, sqrt(dx * dx + dy * dy + dz * dz));
      }
      __syncthreads();
      for (int jn = q; jn < cx; jn += 4)
        rebo_pass1<float>(i, jn, vi, d.nlc_r, d.nll_r, slot, d.bonds, p.pcc);
      __syncthreads();
    }
    for (int e = t; e < total; e += NT) {
      const int x = s0 + eat[e], yn = e - pre[eat[e]];
      double F[3] = {0, 0, 0};
      rebo_pass2_entry<float, MODE, TRACE>(x, yn, d.nlc_r, d.nll_r, slot, d.bonds, F, E, W);
      for (int k = 0; k < 3; k++) d.rebo_fe[3 * slot(x, yn) + k] = F[k];
    }
    __syncthreads();
    if (t < n) {  // atom t, its c entries
      const int x = s0 + t;
      double F[3] = {0, 0, 0};
      for (int yn = 0; yn < c; yn++)
        for (int k = 0; k < 3; k++) F[k] += d.rebo_fe[3 * slot(x, yn) + k];
      for (int k = 0; k < 3; k++) d.Frebo[3 * (size_t)x + k] = F[k];
    }
  }
  if (TRACE) {
    block_sum_trace<NW>(W[0], E, d.rebo_mpart + m, p.Nm);
  } else {
    double part[kAtPart];
    for (int k = 0; k < 9; k++) part[k] = W[k];
    part[9] = E;
    block_sum<kAtPart, NW>(part, d.rebo_mpart + m, p.Nm);
  }
}

// --------------------------------------------------------------- torsion
// The AIREBO torsion term (Stuart, Tutein & Harrison, J. Chem. Phys. 112,
// 6472 (2000)); no counterpart in the reference, whose REBO has none.  Over
// every chain i-j-k-l of bonds (i-j, j-k, k-l shorter than the switch's outer
// radius, i != l), counted once:
//   E_t = w_ij w_jk w_kl s(cos th_ijk) s(cos th_jkl) V(cos om),
//   V(c) = eps (256/405 ((1 + c)/2)^5 - 1/10) = eps (256/405 cos^10(om/2) - 1/10),
// om the chain's dihedral (0 = cis), w the REBO switch fc, s a cubic fade to
// 0 as a bond angle straightens (cos th from -0.995 to -1), where the
// dihedral is undefined.  Chains in FP32, their sums in FP64.
namespace tors {
constexpr float kC0 = -1.0f, kC1 = -0.995f;

__device__ __forceinline__ float angle_switch(float c, float &ds) {
  if (c >= kC1) {
    ds = 0.f;
    return 1.f;
  }
  if (c <= kC0) {
    ds = 0.f;
    return 0.f;
  }
  constexpr float kInv = 1.0f / (kC1 - kC0);
  const float t = (c - kC0) * kInv;
  ds = 6.f * t * (1.f - t) * kInv;
  return t * t * (3.f - 2.f * t);
}

__device__ __forceinline__ void cross(const float *a, const float *b, float *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ float dot(const float *a, const float *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Bond x -> y of the table: vector r_y - r_x (minimum image), 1 / length,
// switch and its slope.
struct Bond {
  float d[3], ir, w, dw;
};

// The chain with bonds b1 = r_j - r_i, b2 = r_k - r_j, b3 = r_l - r_k: its
// energy, and into f the force on i (END) or on j (!END),
// -dE/dr_i = dE/db1 or -dE/dr_j = dE/db2 - dE/db1.  FP32 (the energy and
// forces are summed in FP64).
template <bool END>
__device__ __forceinline__ float term(const Bond &e1, const Bond &e2, const Bond &e3, float eps,
                                      float *f) {
  const float *b1 = e1.d, *b2 = e2.d, *b3 = e3.d;
  float u1[3], u2[3], u3[3];
  for (int a = 0; a < 3; a++) {
    u1[a] = b1[a] * e1.ir;
    u2[a] = b2[a] * e2.ir;
    u3[a] = b3[a] * e3.ir;
    f[a] = 0.f;
  }
  const float cj = -dot(u1, u2), ck = -dot(u2, u3);
  float dsj, dsk;
  const float sj = angle_switch(cj, dsj), sk = angle_switch(ck, dsk);
  if (sj == 0.f || sk == 0.f) return 0.f;
  float n1[3], n2[3];
  cross(b1, b2, n1);
  cross(b2, b3, n2);
  const float q1 = dot(n1, n1), q2 = dot(n2, n2);
  if (q1 == 0.f || q2 == 0.f) return 0.f;
  // 1/|n| by the hardware reciprocal square root (1 ulp)
  const float im1 = __builtin_amdgcn_rsqf(q1), im2 = __builtin_amdgcn_rsqf(q2);
  const float c = fminf(1.f, fmaxf(-1.f, dot(n1, n2) * (im1 * im2)));
  const float hc = 0.5f * (1.f + c), hc4 = (hc * hc) * (hc * hc);
  const float V = eps * (256.f / 405.f * hc4 * hc - 0.1f);
  const float dV = eps * (256.f / 405.f * 2.5f) * hc4;
  const float Wt = e1.w * e2.w * e3.w, S = sj * sk;
  float A[3], B[3];
  for (int a = 0; a < 3; a++) {
    A[a] = (n2[a] * im2 - c * n1[a] * im1) * im1;  // dc/dn1
    B[a] = (n1[a] * im1 - c * n2[a] * im2) * im2;  // dc/dn2
  }
  const float kc = Wt * S * dV, kj = Wt * V * dsj * sk, kk = Wt * V * sj * dsk;
  // dE/db1 = kc (b2 x A) - kj (u2 + cj u1) / r1 + S V dw1 w2 w3 u1
  float g1[3], x1[3];
  cross(b2, A, x1);
  const float kw1 = S * V * e1.dw * e2.w * e3.w;
  for (int a = 0; a < 3; a++)
    g1[a] = kc * x1[a] - kj * (u2[a] + cj * u1[a]) * e1.ir + kw1 * u1[a];
  if (END) {
    for (int a = 0; a < 3; a++) f[a] = g1[a];
  } else {
    // dE/db2 = kc (A x b1 + b3 x B) - kj (u1 + cj u2) / r2 - kk (u3 + ck u2) / r2
    //          + S V w1 dw2 w3 u2
    float xa[3], xb[3];
    cross(A, b1, xa);
    cross(b3, B, xb);
    const float kw2 = S * V * e1.w * e2.dw * e3.w;
    for (int a = 0; a < 3; a++) {
      const float g2 = kc * (xa[a] + xb[a]) - kj * (u1[a] + cj * u2[a]) * e2.ir -
                       kk * (u3[a] + ck * u2[a]) * e2.ir + kw2 * u2[a];
      f[a] = g2 - g1[a];
    }
  }
  return Wt * S * V;
}
}  // namespace tors

constexpr int kTorB = 4;  // bonds per atom the torsion kernel holds (carbon cages: 3)

// One workgroup of 256 threads per molecule, TPA = 256 / NA threads per atom
// (NA = the tile width rounded up to 64): the chains that hold atom x, with x
// at an end (x-j-k-l) or in the middle (i-x-k-l), in a fixed order, dealt
// round robin to x's TPA threads, each evaluated by its thread (so a chain
// is evaluated once for each of its four atoms and nothing is accumulated
// across atoms); x's force is the sum of its threads' in thread order, added
// to the REBO forces.  A quarter of each chain's energy and
// (r_x - r_ref) (x) f_x (r_ref: the central bond's lower-index atom) go to
// the molecule's REBO partials.
template <bool TRACE, int NA>
__global__ __launch_bounds__(256) void k_rebo_tors(AtomicDev d, AtomicParams p, HHRef hh) {
  constexpr int TPA = 256 / NA;
  __shared__ int nbk[NA], bjk[NA][kTorB];
  __shared__ tors::Bond bk[NA][kTorB];
  __shared__ double fq[TPA][NA][3];
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int m = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  if (m >= p.Nm) return;
  const int t = threadIdx.x;
  const int x = t % NA, q = t / NA;  // atom, part
  const double *h = hh_of(hh, d.mol_rep[m]), *hi = h + 9;
  const int s0 = d.mol_start[m], n = d.mol_start[m + 1] - s0;
  if (q == 0 && x < n) {
    const int i = s0 + x;
    int c = 0;
    for (int jn = 0; jn < d.nlc_r[i]; jn++) {
      const int j = d.nll_r[(size_t)i * p.maxr + jn];
      double dx = d.pos[3 * j] - d.pos[3 * i], dy = d.pos[3 * j + 1] - d.pos[3 * i + 1],
             dz = d.pos[3 * j + 2] - d.pos[3 * i + 2];
      mimg(dx, dy, dz, hi, h);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      if (r >= rebo::kDmax) continue;
      if (c == kTorB) {
        atomicAdd(&d.flags[0], 1);  // more bonds than held: counted as an overflow
        break;
      }
      bjk[x][c] = j - s0;  // the REBO list never leaves the molecule
      tors::Bond &b = bk[x][c];
      b.d[0] = (float)dx;
      b.d[1] = (float)dy;
      b.d[2] = (float)dz;
      b.ir = (float)(1.0 / r);
      b.w = (float)rebo::fc(r);
      b.dw = (float)rebo::dfc(r);
      c++;
    }
    nbk[x] = c;
  }
  __syncthreads();
  double F[3] = {0.0, 0.0, 0.0}, E = 0.0, W[TRACE ? 1 : 9];
  for (int k = 0; k < (TRACE ? 1 : 9); k++) W[k] = 0.0;
  const float eps = (float)p.tors_eps;
  auto add = [&](float e4, const float *f, const float *rel) {
    E += 0.25 * (double)e4;
    for (int a = 0; a < 3; a++) F[a] += (double)f[a];
    if (TRACE) {
      W[0] += (double)rel[0] * f[0] + (double)rel[1] * f[1] + (double)rel[2] * f[2];
    } else {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) W[3 * a + b] += (double)rel[a] * f[b];
    }
  };
  if (x < n) {
    int ord = 0;  // the chain's place in x's sequence; this thread takes ord = q mod TPA
    // x at an end: chains x - j - k - l
    for (int a = 0; a < nbk[x]; a++) {
      const int j = bjk[x][a];
      for (int b = 0; b < nbk[j]; b++) {
        const int k = bjk[j][b];
        if (k == x) continue;
        for (int c = 0; c < nbk[k]; c++) {
          const int l = bjk[k][c];
          if (l == j || l == x) continue;
          if (ord++ % TPA != q) continue;
          const tors::Bond e1 = bk[x][a], e2 = bk[j][b], e3 = bk[k][c];
          float f[3];
          const float e = tors::term<true>(e1, e2, e3, eps, f);
          // r_x - r_ref: -b1 (ref j) or -(b1 + b2) (ref k)
          float rel[3];
          for (int u = 0; u < 3; u++) rel[u] = j < k ? -e1.d[u] : -(e1.d[u] + e2.d[u]);
          add(e, f, rel);
        }
      }
    }
    // x in the middle: chains i - x - k - l
    for (int b = 0; b < nbk[x]; b++) {
      const int k = bjk[x][b];
      for (int a = 0; a < nbk[x]; a++) {
        if (a == b) continue;
        const int i = bjk[x][a];
        for (int c = 0; c < nbk[k]; c++) {
          const int l = bjk[k][c];
          if (l == x || l == i) continue;
          if (ord++ % TPA != q) continue;
          tors::Bond e1 = bk[x][a];
          for (int u = 0; u < 3; u++) e1.d[u] = -e1.d[u];  // r_x - r_i
          const tors::Bond e2 = bk[x][b], e3 = bk[k][c];
          float f[3];
          const float e = tors::term<false>(e1, e2, e3, eps, f);
          // r_x - r_ref: 0 (ref x) or -b2 (ref k)
          float rel[3];
          for (int u = 0; u < 3; u++) rel[u] = x < k ? 0.f : -e2.d[u];
          add(e, f, rel);
        }
      }
    }
  }
  for (int a = 0; a < 3; a++) fq[q][x][a] = F[a];
  // the molecule's energy and virial into its REBO partials (written by
  // k_rebo_mol earlier on the stream)
  __shared__ double tot[TRACE ? 2 : kAtPart];
  if (TRACE) {
    const double v[2] = {W[0], E};
    block_sum<2, 4>(v, tot);
  } else {
    double v[kAtPart];
    for (int k = 0; k < 9; k++) v[k] = W[k];
    v[9] = E;
    block_sum<kAtPart, 4>(v, tot);
  }
  if (q == 0 && x < n) {  // block_sum ended with a barrier: fq is complete
    for (int a = 0; a < 3; a++) {
      double s = fq[0][x][a];
      for (int u = 1; u < TPA; u++) s += fq[u][x][a];
      d.Frebo[3 * (size_t)(s0 + x) + a] += s;
    }
  }
  if (t < kAtPart) {
    double add_k;
    if (TRACE)
      add_k = t == 9 ? tot[1] : (t == 0 || t == 4 || t == 8) ? tot[0] / 3.0 : 0.0;
    else
      add_k = tot[t];
    d.rebo_mpart[(size_t)t * p.Nm + m] += add_k;
  }
}

// -------------------------------------------------------------------- LJ
// One group of four neighbour atoms against the lane's atom(s).
// PRED: the reference half-list truncation (pair kept iff j <= jmax(i) for
// the lower index i).  PERPAIR: per-pair minimum image (cells narrower than
// twice the list reach).
// PERPAIR re-images every pair vector, which would fold the far-away
// sentinels of idle lanes and padding sites back into the cell, so that path
// masks them explicitly (lane_ok, jb <= j_last).
template <int NS, bool PRED, bool PERPAIR>
__device__ __forceinline__ void lj_group(const float4 *rb, const int *jb_jmax, int jb0,
                                         int j_last, const bool *lane_ok,
                                         const float *cx, const float *cy, const float *cz,
                                         const int *ia, const int *ia_jmax, bool j_above,
                                         const LJ32 &lj, const float *hf, const float *hif,
                                         const double *pos, const double *h, const double *hi,
                                         float *gx, float *gy, float *gz, float *e,
                                         float *wd) {
#pragma unroll
  for (int u = 0; u < 4; u++) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      float ex = rb[u].x + cx[s], ey = rb[u].y + cy[s], ez = rb[u].z + cz[s];
      if (PERPAIR) {
        float s0 = hif[0] * ex + hif[1] * ey + hif[2] * ez;
        float s1 = hif[3] * ex + hif[4] * ey + hif[5] * ez;
        float s2 = hif[6] * ex + hif[7] * ey + hif[8] * ez;
        s0 -= rintf(s0);
        s1 -= rintf(s1);
        s2 -= rintf(s2);
        ex = hf[0] * s0 + hf[1] * s1 + hf[2] * s2;
        ey = hf[3] * s0 + hf[4] * s1 + hf[5] * s2;
        ez = hf[6] * s0 + hf[7] * s1 + hf[8] * s2;
      }
      const float r2 = ex * ex + ey * ey + ez * ez;
      bool in = r2 <= lj.rc2_hi;
      if (PERPAIR) in = in && lane_ok[s] && jb0 + u <= j_last;
      // within FP32 rounding of the cutoff: the reference's own FP64 test
      if (in && r2 > lj.rc2_lo) in = r2_ref(pos, ia[s], jb0 + u, h, hi) <= lj.rc2d;
      if (PRED) in = in && (j_above ? (jb0 + u <= ia_jmax[s]) : (ia[s] <= jb_jmax[u]));
      if (in) {
        const float fm = lj_fm(r2, lj, e[s]);
        gx[s] += fm * ex;
        gy[s] += fm * ey;
        gz[s] += fm * ez;
        if (PERPAIR) {
          wd[6 * s + 0] += fm * ex * ex;
          wd[6 * s + 1] += fm * ex * ey;
          wd[6 * s + 2] += fm * ex * ez;
          wd[6 * s + 3] += fm * ey * ey;
          wd[6 * s + 4] += fm * ey * ez;
          wd[6 * s + 5] += fm * ez * ez;
        }
      }
    }
  }
}

// Per lane (atom a of molecule i, offset r_a) and neighbour molecule j
// (minimum-image centre vector dm): G_aj = sum_b f_ab, FP32 over the <= apad
// atoms b, promoted to FP64.  Force on a: -48 eps sum_j G_aj (plus its REBO
// force); virial 1/2 sum_j dm (x) G_aj - r_a (x) sum_j G_aj (the
// reference's sum over pairs of d (x) f, AB:465-472, regrouped through the
// symmetric pair set); energy 1/2 per ordered pair.
// TRACE (time steps on the molecule-image path): only the virial trace, the
// part the barostat uses (AB:527, 583), one FP64 FMA per (lane, neighbour)
// instead of nine; the diagonal then carries trace / 3.
// The body for molecule mi (a whole workgroup); k_at_lj is one workgroup per
// molecule.
template <int NS, bool PERPAIR, bool TRACE>
__device__ __forceinline__ void lj_tiles(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                                         int mi) {
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const double *h = hh_of(hh, d.mol_rep[mi]), *hi = h + 9;
  float hf[9], hif[9];
  for (int k = 0; k < 9; k++) {
    hf[k] = (float)h[k];
    hif[k] = (float)hi[k];
  }
  const int apad = p.apad;
  const int s0 = d.mol_start[mi], n = d.mol_start[mi + 1] - s0;
  const double cix = d.center[3 * mi], ciy = d.center[3 * mi + 1], ciz = d.center[3 * mi + 2];
  float rax[NS], ray[NS], raz[NS];
  int ia[NS], iam[NS];  // atom index, its truncation point
  bool lane_ok[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < n) {
      const float4 r = d.off[(size_t)mi * apad + a];
      rax[s] = r.x;
      ray[s] = r.y;
      raz[s] = r.z;
      iam[s] = d.jmax[(size_t)mi * apad + a];
    } else {
      rax[s] = ray[s] = raz[s] = -3.0e5f;
      iam[s] = INT_MAX;
    }
    ia[s] = a < n ? s0 + a : s0;  // idle lanes never pass the cutoff test
    lane_ok[s] = a < n;
  }
  double GA[NS][3], WA[NS][TRACE ? 1 : 9];
  float EA[NS], WD[NS][6];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    for (int k = 0; k < 3; k++) GA[s][k] = 0;
    for (int k = 0; k < (TRACE ? 1 : 9); k++) WA[s][k] = 0;
    for (int k = 0; k < 6; k++) WD[s][k] = 0;
    EA[s] = 0;
  }
  const LJ32 lj = p.lj;
  // quad cull radius: rc (with its FP32 band) + the radius of molecule i
  const float reach_i = sqrtf(lj.rc2_hi) + d.radius[mi] + 0.05f;
  const int i_last = s0 + n - 1;
  const int jmin_i = d.jmax_min[mi];
  const int nq = d.mol_nl_count[mi];
  const int *lq = d.mol_nl + (size_t)mi * p.mol_cap;
  for (int k = wv; k < nq; k += 4) {
    const int mj = uniform(lq[k]);
    double dx = d.center[3 * mj] - cix, dy = d.center[3 * mj + 1] - ciy,
           dz = d.center[3 * mj + 2] - ciz;
    mimg(dx, dy, dz, hi, h);
    const float rcut = sqrtf(lj.rc2) + d.radius[mj] + 0.05f;
    float cx[NS], cy[NS], cz[NS];
    bool act = PERPAIR;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      // dm - r_a in FP64, rounded once per lane (no shared rounding of dm)
      cx[s] = (float)(dx - (double)rax[s]);
      cy[s] = (float)(dy - (double)ray[s]);
      cz[s] = (float)(dz - (double)raz[s]);
      act |= cx[s] * cx[s] + cy[s] * cy[s] + cz[s] * cz[s] < rcut * rcut;
    }
    if (!__any(act)) continue;
    const int t0 = d.mol_start[mj];
    const int j_last = d.mol_start[mj + 1] - 1;
    const bool j_above = mj > mi;
    // truncation can only bite where the lower-index side has a finite jmax
    const bool pred = p.lj_compat && (j_above ? (j_last > jmin_i) : (i_last > d.jmax_min[mj]));
    float gx[NS], gy[NS], gz[NS], e[NS], wd[NS * 6];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      gx[s] = gy[s] = gz[s] = e[s] = 0.f;
      for (int q = 0; q < 6; q++) wd[6 * s + q] = 0.f;
    }
    const float4 *sj = d.off + (size_t)mj * apad;
    const int *jj = d.jmax + (size_t)mj * apad;
    // groups of four atoms of mj within reach of molecule i (all of them on
    // the per-pair image path)
    const int nq = apad >> 2;
    unsigned long long qmask = nq == 64 ? ~0ull : ((1ull << nq) - 1ull);
    if (!PERPAIR) {
      bool near = false;
      if (lane < nq) {
        const float4 c = d.qcen[(size_t)mj * nq + lane];
        const float qx = (float)dx + c.x, qy = (float)dy + c.y, qz = (float)dz + c.z;
        const float rr = reach_i + c.w;
        near = qx * qx + qy * qy + qz * qz < rr * rr;
      }
      qmask = __ballot(near);
    }
    while (qmask) {
      const int b = 4 * __builtin_ctzll(qmask);
      qmask &= qmask - 1ull;
      float4 rb[4];
      int jm[4];
#pragma unroll
      for (int u = 0; u < 4; u++) rb[u] = sj[b + u];
      if (pred) {
#pragma unroll
        for (int u = 0; u < 4; u++) jm[u] = jj[b + u];
        lj_group<NS, true, PERPAIR>(rb, jm, t0 + b, j_last, lane_ok, cx, cy, cz, ia, iam,
                                    j_above, lj, hf, hif,
                                    d.pos, h, hi, gx, gy, gz, e, wd);
      } else {
        lj_group<NS, false, PERPAIR>(rb, jm, t0 + b, j_last, lane_ok, cx, cy, cz, ia, iam,
                                     j_above, lj, hf, hif,
                                     d.pos, h, hi, gx, gy, gz, e, wd);
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double g0 = gx[s], g1 = gy[s], g2 = gz[s];
      GA[s][0] += g0;
      GA[s][1] += g1;
      GA[s][2] += g2;
      if (TRACE) {
        WA[s][0] += dx * g0 + dy * g1 + dz * g2;
      } else if (!PERPAIR) {
        WA[s][0] += dx * g0;
        WA[s][1] += dx * g1;
        WA[s][2] += dx * g2;
        WA[s][3] += dy * g0;
        WA[s][4] += dy * g1;
        WA[s][5] += dy * g2;
        WA[s][6] += dz * g0;
        WA[s][7] += dz * g1;
        WA[s][8] += dz * g2;
      }
      for (int q = 0; q < 6; q++) WD[s][q] += wd[6 * s + q];
      EA[s] += e[s];
    }
  }
  __shared__ double fsum[4][2 * 64][3];
  double acc[10];
  for (int k = 0; k < 10; k++) acc[k] = 0;
  const double c48 = 48.0 * p.eps;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    for (int k = 0; k < 3; k++) fsum[wv][lane + 64 * s][k] = -c48 * GA[s][k];
    if (PERPAIR) {
      // sum over ordered pairs of 1/2 fm d d^T (symmetric)
      const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
      for (int q = 0; q < 9; q++) acc[q] += 0.5 * c48 * (double)WD[s][map[q]];
    } else if (TRACE) {
      const double tr = c48 * (0.5 * WA[s][0] - (rax[s] * GA[s][0] + ray[s] * GA[s][1] +
                                                 raz[s] * GA[s][2]));
      acc[0] += tr / 3.0;
      acc[4] += tr / 3.0;
      acc[8] += tr / 3.0;
    } else {
      const double ra[3] = {rax[s], ray[s], raz[s]};
      for (int x = 0; x < 3; x++)
        for (int y = 0; y < 3; y++)
          acc[3 * x + y] += c48 * (0.5 * WA[s][3 * x + y] - ra[x] * GA[s][y]);
    }
    acc[9] += (double)EA[s];
  }
  __shared__ double red[4][10];
  for (int k = 0; k < 10; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  // per-atom forces: the 4 wavefronts' partial sums in a fixed order
  for (int a = threadIdx.x; a < n; a += 256) {
    const int x = s0 + a;
    for (int k = 0; k < 3; k++) {
      const double f = ((fsum[0][a][k] + fsum[1][a][k]) + fsum[2][a][k]) + fsum[3][a][k];
      d.F[3 * x + k] = d.Frebo[3 * x + k] + f;
    }
  }
  if (threadIdx.x < 10) {
    const int k = threadIdx.x;
    double v = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    if (k == 9) v *= 2.0 * p.eps;  // 1/2 * 4 eps * sum of reduced energies
    d.lj_part[(size_t)k * p.Nm + mi] = v;
  }
}

template <int NS, bool PERPAIR, bool TRACE>
__global__ __launch_bounds__(256) void k_at_lj(AtomicDev d, AtomicParams p, HHRef hh) {
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int bid = blockIdx.x;
  const int mi = (bid % kXcd) * per + bid / kXcd;
  if (mi >= p.Nm) return;
  lj_tiles<NS, PERPAIR, TRACE>(d, p, hh, mi);
}

// ------------------------------------------------- LJ per-atom lists
// Engines on cells with one image per molecule pair keep, per tile slot a of
// molecule i, a list of the atoms within rc + 1 A at the list rebuild, so the
// step evaluates about 1.4 candidates per pair in range instead of the 3.9 of
// the tile loops of k_at_lj.  The list carries the reference's half-list
// truncation (the PRED test of lj_group, applied once at the rebuild).  The
// build stages each neighbour's set-up in LDS and its four wavefronts take
// every neighbour's live four-atom groups by rank (k + r = w mod 4), each
// writing its own segment of the atom's list, which the same wavefront of the
// step then walks.  (Splitting the molecule list itself round robin is
// unbalanced: lattice order can put most near neighbours in one residue of
// k mod 4.)
constexpr int kAlWaves = 4;
constexpr int kAlMolCap = 160;  // the molecule list stride these kernels stage in LDS
// padding entry: neighbour slot 255 (a far-away centre offset in the step's
// table, tile slot 0 of molecule 0), never within the cutoff
constexpr unsigned kAlPad = 255u << 8;

// The lists of molecule mi's atoms, by the whole workgroup.  FULL (the
// scheduled rebuild) also records the replica's cell; a rebuild within a
// step (stale lists) leaves that to k_at_kick2, since the other workgroups
// of the step's LJ kernel still read it.
template <int NS, bool FULL>
__device__ __forceinline__ void al_build(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                                         int mi) {
  // per neighbour k: centre offset (FP64 minimum image), live four-atom
  // groups (0: none within reach), molecule, first atom, atom count, PRED
  __shared__ double dmk[kAlMolCap][3];
  __shared__ unsigned long long qmk[kAlMolCap];
  __shared__ int mjk[kAlMolCap], t0k[kAlMolCap], tnk[kAlMolCap], prk[kAlMolCap];
  constexpr int NA = NS * 64;
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const double *h = hh_of(hh, d.mol_rep[mi]), *hi = h + 9;
  const int apad = p.apad;
  const int s0 = d.mol_start[mi], n = d.mol_start[mi + 1] - s0;
  const double cix = d.center[3 * mi], ciy = d.center[3 * mi + 1], ciz = d.center[3 * mi + 2];
  float rax[NS], ray[NS], raz[NS];
  int ia[NS], iam[NS], cnt[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    if (a < n) {
      const float4 r = d.off[(size_t)mi * apad + a];
      rax[s] = r.x;
      ray[s] = r.y;
      raz[s] = r.z;
      iam[s] = d.jmax[(size_t)mi * apad + a];
    } else {
      rax[s] = ray[s] = raz[s] = -3.0e5f;  // never within reach
      iam[s] = INT_MAX;
    }
    ia[s] = s0 + a;
  }
  const float rl2 = p.al_r2, rl = sqrtf(rl2);
  // the state these lists are built from: tile offsets, centre, cell; the
  // drift since then is tracked by k_at_centers
  if (wv == 0) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int a = lane + 64 * s;
      if (a < n) {
        const double *x = d.pos + 3 * (s0 + a);
        double fx = hi[0] * x[0] + hi[1] * x[1] + hi[2] * x[2];
        double fy = hi[3] * x[0] + hi[4] * x[1] + hi[5] * x[2];
        double fz = hi[6] * x[0] + hi[7] * x[1] + hi[8] * x[2];
        d.al_pos[s0 + a] = make_float4((float)(fx - floor(fx)), (float)(fy - floor(fy)),
                                       (float)(fz - floor(fz)), 1.f);
      }
    }
    const int r = d.mol_rep[mi];
    if (FULL && mi == d.rep_mol[r] && lane < 9) d.al_h[9 * r + lane] = h[lane];
  }
  const float reach_i = rl + d.radius[mi] + 0.05f;
  const int i_last = s0 + n - 1;
  const int jmin_i = d.jmax_min[mi];
  const int nq = d.mol_nl_count[mi];
  const int *lq = d.mol_nl + (size_t)mi * p.mol_cap;
  const int cap = p.al_cap;
  const int nq4 = apad >> 2;
  // neighbour set-up, once per neighbour (wavefront w takes k = w, w + 4, ...)
  for (int k = wv; k < nq; k += kAlWaves) {
    const int mj = uniform(lq[k]);
    double dx = d.center[3 * mj] - cix, dy = d.center[3 * mj + 1] - ciy,
           dz = d.center[3 * mj + 2] - ciz;
    mimg(dx, dy, dz, hi, h);
    const float rcut = rl + d.radius[mj] + 0.05f;
    bool act = false;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const float cx = (float)(dx - (double)rax[s]), cy = (float)(dy - (double)ray[s]),
                  cz = (float)(dz - (double)raz[s]);
      act |= cx * cx + cy * cy + cz * cz < rcut * rcut;
    }
    bool near = false;
    if (lane < nq4) {
      const float4 c = d.qcen[(size_t)mj * nq4 + lane];
      const float qx = (float)dx + c.x, qy = (float)dy + c.y, qz = (float)dz + c.z;
      const float rr = reach_i + c.w;
      near = qx * qx + qy * qy + qz * qz < rr * rr;
    }
    const unsigned long long qm = __ballot(near);
    const bool any_act = __any(act);
    if (lane == 0) {
      const int t0 = d.mol_start[mj], tn = d.mol_start[mj + 1] - t0;
      dmk[k][0] = dx;
      dmk[k][1] = dy;
      dmk[k][2] = dz;
      qmk[k] = any_act ? qm : 0ull;
      mjk[k] = mj;
      t0k[k] = t0;
      tnk[k] = tn;
      const bool j_above = mj > mi;
      prk[k] = p.lj_compat && (j_above ? (t0 + tn - 1 > jmin_i) : (i_last > d.jmax_min[mj]));
    }
  }
  __syncthreads();
  // every pair kept by the list among this wavefront's share: of each
  // neighbour's live groups, those of rank w, w + 4, ...
  auto walk = [&](auto emit) {
    for (int k = 0; k < nq; k++) {
      unsigned long long qm = qmk[k];
      if (!qm) continue;
      const int mj = mjk[k], t0 = t0k[k], tn = tnk[k];
      const bool pred = prk[k], j_above = mj > mi;
      float cx[NS], cy[NS], cz[NS];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        cx[s] = (float)(dmk[k][0] - (double)rax[s]);
        cy[s] = (float)(dmk[k][1] - (double)ray[s]);
        cz[s] = (float)(dmk[k][2] - (double)raz[s]);
      }
      const float4 *sj = d.off + (size_t)mj * apad;
      const int *jj = d.jmax + (size_t)mj * apad;
      // this wavefront's groups: live groups whose rank r among the live ones
      // (ascending) has r + k = w mod 4 (the wavefront that takes a
      // neighbour's extra group rotates with k), picked by one ballot
      bool mine = false;
      if (lane < 64 && ((qm >> lane) & 1ull)) {
        const int r = __popcll(qm & ((1ull << lane) - 1ull));
        mine = ((r + k) & (kAlWaves - 1)) == wv;
      }
      unsigned long long my = __ballot(mine);
      while (my) {
        const int b4 = 4 * __builtin_ctzll(my);
        my &= my - 1ull;
        float4 rb[4];
        int jm[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          rb[u] = sj[b4 + u];  // slots past the atoms hold far-away sentinels
          jm[u] = pred ? jj[b4 + u] : 0;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int b = b4 + u;
          if (b >= tn) break;
#pragma unroll
          for (int s = 0; s < NS; s++) {
            const float ex = rb[u].x + cx[s], ey = rb[u].y + cy[s], ez = rb[u].z + cz[s];
            bool in = ex * ex + ey * ey + ez * ez <= rl2;
            if (pred) in = in && (j_above ? (t0 + b <= iam[s]) : (ia[s] <= jm[u]));
            if (in) emit(s, k << 8 | b);
          }
        }
      }
    }
  };
  // the entries go out eight at a time (one 16-byte block per lane) into
  // this wavefront's segment of the atom's list, padded to whole blocks
  const int nbk = cap >> 3;  // blocks per segment
  uint4 *out = reinterpret_cast<uint4 *>(d.al_e) + ((size_t)mi * kAlWaves + wv) * nbk * NA;
  unsigned bw[NS][4];
  int fill[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    cnt[s] = 0;
    fill[s] = 0;
    for (int k = 0; k < 4; k++) bw[s][k] = 0;
  }
  auto push = [&](int s, unsigned en) {
    // shift the block right by one entry, the new one entering at the top
    bw[s][0] = (bw[s][0] >> 16) | (bw[s][1] << 16);
    bw[s][1] = (bw[s][1] >> 16) | (bw[s][2] << 16);
    bw[s][2] = (bw[s][2] >> 16) | (bw[s][3] << 16);
    bw[s][3] = (bw[s][3] >> 16) | (en << 16);
    if (++fill[s] == 8) {
      if (cnt[s] < nbk)
        out[(size_t)cnt[s] * NA + lane + 64 * s] = make_uint4(bw[s][0], bw[s][1], bw[s][2], bw[s][3]);
      cnt[s]++;
      fill[s] = 0;
    }
  };
  walk([&](int s, int en) { push(s, (unsigned)en); });
#pragma unroll
  for (int s = 0; s < NS; s++) {
    while (fill[s] != 0) push(s, kAlPad);
    d.al_n[((size_t)mi * kAlWaves + wv) * NA + lane + 64 * s] =
        (unsigned short)(cnt[s] < nbk ? cnt[s] : nbk);
    if (cnt[s] > nbk) atomicAdd(&d.flags[4], 1);  // entries lost: the host turns to k_at_lj
  }
}

template <int NS>
__global__ __launch_bounds__(256) void k_at_lj_alist(AtomicDev d, AtomicParams p, HHRef hh) {
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int mi = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  if (mi >= p.Nm) return;
  al_build<NS, true>(d, p, hh, mi);
}


// The lists of molecule mi rebuilt from this step's positions: its row of
// the molecule list, then its atoms' lists, keeping the truncation points of
// the scheduled rebuild (every workgroup of the replica decides alike;
// k_at_kick2 records the rebuilt state).
template <int NS>
__device__ __forceinline__ void al_relist(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                                          int mi) {
  if (threadIdx.x < 64) mol_row<false>(d, p, hh, mi, threadIdx.x);
  __threadfence();
  __syncthreads();
  al_build<NS, false>(d, p, hh, mi);
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0 && mi == d.rep_mol[d.mol_rep[mi]]) atomicAdd(&d.flags[2], 1);
}

template <int NS>
__global__ __launch_bounds__(256) void k_at_lj_relist(AtomicDev d, AtomicParams p, HHRef hh) {
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int mi = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  if (mi >= p.Nm || !d.al_redo[d.mol_rep[mi]]) return;
  al_relist<NS>(d, p, hh, mi);
}

// LJ from the per-atom lists: lane = atom slot of molecule i, wavefront w
// walks quarter w of the atom's blocks (its four build segments one after the
// other); each block is one 16-byte load of eight
// entries, taken four at a time (the neighbours' centre offsets from LDS,
// then their tile slots, then the pairs).  Pair vector
// r_b - r_a = (off_b - off_a) + (centre_j - centre_i) in FP32 with the
// centres' minimum image from FP64; the cutoff is the reference's (FP64
// re-test within 1e-5 of it).  Sums: FP32 over 32 entries, then FP64.
// Virial: 1/2 sum over ordered pairs of fm d (x) d (trace only on time steps).
template <int NS, bool TRACE>
__global__ __launch_bounds__(256) void k_at_lj_al(AtomicDev d, AtomicParams p, HHRef hh) {
  __shared__ float4 dmk[256];  // centre offset of neighbour k, its tile's byte offset
  __shared__ int j0k[256];      // its first atom
  const int per = (p.Nm + kXcd - 1) / kXcd;
  const int mi = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  if (mi >= p.Nm) return;
  constexpr int NA = NS * 64;
  constexpr int NW = TRACE ? 1 : 6;
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const int rep = d.mol_rep[mi];
  // Stale lists (k_at_centers decided for the replica) are rebuilt from this
  // step's positions before the walk: here for one atom per lane; for two
  // the rebuild's registers would cost the walk a wavefront per SIMD, and
  // k_at_lj_relist runs before this kernel instead.
  if (NS == 1 && __builtin_expect(d.al_redo[rep] != 0, 0)) al_relist<NS>(d, p, hh, mi);
  const double *h = hh_of(hh, rep), *hi = h + 9;
  const int apad = p.apad;
  const int s0 = d.mol_start[mi], n = d.mol_start[mi + 1] - s0;
  const int nq = d.mol_nl_count[mi];
  {
    const int *lq = d.mol_nl + (size_t)mi * p.mol_cap;
    const double cix = d.center[3 * mi], ciy = d.center[3 * mi + 1], ciz = d.center[3 * mi + 2];
    for (int t = threadIdx.x; t < nq; t += 256) {
      const int q = lq[t];
      double dx = d.center[3 * q] - cix, dy = d.center[3 * q + 1] - ciy,
             dz = d.center[3 * q + 2] - ciz;
      mimg(dx, dy, dz, hi, h);
      dmk[t] = make_float4((float)dx, (float)dy, (float)dz,
                           __uint_as_float((unsigned)(q * apad) * (unsigned)sizeof(float4)));
      j0k[t] = d.mol_start[q];
    }
    if (threadIdx.x == 0) {
      dmk[kAlPad >> 8] = make_float4(1.0e6f, 1.0e6f, 1.0e6f, __uint_as_float(0u));
      j0k[kAlPad >> 8] = 0;
    }
  }
  __syncthreads();
  const LJ32 lj = p.lj;
  const float4 *off = d.off;
  double GA[NS][3], WA[NS][NW], EA[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int a = lane + 64 * s;
    for (int k = 0; k < 3; k++) GA[s][k] = 0;
    for (int k = 0; k < NW; k++) WA[s][k] = 0;
    EA[s] = 0;
    // The atom's list is its four build segments one after the other; this
    // wavefront walks the quarter [lo, lo + cnt) of their blocks.  (Segment
    // w holds the groups the build's wavefront w took, and their lengths
    // differ from atom to atom.)
    const int nbk = p.al_cap >> 3;
    int p1 = 0, p2 = 0, p3 = 0, tot = 0;
    if (a < n) {
      const unsigned short *cn = d.al_n + (size_t)mi * kAlWaves * NA + a;
      p1 = cn[0];
      p2 = p1 + cn[NA];
      p3 = p2 + cn[2 * NA];
      tot = p3 + cn[3 * NA];
    }
    const int lo = (tot * wv) >> 2, cnt = ((tot * (wv + 1)) >> 2) - lo;
    const uint4 *lst = reinterpret_cast<const uint4 *>(d.al_e) +
                       (size_t)mi * kAlWaves * nbk * NA + a;
    // block t of the concatenation: segment (t >= p1) + (t >= p2) + (t >= p3)
    auto blk_at = [&](int t) {
      const int pre = t >= p3 ? p3 : t >= p2 ? p2 : t >= p1 ? p1 : 0;
      const int seg = (t >= p1) + (t >= p2) + (t >= p3);
      return lst[(size_t)(seg * nbk + (t - pre)) * NA];
    };
    float rx = 0.f, ry = 0.f, rz = 0.f;
    if (a < n) {
      const float4 r = off[(size_t)mi * apad + a];
      rx = r.x;
      ry = r.y;
      rz = r.z;
    }
    const int ia = s0 + a;
    float g[3] = {0.f, 0.f, 0.f}, e = 0.f, w[NW];
    for (int k = 0; k < NW; k++) w[k] = 0.f;
    auto flush = [&]() {
      for (int k = 0; k < 3; k++) {
        GA[s][k] += (double)g[k];
        g[k] = 0.f;
      }
      for (int k = 0; k < NW; k++) {
        WA[s][k] += (double)w[k];
        w[k] = 0.f;
      }
      EA[s] += (double)e;
      e = 0.f;
    };
    const char *offb = reinterpret_cast<const char *>(off);
    uint4 nxt = cnt > 0 ? blk_at(lo) : make_uint4(0u, 0u, 0u, 0u);
    for (int ib = 0; ib < cnt; ib++) {
      // the next block's load is in flight while this one is walked
      const uint4 blk = nxt;
      if (ib + 1 < cnt) nxt = blk_at(lo + ib + 1);
      const unsigned wd[4] = {blk.x, blk.y, blk.z, blk.w};
#pragma unroll
      for (int h4 = 0; h4 < 2; h4++) {
        // four entries, their centre offsets, then their tile slots: each
        // group of loads in flight
        unsigned en[4];
#pragma unroll
        for (int u = 0; u < 4; u++) en[u] = (wd[2 * h4 + (u >> 1)] >> (16 * (u & 1))) & 0xffffu;
        float4 dm[4], rb[4];
#pragma unroll
        for (int u = 0; u < 4; u++) dm[u] = dmk[en[u] >> 8];
#pragma unroll
        for (int u = 0; u < 4; u++)
          rb[u] = *reinterpret_cast<const float4 *>(
              offb + (__float_as_uint(dm[u].w) + ((en[u] & 255u) << 4)));
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const float ex = (rb[u].x - rx) + dm[u].x, ey = (rb[u].y - ry) + dm[u].y,
                      ez = (rb[u].z - rz) + dm[u].z;
          const float r2 = ex * ex + ey * ey + ez * ez;
          bool in = r2 <= lj.rc2_hi;
          // within FP32 rounding of the cutoff: the reference's own FP64 test
          if (in && r2 > lj.rc2_lo)
            in = r2_ref(d.pos, ia, j0k[en[u] >> 8] + (int)(en[u] & 255u), h, hi) <= lj.rc2d;
          if (in) {
            const float fm = lj_fm(r2, lj, e);
            g[0] += fm * ex;
            g[1] += fm * ey;
            g[2] += fm * ez;
            if (TRACE) {
              w[0] += fm * r2;
            } else {
              w[0] += fm * ex * ex;
              w[1] += fm * ex * ey;
              w[2] += fm * ex * ez;
              w[3] += fm * ey * ey;
              w[4] += fm * ey * ez;
              w[5] += fm * ez * ez;
            }
          }
        }
      }
      if ((ib & 3) == 3) flush();  // every 32 entries
    }
    flush();
  }
  __shared__ double fsum[4][2 * 64][3];
  // the virial's 9 components and the energy (time steps: the trace, in
  // acc[0], and the energy)
  constexpr int NR = TRACE ? 2 : 10;
  double acc[NR];
  for (int k = 0; k < NR; k++) acc[k] = 0;
  const double c48 = 48.0 * p.eps;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    for (int k = 0; k < 3; k++) fsum[wv][lane + 64 * s][k] = -c48 * GA[s][k];
    if (TRACE) {
      acc[0] += 0.5 * c48 * WA[s][0];
    } else {
      const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
      for (int q = 0; q < 9; q++) acc[q] += 0.5 * c48 * WA[s][map[q]];
    }
    acc[NR - 1] += EA[s];
  }
  __shared__ double red[4][10];
  for (int k = 0; k < NR; k++) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  for (int a = threadIdx.x; a < n; a += 256) {
    const int x = s0 + a;
    for (int k = 0; k < 3; k++) {
      const double f = ((fsum[0][a][k] + fsum[1][a][k]) + fsum[2][a][k]) + fsum[3][a][k];
      d.F[3 * x + k] = d.Frebo[3 * x + k] + f;
    }
  }
  if (threadIdx.x < 10) {
    const int k = threadIdx.x;
    const int kr = TRACE ? (k == 9 ? 1 : 0) : k;
    double v = ((red[0][kr] + red[1][kr]) + red[2][kr]) + red[3][kr];
    if (k == 9) v *= 2.0 * p.eps;  // 1/2 * 4 eps * sum of reduced energies
    else if (TRACE) v = (k == 0 || k == 4 || k == 8) ? v / 3.0 : 0.0;
    d.lj_part[(size_t)k * p.Nm + mi] = v;
  }
}

// ------------------------------------------------------- kick 2 + partials
__global__ __launch_bounds__(256) void k_at_kick2(AtomicDev d, AtomicParams p, int kick,
                                                  int par) {
  if (kick && p.lj_alist && !p.small_box && threadIdx.x == 0) {
    // lists rebuilt within this step's LJ kernel: their cell, drift zero
    const int r = d.blk_rep[blockIdx.x];
    if (blockIdx.x == d.rep_blk[r] && d.al_redo[r]) {
      const double *hm = d.mid[(size_t)par * p.R + r].h;
      for (int k = 0; k < 9; k++) d.al_h[9 * r + k] = hm[k];
      for (int k = 0; k < kAlDriftSlots; k++) d.al_dmax[kAlDriftSlots * r + k] = 0;
      d.al_redo[r] = 0;
    }
  }
  const int i = d.blk_a0[blockIdx.x] + threadIdx.x;
  // sums of this block: kinetic energy, then over this block's share of its
  // replica's molecules (contiguous, equal counts) the time steps'
  // trace-only partials, which the finish reads per block: AIREBO the LJ +
  // REBO virial trace and the two energies; MM the LJ trace and energy and
  // the 13 bonded components
  double v2[kKePart];
  for (int k = 0; k < kKePart; k++) v2[k] = 0;
  {
    const int b = blockIdx.x, r = d.blk_rep[b];
    const int nbr = d.rep_blk[r + 1] - d.rep_blk[r];
    const int m0 = d.rep_mol[r], nmr = d.rep_mol[r + 1] - m0;
    const int per = (nmr + nbr - 1) / nbr;  // <= 256: a block holds 256 atoms
    const int t = (b - d.rep_blk[r]) * per + (int)threadIdx.x;
    if ((int)threadIdx.x < per && t < nmr) {
      const size_t m = (size_t)m0 + t, nm = p.Nm;
      if (p.ff == kFfMM) {
        v2[1] = d.lj_part[m];
        v2[2] = d.lj_part[9 * nm + m];
        for (int k = 0; k < 13; k++) v2[3 + k] = d.mm.part[k * nm + m];
      } else {
        v2[1] = d.lj_part[m] + d.rebo_mpart[m];
        v2[2] = d.lj_part[9 * nm + m];
        v2[3] = d.rebo_mpart[9 * nm + m];
      }
    }
  }
  if (i < d.blk_a1[blockIdx.x]) {
    double vx = d.vel[3 * i], vy = d.vel[3 * i + 1], vz = d.vel[3 * i + 2];
    if (kick) {
      const AtomicMid &M = d.mid[(size_t)par * p.R + d.blk_rep[blockIdx.x]];
      const double hdt = 0.5 * p.dt, mi = kConv / d.mass[i];
      const double eps2 = M.Vg[0] * M.hi[0] + M.Vg[4] * M.hi[4] + M.Vg[8] * M.hi[8];
      const double sc_v2 = M.sc_nh * exp(-hdt * eps2 / 3.0);
      vx = (vx + hdt * d.F[3 * i] * mi) * sc_v2;
      vy = (vy + hdt * d.F[3 * i + 1] * mi) * sc_v2;
      vz = (vz + hdt * d.F[3 * i + 2] * mi) * sc_v2;
      d.vel[3 * i] = vx;
      d.vel[3 * i + 1] = vy;
      d.vel[3 * i + 2] = vz;
    }
    v2[0] = d.mass[i] * (vx * vx + vy * vy + vz * vz);
  }
  if (p.ff == kFfMM)
    block_sum<kKePart>(v2, d.ke_part + blockIdx.x, d.nblk);
  else
    block_sum<4>(v2, d.ke_part + blockIdx.x, d.nblk);
}

// Fixed-order sums of the partial arrays: s[0] = sum m v^2, s[1..9] = W9,
// s[10] = E of the intramolecular term (REBO, or the MM bonded total),
// s[11] = E_lj, s[12..15] = MM Eb, Ea, Ed, Ei, s[16] = E_coulomb.
// Replica r's sums (one workgroup; the strided order is the same whatever
// the replica's offset, so a replica sums bitwise as it would alone).
constexpr int kAtSums = 17;
// threads that reduce a replica's partials: the finish kernel's workgroup,
// and every integrator block of the kick-drift that folds a finish in (both
// must add in the same order)
constexpr int kFinThreads = 256;
// The partial arrays are component-major (component k of record r at
// k * records + r), so the strided loops below read them coalesced.
// ALL = false (kick-drift blocks that only need the finished scalars, not
// the record): the time-step path of the AIREBO engine leaves the energy
// sums out; every sum it does take is the full reduction's, bit for bit.
template <bool MM, bool ALL = true>
__device__ __forceinline__ void at_reduce(const AtomicDev &d, const AtomicParams &p, int r,
                                          double *sums) {
  double v[kAtSums];
  for (int k = 0; k < kAtSums; k++) v[k] = 0;
  const size_t nb = d.nblk, nm = p.Nm;
  const int b0 = d.rep_blk[r], b1 = d.rep_blk[r + 1];
  for (int b = b0 + threadIdx.x; b < b1; b += kFinThreads) {
    v[0] += d.ke_part[b];
    if (!MM && p.use_sg) {
      for (int k = 0; k < 9; k++) v[1 + k] += d.rebo_part[k * nb + b];
      v[10] += d.rebo_part[9 * nb + b];
    } else if (d.mm.charge) {
      for (int k = 0; k < 9; k++) v[1 + k] += d.mm.coul_part[k * nb + b];
      v[16] += d.mm.coul_part[9 * nb + b];
    }
  }
  const int m0 = d.rep_mol[r], m1 = d.rep_mol[r + 1];
  if (p.lj_trace && !p.use_sg && !p.small_box) {
    // time steps: the LJ and REBO partials hold the virial trace / 3 on the
    // diagonal and zeros off it, so two of their ten components carry all
    // (the same sums as below, in the same order; the per-pair-image LJ
    // kernel of narrow cells keeps the full tensor)
    if (MM) {
      // the integrator blocks' sums of their molecules (k_at_kick2)
      for (int b = b0 + threadIdx.x; b < b1; b += kFinThreads) {
        const double lt = d.ke_part[nb + b];
        double x[13];
#pragma unroll
        for (int k = 0; k < 9; k++) x[k] = d.ke_part[(3 + k) * nb + b];
        v[1] += lt;
        v[5] += lt;
        v[9] += lt;
        for (int k = 0; k < 9; k++) v[1 + k] += x[k];
        if (ALL) {
#pragma unroll
          for (int k = 9; k < 13; k++) x[k] = d.ke_part[(3 + k) * nb + b];
          v[11] += d.ke_part[2 * nb + b];
          for (int k = 0; k < 4; k++) v[12 + k] += x[9 + k];
        }
      }
    } else {
      // the integrator blocks' sums of their molecules (k_at_kick2)
      for (int b = b0 + threadIdx.x; b < b1; b += kFinThreads) {
        const double ct = d.ke_part[nb + b];
        v[1] += ct;
        v[5] += ct;
        v[9] += ct;
        if (ALL) {
          v[11] += d.ke_part[2 * nb + b];
          v[10] += d.ke_part[3 * nb + b];
        }
      }
    }
  } else {
  for (int m = m0 + threadIdx.x; m < m1; m += kFinThreads) {
    double x[MM ? kAtPart + 13 : 2 * kAtPart];
#pragma unroll
    for (int k = 0; k < kAtPart; k++) x[k] = d.lj_part[k * nm + m];
    if (MM) {
#pragma unroll
      for (int k = 0; k < 13; k++) x[kAtPart + k] = d.mm.part[k * nm + m];
    } else if (!p.use_sg) {
#pragma unroll
      for (int k = 0; k < kAtPart; k++) x[kAtPart + k] = d.rebo_mpart[k * nm + m];
    }
    for (int k = 0; k < 9; k++) v[1 + k] += x[k];
    v[11] += x[9];
    if (MM) {
      for (int k = 0; k < 9; k++) v[1 + k] += x[kAtPart + k];
      for (int k = 0; k < 4; k++) v[12 + k] += x[kAtPart + 9 + k];
    } else if (!p.use_sg) {
      for (int k = 0; k < 9; k++) v[1 + k] += x[kAtPart + k];
      v[10] += x[kAtPart + 9];
    }
  }
  }
  __shared__ double tot[kAtSums];
  block_sum<kAtSums, kFinThreads / 64>(v, tot);
  for (int k = 0; k < kAtSums; k++) sums[k] = tot[k];
  if (MM) {
    sums[10] = ((sums[12] + sums[13]) + sums[14]) + sums[15];
    // the reference's MM virial holds the diagonal only (MM:434-441, 749-757)
    if (p.mm_mode == kMmCompat)
      for (int k : {1, 2, 3, 5, 6, 7}) sums[1 + k] = 0.0;
  }
}

__device__ __forceinline__ void at_set_energies(AtomicState &S, const double *s) {
  S.Ep_rebo = s[10];
  S.Ep_lj = s[11];
  S.Eb = s[12];
  S.Ea = s[13];
  S.Ed = s[14];
  S.Ei = s[15];
  S.Ec = s[16];
  S.Ep = (s[10] + s[11]) + s[16];
}

// One workgroup per replica.
template <bool MM>
__global__ __launch_bounds__(kFinThreads) void k_at_observe(AtomicDev d, AtomicParams p,
                                                            int what) {
  const int r = blockIdx.x;
  double s[kAtSums];
  at_reduce<MM>(d, p, r, s);
  if (threadIdx.x != 0) return;
  AtomicState &S = d.state[r];
  if (what & kAtObsKE) S.KE = 0.5 * s[0] / kConv;
  if (what & kAtObsForce) {
    for (int k = 0; k < 9; k++) S.W9[k] = s[1 + k];
    at_set_energies(S, s);
  }
}

// The second half of step_npt (AB:570-590) and the rescale control
// (AB:898-908) of a replica from its sums s, midpoint M, state constants and
// the step's control c: the finished thermostat, cell velocity, energies and
// the velocity rescale factor.
struct AtFinish {
  double h[9], xi, Vg[9], KE, W9[9], Q, W, Pe, T, P, V, scale, Nf;
};
__device__ __forceinline__ void at_finish_math(const AtomicState &S, const AtomicMid &M,
                                               const StepCtrl &c, const AtomicParams &p,
                                               const double *s, AtFinish &o) {
  const double Nf = S.Nf, Q = S.Q, Pe = S.Pe, Wb = S.W;
  const double hdt = 0.5 * p.dt;
  double KE = 0.5 * s[0] / kConv;
  for (int k = 0; k < 9; k++) o.W9[k] = s[1 + k];
  double xi = M.xi;
  for (int k = 0; k < 9; k++) o.Vg[k] = M.Vg[k];
  xi += hdt * (2 * KE - Nf * kB * c.Tt) / Q;
  xi = clampd(xi, -p.clamp_xi, p.clamp_xi);
  const double V2 = fabs(det9(M.h));
  const double P = (2 * KE + o.W9[0] + o.W9[4] + o.W9[8]) / (3 * V2) * kEv2GPa;
  for (int a = 0; a < 3; a++) o.Vg[a * 4] += hdt * V2 * (P - Pe) / (Wb * kEv2GPa);
  for (int a = 0; a < 3; a++) o.Vg[a * 4] = clampd(o.Vg[a * 4], -p.clamp_vg, p.clamp_vg);
  double T = 2 * KE / (Nf * kB);
  double scale = 1.0;
  if ((c.flags & kCtrlRescale) && T > 0.1) {  // AB:898-908
    const double tgt = c.rescale_T > 0.1 ? c.rescale_T : 0.1;
    scale = sqrt(tgt / T);
    KE *= scale * scale;
    T = 2 * KE / (Nf * kB);
    xi = 0;
    if (c.flags & kCtrlRescaleZeroVg)
      for (int k = 0; k < 9; k++) o.Vg[k] = 0;
  }
  for (int k = 0; k < 9; k++) o.h[k] = M.h[k];
  o.xi = xi;
  o.KE = KE;
  o.T = T;
  o.P = P;
  o.V = V2;
  o.scale = scale;
  o.Q = Q;
  o.W = Wb;
  o.Pe = Pe;
  o.Nf = Nf;
}

// The finished state and the step's record (one thread).
__device__ __forceinline__ void at_finish_write(const AtomicDev &d, const AtomicParams &p, int r,
                                                const AtomicMid &M, const StepCtrl &c,
                                                const double *s, const AtFinish &o,
                                                int thermo_slot) {
  AtomicState &S = d.state[r];
  for (int k = 0; k < 9; k++) {
    S.h[k] = M.h[k];
    S.hi[k] = M.hi[k];
    S.Vg[k] = o.Vg[k];
    S.W9[k] = o.W9[k];
  }
  S.xi = o.xi;
  S.Tt = c.Tt;
  S.KE = o.KE;
  at_set_energies(S, s);
  S.scale = o.scale;
  S.step += 1;
  Thermo &t = d.thermo[(size_t)thermo_slot * p.R + r];
  t.Ep = S.Ep;
  t.KE = o.KE;
  t.KEt = o.KE;
  t.KEr = 0;
  t.T = o.T;
  t.P = o.P;
  t.V = o.V;
  t.xi = o.xi;
  t.Ep_rebo = S.Ep_rebo;
  t.Ep_lj = S.Ep_lj;
  for (int k = 0; k < 9; k++) t.h[k] = M.h[k];
  t.Eb = S.Eb;
  t.Ea = S.Ea;
  t.Ed = S.Ed;
  t.Ei = S.Ei;
  t.Ecoul = S.Ec;
}

// One workgroup per replica: the finish of the step whose midpoint is in
// buffer par.
template <bool MM>
__global__ __launch_bounds__(kFinThreads) void k_at_finish(AtomicDev d, AtomicParams p,
                                                           int ctrl_slot, int thermo_slot,
                                                           int par) {
  const int r = blockIdx.x;
  // the state, midpoint and control before the reduction: their latency
  // overlaps it
  const AtomicState &S = d.state[r];
  const AtomicMid M = d.mid[(size_t)par * p.R + r];
  const StepCtrl c = d.ctrl[(size_t)ctrl_slot * p.R + r];
  double s[kAtSums];
  at_reduce<MM>(d, p, r, s);
  if (threadIdx.x != 0) return;
  AtFinish o;
  at_finish_math(S, M, c, p, s, o);
  at_finish_write(d, p, r, M, c, s, o, thermo_slot);
}

// The kick-drift of a step with the previous step's finish folded in: every
// block reduces its replica's partials itself (the finish kernel's
// reduction, the same sums in the same order), derives the finished state,
// rescales its atoms' velocities by the previous step's factor and takes the
// first-half NPT scalars from that state; the replica's first block writes
// the finished state and record and the new midpoint (buffer par; the
// previous step's is in 1 - par).
template <bool MM>
__global__ __launch_bounds__(256) void k_at_kick_drift_fin(AtomicDev d, AtomicParams p,
                                                           int ctrl_slot, int prev_slot,
                                                           int par) {
  const int b = blockIdx.x;
  const int rep = d.blk_rep[b];
  const AtomicState &S = d.state[rep];
  const AtomicMid M = d.mid[(size_t)(1 - par) * p.R + rep];
  const StepCtrl cp = d.ctrl[(size_t)prev_slot * p.R + rep];
  const StepCtrl c = d.ctrl[(size_t)ctrl_slot * p.R + rep];
  double s[kAtSums];
  // the replica's first block takes every sum for the record, the others
  // only those the finished scalars need (uniform per block)
  const bool first = b == d.rep_blk[rep];
  if (first)
    at_reduce<MM>(d, p, rep, s);
  else
    at_reduce<MM, false>(d, p, rep, s);
  AtFinish o;
  at_finish_math(S, M, cp, p, s, o);
  const bool lead = first && threadIdx.x == 0;
  if (lead) at_finish_write(d, p, rep, M, cp, s, o, prev_slot);
  AtHalf1 H;
  at_half1(o, c, p, H);
  if (lead) write_mid(d.mid[(size_t)par * p.R + rep], H);
  const int i = d.blk_a0[b] + threadIdx.x;
  if (i >= d.blk_a1[b]) return;
  kick_drift_atom(d, p, H, i, rep, b, o.scale);
}

// Velocity rescale by each replica's factor (1 where its control did not ask).
__global__ __launch_bounds__(256) void k_at_rescale(AtomicDev d, AtomicParams p) {
  const int i = d.blk_a0[blockIdx.x] + threadIdx.x;
  if (i >= d.blk_a1[blockIdx.x]) return;
  const double s = d.state[d.blk_rep[blockIdx.x]].scale;
  for (int a = 0; a < 3; a++) d.vel[3 * i + a] *= s;
}

inline int ns_of(const AtomicParams &p) { return p.apad <= 64 ? 1 : 2; }

}  // namespace

hipError_t at_launch_apply_pbc(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                               hipStream_t s) {
  k_at_apply_pbc<<<(p.Na + 255) / 256, 256, 0, s>>>(d.pos, d.atom_rep, hh, p.Na);
  return hipGetLastError();
}

hipError_t at_launch_centers(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                             hipStream_t s) {
  dim3 g((p.Nm + 3) / 4);
  if (ns_of(p) == 1)
    k_at_centers<1><<<g, 256, 0, s>>>(d, p, hh);
  else
    k_at_centers<2><<<g, 256, 0, s>>>(d, p, hh);
  return hipGetLastError();
}

hipError_t at_launch_lists(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                           int rebo_list, hipStream_t s) {
  k_at_image_check<<<p.R, 256, 0, s>>>(d, p, hh);
  if (p.mol_cells) k_at_cell_bin<<<(p.Nm + 255) / 256, 256, 0, s>>>(d, p, hh);
  k_at_mol_nlist<<<p.Nm, 64, 0, s>>>(d, p, hh);
  dim3 g((p.Nm + 3) / 4);
  const bool one = ns_of(p) == 1;
  if (rebo_list && p.ff != kFfMM) {
    if (one)
      k_at_rebo_nlist<1><<<g, 256, 0, s>>>(d, p, hh);
    else
      k_at_rebo_nlist<2><<<g, 256, 0, s>>>(d, p, hh);
  }
  if (p.lj_compat) {
    if (one)
      k_at_scan_jmax<1><<<p.Nm, 256, (size_t)p.mol_cap * 64, s>>>(d, p, hh);
    else
      k_at_scan_jmax<2><<<p.Nm, 256, (size_t)p.mol_cap * 128, s>>>(d, p, hh);
  } else {
    const size_t n = (size_t)p.Nm * p.apad;
    k_at_fill_int<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(d.jmax, INT_MAX, n);
    k_at_fill_int<<<(p.Nm + 255) / 256, 256, 0, s>>>(d.jmax_min, INT_MAX, p.Nm);
  }
  if (p.lj_alist && !p.small_box) {
    if (p.mol_cap > kAlMolCap || p.mol_cap > 256 || p.apad > 256) return hipErrorInvalidValue;
    const int per = (p.Nm + kXcd - 1) / kXcd;
    if (one)
      k_at_lj_alist<1><<<per * kXcd, 256, 0, s>>>(d, p, hh);
    else
      k_at_lj_alist<2><<<per * kXcd, 256, 0, s>>>(d, p, hh);
  }
  return hipGetLastError();
}

hipError_t at_launch_jmax_from_list(const AtomicDev &d, const AtomicParams &p,
                                    const int *nlc_l, const int *nll_l, int cap,
                                    hipStream_t s) {
  dim3 g((p.Nm + 3) / 4);
  if (ns_of(p) == 1)
    k_at_jmax_from_list<1><<<g, 256, 0, s>>>(d, p, nlc_l, nll_l, cap);
  else
    k_at_jmax_from_list<2><<<g, 256, 0, s>>>(d, p, nlc_l, nll_l, cap);
  return hipGetLastError();
}

hipError_t at_launch_halflist(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                              int *nlc_l, int *nll_l, int cap, hipStream_t s) {
  dim3 g((p.Nm + 3) / 4);
  if (ns_of(p) == 1)
    k_at_scan_half<1, 1><<<g, 256, 0, s>>>(d, p, hh, nlc_l, nll_l, cap);
  else
    k_at_scan_half<2, 1><<<g, 256, 0, s>>>(d, p, hh, nlc_l, nll_l, cap);
  return hipGetLastError();
}

hipError_t at_launch_rebo(const AtomicDev &d, const AtomicParams &p, HHRef hh, int mode,
                          hipStream_t s) {
  if (p.apad > 256) return hipErrorInvalidValue;  // one thread per atom in the prefix sums
  const size_t fast = rebo_fast_lds(p.apad, p.rebo_slots), slow = rebo_slow_lds(p.apad, p.maxr);
  const int fast_ok = fast <= kReboFastLdsMax;
  const size_t shm = fast_ok && fast > slow ? fast : slow;
  const int per = (p.Nm + kXcd - 1) / kXcd;
  dim3 g(per * kXcd);
  // three wavefronts when asked for and the molecule's atoms fit (with four
  // bond slots per atom, five workgroups per CU within the VGPR budget), else
  // four
  if (p.rebo_nt == 192 && p.apad <= 192) {
    if (mode == kReboPhysical && p.lj_trace)
      k_rebo_mol<kReboPhysical, true, 192><<<g, 192, shm, s>>>(d, p, hh, fast_ok);
    else if (mode == kReboPhysical)
      k_rebo_mol<kReboPhysical, false, 192><<<g, 192, shm, s>>>(d, p, hh, fast_ok);
    else if (p.lj_trace)
      k_rebo_mol<kReboCompat, true, 192><<<g, 192, shm, s>>>(d, p, hh, fast_ok);
    else
      k_rebo_mol<kReboCompat, false, 192><<<g, 192, shm, s>>>(d, p, hh, fast_ok);
  } else {
    if (mode == kReboPhysical && p.lj_trace)
      k_rebo_mol<kReboPhysical, true, 256><<<g, 256, shm, s>>>(d, p, hh, fast_ok);
    else if (mode == kReboPhysical)
      k_rebo_mol<kReboPhysical, false, 256><<<g, 256, shm, s>>>(d, p, hh, fast_ok);
    else if (p.lj_trace)
      k_rebo_mol<kReboCompat, true, 256><<<g, 256, shm, s>>>(d, p, hh, fast_ok);
    else
      k_rebo_mol<kReboCompat, false, 256><<<g, 256, shm, s>>>(d, p, hh, fast_ok);
  }
  if (p.tors_eps != 0.0) {
    // the torsion term over the same molecules (256 threads, 1 to 4 per atom)
    if (p.apad <= 64) {
      if (p.lj_trace)
        k_rebo_tors<true, 64><<<g, 256, 0, s>>>(d, p, hh);
      else
        k_rebo_tors<false, 64><<<g, 256, 0, s>>>(d, p, hh);
    } else if (p.apad <= 128) {
      if (p.lj_trace)
        k_rebo_tors<true, 128><<<g, 256, 0, s>>>(d, p, hh);
      else
        k_rebo_tors<false, 128><<<g, 256, 0, s>>>(d, p, hh);
    } else {
      if (p.lj_trace)
        k_rebo_tors<true, 256><<<g, 256, 0, s>>>(d, p, hh);
      else
        k_rebo_tors<false, 256><<<g, 256, 0, s>>>(d, p, hh);
    }
  }
  return hipGetLastError();
}

hipError_t at_launch_forces(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                            hipStream_t s, hipEvent_t ev_mid, unsigned ev_flags) {
  hipError_t e;
  if (p.ff == kFfMM) {
    // bonded molecular-mechanics terms into Frebo and mm.part
    if ((e = mm_launch_bonded(d, p, hh, s)) != hipSuccess) return e;
  } else if (p.use_sg) {
    // the learned surrogate in place of compute_rebo: Frebo and rebo_part
    if ((e = sg_launch_descr(d, p, d.sgw, d.sgd, s)) != hipSuccess) return e;
    if ((e = sg_launch_mlp(p, d.sgw, d.sgd, s)) != hipSuccess) return e;
    if ((e = sg_launch_force(d, p, d.sgw, d.sgd, d.Frebo, s)) != hipSuccess) return e;
    if ((e = sg_launch_parts(d, d.sgd, d.rebo_part, s)) != hipSuccess) return e;
  } else if ((e = at_launch_rebo(d, p, hh, p.rebo_mode, s)) != hipSuccess) {
    return e;
  }
  if (ev_mid && (e = hipEventRecordWithFlags(ev_mid, s, ev_flags)) != hipSuccess) return e;
  const int per = (p.Nm + kXcd - 1) / kXcd;
  dim3 g(per * kXcd);
  const bool one = ns_of(p) == 1;
  if (p.small_box) {
    if (one)
      k_at_lj<1, true, false><<<g, 256, 0, s>>>(d, p, hh);
    else
      k_at_lj<2, true, false><<<g, 256, 0, s>>>(d, p, hh);
  } else if (p.lj_alist) {
    if (!one) k_at_lj_relist<2><<<g, 256, 0, s>>>(d, p, hh);
    if (p.lj_trace) {
      if (one)
        k_at_lj_al<1, true><<<g, 256, 0, s>>>(d, p, hh);
      else
        k_at_lj_al<2, true><<<g, 256, 0, s>>>(d, p, hh);
    } else {
      if (one)
        k_at_lj_al<1, false><<<g, 256, 0, s>>>(d, p, hh);
      else
        k_at_lj_al<2, false><<<g, 256, 0, s>>>(d, p, hh);
    }
  } else if (p.lj_trace) {
    if (one)
      k_at_lj<1, false, true><<<g, 256, 0, s>>>(d, p, hh);
    else
      k_at_lj<2, false, true><<<g, 256, 0, s>>>(d, p, hh);
  } else {
    if (one)
      k_at_lj<1, false, false><<<g, 256, 0, s>>>(d, p, hh);
    else
      k_at_lj<2, false, false><<<g, 256, 0, s>>>(d, p, hh);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (p.ff == kFfMM && d.mm.charge) return mm_launch_coulomb(d, p, hh, s);
  return hipSuccess;
}

hipError_t at_launch_observe(const AtomicDev &d, const AtomicParams &p, int what,
                             hipStream_t s) {
  k_at_kick2<<<d.nblk, 256, 0, s>>>(d, p, 0, 0);
  if (p.ff == kFfMM)
    k_at_observe<true><<<p.R, kFinThreads, 0, s>>>(d, p, what);
  else
    k_at_observe<false><<<p.R, kFinThreads, 0, s>>>(d, p, what);
  return hipGetLastError();
}

hipError_t at_launch_step(const AtomicDev &d, const AtomicParams &p, int ctrl_slot,
                          int thermo_slot, int rebuild, int rescale, const StepEvents *ev,
                          hipStream_t s, int prev_slot, int fin) {
  hipError_t e;
  const unsigned evf = ev ? ev->flags : 0;
  // midpoint buffers alternate with the slot, so that the one of the step
  // before (slot - 1) is intact while this step's is written
  const int par = ctrl_slot & 1;
  AtomicParams ps = p;
  ps.lj_trace = p.trace_virial;  // steps need the virial trace only
  if (rebuild) {
    // AB:881-887: both lists from the current positions and cell
    if (ev && (e = hipEventRecordWithFlags(ev->l0, s, evf)) != hipSuccess) return e;
    if ((e = at_launch_centers(d, p, hh_state(d), s)) != hipSuccess) return e;
    if ((e = at_launch_lists(d, p, hh_state(d), 1, s)) != hipSuccess) return e;
    if (ev && (e = hipEventRecordWithFlags(ev->l1, s, evf)) != hipSuccess) return e;
  }
  if (prev_slot >= 0) {
    if (rebuild || (prev_slot & 1) == par) return hipErrorInvalidValue;
    if (p.ff == kFfMM)
      k_at_kick_drift_fin<true><<<d.nblk, 256, 0, s>>>(d, ps, ctrl_slot, prev_slot, par);
    else
      k_at_kick_drift_fin<false><<<d.nblk, 256, 0, s>>>(d, ps, ctrl_slot, prev_slot, par);
  } else {
    k_at_kick_drift<<<d.nblk, 256, 0, s>>>(d, p, ctrl_slot, par);
  }
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = at_launch_centers(d, p, hh_mid(d, par, p.R), s)) != hipSuccess) return e;
  if (ev && (e = hipEventRecordWithFlags(ev->f0, s, evf)) != hipSuccess) return e;
  if ((e = at_launch_forces(d, ps, hh_mid(d, par, p.R), s, ev ? ev->fm : nullptr, evf)) !=
      hipSuccess)
    return e;
  if (ev && (e = hipEventRecordWithFlags(ev->f1, s, evf)) != hipSuccess) return e;
  k_at_kick2<<<d.nblk, 256, 0, s>>>(d, p, 1, par);
  if (!fin) return hipGetLastError();  // finished by the next step's kick-drift
  if (p.ff == kFfMM)
    k_at_finish<true><<<p.R, kFinThreads, 0, s>>>(d, ps, ctrl_slot, thermo_slot, par);
  else
    k_at_finish<false><<<p.R, kFinThreads, 0, s>>>(d, ps, ctrl_slot, thermo_slot, par);
  if (rescale) k_at_rescale<<<d.nblk, 256, 0, s>>>(d, p);
  return hipGetLastError();
}

}  // namespace fmd
