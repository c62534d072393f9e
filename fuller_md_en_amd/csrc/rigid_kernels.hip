// rigid_kernels.hip -- gfx950 kernels of the rigid-body LJ engine.
//
//   k_rigid_nlist       symmetric molecule list, one wavefront per molecule,
//                       ballot/popcount compaction keeps ascending order
//                       (nlist_build_sym C1:280-301); from a cell grid
//                       (k_rigid_cell_grid, k_rigid_cell_bin) for large N.
//   k_rigid_sites       lab-frame FP32 site offsets R(q) body (C1:354-364).
//   k_rigid_kick_drift  NPT first half + half-kick + fractional drift + cell
//                       update + quaternion update + sites, wavefront per
//                       molecule (step_npt (A)-(G), C1:512-574).
//   k_rigid_forces      site-site LJ force / torque / virial (forces
//                       C1:346-444): one 256-thread workgroup per molecule,
//                       the 4 wavefronts split its neighbour list, lane = site
//                       of molecule i, the sites of neighbour j are
//                       wavefront-uniform and arrive through scalar loads.
//   k_rigid_kick2       second half-kick + per-block KE / virial partials
//                       (step_npt (I), C1:581-589).
//   k_rigid_finish      fixed-order reduction + NPT second half (J)(K)
//                       (C1:591-598), thermo record, optional rescale factor
//                       (LJ:879-894).
//
// No floating-point atomics anywhere: every sum has a fixed order, so runs are
// bitwise repeatable.
#include <algorithm>
#include <cstddef>

#include "fmd_rigid.hpp"
#include "fmd_wave.hpp"

namespace fmd {

namespace {

constexpr int kXcd = 8;  // MI355X: workgroups are dealt round-robin to 8 XCDs

__device__ __forceinline__ double clampd(double x, double lo, double hi) {
  return x < lo ? lo : (x > hi ? hi : x);
}

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uniform_f(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uniform_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b);
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// The two halves-swaps of gfx950 on a pair of doubles: permlane32 exchanges
// lanes 32-63 of a with lanes 0-31 of b, permlane16 rows 1 and 3 of a with
// rows 0 and 2 of b.  r0 + r1 is then, in each lane, the lane's own value
// of a (lower half / even rows) or b (upper half / odd rows) plus its
// partner's (lane ^ 32 / lane ^ 16): one butterfly stage of two values.
template <bool H32>
__device__ __forceinline__ double swap_add(double a, double b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const int alo = (int)A, ahi = (int)(A >> 32), blo = (int)B, bhi = (int)(B >> 32);
  int l0, l1, h0, h1;
  if (H32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(ahi, bhi, false, false);
    l0 = lo[0], l1 = lo[1], h0 = hi[0], h1 = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(alo, blo, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(ahi, bhi, false, false);
    l0 = lo[0], l1 = lo[1], h0 = hi[0], h1 = hi[1];
  }
  const double r0 = __longlong_as_double(((long long)h0 << 32) | (unsigned int)l0);
  const double r1 = __longlong_as_double(((long long)h1 << 32) | (unsigned int)l1);
  return r0 + r1;
}

// Sums of eight doubles over the 64 lanes, in a fixed order; every lane gets
// all eight.  Transposed butterfly: the exchanges with lane ^ 32, ^ 16, ^ 8
// each halve the values a lane carries (it keeps the half its side of the
// split owns and adds the partner's copy of it), so lane L then holds a
// partial of value (L >> 3) & 7; three DPP pairings inside each group of 8
// complete it, and lane 8k holds sum k.  The first two exchanges are lane
// swaps (no LDS traffic).  Whole wavefront active.
__device__ __forceinline__ void wave_sum8(double *v) {
  const int lane = __lane_id();
  double k4[4], k2[2];
#pragma unroll
  for (int c = 0; c < 4; c++) k4[c] = swap_add<true>(v[c], v[c + 4]);  // c | c + 4
#pragma unroll
  for (int c = 0; c < 2; c++) k2[c] = swap_add<false>(k4[c], k4[c + 2]);
  const bool up = lane & 8;
  double k1 = (up ? k2[1] : k2[0]) + dpp64<0x128>(up ? k2[0] : k2[1]);  // row_ror:8 = lane ^ 8
  k1 += dpp64<0xB1>(k1);   // quad_perm [1,0,3,2]
  k1 += dpp64<0x4E>(k1);   // quad_perm [2,3,0,1]
  k1 += dpp64<0x141>(k1);  // row_half_mirror
#pragma unroll
  for (int c = 0; c < 8; c++) v[c] = readlane64(k1, 8 * c);
}

// Lab-frame offset of site a of molecule i: FP32 for the pair loop and FP64
// (bitwise the reference's lab array) for re-tests.
__device__ __forceinline__ void write_site(int i, int a, const double *q, const double *body,
                                           float4 *sites, double *sites_d, int natom_pad) {
  const double *b = body + 4 * a;
  double o[3];
  if (b[3] != 0.0) {
    o[0] = o[1] = o[2] = 1e6;
  } else {
    lab_exact(q, b, o);
  }
  const size_t t = (size_t)i * natom_pad + a;
  sites[t] = make_float4((float)o[0], (float)o[1], (float)o[2], 0.f);
  sites_d[3 * t] = o[0];
  sites_d[3 * t + 1] = o[1];
  sites_d[3 * t + 2] = o[2];
}

// The sites of molecule i on the lanes of a wavefront (lanes a and a + 64).
template <int NS>
__device__ __forceinline__ void write_sites(int i, const double *q, const double *body,
                                            float4 *sites, double *sites_d, int natom_pad,
                                            int lane) {
#pragma unroll
  for (int s = 0; s < NS; s++) {
    int a = lane + 64 * s;
    if (a < natom_pad) write_site(i, a, q, body, sites, sites_d, natom_pad);
  }
}

// Lab-frame centre of site quad k of molecule i, w = the quad's radius.
__device__ __forceinline__ void write_quad(int i, int k, const double *q, const double *qbody,
                                           float4 *qcen, int nq) {
  double o[3];
  lab_exact(q, qbody + 4 * k, o);
  qcen[(size_t)i * nq + k] =
      make_float4((float)o[0], (float)o[1], (float)o[2], (float)qbody[4 * k + 3]);
}

// The quads of molecule i on lanes k < nq.
__device__ __forceinline__ void write_quads(int i, const double *q, const double *qbody,
                                            float4 *qcen, int nq, int lane) {
  if (lane < nq) write_quad(i, lane, q, qbody, qcen, nq);
}

// ------------------------------------------------------------------ list
// The list's cell grid for the cell hh: cells at least the list radius wide
// along each cell vector, at most max(27, N); the counts cleared.
__global__ __launch_bounds__(256) void k_rigid_cell_grid(const double *__restrict__ hh,
                                                         int4 *__restrict__ cell_dim,
                                                         int *__restrict__ cell_n, RigidParams p) {
  __shared__ int nc;
  if (threadIdx.x == 0) {
    const int4 g = cell_grid(hh, sqrt(p.rlist2) + 1e-3, max(27, p.N));
    *cell_dim = g;
    nc = g.w ? g.x * g.y * g.z : 0;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nc; c += 256) cell_n[c] = 0;
}

__global__ __launch_bounds__(256) void k_rigid_cell_bin(const double *__restrict__ pos,
                                                        const double *__restrict__ hh,
                                                        int4 *cell_dim, int *__restrict__ cell_n,
                                                        int *__restrict__ cell_m, int N) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int4 cd = *cell_dim;
  if (i >= N || !cd.w) return;
  cell_put(i, pos + 3 * i, hh + 9, cd, 0, cell_n, cell_m, &cell_dim->w);
}

// Row i of the symmetric list by one wavefront: the molecules j != i whose
// centre lies within the list radius (the reference's FP64 minimum image and
// r^2), ascending, the first max_neigh kept and all counted.  CELLS: from the
// 27 cells around i's while the grid is usable (the hits take their places by
// rank; a row of more than kCellHitCap hits is scanned), else by a scan of
// every molecule, 64 per pass with ballot compaction.
template <bool CELLS>
__global__ __launch_bounds__(64) void k_rigid_nlist(const double *__restrict__ pos,
                                                    const double *__restrict__ hh,
                                                    int *__restrict__ nl_count,
                                                    int *__restrict__ nl_list,
                                                    int *__restrict__ flags,
                                                    const int4 *__restrict__ cell_dim,
                                                    const int *__restrict__ cell_n,
                                                    const int *__restrict__ cell_m, RigidParams p) {
  const int i = blockIdx.x;
  const int lane = threadIdx.x;
  const double *h = hh, *hi = hh + 9;
  const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
  auto within = [&](int j) {
    double dx = pos[3 * j] - pix, dy = pos[3 * j + 1] - piy, dz = pos[3 * j + 2] - piz;
    mimg_exact(dx, dy, dz, hi, h);
    return norm2_exact(dx, dy, dz) < p.rlist2;
  };
  if (CELLS) {
    const int4 cd = *cell_dim;
    if (cd.w) {
      __shared__ int hit[kCellHitCap], seg[64];
      const int nh = cell_hits(cd, cell_of(pos + 3 * i, hi, cd), 0, cell_n, cell_m, i, lane,
                               within, hit, seg);
      if (nh <= kCellHitCap) {
        cell_place(hit, nh, lane, nl_list + (size_t)i * p.max_neigh, p.max_neigh);
        if (lane == 0) {
          nl_count[i] = nh < p.max_neigh ? nh : p.max_neigh;
          if (nh > p.max_neigh) atomicAdd(&flags[0], 1);
        }
        return;
      }
    }
  }
  int count = 0;
  for (int j0 = 0; j0 < p.N; j0 += 64) {
    const int j = j0 + lane;
    const bool in = j < p.N && j != i && within(j);
    unsigned long long m = __ballot(in);
    int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (in && count + rank < p.max_neigh) nl_list[(size_t)i * p.max_neigh + count + rank] = j;
    count += __popcll(m);
  }
  if (lane == 0) {
    nl_count[i] = count < p.max_neigh ? count : p.max_neigh;
    if (count > p.max_neigh) atomicAdd(&flags[0], 1);
  }
}

// Before the engine replaces its list: it held every pair within the list
// radius at its build, and a pair has closed by at most the two centres'
// displacements plus the cell change along the image (at most the sum of
// the column changes) since then.  A molecule that may have closed the skin
// beyond the force loop's cutoff with any neighbour counts a list fault.
__global__ __launch_bounds__(256) void k_rigid_list_check(const double *__restrict__ pos,
                                                          const double *__restrict__ nl_pos,
                                                          const double *__restrict__ nl_h,
                                                          const double *__restrict__ hh,
                                                          int *__restrict__ flags, RigidParams p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.N || nl_h[9] == 0.0) return;
  const double *h = hh, *hi = hh + 9;
  double dh = 0;
  for (int c = 0; c < 3; c++) {
    const double x = h[c] - nl_h[c], y = h[3 + c] - nl_h[3 + c], z = h[6 + c] - nl_h[6 + c];
    dh += sqrt(x * x + y * y + z * z);
  }
  double dx = pos[3 * i] - nl_pos[3 * i], dy = pos[3 * i + 1] - nl_pos[3 * i + 1],
         dz = pos[3 * i + 2] - nl_pos[3 * i + 2];
  mimg(dx, dy, dz, hi, h);
  const double skin = sqrt(p.rlist2) - sqrt(p.rmcut2);
  if (2.0 * sqrt(dx * dx + dy * dy + dz * dz) + dh > skin) atomicAdd(&flags[1], 1);
}

// The state a new list is built from.
__global__ __launch_bounds__(256) void k_rigid_list_mark(const double *__restrict__ pos,
                                                         double *__restrict__ nl_pos,
                                                         double *__restrict__ nl_h,
                                                         const double *__restrict__ hh, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n)
    for (int k = 0; k < 3; k++) nl_pos[3 * i + k] = pos[3 * i + k];
  if (i < 9) nl_h[i] = hh[i];
  if (i == 9) nl_h[9] = 1.0;
}

// Ownership of a molecule pair: i owns (i, j) iff (j - i) mod N < N / 2 (ties
// to the lower index), so every molecule owns about half of its neighbours
// whatever its index.
__device__ __forceinline__ bool owns(int i, int j, int N) {
  const int dlt = (j - i + N) % N;
  return 2 * dlt < N || (2 * dlt == N && i < j);
}

// From the symmetric list: the owned neighbours of i (hl, in list order) and
// the pairout slots (owner * max_neigh + position in its half list) of the
// pairs owned by i's neighbours, where i's force and torque arrive.  For the
// pair kernel also its work items: the owned pairs by centre distance
// (hl_item).
__global__ __launch_bounds__(64) void k_rigid_halflist(RigidDev d, RigidParams p,
                                                       const double *__restrict__ hh) {
  const int i = blockIdx.x, lane = threadIdx.x;
  const int cnt = d.nl_count[i], cap = p.max_neigh;
  const double *h = hh, *hi = hh + 9;
  int nh = 0, nin = 0;
  for (int k0 = 0; k0 < cnt; k0 += 64) {
    const int k = k0 + lane;
    const int j = k < cnt ? d.nl_list[(size_t)i * cap + k] : -1;
    const bool mine = j >= 0 && owns(i, j, p.N);
    if (p.pair_kernel) {
      // rank of each owned pair of this chunk by centre distance (ties by
      // list position): work item (nh + rank, i) is (slot, j)
      float dist = 3.0e38f;
      if (mine) {
        double dx = d.pos[3 * j] - d.pos[3 * i], dy = d.pos[3 * j + 1] - d.pos[3 * i + 1],
               dz = d.pos[3 * j + 2] - d.pos[3 * i + 2];
        mimg_exact(dx, dy, dz, hi, h);
        dist = (float)(dx * dx + dy * dy + dz * dz);
      }
      const unsigned long long bm = __ballot(mine);
      int rank = 0;
      for (int s = 0; s < 64; s++) {
        const float ds = __shfl(dist, s);
        rank += ds < dist || (ds == dist && s < lane);
      }
      if (mine)
        d.hl_item[(size_t)(nh + rank) * p.N + i] =
            make_int2(i * cap + nh + __popcll(bm & ((1ull << lane) - 1ull)), j);
    }
    bool in = false;
    int slot = 0;
    if (j >= 0 && !mine) {  // position of i among j's owned neighbours
      const int cj = d.nl_count[j];
      int r = 0;
      for (int t = 0; t < cj; t++) {
        const int jj = d.nl_list[(size_t)j * cap + t];
        if (jj == i) {
          in = true;
          break;
        }
        if (owns(j, jj, p.N)) r++;
      }
      slot = j * cap + r;
    }
    unsigned long long bm = __ballot(mine), bi = __ballot(in);
    const unsigned long long below = (1ull << lane) - 1ull;
    if (mine) d.hl_list[(size_t)i * cap + nh + __popcll(bm & below)] = j;
    if (in) d.in_list[(size_t)i * cap + nin + __popcll(bi & below)] = slot;
    nh += __popcll(bm);
    nin += __popcll(bi);
    if (j >= 0 && !mine && !in) atomicAdd(&d.flags[0], 1);  // asymmetric (overflowed) list
  }
  if (p.pair_kernel)  // no pair past i's count
    for (int r = nh + lane; r < cap; r += 64) d.hl_item[(size_t)r * p.N + i] = make_int2(-1, -1);
  if (lane == 0) {
    d.hl_count[i] = nh;
    d.in_count[i] = nin;
  }
}

// Largest owned-pair count over the molecules into ord_ctl[0]: the pair
// kernel's work items are (rank r, molecule i) for r below it.  One workgroup.
__global__ __launch_bounds__(1024) void k_rigid_order_max(const int *__restrict__ hl_count,
                                                          int N, int *__restrict__ ord_ctl) {
  __shared__ int wmax[16];
  int m = 0;
  for (int i = threadIdx.x; i < N; i += 1024) m = max(m, hl_count[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; w++) m = max(m, wmax[w]);
    ord_ctl[0] = m;
  }
}

// ------------------------------------------------------------------ sites
template <int NS>
__global__ __launch_bounds__(256) void k_rigid_sites(RigidDev d, RigidParams p) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= p.N) return;
  double q[4] = {d.qv[4 * i], d.qv[4 * i + 1], d.qv[4 * i + 2], d.qv[4 * i + 3]};
  write_sites<NS>(i, q, d.body, d.sites, d.sites_d, p.natom_pad, lane);
  write_quads(i, q, d.qbody, d.qcen, p.nq, lane);
}

__global__ __launch_bounds__(256) void k_rigid_apply_pbc(double *__restrict__ pos,
                                                         const double *__restrict__ hh, int n) {
#pragma clang fp contract(off)  // the reference's apply_pbc bit for bit (see inv9)
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double *h = hh, *hi = hh + 9;
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

// ------------------------------------------------------- NPT first half
struct Half1 {
  double hi[9], h1[9], hi1[9], Vg1[9];
  double xi1, sc_nh, sc_v;
};

// step_npt (A)(B) and the scaling factors, C1:512-531, with the per-step
// phase controls of LJ:851-857.
__device__ __forceinline__ void npt_half1(const RigidState &S, const StepCtrl &c,
                                          const RigidParams &p, Half1 &o) {
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
  double V = fabs(det9(S.h));
  double KE = S.KEt + S.KEr;
  xi += hdt * (2 * KE - S.Nf * kB * c.Tt) / S.Q;
  xi = clampd(xi, -p.clamp_xi, p.clamp_xi);
  double dP = (2 * S.KEt + S.W9[0] + S.W9[4] + S.W9[8]) / (3 * V) * kEv2GPa - S.Pe;
  for (int a = 0; a < 3; a++) Vg[a * 4] += hdt * V * dP / (S.W * kEv2GPa);
  for (int a = 0; a < 3; a++) Vg[a * 4] = clampd(Vg[a * 4], -p.clamp_vg, p.clamp_vg);
  double eps_tr = Vg[0] * o.hi[0] + Vg[4] * o.hi[4] + Vg[8] * o.hi[8];
  o.sc_nh = exp(-hdt * xi);
  o.sc_v = o.sc_nh * exp(-hdt * eps_tr / 3.0);
  for (int k = 0; k < 9; k++) {
    o.h1[k] = S.h[k] + p.dt * Vg[k];
    o.Vg1[k] = Vg[k];
  }
  inv9(o.h1, o.hi1);
  o.xi1 = xi;
}

// kKdMol molecules per workgroup: threads t < kKdMol integrate molecule t
// (the NPT first-half scalars derived by each of them from the same inputs),
// then the whole workgroup writes the molecules' sites and quads.  Every
// phase is one dependent chain deep per thread.
constexpr int kKdMol = 16;
__global__ __launch_bounds__(256) void k_rigid_kick_drift(RigidDev d, RigidParams p,
                                                          int ctrl_slot) {
  __shared__ double sq[kKdMol][4];
  const int t = threadIdx.x, i0 = blockIdx.x * kKdMol;
  if (t < kKdMol) {
    const RigidState &S = *d.state;
    const StepCtrl c = d.ctrl[ctrl_slot];
    Half1 H;
    npt_half1(S, c, p, H);
    if (blockIdx.x == 0 && t == 0) {
      RigidMid &M = *d.mid;
      for (int k = 0; k < 9; k++) {
        M.h[k] = H.h1[k];
        M.hi[k] = H.hi1[k];
        M.Vg[k] = H.Vg1[k];
      }
      M.xi = H.xi1;
      M.sc_nh = H.sc_nh;
    }
    const int i = i0 + t;
    if (i < p.N) {
      const double hdt = 0.5 * p.dt, cF = kConv / p.Mmol, cT = kConv / p.I0;
      double v[3], w[3], r[3], q[4];
      for (int a = 0; a < 3; a++) {
        v[a] = d.vel[3 * i + a] * H.sc_v + hdt * d.Fv[3 * i + a] * cF;
        w[a] = d.omg[3 * i + a] * H.sc_nh + hdt * d.Tv[3 * i + a] * cT;
        r[a] = d.pos[3 * i + a];
      }
      const double *hi = H.hi;
      double sx = hi[0] * r[0] + hi[1] * r[1] + hi[2] * r[2];
      double sy = hi[3] * r[0] + hi[4] * r[1] + hi[5] * r[2];
      double sz = hi[6] * r[0] + hi[7] * r[1] + hi[8] * r[2];
      double vsx = hi[0] * v[0] + hi[1] * v[1] + hi[2] * v[2];
      double vsy = hi[3] * v[0] + hi[4] * v[1] + hi[5] * v[2];
      double vsz = hi[6] * v[0] + hi[7] * v[1] + hi[8] * v[2];
      sx += p.dt * vsx;
      sy += p.dt * vsy;
      sz += p.dt * vsz;
      sx -= floor(sx);
      sy -= floor(sy);
      sz -= floor(sz);
      const double *h = H.h1;
      d.pos[3 * i] = h[0] * sx + h[1] * sy + h[2] * sz;
      d.pos[3 * i + 1] = h[3] * sx + h[4] * sy + h[5] * sz;
      d.pos[3 * i + 2] = h[6] * sx + h[7] * sy + h[8] * sz;
      for (int a = 0; a < 3; a++) {
        d.vel[3 * i + a] = v[a];
        d.omg[3 * i + a] = w[a];
      }
      for (int a = 0; a < 4; a++) q[a] = d.qv[4 * i + a];
      quat_advance(q, w[0], w[1], w[2], p.dt);
      for (int a = 0; a < 4; a++) {
        d.qv[4 * i + a] = q[a];
        sq[t][a] = q[a];
      }
    }
  }
  __syncthreads();
  const int nm = min(kKdMol, p.N - i0), npad = p.natom_pad, nq = p.nq;
  for (int k = t; k < nm * npad; k += 256) {
    const int m = k / npad;
    write_site(i0 + m, k - m * npad, sq[m], d.body, d.sites, d.sites_d, npad);
  }
  for (int k = t; k < nm * nq; k += 256) {
    const int m = k / nq;
    write_quad(i0 + m, k - m * nq, sq[m], d.qbody, d.qcen, nq);
  }
}

// ------------------------------------------------------------------ forces
// Per lane (site a of molecule i, lab offset r_a) and neighbour j (minimum
// image COM vector dm): G_aj = sum_b f_ab, f_ab = fm(d) d, d = dm + r_b - r_a,
// summed in FP32 over the <= natom sites b, then promoted to FP64.
//   force on i      F_i = -sum_a sum_j G_aj
//   torque          T_i = sum_a r_a x (-sum_j G_aj)               (C1:415-417)
//   virial          W  += 1/2 sum_j dm (x) G_aj - r_a (x) sum_j G_aj
// The virial line is the reference's sum_pairs 1/2 d (x) f (C1:421-423)
// rewritten with the symmetric list: the r_b part of pair (i,a,j,b) equals
// minus the r_a part of its mirror (j,b,i,a), so the per-pair tensor never
// has to be formed.
template <int NS>
__global__ __launch_bounds__(256) void k_rigid_forces(
    const double *__restrict__ pos, const float4 *__restrict__ sites,
    const double *__restrict__ sites_d, const int *__restrict__ nl_count,
    const int *__restrict__ nl_list, const double *__restrict__ hh, double *__restrict__ Fv,
    double *__restrict__ Tv, double *__restrict__ molpart, RigidParams p) {
  // XCD-aware mapping: the workgroups one XCD receives (bid = x mod 8) take a
  // contiguous range of molecules, so neighbours share that XCD's L2.
  const int per = (p.N + kXcd - 1) / kXcd;
  const int bid = blockIdx.x;
  const int i = (bid % kXcd) * per + bid / kXcd;
  if (i >= p.N) return;
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const double *h = hh, *hi = hh + 9;
  const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
  const int npad = p.natom_pad;

  float rax[NS], ray[NS], raz[NS];
  double rad[NS][3];  // FP64 offsets (reference lab values)
#pragma unroll
  for (int s = 0; s < NS; s++) {
    int a = lane + 64 * s;
    if (a < p.natom) {
      const size_t t = (size_t)i * npad + a;
      rad[s][0] = sites_d[3 * t];
      rad[s][1] = sites_d[3 * t + 1];
      rad[s][2] = sites_d[3 * t + 2];
    } else {  // idle lane: far away from everything
      rad[s][0] = rad[s][1] = rad[s][2] = -3.0e5;
    }
    rax[s] = (float)rad[s][0];
    ray[s] = (float)rad[s][1];
    raz[s] = (float)rad[s][2];
  }
  double GA[NS][3], WA[NS][9];
  float EA[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    for (int k = 0; k < 3; k++) GA[s][k] = 0;
    for (int k = 0; k < 9; k++) WA[s][k] = 0;
    EA[s] = 0;
  }
  const LJ32 lj = p.lj;
  const int nn = nl_count[i];
  const int *li = nl_list + (size_t)i * p.max_neigh;
  for (int k = wv; k < nn; k += 4) {
    const int j = uniform(li[k]);
    double dx = pos[3 * j] - pix, dy = pos[3 * j + 1] - piy, dz = pos[3 * j + 2] - piz;
    mimg_exact(dx, dy, dz, hi, h);
    if (dx * dx + dy * dy + dz * dz > p.rmcut2) continue;
    const double dm[3] = {dx, dy, dz};
    float cx[NS], cy[NS], cz[NS];
    bool act = false;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      // dm - r_a formed in FP64 and rounded once: no FP32 rounding of the
      // (up to 18 A) centre vector that all 3600 site pairs would share
      cx[s] = (float)(dx - rad[s][0]);
      cy[s] = (float)(dy - rad[s][1]);
      cz[s] = (float)(dz - rad[s][2]);
      act |= cx[s] * cx[s] + cy[s] * cy[s] + cz[s] * cz[s] < (float)p.cull_r2;
    }
    if (!__any(act)) continue;
    float gx[NS], gy[NS], gz[NS], e[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) gx[s] = gy[s] = gz[s] = e[s] = 0.f;
    const float4 *sj = sites + (size_t)j * npad;
    for (int b = 0; b < npad; b += 4) {
      // the four sites of this group are loaded together (one 64-byte scalar
      // load) before any of the divergent cutoff branches
      float4 rb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) rb[u] = sj[b + u];
#pragma unroll
      for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
          float ex = rb[u].x + cx[s], ey = rb[u].y + cy[s], ez = rb[u].z + cz[s];
          float r2 = ex * ex + ey * ey + ez * ez;
          if (r2 < lj.rc2_hi) {
            bool in = r2 < lj.rc2_lo;
            if (!in)  // within FP32 rounding of the cutoff: the reference's test
              in = pair_r2_exact(dm, sites_d + 3 * ((size_t)j * npad + b + u), rad[s]) <
                   lj.rc2d;
            if (in) {
              float fm = lj_fm(r2, lj, e[s]);
              gx[s] += fm * ex;
              gy[s] += fm * ey;
              gz[s] += fm * ez;
            }
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double c48 = 48.0 * p.eps;
      double g0 = c48 * gx[s], g1 = c48 * gy[s], g2 = c48 * gz[s];
      GA[s][0] += g0;
      GA[s][1] += g1;
      GA[s][2] += g2;
      WA[s][0] += dx * g0;
      WA[s][1] += dx * g1;
      WA[s][2] += dx * g2;
      WA[s][3] += dy * g0;
      WA[s][4] += dy * g1;
      WA[s][5] += dy * g2;
      WA[s][6] += dz * g0;
      WA[s][7] += dz * g1;
      WA[s][8] += dz * g2;
      EA[s] += e[s];
    }
  }
  // Per-lane molecule contributions: F(3) T(3) W(9) E(1).
  double acc[16];
  for (int k = 0; k < 16; k++) acc[k] = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    double r0 = rax[s], r1 = ray[s], r2 = raz[s];
    double f0 = -GA[s][0], f1 = -GA[s][1], f2 = -GA[s][2];
    acc[0] += f0;
    acc[1] += f1;
    acc[2] += f2;
    acc[3] += r1 * f2 - r2 * f1;
    acc[4] += r2 * f0 - r0 * f2;
    acc[5] += r0 * f1 - r1 * f0;
    const double ra[3] = {r0, r1, r2};
    for (int x = 0; x < 3; x++)
      for (int y = 0; y < 3; y++) acc[6 + 3 * x + y] += 0.5 * WA[s][3 * x + y] - ra[x] * GA[s][y];
    acc[15] += (double)EA[s];
  }
  __shared__ double red[4][16];
  for (int k = 0; k < 16; k++) {
    double v = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 16) {
    const int k = threadIdx.x;
    double v = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    if (k < 3)
      Fv[3 * i + k] = v;
    else if (k < 6)
      Tv[3 * i + k - 3] = v;
    else if (k < 15)
      molpart[(size_t)i * kMolPart + (k - 6)] = v;
    else
      molpart[(size_t)i * kMolPart + 9] = 2.0 * p.eps * v;  // 1/2 * 4 eps * sum e
  }
}

// Half-pair variant for time stepping: every molecule pair once, by its
// owner.  Per (lane a, owned neighbour j) the FP32 sum G_aj over j's sites as
// in k_rigid_forces gives i's force and torque; the same G_aj summed over the
// wavefront gives j's: F_j = sum_a G_aj and, since every pair force is along
// its pair vector, tau_j = sum_a r_a x G_aj - dm x F_j (total torque of the
// pair interaction about any point vanishes).  Only neighbour sites in quads
// within reach of molecule i are visited.  The virial is the trace
// sum_pairs d.f (what the pressure needs); energy is per pair, no 1/2.
template <int NS>
__global__ __launch_bounds__(256) void k_rigid_forces_half(
    const double *__restrict__ pos, const float4 *__restrict__ sites,
    const double *__restrict__ sites_d, const float4 *__restrict__ qcen,
    const int *__restrict__ hl_count, const int *__restrict__ hl_list,
    const double *__restrict__ hh, double *__restrict__ Fv, double *__restrict__ Tv,
    double *__restrict__ molpart, double *__restrict__ pairout, RigidParams p) {
  const int per = (p.N + kXcd - 1) / kXcd;
  const int bid = blockIdx.x;
  const int i = (bid % kXcd) * per + bid / kXcd;
  if (i >= p.N) return;
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const double *h = hh, *hi = hh + 9;
  const double pix = pos[3 * i], piy = pos[3 * i + 1], piz = pos[3 * i + 2];
  const int npad = p.natom_pad, nq = p.nq;
  float rax[NS], ray[NS], raz[NS];
  double rad[NS][3];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    int a = lane + 64 * s;
    if (a < p.natom) {
      const size_t t = (size_t)i * npad + a;
      rad[s][0] = sites_d[3 * t];
      rad[s][1] = sites_d[3 * t + 1];
      rad[s][2] = sites_d[3 * t + 2];
    } else {
      rad[s][0] = rad[s][1] = rad[s][2] = -3.0e5;
    }
    rax[s] = (float)rad[s][0];
    ray[s] = (float)rad[s][1];
    raz[s] = (float)rad[s][2];
  }
  double GA[NS][3];
  double EA = 0, TA = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) GA[s][0] = GA[s][1] = GA[s][2] = 0;
  const LJ32 lj = p.lj;
  const double c48 = 48.0 * p.eps;
  const int nn = hl_count[i];
  const int *li = hl_list + (size_t)i * p.max_neigh;
  for (int k = wv; k < nn; k += 4) {
    const int j = uniform(li[k]);
    double dx = pos[3 * j] - pix, dy = pos[3 * j + 1] - piy, dz = pos[3 * j + 2] - piz;
    mimg_exact(dx, dy, dz, hi, h);
    unsigned long long qmask = 0;
    if (dx * dx + dy * dy + dz * dz <= p.rmcut2) {
      bool near = false;
      if (lane < nq) {
        const float4 c = qcen[(size_t)j * nq + lane];
        const float qx = (float)dx + c.x, qy = (float)dy + c.y, qz = (float)dz + c.z;
        const float rr = p.reach + c.w;
        near = qx * qx + qy * qy + qz * qz < rr * rr;
      }
      qmask = __ballot(near);
    }
    double *out = pairout + ((size_t)i * p.max_neigh + k) * kPairOut + 8;
    if (qmask == 0) {
      if (lane < 6) out[lane] = 0.0;
      continue;
    }
    const double dm[3] = {dx, dy, dz};
    float cx[NS], cy[NS], cz[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
      cx[s] = (float)(dx - rad[s][0]);
      cy[s] = (float)(dy - rad[s][1]);
      cz[s] = (float)(dz - rad[s][2]);
    }
    float gx[NS], gy[NS], gz[NS], e = 0.f, tr = 0.f;
#pragma unroll
    for (int s = 0; s < NS; s++) gx[s] = gy[s] = gz[s] = 0.f;
    const float4 *sj = sites + (size_t)j * npad;
    while (qmask) {
      const int q = __builtin_ctzll(qmask);
      qmask &= qmask - 1ull;
      float4 rb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) rb[u] = sj[4 * q + u];
#pragma unroll
      for (int u = 0; u < 4; u++) {
#pragma unroll
        for (int s = 0; s < NS; s++) {
          const float ex = rb[u].x + cx[s], ey = rb[u].y + cy[s], ez = rb[u].z + cz[s];
          const float r2 = ex * ex + ey * ey + ez * ez;
          if (r2 < lj.rc2_hi) {
            bool in = r2 < lj.rc2_lo;
            if (!in)
              in = pair_r2_exact(dm, sites_d + 3 * ((size_t)j * npad + 4 * q + u), rad[s]) <
                   lj.rc2d;
            if (in) {
              const float fm = lj_fm(r2, lj, e);
              tr += fm * r2;
              gx[s] += fm * ex;
              gy[s] += fm * ey;
              gz[s] += fm * ez;
            }
          }
        }
      }
    }
    double fj[3] = {0, 0, 0}, tq[3] = {0, 0, 0};
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const double g0 = c48 * gx[s], g1 = c48 * gy[s], g2 = c48 * gz[s];
      GA[s][0] += g0;
      GA[s][1] += g1;
      GA[s][2] += g2;
      fj[0] += g0;
      fj[1] += g1;
      fj[2] += g2;
      tq[0] += rad[s][1] * g2 - rad[s][2] * g1;
      tq[1] += rad[s][2] * g0 - rad[s][0] * g2;
      tq[2] += rad[s][0] * g1 - rad[s][1] * g0;
    }
    EA += (double)e;
    TA += (double)tr;
    for (int c = 0; c < 3; c++) {
      fj[c] = wave_sum(fj[c]);
      tq[c] = wave_sum(tq[c]);
    }
    if (lane == 0) {
      out[0] = fj[0];
      out[1] = fj[1];
      out[2] = fj[2];
      out[3] = tq[0] - (dy * fj[2] - dz * fj[1]);
      out[4] = tq[1] - (dz * fj[0] - dx * fj[2]);
      out[5] = tq[2] - (dx * fj[1] - dy * fj[0]);
    }
  }
  // i's own share: F(3) T(3) E tr
  double acc[8];
  for (int k = 0; k < 8; k++) acc[k] = 0;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const double f0 = -GA[s][0], f1 = -GA[s][1], f2 = -GA[s][2];
    acc[0] += f0;
    acc[1] += f1;
    acc[2] += f2;
    acc[3] += rad[s][1] * f2 - rad[s][2] * f1;
    acc[4] += rad[s][2] * f0 - rad[s][0] * f2;
    acc[5] += rad[s][0] * f1 - rad[s][1] * f0;
  }
  acc[6] = EA;
  acc[7] = TA;
  __shared__ double red[4][8];
  for (int k = 0; k < 8; k++) {
    double v = wave_sum(acc[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const int k = threadIdx.x;
    const double v = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
    if (k < 3)
      Fv[3 * i + k] = v;
    else if (k < 6)
      Tv[3 * i + k - 3] = v;
    else if (k == 6)
      molpart[(size_t)i * kMolPart + 9] = 4.0 * p.eps * v;
    else {
      // trace only: W = tr/3 on the diagonal (the pressure is all the step uses)
      const double t3 = c48 * v / 3.0;
      double *m = molpart + (size_t)i * kMolPart;
      m[0] = m[4] = m[8] = t3;
      m[1] = m[2] = m[3] = m[5] = m[6] = m[7] = 0.0;
    }
  }
}

// Wavefront-scope ordering of LDS writes before reads by other lanes.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pair-once forces for natom <= 64.  Work item (r, i) is molecule i's r-th
// nearest owned pair (hl_item[r N + i], built at the list rebuild); each
// wavefront takes four items of one rank and neighbouring molecules (similar
// work), the close pairs first, so the costly ones start early and the light
// far ones fill the tail, and the hardware balances the wavefronts.
// Screening, the four items at once: lanes 0-3 take the FP64 minimum image,
// lanes 16k + q test quad q of item k -- some site quad of i within reach of
// j's centre and some quad of j within reach of i's centre; items that fail
// get zero shares, the others are computed one after the other.
// Per pair (i, j): the sites of i within reach of j's centre (set A) and the
// sites of j within reach of i's centre (set B) are compacted through LDS;
// the lanes hold A in P groups of 64/P lanes (P the largest power of two
// with P |A| <= 64, at most 16) and the groups take P sites of B per pass.
// A lane's sums over its B sites (G, the force on j from its i site; energy;
// sum d.f) are summed over the wavefront into the pair's slot of pairout:
// F_i = -sum G, tau_i = -sum r_a x G, F_j = sum G, tau_j = sum r_a x G -
// dm x F_j (every pair force lies along its pair vector).  The cutoff test
// is the same as k_rigid_forces' (FP32 r^2, FP64 re-test inside the band),
// on coordinates in units of sigma.
__global__ __launch_bounds__(256, 8) void k_rigid_pairs(
    const double *__restrict__ pos, const float4 *__restrict__ sites,
    const double *__restrict__ sites_d, const float4 *__restrict__ qcen,
    const int2 *__restrict__ hl_item, const int *__restrict__ ord_ctl,
    const double *__restrict__ hh, double *__restrict__ pairout, RigidParams p) {
  __shared__ float4 sB[4][64 + 16];
  __shared__ int sA[4][64];
  __shared__ double sDm[4][4][3];  // per wavefront: the four items' dm
  __shared__ int2 sIt[4][4];       // ... and (slot, j)
  const int lane = threadIdx.x & 63;
  const int wv = uniform(threadIdx.x >> 6);
  const double *h = hh, *hi = hh + 9;
  const int npad = p.natom_pad, nq = p.nq, cap = p.max_neigh;
  const LJ32 lj = p.lj;
  // pair tests in units of sigma: (r / sigma)^2 gives (sigma / r)^2 by one
  // reciprocal; the FP64 band re-test stays in A
  const float isig = p.isig, rc2_lo = p.rc2_lo_s, rc2_hi = p.rc2_hi_s, rmin2 = p.rmin2_s;
  const double c48s = p.c48s, c48 = 48.0 * p.eps, e4 = 4.0 * p.eps;
  const float reach = p.reach, reach2 = p.reach * p.reach;
  // XCD-local work: workgroups are dealt round-robin to the 8 XCDs, and the
  // ones on XCD x take the molecules of slab x (an eighth of the index range,
  // a slab of the crystal), rank by rank, so an XCD's L2 holds its slab's
  // sites and their neighbours'
  const int R = uniform(ord_ctl[0]);
  const int ns = (p.N + kXcd - 1) / kXcd, xcd = blockIdx.x % kXcd;
  const int lb = blockIdx.x / kXcd, nlb = (gridDim.x - xcd + kXcd - 1) / kXcd;
  const int slab_items = ns * R;
  for (int tb = (lb * 4 + wv) * 4; tb < slab_items; tb += nlb * 16) {
    // lanes 0-3: item tb + lane and its minimum image
    int2 item = make_int2(-1, -1);
    double mx = 0.0, my = 0.0, mz = 0.0;
    bool near = false;
    if (lane < 4 && tb + lane < slab_items) {
      const int u = tb + lane, r = u / ns, i = xcd * ns + (u - r * ns);
      if (i < p.N) item = hl_item[(size_t)r * p.N + i];
      if (item.x >= 0) {
        const int i = item.x / cap, j = item.y;
        mx = pos[3 * j] - pos[3 * i];
        my = pos[3 * j + 1] - pos[3 * i + 1];
        mz = pos[3 * j + 2] - pos[3 * i + 2];
        mimg_exact(mx, my, mz, hi, h);
        near = mx * mx + my * my + mz * mz <= p.rmcut2;
      }
      sDm[wv][lane][0] = mx;
      sDm[wv][lane][1] = my;
      sDm[wv][lane][2] = mz;
      sIt[wv][lane] = item;
    }
    // lane 16k + q: quad q of item k
    const int k = lane >> 4, q = lane & 15;
    const int slot_k = __shfl(item.x, k), j_k = __shfl(item.y, k);
    const bool near_k = __shfl((int)near, k) != 0;
    const float fx = __shfl((float)mx, k), fy = __shfl((float)my, k), fz = __shfl((float)mz, k);
    bool qa = false, qb = false;
    if (near_k && q < nq) {
      const float4 ca = qcen[(size_t)(slot_k / cap) * nq + q];
      const float4 cb = qcen[(size_t)j_k * nq + q];
      const float ax = ca.x - fx, ay = ca.y - fy, az = ca.z - fz, ra = reach + ca.w;
      const float bx = cb.x + fx, by = cb.y + fy, bz = cb.z + fz, rb = reach + cb.w;
      qa = ax * ax + ay * ay + az * az < ra * ra;
      qb = bx * bx + by * by + bz * bz < rb * rb;
    }
    const unsigned long long qA = __ballot(qa), qB = __ballot(qb);
    // items that cannot hold a site pair in range: zero shares (lane 16k + c
    // writes double c of item k)
    const bool live_k = ((qA >> (16 * k)) & 0xFFFFull) && ((qB >> (16 * k)) & 0xFFFFull);
    if (slot_k >= 0 && !live_k) pairout[(size_t)slot_k * kPairOut + q] = 0.0;
    const unsigned long long live = __ballot(live_k && q == 0);  // bit 16k: item k
    wave_lds_sync();
#pragma unroll 1
    for (int kk = 0; kk < 4; kk++) {
      if (!((live >> (16 * kk)) & 1ull)) continue;
      const int slot = uniform(sIt[wv][kk].x), j = uniform(sIt[wv][kk].y);
      const int i = uniform(slot / cap);
      const double dx = uniform_d(sDm[wv][kk][0]), dy = uniform_d(sDm[wv][kk][1]),
                   dz = uniform_d(sDm[wv][kk][2]);
      double *out = pairout + (size_t)slot * kPairOut;
      bool inA = false, inB = false;
      float4 sb = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane < p.natom) {
        const float4 sa = sites[(size_t)i * npad + lane];
        sb = sites[(size_t)j * npad + lane];
        const float gx = (float)dx, gy = (float)dy, gz = (float)dz;
        const float ax = sa.x - gx, ay = sa.y - gy, az = sa.z - gz;
        const float bx = sb.x + gx, by = sb.y + gy, bz = sb.z + gz;
        inA = ax * ax + ay * ay + az * az < reach2;
        inB = bx * bx + by * by + bz * bz < reach2;
      }
      const unsigned long long mA = __ballot(inA), mB = __ballot(inB);
      if (mA == 0 || mB == 0) {
        if (lane < kPairOut) out[lane] = 0.0;
        continue;
      }
      const int nA = __popcll(mA), nB = __popcll(mB);
      if (inA) sA[wv][__popcll(mA & ((1ull << lane) - 1ull))] = lane;
      if (inB)
        sB[wv][__popcll(mB & ((1ull << lane) - 1ull))] =
            make_float4(sb.x * isig, sb.y * isig, sb.z * isig, __int_as_float(lane));
      if (lane < 16) sB[wv][nB + lane] = make_float4(1.0e6f, 1.0e6f, 1.0e6f, 0.f);
      wave_lds_sync();
      int lgG = 6;  // log2 of the lanes per group
      while (lgG > 2 && nA <= (1 << (lgG - 1))) lgG--;
      const int P = 64 >> lgG;
      const int g = lane >> lgG, ta = lane & ((1 << lgG) - 1);
      // lanes past |A| sit far away: every pair test of theirs fails
      const bool act = ta < nA;
      double ra[3] = {0.0, 0.0, 0.0};
      float cx = 1.0e6f, cy = 1.0e6f, cz = 1.0e6f;
      if (act) {
        const double *qd = sites_d + 3 * ((size_t)i * npad + sA[wv][ta]);
        ra[0] = qd[0];
        ra[1] = qd[1];
        ra[2] = qd[2];
        cx = (float)(dx - ra[0]) * isig;
        cy = (float)(dy - ra[1]) * isig;
        cz = (float)(dz - ra[2]) * isig;
      }
      const double dm[3] = {dx, dy, dz};
      // per lane: G / (48 eps / sigma) = sum t (sigma/r)^2 e', t = s6 (s6 - 1/2)
      // with s6 = (sigma/r)^6 (so sum d.f = 48 eps sum t), sum s6 and the
      // in-range pair count n: E = 4 eps (sum t - sum s6 / 2 - n vs4)
      float gx = 0.f, gy = 0.f, gz = 0.f, tr = 0.f, s6 = 0.f;
      int n = 0;
      // no bounds tests: indices past |B| (up to |B| + P - 1) read the padding
#pragma unroll 2
      for (int b0 = 0; b0 < nB; b0 += P) {
        const float4 rb = sB[wv][b0 + g];
        const float ex = rb.x + cx, ey = rb.y + cy, ez = rb.z + cz;
        const float r2 = ex * ex + ey * ey + ez * ez;
        if (r2 < rc2_hi) {
          bool in = r2 < rc2_lo;
          if (!in)
            in = pair_r2_exact(dm, sites_d + 3 * ((size_t)j * npad + __float_as_int(rb.w)),
                               ra) < lj.rc2d;
          if (in) {
            const float sr2 = rcp32(__builtin_amdgcn_fmed3f(r2, rmin2, 3.0e38f));
            const float sr4 = sr2 * sr2;
            const float sr6 = sr2 * sr4;
            const float tt = sr6 * (sr6 - 0.5f);
            const float fm = tt * sr2;
            tr += tt;
            s6 += sr6;
            n++;
            gx += fm * ex;
            gy += fm * ey;
            gz += fm * ez;
          }
        }
      }
      // the next pair's compaction overwrites sA / sB
      wave_lds_sync();
      const double g0 = c48s * gx, g1 = c48s * gy, g2 = c48s * gz;
      double v[8] = {g0, g1, g2, ra[1] * g2 - ra[2] * g1, ra[2] * g0 - ra[0] * g2,
                     ra[0] * g1 - ra[1] * g0,
                     e4 * ((double)tr - 0.5 * (double)s6 - (double)lj.vs4 * n), c48 * tr};
      wave_sum8(v);
      if (lane == 0) {
        out[0] = -v[0];
        out[1] = -v[1];
        out[2] = -v[2];
        out[3] = -v[3];
        out[4] = -v[4];
        out[5] = -v[5];
        out[6] = v[6];
        out[7] = v[7];
        out[8] = v[0];
        out[9] = v[1];
        out[10] = v[2];
        out[11] = v[3] - (dy * v[2] - dz * v[1]);
        out[12] = v[4] - (dz * v[0] - dx * v[2]);
        out[13] = v[5] - (dx * v[1] - dy * v[0]);
      }
    }
  }
}

// ------------------------------------------------------ kick 2 + partials
// Block sums of 12 doubles in a fixed order.
__device__ __forceinline__ void block_sum12(double *v, double *out) {
  __shared__ double red[4][kBlkPart];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int k = 0; k < kBlkPart; k++) {
    double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < kBlkPart) {
    const int k = threadIdx.x;
    out[k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// Second half-kick, one wavefront per molecule.  In pair-once mode the
// molecule's force and torque are its own share plus the shares its
// neighbours computed for it (pairout slots listed in in_list), gathered by
// the lanes and summed in a fixed order.  Block partials of KE, W, E follow.
__global__ __launch_bounds__(256) void k_rigid_kick2(RigidDev d, RigidParams p, int kick) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  double v[kBlkPart];
  for (int k = 0; k < kBlkPart; k++) v[k] = 0;
  if (i < p.N) {
    // the kinematics and the second-half scale factors first: their latency
    // overlaps the gather below
    double vx = 0.0, vy = 0.0, vz = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    if (lane == 0) {
      vx = d.vel[3 * i], vy = d.vel[3 * i + 1], vz = d.vel[3 * i + 2];
      wx = d.omg[3 * i], wy = d.omg[3 * i + 1], wz = d.omg[3 * i + 2];
    }
    const double hdt = 0.5 * p.dt;
    double sc_nh = 1.0, sc_v2 = 1.0;
    if (kick) {
      const RigidMid &M = *d.mid;
      const double eps_tr2 = M.Vg[0] * M.hi[0] + M.Vg[4] * M.hi[4] + M.Vg[8] * M.hi[8];
      sc_nh = M.sc_nh;
      sc_v2 = M.sc_nh * exp(-hdt * eps_tr2 / 3.0);
    }
    double f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (kick && p.pair_kernel) {
      // both sides from the pair slots: owned pairs (F, tau, energy, sum d.f)
      // then the incoming ones (F, tau)
      const int nh = d.hl_count[i], nin = d.in_count[i];
      for (int m = lane; m < nh; m += 64) {
        const double *o = d.pairout + ((size_t)i * p.max_neigh + m) * kPairOut;
        for (int c = 0; c < 8; c++) f[c] += o[c];
      }
      for (int m = lane; m < nin; m += 64) {
        const double *o =
            d.pairout + (size_t)d.in_list[(size_t)i * p.max_neigh + m] * kPairOut + 8;
        for (int c = 0; c < 6; c++) f[c] += o[c];
      }
      wave_sum8(f);
      if (lane == 0) {
        double *mp = d.molpart + (size_t)i * kMolPart;
        mp[0] = mp[4] = mp[8] = f[7] / 3.0;
        mp[1] = mp[2] = mp[3] = mp[5] = mp[6] = mp[7] = 0.0;
        mp[9] = f[6];
      }
    } else if (kick && p.use_half) {
      const int nin = d.in_count[i];
      for (int m = lane; m < nin; m += 64) {
        const double *o =
            d.pairout + (size_t)d.in_list[(size_t)i * p.max_neigh + m] * kPairOut + 8;
        for (int c = 0; c < 6; c++) f[c] += o[c];
      }
      for (int c = 0; c < 6; c++) f[c] = wave_sum(f[c]);
    }
    if (lane == 0) {
      if (kick) {
        double F[3], T[3];
        for (int c = 0; c < 3; c++) {  // the pair kernel leaves both sides in f
          F[c] = (p.pair_kernel ? 0.0 : d.Fv[3 * i + c]) + f[c];
          T[c] = (p.pair_kernel ? 0.0 : d.Tv[3 * i + c]) + f[3 + c];
        }
        if (p.use_half)
          for (int c = 0; c < 3; c++) {
            d.Fv[3 * i + c] = F[c];
            d.Tv[3 * i + c] = T[c];
          }
        const double cF = kConv / p.Mmol, cT = kConv / p.I0;
        vx = (vx + hdt * F[0] * cF) * sc_v2;
        vy = (vy + hdt * F[1] * cF) * sc_v2;
        vz = (vz + hdt * F[2] * cF) * sc_v2;
        wx = (wx + hdt * T[0] * cT) * sc_nh;
        wy = (wy + hdt * T[1] * cT) * sc_nh;
        wz = (wz + hdt * T[2] * cT) * sc_nh;
        d.vel[3 * i] = vx;
        d.vel[3 * i + 1] = vy;
        d.vel[3 * i + 2] = vz;
        d.omg[3 * i] = wx;
        d.omg[3 * i + 1] = wy;
        d.omg[3 * i + 2] = wz;
      }
      v[0] = vx * vx + vy * vy + vz * vz;
      v[1] = wx * wx + wy * wy + wz * wz;
      for (int k = 0; k < kMolPart; k++) v[2 + k] = d.molpart[(size_t)i * kMolPart + k];
    }
  }
  // one molecule per wavefront, its values on lane 0: the block sum is the
  // four wavefronts' lane-0 values in order (component-major blockpart)
  __shared__ double red[4][kBlkPart];
  const int wv = threadIdx.x >> 6;
  if (lane == 0)
    for (int k = 0; k < kBlkPart; k++) red[wv][k] = v[k];
  __syncthreads();
  if (threadIdx.x < kBlkPart) {
    const int k = threadIdx.x;
    d.blockpart[(size_t)k * d.nblk + blockIdx.x] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
  }
}

// Fixed-order reduction of the block partials (one workgroup).
// bp is component-major (component k of block b at k * nblk + b); every
// thread issues the loads of four blocks before adding them.
__device__ __forceinline__ void reduce_blocks(const double *bp, int nblk, double *sums) {
  double v[kBlkPart];
  for (int k = 0; k < kBlkPart; k++) v[k] = 0;
  for (int b = threadIdx.x; b < nblk; b += 4 * 256) {
    double x[4][kBlkPart];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int bu = b + 256 * u;
#pragma unroll
      for (int k = 0; k < kBlkPart; k++) x[u][k] = bu < nblk ? bp[(size_t)k * nblk + bu] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
      for (int k = 0; k < kBlkPart; k++) v[k] += x[u][k];
  }
  __shared__ double tot[kBlkPart];
  block_sum12(v, tot);
  __syncthreads();
  for (int k = 0; k < kBlkPart; k++) sums[k] = tot[k];
}

__global__ __launch_bounds__(256) void k_rigid_observe(RigidDev d, RigidParams p, int what) {
  double sums[kBlkPart];
  reduce_blocks(d.blockpart, d.nblk, sums);
  if (threadIdx.x != 0) return;
  RigidState &S = *d.state;
  if (what & kObsKE) {
    S.KEt = 0.5 * p.Mmol * sums[0] / kConv;
    S.KEr = 0.5 * p.I0 * sums[1] / kConv;
  }
  if (what & kObsForce) {
    for (int k = 0; k < 9; k++) S.W9[k] = sums[2 + k];
    S.Ep = sums[11];
  }
}

// step_npt (J)(K) C1:591-598 plus the monitor quantities of LJ:874-876 and
// the COLD/WARM rescale of LJ:879-894 (the velocities are scaled by
// k_rigid_rescale; KE is rescaled by s^2 here).
__global__ __launch_bounds__(256) void k_rigid_finish(RigidDev d, RigidParams p, int ctrl_slot,
                                                      int thermo_slot) {
  // the state, midpoint and control are loaded before the reduction, so
  // their latency overlaps it
  RigidState &S = *d.state;
  const RigidMid M = *d.mid;
  const StepCtrl c = d.ctrl[ctrl_slot];
  const double Nf = S.Nf, Q = S.Q, Pe = S.Pe, Wb = S.W;
  double sums[kBlkPart];
  reduce_blocks(d.blockpart, d.nblk, sums);
  if (threadIdx.x != 0) return;
  const double hdt = 0.5 * p.dt;
  double KEt = 0.5 * p.Mmol * sums[0] / kConv;
  double KEr = 0.5 * p.I0 * sums[1] / kConv;
  double KE = KEt + KEr;
  double W9[9];
  for (int k = 0; k < 9; k++) W9[k] = sums[2 + k];
  double Ep = sums[11];
  double xi = M.xi, Vg[9];
  for (int k = 0; k < 9; k++) Vg[k] = M.Vg[k];
  xi += hdt * (2 * KE - Nf * kB * c.Tt) / Q;
  xi = clampd(xi, -p.clamp_xi, p.clamp_xi);
  double V2 = fabs(det9(M.h));
  double P = (2 * KEt + W9[0] + W9[4] + W9[8]) / (3 * V2) * kEv2GPa;
  double dP = P - Pe;
  for (int a = 0; a < 3; a++) Vg[a * 4] += hdt * V2 * dP / (Wb * kEv2GPa);
  for (int a = 0; a < 3; a++) Vg[a * 4] = clampd(Vg[a * 4], -p.clamp_vg, p.clamp_vg);
  double T = 2 * KE / (Nf * kB);
  double scale = 1.0;
  if ((c.flags & kCtrlRescale) && T > 0.1) {
    double tgt = c.rescale_T > 0.1 ? c.rescale_T : 0.1;
    scale = sqrt(tgt / T);
    KEt *= scale * scale;
    KEr *= scale * scale;
    KE = KEt + KEr;
    T = 2 * KE / (Nf * kB);
    xi = 0;
    if (c.flags & kCtrlRescaleZeroVg)
      for (int k = 0; k < 9; k++) Vg[k] = 0;
  }
  for (int k = 0; k < 9; k++) {
    S.h[k] = M.h[k];
    S.hi[k] = M.hi[k];
    S.Vg[k] = Vg[k];
    S.W9[k] = W9[k];
  }
  S.xi = xi;
  S.Tt = c.Tt;
  S.KEt = KEt;
  S.KEr = KEr;
  S.Ep = Ep;
  S.scale = scale;
  S.step += 1;
  Thermo &t = d.thermo[thermo_slot];
  t.Ep = Ep;
  t.KE = KE;
  t.KEt = KEt;
  t.KEr = KEr;
  t.T = T;
  t.P = P;
  t.V = V2;
  t.xi = xi;
  t.Ep_rebo = 0;
  t.Ep_lj = Ep;
  t.Eb = t.Ea = t.Ed = t.Ei = t.Ecoul = 0;
  for (int k = 0; k < 9; k++) t.h[k] = M.h[k];
}

__global__ __launch_bounds__(256) void k_rigid_rescale(RigidDev d, RigidParams p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.N) return;
  const double s = d.state->scale;
  for (int a = 0; a < 3; a++) {
    d.vel[3 * i + a] *= s;
    d.omg[3 * i + a] *= s;
  }
}

__global__ __launch_bounds__(256) void k_rigid_reduce_molpart(const double *__restrict__ molpart,
                                                              int n, double *__restrict__ out) {
  double v[kBlkPart];
  for (int k = 0; k < kBlkPart; k++) v[k] = 0;
  for (int i = threadIdx.x; i < n; i += 256)
    for (int k = 0; k < kMolPart; k++) v[k] += molpart[(size_t)i * kMolPart + k];
  __shared__ double tot[kBlkPart];
  block_sum12(v, tot);
  __syncthreads();
  if (threadIdx.x < kMolPart) out[threadIdx.x] = tot[threadIdx.x];
}

inline int ns_of(const RigidParams &p) { return p.natom <= 64 ? 1 : 2; }

}  // namespace

hipError_t rigid_launch_sites(const RigidDev &d, const RigidParams &p, hipStream_t s) {
  dim3 g((p.N + 3) / 4);
  if (ns_of(p) == 1)
    k_rigid_sites<1><<<g, 256, 0, s>>>(d, p);
  else
    k_rigid_sites<2><<<g, 256, 0, s>>>(d, p);
  return hipGetLastError();
}

hipError_t rigid_launch_apply_pbc(const RigidDev &d, const RigidParams &p, const double *hh,
                                  hipStream_t s) {
  k_rigid_apply_pbc<<<(p.N + 255) / 256, 256, 0, s>>>(d.pos, hh, p.N);
  return hipGetLastError();
}

hipError_t rigid_launch_nlist(const RigidDev &d, const RigidParams &p, const double *hh,
                              hipStream_t s) {
  if (d.cell_dim) {  // allocated for N >= cell_min
    k_rigid_cell_grid<<<1, 256, 0, s>>>(hh, d.cell_dim, d.cell_n, p);
    k_rigid_cell_bin<<<(p.N + 255) / 256, 256, 0, s>>>(d.pos, hh, d.cell_dim, d.cell_n,
                                                       d.cell_m, p.N);
    k_rigid_nlist<true><<<p.N, 64, 0, s>>>(d.pos, hh, d.nl_count, d.nl_list, d.flags,
                                           d.cell_dim, d.cell_n, d.cell_m, p);
  } else {
    k_rigid_nlist<false><<<p.N, 64, 0, s>>>(d.pos, hh, d.nl_count, d.nl_list, d.flags,
                                            nullptr, nullptr, nullptr, p);
  }
  return hipGetLastError();
}

hipError_t rigid_launch_rebuild(const RigidDev &d, const RigidParams &p, const double *hh,
                                hipStream_t s) {
  const int g = (p.N + 255) / 256;
  k_rigid_list_check<<<g, 256, 0, s>>>(d.pos, d.nl_pos, d.nl_h, hh, d.flags, p);
  hipError_t e = rigid_launch_nlist(d, p, hh, s);
  if (e != hipSuccess) return e;
  k_rigid_list_mark<<<g, 256, 0, s>>>(d.pos, d.nl_pos, d.nl_h, hh, p.N);
  return hipGetLastError();
}

hipError_t rigid_launch_forces(const RigidDev &d, const RigidParams &p, const double *hh,
                               hipStream_t s) {
  const int per = (p.N + kXcd - 1) / kXcd;
  dim3 g(per * kXcd);
  if (ns_of(p) == 1)
    k_rigid_forces<1><<<g, 256, 0, s>>>(d.pos, d.sites, d.sites_d, d.nl_count, d.nl_list, hh,
                                        d.Fv, d.Tv, d.molpart, p);
  else
    k_rigid_forces<2><<<g, 256, 0, s>>>(d.pos, d.sites, d.sites_d, d.nl_count, d.nl_list, hh,
                                        d.Fv, d.Tv, d.molpart, p);
  return hipGetLastError();
}

hipError_t rigid_launch_halflist(const RigidDev &d, const RigidParams &p, const double *hh,
                                 hipStream_t s) {
  k_rigid_halflist<<<p.N, 64, 0, s>>>(d, p, hh);
  if (p.pair_kernel) k_rigid_order_max<<<1, 1024, 0, s>>>(d.hl_count, p.N, d.ord_ctl);
  return hipGetLastError();
}

static hipError_t rigid_launch_forces_half(const RigidDev &d, const RigidParams &p,
                                           const double *hh, hipStream_t s) {
  if (p.pair_kernel) {
    // 4 work items per wavefront, 16 per workgroup, 8 slabs; the item count
    // is on the device, so the grid covers up to 32 owned pairs per molecule
    // (a C60 crystal owns at most about 30 within the list radius), items
    // past the count are skipped and the kernel's grid-stride loop takes any
    // excess
    const size_t slab = ((size_t)p.N + kXcd - 1) / kXcd * std::min(p.max_neigh, 32);
    dim3 g((unsigned)(kXcd * std::min<size_t>((slab + 15) / 16, 1u << 17)));
    k_rigid_pairs<<<g, 256, 0, s>>>(d.pos, d.sites, d.sites_d, d.qcen, d.hl_item, d.ord_ctl,
                                    hh, d.pairout, p);
    return hipGetLastError();
  }
  const int per = (p.N + kXcd - 1) / kXcd;
  dim3 g(per * kXcd);
  if (ns_of(p) == 1)
    k_rigid_forces_half<1><<<g, 256, 0, s>>>(d.pos, d.sites, d.sites_d, d.qcen, d.hl_count,
                                             d.hl_list, hh, d.Fv, d.Tv, d.molpart, d.pairout, p);
  else
    k_rigid_forces_half<2><<<g, 256, 0, s>>>(d.pos, d.sites, d.sites_d, d.qcen, d.hl_count,
                                             d.hl_list, hh, d.Fv, d.Tv, d.molpart, d.pairout, p);
  return hipGetLastError();
}

hipError_t rigid_launch_observe(const RigidDev &d, const RigidParams &p, int what,
                                hipStream_t s) {
  k_rigid_kick2<<<d.nblk, 256, 0, s>>>(d, p, 0);
  k_rigid_observe<<<1, 256, 0, s>>>(d, p, what);
  return hipGetLastError();
}

hipError_t rigid_launch_step(const RigidDev &d, const RigidParams &p, int ctrl_slot,
                             int thermo_slot, int rebuild, int rescale, const StepEvents *ev,
                             hipStream_t s) {
  hipError_t e;
  const unsigned evf = ev ? ev->flags : 0;
  const double *state_hh = reinterpret_cast<const double *>(
      reinterpret_cast<const char *>(d.state) + offsetof(RigidState, h));
  const double *mid_hh = reinterpret_cast<const double *>(
      reinterpret_cast<const char *>(d.mid) + offsetof(RigidMid, h));
  if (rebuild) {
    if (ev && (e = hipEventRecordWithFlags(ev->l0, s, evf)) != hipSuccess) return e;
    if ((e = rigid_launch_rebuild(d, p, state_hh, s)) != hipSuccess) return e;
    if (p.use_half && (e = rigid_launch_halflist(d, p, state_hh, s)) != hipSuccess) return e;
    if (ev && (e = hipEventRecordWithFlags(ev->l1, s, evf)) != hipSuccess) return e;
  }
  k_rigid_kick_drift<<<(p.N + kKdMol - 1) / kKdMol, 256, 0, s>>>(d, p, ctrl_slot);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (ev && (e = hipEventRecordWithFlags(ev->f0, s, evf)) != hipSuccess) return e;
  if (ev && (e = hipEventRecordWithFlags(ev->fm, s, evf)) != hipSuccess) return e;
  if (p.use_half)
    e = rigid_launch_forces_half(d, p, mid_hh, s);
  else
    e = rigid_launch_forces(d, p, mid_hh, s);
  if (e != hipSuccess) return e;
  if (ev && (e = hipEventRecordWithFlags(ev->f1, s, evf)) != hipSuccess) return e;
  k_rigid_kick2<<<d.nblk, 256, 0, s>>>(d, p, 1);
  k_rigid_finish<<<1, 256, 0, s>>>(d, p, ctrl_slot, thermo_slot);
  if (rescale) k_rigid_rescale<<<(p.N + 255) / 256, 256, 0, s>>>(d, p);
  return hipGetLastError();
}

hipError_t rigid_launch_reduce_molpart(const RigidDev &d, const RigidParams &p, double *out10,
                                       hipStream_t s) {
  k_rigid_reduce_molpart<<<1, 256, 0, s>>>(d.molpart, p.N, out10);
  return hipGetLastError();
}

}  // namespace fmd
