// fmd_cells.hpp -- cell grids for the molecule lists of both engines.
//
// A list row holds the molecules whose centres lie within a reach of the
// row's centre under the fractional minimum image.  Cells at least that reach
// wide along each cell vector (the cell's perpendicular width over the count
// of cells on that axis) hold every such neighbour in the 27 cells around the
// row's own: a centre difference d changes the fractional coordinate along
// axis a by (row a of h^-1) . d, at most |d| / W_a.  With three or more cells
// per axis those 27 cells are distinct, so each candidate is met once; the
// list test itself is the scan's, and the hits take their places by rank, so
// a row from the cells equals the scan's row entry for entry.
#pragma once

#include <hip/hip_runtime.h>

#include "fmd_common.hpp"

namespace fmd {

constexpr int kCellCap = 64;      // molecules per cell; a full cell sends the grid's
                                  // system back to the scan
constexpr int kCellHitCap = 256;  // hits a row gathers from its cells; a row with
                                  // more is scanned

// Cells per axis for a grid of cells at least `reach` wide along each cell
// vector (h row-major, the cell vectors its columns), at most `cap` cells in
// all (the widest axis gives up cells first); w = 1 when every axis has three
// or more.
__device__ inline int4 cell_grid(const double *h, double reach, int cap) {
  int n[3];
  const double V = fabs(det9(h));
  for (int a = 0; a < 3; a++) {
    const int b = (a + 1) % 3, c = (a + 2) % 3;
    const double x = h[3 * 1 + b] * h[3 * 2 + c] - h[3 * 2 + b] * h[3 * 1 + c],
                 y = h[3 * 2 + b] * h[c] - h[b] * h[3 * 2 + c],
                 z = h[b] * h[3 * 1 + c] - h[3 * 1 + b] * h[c];
    const double w = V / sqrt(x * x + y * y + z * z);  // perpendicular width along a
    n[a] = (int)fmin(w / reach, 1024.0);
  }
  while ((long long)n[0] * n[1] * n[2] > cap) {
    const int a = n[0] >= n[1] ? (n[0] >= n[2] ? 0 : 2) : (n[1] >= n[2] ? 1 : 2);
    n[a]--;
  }
  return make_int4(n[0], n[1], n[2], n[0] >= 3 && n[1] >= 3 && n[2] >= 3);
}

// The cell of centre c (fractional coordinates wrapped into [0, 1)), as the
// flat index (ix * ny + iy) * nz + iz.
__device__ __forceinline__ int3 cell_of(const double *c, const double *hi, int4 cd) {
  double s[3];
  for (int a = 0; a < 3; a++) {
    s[a] = hi[3 * a] * c[0] + hi[3 * a + 1] * c[1] + hi[3 * a + 2] * c[2];
    s[a] -= floor(s[a]);
  }
  return make_int3(min((int)(s[0] * cd.x), cd.x - 1), min((int)(s[1] * cd.y), cd.y - 1),
                   min((int)(s[2] * cd.z), cd.z - 1));
}

__device__ __forceinline__ int cell_flat(int3 k, int4 cd) { return (k.x * cd.y + k.y) * cd.z + k.z; }

// One molecule into its cell (cells of the grid start at cell_n + cb,
// cell_m + cb * kCellCap); a full cell clears *ok.
__device__ __forceinline__ void cell_put(int m, const double *c, const double *hi, int4 cd,
                                         int cb, int *cell_n, int *cell_m, int *ok) {
  const int cl = cb + cell_flat(cell_of(c, hi, cd), cd);
  const int slot = atomicAdd(&cell_n[cl], 1);
  if (slot < kCellCap)
    cell_m[(size_t)cl * kCellCap + slot] = m;
  else
    atomicExch(ok, 0);
}

// By one wavefront (a workgroup of 64 lanes): the molecules q != self in the
// 27 cells around cell k for which in(q) holds, into hit[] (the first
// kCellHitCap of them, in cell order).  The cells' members are walked as one
// flat range, 64 per pass.  Returns the number of hits (may exceed the cap).
// seg: 64 ints of LDS.
template <class In>
__device__ __forceinline__ int cell_hits(int4 cd, int3 k, int cb, const int *cell_n,
                                         const int *cell_m, int self, int lane, In in, int *hit,
                                         int *seg) {
  int cl = 0, n = 0;
  if (lane < 27) {
    const int ix = (k.x + lane / 9 - 1 + cd.x) % cd.x, iy = (k.y + (lane / 3) % 3 - 1 + cd.y) % cd.y,
              iz = (k.z + lane % 3 - 1 + cd.z) % cd.z;
    cl = cb + (ix * cd.y + iy) * cd.z + iz;
    n = min(cell_n[cl], kCellCap);
  }
  int incl = n;  // inclusive prefix over the 27 cells
  for (int s = 1; s < 32; s <<= 1) {
    const int t = __shfl_up(incl, s);
    if (lane >= s) incl += t;
  }
  const int total = __shfl(incl, 26);
  if (lane < 27) {
    seg[lane] = incl - n;  // first flat index of cell `lane`
    seg[32 + lane] = cl;
  }
  __syncthreads();
  int nh = 0;
  for (int t0 = 0; t0 < total; t0 += 64) {
    const int t = t0 + lane;
    int q = -1;
    if (t < total) {
      int lo = 0, hi = 26;  // the last cell starting at or before t (a non-empty one)
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (seg[mid] <= t)
          lo = mid;
        else
          hi = mid - 1;
      }
      q = cell_m[(size_t)seg[32 + lo] * kCellCap + (t - seg[lo])];
    }
    const bool h = q >= 0 && q != self && in(q);
    const unsigned long long bits = __ballot(h);
    const int rank = __popcll(bits & ((1ull << lane) - 1ull));
    if (h && nh + rank < kCellHitCap) hit[nh + rank] = q;
    nh += __popcll(bits);
  }
  __syncthreads();
  return nh;
}

// The nh <= kCellHitCap hits into out[] in ascending order, the first cap.
__device__ __forceinline__ void cell_place(const int *hit, int nh, int lane, int *out, int cap) {
  for (int t = lane; t < nh; t += 64) {
    const int q = hit[t];
    int rank = 0;
    for (int u = 0; u < nh; u++) rank += hit[u] < q;
    if (rank < cap) out[rank] = q;
  }
}

}  // namespace fmd
