// fmd_atomic_api.hip -- host side of the atomic (AIREBO) engine: creation,
// state transfer, the stream-ordered step loop and the drop-in entry points
// with the signatures of the reference's AIREBO functions.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

#include "fmd_engine.hpp"

using namespace fmd;

namespace {

// Molecule ranges from mol_id: every molecule's atoms must be one contiguous
// run (the reference builds atoms molecule by molecule, AB:781-786).
int mol_ranges(const int *mol_id, int na, std::vector<int> &start, int &apad) {
  start.clear();
  start.push_back(0);
  std::vector<int> seen;
  int maxn = 0;
  for (int i = 1; i <= na; i++) {
    if (i == na || mol_id[i] != mol_id[i - 1]) {
      start.push_back(i);
      maxn = std::max(maxn, start.back() - start[start.size() - 2]);
    }
  }
  // an id reappearing after its run ended means a non-contiguous molecule
  std::vector<int> ids;
  for (size_t m = 0; m + 1 < start.size(); m++) ids.push_back(mol_id[start[m]]);
  std::vector<int> sorted = ids;
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return set_error(FMD_ERR_ARG, "atomic: the atoms of each molecule must be contiguous");
  if (maxn > 128) return set_error(FMD_ERR_ARG, "atomic: molecules of more than 128 atoms");
  apad = (maxn + 3) & ~3;
  return FMD_OK;
}

// Largest atom distance from its molecule's centre (host, FP64).
double max_radius(const double *pos, const std::vector<int> &start, const double *h) {
  double hi[9];
  inv9(h, hi);
  double rmax = 0;
  for (size_t m = 0; m + 1 < start.size(); m++) {
    const int s0 = start[m], n = start[m + 1] - s0;
    std::vector<double> d(3 * n);
    double c[3] = {0, 0, 0};
    for (int a = 0; a < n; a++) {
      double dx = pos[3 * (s0 + a)] - pos[3 * s0], dy = pos[3 * (s0 + a) + 1] - pos[3 * s0 + 1],
             dz = pos[3 * (s0 + a) + 2] - pos[3 * s0 + 2];
      mimg(dx, dy, dz, hi, h);
      d[3 * a] = dx;
      d[3 * a + 1] = dy;
      d[3 * a + 2] = dz;
      c[0] += dx;
      c[1] += dy;
      c[2] += dz;
    }
    for (int k = 0; k < 3; k++) c[k] /= n;
    for (int a = 0; a < n; a++) {
      const double x = d[3 * a] - c[0], y = d[3 * a + 1] - c[1], z = d[3 * a + 2] - c[2];
      rmax = std::max(rmax, std::sqrt(x * x + y * y + z * z));
    }
  }
  return rmax;
}

void free_atomic_dev(AtomicDev &d) {
  void *ptrs[] = {d.pos,       d.vel,       d.F,        d.Frebo,     d.mass,       d.mol_start,
                  d.center,    d.radius,    d.off,      d.qcen,      d.jmax,       d.jmax_min,
                  d.mol_nl_count,
                  d.mol_nl,    d.nlc_r,     d.nll_r,    d.bonds.bd,  d.bonds.ba,   d.bonds.bp,
                  d.bonds.bq,  d.bonds.bflag, d.bonds.brev, d.rebo_part, d.rebo_mpart,
                  d.lj_part,   d.ke_part,   d.rebo_fe,   d.al_e,       d.al_n,
                  d.al_pos,    d.al_h,      d.al_dmax,  d.al_redo,   d.img,
                  d.cell_dim,  d.cell_base, d.cell_n,   d.cell_m,
                  d.hh,        d.state,     d.mid,      d.ctrl,      d.thermo,     d.flags,
                  d.mol_rep,   d.atom_rep,  d.rep_mol,  d.rep_atom,  d.blk_a0,     d.blk_a1,
                  d.blk_rep,   d.rep_blk};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  memset(&d, 0, sizeof d);
}

template <class T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

// Host image of the replica tables (fmd_atomic.hpp).
struct Layout {
  std::vector<int> mol_start, mol_rep, atom_rep, rep_mol, rep_atom;
  std::vector<int> blk_a0, blk_a1, blk_rep, rep_blk;
  std::vector<int> cell_base;  // per replica: max(27, its molecules) cells
};

// Concatenate the replicas' molecule layouts.
int build_layout(const fmd_atomic_config *c, int R, Layout &L, int &apad) {
  L = Layout();
  L.mol_start.push_back(0);
  L.rep_mol.push_back(0);
  L.rep_atom.push_back(0);
  L.rep_blk.push_back(0);
  L.cell_base.push_back(0);
  apad = 4;
  int a0 = 0;
  for (int r = 0; r < R; r++) {
    std::vector<int> st;
    int ap = 0;
    int rc = mol_ranges(c[r].mol_id, c[r].n_atoms, st, ap);
    if (rc) return rc;
    apad = std::max(apad, ap);
    for (size_t m = 1; m < st.size(); m++) {
      L.mol_start.push_back(a0 + st[m]);
      L.mol_rep.push_back(r);
    }
    for (int i = 0; i < c[r].n_atoms; i++) L.atom_rep.push_back(r);
    for (int b = 0; b < c[r].n_atoms; b += 256) {
      L.blk_a0.push_back(a0 + b);
      L.blk_a1.push_back(a0 + std::min(b + 256, c[r].n_atoms));
      L.blk_rep.push_back(r);
    }
    a0 += c[r].n_atoms;
    L.rep_mol.push_back((int)L.mol_rep.size());
    L.rep_atom.push_back(a0);
    L.rep_blk.push_back((int)L.blk_rep.size());
    const int nm = L.rep_mol[r + 1] - L.rep_mol[r];
    L.cell_base.push_back(L.cell_base.back() + std::max(27, nm));
  }
  return FMD_OK;
}

int upload_layout(AtomicDev &d, const Layout &L) {
  auto up = [](int *dst, const std::vector<int> &v) {
    return hipMemcpy(dst, v.data(), v.size() * sizeof(int), hipMemcpyHostToDevice);
  };
  hipError_t e = up(d.mol_start, L.mol_start);
  if (e == hipSuccess) e = up(d.mol_rep, L.mol_rep);
  if (e == hipSuccess) e = up(d.atom_rep, L.atom_rep);
  if (e == hipSuccess) e = up(d.rep_mol, L.rep_mol);
  if (e == hipSuccess) e = up(d.rep_atom, L.rep_atom);
  if (e == hipSuccess) e = up(d.blk_a0, L.blk_a0);
  if (e == hipSuccess) e = up(d.blk_a1, L.blk_a1);
  if (e == hipSuccess) e = up(d.blk_rep, L.blk_rep);
  if (e == hipSuccess) e = up(d.rep_blk, L.rep_blk);
  if (e == hipSuccess) e = up(d.cell_base, L.cell_base);
  if (e != hipSuccess) return hip_fail(e, "upload replica layout");
  return FMD_OK;
}

int atomic_alloc(AtomicDev &d, const AtomicParams &p, int ring, int nblk) {
  memset(&d, 0, sizeof d);
  const size_t Na = p.Na, Nm = p.Nm, ns = Na * p.maxr, nt = Nm * p.apad, R = p.R;
  d.nblk = nblk;
  hipError_t e = hipSuccess;
#define AL(ptr, n) \
  if (e == hipSuccess) e = dalloc(&(ptr), (n))
  AL(d.pos, 3 * Na);
  AL(d.vel, 3 * Na);
  AL(d.F, 3 * Na);
  AL(d.Frebo, 3 * Na);
  AL(d.mass, Na);
  AL(d.mol_start, Nm + 1);
  AL(d.center, 3 * Nm);
  AL(d.radius, Nm);
  AL(d.off, nt);
  AL(d.qcen, nt / 4);
  AL(d.jmax, nt);
  AL(d.jmax_min, Nm);
  AL(d.mol_nl_count, Nm);
  AL(d.mol_nl, Nm * p.mol_cap);
  AL(d.nlc_r, Na);
  AL(d.nll_r, ns);
  AL(d.bonds.bd, 4 * ns);
  AL(d.bonds.ba, 4 * ns);
  AL(d.bonds.bp, 4 * ns);
  AL(d.bonds.bq, 4 * ns);
  AL(d.bonds.bflag, ns);
  AL(d.bonds.brev, ns);
  AL(d.rebo_fe, 3 * ns);
  {
    const size_t na = p.apad <= 64 ? 64 : 128;  // lanes of a molecule tile
    AL(d.al_e, Nm * 4 * (size_t)p.al_cap * na);
    AL(d.al_n, Nm * 4 * na);
    AL(d.al_pos, Na);
    AL(d.al_h, 9 * R);
    AL(d.al_dmax, kAlDriftSlots * R);
    AL(d.al_redo, R);
  }
  AL(d.img, R);
  AL(d.cell_dim, R);
  AL(d.cell_base, R + 1);
  AL(d.cell_n, (size_t)p.cell_total);
  AL(d.cell_m, (size_t)p.cell_total * kCellCap);
  AL(d.rebo_part, (size_t)d.nblk * kAtPart);
  AL(d.rebo_mpart, Nm * kAtPart);
  AL(d.lj_part, Nm * kAtPart);
  AL(d.ke_part, kKePart * (size_t)d.nblk);
  AL(d.hh, 18);
  AL(d.state, R);
  AL(d.mid, 2 * R);
  AL(d.ctrl, ring * R);
  AL(d.thermo, ring * R);
  AL(d.flags, kAtFlags);
  AL(d.mol_rep, Nm);
  AL(d.atom_rep, Na);
  AL(d.rep_mol, R + 1);
  AL(d.rep_atom, R + 1);
  AL(d.blk_a0, nblk);
  AL(d.blk_a1, nblk);
  AL(d.blk_rep, nblk);
  AL(d.rep_blk, R + 1);
#undef AL
  if (e != hipSuccess) {
    free_atomic_dev(d);
    return hip_fail(e, "hipMalloc (atomic engine)");
  }
  (void)hipMemset(d.flags, 0, kAtFlags * sizeof(int));
  (void)hipMemset(d.al_pos, 0, Na * sizeof(float4));
  (void)hipMemset(d.al_dmax, 0, kAlDriftSlots * R * sizeof(int));
  (void)hipMemset(d.al_redo, 0, R * sizeof(int));
  (void)hipMemset(d.F, 0, 3 * Na * sizeof(double));
  (void)hipMemset(d.state, 0, R * sizeof(AtomicState));
  (void)hipMemset(d.nlc_r, 0, Na * sizeof(int));
  (void)hipMemset(d.mol_nl_count, 0, Nm * sizeof(int));
  return FMD_OK;
}

void fill_params(AtomicParams &p, const fmd_atomic_config &c, int Na, int Nm, int R,
                 int apad) {
  memset(&p, 0, sizeof p);
  p.Na = Na;
  p.Nm = Nm;
  p.R = R;
  p.apad = apad;
  p.maxr = c.rebo_max_neigh > 0 ? c.rebo_max_neigh : 12;
  p.mol_cap = 160;
  p.lj_cap = c.lj_max_neigh;
  p.lj_compat = c.lj_max_neigh > 0;
  p.rebo_mode = c.rebo_mode;
  p.rebo_rlist2 = c.rebo_rlist * c.rebo_rlist;
  const double rl = c.lj.rcut + c.lj_skin;
  p.lj_rlist2 = rl * rl;  // (LJ_RCUT + 2.0)^2, AB:265
  p.mol_rlist = rl + 1.0;  // + 1 A for centre drift between rebuilds
  p.al_r2 = (float)((c.lj.rcut + c.al_skin) * (c.lj.rcut + c.al_skin));  // per-atom lists
  p.al_cap = 384;  // per wavefront segment: about 3.4 times a C60 crystal's share
  p.rebo_slots = c.rebo_lds_slots > 0 ? c.rebo_lds_slots : 4;
  p.rebo_nt = c.rebo_threads == 256 ? 256 : 192;
  p.pcc.on = 0;
  for (int k = 0; k < 5; k++) {
    p.pcc.v[k] = (float)c.rebo_pcc[k];
    p.pcc.on |= c.rebo_pcc[k] != 0.0;
  }
  p.tors_eps = c.torsion_eps;
  p.dt = c.dt;
  p.clamp_xi = c.clamp_xi;
  p.clamp_vg = c.clamp_vg;
  p.lj = make_lj32(c.lj);
  p.eps = c.lj.eps;
}

// The molecule-tile LJ formulation (one centre image per molecule pair)
// equals the per-atom minimum image whenever rc + 2 R_max fits in half the
// narrowest cell width; otherwise every pair takes its own image.
int decide_small_box(const AtomicParams &p, const double *h, double rmax) {
  const double rc = std::sqrt((double)p.lj.rc2d);
  return (rc + 2.0 * rmax + 1.0) >= 0.5 * min_width(h) ? 1 : 0;
}

// The list walk may go through the molecule tiles when every atom pair within
// the list radius is the centres' minimum image plus the two offsets.
int decide_scan_fast(const AtomicParams &p, const double *h, double rmax) {
  return (std::sqrt(p.lj_rlist2) + 2.0 * rmax + 1.0) < 0.5 * min_width(h) ? 1 : 0;
}

}  // namespace

namespace fmd {

int read_flags(fmd_engine *e);

int engine_prepare_atomic(fmd_engine *e) {
  AtomicEngine &A = *e->at;
  const AtomicDev &d = A.d;
  const AtomicParams &p = A.p;
  hipStream_t s = e->stream;
  const HHRef hh = hh_state(d);
  HIPCHK(at_launch_apply_pbc(d, p, hh, s));  // AB:852
  HIPCHK(at_launch_centers(d, p, hh, s));
  HIPCHK(at_launch_lists(d, p, hh, 1, s));
  HIPCHK(at_launch_forces(d, p, hh, s));
  HIPCHK(at_launch_observe(d, p, kAtObsKE | kAtObsForce, s));
  e->rebuilds++;
  e->prepared = true;
  return read_flags(e);
}

// Device flags since the last read into the list-fault counter.  After an
// image fault (the cell became too narrow for per-molecule images while the
// engine stepped with them), the image mode is decided again from the
// current state, so that the following steps take every pair's own image.
// After a per-atom list overflow the engine evaluates LJ by molecule tiles.
int read_flags(fmd_engine *e) {
  AtomicEngine &A = *e->at;
  int f[kAtFlags] = {};
  HIPCHK(hipMemcpyAsync(f, A.d.flags, sizeof f, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  int any = 0;
  for (int k = 0; k < kAtFlags; k++) {
    if (k == 2)
      e->rebuilds += f[k];  // per-atom lists rebuilt within a step
    else
      e->overflow += f[k];
    any |= f[k];
  }
  if (any) HIPCHK(hipMemsetAsync(A.d.flags, 0, sizeof f, e->stream));
  if (f[4] && A.p.lj_alist) {
    A.p.lj_alist = 0;
    engine_drop_graphs(e);
  }
  if (f[3]) {
    for (int r = 0; r < A.p.R; r++) {
      const size_t a0 = A.rep_atom[r], na = A.rep_atom[r + 1] - a0;
      std::vector<double> pos(3 * na);
      AtomicState S;
      HIPCHK(hipMemcpyAsync(pos.data(), A.d.pos + 3 * a0, 3 * na * 8, hipMemcpyDeviceToHost,
                            e->stream));
      HIPCHK(hipMemcpyAsync(&S, A.d.state + r, sizeof S, hipMemcpyDeviceToHost, e->stream));
      HIPCHK(hipStreamSynchronize(e->stream));
      std::vector<int> start;
      for (int m = A.rep_mol[r]; m <= A.rep_mol[r + 1]; m++)
        start.push_back(A.mol_start[m] - (int)a0);
      const double rmax = max_radius(pos.data(), start, S.h);
      A.small[r] = decide_small_box(A.p, S.h, rmax);
      A.fast[r] = decide_scan_fast(A.p, S.h, rmax);
    }
    A.p.small_box = 0;
    A.p.scan_fast = 1;
    for (int x : A.small) A.p.small_box |= x;
    for (int x : A.fast) A.p.scan_fast &= x;
    engine_drop_graphs(e);  // captured launches hold the old parameters
  }
  HIPCHK(hipStreamSynchronize(e->stream));
  return FMD_OK;
}

int engine_run_atomic(fmd_engine *e, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
                      fmd_thermo *thermo) {
  AtomicEngine &A = *e->at;
  const AtomicDev &d = A.d;
  const AtomicParams &p = A.p;
  if (!e->prepared) {
    int rc = engine_prepare_atomic(e);
    if (rc) return rc;
  }
  const bool fold_ok = d.nblk <= kFoldMaxBlocks;
  auto launch = [&](int slot, int rebuild, int rescale, int fold, int fin,
                    const StepEvents *ev) -> hipError_t {
    return at_launch_step(d, p, slot, slot, rebuild, rescale, ev, e->stream,
                          fold_ok && fold ? slot - 1 : -1, fold_ok ? fin : 1);
  };
  // with the learned surrogate on, stop at every judge step (global step a
  // multiple of judge_every) to compare it with full REBO
  int done = 0, every = 0, rc = FMD_OK;
  while (done < nsteps) {
    int n = nsteps - done;
    const bool judge = surrogate_judging(e, &every);
    if (judge) n = std::min(n, every - (int)(((long long)first_step + done) % every));
    rc = run_steps(e, p.R, first_step + done, n, ctrl ? ctrl + (size_t)done * p.R : nullptr,
                   thermo ? thermo + (size_t)done * p.R : nullptr, A.tt.data(), d.ctrl,
                   d.thermo, launch);
    if (rc) return rc;
    done += n;
    if (judge && ((long long)first_step + done) % every == 0 && (rc = surrogate_judge(e)))
      return rc;
  }
  return read_flags(e);
}

int engine_forces_atomic(fmd_engine *e, double *F, double *T, double *W9, double *energies) {
  (void)T;
  return replica_forces(e, 0, F, W9, energies);
}

int replica_forces(fmd_engine *e, int r, double *F, double *W9, double *energies) {
  if (!e->at) return set_error(FMD_ERR_STATE, "engine has no atomic system");
  if (r < 0 || r >= e->at->p.R) return set_error(FMD_ERR_ARG, "replica %d out of range", r);
  if (!e->prepared) {
    int rc = engine_prepare_atomic(e);
    if (rc) return rc;
  }
  const AtomicEngine &A = *e->at;
  const AtomicDev &d = A.d;
  hipStream_t s = e->stream;
  if (A.p.trace_virial && e->steps > 0) {
    // steps kept the LJ virial trace only: the full tensor at the current state
    HIPCHK(at_launch_centers(d, A.p, hh_state(d), s));
    HIPCHK(at_launch_forces(d, A.p, hh_state(d), s));
    HIPCHK(at_launch_observe(d, A.p, kAtObsForce, s));
  }
  const size_t a0 = A.rep_atom[r], na = A.rep_atom[r + 1] - a0;
  if (F) HIPCHK(hipMemcpyAsync(F, d.F + 3 * a0, 3 * na * 8, hipMemcpyDeviceToHost, s));
  AtomicState S;
  HIPCHK(hipMemcpyAsync(&S, d.state + r, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (W9)
    for (int k = 0; k < 9; k++) W9[k] = S.W9[k];
  if (energies) {
    energies[0] = S.Ep;
    energies[1] = S.Ep_rebo;
    energies[2] = S.Ep_lj;
  }
  return FMD_OK;
}

void engine_free_atomic(fmd_engine *e) {
  if (!e->at) return;
  engine_free_surrogate(*e->at);
  engine_free_mm(*e->at);
  free_atomic_dev(e->at->d);
  delete e->at;
  e->at = nullptr;
}

// Per-pair or per-molecule LJ images for positions pos (host) in cell h.
int atomic_image_mode(fmd_engine *e, const double *pos, const double *h) {
  AtomicEngine &A = *e->at;
  const double rmax = max_radius(pos, A.mol_start, h);
  A.p.small_box = decide_small_box(A.p, h, rmax);
  A.p.scan_fast = decide_scan_fast(A.p, h, rmax);
  return FMD_OK;
}

// The reference-layout LJ half list (stride 400) of positions pos in the
// single-system workspace e.
int atomic_halflist(fmd_engine *e, const double *pos, const double *h, const double *hi,
                    int *nlc, int *nll) {
  constexpr int kCap = 400;
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  const size_t Na = A.p.Na;
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  HIPCHK(hipMemcpyAsync(A.d.pos, pos, 3 * Na * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.hh, hh, sizeof hh, hipMemcpyHostToDevice, s));
  int rc = atomic_image_mode(e, pos, h);
  if (rc) return rc;
  int *dl = nullptr, *dc = nullptr;
  HIPCHK(hipMallocAsync((void **)&dl, Na * kCap * sizeof(int), s));
  HIPCHK(hipMallocAsync((void **)&dc, Na * sizeof(int), s));
  HIPCHK(hipMemsetAsync(dl, 0, Na * kCap * sizeof(int), s));
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_lists(A.d, A.p, hh_fixed(A.d.hh), 0, s));
  HIPCHK(at_launch_halflist(A.d, A.p, hh_fixed(A.d.hh), dc, dl, kCap, s));
  HIPCHK(hipMemcpyAsync(nlc, dc, Na * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(nll, dl, Na * kCap * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipFreeAsync(dl, s));
  HIPCHK(hipFreeAsync(dc, s));
  HIPCHK(hipStreamSynchronize(s));
  return FMD_OK;
}

// One NPT step of the single-system workspace e whose positions, cell and
// lists are staged: vel, F, vir9, h and *npt in and out; *Sout the state
// after the step (h^-1 in Sout->hi).
int atomic_step_once(fmd_engine *e, double *vel, double *F, double *vir9, double *h,
                     fmd_npt_state *npt, AtomicState *Sout) {
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)A.p.Na * 8;
  HIPCHK(hipMemcpyAsync(A.d.vel, vel, n3, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.F, F, n3, hipMemcpyHostToDevice, s));
  AtomicState S;
  memset(&S, 0, sizeof S);
  for (int k = 0; k < 9; k++) {
    S.h[k] = h[k];
    S.Vg[k] = npt->Vg[k];
    S.W9[k] = vir9[k];
  }
  inv9(S.h, S.hi);
  S.xi = npt->xi;
  S.Q = npt->Q;
  S.W = npt->W;
  S.Pe = npt->Pe;
  S.Tt = npt->Tt;
  S.Nf = npt->Nf;
  S.scale = 1.0;
  StepCtrl c = {npt->Tt, 0.0, 0, 0};
  HIPCHK(hipMemcpyAsync(A.d.state, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.ctrl, &c, sizeof c, hipMemcpyHostToDevice, s));
  HIPCHK(at_launch_observe(A.d, A.p, kAtObsKE, s));
  HIPCHK(at_launch_step(A.d, A.p, 0, 0, 0, 0, nullptr, s));
  HIPCHK(hipMemcpyAsync(vel, A.d.vel, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(F, A.d.F, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&S, A.d.state, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) {
    h[k] = S.h[k];
    vir9[k] = S.W9[k];
    npt->Vg[k] = S.Vg[k];
  }
  npt->xi = S.xi;
  *Sout = S;
  return FMD_OK;
}

}  // namespace fmd

extern "C" {

void fmd_atomic_config_default(fmd_atomic_config *c) {
  memset(c, 0, sizeof *c);
  fmd_lj_params_airebo(&c->lj);
  c->lj_skin = 2.0;
  c->lj_max_neigh = 400;
  c->rebo_rlist = 2.3;
  c->rebo_max_neigh = 12;
  c->rebo_mode = FMD_REBO_COMPAT;
  c->nlup = 20;
  c->dt = 0.5;
  c->clamp_xi = 0.05;
  c->clamp_vg = 0.005;
  c->al_skin = 1.0;
  c->cell_min_molecules = 2048;
  c->rebo_lds_slots = 4;
  c->rebo_threads = 192;
}

int fmd_atomic_create(const fmd_atomic_config *c, fmd_engine **out) {
  return fmd_batch_create(c, 1, out);
}

int fmd_batch_create(const fmd_atomic_config *c, int R, fmd_engine **out) {
  return atomic_create_ff(c, R, kFfAirebo, out);
}

}  // extern "C"

// R replicas with the intramolecular force field ff (kFfAirebo / kFfMM).
int fmd::atomic_create_ff(const fmd_atomic_config *c, int R, int ff, fmd_engine **out) {
  if (!c || !out || R < 1) return set_error(FMD_ERR_ARG, "fmd_batch_create: bad argument");
  if (!(c[0].al_skin > 0.0 && c[0].al_skin <= c[0].lj_skin + 1.0))
    return set_error(FMD_ERR_ARG, "fmd_batch_create: al_skin must lie in (0, lj_skin + 1]");
  if (c[0].rebo_lds_slots != 0 && (c[0].rebo_lds_slots < 3 || c[0].rebo_lds_slots > 8))
    return set_error(FMD_ERR_ARG, "fmd_batch_create: rebo_lds_slots must lie in [3, 8]");
  if (c[0].rebo_threads != 0 && c[0].rebo_threads != 192 && c[0].rebo_threads != 256)
    return set_error(FMD_ERR_ARG, "fmd_batch_create: rebo_threads must be 192 or 256");
  bool ext = c[0].torsion_eps != 0.0;
  for (int k = 0; k < 5; k++) ext |= c[0].rebo_pcc[k] != 0.0;
  if (ext && (c[0].rebo_mode != FMD_REBO_PHYSICAL || ff != kFfAirebo))
    return set_error(FMD_ERR_ARG,
                     "fmd_batch_create: rebo_pcc and torsion_eps need FMD_REBO_PHYSICAL "
                     "(the reference's REBO has neither term)");
  for (int r = 0; r < R; r++) {
    if (!c[r].mol_id || !c[r].mass || c[r].n_atoms < 1)
      return set_error(FMD_ERR_ARG, "fmd_batch_create: replica %d: bad configuration", r);
    const fmd_atomic_config &a = c[0], &b = c[r];
    if (a.lj.sigma != b.lj.sigma || a.lj.eps != b.lj.eps || a.lj.rcut != b.lj.rcut ||
        a.lj_skin != b.lj_skin || a.lj_max_neigh != b.lj_max_neigh ||
        a.rebo_rlist != b.rebo_rlist || a.rebo_max_neigh != b.rebo_max_neigh ||
        a.rebo_mode != b.rebo_mode || a.nlup != b.nlup || a.dt != b.dt ||
        a.clamp_xi != b.clamp_xi || a.clamp_vg != b.clamp_vg || a.al_skin != b.al_skin ||
        a.rebo_lds_slots != b.rebo_lds_slots || a.rebo_threads != b.rebo_threads ||
        a.torsion_eps != b.torsion_eps || memcmp(a.rebo_pcc, b.rebo_pcc, sizeof a.rebo_pcc))
      return set_error(FMD_ERR_ARG,
                       "fmd_batch_create: replica %d: potential, lists, dt, nlup and clamps "
                       "must equal replica 0's",
                       r);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(FMD_ERR_NODEVICE, "fmd_batch_create: no HIP device");
  Layout L;
  int apad = 0;
  int rc = build_layout(c, R, L, apad);
  if (rc) return rc;
  fmd_engine *e = new fmd_engine();
  e->kind = FMD_KIND_ATOMIC;
  e->at = new AtomicEngine();
  AtomicEngine &A = *e->at;
  e->nlup = c->nlup > 0 ? c->nlup : 20;
  e->use_graph = R > 1;
  fill_params(A.p, *c, L.rep_atom[R], (int)L.mol_rep.size(), R, apad);
  A.p.trace_virial = 1;
  A.p.lj_alist = 1;
  A.p.ff = ff;
  A.p.coul_rc = c->lj.rcut;
  if (ff == kFfMM) A.p.maxr = 1;  // no REBO list
  A.rep_atom = L.rep_atom;
  A.rep_mol = L.rep_mol;
  A.mol_start = L.mol_start;
  A.tt.assign(R, 0.0);
  A.small.assign(R, 0);
  A.fast.assign(R, 0);
  const int ring = std::max(64, 2048 / R);
  rc = engine_init_common(e, ring, R);
  A.p.cell_total = L.cell_base.back();
  A.p.cell_min = std::max(1, c[0].cell_min_molecules);
  for (int r = 0; r < R; r++) A.p.mol_cells |= L.rep_mol[r + 1] - L.rep_mol[r] >= A.p.cell_min;
  if (rc == FMD_OK) rc = atomic_alloc(A.d, A.p, e->ring, (int)L.blk_rep.size());
  if (rc == FMD_OK) rc = upload_layout(A.d, L);
  if (rc == FMD_OK) {
    hipError_t he = hipSuccess;
    for (int r = 0; r < R && he == hipSuccess; r++)
      he = hipMemcpy(A.d.mass + L.rep_atom[r], c[r].mass, c[r].n_atoms * sizeof(double),
                     hipMemcpyHostToDevice);
    if (he != hipSuccess) rc = hip_fail(he, "upload masses");
  }
  if (rc) {
    fmd_engine_destroy(e);
    return rc;
  }
  *out = e;
  return FMD_OK;
}

extern "C" {

int fmd_batch_size(fmd_engine *e, int *R) {
  if (!e || !R || e->kind != FMD_KIND_ATOMIC) return set_error(FMD_ERR_ARG, "not an atomic engine");
  *R = e->at->p.R;
  return FMD_OK;
}

int fmd_batch_set_state(fmd_engine *e, int r, const double *pos, const double *vel,
                        const double *h, const fmd_npt_state *npt) {
  if (!e || e->kind != FMD_KIND_ATOMIC) return set_error(FMD_ERR_ARG, "not an atomic engine");
  AtomicEngine &A = *e->at;
  if (r < 0 || r >= A.p.R) return set_error(FMD_ERR_ARG, "replica %d out of range", r);
  hipStream_t s = e->stream;
  const size_t a0 = A.rep_atom[r], na = A.rep_atom[r + 1] - a0;
  if (pos) {
    HIPCHK(hipMemcpyAsync(A.d.pos + 3 * a0, pos, 3 * na * 8, hipMemcpyHostToDevice, s));
    // new positions: the next list build has no earlier one to check against
    HIPCHK(hipMemsetAsync(A.d.al_pos + a0, 0, na * sizeof(float4), s));
  }
  if (vel) HIPCHK(hipMemcpyAsync(A.d.vel + 3 * a0, vel, 3 * na * 8, hipMemcpyHostToDevice, s));
  AtomicState S;
  HIPCHK(hipMemcpyAsync(&S, A.d.state + r, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (h) {
    for (int k = 0; k < 9; k++) S.h[k] = h[k];
    inv9(S.h, S.hi);
  }
  if (npt) {
    S.xi = npt->xi;
    S.Q = npt->Q;
    for (int k = 0; k < 9; k++) S.Vg[k] = npt->Vg[k];
    S.W = npt->W;
    S.Pe = npt->Pe;
    S.Tt = npt->Tt;
    S.Nf = npt->Nf;
    A.tt[r] = npt->Tt;
    if (A.p.R == 1) e->tt = npt->Tt;
  }
  S.scale = 1.0;
  if (pos) {
    // this replica's molecule ranges, relative to its first atom
    std::vector<int> start;
    for (int m = A.rep_mol[r]; m <= A.rep_mol[r + 1]; m++) start.push_back(A.mol_start[m] - (int)a0);
    const double rmax = max_radius(pos, start, S.h);
    A.small[r] = decide_small_box(A.p, S.h, rmax);
    A.fast[r] = decide_scan_fast(A.p, S.h, rmax);
    A.p.small_box = 0;
    A.p.scan_fast = 1;
    for (int x : A.small) A.p.small_box |= x;
    for (int x : A.fast) A.p.scan_fast &= x;
  }
  HIPCHK(hipMemcpyAsync(A.d.state + r, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  e->prepared = false;
  engine_drop_graphs(e);  // captured launches hold the old parameters
  return FMD_OK;
}

int fmd_batch_get_state(fmd_engine *e, int r, double *pos, double *vel, double *h,
                        fmd_npt_state *npt) {
  if (!e || e->kind != FMD_KIND_ATOMIC) return set_error(FMD_ERR_ARG, "not an atomic engine");
  AtomicEngine &A = *e->at;
  if (r < 0 || r >= A.p.R) return set_error(FMD_ERR_ARG, "replica %d out of range", r);
  hipStream_t s = e->stream;
  const size_t a0 = A.rep_atom[r], na = A.rep_atom[r + 1] - a0;
  if (pos) HIPCHK(hipMemcpyAsync(pos, A.d.pos + 3 * a0, 3 * na * 8, hipMemcpyDeviceToHost, s));
  if (vel) HIPCHK(hipMemcpyAsync(vel, A.d.vel + 3 * a0, 3 * na * 8, hipMemcpyDeviceToHost, s));
  AtomicState S;
  HIPCHK(hipMemcpyAsync(&S, A.d.state + r, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (h)
    for (int k = 0; k < 9; k++) h[k] = S.h[k];
  if (npt) {
    npt->xi = S.xi;
    npt->Q = S.Q;
    for (int k = 0; k < 9; k++) npt->Vg[k] = S.Vg[k];
    npt->W = S.W;
    npt->Pe = S.Pe;
    npt->Tt = S.Tt;
    npt->Nf = S.Nf;
  }
  return FMD_OK;
}

int fmd_batch_forces(fmd_engine *e, int r, double *F, double *W9, double *energies) {
  if (!e || e->kind != FMD_KIND_ATOMIC) return set_error(FMD_ERR_ARG, "not an atomic engine");
  return replica_forces(e, r, F, W9, energies);
}

int fmd_atomic_set_state(fmd_engine *e, const double *pos, const double *vel, const double *h,
                         const fmd_npt_state *npt) {
  return fmd_batch_set_state(e, 0, pos, vel, h, npt);
}

int fmd_atomic_get_state(fmd_engine *e, double *pos, double *vel, double *h,
                         fmd_npt_state *npt) {
  return fmd_batch_get_state(e, 0, pos, vel, h, npt);
}

}  // extern "C"

// ------------------------------------------------------- AIREBO drop-ins
namespace {

constexpr int kRefMaxRebo = 12, kRefMaxLj = 400;  // AB:126-127

struct AtomicWorkspace {
  fmd_engine *e = nullptr;
  std::vector<int> mol_id;
};
AtomicWorkspace g_aws;

// Workspace engine for this molecule layout (re-created when it changes).
int atomic_workspace(int na, const int *mol_id, int rebo_mode, fmd_engine **out) {
  if (!mol_id || na < 1) return set_error(FMD_ERR_ARG, "AIREBO drop-in: need mol_id and Na >= 1");
  bool same = g_aws.e && (int)g_aws.mol_id.size() == na &&
              std::equal(g_aws.mol_id.begin(), g_aws.mol_id.end(), mol_id);
  if (!same) {
    if (g_aws.e) fmd_engine_destroy(g_aws.e);
    g_aws = AtomicWorkspace();
    std::vector<double> mass(na, kMassC);
    fmd_atomic_config c;
    fmd_atomic_config_default(&c);
    c.n_atoms = na;
    c.mol_id = mol_id;
    c.mass = mass.data();
    int rc = fmd_atomic_create(&c, &g_aws.e);
    if (rc) return rc;
    g_aws.mol_id.assign(mol_id, mol_id + na);
  }
  g_aws.e->at->p.rebo_mode = rebo_mode;
  g_aws.e->at->p.trace_virial = 0;  // the drop-ins return the full virial tensor
  g_aws.e->at->p.lj_alist = 0;      // and evaluate the given lists' pairs each call
  *out = g_aws.e;
  return FMD_OK;
}

// Upload pos and the (h, h^-1) pair, decide the image mode.
int stage_positions(fmd_engine *e, const double *pos, const double *h, const double *hi) {
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  HIPCHK(hipMemcpyAsync(A.d.pos, pos, 3 * (size_t)A.p.Na * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.hh, hh, sizeof hh, hipMemcpyHostToDevice, s));
  return atomic_image_mode(e, pos, h);
}

// Caller's REBO list (stride 12) and the truncation points of its LJ half
// list (stride 400) onto the device.
int stage_lists(fmd_engine *e, const int *nlc_r, const int *nll_r, const int *nlc_l,
                const int *nll_l) {
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  const size_t na = A.p.Na;
  HIPCHK(hipMemcpyAsync(A.d.nlc_r, nlc_r, na * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.nll_r, nll_r, na * kRefMaxRebo * sizeof(int), hipMemcpyHostToDevice,
                        s));
  int *dl = nullptr;
  int *dc = nullptr;
  HIPCHK(hipMallocAsync((void **)&dl, na * kRefMaxLj * sizeof(int), s));
  HIPCHK(hipMallocAsync((void **)&dc, na * sizeof(int), s));
  HIPCHK(hipMemcpyAsync(dc, nlc_l, na * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dl, nll_l, na * kRefMaxLj * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(at_launch_jmax_from_list(A.d, A.p, dc, dl, kRefMaxLj, s));
  HIPCHK(hipFreeAsync(dl, s));
  HIPCHK(hipFreeAsync(dc, s));
  return FMD_OK;
}

}  // namespace

extern "C" {

int fmd_airebo_nlist_rebo(const double *pos, const double *h, const double *hi, int Na,
                          const int *mol_id, int *nlc, int *nll) {
  if (!pos || !h || !hi || !nlc || !nll) return set_error(FMD_ERR_ARG, "nlist_rebo: null argument");
  fmd_engine *e;
  int rc = atomic_workspace(Na, mol_id, FMD_REBO_COMPAT, &e);
  if (rc || (rc = stage_positions(e, pos, h, hi))) return rc;
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  HIPCHK(hipMemsetAsync(A.d.nll_r, 0, (size_t)Na * kRefMaxRebo * sizeof(int), s));
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_lists(A.d, A.p, hh_fixed(A.d.hh), 1, s));
  HIPCHK(hipMemcpyAsync(nlc, A.d.nlc_r, (size_t)Na * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(nll, A.d.nll_r, (size_t)Na * kRefMaxRebo * sizeof(int),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FMD_OK;
}

int fmd_airebo_nlist_lj(const double *pos, const double *h, const double *hi, int Na,
                        const int *mol_id, int *nlc, int *nll) {
  if (!pos || !h || !hi || !nlc || !nll) return set_error(FMD_ERR_ARG, "nlist_lj: null argument");
  fmd_engine *e;
  int rc = atomic_workspace(Na, mol_id, FMD_REBO_COMPAT, &e);
  if (rc) return rc;
  return atomic_halflist(e, pos, h, hi, nlc, nll);
}

int fmd_airebo_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                              const double *hi, const int *nlc_r, const int *nll_r,
                              const int *nlc_l, const int *nll_l, int Na, const int *mol_id,
                              int rebo_mode, double *out3) {
  if (!F || !vir9 || !pos || !h || !hi || !nlc_r || !nll_r || !nlc_l || !nll_l)
    return set_error(FMD_ERR_ARG, "compute_forces: null argument");
  fmd_engine *e;
  int rc = atomic_workspace(Na, mol_id, rebo_mode, &e);
  if (rc || (rc = stage_positions(e, pos, h, hi))) return rc;
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  const int lj_compat = A.p.lj_compat;
  A.p.lj_compat = 0;  // the molecule list only; jmax comes from the caller's list
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_lists(A.d, A.p, hh_fixed(A.d.hh), 0, s));
  A.p.lj_compat = lj_compat;
  if ((rc = stage_lists(e, nlc_r, nll_r, nlc_l, nll_l))) return rc;
  HIPCHK(at_launch_forces(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_observe(A.d, A.p, kAtObsForce, s));
  AtomicState S;
  HIPCHK(hipMemcpyAsync(F, A.d.F, 3 * (size_t)Na * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&S, A.d.state, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) vir9[k] = S.W9[k];
  if (out3) {
    out3[0] = S.Ep;
    out3[1] = S.Ep_rebo;
    out3[2] = S.Ep_lj;
  }
  return FMD_OK;
}

int fmd_airebo_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                        double *hi, const double *mass, int Na, double dt,
                        fmd_npt_state *npt, const int *nlc_r, const int *nll_r,
                        const int *nlc_l, const int *nll_l, const int *mol_id,
                        int rebo_mode, double *out4) {
  if (!pos || !vel || !F || !vir9 || !h || !hi || !mass || !npt || !nlc_r || !nll_r ||
      !nlc_l || !nll_l)
    return set_error(FMD_ERR_ARG, "step_npt: null argument");
  fmd_engine *e;
  int rc = atomic_workspace(Na, mol_id, rebo_mode, &e);
  if (rc || (rc = stage_positions(e, pos, h, hi))) return rc;
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  A.p.dt = dt;
  HIPCHK(hipMemcpyAsync(A.d.mass, mass, (size_t)Na * 8, hipMemcpyHostToDevice, s));
  // the molecule list from the current positions, then the caller's lists
  const int lj_compat = A.p.lj_compat;
  A.p.lj_compat = 0;
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_lists(A.d, A.p, hh_fixed(A.d.hh), 0, s));
  A.p.lj_compat = lj_compat;
  if ((rc = stage_lists(e, nlc_r, nll_r, nlc_l, nll_l))) return rc;
  AtomicState S;
  if ((rc = atomic_step_once(e, vel, F, vir9, h, npt, &S))) return rc;
  HIPCHK(hipMemcpyAsync(pos, A.d.pos, 3 * (size_t)Na * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) hi[k] = S.hi[k];
  if (out4) {
    out4[0] = S.Ep;
    out4[1] = S.Ep_rebo;
    out4[2] = S.Ep_lj;
    out4[3] = S.KE;
  }
  return FMD_OK;
}

}  // extern "C"
