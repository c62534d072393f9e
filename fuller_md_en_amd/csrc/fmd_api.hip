// fmd_api.hip -- host implementation of the C ABI in include/fmd.h:
// the resident engine (HBM-resident state, stream-ordered steps) and the
// drop-in entry points that mirror the reference's function signatures.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fmd.h"
#include "fmd_engine.hpp"

using namespace fmd;

namespace fmd {

static thread_local std::string g_err;

int set_error(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_fail(hipError_t e, const char *what) {
  return set_error(FMD_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

LJ32 make_lj32(const fmd_lj_params &lj) {
  LJ32 p;
  p.sig2 = (float)(lj.sigma * lj.sigma);
  p.rc2 = (float)(lj.rcut * lj.rcut);
  p.rc2d = lj.rcut * lj.rcut;
  p.rc2_lo = (float)(p.rc2d * (1.0 - 1e-5));
  p.rc2_hi = (float)(p.rc2d * (1.0 + 1e-5));
  p.c48 = (float)(48.0 * lj.eps);
  p.c24 = (float)(24.0 * lj.eps);
  p.vs4 = (float)(lj.vshift / (4.0 * lj.eps));
  p.rmin2 = (float)lj.rmin2;
  return p;
}

}  // namespace fmd


// Device address of RigidState::h (h then h^-1).
static const double *state_hh(const RigidDev &d) {
  return reinterpret_cast<const double *>(reinterpret_cast<const char *>(d.state) +
                                          offsetof(RigidState, h));
}

extern "C" {

const char *fmd_last_error(void) { return g_err.c_str(); }
int fmd_version(void) { return 100; }

int fmd_device_info(char *name, int cap, int *cu_count, double *hbm_bytes) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t pr;
  HIPCHK(hipGetDeviceProperties(&pr, dev));
  if (name && cap > 0) snprintf(name, cap, "%s (%s)", pr.name, pr.gcnArchName);
  if (cu_count) *cu_count = pr.multiProcessorCount;
  if (hbm_bytes) *hbm_bytes = (double)pr.totalGlobalMem;
  return FMD_OK;
}

void fmd_lj_params_rigid(fmd_lj_params *p) {
  const double sigma = 3.431, eps = 2.635e-3;  // C1:93-96
  const double sr = 1.0 / 3.0, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;  // C1:101-104
  p->sigma = sigma;
  p->eps = eps;
  p->rcut = 3.0 * sigma;
  p->vshift = 4.0 * eps * (sr6 * sr6 - sr6);
  p->rmin2 = 0.25;
}

void fmd_lj_params_airebo(fmd_lj_params *p) {
  const double sigma = 3.40, eps = 2.84e-3;  // AB:119-124
  const double rc = 3.0 * sigma, sr = sigma / rc, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
  p->sigma = sigma;
  p->eps = eps;
  p->rcut = rc;
  p->vshift = 4 * eps * (sr6 * sr6 - sr6);
  p->rmin2 = 0.25;
}

void fmd_lj_params_mm(fmd_lj_params *p) {
  const double sigma = 3.431, eps = 2.635e-3;  // MM:111-117 (s / rc written as 1/3)
  const double rc = 3.0 * sigma, sr = 1.0 / 3.0, sr2 = sr * sr, sr6 = sr2 * sr2 * sr2;
  p->sigma = sigma;
  p->eps = eps;
  p->rcut = rc;
  p->vshift = 4.0 * eps * (sr6 * sr6 - sr6);
  p->rmin2 = 0.25;
}

void fmd_make_npt(fmd_npt_state *s, double T, double Pe, int Nf) {
  s->xi = 0;
  s->Q = std::max(Nf * kB * T * 100.0 * 100.0, 1e-20);
  for (int i = 0; i < 9; i++) s->Vg[i] = 0;
  s->W = std::max((Nf + 9) * kB * T * 1000.0 * 1000.0, 1e-20);
  s->Pe = Pe;
  s->Tt = T;
  s->Nf = Nf;
}

void fmd_rigid_config_default(fmd_rigid_config *c, double rmol) {
  memset(c, 0, sizeof *c);
  fmd_lj_params_rigid(&c->lj);
  c->rmcut = c->lj.rcut + 2 * rmol + 1.0;  // LJ:716
  c->skin = 3.0;                           // C1:282
  c->max_neigh = 80;                       // C1:121
  c->nlup = 25;                            // C1:617
  c->dt = 1.0;
  c->clamp_xi = 0.1;
  c->clamp_vg = 0.01;
  c->half_pairs = 1;
  c->cell_min_molecules = 3000;  // the scan is faster below about 2,500
}

}  // extern "C"

// ------------------------------------------------------------ rigid engine
namespace fmd {

static void free_rigid(RigidDev &d) {
  void *ptrs[] = {d.pos,      d.vel,      d.omg,      d.qv,       d.Fv,       d.Tv,
                  d.sites,    d.sites_d,  d.body,     d.qbody,    d.qcen,     d.hh,
                  d.nl_count, d.nl_list,  d.hl_count, d.hl_list,  d.in_count, d.in_list,
                  d.pairout,  d.flags,    d.molpart,  d.blockpart, d.state,   d.mid,
                  d.ctrl,     d.thermo,   d.hl_item,  d.ord_ctl,  d.nl_pos,   d.nl_h,
                  d.cell_dim, d.cell_n,   d.cell_m};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  memset(&d, 0, sizeof d);
}

template <class T>
static hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

int rigid_alloc(RigidDev &d, const RigidParams &p, int ring) {
  memset(&d, 0, sizeof d);
  const size_t N = p.N;
  d.nblk = (p.N + 3) / 4;  // k_rigid_kick2: one wavefront per molecule
  hipError_t e = hipSuccess;
#define AL(ptr, n) \
  if (e == hipSuccess) e = dalloc(&(ptr), (n))
  AL(d.pos, 3 * N);
  AL(d.vel, 3 * N);
  AL(d.omg, 3 * N);
  AL(d.qv, 4 * N);
  AL(d.Fv, 3 * N);
  AL(d.Tv, 3 * N);
  AL(d.sites, N * p.natom_pad);
  AL(d.sites_d, 3 * N * p.natom_pad);
  AL(d.body, 4 * (size_t)p.natom_pad);
  AL(d.hh, 18);
  AL(d.nl_count, N);
  AL(d.nl_list, N * p.max_neigh);
  AL(d.nl_pos, 3 * N);
  AL(d.nl_h, 10);
  if (p.use_half) {
    AL(d.qbody, 4 * (size_t)p.nq);
    AL(d.qcen, N * p.nq);
    AL(d.hl_count, N);
    AL(d.hl_list, N * p.max_neigh);
    AL(d.in_count, N);
    AL(d.in_list, N * p.max_neigh);
    AL(d.pairout, kPairOut * N * p.max_neigh);
  }
  if (p.pair_kernel) {
    AL(d.hl_item, N * p.max_neigh);
    AL(d.ord_ctl, 1);
  }
  if (p.cell_min > 0 && p.N >= p.cell_min) {
    const size_t nc = std::max<size_t>(27, N);
    AL(d.cell_dim, 1);
    AL(d.cell_n, nc);
    AL(d.cell_m, nc * kCellCap);
  }
  AL(d.flags, 4);
  AL(d.molpart, N * kMolPart);
  AL(d.blockpart, (size_t)d.nblk * kBlkPart);
  AL(d.state, 1);
  AL(d.mid, 1);
  AL(d.ctrl, ring);
  AL(d.thermo, ring);
#undef AL
  if (e != hipSuccess) {
    free_rigid(d);
    return hip_fail(e, "hipMalloc (rigid engine)");
  }
  (void)hipMemset(d.flags, 0, 4 * sizeof(int));
  (void)hipMemset(d.nl_count, 0, N * sizeof(int));
  (void)hipMemset(d.nl_h, 0, 10 * sizeof(double));
  if (d.ord_ctl) (void)hipMemset(d.ord_ctl, 0, sizeof(int));
  (void)hipMemset(d.Fv, 0, 3 * N * sizeof(double));
  (void)hipMemset(d.Tv, 0, 3 * N * sizeof(double));
  (void)hipMemset(d.state, 0, sizeof(RigidState));
  return FMD_OK;
}

// Body sites regrouped into spatially compact quads (groups of 4 consecutive
// device sites) for the half-pair kernel's per-quad cull.  Greedy: the site
// farthest from the centroid of those left, then its 3 nearest left.  The
// order is internal to the device: sums over sites do not depend on it.
static void quad_order(const double *body, int natom, int *perm) {
  std::vector<int> left(natom);
  for (int a = 0; a < natom; a++) left[a] = a;
  int np = 0;
  auto d2 = [&](int a, const double *c) {
    double x = body[3 * a] - c[0], y = body[3 * a + 1] - c[1], z = body[3 * a + 2] - c[2];
    return x * x + y * y + z * z;
  };
  while (!left.empty()) {
    double c[3] = {0, 0, 0};
    for (int a : left)
      for (int k = 0; k < 3; k++) c[k] += body[3 * a + k] / left.size();
    size_t far = 0;
    for (size_t t = 1; t < left.size(); t++)
      if (d2(left[t], c) > d2(left[far], c)) far = t;
    const int s0 = left[far];
    std::stable_sort(left.begin(), left.end(), [&](int x, int y) {
      return d2(x, body + 3 * s0) < d2(y, body + 3 * s0);
    });
    const size_t take = std::min<size_t>(4, left.size());
    for (size_t t = 0; t < take; t++) perm[np++] = left[t];
    left.erase(left.begin(), left.begin() + take);
  }
}

// body4 receives 4 * natom_pad doubles (at most 4 * 128); no heap memory is
// used unless qbody is asked for.
int rigid_params(RigidParams &p, int N, int natom, const double *body, double I0, double Mmol,
                 const fmd_lj_params &lj, double rmcut2, double rlist, int max_neigh,
                 double dt, double cxi, double cvg, double *body4,
                 std::vector<double> *qbody) {
  if (N < 1 || natom < 1 || natom > 128 || max_neigh < 1)
    return set_error(FMD_ERR_ARG, "rigid: need N >= 1, 1 <= natom <= 128, max_neigh >= 1");
  memset(&p, 0, sizeof p);
  p.N = N;
  p.natom = natom;
  p.natom_pad = (natom + 3) & ~3;
  p.max_neigh = max_neigh;
  p.rmcut2 = rmcut2;
  p.rlist2 = rlist * rlist;
  p.dt = dt;
  p.I0 = I0;
  p.Mmol = Mmol;
  p.clamp_xi = cxi;
  p.clamp_vg = cvg;
  p.lj = make_lj32(lj);
  p.eps = lj.eps;
  p.isig = (float)(1.0 / std::sqrt((double)p.lj.sig2));
  p.rc2_lo_s = p.lj.rc2_lo / p.lj.sig2;
  p.rc2_hi_s = p.lj.rc2_hi / p.lj.sig2;
  p.rmin2_s = p.lj.rmin2 / p.lj.sig2;
  p.c48s = 48.0 * lj.eps / std::sqrt((double)p.lj.sig2);
  int perm[128];
  for (int a = 0; a < natom; a++) perm[a] = a;
  if (qbody) quad_order(body, natom, perm);
  double rb = 0;
  for (int k = 0; k < 4 * p.natom_pad; k++) body4[k] = 0.0;
  for (int a = 0; a < p.natom_pad; a++) {
    if (a < natom) {
      const double *b = body + 3 * perm[a];
      for (int k = 0; k < 3; k++) body4[4 * a + k] = b[k];
      rb = std::max(rb, std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]));
    } else {
      body4[4 * a + 3] = 1.0;
    }
  }
  const double cr = lj.rcut + rb + 0.05;  // FP32 margin
  p.cull_r2 = cr * cr;
  if (qbody) {
    p.use_half = 1;
    p.pair_kernel = natom <= 64;
    p.nq = p.natom_pad / 4;
    p.reach = (float)cr;
    qbody->assign(4 * (size_t)p.nq, 0.0);
    for (int q = 0; q < p.nq; q++) {
      double c[3] = {0, 0, 0}, r = 0;
      int n = 0;
      for (int a = 4 * q; a < 4 * q + 4 && a < natom; a++, n++)
        for (int k = 0; k < 3; k++) c[k] += body4[4 * a + k];
      for (int k = 0; k < 3; k++) c[k] /= n;
      for (int a = 4 * q; a < 4 * q + 4 && a < natom; a++) {
        double x = body4[4 * a] - c[0], y = body4[4 * a + 1] - c[1], z = body4[4 * a + 2] - c[2];
        r = std::max(r, std::sqrt(x * x + y * y + z * z));
      }
      for (int k = 0; k < 3; k++) (*qbody)[4 * q + k] = c[k];
      (*qbody)[4 * q + 3] = r;
    }
  }
  return FMD_OK;
}

}  // namespace fmd



int fmd::engine_init_common(fmd_engine *e, int ring, int R) {
  e->ring = ring;
  HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
  HIPCHK(hipHostMalloc((void **)&e->h_ctrl, (size_t)ring * R * sizeof(StepCtrl)));
  HIPCHK(hipHostMalloc((void **)&e->h_thermo, (size_t)ring * R * sizeof(Thermo)));
  return FMD_OK;
}

extern "C" {

int fmd_rigid_create(const fmd_rigid_config *c, fmd_engine **out) {
  if (!c || !out || !c->body) return set_error(FMD_ERR_ARG, "fmd_rigid_create: null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_error(FMD_ERR_NODEVICE, "fmd_rigid_create: no HIP device");
  fmd_engine *e = new fmd_engine();
  e->kind = FMD_KIND_RIGID;
  std::vector<double> body4(4 * 128), qbody;
  int rc = rigid_params(e->rp, c->n_mol, c->natom, c->body, c->I0, c->Mmol, c->lj,
                        c->rmcut * c->rmcut, c->rmcut + c->skin, c->max_neigh, c->dt,
                        c->clamp_xi, c->clamp_vg, body4.data(),
                        c->half_pairs ? &qbody : nullptr);
  if (c->half_pairs == 2) e->rp.pair_kernel = 0;  // the per-molecule pair-once kernel
  e->rp.cell_min = std::max(0, c->cell_min_molecules);
  if (rc == FMD_OK) rc = engine_init_common(e, 4096, 1);
  if (rc == FMD_OK) rc = rigid_alloc(e->rd, e->rp, e->ring);
  if (rc != FMD_OK) {
    fmd_engine_destroy(e);
    return rc;
  }
  e->nlup = c->nlup > 0 ? c->nlup : 25;
  hipError_t he = hipMemcpy(e->rd.body, body4.data(), 4 * e->rp.natom_pad * sizeof(double),
                            hipMemcpyHostToDevice);
  if (he == hipSuccess && e->rp.use_half)
    he = hipMemcpy(e->rd.qbody, qbody.data(), qbody.size() * sizeof(double),
                   hipMemcpyHostToDevice);
  if (he != hipSuccess) {
    fmd_engine_destroy(e);
    return hip_fail(he, "upload body");
  }
  *out = e;
  return FMD_OK;
}

int fmd_rigid_set_state(fmd_engine *e, const double *pos, const double *vel, const double *qv,
                        const double *omg, const double *h, const fmd_npt_state *npt) {
  if (!e || e->kind != FMD_KIND_RIGID) return set_error(FMD_ERR_ARG, "not a rigid engine");
  RigidDev &d = e->rd;
  const size_t N = e->rp.N;
  hipStream_t s = e->stream;
  if (pos) HIPCHK(hipMemcpyAsync(d.pos, pos, 3 * N * 8, hipMemcpyHostToDevice, s));
  if (vel) HIPCHK(hipMemcpyAsync(d.vel, vel, 3 * N * 8, hipMemcpyHostToDevice, s));
  if (qv) HIPCHK(hipMemcpyAsync(d.qv, qv, 4 * N * 8, hipMemcpyHostToDevice, s));
  if (omg) HIPCHK(hipMemcpyAsync(d.omg, omg, 3 * N * 8, hipMemcpyHostToDevice, s));
  RigidState S;
  HIPCHK(hipMemcpyAsync(&S, d.state, sizeof S, hipMemcpyDeviceToHost, s));
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
    e->tt = npt->Tt;
  }
  S.scale = 1.0;
  HIPCHK(hipMemcpyAsync(d.state, &S, sizeof S, hipMemcpyHostToDevice, s));
  // a new state: the next list build has no previous one to check
  if (pos || h) HIPCHK(hipMemsetAsync(d.nl_h + 9, 0, sizeof(double), s));
  HIPCHK(hipStreamSynchronize(s));
  e->prepared = false;
  return FMD_OK;
}

int fmd_rigid_get_state(fmd_engine *e, double *pos, double *vel, double *qv, double *omg,
                        double *h, fmd_npt_state *npt) {
  if (!e || e->kind != FMD_KIND_RIGID) return set_error(FMD_ERR_ARG, "not a rigid engine");
  RigidDev &d = e->rd;
  const size_t N = e->rp.N;
  hipStream_t s = e->stream;
  if (pos) HIPCHK(hipMemcpyAsync(pos, d.pos, 3 * N * 8, hipMemcpyDeviceToHost, s));
  if (vel) HIPCHK(hipMemcpyAsync(vel, d.vel, 3 * N * 8, hipMemcpyDeviceToHost, s));
  if (qv) HIPCHK(hipMemcpyAsync(qv, d.qv, 4 * N * 8, hipMemcpyDeviceToHost, s));
  if (omg) HIPCHK(hipMemcpyAsync(omg, d.omg, 3 * N * 8, hipMemcpyDeviceToHost, s));
  RigidState S;
  HIPCHK(hipMemcpyAsync(&S, d.state, sizeof S, hipMemcpyDeviceToHost, s));
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

}  // extern "C"

// ------------------------------------------------------ step sequencing
namespace fmd {

int engine_graph(fmd_engine *e, const std::string &key, const std::function<int()> &launch,
                 hipGraphExec_t *out) {
  for (auto &g : e->graphs)
    if (g.first == key) {
      *out = g.second;
      return FMD_OK;
    }
  if (e->graphs.size() >= 32) engine_drop_graphs(e);
  HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  int rc = launch();
  hipGraph_t graph = nullptr;
  hipError_t he = hipStreamEndCapture(e->stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (he != hipSuccess) return hip_fail(he, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  he = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (he != hipSuccess) return hip_fail(he, "hipGraphInstantiate");
  e->graphs.emplace_back(key, exec);
  *out = exec;
  return FMD_OK;
}

void engine_drop_graphs(fmd_engine *e) {
  for (auto &g : e->graphs) (void)hipGraphExecDestroy(g.second);
  e->graphs.clear();
}

int run_steps(fmd_engine *e, int R, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
              fmd_thermo *thermo, double *tt, StepCtrl *d_ctrl, Thermo *d_thermo,
              const StepLaunch &launch) {
  hipStream_t s = e->stream;
  std::vector<char> rebuild(e->ring), rescale(e->ring);
  for (int c0 = 0; c0 < nsteps; c0 += e->ring) {
    const int n = std::min(e->ring, nsteps - c0);
    for (int k = 0; k < n; k++) {
      const long long g = (long long)first_step + c0 + k + 1;
      int any_rebuild = 0, any_rescale = 0;
      for (int r = 0; r < R; r++) {
        StepCtrl &c = e->h_ctrl[(size_t)k * R + r];
        if (ctrl) {
          const fmd_step_ctrl &u = ctrl[(size_t)(c0 + k) * R + r];
          c.Tt = u.Tt;
          c.rescale_T = u.rescale_T;
          c.flags = u.flags;
        } else {
          c.Tt = tt[r];
          c.rescale_T = 0;
          c.flags = 0;
        }
        c.pad = 0;
        any_rebuild |= c.flags & FMD_CTRL_REBUILD;
        any_rescale |= c.flags & FMD_CTRL_RESCALE;
      }
      rebuild[k] = (g % e->nlup == 0) || any_rebuild;
      rescale[k] = any_rescale != 0;
    }
    if (ctrl)
      for (int r = 0; r < R; r++) tt[r] = ctrl[(size_t)(c0 + n - 1) * R + r].Tt;
    const size_t cb = (size_t)R * sizeof(StepCtrl), tb = (size_t)R * sizeof(Thermo);
    while (e->timing && (int)e->ev.size() < kEvPerStep * n) {
      hipEvent_t ev;
      HIPCHK(hipEventCreate(&ev));
      e->ev.push_back(ev);
    }
    // slot k is timed on every t_stride-th slot and on list rebuilds (a
    // graph replay rebuilds at its slot 0 only); events order the stream and
    // write back caches, so sparse sampling keeps the timed run unperturbed
    auto timed = [&](int k, int rb) { return e->timing && (k % e->t_stride == 0 || rb); };
    // events of step slot k (graph replays reuse slots 0 .. K-1)
    auto step_events = [&](int k, int rb, unsigned flags, StepEvents *out) -> const StepEvents * {
      if (!timed(k, rb)) return nullptr;
      const hipEvent_t *v = &e->ev[(size_t)kEvPerStep * k];
      *out = StepEvents{v[0], v[1], v[2], v[3], v[4], flags};
      return out;
    };
    auto elapsed = [&](hipEvent_t a, hipEvent_t b, int phase) -> int {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, a, b));
      e->t_sum[phase] += ms;
      e->t_n[phase]++;
      return FMD_OK;
    };
    // durations of slots 0 .. m-1, whose steps were global steps k0 .. k0+m-1
    // of this chunk
    auto take_times = [&](int k0, int m) -> int {
      for (int k = 0; k < m; k++) {
        if (!timed(k, rebuild[k0 + k])) continue;
        const hipEvent_t *v = &e->ev[(size_t)kEvPerStep * k];
        if (elapsed(v[0], v[2], kPhaseForce) || elapsed(v[0], v[1], kPhaseIntra) ||
            elapsed(v[1], v[2], kPhaseLJ))
          return FMD_ERR_HIP;
        if (rebuild[k0 + k] && elapsed(v[3], v[4], kPhaseList)) return FMD_ERR_HIP;
      }
      return FMD_OK;
    };
    if (!e->use_graph) {
      HIPCHK(hipMemcpyAsync(d_ctrl, e->h_ctrl, n * cb, hipMemcpyHostToDevice, s));
      // a step's finish is folded into the next one's launch within the
      // chunk, except before a list rebuild
      for (int k = 0; k < n; k++) {
        StepEvents sev;
        const int fold = k > 0 && !rebuild[k], fin = k == n - 1 || rebuild[k + 1];
        HIPCHK(launch(k, rebuild[k], rescale[k], fold, fin,
                      step_events(k, rebuild[k], hipEventRecordDefault, &sev)));
      }
      if (thermo)
        HIPCHK(hipMemcpyAsync(e->h_thermo, d_thermo, n * tb, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      if (e->timing && take_times(0, n)) return FMD_ERR_HIP;
    } else {
      // runs of steps with a list rebuild only at their first step, each
      // replayed from a graph captured once per (length, rebuild, rescale)
      // pattern; controls go to slots [0, K) before and records come back
      // from them after every replay (stream order keeps them apart)
      constexpr int kMaxGraphSteps = 64;
      for (int k = 0; k < n;) {
        int K = 1;
        while (k + K < n && !rebuild[k + K] && K < kMaxGraphSteps) K++;
        std::string key = std::to_string(K) + (rebuild[k] ? "R" : "-") +
                          (e->timing ? "T" + std::to_string(e->t_stride) : "-");
        for (int j = 0; j < K; j++) key += rescale[k + j] ? '1' : '0';
        hipGraphExec_t exec;
        int rc = engine_graph(e, key, [&]() -> int {
          for (int j = 0; j < K; j++) {
            StepEvents sev;
            const int rb = j == 0 ? rebuild[k] : 0;
            HIPCHK(launch(j, rb, rescale[k + j], j > 0, j == K - 1,
                          step_events(j, rb, hipEventRecordExternal, &sev)));
          }
          return FMD_OK;
        }, &exec);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(d_ctrl, e->h_ctrl + (size_t)k * R, K * cb, hipMemcpyHostToDevice,
                              s));
        HIPCHK(hipGraphLaunch(exec, s));
        if (thermo)
          HIPCHK(hipMemcpyAsync(e->h_thermo + (size_t)k * R, d_thermo, K * tb,
                                hipMemcpyDeviceToHost, s));
        if (e->timing) {
          HIPCHK(hipStreamSynchronize(s));
          if (take_times(k, K)) return FMD_ERR_HIP;
        }
        k += K;
      }
      HIPCHK(hipStreamSynchronize(s));
    }
    for (int k = 0; k < n; k++) e->rebuilds += rebuild[k];
    if (thermo) memcpy(thermo + (size_t)c0 * R, e->h_thermo, n * tb);
    e->steps += n;
  }
  return FMD_OK;
}

}  // namespace fmd

static int rigid_prepare(fmd_engine *e) {
  RigidDev &d = e->rd;
  const RigidParams &p = e->rp;
  hipStream_t s = e->stream;
  const double *hh = state_hh(d);
  HIPCHK(rigid_launch_apply_pbc(d, p, hh, s));
  HIPCHK(rigid_launch_rebuild(d, p, hh, s));
  if (p.use_half) HIPCHK(rigid_launch_halflist(d, p, hh, s));
  HIPCHK(rigid_launch_sites(d, p, s));
  HIPCHK(rigid_launch_forces(d, p, hh, s));
  HIPCHK(rigid_launch_observe(d, p, kObsKE | kObsForce, s));
  e->rebuilds++;
  e->prepared = true;
  return FMD_OK;
}

// List overflows and list faults since the last read, into the fault counter.
static int read_overflow(fmd_engine *e, int *flags_dev) {
  int f[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(f, flags_dev, sizeof f, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  if (f[0] || f[1]) {
    e->overflow += f[0] + f[1];
    HIPCHK(hipMemsetAsync(flags_dev, 0, sizeof f, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
  }
  return FMD_OK;
}

static int rigid_run(fmd_engine *e, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
                     fmd_thermo *thermo) {
  RigidDev &d = e->rd;
  const RigidParams &p = e->rp;
  if (!e->prepared) {
    int rc = rigid_prepare(e);
    if (rc) return rc;
  }
  auto launch = [&](int slot, int rebuild, int rescale, int, int,
                    const StepEvents *ev) -> hipError_t {
    return rigid_launch_step(d, p, slot, slot, rebuild, rescale, ev, e->stream);
  };
  int rc = run_steps(e, 1, first_step, nsteps, ctrl, thermo, &e->tt, d.ctrl, d.thermo, launch);
  if (rc) return rc;
  return read_overflow(e, d.flags);
}

extern "C" {

int fmd_engine_prepare(fmd_engine *e) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  if (e->kind == FMD_KIND_RIGID) {
    int rc = rigid_prepare(e);
    if (rc) return rc;
    return read_overflow(e, e->rd.flags);
  }
  return engine_prepare_atomic(e);
}

int fmd_engine_run(fmd_engine *e, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
                   fmd_thermo *thermo) {
  if (!e || nsteps < 0) return set_error(FMD_ERR_ARG, "fmd_engine_run: bad argument");
  static_assert(sizeof(fmd_thermo) == sizeof(Thermo), "thermo layout");
  static_assert(sizeof(fmd_step_ctrl) == sizeof(StepCtrl), "ctrl layout");
  if (nsteps == 0) return FMD_OK;
  if (e->kind == FMD_KIND_RIGID) return rigid_run(e, first_step, nsteps, ctrl, thermo);
  return engine_run_atomic(e, first_step, nsteps, ctrl, thermo);
}

int fmd_engine_forces(fmd_engine *e, double *F, double *T, double *W9, double *energies) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  if (e->kind != FMD_KIND_RIGID) return engine_forces_atomic(e, F, T, W9, energies);
  if (!e->prepared) {
    int rc = rigid_prepare(e);
    if (rc) return rc;
  }
  RigidDev &d = e->rd;
  const size_t N = e->rp.N;
  hipStream_t s = e->stream;
  if (e->rp.use_half && e->steps > 0) {
    // steps keep only the virial trace: the full tensor from the both-sides
    // kernel at the current state
    double out10[kMolPart];
    HIPCHK(rigid_launch_forces(d, e->rp, state_hh(d), s));
    HIPCHK(rigid_launch_reduce_molpart(d, e->rp, d.blockpart, s));
    if (F) HIPCHK(hipMemcpyAsync(F, d.Fv, 3 * N * 8, hipMemcpyDeviceToHost, s));
    if (T) HIPCHK(hipMemcpyAsync(T, d.Tv, 3 * N * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(out10, d.blockpart, sizeof out10, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (W9)
      for (int k = 0; k < 9; k++) W9[k] = out10[k];
    if (energies) {
      energies[0] = out10[9];
      energies[1] = 0;
      energies[2] = out10[9];
    }
    return FMD_OK;
  }
  if (F) HIPCHK(hipMemcpyAsync(F, d.Fv, 3 * N * 8, hipMemcpyDeviceToHost, s));
  if (T) HIPCHK(hipMemcpyAsync(T, d.Tv, 3 * N * 8, hipMemcpyDeviceToHost, s));
  RigidState S;
  HIPCHK(hipMemcpyAsync(&S, d.state, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (W9)
    for (int k = 0; k < 9; k++) W9[k] = S.W9[k];
  if (energies) {
    energies[0] = S.Ep;
    energies[1] = 0;
    energies[2] = S.Ep;
  }
  return FMD_OK;
}

int fmd_engine_counters(fmd_engine *e, long long *out3) {
  if (!e || !out3) return set_error(FMD_ERR_ARG, "null argument");
  out3[0] = e->overflow;
  out3[1] = e->rebuilds;
  out3[2] = e->steps;
  return FMD_OK;
}

int fmd_engine_set_timing(fmd_engine *e, int on) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  e->timing = on != 0;
  e->t_stride = on > 1 ? on : 1;
  for (int k = 0; k < 4; k++) {
    e->t_sum[k] = 0;
    e->t_n[k] = 0;
  }
  return FMD_OK;
}

int fmd_engine_set_graph(fmd_engine *e, int on) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  e->use_graph = on != 0;
  return FMD_OK;
}

int fmd_engine_kernel_time(fmd_engine *e, double *mean_ms, int *n) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  if (mean_ms) *mean_ms = e->t_n[kPhaseForce] ? e->t_sum[kPhaseForce] / e->t_n[kPhaseForce] : 0.0;
  if (n) *n = e->t_n[kPhaseForce];
  for (int k = 0; k < 4; k++) {
    e->t_sum[k] = 0;
    e->t_n[k] = 0;
  }
  return FMD_OK;
}

int fmd_engine_phase_times(fmd_engine *e, double *mean_ms4, int *n4) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  for (int k = 0; k < 4; k++) {
    if (mean_ms4) mean_ms4[k] = e->t_n[k] ? e->t_sum[k] / e->t_n[k] : 0.0;
    if (n4) n4[k] = e->t_n[k];
    e->t_sum[k] = 0;
    e->t_n[k] = 0;
  }
  return FMD_OK;
}

int fmd_engine_sync(fmd_engine *e) {
  if (!e) return set_error(FMD_ERR_ARG, "null engine");
  HIPCHK(hipStreamSynchronize(e->stream));
  return FMD_OK;
}

void fmd_engine_destroy(fmd_engine *e) {
  if (!e) return;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  engine_drop_graphs(e);
  if (e->kind == FMD_KIND_RIGID) free_rigid(e->rd);
  engine_free_atomic(e);
  for (hipEvent_t ev : e->ev) (void)hipEventDestroy(ev);
  if (e->h_ctrl) (void)hipHostFree(e->h_ctrl);
  if (e->h_thermo) (void)hipHostFree(e->h_thermo);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"

// ------------------------------------------------------ rigid drop-ins
// One per-process workspace engine, re-created when the shape changes.
namespace {

struct RigidWorkspace {
  fmd_engine *e = nullptr;
  int N = 0, natom = 0, max_neigh = 0;
};
RigidWorkspace g_rws;

int rigid_workspace(int N, int natom, int max_neigh, const double *body, double I0, double Mmol,
                    const fmd_lj_params *lj, double rmcut2, double dt, fmd_engine **out) {
  fmd_lj_params ljd;
  if (!lj) {
    fmd_lj_params_rigid(&ljd);
    lj = &ljd;
  }
  if (!g_rws.e || g_rws.N != N || g_rws.natom != natom || g_rws.max_neigh != max_neigh) {
    if (g_rws.e) fmd_engine_destroy(g_rws.e);
    g_rws = RigidWorkspace();
    fmd_rigid_config c;
    fmd_rigid_config_default(&c, 0.0);
    c.half_pairs = 0;
    c.n_mol = N;
    c.natom = natom;
    c.body = body;
    c.I0 = I0;
    c.Mmol = Mmol;
    c.lj = *lj;
    c.max_neigh = max_neigh;
    int rc = fmd_rigid_create(&c, &g_rws.e);
    if (rc) return rc;
    g_rws.N = N;
    g_rws.natom = natom;
    g_rws.max_neigh = max_neigh;
  }
  fmd_engine *e = g_rws.e;
  std::vector<double> body4(4 * 128);
  const int cell_min = e->rp.cell_min;  // as allocated
  int rc = rigid_params(e->rp, N, natom, body, I0, Mmol, *lj, rmcut2, 0.0, max_neigh, dt, 0.1,
                        0.01, body4.data(), nullptr);
  if (rc) return rc;
  e->rp.cell_min = cell_min;
  HIPCHK(hipMemcpyAsync(e->rd.body, body4.data(), 4 * e->rp.natom_pad * sizeof(double),
                        hipMemcpyHostToDevice, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  *out = e;
  return FMD_OK;
}

int upload_lists(fmd_engine *e, const int *nl_count, const int *nl_list) {
  const size_t N = e->rp.N;
  HIPCHK(hipMemcpyAsync(e->rd.nl_count, nl_count, N * sizeof(int), hipMemcpyHostToDevice,
                        e->stream));
  HIPCHK(hipMemcpyAsync(e->rd.nl_list, nl_list, N * e->rp.max_neigh * sizeof(int),
                        hipMemcpyHostToDevice, e->stream));
  return FMD_OK;
}

// The reference's lab scratch (C1:354-364), filled on the host in FP64.
void host_lab(const double *qv, const double *body, int N, int natom, double *lab) {
  for (int i = 0; i < N; i++) {
    double R[9];
    q2R(&qv[i * 4], R);
    for (int a = 0; a < natom; a++) {
      const double *b = body + 3 * a;
      double *o = lab + (size_t)i * natom * 3 + a * 3;
      o[0] = R[0] * b[0] + R[1] * b[1] + R[2] * b[2];
      o[1] = R[3] * b[0] + R[4] * b[1] + R[5] * b[2];
      o[2] = R[6] * b[0] + R[7] * b[1] + R[8] * b[2];
    }
  }
}

}  // namespace

extern "C" {

int fmd_rigid_nlist(const double *pos, const double *h, const double *hi, int N, double rmcut,
                    int max_neigh, int *nl_count, int *nl_list, int *overflow) {
  if (!pos || !h || !hi || !nl_count || !nl_list || N < 1 || max_neigh < 1)
    return set_error(FMD_ERR_ARG, "fmd_rigid_nlist: bad argument");
  const double one[3] = {0, 0, 0};
  fmd_engine *e;
  int rc = rigid_workspace(N, 1, max_neigh, one, 1.0, 1.0, nullptr, 0.0, 1.0, &e);
  if (rc) return rc;
  e->rp.rlist2 = (rmcut + 3.0) * (rmcut + 3.0);  // C1:282
  RigidDev &d = e->rd;
  hipStream_t s = e->stream;
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  HIPCHK(hipMemcpyAsync(d.pos, pos, 3 * (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.hh, hh, sizeof hh, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemsetAsync(d.flags, 0, sizeof(int), s));
  HIPCHK(rigid_launch_nlist(d, e->rp, d.hh, s));
  int ovf = 0;
  HIPCHK(hipMemcpyAsync(nl_count, d.nl_count, N * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(nl_list, d.nl_list, (size_t)N * max_neigh * sizeof(int),
                        hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&ovf, d.flags, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (overflow) *overflow = ovf;
  return FMD_OK;
}

int fmd_rigid_forces(double *Fv, double *Tv, double *Wm9, const double *pos, const double *qv,
                     const double *body, const double *h, const double *hi,
                     const int *nl_count, const int *nl_list, int max_neigh, int N, int natom,
                     double rmcut2, const fmd_lj_params *lj, double *lab, double *Ep) {
  if (!Fv || !Tv || !Wm9 || !pos || !qv || !body || !h || !hi || !nl_count || !nl_list)
    return set_error(FMD_ERR_ARG, "fmd_rigid_forces: null argument");
  fmd_engine *e;
  int rc = rigid_workspace(N, natom, max_neigh, body, 1.0, 1.0, lj, rmcut2, 1.0, &e);
  if (rc) return rc;
  RigidDev &d = e->rd;
  hipStream_t s = e->stream;
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  HIPCHK(hipMemcpyAsync(d.pos, pos, 3 * (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.qv, qv, 4 * (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.hh, hh, sizeof hh, hipMemcpyHostToDevice, s));
  if ((rc = upload_lists(e, nl_count, nl_list))) return rc;
  HIPCHK(rigid_launch_sites(d, e->rp, s));
  HIPCHK(rigid_launch_forces(d, e->rp, d.hh, s));
  HIPCHK(rigid_launch_reduce_molpart(d, e->rp, d.blockpart, s));
  double out10[kMolPart];
  HIPCHK(hipMemcpyAsync(Fv, d.Fv, 3 * (size_t)N * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(Tv, d.Tv, 3 * (size_t)N * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(out10, d.blockpart, sizeof out10, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) Wm9[k] = out10[k];
  if (Ep) *Ep = out10[9];
  if (lab) host_lab(qv, body, N, natom, lab);
  return FMD_OK;
}

int fmd_rigid_step_npt(double *pos, double *vel, double *qv, double *omg, double *Fv,
                       double *Tv, double *Wm9, double *h, double *hi, const double *body,
                       double I0, double Mmol, int N, int natom, double rmcut2, double dt,
                       fmd_npt_state *npt, const int *nl_count, const int *nl_list,
                       int max_neigh, const fmd_lj_params *lj, double *Ep, double *KE) {
  if (!pos || !vel || !qv || !omg || !Fv || !Tv || !Wm9 || !h || !hi || !body || !npt ||
      !nl_count || !nl_list)
    return set_error(FMD_ERR_ARG, "fmd_rigid_step_npt: null argument");
  fmd_engine *e;
  int rc = rigid_workspace(N, natom, max_neigh, body, I0, Mmol, lj, rmcut2, dt, &e);
  if (rc) return rc;
  RigidDev &d = e->rd;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)N * 8;
  HIPCHK(hipMemcpyAsync(d.pos, pos, n3, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.vel, vel, n3, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.qv, qv, 4 * (size_t)N * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.omg, omg, n3, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.Fv, Fv, n3, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.Tv, Tv, n3, hipMemcpyHostToDevice, s));
  if ((rc = upload_lists(e, nl_count, nl_list))) return rc;
  RigidState S;
  memset(&S, 0, sizeof S);
  for (int k = 0; k < 9; k++) {
    S.h[k] = h[k];
    S.Vg[k] = npt->Vg[k];
    S.W9[k] = Wm9[k];
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
  HIPCHK(hipMemcpyAsync(d.state, &S, sizeof S, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.ctrl, &c, sizeof c, hipMemcpyHostToDevice, s));
  HIPCHK(rigid_launch_observe(d, e->rp, kObsKE, s));
  HIPCHK(rigid_launch_step(d, e->rp, 0, 0, 0, 0, nullptr, s));
  HIPCHK(hipMemcpyAsync(pos, d.pos, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(vel, d.vel, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(qv, d.qv, 4 * (size_t)N * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(omg, d.omg, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(Fv, d.Fv, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(Tv, d.Tv, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&S, d.state, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) {
    h[k] = S.h[k];
    hi[k] = S.hi[k];
    Wm9[k] = S.W9[k];
    npt->Vg[k] = S.Vg[k];
  }
  npt->xi = S.xi;
  if (Ep) *Ep = S.Ep;
  if (KE) *KE = S.KEt + S.KEr;
  return FMD_OK;
}

}  // extern "C"

// ------------------------------------------- rigid forces, no allocation
// fmd_rigid_forces_noalloc keeps every device buffer it needs in `scratch`,
// one file-scope array per device, instead of the workspace engine.  The
// array's name is a host-side stand-in: host code uses only the device
// address that hipGetSymbolAddress returns for the current device.
namespace {

constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;  // 16 MiB per device
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kNoAllocMaxDevices];  // scratch's address on each device

// Device address of `scratch` on the current device, looked up once per device.
int scratch_address(unsigned char **out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return set_error(FMD_ERR_ARG, "fmd_rigid_forces_noalloc: device %d beyond %d", dev,
                     kNoAllocMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_ptr[dev];
  return FMD_OK;
}

// Carves consecutive 256-byte aligned pieces out of a byte range; with a
// null base it only counts.
struct Carve {
  unsigned char *base;
  size_t used = 0;
  template <class T>
  T *take(size_t n) {
    T *p = base ? reinterpret_cast<T *>(base + used) : nullptr;
    used += (n * sizeof(T) + 255) & ~size_t(255);
    return p;
  }
};

void carve_rigid(Carve &c, RigidDev &d, double *&hh, double *&out10, int N, int natom_pad,
                 int max_neigh) {
  const size_t n = (size_t)N;
  hh = c.take<double>(18);
  out10 = c.take<double>(kMolPart);
  d.body = c.take<double>(4 * (size_t)natom_pad);
  d.pos = c.take<double>(3 * n);
  d.qv = c.take<double>(4 * n);
  d.nl_count = c.take<int>(n);
  d.nl_list = c.take<int>(n * max_neigh);
  d.sites = c.take<float4>(n * natom_pad);
  d.sites_d = c.take<double>(3 * n * natom_pad);
  d.Fv = c.take<double>(3 * n);
  d.Tv = c.take<double>(3 * n);
  d.molpart = c.take<double>(n * kMolPart);
}

}  // namespace

extern "C" {

int fmd_rigid_forces_noalloc(double *Fv, double *Tv, double *Wm9, const double *pos,
                             const double *qv, const double *body, const double *h,
                             const double *hi, const int *nl_count, const int *nl_list,
                             int max_neigh, int N, int natom, double rmcut2,
                             const fmd_lj_params *lj, double *lab, double *Ep) {
  if (!Fv || !Tv || !Wm9 || !pos || !qv || !body || !h || !hi || !nl_count || !nl_list)
    return set_error(FMD_ERR_ARG, "fmd_rigid_forces_noalloc: null argument");
  fmd_lj_params ljd;
  if (!lj) {
    fmd_lj_params_rigid(&ljd);
    lj = &ljd;
  }
  RigidParams p;
  double body4[4 * 128];
  int rc = rigid_params(p, N, natom, body, 1.0, 1.0, *lj, rmcut2, 0.0, max_neigh, 1.0, 0.1,
                        0.01, body4, nullptr);
  if (rc) return rc;
  RigidDev d;
  memset(&d, 0, sizeof d);
  double *hh_d = nullptr, *out10_d = nullptr;
  Carve count{nullptr};
  carve_rigid(count, d, hh_d, out10_d, N, p.natom_pad, max_neigh);
  if (count.used > kNoAllocScratchBytes)
    return set_error(FMD_ERR_CAPACITY,
                     "fmd_rigid_forces_noalloc: N=%d natom=%d max_neigh=%d needs %zu bytes of "
                     "scratch, more than its %zu",
                     N, natom, max_neigh, count.used, kNoAllocScratchBytes);
  unsigned char *base = nullptr;
  if ((rc = scratch_address(&base))) return rc;
  Carve carve{base};
  carve_rigid(carve, d, hh_d, out10_d, N, p.natom_pad, max_neigh);
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  hipStream_t s = 0;
  const size_t n = (size_t)N;
  HIPCHK(hipMemcpyAsync(hh_d, hh, sizeof hh, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.body, body4, 4 * (size_t)p.natom_pad * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.pos, pos, 3 * n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.qv, qv, 4 * n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.nl_count, nl_count, n * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(d.nl_list, nl_list, n * max_neigh * sizeof(int), hipMemcpyHostToDevice,
                        s));
  HIPCHK(rigid_launch_sites(d, p, s));
  HIPCHK(rigid_launch_forces(d, p, hh_d, s));
  HIPCHK(rigid_launch_reduce_molpart(d, p, out10_d, s));
  double out10[kMolPart];
  HIPCHK(hipMemcpyAsync(Fv, d.Fv, 3 * n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(Tv, d.Tv, 3 * n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(out10, out10_d, sizeof out10, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) Wm9[k] = out10[k];
  if (Ep) *Ep = out10[9];
  if (lab) host_lab(qv, body, N, natom, lab);
  return FMD_OK;
}

}  // extern "C"
