// fmd_surrogate_api.hip -- host side of the learned REBO surrogate: parameters,
// training frames generated by the engine itself (positions + exact-gradient
// REBO forces), GPU training (Adam over the force-matching loss), evaluation,
// and the judge that re-runs full REBO every K steps and falls back to it
// when the force RMSE exceeds a threshold.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "fmd_engine.hpp"
#include "fmd_surrogate.hpp"

namespace fmd {

struct SgFrame {
  double *pos = nullptr, *F = nullptr, *hh = nullptr;  // device: Na x 3, Na x 3, 18
};

struct Surrogate {
  float *m = nullptr, *v = nullptr, *grad = nullptr, *partials = nullptr;
  int nwg = 0;
  double *Fm = nullptr;        // model forces (training, judge)
  double *part = nullptr;      // per-block squared errors
  double *loss = nullptr;      // per-iteration MSE (training)
  int loss_cap = 0;
  double *pos_save = nullptr;  // engine positions while frames are loaded
  std::vector<SgFrame> frames;
  int t = 0;                   // Adam step
  int enabled = 0, judge_every = 100;
  double threshold = 0.05, last_rmse = -1.0;
  int judged = 0, fallbacks = 0;
};

namespace {

__global__ void k_sum_parts(const double *part, int n, double scale, double *out) {
  __shared__ double red[256];
  double acc = 0;
  for (int k = threadIdx.x; k < n; k += 256) acc += part[k];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0] * scale;
}

template <class T>
hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}

AtomicEngine *atomic_of(fmd_engine *e) {
  if (!e || e->kind != FMD_KIND_ATOMIC || !e->at) return nullptr;
  return e->at;
}

int need_sg(fmd_engine *e, AtomicEngine **A) {
  *A = atomic_of(e);
  if (!*A) return set_error(FMD_ERR_ARG, "surrogate: not an atomic engine");
  if (!(*A)->sg) return set_error(FMD_ERR_STATE, "surrogate: call fmd_sg_attach first");
  return FMD_OK;
}

int need_frame(AtomicEngine *A, int frame) {
  if (frame < 0 || frame >= (int)A->sg->frames.size())
    return set_error(FMD_ERR_ARG, "surrogate: frame %d of %d", frame, (int)A->sg->frames.size());
  return FMD_OK;
}

// Load a stored frame's positions into the engine (engine positions saved
// first) and compute the molecule tiles for its cell.
int load_frame(fmd_engine *e, int frame) {
  AtomicEngine &A = *e->at;
  Surrogate &S = *A.sg;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)A.p.Na;
  HIPCHK(hipMemcpyAsync(A.d.pos, S.frames[frame].pos, n3 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(S.frames[frame].hh), s));
  return FMD_OK;
}

int save_pos(fmd_engine *e) {
  AtomicEngine &A = *e->at;
  HIPCHK(hipMemcpyAsync(A.sg->pos_save, A.d.pos, 3 * (size_t)A.p.Na * 8,
                        hipMemcpyDeviceToDevice, e->stream));
  return FMD_OK;
}

int restore_pos(fmd_engine *e) {
  AtomicEngine &A = *e->at;
  HIPCHK(hipMemcpyAsync(A.d.pos, A.sg->pos_save, 3 * (size_t)A.p.Na * 8,
                        hipMemcpyDeviceToDevice, e->stream));
  HIPCHK(hipMemsetAsync(A.d.al_pos, 0, (size_t)A.p.Na * sizeof(float4), e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  e->prepared = false;  // tiles, lists and forces are recomputed before the next step
  return FMD_OK;
}

// Model forces of the loaded tiles into out (descriptor, MLP, forces).
int model_forces(fmd_engine *e, double *out) {
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  HIPCHK(sg_launch_descr(A.d, A.p, A.d.sgw, A.d.sgd, s));
  HIPCHK(sg_launch_mlp(A.p, A.d.sgw, A.d.sgd, s));
  HIPCHK(sg_launch_force(A.d, A.p, A.d.sgw, A.d.sgd, out, s));
  return FMD_OK;
}

// One loss + gradient evaluation on the loaded frame; MSE into *loss_dev.
int loss_grad_pass(fmd_engine *e, int frame, double *loss_dev) {
  AtomicEngine &A = *e->at;
  Surrogate &S = *A.sg;
  hipStream_t s = e->stream;
  const double *Fref = S.frames[frame].F;
  int rc = model_forces(e, S.Fm);
  if (rc) return rc;
  HIPCHK(sg_launch_sqerr(A.d, S.Fm, Fref, S.part, s));
  k_sum_parts<<<1, 256, 0, s>>>(S.part, A.d.nblk, 1.0 / (3.0 * A.p.Na), loss_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(sg_launch_u(A.d, A.p, A.d.sgw, A.d.sgd, S.Fm, Fref, s));
  HIPCHK(sg_launch_train(A.p, A.d.sgw, A.d.sgd, S.partials, s));
  HIPCHK(sg_launch_reduce_grad(S.partials, S.nwg, (float)(-2.0 / (3.0 * A.p.Na)), S.grad, s));
  return FMD_OK;
}

void free_sg(AtomicEngine &A) {
  if (!A.sg) return;
  Surrogate &S = *A.sg;
  void *ptrs[] = {A.d.sgw.w1, A.d.sgw.w1t, A.d.sgw.w2, A.d.sgw.w2t, A.d.sgw.w3, A.d.sgw.w3t,
                  A.d.sgw.theta, A.d.sgw.mu, A.d.sgw.sd, A.d.sgd.feat, A.d.sgd.mom, A.d.sgd.nbc,
                  A.d.sgd.nbl, A.d.sgd.g, A.d.sgd.ei, A.d.sgd.vir, A.d.sgd.u, A.d.sgd.F,
                  A.d.sgd.flags, S.m, S.v, S.grad, S.partials, S.Fm, S.part, S.loss, S.pos_save};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &f : S.frames) {
    (void)hipFree(f.pos);
    (void)hipFree(f.F);
    (void)hipFree(f.hh);
  }
  memset(&A.d.sgw, 0, sizeof A.d.sgw);
  memset(&A.d.sgd, 0, sizeof A.d.sgd);
  A.p.use_sg = 0;
  delete A.sg;
  A.sg = nullptr;
}

}  // namespace

void engine_free_surrogate(AtomicEngine &A) { free_sg(A); }

// Full REBO (exact gradient) at the current positions against the surrogate
// forces of the step just taken; fall back to REBO above the threshold.
int surrogate_judge(fmd_engine *e) {
  AtomicEngine &A = *e->at;
  Surrogate &S = *A.sg;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)A.p.Na;
  HIPCHK(hipMemcpyAsync(S.Fm, A.d.Frebo, n3 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(at_launch_centers(A.d, A.p, hh_state(A.d), s));
  HIPCHK(at_launch_rebo(A.d, A.p, hh_state(A.d), kReboPhysical, s));
  HIPCHK(sg_launch_sqerr(A.d, S.Fm, A.d.Frebo, S.part, s));
  double mse = 0;
  k_sum_parts<<<1, 256, 0, s>>>(S.part, A.d.nblk, 1.0 / (3.0 * A.p.Na), S.loss);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(&mse, S.loss, sizeof mse, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  S.last_rmse = std::sqrt(mse);
  S.judged++;
  if (S.last_rmse > S.threshold) {
    // fallback: REBO from here on, forces and observables recomputed with it
    S.fallbacks++;
    S.enabled = 0;
    A.p.use_sg = 0;
    engine_drop_graphs(e);
    HIPCHK(at_launch_forces(A.d, A.p, hh_state(A.d), s));
    HIPCHK(at_launch_observe(A.d, A.p, kAtObsForce, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  return FMD_OK;
}

bool surrogate_judging(fmd_engine *e, int *every) {
  AtomicEngine *A = e->at;
  if (!A || !A->sg || !A->sg->enabled || A->sg->judge_every <= 0) return false;
  *every = A->sg->judge_every;
  return true;
}

}  // namespace fmd

using namespace fmd;

extern "C" {

int fmd_sg_param_count(void) { return kSgNParam; }
int fmd_sg_feature_count(void) { return kSgNF; }

int fmd_sg_attach(fmd_engine *e, int seed) {
  AtomicEngine *A = atomic_of(e);
  if (!A) return set_error(FMD_ERR_ARG, "fmd_sg_attach: not an atomic engine");
  if (A->p.ff != kFfAirebo)
    return set_error(FMD_ERR_ARG, "fmd_sg_attach: the surrogate stands in for REBO only");
  if (A->sg) return FMD_OK;
  A->sg = new Surrogate();
  Surrogate &S = *A->sg;
  SgWeights &w = A->d.sgw;
  SgDev &g = A->d.sgd;
  const size_t na = A->p.Na;
  S.nwg = sg_train_groups(A->p.Na);
  hipError_t he = hipSuccess;
#define AL(ptr, n) \
  if (he == hipSuccess) he = dalloc(&(ptr), (n))
  AL(w.w1, kSgKF * kSgH);
  AL(w.w1t, kSgKF * kSgH);
  AL(w.w2, kSgH * kSgH);
  AL(w.w2t, kSgH * kSgH);
  AL(w.w3, kSgH * kSgH);
  AL(w.w3t, kSgH * kSgH);
  AL(w.theta, kSgNParam);
  AL(w.mu, kSgNF);
  AL(w.sd, kSgNF);
  AL(g.feat, na * kSgKF);
  AL(g.mom, na * kSgMom);
  AL(g.nbc, na);
  AL(g.nbl, na * kSgNbMax);
  AL(g.g, na * kSgNF);
  AL(g.ei, na);
  AL(g.vir, na * 9);
  AL(g.u, na * kSgNF);
  AL(g.F, 3 * na);
  AL(g.flags, 4);
  AL(S.m, kSgNParam);
  AL(S.v, kSgNParam);
  AL(S.grad, kSgNParam);
  AL(S.partials, (size_t)S.nwg * kSgNParam);
  AL(S.Fm, 3 * na);
  AL(S.part, (size_t)A->d.nblk);
  AL(S.loss, 1);
  AL(S.pos_save, 3 * na);
#undef AL
  if (he != hipSuccess) {
    free_sg(*A);
    return hip_fail(he, "hipMalloc (surrogate)");
  }
  S.loss_cap = 1;
  // Glorot-uniform weights, zero biases, identity normalisation
  std::vector<float> th(kSgNParam, 0.f), mu(kSgNF, 0.f), sd(kSgNF, 1.f);
  std::mt19937 rng((unsigned)seed);
  std::uniform_real_distribution<float> U(-1.f, 1.f);
  auto glorot = [&](int off, int a, int b) {
    const float sc = std::sqrt(6.0f / (a + b));
    for (int k = 0; k < a * b; k++) th[off + k] = U(rng) * sc;
  };
  glorot(kSgOffW1, kSgNF, kSgH);
  glorot(kSgOffW2, kSgH, kSgH);
  glorot(kSgOffW3, kSgH, kSgH);
  glorot(kSgOffW4, kSgH, 1);
  int rc = fmd_sg_set_params(e, th.data(), mu.data(), sd.data());
  if (rc) return rc;
  HIPCHK(hipMemset(g.flags, 0, 4 * sizeof(int)));
  return FMD_OK;
}

int fmd_sg_set_params(fmd_engine *e, const float *theta, const float *mu, const float *sd) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  hipStream_t s = e->stream;
  SgWeights &w = A->d.sgw;
  if (theta) HIPCHK(hipMemcpyAsync(w.theta, theta, kSgNParam * 4, hipMemcpyHostToDevice, s));
  if (mu) HIPCHK(hipMemcpyAsync(w.mu, mu, kSgNF * 4, hipMemcpyHostToDevice, s));
  if (sd) HIPCHK(hipMemcpyAsync(w.sd, sd, kSgNF * 4, hipMemcpyHostToDevice, s));
  HIPCHK(sg_launch_pack(w, s));
  HIPCHK(hipMemsetAsync(A->sg->m, 0, kSgNParam * 4, s));
  HIPCHK(hipMemsetAsync(A->sg->v, 0, kSgNParam * 4, s));
  HIPCHK(hipStreamSynchronize(s));
  A->sg->t = 0;
  e->prepared = false;
  return FMD_OK;
}

int fmd_sg_get_params(fmd_engine *e, float *theta, float *mu, float *sd) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  hipStream_t s = e->stream;
  const SgWeights &w = A->d.sgw;
  if (theta) HIPCHK(hipMemcpyAsync(theta, w.theta, kSgNParam * 4, hipMemcpyDeviceToHost, s));
  if (mu) HIPCHK(hipMemcpyAsync(mu, w.mu, kSgNF * 4, hipMemcpyDeviceToHost, s));
  if (sd) HIPCHK(hipMemcpyAsync(sd, w.sd, kSgNF * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FMD_OK;
}

int fmd_sg_collect(fmd_engine *e) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  if (A->p.R != 1) return set_error(FMD_ERR_ARG, "fmd_sg_collect: single-system engines only");
  if (!e->prepared && (rc = engine_prepare_atomic(e))) return rc;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)A->p.Na;
  SgFrame f;
  HIPCHK(dalloc(&f.pos, n3));
  HIPCHK(dalloc(&f.F, n3));
  HIPCHK(dalloc(&f.hh, 18));
  // exact-gradient REBO forces at the current positions (lists rebuilt)
  HIPCHK(at_launch_centers(A->d, A->p, hh_state(A->d), s));
  HIPCHK(at_launch_lists(A->d, A->p, hh_state(A->d), 1, s));
  HIPCHK(at_launch_rebo(A->d, A->p, hh_state(A->d), kReboPhysical, s));
  HIPCHK(hipMemcpyAsync(f.pos, A->d.pos, n3 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(f.F, A->d.Frebo, n3 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(f.hh, hh_state(A->d).p, 18 * 8, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  A->sg->frames.push_back(f);
  e->prepared = false;  // Frebo / partials were used as scratch
  return FMD_OK;
}

int fmd_sg_frame_count(fmd_engine *e, int *n) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  *n = (int)A->sg->frames.size();
  return FMD_OK;
}

int fmd_sg_fit_normalisation(fmd_engine *e) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  Surrogate &S = *A->sg;
  if (S.frames.empty()) return set_error(FMD_ERR_STATE, "fmd_sg_fit_normalisation: no frames");
  hipStream_t s = e->stream;
  const size_t na = A->p.Na;
  std::vector<float> mu0(kSgNF, 0.f), sd0(kSgNF, 1.f);
  HIPCHK(hipMemcpyAsync(A->d.sgw.mu, mu0.data(), kSgNF * 4, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A->d.sgw.sd, sd0.data(), kSgNF * 4, hipMemcpyHostToDevice, s));
  if ((rc = save_pos(e))) return rc;
  std::vector<double> sum(kSgNF, 0.0), sum2(kSgNF, 0.0);
  std::vector<unsigned short> feat(na * kSgKF);
  double cnt = 0;
  for (size_t f = 0; f < S.frames.size(); f++) {
    if ((rc = load_frame(e, (int)f))) return rc;
    HIPCHK(sg_launch_descr(A->d, A->p, A->d.sgw, A->d.sgd, s));
    HIPCHK(hipMemcpyAsync(feat.data(), A->d.sgd.feat, feat.size() * 2, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (size_t i = 0; i < na; i++)
      for (int k = 0; k < kSgNF; k++) {
        unsigned int u = (unsigned int)feat[i * kSgKF + k] << 16;
        float x;
        memcpy(&x, &u, 4);
        sum[k] += x;
        sum2[k] += (double)x * x;
      }
    cnt += na;
  }
  std::vector<float> mu(kSgNF), sd(kSgNF);
  for (int k = 0; k < kSgNF; k++) {
    mu[k] = (float)(sum[k] / cnt);
    const double var = std::max(sum2[k] / cnt - (sum[k] / cnt) * (sum[k] / cnt), 0.0);
    sd[k] = (float)std::max(std::sqrt(var), 1e-3 * std::max(1.0, std::fabs(sum[k] / cnt)));
  }
  if ((rc = restore_pos(e))) return rc;
  return fmd_sg_set_params(e, nullptr, mu.data(), sd.data());
}

int fmd_sg_train(fmd_engine *e, int iters, double lr, double *loss) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  Surrogate &S = *A->sg;
  if (S.frames.empty()) return set_error(FMD_ERR_STATE, "fmd_sg_train: no frames");
  if (iters < 1) return FMD_OK;
  hipStream_t s = e->stream;
  if (S.loss_cap < iters) {
    HIPCHK(hipFree(S.loss));
    HIPCHK(dalloc(&S.loss, iters));
    S.loss_cap = iters;
  }
  if ((rc = save_pos(e))) return rc;
  const int nf = (int)S.frames.size();
  for (int it = 0; it < iters; it++) {
    const int f = it % nf;
    if ((rc = load_frame(e, f))) return rc;
    if ((rc = loss_grad_pass(e, f, S.loss + it))) return rc;
    HIPCHK(sg_launch_adam(A->d.sgw, S.grad, S.m, S.v, ++S.t, (float)lr, s));
    HIPCHK(sg_launch_pack(A->d.sgw, s));
  }
  if (loss) HIPCHK(hipMemcpyAsync(loss, S.loss, iters * sizeof(double), hipMemcpyDeviceToHost, s));
  return restore_pos(e);
}

int fmd_sg_loss_grad(fmd_engine *e, int frame, double *loss, float *grad) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc || (rc = need_frame(A, frame))) return rc;
  hipStream_t s = e->stream;
  if ((rc = save_pos(e)) || (rc = load_frame(e, frame)) ||
      (rc = loss_grad_pass(e, frame, A->sg->loss)))
    return rc;
  if (loss) HIPCHK(hipMemcpyAsync(loss, A->sg->loss, sizeof(double), hipMemcpyDeviceToHost, s));
  if (grad) HIPCHK(hipMemcpyAsync(grad, A->sg->grad, kSgNParam * 4, hipMemcpyDeviceToHost, s));
  return restore_pos(e);
}

int fmd_sg_descriptor(fmd_engine *e, int frame, float *feat) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc || (rc = need_frame(A, frame))) return rc;
  hipStream_t s = e->stream;
  const size_t na = A->p.Na;
  std::vector<unsigned short> raw(na * kSgKF);
  if ((rc = save_pos(e)) || (rc = load_frame(e, frame))) return rc;
  HIPCHK(sg_launch_descr(A->d, A->p, A->d.sgw, A->d.sgd, s));
  HIPCHK(hipMemcpyAsync(raw.data(), A->d.sgd.feat, raw.size() * 2, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t i = 0; i < na; i++)
    for (int k = 0; k < kSgNF; k++) {
      unsigned int u = (unsigned int)raw[i * kSgKF + k] << 16;
      memcpy(&feat[i * kSgNF + k], &u, 4);
    }
  return restore_pos(e);
}

int fmd_sg_evaluate(fmd_engine *e, int frame, double *F, double *W9, double *E, double *rmse) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  Surrogate &S = *A->sg;
  hipStream_t s = e->stream;
  const size_t na = A->p.Na;
  if ((rc = save_pos(e))) return rc;
  if (frame >= 0) {
    if ((rc = need_frame(A, frame)) || (rc = load_frame(e, frame))) return rc;
  } else {
    HIPCHK(at_launch_centers(A->d, A->p, hh_state(A->d), s));
  }
  if ((rc = model_forces(e, S.Fm))) return rc;
  double mse = -1.0;
  if (frame >= 0) {
    HIPCHK(sg_launch_sqerr(A->d, S.Fm, S.frames[frame].F, S.part, s));
    k_sum_parts<<<1, 256, 0, s>>>(S.part, A->d.nblk, 1.0 / (3.0 * na), S.loss);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&mse, S.loss, sizeof mse, hipMemcpyDeviceToHost, s));
  }
  std::vector<float> vir(na * 9), ei(na);
  if (F) HIPCHK(hipMemcpyAsync(F, S.Fm, 3 * na * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(vir.data(), A->d.sgd.vir, vir.size() * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(ei.data(), A->d.sgd.ei, ei.size() * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (W9)
    for (int k = 0; k < 9; k++) {
      double acc = 0;
      for (size_t i = 0; i < na; i++) acc += vir[9 * i + k];
      W9[k] = acc;
    }
  if (E) {
    double acc = 0;
    for (size_t i = 0; i < na; i++) acc += ei[i];
    *E = acc;
  }
  if (rmse) *rmse = mse >= 0 ? std::sqrt(mse) : -1.0;
  return restore_pos(e);
}

int fmd_sg_enable(fmd_engine *e, int on, int judge_every, double threshold) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  if (on && A->p.rebo_mode != FMD_REBO_PHYSICAL)
    return set_error(FMD_ERR_ARG,
                     "fmd_sg_enable: the surrogate stands in for the exact-gradient REBO "
                     "(FMD_REBO_PHYSICAL engines)");
  Surrogate &S = *A->sg;
  S.enabled = on != 0;
  S.judge_every = judge_every;
  S.threshold = threshold;
  A->p.use_sg = S.enabled;
  engine_drop_graphs(e);
  e->prepared = false;
  return FMD_OK;
}

int fmd_sg_stats(fmd_engine *e, double *last_rmse, int *judged, int *fallbacks, int *active) {
  AtomicEngine *A;
  int rc = need_sg(e, &A);
  if (rc) return rc;
  const Surrogate &S = *A->sg;
  if (last_rmse) *last_rmse = S.last_rmse;
  if (judged) *judged = S.judged;
  if (fallbacks) *fallbacks = S.fallbacks;
  if (active) *active = A->p.use_sg;
  return FMD_OK;
}

}  // extern "C"
