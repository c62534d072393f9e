// fmd_engine.hpp -- the opaque fmd_engine of the C ABI and the helpers shared
// by the rigid and atomic halves of the host implementation.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>
#include <vector>

#include "../../include/fmd.h"
#include "fmd_atomic.hpp"
#include "fmd_rigid.hpp"

namespace fmd {

int set_error(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
LJ32 make_lj32(const fmd_lj_params &lj);

}  // namespace fmd

#define HIPCHK(expr)                                        \
  do {                                                      \
    hipError_t _e = (expr);                                 \
    if (_e != hipSuccess) return fmd::hip_fail(_e, #expr);  \
  } while (0)

struct fmd_engine {
  int kind = 0;
  hipStream_t stream = nullptr;
  int ring = 0;  // capacity of the per-step control / record rings
  int nlup = 25;
  double tt = 0;  // host mirror of the target temperature
  bool prepared = false;
  long long overflow = 0, rebuilds = 0, steps = 0;
  fmd::StepCtrl *h_ctrl = nullptr;
  fmd::Thermo *h_thermo = nullptr;
  // rigid
  fmd::RigidParams rp{};
  fmd::RigidDev rd{};
  // atomic
  fmd::AtomicEngine *at = nullptr;
  // timing of the step phases (HIP events on `stream`): kPhase* sums of the
  // durations (ms) and their counts since the last read
  bool timing = false;
  int t_stride = 1;            // events on every t_stride-th step slot (and on rebuilds)
  std::vector<hipEvent_t> ev;  // kEvPerStep per step of a run chunk
  double t_sum[4] = {0, 0, 0, 0};
  int t_n[4] = {0, 0, 0, 0};
  // hipGraph replay of step sequences: executable graphs keyed by the step
  // pattern they were captured from
  bool use_graph = false;
  std::vector<std::pair<std::string, hipGraphExec_t>> graphs;
};

namespace fmd {
// Timed phases: the force phase, its intramolecular part (REBO, surrogate or
// bonded terms; none for rigid molecules), its intermolecular LJ part, and a
// list rebuild.
enum : int { kPhaseForce = 0, kPhaseIntra = 1, kPhaseLJ = 2, kPhaseList = 3 };
constexpr int kEvPerStep = 5;  // StepEvents f0, fm, f1, l0, l1
int engine_init_common(fmd_engine *e, int ring, int R);
// Launch of one step into control / record slot `slot`; ev: the step's
// events, or null when not timed.  fold: the step of slot - 1 was launched
// with fin = 0 and this launch completes it; fin = 0: the next launch (slot
// + 1, no list rebuild) completes this step.  Engines that finish every
// step on its own ignore both.
using StepLaunch = std::function<hipError_t(int slot, int rebuild, int rescale, int fold,
                                            int fin, const StepEvents *ev)>;
// Steps first_step+1 .. first_step+nsteps of R replicas: controls
// ctrl[k * R + r] (null: hold each replica's target tt[r]), records
// thermo[k * R + r]; list rebuild every e->nlup global steps or on request.
// Replays captured graphs when e->use_graph.
int run_steps(fmd_engine *e, int R, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
              fmd_thermo *thermo, double *tt, StepCtrl *d_ctrl, Thermo *d_thermo,
              const StepLaunch &launch);
// The executable graph for `key`, captured from `launch` (which enqueues on
// e->stream) the first time the key is seen.
int engine_graph(fmd_engine *e, const std::string &key, const std::function<int()> &launch,
                 hipGraphExec_t *out);
void engine_drop_graphs(fmd_engine *e);
// Atomic (AIREBO) engine, fmd_atomic_api.hip.
int engine_prepare_atomic(fmd_engine *e);
int engine_run_atomic(fmd_engine *e, int first_step, int nsteps, const fmd_step_ctrl *ctrl,
                      fmd_thermo *thermo);
int engine_forces_atomic(fmd_engine *e, double *F, double *T, double *W9, double *energies);
int replica_forces(fmd_engine *e, int r, double *F, double *W9, double *energies);
// Learned surrogate (fmd_surrogate_api.hip).
void engine_free_surrogate(AtomicEngine &A);
bool surrogate_judging(fmd_engine *e, int *every);
int surrogate_judge(fmd_engine *e);
void engine_free_atomic(fmd_engine *e);
// Atomic engine creation with the intramolecular force field ff
// (kFfAirebo: fmd_batch_create; kFfMM: fmd_mm_create).
int atomic_create_ff(const fmd_atomic_config *c, int R, int ff, fmd_engine **out);
// Drop-in helpers on a single-system workspace engine (fmd_atomic_api.hip).
int atomic_image_mode(fmd_engine *e, const double *pos, const double *h);
int atomic_halflist(fmd_engine *e, const double *pos, const double *h, const double *hi,
                    int *nlc, int *nll);
int atomic_step_once(fmd_engine *e, double *vel, double *F, double *vir9, double *h,
                     fmd_npt_state *npt, AtomicState *Sout);
// Molecular-mechanics device buffers (fmd_mm_api.hip).
void engine_free_mm(AtomicEngine &A);
}  // namespace fmd
