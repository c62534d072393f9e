// fmd_rigid.hpp -- device data layout and launchers of the rigid-body engine
// (site-site LJ between rigid molecules; reference
// src/fuller_LJ_npt_md_core_serial.cpp "C1" and
// src/fuller_LJ_npt_md_serial_omp_acc.cpp "LJ").
//
// HBM layout (all per-molecule arrays in the reference's flat AoS order):
//   pos, vel, omg, Fv, Tv   double[N*3]
//   qv                      double[N*4]           [w,x,y,z]
//   sites                   float4[N*natom_pad]   lab-frame site offsets
//                                                 R(q_i) body_a, FP32; padding
//                                                 sites sit at +1e6 A
//   nl_count / nl_list      int[N] / int[N*max_neigh]  symmetric full list
//   cell_dim / cell_n / cell_m  int4 / int[C] / int[C*kCellCap]  the list's
//                           cell grid (C = max(27, N) cells) when N >= cell_min
//   hl_* / in_*             int, same shapes          owned pairs / incoming slots
//   pairout                 double[N*max_neigh*16]    per owned pair, both sides
//   hl_item                 int2[max_neigh*N]         pair kernel work items, rank-major
//   molpart                 double[N*kMolPart]    per-molecule W9, E partials
//   blockpart               double[kBlkPart*nblk] per-block KE / W / E sums,
//                           component-major
//   RigidState              NPT + cell, FP64 (one per system)
//   RigidMid                first-half results of the current step
#pragma once

#include <hip/hip_runtime.h>

#include "fmd_cells.hpp"
#include "fmd_common.hpp"

namespace fmd {

constexpr int kMolPart = 10;  // W9 + E
constexpr int kBlkPart = 12;  // sum v^2, sum w^2, W9, E
// Per owned molecule pair (slot owner * max_neigh + position in its half
// list): [0..2] F_i [3..5] tau_i [6] energy [7] sum d.f  (owner side, written
// by the pair kernel) and [8..10] F_j [11..13] tau_j (the other molecule).
constexpr int kPairOut = 16;

// Per-step control word (host-decided phase logic of LJ:848-894).
enum : int {
  kCtrlZeroXiVg = 1,      // xi = 0, Vg = 0 before the step (first WARM step)
  kCtrlZeroVg = 2,        // Vg = 0 before the step (COLD phase)
  kCtrlRescale = 4,       // velocity rescale to rescale_T after the step
  kCtrlRescaleZeroVg = 8, // ... and Vg = 0 (COLD)
  kCtrlRebuild = 16,      // rebuild the molecule list before the step
};

struct StepCtrl {
  double Tt;         // target temperature of this step
  double rescale_T;  // target of the velocity rescale
  int flags;
  int pad;
};

struct RigidState {
  double h[9], hi[9];  // kept adjacent: kernels take them as one 18-double block
  double xi, Vg[9], Q, W, Pe, Tt;
  double KEt, KEr, Ep, W9[9];  // observables of the last completed step
  double scale;                // last velocity rescale factor (1 if none)
  int Nf, step;
};

struct RigidMid {
  double h[9], hi[9];
  double xi, Vg[9];
  double sc_nh;
};

// One monitor record per step, as LJ:874-876 / 896 compute them.
struct Thermo {
  double Ep, KE, KEt, KEr, T, P, V, xi, Ep_rebo, Ep_lj;
  double h[9];
  double Eb, Ea, Ed, Ei, Ecoul;  // molecular-mechanics terms (0 otherwise)
};

struct RigidParams {
  int N, natom, natom_pad, max_neigh;
  int nq;             // site quads per molecule (natom_pad / 4)
  int use_half;       // time stepping with the half-pair kernel (trace virial)
  int pair_kernel;    // ... as one wavefront per owned pair (natom <= 64)
  double rmcut2;      // molecule-pair cutoff^2 inside the force loop
  double rlist2;      // list radius^2 (rmcut + skin)^2
  double dt, I0, Mmol;
  double clamp_xi, clamp_vg;
  double cull_r2;     // (rc + max site radius)^2: per-lane neighbour cull
  float reach;        // rc + max site radius + FP32 margin: quad cull radius
  LJ32 lj;
  double eps;         // LJ epsilon (energy scale of the reduced sums)
  // the pair kernel's tests in units of sigma: 1/sigma, the cutoff band and
  // the r^2 clamp over sigma^2, and 48 eps / sigma
  float isig, rc2_lo_s, rc2_hi_s, rmin2_s;
  double c48s;
  int cell_min;       // lists from a cell grid for N >= cell_min (0: never)
};

// Device buffers of one rigid system.
struct RigidDev {
  double *pos, *vel, *omg, *qv, *Fv, *Tv;
  float4 *sites;
  double *sites_d;  // N*natom_pad*3: the same offsets in FP64, bitwise the
                    // reference's lab array, for cutoff re-tests
  double *body;  // natom_pad x 4 doubles: x, y, z, padding flag (1 = padding)
  double *qbody; // nq x 4 doubles: body-frame quad centre and radius
  float4 *qcen;  // N*nq: lab-frame quad centre offsets (w = quad radius)
  double *hh;    // 18 doubles: h, h^-1 for the stateless entry points
  int *nl_count, *nl_list, *flags;  // flags[0]: list overflow count, [1]: list
                                    // faults (a pair could have closed the skin)
  // the list's cell grid (cells per axis, w = 1 while usable), each cell's
  // molecule count and ids; allocated when N >= cell_min
  int4 *cell_dim;
  int *cell_n, *cell_m;
  // molecule centres and cell (h, then 1 once set) at the last list build
  double *nl_pos, *nl_h;
  // pairs owned by each molecule (half list) and, per molecule, the slots of
  // the pairs owned by its neighbours that carry its force / torque
  int *hl_count, *hl_list, *in_count, *in_list;
  double *pairout;  // N*max_neigh*kPairOut: per owned pair, both sides
  // pair kernel work items: hl_item[r * N + i] = (slot, j) of molecule i's
  // r-th nearest owned pair, (-1, -1) past its count; ord_ctl[0] = the
  // largest owned-pair count
  int2 *hl_item;
  int *ord_ctl;
  double *molpart, *blockpart;
  RigidState *state;
  RigidMid *mid;
  StepCtrl *ctrl;     // ring of per-step controls
  Thermo *thermo;     // ring of per-step records
  int nblk;           // blocks of k_rigid_kick2 (4 molecules each): partials in blockpart
};

// Launchers (rigid_kernels.hip).  All enqueue on `s` and return the HIP
// status of the launch.
// `hh` points at 18 device doubles (h then h^-1).
hipError_t rigid_launch_sites(const RigidDev &d, const RigidParams &p, hipStream_t s);
hipError_t rigid_launch_apply_pbc(const RigidDev &d, const RigidParams &p, const double *hh,
                                  hipStream_t s);
hipError_t rigid_launch_nlist(const RigidDev &d, const RigidParams &p, const double *hh,
                              hipStream_t s);
// The engine's rebuild: checks the list being replaced (flags[1]), builds the
// new one and records the state it was built from.
hipError_t rigid_launch_rebuild(const RigidDev &d, const RigidParams &p, const double *hh,
                                hipStream_t s);
// Half list and incoming slots from the symmetric list; with the pair kernel
// also its work order (centre distances from the cell `hh`).
hipError_t rigid_launch_halflist(const RigidDev &d, const RigidParams &p, const double *hh,
                                 hipStream_t s);
hipError_t rigid_launch_forces(const RigidDev &d, const RigidParams &p, const double *hh,
                               hipStream_t s);
// Observables of the current state into RigidState: what = kObsKE | kObsForce.
enum : int { kObsKE = 1, kObsForce = 2 };
hipError_t rigid_launch_observe(const RigidDev &d, const RigidParams &p, int what,
                                hipStream_t s);
// One NPT step (LJ:847-894 loop body without I/O): optional list rebuild,
// kick-drift + sites, forces, kick + reductions + NPT second half, optional
// velocity rescale.  Controls come from ctrl[ctrl_slot]; the record goes to
// thermo[thermo_slot].
// ev (may be null): events around the force kernel (f0 and fm before it, f1
// after) and the list rebuild.
hipError_t rigid_launch_step(const RigidDev &d, const RigidParams &p, int ctrl_slot,
                             int thermo_slot, int rebuild, int rescale, const StepEvents *ev,
                             hipStream_t s);
// Sum of molpart (W9, E) into out10 on the device.
hipError_t rigid_launch_reduce_molpart(const RigidDev &d, const RigidParams &p, double *out10,
                                       hipStream_t s);

}  // namespace fmd
