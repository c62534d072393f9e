// fmd_atomic.hpp -- device data layout and launchers of the atomic (AIREBO)
// engine: REBO bond-order forces within molecules plus shifted LJ between
// molecules (reference src/fuller_airebo_npt_md_serial_omp_acc.cpp, "AB").
//
// HBM layout:
//   pos, vel, F, Frebo     double[Na*3]   reference AoS order
//   mass                   double[Na]
//   mol_start              int[Nm+1]      atoms of molecule m are the contiguous
//                                         range [mol_start[m], mol_start[m+1])
//   center                 double[Nm*3]   molecule centre (mean of the
//                                         unwrapped atoms), FP64
//   radius                 float[Nm]      max |atom - centre|
//   off                    float4[Nm*apad] atom offsets from the centre, one
//                                         padded tile per molecule (padding
//                                         sites at +1e6 A)
//   qcen                   float4[Nm*apad/4] groups of four tile slots:
//                                         centre offset and radius
//   jmax                   int[Nm*apad]   per atom: largest j kept in its
//                                         400-entry half list (INT_MAX if
//                                         the list is not truncated)
//   jmax_min               int[Nm]
//   mol_nl_count / mol_nl  molecule neighbour list (ascending)
//   nlc_r / nll_r          REBO list, reference layout (stride maxr = 12)
//   al_e / al_n            per-atom LJ lists (engines): for tile slot a of
//                          molecule i, 16-bit entries k << 8 | b (atom slot b
//                          of the k-th neighbour in mol_nl) of the pairs
//                          within rc + 1 A that the reference's truncated half
//                          list keeps, in four segments (one per build
//                          wavefront) of blocks of eight (16 bytes, padded
//                          with a never-near entry): block c of segment w of
//                          slot a at ((i * 4 + w) * al_cap / 8 + c) * NA + a
//                          (NA = 64 NS); al_n counts blocks per segment; the
//                          walk splits each atom's blocks evenly over its
//                          four wavefronts
//   bonds, rebo_fe         per directed REBO entry (fmd_rebo_passes.hpp), used
//                          for molecules whose entries do not fit the LDS
//   rebo_mpart, lj_part    fixed-order partial sums per molecule, and
//   rebo_part, ke_part     per integrator block (rebo_part: the surrogate's;
//                          ke_part, kKePart components: kinetic energy,
//                          then sums over the block's share of its replica's
//                          molecules of the trace-only partials: AIREBO the
//                          LJ + REBO trace, the LJ energy and the REBO
//                          energy; MM the LJ trace, the LJ energy and the 13
//                          bonded components), component-major (component k
//                          of record r at k * count + r)
//
// Replicas: R independent systems share every array, concatenated replica
// after replica (atoms, molecules).  Molecule and atom ranges of replica r
// are [rep_mol[r], rep_mol[r+1]) and [rep_atom[r], rep_atom[r+1]); mol_rep /
// atom_rep map back.  Per-atom integrator blocks never straddle replicas:
// block b covers atoms [blk_a0[b], blk_a1[b]) of replica blk_rep[b], and
// replica r owns blocks [rep_blk[r], rep_blk[r+1]).  state: one per
// replica; mid: two buffers of R (steps alternate between them); ctrl /
// thermo rings: slot-major, R entries per step slot.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "fmd_cells.hpp"
#include "fmd_common.hpp"
#include "fmd_mm.hpp"
#include "fmd_rebo_passes.hpp"
#include "fmd_rigid.hpp"  // StepCtrl, Thermo
#include "fmd_sg_types.hpp"

namespace fmd {

constexpr int kAtPart = 10;  // W9 + E
constexpr int kKePart = 16;  // AtomicDev::ke_part components per block
// Steps fold the previous step's finish into their kick-drift only up to
// this many integrator blocks: every block repeats the finish, which pays
// while the whole grid is resident at once (the 8x8x8 C60 crystal: 480
// blocks) and costs more than the separate finish once blocks queue up
// (64 replicas, 7,100 blocks: 2.210 -> 2.241 ms per step).  Either way the
// results are the same bit for bit.
constexpr int kFoldMaxBlocks = 1024;
constexpr int kAtFlags = 5;  // AtomicDev::flags
constexpr int kAlDriftSlots = 8;  // AtomicDev::al_dmax entries per replica
constexpr float kAlDriftFloor = 0.125f;  // drifts below it count as it (A)
// image modes of a replica at a list rebuild (AtomicDev::img)
enum : int {
  kImgMol = 1,   // atom pairs within rc + 1 A are the centres' image + offsets
  kImgScan = 2,  // the same within the half-list radius: FP32 prefiltered walk
};

struct AtomicState {
  double h[9], hi[9];  // adjacent: kernels take them as one 18-double block
  double xi, Vg[9], Q, W, Pe, Tt;
  double KE, Ep, Ep_rebo, Ep_lj, W9[9];
  double Eb, Ea, Ed, Ei, Ec;  // molecular-mechanics terms (Ep_rebo = Eb+Ea+Ed+Ei)
  double scale;
  int Nf, step;
};

struct AtomicMid {
  double h[9], hi[9];
  double xi, Vg[9];
  double sc_nh;
};

struct AtomicParams {
  int Na, Nm;      // totals over all replicas
  int R;           // replicas
  int apad;        // atoms per molecule tile (max molecule size, rounded up to 4)
  int maxr;        // REBO list stride (12, AB:126)
  int mol_cap;     // molecule list stride
  int lj_cap;      // half-list cap of the reference (400, AB:127)
  int lj_compat;   // 1: pairs limited by the reference's truncated half list
  int rebo_mode;   // kReboCompat / kReboPhysical
  int small_box;   // 1: per-pair minimum image (a cell narrower than the list)
  int use_sg;      // 1: the learned surrogate stands in for the REBO term
  int ff;          // kFfAirebo / kFfMM: the intramolecular force field
  int mm_mode;     // kMmCompat / kMmPhysical (ff == kFfMM)
  int trace_virial;  // 1: time steps keep only the LJ virial trace (engines)
  int lj_trace;      // set per launch: this force evaluation keeps only the
                     // virial trace (LJ and REBO)
  int lj_alist;    // 1: LJ from per-atom lists (engines; cells with one image
                   // per molecule pair); 0: every tile pair each step (drop-ins)
  int al_cap;      // per-atom list capacity per wavefront segment (entries)
  int rebo_slots;  // REBO kernel: bond entries per atom in LDS
  int rebo_nt;     // REBO kernel: workgroup size (192 used for apad <= 192)
  // REBO-II / AIREBO terms the reference does not have (exact-gradient REBO
  // only; zero: off): the P_ij spline of the bond order, and the torsion
  // term's well depth (eV, k_rebo_tors)
  rebo::ReboPcc pcc;
  double tors_eps;
  int cell_total;  // molecule-list cells over all replicas
  int mol_cells;   // 1: some replica builds its molecule list from cells
  int cell_min;    // the molecules a replica needs for that
  float al_r2;     // per-atom list radius^2: (rc + 1 A)^2
  int scan_fast;   // 1: list walk prefiltered in FP32 through the molecule
                   // tiles (cell wide enough for the list radius + 2 R_max)
  double rebo_rlist2;  // REBO list radius^2 (2.3^2, AB:253)
  double lj_rlist2;    // LJ half-list radius^2 ((rc + 2)^2, AB:265)
  double mol_rlist;    // molecule list radius beyond the two molecule radii
  double dt, clamp_xi, clamp_vg;
  double coul_rc;      // Coulomb cutoff (MM: the LJ cutoff, MM:119)
  LJ32 lj;
  double eps;
};

struct AtomicDev {
  double *pos, *vel, *F, *Frebo, *mass;
  int *mol_start;
  double *center;
  float *radius;
  float4 *off;
  float4 *qcen;  // Nm*apad/4: centre offset and radius of each 4-slot group
  int *jmax, *jmax_min;
  int *mol_nl_count, *mol_nl;
  int *nlc_r, *nll_r;
  ReboBonds bonds;
  double *rebo_part, *rebo_mpart, *lj_part, *ke_part;
  double *rebo_fe;  // 3 x Na*maxr: per-entry REBO forces (global-memory REBO path)
  unsigned short *al_e;  // per-atom LJ lists: [Nm][4][al_cap / 8][NA] blocks of 8 entries
  unsigned short *al_n;  // segment lengths in blocks: [Nm][4][NA]  (NA = 64 * NS)
  // per-atom list staleness: at the last list build, each atom's fractional
  // position (al_pos, w = 1 once set) and each replica's cell (al_h);
  // al_dmax[kAlDriftSlots r + k], the largest atom drift |h (s - s0)| since
  // then over replica r's integrator blocks b = k mod kAlDriftSlots (float
  // bits, drifts below kAlDriftFloor not recorded; reset by the image check
  // that opens every scheduled rebuild, and by k_at_kick2 after a rebuild
  // within a step)
  float4 *al_pos;
  double *al_h;
  int *al_dmax;
  int *al_redo;  // per replica: lists to rebuild within this step (k_at_centers)
  int *img;  // per replica, decided at every list rebuild: kImgMol | kImgScan
  // molecule-list cells, per replica at every list rebuild: the grid (cells
  // per cell vector, w = 1 when the molecules were binned, 0 for a scan of
  // every molecule), the first of the replica's cells (cell_base[R] = all),
  // each cell's molecule count and ids (kCellCap per cell)
  int4 *cell_dim;
  int *cell_base, *cell_n, *cell_m;
  double *hh;  // 18 doubles for the stateless entry points
  AtomicState *state;
  AtomicMid *mid;
  StepCtrl *ctrl;
  Thermo *thermo;
  int *flags;  // [0] REBO list overflow, [1] molecule list overflow,
               // [2] a step whose per-atom lists had gone stale (the LJ
               // kernel rebuilt them from that step's positions), [3] image
               // fault: the cell became too narrow for per-molecule images
               // while the engine stepped with them, [4] per-atom list
               // overflow (the engine then turns to the tile kernel)
  int *mol_rep, *atom_rep, *rep_mol, *rep_atom;
  int *blk_a0, *blk_a1, *blk_rep, *rep_blk;
  int nblk;    // per-atom integrator blocks (<= 256 atoms of one replica each)
  SgWeights sgw;  // learned surrogate (allocated when attached)
  SgDev sgd;
  MmDev mm;       // molecular-mechanics terms (ff == kFfMM)
};

// h and h^-1 of each replica: 18 doubles at p + r * stride (stride 0: one
// cell for all, the stateless entry points).
struct HHRef {
  const double *p;
  int stride;
};
inline HHRef hh_state(const AtomicDev &d) {
  static_assert(sizeof(AtomicState) % 8 == 0, "AtomicState stride");
  return {reinterpret_cast<const double *>(reinterpret_cast<const char *>(d.state) +
                                           offsetof(AtomicState, h)),
          (int)(sizeof(AtomicState) / 8)};
}
// the cells of midpoint buffer par (d.mid holds two of R replicas each)
inline HHRef hh_mid(const AtomicDev &d, int par, int R) {
  static_assert(sizeof(AtomicMid) % 8 == 0, "AtomicMid stride");
  const AtomicMid *m = d.mid + (size_t)par * R;
  return {reinterpret_cast<const double *>(reinterpret_cast<const char *>(m) +
                                           offsetof(AtomicMid, h)),
          (int)(sizeof(AtomicMid) / 8)};
}
inline HHRef hh_fixed(const double *hh) { return {hh, 0}; }

// Launchers (atomic_kernels.hip).  hh: the cells (HHRef).
hipError_t at_launch_apply_pbc(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                               hipStream_t s);
// Molecule centres, radii and offset tiles from pos.
hipError_t at_launch_centers(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                             hipStream_t s);
// Molecule list, REBO list and (compat) truncation points; needs centres.
hipError_t at_launch_lists(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                           int rebo_list, hipStream_t s);
// REBO + LJ forces into F (and the partial sums); needs centres and lists.
// ev_mid (may be null) is recorded between the intramolecular part (REBO,
// surrogate or bonded terms) and the intermolecular LJ, with ev_flags.
hipError_t at_launch_forces(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                            hipStream_t s, hipEvent_t ev_mid = nullptr,
                            unsigned ev_flags = 0);
enum : int { kAtObsKE = 1, kAtObsForce = 2 };
hipError_t at_launch_observe(const AtomicDev &d, const AtomicParams &p, int what,
                             hipStream_t s);
// One NPT step of every replica; controls from ctrl[ctrl_slot * R + r],
// records to thermo[thermo_slot * R + r].  rescale: launch the velocity
// rescale (it acts on the replicas whose control asks for it).
// ev (may be null): events around the force phase, its LJ part and the list
// rebuild.  prev_slot >= 0: the step of slot prev_slot (= ctrl_slot - 1, the
// step just before, no list rebuild in this one) was launched with fin = 0
// and its finish (thermostat, barostat, record, rescale) is folded into this
// step's kick-drift; fin = 0 leaves this step's finish to the next step.
// The results are bitwise those of finishing every step on its own.
hipError_t at_launch_step(const AtomicDev &d, const AtomicParams &p, int ctrl_slot,
                          int thermo_slot, int rebuild, int rescale, const StepEvents *ev,
                          hipStream_t s, int prev_slot = -1, int fin = 1);
// REBO forces alone into Frebo and rebo_mpart, in the given mode (needs
// centres and the REBO list).
hipError_t at_launch_rebo(const AtomicDev &d, const AtomicParams &p, HHRef hh, int mode,
                          hipStream_t s);
// Molecular-mechanics bonded forces into Frebo and mm.part (mm_kernels.hip).
hipError_t mm_launch_bonded(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                            hipStream_t s);
// Coulomb forces added into F, partials into mm.coul_part (needs charges).
hipError_t mm_launch_coulomb(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                             hipStream_t s);
// jmax from a reference-layout half list (stride cap) for the drop-in path.
hipError_t at_launch_jmax_from_list(const AtomicDev &d, const AtomicParams &p,
                                    const int *nlc_l, const int *nll_l, int cap,
                                    hipStream_t s);
// Reference-layout half list (AB:263-272) built on the GPU.
hipError_t at_launch_halflist(const AtomicDev &d, const AtomicParams &p, HHRef hh,
                              int *nlc_l, int *nll_l, int cap, hipStream_t s);

struct Surrogate;  // fmd_surrogate_api.hip
struct MmHost;     // fmd_mm_api.hip

struct AtomicEngine {
  AtomicParams p{};
  AtomicDev d{};
  Surrogate *sg = nullptr;
  MmHost *mmh = nullptr;  // host image of d.mm (molecular mechanics)
  // host copies of the replica tables
  std::vector<int> rep_atom, rep_mol, mol_start;
  std::vector<double> tt;  // per replica: target temperature (host mirror)
  std::vector<int> small;  // per replica: cell narrower than the LJ reach
  std::vector<int> fast;   // per replica: FP32-prefiltered list walk allowed
};

}  // namespace fmd
