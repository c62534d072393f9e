// fmd_mm.hpp -- device layout of the molecular-mechanics force field
// (reference src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp, "MM"), the
// intramolecular term of an atomic engine whose AtomicParams::ff is kFfMM.
//
// Terms are kept per molecule template (local atom indices); molecules with
// identical terms share one template, so a crystal of one species reads one
// copy that stays in L2.  Unified term index u of a template:
//   [0, nb)                    bonds       1 force slot  (force on atom 0)
//   [nb, nb+na)                angles      3 slots (f_i, f_k, f_i + f_k)
//   [nb+na, nb+na+nd)          dihedrals   4 slots (atoms 0..3)
//   [nb+na+nd, nb+na+nd+ni)    impropers   4 slots (c1, c0, c2, c3)
// Slots are numbered in the same order (the reference's order of force
// accumulation, MM:378-710).  Each atom's incidence list names its slots in
// ascending order with a sign bit, so summing it reproduces the reference's
// per-atom sum sequence.  Chunks cut the term range (at term boundaries, at a
// multiple of the workgroup size where possible) so that one chunk's slots,
// the molecule's atoms and its incidence lists fit the workgroup's LDS.
#pragma once

#include <hip/hip_runtime.h>

namespace fmd {

enum : int { kFfAirebo = 0, kFfMM = 1 };
enum : int { kMmCompat = 0, kMmPhysical = 1 };
constexpr int kMmPart = 13;        // per molecule: W9, Eb, Ea, Ed, Ei
constexpr int kMmThreads = 256;    // workgroup of the bonded kernel
constexpr int kMmMaxAtoms = 128;   // atoms of one molecule (a thread quad per atom, 2 rounds)
constexpr double kCoulombK = 14.3996;  // eV A / e^2 (MM:117)
constexpr int kMmLds = 53248;      // LDS budget of one bonded workgroup (bytes): three
                                   // workgroups per CU (160 KB of LDS)

struct MmTmpl {
  int natom;
  int nb, na, nd, ni;  // term counts
  int u0;              // first term in MmDev::tq / tp / tm
  int inc0, ninc;      // its incidence entries: inc[inc0, inc0 + ninc)
  int cut0;            // its cuts: inc_cut[cut0 + c * natom + a], c = 0..nch
  int ch0, nch;        // its chunks: ch_u / ch_s [ch0, ch0 + nch]
};

struct MmDev {
  int ntmpl = 0;
  int slots = 0;         // largest chunk, in slots (LDS: 24 B each)
  int max_natom = 0;
  int max_inc = 0;       // most incidence entries of a template
  MmTmpl *tmpl = nullptr;
  int *mol_tmpl = nullptr;   // per molecule
  int4 *tq = nullptr;        // atoms of each term (local)
  double2 *tp = nullptr;     // (force constant, reference value)
  double2 *tg = nullptr;     // torsions: (cos gamma, sin gamma)
  int *tm = nullptr;         // torsion multiplicity
  int *inc = nullptr;        // slot << 1 | negate, atom-major per template
  int *inc_cut = nullptr;    // atom a's entries of chunk c: [cut[c][a], cut[c+1][a])
                             // (relative to the template's inc0)
  int *ch_u = nullptr;       // chunk boundaries in terms (template-relative)
  int *ch_s = nullptr;       // the same boundaries in slots
  double *part = nullptr;    // kMmPart x molecules (component-major)
  double *charge = nullptr;  // per atom, or null (no Coulomb term)
  int *atom_mol = nullptr;   // per atom (with charges)
  double *coul_part = nullptr;  // kAtPart x integrator blocks: W9, E
};

}  // namespace fmd
