// fmd_mm_api.hip -- host side of the molecular-mechanics force field
// (src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp, "MM"): per-molecule term
// templates built from a flat topology, the resident MM engine (an atomic
// engine with ff = kFfMM) and the drop-in entry points with the signatures of
// the reference's MM functions.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "fmd_engine.hpp"

namespace fmd {

// Host image of MmDev (fmd_mm.hpp).
struct MmHost {
  std::vector<MmTmpl> tmpl;
  std::vector<int> mol_tmpl;
  std::vector<int4> tq;
  std::vector<double2> tp, tg;
  std::vector<int> tm, inc, inc_cut, ch_u, ch_s;
  int slots = 0, max_natom = 0, max_inc = 0;
  std::vector<double> charge;  // empty: no Coulomb term
  std::vector<int> atom_mol;
  double *owned_charge = nullptr;  // device buffers behind MmDev::charge / atom_mol
  int *owned_atom_mol = nullptr;
};

}  // namespace fmd

using namespace fmd;

namespace {

struct Term {
  int4 q;
  double2 p;
  int mult;
};

bool term_eq(const Term &a, const Term &b) {
  return a.q.x == b.q.x && a.q.y == b.q.y && a.q.z == b.q.z && a.q.w == b.q.w &&
         std::memcmp(&a.p, &b.p, sizeof a.p) == 0 && a.mult == b.mult;
}

uint64_t fnv(uint64_t h, const void *p, size_t n) {
  const unsigned char *c = static_cast<const unsigned char *>(p);
  for (size_t i = 0; i < n; i++) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

// One molecule's terms by kind: bonds, angles, dihedrals, impropers.
using MolTerms = std::array<std::vector<Term>, 4>;

bool same_terms(const MolTerms &a, const MolTerms &b) {
  for (int k = 0; k < 4; k++) {
    if (a[k].size() != b[k].size()) return false;
    for (size_t t = 0; t < a[k].size(); t++)
      if (!term_eq(a[k][t], b[k][t])) return false;
  }
  return true;
}

// What chunking needs of a template beyond MmTmpl.
struct TmplBuild {
  std::vector<int> nslot;       // slots of each unified term
  std::vector<int> atom_start;  // natom + 1 starts of the atoms' entries (relative)
};

// Appends a new template for terms T of a natom-atom molecule: its terms and
// incidence lists (slot << 1 | negate, ascending per atom); chunks later.
void add_template(MmHost &H, const MolTerms &T, int natom, std::vector<TmplBuild> &tb) {
  MmTmpl t{};
  t.natom = natom;
  t.nb = (int)T[0].size();
  t.na = (int)T[1].size();
  t.nd = (int)T[2].size();
  t.ni = (int)T[3].size();
  t.u0 = (int)H.tq.size();
  TmplBuild b;
  std::vector<std::vector<int>> lists(natom);
  int slot = 0;
  for (int k = 0; k < 4; k++) {
    for (const Term &x : T[k]) {
      H.tq.push_back(x.q);
      H.tp.push_back(x.p);
      H.tg.push_back(k >= 2 ? make_double2(std::cos(x.p.y), std::sin(x.p.y))
                            : make_double2(0.0, 0.0));
      H.tm.push_back(x.mult);
      if (k == 0) {  // force on q.x, its negative on q.y
        lists[x.q.x].push_back(slot << 1);
        lists[x.q.y].push_back(slot << 1 | 1);
        b.nslot.push_back(1);
      } else if (k == 1) {  // f_i, f_k, -(f_i + f_k) on the centre
        lists[x.q.x].push_back(slot << 1);
        lists[x.q.z].push_back((slot + 1) << 1);
        lists[x.q.y].push_back((slot + 2) << 1 | 1);
        b.nslot.push_back(3);
      } else {  // four forces in atom order
        lists[x.q.x].push_back(slot << 1);
        lists[x.q.y].push_back((slot + 1) << 1);
        lists[x.q.z].push_back((slot + 2) << 1);
        lists[x.q.w].push_back((slot + 3) << 1);
        b.nslot.push_back(4);
      }
      slot += b.nslot.back();
    }
  }
  t.inc0 = (int)H.inc.size();
  for (int a = 0; a < natom; a++) {
    b.atom_start.push_back((int)H.inc.size() - t.inc0);
    H.inc.insert(H.inc.end(), lists[a].begin(), lists[a].end());
  }
  t.ninc = (int)H.inc.size() - t.inc0;
  b.atom_start.push_back(t.ninc);
  H.max_natom = std::max(H.max_natom, natom);
  H.max_inc = std::max(H.max_inc, t.ninc);
  H.tmpl.push_back(t);
  tb.push_back(std::move(b));
}

// Chunks of every template: at most cap slots each, cut at a multiple of
// kMmThreads terms from the chunk's start where that is possible (so that no
// pass over the terms runs with a nearly empty workgroup), and the per-atom
// incidence cuts at the chunk boundaries.
void chunk_templates(MmHost &H, const std::vector<TmplBuild> &tb, int cap) {
  H.slots = 0;
  for (size_t ti = 0; ti < H.tmpl.size(); ti++) {
    MmTmpl &t = H.tmpl[ti];
    const std::vector<int> &nslot = tb[ti].nslot;
    const int nu = (int)nslot.size();
    t.ch0 = (int)H.ch_u.size();
    H.ch_u.push_back(0);
    H.ch_s.push_back(0);
    int u = 0, s = 0, nch = 0;
    while (u < nu) {
      int u1 = u, s1 = s;
      while (u1 < nu && s1 + nslot[u1] - s <= cap) s1 += nslot[u1++];
      if (u1 < nu && u1 - u >= kMmThreads) {
        u1 = u + (u1 - u) / kMmThreads * kMmThreads;
        s1 = s;
        for (int x = u; x < u1; x++) s1 += nslot[x];
      }
      H.slots = std::max(H.slots, s1 - s);
      u = u1;
      s = s1;
      H.ch_u.push_back(u);
      H.ch_s.push_back(s);
      nch++;
    }
    t.nch = nch;
    // cut[c][a]: first entry of atom a whose slot is at or past chunk c's
    // first slot (c = nch: the end of atom a's entries)
    t.cut0 = (int)H.inc_cut.size();
    for (int c = 0; c <= nch; c++) {
      const int sc = H.ch_s[t.ch0 + c];
      for (int a = 0; a < t.natom; a++) {
        int e = tb[ti].atom_start[a];
        const int e1 = tb[ti].atom_start[a + 1];
        while (e < e1 && (H.inc[t.inc0 + e] >> 1) < sc) e++;
        H.inc_cut.push_back(c == nch ? e1 : e);
      }
    }
  }
}

// Templates of every molecule from the flat topology: each term goes to the
// molecule of its first atom (all its atoms must belong to it), keeping the
// terms' order; molecules with identical terms share a template.
int mm_build_host(const fmd_mm_topology &ft, const std::vector<int> &mol_start, int Na,
                  MmHost &H) {
  H = MmHost();
  const int Nm = (int)mol_start.size() - 1;
  std::vector<int> mol_of(Na);
  for (int m = 0; m < Nm; m++) {
    if (mol_start[m + 1] - mol_start[m] > kMmMaxAtoms)
      return set_error(FMD_ERR_CAPACITY, "MM: molecules of more than %d atoms", kMmMaxAtoms);
    for (int a = mol_start[m]; a < mol_start[m + 1]; a++) mol_of[a] = m;
  }
  std::vector<MolTerms> per(Nm);
  const int counts[4] = {ft.Nb, ft.Nang, ft.Ndih, ft.Nimp};
  for (int k = 0; k < 4; k++) {
    if (counts[k] < 0) return set_error(FMD_ERR_ARG, "MM: negative term count");
    for (int t = 0; t < counts[k]; t++) {
      Term x{};
      int n = 0;
      int at[4] = {0, 0, 0, 0};
      if (k == 0) {
        if (!ft.b_i0 || !ft.b_i1 || !ft.b_kb || !ft.b_r0) return set_error(FMD_ERR_ARG, "MM: null bond array");
        at[0] = ft.b_i0[t];
        at[1] = ft.b_i1[t];
        n = 2;
        x.p = make_double2(ft.b_kb[t], ft.b_r0[t]);
      } else if (k == 1) {
        if (!ft.ang_i0 || !ft.ang_i1 || !ft.ang_i2 || !ft.ang_kth || !ft.ang_th0)
          return set_error(FMD_ERR_ARG, "MM: null angle array");
        at[0] = ft.ang_i0[t];
        at[1] = ft.ang_i1[t];
        at[2] = ft.ang_i2[t];
        n = 3;
        x.p = make_double2(ft.ang_kth[t], ft.ang_th0[t]);
      } else if (k == 2) {
        if (!ft.dih_i0 || !ft.dih_i1 || !ft.dih_i2 || !ft.dih_i3 || !ft.dih_Vn || !ft.dih_mult ||
            !ft.dih_gamma)
          return set_error(FMD_ERR_ARG, "MM: null dihedral array");
        at[0] = ft.dih_i0[t];
        at[1] = ft.dih_i1[t];
        at[2] = ft.dih_i2[t];
        at[3] = ft.dih_i3[t];
        n = 4;
        x.p = make_double2(ft.dih_Vn[t], ft.dih_gamma[t]);
        x.mult = ft.dih_mult[t];
      } else {
        if (!ft.imp_i0 || !ft.imp_i1 || !ft.imp_i2 || !ft.imp_i3 || !ft.imp_ki || !ft.imp_gamma)
          return set_error(FMD_ERR_ARG, "MM: null improper array");
        // the improper's torsion runs c1-c0-c2-c3 (MM:612-615, 634)
        at[0] = ft.imp_i1[t];
        at[1] = ft.imp_i0[t];
        at[2] = ft.imp_i2[t];
        at[3] = ft.imp_i3[t];
        n = 4;
        x.p = make_double2(ft.imp_ki[t], ft.imp_gamma[t]);
        x.mult = 2;  // MM:628-629
      }
      for (int q = 0; q < n; q++)
        if (at[q] < 0 || at[q] >= Na)
          return set_error(FMD_ERR_ARG, "MM: term %d of kind %d names atom %d of %d", t, k, at[q],
                           Na);
      const int m = mol_of[at[0]];
      for (int q = 1; q < n; q++)
        if (mol_of[at[q]] != m)
          return set_error(FMD_ERR_ARG, "MM: term %d of kind %d spans two molecules", t, k);
      const int s0 = mol_start[m];
      x.q = make_int4(at[0] - s0, at[1] - s0, at[2] - s0, at[3] - s0);
      if (n < 4) x.q.w = 0;
      if (n < 3) x.q.z = 0;
      per[m][k].push_back(x);
    }
  }
  std::unordered_map<uint64_t, std::vector<int>> seen;  // hash -> templates
  std::vector<int> first_of_tmpl;
  std::vector<TmplBuild> tb;
  H.mol_tmpl.assign(Nm, -1);
  for (int m = 0; m < Nm; m++) {
    const int natom = mol_start[m + 1] - mol_start[m];
    uint64_t h = fnv(1469598103934665603ull, &natom, sizeof natom);
    for (int k = 0; k < 4; k++) {
      const size_t n = per[m][k].size();
      h = fnv(h, &n, sizeof n);
      for (const Term &x : per[m][k]) {
        h = fnv(h, &x.q, sizeof x.q);
        h = fnv(h, &x.p, sizeof x.p);
        h = fnv(h, &x.mult, sizeof x.mult);
      }
    }
    int found = -1;
    for (int cand : seen[h]) {
      const int m2 = first_of_tmpl[cand];
      if (mol_start[m2 + 1] - mol_start[m2] == natom && same_terms(per[m2], per[m])) {
        found = cand;
        break;
      }
    }
    if (found < 0) {
      found = (int)H.tmpl.size();
      add_template(H, per[m], natom, tb);
      first_of_tmpl.push_back(m);
      seen[h].push_back(found);
    }
    H.mol_tmpl[m] = found;
  }
  // LDS of one workgroup: atoms (24 B each), incidence entries (4 B), slots
  const int cap = (kMmLds - 24 * H.max_natom - 4 * H.max_inc) / 24;
  if (cap < 64) return set_error(FMD_ERR_CAPACITY, "MM: molecule topology too large for LDS");
  chunk_templates(H, tb, cap);
  return FMD_OK;
}

template <class T>
hipError_t upload(T **dst, const std::vector<T> &v) {
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  hipError_t e = hipMalloc((void **)dst, std::max<size_t>(v.size(), 1) * sizeof(T));
  if (e == hipSuccess && !v.empty())
    e = hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

bool same_image(const MmHost &a, const MmHost &b) {
  auto eqv = [](const auto &x, const auto &y) {
    return x.size() == y.size() &&
           (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(x[0])) == 0);
  };
  return eqv(a.tmpl, b.tmpl) && eqv(a.mol_tmpl, b.mol_tmpl) && eqv(a.tq, b.tq) &&
         eqv(a.tp, b.tp) && eqv(a.tg, b.tg) && eqv(a.tm, b.tm) && eqv(a.inc, b.inc) &&
         eqv(a.inc_cut, b.inc_cut) && eqv(a.ch_u, b.ch_u) && eqv(a.ch_s, b.ch_s);
}

// The templates of topology ft onto the device (skipped when unchanged).
int mm_set_topology(AtomicEngine &A, const fmd_mm_topology &ft) {
  MmHost H;
  int rc = mm_build_host(ft, A.mol_start, A.p.Na, H);
  if (rc) return rc;
  if (!A.mmh) A.mmh = new MmHost();
  MmHost &cur = *A.mmh;
  if (A.d.mm.tmpl && same_image(cur, H)) return FMD_OK;
  MmDev &m = A.d.mm;
  hipError_t e = upload(&m.tmpl, H.tmpl);
  if (e == hipSuccess) e = upload(&m.mol_tmpl, H.mol_tmpl);
  if (e == hipSuccess) e = upload(&m.tq, H.tq);
  if (e == hipSuccess) e = upload(&m.tp, H.tp);
  if (e == hipSuccess) e = upload(&m.tg, H.tg);
  if (e == hipSuccess) e = upload(&m.tm, H.tm);
  if (e == hipSuccess) e = upload(&m.inc, H.inc);
  if (e == hipSuccess) e = upload(&m.inc_cut, H.inc_cut);
  if (e == hipSuccess) e = upload(&m.ch_u, H.ch_u);
  if (e == hipSuccess) e = upload(&m.ch_s, H.ch_s);
  if (e == hipSuccess && !m.part)
    e = hipMalloc((void **)&m.part, (size_t)A.p.Nm * kMmPart * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "upload MM templates");
  m.ntmpl = (int)H.tmpl.size();
  m.slots = H.slots;
  m.max_natom = H.max_natom;
  m.max_inc = H.max_inc;
  H.charge = std::move(cur.charge);
  H.atom_mol = std::move(cur.atom_mol);
  H.owned_charge = cur.owned_charge;
  H.owned_atom_mol = cur.owned_atom_mol;
  cur = std::move(H);
  return FMD_OK;
}

// Charges for the Coulomb term (MM:766-787); none when every |q| <= 1e-10,
// as the reference decides (MM:1111).
int mm_set_charges(AtomicEngine &A, const double *charge) {
  MmDev &m = A.d.mm;
  bool any = false;
  for (int i = 0; charge && i < A.p.Na && !any; i++) any = std::fabs(charge[i]) > 1e-10;
  if (!any) {
    m.charge = nullptr;  // the buffers stay allocated for a later call
    return FMD_OK;
  }
  if (!A.mmh) A.mmh = new MmHost();
  MmHost &H = *A.mmh;
  H.charge.assign(charge, charge + A.p.Na);
  H.atom_mol.resize(A.p.Na);
  for (size_t mo = 0; mo + 1 < A.mol_start.size(); mo++)
    for (int a = A.mol_start[mo]; a < A.mol_start[mo + 1]; a++) H.atom_mol[a] = (int)mo;
  double *q = nullptr;
  int *am = nullptr;
  hipError_t e = hipMalloc((void **)&q, A.p.Na * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&am, A.p.Na * sizeof(int));
  if (e == hipSuccess)
    e = hipMemcpy(q, H.charge.data(), A.p.Na * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(am, H.atom_mol.data(), A.p.Na * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && !m.coul_part)
    e = hipMalloc((void **)&m.coul_part, (size_t)A.d.nblk * kAtPart * sizeof(double));
  if (e != hipSuccess) {
    if (q) (void)hipFree(q);
    if (am) (void)hipFree(am);
    return hip_fail(e, "upload MM charges");
  }
  if (H.owned_charge) (void)hipFree(H.owned_charge);
  if (H.owned_atom_mol) (void)hipFree(H.owned_atom_mol);
  H.owned_charge = q;
  H.owned_atom_mol = am;
  m.charge = q;
  m.atom_mol = am;
  return FMD_OK;
}

}  // namespace

namespace fmd {

void engine_free_mm(AtomicEngine &A) {
  MmDev &m = A.d.mm;
  void *ptrs[] = {m.tmpl, m.mol_tmpl, m.tq, m.tp, m.tg, m.tm, m.inc, m.inc_cut, m.ch_u, m.ch_s,
                  m.part, m.coul_part};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (A.mmh) {
    if (A.mmh->owned_charge) (void)hipFree(A.mmh->owned_charge);
    if (A.mmh->owned_atom_mol) (void)hipFree(A.mmh->owned_atom_mol);
    delete A.mmh;
    A.mmh = nullptr;
  }
  m = MmDev();
}

}  // namespace fmd

namespace {

// The atomic-engine configuration of an MM configuration (carbon masses).
fmd_atomic_config atomic_config_of(const fmd_mm_config &c, std::vector<double> &mass) {
  fmd_atomic_config a;
  fmd_atomic_config_default(&a);
  mass.assign(c.n_atoms, kMassC);
  a.n_atoms = c.n_atoms;
  a.mol_id = c.mol_id;
  a.mass = mass.data();
  a.lj = c.lj;
  a.lj_skin = c.lj_skin;
  a.lj_max_neigh = c.lj_max_neigh;
  a.nlup = c.nlup;
  a.dt = c.dt;
  a.clamp_xi = c.clamp_xi;
  a.clamp_vg = c.clamp_vg;
  return a;
}

AtomicEngine *mm_of(fmd_engine *e) {
  return e && e->kind == FMD_KIND_ATOMIC && e->at && e->at->p.ff == kFfMM ? e->at : nullptr;
}

// ------------------------------------------------------------ drop-ins
constexpr int kRefMaxLj = 400;  // MM:119

struct MmWorkspace {
  fmd_engine *e = nullptr;
  std::vector<int> mol_id;
};
MmWorkspace g_mws;

// Workspace MM engine for this molecule layout (re-created when it changes).
int mm_workspace(int na, const int *mol_id, int mode, fmd_engine **out) {
  if (!mol_id || na < 1) return set_error(FMD_ERR_ARG, "MM drop-in: need mol_id and Na >= 1");
  if (mode != FMD_MM_COMPAT && mode != FMD_MM_PHYSICAL)
    return set_error(FMD_ERR_ARG, "MM drop-in: mm_mode must be FMD_MM_COMPAT or FMD_MM_PHYSICAL");
  bool same = g_mws.e && (int)g_mws.mol_id.size() == na &&
              std::equal(g_mws.mol_id.begin(), g_mws.mol_id.end(), mol_id);
  if (!same) {
    if (g_mws.e) fmd_engine_destroy(g_mws.e);
    g_mws = MmWorkspace();
    fmd_mm_config c;
    fmd_mm_config_default(&c);
    c.n_atoms = na;
    c.mol_id = mol_id;
    std::vector<double> mass;
    fmd_atomic_config a = atomic_config_of(c, mass);
    int rc = atomic_create_ff(&a, 1, kFfMM, &g_mws.e);
    if (rc) return rc;
    g_mws.mol_id.assign(mol_id, mol_id + na);
  }
  g_mws.e->at->p.mm_mode = mode;
  g_mws.e->at->p.trace_virial = 0;  // the drop-ins return the full virial
  g_mws.e->at->p.lj_alist = 0;      // and evaluate the given lists' pairs each call
  *out = g_mws.e;
  return FMD_OK;
}

// Positions, cell, topology, charges and the caller's LJ half list (through
// its truncation points) onto the workspace; molecule list from pos.
int mm_stage(fmd_engine *e, const double *pos, const double *h, const double *hi,
             const fmd_mm_topology *ft, const int *nlc, const int *nll, const double *charge,
             int has_charge) {
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  int rc = mm_set_topology(A, *ft);
  if (rc || (rc = mm_set_charges(A, has_charge ? charge : nullptr))) return rc;
  double hh[18];
  for (int k = 0; k < 9; k++) {
    hh[k] = h[k];
    hh[9 + k] = hi[k];
  }
  HIPCHK(hipMemcpyAsync(A.d.pos, pos, 3 * (size_t)A.p.Na * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(A.d.hh, hh, sizeof hh, hipMemcpyHostToDevice, s));
  if ((rc = atomic_image_mode(e, pos, h))) return rc;
  const int lj_compat = A.p.lj_compat;
  A.p.lj_compat = 0;  // the molecule list only; jmax comes from the caller's list
  HIPCHK(at_launch_centers(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_lists(A.d, A.p, hh_fixed(A.d.hh), 0, s));
  A.p.lj_compat = lj_compat;
  const size_t na = A.p.Na;
  int *dl = nullptr, *dc = nullptr;
  HIPCHK(hipMallocAsync((void **)&dl, na * kRefMaxLj * sizeof(int), s));
  HIPCHK(hipMallocAsync((void **)&dc, na * sizeof(int), s));
  HIPCHK(hipMemcpyAsync(dc, nlc, na * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(dl, nll, na * kRefMaxLj * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(at_launch_jmax_from_list(A.d, A.p, dc, dl, kRefMaxLj, s));
  HIPCHK(hipFreeAsync(dl, s));
  HIPCHK(hipFreeAsync(dc, s));
  return FMD_OK;
}

void energies7(const AtomicState &S, double *out) {
  out[0] = S.Eb;
  out[1] = S.Ea;
  out[2] = S.Ed;
  out[3] = S.Ei;
  out[4] = S.Ep_lj;
  out[5] = S.Ec;
  out[6] = S.Ep;
}

}  // namespace

extern "C" {

void fmd_mm_config_default(fmd_mm_config *c) {
  memset(c, 0, sizeof *c);
  fmd_lj_params_mm(&c->lj);
  c->lj_skin = 2.0;
  c->lj_max_neigh = 400;
  c->nlup = 20;
  c->dt = 0.1;
  c->clamp_xi = 0.05;
  c->clamp_vg = 0.005;
  c->mm_mode = FMD_MM_COMPAT;
}

int fmd_mm_create(const fmd_mm_config *c, fmd_engine **out) {
  if (!c || !out || !c->ft || !c->mol_id || c->n_atoms < 1)
    return set_error(FMD_ERR_ARG, "fmd_mm_create: bad argument");
  if (c->mm_mode != FMD_MM_COMPAT && c->mm_mode != FMD_MM_PHYSICAL)
    return set_error(FMD_ERR_ARG, "fmd_mm_create: mm_mode must be FMD_MM_COMPAT or FMD_MM_PHYSICAL");
  std::vector<double> mass;
  fmd_atomic_config a = atomic_config_of(*c, mass);
  fmd_engine *e = nullptr;
  int rc = atomic_create_ff(&a, 1, kFfMM, &e);
  if (rc) return rc;
  AtomicEngine &A = *e->at;
  A.p.mm_mode = c->mm_mode;
  rc = mm_set_topology(A, *c->ft);
  if (rc == FMD_OK) rc = mm_set_charges(A, c->charge);
  if (rc) {
    fmd_engine_destroy(e);
    return rc;
  }
  *out = e;
  return FMD_OK;
}

int fmd_mm_energies(fmd_engine *e, double *out7) {
  AtomicEngine *A = mm_of(e);
  if (!A || !out7) return set_error(FMD_ERR_ARG, "fmd_mm_energies: not an MM engine");
  AtomicState S;
  HIPCHK(hipMemcpyAsync(&S, A->d.state, sizeof S, hipMemcpyDeviceToHost, e->stream));
  HIPCHK(hipStreamSynchronize(e->stream));
  energies7(S, out7);
  return FMD_OK;
}

int fmd_mm_nlist_lj(const double *pos, const double *h, const double *hi, int Na,
                    const int *mol_id, int *nlc, int *nll) {
  if (!pos || !h || !hi || !nlc || !nll) return set_error(FMD_ERR_ARG, "mm nlist_lj: null argument");
  fmd_engine *e;
  int rc = mm_workspace(Na, mol_id, FMD_MM_COMPAT, &e);
  if (rc) return rc;
  return atomic_halflist(e, pos, h, hi, nlc, nll);
}

int fmd_mm_compute_forces(double *F, double *vir9, const double *pos, const double *h,
                          const double *hi, const fmd_mm_topology *ft, const int *nlc,
                          const int *nll, const double *charge, const int *mol_id,
                          int has_charge, int Na, int mm_mode, double *out7) {
  if (!F || !vir9 || !pos || !h || !hi || !ft || !nlc || !nll || (has_charge && !charge))
    return set_error(FMD_ERR_ARG, "mm compute_forces: null argument");
  fmd_engine *e;
  int rc = mm_workspace(Na, mol_id, mm_mode, &e);
  if (rc || (rc = mm_stage(e, pos, h, hi, ft, nlc, nll, charge, has_charge))) return rc;
  AtomicEngine &A = *e->at;
  hipStream_t s = e->stream;
  HIPCHK(at_launch_forces(A.d, A.p, hh_fixed(A.d.hh), s));
  HIPCHK(at_launch_observe(A.d, A.p, kAtObsForce, s));
  AtomicState S;
  HIPCHK(hipMemcpyAsync(F, A.d.F, 3 * (size_t)Na * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(&S, A.d.state, sizeof S, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) vir9[k] = S.W9[k];
  if (out7) energies7(S, out7);
  return FMD_OK;
}

int fmd_mm_step_npt(double *pos, double *vel, double *F, double *vir9, double *h,
                    double *hi, int Na, double dt, fmd_npt_state *npt,
                    const fmd_mm_topology *ft, const int *nlc, const int *nll,
                    const double *charge, const int *mol_id, int has_charge, int mm_mode,
                    double *out8) {
  if (!pos || !vel || !F || !vir9 || !h || !hi || !npt || !ft || !nlc || !nll ||
      (has_charge && !charge))
    return set_error(FMD_ERR_ARG, "mm step_npt: null argument");
  fmd_engine *e;
  int rc = mm_workspace(Na, mol_id, mm_mode, &e);
  if (rc || (rc = mm_stage(e, pos, h, hi, ft, nlc, nll, charge, has_charge))) return rc;
  AtomicEngine &A = *e->at;
  A.p.dt = dt;
  AtomicState S;
  if ((rc = atomic_step_once(e, vel, F, vir9, h, npt, &S))) return rc;
  hipStream_t s = e->stream;
  const size_t n3 = 3 * (size_t)Na * 8;
  HIPCHK(hipMemcpyAsync(pos, A.d.pos, n3, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < 9; k++) hi[k] = S.hi[k];
  if (out8) {
    energies7(S, out8);
    out8[7] = S.KE;
  }
  return FMD_OK;
}

}  // extern "C"
