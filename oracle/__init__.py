"""FP64 CPU oracle for the fuller_md force / integrator path (test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package.  It is the checker the HIP kernels are
compared against, never the thing measured or shipped.

``oracle.orc``  ctypes bindings of ``liboracle.so`` (our C restatement,
                 ``fmd_oracle.c``), with numpy-friendly wrappers below.
``oracle.ref``  bindings of the reference build in ``oracle/_ref`` (only where
                 it has been built; see ``oracle/Makefile``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

CONV = 9.64853321e-3
KB = 8.617333262e-5
EV2GPA = 160.21766208

REBO_COMPAT = 0
REBO_PHYSICAL = 1
MM_COMPAT = 0
MM_PHYSICAL = 1
KCAL2EV = 1.0 / 23.06054783
MM_K4_DEFAULT = (469.0 * KCAL2EV, 63.0 * KCAL2EV, 14.5 * KCAL2EV, 15.0 * KCAL2EV)
MM_RLIST = 3.0 * 3.431 + 2.0


class LJParams(C.Structure):
    _fields_ = [("sig2", C.c_double), ("eps", C.c_double), ("rc2", C.c_double),
                ("vshift", C.c_double)]


class NPT(C.Structure):
    """orc_npt == NPTState (C1:478, AB:511)."""
    _fields_ = [("xi", C.c_double), ("Q", C.c_double), ("Vg", C.c_double * 9),
                ("W", C.c_double), ("Pe", C.c_double), ("Tt", C.c_double),
                ("Nf", C.c_int)]

    def to_array(self) -> np.ndarray:
        """Flat 15-double layout [xi, Q, Vg0..8, W, Pe, Tt, Nf]."""
        return np.array([self.xi, self.Q, *self.Vg, self.W, self.Pe, self.Tt,
                         float(self.Nf)])

    @classmethod
    def from_array(cls, a) -> "NPT":
        s = cls()
        s.xi, s.Q = a[0], a[1]
        for i in range(9):
            s.Vg[i] = a[2 + i]
        s.W, s.Pe, s.Tt, s.Nf = a[11], a[12], a[13], int(a[14])
        return s


class MMTopo(C.Structure):
    """orc_mm_topo: the flat topology arrays of build_flat_topology (MM:228-239)."""
    _fields_ = [("Nb", C.c_int), ("Nang", C.c_int), ("Ndih", C.c_int), ("Nimp", C.c_int),
                ("b", C.c_void_p), ("bp", C.c_void_p), ("a", C.c_void_p),
                ("ap", C.c_void_p), ("d", C.c_void_p), ("dp", C.c_void_p),
                ("dm", C.c_void_p), ("im", C.c_void_p), ("ip", C.c_void_p)]


def build(force: bool = False) -> str:
    """Compile liboracle.so (and, when the reference tree is present, _ref)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "liboracle"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d = C.c_double
_i = C.c_int


def _load():
    lib = C.CDLL(build())
    sig = {
        "orc_mat_det9": (_d, [_dp]),
        "orc_mat_inv9": (None, [_dp, _dp]),
        "orc_mimg": (None, [_dp, _dp, _dp]),
        "orc_apply_pbc": (None, [_dp, _dp, _dp, _i]),
        "orc_q2R": (None, [_dp, _dp]),
        "orc_lj_params_rigid": (None, [C.POINTER(LJParams)]),
        "orc_lj_params_airebo": (None, [C.POINTER(LJParams)]),
        "orc_rigid_nlist": (None, [_dp, _dp, _dp, _i, _d, _i, _ip, _ip]),
        "orc_rigid_lab": (None, [_dp, _dp, _i, _i, _dp]),
        "orc_rigid_forces": (_d, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip,
                                  _i, _i, _i, _d, C.POINTER(LJParams), _dp]),
        "orc_ke_trans": (_d, [_dp, _i, _d]),
        "orc_ke_rot": (_d, [_dp, _i, _d]),
        "orc_inst_T": (_d, [_d, _i]),
        "orc_inst_P": (_d, [_dp, _d, _d]),
        "orc_make_npt": (None, [C.POINTER(NPT), _d, _d, _i]),
        "orc_rigid_step_npt": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                      _d, _d, _i, _i, _d, _d, C.POINTER(NPT), _ip, _ip,
                                      _i, C.POINTER(LJParams), _dp, _dp]),
        "orc_rebo_fc": (_d, [_d]), "orc_rebo_dfc": (_d, [_d]),
        "orc_rebo_vr": (_d, [_d]), "orc_rebo_dvr": (_d, [_d]),
        "orc_rebo_va": (_d, [_d]), "orc_rebo_dva": (_d, [_d]),
        "orc_rebo_G": (_d, [_d]), "orc_rebo_dG": (_d, [_d]),
        "orc_build_nlist_rebo": (None, [_dp, _dp, _dp, _i, _ip, _d, _i, _ip, _ip]),
        "orc_build_nlist_lj": (None, [_dp, _dp, _dp, _i, _ip, _d, _i, _ip, _ip]),
        "orc_compute_rebo": (_d, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i, _i]),
        "orc_rebo_ext": (_d, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i, _dp, _d, _dp]),
        "orc_compute_lj": (_d, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i,
                                C.POINTER(LJParams)]),
        "orc_airebo_forces": (None, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _ip, _ip, _i,
                                     _i, _i, C.POINTER(LJParams), _dp]),
        "orc_ke_total": (_d, [_dp, _dp, _i]),
        "orc_airebo_step_npt": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _d,
                                       C.POINTER(NPT), _ip, _ip, _i, _ip, _ip, _i, _i,
                                       C.POINTER(LJParams), _dp]),
        "orc_lj_params_mm": (None, [C.POINTER(LJParams)]),
        "orc_mm_topology_counts": (None, [_i, _ip, _i, _ip]),
        "orc_mm_topology": (None, [_dp, _i, _ip, _i, _i, _dp, _ip, _dp, _ip, _dp, _ip, _dp,
                                   _ip, _ip, _dp, _ip]),
        "orc_mm_forces": (None, [_dp, _dp, _dp, _dp, _dp, C.POINTER(MMTopo), _ip, _ip, _i,
                                 _dp, _ip, _i, _i, _i, _dp]),
        "orc_mm_step_npt": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _i, _d, C.POINTER(NPT),
                                   C.POINTER(MMTopo), _ip, _ip, _i, _dp, _ip, _i, _i, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


def lj_params(kind: str = "rigid") -> LJParams:
    p = LJParams()
    if kind == "rigid":
        lib().orc_lj_params_rigid(C.byref(p))
    else:
        lib().orc_lj_params_airebo(C.byref(p))
    return p


def lj_params_custom(sigma: float, eps: float, rc: float) -> LJParams:
    sr = sigma / rc
    sr6 = sr ** 6
    return LJParams(sigma * sigma, eps, rc * rc, 4.0 * eps * (sr6 * sr6 - sr6))


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def inv9(h):
    h = _c(h)
    hi = np.zeros(9)
    lib().orc_mat_inv9(h, hi)
    return hi


def det9(h):
    return lib().orc_mat_det9(_c(h))


def rigid_nlist(pos, h, rmcut, max_neigh=80):
    pos, h = _c(pos).reshape(-1), _c(h)
    n = pos.size // 3
    hi = inv9(h)
    cnt = np.zeros(n, np.int32)
    lst = np.zeros(n * max_neigh, np.int32)
    lib().orc_rigid_nlist(pos, h, hi, n, rmcut, max_neigh, cnt, lst)
    return cnt, lst.reshape(n, max_neigh)


def rigid_forces(pos, qv, body, h, nl_count, nl_list, rmcut2, params=None):
    """orc_rigid_forces -> (Ep, Fv[N,3], Tv[N,3], Wm9[9])."""
    params = params or lj_params("rigid")
    pos, qv, body, h = _c(pos).reshape(-1), _c(qv).reshape(-1), _c(body).reshape(-1), _c(h)
    nl_count = _c(nl_count, np.int32)
    nl_list = _c(nl_list, np.int32)
    n = pos.size // 3
    natom = body.size // 3
    max_neigh = nl_list.size // n
    hi = inv9(h)
    Fv = np.zeros(n * 3)
    Tv = np.zeros(n * 3)
    W = np.zeros(9)
    lab = np.zeros(n * natom * 3)
    ep = lib().orc_rigid_forces(Fv, Tv, W, pos, qv, body, h, hi, nl_count,
                                nl_list.reshape(-1), max_neigh, n, natom, rmcut2,
                                C.byref(params), lab)
    return ep, Fv.reshape(n, 3), Tv.reshape(n, 3), W


def nlist_rebo(pos, h, mol_id, rcut=2.3, max_neigh=12):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    cnt = np.zeros(na, np.int32)
    lst = np.zeros(na * max_neigh, np.int32)
    lib().orc_build_nlist_rebo(pos, h, inv9(h), na, _c(mol_id, np.int32), rcut, max_neigh,
                               cnt, lst)
    return cnt, lst.reshape(na, max_neigh)


def nlist_lj(pos, h, mol_id, rlist=12.2, max_neigh=400):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    cnt = np.zeros(na, np.int32)
    lst = np.zeros(na * max_neigh, np.int32)
    lib().orc_build_nlist_lj(pos, h, inv9(h), na, _c(mol_id, np.int32), rlist, max_neigh,
                             cnt, lst)
    return cnt, lst.reshape(na, max_neigh)


def compute_rebo(pos, h, nlc, nll, mode=REBO_COMPAT):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nll = _c(nll, np.int32)
    F = np.zeros(na * 3)
    W = np.zeros(9)
    e = lib().orc_compute_rebo(F, W, pos, h, inv9(h), _c(nlc, np.int32), nll.reshape(-1),
                               nll.size // na, na, mode)
    return e, F.reshape(na, 3), W


# AIREBO carbon parameters of the two extension terms: P_CC(N_C, N_H = 0)
# at N_C = 0..4 and the CCCC torsion well depth (eV)
AIREBO_PCC = (0.0, 0.0, -0.0276030, 0.0, 0.0)
AIREBO_TORSION_EPS = 0.3079


def rebo_ext(pos, h, nlc, nll, pcc=AIREBO_PCC, eps_t=AIREBO_TORSION_EPS):
    """orc_rebo_ext: exact-gradient REBO with the P_ij spline and the torsion
    term -> (E total, F[na, 3], W9, E torsion)."""
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nll = _c(nll, np.int32)
    F = np.zeros(na * 3)
    W = np.zeros(9)
    et = np.zeros(1)
    e = lib().orc_rebo_ext(F, W, pos, h, inv9(h), _c(nlc, np.int32), nll.reshape(-1),
                           nll.size // na, na, _c(pcc), float(eps_t), et)
    return e, F.reshape(na, 3), W, float(et[0])


def compute_lj(pos, h, nlc, nll, params=None):
    params = params or lj_params("airebo")
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nll = _c(nll, np.int32)
    F = np.zeros(na * 3)
    W = np.zeros(9)
    e = lib().orc_compute_lj(F, W, pos, h, inv9(h), _c(nlc, np.int32), nll.reshape(-1),
                             nll.size // na, na, C.byref(params))
    return e, F.reshape(na, 3), W


def airebo_forces(pos, h, lists, mode=REBO_COMPAT, params=None):
    """lists = (nlc_r, nll_r, nlc_l, nll_l) -> ((Ep, Er, El), F[Na,3], W[9])."""
    params = params or lj_params("airebo")
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nlc_r, nll_r, nlc_l, nll_l = lists
    nll_r = _c(nll_r, np.int32)
    nll_l = _c(nll_l, np.int32)
    F = np.zeros(na * 3)
    W = np.zeros(9)
    out = np.zeros(3)
    lib().orc_airebo_forces(F, W, pos, h, inv9(h), _c(nlc_r, np.int32), nll_r.reshape(-1),
                            nll_r.size // na, _c(nlc_l, np.int32), nll_l.reshape(-1),
                            nll_l.size // na, na, mode, C.byref(params), out)
    return tuple(out), F.reshape(na, 3), W


# ------------------------------------------------------ molecular mechanics
def mm_topology(coords, bonds, nmol, k4=MM_K4_DEFAULT):
    """build_flat_topology (MM:254-334) of nmol copies: dict of flat arrays
    b (Nb,2), bp (Nb,{kb,r0}), a (Nang,3), ap, d (Ndih,4), dp, dm, i (Nimp,4),
    ip, mol_id."""
    coords = _c(coords).reshape(-1)
    bonds = _c(bonds, np.int32).reshape(-1)
    natom, nbond = coords.size // 3, bonds.size // 2
    cnt = np.zeros(4, np.int32)
    lib().orc_mm_topology_counts(natom, bonds, nbond, cnt)
    nb, na_, nd, ni = (int(x) * nmol for x in cnt)
    t = dict(b=np.zeros((nb, 2), np.int32), bp=np.zeros((nb, 2)),
             a=np.zeros((na_, 3), np.int32), ap=np.zeros((na_, 2)),
             d=np.zeros((nd, 4), np.int32), dp=np.zeros((nd, 2)),
             dm=np.zeros(nd, np.int32), i=np.zeros((ni, 4), np.int32),
             ip=np.zeros((ni, 2)), mol_id=np.zeros(nmol * natom, np.int32))
    lib().orc_mm_topology(coords, natom, bonds, nbond, nmol, _c(k4), t["b"].reshape(-1),
                          t["bp"].reshape(-1), t["a"].reshape(-1), t["ap"].reshape(-1),
                          t["d"].reshape(-1), t["dp"].reshape(-1), t["dm"],
                          t["i"].reshape(-1), t["ip"].reshape(-1), t["mol_id"])
    return t


def _mm_struct(t):
    arrs = {k: _c(t[k], np.int32 if k in ("b", "a", "d", "dm", "i") else np.float64)
            for k in ("b", "bp", "a", "ap", "d", "dp", "dm", "i", "ip")}
    s = MMTopo(len(arrs["b"]), len(arrs["a"]), len(arrs["d"]), len(arrs["i"]),
               *(arrs[k].ctypes.data for k in ("b", "bp", "a", "ap", "d", "dp", "dm", "i",
                                               "ip")))
    return s, arrs  # keep arrs alive with s


def mm_nlist(pos, h, mol_id, max_neigh=400):
    """build_nlist_lj (MM:337-349): half list within rc + 2 A."""
    return nlist_lj(pos, h, mol_id, rlist=MM_RLIST, max_neigh=max_neigh)


def mm_forces(pos, h, topo, nlc, nll, charge=None, mode=MM_COMPAT):
    """compute_forces (MM:351-788) -> (out7 {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot},
    F[Na,3], W[9])."""
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nll = _c(nll, np.int32)
    q = np.zeros(na) if charge is None else _c(charge)
    has_q = int(bool(np.any(np.abs(q) > 1e-10)))
    s, keep = _mm_struct(topo)
    F, W, out = np.zeros(na * 3), np.zeros(9), np.zeros(7)
    lib().orc_mm_forces(F, W, pos, h, inv9(h), C.byref(s), _c(nlc, np.int32),
                        nll.reshape(-1), nll.size // na, q, _c(topo["mol_id"], np.int32),
                        has_q, na, mode, out)
    return out, F.reshape(na, 3), W


def mm_step_npt(pos, vel, F, W, h, npt, dt, topo, nlc, nll, charge=None, mode=MM_COMPAT):
    """step_npt (MM:812-890) in place on copies; returns
    (pos, vel, F, W, h, npt, out8 {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot, KE})."""
    pos, vel, F, W, h = (_c(x).reshape(-1).copy() for x in (pos, vel, F, W, h))
    na = pos.size // 3
    hi = inv9(h)
    nll = _c(nll, np.int32)
    q = np.zeros(na) if charge is None else _c(charge)
    has_q = int(bool(np.any(np.abs(q) > 1e-10)))
    s, keep = _mm_struct(topo)
    out = np.zeros(8)
    lib().orc_mm_step_npt(pos, vel, F, W, h, hi, na, dt, C.byref(npt), C.byref(s),
                          _c(nlc, np.int32), nll.reshape(-1), nll.size // na, q,
                          _c(topo["mol_id"], np.int32), has_q, mode, out)
    return pos.reshape(na, 3), vel.reshape(na, 3), F.reshape(na, 3), W, h, npt, out
