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
        "orc_compute_lj": (_d, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i,
                                C.POINTER(LJParams)]),
        "orc_airebo_forces": (None, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _ip, _ip, _i,
                                     _i, _i, C.POINTER(LJParams), _dp]),
        "orc_ke_total": (_d, [_dp, _dp, _i]),
        "orc_airebo_step_npt": (None, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _d,
                                       C.POINTER(NPT), _ip, _ip, _i, _ip, _ip, _i, _i,
                                       C.POINTER(LJParams), _dp]),
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
