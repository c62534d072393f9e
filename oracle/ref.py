"""ctypes bindings of the reference build in ``oracle/_ref`` (test infrastructure).

The shared libraries are compiled by ``oracle/Makefile`` from the reference
sources where they lie; nothing of the reference is kept in this repository.
``available()`` is False wherever they have not been built (e.g. where the
reference tree is absent); callers then skip.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import REF_DIR, inv9

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d, _i = C.c_double, C.c_int


def path(name: str) -> str:
    return os.path.join(REF_DIR, name)


def available(which: str = "all") -> bool:
    names = {"core": ["libref_core.so"], "lj": ["libref_lj.so"],
             "airebo": ["libref_airebo.so"], "mm": ["libref_mm.so"]}
    want = sum(names.values(), []) if which == "all" else names[which]
    return all(os.path.exists(path(n)) for n in want)


_libs: dict = {}


def core():
    if "core" not in _libs:
        L = C.CDLL(path("libref_core.so"))
        L.refc_rmcut.restype = _d
        L.refc_generate_c60.argtypes = [_dp, _dp]
        L.refc_make_fcc.argtypes = [_d, _i, _dp, _dp]
        L.refc_nlist.argtypes = [_dp, _dp, _dp, _i, _d, _ip, _ip]
        L.refc_forces.restype = _d
        L.refc_forces.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i,
                                  _d, _dp]
        L.refc_make_npt.argtypes = [_d, _d, _i, _dp]
        L.refc_step_npt.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _d,
                                    _i, _i, _d, _d, _dp, _ip, _ip, _dp, _dp]
        L.refc_init.argtypes = [_i, C.c_uint, _d, _dp, _dp, _dp, _dp, _dp]
        L.refc_init.restype = _i
        L.refc_normal_stream.argtypes = [C.c_uint, _i, _dp]
        L.refc_main_loop.argtypes = [_i, C.c_uint, _i, _dp]
        L.refc_main_loop.restype = _i
        L.refc_apply_pbc.argtypes = [_dp, _dp, _dp, _i]
        L.refc_mat_inv9.argtypes = [_dp, _dp]
        _libs["core"] = L
    return _libs["core"]


def lj():
    if "lj" not in _libs:
        L = C.CDLL(path("libref_lj.so"))
        L.reflj_load_cc1.argtypes = [C.c_char_p, _dp, _dp]
        L.reflj_default_a0.restype = _d
        L.reflj_default_a0.argtypes = [_d, C.c_char_p, _d]
        L.reflj_make_crystal.argtypes = [C.c_char_p, _d, _i, _dp, _dp]
        L.reflj_apply_pbc.argtypes = [_dp, _dp, _dp, _i]
        L.reflj_nlist.argtypes = [_dp, _dp, _dp, _i, _d, _ip, _ip]
        L.reflj_forces.restype = _d
        L.reflj_forces.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i,
                                   _d, _dp]
        L.reflj_step_npt.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d,
                                     _d, _i, _i, _d, _d, _dp, _ip, _ip, _dp, _dp]
        L.reflj_make_npt.argtypes = [_d, _d, _i, _dp]
        L.reflj_restart_filename.argtypes = [C.c_char_p, _i, _i, C.c_char_p, _i]
        _libs["lj"] = L
    return _libs["lj"]


def airebo(omp: bool = False):
    """The AIREBO harness: the serial build, or (omp) the OpenMP build."""
    key = "airebo_omp" if omp else "airebo"
    if key not in _libs:
        L = C.CDLL(path("libref_airebo_omp.so" if omp else "libref_airebo.so"))
        L.refab_load_cc1.argtypes = [C.c_char_p, _dp, _dp]
        L.refab_nlist_rebo.argtypes = [_dp, _dp, _dp, _i, _ip, _ip, _ip]
        L.refab_nlist_lj.argtypes = [_dp, _dp, _dp, _i, _ip, _ip, _ip]
        L.refab_compute_rebo.restype = _d
        L.refab_compute_rebo.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i]
        L.refab_compute_lj.restype = _d
        L.refab_compute_lj.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _i]
        L.refab_compute_forces.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _i,
                                           _dp]
        L.refab_make_npt.argtypes = [_d, _d, _i, _dp]
        L.refab_step_npt.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _d, _dp, _ip,
                                     _ip, _ip, _ip, _dp]
        L.refab_apply_pbc.argtypes = [_dp, _dp, _dp, _i]
        L.refab_mat_inv9.argtypes = [_dp, _dp]
        _libs[key] = L
    return _libs[key]


def mm():
    if "mm" not in _libs:
        L = C.CDLL(path("libref_mm.so"))
        L.refmm_load_cc1.argtypes = [C.c_char_p, _dp, _ip, _i, _dp]
        L.refmm_build.argtypes = [C.c_char_p, _i, _dp, _ip]
        L.refmm_get.argtypes = [_ip, _dp, _ip, _dp, _ip, _dp, _ip, _ip, _dp, _ip, _i]
        L.refmm_nlist_lj.argtypes = [_dp, _dp, _dp, _i, _ip, _ip]
        L.refmm_compute_forces.argtypes = [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _dp, _i, _i,
                                           _dp]
        L.refmm_make_npt.argtypes = [_d, _d, _i, _dp]
        L.refmm_step_npt.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, _i, _d, _dp, _ip, _ip,
                                     _dp, _i, _dp]
        L.refmm_apply_pbc.argtypes = [_dp, _dp, _dp, _i]
        L.refmm_mat_inv9.argtypes = [_dp, _dp]
        _libs["mm"] = L
    return _libs["mm"]


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


# --------------------------------------------------------------- core (C1)
def core_c60():
    coords = np.zeros(180)
    props = np.zeros(3)
    core().refc_generate_c60(coords, props)
    return coords.reshape(60, 3), {"I0": props[0], "Mmol": props[1], "Rmol": props[2]}


def core_init(nc, seed=42, T=300.0):
    n = 4 * nc ** 3
    pos, vel, omg = np.zeros(n * 3), np.zeros(n * 3), np.zeros(n * 3)
    qv, h = np.zeros(n * 4), np.zeros(9)
    core().refc_init(nc, seed, T, pos, vel, omg, qv, h)
    return pos.reshape(n, 3), vel.reshape(n, 3), omg.reshape(n, 3), qv.reshape(n, 4), h


def core_nlist(pos, h, rmcut=None):
    pos, h = _c(pos).reshape(-1), _c(h)
    n = pos.size // 3
    rmcut = core().refc_rmcut() if rmcut is None else rmcut
    cnt = np.zeros(n, np.int32)
    lst = np.zeros(n * 80, np.int32)
    core().refc_nlist(pos, h, inv9(h), n, rmcut, cnt, lst)
    return cnt, lst.reshape(n, 80)


def core_forces(pos, qv, body, h, nl_count, nl_list, rmcut2=None):
    pos, qv, body, h = _c(pos).reshape(-1), _c(qv).reshape(-1), _c(body).reshape(-1), _c(h)
    n = pos.size // 3
    natom = body.size // 3
    if rmcut2 is None:
        rmcut2 = core().refc_rmcut() ** 2
    Fv, Tv, W = np.zeros(n * 3), np.zeros(n * 3), np.zeros(9)
    lab = np.zeros(n * natom * 3)
    ep = core().refc_forces(Fv, Tv, W, pos, qv, body, h, inv9(h), _c(nl_count, np.int32),
                            _c(nl_list, np.int32).reshape(-1), n, natom, rmcut2, lab)
    return ep, Fv.reshape(n, 3), Tv.reshape(n, 3), W


# ------------------------------------------------------------------ AB
def ab_nlists(pos, h, mol_id):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    hi = inv9(h)
    mol_id = _c(mol_id, np.int32)
    nlc_r = np.zeros(na, np.int32)
    nll_r = np.zeros(na * 12, np.int32)
    nlc_l = np.zeros(na, np.int32)
    nll_l = np.zeros(na * 400, np.int32)
    airebo().refab_nlist_rebo(pos, h, hi, na, mol_id, nlc_r, nll_r)
    airebo().refab_nlist_lj(pos, h, hi, na, mol_id, nlc_l, nll_l)
    return nlc_r, nll_r.reshape(na, 12), nlc_l, nll_l.reshape(na, 400)


def ab_compute_forces(pos, h, lists):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nlc_r, nll_r, nlc_l, nll_l = (_c(x, np.int32).reshape(-1) for x in lists)
    F, W, out = np.zeros(na * 3), np.zeros(9), np.zeros(3)
    airebo().refab_compute_forces(F, W, pos, h, inv9(h), nlc_r, nll_r, nlc_l, nll_l, na, out)
    return tuple(out), F.reshape(na, 3), W


# ------------------------------------------------------------------ MM
def mm_load_cc1(path_cc1):
    coords = np.zeros(3 * 256)
    bonds = np.zeros(2 * 1024, np.int32)
    props = np.zeros(4)
    n = mm().refmm_load_cc1(path_cc1.encode(), coords, bonds, 1024, props)
    nb = int(props[1])
    return coords[:3 * n].reshape(n, 3), bonds[:2 * nb].reshape(nb, 2), props


def mm_build(path_cc1, nmol, k4):
    """The reference flat topology of nmol copies (kept inside the library for
    mm_compute_forces / mm_step_npt) as a dict of arrays."""
    counts = np.zeros(4, np.int32)
    mm().refmm_build(path_cc1.encode(), nmol, _c(k4), counts)
    nb, na_, nd, ni = (int(x) for x in counts)
    natom = mm_load_cc1(path_cc1)[0].shape[0]
    t = dict(b=np.zeros((nb, 2), np.int32), bp=np.zeros((nb, 2)),
             a=np.zeros((na_, 3), np.int32), ap=np.zeros((na_, 2)),
             d=np.zeros((nd, 4), np.int32), dp=np.zeros((nd, 2)),
             dm=np.zeros(nd, np.int32), i=np.zeros((ni, 4), np.int32),
             ip=np.zeros((ni, 2)), mol_id=np.zeros(nmol * natom, np.int32))
    mm().refmm_get(t["b"].reshape(-1), t["bp"].reshape(-1), t["a"].reshape(-1),
                   t["ap"].reshape(-1), t["d"].reshape(-1), t["dp"].reshape(-1), t["dm"],
                   t["i"].reshape(-1), t["ip"].reshape(-1), t["mol_id"], nmol * natom)
    return t


def mm_nlist(pos, h):
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    nlc = np.zeros(na, np.int32)
    nll = np.zeros(na * 400, np.int32)
    mm().refmm_nlist_lj(pos, h, inv9(h), na, nlc, nll)
    return nlc, nll.reshape(na, 400)


def mm_compute_forces(pos, h, nlc, nll, charge=None):
    """Over the topology of the last mm_build: (out7, F, vir9)."""
    pos, h = _c(pos).reshape(-1), _c(h)
    na = pos.size // 3
    q = np.zeros(na) if charge is None else _c(charge)
    F, W, out = np.zeros(na * 3), np.zeros(9), np.zeros(7)
    mm().refmm_compute_forces(F, W, pos, h, inv9(h), _c(nlc, np.int32),
                              _c(nll, np.int32).reshape(-1), q,
                              int(charge is not None and bool(np.any(np.abs(q) > 1e-10))),
                              na, out)
    return out, F.reshape(na, 3), W
