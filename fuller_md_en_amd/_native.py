"""ctypes binding of libfmd.so (the C ABI declared in include/fmd.h).

The library is built in-tree by ``fuller_md_en_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfmd.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "fmd.h")

FMD_OK = 0


class FmdError(RuntimeError):
    pass


class LJParams(C.Structure):
    _fields_ = [("sigma", C.c_double), ("eps", C.c_double), ("rcut", C.c_double),
                ("vshift", C.c_double), ("rmin2", C.c_double)]


class NPTState(C.Structure):
    """NPTState of the reference (C1:478)."""
    _fields_ = [("xi", C.c_double), ("Q", C.c_double), ("Vg", C.c_double * 9),
                ("W", C.c_double), ("Pe", C.c_double), ("Tt", C.c_double), ("Nf", C.c_int)]

    def to_array(self) -> np.ndarray:
        return np.array([self.xi, self.Q, *self.Vg, self.W, self.Pe, self.Tt, float(self.Nf)])

    @classmethod
    def from_array(cls, a) -> "NPTState":
        s = cls()
        s.xi, s.Q = a[0], a[1]
        for k in range(9):
            s.Vg[k] = a[2 + k]
        s.W, s.Pe, s.Tt, s.Nf = a[11], a[12], a[13], int(a[14])
        return s


class RigidConfig(C.Structure):
    _fields_ = [("n_mol", C.c_int), ("natom", C.c_int), ("body", C.POINTER(C.c_double)),
                ("I0", C.c_double), ("Mmol", C.c_double), ("lj", LJParams),
                ("rmcut", C.c_double), ("skin", C.c_double), ("max_neigh", C.c_int),
                ("nlup", C.c_int), ("dt", C.c_double), ("clamp_xi", C.c_double),
                ("clamp_vg", C.c_double), ("half_pairs", C.c_int),
                ("cell_min_molecules", C.c_int)]


class AtomicConfig(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("mol_id", C.POINTER(C.c_int)),
                ("mass", C.POINTER(C.c_double)), ("lj", LJParams), ("lj_skin", C.c_double),
                ("lj_max_neigh", C.c_int), ("rebo_rlist", C.c_double),
                ("rebo_max_neigh", C.c_int), ("rebo_mode", C.c_int), ("nlup", C.c_int),
                ("dt", C.c_double), ("clamp_xi", C.c_double), ("clamp_vg", C.c_double),
                ("al_skin", C.c_double), ("cell_min_molecules", C.c_int),
                ("rebo_lds_slots", C.c_int), ("rebo_threads", C.c_int),
                ("rebo_pcc", C.c_double * 5), ("torsion_eps", C.c_double)]


class StepCtrl(C.Structure):
    _fields_ = [("Tt", C.c_double), ("rescale_T", C.c_double), ("flags", C.c_int),
                ("pad", C.c_int)]


class Thermo(C.Structure):
    _fields_ = [("Ep", C.c_double), ("KE", C.c_double), ("KEt", C.c_double),
                ("KEr", C.c_double), ("T", C.c_double), ("P", C.c_double), ("V", C.c_double),
                ("xi", C.c_double), ("Ep_rebo", C.c_double), ("Ep_lj", C.c_double),
                ("h", C.c_double * 9), ("Eb", C.c_double), ("Ea", C.c_double),
                ("Ed", C.c_double), ("Ei", C.c_double), ("Ecoul", C.c_double)]


_IP, _DP = C.POINTER(C.c_int), C.POINTER(C.c_double)


class MMTopology(C.Structure):
    """fmd_mm_topology == the reference's FlatTopology (MM:224-239)."""
    _fields_ = [("Nb", C.c_int), ("b_i0", _IP), ("b_i1", _IP), ("b_kb", _DP), ("b_r0", _DP),
                ("Nang", C.c_int), ("ang_i0", _IP), ("ang_i1", _IP), ("ang_i2", _IP),
                ("ang_kth", _DP), ("ang_th0", _DP),
                ("Ndih", C.c_int), ("dih_i0", _IP), ("dih_i1", _IP), ("dih_i2", _IP),
                ("dih_i3", _IP), ("dih_Vn", _DP), ("dih_mult", _IP), ("dih_gamma", _DP),
                ("Nimp", C.c_int), ("imp_i0", _IP), ("imp_i1", _IP), ("imp_i2", _IP),
                ("imp_i3", _IP), ("imp_ki", _DP), ("imp_gamma", _DP), ("mol_id", _IP)]


class MMConfig(C.Structure):
    _fields_ = [("n_atoms", C.c_int), ("mol_id", _IP), ("ft", C.POINTER(MMTopology)),
                ("charge", _DP), ("lj", LJParams), ("lj_skin", C.c_double),
                ("lj_max_neigh", C.c_int), ("nlup", C.c_int), ("dt", C.c_double),
                ("clamp_xi", C.c_double), ("clamp_vg", C.c_double), ("mm_mode", C.c_int)]


CTRL_ZERO_XI_VG = 1
CTRL_ZERO_VG = 2
CTRL_RESCALE = 4
CTRL_RESCALE_ZERO_VG = 8
CTRL_REBUILD = 16

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d, _i, _vp = C.c_double, C.c_int, C.c_void_p


_SIGS = {
    "fmd_last_error": (C.c_char_p, []),
    "fmd_version": (_i, []),
    "fmd_device_info": (_i, [C.c_char_p, _i, C.POINTER(_i), C.POINTER(_d)]),
    "fmd_lj_params_rigid": (None, [C.POINTER(LJParams)]),
    "fmd_lj_params_airebo": (None, [C.POINTER(LJParams)]),
    "fmd_lj_params_mm": (None, [C.POINTER(LJParams)]),
    "fmd_load_cc1_bonds": (_i, [C.c_char_p, _ip, _i, C.POINTER(_i)]),
    "fmd_mm_topology_counts": (_i, [_i, _ip, _i, _ip]),
    "fmd_mm_build_topology": (_i, [_dp, _i, _ip, _i, _i, _d, _d, _d, _d,
                                   C.POINTER(MMTopology)]),
    "fmd_mm_nlist_lj": (_i, [_dp, _dp, _dp, _i, _ip, _ip, _ip]),
    "fmd_mm_compute_forces": (_i, [_dp, _dp, _dp, _dp, _dp, C.POINTER(MMTopology), _ip, _ip,
                                   _vp, _ip, _i, _i, _i, _dp]),
    "fmd_mm_step_npt": (_i, [_dp, _dp, _dp, _dp, _dp, _dp, _i, _d, C.POINTER(NPTState),
                             C.POINTER(MMTopology), _ip, _ip, _vp, _ip, _i, _i, _dp]),
    "fmd_mm_config_default": (None, [C.POINTER(MMConfig)]),
    "fmd_mm_create": (_i, [C.POINTER(MMConfig), C.POINTER(_vp)]),
    "fmd_mm_energies": (_i, [_vp, _dp]),
    "fmd_make_npt": (None, [C.POINTER(NPTState), _d, _d, _i]),
    "fmd_rigid_nlist": (_i, [_dp, _dp, _dp, _i, _d, _i, _ip, _ip, C.POINTER(_i)]),
    "fmd_rigid_forces": (_i, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _i, _i, _i,
                              _d, C.POINTER(LJParams), _vp, C.POINTER(_d)]),
    "fmd_rigid_forces_noalloc": (_i, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _ip, _ip, _i,
                                      _i, _i, _d, C.POINTER(LJParams), _vp, C.POINTER(_d)]),
    "fmd_rigid_step_npt": (_i, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _d, _d, _i,
                                _i, _d, _d, C.POINTER(NPTState), _ip, _ip, _i,
                                C.POINTER(LJParams), C.POINTER(_d), C.POINTER(_d)]),
    "fmd_rigid_config_default": (None, [C.POINTER(RigidConfig), _d]),
    "fmd_rigid_create": (_i, [C.POINTER(RigidConfig), C.POINTER(_vp)]),
    "fmd_rigid_set_state": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_rigid_get_state": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_airebo_nlist_rebo": (_i, [_dp, _dp, _dp, _i, _ip, _ip, _ip]),
    "fmd_airebo_nlist_lj": (_i, [_dp, _dp, _dp, _i, _ip, _ip, _ip]),
    "fmd_airebo_compute_forces": (_i, [_dp, _dp, _dp, _dp, _dp, _ip, _ip, _ip, _ip, _i, _ip,
                                       _i, _dp]),
    "fmd_airebo_step_npt": (_i, [_dp, _dp, _dp, _dp, _dp, _dp, _dp, _i, _d,
                                 C.POINTER(NPTState), _ip, _ip, _ip, _ip, _ip, _i, _dp]),
    "fmd_atomic_config_default": (None, [C.POINTER(AtomicConfig)]),
    "fmd_atomic_create": (_i, [C.POINTER(AtomicConfig), C.POINTER(_vp)]),
    "fmd_atomic_set_state": (_i, [_vp, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_atomic_get_state": (_i, [_vp, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_batch_create": (_i, [_vp, _i, C.POINTER(_vp)]),
    "fmd_batch_size": (_i, [_vp, C.POINTER(_i)]),
    "fmd_batch_set_state": (_i, [_vp, _i, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_batch_get_state": (_i, [_vp, _i, _vp, _vp, _vp, C.POINTER(NPTState)]),
    "fmd_batch_forces": (_i, [_vp, _i, _vp, _vp, _vp]),
    "fmd_engine_set_graph": (_i, [_vp, _i]),
    "fmd_sg_param_count": (_i, []),
    "fmd_sg_feature_count": (_i, []),
    "fmd_sg_attach": (_i, [_vp, _i]),
    "fmd_sg_set_params": (_i, [_vp, _vp, _vp, _vp]),
    "fmd_sg_get_params": (_i, [_vp, _vp, _vp, _vp]),
    "fmd_sg_collect": (_i, [_vp]),
    "fmd_sg_frame_count": (_i, [_vp, C.POINTER(_i)]),
    "fmd_sg_fit_normalisation": (_i, [_vp]),
    "fmd_sg_train": (_i, [_vp, _i, _d, _vp]),
    "fmd_sg_loss_grad": (_i, [_vp, _i, _vp, _vp]),
    "fmd_sg_descriptor": (_i, [_vp, _i, _vp]),
    "fmd_sg_evaluate": (_i, [_vp, _i, _vp, _vp, _vp, _vp]),
    "fmd_sg_enable": (_i, [_vp, _i, _i, _d]),
    "fmd_sg_stats": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fmd_engine_prepare": (_i, [_vp]),
    "fmd_engine_run": (_i, [_vp, _i, _i, _vp, _vp]),
    "fmd_engine_forces": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "fmd_engine_counters": (_i, [_vp, C.POINTER(C.c_longlong)]),
    "fmd_engine_set_timing": (_i, [_vp, _i]),
    "fmd_engine_kernel_time": (_i, [_vp, C.POINTER(_d), C.POINTER(_i)]),
    "fmd_engine_phase_times": (_i, [_vp, _dp, _ip]),
    "fmd_engine_sync": (_i, [_vp]),
    "fmd_engine_destroy": (None, [_vp]),
}

_LIB = None


def lib() -> C.CDLL:
    """Load libfmd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise FmdError(f"{LIB_PATH} is missing: run `python -m fuller_md_en_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != FMD_OK:
        raise FmdError(f"libfmd error {rc}: {lib().fmd_last_error().decode()}")


def declared_symbols() -> list[str]:
    """Function names declared in include/fmd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fmd_[a-z0-9_]+)\s*\(", text)))


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
