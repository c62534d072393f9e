"""AIREBO path (REBO bond order + intermolecular LJ), computed on the GPU.

The functions keep the names and argument meaning of the reference's
(src/fuller_airebo_npt_md_serial_omp_acc.cpp "AB"): build_nlist_rebo,
build_nlist_lj, compute_forces, step_npt.  ``AtomicEngine`` is the
HBM-resident fast path.  ``mode``: REBO_COMPAT (the reference's forces and
pair-only virial) or REBO_PHYSICAL (exact gradient, full virial).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from ._native import NPTState  # noqa: F401

REBO_COMPAT = 0
REBO_PHYSICAL = 1
MAX_REBO_NEIGH = 12  # AB:126
MAX_LJ_NEIGH = 400  # AB:127
# AIREBO carbon values of the two REBO-II / AIREBO terms the reference does
# not have (engines, REBO_PHYSICAL only): P_CC(N, N_H = 0) at N = 0..4 and the
# torsion well depth (eV)
AIREBO_PCC = (0.0, 0.0, -0.0276030, 0.0, 0.0)
AIREBO_TORSION_EPS = 0.3079


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def make_npt(T: float, Pe: float, Na: int) -> N.NPTState:
    """make_npt (AB:512-515): Nf = 3 Na - 3."""
    s = N.NPTState()
    N.lib().fmd_make_npt(C.byref(s), T, Pe, 3 * Na - 3)
    return s


def build_nlist_rebo(pos, h, hi, mol_id):
    """build_nlist_rebo (AB:251-261) -> (nlc[Na], nll[Na, 12])."""
    pos = _c(pos).reshape(-1)
    na = pos.size // 3
    nlc = np.zeros(na, np.int32)
    nll = np.zeros(na * MAX_REBO_NEIGH, np.int32)
    N.check(N.lib().fmd_airebo_nlist_rebo(pos, _c(h), _c(hi), na, _c(mol_id, np.int32), nlc,
                                          nll))
    return nlc, nll.reshape(na, MAX_REBO_NEIGH)


def build_nlist_lj(pos, h, hi, mol_id):
    """build_nlist_lj (AB:263-272) -> (nlc[Na], nll[Na, 400])."""
    pos = _c(pos).reshape(-1)
    na = pos.size // 3
    nlc = np.zeros(na, np.int32)
    nll = np.zeros(na * MAX_LJ_NEIGH, np.int32)
    N.check(N.lib().fmd_airebo_nlist_lj(pos, _c(h), _c(hi), na, _c(mol_id, np.int32), nlc, nll))
    return nlc, nll.reshape(na, MAX_LJ_NEIGH)


def compute_forces(pos, h, hi, lists, mol_id, mode=REBO_COMPAT):
    """compute_forces (AB:477-495) -> ((Ep, Ep_rebo, Ep_lj), F[Na,3], vir9[9])."""
    pos = _c(pos).reshape(-1)
    na = pos.size // 3
    nlc_r, nll_r, nlc_l, nll_l = (_c(x, np.int32).reshape(-1) for x in lists)
    F, W, out = np.zeros(na * 3), np.zeros(9), np.zeros(3)
    N.check(N.lib().fmd_airebo_compute_forces(F, W, pos, _c(h), _c(hi), nlc_r, nll_r, nlc_l,
                                              nll_l, na, _c(mol_id, np.int32), mode, out))
    return tuple(out), F.reshape(na, 3), W


def step_npt(pos, vel, F, vir9, h, hi, mass, dt, npt, lists, mol_id, mode=REBO_COMPAT):
    """step_npt (AB:518-590): in place like the reference; returns
    (Ep, Ep_rebo, Ep_lj, KE)."""
    for a in (pos, vel, F, vir9, h, hi):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
            raise TypeError("step_npt updates its arrays in place: pass C-contiguous float64")
    na = pos.size // 3
    nlc_r, nll_r, nlc_l, nll_l = (_c(x, np.int32).reshape(-1) for x in lists)
    out = np.zeros(4)
    N.check(N.lib().fmd_airebo_step_npt(pos.reshape(-1), vel.reshape(-1), F.reshape(-1), vir9, h,
                                        hi, _c(mass), na, dt, C.byref(npt), nlc_r, nll_r, nlc_l,
                                        nll_l, _c(mol_id, np.int32), mode, out))
    return tuple(out)


def _extension_terms(cfg, pcc, torsion_eps):
    """The REBO-II P_ij knots (5 values, P_CC at N = 0..4) and the AIREBO
    torsion well depth into an engine configuration (None / 0: off)."""
    if pcc is not None:
        if len(pcc) != 5:
            raise ValueError("pcc: P_CC at N = 0..4 (5 values)")
        for k in range(5):
            cfg.rebo_pcc[k] = float(pcc[k])
    cfg.torsion_eps = float(torsion_eps)


class AtomicEngine:
    """HBM-resident AIREBO NPT engine (fmd_atomic_create / fmd_engine_*)."""

    def __init__(self, mol_id, mass, dt=0.5, mode=REBO_COMPAT, lj_max_neigh=MAX_LJ_NEIGH,
                 nlup=20, lj=None, rebo_rlist=None, clamp_vg=None, al_skin=None,
                 cell_min_molecules=None, rebo_slots=None, rebo_threads=None, pcc=None,
                 torsion_eps=0.0):
        self._mol = _c(mol_id, np.int32)
        self._mass = _c(mass)
        cfg = N.AtomicConfig()
        N.lib().fmd_atomic_config_default(C.byref(cfg))
        cfg.n_atoms = len(self._mol)
        cfg.mol_id = self._mol.ctypes.data_as(C.POINTER(C.c_int))
        cfg.mass = self._mass.ctypes.data_as(C.POINTER(C.c_double))
        cfg.dt, cfg.rebo_mode, cfg.lj_max_neigh, cfg.nlup = dt, mode, lj_max_neigh, nlup
        if lj is not None:
            cfg.lj = lj
        if rebo_rlist is not None:
            cfg.rebo_rlist = rebo_rlist
        if clamp_vg is not None:
            cfg.clamp_vg = clamp_vg
        if al_skin is not None:
            cfg.al_skin = al_skin
        if cell_min_molecules is not None:
            cfg.cell_min_molecules = cell_min_molecules
        if rebo_slots is not None:
            cfg.rebo_lds_slots = int(rebo_slots)
        if rebo_threads is not None:
            cfg.rebo_threads = int(rebo_threads)
        _extension_terms(cfg, pcc, torsion_eps)
        self.cfg = cfg
        self.n = len(self._mol)
        h = C.c_void_p()
        N.check(N.lib().fmd_atomic_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            N.lib().fmd_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos=None, vel=None, h=None, npt=None):
        arrs = [None if a is None else _c(a).reshape(-1) for a in (pos, vel, h)]
        N.check(N.lib().fmd_atomic_set_state(self._h, *[N.ptr(a) for a in arrs],
                                             C.byref(npt) if npt is not None else None))

    def get_state(self):
        pos, vel, h = np.zeros((self.n, 3)), np.zeros((self.n, 3)), np.zeros(9)
        npt = N.NPTState()
        N.check(N.lib().fmd_atomic_get_state(self._h, N.ptr(pos), N.ptr(vel), N.ptr(h),
                                             C.byref(npt)))
        return dict(pos=pos, vel=vel, h=h, npt=npt)

    def prepare(self):
        N.check(N.lib().fmd_engine_prepare(self._h))

    def forces(self):
        F, W, E = np.zeros((self.n, 3)), np.zeros(9), np.zeros(3)
        N.check(N.lib().fmd_engine_forces(self._h, N.ptr(F), None, N.ptr(W), N.ptr(E)))
        return tuple(E), F, W

    def run(self, first_step, nsteps, ctrl=None, thermo=True):
        ctrl_arr = (N.StepCtrl * nsteps)(*ctrl) if ctrl is not None else None
        th = (N.Thermo * nsteps)() if thermo else None
        N.check(N.lib().fmd_engine_run(self._h, first_step, nsteps,
                                       C.cast(ctrl_arr, C.c_void_p) if ctrl_arr else None,
                                       C.cast(th, C.c_void_p) if th else None))
        return th

    def counters(self):
        out = (C.c_longlong * 3)()
        N.check(N.lib().fmd_engine_counters(self._h, out))
        return list(out)

    def set_timing(self, on=True, stride=1):
        """Event timing of the step phases on every `stride`-th step (and on
        list rebuilds)."""
        N.check(N.lib().fmd_engine_set_timing(self._h, stride if on else 0))

    def kernel_time(self):
        ms, n = C.c_double(0), C.c_int(0)
        N.check(N.lib().fmd_engine_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def phase_times(self):
        """Mean ms and counts of the timed phases since the last read:
        {"force", "intra", "lj", "list"} -> (ms, n)."""
        ms, n = np.zeros(4), np.zeros(4, np.int32)
        N.check(N.lib().fmd_engine_phase_times(self._h, ms, n))
        return {k: (float(ms[i]), int(n[i])) for i, k in
                enumerate(("force", "intra", "lj", "list"))}

    def sync(self):
        N.check(N.lib().fmd_engine_sync(self._h))

    def set_graph(self, on=True):
        N.check(N.lib().fmd_engine_set_graph(self._h, 1 if on else 0))


class ReplicaBatch:
    """R independent AIREBO NPT replicas in one engine (fmd_batch_*): every
    kernel launch of a step covers all replicas; steps replay from captured
    HIP graphs.  replicas: list of (mol_id, mass) per replica."""

    def __init__(self, replicas, dt=0.5, mode=REBO_COMPAT, lj_max_neigh=MAX_LJ_NEIGH, nlup=20,
                 lj=None, graph=True, al_skin=None, rebo_slots=None, rebo_threads=None,
                 pcc=None, torsion_eps=0.0):
        self._keep = []
        cfgs = (N.AtomicConfig * len(replicas))()
        for r, (mol_id, mass) in enumerate(replicas):
            mol, m = _c(mol_id, np.int32), _c(mass)
            self._keep += [mol, m]
            N.lib().fmd_atomic_config_default(C.byref(cfgs[r]))
            cfgs[r].n_atoms = len(mol)
            cfgs[r].mol_id = mol.ctypes.data_as(C.POINTER(C.c_int))
            cfgs[r].mass = m.ctypes.data_as(C.POINTER(C.c_double))
            cfgs[r].dt, cfgs[r].rebo_mode = dt, mode
            cfgs[r].lj_max_neigh, cfgs[r].nlup = lj_max_neigh, nlup
            if lj is not None:
                cfgs[r].lj = lj
            if al_skin is not None:
                cfgs[r].al_skin = al_skin
            if rebo_slots is not None:
                cfgs[r].rebo_lds_slots = int(rebo_slots)
            if rebo_threads is not None:
                cfgs[r].rebo_threads = int(rebo_threads)
            _extension_terms(cfgs[r], pcc, torsion_eps)
        self.sizes = [len(mol_id) for mol_id, _ in replicas]
        self.R = len(replicas)
        h = C.c_void_p()
        N.check(N.lib().fmd_batch_create(C.cast(cfgs, C.c_void_p), self.R, C.byref(h)))
        self._h = h
        N.check(N.lib().fmd_engine_set_graph(self._h, 1 if graph else 0))

    def close(self):
        if self._h:
            N.lib().fmd_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, r, pos=None, vel=None, h=None, npt=None):
        arrs = [None if a is None else _c(a).reshape(-1) for a in (pos, vel, h)]
        N.check(N.lib().fmd_batch_set_state(self._h, r, *[N.ptr(a) for a in arrs],
                                            C.byref(npt) if npt is not None else None))

    def get_state(self, r):
        n = self.sizes[r]
        pos, vel, h = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(9)
        npt = N.NPTState()
        N.check(N.lib().fmd_batch_get_state(self._h, r, N.ptr(pos), N.ptr(vel), N.ptr(h),
                                            C.byref(npt)))
        return dict(pos=pos, vel=vel, h=h, npt=npt)

    def prepare(self):
        N.check(N.lib().fmd_engine_prepare(self._h))

    def forces(self, r):
        F, W, E = np.zeros((self.sizes[r], 3)), np.zeros(9), np.zeros(3)
        N.check(N.lib().fmd_batch_forces(self._h, r, N.ptr(F), N.ptr(W), N.ptr(E)))
        return tuple(E), F, W

    def run(self, first_step, nsteps, ctrl=None, thermo=True):
        """ctrl: nsteps x R StepCtrl (row-major) or None.  Returns the records
        as an (nsteps, R) object array of Thermo, or None."""
        ctrl_arr = None
        if ctrl is not None:
            flat = [c for row in ctrl for c in row]
            ctrl_arr = (N.StepCtrl * len(flat))(*flat)
        th = (N.Thermo * (nsteps * self.R))() if thermo else None
        N.check(N.lib().fmd_engine_run(self._h, first_step, nsteps,
                                       C.cast(ctrl_arr, C.c_void_p) if ctrl_arr else None,
                                       C.cast(th, C.c_void_p) if th else None))
        if th is None:
            return None
        return [[th[k * self.R + r] for r in range(self.R)] for k in range(nsteps)]

    def counters(self):
        out = (C.c_longlong * 3)()
        N.check(N.lib().fmd_engine_counters(self._h, out))
        return list(out)

    def set_timing(self, on=True, stride=1):
        """Event timing of the step phases on every `stride`-th step (and on
        list rebuilds)."""
        N.check(N.lib().fmd_engine_set_timing(self._h, stride if on else 0))

    def kernel_time(self):
        ms, n = C.c_double(0), C.c_int(0)
        N.check(N.lib().fmd_engine_kernel_time(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def phase_times(self):
        """Mean ms and counts of the timed phases since the last read:
        {"force", "intra", "lj", "list"} -> (ms, n)."""
        ms, n = np.zeros(4), np.zeros(4, np.int32)
        N.check(N.lib().fmd_engine_phase_times(self._h, ms, n))
        return {k: (float(ms[i]), int(n[i])) for i, k in
                enumerate(("force", "intra", "lj", "list"))}

    def sync(self):
        N.check(N.lib().fmd_engine_sync(self._h))
