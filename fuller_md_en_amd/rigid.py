"""Rigid-body LJ path: the reference's functions, computed on the GPU.

Each function keeps the name, argument meaning and in-place behaviour of the
reference function it replaces (src/fuller_LJ_npt_md_serial_omp_acc.cpp "LJ",
src/fuller_LJ_npt_md_core_serial.cpp "C1"); arrays are numpy float64 / int32
in the reference layouts.  ``RigidEngine`` is the HBM-resident fast path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

MAX_NEIGH = 80  # C1:121


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def lj_params_rigid() -> N.LJParams:
    p = N.LJParams()
    N.lib().fmd_lj_params_rigid(C.byref(p))
    return p


def lj_params(sigma: float, eps: float, rcut: float, rmin2: float = 0.25) -> N.LJParams:
    sr6 = (sigma / rcut) ** 6
    return N.LJParams(sigma, eps, rcut, 4.0 * eps * (sr6 * sr6 - sr6), rmin2)


def make_npt(T: float, Pe: float, N_mol: int) -> N.NPTState:
    """make_npt (C1:480-488): Nf = 6N - 3."""
    s = N.NPTState()
    N.lib().fmd_make_npt(C.byref(s), T, Pe, 6 * N_mol - 3)
    return s


def nlist_build_sym(pos, h, hi, rmcut, max_neigh=MAX_NEIGH):
    """nlist_build_sym (LJ:257-268) -> (nl_count[N], nl_list[N, max_neigh])."""
    pos, h, hi = _c(pos).reshape(-1), _c(h), _c(hi)
    n = pos.size // 3
    cnt = np.zeros(n, np.int32)
    lst = np.zeros(n * max_neigh, np.int32)
    ovf = C.c_int(0)
    N.check(N.lib().fmd_rigid_nlist(pos, h, hi, n, rmcut, max_neigh, cnt, lst, C.byref(ovf)))
    return cnt, lst.reshape(n, max_neigh)


def forces(pos, qv, body, h, hi, nl_count, nl_list, rmcut2, lj=None, want_lab=False):
    """forces (LJ:271-400) -> (Ep, Fv[N,3], Tv[N,3], Wm9[9]) (+ lab if want_lab)."""
    pos, qv, body, h, hi = (_c(x).reshape(-1) for x in (pos, qv, body, h, hi))
    nl_count = _c(nl_count, np.int32)
    nl_list = _c(nl_list, np.int32)
    n = pos.size // 3
    natom = body.size // 3
    max_neigh = nl_list.size // n
    Fv, Tv, W = np.zeros(n * 3), np.zeros(n * 3), np.zeros(9)
    lab = np.zeros(n * natom * 3) if want_lab else None
    ep = C.c_double(0)
    N.check(N.lib().fmd_rigid_forces(Fv, Tv, W, pos, qv, body, h, hi, nl_count,
                                     nl_list.reshape(-1), max_neigh, n, natom, rmcut2,
                                     C.byref(lj) if lj is not None else None, N.ptr(lab),
                                     C.byref(ep)))
    out = (ep.value, Fv.reshape(n, 3), Tv.reshape(n, 3), W)
    return out + (lab.reshape(n, natom, 3),) if want_lab else out


def step_npt(pos, vel, qv, omg, Fv, Tv, Wm9, h, hi, body, I0, Mmol, rmcut2, dt, npt,
             nl_count, nl_list, lj=None):
    """step_npt (LJ:434-538): updates every array and ``npt`` in place like the
    reference; returns (Ep, KE).  Arrays must be C-contiguous float64."""
    for a in (pos, vel, qv, omg, Fv, Tv, Wm9, h, hi):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
            raise TypeError("step_npt updates its arrays in place: pass C-contiguous float64")
    body = _c(body).reshape(-1)
    nl_count = _c(nl_count, np.int32)
    nl_list = _c(nl_list, np.int32)
    n = pos.size // 3
    natom = body.size // 3
    ep, ke = C.c_double(0), C.c_double(0)
    N.check(N.lib().fmd_rigid_step_npt(
        pos.reshape(-1), vel.reshape(-1), qv.reshape(-1), omg.reshape(-1), Fv.reshape(-1),
        Tv.reshape(-1), Wm9, h, hi, body, I0, Mmol, n, natom, rmcut2, dt, C.byref(npt),
        nl_count, nl_list.reshape(-1), nl_list.size // n,
        C.byref(lj) if lj is not None else None, C.byref(ep), C.byref(ke)))
    return ep.value, ke.value


class RigidEngine:
    """HBM-resident rigid-body NPT engine (fmd_rigid_create / fmd_engine_*)."""

    def __init__(self, body, n_mol, I0, Mmol, rmol, dt=1.0, lj=None, nlup=25,
                 max_neigh=MAX_NEIGH, rmcut=None, skin=3.0, half_pairs=1,
                 cell_min_molecules=None):
        self.body = _c(body).reshape(-1, 3)
        cfg = N.RigidConfig()
        N.lib().fmd_rigid_config_default(C.byref(cfg), rmol)
        self._body_flat = _c(self.body).reshape(-1)
        cfg.n_mol = n_mol
        cfg.natom = len(self.body)
        cfg.body = self._body_flat.ctypes.data_as(C.POINTER(C.c_double))
        cfg.I0, cfg.Mmol, cfg.dt, cfg.nlup, cfg.max_neigh = I0, Mmol, dt, nlup, max_neigh
        cfg.skin = skin
        cfg.half_pairs = int(half_pairs)
        if cell_min_molecules is not None:
            cfg.cell_min_molecules = int(cell_min_molecules)
        if lj is not None:
            cfg.lj = lj
            cfg.rmcut = lj.rcut + 2 * rmol + 1.0
        if rmcut is not None:
            cfg.rmcut = rmcut
        self.cfg = cfg
        self.n = n_mol
        h = C.c_void_p()
        N.check(N.lib().fmd_rigid_create(C.byref(cfg), C.byref(h)))
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

    def set_state(self, pos=None, vel=None, qv=None, omg=None, h=None, npt=None):
        arrs = [None if a is None else _c(a).reshape(-1) for a in (pos, vel, qv, omg, h)]
        N.check(N.lib().fmd_rigid_set_state(self._h, *[N.ptr(a) for a in arrs],
                                            C.byref(npt) if npt is not None else None))

    def get_state(self):
        n = self.n
        pos, vel, omg = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        qv, h = np.zeros((n, 4)), np.zeros(9)
        npt = N.NPTState()
        N.check(N.lib().fmd_rigid_get_state(self._h, N.ptr(pos), N.ptr(vel), N.ptr(qv),
                                            N.ptr(omg), N.ptr(h), C.byref(npt)))
        return dict(pos=pos, vel=vel, qv=qv, omg=omg, h=h, npt=npt)

    def prepare(self):
        N.check(N.lib().fmd_engine_prepare(self._h))

    def forces(self):
        n = self.n
        F, T, W, E = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(9), np.zeros(3)
        N.check(N.lib().fmd_engine_forces(self._h, N.ptr(F), N.ptr(T), N.ptr(W), N.ptr(E)))
        return E[0], F, T, W

    def run(self, first_step, nsteps, ctrl=None, thermo=True):
        ctrl_arr = None
        if ctrl is not None:
            ctrl_arr = (N.StepCtrl * nsteps)(*ctrl)
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

    def set_graph(self, on=True):
        N.check(N.lib().fmd_engine_set_graph(self._h, 1 if on else 0))

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
