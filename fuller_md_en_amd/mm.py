"""Molecular-mechanics path (bonds, angles, dihedrals, impropers within each
molecule; shifted LJ and cut Coulomb between molecules), computed on the GPU.

The functions keep the names and argument meaning of the reference's
(src/fuller_LJ_npt_mmmd_serial_omp_acc.cpp "MM"): load_cc1 (with its bond
list), build_flat_topology, build_nlist_lj, compute_forces, step_npt.
``MMEngine`` is the HBM-resident fast path.  ``mode``: MM_COMPAT (the
reference's forces, which are not the gradient of its energy, and its
diagonal virial) or MM_PHYSICAL (exact gradient, full virial).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from . import system
from .airebo import AtomicEngine
from ._native import NPTState  # noqa: F401

MM_COMPAT = 0
MM_PHYSICAL = 1
MAX_LJ_NEIGH = 400  # MM:119
KCAL2EV = 1.0 / 23.06054783  # MM:103
# --ff_kb / --ff_kth / --ff_v2 / --ff_kimp defaults in kcal/mol (MM:1039-1042)
FF_DEFAULTS_KCAL = (469.0, 63.0, 14.5, 15.0)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def lj_params():
    p = N.LJParams()
    N.lib().fmd_lj_params_mm(C.byref(p))
    return p


def load_cc1_bonds(path: str) -> np.ndarray:
    """The bonds of a .cc1 file as load_cc1 collects them (MM:160-183):
    (nbond, 2) 0-based pairs i < j, ascending."""
    cap = 1024
    out = np.zeros(2 * cap, np.int32)
    nb = C.c_int(0)
    N.check(N.lib().fmd_load_cc1_bonds(path.encode(), out, cap, C.byref(nb)))
    return out[: 2 * nb.value].reshape(nb.value, 2).copy()


class FlatTopology:
    """The flat topology arrays of build_flat_topology (MM:226-237), held as
    numpy arrays; ``struct`` is the fmd_mm_topology view of them."""

    KEYS = ("b_i0", "b_i1", "b_kb", "b_r0", "ang_i0", "ang_i1", "ang_i2", "ang_kth",
            "ang_th0", "dih_i0", "dih_i1", "dih_i2", "dih_i3", "dih_Vn", "dih_mult",
            "dih_gamma", "imp_i0", "imp_i1", "imp_i2", "imp_i3", "imp_ki", "imp_gamma",
            "mol_id")

    def __init__(self, counts, natom_total, arrays=None):
        nb, na, nd, ni = counts
        size = {"b": nb, "ang": na, "dih": nd, "imp": ni}
        self.a = {}
        for k in self.KEYS:
            n = natom_total if k == "mol_id" else size[k.split("_")[0]]
            is_int = k == "mol_id" or k == "dih_mult" or k[-2] == "i" and k[-1].isdigit()
            dt = np.int32 if is_int else np.float64
            self.a[k] = (_c(arrays[k], dt).reshape(-1) if arrays is not None
                         else np.zeros(n, dt))
        self.counts = (nb, na, nd, ni)
        self.struct = self._view()

    def _view(self):
        s = N.MMTopology()
        s.Nb, s.Nang, s.Ndih, s.Nimp = self.counts
        for k in self.KEYS:
            a = self.a[k]
            setattr(s, k, a.ctypes.data_as(N._IP if a.dtype == np.int32 else N._DP))
        return s

    def __getitem__(self, k):
        return self.a[k]


def build_flat_topology(coords, bonds, Nmol, kb=None, kth=None, v2_dih=None, k_imp=None):
    """build_flat_topology (MM:254-334) of Nmol copies of a molecule; force
    constants in eV units (default: the reference's defaults)."""
    d = [x * KCAL2EV for x in FF_DEFAULTS_KCAL]
    k = [d[i] if v is None else v for i, v in enumerate((kb, kth, v2_dih, k_imp))]
    coords = _c(coords).reshape(-1)
    bonds = _c(bonds, np.int32).reshape(-1)
    natom = coords.size // 3
    cnt = np.zeros(4, np.int32)
    N.check(N.lib().fmd_mm_topology_counts(natom, bonds, bonds.size // 2, cnt))
    ft = FlatTopology(tuple(int(x) * Nmol for x in cnt), natom * Nmol)
    N.check(N.lib().fmd_mm_build_topology(coords, natom, bonds, bonds.size // 2, Nmol, *k,
                                          C.byref(ft.struct)))
    return ft


def build_nlist_lj(pos, h, hi, mol_id):
    """build_nlist_lj (MM:337-349) -> (nlc[Na], nll[Na, 400])."""
    pos = _c(pos).reshape(-1)
    na = pos.size // 3
    nlc = np.zeros(na, np.int32)
    nll = np.zeros(na * MAX_LJ_NEIGH, np.int32)
    N.check(N.lib().fmd_mm_nlist_lj(pos, _c(h), _c(hi), na, _c(mol_id, np.int32), nlc, nll))
    return nlc, nll.reshape(na, MAX_LJ_NEIGH)


def _charge(charge, na):
    if charge is None:
        return None, 0
    q = _c(charge)
    return q, int(bool(np.any(np.abs(q) > 1e-10)))


def compute_forces(pos, h, hi, ft: FlatTopology, nlc, nll, charge=None, mol_id=None,
                   mode=MM_COMPAT):
    """compute_forces (MM:351-788) -> (out7 {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot},
    F[Na,3], vir9[9])."""
    pos = _c(pos).reshape(-1)
    na = pos.size // 3
    mol = _c(ft["mol_id"] if mol_id is None else mol_id, np.int32)
    q, has_q = _charge(charge, na)
    F, W, out = np.zeros(na * 3), np.zeros(9), np.zeros(7)
    N.check(N.lib().fmd_mm_compute_forces(F, W, pos, _c(h), _c(hi), C.byref(ft.struct),
                                          _c(nlc, np.int32), _c(nll, np.int32).reshape(-1),
                                          N.ptr(q), mol, has_q, na, mode, out))
    return out, F.reshape(na, 3), W


def step_npt(pos, vel, F, vir9, h, hi, dt, npt, ft: FlatTopology, nlc, nll, charge=None,
             mol_id=None, mode=MM_COMPAT):
    """step_npt (MM:812-890): in place like the reference; returns
    out8 {Eb, Ea, Ed, Ei, Elj, Ecoul, Etot, KE}."""
    for a in (pos, vel, F, vir9, h, hi):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.c_contiguous):
            raise TypeError("step_npt updates its arrays in place: pass C-contiguous float64")
    na = pos.size // 3
    mol = _c(ft["mol_id"] if mol_id is None else mol_id, np.int32)
    q, has_q = _charge(charge, na)
    out = np.zeros(8)
    N.check(N.lib().fmd_mm_step_npt(pos.reshape(-1), vel.reshape(-1), F.reshape(-1), vir9, h, hi,
                                    na, dt, C.byref(npt), C.byref(ft.struct),
                                    _c(nlc, np.int32), _c(nll, np.int32).reshape(-1), N.ptr(q),
                                    mol, has_q, mode, out))
    return out


class MMEngine(AtomicEngine):
    """HBM-resident molecular-mechanics NPT engine (fmd_mm_create /
    fmd_engine_*): the atomic engine's state, step loop and LJ with the MM
    bonded field as the intramolecular term."""

    def __init__(self, ft: FlatTopology, mol_id=None, dt=0.1, mode=MM_COMPAT, charge=None,
                 lj_max_neigh=MAX_LJ_NEIGH, nlup=20):
        self._ft = ft
        self._mol = _c(ft["mol_id"] if mol_id is None else mol_id, np.int32)
        self._q = None if charge is None else _c(charge)
        cfg = N.MMConfig()
        N.lib().fmd_mm_config_default(C.byref(cfg))
        cfg.n_atoms = len(self._mol)
        cfg.mol_id = self._mol.ctypes.data_as(N._IP)
        cfg.ft = C.pointer(ft.struct)
        cfg.charge = self._q.ctypes.data_as(N._DP) if self._q is not None else None
        cfg.dt, cfg.mm_mode, cfg.lj_max_neigh, cfg.nlup = dt, mode, lj_max_neigh, nlup
        self.cfg = cfg
        self.n = len(self._mol)
        h = C.c_void_p()
        N.check(N.lib().fmd_mm_create(C.byref(cfg), C.byref(h)))
        self._h = h

    def energies(self):
        """{Eb, Ea, Ed, Ei, Elj, Ecoul, Etot} of the last force evaluation."""
        out = np.zeros(7)
        N.check(N.lib().fmd_mm_energies(self._h, out))
        return out


def mm_system(nc: int = 3, spec: str = "C60", crystal: str = "FCC", T: float = 298.0,
              seed: int = 42, init_scale: float = 1.0, jitter: float = 0.0,
              jitter_seed: int = 12345, libdir: str = system.LIBDIR, k4=None):
    """The MM program's initial system (MM:1085-1141): atoms at lattice point +
    .cc1 coordinates, mol_id = molecule (0-based), its flat topology, carbon
    masses, per-atom Maxwell velocities from `seed`; optional uniform jitter
    of +-jitter A per coordinate from its own generator."""
    path, label = system.resolve_fullerene(spec, libdir)
    mol = system.load_cc1(path, label)
    bonds = load_cc1_bonds(path)
    a0 = system.default_a0(mol.dmol, crystal, init_scale)
    cen, h = system.make_crystal(crystal, a0, nc)
    pos = (cen[:, None, :] + mol.coords[None]).reshape(-1, 3)
    if jitter:
        pos = pos + np.random.default_rng(jitter_seed).uniform(-jitter, jitter, pos.shape)
    ft = build_flat_topology(mol.coords, bonds, len(cen), *(k4 or ()))
    mass = np.full(len(pos), system.MASS_C)
    vel = system.atomic_init_velocities(mass, seed, T)
    return dict(mol=mol, bonds=bonds, a0=a0, pos=pos, h=h, mol_id=ft["mol_id"].copy(), ft=ft,
                vel=vel, n_mol=len(cen))
