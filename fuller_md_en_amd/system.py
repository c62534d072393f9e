"""Host-side system construction (the reference's init code), through libfmd.

``load_cc1``, ``resolve_fullerene``, ``default_a0``, ``make_crystal`` and the
seeded initial states follow LJ:178-254 / 790-801 and AB:763-815; the seeded
draws come from libstdc++'s std::mt19937 + std::normal_distribution inside
libfmd.so, i.e. the same stream the reference programs draw.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from . import _native as N

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "FullereneLib")
MASS_C = 12.011


def _sig():
    L = N.lib()
    _dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.fmd_load_cc1.argtypes = [C.c_char_p, _dp, C.POINTER(C.c_int), _dp]
    L.fmd_generate_c60.argtypes = [_dp, _dp]
    L.fmd_resolve_fullerene.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int,
                                        C.c_char_p, C.c_int]
    L.fmd_default_a0.restype = C.c_double
    L.fmd_default_a0.argtypes = [C.c_double, C.c_char_p, C.c_double]
    L.fmd_make_crystal.argtypes = [C.c_char_p, C.c_double, C.c_int, _dp, _dp]
    L.fmd_rigid_init_state.argtypes = [C.c_int, C.c_uint, C.c_double, C.c_double, C.c_double,
                                       _dp, _dp, _dp]
    L.fmd_atomic_init_velocities.argtypes = [C.c_int, C.c_uint, C.c_double, _dp, _dp]
    L.fmd_normal_stream.argtypes = [C.c_uint, C.c_int, _dp]
    return L


@dataclass
class Molecule:
    label: str
    coords: np.ndarray  # natom x 3, centred
    rmol: float
    dmol: float
    I0: float
    mass: float

    @property
    def natom(self) -> int:
        return len(self.coords)


def resolve_fullerene(spec: str = "C60", libdir: str = LIBDIR) -> tuple[str, str]:
    path = C.create_string_buffer(1024)
    label = C.create_string_buffer(128)
    N.check(_sig().fmd_resolve_fullerene(spec.encode(), libdir.encode(), path, 1024, label, 128))
    return path.value.decode(), label.value.decode()


def load_cc1(path: str, label: str = "") -> Molecule:
    coords = np.zeros(84 * 3)
    props = np.zeros(4)
    n = C.c_int(0)
    N.check(_sig().fmd_load_cc1(path.encode(), coords, C.byref(n), props))
    return Molecule(label, coords[: n.value * 3].reshape(n.value, 3).copy(), *props)


def c60_analytic() -> Molecule:
    """generate_c60 (C1:228-252): the core program's analytic C60."""
    coords, props = np.zeros(180), np.zeros(4)
    N.check(_sig().fmd_generate_c60(coords, props))
    return Molecule("C60 (analytic)", coords.reshape(60, 3), *props)


def molecule(spec: str = "C60", libdir: str = LIBDIR) -> Molecule:
    path, label = resolve_fullerene(spec, libdir)
    return load_cc1(path, label)


def default_a0(dmax: float, crystal: str = "FCC", scale: float = 1.0) -> float:
    return _sig().fmd_default_a0(dmax, crystal.encode(), scale)


def n_molecules(crystal: str, nc: int) -> int:
    return (4 if crystal.upper() == "FCC" else 2) * nc ** 3


def make_crystal(crystal: str, a: float, nc: int):
    nmax = n_molecules(crystal, nc)
    pos = np.zeros(nmax * 3)
    h = np.zeros(9)
    n = _sig().fmd_make_crystal(crystal.encode(), a, nc, pos, h)
    if n < 0:
        N.check(n)
    return pos[: n * 3].reshape(n, 3), h


def rigid_init_state(n: int, seed: int, T: float, mol: Molecule):
    vel, omg, qv = np.zeros(n * 3), np.zeros(n * 3), np.zeros(n * 4)
    N.check(_sig().fmd_rigid_init_state(n, seed, T, mol.mass, mol.I0, vel, omg, qv))
    return vel.reshape(n, 3), omg.reshape(n, 3), qv.reshape(n, 4)


def atomic_init_velocities(mass: np.ndarray, seed: int, T: float):
    mass = np.ascontiguousarray(mass, dtype=np.float64)
    vel = np.zeros(len(mass) * 3)
    N.check(_sig().fmd_atomic_init_velocities(len(mass), seed, T, mass, vel))
    return vel.reshape(-1, 3)


def normal_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    N.check(_sig().fmd_normal_stream(seed, n, out))
    return out


def rigid_system(nc: int = 3, spec: str = "C60", crystal: str = "FCC", T: float = 298.0,
                 seed: int = 42, init_scale: float = 1.0, libdir: str = LIBDIR):
    """The rigid program's initial system (LJ:713-801)."""
    mol = molecule(spec, libdir)
    a0 = default_a0(mol.dmol, crystal, init_scale)
    pos, h = make_crystal(crystal, a0, nc)
    vel, omg, qv = rigid_init_state(len(pos), seed, T, mol)
    return dict(mol=mol, a0=a0, pos=pos, h=h, vel=vel, omg=omg, qv=qv)


def core_system(nc: int = 3, seed: int = 42, T: float = 300.0, a0: float = 14.17):
    """The core program's initial system (C1:640-680): analytic C60, FCC at
    a0 = 14.17 A, seeded velocities / orientations at 300 K."""
    mol = c60_analytic()
    pos, h = make_crystal("FCC", a0, nc)
    vel, omg, qv = rigid_init_state(len(pos), seed, T, mol)
    return dict(mol=mol, a0=a0, pos=pos, h=h, vel=vel, omg=omg, qv=qv)


def airebo_system(nc: int = 3, spec: str = "C60", crystal: str = "FCC", T: float = 300.0,
                  seed: int = 42, init_scale: float = 1.0, jitter: float = 0.0,
                  jitter_seed: int = 12345, libdir: str = LIBDIR):
    """The AIREBO program's initial system (AB:763-815): atoms at lattice
    point + .cc1 coordinates (no rotation), mol_id = molecule + 1, carbon
    masses, per-atom Maxwell velocities from `seed`; optional uniform jitter
    of +-jitter A per coordinate from its own generator."""
    mol = molecule(spec, libdir)
    a0 = default_a0(mol.dmol, crystal, init_scale)
    cen, h = make_crystal(crystal, a0, nc)
    pos = (cen[:, None, :] + mol.coords[None]).reshape(-1, 3)
    if jitter:
        pos = pos + np.random.default_rng(jitter_seed).uniform(-jitter, jitter, pos.shape)
    mol_id = np.repeat(np.arange(1, len(cen) + 1), mol.natom).astype(np.int32)
    mass = np.full(len(pos), MASS_C)
    vel = atomic_init_velocities(mass, seed, T)
    return dict(mol=mol, a0=a0, pos=pos, h=h, mol_id=mol_id, mass=mass, vel=vel,
                n_mol=len(cen))
