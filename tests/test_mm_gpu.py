"""Molecular mechanics (bonded terms + intermolecular LJ + Coulomb) on the
GPU against the reference build's golden vectors and the FP64 oracle.

Tolerances: forces / virial max|a-b| / max|b| <= 1e-5 (the LJ pair term is
FP32), energies relative <= 1e-5, lists exact.  The bonded terms alone are
FP64 in the reference's operation order: against the oracle to 1e-12.
"""
import numpy as np
import pytest

import oracle
from conftest import golden, rel_max
from fuller_md_en_amd import _native as N
from fuller_md_en_amd import mm, system

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _ft_from_golden(g, nmol):
    return mm.build_flat_topology(g["coords"], g["bonds"], nmol, *g["k4"])


def _otopo(ft):
    """The oracle's dict form of a FlatTopology."""
    a = ft.a
    return dict(b=np.c_[a["b_i0"], a["b_i1"]], bp=np.c_[a["b_kb"], a["b_r0"]],
                a=np.c_[a["ang_i0"], a["ang_i1"], a["ang_i2"]],
                ap=np.c_[a["ang_kth"], a["ang_th0"]],
                d=np.c_[a["dih_i0"], a["dih_i1"], a["dih_i2"], a["dih_i3"]],
                dp=np.c_[a["dih_Vn"], a["dih_gamma"]], dm=a["dih_mult"],
                i=np.c_[a["imp_i0"], a["imp_i1"], a["imp_i2"], a["imp_i3"]],
                ip=np.c_[a["imp_ki"], a["imp_gamma"]], mol_id=a["mol_id"])


def test_lj_list_equals_reference():
    g = golden("mm_c60_nc2.npz")
    cl, ll = mm.build_nlist_lj(g["pos"], g["h"], oracle.inv9(g["h"]), g["mol_id"])
    assert (cl == g["nlc"]).all()
    last = np.array([ll[i, cl[i] - 1] if cl[i] else -1 for i in range(len(cl))])
    assert (last == g["lj_last"]).all()
    ocl, oll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    for i in range(len(cl)):
        assert (ll[i, :cl[i]] == oll[i, :cl[i]]).all()


@pytest.mark.parametrize("name", ["mm_c60_nc1.npz", "mm_c60_nc2.npz", "mm_c70_hcp.npz"])
def test_compute_forces_compat_matches_reference(name):
    g = golden(name)
    print(f"case {name}", flush=True)
    nmol = int(g["mol_id"].max()) + 1
    src = g if "coords" in g else golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(src, nmol)
    cl, ll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    E, F, W = mm.compute_forces(g["pos"], g["h"], oracle.inv9(g["h"]), ft, cl, ll)
    assert rel_max(F, g["F"]) < TOL, rel_max(F, g["F"])
    assert rel_max(W, g["W"]) < TOL
    assert np.count_nonzero(W[[1, 2, 3, 5, 6, 7]]) == 0  # diagonal virial, as the reference's
    for k in range(4):  # bonded energies: FP64 in the reference's order
        assert rel(E[k], g["E"][k]) < 1e-11, (k, E[k], g["E"][k])
    for k in (4, 6):
        assert rel(E[k], g["E"][k]) < TOL
    assert E[5] == 0.0


def test_bonded_forces_fp64_against_oracle():
    """One molecule alone (no LJ pairs): bonded forces and energies in FP64
    with the reference's operation order, compat and physical."""
    g = golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(g, 1)
    rng = np.random.default_rng(4)
    pos = g["coords"] + 20.0 + rng.uniform(-0.05, 0.05, g["coords"].shape)
    h = np.diag([40.0] * 3).reshape(-1)
    cl = np.zeros(60, np.int32)
    ll = np.zeros((60, 400), np.int32)
    for mode in (mm.MM_COMPAT, mm.MM_PHYSICAL):
        E, F, W = mm.compute_forces(pos, h, oracle.inv9(h), ft, cl, ll, mode=mode)
        E0, F0, W0 = oracle.mm_forces(pos, h, _otopo(ft), cl, ll, mode=mode)
        assert rel_max(F, F0) < 1e-12, rel_max(F, F0)
        assert rel_max(W, W0) < 1e-12
        assert rel_max(E, E0) < 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_compute_forces_with_charges_against_oracle(mode):
    g = golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(g, 4)
    cl, ll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    E, F, W = mm.compute_forces(g["pos"], g["h"], oracle.inv9(g["h"]), ft, cl, ll,
                                charge=g["charge"], mode=mode)
    E0, F0, W0 = oracle.mm_forces(g["pos"], g["h"], _otopo(ft), cl, ll, g["charge"], mode)
    if mode == 0:
        assert rel_max(F, g["F_q"]) < TOL and rel_max(W, g["W_q"]) < TOL
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    assert abs(E0[5]) > 1e-3
    for k in range(7):
        assert rel(E[k], E0[k]) < TOL, (k, E[k], E0[k])


def test_mixed_species_topology_against_oracle():
    """Two C60 and two C70 molecules in one flat topology (templates shared
    per species), physical mode, against the oracle."""
    parts, pos, mol, off = [], [], [], 0
    for k, spec in enumerate(["C60", "C70", "C60", "C70"]):
        path, _ = system.resolve_fullerene(spec)
        m = system.load_cc1(path)
        ft1 = mm.build_flat_topology(m.coords, mm.load_cc1_bonds(path), 1)
        parts.append((ft1, off))
        pos.append(m.coords + np.array([(k % 2) * 11.0, (k // 2) * 11.0, 5.0]))
        mol += [k] * m.natom
        off += m.natom
    na = off
    arr = {}
    for key in mm.FlatTopology.KEYS:
        if key == "mol_id":
            arr[key] = np.array(mol, np.int32)
            continue
        chunks = []
        for ft1, o in parts:
            a = ft1[key]
            is_idx = a.dtype == np.int32 and key != "dih_mult"
            chunks.append(a + o if is_idx else a)
        arr[key] = np.concatenate(chunks)
    counts = (len(arr["b_i0"]), len(arr["ang_i0"]), len(arr["dih_i0"]), len(arr["imp_i0"]))
    ft = mm.FlatTopology(counts, na, arr)
    rng = np.random.default_rng(8)
    pos = np.concatenate(pos) + rng.uniform(-0.03, 0.03, (na, 3))
    h = np.diag([22.0, 22.0, 16.0]).reshape(-1)
    cl, ll = oracle.mm_nlist(pos, h, arr["mol_id"])
    E, F, W = mm.compute_forces(pos, h, oracle.inv9(h), ft, cl, ll, mode=mm.MM_PHYSICAL)
    E0, F0, W0 = oracle.mm_forces(pos, h, _otopo(ft), cl, ll, mode=oracle.MM_PHYSICAL)
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for k in range(7):
        assert rel(E[k], E0[k]) < TOL


def test_term_across_molecules_is_rejected():
    g = golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(g, 4)
    ft["b_i1"][0] = 61  # bond 0 now joins molecules 0 and 1
    cl, ll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    with pytest.raises(N.FmdError, match="spans two molecules"):
        mm.compute_forces(g["pos"], g["h"], oracle.inv9(g["h"]), ft, cl, ll)


def test_step_npt_dropin_matches_reference():
    g = golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(g, 4)
    cl, ll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    P, V, F, W = g["pos"].copy(), g["vel"].copy(), g["F"].copy(), g["W"].copy()
    h, hi = g["h"].copy(), oracle.inv9(g["h"])
    npt = mm.NPTState.from_array(g["npt0"])
    traj = np.array([mm.step_npt(P, V, F, W, h, hi, float(g["dt"]), npt, ft, cl, ll)
                     for _ in range(5)])
    for k in (0, 1, 2, 3, 4, 6, 7):
        assert rel_max(traj[:, k], g["traj"][:, k]) < TOL, k
    assert np.abs(P - g["pos_end"]).max() < 1e-9
    assert rel_max(V, g["vel_end"]) < 1e-5
    assert rel_max(F, g["F_end"]) < TOL
    assert np.abs(h - g["h_end"]).max() < 1e-11


def _engine(s, mode=mm.MM_COMPAT, vel=None, T=298.0, charge=None):
    from fuller_md_en_amd import airebo
    eng = mm.MMEngine(s["ft"], mode=mode, charge=charge)
    eng.set_state(pos=s["pos"], vel=np.zeros_like(s["pos"]) if vel is None else vel, h=s["h"],
                  npt=airebo.make_npt(T, 0.0, len(s["pos"])))
    return eng


def _wrapped(s):
    p = s["pos"].reshape(-1).copy()
    oracle.lib().orc_apply_pbc(p, s["h"], oracle.inv9(s["h"]), len(s["pos"]))
    s = dict(s)
    s["pos"] = p.reshape(-1, 3)
    return s


@pytest.mark.parametrize("spec,lattice,nc,mode", [("C60", "FCC", 3, 0), ("C60", "FCC", 3, 1),
                                                  ("C70", "HCP", 3, 1), ("C84:20", "BCC", 2, 0)])
def test_engine_forces_against_oracle(spec, lattice, nc, mode):
    """Resident engine (molecule-image LJ path where the cell allows) against
    the oracle on the same truncated half list."""
    print(f"case {spec} {lattice} nc={nc} mode={mode}", flush=True)
    s = _wrapped(mm.mm_system(nc, spec, lattice, jitter=0.03, jitter_seed=3))
    cl, ll = oracle.mm_nlist(s["pos"], s["h"], s["mol_id"])
    E0, F0, W0 = oracle.mm_forces(s["pos"], s["h"], _otopo(s["ft"]), cl, ll, mode=mode)
    eng = _engine(s, mode=mode)
    eng.prepare()
    E3, F, W = eng.forces()
    E = eng.energies()
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for k in range(7):
        assert rel(E[k], E0[k]) < TOL, (k, E[k], E0[k])
    assert rel(E3[0], E0[6]) < TOL and rel(E3[1], E0[:4].sum()) < TOL
    assert eng.counters()[0] == 0
    eng.close()


def test_engine_trajectory_matches_reference():
    g = golden("mm_c60_nc1.npz")
    ft = _ft_from_golden(g, 4)
    eng = mm.MMEngine(ft)
    eng.set_state(pos=g["pos"], vel=g["vel"], h=g["h"], npt=mm.NPTState.from_array(g["npt0"]))
    th = eng.run(0, 5)
    got = np.array([[t.Eb, t.Ea, t.Ed, t.Ei, t.Ep_lj, t.Ecoul, t.Ep, t.KE] for t in th])
    for k in (0, 1, 2, 3, 4, 6, 7):
        assert rel_max(got[:, k], g["traj"][:, k]) < TOL, k
    st = eng.get_state()
    assert np.abs(st["pos"] - g["pos_end"]).max() < 1e-9
    assert rel_max(st["vel"], g["vel_end"]) < 1e-5
    eng.close()


def test_engine_physical_trajectory_against_oracle():
    """30 steps of the physical field (list rebuilt at step 20) against the
    oracle's step_npt with the same rebuild schedule; the physical field keeps
    the temperature bounded, where the reference's forces heat the crystal."""
    s = _wrapped(mm.mm_system(2, "C60", "FCC", T=300.0, jitter=0.01, jitter_seed=9))
    eng = _engine(s, mode=mm.MM_PHYSICAL, vel=s["vel"], T=300.0)
    th = eng.run(0, 30)
    P, V, h = s["pos"].copy(), s["vel"].copy(), s["h"].copy()
    topo = _otopo(s["ft"])
    cl, ll = oracle.mm_nlist(P, h, s["mol_id"])
    _, F, W = oracle.mm_forces(P, h, topo, cl, ll, mode=oracle.MM_PHYSICAL)
    npt = oracle.NPT()
    oracle.lib().orc_make_npt(npt, 300.0, 0.0, 3 * len(P) - 3)
    ref = []
    for k in range(1, 31):
        if k % 20 == 0:
            cl, ll = oracle.mm_nlist(P, h, s["mol_id"])
        P, V, F, W, h, npt, out = oracle.mm_step_npt(P, V, F, W, h, npt, 0.1, topo, cl, ll,
                                                     mode=oracle.MM_PHYSICAL)
        ref.append(out)
    ref = np.array(ref)
    got = np.array([[t.Eb, t.Ea, t.Ed, t.Ei, t.Ep_lj, t.Ecoul, t.Ep, t.KE] for t in th])
    for k in (0, 1, 2, 3, 4, 6, 7):
        assert rel_max(got[:, k], ref[:, k]) < 1e-4, k
    st = eng.get_state()
    assert np.abs(st["pos"] - P).max() < 1e-6
    assert max(t.T for t in th) < 600.0
    eng.close()
