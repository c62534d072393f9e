"""AIREBO (REBO bond order + intermolecular LJ) on the GPU against the
reference build's golden vectors and the FP64 oracle.

Tolerances: forces / virial max|a-b| / max|b| <= 1e-5, energies relative
<= 1e-5, lists exact.  Cases: 1x1x1 and 2x2x2 C60 crystals jittered by
+-0.01 A (per-pair images: the cells are narrower than the LJ reach), and
3x3x3 (molecule-image path) built on the fly.
"""
import numpy as np
import pytest

import oracle
from conftest import golden, rel_max, unpack_list
from fuller_md_en_amd import airebo

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return abs(a - b) / abs(b)


def crystal(nc, seed, jitter):
    """Jittered C60 FCC crystal the way the AIREBO program builds it."""
    from fuller_md_en_amd import system
    mol = system.molecule("C60")
    a0 = system.default_a0(mol.dmol, "FCC", 1.0)
    cen, h = system.make_crystal("FCC", a0, nc)
    pos = (cen[:, None, :] + mol.coords[None]).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape)
    mol_id = np.repeat(np.arange(1, len(cen) + 1), mol.natom).astype(np.int32)
    p = pos.reshape(-1).copy()
    oracle.lib().orc_apply_pbc(p, h, oracle.inv9(h), len(pos))
    return p.reshape(-1, 3), h, mol_id


def test_rebo_and_lj_lists_equal_reference():
    g = golden("airebo_nc1.npz")
    hi = oracle.inv9(g["h"])
    cr, lr = airebo.build_nlist_rebo(g["pos"], g["h"], hi, g["mol_id"])
    cl, ll = airebo.build_nlist_lj(g["pos"], g["h"], hi, g["mol_id"])
    assert (cr == g["nlc_r"]).all() and (cl == g["nlc_l"]).all()
    rr = unpack_list(g["nlc_r"], g["nll_r"], 12)
    rl = unpack_list(g["nlc_l"], g["nll_l"], 400)
    for i in range(len(cr)):  # entries past the count are unspecified
        assert (lr[i, :cr[i]] == rr[i, :cr[i]]).all()
        assert (ll[i, :cl[i]] == rl[i, :cl[i]]).all()


def test_lj_truncated_list_equals_reference():
    g = golden("airebo_nc2.npz")
    cl, ll = airebo.build_nlist_lj(g["pos"], g["h"], oracle.inv9(g["h"]), g["mol_id"])
    assert (cl == g["nlc_l"]).all()
    last = np.array([ll[i, cl[i] - 1] if cl[i] else -1 for i in range(len(cl))])
    assert (last == g["lj_last"]).all()
    ocl, oll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    assert (ll == oll).all()


def test_lj_half_list_wide_cell_equals_oracle():
    """3x3x3 (cell wide enough for the FP32-prefiltered walk through the
    molecule tiles): the truncated half list equals the oracle's, entry for
    entry."""
    pos, h, mol = crystal(3, 23, 0.1)
    cl, ll = airebo.build_nlist_lj(pos, h, oracle.inv9(h), mol)
    ocl, oll = oracle.nlist_lj(pos, h, mol)
    assert (cl == ocl).all()
    for i in range(len(cl)):
        assert (ll[i, :cl[i]] == oll[i, :cl[i]]).all()
    assert (cl == 400).any()  # truncation active


@pytest.mark.parametrize("name", ["airebo_nc1.npz", "airebo_nc2.npz"])
def test_compute_forces_compat_matches_reference(name):
    g = golden(name)
    print(f"case {name}", flush=True)
    cr, lr = oracle.nlist_rebo(g["pos"], g["h"], g["mol_id"])
    cl, ll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    E, F, W = airebo.compute_forces(g["pos"], g["h"], oracle.inv9(g["h"]), (cr, lr, cl, ll),
                                    g["mol_id"], airebo.REBO_COMPAT)
    assert rel_max(F, g["F"]) < TOL
    assert rel_max(W, g["W"]) < TOL
    for a, b in zip(E, g["E"]):
        assert rel(a, b) < TOL


@pytest.mark.parametrize("nc,jit", [(2, 0.01), (2, 0.1)])
def test_compute_forces_physical_matches_oracle(nc, jit):
    pos, h, mol = crystal(nc, 31 + nc, jit)
    lists = (*oracle.nlist_rebo(pos, h, mol), *oracle.nlist_lj(pos, h, mol))
    E0, F0, W0 = oracle.airebo_forces(pos, h, lists, oracle.REBO_PHYSICAL)
    E, F, W = airebo.compute_forces(pos, h, oracle.inv9(h), lists, mol, airebo.REBO_PHYSICAL)
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for a, b in zip(E, E0):
        assert rel(a, b) < TOL


def test_step_npt_dropin_matches_reference():
    g = golden("airebo_nc1.npz")
    cr, lr = oracle.nlist_rebo(g["pos"], g["h"], g["mol_id"])
    cl, ll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    P, V, F, W = g["pos"].copy(), g["vel"].copy(), g["F"].copy(), g["W"].copy()
    h, hi = g["h"].copy(), oracle.inv9(g["h"])
    npt = airebo.NPTState.from_array(g["npt0"])
    traj = []
    for _ in range(5):
        traj.append(airebo.step_npt(P, V, F, W, h, hi, g["mass"], float(g["dt"]), npt,
                                    (cr, lr, cl, ll), g["mol_id"]))
    traj = np.array(traj)
    for k in range(4):
        assert rel_max(traj[:, k], g["traj"][:, k]) < TOL
    assert np.abs(P - g["pos_end"]).max() < 1e-8
    assert rel_max(V, g["vel_end"]) < 1e-5
    assert rel_max(F, g["F_end"]) < TOL
    assert np.abs(h - g["h_end"]).max() < 1e-10


def _engine(pos, h, mol, vel=None, mode=airebo.REBO_COMPAT, T=300.0):
    mass = np.full(len(pos), 12.011)
    eng = airebo.AtomicEngine(mol, mass, mode=mode)
    eng.set_state(pos=pos, vel=np.zeros_like(pos) if vel is None else vel, h=h,
                  npt=airebo.make_npt(T, 0.0, len(pos)))
    return eng


@pytest.mark.parametrize("mode", [0, 1])
def test_engine_forces_nc3_against_oracle(mode):
    """3x3x3 (6480 atoms): molecule-image LJ path, truncation active."""
    pos, h, mol = crystal(3, 7, 0.01)
    lists = (*oracle.nlist_rebo(pos, h, mol), *oracle.nlist_lj(pos, h, mol))
    E0, F0, W0 = oracle.airebo_forces(pos, h, lists, mode)
    eng = _engine(pos, h, mol, mode=mode)
    eng.prepare()
    E, F, W = eng.forces()
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for a, b in zip(E, E0):
        assert rel(a, b) < TOL
    assert eng.counters()[0] == 0
    eng.close()


@pytest.mark.parametrize("spec,lattice,nc", [("C70", "HCP", 3), ("C60", "BCC", 3),
                                             ("C84:20", "HCP", 2)])
def test_engine_forces_other_lattices_against_oracle(spec, lattice, nc):
    """HCP and BCC crystals of other species (the reference's lattice and .cc1
    builders, +-0.02 A jitter), exact-gradient REBO + LJ with the truncated
    half list, against the FP64 oracle on the same lists."""
    from fuller_md_en_amd import system
    print(f"case {spec} {lattice} nc={nc}", flush=True)
    s = system.airebo_system(nc, spec, lattice, jitter=0.02, jitter_seed=5)
    p = s["pos"].reshape(-1).copy()
    oracle.lib().orc_apply_pbc(p, s["h"], oracle.inv9(s["h"]), len(s["pos"]))
    pos, h, mol = p.reshape(-1, 3), s["h"], s["mol_id"]
    lists = (*oracle.nlist_rebo(pos, h, mol), *oracle.nlist_lj(pos, h, mol))
    E0, F0, W0 = oracle.airebo_forces(pos, h, lists, oracle.REBO_PHYSICAL)
    eng = _engine(pos, h, mol, mode=airebo.REBO_PHYSICAL)
    eng.prepare()
    E, F, W = eng.forces()
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for a, b in zip(E, E0):
        assert rel(a, b) < TOL
    assert eng.counters()[0] == 0
    eng.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_rebo_lds_and_global_paths_agree(mode):
    """The REBO kernel stages a molecule's bond entries in LDS when its list
    has at most two entries per thread and its bonds fit the LDS budget, and
    works from the global entry arrays otherwise.  A 2.6 A REBO list (nine
    entries per C60 atom, 540 per molecule against 384 gathered by 192
    threads; the extra ones beyond the 2.0 A switch carry no force) sends
    every molecule down the global path; forces, virial and energies must
    equal those of the 2.3 A list (LDS path) up to summation order, and both
    the oracle's."""
    pos, h, mol = crystal(2, 11, 0.02)
    lists = (*oracle.nlist_rebo(pos, h, mol), *oracle.nlist_lj(pos, h, mol))
    E0, F0, W0 = oracle.airebo_forces(pos, h, lists, mode)
    out = []
    for rl in (2.3, 2.6):
        print(f"case mode={mode} rebo_rlist={rl}", flush=True)
        eng = airebo.AtomicEngine(mol, np.full(len(pos), 12.011), mode=mode, rebo_rlist=rl)
        eng.set_state(pos=pos, vel=np.zeros_like(pos), h=h,
                      npt=airebo.make_npt(300.0, 0.0, len(pos)))
        eng.prepare()
        out.append(eng.forces())
        assert eng.counters()[0] == 0
        eng.close()
    (Ea, Fa, Wa), (Eb, Fb, Wb) = out
    assert rel_max(Fa, Fb) < 1e-12 and rel_max(Wa, Wb) < 1e-12
    for a, b in zip(Ea, Eb):
        assert rel(a, b) < 1e-12
    assert rel_max(Fa, F0) < TOL and rel_max(Wa, W0) < TOL


@pytest.mark.parametrize("slots,threads", [(3, 192), (6, 256), (8, 192)])
def test_rebo_kernel_parameters_keep_forces(slots, threads):
    """The REBO kernel's tunables -- bond entries per atom held in LDS and the
    workgroup size -- change which threads sum which energy and virial shares,
    never an atom's force: forces equal the default's bit for bit, energies and
    virial to summation order.  Three slots hold a C60's three bonds per atom
    exactly."""
    pos, h, mol = crystal(2, 11, 0.02)
    out = []
    for sl, th in ((None, None), (slots, threads)):
        print(f"case rebo_slots={sl} rebo_threads={th}", flush=True)
        eng = airebo.AtomicEngine(mol, np.full(len(pos), 12.011), mode=airebo.REBO_PHYSICAL,
                                  rebo_slots=sl, rebo_threads=th)
        eng.set_state(pos=pos, vel=np.zeros_like(pos), h=h,
                      npt=airebo.make_npt(300.0, 0.0, len(pos)))
        eng.prepare()
        out.append(eng.forces())
        eng.close()
    (Ea, Fa, Wa), (Eb, Fb, Wb) = out
    assert np.array_equal(Fa, Fb) and rel_max(Wa, Wb) < 1e-12
    for a, b in zip(Ea, Eb):
        assert rel(a, b) < 1e-12


def _frozen_npt(n):
    """NPT state whose thermostat and barostat masses (1e30) keep xi and the
    cell velocity where they are set."""
    npt = airebo.make_npt(300.0, 0.0, n)
    npt.Q = npt.W = 1e30
    return npt


@pytest.mark.parametrize("spec,expand", [("C60", 1.1), ("C70", 1.0)])
def test_stale_lj_lists_are_rebuilt_within_the_step(spec, expand):
    """The per-atom LJ lists hold the pairs within rc + 1 A at their build.
    Molecules of a 3x3x3 crystal (C60: lattice constant 1.1 times the
    default; C70, whose default is wider: 1.0)
    given random rigid velocities of up to 0.05 A/fs per component (cell and
    thermostat frozen, REBO bonds unchanged) close in on each other by more
    than that skin before the scheduled rebuild at step 20: the engine
    rebuilds the lists within the steps where a pair may have closed the
    skin (counted with the rebuilds), so its forces at step 19 equal the
    oracle's on complete lists of that state, while the pairs within
    rc + 1 A of the initial state miss some (6e-4 of the largest force).  At
    1/20 of the speed nothing is rebuilt.  LJ cut at 5 A, where the pairs
    entering the cutoff carry forces well above the tolerance; untruncated
    lists (lj_max_neigh = 0), so that both sides mean every pair within rc.
    C70 takes two atoms per lane, whose lists are rebuilt by a kernel of
    their own before the walk."""
    from fuller_md_en_amd import rigid, system
    m0 = system.molecule(spec)
    cen, h = system.make_crystal("FCC", expand * system.default_a0(m0.dmol, "FCC", 1.0), 3)
    pos = (cen[:, None, :] + m0.coords[None]).reshape(-1, 3)
    pos = pos + np.random.default_rng(7).uniform(-0.01, 0.01, pos.shape)
    p = pos.reshape(-1).copy()
    oracle.lib().orc_apply_pbc(p, h, oracle.inv9(h), len(pos))
    pos = p.reshape(-1, 3)
    mol = np.repeat(np.arange(1, len(cen) + 1), m0.natom).astype(np.int32)
    n = len(pos)
    lj, lj_o = rigid.lj_params(3.4, 0.00284, 5.0), oracle.lj_params_custom(3.4, 0.00284, 5.0)
    vm = np.random.default_rng(3).uniform(-0.05, 0.05, (mol.max(), 3))
    old = (*oracle.nlist_rebo(pos, h, mol),
           *oracle.nlist_lj(pos, h, mol, rlist=6.0, max_neigh=2000))
    for scale, stale in ((0.05, False), (1.0, True)):
        print(f"case {spec} speed scale={scale}", flush=True)
        eng = airebo.AtomicEngine(mol, np.full(n, 12.011), mode=airebo.REBO_PHYSICAL,
                                  lj_max_neigh=0, lj=lj)
        eng.set_state(pos=pos, vel=vm[mol - 1] * scale, h=h, npt=_frozen_npt(n))
        eng.run(0, 19)  # no scheduled rebuild after the first
        faults, rebuilds = eng.counters()[:2]
        assert faults == 0 and (rebuilds > 1) == stale
        st = eng.get_state()
        E, F, W = eng.forces()
        fresh = (*oracle.nlist_rebo(st["pos"], st["h"], mol),
                 *oracle.nlist_lj(st["pos"], st["h"], mol, rlist=7.0, max_neigh=2000))
        assert (fresh[2] < 2000).all()
        E0, F0, W0 = oracle.airebo_forces(st["pos"], st["h"], fresh, oracle.REBO_PHYSICAL,
                                          params=lj_o)
        assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
        for a, b in zip(E, E0):
            assert rel(a, b) < TOL
        if stale:
            E1, F1, W1 = oracle.airebo_forces(st["pos"], st["h"], old, oracle.REBO_PHYSICAL,
                                              params=lj_o)
            assert rel_max(F1, F0) > 20 * TOL
        eng.close()


def _radius(pos, h, mol):
    """Largest atom distance from its molecule's centre (minimum image)."""
    hi = oracle.inv9(h).reshape(3, 3)
    hm = h.reshape(3, 3)
    out = 0.0
    for m in range(1, mol.max() + 1):
        x = pos[mol == m]
        d = x - x[0]
        d -= (np.round(d @ hi.T)) @ hm.T
        c = x[0] + d.mean(axis=0)
        out = max(out, np.linalg.norm(x[0] + d - c, axis=1).max())
    return out


def test_image_modes_follow_a_compressed_cell():
    """The list walk takes every molecule pair's image from the molecule
    centres in FP32 only while the cell's half width exceeds the list radius
    (rc + 2 A) + 2 R_max + 1 A; the engine decides this on the device at every
    rebuild.  A 3x3x3 crystal starts 0.1 A above that bound and is compressed
    at a constant cell velocity of -0.08 A/fs (barostat frozen) to below it
    by the rebuild before step 20.  The step-20 forces equal the oracle's on
    the reference's truncated lists of the rebuild state.  On the way the
    cell columns shrink by 0.04 A a step, which makes the per-atom lists
    stale every few steps (rebuilt within the step)."""
    from fuller_md_en_amd import system
    m0 = system.molecule("C60")
    rm = np.linalg.norm(m0.coords - m0.coords.mean(axis=0), axis=1).max()
    a0 = (12.2 + 2 * rm + 1.0 + 0.1) / 1.5
    cen, h = system.make_crystal("FCC", a0, 3)
    pos = (cen[:, None, :] + m0.coords[None]).reshape(-1, 3)
    pos = pos + np.random.default_rng(31).uniform(-0.01, 0.01, pos.shape)
    p = pos.reshape(-1).copy()
    oracle.lib().orc_apply_pbc(p, h, oracle.inv9(h), len(pos))
    pos = p.reshape(-1, 3)
    mol = np.repeat(np.arange(1, len(cen) + 1), m0.natom).astype(np.int32)
    n = len(pos)
    eng = airebo.AtomicEngine(mol, np.full(n, 12.011), mode=airebo.REBO_COMPAT, clamp_vg=0.1)
    npt = _frozen_npt(n)
    npt.Vg[0] = npt.Vg[4] = npt.Vg[8] = -0.08
    eng.set_state(pos=pos, vel=np.zeros_like(pos), h=h, npt=npt)
    eng.run(0, 19)
    S = eng.get_state()
    bound = lambda x, hh: 12.2 + 2 * _radius(x, hh, mol) + 1.0  # noqa: E731
    assert 0.5 * h[0] > bound(pos, h) and 0.5 * S["h"][0] < bound(S["pos"], S["h"])
    lists = (*oracle.nlist_rebo(S["pos"], S["h"], mol), *oracle.nlist_lj(S["pos"], S["h"], mol))
    assert (lists[2] == 400).any()  # truncation active
    eng.run(19, 1)  # step 20: both lists rebuilt from S
    S2 = eng.get_state()
    E, F, W = eng.forces()
    E0, F0, W0 = oracle.airebo_forces(S2["pos"], S2["h"], lists, oracle.REBO_COMPAT)
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    for a, b in zip(E, E0):
        assert rel(a, b) < TOL
    faults, rebuilds = eng.counters()[:2]
    assert faults == 0 and rebuilds > 2  # the scheduled ones: the first and step 20
    eng.close()


def test_engine_trajectory_across_rebuild_matches_reference():
    """25 NPT steps of a 3x3x3 crystal (per-atom LJ lists, molecule images)
    with both lists rebuilt before step 20, against the reference build."""
    g = golden("airebo_rebuild_nc3.npz")
    pos, vel = g["pos"].astype(np.float64), g["vel"].astype(np.float64)
    eng = _engine(pos, g["h"], g["mol_id"], vel=vel)
    eng.set_state(npt=airebo.NPTState.from_array(g["npt0"]))
    th = eng.run(0, 25)
    ep = np.array([[t.Ep, t.Ep_rebo, t.Ep_lj, t.KE] for t in th])
    for k in range(4):
        assert rel_max(ep[:, k], g["traj"][:, k]) < TOL
    st = eng.get_state()
    L0 = np.diag(g["h"].reshape(3, 3))
    d = st["pos"] - pos
    d -= np.round(d / L0) * L0
    assert np.abs(d - g["dpos_end"]).max() < 1e-6
    assert rel_max(st["vel"], g["vel_end"]) < 1e-4
    assert np.abs(st["h"] - g["h_end"]).max() < 1e-9
    assert eng.counters()[0] == 0 and eng.counters()[1] >= 2
    eng.close()


def test_engine_trajectory_matches_reference():
    g = golden("airebo_nc1.npz")
    eng = _engine(g["pos"], g["h"], g["mol_id"], vel=g["vel"])
    eng.set_state(npt=airebo.NPTState.from_array(g["npt0"]))
    th = eng.run(0, 5)
    ep = np.array([[t.Ep, t.Ep_rebo, t.Ep_lj, t.KE] for t in th])
    for k in range(4):
        assert rel_max(ep[:, k], g["traj"][:, k]) < TOL
    st = eng.get_state()
    assert np.abs(st["pos"] - g["pos_end"]).max() < 1e-8
    assert rel_max(st["vel"], g["vel_end"]) < 1e-5
    eng.close()


def test_engine_deterministic_and_list_rebuild():
    pos, h, mol = crystal(3, 11, 0.01)
    from fuller_md_en_amd import system
    vel = system.atomic_init_velocities(np.full(len(pos), 12.011), 42, 300.0)
    out = []
    for _ in range(2):
        eng = _engine(pos, h, mol, vel=vel)
        th = eng.run(0, 25)  # crosses the rebuild at step 20
        out.append((eng.get_state()["pos"].copy(), th[24].Ep))
        assert eng.counters()[1] >= 2
        eng.close()
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_molecule_lists_from_cells_equal_the_scan():
    """Replicas of at least cell_min_molecules molecules bin the molecule
    centres into cells (at least one pair reach wide along each cell vector)
    and take each molecule's list from its 27 cells, sorted; below that every
    molecule is scanned.  Both give the same lists, so on a 5x5x5 crystal
    (three cells per axis) the forces after the first build and 25 steps
    across the step-20 rebuild are bitwise equal."""
    from fuller_md_en_amd import system
    s = system.airebo_system(5, "C60", T=300.0, seed=42)
    # three cells per axis: a third of the box holds the pair reach of the
    # molecule lists (rc + list skin + 1 A drift + two molecule radii)
    reach = 10.2 + 2.0 + 1.0 + 2.0 * _radius(s["pos"], s["h"], s["mol_id"])
    assert s["h"][0] / 3.0 >= reach and s["h"][0] / 4.0 < reach
    out = []
    for cmin in (1, 1 << 30):
        print(f"case cell_min_molecules={cmin}", flush=True)
        eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL,
                                  cell_min_molecules=cmin)
        eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"],
                      npt=airebo.make_npt(300.0, 0.0, len(s["pos"])))
        eng.prepare()
        E, F, W = eng.forces()
        th = eng.run(0, 25)
        out.append((F, np.array([t.Ep for t in th]), eng.get_state()["pos"].copy()))
        assert eng.counters()[0] == 0
        eng.close()
    (Fa, ea, pa), (Fb, eb, pb) = out
    assert np.array_equal(Fa, Fb) and np.array_equal(ea, eb) and np.array_equal(pa, pb)

