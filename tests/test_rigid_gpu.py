"""Rigid-body LJ on the GPU against the reference build's golden vectors and
the FP64 oracle.

Tolerances (FP32 pair arithmetic, FP64 accumulation / integrator):
  forces, torques, virial   max|a-b| / max|b|  <= 1e-5
  energies                  |a-b| / |b|        <= 1e-5
  neighbour lists           exact
"""
import numpy as np
import pytest

import oracle
from conftest import golden, rel_max
from fuller_md_en_amd import rigid

pytestmark = pytest.mark.gpu

TOL = 1e-5
CC1 = ["C60", "C70", "C84", "C72_HCP", "C76Td_BCC", "C84Td_HCP", "C60_BCC"]
RIGID = ["rigid_core_nc2.npz", "rigid_core_nc3.npz"] + [f"rigid_cc1_{t}.npz" for t in CC1]


def rel(a, b):
    return abs(a - b) / abs(b)


@pytest.mark.parametrize("name", RIGID)
def test_nlist_equals_reference(name):
    g = golden(name)
    print(f"case {name}", flush=True)
    cnt, lst = rigid.nlist_build_sym(g["pos"], g["h"], oracle.inv9(g["h"]), float(g["rmcut"]))
    assert (cnt == g["nl_count"]).all()
    for i in range(len(cnt)):
        assert (lst[i, :cnt[i]] == g["nl_list"][i, :cnt[i]]).all()


@pytest.mark.parametrize("name", RIGID)
def test_forces_match_reference(name):
    g = golden(name)
    print(f"case {name}", flush=True)
    ep, F, T, W = rigid.forces(g["pos"], g["qv"], g["body"], g["h"], oracle.inv9(g["h"]),
                               g["nl_count"], g["nl_list"], float(g["rmcut"]) ** 2)
    assert rel_max(F, g["Fv"]) < TOL
    assert rel_max(T, g["Tv"]) < TOL
    assert rel_max(W, g["Wm9"]) < TOL
    assert rel(ep, float(g["Ep"])) < TOL


def test_forces_compressed_lattice_against_oracle():
    """Strongly compressed crystal (a = 0.8 a0): repulsive wall, large forces."""
    g = golden("rigid_cc1_C60.npz")
    pos = g["pos"] * 0.8
    h = g["h"] * 0.8
    rmcut = float(g["rmcut"])
    cnt, lst = oracle.rigid_nlist(pos, h, rmcut)
    e0, F0, T0, W0 = oracle.rigid_forces(pos, g["qv"], g["body"], h, cnt, lst, rmcut ** 2)
    e1, F1, T1, W1 = rigid.forces(pos, g["qv"], g["body"], h, oracle.inv9(h), cnt, lst,
                                  rmcut ** 2)
    assert rel_max(F1, F0) < TOL and rel_max(T1, T0) < TOL and rel_max(W1, W0) < TOL
    assert rel(e1, e0) < TOL


def test_forces_lab_scratch():
    g = golden("rigid_core_nc2.npz")
    out = rigid.forces(g["pos"], g["qv"], g["body"], g["h"], oracle.inv9(g["h"]),
                       g["nl_count"], g["nl_list"], float(g["rmcut"]) ** 2, want_lab=True)
    lab = np.zeros(len(g["pos"]) * 60 * 3)
    oracle.lib().orc_rigid_lab(g["qv"].reshape(-1).copy(), g["body"].reshape(-1).copy(),
                               len(g["pos"]), 60, lab)
    assert np.abs(out[4].reshape(-1) - lab).max() < 1e-12


@pytest.mark.parametrize("tag", ["C60", "C70", "C84"])
def test_step_npt_dropin(tag):
    g = golden(f"rigid_cc1_{tag}.npz")
    print(f"case {tag}", flush=True)
    P, V = g["pos"].copy(), g["vel"].copy()
    Q, O = g["qv"].copy(), g["omg"].copy()
    F, T, W = g["Fv"].copy(), g["Tv"].copy(), g["Wm9"].copy()
    h, hi = g["h"].copy(), np.zeros(9)
    from fuller_md_en_amd._native import NPTState
    npt = NPTState.from_array(g["npt0"])
    eps = []
    for _ in range(3):
        ep, ke = rigid.step_npt(P, V, Q, O, F, T, W, h, hi, g["body"], float(g["I0"]),
                                float(g["Mmol"]), float(g["rmcut"]) ** 2, 1.0, npt,
                                g["nl_count"], g["nl_list"])
        eps.append(ep)
    assert max(rel(a, b) for a, b in zip(eps, g["traj_Ep"])) < TOL
    assert np.abs(P - g["pos_end"]).max() < 1e-8
    assert rel_max(V, g["vel_end"]) < 1e-6
    assert rel_max(O, g["omg_end"]) < 1e-6
    assert np.abs(Q - g["qv_end"]).max() < 1e-9
    assert np.abs(h - g["h_end"]).max() < 1e-10
    assert rel_max(F, g["Fv_end"]) < TOL
    assert rel_max(npt.to_array()[:11], g["npt_end"][:11]) < 1e-5


def _core_engine(g):
    eng = rigid.RigidEngine(g["body"], len(g["pos"]), float(g["I0"]), float(g["Mmol"]),
                            rmol=3.55, rmcut=float(g["rmcut"]))
    from fuller_md_en_amd._native import NPTState
    eng.set_state(pos=g["pos"], vel=g["vel"], qv=g["qv"], omg=g["omg"], h=g["h"],
                  npt=NPTState.from_array(g["npt0"]))
    return eng


def test_engine_trajectory_matches_reference():
    """30 NPT steps of the C1 core (nc=2) with list rebuild at step 25."""
    g = golden("rigid_core_nc2.npz")
    eng = _core_engine(g)
    th = eng.run(0, int(g["nsteps"]))
    ep = np.array([t.Ep for t in th])
    ke = np.array([t.KE for t in th])
    T = np.array([t.T for t in th])
    P = np.array([t.P for t in th])
    assert rel_max(ep, g["traj_Ep"]) < TOL
    assert rel_max(ke, g["traj_KE"]) < TOL
    assert rel_max(T, g["traj_T"]) < TOL
    assert np.abs(P - g["traj_P"]).max() < 1e-4 * max(np.abs(g["traj_P"]).max(), 0.01)
    st = eng.get_state()
    assert np.abs(st["pos"] - g["pos_end"]).max() < 1e-7
    assert rel_max(st["vel"], g["vel_end"]) < 1e-5
    assert np.abs(st["h"] - g["h_end"]).max() < 1e-9
    assert eng.counters()[0] == 0  # no list overflow
    eng.close()


@pytest.mark.parametrize("half", [1, 2, 0])
@pytest.mark.parametrize("tag", CC1)
def test_engine_steps_match_reference(tag, half):
    """Resident engine, every pair mode (1: wavefront per owned pair for up to
    64 sites, else per molecule; 2: per molecule; 0: both sides), against the
    reference's 3 NPT steps.

    C70 leaves a last site quad with two real sites; C84 (84 sites) runs the
    two-sites-per-lane kernels; the HCP and BCC cases (C72, C76-Td, C84 No.20,
    C60) cover the other lattices of the reference programs."""
    g = golden(f"rigid_cc1_{tag}.npz")
    print(f"case {tag} half={half}", flush=True)
    from fuller_md_en_amd._native import NPTState
    eng = rigid.RigidEngine(g["body"], len(g["pos"]), float(g["I0"]), float(g["Mmol"]),
                            rmol=float(g["Rmol"]), rmcut=float(g["rmcut"]), half_pairs=half)
    eng.set_state(pos=g["pos"], vel=g["vel"], qv=g["qv"], omg=g["omg"], h=g["h"],
                  npt=NPTState.from_array(g["npt0"]))
    th = eng.run(0, 3)
    assert max(rel(t.Ep, b) for t, b in zip(th, g["traj_Ep"])) < TOL
    assert max(rel(t.KE, b) for t, b in zip(th, g["traj_KE"])) < TOL
    st = eng.get_state()
    assert np.abs(st["pos"] - g["pos_end"]).max() < 1e-8
    assert rel_max(st["vel"], g["vel_end"]) < 1e-6
    assert rel_max(st["omg"], g["omg_end"]) < 1e-6
    assert np.abs(st["qv"] - g["qv_end"]).max() < 1e-9
    assert rel_max(st["h"], g["h_end"]) < 1e-10  # the cell follows the FP32 virial
    ep, F, T, W = eng.forces()
    assert rel_max(F, g["Fv_end"]) < TOL and rel_max(T, g["Tv_end"]) < TOL
    eng.close()


def test_engine_half_pairs_equal_full_pairs():
    """108 molecules, 20 steps with list rebuilds: the pair-once kernels (per
    owned pair, and per molecule: own share + gathered neighbour shares, trace
    virial) against the both-sides kernel."""
    g = golden("rigid_core_nc3.npz")
    out = []
    for half in (1, 2, 0):
        eng = rigid.RigidEngine(g["body"], len(g["pos"]), float(g["I0"]), float(g["Mmol"]),
                                rmol=3.55, rmcut=float(g["rmcut"]), nlup=10, half_pairs=half)
        eng.set_state(pos=g["pos"], vel=g["vel"], qv=g["qv"], omg=g["omg"], h=g["h"],
                      npt=rigid.make_npt(300.0, 0.0, len(g["pos"])))
        th = eng.run(0, 20)
        out.append((eng.get_state(), np.array([[t.Ep, t.KE, t.P] for t in th])))
        assert eng.counters()[0] == 0
        eng.close()
    s2, t2 = out[-1]
    for s1, t1 in out[:-1]:
        for k in range(3):
            assert rel_max(t1[:, k], t2[:, k]) < 1e-6
        assert np.abs(s1["pos"] - s2["pos"]).max() < 1e-8
        assert rel_max(s1["vel"], s2["vel"]) < 1e-6 and rel_max(s1["omg"], s2["omg"]) < 1e-6


def test_engine_deterministic():
    g = golden("rigid_core_nc2.npz")
    runs = []
    for _ in range(2):
        eng = _core_engine(g)
        eng.run(0, 12)
        st = eng.get_state()
        runs.append(np.concatenate([st["pos"].ravel(), st["vel"].ravel(), st["qv"].ravel()]))
        eng.close()
    assert np.array_equal(runs[0], runs[1])


def test_engine_forces_newton3():
    g = golden("rigid_core_nc3.npz")
    eng = _core_engine({**g, "npt0": rigid.make_npt(300.0, 0.0, len(g["pos"])).to_array()})
    eng.prepare()
    ep, F, T, W = eng.forces()
    assert rel_max(F, g["Fv"]) < TOL and rel(ep, float(g["Ep"])) < TOL
    # momentum conservation of the pair forces, within FP32 rounding
    assert np.abs(F.sum(axis=0)).max() < 1e-5 * np.abs(F).max() * np.sqrt(len(F))
    assert rel_max(W, W.reshape(3, 3).T.ravel()) < 1e-6  # LJ virial is symmetric
    eng.close()


def test_engine_counts_a_stale_list():
    """The molecule list holds the pairs within rmcut + skin at its build.
    When the next rebuild comes (step 25), a molecule whose pairs may have
    closed in by more than the skin since (twice its centre displacement plus
    the cell change) is counted as a list fault.  The nc=3 state at 300 K
    with the 3 A skin counts none; at ten times the speeds with a 0.1 A skin
    it counts.  Thermostat and barostat frozen (masses 1e30)."""
    g = golden("rigid_core_nc3.npz")
    n = len(g["pos"])
    for skin, speed, faults in ((3.0, 1.0, False), (0.1, 10.0, True)):
        print(f"case skin={skin} speed x{speed}", flush=True)
        eng = rigid.RigidEngine(g["body"], n, float(g["I0"]), float(g["Mmol"]), rmol=3.55,
                                rmcut=float(g["rmcut"]), skin=skin)
        npt = rigid.make_npt(300.0, 0.0, n)
        npt.Q = npt.W = 1e30
        eng.set_state(pos=g["pos"], vel=g["vel"] * speed, qv=g["qv"], omg=g["omg"], h=g["h"],
                      npt=npt)
        eng.run(0, 25)
        assert (eng.counters()[0] > 0) == faults
        eng.close()


def test_lists_from_cells_equal_the_scan():
    """Systems of at least cell_min_molecules molecules bin the centres into
    cells at least the list radius wide and take each list row from the 27
    cells around its own, sorted; below that every molecule is scanned.  On a
    5x5x5 C60 crystal (500 molecules, three cells per axis) both give the same
    lists, so 60 steps across the rebuilds at steps 25 and 50 are bitwise
    equal.  The stateless list of a 10x10x10 crystal (4000 molecules: from
    the cells by default) equals the reference build's."""
    from fuller_md_en_amd import system
    mol = system.molecule("C60")
    s = system.rigid_system(nc=5, spec="C60", T=300.0, seed=42)
    n = len(s["pos"])
    rmcut = 3.0 * 3.431 + 2.0 * mol.rmol + 1.0
    assert s["h"][0] / 3.0 >= rmcut + 3.0 > s["h"][0] / 4.0  # three cells per axis
    out = []
    for cmin in (1, 1 << 30):
        print(f"case nc=5 cell_min_molecules={cmin}", flush=True)
        eng = rigid.RigidEngine(mol.coords, n, mol.I0, mol.mass, rmol=mol.rmol,
                                cell_min_molecules=cmin)
        eng.set_state(pos=s["pos"], vel=s["vel"], qv=s["qv"], omg=s["omg"], h=s["h"],
                      npt=rigid.make_npt(300.0, 0.0, n))
        th = eng.run(0, 60)
        out.append((np.array([[t.Ep, t.P] for t in th]), eng.get_state()["pos"].copy()))
        assert eng.counters()[0] == 0
        eng.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    print("case stateless nc=10", flush=True)
    s = system.rigid_system(nc=10, spec="C60", T=300.0, seed=42)
    cnt, lst = rigid.nlist_build_sym(s["pos"], s["h"], oracle.inv9(s["h"]), rmcut)
    c0, l0 = oracle.rigid_nlist(s["pos"], s["h"], rmcut)
    assert (cnt == c0).all() and cnt.min() > 0
    for i in range(len(cnt)):
        assert (lst[i, :cnt[i]] == l0[i, :cnt[i]]).all()
