"""The FP64 oracle against the golden vectors written by the reference build.

The oracle follows the reference's arithmetic order, so agreement is expected
to the last few ulps (rtol 1e-12 leaves room for libm differences only).
"""
import numpy as np
import pytest

import oracle
from conftest import golden, unpack_list

RT = 1e-12


def close(a, b, rt=RT):
    a, b = np.asarray(a, float), np.asarray(b, float)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rt * scale, (np.abs(a - b).max(), scale)


CC1 = ["C60", "C70", "C84", "C72_HCP", "C76Td_BCC", "C84Td_HCP", "C60_BCC"]


@pytest.mark.parametrize("name", ["rigid_core_nc2.npz", "rigid_core_nc3.npz"] +
                         [f"rigid_cc1_{t}.npz" for t in CC1])
def test_rigid_list_and_forces(name):
    g = golden(name)
    cnt, lst = oracle.rigid_nlist(g["pos"], g["h"], float(g["rmcut"]))
    assert (cnt == g["nl_count"]).all()
    for i in range(len(cnt)):
        assert (lst[i, :cnt[i]] == g["nl_list"][i, :cnt[i]]).all()
    ep, F, T, W = oracle.rigid_forces(g["pos"], g["qv"], g["body"], g["h"], cnt, lst,
                                      float(g["rmcut"]) ** 2)
    close(ep, g["Ep"])
    close(F, g["Fv"])
    close(T, g["Tv"])
    close(W, g["Wm9"])


def _rigid_traj(g, nsteps, rebuild_every):
    L = oracle.lib()
    n = len(g["pos"])
    natom = len(g["body"])
    P, V = g["pos"].reshape(-1).copy(), g["vel"].reshape(-1).copy()
    Q, O = g["qv"].reshape(-1).copy(), g["omg"].reshape(-1).copy()
    F, T, W = g["Fv"].reshape(-1).copy(), g["Tv"].reshape(-1).copy(), g["Wm9"].copy()
    h, hi = g["h"].copy(), np.zeros(9)
    npt = oracle.NPT.from_array(g["npt0"])
    cnt, lst = g["nl_count"].copy(), g["nl_list"].reshape(-1).copy()
    lab = np.zeros(n * natom * 3)
    out = np.zeros(2)
    p = oracle.lj_params("rigid")
    rmcut = float(g["rmcut"])
    eps, kes = [], []
    for step in range(1, nsteps + 1):
        if rebuild_every and step % rebuild_every == 0:
            oracle.lib().orc_mat_inv9(h, hi)
            L.orc_rigid_nlist(P, h, hi, n, rmcut, 80, cnt, lst)
        L.orc_rigid_step_npt(P, V, Q, O, F, T, W, h, hi, g["body"].reshape(-1).copy(),
                             float(g["I0"]), float(g["Mmol"]), n, natom, rmcut ** 2, 1.0,
                             npt, cnt, lst, 80, p, lab, out)
        eps.append(out[0])
        kes.append(out[1])
    return dict(P=P, V=V, Q=Q, O=O, F=F, T=T, W=W, h=h, npt=npt, eps=eps, kes=kes)


def test_rigid_core_trajectory():
    g = golden("rigid_core_nc2.npz")
    r = _rigid_traj(g, int(g["nsteps"]), 25)
    close(r["eps"], g["traj_Ep"], 1e-11)
    close(r["kes"], g["traj_KE"], 1e-11)
    close(r["P"], g["pos_end"].reshape(-1), 1e-11)
    close(r["V"], g["vel_end"].reshape(-1), 1e-9)
    close(r["Q"], g["qv_end"].reshape(-1), 1e-11)
    close(r["h"], g["h_end"], 1e-12)
    close(r["npt"].to_array(), g["npt_end"], 1e-9)


@pytest.mark.parametrize("tag", CC1)
def test_rigid_cc1_steps(tag):
    g = golden(f"rigid_cc1_{tag}.npz")
    r = _rigid_traj(g, 3, 0)
    close(r["eps"], g["traj_Ep"], 1e-11)
    close(r["P"], g["pos_end"].reshape(-1), 1e-11)
    close(r["O"], g["omg_end"].reshape(-1), 1e-9)


def test_airebo_nc1_lists_and_forces():
    g = golden("airebo_nc1.npz")
    cr, lr = oracle.nlist_rebo(g["pos"], g["h"], g["mol_id"])
    cl, ll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    assert (cr == g["nlc_r"]).all() and (cl == g["nlc_l"]).all()
    assert (lr == unpack_list(g["nlc_r"], g["nll_r"], 12)).all()
    assert (ll == unpack_list(g["nlc_l"], g["nll_l"], 400)).all()
    er, Fr, Wr = oracle.compute_rebo(g["pos"], g["h"], cr, lr, oracle.REBO_COMPAT)
    close(er, g["E_rebo"])
    close(Fr, g["F_rebo"])
    close(Wr, g["W_rebo"])
    el, Fl, Wl = oracle.compute_lj(g["pos"], g["h"], cl, ll)
    close(el, g["E_lj"])
    close(Fl, g["F_lj"])
    close(Wl, g["W_lj"])
    E, F, W = oracle.airebo_forces(g["pos"], g["h"], (cr, lr, cl, ll))
    close(E, g["E"])
    close(F, g["F"])
    close(W, g["W"])


def test_airebo_nc1_steps():
    g = golden("airebo_nc1.npz")
    L = oracle.lib()
    na = len(g["pos"])
    cr, lr = oracle.nlist_rebo(g["pos"], g["h"], g["mol_id"])
    cl, ll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    P, V = g["pos"].reshape(-1).copy(), g["vel"].reshape(-1).copy()
    F, W = g["F"].reshape(-1).copy(), g["W"].copy()
    h, hi = g["h"].copy(), oracle.inv9(g["h"])
    npt = oracle.NPT.from_array(g["npt0"])
    out = np.zeros(4)
    p = oracle.lj_params("airebo")
    traj = []
    for _ in range(5):
        L.orc_airebo_step_npt(P, V, F, W, h, hi, g["mass"], na, float(g["dt"]), npt, cr,
                              lr.reshape(-1), 12, cl, ll.reshape(-1), 400,
                              oracle.REBO_COMPAT, p, out)
        traj.append(out.copy())
    close(np.array(traj), g["traj"], 1e-11)
    close(P, g["pos_end"].reshape(-1), 1e-11)
    close(V, g["vel_end"].reshape(-1), 1e-9)
    close(npt.to_array(), g["npt_end"], 1e-9)


def test_airebo_nc2_truncated_lj():
    g = golden("airebo_nc2.npz")
    cr, lr = oracle.nlist_rebo(g["pos"], g["h"], g["mol_id"])
    cl, ll = oracle.nlist_lj(g["pos"], g["h"], g["mol_id"])
    assert (cl == g["nlc_l"]).all() and (cr == g["nlc_r"]).all()
    last = np.array([ll[i, cl[i] - 1] if cl[i] else -1 for i in range(len(cl))])
    assert (last == g["lj_last"]).all()
    assert (cl == 400).sum() > 0  # the truncation is exercised
    E, F, W = oracle.airebo_forces(g["pos"], g["h"], (cr, lr, cl, ll))
    close(E, g["E"])
    close(F, g["F"])
    close(W, g["W"])


@pytest.mark.parametrize("tag", ["C72_HCP", "C76Td_BCC", "C84Td_HCP", "C60_BCC"])
def test_host_lattices_equal_reference(tag):
    """The host crystal builders and lattice constant (LJ:229-254) against
    the reference program's for HCP and BCC, and the .cc1 loader's body
    frame for other species (LJ:178-207)."""
    from fuller_md_en_amd import system
    g = golden(f"rigid_cc1_{tag}.npz")
    st = str(g["crystal"])
    # the reference's --fullerene spellings (LJ:210-226)
    mol = system.molecule({"C72_HCP": "C72:D6d", "C76Td_BCC": "C76:Td",
                           "C84Td_HCP": "C84:20", "C60_BCC": "C60"}[tag])
    assert mol.natom == int(g["natom"])
    close(mol.coords, g["body"], 1e-14)
    a0 = system.default_a0(mol.dmol, st, 1.0)
    close(a0, float(g["a0"]), 1e-14)
    pos, h = system.make_crystal(st, a0, int(g["nc"]))
    assert pos.shape == g["pos"].shape
    close(pos, g["pos"], 1e-14)
    close(h, g["h"], 1e-14)


# ------------------------------------------------------ molecular mechanics
def _mm_topo(g, nmol):
    t = oracle.mm_topology(g["coords"], g["bonds"], nmol, g["k4"])
    for k in ("b", "bp", "a", "ap", "d", "dp", "dm", "i", "ip"):
        n = len(t[k]) // nmol
        # molecule 0 bitwise as the reference builds it (MM:254-334) ...
        assert np.array_equal(t[k][:n], g["t_" + k]), k
        # ... and the replicas shifted by natom per molecule (MM:314-332)
        if k in ("b", "a", "d", "i"):
            natom = len(g["coords"])
            shift = (np.arange(nmol) * natom).repeat(n)[:, None]
            assert np.array_equal(t[k], np.tile(g["t_" + k], (nmol, 1)) + shift), k
    return t


@pytest.mark.parametrize("name", ["mm_c60_nc1.npz", "mm_c70_hcp.npz"])
def test_mm_topology_list_and_forces(name):
    g = golden(name)
    nmol = int(g["mol_id"].max()) + 1
    t = _mm_topo(g, nmol)
    assert np.array_equal(t["mol_id"], g["mol_id"])
    cnt, lst = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    assert (cnt == g["nlc"]).all()
    if "nll" in g:
        assert (lst == unpack_list(g["nlc"], g["nll"], 400)).all()
    E, F, W = oracle.mm_forces(g["pos"], g["h"], t, cnt, lst)
    close(E, g["E"])
    close(F, g["F"])
    close(W, g["W"])
    assert np.count_nonzero(W - np.diag(np.diag(W.reshape(3, 3))).reshape(-1)) == 0


def test_mm_coulomb_and_steps():
    g = golden("mm_c60_nc1.npz")
    t = _mm_topo(g, 4)
    cnt, lst = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    E, F, W = oracle.mm_forces(g["pos"], g["h"], t, cnt, lst, charge=g["charge"])
    assert abs(E[5]) > 1e-3  # the Coulomb term is exercised
    close(E, g["E_q"])
    close(F, g["F_q"])
    close(W, g["W_q"])
    P, V, F, W, h = g["pos"], g["vel"], g["F"], g["W"], g["h"]
    npt = oracle.NPT.from_array(g["npt0"])
    traj = []
    for _ in range(5):
        P, V, F, W, h, npt, out = oracle.mm_step_npt(P, V, F, W, h, npt, float(g["dt"]), t,
                                                     cnt, lst)
        traj.append(out)
    close(np.array(traj), g["traj"], 1e-11)
    close(P, g["pos_end"], 1e-11)
    close(V, g["vel_end"], 1e-9)
    close(h, g["h_end"], 1e-12)
    close(npt.to_array(), g["npt_end"], 1e-9)


def test_mm_nc2_truncated_lj():
    g = golden("mm_c60_nc2.npz")
    nmol = int(g["mol_id"].max()) + 1
    coords = golden("mm_c60_nc1.npz")
    t = oracle.mm_topology(coords["coords"], coords["bonds"], nmol, coords["k4"])
    cl, ll = oracle.mm_nlist(g["pos"], g["h"], g["mol_id"])
    assert (cl == g["nlc"]).all()
    last = np.array([ll[i, cl[i] - 1] if cl[i] else -1 for i in range(len(cl))])
    assert (last == g["lj_last"]).all()
    assert (cl == 400).sum() > 0  # the truncation is exercised
    E, F, W = oracle.mm_forces(g["pos"], g["h"], t, cl, ll)
    close(E, g["E"])
    close(F, g["F"])
    close(W, g["W"])
