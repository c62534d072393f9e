"""What the mathematics fixes about the oracle, independent of the reference:
finite-difference gradients, strain-derivative virials, closed forms.

These pin the oracle's physical REBO mode (which the reference does not have)
and document the reference's compat REBO forces not being the gradient of its
own energy (its angular force has the opposite sign and lacks d cos / d r_j).
"""
import numpy as np
import pytest

import oracle
from conftest import golden


def _sys(jit=0.05, seed=3):
    g = golden("airebo_nc1.npz")
    rng = np.random.default_rng(seed)
    pos = g["pos"] + rng.uniform(-jit, jit, g["pos"].shape)
    return pos, g["h"].copy(), g["mol_id"]


def _rebo_E(pos, h, mol, mode=oracle.REBO_PHYSICAL):
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    return oracle.compute_rebo(pos, h, cr, lr, mode)


def test_dimer_closed_form():
    """Two bonded atoms (b_ij = b_ji = 1): E = fc(r) (V_R - V_A), force along the bond."""
    L = oracle.lib()
    h = np.diag([50.0] * 3).reshape(-1)
    for r in (1.2, 1.45, 1.85):
        pos = np.array([[10.0, 10.0, 10.0], [10.0 + r, 10.0, 10.0]])
        e, F, _ = _rebo_E(pos, h, np.array([1, 1], np.int32))
        assert abs(e - L.orc_rebo_fc(r) * (L.orc_rebo_vr(r) - L.orc_rebo_va(r))) < 1e-12 * abs(e)
        dE = (L.orc_rebo_dfc(r) * (L.orc_rebo_vr(r) - L.orc_rebo_va(r)) +
              L.orc_rebo_fc(r) * (L.orc_rebo_dvr(r) - L.orc_rebo_dva(r)))
        assert abs(F[1, 0] + dE) < 1e-10 * max(abs(dE), 1) and abs(F[0, 0] - dE) < 1e-10 * max(abs(dE), 1)


def test_scalar_derivatives_by_finite_difference():
    L = oracle.lib()
    for f, df in (("vr", "dvr"), ("va", "dva"), ("fc", "dfc")):
        for r in (1.3, 1.75, 1.9):
            d = 1e-6
            num = (getattr(L, f"orc_rebo_{f}")(r + d) - getattr(L, f"orc_rebo_{f}")(r - d)) / (2 * d)
            assert abs(num - getattr(L, f"orc_rebo_{df}")(r)) < 1e-6 * max(1, abs(num))
    for x in (-0.9, -0.5, 0.0, 0.7):
        d = 1e-7
        num = (L.orc_rebo_G(x + d) - L.orc_rebo_G(x - d)) / (2 * d)
        assert abs(num - L.orc_rebo_dG(x)) < 1e-6 * max(1, abs(num))


def test_rebo_physical_forces_are_the_gradient():
    pos, h, mol = _sys()
    e, F, _ = _rebo_E(pos, h, mol)
    rng = np.random.default_rng(0)
    for i in rng.choice(len(pos), 6, replace=False):
        for a in range(3):
            d = 1e-5
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += d
            p2[i, a] -= d
            num = -(_rebo_E(p1, h, mol)[0] - _rebo_E(p2, h, mol)[0]) / (2 * d)
            assert abs(num - F[i, a]) < 1e-5 * max(1.0, abs(num))


def test_rebo_compat_forces_are_not_the_gradient():
    """The reference's REBO forces (compat mode) differ from -dE/dx: the angular term
    carries the opposite sign and no d cos / d r_partner term (AB:385-386, 413-419)."""
    pos, h, mol = _sys()
    _, Fp, _ = _rebo_E(pos, h, mol, oracle.REBO_PHYSICAL)
    _, Fc, _ = _rebo_E(pos, h, mol, oracle.REBO_COMPAT)
    assert np.abs(Fc - Fp).max() > 0.05 * np.abs(Fp).max()


@pytest.mark.parametrize("ab", [(0, 0), (1, 1), (0, 1), (2, 1)])
def test_rebo_physical_virial_is_strain_derivative(ab):
    """W_ab = -dE/d eps_ab for the homogeneous deformation r -> (1 + eps) r."""
    pos, h, mol = _sys()
    _, _, W = _rebo_E(pos, h, mol)
    a, b = ab
    d = 1e-6

    def E_at(eps):
        M = np.eye(3)
        M[a, b] += eps
        return _rebo_E(pos @ M.T, (M @ h.reshape(3, 3)).reshape(-1), mol)[0]

    num = -(E_at(d) - E_at(-d)) / (2 * d)
    assert abs(num - W[3 * a + b]) < 1e-5 * max(1.0, np.abs(W).max())


def test_lj_exact_forces_gradient_and_virial():
    g = golden("airebo_nc2.npz")
    pos, h, mol = g["pos"], g["h"], g["mol_id"]
    cl, ll = oracle.nlist_lj(pos, h, mol, max_neigh=2000)  # no truncation

    def E(p, hh):
        return oracle.compute_lj(p, hh, cl, ll)[0]

    e, F, W = oracle.compute_lj(pos, h, cl, ll)
    for i in (0, 500, 1919):
        for a in range(3):
            d = 1e-5
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += d
            p2[i, a] -= d
            num = -(E(p1, h) - E(p2, h)) / (2 * d)
            assert abs(num - F[i, a]) < 1e-6 * max(1.0, np.abs(F).max())
    d = 1e-6
    M = np.eye(3) * (1 + d)
    Mm = np.eye(3) * (1 - d)
    num_tr = -(E(pos @ M.T, (M @ h.reshape(3, 3)).reshape(-1)) -
               E(pos @ Mm.T, (Mm @ h.reshape(3, 3)).reshape(-1))) / (2 * d)
    assert abs(num_tr - (W[0] + W[4] + W[8])) < 1e-5 * abs(num_tr)


def test_rigid_forces_are_the_gradient():
    # nc=3: at nc=2 the cell edge is 2a, second-shell neighbours sit exactly at
    # L/2 and the molecule-level minimum image (C1:387) switches under any
    # displacement, so the reference energy is not differentiable there.
    g = golden("rigid_core_nc3.npz")
    rmcut = float(g["rmcut"])

    def E(pos, qv):
        return oracle.rigid_forces(pos, qv, g["body"], g["h"], g["nl_count"], g["nl_list"],
                                   rmcut ** 2)[0]

    _, F, T, _ = oracle.rigid_forces(g["pos"], g["qv"], g["body"], g["h"], g["nl_count"],
                                     g["nl_list"], rmcut ** 2)
    for i in (0, 57):
        for a in range(3):
            d = 1e-5
            p1, p2 = g["pos"].copy(), g["pos"].copy()
            p1[i, a] += d
            p2[i, a] -= d
            num = -(E(p1, g["qv"]) - E(p2, g["qv"])) / (2 * d)
            assert abs(num - F[i, a]) < 1e-6 * max(1.0, np.abs(F).max())
        # torque: -dE/dtheta for a small rotation of molecule i about each axis
        for a in range(3):
            th = 1e-6
            axis = np.zeros(3)
            axis[a] = 1.0

            def rot(q, ang):
                dq = np.array([np.cos(ang / 2), *(np.sin(ang / 2) * axis)])
                w1, x1, y1, z1 = dq
                w2, x2, y2, z2 = q
                return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                                 w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                                 w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                                 w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])
            q1, q2 = g["qv"].copy(), g["qv"].copy()
            q1[i] = rot(q1[i], th)
            q2[i] = rot(q2[i], -th)
            num = -(E(g["pos"], q1) - E(g["pos"], q2)) / (2 * th)
            assert abs(num - T[i, a]) < 1e-5 * max(1.0, np.abs(T).max())


# ------------------------------------------------------ molecular mechanics
# The reference's MM forces (compat) are not the gradient of its energy: bond
# and angle forces are +dE/dr (MM:391, 468) and the middle-atom
# torsion forces take the dot-product coefficients with the opposite sign
# (MM:528-529, 653, 672).  The physical mode is pinned here by the mathematics.
def _mm_sys(k4=None, jit=0.05, seed=11, shear=True):
    g = golden("mm_c60_nc1.npz")
    rng = np.random.default_rng(seed)
    pos = g["pos"] + rng.uniform(-jit, jit, g["pos"].shape)
    h = g["h"].copy()
    if shear:
        h[1] = 0.4  # a non-orthogonal cell
    t = oracle.mm_topology(g["coords"], g["bonds"], 4, g["k4"] if k4 is None else k4)
    cnt, lst = oracle.mm_nlist(pos, h, t["mol_id"])
    return pos, h, t, cnt, lst, rng


def _mm_fd(pos, h, t, cnt, lst, mode, charge=None, atoms=range(0, 240, 13), d=1e-5):
    G = np.zeros((len(atoms), 3))
    for r, a in enumerate(atoms):
        for x in range(3):
            pp, pm = pos.copy(), pos.copy()
            pp[a, x] += d
            pm[a, x] -= d
            G[r, x] = (oracle.mm_forces(pp, h, t, cnt, lst, charge, mode)[0][6] -
                       oracle.mm_forces(pm, h, t, cnt, lst, charge, mode)[0][6]) / (2 * d)
    return G


def test_mm_physical_forces_are_the_gradient():
    pos, h, t, cnt, lst, rng = _mm_sys()
    q = rng.uniform(-0.05, 0.05, len(pos))
    atoms = list(range(0, 240, 13))
    for charge in (None, q):
        E, F, W = oracle.mm_forces(pos, h, t, cnt, lst, charge, oracle.MM_PHYSICAL)
        G = _mm_fd(pos, h, t, cnt, lst, oracle.MM_PHYSICAL, charge, atoms)
        assert np.abs(F[atoms] + G).max() < 1e-7 * np.abs(F).max()


@pytest.mark.parametrize("ab", [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)])
def test_mm_physical_virial_is_strain_derivative(ab):
    """W_ab = -dE/de_ab under r -> (1 + e) r, h -> (1 + e) h (lists fixed)."""
    pos, h, t, cnt, lst, rng = _mm_sys()
    q = rng.uniform(-0.05, 0.05, len(pos))
    E, F, W = oracle.mm_forces(pos, h, t, cnt, lst, q, oracle.MM_PHYSICAL)
    a, b = ab
    d = 1e-5

    def energy(e):
        M = np.eye(3)
        M[a, b] += e
        return oracle.mm_forces(pos @ M.T, (M @ h.reshape(3, 3)).reshape(-1), t, cnt, lst, q,
                                oracle.MM_PHYSICAL)[0][6]

    wn = -(energy(d) - energy(-d)) / (2 * d)
    assert abs(W[3 * a + b] - wn) < 1e-7 * np.abs(W).max()


def test_mm_compat_forces_against_the_gradient():
    """Term by term (one force constant at a time): bond and angle forces are
    exactly +gradient; torsions are neither -gradient nor +gradient, while
    the physical mode is -gradient for every term."""
    k = np.array(oracle.MM_K4_DEFAULT)
    for term in range(4):
        k4 = np.zeros(4)
        k4[term] = k[term]
        pos, h, t, cnt, lst, _ = _mm_sys(k4=k4, shear=False)
        # one molecule in a large cell: no LJ pairs
        sel = t["mol_id"] == 0
        p1 = pos[sel] - pos[sel].mean(axis=0) + 30.0
        h1 = np.diag([60.0] * 3).reshape(-1)
        t1 = oracle.mm_topology(golden("mm_c60_nc1.npz")["coords"],
                                golden("mm_c60_nc1.npz")["bonds"], 1, k4)
        c1, l1 = oracle.mm_nlist(p1, h1, t1["mol_id"])
        assert c1.sum() == 0
        atoms = list(range(0, 60, 7))
        G = _mm_fd(p1, h1, t1, c1, l1, oracle.MM_COMPAT, None, atoms)
        Fc = oracle.mm_forces(p1, h1, t1, c1, l1, None, oracle.MM_COMPAT)[1][atoms]
        Fp = oracle.mm_forces(p1, h1, t1, c1, l1, None, oracle.MM_PHYSICAL)[1][atoms]
        s = np.abs(G).max()
        assert np.abs(Fp + G).max() < 1e-6 * s
        if term < 2:
            assert np.abs(Fc - G).max() < 1e-6 * s
        else:
            assert np.abs(Fc - G).max() > 0.1 * s and np.abs(Fc + G).max() > 0.1 * s


def test_mm_newton_third_law_and_translation():
    """Every term's forces sum to zero in both modes; energies equal between
    modes and invariant under a rigid translation across the boundary."""
    pos, h, t, cnt, lst, rng = _mm_sys()
    q = rng.uniform(-0.05, 0.05, len(pos))
    Ec, Fc, _ = oracle.mm_forces(pos, h, t, cnt, lst, q, oracle.MM_COMPAT)
    Ep, Fp, _ = oracle.mm_forces(pos, h, t, cnt, lst, q, oracle.MM_PHYSICAL)
    assert np.array_equal(Ec, Ep)
    assert np.abs(Fc.sum(axis=0)).max() < 1e-10
    assert np.abs(Fp.sum(axis=0)).max() < 1e-10
    shift = np.array([3.3, -7.1, 5.2])
    Es = oracle.mm_forces(pos + shift, h, t, cnt, lst, q, oracle.MM_PHYSICAL)[0]
    assert np.abs(Es - Ep).max() < 1e-9 * np.abs(Ep).max()
