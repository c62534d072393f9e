"""The oracle's REBO-II / AIREBO extension terms (orc_rebo_ext), which the
reference does not have: the P_ij correction of the bond order and the
torsion term.  Pinned by what the mathematics fixes -- equality with the
exact-gradient REBO when both are off, finite-difference gradients through
the switch regions, the strain-derivative virial, and closed forms of a
three-coordinated centre and of planar, twisted and stretched chains.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import golden

EPS_T = oracle.AIREBO_TORSION_EPS
PCC = oracle.AIREBO_PCC


def _stretched(scale=1.25, jit=0.05, seed=5):
    """One C60 of the 1x1x1 crystal scaled so that its bonds (1.75-1.81 A) lie
    in the 1.7-2.0 A switch: fc < 1, so dP/dN and the torsion weights'
    derivatives carry force; the other three molecules as they are."""
    g = golden("airebo_nc1.npz")
    pos, h, mol = g["pos"].copy(), g["h"].copy(), g["mol_id"]
    m1 = mol == 1
    L = np.diag(h.reshape(3, 3))  # the cubic cell
    d = pos[m1] - pos[m1][0]
    d -= L * np.round(d / L)  # the molecule whole, across the periodic boundary
    c = d.mean(0)
    pos[m1] = pos[m1][0] + c + scale * (d - c)
    rng = np.random.default_rng(seed)
    return pos + rng.uniform(-jit, jit, pos.shape), h, mol


def _E(pos, h, mol, pcc=PCC, eps_t=EPS_T):
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    return oracle.rebo_ext(pos, h, cr, lr, pcc, eps_t)


def test_extension_off_equals_the_exact_gradient_rebo():
    pos, h, mol = _stretched()
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    e0, F0, W0 = oracle.compute_rebo(pos, h, cr, lr, oracle.REBO_PHYSICAL)
    e1, F1, W1, et = oracle.rebo_ext(pos, h, cr, lr, (0.0,) * 5, 0.0)
    assert et == 0.0
    assert abs(e1 - e0) < 1e-13 * abs(e0)
    assert np.abs(F1 - F0).max() < 1e-12 * np.abs(F0).max()
    assert np.abs(W1 - W0).max() < 1e-12 * np.abs(W0).max()


def test_stretched_molecule_exercises_the_switches():
    pos, h, mol = _stretched()
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    L = oracle.lib()
    r = []
    for i in np.flatnonzero(mol == 1):
        for j in lr[i, : cr[i]]:
            d = pos[j] - pos[i]
            r.append(np.sqrt(d @ d))
    fc = np.array([L.orc_rebo_fc(x) for x in r])
    assert ((fc > 0.05) & (fc < 0.95)).sum() > 100


@pytest.mark.parametrize("terms", ["pcc", "torsion", "both"])
def test_extension_forces_are_the_gradient(terms):
    pos, h, mol = _stretched()
    pcc = PCC if terms in ("pcc", "both") else (0.0,) * 5
    eps_t = EPS_T if terms in ("torsion", "both") else 0.0
    e, F, _, _ = _E(pos, h, mol, pcc, eps_t)
    rng = np.random.default_rng(1)
    # atoms of the stretched molecule (switch regions) and of an intact one
    atoms = list(rng.choice(np.flatnonzero(mol == 1), 6, replace=False)) + \
        list(rng.choice(np.flatnonzero(mol == 2), 2, replace=False))
    for i in atoms:
        for a in range(3):
            d = 1e-5
            p1, p2 = pos.copy(), pos.copy()
            p1[i, a] += d
            p2[i, a] -= d
            num = -(_E(p1, h, mol, pcc, eps_t)[0] - _E(p2, h, mol, pcc, eps_t)[0]) / (2 * d)
            assert abs(num - F[i, a]) < 1e-6 * max(1.0, abs(num)), (i, a, num, F[i, a])


@pytest.mark.parametrize("ab", [(0, 0), (1, 1), (0, 2), (2, 1)])
def test_extension_virial_is_strain_derivative(ab):
    pos, h, mol = _stretched()
    _, _, W, _ = _E(pos, h, mol)
    a, b = ab
    d = 1e-6

    def E_at(eps):
        M = np.eye(3)
        M[a, b] += eps
        return _E(pos @ M.T, (M @ h.reshape(3, 3)).reshape(-1), mol)[0]

    num = -(E_at(d) - E_at(-d)) / (2 * d)
    assert abs(num - W[3 * a + b]) < 1e-6 * max(1.0, np.abs(W).max())


def _chain(omega_deg, r=(1.42, 1.42, 1.42), theta_deg=(120.0, 120.0)):
    """Atoms i, j, k, l with bonds i-j, j-k, k-l of lengths r, bond angles
    theta at j and k, dihedral omega (0 = cis), in a 40 A box."""
    tj, tk = np.radians(theta_deg)
    om = np.radians(omega_deg)
    j = np.zeros(3)
    k = np.array([r[1], 0.0, 0.0])
    i = j + r[0] * np.array([np.cos(tj), np.sin(tj), 0.0])
    l = k + r[2] * np.array([-np.cos(tk), np.sin(tk) * np.cos(om), np.sin(tk) * np.sin(om)])
    pos = np.array([i, j, k, l]) + 20.0
    h = np.diag([40.0] * 3).reshape(-1)
    return pos, h, np.ones(4, np.int32)


@pytest.mark.parametrize("omega", [0.0, 35.0, 90.0, 150.0, 180.0])
def test_torsion_of_a_chain_closed_form(omega):
    """One chain, all bonds inside the switch's inner radius:
    E_t = eps (256/405 cos^10(omega/2) - 1/10); the trans chain (omega = 180)
    sits at -eps/10, the cis one at eps (256/405 - 1/10)."""
    pos, h, mol = _chain(omega)
    _, _, _, et = _E(pos, h, mol, (0.0,) * 5, EPS_T)
    ref = EPS_T * (256.0 / 405.0 * math.cos(math.radians(omega) / 2) ** 10 - 0.1)
    assert abs(et - ref) < 1e-12


def test_torsion_weights_and_angle_switch():
    L = oracle.lib()
    # a stretched central bond scales the term by its switch value
    pos, h, mol = _chain(40.0, r=(1.42, 1.85, 1.42))
    _, _, _, et = _E(pos, h, mol, (0.0,) * 5, EPS_T)
    ref = EPS_T * (256.0 / 405.0 * math.cos(math.radians(40.0) / 2) ** 10 - 0.1)
    assert abs(et - L.orc_rebo_fc(1.85) * ref) < 1e-12
    # a straight bond angle (no dihedral) contributes nothing
    pos, h, mol = _chain(40.0, theta_deg=(180.0, 120.0))
    _, _, _, et = _E(pos, h, mol, (0.0,) * 5, EPS_T)
    assert et == 0.0


def test_pcc_of_a_three_coordinated_centre_closed_form():
    """A planar centre c with three neighbours at 1.42 A and 120 degrees: each
    bond's order from c sees N = 2 and two angles of cos = -1/2, from the
    neighbour N = 0, so E(P) - E(0) = 3 (-1/2) VA [(1 + 2 G + P(2))^-1/2 -
    (1 + 2 G)^-1/2] with P(2) the knot value."""
    L = oracle.lib()
    r = 1.42
    c = np.array([20.0, 20.0, 20.0])
    pos = np.array([c] + [c + r * np.array([math.cos(a), math.sin(a), 0.0])
                          for a in np.radians([90.0, 210.0, 330.0])])
    h = np.diag([40.0] * 3).reshape(-1)
    mol = np.ones(4, np.int32)
    e1 = _E(pos, h, mol, PCC, 0.0)[0]
    e0 = _E(pos, h, mol, (0.0,) * 5, 0.0)[0]
    G = L.orc_rebo_G(-0.5)
    ref = 3 * (-0.5) * L.orc_rebo_va(r) * ((1 + 2 * G + PCC[2]) ** -0.5 - (1 + 2 * G) ** -0.5)
    assert abs((e1 - e0) - ref) < 1e-10


def test_pcc_spline_between_the_knots():
    """One neighbour bond stretched to fc = 1/2 puts the others' N at 1.5,
    midway between the knots: P = (P(1) + P(2)) / 2 (the Hermite cubic with
    zero knot slopes is symmetric about the midpoint), and the force is still
    the gradient."""
    L = oracle.lib()
    r = 1.42
    c = np.array([20.0, 20.0, 20.0])
    pos = np.array([c] + [c + rr * np.array([math.cos(a), math.sin(a), 0.0])
                          for rr, a in zip((r, r, 1.85), np.radians([90.0, 210.0, 330.0]))])
    h = np.diag([40.0] * 3).reshape(-1)
    mol = np.ones(4, np.int32)
    assert abs(L.orc_rebo_fc(1.85) - 0.5) < 1e-12
    pcc2 = (0.0, 0.0, -0.08, 0.0, 0.0)
    e1, F1, _, _ = _E(pos, h, mol, pcc2, 0.0)
    e0 = _E(pos, h, mol, (0.0,) * 5, 0.0)[0]
    # bonds 1 and 2 (from c): N = 1 + 1/2, two angles of cos -1/2 (weights 1, 1/2);
    # bond 3 (from c): N = 2, P(2); the neighbours' own orders: N = 0
    G = L.orc_rebo_G(-0.5)
    va, fc3 = L.orc_rebo_va(r), 0.5
    p15 = 0.5 * (pcc2[1] + pcc2[2])
    d12 = (1 + 1.5 * G + p15) ** -0.5 - (1 + 1.5 * G) ** -0.5
    d3 = (1 + 2 * G + pcc2[2]) ** -0.5 - (1 + 2 * G) ** -0.5
    ref = 2 * (-0.5) * va * d12 + fc3 * (-0.5) * L.orc_rebo_va(1.85) * d3
    assert abs((e1 - e0) - ref) < 1e-10
    for a in range(3):
        d = 1e-6
        p1, p2 = pos.copy(), pos.copy()
        p1[3, a] += d
        p2[3, a] -= d
        num = -(_E(p1, h, mol, pcc2, 0.0)[0] - _E(p2, h, mol, pcc2, 0.0)[0]) / (2 * d)
        assert abs(num - F1[3, a]) < 1e-6 * max(1.0, abs(num))
