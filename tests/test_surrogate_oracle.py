"""The FP64 oracle of the learned REBO surrogate pinned by mathematics:
forces are minus the finite-difference gradient of its energy, the virial is
minus the strain derivative, the loss gradient is the finite-difference
gradient of the force loss, the invariants equal their angular-sum closed
forms, and the descriptor is rotation and permutation invariant."""
import numpy as np
import pytest

from oracle import surrogate as S
from fuller_md_en_amd import system


def _two_molecules(seed=0, amp=0.08):
    rng = np.random.default_rng(seed)
    c60 = system.molecule("C60").coords
    c70 = system.molecule("C70").coords
    x = np.concatenate([c60 + rng.normal(0, amp, c60.shape),
                        c70 + 20.0 + rng.normal(0, amp, c70.shape)])
    return x, np.array([0, 60, 130])


def _params(seed=1):
    x, ms = _two_molecules(5)
    D = S.descriptor(x, ms)
    return S.init_params(seed, D.mean(0), D.std(0) + 1e-3)


def test_forces_are_minus_energy_gradient():
    x, ms = _two_molecules()
    p = _params()
    E, F, W, _ = S.energy_forces(p, x, ms)
    eps = 1e-5
    for a in (0, 17, 59, 60, 101):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[a, c] += eps
            xm[a, c] -= eps
            fd = -(S.energy_forces(p, xp, ms)[0] - S.energy_forces(p, xm, ms)[0]) / (2 * eps)
            assert abs(fd - F[a, c]) < 1e-6 * max(1.0, abs(fd))
    # no net force or torque from an internal energy
    assert np.abs(F.sum(0)).max() < 1e-9
    assert np.abs(np.cross(x, F).sum(0)).max() < 1e-8


def test_virial_is_minus_strain_derivative():
    x, ms = _two_molecules(2)
    p = _params()
    _, _, W, _ = S.energy_forces(p, x, ms)
    eps = 1e-6
    for a in range(3):
        for b in range(3):
            e = np.zeros((3, 3))
            e[a, b] = eps
            Ep = S.energy_forces(p, x @ (np.eye(3) + e).T, ms)[0]
            Em = S.energy_forces(p, x @ (np.eye(3) - e).T, ms)[0]
            assert abs(-(Ep - Em) / (2 * eps) - W[3 * a + b]) < 1e-5 * max(1.0, np.abs(W).max())


def test_loss_gradient_matches_finite_differences():
    x, ms = _two_molecules(3)
    p = _params()
    Fref = np.random.default_rng(9).normal(0, 1, x.shape)
    L, G, _ = S.loss_grad(p, x, ms, Fref)
    rng = np.random.default_rng(4)
    eps = 1e-6
    for name in S.TRAINED:
        for _ in range(3):
            idx = tuple(rng.integers(0, s) for s in p[name].shape)
            q = {k: v.copy() for k, v in p.items()}
            q[name][idx] += eps
            Lp = S.loss_grad(q, x, ms, Fref)[0]
            q[name][idx] -= 2 * eps
            Lm = S.loss_grad(q, x, ms, Fref)[0]
            fd = (Lp - Lm) / (2 * eps)
            assert abs(fd - G[name][idx]) < 1e-6 * max(1e-3, abs(fd)), (name, idx, fd, G[name][idx])


def test_invariants_equal_angular_sums():
    rng = np.random.default_rng(7)
    x = rng.normal(0, 1.2, (9, 3))
    ms = np.array([0, 9])
    D = S.descriptor(x, ms)
    for i in range(9):
        d = np.delete(x, i, 0) - x[i]
        r = np.linalg.norm(d, axis=1)
        u = d / r[:, None]
        cos = u @ u.T
        b, _ = S.radial(r)
        for k in range(S.NV):
            w = b[:, k]
            assert np.isclose(D[i, S.NB + k], w @ cos @ w, rtol=1e-12, atol=1e-12)
            assert np.isclose(D[i, S.NB + S.NV + k], w @ cos ** 2 @ w, rtol=1e-12, atol=1e-12)
        for ch, (w, _) in enumerate(S.bond_channels(r)):
            for l in range(1, S.LB + 1):
                assert np.isclose(D[i, S.NB + 2 * S.NV + ch * S.LB + l - 1], w @ cos ** l @ w,
                                  rtol=1e-12, atol=1e-12)


def test_descriptor_rotation_and_permutation_invariant():
    x, ms = _two_molecules(6)
    D = S.descriptor(x, ms)
    q = np.linalg.qr(np.random.default_rng(1).normal(size=(3, 3)))[0]
    assert np.allclose(S.descriptor(x @ q.T + 3.0, ms), D, rtol=1e-10, atol=1e-10)
    perm = np.concatenate([np.random.default_rng(2).permutation(60), 60 + np.arange(70)])
    assert np.allclose(S.descriptor(x[perm], ms), D[perm], rtol=1e-10, atol=1e-10)


def test_chebyshev_radial_basis_closed_form():
    r = np.linspace(1.0, 5.9, 50)
    b, db = S.radial(r)
    s0 = np.exp(-5 / 1.5)
    xx = 2 * (np.exp(-(r - 1) / 1.5) - s0) / (1 - s0) - 1
    fc = 0.5 * (np.cos(np.pi * r / 6) + 1)
    for k in range(S.NB):
        assert np.allclose(b[:, k], np.cos(k * np.arccos(np.clip(xx, -1, 1))) * fc, atol=1e-12)
    eps = 1e-6
    fd = (S.radial(r + eps)[0] - S.radial(r - eps)[0]) / (2 * eps)
    assert np.allclose(fd, db, atol=1e-6)


def test_bf16_rounding():
    a = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0])
    r = S.to_bf16(a)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125  # ties to even
    assert abs(r[3] + 3.140625) < 1e-12
