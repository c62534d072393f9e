"""FP64 numpy oracle of the learned REBO surrogate (test infrastructure).

Only ``tests/`` and ``bench.py``'s ``cpu_baseline`` leg may import this module;
the GPU kernels in ``fuller_md_en_amd/csrc/surrogate_kernels.hip`` are checked
against it.  The surrogate has no counterpart in the reference: it is a
per-atom energy model E = sum_i MLP(D_i) whose forces are exact gradients,
standing in for the intramolecular REBO term (compute_rebo AB:275-421).

Descriptor of atom i (same-molecule neighbours j, d = x_j - x_i, r = |d|,
u = d / r):
  radial basis b_k(r) = T_k(x(r)) fc(r), k < 16: Chebyshev polynomials of
      x = 2 (s - s0) / (1 - s0) - 1, s = exp(-(r - 1) / 1.5), s0 = s(6 A),
      cut by fc(r) = (cos(pi r / 6) + 1) / 2 for r < 6 A;
  A_k = sum_j b_k                               (16)
  B_k = |sum_j b_k u|^2,              k < 8     (8)
  C_k = ||sum_j b_k u (x) u||_F^2,    k < 8     (8)
  bonded shell w(r) = (cos(pi t) + 1) / 2, t = (r - 1.8) / 0.4 clipped to
      [0, 1], channels w and w (r - 1.42):
  I_cl = ||sum_j w_c u^(x)l||^2,   l = 1..4     (8)
40 features, normalised as (D - mu) / sd, then 40 -> 64 -> 64 -> 64 -> 1 with
tanh.  Every invariant ||sum_j w_j u_j^(x)l||^2 equals
sum_jj' w_j w_j' cos^l(theta_jj'), so B, C and I are angular sums.

Gradient of an invariant of rank l with respect to one pair vector d:
  d/dd ||M||^2 = 2 w' M(u..u) u + 2 l (w / r) (M(u..u, .) - M(u..u) u)
(M(u..u, .) contracts all but the last index with u).
"""
from __future__ import annotations

import numpy as np

RC = 6.0        # descriptor cutoff (A)
NB = 16         # radial bases
NV = 8          # radial channels carrying the l = 1, 2 moments
RB0, RB1 = 1.8, 2.2   # bonded-shell switch (A)
R0B = 1.42      # bond length of the second bonded channel
LB = 4          # bonded-shell ranks 1..LB
NF = NB + 2 * NV + 2 * LB   # 40 features
H = 64          # hidden width


def radial(r):
    """b_k(r), db_k/dr for k < NB."""
    s0 = np.exp(-(RC - 1.0) / 1.5)
    s = np.exp(-(r - 1.0) / 1.5)
    x = 2.0 * (s - s0) / (1.0 - s0) - 1.0
    dx = 2.0 / (1.0 - s0) * (-s / 1.5)
    T = np.zeros(r.shape + (NB,))
    dT = np.zeros_like(T)
    T[..., 0] = 1.0
    T[..., 1] = x
    dT[..., 1] = 1.0
    for k in range(2, NB):
        T[..., k] = 2 * x * T[..., k - 1] - T[..., k - 2]
        dT[..., k] = 2 * T[..., k - 1] + 2 * x * dT[..., k - 1] - dT[..., k - 2]
    inside = r < RC
    fc = np.where(inside, 0.5 * (np.cos(np.pi * r / RC) + 1.0), 0.0)
    dfc = np.where(inside, -0.5 * np.pi / RC * np.sin(np.pi * r / RC), 0.0)
    b = T * fc[..., None]
    db = dT * (dx * fc)[..., None] + T * dfc[..., None]
    return b, db


def bond_channels(r):
    """[(w, dw/dr), (w (r - R0B), d/dr)] of the bonded shell."""
    t = np.clip((r - RB0) / (RB1 - RB0), 0.0, 1.0)
    w = 0.5 * (np.cos(np.pi * t) + 1.0)
    dw = np.where((r > RB0) & (r < RB1), -0.5 * np.pi / (RB1 - RB0) * np.sin(np.pi * t), 0.0)
    return [(w, dw), (w * (r - R0B), dw * (r - R0B) + w)]


def pairs(x, mol_start):
    """Ordered same-molecule pairs within RC: I, J, d = x_J - x_I, r."""
    Is, Js = [], []
    for m in range(len(mol_start) - 1):
        a0, a1 = mol_start[m], mol_start[m + 1]
        idx = np.arange(a0, a1)
        i, j = np.meshgrid(idx, idx, indexing="ij")
        keep = i != j
        Is.append(i[keep])
        Js.append(j[keep])
    I = np.concatenate(Is)
    J = np.concatenate(Js)
    d = x[J] - x[I]
    r = np.sqrt((d * d).sum(1))
    k = r < RC
    return I[k], J[k], d[k], r[k]


def _rank(u, l):
    t = u
    for _ in range(l - 1):
        t = np.einsum("p...,pc->p...c", t, u)
    return t


def descriptor(x, mol_start, jac=False):
    """D [n, NF]; with jac, also (I, J, c) with c[p] = dD_I / dd_p, [P, NF, 3]."""
    n = len(x)
    I, J, d, r = pairs(np.asarray(x, float), np.asarray(mol_start))
    u = d / r[:, None]
    b, db = radial(r)
    feats, cols = [], []
    A = np.zeros((n, NB))
    np.add.at(A, I, b)
    feats.append(A)
    cols.append(db[:, :, None] * u[:, None, :])
    for l in (1, 2):
        ut = _rank(u, l)
        bl = b[:, :NV].reshape((-1, NV) + (1,) * l)
        M = np.zeros((n, NV) + (3,) * l)
        np.add.at(M, I, bl * ut[:, None])
        feats.append((M.reshape(n, NV, -1) ** 2).sum(-1))
        if jac:
            Mu = M[I]
            for _ in range(l - 1):
                Mu = np.einsum("pk...c,pc->pk...", Mu, u)
            Mll = (Mu * u[:, None, :]).sum(-1)
            cols.append(2 * db[:, :NV, None] * Mll[:, :, None] * u[:, None, :] +
                        2 * l * (b[:, :NV] / r[:, None])[:, :, None] *
                        (Mu - Mll[:, :, None] * u[:, None, :]))
    for w, dw in bond_channels(r):
        for l in range(1, LB + 1):
            ut = _rank(u, l)
            M = np.zeros((n,) + (3,) * l)
            np.add.at(M, I, w.reshape((-1,) + (1,) * l) * ut)
            feats.append((M.reshape(n, -1) ** 2).sum(1)[:, None])
            if jac:
                Mu = M[I]
                for _ in range(l - 1):
                    Mu = np.einsum("p...c,pc->p...", Mu, u)
                Mll = (Mu * u).sum(-1)
                cols.append((2 * dw[:, None] * Mll[:, None] * u +
                             2 * l * (w / r)[:, None] * (Mu - Mll[:, None] * u))[:, None, :])
    D = np.concatenate(feats, 1)
    if not jac:
        return D
    return D, (I, J, np.concatenate(cols, 1))


# ------------------------------------------------------------------ network
def to_bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0, mu=None, sd=None):
    rng = np.random.default_rng(seed)

    def lin(a, b):
        return rng.uniform(-1, 1, (a, b)) * np.sqrt(6.0 / (a + b))
    p = {"W1": lin(NF, H), "b1": np.zeros(H), "W2": lin(H, H), "b2": np.zeros(H),
         "W3": lin(H, H), "b3": np.zeros(H), "w4": lin(H, 1)[:, 0], "b4": np.zeros(1)}
    p["mu"] = np.zeros(NF) if mu is None else np.asarray(mu, float)
    p["sd"] = np.ones(NF) if sd is None else np.asarray(sd, float)
    return p


TRAINED = ("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4")


def mlp(p, D, bf16=False):
    """Per-atom energies and dE_i/dD_i.  bf16: round weights and every
    GEMM operand to bfloat16 as the MFMA kernels do (FP32 accumulation is
    emulated in FP64)."""
    q = (lambda a: to_bf16(a)) if bf16 else (lambda a: a)
    h0 = (D - p["mu"]) / p["sd"]
    z1 = q(h0) @ q(p["W1"]) + p["b1"]
    h1 = np.tanh(z1)
    z2 = q(h1) @ q(p["W2"]) + p["b2"]
    h2 = np.tanh(z2)
    z3 = q(h2) @ q(p["W3"]) + p["b3"]
    h3 = np.tanh(z3)
    E = h3 @ q(p["w4"]) + p["b4"][0]
    d3 = q(p["w4"])[None, :] * (1 - h3 * h3)
    d2 = (q(d3) @ q(p["W3"]).T) * (1 - h2 * h2)
    d1 = (q(d2) @ q(p["W2"]).T) * (1 - h1 * h1)
    g = (q(d1) @ q(p["W1"]).T) / p["sd"]
    return E, g


def energy_forces(p, x, mol_start, bf16=False):
    """Total energy, forces F = -dE/dx [n, 3] and virial W = sum_p d (x) f_p."""
    D, (I, J, c) = descriptor(x, mol_start, jac=True)
    E, g = mlp(p, D, bf16)
    f = -np.einsum("pk,pkc->pc", g[I], c)      # -dE/dd_p: force on J from pair p
    F = np.zeros((len(x), 3))
    np.add.at(F, J, f)
    np.add.at(F, I, -f)
    d = x[J] - x[I]
    W = np.einsum("pa,pb->ab", d, f).reshape(9)
    return E.sum(), F, W, E


def loss_grad(p, x, mol_start, Fref, bf16=False):
    """L = mean (F - Fref)^2 over the 3n components and dL/dtheta for the
    trained parameters (forward-mode tangent of dE/dD along the residual's
    descriptor projection, reverse-mode through it).  bf16: the forces of
    the loss with the MFMA kernels' BF16 rounding (the gradient sweep stays
    FP64)."""
    D, (I, J, c) = descriptor(x, mol_start, jac=True)
    n = len(x)
    E, g = mlp(p, D, bf16)
    f = -np.einsum("pk,pkc->pc", g[I], c)
    F = np.zeros((n, 3))
    np.add.at(F, J, f)
    np.add.at(F, I, -f)
    e = F - Fref
    L = (e * e).mean()
    # u_i = sum_p c_p (e_J - e_I), in normalised-feature units
    u = np.zeros((n, NF))
    np.add.at(u, I, np.einsum("pkc,pc->pk", c, e[J] - e[I]))
    ut = u / p["sd"]
    h0 = (D - p["mu"]) / p["sd"]
    W1, W2, W3, w4 = p["W1"], p["W2"], p["W3"], p["w4"]
    z1 = h0 @ W1 + p["b1"]; h1 = np.tanh(z1)
    z2 = h1 @ W2 + p["b2"]; h2 = np.tanh(z2)
    z3 = h2 @ W3 + p["b3"]; h3 = np.tanh(z3)
    t0 = ut
    s1 = t0 @ W1; t1 = (1 - h1 * h1) * s1
    s2 = t1 @ W2; t2 = (1 - h2 * h2) * s2
    s3 = t2 @ W3; t3 = (1 - h3 * h3) * s3
    G = {"w4": t3.sum(0), "b4": np.zeros(1)}
    lt = np.broadcast_to(w4, h3.shape)
    lp = np.zeros_like(h3)
    for k, (h, s, tin, hin, W) in enumerate(((h3, s3, t2, h2, W3), (h2, s2, t1, h1, W2),
                                             (h1, s1, t0, h0, W1))):
        zt = lt * (1 - h * h)
        lp = lp + lt * (-2 * h * s)
        zp = lp * (1 - h * h)
        name = 3 - k
        G[f"W{name}"] = tin.T @ zt + hin.T @ zp
        G[f"b{name}"] = zp.sum(0)
        lt = zt @ W.T
        lp = zp @ W.T
    scale = -2.0 / (3 * n)
    return L, {k: scale * v for k, v in G.items()}, F
