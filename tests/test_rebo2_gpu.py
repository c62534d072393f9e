"""The REBO-II P_ij spline and the AIREBO torsion term on the GPU (engines,
exact-gradient REBO), against the FP64 oracle's independent formulation
(orc_rebo_ext, pinned by FD gradients and closed forms in
test_oracle_rebo2.py).  The reference has neither term, so these are not
reference parity; the GPU bar is the oracle's, max |a - b| / max |b| <= 1e-5
for forces and virial, 1e-5 relative for energies.

The configurations scale one molecule of a jittered 2x2x2 crystal by 1.25,
so that its bonds (1.75-1.81 A) lie in the 1.7-2.0 A switch, where the P_ij
slope and the torsion weights' derivatives carry force.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_max
from fuller_md_en_amd import airebo, rigid
from test_airebo_gpu import crystal, rel

pytestmark = pytest.mark.gpu

TOL = 1e-5
PCC = airebo.AIREBO_PCC
EPS_T = airebo.AIREBO_TORSION_EPS


def stretched(nc=2, seed=11, jitter=0.02, scale=1.25):
    pos, h, mol = crystal(nc, seed, jitter)
    m1 = mol == 1
    L = np.diag(h.reshape(3, 3))
    d = pos[m1] - pos[m1][0]
    d -= L * np.round(d / L)
    c = d.mean(0)
    pos = pos.copy()
    p = pos[m1][0] + c + scale * (d - c)
    pos[m1] = p - L * np.floor(p / L)
    return pos, h, mol


def engine(pos, h, mol, lj=None, **kw):
    eng = airebo.AtomicEngine(mol, np.full(len(pos), 12.011), mode=airebo.REBO_PHYSICAL,
                              lj=lj, **kw)
    eng.set_state(pos=pos, vel=np.zeros_like(pos), h=h,
                  npt=airebo.make_npt(300.0, 0.0, len(pos)))
    eng.prepare()
    return eng


@pytest.mark.parametrize("terms", ["pcc", "torsion", "both"])
def test_extension_forces_against_oracle(terms):
    """REBO alone (LJ well depth 1e-12 eV): the stretched molecule's switch
    regions and the intact ones."""
    pcc = PCC if terms in ("pcc", "both") else None
    eps_t = EPS_T if terms in ("torsion", "both") else 0.0
    print(f"case terms={terms}", flush=True)
    pos, h, mol = stretched()
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    E0, F0, W0, Et0 = oracle.rebo_ext(pos, h, cr, lr, pcc or (0.0,) * 5, eps_t)
    eng = engine(pos, h, mol, lj=rigid.lj_params(3.4, 1e-12, 10.2), pcc=pcc,
                 torsion_eps=eps_t)
    E, F, W = eng.forces()
    assert eng.counters()[0] == 0
    eng.close()
    assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
    assert rel(E[1], E0) < TOL  # Ep_rebo holds the torsion energy
    if eps_t:
        assert Et0 > 1.0  # the term is not negligible here


def test_extension_with_lj_against_oracle():
    """The whole force field with both terms: REBO-II extension + LJ with the
    reference's truncated half list, on a jittered 2x2x2 crystal."""
    pos, h, mol = crystal(2, 13, 0.05)
    cr, lr = oracle.nlist_rebo(pos, h, mol)
    cl, ll = oracle.nlist_lj(pos, h, mol)
    Er, Fr, Wr, _ = oracle.rebo_ext(pos, h, cr, lr, PCC, EPS_T)
    El, Fl, Wl = oracle.compute_lj(pos, h, cl, ll)
    eng = engine(pos, h, mol, pcc=PCC, torsion_eps=EPS_T)
    E, F, W = eng.forces()
    eng.close()
    assert rel_max(F, Fr + Fl) < TOL and rel_max(W, Wr + Wl) < TOL
    for a, b in zip(E, (Er + El, Er, El)):
        assert rel(a, b) < TOL


def test_extension_lds_and_global_rebo_paths_agree():
    """A 2.6 A REBO list sends every molecule down the REBO kernel's global
    path; the P_ij slope travels there in the entry arrays instead of
    registers, with the same results."""
    pos, h, mol = stretched()
    out = []
    for rl in (2.3, 2.6):
        print(f"case rebo_rlist={rl}", flush=True)
        eng = engine(pos, h, mol, rebo_rlist=rl, pcc=PCC, torsion_eps=EPS_T)
        out.append(eng.forces())
        eng.close()
    (Ea, Fa, Wa), (Eb, Fb, Wb) = out
    assert rel_max(Fa, Fb) < 1e-12 and rel_max(Wa, Wb) < 1e-12
    for a, b in zip(Ea, Eb):
        assert rel(a, b) < 1e-12


def test_zero_extension_is_the_reference_potential():
    """Knots and well depth all zero leave the forces bit for bit those of
    the engine without the terms."""
    pos, h, mol = stretched()
    a = engine(pos, h, mol)
    b = engine(pos, h, mol, pcc=(0.0,) * 5, torsion_eps=0.0)
    (Ea, Fa, Wa), (Eb, Fb, Wb) = a.forces(), b.forces()
    a.close()
    b.close()
    assert np.array_equal(Fa, Fb) and np.array_equal(Wa, Wb) and Ea == Eb


def test_extension_needs_the_exact_gradient_mode():
    pos, h, mol = crystal(1, 3, 0.01)
    with pytest.raises(RuntimeError, match="FMD_REBO_PHYSICAL"):
        airebo.AtomicEngine(mol, np.full(len(pos), 12.011), mode=airebo.REBO_COMPAT,
                            torsion_eps=EPS_T)


def test_extension_conserves_energy():
    """With the thermostat and barostat frozen and the LJ list untruncated,
    400 steps of 0.5 fs from a 300 K crystal keep Ep + KE as well with both
    terms as without them (their forces are the gradient of their energy):
    the largest excursion from the start stays within twice the plain
    engine's, or 2e-5 eV per atom."""
    from test_airebo_gpu import _frozen_npt
    pos, h, mol = crystal(2, 17, 0.02)
    n = len(pos)
    vel = np.random.default_rng(5).normal(0.0, 0.0045, (n, 3))
    vel -= vel.mean(0)
    drift = []
    for pcc, eps_t in ((None, 0.0), (PCC, EPS_T)):
        print(f"case torsion_eps={eps_t}", flush=True)
        eng = airebo.AtomicEngine(mol, np.full(n, 12.011), mode=airebo.REBO_PHYSICAL,
                                  lj_max_neigh=0, pcc=pcc, torsion_eps=eps_t)
        eng.set_state(pos=pos, vel=vel, h=h, npt=_frozen_npt(n))
        th = eng.run(0, 400)
        eng.close()
        tot = np.array([t.Ep + t.KE for t in th])
        assert np.isfinite(tot).all()
        drift.append(np.abs(tot - tot[0]).max() / n)
    print(f"max |E - E0| per atom: plain {drift[0]:.2e}, with the terms {drift[1]:.2e} eV")
    assert drift[1] < max(2 * drift[0], 2e-5)
