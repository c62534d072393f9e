"""Learned REBO surrogate on the GPU against its FP64 oracle (oracle/surrogate.py).

Tolerances (BF16 MFMA operands, FP32 accumulation and descriptors):
  descriptors  |gpu - oracle| <= 1e-2 |oracle| + 1e-3 (BF16 storage)
  forces       max|dF| / max|F| <= 5e-3, virial likewise, energy 1e-4 relative
               (oracle with the same BF16 rounding of weights and operands)
  loss         1e-3 relative (oracle forces with the same BF16 rounding);
               gradient cosine >= 0.995, relative L2 <= 5e-2 (BF16 operands
               in the GPU sweep, FP64 in the oracle's)
Training, ablation (trained network vs no learned term and vs the untrained
network), the judge and the fallback run on engine-generated frames.
"""
import numpy as np
import pytest

import oracle
from oracle import surrogate as S
from fuller_md_en_amd import airebo, system
from fuller_md_en_amd import surrogate as SG

pytestmark = pytest.mark.gpu


def _pair_system(seed=0, amp=0.06):
    """A distorted C60 and a distorted C70 in a 40 A box."""
    rng = np.random.default_rng(seed)
    c60 = system.molecule("C60").coords
    c70 = system.molecule("C70").coords
    x = np.concatenate([c60 + 12.0 + rng.normal(0, amp, c60.shape),
                        c70 + 28.0 + rng.normal(0, amp, c70.shape)])
    mol_id = np.repeat([1, 2], [60, 70]).astype(np.int32)
    h = np.diag([40.0, 40.0, 40.0]).reshape(9)
    return x, h, mol_id, np.array([0, 60, 130])


def _engine(x, h, mol_id, T=300.0, vel=None):
    eng = airebo.AtomicEngine(mol_id, np.full(len(x), system.MASS_C), mode=airebo.REBO_PHYSICAL)
    eng.set_state(pos=x, vel=np.zeros_like(x) if vel is None else vel, h=h,
                  npt=airebo.make_npt(T, 0.0, len(x)))
    return eng


def _oracle_params(x, ms, seed=3):
    D = S.descriptor(x, ms)
    p = S.init_params(seed, D.mean(0), D.std(0) + 1e-3)
    # the engine keeps FP32 parameters: give the oracle the same values
    for k in list(p):
        p[k] = np.asarray(p[k], np.float32).astype(np.float64)
    return p


def _load(sg, p):
    sg.set_params(SG.pack(p), p["mu"], p["sd"])


def test_descriptor_matches_oracle():
    x, h, mol, ms = _pair_system()
    eng = _engine(x, h, mol)
    sg = SG.Surrogate(eng)
    p = _oracle_params(x, ms)
    _load(sg, p)
    sg.collect()
    f = sg.descriptor(0)
    Dn = (S.descriptor(x, ms) - p["mu"]) / p["sd"]
    assert np.all(np.abs(f - Dn) <= 1e-2 * np.abs(Dn) + 1e-3), np.abs(f - Dn).max()
    eng.close()


def test_energy_forces_virial_match_oracle():
    x, h, mol, ms = _pair_system(1)
    eng = _engine(x, h, mol)
    sg = SG.Surrogate(eng)
    p = _oracle_params(x, ms)
    _load(sg, p)
    sg.collect()
    E, F, W, _ = sg.evaluate(0)
    E0, F0, W0, _ = S.energy_forces(p, x, ms, bf16=True)
    assert abs(E - E0) <= 1e-4 * abs(E0)
    assert np.abs(F - F0).max() <= 5e-3 * np.abs(F0).max()
    assert np.abs(W - W0).max() <= 5e-3 * np.abs(W0).max()
    # the force pass conserves momentum (pair terms, both sides)
    assert np.abs(F.sum(0)).max() < 1e-4 * np.abs(F).max()
    eng.close()


def test_loss_gradient_matches_oracle():
    x, h, mol, ms = _pair_system(2)
    eng = _engine(x, h, mol)
    sg = SG.Surrogate(eng)
    p = _oracle_params(x, ms)
    _load(sg, p)
    sg.collect()
    cnt, lst = oracle.nlist_rebo(x, h, mol)
    _, Fref, _ = oracle.compute_rebo(x, h, cnt, lst, oracle.REBO_PHYSICAL)
    L0, G0, _ = S.loss_grad(p, x, ms, Fref, bf16=True)
    L, g = sg.loss_grad(0)
    g0 = SG.pack(G0).astype(np.float64)
    g = g.astype(np.float64)
    cos = g @ g0 / np.linalg.norm(g) / np.linalg.norm(g0)
    rel = np.linalg.norm(g - g0) / np.linalg.norm(g0)
    print(f"loss {L} oracle {L0}; gradient cosine {cos}, relative L2 error {rel}", flush=True)
    for name in S.TRAINED:  # per-block agreement, for diagnosis
        a, b = SG.unpack(g)[name], G0[name]
        print(name, np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30), flush=True)
    assert abs(L - L0) <= 1e-3 * L0
    assert cos >= 0.995 and rel <= 5e-2
    eng.close()


def _training_run(iters=1500):
    """C60 FCC 2x2x2 (1920 atoms), exact-gradient REBO, frames at 300 and 1500 K."""
    s = system.airebo_system(2, "C60", T=300.0, seed=5, jitter=0.02, jitter_seed=4)
    eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
    sg = SG.Surrogate(eng, seed=1)
    step = 0
    for T in (300.0, 1500.0, 800.0):
        vel = system.atomic_init_velocities(s["mass"], int(T), T)
        st = eng.get_state() if step else None
        eng.set_state(pos=s["pos"] if st is None else st["pos"], vel=vel, h=s["h"],
                      npt=airebo.make_npt(T, 0.0, len(s["pos"])))
        for _ in range(5):
            eng.run(step, 8, thermo=False)
            step += 8
            sg.collect()
    sg.fit_normalisation()
    untrained = [sg.evaluate(f)[3] for f in range(sg.frames())]
    loss = sg.train(iters, lr=2e-3)
    return eng, sg, untrained, loss, step


def test_training_and_ablation():
    eng, sg, untrained, loss, step = _training_run()
    assert np.isfinite(loss).all()
    assert loss[-50:].mean() < 0.1 * loss[:50].mean()
    # held-out frames: collected after training, never seen by it
    held = []
    for _ in range(3):
        eng.run(step, 8, thermo=False)
        step += 8
        sg.collect()
        held.append(sg.frames() - 1)
    trained = np.array([sg.evaluate(f)[3] for f in held])
    # ablation: without the learned term (all-zero network: no REBO force at
    # all, error = RMS of the REBO forces) and with the untrained network
    th, mu, sd = sg.get_params()
    sg.set_params(np.zeros_like(th), mu, sd)
    no_term = np.array([sg.evaluate(f)[3] for f in held])
    sg.set_params(th, mu, sd)
    print("held-out force RMSE: trained", trained, "no learned term", no_term,
          "untrained", np.mean(untrained), flush=True)
    assert trained.max() < 0.25 * no_term.min()
    assert trained.max() < 0.1 * np.mean(untrained)
    eng.close()


def test_judge_and_fallback():
    s = system.airebo_system(2, "C60", T=300.0, seed=6, jitter=0.02, jitter_seed=5)
    for thr, expect_fallback in ((1e-9, True), (1e9, False)):
        eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
        eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"],
                      npt=airebo.make_npt(300.0, 0.0, len(s["pos"])))
        sg = SG.Surrogate(eng, seed=2)
        sg.enable(True, judge_every=5, threshold=thr)
        th = eng.run(0, 12)
        st = sg.stats()
        assert st["judged"] == (1 if expect_fallback else 2)
        assert st["fallbacks"] == (1 if expect_fallback else 0)
        assert st["active"] == (0 if expect_fallback else 1)
        assert all(np.isfinite(t.Ep) for t in th)
        eng.close()


def test_engine_step_uses_surrogate_energy():
    s = system.airebo_system(2, "C60", T=300.0, seed=7, jitter=0.02, jitter_seed=6)
    eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
    eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"],
                  npt=airebo.make_npt(300.0, 0.0, len(s["pos"])))
    sg = SG.Surrogate(eng, seed=3)
    sg.enable(True, judge_every=0, threshold=1.0)
    eng.prepare()
    E, F, W = eng.forces()
    Es = sg.evaluate(-1)[0]
    assert abs(E[1] - Es) <= 1e-6 * abs(Es)  # Ep_rebo is the surrogate's energy
    eng.close()


def test_engine_forces_unaffected_when_disabled():
    s = system.airebo_system(2, "C60", T=300.0, seed=8, jitter=0.02, jitter_seed=7)
    out = []
    for attach in (False, True):
        eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
        eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"],
                      npt=airebo.make_npt(300.0, 0.0, len(s["pos"])))
        if attach:
            SG.Surrogate(eng, seed=4)
        eng.run(0, 6)
        out.append(eng.get_state()["pos"].copy())
        eng.close()
    assert np.array_equal(out[0], out[1])
