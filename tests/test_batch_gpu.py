"""Batched AIREBO replicas (fmd_batch_*): R independent NPT systems of
uneven size, species, temperature and pressure in one engine.

* Each replica of a batch evolves bit for bit as the same system run alone
  in its own engine (same kernels, replica-local reduction order), across a
  list rebuild, with graph replay on and off: a batch with tiles wider than
  64 atoms (two atoms per lane) and a batch of C60 replicas (one per lane).
* Forces of every replica against the FP64 oracle (compat REBO), 1e-5.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_max
from fuller_md_en_amd import airebo, system

pytestmark = pytest.mark.gpu

TOL = 1e-5

# (species, nc, T, P, jitter seed): C70 tiles are wider than C60's, so the
# batch pads every molecule tile to 72 atoms (two atoms per lane)
SPECS = [("C60", 3, 300.0, 0.0, 1), ("C70", 3, 900.0, 2.0, 2), ("C60", 3, 1500.0, 5.0, 3)]


def _systems():
    out = []
    for spec, nc, T, P, js in SPECS:
        s = system.airebo_system(nc, spec, T=T, seed=40 + js, jitter=0.01, jitter_seed=js)
        s["T"], s["P"] = T, P
        out.append(s)
    return out


def _systems_c60():
    """Two C60 replicas of different size, temperature and pressure."""
    out = []
    for nc, T, P, js in ((3, 300.0, 0.0, 1), (4, 1200.0, 3.0, 4)):
        s = system.airebo_system(nc, "C60", T=T, seed=60 + js, jitter=0.01, jitter_seed=js)
        s["T"], s["P"] = T, P
        out.append(s)
    return out


def _npt(s):
    return airebo.make_npt(s["T"], s["P"], len(s["pos"]))


def _run_batch(systems, nsteps, graph):
    b = airebo.ReplicaBatch([(s["mol_id"], s["mass"]) for s in systems], graph=graph)
    for r, s in enumerate(systems):
        b.set_state(r, pos=s["pos"], vel=s["vel"], h=s["h"], npt=_npt(s))
    th = b.run(0, nsteps)
    st = [b.get_state(r) for r in range(len(systems))]
    ep = np.array([[[t.Ep, t.KE, t.P, t.V] for t in row] for row in th])
    assert b.counters()[0] == 0
    b.close()
    return st, ep


@pytest.mark.parametrize("specs", ["mixed", "c60"])
def test_batch_replicas_equal_single_runs(specs):
    systems = _systems()
    if specs == "c60":
        systems = [systems[0], _systems_c60()[1]]
    nsteps = 25  # crosses the list rebuild at step 20
    st_b, ep_b = _run_batch(systems, nsteps, graph=True)
    for r, s in enumerate(systems):
        print(f"case {specs} replica {r}", flush=True)
        eng = airebo.AtomicEngine(s["mol_id"], s["mass"])
        eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"], npt=_npt(s))
        th = eng.run(0, nsteps)
        st = eng.get_state()
        eng.close()
        ep = np.array([[t.Ep, t.KE, t.P, t.V] for t in th])
        assert np.array_equal(ep, ep_b[:, r]), r
        assert np.array_equal(st["pos"], st_b[r]["pos"]) and np.array_equal(st["vel"], st_b[r]["vel"])
        assert np.array_equal(st["h"], st_b[r]["h"])


def test_batch_graph_replay_equals_launches():
    systems = _systems()[:2]
    st1, ep1 = _run_batch(systems, 23, graph=True)
    st2, ep2 = _run_batch(systems, 23, graph=False)
    assert np.array_equal(ep1, ep2)
    for a, b in zip(st1, st2):
        assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])


def _thermo_row(t):
    return [t.Ep, t.KE, t.T, t.P, t.V, t.xi] + list(t.h)


@pytest.mark.parametrize("graph", [True, False])
def test_folded_finish_equals_step_by_step(graph):
    """Within a run each step's finish (thermostat, barostat, record, velocity
    rescale) is folded into the next step's kick-drift; one call per step
    finishes every step on its own.  Both give the same records and state bit
    for bit, with a rescale of one replica inside the run (global steps
    11-25, a list rebuild at step 20)."""
    from fuller_md_en_amd import _native as N
    systems = _systems_c60()
    nsteps, R = 15, 2
    ctrl = [[N.StepCtrl(s["T"], 0.0, 0, 0) for s in systems] for _ in range(nsteps)]
    ctrl[4][0] = N.StepCtrl(systems[0]["T"], 450.0, N.CTRL_RESCALE, 0)
    ctrl[12][1] = N.StepCtrl(systems[1]["T"], 700.0, N.CTRL_RESCALE | N.CTRL_RESCALE_ZERO_VG, 0)
    out = []
    for per_step in (False, True):
        print(f"case graph={graph} per_step={per_step}", flush=True)
        b = airebo.ReplicaBatch([(s["mol_id"], s["mass"]) for s in systems], graph=graph)
        for r, s in enumerate(systems):
            b.set_state(r, pos=s["pos"], vel=s["vel"], h=s["h"], npt=_npt(s))
        if per_step:
            th = [b.run(10 + k, 1, ctrl=[ctrl[k]])[0] for k in range(nsteps)]
        else:
            th = b.run(10, nsteps, ctrl=ctrl)
        ep = np.array([[_thermo_row(t) for t in row] for row in th])
        out.append((ep, [b.get_state(r) for r in range(R)]))
        assert b.counters()[0] == 0
        b.close()
    (ep1, st1), (ep2, st2) = out
    assert np.array_equal(ep1, ep2)
    for a, b in zip(st1, st2):
        assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
        assert np.array_equal(a["h"], b["h"])
        assert a["npt"].xi == b["npt"].xi and list(a["npt"].Vg) == list(b["npt"].Vg)
    # the rescaled steps' records carry the targets
    assert abs(ep1[4, 0, 2] - 450.0) < 1e-9 and abs(ep1[12, 1, 2] - 700.0) < 1e-9


def test_batch_replica_forces_match_oracle():
    systems = _systems()
    b = airebo.ReplicaBatch([(s["mol_id"], s["mass"]) for s in systems])
    for r, s in enumerate(systems):
        b.set_state(r, pos=s["pos"], vel=s["vel"], h=s["h"], npt=_npt(s))
    b.prepare()
    for r, s in enumerate(systems):
        print(f"replica {r} {SPECS[r]}", flush=True)
        pos = s["pos"].reshape(-1).copy()
        oracle.lib().orc_apply_pbc(pos, s["h"], oracle.inv9(s["h"]), len(s["pos"]))
        pos = pos.reshape(-1, 3)
        lists = (*oracle.nlist_rebo(pos, s["h"], s["mol_id"]),
                 *oracle.nlist_lj(pos, s["h"], s["mol_id"]))
        E0, F0, W0 = oracle.airebo_forces(pos, s["h"], lists, oracle.REBO_COMPAT)
        E, F, W = b.forces(r)
        assert rel_max(F, F0) < TOL and rel_max(W, W0) < TOL
        for x, y in zip(E, E0):
            assert abs(x - y) / abs(y) < TOL
    b.close()
