"""Autotuner on the GPU: the tuned rigid-engine parameters are at least as
fast as the defaults and give the same trajectory (within FP rounding)."""
import time

import numpy as np
import pytest

from conftest import golden
from fuller_md_en_amd import autotune, rigid

pytestmark = pytest.mark.gpu


def _build(g):
    def build(params):
        eng = rigid.RigidEngine(g["body"], len(g["pos"]), float(g["I0"]), float(g["Mmol"]),
                                rmol=3.55, rmcut=float(g["rmcut"]), skin=params["skin"],
                                nlup=params["nlup"], half_pairs=int(params["half_pairs"]))
        eng.set_state(pos=g["pos"], vel=g["vel"], qv=g["qv"], omg=g["omg"], h=g["h"],
                      npt=rigid.make_npt(300.0, 0.0, len(g["pos"])))
        return eng
    return build


def test_tuned_rigid_parameters(tmp_path):
    g = golden("rigid_core_nc3.npz")
    run = autotune.engine_runner(_build(g), steps=120, warmup=10)
    entry = autotune.tune("test|rigid|nc3", run, autotune.RIGID_SPACE,
                          path=str(tmp_path / "tune.json"))
    print(entry["params"], entry["seconds"], entry["default_seconds"], flush=True)
    assert entry["tried"][0]["valid"] and entry["seconds"] <= entry["default_seconds"]
    # re-measure both: the tuned set is not slower and gives the same results
    default = {k: v[0] for k, v in autotune.RIGID_SPACE.items()}
    a = min((run(default) for _ in range(2)), key=lambda r: r["seconds"])
    b = min((run(entry["params"]) for _ in range(2)), key=lambda r: r["seconds"])
    assert autotune.same_results(a, b, 1e-7)
    assert b["seconds"] <= 1.15 * a["seconds"]
