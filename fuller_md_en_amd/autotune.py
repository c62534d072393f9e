"""Autotuner for the engines' launch and list parameters, with a JSON cache.

For a given system the tuner times every candidate parameter set on the GPU
(warm-up, then a timed window of steps from the same initial state), keeps
only the candidates whose results equal the default's over the window
(final positions and per-step energies within `rtol` -- list and ownership
choices change the summation order, never the pair set), and stores the
fastest under a key of device, engine kind and system shape.

Tunables:
  rigid    skin (list radius beyond the force cutoff), nlup (rebuild
           interval), half_pairs (1 pair-once per owned pair, 2 pair-once
           per molecule, 0 both sides), graph (step replay)
  atomic   al_skin (the per-atom LJ lists' radius beyond the cutoff: a
           thinner skin walks fewer candidates per step but rebuilds the
           lists within steps more often), graph (step replay); then, with
           those fixed, the REBO kernel's LDS budget (bond entries per atom:
           fewer fit more workgroups per CU, more keep hot molecules off the
           global-memory path) and workgroup size (three or four
           wavefronts); the half-list radius and rebuild interval are part
           of the reference's truncated-half-list semantics and stay

A space is one dict of axes (every combination is tried) or a list of such
stages: each stage tries its axes' combinations with the earlier stages'
choices fixed.  A candidate that counts a list fault is rejected.
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np

CACHE_ENV = "FMD_TUNE_CACHE"


def cache_path() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(
        os.path.expanduser("~"), ".cache", "fuller_md_en_amd", "tune.json")


def load_cache(path: str | None = None) -> dict:
    path = path or cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store(key: str, entry: dict, path: str | None = None) -> None:
    path = path or cache_path()
    data = load_cache(path)
    data[key] = entry
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


RIGID_SPACE = {"skin": [3.0, 2.0, 1.5], "nlup": [25, 10, 50], "half_pairs": [1, 2, 0],
               "graph": [0, 1]}
ATOMIC_SPACE = [{"al_skin": [1.0, 0.75, 0.6, 0.5], "graph": [0, 1]},
                {"rebo_slots": [4, 6], "rebo_threads": [192, 256]}]


def stages(space) -> list:
    return list(space) if isinstance(space, (list, tuple)) else [space]


def defaults(space) -> dict:
    """The first value of every axis of every stage."""
    out = {}
    for st in stages(space):
        out.update({k: v[0] for k, v in st.items()})
    return out


def candidates(space: dict):
    """All combinations, the default (first value of every axis) first."""
    keys = list(space)
    for vals in itertools.product(*(space[k] for k in keys)):
        yield dict(zip(keys, vals))


def same_results(ref: dict, res: dict, rtol: float) -> bool:
    """Equal final positions and per-step energies within rtol (relative to
    the largest magnitude)."""
    for k in ("pos", "ep"):
        a, b = np.asarray(ref[k]), np.asarray(res[k])
        if a.shape != b.shape or not np.all(np.isfinite(b)):
            return False
        if np.abs(a - b).max() > rtol * max(np.abs(a).max(), 1e-30):
            return False
    return True


def tune(key: str, run, space, rtol: float = 1e-7, path: str | None = None,
         use_cache: bool = True) -> dict:
    """run(params) -> {"seconds": s, "pos": final positions, "ep": energies,
    "overflow": list overflows}.  Returns the cache entry
    {"params", "seconds", "default_seconds", "tried"}."""
    if use_cache:
        hit = load_cache(path).get(key)
        if hit:
            return hit
    tried = []
    ref = None
    best = None
    fixed = defaults(space)
    for st in stages(space):
        for sub in candidates(st):
            params = {**fixed, **sub}
            res = run(params)
            ok = res.get("overflow", 0) == 0
            if ref is None:
                ref = res
            else:
                ok = ok and same_results(ref, res, rtol)
            tried.append({"params": params, "seconds": res["seconds"], "valid": bool(ok)})
            if ok and (best is None or res["seconds"] < best[1]):
                best = (params, res["seconds"])
        fixed = dict(best[0])
    entry = {"params": best[0], "seconds": best[1], "default_seconds": tried[0]["seconds"],
             "tried": tried}
    store(key, entry, path)
    return entry


def device_name() -> str:
    from . import _native as N
    import ctypes as C
    buf = C.create_string_buffer(256)
    cu, hbm = C.c_int(0), C.c_double(0)
    try:
        N.check(N.lib().fmd_device_info(buf, 256, C.byref(cu), C.byref(hbm)))
        return f"{buf.value.decode(errors='replace')}/{cu.value}CU"
    except Exception:  # no device: the key still distinguishes the host
        return "unknown"


def engine_runner(build, steps: int = 200, warmup: int = 20):
    """run(params) for tune(): build(params) -> an engine (RigidEngine or
    AtomicEngine) in the same initial state on every call."""
    def run(params):
        eng = build(params)
        eng.set_graph(bool(params.get("graph", 0)))
        eng.prepare()
        eng.run(0, warmup, thermo=False)
        eng.sync()
        t0 = time.perf_counter()
        th = eng.run(warmup, steps)
        eng.sync()
        dt = time.perf_counter() - t0
        out = {"seconds": dt, "pos": eng.get_state()["pos"].copy(),
               "ep": np.array([t.Ep for t in th]), "overflow": eng.counters()[0]}
        eng.close()
        return out
    return run

