"""Benchmark of the MI355X fuller_md force engine on one GPU.

Default workload: flexible AIREBO (REBO bond order + intermolecular LJ with the
reference's 400-entry half-list semantics), C60 FCC 8x8x8 = 2048 molecules /
122,880 atoms at the reference lattice constant, seed-42 Maxwell velocities at
300 K, NPT at 0 GPa, dt = 0.5 fs, lists rebuilt every 20 steps, exact-gradient
REBO (the reference's own REBO forces are not a gradient and heat the crystal
without bound), the autotuner's parameters for this system (per-atom list skin,
step replay; cached per device and system, --no-tune for the defaults).  One
"step" = one full NPT step, list rebuilds included; value = atoms x steps / s
over the timed steps, inputs resident in HBM.  vs_baseline stays null: no
published number exists for this metric on this configuration.

    python bench.py [--workload airebo|rigid|core|batch|surrogate|mm|copy]
                    [--steps K] [--warmup W] [--gpus 1] [--no-cpu]

Prints one JSON line.  Extra objects:
  roofline      the dominant kernel's achieved rate (frozen algorithmic count
                / its mean launch time from HIP events on the engine's stream)
                against the MI355X peak; "kernels" lists the others.
  rebuild       list rebuilds inside the timed window, their measured mean
                cost, and the step time with exactly one rebuild per nlup.
  fp64_error    (AIREBO) forces, energies, virial and pressure of the engine
                against the reference build (its FP64 compute_forces on the
                same state and lists) and against the FP64 oracle of the
                exact-gradient mode.
  cpu_baseline  the reference build timed on this host's cores on a bounded
                sample of the same workload (kind "reference").

The workloads other than the default are the remaining configurations:
rigid (C60 FCC 10^3 rigid LJ), core (the core program, C60 FCC 3^3), batch
(64 AIREBO replicas), surrogate (learned REBO), mm (molecular mechanics) and
copy (a plain device copy of known size, the calibration of the HBM byte
counters).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import re
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atom-steps/s (AIREBO & LJ) vs host OpenMP; surrogate speedup at force RMSE"
SEED = 42
T_K = 300.0
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
LIBDIR = os.path.join(ROOT, "fuller_md_en_amd", "FullereneLib")
EV2GPA = 160.21766208

# ---- frozen algorithmic counts (per unit of work) -------------------------
# AIREBO intermolecular LJ: 42 FLOP per kept in-cutoff pair (8 distance, 13
# scalar, 3 force, 6 accumulation, 12 virial); the reference's 400-entry
# half list keeps 19.08 M in-cutoff pairs of the C60 FCC 8^3 crystal, 155.3
# per atom (6.5% of the true pairs are truncated).
FLOP_PER_ATOM_LJ = 42.0 * 155.3
# REBO, exact-gradient mode, per bond of a C60 crystal (three bonds per atom,
# all inside the switch's inner radius, so two side bonds at each end): bond
# vector, minimum image, length and reciprocal 42; pair terms 25 (+4 exp);
# unit vector 3; four angles: cosine 5 + G, dG and the G sum 12 each (68);
# bond orders at both ends 8 (+2 rsqrt); energy and pair-force scalar 10;
# pair force on both atoms and the virial tensor 27; four angular force
# triples 28 each (112); the exact gradient's partner term and triple virial
# 18 each (72): 370 FLOP (and 7 transcendentals) per bond, 1.5 bonds per atom.
FLOP_PER_ATOM_REBO = 370.0 * 1.5
# Rigid LJ: 69 FLOP per unique in-cutoff site pair (11 displacement + r^2, 13
# LJ scalar incl. energy, 3 force, 6 force accumulation, 24 torque, 12 virial)
# x 163.2 unique pairs per site (C60 FCC, a = 14.17 A); 162.6 for the core
# program's analytic C60 at a = 14.17 A (2,114,534 directed pairs / 6480 / 2).
FLOP_PER_SITE_RIGID = 69.0 * 163.2
FLOP_PER_SITE_CORE = 69.0 * (2114534 / 6480 / 2)
# Molecular mechanics intermolecular LJ: 42 FLOP x 169.6 unique in-cutoff
# pairs per atom (C60 FCC at the default a0, exact list).
FLOP_PER_ATOM_MM_LJ = 42.0 * 169.6

# HBM bytes per launch of the dominant kernels, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what the same
# counters read for `--workload copy` (a 1 GiB device copy): FETCH x 2.0,
# WRITE x 1.0 (profiles/pmc_traffic_*.md).
RIGID_TRAFFIC = {1: (9710.9 * 2.0 + 23524.8 * 1.0) * 1024.0,
                 2: (11032.8 * 2.0 + 5618.9 * 1.0) * 1024.0}
AIREBO_LJ_TRAFFIC = (62639.7 * 2.0 + 3040.3 * 1.0) * 1024.0  # tuned 0.5 A list skin
AIREBO_REBO_TRAFFIC = (4766.0 * 2.0 + 3040.0 * 1.0) * 1024.0
RIGID_KERNEL = {1: "k_rigid_pairs", 2: "k_rigid_forces_half", 0: "k_rigid_forces"}
TIMING_STRIDE = 10


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=0,
                    help="timed steps (default: 2000 airebo, 10000 rigid, 1000 core, 400 batch)")
    ap.add_argument("--warmup", type=int, default=-1,
                    help="untimed steps first (default: 100, batch 20)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline leg (and the FP64 error check)")
    ap.add_argument("--no-tune", action="store_true",
                    help="rigid workload: default parameters instead of the autotuner's")
    ap.add_argument("--workload", default="airebo",
                    choices=["airebo", "rigid", "core", "batch", "surrogate", "mm", "copy"])
    ap.add_argument("--nc", type=int, default=0, help="crystal repeats (default per workload)")
    ap.add_argument("--mm", choices=["compat", "physical"], default="physical",
                    help="mm workload: the reference's forces (which heat the crystal without "
                         "bound) or the exact gradient")
    ap.add_argument("--al-skin", type=float, default=0.0,
                    help="airebo workload: per-atom LJ list skin (default: the autotuner's)")
    ap.add_argument("--cell-min", type=int, default=None,
                    help="airebo / rigid workloads: molecule lists from a cell grid for systems "
                         "of at least this many molecules (default: the engine's, 2048 AIREBO, "
                         "3000 rigid; a larger value keeps the scan of every molecule)")
    ap.add_argument("--rebo", choices=["compat", "physical"], default="physical",
                    help="REBO force mode of the timed run")
    ap.add_argument("--rebo-ext", action="store_true",
                    help="airebo workload: add the REBO-II P_ij spline and the AIREBO torsion "
                         "term (AIREBO carbon values; no reference counterpart, no CPU leg)")
    ap.add_argument("--sg-iters", type=int, default=30000,
                    help="surrogate workload: training iterations")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ helpers
def host_info():
    """CPU model, logical CPUs usable here, physical cores, and the thread
    share this process is given (OMP_NUM_THREADS when the host sets it)."""
    model, phys = "?", set()
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name":
                model = v
            elif k in ("physical id", "core id"):
                cur[k] = v
            if not line.strip() and cur:
                phys.add((cur.get("physical id"), cur.get("core id")))
                cur = {}
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or aff
    return {"cpu_model": model, "affinity_cpus": aff, "physical_cores": len(phys) or None,
            "thread_share": min(share, aff)}


def thread_counts(share):
    ts, t = [], 1
    while t < share:
        ts.append(t)
        t *= 2
    return ts + [share]


def timed_run(eng, warmup, steps, nlup):
    """Warm-up, then exactly `steps` timed steps bracketed by device syncs;
    per-phase event timing over the timed steps."""
    eng.prepare()
    if warmup:
        eng.run(0, warmup, thermo=False)
    eng.sync()
    # HIP events on every 10th step and on every list rebuild: each event
    # writes back the caches, so timing every step would slow the run itself
    eng.set_timing(True, stride=TIMING_STRIDE)
    t0 = time.perf_counter()
    th = eng.run(warmup, steps, thermo=True)
    eng.sync()
    dt = time.perf_counter() - t0
    ph = eng.phase_times()
    n_rb = sum(1 for g in range(warmup + 1, warmup + steps + 1) if g % nlup == 0)
    list_ms = ph["list"][0]
    amort = (dt * 1e3 - n_rb * list_ms + steps * list_ms / nlup) / steps
    rebuild = {"nlup": nlup, "in_window": n_rb, "ms_per_rebuild": list_ms,
               "ms_per_step_amortized": amort}
    return dt, th, ph, rebuild


def kernel_rate(name, flop_per_launch, ms, peak, bound="valu", traffic=None):
    ach = flop_per_launch / (ms * 1e-3) / 1e12 if ms > 0 else None
    out = {"kernel": name, "bound": bound, "achieved": ach, "peak": peak, "unit": "TFLOP/s",
           "frac": ach / peak if ach else None, "traffic": traffic, "kernel_ms": ms}
    if traffic and ms > 0:  # the same launch against the HBM roof
        out["traffic_GB_s"] = traffic / (ms * 1e-3) / 1e9
        out["hbm_frac"] = out["traffic_GB_s"] / PEAK_HBM_GBS
    return out


def line(workload_cfg, value, steps, warmup, dt, dtype, roofline, **extra):
    res = {"metric": METRIC, "value": value, "unit": "atom-steps/s", "n_gpus": 1,
           "steps": steps, "warmup": warmup, "ms_per_step": dt * 1e3 / steps,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": workload_cfg, "roofline": roofline}
    res.update(extra)
    return res


def done_secs(out):
    m = re.search(r"Done \(?([\d.]+) ?s", out)
    return float(m.group(1)) if m else None


# ----------------------------------------------------------- AIREBO (default)
def rel_err(a, b):
    import numpy as np
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def airebo_cpu_leg(s, nc):
    """The reference AIREBO build on the 8^3 state, and the FP64 checks.

    Timing (kind "reference", the reference's own compiled functions through
    oracle/_ref's harness): its step_npt (forces + integrator, AB:518-590) on
    the 8^3 state and lists, serial and OpenMP over a thread sweep, plus its
    O(N^2) list builders (AB:251-272), which run every 20 steps serially in
    both builds; timed at 3^3 and 4^3 and scaled by the pair count N(N-1)/2
    to 8^3 (a full 8^3 build takes about 100 s).  The lists handed to the
    reference are the GPU drop-in's, which equal the reference builder's.

    Errors: the engine's forces at the same state against the reference's
    FP64 compute_forces (compat mode, the reference's own arithmetic) and
    against the FP64 oracle (exact-gradient mode); the pressure error uses
    the virial trace, P = (2 KE + tr W) / 3V."""
    import numpy as np

    from fuller_md_en_amd import airebo, system
    from oracle import ref
    import oracle
    if not ref.available("airebo"):
        return None, None
    L = ref.airebo()
    na = len(s["pos"])
    mol_id, mass = s["mol_id"], s["mass"]
    h = s["h"].copy()
    hi = np.zeros(9)
    L.refab_mat_inv9(h, hi)
    pos = s["pos"].reshape(-1).copy()
    L.refab_apply_pbc(pos, h, hi, na)
    lr = airebo.build_nlist_rebo(pos, h, hi, mol_id)
    ll = airebo.build_nlist_lj(pos, h, hi, mol_id)
    lists = tuple(np.ascontiguousarray(x.reshape(-1)) for x in (lr[0], lr[1], ll[0], ll[1]))

    # ---- FP64 errors of the engine
    F0, W0, E0 = np.zeros(3 * na), np.zeros(9), np.zeros(3)
    t0 = time.perf_counter()
    L.refab_compute_forces(F0, W0, pos, h, hi, *lists, na, E0)
    t_force = time.perf_counter() - t0
    vel = np.ascontiguousarray(s["vel"].reshape(-1))
    ke = 0.5 * float((mass * (vel.reshape(-1, 3) ** 2).sum(axis=1)).sum()) / 9.64853321e-3
    V = abs(float(np.linalg.det(h.reshape(3, 3))))

    def errs(E, F, W, Er, Fr, Wr):
        P = (2 * ke + W[0] + W[4] + W[8]) / (3 * V) * EV2GPA
        Pr = (2 * ke + Wr[0] + Wr[4] + Wr[8]) / (3 * V) * EV2GPA
        return {"force_max_rel": rel_err(F, Fr),
                "force_l2_rel": float(np.linalg.norm(F - Fr) / np.linalg.norm(Fr)),
                "E_rel": abs(E[0] - Er[0]) / abs(Er[0]),
                "E_rebo_rel": abs(E[1] - Er[1]) / abs(Er[1]),
                "E_lj_rel": abs(E[2] - Er[2]) / abs(Er[2]),
                "virial_max_rel": rel_err(W, Wr),
                "pressure_rel": abs(P - Pr) / max(abs(Pr), 0.01), "P_GPa": Pr}

    out = {}
    for mode, name in ((airebo.REBO_COMPAT, "compat_vs_reference_build"),
                       (airebo.REBO_PHYSICAL, "physical_vs_fp64_oracle")):
        eng = airebo.AtomicEngine(mol_id, mass, dt=0.5, mode=mode)
        # the engine wraps the positions into the cell itself, as the
        # reference does before its first forces
        eng.set_state(pos=s["pos"], vel=vel, h=s["h"], npt=airebo.make_npt(T_K, 0.0, na))
        eng.prepare()
        E, F, W = eng.forces()
        eng.close()
        if mode == airebo.REBO_COMPAT:
            out[name] = errs(E, F.reshape(-1), W, E0, F0, W0)
        else:
            Eo, Fo, Wo = oracle.airebo_forces(pos, h, lists, mode=oracle.REBO_PHYSICAL)
            out[name] = errs(E, F.reshape(-1), W, Eo, Fo.reshape(-1), Wo)
    out["state"] = f"C60 FCC {nc}^3, seed-42 initial state wrapped into the cell"

    # ---- reference step time, serial
    def step_time(lib, nrep):
        best = None
        for _ in range(nrep):
            P, Vv, Fs, Ws = pos.copy(), vel.copy(), F0.copy(), W0.copy()
            hh, hhi = h.copy(), hi.copy()
            st = np.zeros(15)
            lib.refab_make_npt(T_K, 0.0, na, st)
            o = np.zeros(4)
            t1 = time.perf_counter()
            lib.refab_step_npt(P, Vv, Fs, Ws, hh, hhi, mass, na, 0.5, st, *lists, o)
            t = time.perf_counter() - t1
            best = t if best is None else min(best, t)
        return best

    t_serial = step_time(L, 2)
    # ---- list build (serial in both builds), scaled from 3^3 and 4^3
    lb = {}
    for k in (3, 4):
        sk = system.airebo_system(k, "C60", T=T_K, seed=SEED)
        n = len(sk["pos"])
        pk = sk["pos"].reshape(-1).copy()
        hk, hik = sk["h"].copy(), np.zeros(9)
        L.refab_mat_inv9(hk, hik)
        c1, l1 = np.zeros(n, np.int32), np.zeros(n * 12, np.int32)
        c2, l2 = np.zeros(n, np.int32), np.zeros(n * 400, np.int32)
        t1 = time.perf_counter()
        L.refab_nlist_rebo(pk, hk, hik, n, sk["mol_id"], c1, l1)
        L.refab_nlist_lj(pk, hk, hik, n, sk["mol_id"], c2, l2)
        lb[k] = (time.perf_counter() - t1) / (n * (n - 1) / 2)
    t_list = lb[4] * na * (na - 1) / 2
    per_serial = t_serial + t_list / 20
    # ---- OpenMP build over a thread sweep (stops once it is clearly slower)
    info = host_info()
    omp = {}
    if ref.available("airebo") and os.path.exists(ref.path("libref_airebo_omp.so")):
        Lo = ref.airebo(omp=True)
        gomp = C.CDLL("libgomp.so.1")
        best = None
        for t in thread_counts(info["thread_share"]):
            gomp.omp_set_num_threads(t)
            ts = step_time(Lo, 1)
            omp[t] = ts
            best = ts if best is None else min(best, ts)
            if ts > 1.5 * best and t >= 4:
                break
    t_omp = min(omp.values()) if omp else None
    per_omp = t_omp + t_list / 20 if omp else None
    use_omp = per_omp is not None and per_omp < per_serial
    per = per_omp if use_omp else per_serial
    t_best = min(omp, key=omp.get) if use_omp else 1
    cpu = {"value": na / per, "unit": "atom-steps/s", "cores": t_best, "kind": "reference",
           "sample": (f"reference AIREBO functions on the C60 FCC {nc}^3 state (oracle/_ref "
                      f"harness): step_npt {t_serial:.2f} s serial"
                      + (f", OpenMP best {t_omp:.2f} s at {min(omp, key=omp.get)} threads"
                         if omp else "")
                      + f"; serial list builds scaled by N^2 from 4^3 to {na} atoms: "
                      f"{t_list:.1f} s every 20 steps; per step "
                      f"{per:.2f} s ({'OpenMP' if use_omp else 'serial'} build)"),
           "serial_atom_steps_s": na / per_serial,
           "openmp_atom_steps_s": {str(t): na / (v + t_list / 20) for t, v in omp.items()},
           "compute_forces_s": t_force, "list_build_s": t_list,
           "list_s_per_pair_3c_4c": [lb[3], lb[4]], **info}
    return cpu, out


def airebo_engine(s, mode, params, cell_min=None, ext=False):
    from fuller_md_en_amd import airebo
    eng = airebo.AtomicEngine(s["mol_id"], s["mass"], dt=0.5, mode=mode,
                              al_skin=params["al_skin"], cell_min_molecules=cell_min,
                              rebo_slots=params.get("rebo_slots"),
                              rebo_threads=params.get("rebo_threads"),
                              pcc=airebo.AIREBO_PCC if ext else None,
                              torsion_eps=airebo.AIREBO_TORSION_EPS if ext else 0.0)
    eng.set_graph(bool(params["graph"]))
    eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"],
                  npt=airebo.make_npt(T_K, 0.0, len(s["pos"])))
    return eng


def airebo_params(args, s, nc, mode):
    """The per-atom list skin and step replay: --al-skin, the defaults
    (--no-tune), or the autotuner's choice for this system (cached per device,
    system, REBO mode, temperature and pressure)."""
    from fuller_md_en_amd import autotune
    default = autotune.defaults(autotune.ATOMIC_SPACE)
    if args.al_skin:
        return {**default, "al_skin": args.al_skin}, None
    if args.no_tune:
        return default, None
    key = (f"{autotune.device_name()}|airebo|C60-FCC{nc}|{args.rebo}|T={T_K:g}|P=0|"
           f"seed={SEED}")
    entry = autotune.tune(key, autotune.engine_runner(lambda p: airebo_engine(s, mode, p),
                                                      steps=200, warmup=20),
                          autotune.ATOMIC_SPACE)
    return entry["params"], entry


def main_airebo(args):
    import numpy as np
    from fuller_md_en_amd import airebo, system
    nc = args.nc or 8
    s = system.airebo_system(nc, "C60", T=T_K, seed=SEED)
    na, nm = len(s["pos"]), s["n_mol"]
    mode = airebo.REBO_PHYSICAL if args.rebo == "physical" else airebo.REBO_COMPAT
    if args.rebo_ext and mode != airebo.REBO_PHYSICAL:
        raise SystemExit("--rebo-ext needs --rebo physical")
    params, tuned = airebo_params(args, s, nc, mode)
    eng = airebo_engine(s, mode, params, args.cell_min, ext=args.rebo_ext)
    dt, th, ph, rebuild = timed_run(eng, args.warmup, args.steps, 20)
    faults, rebuilds = eng.counters()[:2]
    eng.close()
    value = na * args.steps / dt
    lj_ms, rebo_ms = ph["lj"][0], ph["intra"][0]
    lj = kernel_rate("k_at_lj_al", FLOP_PER_ATOM_LJ * na, lj_ms, PEAK_FP32_TFLOPS,
                     traffic=AIREBO_LJ_TRAFFIC)
    lj["launches"] = ph["lj"][1]
    rebo = kernel_rate("k_rebo_mol + k_rebo_tors" if args.rebo_ext else "k_rebo_mol",
                       FLOP_PER_ATOM_REBO * na, rebo_ms, PEAK_FP32_TFLOPS,
                       traffic=None if args.rebo_ext else AIREBO_REBO_TRAFFIC)
    force = kernel_rate("force phase (k_rebo_mol + k_at_lj_al)",
                        (FLOP_PER_ATOM_LJ + FLOP_PER_ATOM_REBO) * na, ph["force"][0],
                        PEAK_FP32_TFLOPS)
    last = th[args.steps - 1]
    ext = (" with the REBO-II P_ij spline and the AIREBO torsion term (no reference "
           "counterpart)") if args.rebo_ext else ""
    res = line({"workload": f"AIREBO C60 FCC {nc}x{nc}x{nc} ({nm} mol, {na} atoms), REBO "
                            f"{args.rebo}{ext} + LJ (400-entry half-list semantics), 300 K, "
                            "0 GPa, dt=0.5 fs, NPT, lists every 20 steps", "atoms": na},
               value, args.steps, args.warmup, dt, "fp32 pair/bond-order, fp64 integrator",
               lj, kernels=[rebo, force], rebuild=rebuild,
               value_amortized=na / (rebuild["ms_per_step_amortized"] * 1e-3),
               list_faults=faults, list_rebuilds=rebuilds, params=params,
               tuned=tuned and {k: tuned[k] for k in ("seconds", "default_seconds", "tried")},
               final={"T": last.T, "P": last.P, "E_rebo_per_mol": last.Ep_rebo / nm,
                      "E_lj_per_mol": last.Ep_lj / nm})
    if not args.no_cpu and not args.rebo_ext:  # the reference has neither term
        cpu, err = airebo_cpu_leg(s, nc)
        res["cpu_baseline"] = cpu
        res["fp64_error"] = err
        if cpu:
            res["speedup_vs_cpu_baseline"] = value / cpu["value"]
    print(json.dumps(res))


# ------------------------------------------------------------------ rigid
def rigid_engine(params, nc=10, a=14.17, cell_min=None):
    from fuller_md_en_amd import rigid, system
    mol = system.molecule("C60")
    scale = a / system.default_a0(mol.dmol, "FCC", 1.0)
    s = system.rigid_system(nc=nc, spec="C60", T=T_K, seed=SEED, init_scale=scale)
    eng = rigid.RigidEngine(mol.coords, len(s["pos"]), mol.I0, mol.mass, rmol=mol.rmol,
                            dt=1.0, skin=params["skin"], nlup=params["nlup"],
                            half_pairs=int(params["half_pairs"]), cell_min_molecules=cell_min)
    eng.set_graph(bool(params["graph"]))
    eng.set_state(pos=s["pos"], vel=s["vel"], qv=s["qv"], omg=s["omg"], h=s["h"],
                  npt=rigid.make_npt(T_K, 0.0, len(s["pos"])))
    return eng, mol, s, scale


def tuned_params(args, nc):
    from fuller_md_en_amd import autotune
    if args.no_tune:
        return {k: v[0] for k, v in autotune.RIGID_SPACE.items()}, None
    key = f"{autotune.device_name()}|rigid|C60-FCC{nc}|a=14.17|T={T_K:g}|P=0"
    entry = autotune.tune(key, autotune.engine_runner(lambda p: rigid_engine(p, nc)[0],
                                                     steps=200, warmup=20),
                          autotune.RIGID_SPACE)
    return entry["params"], entry


def omp_sweep(exe_args, sites, steps_at_1, budget_s=25.0):
    """The reference OpenMP binary over OMP_NUM_THREADS = 1, 2, 4, ... up to
    the host's thread share; steps grow with the thread count so that each
    run takes about the same time."""
    info = host_info()
    res, used = {}, 0.0
    for t in thread_counts(info["thread_share"]):
        # the reference programs need 4 steps at least (their averaging
        # window starts at nsteps - nsteps/4 and must end after it starts)
        steps = max(4, steps_at_1 * t)
        env = dict(os.environ, OMP_NUM_THREADS=str(t), OMP_PROC_BIND="close", OMP_PLACES="cores")
        t0 = time.perf_counter()
        out = subprocess.run([*exe_args(steps)], env=env, capture_output=True, text=True,
                             timeout=600, cwd="/tmp").stdout
        used += time.perf_counter() - t0
        secs = done_secs(out)
        if secs:
            res[t] = sites * steps / secs
        if used > budget_s:
            break
    return res, info


def rigid_cpu_baseline(nc, scale):
    exe = os.path.join(ROOT, "oracle", "_ref", "fuller_LJ_npt_md_omp")
    if not os.path.exists(exe):
        return None
    sites = 4 * nc ** 3 * 60

    def args(steps):
        return [exe, f"--cell={nc}", f"--temp={T_K:g}", "--pres=0", f"--step={steps}",
                f"--seed={SEED}", f"--init_scale={scale:.9f}", f"--libdir={LIBDIR}",
                f"--mon={steps}"]
    sweep, info = omp_sweep(args, sites, 1)
    if not sweep:
        return None
    t_best = max(sweep, key=sweep.get)
    return {"value": sweep[t_best], "unit": "atom-steps/s", "cores": t_best, "kind": "reference",
            "sample": f"fuller_LJ_npt_md_omp --cell={nc} (a=14.17 A, seed 42), OMP_NUM_THREADS "
                      f"swept over {sorted(sweep)} with t steps at t threads (no list rebuild "
                      "in so few steps: its O(N^2) build every 25 steps adds about 2%), its own "
                      "'Done' timer; value = the best thread count",
            "openmp_atom_steps_s": {str(t): v for t, v in sweep.items()}, **info}


def main_rigid(args):
    nc = args.nc or 10
    params, entry = tuned_params(args, nc)
    eng, mol, s, scale = rigid_engine(params, nc, cell_min=args.cell_min)
    n_sites = len(s["pos"]) * mol.natom
    dt, th, ph, rebuild = timed_run(eng, args.warmup, args.steps, int(params["nlup"]))
    faults = eng.counters()[0]
    eng.close()
    value = n_sites * args.steps / dt
    hp = int(params["half_pairs"])
    roof = kernel_rate(RIGID_KERNEL[hp], FLOP_PER_SITE_RIGID * n_sites, ph["force"][0],
                       PEAK_FP32_TFLOPS, traffic=RIGID_TRAFFIC.get(hp))
    roof["launches"] = ph["force"][1]
    last = th[args.steps - 1]
    res = line({"workload": f"rigid-LJ C60 FCC {nc}x{nc}x{nc} ({len(s['pos'])} mol, {n_sites} "
                            "sites), a=14.17 A, 300 K, 0 GPa, dt=1 fs, NPT",
                "sites": n_sites, "tuned": params,
                "untuned_ms_per_step": (entry["default_seconds"] * 1e3 / 200 if entry else None)},
               value, args.steps, args.warmup, dt, "fp32 pair forces, fp64 integrator/virial",
               roof, rebuild=rebuild,
               value_amortized=n_sites / (rebuild["ms_per_step_amortized"] * 1e-3),
               list_faults=faults,
               final={"T": last.T, "P": last.P, "Ep_per_mol": last.Ep / len(s["pos"])})
    if not args.no_cpu:
        res["cpu_baseline"] = rigid_cpu_baseline(nc, scale)
        if res["cpu_baseline"]:
            res["speedup_vs_cpu_baseline"] = value / res["cpu_baseline"]["value"]
    print(json.dumps(res))


# ------------------------------------------------------------------- core
def main_core(args):
    """The core program (C1): analytic C60, FCC 3^3 at a0 = 14.17 A, 300 K,
    0 GPa, dt = 1 fs, list every 25 steps (config 1)."""
    import numpy as np
    from fuller_md_en_amd import rigid, system
    nc = args.nc or 3
    s = system.core_system(nc, SEED)
    mol = s["mol"]
    n = len(s["pos"])
    eng = rigid.RigidEngine(mol.coords, n, mol.I0, mol.mass, rmol=3.55, nlup=25)
    eng.set_graph(True)
    eng.set_state(pos=s["pos"], vel=s["vel"], qv=s["qv"], omg=s["omg"], h=s["h"],
                  npt=rigid.make_npt(T_K, 0.0, n))
    dt, th, ph, rebuild = timed_run(eng, args.warmup, args.steps, 25)
    eng.close()
    sites = n * 60
    value = sites * args.steps / dt
    roof = kernel_rate("k_rigid_pairs", FLOP_PER_SITE_CORE * sites, ph["force"][0],
                       PEAK_FP32_TFLOPS)
    roof["launches"] = ph["force"][1]
    ser = np.array([[t.T, t.P, t.h[0] / nc, t.Ep / n] for t in th])
    res = line({"workload": f"core program: analytic C60 FCC {nc}x{nc}x{nc} ({n} mol, {sites} "
                            "sites), a0=14.17 A, 300 K, 0 GPa, dt=1 fs, NPT, graph-replayed "
                            "steps", "sites": sites},
               value, args.steps, args.warmup, dt, "fp32 pair forces, fp64 integrator/virial",
               roof, rebuild=rebuild,
               averages_last_quarter=dict(zip(["T", "P", "a", "Ecoh"],
                                              ser[-(args.steps // 4):].mean(axis=0).tolist())))
    if not args.no_cpu:
        exe = os.path.join(ROOT, "oracle", "_ref", "fuller_LJ_core_omp")
        if os.path.exists(exe):
            info = host_info()
            t = info["thread_share"]
            env = dict(os.environ, OMP_NUM_THREADS=str(t), OMP_PROC_BIND="close",
                       OMP_PLACES="cores")
            out = subprocess.run([exe, str(nc)], env=env, capture_output=True, text=True,
                                 timeout=600, cwd="/tmp").stdout
            secs = done_secs(out)
            if secs:
                res["cpu_baseline"] = {
                    "value": sites * 1000 / secs, "unit": "atom-steps/s", "cores": t,
                    "kind": "reference",
                    "sample": f"fuller_LJ_core_omp {nc} (its fixed 1000 steps), "
                              f"OMP_NUM_THREADS={t}, its own 'Done' timer ({secs:.1f} s); the "
                              "rigid OpenMP build scales near linearly, so all of the host "
                              "share is its best", **info}
                res["speedup_vs_cpu_baseline"] = value / res["cpu_baseline"]["value"]
    print(json.dumps(res))


# --------------------------------------------------------------------- mm
def mm_cpu_baseline(steps):
    """Reference serial MM build (its OpenMP build anti-scales), bounded
    sample at 5x5x5 (30,000 atoms)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "fuller_LJ_npt_mmmd_serial")
    if not os.path.exists(exe):
        return None
    cmd = [exe, "--cell=5", "--temp=300", "--pres=0", f"--step={steps}", "--seed=42",
           f"--libdir={LIBDIR}", f"--mon={steps}"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=900, cwd="/tmp").stdout
    secs = done_secs(out)
    if not secs:
        return None
    return {"value": 30000 * steps / secs, "unit": "atom-steps/s", "cores": 1,
            "kind": "reference",
            "sample": f"fuller_LJ_npt_mmmd_serial --cell=5 --step={steps} (30,000 atoms, "
                      f"1 core; its O(N^2) list at every 20th step included), "
                      f"'Done' timer {secs:.1f} s"}


def main_mm(args):
    from fuller_md_en_amd import airebo, mm
    nc = args.nc or 8
    s = mm.mm_system(nc, "C60", "FCC", T=T_K, seed=SEED)
    na, nm = len(s["pos"]), s["n_mol"]
    mode = mm.MM_COMPAT if args.mm == "compat" else mm.MM_PHYSICAL
    eng = mm.MMEngine(s["ft"], dt=0.1, mode=mode)
    eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"], npt=airebo.make_npt(T_K, 0.0, na))
    dt, th, ph, rebuild = timed_run(eng, args.warmup, args.steps, 20)
    eng.close()
    value = na * args.steps / dt
    lj = kernel_rate("k_at_lj_al (MM parameters)", FLOP_PER_ATOM_MM_LJ * na, ph["lj"][0],
                     PEAK_FP32_TFLOPS)
    bonded = {"kernel": "k_mm_bonded", "kernel_ms": ph["intra"][0], "bound": "fp64 valu",
              "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s"}
    last = th[args.steps - 1]
    res = line({"workload": f"MM C60 FCC {nc}x{nc}x{nc} ({nm} mol, {na} atoms), bonded "
                            f"{args.mm} (FP64) + LJ (400-entry half-list semantics), 300 K, "
                            "0 GPa, dt=0.1 fs, NPT", "atoms": na},
               value, args.steps, args.warmup, dt, "fp64 bonded, fp32 LJ", lj,
               kernels=[bonded], rebuild=rebuild,
               final={"T": last.T, "P": last.P, "E_bond_per_mol": last.Eb / nm,
                      "E_angle_per_mol": last.Ea / nm,
                      "E_dih_per_mol": (last.Ed + last.Ei) / nm,
                      "E_lj_per_mol": last.Ep_lj / nm})
    if not args.no_cpu:
        res["cpu_baseline"] = mm_cpu_baseline(40)
    print(json.dumps(res))


# ------------------------------------------------------------------ batch
def batch_replicas(seed=2024, n=64):
    """Seeded replica table: nc in {3,4,5,6}, C60-Ih or C70-D5h, T ~ U(100, 1500) K,
    P ~ U(0, 5) GPa."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [dict(nc=int(rng.integers(3, 7)), spec=("C60", "C70")[int(rng.integers(0, 2))],
                 T=float(rng.uniform(100.0, 1500.0)), P=float(rng.uniform(0.0, 5.0)))
            for _ in range(n)]


def batch_cpu_baseline(table, budget_s=20.0):
    """Reference serial AIREBO, one process per replica run one after another
    (the reference has no batching): replicas of the table in order, 10 steps
    each, until about budget_s of CPU time is spent."""
    exe = os.path.join(ROOT, "oracle", "_ref", "fuller_airebo_npt_md_serial")
    if not os.path.exists(exe):
        return None
    work, secs, used = 0.0, 0.0, []
    for r, rep in enumerate(table):
        steps = 10
        cmd = [exe, f"--cell={rep['nc']}", f"--temp={rep['T']:.3f}", f"--pres={rep['P']:.3f}",
               f"--fullerene={rep['spec']}", f"--step={steps}", "--seed=42",
               f"--libdir={LIBDIR}", f"--mon={steps}"]
        t0 = time.perf_counter()
        out = subprocess.run(cmd, capture_output=True, text=True, timeout=600, cwd="/tmp").stdout
        wall = time.perf_counter() - t0
        na = re.search(r"Natom=(\d+)", out)
        if not done_secs(out) or not na:
            return None
        work += int(na.group(1)) * steps
        secs += wall
        used.append(r)
        if secs > budget_s:
            break
    return {"value": work / secs, "unit": "atom-steps/s", "cores": 1, "kind": "reference",
            "sample": f"fuller_airebo_npt_md_serial, replicas {used[0]}..{used[-1]} of the same "
                      f"table run one after another, 10 steps each, process wall time "
                      f"{secs:.1f} s"}


def main_batch(args):
    import numpy as np
    from fuller_md_en_amd import airebo, system
    table = batch_replicas()
    systems = [system.airebo_system(rep["nc"], rep["spec"], T=rep["T"], seed=1000 + r)
               for r, rep in enumerate(table)]
    na = sum(len(s["pos"]) for s in systems)
    mode = airebo.REBO_PHYSICAL if args.rebo == "physical" else airebo.REBO_COMPAT
    b = airebo.ReplicaBatch([(s["mol_id"], s["mass"]) for s in systems], mode=mode)
    for r, (s, rep) in enumerate(zip(systems, table)):
        b.set_state(r, pos=s["pos"], vel=s["vel"], h=s["h"],
                    npt=airebo.make_npt(rep["T"], rep["P"], len(s["pos"])))
    dt, th, ph, rebuild = timed_run(b, args.warmup, args.steps, 20)
    faults = b.counters()[0]
    value = na * args.steps / dt
    lj = kernel_rate("k_at_lj_al (all replicas)", FLOP_PER_ATOM_LJ * na, ph["lj"][0],
                     PEAK_FP32_TFLOPS)
    rebo = kernel_rate("k_rebo_mol (all replicas)", FLOP_PER_ATOM_REBO * na, ph["intra"][0],
                       PEAK_FP32_TFLOPS)
    Ts = np.array([t.T for t in th[-1]])
    res = line({"workload": f"{len(table)} AIREBO NPT replicas (C60-Ih / C70-D5h FCC, nc 3..6, "
                            "T 100-1500 K, P 0-5 GPa, seeded table), REBO "
                            f"{args.rebo} + LJ, dt=0.5 fs, graph-replayed steps",
                "atoms": na, "replicas": len(table),
                "sizes": sorted({len(s["pos"]) for s in systems})},
               value, args.steps, args.warmup, dt, "fp32 pair/bond-order, fp64 integrator", lj,
               kernels=[rebo], rebuild=rebuild, list_faults=faults,
               final={"T_min": float(Ts.min()), "T_max": float(Ts.max())})
    if not args.no_cpu:
        res["cpu_baseline"] = batch_cpu_baseline(table)
    print(json.dumps(res))


# -------------------------------------------------------------- surrogate
def main_surrogate(args):
    """C60 FCC 8x8x8 (122,880 atoms), exact-gradient REBO + LJ: the learned
    surrogate in place of REBO (judged against full REBO every 100 steps,
    fallback above 0.05 eV/A) against the same engine with REBO."""
    import numpy as np
    from fuller_md_en_amd import airebo, system
    from fuller_md_en_amd import surrogate as SG
    t0 = time.perf_counter()
    th, mu, sd, loss, nframes = SG.train_from_engine_runs(iters=args.sg_iters, lr=3e-3, log=log)
    t_train = time.perf_counter() - t0
    nc = args.nc or 8
    s = system.airebo_system(nc, "C60", T=T_K, seed=SEED)
    na = len(s["pos"])
    res = {}
    for mode in ("rebo", "surrogate"):
        eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
        eng.set_state(pos=s["pos"], vel=s["vel"], h=s["h"], npt=airebo.make_npt(T_K, 0.0, na))
        sg = SG.Surrogate(eng)
        sg.set_params(th, mu, sd)
        if mode == "surrogate":
            sg.enable(True, judge_every=100, threshold=0.05)
        dt, _, ph, rebuild = timed_run(eng, args.warmup, args.steps, 20)
        res[mode] = dict(dt=dt, ph=ph, stats=sg.stats())
        log(f"{mode}: {dt * 1e3 / args.steps:.3f} ms/step, phases {ph}")
        eng.close()
    sgr, reb = res["surrogate"], res["rebo"]
    value = na * args.steps / sgr["dt"]
    intra_ms = sgr["ph"]["intra"][0]
    # the part a zero-cost REBO replacement could remove
    ceiling = reb["dt"] / (reb["dt"] - reb["ph"]["intra"][0] * 1e-3 * args.steps)
    roof = {"kernel": "surrogate (k_sg_descr + k_sg_mlp + k_sg_force)", "bound": "mfma+valu",
            "achieved": SG.FLOP_PER_ATOM * na / (intra_ms * 1e-3) / 1e12 if intra_ms else None,
            "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s", "traffic": None, "kernel_ms": intra_ms}
    roof["frac"] = roof["achieved"] / PEAK_BF16_TFLOPS if roof["achieved"] else None
    out = line({"workload": f"C60 FCC {nc}x{nc}x{nc} ({na} atoms), learned REBO surrogate "
                            "(40 descriptors, 3x64 tanh, BF16 MFMA) + LJ, judged against "
                            "exact-gradient REBO every 100 steps (fallback > 0.05 eV/A), "
                            "300 K, 0 GPa, dt=0.5 fs", "atoms": na},
               value, args.steps, args.warmup, sgr["dt"], "bf16 MFMA surrogate / fp32 LJ", roof,
               speedup_vs_gpu_airebo=reb["dt"] / sgr["dt"],
               amdahl_ceiling_vs_gpu_airebo=ceiling,
               intra_phase_ms={"rebo": reb["ph"]["intra"][0], "surrogate": intra_ms},
               judge=sgr["stats"],
               training={"frames": nframes, "iterations": int(len(loss)), "seconds": t_train,
                         "final_train_rmse": float(np.sqrt(loss[-50:].mean()))})
    print(json.dumps(out))


# ------------------------------------------------------------------- copy
def main_copy(args):
    """A plain device copy of 1 GiB (torch), the calibration of FETCH_SIZE /
    WRITE_SIZE: each counter's reading per dispatch against the 1 GiB read
    and written."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps)):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(1, args.steps)
    print(json.dumps({"workload": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms": dt * 1e3, "GB_s": 8 * n / dt / 1e9}))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): every GPU box "
              "here has one GPU, and the multi-GPU decomposition is design only.",
              file=sys.stderr)
        sys.exit(2)
    if not args.steps:
        args.steps = {"airebo": 2000, "rigid": 10000, "core": 1000, "batch": 400,
                      "surrogate": 1000, "mm": 2000, "copy": 10}[args.workload]
    if args.warmup < 0:
        args.warmup = 20 if args.workload == "batch" else 100
    {"airebo": main_airebo, "rigid": main_rigid, "core": main_core, "batch": main_batch,
     "surrogate": main_surrogate, "mm": main_mm, "copy": main_copy}[args.workload](args)


if __name__ == "__main__":
    main()
