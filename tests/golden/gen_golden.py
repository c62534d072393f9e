"""Write the golden vectors under tests/golden/ from the reference build.

Run in a container where the reference tree exists:

    make -C oracle ref && python tests/golden/gen_golden.py

It calls the reference functions through the C-ABI harnesses in oracle/_ref
(compiled from the reference sources where they lie, see oracle/Makefile) on
small seeded inputs and stores inputs and outputs as .npz (no pickles).  Only
numbers are committed; nothing of the reference source is.

Fixtures
  rng_normal.npz          first 256 draws of std::normal_distribution over
                          std::mt19937(42) (libstdc++), the stream the initial
                          velocities / orientations come from (C1:666-675).
  c60_analytic.npz        generate_c60 (C1:228-252).
  rigid_core_nc2.npz      C1 initial state, nc=2 (32 molecules, seed 42, 300 K),
                          list, forces, and a 30-step NPT trajectory with the
                          list rebuilt every 25 steps as C1's main does.
  rigid_core_nc3.npz      C1 initial state at nc=3 (config 1, 108 molecules):
                          list and forces.
  rigid_cc1_<X>.npz       LJ program functions on .cc1 molecules (C60, C70,
                          C84 No.01 in FCC; C72 and C84 No.20 in HCP, C76-Td
                          and C60 in BCC) with seeded random orientations:
                          list, forces, 3 NPT steps.
  airebo_nc1.npz          AB on a jittered 1x1x1 C60 crystal (240 atoms):
                          REBO / LJ lists, compute_rebo, compute_lj,
                          compute_forces and 5 NPT steps.
  airebo_nc2.npz          AB forces at 2x2x2 (1920 atoms), where the 400-entry
                          LJ half list truncates (inputs and outputs only).
  mm_c60_nc1.npz          MM on a jittered 1x1x1 C60 crystal (240 atoms): the
                          one-molecule topology of build_flat_topology, LJ
                          list, compute_forces without and with (random)
                          charges, and 5 NPT steps at dt = 0.1 fs.
  mm_c60_nc2.npz          MM forces at 2x2x2 (1920 atoms; the 400-entry half
                          list truncates): inputs and outputs only.
  mm_c70_hcp.npz          MM on C70 in a 2x2x2 HCP cell (1120 atoms): topology
                          and forces.
  airebo_cfg3_nc5.npz     AB at config 3's size, C60 FCC 5x5x5 (30,000 atoms,
                          +-0.01 A jitter): list counts and last LJ entries,
                          compute_forces, and 5 NPT steps (float32 inputs
                          and forces, stored exactly).
  airebo_rebuild_nc3.npz  AB on a jittered 3x3x3 C60 crystal (6480 atoms): 25
                          NPT steps with both lists rebuilt before step 20 as
                          AB's main loop does (AB:881-887); per-step energies
                          and the end state.
  core_ensemble_nc3.npz   the core program's main loop (C1:683-733) at nc=3
                          (config 1: 108 molecules, 1000 steps) for seeds
                          42..49: per-step T, P, a, Ecoh of seed 42 (the
                          program's own run, checked against the printed rows
                          and Avg line of the reference binary) and the
                          averages over steps 750..1000 (C1:624, 720) of every
                          seed, whose spread is the statistical error of one
                          run's average.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import ref  # noqa: E402

LIBDIR = os.path.join(os.path.dirname(os.path.dirname(HERE)), "fuller_md_en_amd",
                      "FullereneLib")


def save(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


def gen_rng():
    out = np.zeros(256)
    ref.core().refc_normal_stream(42, 256, out)
    save("rng_normal.npz", seed=42, draws=out)


def gen_core():
    body, props = ref.core_c60()
    save("c60_analytic.npz", coords=body, I0=props["I0"], Mmol=props["Mmol"], Rmol=props["Rmol"])
    L = ref.core()
    rmcut = L.refc_rmcut()
    for nc, nsteps in ((2, 30), (3, 0)):
        pos, vel, omg, qv, h = ref.core_init(nc)
        n = len(pos)
        hi = oracle.inv9(h)
        p = pos.reshape(-1).copy()
        L.refc_apply_pbc(p, h, hi, n)
        pos = p.reshape(n, 3)
        cnt, lst = ref.core_nlist(pos, h)
        ep, Fv, Tv, W = ref.core_forces(pos, qv, body, h, cnt, lst)
        rec = dict(nc=nc, pos=pos, vel=vel, omg=omg, qv=qv, h=h, body=body, I0=props["I0"],
                   Mmol=props["Mmol"], rmcut=rmcut, nl_count=cnt, nl_list=lst, Ep=ep, Fv=Fv,
                   Tv=Tv, Wm9=W)
        if nsteps:
            state = np.zeros(15)
            L.refc_make_npt(300.0, 0.0, n, state)
            rec["npt0"] = state.copy()
            P, V, Q, O = (pos.reshape(-1).copy(), vel.reshape(-1).copy(),
                          qv.reshape(-1).copy(), omg.reshape(-1).copy())
            F, T, Wm, hh = Fv.reshape(-1).copy(), Tv.reshape(-1).copy(), W.copy(), h.copy()
            hi2 = np.zeros(9)
            lab = np.zeros(n * 60 * 3)
            c, l = cnt.copy(), lst.reshape(-1).copy()
            eps, kes, Ts, Ps = [], [], [], []
            out = np.zeros(2)
            for g in range(1, nsteps + 1):
                if g % 25 == 0:  # C1:708-711
                    L.refc_mat_inv9(hh, hi2)
                    L.refc_nlist(P, hh, hi2, n, rmcut, c, l)
                L.refc_step_npt(P, V, Q, O, F, T, Wm, hh, hi2, body.reshape(-1), props["I0"],
                                props["Mmol"], n, 60, rmcut ** 2, 1.0, state, c, l, lab, out)
                eps.append(out[0])
                kes.append(out[1])
                kt = oracle.lib().orc_ke_trans(V, n, props["Mmol"])
                Ts.append(oracle.lib().orc_inst_T(out[1], int(state[14])))
                Ps.append(oracle.lib().orc_inst_P(Wm, kt, abs(oracle.det9(hh))))
            rec.update(nsteps=nsteps, traj_Ep=eps, traj_KE=kes, traj_T=Ts, traj_P=Ps,
                       pos_end=P.reshape(n, 3), vel_end=V.reshape(n, 3), qv_end=Q.reshape(n, 4),
                       omg_end=O.reshape(n, 3), h_end=hh, npt_end=state, Fv_end=F.reshape(n, 3),
                       Tv_end=T.reshape(n, 3), Wm9_end=Wm)
        save(f"rigid_core_nc{nc}.npz", **rec)


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


CC1_CASES = [("C60", "C60-76/C60-Ih.cc1", "FCC"), ("C70", "C60-76/C70-D5h.cc1", "FCC"),
             ("C84", "C84/C84-No.01-D2.cc1", "FCC")]
# other lattices and species (written by `gen_golden.py lattices`)
LATTICE_CASES = [("C72_HCP", "C60-76/C72-D6d.cc1", "HCP"),
                 ("C76Td_BCC", "C60-76/C76-Td.cc1", "BCC"),
                 ("C84Td_HCP", "C84/C84-No.20-Td.cc1", "HCP"),
                 ("C60_BCC", "C60-76/C60-Ih.cc1", "BCC")]


def gen_cc1(cases=CC1_CASES):
    L = ref.lj()
    for tag, rel, st in cases:
        coords = np.zeros(84 * 3)
        props = np.zeros(5)
        natom = L.reflj_load_cc1(os.path.join(LIBDIR, rel).encode(), coords, props)
        body = coords[: natom * 3].reshape(natom, 3).copy()
        rmol, dmol, I0, Mmol = props[1], props[2], props[3], props[4]
        a0 = L.reflj_default_a0(dmol, st.encode(), 1.0)
        nc = 2
        pos = np.zeros(8 * nc ** 3 * 3)
        h = np.zeros(9)
        n = L.reflj_make_crystal(st.encode(), a0, nc, pos, h)
        pos = pos[: 3 * n].copy()
        if st != "FCC":  # the program wraps before its first forces (LJ:827)
            L.reflj_apply_pbc(pos, h, oracle.inv9(h), n)
        rng = np.random.default_rng(2024 + natom)
        qv = random_quats(rng, n)
        sv = np.sqrt(oracle.KB * 300.0 * oracle.CONV / Mmol)
        sw = np.sqrt(oracle.KB * 300.0 * oracle.CONV / I0)
        vel = sv * rng.normal(size=(n, 3))
        vel -= vel.mean(axis=0)
        omg = sw * rng.normal(size=(n, 3))
        rmcut = 3.0 * 3.431 + 2 * rmol + 1.0  # LJ:716
        hi = oracle.inv9(h)
        cnt = np.zeros(n, np.int32)
        lst = np.zeros(n * 80, np.int32)
        L.reflj_nlist(pos, h, hi, n, rmcut, cnt, lst)
        Fv, Tv, W = np.zeros(n * 3), np.zeros(n * 3), np.zeros(9)
        lab = np.zeros(n * natom * 3)
        ep = L.reflj_forces(Fv, Tv, W, pos, qv.reshape(-1), body.reshape(-1), h, hi, cnt, lst,
                            n, natom, rmcut ** 2, lab)
        rec = dict(natom=natom, body=body, I0=I0, Mmol=Mmol, Rmol=rmol, Dmol=dmol, a0=a0, nc=nc,
                   crystal=st,
                   pos=pos.reshape(n, 3), qv=qv, vel=vel, omg=omg, h=h, rmcut=rmcut,
                   nl_count=cnt, nl_list=lst.reshape(n, 80), Ep=ep, Fv=Fv.reshape(n, 3),
                   Tv=Tv.reshape(n, 3), Wm9=W)
        state = np.zeros(15)
        L.reflj_make_npt(300.0, 0.0, n, state)
        rec["npt0"] = state.copy()
        P, V, Q, O = pos.copy(), vel.reshape(-1).copy(), qv.reshape(-1).copy(), omg.reshape(-1).copy()
        F, T, Wm, hh, hi2 = Fv.copy(), Tv.copy(), W.copy(), h.copy(), np.zeros(9)
        out = np.zeros(2)
        eps, kes = [], []
        for _ in range(3):
            L.reflj_step_npt(P, V, Q, O, F, T, Wm, hh, hi2, body.reshape(-1), I0, Mmol, n, natom,
                             rmcut ** 2, 1.0, state, cnt, lst, lab, out)
            eps.append(out[0])
            kes.append(out[1])
        rec.update(traj_Ep=eps, traj_KE=kes, pos_end=P.reshape(n, 3), vel_end=V.reshape(n, 3),
                   qv_end=Q.reshape(n, 4), omg_end=O.reshape(n, 3), h_end=hh, npt_end=state,
                   Fv_end=F.reshape(n, 3), Tv_end=T.reshape(n, 3), Wm9_end=Wm)
        save(f"rigid_cc1_{tag}.npz", **rec)


def airebo_system(nc, seed=12345, jitter=0.01):
    L = ref.airebo()
    coords = np.zeros(84 * 3)
    props = np.zeros(3)
    natom = L.refab_load_cc1(os.path.join(LIBDIR, "C60-76/C60-Ih.cc1").encode(), coords, props)
    coords = coords[: natom * 3].reshape(natom, 3)
    a0 = props[2] * np.sqrt(2.0) * 1.4  # default_a0 FCC (AB:246-248)
    bas = np.array([[0, 0, 0], [.5, .5, 0], [.5, 0, .5], [0, .5, .5]]) * a0
    cen = np.array([np.array([x, y, z]) * a0 + bas[b] for x in range(nc) for y in range(nc)
                    for z in range(nc) for b in range(4)])
    nm = len(cen)
    pos = (cen[:, None, :] + coords[None]).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape)
    h = np.diag([nc * a0] * 3).reshape(-1).astype(float)
    mol = np.repeat(np.arange(1, nm + 1), natom).astype(np.int32)
    p = pos.reshape(-1).copy()
    L.refab_apply_pbc(p, h, oracle.inv9(h), len(pos))
    return p.reshape(-1, 3), h, mol, a0


def pack_list(cnt, lst):
    return np.concatenate([lst[i, :cnt[i]] for i in range(len(cnt))]).astype(np.int32)


def gen_airebo():
    L = ref.airebo()
    # nc = 1: everything, including lists and a short trajectory
    pos, h, mol, a0 = airebo_system(1)
    na = len(pos)
    lists = ref.ab_nlists(pos, h, mol)
    nlc_r, nll_r, nlc_l, nll_l = lists
    hi = oracle.inv9(h)
    Fr, Wr = np.zeros(na * 3), np.zeros(9)
    er = L.refab_compute_rebo(Fr, Wr, pos.reshape(-1), h, hi, nlc_r, nll_r.reshape(-1), na)
    Fl, Wl = np.zeros(na * 3), np.zeros(9)
    el = L.refab_compute_lj(Fl, Wl, pos.reshape(-1), h, hi, nlc_l, nll_l.reshape(-1), na)
    E, F, W = ref.ab_compute_forces(pos, h, lists)
    rng = np.random.default_rng(99)
    mass = np.full(na, 12.011)
    vel = np.sqrt(oracle.KB * 300.0 * oracle.CONV / 12.011) * rng.normal(size=(na, 3))
    vel -= vel.mean(axis=0)
    state = np.zeros(15)
    L.refab_make_npt(300.0, 0.0, na, state)
    npt0 = state.copy()
    P, V, FF, WW, hh, hi2 = (pos.reshape(-1).copy(), vel.reshape(-1).copy(), F.reshape(-1).copy(),
                             W.copy(), h.copy(), hi.copy())
    out = np.zeros(4)
    traj = []
    r_cnt, r_lst, l_cnt, l_lst = (x.reshape(-1).copy() for x in lists)
    for _ in range(5):
        L.refab_step_npt(P, V, FF, WW, hh, hi2, mass, na, 0.5, state, r_cnt, r_lst, l_cnt, l_lst,
                         out)
        traj.append(out.copy())
    save("airebo_nc1.npz", pos=pos, h=h, mol_id=mol, a0=a0, nlc_r=nlc_r,
         nll_r=pack_list(nlc_r, nll_r), nlc_l=nlc_l, nll_l=pack_list(nlc_l, nll_l),
         E_rebo=er, F_rebo=Fr.reshape(na, 3), W_rebo=Wr, E_lj=el, F_lj=Fl.reshape(na, 3),
         W_lj=Wl, E=np.array(E), F=F, W=W, vel=vel, mass=mass, npt0=npt0, dt=0.5,
         traj=np.array(traj), pos_end=P.reshape(na, 3), vel_end=V.reshape(na, 3),
         F_end=FF.reshape(na, 3), W_end=WW, h_end=hh, npt_end=state)
    # nc = 2: the 400-entry truncation is active (inputs and outputs only)
    pos, h, mol, a0 = airebo_system(2, seed=777)
    lists = ref.ab_nlists(pos, h, mol)
    E, F, W = ref.ab_compute_forces(pos, h, lists)
    save("airebo_nc2.npz", pos=pos, h=h, mol_id=mol, a0=a0, nlc_r=lists[0], nlc_l=lists[2],
         lj_last=np.array([lists[3][i, lists[2][i] - 1] if lists[2][i] else -1
                           for i in range(len(pos))], np.int32),
         E=np.array(E), F=F, W=W)


def gen_airebo_cfg3():
    """Config 3 at its size: C60 FCC 5x5x5 (30,000 atoms), +-0.01 A jitter
    (numpy PCG64 seed 12345); positions and velocities rounded to float32 so
    that the fixture stores them exactly in half the bytes."""
    L = ref.airebo()
    pos, h, mol, a0 = airebo_system(5)
    pos = pos.astype(np.float32).astype(np.float64)
    na = len(pos)
    lists = ref.ab_nlists(pos, h, mol)
    nlc_r, nll_r, nlc_l, nll_l = lists
    E, F, W = ref.ab_compute_forces(pos, h, lists)
    rng = np.random.default_rng(99)
    vel = np.sqrt(oracle.KB * 300.0 * oracle.CONV / 12.011) * rng.normal(size=(na, 3))
    vel = (vel - vel.mean(axis=0)).astype(np.float32).astype(np.float64)
    mass = np.full(na, 12.011)
    state = np.zeros(15)
    L.refab_make_npt(300.0, 0.0, na, state)
    npt0 = state.copy()
    P, V, FF, WW, hh, hi = (pos.reshape(-1).copy(), vel.reshape(-1).copy(), F.reshape(-1).copy(),
                            W.copy(), h.copy(), oracle.inv9(h))
    out, traj = np.zeros(4), []
    r_cnt, r_lst, l_cnt, l_lst = (x.reshape(-1).copy() for x in lists)
    for _ in range(5):
        L.refab_step_npt(P, V, FF, WW, hh, hi, mass, na, 0.5, state, r_cnt, r_lst, l_cnt, l_lst,
                         out)
        traj.append(out.copy())
    # displacements by the minimum image of the initial cubic cell
    L0 = np.diag(h.reshape(3, 3))
    dpos = P.reshape(na, 3) - pos
    dpos -= np.round(dpos / L0) * L0
    save("airebo_cfg3_nc5.npz", pos=pos.astype(np.float32), vel=vel.astype(np.float32), h=h,
         a0=a0, nlc_r=nlc_r.astype(np.int8), nlc_l=nlc_l.astype(np.int16),
         lj_last=np.array([nll_l[i, nlc_l[i] - 1] if nlc_l[i] else -1 for i in range(na)],
                          np.int32),
         E=np.array(E), F=F.astype(np.float32), W=W, npt0=npt0, traj=np.array(traj),
         dpos_end=dpos.astype(np.float32), h_end=hh, npt_end=state)


def gen_airebo_rebuild():
    """C60 FCC 3x3x3 (6480 atoms, +-0.01 A jitter, seed 4242), 25 NPT steps of
    0.5 fs at 300 K with both lists rebuilt before step 20 (AB:881-887): the
    per-step energies and the end state.  Inputs and end state rounded to
    float32 for storage; the reference ran on the rounded inputs."""
    L = ref.airebo()
    pos, h, mol, a0 = airebo_system(3, seed=4242)
    pos = pos.astype(np.float32).astype(np.float64)
    na = len(pos)
    rng = np.random.default_rng(7)
    vel = np.sqrt(oracle.KB * 300.0 * oracle.CONV / 12.011) * rng.normal(size=(na, 3))
    vel = (vel - vel.mean(axis=0)).astype(np.float32).astype(np.float64)
    mass = np.full(na, 12.011)
    lists = ref.ab_nlists(pos, h, mol)
    E, F, W = ref.ab_compute_forces(pos, h, lists)
    state = np.zeros(15)
    L.refab_make_npt(300.0, 0.0, na, state)
    npt0 = state.copy()
    P, V, FF, WW, hh, hi = (pos.reshape(-1).copy(), vel.reshape(-1).copy(),
                            F.reshape(-1).copy(), W.copy(), h.copy(), oracle.inv9(h))
    out, traj = np.zeros(4), []
    r_cnt, r_lst, l_cnt, l_lst = (x.reshape(-1).copy() for x in lists)
    for g in range(1, 26):
        if g % 20 == 0:
            L.refab_mat_inv9(hh, hi)
            nl = ref.ab_nlists(P.reshape(na, 3), hh, mol)
            r_cnt, r_lst, l_cnt, l_lst = (x.reshape(-1).copy() for x in nl)
        L.refab_step_npt(P, V, FF, WW, hh, hi, mass, na, 0.5, state, r_cnt, r_lst, l_cnt,
                         l_lst, out)
        traj.append(out.copy())
    L0 = np.diag(h.reshape(3, 3))
    dpos = P.reshape(na, 3) - pos
    dpos -= np.round(dpos / L0) * L0
    save("airebo_rebuild_nc3.npz", pos=pos.astype(np.float32), vel=vel.astype(np.float32),
         h=h, mol_id=mol, npt0=npt0, traj=np.array(traj), dpos_end=dpos,
         vel_end=V.reshape(na, 3), h_end=hh, npt_end=state)


MM_K4 = np.array([469.0, 63.0, 14.5, 15.0]) / 23.06054783  # MM:1039-1042


def mm_system(spec, crystal, nc, seed, jitter=0.02):
    path = os.path.join(LIBDIR, spec)
    coords, bonds, props = ref.mm_load_cc1(path)
    L = ref.lj()
    a0 = L.reflj_default_a0(props[3], crystal.encode(), 1.0)
    nmax = 4 * nc ** 3
    cen, h = np.zeros(3 * nmax), np.zeros(9)
    L.reflj_make_crystal(crystal.encode(), a0, nc, cen, h)
    nm = (4 if crystal == "FCC" else 2) * nc ** 3
    cen = cen[: 3 * nm].reshape(nm, 3)
    pos = (cen[:, None, :] + coords[None]).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter, jitter, pos.shape)
    p = pos.reshape(-1).copy()
    ref.mm().refmm_apply_pbc(p, h, oracle.inv9(h), len(pos))
    topo = ref.mm_build(path, nm, MM_K4)
    return path, coords, bonds, p.reshape(-1, 3), h, nm, topo, rng


def mm_template(topo, nm):
    """Terms of molecule 0 (build_flat_topology replicates them, MM:314-332)."""
    out = {}
    for k in ("b", "bp", "a", "ap", "d", "dp", "dm", "i", "ip"):
        out["t_" + k] = topo[k][: len(topo[k]) // nm]
    return out


def gen_mm():
    L = ref.mm()
    # C60 FCC 1x1x1: everything
    path, coords, bonds, pos, h, nm, topo, rng = mm_system("C60-76/C60-Ih.cc1", "FCC", 1, 2024)
    na = len(pos)
    nlc, nll = ref.mm_nlist(pos, h)
    out, F, W = ref.mm_compute_forces(pos, h, nlc, nll)
    q = rng.uniform(-0.1, 0.1, na)
    out_q, F_q, W_q = ref.mm_compute_forces(pos, h, nlc, nll, charge=q)
    vel = np.sqrt(oracle.KB * 300.0 * oracle.CONV / 12.011) * rng.normal(size=(na, 3))
    vel -= vel.mean(axis=0)
    state = np.zeros(15)
    L.refmm_make_npt(300.0, 0.0, na, state)
    npt0 = state.copy()
    hi = oracle.inv9(h)
    P, V, FF, WW, hh, hi2 = (pos.reshape(-1).copy(), vel.reshape(-1).copy(),
                             F.reshape(-1).copy(), W.copy(), h.copy(), hi.copy())
    traj = []
    o8 = np.zeros(8)
    for _ in range(5):
        L.refmm_step_npt(P, V, FF, WW, hh, hi2, na, 0.1, state, nlc, nll.reshape(-1),
                         np.zeros(na), 0, o8)
        traj.append(o8.copy())
    save("mm_c60_nc1.npz", coords=coords, bonds=bonds, k4=MM_K4, pos=pos, h=h,
         mol_id=topo["mol_id"], nlc=nlc, nll=pack_list(nlc, nll), E=out, F=F, W=W, charge=q,
         E_q=out_q, F_q=F_q, W_q=W_q, vel=vel, npt0=npt0, dt=0.1, traj=np.array(traj),
         pos_end=P.reshape(na, 3), vel_end=V.reshape(na, 3), F_end=FF.reshape(na, 3),
         W_end=WW, h_end=hh, npt_end=state, **mm_template(topo, nm))
    # C60 FCC 2x2x2: the 400-entry truncation is active
    path, coords, bonds, pos, h, nm, topo, rng = mm_system("C60-76/C60-Ih.cc1", "FCC", 2, 99)
    nlc, nll = ref.mm_nlist(pos, h)
    out, F, W = ref.mm_compute_forces(pos, h, nlc, nll)
    save("mm_c60_nc2.npz", pos=pos, h=h, mol_id=topo["mol_id"], nlc=nlc,
         lj_last=np.array([nll[i, nlc[i] - 1] if nlc[i] else -1 for i in range(len(pos))],
                          np.int32), E=out, F=F, W=W)
    # C70 HCP 2x2x2: another species, non-orthogonal cell
    path, coords, bonds, pos, h, nm, topo, rng = mm_system("C60-76/C70-D5h.cc1", "HCP", 2, 7)
    nlc, nll = ref.mm_nlist(pos, h)
    out, F, W = ref.mm_compute_forces(pos, h, nlc, nll)
    save("mm_c70_hcp.npz", coords=coords, bonds=bonds, k4=MM_K4, pos=pos, h=h,
         mol_id=topo["mol_id"], nlc=nlc, E=out, F=F, W=W, **mm_template(topo, nm))


def _core_run(seed, nc=3, nsteps=1000):
    series = np.zeros(4 * nsteps)
    ref.core().refc_main_loop(nc, seed, nsteps, series)
    return series.reshape(nsteps, 4)


def gen_core_ensemble():
    import re
    import subprocess
    from multiprocessing import Pool
    seeds = list(range(42, 50))
    with Pool(8) as pool:
        runs = pool.map(_core_run, seeds)
    avg_from = 1000 - 1000 // 4  # C1:624
    avgs = np.array([r[avg_from - 1:].mean(axis=0) for r in runs])
    # the harness loop is the program's: its rows and Avg line equal the binary's
    out = subprocess.run([ref.path("fuller_LJ_core_omp"), "3"], capture_output=True, text=True,
                         check=True, env=dict(os.environ, OMP_NUM_THREADS="8")).stdout
    rows = [l.split() for l in out.splitlines() if re.match(r"\s+\d+\s+[-\d.]+\s", l)]
    for r in rows:
        g = int(r[0])
        want = [f"{runs[0][g - 1, 0]:.1f}", f"{runs[0][g - 1, 1]:.3f}", f"{runs[0][g - 1, 2]:.3f}",
                f"{runs[0][g - 1, 3]:.5f}"]
        assert r[1:5] == want, (r, want)
    m = re.search(r"Avg\((\d+)\): T=(\S+) P=(\S+) a=(\S+) Ecoh=(\S+)", out)
    assert int(m.group(1)) == 1000 - avg_from + 1
    assert [float(x) for x in m.group(2, 3, 4, 5)] == [
        float(f"{avgs[0, 0]:.2f}"), float(f"{avgs[0, 1]:.4f}"), float(f"{avgs[0, 2]:.4f}"),
        float(f"{avgs[0, 3]:.5f}")]
    save("core_ensemble_nc3.npz", seeds=np.array(seeds), avg_from=avg_from,
         series42=runs[0], avgs=avgs, columns=np.array(["T", "P", "a", "Ecoh"]))


if __name__ == "__main__":
    if not ref.available():
        sys.exit("oracle/_ref is not built: run `make -C oracle ref` where the reference exists")
    if sys.argv[1:] == ["lattices"]:
        gen_cc1(LATTICE_CASES)
        sys.exit(0)
    if sys.argv[1:] == ["airebo_cfg3"]:
        gen_airebo_cfg3()
        sys.exit(0)
    if sys.argv[1:] == ["airebo_rebuild"]:
        gen_airebo_rebuild()
        sys.exit(0)
    if sys.argv[1:] == ["core_ensemble"]:
        gen_core_ensemble()
        sys.exit(0)
    if sys.argv[1:] == ["mm"]:
        gen_mm()
        sys.exit(0)
    gen_rng()
    gen_core()
    gen_cc1()
    gen_airebo()
    gen_mm()
