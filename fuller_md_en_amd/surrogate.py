"""Learned REBO surrogate (fmd_sg_* in include/fmd.h) attached to an AIREBO engine.

A per-atom energy model E = sum_i MLP(D_i) over 40 invariant same-molecule
descriptors (6 A), three tanh layers of width 64 on BF16 MFMA, forces as exact
gradients.  It stands in for the exact-gradient REBO term (FMD_REBO_PHYSICAL)
after training on frames the engine generates; while enabled, full REBO
re-runs every ``judge_every`` steps and the engine falls back to it when the
force RMSE exceeds ``threshold``.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

NPARAM = None
NFEAT = None
H = 64
# Matrix-core work per atom-step: the network's forward products
# 2 (40*64 + 64*64 + 64*64 + 64) and the same again for the input gradient
# that the forces need (the descriptor and force projections run on the
# vector units and are not counted).
FLOP_PER_ATOM = 2 * 2 * (40 * H + H * H + H * H + H)
# parameter layout of theta (offsets into the flat FP32 vector)
_SHAPES = None


def _layout():
    global NPARAM, NFEAT, _SHAPES
    if NPARAM is None:
        NPARAM = N.lib().fmd_sg_param_count()
        NFEAT = N.lib().fmd_sg_feature_count()
        _SHAPES = [("W1", (NFEAT, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)),
                   ("W3", (H, H)), ("b3", (H,)), ("w4", (H,)), ("b4", (1,))]
    return NPARAM, NFEAT


def unpack(theta):
    """Flat theta -> dict of arrays (W1, b1, ..., b4)."""
    _layout()
    out, k = {}, 0
    for name, shape in _SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[k:k + n], np.float64).reshape(shape)
        k += n
    return out


def pack(params):
    """dict of arrays -> flat FP32 theta."""
    _layout()
    return np.concatenate([np.asarray(params[name], np.float32).reshape(-1)
                           for name, _ in _SHAPES]).astype(np.float32)


class Surrogate:
    def __init__(self, engine, seed: int = 0):
        _layout()
        self.engine = engine
        self._h = engine._h
        N.check(N.lib().fmd_sg_attach(self._h, seed))

    def set_params(self, theta=None, mu=None, sd=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (theta, mu, sd)]
        N.check(N.lib().fmd_sg_set_params(self._h, *[None if a is None else a.ctypes.data
                                                      for a in arrs]))

    def get_params(self):
        th = np.zeros(NPARAM, np.float32)
        mu = np.zeros(NFEAT, np.float32)
        sd = np.zeros(NFEAT, np.float32)
        N.check(N.lib().fmd_sg_get_params(self._h, th.ctypes.data, mu.ctypes.data, sd.ctypes.data))
        return th, mu, sd

    def collect(self):
        N.check(N.lib().fmd_sg_collect(self._h))

    def frames(self) -> int:
        n = C.c_int(0)
        N.check(N.lib().fmd_sg_frame_count(self._h, C.byref(n)))
        return n.value

    def fit_normalisation(self):
        N.check(N.lib().fmd_sg_fit_normalisation(self._h))

    def train(self, iters: int, lr: float = 1e-3):
        loss = np.zeros(iters)
        N.check(N.lib().fmd_sg_train(self._h, iters, lr, loss.ctypes.data))
        return loss

    def loss_grad(self, frame: int = 0):
        loss = np.zeros(1)
        grad = np.zeros(NPARAM, np.float32)
        N.check(N.lib().fmd_sg_loss_grad(self._h, frame, loss.ctypes.data, grad.ctypes.data))
        return loss[0], grad

    def descriptor(self, frame: int = 0):
        na = self.engine.n
        feat = np.zeros((na, NFEAT), np.float32)
        N.check(N.lib().fmd_sg_descriptor(self._h, frame, feat.ctypes.data))
        return feat

    def evaluate(self, frame: int = -1):
        """(E, F [Na, 3], W9, rmse vs the frame's REBO forces or -1)."""
        na = self.engine.n
        F = np.zeros((na, 3))
        W = np.zeros(9)
        E = np.zeros(1)
        r = np.zeros(1)
        N.check(N.lib().fmd_sg_evaluate(self._h, frame, F.ctypes.data, W.ctypes.data,
                                        E.ctypes.data, r.ctypes.data))
        return E[0], F, W, r[0]

    def enable(self, on: bool = True, judge_every: int = 100, threshold: float = 0.05):
        N.check(N.lib().fmd_sg_enable(self._h, 1 if on else 0, judge_every, threshold))

    def stats(self):
        r = C.c_double(0)
        j, f, a = C.c_int(0), C.c_int(0), C.c_int(0)
        N.check(N.lib().fmd_sg_stats(self._h, C.byref(r), C.byref(j), C.byref(f), C.byref(a)))
        return dict(last_rmse=r.value, judged=j.value, fallbacks=f.value, active=a.value)

    def save(self, path: str):
        th, mu, sd = self.get_params()
        np.savez(path, theta=th, mu=mu, sd=sd)

    def load(self, path: str):
        z = np.load(path)
        self.set_params(z["theta"], z["mu"], z["sd"])


def train_from_engine_runs(nc: int = 3, temps=(300.0, 800.0, 1400.0, 2000.0), frames_per_t: int = 12,
                           stride: int = 10, iters: int = 4000, lr: float = 2e-3, seed: int = 0,
                           log=None):
    """Train a surrogate on frames the engine generates: C60 FCC nc^3 with
    exact-gradient REBO, NPT runs at each temperature (velocities redrawn),
    one frame every `stride` steps.  Learning rate lr, decayed x0.3 at 60 %
    and 85 % of the iterations.  Returns (theta, mu, sd, loss history,
    frames)."""
    from . import airebo, system
    s = system.airebo_system(nc, "C60", T=temps[0], seed=seed + 11, jitter=0.02,
                             jitter_seed=seed + 3)
    eng = airebo.AtomicEngine(s["mol_id"], s["mass"], mode=airebo.REBO_PHYSICAL)
    sg = Surrogate(eng, seed=seed)
    step, pos = 0, s["pos"]
    for T in temps:
        vel = system.atomic_init_velocities(s["mass"], seed + int(T), T)
        eng.set_state(pos=pos, vel=vel, h=s["h"], npt=airebo.make_npt(T, 0.0, len(pos)))
        eng.run(step, 40, thermo=False)  # settle at the new temperature
        step += 40
        for _ in range(frames_per_t):
            eng.run(step, stride, thermo=False)
            step += stride
            sg.collect()
        pos = eng.get_state()["pos"]
    sg.fit_normalisation()
    losses = []
    cuts = [int(iters * 0.6), int(iters * 0.85), iters]
    start, rate = 0, lr
    for cut in cuts:
        losses.append(sg.train(cut - start, rate))
        if log:
            log(f"surrogate training: {cut} iterations, force RMSE {np.sqrt(losses[-1][-20:].mean()):.4f} eV/A")
        start, rate = cut, rate * 0.3
    th, mu, sd = sg.get_params()
    n = sg.frames()
    eng.close()
    return th, mu, sd, np.concatenate(losses), n
