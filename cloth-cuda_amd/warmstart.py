"""Frame loops around the learned warm start (clh_ws_* through clothhip).

The solver's first iterate of frame n+1 is one of
  * "plain":   x_n, as the reference starts every frame (src/pd_solver.cu:528);
  * "extrap":  the second-order extrapolation 3 x_n - 3 x_{n-1} + x_{n-2};
  * "learned": the extrapolation plus the network's change of acceleration;
and the frame then iterates to the solver's residual tolerance, so every
mode ends at the same accuracy. All state stays on the device: the previous
positions live in three rotating buffers, one device-to-device copy per frame.

The plain method is both the arbiter and the fallback: with a fallback solver
(the same configuration, plain start, full iteration cap) a warm-started
frame runs under the warm solver's own smaller cap, and a frame that misses
the tolerance within it is redone from (x_n, v_n) by the fallback solver, so
its result is the plain frame's. After `strikes` consecutive fallbacks the
warm start pauses for `cooldown` frames. Before it is used, `select_gain`
compares the warm starts with the plain start on held-out frames; a frame
loop without a warm start (ws None) runs plain frames in every mode.
"""
from __future__ import annotations

import time

import numpy as np

import clothhip as ch

MODES = ("plain", "extrap", "learned")


class FrameLoop:
    """Steps `scene` with `solver`, keeping the position history the warm
    start reads."""

    def __init__(self, scene, solver, ws=None, cloth_accel=None, fallback=None, strikes=3,
                 cooldown=20):
        self.scene, self.solver, self.ws = scene, solver, ws
        self.fallback, self.strikes, self.cooldown = fallback, strikes, cooldown
        self.N = scene.N
        nbytes = self.N * 3 * 4
        self.hist = [ch.DeviceArray(nbytes) for _ in range(3)]
        self.guess = ch.DeviceArray(nbytes)
        self.vel = ch.DeviceArray(nbytes) if fallback is not None else None
        self.args = scene.step_args()
        self.accel = None
        if cloth_accel is not None:
            self.accel = ch.DeviceArray.from_host(np.ascontiguousarray(cloth_accel, np.float32))
            self.args.d_cloth_accel = self.accel.ptr
        self.weight = ch.DeviceArray(max(getattr(scene, "C", 1), 1) * 4)
        self.kept = self.offered = 0
        self.fallbacks = self.paused_frames = 0
        self._misses = self._pause = 0
        self.n = 0
        self.restart()

    def _copy(self, dst_ptr, src_ptr):
        ch.check(ch.hip().clh_memcpy(dst_ptr, src_ptr, self.N * 3 * 4, ch.CLH_D2D), "clh_memcpy")

    def restart(self):
        """History = the current state repeated (a cloth at rest)."""
        for h in self.hist:
            self._copy(h.ptr, self.scene.device("pos"))
        self.n = 0

    def _plain(self, solver):
        self.args.d_init_guess = None
        solver.step(self.args)
        return solver.stats()

    def step(self, mode="plain", dataset=False):
        """One frame; returns the iterations it took (a fallback frame counts
        both attempts). With dataset, the frame's (history, converged
        positions) samples go to the warm start's training set."""
        n = self.n
        xm2, xm1, x = self.hist[(n + 1) % 3], self.hist[(n + 2) % 3], self.hist[n % 3]
        self._copy(x.ptr, self.scene.device("pos"))
        warm = mode != "plain" and self.ws is not None and self._pause == 0
        if mode != "plain" and self.ws is not None and self._pause > 0:
            self._pause -= 1
            self.paused_frames += 1
        used = self.solver
        if not warm:
            used = self.fallback if self.fallback is not None else self.solver
            iters = self._plain(used).iterations
        else:
            if self.fallback is not None:
                self._copy(self.vel.ptr, self.scene.device("vel"))
            self.ws.predict(xm2, xm1, x, self.guess, self.accel, use_network=(mode == "learned"))
            self.args.d_init_guess = self.guess.ptr
            self.solver.step(self.args)
            st = self.solver.stats()
            iters = st.iterations
            if self.fallback is not None:
                if st.converged:
                    self._misses = 0
                else:
                    # the plain frame from the same state replaces it
                    self._copy(self.scene.device("pos"), x.ptr)
                    self._copy(self.scene.device("vel"), self.vel.ptr)
                    used = self.fallback
                    iters += self._plain(used).iterations
                    self.fallbacks += 1
                    self._misses += 1
                    if self._misses >= self.strikes:
                        self._pause, self._misses = self.cooldown, 0
        if dataset:
            # only cloths whose frame met the tolerance become training samples
            tol = used.cfg.tolerance
            r = used.cloth_residuals()
            wgt = np.where(r <= tol, 1.0, 0.0).astype(np.float32) if tol > 0 else \
                np.ones(len(r), np.float32)
            self.weight.upload(wgt)
            self.ws.dataset_add(xm2, xm1, x, self.scene.device("pos"), self.accel, self.weight)
            self.kept += int(wgt.sum())
            self.offered += len(wgt)
        self.n += 1
        return iters

    def run(self, frames, mode="plain", dataset=False):
        """frames frames; returns (iterations per frame, seconds)."""
        its = []
        ch.check(ch.hip().clh_synchronize(), "sync")
        t0 = time.perf_counter()
        for _ in range(frames):
            its.append(self.step(mode, dataset))
        ch.check(ch.hip().clh_synchronize(), "sync")
        return its, time.perf_counter() - t0


def select_gain(scene, solver, ws, gains=(0.0, 0.25, 0.5, 1.0), warmup=3, frames=6,
                cloth_accel=None, plain=True):
    """Chooses the factor on the network's correction by the iterations it
    saves on held-out frames: from the same state, each gain runs `warmup`
    plain frames (to fill the position history) and then `frames` learned
    frames on `scene` with `solver`, which should iterate to the tolerance
    the warm start is used with. Gain 0 is the extrapolation alone, so on
    these frames the chosen start is never slower than it. With `plain`, the
    plain start runs the same frames too and is the choice (returned as gain
    None) when no gain needs fewer iterations: the warm start is then not to
    be used. Sets the chosen gain on `ws` (its output scale times the gain;
    unchanged for None), restores the scene's state and returns
    (gain, {gain or "plain": mean iterations per frame})."""
    params, scale = ws.params()
    pos0, vel0 = scene.get_state()
    res = {}
    if plain:
        loop = FrameLoop(scene, solver, None, cloth_accel)
        loop.run(warmup, mode="plain")
        its, _ = loop.run(frames, mode="plain")
        res["plain"] = float(np.mean(its))
    for g in gains:
        scene.set_state(pos0, vel0)
        ws.set_params(params, scale * g)
        loop = FrameLoop(scene, solver, ws, cloth_accel)
        loop.run(warmup, mode="plain")
        its, _ = loop.run(frames, mode="learned")
        res[g] = float(np.mean(its))
    best = min(gains, key=lambda g: res[g])
    if plain and res["plain"] <= res[best]:
        best = None
    ws.set_params(params, scale * (1.0 if best is None else best))
    scene.set_state(pos0, vel0)
    return best, res
