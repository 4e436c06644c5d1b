"""The HIP path against the golden fixtures (tests/golden/): one iteration
of each mode from the stored 8x8 state, the 32x32 trajectory of PD mode, and
the warm-start network's loss and gradient on the stored samples."""
import os

import numpy as np
import pytest

import clothhip as ch
import ws_oracle as wo
from helpers import CFG1
from test_golden import MG, load

pytestmark = pytest.mark.gpu


def frame(spec, cfg, pos, vel, trace_iters=0):
    s = ch.Scene(**spec)
    s.upload()
    s.set_state(pos, vel)
    sol = ch.Solver(ch.config(**cfg), s.N, s.T, s.E)
    args = s.step_args()
    tr = None
    if trace_iters:
        tr = ch.DeviceArray(trace_iters * s.N * 12)
        args.d_trace = tr.ptr
    sol.step(args)
    x, v = s.get_state()
    return x, v, (tr.download(np.float32, (trace_iters, s.N, 3)) if tr else None)


def test_first_iterate_matches_fixture():
    g = load("iteration1_8x8.npz")
    dt = MG.ITER8["dt"]
    _, _, tr = frame(MG.ITER8, dict(mode=ch.MODE_REF_CUDA, max_iterations=1, gravity=-9.81, dt=dt,
                                    use_graph=0), g["pos"], g["vel"], 1)
    assert np.abs(tr[0] - g["ref_cuda_x1"]).max() < 1e-5
    _, _, tr = frame(MG.ITER8, dict(mode=ch.MODE_PD, max_iterations=1, gravity=-9.81, dt=dt,
                                    rho=0.99, cheb_delay=1, use_graph=0), g["pos"], g["vel"], 1)
    assert np.abs(tr[0] - g["pd_f32_x1"]).max() < 1e-6
    assert np.abs(tr[0] - g["pd_f64_x1"]).max() < 1e-6


def test_trajectory_matches_fixture():
    g = load("trajectory_32x32.npz")
    s = ch.Scene(**CFG1)
    s.upload()
    sol = ch.Solver(ch.config(mode=ch.MODE_PD, max_iterations=10, gravity=-9.81, dt=CFG1["dt"],
                              rho=0.99, cheb_delay=1), s.N, s.T, s.E)
    args = s.step_args()
    for f in range(1, max(MG.TRAJ_FRAMES) + 1):
        sol.step(args)
        if f in MG.TRAJ_FRAMES:
            x, _ = s.get_state()
            assert np.abs(x - g[f"pd_f64_frame{f}"]).max() < 1e-5, f
            assert np.abs(x - g[f"pd_f32_frame{f}"]).max() < 1e-5, f


def test_mlp_loss_grad_matches_fixture():
    g = load("mlp_64_samples.npz")
    s = ch.Scene(**CFG1)
    s.upload()
    ws = ch.WarmStart.for_scene(s, CFG1["dt"], -9.81)
    ws.set_params(g["params"], 1.0)
    ws.dataset_append(g["features_bf16"], g["targets"])
    loss, grad = ws.loss_grad(0, 64)
    assert abs(loss - float(g["loss_emulated"])) <= 1e-4 * float(g["loss_emulated"])
    ref = g["grad_emulated"]
    assert np.linalg.norm(grad - ref) / np.linalg.norm(ref) < 2e-3
    # and the exact (FP64, no BF16) gradient is within BF16 rounding of it
    assert np.linalg.norm(grad - g["grad_exact"]) / np.linalg.norm(g["grad_exact"]) < 3e-2
