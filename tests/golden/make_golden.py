"""Writes the golden fixtures of this directory (run from the repository
root: python tests/golden/make_golden.py).

The reference ships no fixtures or known-answer data, so these vectors come
from this repository's host C++ API (mesh generator, topology, rest state,
bend weights, Jacobi diagonal, pins) and its oracle (oracle/pd_oracle.cpp,
oracle/ws_oracle.py), single-threaded. They pin both against drift; the
tests that read them check the HIP path against the same vectors.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (os.path.join(ROOT, "cloth-cuda_amd"), os.path.join(ROOT, "oracle"),
          os.path.join(ROOT, "tests")):
    sys.path.insert(0, p)
os.environ.setdefault("CLOTH_QUIET", "1")

import clothhip as ch  # noqa: E402
import pyoracle as po  # noqa: E402
import ws_oracle as wo  # noqa: E402
from helpers import CFG1, oracle_mesh, scene_arrays  # noqa: E402

SCENE_ARRAYS = ["rest_pos", "pos0", "tris", "edges", "Dm_inv", "rest_area", "mass", "bend_quads",
                "bend_w", "bend_n", "jacobi_diag", "pin_idx", "pin_target"]

SMALL = {f"grid4_type{t}": dict(nrows=4, ncols=4, size=0.25, type=t, seed=1, stretch_k=1e3,
                                bend_k=1e-2, pin_mode=ch.PIN_CORNERS, pos_jitter=0.0, dt=1 / 60)
         for t in range(4)}
SMALL["irregular6"] = dict(nrows=6, ncols=6, size=0.2, type=ch.MESH_IRREGULAR, rest_jitter=0.3,
                           seed=9, stretch_k=5e2, stretch_k_max=5e3, bend_k=1e-2,
                           pin_mode=ch.PIN_TOP_ROW, pos_jitter=1e-3, dt=1 / 60)
ITER8 = dict(nrows=8, ncols=8, size=1 / 7, type=0, seed=8, stretch_k=1e3, bend_k=1e-2,
             pin_mode=ch.PIN_CORNERS, pos_jitter=1e-3, dt=1 / 60)
TRAJ_FRAMES = (1, 5, 10)


def scenes():
    out = {}
    for name, spec in SMALL.items():
        s = ch.Scene(**spec)
        out[name] = {k: s.get(k) for k in SCENE_ARRAYS}
    return out


def one_iteration():
    """x_1 of one frame from a seeded 8x8 state with a seeded velocity:
    reference mode (FP32) and PD mode (FP32 storage and FP64)."""
    s = ch.Scene(**ITER8)
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    pos = arr["pos0"]
    vel = np.random.default_rng(8).normal(scale=0.05, size=pos.shape).astype(np.float32)
    out = {"pos": pos, "vel": vel}
    p = po.params(mode=po.REF_CUDA, iterations=1, gravity=-9.81, dt=ITER8["dt"], threads=1)
    _, _, tr, _, _ = po.step(m, p, pos, vel, "f32", trace=True)
    out["ref_cuda_x1"] = tr[0]
    for prec in ("f32", "f64"):
        p = po.params(mode=po.PD, iterations=1, gravity=-9.81, dt=ITER8["dt"], rho=0.99,
                      cheb_delay=1, threads=1)
        _, _, tr, _, _ = po.step(m, p, pos, vel, prec, trace=True)
        out[f"pd_{prec}_x1"] = np.asarray(tr[0], np.float64 if prec == "f64" else np.float32)
    return out


def trajectories():
    """Positions after frames 1, 5 and 10 of the 32x32 workload in each mode."""
    s = ch.Scene(**CFG1)
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    out = {}
    for tag, mode, prec in (("ref_cuda_f32", po.REF_CUDA, "f32"), ("pd_f32", po.PD, "f32"),
                            ("pd_f64", po.PD, "f64")):
        p = po.params(mode=mode, iterations=10, gravity=-9.81, dt=CFG1["dt"], rho=0.99,
                      cheb_delay=1, threads=1)
        x, v = arr["pos0"].copy(), np.zeros((s.N, 3), np.float32)
        if prec == "f64":
            x, v = x.astype(np.float64), v.astype(np.float64)
        for f in range(1, max(TRAJ_FRAMES) + 1):
            x, v, *_ = po.step(m, p, x, v, prec)
            if f in TRAJ_FRAMES:
                out[f"{tag}_frame{f}"] = x.copy()
    return out


def mlp():
    """Loss and gradient of the warm-start network on 64 seeded samples."""
    rng = np.random.default_rng(64)
    p = rng.normal(scale=0.3, size=wo.N_PARAMS).astype(np.float32)
    p[wo.O_W3 + 3 * 64:wo.O_B3] = 0.0
    p[wo.O_B3 + 3:] = 0.0
    F = wo.bf16_bits(rng.normal(size=(64, 16)).astype(np.float32))
    T = np.concatenate([rng.normal(size=(64, 3)), (rng.uniform(size=(64, 1)) > 0.1)],
                       axis=1).astype(np.float32)
    Fv = wo.bf16_value(F).astype(np.float64)
    loss_x, g_x = wo.loss_grad(p, Fv, T, emulate=False)
    loss_e, g_e = wo.loss_grad(p, Fv, T, emulate=True)
    return {"params": p, "features_bf16": F, "targets": T, "loss_exact": np.float64(loss_x),
            "grad_exact": g_x, "loss_emulated": np.float64(loss_e), "grad_emulated": g_e,
            "out_exact": wo.forward(p, Fv, emulate=False)}


def main():
    for name, arrays in scenes().items():
        np.savez_compressed(os.path.join(HERE, f"scene_{name}.npz"), **arrays)
    np.savez_compressed(os.path.join(HERE, "iteration1_8x8.npz"), **one_iteration())
    np.savez_compressed(os.path.join(HERE, "trajectory_32x32.npz"), **trajectories())
    np.savez_compressed(os.path.join(HERE, "mlp_64_samples.npz"), **mlp())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
