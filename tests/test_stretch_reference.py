"""CpuStretchReferenceSolver (include/stretch_reference.h), the reference's CPU
backend (src/stretch_reference.cpp:140-280): stretch only, 50 iterations,
exact global solves with the pinned rows and columns zeroed. Checked against
the oracle's dense FP64 Cholesky of the same system (OR_EXACT with the
reference's pin treatment and no bending), against the reference's Test 6
bounds (src/tests/test_pd_constraints.cpp:367-390), for factor-once against
refactor-every-iteration, and through sim_cloth --stretch-backend cpu-ref.
The solver steps the host state, so none of this needs a device."""
import os
import subprocess

import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
from helpers import CFG1, extent, oracle_mesh, scene_arrays


def oracle_frames(arr, pos, vel, frames, dt, gravity, damping):
    """FP64 frames of the dense exact oracle, the state stored in FP32 between
    frames as the solver's host state (mesh.pos_cpu / vel_cpu) is."""
    m = oracle_mesh(arr, bend=False)
    p = po.params(mode=po.EXACT, iterations=50, gravity=gravity, dt=dt, damping=damping,
                  exact_pins=0, exact_bend=0)
    out = []
    for _ in range(frames):
        pos, vel, *_ = po.step(m, p, pos, vel, "f64")
        pos, vel = pos.astype(np.float32), vel.astype(np.float32)
        out.append((pos.copy(), vel.copy()))
    return out


def cpu_ref_frames(scene, pos, vel, frames, gravity, damping, refactor=False):
    scene.set_cpu_state(pos, vel)
    r = ch.CpuRefSolver(scene, gravity=gravity, damping=damping, refactor=refactor)
    out = []
    for _ in range(frames):
        r.step()
        out.append(scene.get_cpu_state())
    n = r.factorizations()
    r.close()
    return out, n


def test_cfg1_frames_match_dense_exact_oracle():
    # the 32x32 workload (bending ignored by this backend, as in the
    # reference), from its jittered start, corners pinned
    s = ch.Scene(**CFG1)
    arr = scene_arrays(s)
    pos, vel = arr["pos0"], np.zeros_like(arr["pos0"])
    frames = 10
    got, nfact = cpu_ref_frames(s, pos, vel, frames, -9.81, 0.0)
    assert nfact == 1  # one factorization for the whole run: the pins never change
    # frame by frame from the same FP32 state: the two FP64 solves agree to
    # their rounding; the state is then stored in FP32 by both
    ref = oracle_frames(arr, pos, vel, 1, 1 / 60, -9.81, 0.0)
    e = extent(arr["rest_pos"])
    assert np.abs(got[0][0] - ref[0][0]).max() < 1e-6 * e
    for f in range(1, frames):
        one = oracle_frames(arr, got[f - 1][0], got[f - 1][1], 1, 1 / 60, -9.81, 0.0)
        assert np.abs(got[f][0] - one[0][0]).max() < 1e-6 * e, f
        assert np.abs(got[f][1] - one[0][1]).max() < 1e-6 * e * 60, f
    # free run of both for the 10 frames: the per-frame differences (the
    # solver takes D from the mesh's FP32 dF/dx, as the reference does, the
    # oracle from Dm^-1 in FP64) accumulate, about 1.2e-6 of the extent per
    # frame; the bound is the multi-frame one of the workloads, 1e-4
    ref = oracle_frames(arr, pos, vel, frames, 1 / 60, -9.81, 0.0)
    assert np.abs(got[-1][0] - ref[-1][0]).max() < 1e-4 * e
    # pins hold their targets; the cloth moved
    pins = arr["pin_idx"]
    np.testing.assert_array_equal(got[-1][0][pins], arr["pin_target"])
    assert np.abs(got[-1][0] - pos).max() > 1e-3


def test_reference_test6_scene():
    # Test 6: 20x20 type 3, size 0.05, stretch 100, top row pinned, dt 0.01,
    # g -9.8, damping 0.99, 20 steps from rest: y and z ranges > 0.02
    s = ch.Scene(nrows=20, ncols=20, size=0.05, type=3, stretch_k=100.0, bend_k=0.0,
                 pin_mode=ch.PIN_TOP_ROW, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    rest = arr["rest_pos"]
    got, _ = cpu_ref_frames(s, rest, np.zeros_like(rest), 20, -9.8, 0.99)
    pos = got[-1][0]
    assert np.ptp(pos[:, 1]) > 0.02 and np.ptp(pos[:, 2]) > 0.02
    ref = oracle_frames(arr, rest, np.zeros_like(rest), 20, 0.01, -9.8, 0.99)
    assert np.abs(pos - ref[-1][0]).max() < 1e-4 * extent(rest)
    assert np.abs(pos[:20, 1]).max() == 0.0  # the pinned row stays at y = 0


def test_reference_pin_semantics_differ_from_correct_elimination():
    # the reference zeroes the pinned columns without moving K_up x_p to the
    # right-hand side (src/stretch_reference.cpp:225-255): its result must
    # match that treatment and differ from the correct one where the pins sit
    # away from the origin
    s = ch.Scene(nrows=6, ncols=6, size=0.2, type=0, stretch_k=50.0, bend_k=0.0,
                 pin_mode=ch.PIN_CORNERS, pos_jitter=0.0, dt=1 / 60)
    arr = scene_arrays(s)
    rest = arr["rest_pos"]
    got, _ = cpu_ref_frames(s, rest, np.zeros_like(rest), 1, -9.81, 0.0)
    m = oracle_mesh(arr, bend=False)
    p = dict(mode=po.EXACT, iterations=50, gravity=-9.81, dt=1 / 60, exact_bend=0)
    x_ref, *_ = po.step(m, po.params(exact_pins=0, **p), rest, np.zeros_like(rest), "f64")
    x_cor, *_ = po.step(m, po.params(exact_pins=1, **p), rest, np.zeros_like(rest), "f64")
    assert np.abs(got[0][0] - x_ref).max() < 1e-6
    assert np.abs(got[0][0] - x_cor).max() > 1e-3


def test_factor_once_equals_refactor_every_iteration():
    s = ch.Scene(nrows=12, ncols=12, size=1 / 11, type=ch.MESH_IRREGULAR, rest_jitter=0.3,
                 seed=8, stretch_k=2e2, stretch_k_max=2e3, bend_k=0.0, pin_mode=ch.PIN_TOP_ROW,
                 pos_jitter=1e-3, dt=1 / 60)
    arr = scene_arrays(s)
    pos, vel = arr["pos0"], np.zeros_like(arr["pos0"])
    a, na = cpu_ref_frames(s, pos, vel, 3, -9.81, 0.0)
    b, nb = cpu_ref_frames(s, pos, vel, 3, -9.81, 0.0, refactor=True)
    assert na == 1 and nb == 3 * 50
    for (pa, va), (pb, vb) in zip(a, b):
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(va, vb)
    # the irregular mesh, per-triangle stiffness and all: against the oracle
    ref = oracle_frames(arr, pos, vel, 3, 1 / 60, -9.81, 0.0)
    assert np.abs(a[-1][0] - ref[-1][0]).max() < 1e-5 * extent(arr["rest_pos"])


EXE = os.path.join(os.path.dirname(ch.LIB_HIP), "sim_cloth")


def test_sim_cloth_cpu_ref_backend(tmp_path):
    # the reference tool's CPU backend run: no device needed, PLY frames of
    # the host state, pinned row at y = 0, free vertices fall
    out = tmp_path / "frames"
    out.mkdir()
    r = subprocess.run([EXE, "20", "20", "0.05", "--steps", "20", "--stiffness", "100",
                        "--stretch-backend", "cpu-ref", "--damping", "0.99", "--export", str(out),
                        "--verbose"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "backend: cpu-ref" in r.stdout and "Final energy:" in r.stdout
    lines = (out / "frame_final.ply").read_text().splitlines()
    nv = int(lines[2].split()[-1])
    xyz = np.array([[float(t) for t in l.split()] for l in lines[9:9 + nv]])
    assert np.abs(xyz[:20, 1]).max() < 1e-6
    assert np.ptp(xyz[:, 1]) > 0.02 and np.ptp(xyz[:, 2]) > 0.02
