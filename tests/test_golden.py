"""The golden fixtures (tests/golden/, written by make_golden.py): the host
C++ API and the oracle reproduce them exactly. The reference holds no
fixtures of its own; these pin this repository's restatements against drift
and give the GPU tests fixed vectors."""
import os

import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
import ws_oracle as wo
from helpers import CFG1, oracle_mesh, scene_arrays

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def make_golden():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


MG = make_golden()


@pytest.mark.parametrize("name", sorted(MG.SMALL))
def test_scene_fixtures_bit_exact(name):
    want = load(f"scene_{name}.npz")
    s = ch.Scene(**MG.SMALL[name])
    for k, v in want.items():
        np.testing.assert_array_equal(s.get(k), v, err_msg=k)


def test_one_iteration_fixture():
    want = load("iteration1_8x8.npz")
    got = MG.one_iteration()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_trajectory_fixture():
    want = load("trajectory_32x32.npz")
    got = MG.trajectories()
    for k in want:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    # the modes relate as they must: PD in FP32 storage stays within 1e-5 of
    # the extent of FP64, the reference mode's update does not (it drifts)
    L = 1.0
    assert np.abs(want["pd_f32_frame10"] - want["pd_f64_frame10"]).max() < 1e-5 * L
    assert np.abs(want["ref_cuda_f32_frame10"] - want["pd_f64_frame10"]).max() > 1e-3


def test_mlp_fixture():
    want = load("mlp_64_samples.npz")
    Fv = wo.bf16_value(want["features_bf16"]).astype(np.float64)
    for mode, em in (("exact", False), ("emulated", True)):
        loss, g = wo.loss_grad(want["params"], Fv, want["targets"], emulate=em)
        np.testing.assert_allclose(loss, want[f"loss_{mode}"], rtol=1e-12)
        np.testing.assert_allclose(g, want[f"grad_{mode}"], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(wo.forward(want["params"], Fv, emulate=False), want["out_exact"],
                               rtol=1e-12)
