"""Lanes per own vertex in the gathers of the one-iteration kernel (1, 2 or
4, from the workgroup size of 256, 512 or 1024 threads over the tile size). A vertex's gather is four
interleaved partial sums combined as (P0 + P1) + (P2 + P3) whatever the lane
count, and the element and gather order follows the global element id, so
fixed-length frames must have the same bits for every tile and workgroup
size. The two-iteration and persistent kernels (one lane per vertex) are
compared with it bit for bit in their own tests."""
import numpy as np
import pytest

import clothhip as ch
from helpers import CFG1

pytestmark = pytest.mark.gpu

CASES = {
    "cfg1_corners": dict(CFG1),
    "irregular24_top": dict(nrows=24, ncols=24, size=1 / 23, type=ch.MESH_IRREGULAR,
                            rest_jitter=0.3, seed=11, stretch_k=5e2, stretch_k_max=5e3,
                            bend_k=1e-2, pin_mode=ch.PIN_TOP_ROW, pos_jitter=1e-3, dt=1 / 60),
    "cross20_corners": dict(nrows=20, ncols=20, size=0.05, type=3, stretch_k=100.0, bend_k=1.0,
                            pin_mode=ch.PIN_CORNERS, pos_jitter=1e-3, dt=0.01),
    "checker40_no_bend": dict(CFG1, nrows=40, ncols=40, size=1 / 39, type=1, bend_k=0.0),
}

# (tile_verts, block_threads) -> lanes per vertex
SHAPES = {(256, 256): 1, (128, 256): 2, (256, 512): 2, (64, 256): 4, (128, 512): 4,
          (32, 512): 4, (256, 1024): 4, (64, 1024): 4}


def run(spec, shape, regions, frames=3, K=17, wind=None):
    s = ch.Scene(**spec)
    s.upload()
    c = dict(mode=ch.MODE_PD, gravity=-9.81, dt=spec["dt"], rho=0.99, cheb_delay=1,
             kernel_variant=ch.KERNEL_FUSED, max_iterations=K, element_regions=regions,
             tile_verts=shape[0], block_threads=shape[1])
    if wind is not None:
        c["ext_accel"] = wind
    sol = ch.Solver(ch.config(**c), s.N, s.T, s.E)
    args = s.step_args()
    for _ in range(frames):
        sol.step(args)
    return s.get_state(), sol.plan_info()


@pytest.mark.parametrize("regions", [1, 2])
@pytest.mark.parametrize("case", list(CASES))
def test_fixed_frames_equal_for_every_lane_count(case, regions):
    spec = CASES[case]
    ref = None
    seen = set()
    for shape, lanes in SHAPES.items():
        (pos, vel), info = run(spec, shape, regions, wind=(0.3, 0.0, -0.7))
        assert info.tile_verts == shape[0] and info.block_threads == shape[1]
        assert info.gather_lanes == lanes
        seen.add(lanes)
        if ref is None:
            ref = (pos, vel)
            continue
        np.testing.assert_array_equal(pos, ref[0], err_msg=f"positions at {shape}")
        np.testing.assert_array_equal(vel, ref[1], err_msg=f"velocities at {shape}")
    assert seen == {1, 2, 4}


def test_reference_mode_lane_counts_agree():
    """The reference update (FP32 positions, unfused multiply-adds) takes the
    same partial sums."""
    spec = CASES["irregular24_top"]
    outs = []
    for shape in ((256, 256), (128, 256), (64, 256)):
        s = ch.Scene(**spec)
        s.upload()
        sol = ch.Solver(ch.config(mode=ch.MODE_REF_CUDA, gravity=-9.81, dt=spec["dt"],
                                  max_iterations=12, kernel_variant=ch.KERNEL_FUSED,
                                  tile_verts=shape[0], block_threads=shape[1]), s.N, s.T, s.E)
        args = s.step_args()
        for _ in range(2):
            sol.step(args)
        outs.append(s.get_state())
    for pos, vel in outs[1:]:
        np.testing.assert_array_equal(pos, outs[0][0])
        np.testing.assert_array_equal(vel, outs[0][1])


def test_two_iteration_setting_at_1024_threads():
    """The two-iteration kernel runs at 256, 512 and 1024 threads: at 1024 and
    512 the frames run two iterations per launch, with the bits of one
    iteration per launch."""
    spec = CASES["irregular24_top"]
    outs = []
    for th, ipl in ((1024, 2), (512, 2), (1024, 1)):
        s = ch.Scene(**spec)
        s.upload()
        sol = ch.Solver(ch.config(mode=ch.MODE_PD, gravity=-9.81, dt=spec["dt"], rho=0.99,
                                  cheb_delay=1, kernel_variant=ch.KERNEL_FUSED,
                                  max_iterations=16, iters_per_launch=ipl, tile_verts=128,
                                  block_threads=th), s.N, s.T, s.E)
        args = s.step_args()
        for _ in range(2):
            sol.step(args)
        info = sol.plan_info()
        assert info.block_threads == th
        assert info.iters_per_launch == ipl
        outs.append(s.get_state())
    for pos, vel in outs[1:]:
        np.testing.assert_array_equal(outs[0][0], pos)
        np.testing.assert_array_equal(outs[0][1], vel)


def test_block_threads_validated():
    s = ch.Scene(**CFG1)
    with pytest.raises(ch.ClothHipError, match="block_threads"):
        ch.Solver(ch.config(mode=ch.MODE_PD, block_threads=768), s.N, s.T, s.E)
