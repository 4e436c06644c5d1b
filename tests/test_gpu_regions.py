"""Separate triangle and hinge LDS regions (element_regions = 2): the fused
kernel evaluates triangles and hinges in one phase and sums both after one
barrier, in the same order as with one shared region, so frames are the
same bit for bit: fixed-length frames, frames to a tolerance (check launches
fall back to the shared region where their larger layout does not fit),
batched cloths, several tile and workgroup sizes."""
import numpy as np
import pytest

import clothhip as ch
from helpers import CFG2, CFG2_WIND
from test_batch import SPECS
from test_gpu_pair import CASES, SHAPES, assert_same, run

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("case", list(CASES))
def test_fixed_frames_equal_shared_region(case, shape):
    spec = CASES[case]
    one = run(spec, 1, 3, shape, max_iterations=20, element_regions=1)
    two = run(spec, 1, 3, shape, max_iterations=20, element_regions=2)
    assert one[3].element_regions == 1
    # separate regions wherever their layout fits the LDS a workgroup may use
    assert two[3].element_regions == 2 or two[3].split_lds_bytes > two[3].lds_limit
    assert_same(one, two)


@pytest.mark.parametrize("case", ["cfg1_corners", "irregular24_top"])
def test_tolerance_frames_equal_shared_region(case):
    spec = CASES[case]
    kw = dict(max_iterations=3000, tolerance=1e-6, check_every=10, rho=0.9999)
    one = run(spec, 1, 3, element_regions=1, **kw)
    two = run(spec, 1, 3, element_regions=2, **kw)
    assert two[3].element_regions == 2
    assert_same(one, two)


def test_batched_cloths_equal_shared_region():
    b = ch.Scene(specs=SPECS)
    b.upload()
    pos0, vel0 = b.get_state()
    acc = ch.DeviceArray.from_host(
        np.random.default_rng(7).uniform(-3.0, 3.0, size=(len(SPECS), 3)).astype(np.float32))
    out = []
    for reg in (1, 2):
        b.set_state(pos0, vel0)
        c = ch.config(mode=ch.MODE_PD, max_iterations=14, gravity=-9.81, dt=1 / 60, rho=0.999,
                      cheb_delay=1, element_regions=reg)
        sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
        args = b.step_args()
        args.d_cloth_accel = acc.ptr
        for _ in range(2):
            sol.step(args)
        assert sol.plan_info().element_regions == reg
        out.append(b.get_state())
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_workload_frame_256():
    kw = dict(max_iterations=80000, tolerance=1e-5, check_every=50, rho=0.999999,
              ext_accel=CFG2_WIND, kernel_variant=ch.KERNEL_AUTO)
    one = run(dict(CFG2), 1, 1, (128, 512), element_regions=1, **kw)
    two = run(dict(CFG2), 1, 1, (128, 512), element_regions=2, **kw)
    assert two[3].element_regions == 2
    assert_same(one, two)


def test_reference_mode_ignores_regions():
    a = run(CASES["cfg1_corners"], 1, 2, mode=ch.MODE_REF_CUDA, max_iterations=10,
            element_regions=1)
    b = run(CASES["cfg1_corners"], 1, 2, mode=ch.MODE_REF_CUDA, max_iterations=10,
            element_regions=2)
    assert b[3].element_regions == 1
    assert_same(a, b)


def test_element_regions_validated():
    s = ch.Scene(**CASES["cfg1_corners"])
    with pytest.raises(ch.ClothHipError, match="element_regions"):
        ch.Solver(ch.config(mode=ch.MODE_PD, element_regions=3), s.N, s.T, s.E)
