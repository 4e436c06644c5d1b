"""clh_step_noalloc, the entry point that allocates nothing: through the C
ABI, its frames equal clh_solver_step's with the unfused kernels bit for bit,
in both modes, on two calls in a row whose second mesh is smaller (the
static scratch array keeps the first call's contents). Its refusals, which
need no device, are in test_abi.py."""
import ctypes as C

import numpy as np
import pytest

import clothhip as ch
from helpers import CFG1

pytestmark = pytest.mark.gpu

BIG = dict(CFG1, nrows=40, ncols=40, size=1 / 39, seed=21, bend_k=1e-2, pos_jitter=1e-3)
SMALL = dict(nrows=16, ncols=16, size=1 / 15, type=ch.MESH_IRREGULAR, rest_jitter=0.3, seed=22,
             stretch_k=5e2, stretch_k_max=5e3, bend_k=1e-2, pin_mode=ch.PIN_TOP_ROW,
             pos_jitter=1e-3, dt=1 / 60)


def scene(spec):
    s = ch.Scene(**spec)
    s.upload()
    return s


def noalloc(cfg, s, args):
    st = ch.hip().clh_step_noalloc(C.byref(cfg), s.N, s.T, s.E, C.byref(args), None)
    ch.check(st, "clh_step_noalloc")


def frames_both_ways(spec, cfg_kw, frames, wind=None):
    """Positions and velocities after `frames` frames from the same start,
    stepped by clh_step_noalloc and by an unfused clh_solver_step."""
    out = []
    for use_noalloc in (True, False):
        s = scene(spec)
        args = s.step_args()
        acc = None
        if wind is not None:
            acc = ch.DeviceArray.from_host(np.asarray(wind, np.float32))
            args.d_cloth_accel = acc.ptr
        cfg = ch.config(kernel_variant=ch.KERNEL_UNFUSED, **cfg_kw)
        if use_noalloc:
            for _ in range(frames):
                noalloc(cfg, s, args)
        else:
            solver = ch.Solver(cfg, s.N, s.T, s.E)
            for _ in range(frames):
                solver.step(args)
            solver.close()
        out.append(s.get_state())
        s.close()
    return out


@pytest.mark.parametrize("mode", ["pd", "ref"])
def test_noalloc_equals_unfused_solver_two_calls(mode):
    if mode == "pd":
        kw = dict(mode=ch.MODE_PD, max_iterations=24, rho=0.99, cheb_delay=2, gravity=-9.81,
                  dt=1 / 60)
    else:
        kw = dict(mode=ch.MODE_REF_CUDA, max_iterations=10, gravity=-9.81, dt=1 / 60)
    print(f"noalloc {mode}: 40x40 then 16x16", flush=True)
    # first the larger mesh, then a smaller one over what it left in scratch
    (pa, va), (pb, vb) = frames_both_ways(BIG, kw, 3, wind=[[0.0, 0.0, 5.0]])
    assert np.all(np.isfinite(pa))
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
    (pa, va), (pb, vb) = frames_both_ways(SMALL, kw, 2)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)
