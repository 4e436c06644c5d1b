"""Two PD iterations per launch (iters_per_launch = 2, the k_fused_pair
kernel): each tile also updates its ring-1 halo in the first iteration, from
ring-1 and ring-2 positions, with the element records, gather order and
update function of the vertex's own tile. Frames must therefore equal the
single-iteration launches bit for bit (positions, velocities, iteration
counts, residual histories): fixed-length frames with an even and an odd
iteration count, frames to a tolerance with restarts, graph-captured and
direct chunks, batched cloths with per-cloth wind, several tile and
workgroup sizes, and the 256x256 workload frame of the benchmark."""
import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
from helpers import CFG1, CFG2, CFG2_WIND, extent, oracle_mesh, rel_err, scene_arrays
from test_batch import SPECS

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

SHAPES = [(64, 256), (128, 256), (128, 512), (256, 512), (128, 1024), (256, 1024)]


def run(spec, ipl, frames, shape=None, accel=None, **kw):
    s = ch.Scene(**spec)
    s.upload()
    c = dict(mode=ch.MODE_PD, gravity=-9.81, dt=spec["dt"], rho=0.99, cheb_delay=1,
             kernel_variant=ch.KERNEL_FUSED, iters_per_launch=ipl)
    if shape:
        c["tile_verts"], c["block_threads"] = shape
    c.update(kw)
    sol = ch.Solver(ch.config(**c), s.N, s.T, s.E)
    args = s.step_args()
    acc = None
    if accel is not None:
        acc = ch.DeviceArray.from_host(np.ascontiguousarray(accel, np.float32).reshape(1, 3))
        args.d_cloth_accel = acc.ptr
    its, hist = [], []
    for _ in range(frames):
        sol.step(args)
        st = sol.stats()
        its.append((st.iterations, st.converged))
        hist.append(sol.residual_history(4096))
    return s.get_state(), its, hist, sol.plan_info()


def assert_same(a, b):
    (pa, va), ita, ha, _ = a
    (pb, vb), itb, hb, _ = b
    assert ita == itb
    for x, y in zip(ha, hb):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("K", [20, 21])
@pytest.mark.parametrize("case", list(CASES))
def test_fixed_frames_equal_single_iteration_launches(case, K, shape):
    spec = CASES[case]
    one = run(spec, 1, 3, shape, max_iterations=K)
    two = run(spec, 2, 3, shape, max_iterations=K)
    assert one[3].iters_per_launch == 1
    # the two-iteration path runs wherever its tables fit in the LDS a
    # workgroup may use
    assert two[3].pair == 1
    assert two[3].iters_per_launch == 2 or two[3].pair_lds_bytes > two[3].lds_limit
    assert_same(one, two)


@pytest.mark.parametrize("use_graph", [0, 1])
@pytest.mark.parametrize("case", ["cfg1_corners", "irregular24_top"])
def test_tolerance_frames_equal_single_iteration_launches(case, use_graph):
    """Device-controlled frames: every check falls on a pair's second
    iteration, the stop and the Chebyshev restarts (rho close to 1 makes the
    residual grow between checks) happen at the same iterations."""
    spec = CASES[case]
    kw = dict(max_iterations=3000, tolerance=1e-6, check_every=10, rho=0.9999,
              use_graph=use_graph)
    one = run(spec, 1, 3, **kw)
    two = run(spec, 2, 3, **kw)
    assert two[3].iters_per_launch == 2
    assert_same(one, two)


def test_odd_check_period_falls_back_to_single_launches():
    """With an odd check period a check would fall on a pair's first
    iteration; such frames run one iteration per launch and stay equal."""
    kw = dict(max_iterations=600, tolerance=1e-6, check_every=7)
    assert_same(run(CFG1, 1, 2, **kw), run(CFG1, 2, 2, **kw))


def test_odd_max_iterations_to_tolerance():
    """A frame capped at an odd iteration count ends with a single launch
    whose check stops it."""
    kw = dict(max_iterations=37, tolerance=1e-12, check_every=10)
    one, two = run(CFG1, 1, 2, **kw), run(CFG1, 2, 2, **kw)
    assert [i for i, _ in two[1]] == [37, 37]
    assert_same(one, two)


def test_cloth_accel_and_pins():
    accel = np.array([0.7, -1.2, 2.5], np.float32)
    spec = CASES["irregular24_top"]
    assert_same(run(spec, 1, 3, accel=accel, max_iterations=16),
                run(spec, 2, 3, accel=accel, max_iterations=16))


def run_batch(specs, ipl, tol, relax=0.0):
    b = ch.Scene(specs=specs)
    b.upload()
    acc = ch.DeviceArray.from_host(
        np.random.default_rng(7).uniform(-3.0, 3.0, size=(len(specs), 3)).astype(np.float32))
    c = ch.config(mode=ch.MODE_PD, max_iterations=4000 if tol else 14, gravity=-9.81, dt=1 / 60,
                  rho=0.999, cheb_delay=1, tolerance=tol, check_every=20, iters_per_launch=ipl,
                  relax=relax)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    args.d_cloth_accel = acc.ptr
    its = []
    for _ in range(2):
        sol.step(args)
        its.append(sol.stats().iterations)
    assert sol.plan_info().iters_per_launch == ipl
    return b.get_state(), its, sol.cloth_residuals() if tol else None


@pytest.mark.parametrize("tol,relax", [(0.0, 0.0), (1e-5, 0.0), (0.0, ch.RELAX_AUTO),
                                       (1e-5, ch.RELAX_AUTO)])
def test_batched_cloths(tol, relax):
    """The batch of the batching tests, per-cloth wind. To a tolerance with
    the plain Jacobi step the single-pin irregular cloth is left out: it
    folds and stalls near 4e-2 (in the oracle too). With each cloth's own
    damped step (CLH_RELAX_AUTO) it is kept, and every cloth converges."""
    specs = SPECS if (tol == 0.0 or relax != 0.0) else SPECS[:3]
    (p1, v1), it1, r1 = run_batch(specs, 1, tol, relax)
    (p2, v2), it2, r2 = run_batch(specs, 2, tol, relax)
    assert it1 == it2
    np.testing.assert_array_equal(p1, p2)
    np.testing.assert_array_equal(v1, v2)
    if tol:
        np.testing.assert_array_equal(r1, r2)
        assert np.all(r2 <= tol)


@pytest.mark.parametrize("shape,schedule", [
    ((128, 256), {}),
    ((256, 1024), dict(relax=ch.RELAX_AUTO, cheb_delay=10, restart_rule=ch.RESTART_SEQ_MIN,
                       restart_factor=10.0)),
])
def test_workload_frame_256(shape, schedule):
    """The benchmark's frame: 256x256, wind, rho 0.999999, relative residual
    1e-5 checked every 50 iterations, over a thousand iterations; also at the
    benchmark's damped schedule in 1024-thread workgroups."""
    spec = dict(CFG2)
    kw = dict(max_iterations=80000, tolerance=1e-5, check_every=50, rho=0.999999,
              ext_accel=CFG2_WIND, kernel_variant=ch.KERNEL_AUTO)
    kw.update(schedule)
    one = run(spec, 1, 1, shape, **kw)
    two = run(spec, 2, 1, shape, **kw)
    assert two[3].iters_per_launch == 2
    assert one[1][0][1] == 1 and one[1][0][0] > 1000
    assert_same(one, two)


def test_against_oracle_fp32():
    """One 20-iteration frame of two-iteration launches against the FP32
    oracle from the same state: 1e-5 of the extent."""
    s = ch.Scene(**CFG1)
    s.upload()
    arr = scene_arrays(s)
    pos, vel = s.get_state()
    sol = ch.Solver(ch.config(mode=ch.MODE_PD, max_iterations=20, gravity=-9.81,
                              dt=CFG1["dt"], rho=0.99, cheb_delay=1, iters_per_launch=2),
                    s.N, s.T, s.E)
    sol.step(s.step_args())
    assert sol.plan_info().iters_per_launch == 2
    p = po.params(mode=po.PD, iterations=20, gravity=-9.81, dt=CFG1["dt"], rho=0.99,
                  cheb_delay=1)
    x, *_ = po.step(oracle_mesh(arr), p, pos, vel, "f32")
    assert rel_err(s.get_state()[0], x, extent(pos)) < 1e-5


def test_timing_uses_the_launch_in_use():
    s = ch.Scene(**CFG1)
    s.upload()
    for ipl in (1, 2):
        sol = ch.Solver(ch.config(mode=ch.MODE_PD, max_iterations=10, gravity=-9.81,
                                  dt=CFG1["dt"], rho=0.99, cheb_delay=1, iters_per_launch=ipl),
                        s.N, s.T, s.E)
        args = s.step_args()
        sol.step(args)
        pos = s.get_state()[0].copy()
        assert sol.time_iteration_kernel(args, 20) > 0.0
        assert sol.plan_info().iters_per_launch == ipl
        # timing writes solver scratch only
        np.testing.assert_array_equal(s.get_state()[0], pos)


def test_autotune_chooses_between_one_and_two():
    """With iters_per_launch = 2 the autotuner also times the two-iteration
    kernel (per tile size: one iteration with one or two element regions at
    256, 512 and 1024 threads, two iterations at 256, 512 and 1024; 36
    candidates; a pair shape runs where its workgroup holds a tile's own and
    ring-1 vertices),
    keeps the shortest time per iteration, and the frames equal the
    single-iteration solver's whichever wins."""
    spec = CASES["irregular24_top"]
    one = run(spec, 1, 2, max_iterations=20)
    s = ch.Scene(**spec)
    s.upload()
    sol = ch.Solver(ch.config(mode=ch.MODE_PD, max_iterations=20, gravity=-9.81, dt=spec["dt"],
                              rho=0.99, cheb_delay=1, iters_per_launch=2), s.N, s.T, s.E)
    args = s.step_args()
    rep = sol.autotune(args, 5)
    cands = rep.candidates()
    assert rep.num_candidates == 36
    assert sorted({c["block_threads"] for c in cands if c["iters_per_launch"] == 2}) == \
        [256, 512, 1024]
    assert sorted({c["iters_per_launch"] for c in cands}) == [1, 2]
    timed = [c["us_per_iteration"] for c in cands if c["us_per_iteration"] > 0]
    assert rep.us_per_iteration[rep.best] == min(timed)
    assert any(c["iters_per_launch"] == 2 and c["us_per_iteration"] > 0 for c in cands)
    assert sol.plan_info().iters_per_launch == cands[rep.best]["iters_per_launch"]
    for _ in range(2):
        sol.step(args)
    np.testing.assert_array_equal(s.get_state()[0], one[0][0])
    np.testing.assert_array_equal(s.get_state()[1], one[0][1])


def test_iters_per_launch_validated():
    s = ch.Scene(**CFG1)
    with pytest.raises(ch.ClothHipError, match="iters_per_launch"):
        ch.Solver(ch.config(mode=ch.MODE_PD, iters_per_launch=3), s.N, s.T, s.E)
