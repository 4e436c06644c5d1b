"""Batched cloths on the GPU (clh_batch_create, one fused launch per
iteration for every cloth): each cloth of a batch equals the same cloth
solved alone, bit for bit, in both modes and with per-cloth wind; frames run
to a tolerance stop only when every cloth meets it; the autotuner's choice
is cached and leaves results unchanged."""
import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
from helpers import cfg4_specs, extent, jacobi_bound_np, oracle_mesh, scene_arrays
from test_batch import SPECS

pytestmark = pytest.mark.gpu

DT = 1 / 60
RNG = np.random.default_rng(7)
ACCEL = RNG.uniform(-3.0, 3.0, size=(len(SPECS), 3)).astype(np.float32)


def batch_scene(specs=SPECS):
    b = ch.Scene(specs=specs)
    b.upload()
    return b


def cfg(mode, K, **kw):
    c = dict(mode=mode, max_iterations=K, gravity=-9.81, dt=DT, rho=0.99, cheb_delay=1)
    c.update(kw)
    return ch.config(**c)


def run_batch(b, c, frames, accel=None, solver=None):
    sol = solver or ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    acc = None
    if accel is not None:
        acc = ch.DeviceArray.from_host(np.ascontiguousarray(accel, np.float32))
        args.d_cloth_accel = acc.ptr
    for _ in range(frames):
        sol.step(args)
    pos, vel = b.get_state()
    return pos, vel, sol


def run_solo(spec, c, frames, accel=None, variant=ch.KERNEL_AUTO):
    s = ch.Scene(**spec)
    s.upload()
    c.kernel_variant = variant
    sol = ch.Solver(c, s.N, s.T, s.E)
    args = s.step_args()
    acc = None
    if accel is not None:
        acc = ch.DeviceArray.from_host(np.ascontiguousarray(accel, np.float32).reshape(1, 3))
        args.d_cloth_accel = acc.ptr
    for _ in range(frames):
        sol.step(args)
    return s.get_state()


@pytest.mark.parametrize("mode,kw", [
    (ch.MODE_PD, {}),
    (ch.MODE_REF_CUDA, {}),
    (ch.MODE_PD, dict(relax=ch.RELAX_AUTO, iters_per_launch=1)),
    (ch.MODE_PD, dict(relax=ch.RELAX_AUTO, iters_per_launch=2)),
], ids=["pd", "ref_cuda", "pd_relaxed", "pd_relaxed_pairs"])
def test_batch_equals_solo_bitwise(mode, kw):
    """With CLH_RELAX_AUTO each cloth of a batch takes the step scale of its
    own bound (here from 1 for the unbent cloth to 0.63 for the folding
    one), so it still equals its solo frame bit for bit, in one- and
    two-iteration launches."""
    b = batch_scene()
    K, F = 12, 3
    pos, vel, _ = run_batch(b, cfg(mode, K, **kw), F, ACCEL)
    for c, spec in enumerate(SPECS):
        sp, sv = run_solo(spec, cfg(mode, K, **kw), F, ACCEL[c])
        lo, hi = b.vert_off[c], b.vert_off[c + 1]
        np.testing.assert_array_equal(pos[lo:hi], sp, err_msg=f"positions of cloth {c}")
        np.testing.assert_array_equal(vel[lo:hi], sv, err_msg=f"velocities of cloth {c}")


def test_cloth_accel_unfused_equals_fused():
    """The per-cloth acceleration of a single cloth enters the unfused
    prologue like the fused one (bitwise) and equals the same uniform
    acceleration given through the config."""
    spec = SPECS[1]
    a = ACCEL[1]
    K, F = 10, 2
    p_f, _ = run_solo(spec, cfg(ch.MODE_PD, K), F, a, ch.KERNEL_FUSED)
    p_u, _ = run_solo(spec, cfg(ch.MODE_PD, K), F, a, ch.KERNEL_UNFUSED)
    p_c, _ = run_solo(spec, cfg(ch.MODE_PD, K, ext_accel=tuple(a)), F, None, ch.KERNEL_UNFUSED)
    np.testing.assert_allclose(p_f, p_u, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(p_u, p_c)


def test_batch_against_oracle_fp64():
    """Each cloth of a 2-frame batched PD run within 1e-5 of its extent of the
    FP64 oracle run on that cloth alone with its wind. (The single-pin
    irregular cloth folds in its third frame: there the FP32 and FP64 oracles
    themselves part by 1e-3 of the extent, as a triangle passes through
    det F = 0 where the polar rotation jumps.)"""
    b = batch_scene()
    K, F = 15, 2
    pos0, vel0 = b.get_state()
    pos, _, _ = run_batch(b, cfg(ch.MODE_PD, K), F, ACCEL)
    for c, spec in enumerate(SPECS):
        s = ch.Scene(**spec)
        arr = scene_arrays(s)
        m = oracle_mesh(arr)
        p = po.params(mode=po.PD, iterations=K, gravity=-9.81, dt=DT, rho=0.99, cheb_delay=1,
                      ext_accel=tuple(float(x) for x in ACCEL[c]))
        x, v = arr["pos0"].astype(np.float64), np.zeros((s.N, 3))
        for _ in range(F):
            x, v, *_ = po.step(m, p, x, v, "f64")
        lo, hi = b.vert_off[c], b.vert_off[c + 1]
        err = np.abs(pos[lo:hi] - x).max() / max(extent(arr["pos0"]), 1e-3)
        assert err < 1e-5, (c, err)


def test_batch_tolerance_every_cloth_converges():
    """Frames to a relative residual of 1e-5: the batch stops only when every
    cloth meets it. (The single-pin irregular cloth is left out here: with
    the plain Jacobi step, Chebyshev stalls on it near 4e-2 after it folds,
    in the oracle too; the damped step brings it back in the next test.)"""
    specs = SPECS[:3]
    b = batch_scene(specs)
    tol = 1e-5
    c = cfg(ch.MODE_PD, 20000, tolerance=tol, check_every=20, rho=0.999)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    acc = ch.DeviceArray.from_host(ACCEL[:3])
    args.d_cloth_accel = acc.ptr
    for _ in range(2):
        sol.step(args)
        st = sol.stats()
        r = sol.cloth_residuals()
        assert st.converged == 1 and len(r) == len(specs)
        assert np.all(r <= tol), r
        # the batch stops at the first check where the slowest cloth converged
        assert st.iterations % 20 == 0 or st.iterations == 20000


def test_batch_relaxed_every_cloth_converges():
    """The damped Jacobi step with each cloth's own bound (CLH_RELAX_AUTO):
    all four cloths, the single-pin irregular one included (Gershgorin bound
    3.145 of D^-1 K, step scale 0.636), reach a relative residual of 1e-5 in
    every one of four frames, as in the FP32 oracle with that scale (500,
    760, 920 and 1,040 iterations for the folding cloth). The stats report
    the largest bound and the smallest scale of the batch."""
    b = batch_scene(SPECS)
    tol = 1e-5
    c = cfg(ch.MODE_PD, 20000, tolerance=tol, check_every=20, rho=0.999,
            relax=ch.RELAX_AUTO)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    acc = ch.DeviceArray.from_host(ACCEL)
    args.d_cloth_accel = acc.ptr
    for f in range(4):
        sol.step(args)
        st, it = sol.cloth_status()
        r = sol.cloth_residuals()
        print(f"frame {f}: states {st.tolist()}, iterations {it.tolist()}, "
              f"residuals {np.array2string(r, precision=2)}", flush=True)
        assert np.all(st == ch.CLOTH_CONVERGED), st
        assert np.all(r <= tol), r
        assert sol.stats().converged == 1
        assert it[3] <= 2000, it
    pos, vel = b.get_state()
    assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
    bounds = [jacobi_bound_np(scene_arrays(ch.Scene(**spec)), DT) for spec in SPECS]
    stats = sol.stats()
    assert abs(stats.jacobi_bound - max(bounds)) <= 1e-12 * max(bounds)
    rho = float(np.float32(0.999))
    want = min(np.float32((1.0 + rho) / m) if m > 1.0 + rho else np.float32(1.0) for m in bounds)
    assert stats.relax == want


@pytest.mark.parametrize("relax", [0.0, ch.RELAX_AUTO], ids=["plain", "damped"])
def test_batch_tolerance_cloth_follows_its_solo_frame(relax):
    """Per-cloth control: in frames to a tolerance each cloth of a batch has
    its own Chebyshev sequence, restarts and convergence test and is frozen
    once it converges, so it follows the frames it would follow alone: the
    same stopping iteration and the same positions, bit for bit, although
    the cloths converge at different iterations. With the damped step each
    cloth also takes the step scale of its own bound, and the single-pin
    irregular cloth (left out with the plain step, where it stalls) is in."""
    specs = SPECS if relax else SPECS[:3]
    b = batch_scene(specs)
    tol = 1e-5
    kw = dict(tolerance=tol, check_every=20, rho=0.999, relax=relax)
    c = cfg(ch.MODE_PD, 20000, **kw)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    acc = ch.DeviceArray.from_host(ACCEL[:len(specs)])
    args.d_cloth_accel = acc.ptr
    iters = []
    for _ in range(2):
        sol.step(args)
        st, it = sol.cloth_status()
        assert np.all(st == ch.CLOTH_CONVERGED)
        assert sol.stats().iterations == it.max()
        iters.append(it)
    pos, vel = b.get_state()
    assert len(set(iters[0].tolist())) > 1  # the cloths stop at different checks
    for i, spec in enumerate(specs):
        s = ch.Scene(**spec)
        s.upload()
        solo = ch.Solver(cfg(ch.MODE_PD, 20000, **kw), s.N, s.T, s.E)
        a1 = s.step_args()
        acc1 = ch.DeviceArray.from_host(ACCEL[i:i + 1])
        a1.d_cloth_accel = acc1.ptr
        for f in range(2):
            solo.step(a1)
            assert solo.stats().iterations == iters[f][i], (i, f)
        sp, sv = s.get_state()
        lo, hi = b.vert_off[i], b.vert_off[i + 1]
        np.testing.assert_array_equal(pos[lo:hi], sp)
        np.testing.assert_array_equal(vel[lo:hi], sv)


def test_batch_with_folding_cloth_ends_early():
    """The single-pin irregular cloth folds, and Jacobi with Chebyshev then
    stalls on it near 4e-2 in the oracle, or diverges: where bending
    dominates, Jacobi modes with eigenvalues below -1 grow, and on the GPU
    this cloth's residual grows about 1e6 times per check from its tenth
    check on, alone as in a batch. The cloth stops as stalled (no 10 % drop
    of its relative residual in 100 checks, 2,000 iterations) or diverged
    (1000 times its best), and a diverged cloth returns to x_n with zero
    velocity; the other cloths converge and are frozen, the frame raises no
    error, and it ends long before the iteration cap."""
    b = batch_scene(SPECS)
    tol = 1e-5
    c = cfg(ch.MODE_PD, 20000, tolerance=tol, check_every=20, rho=0.999, stall_checks=100)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    args = b.step_args()
    acc = ch.DeviceArray.from_host(ACCEL)
    args.d_cloth_accel = acc.ptr
    for f in range(3):
        x_n, _ = b.get_state()
        sol.step(args)
        st, it = sol.cloth_status()
        r = sol.cloth_residuals()
        print(f"frame {f}: states {st.tolist()}, iterations {it.tolist()}, "
              f"residuals {np.array2string(r, precision=2)}", flush=True)
        pos, vel = b.get_state()
        assert np.all(np.isfinite(pos)) and np.all(np.isfinite(vel))
        assert np.all(st != ch.CLOTH_RUNNING)
        assert np.all(r[st == ch.CLOTH_CONVERGED] <= tol)
        assert np.all(st[:3] == ch.CLOTH_CONVERGED)
        assert sol.stats().iterations < 20000
        assert sol.stats().converged == int(np.all(st == ch.CLOTH_CONVERGED))
        for c in np.where(st == ch.CLOTH_NONFINITE)[0]:
            lo, hi = b.vert_off[c], b.vert_off[c + 1]
            np.testing.assert_array_equal(pos[lo:hi], x_n[lo:hi])
            assert np.all(vel[lo:hi] == 0.0)


def test_autotune_caches_and_keeps_results():
    b = batch_scene()
    c = cfg(ch.MODE_PD, 10)
    pos0, vel0 = b.get_state()
    ref, _, _ = run_batch(b, c, 2, ACCEL)
    b.set_state(pos0, vel0)
    sol = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    rep = sol.autotune(b.step_args(), 10)
    # 4 tile sizes x 3 workgroup sizes x 1 or 2 element regions (PD)
    assert rep.num_candidates == 24 and rep.from_cache == 0
    assert sorted({cnd["block_threads"] for cnd in rep.candidates()}) == [256, 512, 1024]
    cands = rep.candidates()
    times = [cnd["us_per_iteration"] for cnd in cands]
    assert rep.us_per_iteration[rep.best] == min(t for t in times if t > 0)
    assert all(cnd["iters_per_launch"] == 1 for cnd in cands)
    assert sorted({cnd["element_regions"] for cnd in cands}) == [1, 2]
    info = sol.plan_info()
    best = cands[rep.best]
    assert (info.tile_verts, info.block_threads, info.iters_per_launch, info.element_regions) == (
        best["tile_verts"], best["block_threads"], best["iters_per_launch"],
        best["element_regions"])
    # whatever shape wins, the frames are those of the default solver, bit for bit
    got, _, _ = run_batch(b, c, 2, ACCEL, solver=sol)
    np.testing.assert_array_equal(got, ref)
    sol2 = ch.Solver(c, b.N, b.T, b.E, cloth_offsets=b.vert_off)
    rep2 = sol2.autotune(b.step_args(), 10)
    assert rep2.from_cache == 1 and rep2.key == rep.key
    assert (rep2.tile_verts[0], rep2.block_threads[0], rep2.iters_per_launch[0],
            rep2.element_regions[0]) == (best["tile_verts"], best["block_threads"],
                                         best["iters_per_launch"], best["element_regions"])


def test_batch_rejects_unfused_and_bad_offsets():
    b = ch.Scene(specs=SPECS)
    with pytest.raises(ch.ClothHipError, match="fused"):
        ch.Solver(cfg(ch.MODE_PD, 5, kernel_variant=ch.KERNEL_UNFUSED), b.N, b.T, b.E,
                  cloth_offsets=b.vert_off)
    bad = b.vert_off.copy()
    bad[-1] += 1
    with pytest.raises(ch.ClothHipError, match="offsets"):
        ch.Solver(cfg(ch.MODE_PD, 5), b.N, b.T, b.E, cloth_offsets=bad)


def test_uneven_batch_sample_equals_solo():
    """32 cloths of the batched workload's distribution (sides up to 64), 15
    iterations: a sample of cloths equal to their solo runs bit for bit."""
    specs = cfg4_specs(32, side_max=64)
    b = batch_scene(specs)
    c = cfg(ch.MODE_PD, 15)
    pos, _, _ = run_batch(b, c, 2)
    for i in (0, 7, 19, 31):
        sp, _ = run_solo(specs[i], cfg(ch.MODE_PD, 15), 2)
        np.testing.assert_array_equal(pos[b.vert_off[i]:b.vert_off[i + 1]], sp)
