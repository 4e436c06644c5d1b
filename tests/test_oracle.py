"""Pins of the oracle (oracle/pd_oracle.cpp) against what fixes it
independently of its own formulas: hand-derived known answers of the
reference kernels, closed forms, invariants, numpy's SVD, brute-force exact
solves, and the reference's own test bounds
(src/tests/test_pd_constraints.cpp:151-390)."""
import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
from helpers import CFG1, oracle_mesh, scene_arrays


def embed(F2d, Q=None):
    """3x2 F = [e_x e_y] F2d (rotated by Q) as [col0, col1]."""
    F = np.zeros((3, 2))
    F[:2, :] = F2d
    if Q is not None:
        F = Q @ F
    return np.concatenate([F[:, 0], F[:, 1]])


def as3x2(v):
    return np.stack([v[:3], v[3:]], axis=1)


def rand_rot(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


# ---- stretch projections ----------------------------------------------------

def test_polar_known_answer_shear():
    # closest rotation to the shear [[1,1],[0,1]]: (cos, sin) ~ (2, -1)
    R = as3x2(po.polar_rotation(embed(np.array([[1.0, 1.0], [0.0, 1.0]]))))
    c, s = 2 / np.sqrt(5), -1 / np.sqrt(5)
    np.testing.assert_allclose(R[:2], [[c, -s], [s, c]], atol=1e-12)
    np.testing.assert_allclose(R[2], 0, atol=1e-12)


def test_polar_matches_numpy_svd():
    rng = np.random.default_rng(0)
    for _ in range(200):
        F = rng.normal(size=(3, 2))
        R = as3x2(po.polar_rotation(np.concatenate([F[:, 0], F[:, 1]])))
        # projection onto the plane of F, then the SO(2)-closest factor
        B, _ = np.linalg.qr(F)
        if np.dot(B[:, 0], F[:, 0]) < 0:
            B[:, 0] = -B[:, 0]
        if np.dot(B[:, 1], F[:, 1] - np.dot(F[:, 1], B[:, 0]) * B[:, 0]) < 0:
            B[:, 1] = -B[:, 1]
        F2 = B.T @ F
        U, S, Vt = np.linalg.svd(F2)
        R2 = U @ Vt
        if np.linalg.det(R2) < 0:
            U[:, 1] = -U[:, 1]
            R2 = U @ Vt
        np.testing.assert_allclose(R, B @ R2, atol=1e-9)


def test_polar_invariants():
    rng = np.random.default_rng(1)
    for _ in range(100):
        F = rng.normal(size=(3, 2))
        v = np.concatenate([F[:, 0], F[:, 1]])
        R = as3x2(po.polar_rotation(v))
        np.testing.assert_allclose(R.T @ R, np.eye(2), atol=1e-10)        # Stiefel
        Q = rand_rot(rng)
        RQ = as3x2(po.polar_rotation(np.concatenate([(Q @ F)[:, 0], (Q @ F)[:, 1]])))
        np.testing.assert_allclose(RQ, Q @ R, atol=1e-9)                  # equivariance
        Rs = as3x2(po.polar_rotation(3.7 * v))
        np.testing.assert_allclose(Rs, R, atol=1e-10)                     # scale invariance
        O, _ = np.linalg.qr(F)                                            # orthonormal F
        np.testing.assert_allclose(as3x2(po.polar_rotation(np.concatenate([O[:, 0], O[:, 1]]))),
                                   O, atol=1e-10)
    # f32 build agrees with f64 to rounding
    F = rng.normal(size=6)
    np.testing.assert_allclose(po.polar_rotation(F, "f32"), po.polar_rotation(F), atol=1e-6)


def test_ref_cuda_rotation_is_transposed_polar():
    # hand evaluation of src/pd_solver.cu:118-204 for F2d = [[1,1],[0,1]]:
    # A = [[1,1],[1,2]], v1 ~ (1, 1.618), and R2d = V W^T is the transpose of
    # the polar factor (the reference defect A1)
    R = as3x2(po.ref_cuda_rotation(embed(np.array([[1.0, 1.0], [0.0, 1.0]]))))
    c, s = 2 / np.sqrt(5), 1 / np.sqrt(5)
    np.testing.assert_allclose(R[:2], [[c, -s], [s, c]], atol=1e-6)
    # pure rotations and symmetric stretches are where the two agree
    th = 0.3
    rot = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    for F2 in (rot * 1.3, np.diag([2.0, 0.5]) @ np.eye(2)):
        np.testing.assert_allclose(po.ref_cuda_rotation(embed(F2)), po.polar_rotation(embed(F2)),
                                   atol=1e-6)


def test_ref_cuda_identity_fallback_branch():
    # |A01| < 1e-10 takes V = I (pd_solver.cu:131-135) while sigma1 is the
    # larger singular value: for F2d = diag(1, 2), w1 = (1/2, 0), w2 = (0, 2)
    # and R2d = V W^T = diag(0.5, 2), not a rotation (the reference defect A2)
    R = as3x2(po.ref_cuda_rotation(embed(np.diag([1.0, 2.0]))))
    np.testing.assert_allclose(R[:2], np.diag([0.5, 2.0]), atol=1e-6)


# ---- frame level ------------------------------------------------------------

def small_scene(**kw):
    base = dict(CFG1)
    base.update(nrows=8, ncols=8, size=1 / 7)
    base.update(kw)
    s = ch.Scene(**base)
    return s, scene_arrays(s)


def test_free_fall_is_exact():
    # zero stiffness, no bend: diag = mass and x_{n+1} = y = x + h v + h^2 g
    # in both modes
    s, arr = small_scene(pin_mode=ch.PIN_NONE, bend_k=0.0, stretch_k=0.0)
    np.testing.assert_array_equal(arr["jacobi_diag"], arr["mass"])
    m = oracle_mesh(arr, bend=False)
    rng = np.random.default_rng(2)
    pos = arr["pos0"].astype(np.float64)
    vel = rng.normal(size=pos.shape)
    h = float(np.float32(1 / 60))                     # the oracle takes FP32 dt and g
    y = pos + h * vel + h * h * np.array([0.0, float(np.float32(-9.81)), 0.0])
    for mode in (po.PD, po.REF_CUDA):
        x, v, *_ = po.step(m, po.params(mode=mode, iterations=7, gravity=-9.81, dt=1 / 60,
                                        rho=0.99), pos, vel, "f64")
        np.testing.assert_allclose(x, y, rtol=1e-14, atol=1e-14)


def test_ref_cuda_contracts_toward_origin():
    # reference defect A3: at rest with zero gravity, an interior vertex's
    # first update is (m y + h^2 sum wA D^T R)/diag with R = I, whose
    # stretch sum is the (zero) Laplacian of the rest positions, so
    # x_1 = m x / diag. PD mode keeps the rest state.
    s, arr = small_scene(pin_mode=ch.PIN_NONE, bend_k=0.0, pos_jitter=0.0)
    m = oracle_mesh(arr, bend=False)
    pos = arr["rest_pos"].astype(np.float64)
    vel = np.zeros_like(pos)
    x, v, tr, *_ = po.step(m, po.params(mode=po.REF_CUDA, iterations=1, gravity=0.0, dt=1 / 60),
                           pos, vel, "f64", trace=True)
    interior = np.array([i * 8 + j for i in range(1, 7) for j in range(1, 7)])
    expected = (arr["mass"][:, None] * pos / arr["jacobi_diag"][:, None])[interior]
    # the stretch sum cancels only to the FP32 rounding of the inputs (Dm_inv,
    # areas): terms of size h^2 wA |G| ~ 1e-3 leave ~1e-9
    np.testing.assert_allclose(tr[0][interior], expected, rtol=0, atol=1e-8)
    assert np.abs(tr[0][interior]).max() < 1e-2 * np.abs(pos[interior]).max()
    xpd, *_ = po.step(m, po.params(mode=po.PD, iterations=5, gravity=0.0, dt=1 / 60), pos, vel,
                      "f64")
    # PD stays at rest up to the FP32 rounding of Dm_inv (F = I to ~1e-7)
    np.testing.assert_allclose(xpd, pos, atol=1e-7)


def random_state(arr, scale, seed):
    rng = np.random.default_rng(seed)
    x = arr["rest_pos"].astype(np.float64) + scale * rng.normal(size=arr["rest_pos"].shape)
    y = x + 0.01 * scale * rng.normal(size=x.shape)
    return x, y


def test_residual_momentum_and_translation():
    # internal forces sum to zero (sum_i D_i = 0 and sum_i w_i ~ 0), so with
    # no pins sum_v r_v = sum_v m_v (y_v - x_v); r is translation invariant
    s, arr = small_scene(pin_mode=ch.PIN_NONE, bend_k=1e-2)
    m = oracle_mesh(arr)
    x, y = random_state(arr, 0.02, 3)
    r, r2, b2 = po.residual_f64(m, 1 / 60, x, y)
    inertia = (arr["mass"][:, None] * (y - x)).sum(0)
    scale = np.abs(r).sum()
    # bend weights are FP32, so sum_i w_i is zero only to FP32 rounding
    np.testing.assert_allclose(r.sum(0), inertia, atol=1e-6 * scale)
    c = np.array([0.3, -1.2, 2.5])
    r_t, *_ = po.residual_f64(m, 1 / 60, x + c, y + c)
    np.testing.assert_allclose(r_t, r, atol=1e-6 * np.abs(r).max())


def test_jacobi_chebyshev_converges_to_exact_local_global_fixed_point():
    # the PD fixed point does not depend on the global solver: Jacobi +
    # Chebyshev (one sweep per local step) and exact Cholesky solves per
    # local step must reach the same state; the exact path also has zero
    # residual as evaluated by the per-element gather
    s, arr = small_scene(nrows=6, ncols=6, size=0.2, pin_mode=ch.PIN_CORNERS, bend_k=1e-2,
                         stretch_k=50.0, pos_jitter=1e-3)
    m = oracle_mesh(arr)
    pos = arr["pos0"].astype(np.float64)
    vel = np.zeros_like(pos)
    common = dict(gravity=-9.81, dt=1 / 60)
    xe, ve, *_ = po.step(m, po.params(mode=po.EXACT, iterations=300, exact_bend=1, exact_pins=1,
                                      **common), pos, vel, "f64")
    xj, vj, *_ = po.step(m, po.params(mode=po.PD, iterations=20000, rho=0.95, **common), pos, vel,
                         "f64")
    assert np.abs(xj - xe).max() < 1e-9
    y = po.prediction(pos, vel, 1 / 60, -9.81)
    r, r2, b2 = po.residual_f64(m, 1 / 60, xe, y)
    assert np.sqrt(r2 / b2) < 1e-9


def test_exact_mode_reference_pin_semantics_differ():
    # the reference CPU path zeroes pinned rows and columns without moving
    # K_fp x_p to the right-hand side (src/stretch_reference.cpp:225-251):
    # unpinned neighbours see the pins at the origin. With the top corner at
    # x = 1 the two pin treatments must differ, and only the corrected one
    # has a zero residual.
    s, arr = small_scene(nrows=6, ncols=6, size=0.2, pin_mode=ch.PIN_CORNERS, bend_k=0.0,
                         stretch_k=50.0, pos_jitter=0.0)
    m = oracle_mesh(arr, bend=False)
    pos = arr["pos0"].astype(np.float64)
    vel = np.zeros_like(pos)
    p = dict(mode=po.EXACT, iterations=200, gravity=-9.81, dt=1 / 60)
    x_ref, *_ = po.step(m, po.params(exact_pins=0, **p), pos, vel, "f64")
    x_cor, *_ = po.step(m, po.params(exact_pins=1, **p), pos, vel, "f64")
    assert np.abs(x_ref - x_cor).max() > 1e-3
    y = po.prediction(pos, vel, 1 / 60, -9.81)
    assert np.sqrt(po.residual_f64(m, 1 / 60, x_cor, y)[1]) < 1e-10 * 1e3


def test_fp32_and_fp64_pd_agree():
    s = ch.Scene(**CFG1)
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    pos, vel = arr["pos0"], np.zeros_like(arr["pos0"])
    p = po.params(mode=po.PD, iterations=20, rho=0.99, gravity=-9.81, dt=1 / 60)
    for _ in range(3):
        x32, v32, *_ = po.step(m, p, pos, vel, "f32")
        x64, v64, *_ = po.step(m, p, pos, vel, "f64")
        assert np.abs(x32 - x64).max() < 1e-5
        pos, vel = x64, v64


def test_setup_matches_fp64():
    s = ch.Scene(nrows=20, ncols=20, size=1 / 19, type=ch.MESH_IRREGULAR, rest_jitter=0.3, seed=6,
                 bend_k=1e-2, stretch_k=5e2, stretch_k_max=5e3, pin_mode=ch.PIN_TOP_ROW)
    arr = scene_arrays(s)
    w, n = po.bend_weights_f64(arr["rest_pos"], arr["bend_quads"])
    np.testing.assert_allclose(arr["bend_w"], w, rtol=1e-4, atol=1e-4 * np.abs(w).max())
    assert np.abs(arr["bend_n"] - n).max() < 1e-5 * np.abs(w).max()
    d = po.jacobi_diag_f64(oracle_mesh(arr), 1 / 60)
    np.testing.assert_allclose(arr["jacobi_diag"], d, rtol=1e-5)
    k = arr["stretch_k"]
    assert k.min() >= 5e2 and k.max() < 5e3 and np.unique(k).size > 100


# ---- the reference's own solver tests (test_pd_constraints.cpp) ---------------

def rows_cylinder(arr, n, R):
    pos = arr["rest_pos"].astype(np.float32).copy()
    for i in range(n):
        th = (i - 4.5) * 0.1 / R
        for j in range(n):
            idx = i * n + j
            pos[idx, 1] = R * (1 - np.cos(th))
            pos[idx, 2] = R * np.sin(th)
    return pos


def run_frames(m, p, pos, frames, each=None):
    vel = np.zeros_like(pos)
    for _ in range(frames):
        pos, vel, *_ = po.step(m, p, pos, vel, "f32")
        if each:
            each(pos, vel)
    return pos, vel


def ke(arr, vel):
    return float(0.5 * (arr["mass"] * (vel.astype(np.float64) ** 2).sum(1)).sum())


def test_reference_test1_static_rest():
    # Test 1 (:151-187): 10x10 type 3, stretch 100, bend 1, corners, g = 0
    s = ch.Scene(nrows=10, ncols=10, size=0.1, type=3, stretch_k=100.0, bend_k=1.0,
                 pin_mode=ch.PIN_CORNERS, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    p = po.params(mode=po.REF_CUDA, iterations=100, gravity=0.0, dt=0.01)
    pos, vel = run_frames(m, p, arr["rest_pos"].copy(), 100)
    assert np.linalg.norm(vel, axis=1).max() < 10 and ke(arr, vel) < 10


def test_reference_test2_isometric_bend():
    # Test 2 (:189-230): 10x10 type 0, stretch 1000, no pins, cylinder R = 1
    s = ch.Scene(nrows=10, ncols=10, size=0.1, type=0, stretch_k=1000.0, bend_k=0.0,
                 pin_mode=ch.PIN_NONE, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    m = oracle_mesh(arr, bend=False)
    p = po.params(mode=po.REF_CUDA, iterations=50, gravity=0.0, dt=0.01)
    pos, vel = run_frames(m, p, rows_cylinder(arr, 10, 1.0), 50)
    assert np.linalg.norm(vel, axis=1).max() < 20


def test_reference_test3_bend_only_scale():
    # Test 3 (:232-268): bend only, x1.5 scaled start
    s = ch.Scene(nrows=10, ncols=10, size=0.1, type=0, stretch_k=0.0, bend_k=1.0,
                 pin_mode=ch.PIN_NONE, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    m = oracle_mesh(arr, stretch=False)
    p = po.params(mode=po.REF_CUDA, iterations=50, gravity=0.0, dt=0.01)
    pos, vel = run_frames(m, p, arr["rest_pos"] * 1.5, 50)
    assert np.linalg.norm(vel, axis=1).max() < 1e-2


def bounded_run(arr, m, p, start, frames):
    rp = arr["rest_pos"]
    init_d = np.linalg.norm(start - rp, axis=1).max()
    stats = {"ke": 0.0, "disp": init_d}

    def each(pos, vel):
        stats["ke"] = max(stats["ke"], ke(arr, vel))
        stats["disp"] = max(stats["disp"], np.linalg.norm(pos - rp, axis=1).max())
    run_frames(m, p, start, frames, each)
    return init_d, stats


def test_reference_test4_stretched_bounded():
    # Test 4 (:270-313): stretch 100, dt 0.005, 200 iterations, corners
    s = ch.Scene(nrows=10, ncols=10, size=0.1, type=0, stretch_k=100.0, bend_k=0.0,
                 pin_mode=ch.PIN_CORNERS, pos_jitter=0.0, dt=0.005)
    arr = scene_arrays(s)
    m = oracle_mesh(arr, bend=False)
    p = po.params(mode=po.REF_CUDA, iterations=200, gravity=0.0, dt=0.005)
    init_d, st = bounded_run(arr, m, p, arr["rest_pos"] * 1.5, 200)
    assert st["ke"] < 5000 and st["disp"] < init_d * 3


def test_reference_test5_bent_bounded():
    # Test 5 (:315-365): stretch 10, bend 0.5, cylinder R = 0.5
    s = ch.Scene(nrows=10, ncols=10, size=0.1, type=0, stretch_k=10.0, bend_k=0.5,
                 pin_mode=ch.PIN_CORNERS, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    p = po.params(mode=po.REF_CUDA, iterations=200, gravity=0.0, dt=0.01)
    init_d, st = bounded_run(arr, m, p, rows_cylinder(arr, 10, 0.5), 200)
    assert st["ke"] < 500 and st["disp"] < init_d * 3


def test_reference_test6_cpu_path_keeps_surface():
    # Test 6 (:367-390): CPU path, 20x20 type 3, stretch 100, top row,
    # dt 0.01, g -9.8, damping 0.99, 20 frames: y and z ranges > 0.02
    s = ch.Scene(nrows=20, ncols=20, size=0.05, type=3, stretch_k=100.0, bend_k=0.0,
                 pin_mode=ch.PIN_TOP_ROW, pos_jitter=0.0, dt=0.01)
    arr = scene_arrays(s)
    m = oracle_mesh(arr, bend=False)
    p = po.params(mode=po.EXACT, iterations=50, gravity=-9.8, dt=0.01, damping=0.99)
    pos, vel = arr["rest_pos"].copy(), np.zeros_like(arr["rest_pos"])
    for _ in range(20):
        pos, vel, *_ = po.step(m, p, pos, vel, "f32")
    assert np.ptp(pos[:, 1]) > 0.02 and np.ptp(pos[:, 2]) > 0.02


def test_restart_rule_floors_rho():
    """Every third Chebyshev restart doubles 1 - rho, down to rho = 0 (plain
    Jacobi). A 12x12 cloth checked every 5 iterations from rho = 0.999999
    restarts hundreds of times; more than 60 restarts would take rho below 0,
    where the weights 2 / (2 - rho^2) pass 2 and the iterate diverges (a rule
    without the floor ends this run non-finite). With the floor it converges
    to the round-off level."""
    s = ch.Scene(**dict(CFG1, nrows=12, ncols=12, size=1 / 11))
    arr = scene_arrays(s)
    m = oracle_mesh(arr)
    pos = arr["pos0"].reshape(-1, 3).astype(np.float32)
    p = po.params(mode=po.PD, iterations=4000, gravity=-9.81, dt=1 / 60, rho=0.999999,
                  cheb_delay=1, tolerance=1e-12, check_every=5, threads=1)
    x, v, _, rs, n = po.step(m, p, pos, np.zeros_like(pos), "f32", resid=True)
    checks = rs[4::5, 0]
    restarts = int(np.count_nonzero(checks[1:] > checks[:-1]))
    assert n == 4000 and restarts > 60
    assert np.isfinite(x).all() and np.isfinite(v).all() and np.isfinite(rs).all()
    assert np.sqrt(rs[-1, 0] / rs[-1, 1]) < 1e-9
