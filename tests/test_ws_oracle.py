"""The CPU restatement of the learned warm start (oracle/ws_oracle.py), pinned
by mathematics before it checks the GPU: central finite differences of the
loss for the gradient, per-sample loops for the network's orientation, the
closed form of Adam's first step, known BF16 roundings, and invariances of
the features (translation, constant velocity)."""
import numpy as np
import pytest

import ws_oracle as wo

RNG = np.random.default_rng(5)


def random_params(scale=0.3):
    p = RNG.normal(scale=scale, size=wo.N_PARAMS)
    p[wo.O_W3 + wo.OUT * wo.HID:wo.O_B3] = 0.0
    p[wo.O_B3 + wo.OUT:] = 0.0
    return p


def test_bf16_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 0.0], np.float32)
    # 1 + 2^-8 is a tie between 0x3F80 and 0x3F81 -> even 0x3F80;
    # 1 + 3 * 2^-8 is a tie between 0x3F81 and 0x3F82 -> even 0x3F82
    assert list(wo.bf16_bits(vals)) == [0x3F80, 0x3F80, 0x3F82, 0x3F80, 0xC020, 0x0000]
    x = RNG.normal(size=1000).astype(np.float32)
    err = np.abs(wo.bf16_value(wo.bf16_bits(x)) - x) / np.abs(x)
    assert err.max() <= 2 ** -8  # half an ulp of an 8-bit significand


def test_forward_orientation_matches_per_sample_loops():
    p = random_params()
    F = RNG.normal(size=(3, 16))
    W1, b1, W2, b2, W3, b3 = wo.unpack(p)
    o = wo.forward(p, F, emulate=False)
    silu = lambda z: z / (1 + np.exp(-z))
    for s in range(3):
        h1 = [silu(sum(W1[j, i] * F[s, i] for i in range(16)) + b1[j]) for j in range(64)]
        h2 = [silu(sum(W2[j, i] * h1[i] for i in range(64)) + b2[j]) for j in range(64)]
        out = [sum(W3[k, i] * h2[i] for i in range(64)) + b3[k] for k in range(3)]
        np.testing.assert_allclose(o[s], out, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences():
    p = random_params()
    F = RNG.normal(size=(32, 16))
    T = np.concatenate([RNG.normal(size=(32, 3)), (RNG.uniform(size=(32, 1)) > 0.2)], axis=1)
    _, g = wo.loss_grad(p, F, T, emulate=False)
    idx = np.concatenate([RNG.choice(wo.N_PARAMS, 300, replace=False),
                          np.arange(wo.O_B1, wo.O_W2), np.arange(wo.O_B3, wo.O_B3 + 3)])
    h = 1e-6
    for i in idx:
        if (wo.O_W3 + 3 * 64 <= i < wo.O_B3) or i >= wo.O_B3 + 3:
            assert g[i] == 0.0
            continue
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (wo.loss_grad(pp, F, T, emulate=False)[0] - wo.loss_grad(pm, F, T, emulate=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_emulated_gradient_close_to_exact():
    p = random_params()
    F = wo.bf16(RNG.normal(size=(64, 16)))
    T = np.concatenate([RNG.normal(size=(64, 3)), np.ones((64, 1))], axis=1)
    _, ge = wo.loss_grad(p, F, T, emulate=True)
    _, gx = wo.loss_grad(p, F, T, emulate=False)
    assert np.linalg.norm(ge - gx) / np.linalg.norm(gx) < 2e-2


def test_adam_first_step_closed_form():
    p = random_params()
    g = RNG.normal(size=wo.N_PARAMS)
    z = np.zeros(wo.N_PARAMS)
    p1, m1, v1 = wo.adam(p, z, z, g, 1, lr=1e-3)
    free = np.ones(wo.N_PARAMS, bool)
    free[wo.O_W3 + 3 * 64:wo.O_B3] = False
    free[wo.O_B3 + 3:] = False
    # t = 1: m^ = g, v^ = g^2, so the step is lr g / (|g| + eps)
    np.testing.assert_allclose((p - p1)[free], 1e-3 * g[free] / (np.abs(g[free]) + 1e-8), rtol=1e-10)
    assert np.array_equal(p1[~free], p[~free])


def grid(n):
    xs, zs = np.meshgrid(np.arange(n) / (n - 1), np.arange(n) / (n - 1))
    pos = np.stack([xs.ravel(), np.zeros(n * n), zs.ravel()], 1)
    tris = []
    for i in range(n - 1):
        for j in range(n - 1):
            v0, v1, v2, v3 = i * n + j, i * n + j + 1, (i + 1) * n + j, (i + 1) * n + j + 1
            tris += [(v0, v1, v3), (v0, v3, v2)]
    return pos.astype(np.float32), np.array(tris, np.int32)


def test_features_invariances():
    pos, tris = grid(6)
    N = len(pos)
    nb = wo.neighbours(N, tris)
    assert all(len(l) >= 2 for l in nb) and nb[0] == [1, 6, 7]
    dt = 1 / 60
    pinned = np.zeros(N, bool)
    pinned[0] = True
    xm2 = pos + RNG.normal(scale=1e-3, size=pos.shape).astype(np.float32)
    xm1 = xm2 + RNG.normal(scale=1e-3, size=pos.shape).astype(np.float32)
    x = xm1 + RNG.normal(scale=1e-3, size=pos.shape).astype(np.float32)
    f = wo.features(xm2, xm1, x, dt, (0.0, -9.81, 5.0), pinned, nb)
    # translation of every frame by the same vector leaves the features alone
    c = np.float32([0.25, -0.5, 0.125])
    f2 = wo.features(xm2 + c, xm1 + c, x + c, dt, (0.0, -9.81, 5.0), pinned, nb)
    np.testing.assert_allclose(f2, f, atol=2e-3)
    np.testing.assert_allclose(f[:, 6:9], np.tile([0.0, -1.0, 5.0 / 9.81], (N, 1)), rtol=1e-6)
    assert f[0, 9] == 1.0 and np.all(f[1:, 9] == 0.0)
    # constant velocity of the whole cloth: zero acceleration, zero 1-ring terms
    vel = np.float32([1e-3, -2e-3, 5e-4])
    g = wo.features(pos, pos + vel, pos + 2 * vel, dt, (0.0, 0.0, 0.0), pinned, nb)
    L = dt * dt * 9.81
    np.testing.assert_allclose(g[:, 0:3], np.tile(vel / L, (N, 1)), rtol=1e-3)
    assert np.abs(g[:, 3:6]).max() < 1e-3 and np.abs(g[:, 10:16]).max() < 1e-3


def test_targets_of_the_extrapolation_are_zero():
    pos, _ = grid(4)
    xm2 = pos
    xm1 = pos + 1e-3
    x = xm1 + 2e-3
    xnext = wo.extrapolation(xm2, xm1, x)  # constant acceleration continued
    f32 = np.float32
    inv_L = f32(1) / (f32(1 / 60) * f32(1 / 60) * f32(9.81))
    ac = ((x - xm1) - (xm1 - xm2)) * inv_L
    t = wo.targets(xm1, x, xnext, 1 / 60, ac.astype(f32), np.zeros(len(pos), bool), 1.0)
    assert np.abs(t[:, :3]).max() < 1e-2 and np.all(t[:, 3] == 1.0)
