"""The learned warm start on the GPU (clh_ws_* through the C ABI) against its
CPU restatement (oracle/ws_oracle.py): features bit for bit in BF16 (one ulp
where FP32 sums differ in order), the MFMA forward pass, the training loss
and gradient, the Adam step; then training on solver-generated frames and
the iterations it saves at the same residual tolerance."""
import numpy as np
import pytest

import clothhip as ch
import ws_oracle as wo
from helpers import CFG1, cfg5_specs
from warmstart import FrameLoop, select_gain

pytestmark = pytest.mark.gpu

DT = 1 / 60
G = -9.81
WIND = (0.0, 0.0, 5.0)
RNG = np.random.default_rng(11)


def scene_and_history(spec=CFG1, scale=2e-3):
    s = ch.Scene(**spec)
    s.upload()
    pos, _ = s.get_state()
    xm2 = pos + RNG.normal(scale=scale, size=pos.shape).astype(np.float32)
    xm1 = xm2 + RNG.normal(scale=scale, size=pos.shape).astype(np.float32)
    x = xm1 + RNG.normal(scale=scale, size=pos.shape).astype(np.float32)
    return s, [xm2, xm1, x]


def pinned_mask(s):
    m = np.zeros(s.N, bool)
    m[s.get("pin_idx")] = True
    return m


def random_params(scale=0.3):
    p = RNG.normal(scale=scale, size=wo.N_PARAMS).astype(np.float32)
    p[wo.O_W3 + 3 * 64:wo.O_B3] = 0.0
    p[wo.O_B3 + 3:] = 0.0
    return p


def oracle_features(s, hist):
    nb = wo.neighbours(s.N, s.get("tris"))
    acc = np.float32([WIND[0], G + WIND[1], WIND[2]])
    return wo.features(*hist, DT, acc, pinned_mask(s), nb)


@pytest.mark.parametrize("ext", [True, False])
def test_features_match_oracle_bf16(ext):
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND, ext_feature=ext)
    dev = [ch.DeviceArray.from_host(h) for h in hist]
    got = ws.features(*dev).astype(np.int32)
    want_f = oracle_features(s, hist)
    if not ext:
        want_f[:, 6:9] = 0.0
    want = wo.bf16_bits(want_f).astype(np.int32)
    diff = np.abs(got - want)
    assert diff.max() <= 1, diff.max()
    assert np.mean(diff == 0) > 0.99


def test_forward_matches_oracle():
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND, ext_feature=True)
    ws.set_smoothing(0)
    p = random_params()
    so = 0.7
    ws.set_params(p, so)
    dev = [ch.DeviceArray.from_host(h) for h in hist]
    guess = ch.DeviceArray(s.N * 12)
    ws.predict(*dev, guess)
    g = guess.download(np.float32, (s.N, 3)).astype(np.float64)
    F = wo.bf16_value(ws.features(*dev)).astype(np.float64)
    o = wo.forward(p, F, emulate=True)
    L = DT * DT * wo.G0
    ext = wo.extrapolation(*hist).astype(np.float64)
    free = ~pinned_mask(s)
    o_gpu = (g - ext) / (L * so)
    err = np.abs(o_gpu[free] - o[free])
    assert err.max() < 5e-3 * max(1.0, np.abs(o).max()), err.max()
    assert np.median(err) < 1e-4 * max(1.0, np.abs(o).max())
    np.testing.assert_array_equal(g[~free], hist[2][~free])
    # the network off: the extrapolation itself
    ws.predict(*dev, guess, use_network=False)
    e = guess.download(np.float32, (s.N, 3))
    np.testing.assert_array_equal(e[free], wo.extrapolation(*hist)[free])


def test_smoothing_passes_match_oracle():
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND)
    p = random_params()
    ws.set_params(p, 0.5)
    dev = [ch.DeviceArray.from_host(h) for h in hist]
    raw, sm = ch.DeviceArray(s.N * 12), ch.DeviceArray(s.N * 12)
    ws.set_smoothing(0)
    ws.predict(*dev, raw)
    ws.set_smoothing(3)
    ws.predict(*dev, sm)
    ext = wo.extrapolation(*hist)
    pin = pinned_mask(s)
    corr = raw.download(np.float32, (s.N, 3)) - ext
    want = ext + wo.smooth(corr, wo.neighbours(s.N, s.get("tris")), pin, 3)
    got = sm.download(np.float32, (s.N, 3))
    np.testing.assert_allclose(got[~pin], want[~pin], rtol=0, atol=2e-6)


def test_step_filter_matches_oracle():
    """x_n + gain * S^m(x_hat - x_n): gain 0 is the plain start bit for bit,
    (0 passes, gain 1) leaves the prediction unchanged, and 3 passes at gain
    0.5 equal the oracle's smoothing of the step."""
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND)
    ws.set_params(random_params(), 0.5)
    dev = [ch.DeviceArray.from_host(h) for h in hist]
    raw, out = ch.DeviceArray(s.N * 12), ch.DeviceArray(s.N * 12)
    ws.predict(*dev, raw)
    xhat = raw.download(np.float32, (s.N, 3))
    x = hist[2]
    pin = pinned_mask(s)
    ws.set_step_filter(0, 0.0)
    ws.predict(*dev, out)
    np.testing.assert_array_equal(out.download(np.float32, (s.N, 3)), x)
    ws.set_step_filter(0, 1.0)
    ws.predict(*dev, out)
    np.testing.assert_array_equal(out.download(np.float32, (s.N, 3)), xhat)
    ws.set_step_filter(3, 0.5)
    ws.predict(*dev, out)
    step = np.where(pin[:, None], 0.0, xhat - x).astype(np.float32)
    sm = wo.smooth(step, wo.neighbours(s.N, s.get("tris")), pin, 3)
    want = x + np.float32(0.5) * sm
    got = out.download(np.float32, (s.N, 3))
    np.testing.assert_allclose(got[~pin], want[~pin], rtol=0, atol=2e-7)
    np.testing.assert_array_equal(got[pin], x[pin])
    with pytest.raises(ch.ClothHipError):
        ws.set_step_filter(-1, 1.0)


def test_untrained_network_is_the_extrapolation():
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND)
    dev = [ch.DeviceArray.from_host(h) for h in hist]
    a, b = ch.DeviceArray(s.N * 12), ch.DeviceArray(s.N * 12)
    ws.predict(*dev, a)
    ws.predict(*dev, b, use_network=False)
    np.testing.assert_array_equal(a.download(np.float32, (s.N, 3)), b.download(np.float32, (s.N, 3)))


def test_dataset_and_gradient_match_oracle():
    s, hist = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND, ext_feature=True)
    xnext = hist[2] + RNG.normal(scale=2e-3, size=hist[2].shape).astype(np.float32)
    dev = [ch.DeviceArray.from_host(h) for h in hist + [xnext]]
    ws.dataset_add(*dev)
    n = ws.dataset_size()
    assert n == -(-s.N // 16) * 16
    F, T = ws.dataset_read(0, n)
    f32 = np.float32
    inv_L = f32(1) / (f32(DT) * f32(DT) * f32(9.81))
    ac = ((hist[2] - hist[1]) - (hist[1] - hist[0])) * inv_L
    t_want = wo.targets(hist[1], hist[2], xnext, DT, ac.astype(f32), pinned_mask(s), 1.0)
    np.testing.assert_allclose(T[:s.N], t_want, rtol=1e-5, atol=1e-4)
    assert np.all(T[s.N:] == 0) and np.all(F[s.N:] == 0)
    so = ws.fit_output_scale()
    F, T = ws.dataset_read(0, n)
    w = T[:, 3:4]
    assert abs(np.sum(w * T[:, :3] ** 2) / np.sum(3 * w) - 1.0) < 1e-4
    p = random_params()
    ws.set_params(p, so)
    for first, count in [(0, 256), (0, n), (512, 128)]:
        loss, g = ws.loss_grad(first, count)
        Fv = wo.bf16_value(F[first:first + count]).astype(np.float64)
        l_ref, g_ref = wo.loss_grad(p, Fv, T[first:first + count], emulate=True)
        assert abs(loss - l_ref) <= 1e-4 * l_ref, (loss, l_ref)
        for lo, hi in [(wo.O_W1, wo.O_B1), (wo.O_B1, wo.O_W2), (wo.O_W2, wo.O_B2),
                       (wo.O_B2, wo.O_W3), (wo.O_W3, wo.O_B3), (wo.O_B3, wo.N_PARAMS)]:
            rel = np.linalg.norm(g[lo:hi] - g_ref[lo:hi]) / max(np.linalg.norm(g_ref[lo:hi]), 1e-30)
            assert rel < 2e-3, (first, count, lo, rel)


def test_adam_step_matches_oracle():
    s, _ = scene_and_history()
    ws = ch.WarmStart.for_scene(s, DT, G, WIND)
    p = random_params()
    ws.set_params(p, 1.0)
    g1 = RNG.normal(size=wo.N_PARAMS).astype(np.float32)
    g2 = RNG.normal(size=wo.N_PARAMS).astype(np.float32)
    ws.adam_step(g1, 1e-2)
    ws.adam_step(g2, 1e-2)
    got, _ = ws.params()
    z = np.zeros(wo.N_PARAMS)
    p1, m1, v1 = wo.adam(p, z, z, g1, 1, 1e-2)
    p2, _, _ = wo.adam(p1, m1, v1, g2, 2, 1e-2)
    np.testing.assert_allclose(got, p2, rtol=1e-5, atol=1e-6)


def solver_cfg(tol=1e-5, cap=20000):
    return ch.config(mode=ch.MODE_PD, max_iterations=cap, tolerance=tol, check_every=10,
                     rho=0.999, cheb_delay=1, gravity=G, dt=DT)


def test_training_saves_iterations():
    """Train on frames the batched solver generates for 32 cloths of 24x24
    with random wind (cloth-frames that miss the tolerance are left out),
    then run held-out cloths (other winds) from rest in the three modes to
    the same tolerance: the learned start needs at least 1.5x fewer
    iterations than the plain start x_n over frames 2..29."""
    specs, accel = cfg5_specs(32, seed=100, n=24)
    b = ch.Scene(specs=specs)
    b.upload()
    sol = ch.Solver(solver_cfg(cap=4000), b.N, b.T, b.E, cloth_offsets=b.vert_off)
    ws = ch.WarmStart.for_scene(b, DT, G, seed=3)
    loop = FrameLoop(b, sol, ws, accel)
    loop.run(30, mode="extrap", dataset=True)
    assert loop.kept > 0.2 * loop.offered, (loop.kept, loop.offered)
    so = ws.fit_output_scale()
    assert so > 0
    losses = ws.train(3000, 8192, 2e-3)
    assert np.mean(losses[-100:]) < 0.5 * np.mean(losses[:20]), (losses[:5], losses[-5:])

    hspecs, haccel = cfg5_specs(2, seed=900, n=24)
    haccel = np.concatenate([haccel, np.float32([[0.0, 0.0, 5.0]])])
    res = {m: [] for m in ("plain", "extrap", "learned")}
    for k, wind in enumerate(haccel):
        for mode in res:
            s = ch.Scene(**dict(hspecs[min(k, 1)], seed=4242 + k))
            s.upload()
            sol1 = ch.Solver(solver_cfg(), s.N, s.T, s.E)
            ws1 = ch.WarmStart.for_scene(s, DT, G)
            ws1.set_params(*ws.params())
            its, _ = FrameLoop(s, sol1, ws1, wind.reshape(1, 3)).run(30, mode=mode)
            res[mode].append(float(np.mean(its[2:])))
    print("mean iterations per frame:", res)
    assert np.sum(res["plain"]) / np.sum(res["learned"]) >= 1.5, res


def test_select_gain_on_held_out_frames():
    """The gain on the network's correction is chosen by the iterations on
    held-out frames: the choice is the argmin of the measured table, the warm
    start keeps its output scale times the gain, the scene's state is
    restored, and gain 0 runs the extrapolation's frames exactly."""
    specs, accel = cfg5_specs(8, seed=101, n=16)
    b = ch.Scene(specs=specs)
    b.upload()
    sol = ch.Solver(solver_cfg(cap=4000), b.N, b.T, b.E, cloth_offsets=b.vert_off)
    ws = ch.WarmStart.for_scene(b, DT, G, seed=3)
    FrameLoop(b, sol, ws, accel).run(12, mode="extrap", dataset=True)
    ws.fit_output_scale()
    ws.train(300, 4096, 2e-3)
    params, scale = ws.params()

    vspecs, vaccel = cfg5_specs(4, seed=901, n=16, cloth_seed=7000)
    v = ch.Scene(specs=vspecs)
    v.upload()
    vsol = ch.Solver(solver_cfg(cap=4000), v.N, v.T, v.E, cloth_offsets=v.vert_off)
    vws = ch.WarmStart.for_scene(v, DT, G)
    vws.set_params(params, scale)
    pos0, vel0 = v.get_state()
    gain, table = select_gain(v, vsol, vws, gains=(0.0, 1.0), warmup=2, frames=4,
                              cloth_accel=vaccel)
    assert set(table) == {0.0, 1.0, "plain"}
    assert table["plain" if gain is None else gain] == min(table.values())
    np.testing.assert_array_equal(vws.params()[0], params)
    assert vws.params()[1] == np.float32(scale * (1.0 if gain is None else gain))
    np.testing.assert_array_equal(v.get_state()[0], pos0)
    # gain 0: the learned frames are the extrapolation's
    vws.set_params(params, 0.0)
    its_g0 = FrameLoop(v, vsol, vws, vaccel)
    its_g0.run(2, mode="plain")
    a, _ = its_g0.run(4, mode="learned")
    v.set_state(pos0, vel0)
    ex = FrameLoop(v, vsol, vws, vaccel)
    ex.run(2, mode="plain")
    e, _ = ex.run(4, mode="extrap")
    assert a == e
    assert table[0.0] == float(np.mean(e))


def test_fallback_frames_equal_plain_frames():
    """A warm solver whose cap no frame can meet: every warm-started frame is
    redone from (x_n, v_n) by the plain fallback solver, so the trajectory is
    the plain one bit for bit; repeated misses pause the warm start."""
    spec = dict(CFG1, nrows=24, ncols=24, size=1 / 23, pin_mode=ch.PIN_TOP_ROW)
    wind = np.float32([[0.0, 0.0, 5.0]])
    a = ch.Scene(**spec)
    a.upload()
    warm = ch.Solver(solver_cfg(cap=5), a.N, a.T, a.E)
    fb = ch.Solver(solver_cfg(), a.N, a.T, a.E)
    ws = ch.WarmStart.for_scene(a, DT, G)
    loop = FrameLoop(a, warm, ws, wind, fallback=fb, strikes=3, cooldown=2)
    its, _ = loop.run(8, mode="learned")
    assert loop.fallbacks == 6 and loop.paused_frames == 2, (loop.fallbacks, loop.paused_frames)
    b = ch.Scene(**spec)
    b.upload()
    plain = ch.Solver(solver_cfg(), b.N, b.T, b.E)
    its_p, _ = FrameLoop(b, plain, None, wind).run(8, mode="plain")
    pa, va = a.get_state()
    pb, vb = b.get_state()
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(va, vb)


def test_start_residual_is_the_frames_first_check():
    """clh_solver_start_residual of a start equals the residual the frame
    records at its first check (iteration 0 with check_every 1), and leaves
    the mesh state alone."""
    spec = dict(CFG1, nrows=24, ncols=24, size=1 / 23, pin_mode=ch.PIN_TOP_ROW)
    s = ch.Scene(**spec)
    s.upload()
    pos0, vel0 = s.get_state()
    cfg = ch.config(mode=ch.MODE_PD, max_iterations=3, tolerance=1e-30, check_every=1,
                    rho=0.999, cheb_delay=1, gravity=G, dt=DT)
    sol = ch.Solver(cfg, s.N, s.T, s.E)
    args = s.step_args()
    guess = ch.DeviceArray.from_host(pos0 + np.float32(1e-3))
    for g in (None, guess.ptr):
        args.d_init_guess = g
        r = sol.start_residual(args)
        p1, v1 = s.get_state()
        np.testing.assert_array_equal(p1, pos0)
        np.testing.assert_array_equal(v1, vel0)
        sol.step(args)
        h = sol.residual_history(8)
        assert h[0] == r, (h[0], r)
        s.set_state(pos0, vel0)


def test_select_gain_keeps_the_plain_start_when_it_wins():
    """A correction scaled far beyond the trained one (gain 40) cannot beat
    the plain start on held-out frames: the choice is the plain start (gain
    None), the warm start's parameters are left as they were, and a frame
    loop without a warm start runs the plain frames exactly."""
    specs, accel = cfg5_specs(8, seed=101, n=16)
    b = ch.Scene(specs=specs)
    b.upload()
    sol = ch.Solver(solver_cfg(cap=4000), b.N, b.T, b.E, cloth_offsets=b.vert_off)
    ws = ch.WarmStart.for_scene(b, DT, G, seed=3)
    FrameLoop(b, sol, ws, accel).run(12, mode="extrap", dataset=True)
    ws.fit_output_scale()
    ws.train(300, 4096, 2e-3)
    params, scale = ws.params()

    vspecs, vaccel = cfg5_specs(4, seed=901, n=16, cloth_seed=7000)
    v = ch.Scene(specs=vspecs)
    v.upload()
    vsol = ch.Solver(solver_cfg(cap=4000), v.N, v.T, v.E, cloth_offsets=v.vert_off)
    vws = ch.WarmStart.for_scene(v, DT, G)
    vws.set_params(params, scale)
    pos0, vel0 = v.get_state()
    gain, table = select_gain(v, vsol, vws, gains=(40.0,), warmup=2, frames=4,
                              cloth_accel=vaccel)
    assert table["plain"] < table[40.0]
    assert gain is None
    np.testing.assert_array_equal(vws.params()[0], params)
    assert vws.params()[1] == np.float32(scale)
    np.testing.assert_array_equal(v.get_state()[0], pos0)
    # no warm start: every mode runs the plain frames
    a, _ = FrameLoop(v, vsol, None, vaccel).run(4, mode="learned")
    v.set_state(pos0, vel0)
    p, _ = FrameLoop(v, vsol, None, vaccel).run(4, mode="plain")
    assert a == p
    assert table["plain"] > 0
