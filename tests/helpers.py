"""Shared helpers of the test suite: building scenes through the C++ API and
running the same frame on the HIP solver (C ABI) and on the oracle."""
import numpy as np

import clothhip as ch

ARRAYS = ["pos0", "rest_pos", "tris", "Dm_inv", "rest_area", "mass", "stretch_k", "jacobi_diag",
          "bend_quads", "bend_w", "bend_n", "bend_k", "pin_idx", "pin_target"]

# The five workloads (reference axes: the cloth lies in x-z, gravity on y;
# a z-up "wind along +y" is reference +z). Wind 0.5 N/m^2 over 0.1 kg/m^2
# is an acceleration of 5 m/s^2.
CFG1 = dict(nrows=32, ncols=32, size=1 / 31, type=0, seed=1, density=0.1, stretch_k=1e3,
            bend_k=1e-2, pin_mode=ch.PIN_CORNERS, pos_jitter=1e-4, dt=1 / 60)
CFG2 = dict(nrows=256, ncols=256, size=1 / 255, type=0, seed=2, density=0.1, stretch_k=1e3,
            bend_k=1e-2, pin_mode=ch.PIN_TOP_ROW, pos_jitter=1e-3, dt=1 / 60)
CFG3 = dict(nrows=1024, ncols=1024, size=1 / 1023, type=ch.MESH_IRREGULAR, rest_jitter=0.3,
            seed=3, density=0.1, stretch_k=5e2, stretch_k_max=5e3, bend_k=1e-2,
            pin_mode=ch.PIN_TOP_ROW, pos_jitter=0.0, dt=1 / 60)
CFG2_WIND = (0.0, 0.0, 5.0)


def cfg4_specs(count=512, seed=4, side_max=256, dt=1 / 60):
    """The batched workload: cloth sides round(16 * 16^U) (capped at
    side_max), stretch log-uniform in [1e2, 1e4], bend log-uniform in
    [1e-3, 1e-1], 1-4 random boundary pins, each cloth 1 m wide, +-1e-3 m
    jitter, seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(min(side_max, round(16 * 16 ** rng.uniform())))
        out.append(dict(nrows=n, ncols=n, size=1 / (n - 1), type=0, seed=1000 + i, density=0.1,
                        stretch_k=float(np.exp(rng.uniform(np.log(1e2), np.log(1e4)))),
                        bend_k=float(np.exp(rng.uniform(np.log(1e-3), np.log(1e-1)))),
                        pin_mode=ch.PIN_RANDOM_BOUNDARY, pin_count=int(rng.integers(1, 5)),
                        pos_jitter=1e-3, dt=dt))
    return out


def scene_arrays(scene):
    out = {}
    for k in ARRAYS:
        out[k] = scene.get(k)
    return out


def oracle_mesh(arrays, stretch=True, bend=True):
    # imported here: the workload definitions above are also read by the
    # benchmark's timed legs, which do not touch the oracle
    import pyoracle as po
    return po.Mesh(arrays, stretch=stretch, bend=bend)


def rel_err(a, b, scale=None):
    """max |a - b| over the scale (default: max |b|, at least 1)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    s = scale if scale is not None else max(np.abs(b).max(), 1.0)
    return float(np.abs(a - b).max() / s)


def extent(pos):
    pos = np.asarray(pos).reshape(-1, 3)
    return float(np.max(pos.max(0) - pos.min(0)))


def cfg5_specs(count=64, seed=5, n=256, cloth_seed=5000, wind_max=2.0):
    """The learned-warm-start workload: count n x n cloths with the 256x256
    cloth's material and pins (stretch 1e3, bend 1e-2, density 0.1, top edge
    pinned, +-1e-3 m jitter, cloth i jittered with seed cloth_seed + i), each
    with a wind of U(0, wind_max) N/m^2 (default 2, an acceleration of up to
    20 m/s^2) in a uniformly random direction. Returns the specs and the
    per-cloth accelerations [count x 3]."""
    rng = np.random.default_rng(seed)
    specs, accel = [], []
    for i in range(count):
        specs.append(dict(CFG2, nrows=n, ncols=n, size=1 / (n - 1), seed=cloth_seed + i))
        d = rng.normal(size=3)
        accel.append(d / np.linalg.norm(d) * rng.uniform(0.0, wind_max) / CFG2["density"])
    return specs, np.asarray(accel, np.float32)


def jacobi_bound_np(arrays, dt, stretch=True, bend=True, signed=False):
    """Per-element Gershgorin bound of D^-1 K (K = M + h^2 (sum_t k A G^T G +
    sum_e k w w^T), D the Jacobi diagonal the solver divides by): max over
    vertices of (m_v + sum over its elements of the absolute entries of the
    element block's row v) / D_v. signed=True sums the entries with their
    signs instead (a plausible mistake, for the tests)."""
    f = (lambda z: z) if signed else np.abs
    N = len(arrays["mass"])
    h2 = float(np.float32(dt)) ** 2  # dt is an FP32 setting
    row = np.abs(arrays["mass"].astype(np.float64))
    if stretch and len(arrays["tris"]):
        t = arrays["tris"].reshape(-1, 3)
        G = arrays["Dm_inv"].reshape(-1, 4).astype(np.float64)
        g00, g10, g01, g11 = G[:, 0], G[:, 1], G[:, 2], G[:, 3]
        D0 = np.stack([-g00 - g10, g00, g10], 1)
        D1 = np.stack([-g01 - g11, g01, g11], 1)
        w = np.abs(h2 * arrays["stretch_k"].astype(np.float64) * arrays["rest_area"].astype(np.float64))
        for c in range(3):
            acc = sum(f(D0[:, c] * D0[:, b] + D1[:, c] * D1[:, b]) for b in range(3))
            np.add.at(row, t[:, c], w * acc)
    if bend and len(arrays["bend_quads"]):
        q = arrays["bend_quads"].reshape(-1, 4)
        bw = arrays["bend_w"].reshape(-1, 4).astype(np.float64)
        kb = np.abs(h2 * arrays["bend_k"].astype(np.float64))
        sw = f(bw).sum(1) if not signed else bw.sum(1)
        for i in range(4):
            np.add.at(row, q[:, i], kb * f(bw[:, i]) * sw)
    d = arrays["jacobi_diag"].astype(np.float64)
    ok = d > 1e-10
    return float((row[ok] / d[ok]).max())
