"""Mathematical pins of the setup constants the solver consumes: the
cotangent bend weights and rest curvature norms (src/sim_constraints.cpp:48-127),
the Jacobi diagonal (src/sim_constraints.cpp:133-192) and the warm start's
1-ring smoothing. Each check comes from a property the formula must have,
computed by an independent route, not from a restatement of the formula:

- cotangents from vectors (u.v / |u x v|) instead of side lengths and Heron
  areas; closed forms on the type-0 grid;
- a flat hinge annihilates its rest positions (sum w_i X_i = 0), and a hinge
  folded by phi about its shared edge of length L has |sum w_i x_i| =
  2 L sin(phi / 2) whatever the wing triangles;
- the diagonal of K = M + h^2 Hess(E) from the quadratic energy
  E(x) = 1/2 sum k A |Ds(x) Dm^-1|^2 + 1/2 sum k |sum w_i x_i|^2, assembled
  densely by polarization, E(e_i + e_j) - E(e_i) - E(e_j);
- smoothing preserves constants, keeps pins at zero and equals a brute-force
  1-ring mean built from the triangles' edges.

Every group also checks that a plausible mistake (paired cotangents
swapped, h^2 or the bend term dropped, a vertex left out of its own mean)
fails the same assertion."""
import numpy as np
import pytest

import clothhip as ch
import pyoracle as po
import ws_oracle as wo
from helpers import oracle_mesh, scene_arrays


def cot_at(p, q, r):
    """cot of the angle at p in triangle (p, q, r), from vectors."""
    u, v = q - p, r - p
    return np.dot(u, v) / np.linalg.norm(np.cross(u, v))


def weights_from_vectors(X):
    """Hinge weights of quad (v0, v1, v2, v3), shared edge (v0, v1), from the
    angles at the shared edge's endpoints: v0 takes the cotangents at v1,
    v1 those at v0 (the pairing that annihilates a flat hinge)."""
    a1_0, a0_0 = cot_at(X[1], X[0], X[2]), cot_at(X[0], X[1], X[2])
    a1_1, a0_1 = cot_at(X[1], X[0], X[3]), cot_at(X[0], X[1], X[3])
    return np.array([a1_0 + a1_1, a0_0 + a0_1, -(a1_0 + a0_0), -(a1_1 + a0_1)])


def weights_swapped(X):
    """The plausible mistake: each endpoint takes its own angles' cotangents."""
    a1_0, a0_0 = cot_at(X[1], X[0], X[2]), cot_at(X[0], X[1], X[2])
    a1_1, a0_1 = cot_at(X[1], X[0], X[3]), cot_at(X[0], X[1], X[3])
    return np.array([a0_0 + a0_1, a1_0 + a1_1, -(a1_0 + a0_0), -(a1_1 + a0_1)])


def irregular(n=12, seed=6, **kw):
    spec = dict(nrows=n, ncols=n, size=1 / (n - 1), type=ch.MESH_IRREGULAR, rest_jitter=0.3,
                seed=seed, bend_k=1e-2, stretch_k=5e2, stretch_k_max=5e3,
                pin_mode=ch.PIN_TOP_ROW, dt=1 / 60)
    spec.update(kw)
    s = ch.Scene(**spec)
    return s, scene_arrays(s)


# ---- cotangent bend weights ---------------------------------------------------

def test_type0_grid_weights_closed_form():
    # "\" diagonals: a cell-diagonal hinge has two right isosceles wings with
    # 45 degrees at both endpoints (all four cotangents 1): w = (2, 2, -2, -2).
    # An axis-edge hinge has the right angle at one endpoint in one wing and
    # at the other endpoint in the other: w = (1, 1, -1, -1).
    for n in (4, 9):
        s = ch.Scene(nrows=n, ncols=n, size=1 / (n - 1), type=0, bend_k=1e-2, pin_mode=ch.PIN_NONE)
        arr = scene_arrays(s)
        X = arr["rest_pos"].astype(np.float64)
        q = arr["bend_quads"]
        shared = np.linalg.norm(X[q[:, 1]] - X[q[:, 0]], axis=1) * (n - 1)
        diag = np.isclose(shared, np.sqrt(2))
        axis = np.isclose(shared, 1.0)
        assert diag.sum() == (n - 1) ** 2 and axis.sum() == 2 * (n - 1) * (n - 2)
        wo64, _ = po.bend_weights_f64(arr["rest_pos"], q)
        for w in (arr["bend_w"].astype(np.float64), wo64):
            np.testing.assert_allclose(w[diag], np.tile([2, 2, -2, -2], (diag.sum(), 1)), atol=2e-5)
            np.testing.assert_allclose(w[axis], np.tile([1, 1, -1, -1], (axis.sum(), 1)), atol=2e-5)


def test_weights_equal_vector_cotangents_on_every_mesh_type():
    meshes = [ch.Scene(nrows=6, ncols=7, size=0.2, type=t, bend_k=1e-2, pin_mode=ch.PIN_NONE)
              for t in (0, 1, 2, 3)]
    meshes.append(irregular()[0])
    for s in meshes:
        arr = scene_arrays(s)
        X = arr["rest_pos"].astype(np.float64)
        q = arr["bend_quads"]
        want = np.array([weights_from_vectors(X[qq]) for qq in q])
        scale = np.abs(want).max()
        np.testing.assert_allclose(arr["bend_w"], want, atol=2e-5 * scale)
        wo64, _ = po.bend_weights_f64(arr["rest_pos"], q)
        np.testing.assert_allclose(wo64, want, atol=1e-9 * scale)


def test_flat_irregular_hinges_annihilate_rest_positions():
    s, arr = irregular()
    X = arr["rest_pos"].astype(np.float64)
    q = arr["bend_quads"]
    wo64, n64 = po.bend_weights_f64(arr["rest_pos"], q)
    for w in (arr["bend_w"].astype(np.float64), wo64):
        e = np.einsum("ei,eid->ed", w, X[q])
        bound = np.einsum("ei,ei->e", np.abs(w), np.linalg.norm(X[q], axis=2))
        assert np.all(np.linalg.norm(e, axis=1) <= 1e-6 * bound)
        np.testing.assert_allclose(w.sum(1), 0, atol=1e-6 * np.abs(w).max())
    # n_rest is the FP32 norm of the same sum: rounding noise
    assert np.abs(arr["bend_n"]).max() < 1e-6 * np.abs(wo64).max()
    assert np.abs(n64).max() < 1e-12 * np.abs(wo64).max()
    # the mistake: swapped pairs leave a curvature on a flat irregular mesh
    bad = np.array([weights_swapped(X[qq]) for qq in q])
    e_bad = np.linalg.norm(np.einsum("ei,eid->ed", bad, X[q]), axis=1)
    bound_bad = np.einsum("ei,ei->e", np.abs(bad), np.linalg.norm(X[q], axis=2))
    assert np.any(e_bad > 1e-3 * bound_bad)


def folded_hinge(rng, phi):
    """Shared edge (v0, v1) on the x axis, wing v2 in the x-y plane at
    positive y, wing v3 rotated by pi - phi from v2's plane about the edge
    (phi = 0 is flat, the wings on opposite sides)."""
    L = rng.uniform(0.5, 2.0)
    X0 = np.array([0.0, 0.0, 0.0])
    X1 = np.array([L, 0.0, 0.0])
    X2 = np.array([rng.uniform(-0.5, 1.5) * L, rng.uniform(0.3, 1.5) * L, 0.0])
    a, h = rng.uniform(-0.5, 1.5) * L, rng.uniform(0.3, 1.5) * L
    th = np.pi - phi
    X3 = np.array([a, h * np.cos(th), h * np.sin(th)])
    return L, np.array([X0, X1, X2, X3])


def test_folded_hinge_curvature_closed_form():
    # sum w_i x_i = -L (n0 + n1), n_j the unit in-plane normals of the shared
    # edge toward each wing: |.| = 2 L sin(phi / 2) for a fold by phi. The
    # weights are intrinsic (side lengths), so the flat layout gives them.
    rng = np.random.default_rng(11)
    quad = np.array([[0, 1, 2, 3]], np.int32)
    caught = 0
    for _ in range(50):
        phi = rng.uniform(0.0, 3.0)
        L, Xf = folded_hinge(rng, phi)
        Xflat = Xf.copy()
        # the same wing shapes unfolded into the plane
        h3 = np.hypot(Xf[3, 1], Xf[3, 2])
        Xflat[3] = [Xf[3, 0], -h3, 0.0]
        w, n = po.bend_weights_f64(Xflat.astype(np.float32), quad)
        e = w[0] @ Xf
        np.testing.assert_allclose(np.linalg.norm(e), 2 * L * np.sin(phi / 2), rtol=1e-5, atol=1e-6)
        # perpendicular to the shared edge
        assert abs(e[0]) < 1e-5 * L
        # the oracle's n_rest of the folded rest shape is the same norm
        _, nf = po.bend_weights_f64(Xf.astype(np.float32), quad)
        np.testing.assert_allclose(nf[0], 2 * L * np.sin(phi / 2), rtol=1e-5, atol=1e-6)
        bad = weights_swapped(Xflat.astype(np.float32).astype(np.float64)) @ Xf
        if abs(Xf[2, 0] - Xf[3, 0]) > 0.05 * L or abs(Xf[2, 0] - 0.5 * L) > 0.05 * L:
            assert abs(np.linalg.norm(bad) - 2 * L * np.sin(phi / 2)) > 1e-4 * L
            caught += 1
    assert caught > 25


def write_obj(path, X, tris):
    with open(path, "w") as f:
        for x in X:
            f.write(f"v {x[0]:.9g} {x[1]:.9g} {x[2]:.9g}\n")
        for t in tris:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def test_product_rest_curvature_of_a_folded_strip(tmp_path):
    # a strip of 2 x k quads folded along its middle line by phi: every hinge
    # on the fold line has n_rest = 2 L sin(phi / 2) with L the grid spacing,
    # every other hinge lies in one plane and has n_rest ~ 0
    k, s, phi = 6, 0.1, 0.7
    X, idx = [], {}
    for i in range(3):           # rows: 0 and 2 are the wings, 1 the fold line
        for j in range(k + 1):
            y = (i - 1) * s
            p = np.array([j * s, 0.0, y]) if i <= 1 else \
                np.array([j * s, s * np.sin(phi), s * np.cos(phi)])
            idx[i, j] = len(X)
            X.append(p)
    tris = []
    for i in range(2):
        for j in range(k):
            a, b, c, d = idx[i, j], idx[i, j + 1], idx[i + 1, j], idx[i + 1, j + 1]
            tris += [(a, b, d), (a, d, c)]
    path = tmp_path / "fold.obj"
    write_obj(path, np.array(X), tris)
    sc = ch.Scene(obj_path=path, ncols=k + 1, bend_k=1e-2, pin_mode=ch.PIN_NONE, dt=1 / 60)
    arr = scene_arrays(sc)
    R = arr["rest_pos"].astype(np.float64)
    q = arr["bend_quads"]
    on_fold = np.array([abs(R[a, 2] - 0.0) < 1e-6 and abs(R[b, 2] - 0.0) < 1e-6 and
                        abs(R[a, 1]) < 1e-6 and abs(R[b, 1]) < 1e-6 for a, b in q[:, :2]])
    assert on_fold.sum() == k
    np.testing.assert_allclose(arr["bend_n"][on_fold], 2 * s * np.sin(phi / 2), rtol=2e-5)
    assert np.abs(arr["bend_n"][~on_fold]).max() < 1e-5


# ---- Jacobi diagonal ------------------------------------------------------------

def quadratic_energy(arr, x, h2=None, bend=True):
    """1/2 sum k A |Ds Dm^-1|^2 + 1/2 sum k |sum w_i x_i|^2 of positions x [N x 3]."""
    t = arr["tris"]
    M = arr["Dm_inv"].astype(np.float64).reshape(-1, 2, 2).transpose(0, 2, 1)  # column-major
    Ds = np.stack([x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]]], axis=-1)
    F = Ds @ M
    kA = arr["stretch_k"].astype(np.float64) * arr["rest_area"].astype(np.float64)
    E = 0.5 * np.sum(kA * np.sum(F * F, axis=(1, 2)))
    if bend and len(arr["bend_quads"]):
        q = arr["bend_quads"]
        e = np.einsum("ei,eid->ed", arr["bend_w"].astype(np.float64), x[q])
        E += 0.5 * np.sum(arr["bend_k"].astype(np.float64) * np.sum(e * e, axis=1))
    return E


def dense_K(arr, dt, bend=True):
    """K = M + h^2 Hess(E), assembled by polarization of the quadratic energy."""
    N = len(arr["mass"])
    n = 3 * N
    unit = np.eye(n).reshape(n, N, 3)
    Ei = np.array([quadratic_energy(arr, unit[i], bend=bend) for i in range(n)])
    H = np.zeros((n, n))
    for i in range(n):
        for j in range(i, n):
            H[i, j] = H[j, i] = quadratic_energy(arr, unit[i] + unit[j], bend=bend) - Ei[i] - Ei[j]
        H[i, i] = 2 * Ei[i]
    h2 = float(np.float32(dt)) ** 2
    return np.diag(np.repeat(arr["mass"].astype(np.float64), 3)) + h2 * H


@pytest.mark.parametrize("bend", [True, False])
def test_jacobi_diag_is_diagonal_of_dense_K(bend):
    s, arr = irregular(n=6, seed=9, bend_k=1e-2 if bend else 0.0)
    assert (len(arr["bend_quads"]) > 0) == bend
    dt = 1 / 60
    K = dense_K(arr, dt)
    N = len(arr["mass"])
    # K is symmetric, decoupled in x, y, z with one scalar matrix, and a
    # translation moves only the inertia (sum_j K_ij = m_i per component)
    np.testing.assert_allclose(K, K.T, atol=1e-15)
    blocks = K.reshape(N, 3, N, 3)
    for a in range(3):
        for b in range(3):
            if a != b:
                assert np.abs(blocks[:, a, :, b]).max() < 1e-12
        np.testing.assert_allclose(blocks[:, a, :, a], blocks[:, 0, :, 0], atol=1e-12)
    np.testing.assert_allclose(blocks[:, 0, :, 0].sum(1), arr["mass"], rtol=0,
                               atol=1e-9 * np.abs(K).max())
    want = np.diag(blocks[:, 0, :, 0])
    np.testing.assert_allclose(arr["jacobi_diag"], want, rtol=2e-6)
    np.testing.assert_allclose(po.jacobi_diag_f64(oracle_mesh(arr), dt), want, rtol=1e-12)
    # the mistakes: h^2 dropped, or the bend term dropped
    h2 = float(np.float32(dt)) ** 2
    stiff = want - arr["mass"]
    assert not np.allclose(arr["mass"] + stiff / h2, want, rtol=2e-6)
    if bend:
        K_nobend = dense_K(arr, dt, bend=False)
        assert not np.allclose(np.diag(K_nobend)[::3], want, rtol=2e-6)


def test_type0_jacobi_diag_against_dense_K():
    # the 32x32 workload's material and pins on an 8x8 type-0 grid
    s = ch.Scene(nrows=8, ncols=8, size=1 / 7, type=0, seed=1, stretch_k=1e3, bend_k=1e-2,
                 pin_mode=ch.PIN_CORNERS, dt=1 / 60)
    arr = scene_arrays(s)
    K = dense_K(arr, 1 / 60)
    np.testing.assert_allclose(arr["jacobi_diag"], np.diag(K)[::3], rtol=2e-6)


# ---- warm-start smoothing ----------------------------------------------------------

def ring_mean_brute(c, tris, pinned):
    N = len(c)
    nb = [set() for _ in range(N)]
    for t in tris:
        for a in range(3):
            for b in range(3):
                if a != b:
                    nb[t[a]].add(int(t[b]))
    out = np.zeros_like(c, dtype=np.float64)
    for v in range(N):
        if pinned[v]:
            continue
        ids = [v] + sorted(nb[v])
        out[v] = c[ids].astype(np.float64).mean(0)
    return out


def test_smoothing_pins():
    s, arr = irregular(n=9, seed=4)
    tris = arr["tris"]
    N = len(arr["mass"])
    nb = wo.neighbours(N, tris)
    pinned = np.zeros(N, bool)
    pinned[arr["pin_idx"]] = True
    rng = np.random.default_rng(3)
    # constants are fixed points without pins
    const = np.tile(np.float32([0.3, -1.5, 2.0]), (N, 1))
    out = wo.smooth(const, nb, np.zeros(N, bool), 4)
    np.testing.assert_allclose(out, const, rtol=3e-7)
    # pinned vertices stay zero, one pass is the 1-ring mean (vertex included)
    c = rng.normal(size=(N, 3)).astype(np.float32)
    one = wo.smooth(c, nb, pinned, 1)
    assert np.all(one[pinned] == 0)
    c0 = c.copy()
    c0[pinned] = 0
    want = ring_mean_brute(c0, tris, pinned)
    np.testing.assert_allclose(one, want, rtol=0, atol=1e-6)
    # two passes are the brute-force mean applied twice
    np.testing.assert_allclose(wo.smooth(c, nb, pinned, 2), ring_mean_brute(want, tris, pinned),
                               rtol=0, atol=2e-6)
    # the mistake: a mean that leaves the vertex itself out
    excl = np.array([c0[sorted(set(nb[v]))].mean(0) if not pinned[v] else np.zeros(3)
                     for v in range(N)])
    assert np.abs(excl - one).max() > 1e-2
