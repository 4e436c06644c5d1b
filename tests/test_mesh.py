"""ClothMesh / generator checks: the reference's test_mesh.cpp assertions
(src/tests/test_mesh.cpp:13-70), closed forms of the grid generator
(src/mesh_generator.cpp:6-91) and an FP64 recomputation of the rest state."""
import os
import tempfile

import numpy as np
import pytest

import clothhip as ch
import pyoracle as po


def scene(**kw):
    base = dict(nrows=10, ncols=10, size=0.1, type=0, density=0.1, stretch_k=100.0, bend_k=0.0,
                pin_mode=ch.PIN_NONE, pos_jitter=0.0)
    base.update(kw)
    return ch.Scene(**base)


def test_mass_conservation():
    # test_mesh.cpp:13-21: total mass = density * total area to 1e-5
    s = scene()
    area, mass = s.get("rest_area"), s.get("mass")
    expected = np.float32(0.1) * np.float32(area.sum(dtype=np.float32))
    assert abs(mass.sum(dtype=np.float32) - expected) / (expected + 1e-12) < 1e-5


def test_dm_inv_times_dm_is_identity():
    # test_mesh.cpp:23-45
    s = scene()
    dminv = s.get("Dm_inv")        # column-major [m00, m10, m01, m11]
    dm = s.get("deltaUV")
    for t in range(s.T):
        G = np.array([[dminv[t, 0], dminv[t, 2]], [dminv[t, 1], dminv[t, 3]]])
        D = np.array([[dm[t, 0], dm[t, 2]], [dm[t, 1], dm[t, 3]]])
        assert np.linalg.norm(G @ D - np.eye(2)) < 1e-4


def test_dF_dx_maps_rest_pose_to_F():
    # test_mesh.cpp:47-70: dF_dx * x = vec(Ds * Dm_inv)
    s = scene()
    X, tris, dminv, dfdx = s.get("rest_pos"), s.get("tris"), s.get("Dm_inv"), s.get("dF_dx")
    for t in range(10):
        i0, i1, i2 = tris[t]
        x = np.concatenate([X[i0], X[i1], X[i2]])
        Fvec = dfdx[t].reshape(6, 9) @ x
        Ds = np.stack([X[i1] - X[i0], X[i2] - X[i0]], axis=1)
        G = np.array([[dminv[t, 0], dminv[t, 2]], [dminv[t, 1], dminv[t, 3]]])
        F = Ds @ G
        assert np.linalg.norm(Fvec - np.concatenate([F[:, 0], F[:, 1]])) < 1e-4


@pytest.mark.parametrize("n,s_", [(3, 1.0), (10, 0.1), (32, 1 / 31)])
def test_grid_closed_forms(n, s_):
    s = scene(nrows=n, ncols=n, size=s_)
    X, tris = s.get("rest_pos"), s.get("tris")
    assert s.N == n * n and s.T == 2 * (n - 1) ** 2
    i, j = np.divmod(np.arange(n * n), n)
    np.testing.assert_array_equal(X[:, 0], (j.astype(np.float32) * np.float32(s_)))
    np.testing.assert_array_equal(X[:, 1], 0.0)
    np.testing.assert_array_equal(X[:, 2], (i.astype(np.float32) * np.float32(s_)))
    # type 0: cell (i, j) -> (a, b, d), (a, d, c) with a = i*n + j
    ci, cj = np.divmod(np.arange((n - 1) ** 2), n - 1)
    a = ci * n + cj
    exp = np.stack([np.stack([a, a + 1, a + n + 1], 1), np.stack([a, a + n + 1, a + n], 1)], 1)
    np.testing.assert_array_equal(tris, exp.reshape(-1, 3))
    # every triangle is a right triangle of area s^2/2; interior mass rho s^2
    np.testing.assert_allclose(s.get("rest_area"), 0.5 * s_ * s_, rtol=1e-5)
    m = s.get("mass").reshape(n, n)
    np.testing.assert_allclose(m[1:-1, 1:-1], 0.1 * s_ * s_, rtol=1e-5)


def test_other_types():
    n = 4
    for typ, per_cell, nv in [(1, 2, n * n), (2, 2, n * n), (3, 4, n * n + (n - 1) ** 2)]:
        s = scene(nrows=n, ncols=n, size=1.0, type=typ)
        assert s.T == per_cell * (n - 1) ** 2 and s.N == nv
        tris = s.get("tris")
        if typ == 2:
            np.testing.assert_array_equal(tris[:2], [[0, 1, 4], [1, 5, 4]])
        if typ == 1:
            # checkerboard: cell (0,0) "\" then cell (0,1) "/"
            np.testing.assert_array_equal(tris[:4], [[0, 1, 5], [0, 5, 4], [1, 2, 5], [2, 6, 5]])
        if typ == 3:
            c = n * n  # first cell-centre vertex
            np.testing.assert_array_equal(tris[:4], [[0, 1, c], [1, 5, c], [5, 4, c], [4, 0, c]])
            np.testing.assert_allclose(s.get("rest_pos")[c], [0.5, 0.0, 0.5])


def test_irregular_generator():
    n, j = 16, 0.3
    a = scene(nrows=n, ncols=n, size=1 / (n - 1), type=ch.MESH_IRREGULAR, rest_jitter=j, seed=7)
    b = scene(nrows=n, ncols=n, size=1 / (n - 1), type=ch.MESH_IRREGULAR, rest_jitter=j, seed=7)
    c = scene(nrows=n, ncols=n, size=1 / (n - 1), type=ch.MESH_IRREGULAR, rest_jitter=j, seed=8)
    Xa, Ta = a.get("rest_pos"), a.get("tris")
    np.testing.assert_array_equal(Xa, b.get("rest_pos"))
    np.testing.assert_array_equal(Ta, b.get("tris"))
    assert not np.array_equal(Ta, c.get("tris"))
    grid = scene(nrows=n, ncols=n, size=1 / (n - 1)).get("rest_pos")
    d = (Xa - grid).reshape(n, n, 3)
    # boundary fixed, interior within +-jitter*size in x and z, y untouched
    assert np.all(d[0] == 0) and np.all(d[-1] == 0) and np.all(d[:, 0] == 0) and np.all(d[:, -1] == 0)
    assert np.abs(d[..., [0, 2]]).max() <= j / (n - 1) * (1 + 1e-6)
    assert np.abs(d[1:-1, 1:-1][..., [0, 2]]).max() > 0.5 * j / (n - 1)
    np.testing.assert_array_equal(d[..., 1], 0)
    # both diagonals occur; each quad split into 2 triangles of positive area
    assert a.T == 2 * (n - 1) ** 2
    area = a.get("rest_area")
    assert area.min() > 0 and abs(area.sum() - 1.0) < 1e-5


def test_rest_state_matches_fp64_recomputation():
    s = scene(nrows=24, ncols=24, size=1 / 23, type=ch.MESH_IRREGULAR, rest_jitter=0.3, seed=5)
    dm, ar, ms = po.rest_state_f64(s.get("rest_pos"), s.get("tris"), 0.1)
    np.testing.assert_allclose(s.get("Dm_inv"), dm, rtol=2e-5, atol=2e-5 * np.abs(dm).max())
    np.testing.assert_allclose(s.get("rest_area"), ar, rtol=1e-5)
    np.testing.assert_allclose(s.get("mass"), ms, rtol=1e-5)


def test_obj_loading(tmp_path):
    # fan triangulation of polygons around their first corner and 1-based
    # "v/t/n" indices (src/cloth_mesh.cpp:89-132)
    path = tmp_path / "shape.obj"
    path.write_text("# comment\nv 0 0 0\nv 1 0 0\nv 1 0 1\nv 0 0 1\nv 2 0 0\n"
                    "f 1/1/1 2/2/2 3/3/3 4/4/4\nf 2 5 3\n")
    s = ch.Scene(obj_path=path, pin_mode=ch.PIN_NONE, pos_jitter=0.0, bend_k=0.0, ncols=2)
    assert (s.N, s.T) == (5, 3)
    np.testing.assert_array_equal(s.get("tris"), [[0, 1, 2], [0, 2, 3], [1, 4, 2]])
    np.testing.assert_allclose(s.get("rest_area"), [0.5, 0.5, 0.5], rtol=1e-6)
    missing = tmp_path / "missing.obj"
    with pytest.raises(ch.ClothHipError):
        ch.Scene(obj_path=missing)
