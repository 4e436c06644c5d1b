"""MeshTopology checks: the reference's test_topology.cpp assertions
(src/tests/test_topology.cpp:17-62), closed-form edge counts, Euler
characteristic, and the edge / hinge numbering."""
import numpy as np
import pytest

import clothhip as ch


def scene(n, typ=0, bend=0.0, **kw):
    return ch.Scene(nrows=n, ncols=n, size=1.0, type=typ, bend_k=bend, pin_mode=ch.PIN_NONE,
                    pos_jitter=0.0, **kw)


def inner(s):
    return int((s.get("edges")[:, 3] != -1).sum())


def test_reference_inner_edge_counts():
    # test_topology.cpp:24-26
    assert inner(scene(3)) == 8
    assert inner(scene(4)) == 21
    assert inner(scene(10)) == 225


def test_reference_types_3x3():
    # test_topology.cpp:29-36
    for typ in (0, 1, 2):
        assert inner(scene(3, typ)) == 8


def test_edges_have_distinct_sorted_vertices():
    # test_topology.cpp:38-45 (v0 != v1), plus v0 < v1 by construction
    e = scene(5, 3).get("edges")
    assert np.all(e[:, 0] < e[:, 1])


def test_vert_to_tris_covers_triangles():
    # test_topology.cpp:47-62
    s = scene(4)
    tris, ptr, idx = s.get("tris"), s.get("vt_ptr"), s.get("vt_idx")
    for t in range(s.T):
        for v in tris[t]:
            assert t in idx[ptr[v]:ptr[v + 1]]


@pytest.mark.parametrize("n", [2, 3, 7, 32])
@pytest.mark.parametrize("typ", [0, 1, 2])
def test_closed_form_counts(n, typ):
    s = scene(n, typ)
    e = s.get("edges")
    assert len(e) == (n - 1) * (3 * n - 1)
    assert inner(s) == (n - 1) * (3 * n - 5)
    assert s.N - len(e) + s.T == 1          # Euler characteristic of a disc


def test_cross_pattern_counts():
    n = 6
    s = scene(n, 3)
    c = (n - 1) ** 2
    assert s.N - len(s.get("edges")) + s.T == 1
    assert s.T == 4 * c


def test_edge_set_matches_triangles():
    s = ch.Scene(nrows=12, ncols=12, size=1 / 11, type=ch.MESH_IRREGULAR, rest_jitter=0.3, seed=4,
                 bend_k=1e-2, pin_mode=ch.PIN_NONE, pos_jitter=0.0)
    tris, e = s.get("tris"), s.get("edges")
    owners = {}
    for t, tri in enumerate(tris):
        for c in range(3):
            a, b = sorted((int(tri[c]), int(tri[(c + 1) % 3])))
            owners.setdefault((a, b), []).append(t)
    got = {(int(a), int(b)): (int(ta), int(tb)) for a, b, ta, tb in e}
    assert set(got) == set(owners)
    for k, ts in owners.items():
        assert got[k] == (ts[0], ts[1] if len(ts) > 1 else -1)
    # rest lengths
    X = s.get("rest_pos")
    np.testing.assert_allclose(s.get("edge_len"), np.linalg.norm(X[e[:, 1]] - X[e[:, 0]], axis=1),
                               rtol=1e-6)


def test_tri_to_tris_symmetric():
    s = scene(6, 1)
    tt, tris = s.get("tri_to_tris"), s.get("tris")
    for t in range(s.T):
        for c in range(3):
            u = tt[t, c]
            if u < 0:
                continue
            assert t in tt[u]
            # the neighbour across the edge opposite corner c shares that edge
            shared = set(tris[t]) - {tris[t, c]}
            assert shared <= set(tris[u])


def test_hinges_follow_inner_edges():
    # build_bend walks the edge list, skipping boundary edges
    # (src/sim_constraints.cpp:57-69); hinge i is the i-th inner edge
    s = scene(9, 0, bend=1.0)
    e = s.get("edges")
    e_in = e[e[:, 3] != -1]
    q = s.get("bend_quads")
    assert s.E == len(e_in)
    np.testing.assert_array_equal(q[:, :2], e_in[:, :2])
    tris = s.get("tris")
    for k in range(0, s.E, 7):
        v2 = set(tris[e_in[k, 2]]) - set(e_in[k, :2])
        v3 = set(tris[e_in[k, 3]]) - set(e_in[k, :2])
        assert {q[k, 2]} == v2 and {q[k, 3]} == v3


def test_edge_order_is_deterministic():
    a = scene(16).get("edges")
    b = scene(16).get("edges")
    np.testing.assert_array_equal(a, b)
