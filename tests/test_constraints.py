"""Constraints checks: the reference's test_constraints.cpp assertions
(src/tests/test_constraints.cpp:10-26) and the seeded pin selection."""
import numpy as np

import clothhip as ch


def test_pin_top_row():
    s = ch.Scene(nrows=10, ncols=10, size=0.1, pin_mode=ch.PIN_TOP_ROW, pos_jitter=0.0)
    idx = s.get("pin_idx")
    assert len(idx) == 10 and idx[0] == 0 and idx[9] == 9
    np.testing.assert_array_equal(idx, np.arange(10))
    np.testing.assert_array_equal(s.get("pin_target"), s.get("rest_pos")[:10])


def test_pin_corners():
    s = ch.Scene(nrows=10, ncols=10, size=0.1, pin_mode=ch.PIN_CORNERS, pos_jitter=0.0)
    np.testing.assert_array_equal(s.get("pin_idx"), [0, 9])
    np.testing.assert_array_equal(s.get("pin_target"), s.get("rest_pos")[[0, 9]])


def test_random_boundary_pins():
    n = 20
    for seed in range(5):
        s = ch.Scene(nrows=n, ncols=n, size=1 / (n - 1), pin_mode=ch.PIN_RANDOM_BOUNDARY,
                     pin_count=4, seed=seed, pos_jitter=0.0)
        idx = s.get("pin_idx")
        assert len(set(idx.tolist())) == 4
        i, j = np.divmod(idx, n)
        assert np.all((i == 0) | (j == 0) | (i == n - 1) | (j == n - 1))
        again = ch.Scene(nrows=n, ncols=n, size=1 / (n - 1), pin_mode=ch.PIN_RANDOM_BOUNDARY,
                         pin_count=4, seed=seed, pos_jitter=0.0)
        np.testing.assert_array_equal(idx, again.get("pin_idx"))


def test_initial_jitter_spares_pins():
    s = ch.Scene(nrows=16, ncols=16, size=1 / 15, pin_mode=ch.PIN_TOP_ROW, pos_jitter=1e-3, seed=9)
    p0, X = s.get("pos0"), s.get("rest_pos")
    d = p0 - X
    np.testing.assert_array_equal(d[:16], 0)
    np.testing.assert_array_equal(d[:, [0, 2]], 0)
    assert np.abs(d[16:, 1]).max() <= 1e-3 and np.abs(d[16:, 1]).max() > 5e-4
