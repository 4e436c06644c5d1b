"""The Gershgorin bound behind CLH_RELAX_AUTO: every eigenvalue of D^-1 K lies
at or below the per-element bound max_v (m_v + sum over v's elements of
|row v of the element block|) / D_v, so the damped Jacobi step gamma D^-1 with
gamma = min(1, (1 + rho) / bound) keeps every eigenvalue of the iteration
matrix I - gamma D^-1 K in [-rho, 1). Checked against the exact largest
eigenvalue of D^-1 K, with K = M + h^2 Hess(E) assembled densely by
polarization of the quadratic energy (an independent route to K), and
against the mistakes a restatement could make: entries summed with their
signs (the rows of K then sum to the mass: translations cost only inertia),
or the bend term left out."""
import numpy as np
import pytest

import clothhip as ch
from helpers import jacobi_bound_np, scene_arrays
from test_oracle_pins import dense_K, irregular


def exact_mu(arr, dt, bend):
    K = dense_K(arr, dt, bend=bend)[::3, ::3]  # x, y, z share one scalar matrix
    d = arr["jacobi_diag"].astype(np.float64)
    return float(np.linalg.eigvalsh(K / np.sqrt(np.outer(d, d))).max()), K, d


@pytest.mark.parametrize("bend_k", [0.0, 1e-2, 100.0, 1000.0])
def test_bound_covers_the_jacobi_spectrum(bend_k):
    s, arr = irregular(n=6, seed=9, bend_k=bend_k)
    dt = 1 / 60
    mu, K, d = exact_mu(arr, dt, bend=bend_k > 0)
    bound = jacobi_bound_np(arr, dt)
    assembled = float((np.abs(K).sum(1) / d).max())  # Gershgorin of the assembled rows
    assert mu <= assembled * (1 + 1e-9) <= bound * (1 + 1e-9)
    # the damped iteration matrix's eigenvalues lie in [-rho, 1)
    rho = 0.999999
    gamma = min(1.0, (1 + rho) / bound)
    assert 1 - gamma * mu >= -rho - 1e-12
    # the mistakes give bounds below the spectrum
    assert jacobi_bound_np(arr, dt, signed=True) < mu
    if bend_k >= 100.0:
        assert jacobi_bound_np(arr, dt, bend=False) < mu


def test_plain_jacobi_diverges_where_the_bound_says():
    # obtuse triangles of an irregular mesh, and stiff bending on a regular
    # one, put the largest eigenvalue of D^-1 K above 2 (an eigenvalue of the
    # plain Jacobi iteration below -1); the stretch-only regular grid stays
    # below 2
    _, arr = irregular(n=6, seed=9, bend_k=0.0)
    assert exact_mu(arr, 1 / 60, bend=False)[0] > 2.0
    for bend_k, above in ((1e-2, False), (1.0, True)):
        s = ch.Scene(nrows=8, ncols=8, size=1 / 7, type=0, seed=1, stretch_k=1e3, bend_k=bend_k,
                     pin_mode=ch.PIN_CORNERS, dt=1 / 60)
        arr = scene_arrays(s)
        mu = exact_mu(arr, 1 / 60, bend=True)[0]
        assert (mu > 2.0) == above, (bend_k, mu)
        assert mu <= jacobi_bound_np(arr, 1 / 60) * (1 + 1e-9)
    assert jacobi_bound_np(arr, 1 / 60, bend=False) < 2.0


def test_irregular_workload_needs_the_damped_step():
    """The 1024x1024 workload's mesh family (interior vertices jittered by
    0.3 of the spacing, a random diagonal per quad, stretch log-uniform in
    [5e2, 5e3]) at 64x64: obtuse triangles put the bound of D^-1 K near 4,
    and one 30-iteration frame from rest with the plain Jacobi step under
    Chebyshev weights (rho 0.99) grows its residual to above 1e3 and moves
    vertices by hundreds of metres (the 1024x1024 mesh's bound is 4.63 and
    its frame reaches 8e4 m); the damped step gamma = (1 + rho) / bound
    keeps the frame's residual below 1e-4 and its motion below a millimetre."""
    import pyoracle as po
    from helpers import CFG3, oracle_mesh
    spec = dict(CFG3, nrows=64, ncols=64, size=1 / 63)
    arr = scene_arrays(ch.Scene(**spec))
    m = oracle_mesh(arr)
    mu = jacobi_bound_np(arr, spec["dt"])
    assert mu > 3.5, mu
    rho = float(np.float32(0.99))
    pos0 = arr["pos0"]
    out = {}
    for name, relax in (("plain", 1.0), ("damped", (1.0 + rho) / mu)):
        p = po.params(mode=po.PD, iterations=30, gravity=-9.81, dt=spec["dt"], rho=0.99,
                      cheb_delay=1, relax=relax)
        x, _, _, rs, n = po.step(m, p, pos0, np.zeros_like(pos0), "f32", resid=True)
        out[name] = (np.sqrt(rs[n - 1, 0] / rs[n - 1, 1]), np.abs(x - pos0).max())
    assert out["plain"][0] > 1e3 and out["plain"][1] > 1.0, out
    assert out["damped"][0] < 1e-4 and out["damped"][1] < 1e-3, out
