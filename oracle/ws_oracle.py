"""CPU restatement of the learned warm start (test infrastructure only: the
checker of the HIP kernels in cloth-cuda_amd/csrc/ws_mlp.hip; only tests/,
the smoke check and bench.py's CPU leg use it).

The warm start has no counterpart in the reference (src/pd_solver.cu:528
starts every frame from x_n); what it must reproduce is the definition the
library states: the features, the network, the loss and its gradient, and
Adam. Two evaluations of the network:
  * exact: float64 throughout;
  * emulate: float64 with the library's BF16 roundings applied where the
    kernels round (weights, features, hidden activations and the backward
    signals that enter BF16 MFMA operands), so the GPU result differs from it
    only by FP32 accumulation order.
The exact gradient is pinned by central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

F_IN, HID, OUT_ROWS, OUT = 16, 64, 16, 3
G0 = 9.81

# flat parameter layout (row-major)
O_W1 = 0
O_B1 = O_W1 + HID * F_IN
O_W2 = O_B1 + HID
O_B2 = O_W2 + HID * HID
O_W3 = O_B2 + HID
O_B3 = O_W3 + OUT_ROWS * HID
N_PARAMS = O_B3 + OUT_ROWS


def bf16_bits(x):
    """float32 -> BF16 bits, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16(x):
    """Round to BF16 and back (as float64)."""
    return bf16_value(bf16_bits(np.asarray(x, np.float32))).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    W1 = p[O_W1:O_B1].reshape(HID, F_IN)
    b1 = p[O_B1:O_W2]
    W2 = p[O_W2:O_B2].reshape(HID, HID)
    b2 = p[O_B2:O_W3]
    W3 = p[O_W3:O_B3].reshape(OUT_ROWS, HID)[:OUT]
    b3 = p[O_B3:][:OUT]
    return W1, b1, W2, b2, W3, b3


def neighbours(N, tris):
    """Sorted unique 1-ring of each vertex (the triangles' edges)."""
    nb = [set() for _ in range(N)]
    for t in np.asarray(tris).reshape(-1, 3):
        for i in range(3):
            for j in range(3):
                if i != j:
                    nb[int(t[i])].add(int(t[j]))
    return [sorted(s - {v}) for v, s in enumerate(nb)]


def features(xm2, xm1, x, dt, acc_uniform, pinned, nb, cloth_accel=None, vcloth=None):
    """The 16 features per vertex in float32, in the kernel's operation order
    (before BF16 rounding)."""
    f32 = np.float32
    xm2, xm1, x = (np.asarray(a, f32).reshape(-1, 3) for a in (xm2, xm1, x))
    inv_L = f32(1.0) / (f32(dt) * f32(dt) * f32(G0))
    d1 = x - xm1
    d0 = xm1 - xm2
    u = d1 * inv_L
    ac = (d1 - d0) * inv_L
    N = x.shape[0]
    out = np.zeros((N, F_IN), f32)
    for v in range(N):
        su = np.zeros(3, f32)
        sa = np.zeros(3, f32)
        for w in nb[v]:
            e1 = x[w] - xm1[w]
            e0 = xm1[w] - xm2[w]
            su = su + e1 * inv_L
            sa = sa + (e1 - e0) * inv_L
        deg = len(nb[v])
        if deg:
            inv_deg = f32(1.0) / f32(deg)
            gu = su * inv_deg - u[v]
            ga = sa * inv_deg - ac[v]
        else:
            gu = ga = np.zeros(3, f32)
        ext = np.asarray(acc_uniform, f32).copy()
        if cloth_accel is not None:
            ext = ext + np.asarray(cloth_accel, f32).reshape(-1, 3)[vcloth[v] if vcloth is not None else 0]
        out[v, 0:3] = u[v]
        out[v, 3:6] = ac[v]
        out[v, 6:9] = ext / f32(G0)
        out[v, 9] = 1.0 if pinned[v] else 0.0
        out[v, 10:13] = gu
        out[v, 13:16] = ga
    return out


def targets(xm1, x, xnext, dt, ac, pinned, out_scale):
    """(change of acceleration / s_o, weight) per vertex, float32."""
    f32 = np.float32
    xm1, x, xnext = (np.asarray(a, f32).reshape(-1, 3) for a in (xm1, x, xnext))
    inv_L = f32(1.0) / (f32(dt) * f32(dt) * f32(G0))
    d1 = x - xm1
    d2 = xnext - x
    t = ((d2 - d1) * inv_L - ac) * (f32(1.0) / f32(out_scale))
    w = np.where(np.asarray(pinned, bool), 0.0, 1.0).astype(f32)
    return np.concatenate([t, w[:, None]], axis=1)


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, F, emulate=True):
    """Network output o [S x 3] (float64) for features F [S x 16] given as
    float values (already BF16 when they come from the library)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    F = np.asarray(F, np.float64)
    z1 = F @ r(W1).T + b1
    h1 = r(_silu(z1))
    z2 = h1 @ r(W2).T + b2
    h2 = r(_silu(z2))
    return h2 @ r(W3).T + b3


def loss_grad(p, F, T, emulate=True):
    """Loss sum_s w_s |o_s - t_s|^2 / (3 S) and its gradient in the flat
    layout (float64)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = bf16 if emulate else (lambda a: np.asarray(a, np.float64))
    F = np.asarray(F, np.float64)
    T = np.asarray(T, np.float64)
    S = F.shape[0]
    inv_norm = 1.0 / (3.0 * S)
    if emulate:
        inv_norm = float(np.float32(1.0) / np.float32(3.0 * S))
    z1 = F @ r(W1).T + b1
    h1 = r(_silu(z1))
    z2 = h1 @ r(W2).T + b2
    h2 = r(_silu(z2))
    o = h2 @ r(W3).T + b3
    e = o - T[:, :3]
    w = T[:, 3:4]
    loss = float(np.sum(w * e * e) * inv_norm)
    d = 2.0 * w * e * inv_norm                  # dL/do
    dq = r(d)
    dz2 = (dq @ r(W3)) * _dsilu(z2)
    dz2q = r(dz2)
    dz1 = (dz2q @ r(W2)) * _dsilu(z1)
    dz1q = r(dz1)
    g = np.zeros(N_PARAMS)
    g[O_W1:O_B1] = (dz1q.T @ F).ravel()
    g[O_B1:O_W2] = dz1.sum(0)
    g[O_W2:O_B2] = (dz2q.T @ h1).ravel()
    g[O_B2:O_W3] = dz2.sum(0)
    gW3 = np.zeros((OUT_ROWS, HID))
    gW3[:OUT] = dq.T @ h2
    g[O_W3:O_B3] = gW3.ravel()
    g[O_B3:O_B3 + OUT] = d.sum(0)
    return loss, g


def adam(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t counts from 1); rows 3.. of W3 and b3 stay fixed."""
    p, m, v, g = (np.asarray(a, np.float64).copy() for a in (p, m, v, g))
    frozen = np.zeros(N_PARAMS, bool)
    frozen[O_W3 + OUT * HID:O_B3] = True
    frozen[O_B3 + OUT:] = True
    m2 = b1 * m + (1 - b1) * g
    v2 = b2 * v + (1 - b2) * g * g
    upd = lr * (m2 / (1 - b1 ** t)) / (np.sqrt(v2 / (1 - b2 ** t)) + eps)
    p2 = np.where(frozen, p, p - upd)
    return p2, np.where(frozen, m, m2), np.where(frozen, v, v2)


def extrapolation(xm2, xm1, x):
    """x + d1 + (d1 - d0) in float32, the kernel's order."""
    f32 = np.float32
    xm2, xm1, x = (np.asarray(a, f32).reshape(-1, 3) for a in (xm2, xm1, x))
    d1 = x - xm1
    d0 = xm1 - xm2
    return x + d1 + (d1 - d0)


def smooth(corr, nb, pinned, passes):
    """passes of the mean over each vertex and its 1-ring, pinned vertices 0
    (float32, neighbours summed in ascending order after the vertex)."""
    c = np.asarray(corr, np.float32).reshape(-1, 3).copy()
    pinned = np.asarray(pinned, bool)
    c[pinned] = 0.0
    for _ in range(passes):
        out = np.zeros_like(c)
        for v in range(len(c)):
            if pinned[v]:
                continue
            s = c[v].copy()
            for w in nb[v]:
                s = s + c[w]
            out[v] = s * (np.float32(1.0) / np.float32(1 + len(nb[v])))
        c = out
    return c
