"""CPU references for the learned cull's GPU kernels (test infrastructure).

* exact_distance: unsigned point-triangle distance by a different route than
  the library's Voronoi-region walk -- project onto the triangle's plane, keep
  the projection if its barycentric coordinates are all >= 0, otherwise take
  the nearest of the three edge segments -- vectorised in float64.
* forward / backward: the MLP in numpy, FP32 with the library's BF16 rounding
  points (weights of the 128x128 layers, activations fed to them, layer
  deltas fed to the matrix products), so that the GPU's matrix-core kernels
  are compared against the same arithmetic up to summation order.
"""
import numpy as np

H = 128
OFF_W1, OFF_B1 = 0, 384
OFF_W2, OFF_B2 = 512, 512 + H * H
OFF_W3, OFF_B3 = OFF_B2 + H, OFF_B2 + H + H * H
OFF_W4, OFF_B4 = OFF_B3 + H, OFF_B3 + 2 * H
NUM_PARAMS = OFF_B4 + 1
OUT_SCALE = 100.0


def _seg_dist2(p, a, b):
    ab = b - a
    t = np.clip(np.einsum("...k,...k->...", p - a, ab) / np.maximum(np.einsum("...k,...k->...", ab, ab), 1e-30), 0, 1)
    d = p - (a + t[..., None] * ab)
    return np.einsum("...k,...k->...", d, d)


def exact_distance(points, tris):
    """points (n, 3), tris (m, 3, 3) -> (n,) float64 distances."""
    p = points.astype(np.float64)[:, None, :]
    a, b, c = (tris[:, i].astype(np.float64)[None] for i in range(3))
    n = np.cross(b - a, c - a)
    nn = np.einsum("...k,...k->...", n, n)
    s = np.einsum("...k,...k->...", p - a, n) / np.maximum(nn, 1e-30)
    q = p - s[..., None] * n
    # barycentric coordinates of the projection
    w0 = np.einsum("...k,...k->...", np.cross(b - q, c - q), n)
    w1 = np.einsum("...k,...k->...", np.cross(c - q, a - q), n)
    w2 = np.einsum("...k,...k->...", np.cross(a - q, b - q), n)
    inside = (w0 >= 0) & (w1 >= 0) & (w2 >= 0) & (nn > 0)
    d_plane = s * s * nn
    d_edge = np.minimum(np.minimum(_seg_dist2(p, a, b), _seg_dist2(p, b, c)), _seg_dist2(p, c, a))
    d2 = np.where(inside, d_plane, d_edge)
    return np.sqrt(d2.min(axis=1))


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def silu(z):
    return z / (1 + np.exp(-z))


def silu_grad(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def unpack(P, dtype=np.float32):
    P = np.asarray(P, dtype)
    return (P[OFF_W1:OFF_B1].reshape(H, 3), P[OFF_B1:OFF_W2], P[OFF_W2:OFF_B2].reshape(H, H), P[OFF_B2:OFF_W3],
            P[OFF_W3:OFF_B3].reshape(H, H), P[OFF_B3:OFF_W4], P[OFF_W4:OFF_B4], P[OFF_B4])


def scale_inputs(points, scale):
    cx, cy, cz, hx, hy, hz = scale
    p = np.asarray(points, np.float32)
    return np.stack([(p[:, 0] - cx) / hx, (p[:, 1] - cy) / hy, (p[:, 2] - cz) / hz], 1).astype(np.float32)


def forward(P, x, rnd=bf16, dtype=np.float32):
    """x: scaled inputs (n, 3).  Returns (y in units of 100 uu, cache).
    rnd is applied where the GPU rounds to bfloat16 (identity: exact math)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(P, dtype)
    x = np.asarray(x, dtype)
    z1 = (x @ W1.T + b1).astype(dtype)
    a1 = rnd(silu(z1))
    z2 = (a1 @ rnd(W2).T + b2).astype(dtype)
    a2 = rnd(silu(z2))
    z3 = (a2 @ rnd(W3).T + b3).astype(dtype)
    a3 = rnd(silu(z3))
    y = (a3 @ w4 + b4).astype(dtype)
    return y, (x, z1, a1, z2, a2, z3, a3)


def loss_and_grads(P, x, labels_uu, clamp_uu, rnd=bf16, dtype=np.float32, over_weight=1.0):
    """Mean L1 loss (units of 100 uu) and its gradient, with the GPU's
    rounding of the deltas to bfloat16 (rnd = identity, dtype = float64:
    the exact gradient)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(P, dtype)
    y, (x, z1, a1, z2, a2, z3, a3) = forward(P, x, rnd, dtype)
    t = np.minimum(labels_uu, clamp_uu) / OUT_SCALE
    e = y - t
    n = len(y)
    w = np.where(e > 0, over_weight, 1.0).astype(dtype)
    loss = float((np.abs(e) * w).mean())
    gy = (np.sign(e) * w).astype(dtype) / n
    g = np.zeros(NUM_PARAMS, np.float64)
    g[OFF_W4:OFF_B4] = gy @ a3
    g[OFF_B4] = gy.sum()
    d3 = rnd(gy[:, None] * w4[None, :] * silu_grad(z3))
    g[OFF_W3:OFF_B3] = (d3.T.astype(np.float64) @ a2).ravel()
    g[OFF_B3:OFF_W4] = d3.sum(0)
    d2 = rnd((d3 @ rnd(W3)) * silu_grad(z2))
    g[OFF_W2:OFF_B2] = (d2.T.astype(np.float64) @ a1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = rnd((d2 @ rnd(W2)) * silu_grad(z1))
    g[OFF_W1:OFF_B1] = (d1.T.astype(np.float64) @ x).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, g.astype(dtype)
