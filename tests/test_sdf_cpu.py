"""The learned cull's exact distance and network references on the CPU:
the numpy point-triangle distance against closed forms, the library's FP32
point-triangle query (compiled for the host) against it, and the BF16
rounding helper against its definition."""
import ctypes as C

import numpy as np

import sdf_ref as S


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_reference_distance_closed_forms():
    tri = np.array([[[0, 0, 0], [10, 0, 0], [0, 10, 0]]], np.float32)
    pts = np.array([[2, 2, 5],        # above the interior: plane distance
                    [-3, -4, 0],      # beyond vertex a
                    [5, -2, 0],       # beside edge ab
                    [10, 10, 0],      # beyond edge bc: distance to its midpoint
                    [12, 0, 0]], np.float32)
    want = [5.0, 5.0, 2.0, np.sqrt(50.0), 2.0]
    np.testing.assert_allclose(S.exact_distance(pts, tri), want, rtol=1e-9)


def test_library_point_triangle_distance_matches_reference(cpucheck):
    lib, _ = cpucheck
    rng = np.random.default_rng(5)
    n = 20000
    tris = rng.normal(0, 100, (n, 3, 3)).astype(np.float32)
    pts = rng.normal(0, 150, (n, 3)).astype(np.float32)
    # degenerate-ish triangles too: nearly collinear vertices
    tris[:500, 2] = tris[:500, 0] + 0.5 * (tris[:500, 1] - tris[:500, 0]) + rng.normal(0, 1e-2, (500, 3))
    out = np.zeros(n, np.float32)
    lib.cc_point_tri_dist2(_fp(np.ascontiguousarray(pts)), _fp(np.ascontiguousarray(tris)), n, _fp(out))
    ref = np.array([S.exact_distance(pts[i:i + 1], tris[i:i + 1])[0] for i in range(n)])
    np.testing.assert_allclose(np.sqrt(out), ref, rtol=2e-4, atol=2e-3)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-5], np.float32)
    r = S.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie rounds to even
    assert r[2] == np.float32(1.0 + 2 ** -7)
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-5) / 3.0e-5 < 2 ** -8


def test_reference_backward_matches_finite_differences():
    """The CPU backward's derivation, in float64 without rounding, against
    central differences on every parameter group."""
    rng = np.random.default_rng(8)
    P = np.zeros(S.NUM_PARAMS)
    P[S.OFF_W1:S.OFF_B1] = rng.normal(0, 1, 384)
    P[S.OFF_B1:S.OFF_W2] = rng.normal(0, 0.3, 128)
    P[S.OFF_W2:S.OFF_B2] = rng.normal(0, 1 / np.sqrt(128), 128 * 128)
    P[S.OFF_B2:S.OFF_W3] = rng.normal(0, 0.1, 128)
    P[S.OFF_W3:S.OFF_B3] = rng.normal(0, 1 / np.sqrt(128), 128 * 128)
    P[S.OFF_B3:S.OFF_W4] = rng.normal(0, 0.1, 128)
    P[S.OFF_W4:S.OFF_B4] = rng.normal(0, 1 / np.sqrt(128), 128)
    P[S.OFF_B4] = 3.0
    x = rng.uniform(-1, 1, (64, 3))
    labels = rng.uniform(0, 900, 64)
    ident = lambda v: v  # noqa: E731
    loss, g = S.loss_and_grads(P, x, labels, 1000.0, rnd=ident, dtype=np.float64)
    for k in (S.OFF_W1 + 7, S.OFF_B1 + 3, S.OFF_W2 + 1000, S.OFF_B2 + 9, S.OFF_W3 + 5000, S.OFF_B3 + 100,
              S.OFF_W4 + 12, S.OFF_B4):
        h = 1e-6
        pp = P.copy()
        pp[k] += h
        lp, _ = S.loss_and_grads(pp, x, labels, 1000.0, rnd=ident, dtype=np.float64)
        pp[k] -= 2 * h
        lm, _ = S.loss_and_grads(pp, x, labels, 1000.0, rnd=ident, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(g[k]), (k, fd, g[k])
