"""Geometric building blocks of the tick, compiled for the host, against the
same routines of the reference's Bullet (the oracle library), on the same
inputs:

* the suspension ray against a convex body (btSubsimplexConvexCast, a zero
  radius sphere cast at the car's box and at the ball's sphere),
* the internal-edge table of every soccar collision mesh
  (btGenerateInternalEdgeInfo), which decides whether a contact normal near a
  triangle edge is bent.

The convex cast is iterative and stops once the simplex is within 1e-2 Bullet
units of the body, so a rounding difference in the first steps (the reference
normalises with the SSE reciprocal square root estimate) can stop it at
another iteration.  Hit or miss must agree on every ray; fractions agree
to 1e-6 on most rays and within the cast's own stopping bound on all.
"""
import ctypes as C

import numpy as np
import pytest

BOX = [0.5905, 0.4025, 0.1789, 0.0386591]  # implicit half extents + margin (Bullet units)
SPHERE = [0.9125, 0, 0, 0]                 # the ball's radius (Bullet units)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _cast_pairs(oracle, lib, kind, n, seed):
    rng = np.random.default_rng(seed)
    dims = np.array(BOX if kind == 0 else SPHERE, np.float32)
    out = []
    for _ in range(n):
        rot, org = _rot(rng), rng.uniform(-1, 1, 3)
        xf = np.concatenate([rot.ravel(), org]).astype(np.float32)
        # aim at a point around the body from 1.2 to 2.5 units away, ending
        # short of, at, or past the aim point
        aim = org + rot @ (rng.uniform(-1.2, 1.2, 3) * np.maximum(dims[:3] + dims[3], dims[0]))
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a = (aim - d * rng.uniform(1.2, 2.5)).astype(np.float32)
        b = (aim + d * rng.uniform(-1.0, 1.0)).astype(np.float32)
        ro = np.zeros(5, np.float32)
        eo = np.zeros(5, np.float32)
        oracle.ref_ray_convex(kind, _fp(dims), _fp(xf), _fp(a), _fp(b), _fp(ro))
        lib.cc_ray_convex(kind, _fp(dims), _fp(xf), _fp(a), _fp(b), _fp(eo))
        out.append((ro, eo, float(np.linalg.norm(b - a))))
    return out


@pytest.mark.parametrize("kind,exact_share", [(0, 0.85), (1, 0.85)], ids=["car box", "ball sphere"])
def test_ray_cast_against_convex_matches_bullet(oracle, cpucheck, kind, exact_share):
    lib, _ = cpucheck
    pairs = _cast_pairs(oracle, lib, kind, 3000, seed=11 + kind)
    hits = [(r, e, L) for r, e, L in pairs if r[0]]
    assert all(r[0] == e[0] for r, e, _ in pairs), "hit / miss disagree"
    assert len(hits) > 500
    exact = np.mean([abs(r[1] - e[1]) <= 1e-6 for r, e, _ in hits])
    assert exact >= exact_share, f"only {exact:.2f} of hits agree to 1e-6"
    # a cast that stopped at another iteration is still within the stopping
    # distance (1e-2) of the body, measured along the ray over its length
    for r, e, L in hits:
        assert abs(r[1] - e[1]) * L <= 0.1, (r, e)
        assert np.dot(r[2:], e[2:]) > 0.9, (r, e)


def test_ray_cast_misses_and_grazes(oracle, cpucheck):
    """Rays that start inside, end short of, run parallel past, or point away
    from the box."""
    lib, _ = cpucheck
    dims = np.array(BOX, np.float32)
    xf = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32)
    rays = [((0, 0, 0), (0, 0, -1)),          # starts inside
            ((0, 0, 1), (0, 0, 0.3)),         # ends short of the top face
            ((0, 0, 1), (0, 0, 0.2)),         # just reaches it
            ((0.7, 0, 1), (0.7, 0, -1)),      # passes beside the box
            ((0.62, 0, 1), (0.62, 0, -1)),    # within the margin of the side
            ((0, 0, 1), (0, 0, 2)),           # points away
            ((-2, 0.1, 0.1), (2, 0.1, 0.1))]  # straight through
    for a, b in rays:
        a = np.array(a, np.float32)
        b = np.array(b, np.float32)
        ro = np.zeros(5, np.float32)
        eo = np.zeros(5, np.float32)
        oracle.ref_ray_convex(0, _fp(dims), _fp(xf), _fp(a), _fp(b), _fp(ro))
        lib.cc_ray_convex(0, _fp(dims), _fp(xf), _fp(a), _fp(b), _fp(eo))
        assert ro[0] == eo[0], (a, b, ro, eo)
        if ro[0]:
            np.testing.assert_allclose(eo[1:], ro[1:], atol=1e-6)


def test_internal_edge_table_matches_bullet(oracle, cpucheck):
    """Per triangle of every mesh: the edge angles and the convex / swap flags
    the host builds equal the ones Bullet stored, the flags and the zero /
    2*pi markers exactly, the angles to a few float ulps."""
    lib, w = cpucheck
    n_max = 200000
    buf = np.zeros(n_max * 14, np.float32)
    n = lib.cc_tri_info(w, _fp(buf), n_max)
    mine = buf[:n * 14].reshape(n, 14)
    by_verts = {row[:9].tobytes(): row for row in mine}
    matched = 0
    for m in range(16):
        rb = np.zeros(100000 * 13, np.float32)
        k = oracle.ref_mesh_tri_info(m, _fp(rb), 100000)
        assert k > 0
        for r in rb[:k * 13].reshape(k, 13):
            row = by_verts.get(r[:9].tobytes())
            assert row is not None, f"mesh {m}: a triangle of the reference is missing"
            matched += 1
            assert row[12:13].view(np.uint32)[0] == r[12:13].view(np.uint32)[0], f"mesh {m}: flags"
            for e in range(3):
                a, b = float(row[9 + e]), float(r[9 + e])
                assert (a == 0) == (b == 0), f"mesh {m}: flat edge marker"
                assert abs(a - b) <= 2e-6 * max(1.0, abs(b)), f"mesh {m}: angle {a} vs {b}"
    assert matched == n
