"""Synthetic arena states for benchmarks and large-batch tests.

random_field_states draws the random-state workload: per arena, the ball
uniform inside the field volume (|x| < 4096, |y| < 5120, 93 < z < 2000 uu) with
velocity uniform in a 4000 uu/s ball and angular velocity uniform in a 6 rad/s
ball; cars resting on the floor (z 17) at uniform x/y inside the field, at
least 300 uu from the ball and from each other, uniform yaw, boost uniform
0-100; each of the 34 pads active with probability 1/2, else with a cooldown
uniform in (0, 4] s (small) or (0, 10] s (big).  Draw order: numpy PCG64
seeded with `seed`; ball, then cars, then pads, arena by arena in bulk."""
import warnings

import numpy as np

from .records import NUM_PADS

NUM_BIG_PADS = 6  # pads 0-5 are the big ones (RLConst.h BoostPads, big first)


def _ball_dirs(rng, n, radius):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * rng.uniform(0, 1, size=(n, 1)) ** (1 / 3)
    return d * r


def _view(arr):
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        return np.ctypeslib.as_array(arr)


def mixed_team_sizes(n, seed):
    """Team size per arena for the mixed workload: 1v1 for half of the arenas,
    2v2 for 30%, 3v3 for the rest, shuffled by numpy PCG64 seeded with `seed`."""
    t = np.array([1] * (n // 2) + [2] * (3 * n // 10) + [3] * (n - n // 2 - 3 * n // 10))
    np.random.default_rng(seed).shuffle(t)
    return [int(x) for x in t]


def random_field_states(batch, seed):
    """Applies the random-state workload to every arena of `batch`."""
    rng = np.random.default_rng(seed)
    n = batch.num_arenas
    balls = batch.GetAllBallStates()
    bv = _view(balls)
    bpos = np.stack([rng.uniform(-4096, 4096, n), rng.uniform(-5120, 5120, n), rng.uniform(93, 2000, n)], 1)
    bvel = _ball_dirs(rng, n, 4000.0)
    bang = _ball_dirs(rng, n, 6.0)
    for k, arr in (("pos", bpos), ("vel", bvel), ("ang_vel", bang)):
        for j, c in enumerate("xyz"):
            bv[k][c][:] = arr[:, j]
    batch.SetAllBallStates(balls)

    cars = batch.GetAllCarStates()
    cv = _view(cars)
    nc = batch.num_cars
    pos = np.zeros((nc, 2))
    arena_of = np.repeat(np.arange(n), [2 * t for t in batch.team_size])
    first = np.concatenate([[0], np.cumsum([2 * t for t in batch.team_size])[:-1]])
    todo = np.arange(nc)
    while todo.size:
        pos[todo] = np.stack([rng.uniform(-3900, 3900, todo.size), rng.uniform(-4900, 4900, todo.size)], 1)
        bad = np.linalg.norm(pos[todo] - bpos[arena_of[todo], :2], axis=1) < 300
        # cars of one arena keep 300 uu apart (checked against the earlier
        # slots' current positions)
        for k in range(1, 6):
            j = todo - k
            ok = j >= first[arena_of[todo]]
            near = np.zeros(todo.size, bool)
            near[ok] = np.linalg.norm(pos[todo[ok]] - pos[j[ok]], axis=1) < 300
            bad |= near
        todo = todo[bad]
    yaw = rng.uniform(-np.pi, np.pi, nc)
    cy, sy = np.cos(yaw), np.sin(yaw)
    cv["pos"]["x"][:], cv["pos"]["y"][:], cv["pos"]["z"][:] = pos[:, 0], pos[:, 1], 17.0
    cv["rot"]["forward"]["x"][:], cv["rot"]["forward"]["y"][:], cv["rot"]["forward"]["z"][:] = cy, sy, 0
    cv["rot"]["right"]["x"][:], cv["rot"]["right"]["y"][:], cv["rot"]["right"]["z"][:] = -sy, cy, 0
    cv["rot"]["up"]["x"][:], cv["rot"]["up"]["y"][:], cv["rot"]["up"]["z"][:] = 0, 0, 1
    for k in ("vel", "ang_vel"):
        for c in "xyz":
            cv[k][c][:] = 0
    cv["boost"][:] = rng.uniform(0, 100, nc)
    cv["is_on_ground"][:] = 1
    batch.SetAllCarStates(cars)

    pads = batch.GetAllBoostPadStates()
    pv = _view(pads)
    active = rng.uniform(size=n * NUM_PADS) < 0.5
    big = (np.arange(n * NUM_PADS) % NUM_PADS) < NUM_BIG_PADS
    cd = np.where(big, 10.0, 4.0) * (1 - rng.uniform(size=n * NUM_PADS))  # (0, max]
    pv["is_active"][:] = active
    pv["cooldown"][:] = np.where(active, 0.0, cd)
    pv["prev_locked_car_id"][:] = 0
    batch.SetAllBoostPadStates(pads)


def soccar_triangles(mesh_dir=None):
    """(ntri, 3, 3) float32 vertices (uu) of the 16 soccar collision meshes
    (CMF files: int32 triangle count, int32 vertex count, int32 indices,
    float32 vertices in Bullet units, 50 uu each; CollisionMeshFile.cpp:11-61)."""
    import os
    from .batch import MESH_DIR
    mesh_dir = mesh_dir or MESH_DIR
    out = []
    for m in range(16):
        raw = open(os.path.join(mesh_dir, f"mesh_{m}.cmf"), "rb").read()
        nt, nv = np.frombuffer(raw[:8], np.int32)
        idx = np.frombuffer(raw[8:8 + 12 * nt], np.int32).reshape(nt, 3)
        v = np.frombuffer(raw[8 + 12 * nt:8 + 12 * nt + 12 * nv], np.float32).reshape(nv, 3)
        out.append(v[idx])
    return (np.concatenate(out) * np.float32(50.0)).astype(np.float32)


def _frames(up, rng):
    """Random orthonormal (forward, right, up) frames with the given up axes.
    Right = up x forward, as the car records' rotation has it for up = z."""
    n = len(up)
    t = rng.normal(size=(n, 3))
    fwd = t - np.sum(t * up, 1, keepdims=True) * up
    fwd /= np.linalg.norm(fwd, axis=1, keepdims=True)
    right = np.cross(up, fwd)
    return fwd, right


def stress_states(batch, seed):
    """The contact-heavy workload, applied to every arena of a 1v1 batch:
    the random-state workload first, then
      - 30% of arenas: the ball within 200 uu of a corner, ramp or goal-mouth
        surface (a uniform point of an area-weighted mesh triangle below
        z = 500 uu, or of the goal mouths |y| > 4900, |x| < 1300; centre at
        radius + U(0, 200) uu along the triangle's normal, into the field);
      - 20% of cars: on a side wall, the ceiling or a back wall, up axis
        along the surface's inward normal, 17 uu off the surface, random
        heading in the surface plane, at rest;
      - 10% of arenas: the orange car overlapping the blue car's spawn
        (within 60 uu, random yaw) for bump and demolition contacts.
    Draw order: numpy PCG64 seeded with `seed`, after the random-state
    workload's own draws with the same seed."""
    random_field_states(batch, seed)
    rng = np.random.default_rng(seed + 1000003)
    n = batch.num_arenas
    if any(t != 1 for t in batch.team_size):
        raise ValueError("stress_states expects 1v1 arenas")
    tris = soccar_triangles()
    a, b, c = tris[:, 0].astype(np.float64), tris[:, 1].astype(np.float64), tris[:, 2].astype(np.float64)
    cen = (a + b + c) / 3
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.maximum(2 * area, 1e-12)[:, None]
    pick = ((cen[:, 2] < 500) | ((np.abs(cen[:, 1]) > 4900) & (np.abs(cen[:, 0]) < 1300))) & (area > 1.0)
    ids = np.nonzero(pick)[0]

    balls = batch.GetAllBallStates()
    bv = _view(balls)
    sel = np.nonzero(rng.uniform(size=n) < 0.3)[0]
    t = ids[rng.choice(len(ids), size=len(sel), p=area[ids] / area[ids].sum())]
    r1, r2 = np.sqrt(rng.uniform(size=len(sel))), rng.uniform(size=len(sel))
    p = (1 - r1)[:, None] * a[t] + (r1 * (1 - r2))[:, None] * b[t] + (r1 * r2)[:, None] * c[t]
    nn = nrm[t]
    inward = np.sum(nn * (np.array([0.0, 0.0, 1000.0]) - p), 1) < 0
    nn[inward] *= -1
    p = p + nn * (91.25 + rng.uniform(0, 200, len(sel)))[:, None]
    for j, k in enumerate("xyz"):
        bv["pos"][k][sel] = p[:, j]
    batch.SetAllBallStates(balls)

    cars = batch.GetAllCarStates()
    cv = _view(cars)
    nc = batch.num_cars
    on_wall = np.nonzero(rng.uniform(size=nc) < 0.2)[0]
    kind = rng.integers(0, 3, len(on_wall))      # 0 side wall, 1 ceiling, 2 back wall
    side = rng.choice([-1.0, 1.0], len(on_wall))
    pos = np.zeros((len(on_wall), 3))
    up = np.zeros((len(on_wall), 3))
    w0 = kind == 0
    pos[w0] = np.stack([side[w0] * (4096 - 17), rng.uniform(-4000, 4000, w0.sum()), rng.uniform(300, 1700, w0.sum())], 1)
    up[w0] = np.stack([-side[w0], np.zeros(w0.sum()), np.zeros(w0.sum())], 1)
    w1 = kind == 1
    pos[w1] = np.stack([rng.uniform(-3800, 3800, w1.sum()), rng.uniform(-4800, 4800, w1.sum()),
                        np.full(w1.sum(), 2044 - 17)], 1)
    up[w1] = np.array([0.0, 0.0, -1.0])
    w2 = kind == 2
    xs = rng.uniform(1100, 3400, w2.sum()) * rng.choice([-1.0, 1.0], w2.sum())
    pos[w2] = np.stack([xs, side[w2] * (5120 - 17), rng.uniform(400, 1600, w2.sum())], 1)
    up[w2] = np.stack([np.zeros(w2.sum()), -side[w2], np.zeros(w2.sum())], 1)
    fwd, right = _frames(up, rng)
    for j, k in enumerate("xyz"):
        cv["pos"][k][on_wall] = pos[:, j]
        cv["rot"]["forward"][k][on_wall] = fwd[:, j]
        cv["rot"]["right"][k][on_wall] = right[:, j]
        cv["rot"]["up"][k][on_wall] = up[:, j]
    cv["is_on_ground"][on_wall] = 0

    overlap = np.nonzero(rng.uniform(size=n) < 0.1)[0]
    blue, orange = 2 * overlap, 2 * overlap + 1
    ang = rng.uniform(0, 2 * np.pi, len(overlap))
    rad = rng.uniform(0, 60, len(overlap))
    yaw = rng.uniform(-np.pi, np.pi, len(overlap))
    cv["pos"]["x"][orange] = cv["pos"]["x"][blue] + rad * np.cos(ang)
    cv["pos"]["y"][orange] = cv["pos"]["y"][blue] + rad * np.sin(ang)
    cv["pos"]["z"][orange] = cv["pos"]["z"][blue]
    for k, v in (("x", np.cos(yaw)), ("y", np.sin(yaw)), ("z", 0.0)):
        cv["rot"]["forward"][k][orange] = v
    for k, v in (("x", -np.sin(yaw)), ("y", np.cos(yaw)), ("z", 0.0)):
        cv["rot"]["right"][k][orange] = v
    for k, v in (("x", 0.0), ("y", 0.0), ("z", 1.0)):
        cv["rot"]["up"][k][orange] = v
    cv["is_on_ground"][orange] = cv["is_on_ground"][blue]
    batch.SetAllCarStates(cars)
