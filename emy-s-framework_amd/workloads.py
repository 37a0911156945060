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
    todo = np.arange(nc)
    while todo.size:
        pos[todo] = np.stack([rng.uniform(-3900, 3900, todo.size), rng.uniform(-4900, 4900, todo.size)], 1)
        bad = np.linalg.norm(pos[todo] - bpos[arena_of[todo], :2], axis=1) < 300
        # cars of one arena keep 300 uu apart (checked against earlier slots)
        first = np.concatenate([[0], np.cumsum([2 * t for t in batch.team_size])[:-1]])
        for i in todo[~bad]:
            a = arena_of[i]
            for j in range(first[a], i):
                if np.linalg.norm(pos[i] - pos[j]) < 300:
                    bad[np.searchsorted(todo, i)] = True
                    break
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
