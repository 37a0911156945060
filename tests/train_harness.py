"""Random reward tables and random public records that reach every branch
of the training rewards (contacts, demos, touches at the step boundary, the
ball at centre, flicks near the ball, cars stacked on it), for checking the
engine's reward code against oracle/train_spec.py."""
import numpy as np

import rsxlib as R

E = R.engine
N_IDS = 29


def tables(seed):
    """Reward tables that together hold every reward id, with random weights,
    parameters and zero-sum settings."""
    rng = np.random.default_rng(seed)
    ids = list(rng.permutation(N_IDS))
    out = []
    while ids:
        k = min(16, len(ids))
        t = []
        for rid in ids[:k]:
            params = [float(rng.uniform(0, 1.5)), float(rng.uniform(200, 2000))] + [0.0] * 8
            if rid == 7:  # STRONG_TOUCH: least change, full-reward change
                params[:2] = [float(rng.uniform(0, 400)), float(rng.uniform(100, 1500))]
            zs = bool(rng.uniform() < 0.4)
            t.append(dict(id=int(rid), weight=float(rng.uniform(-3, 3)), params=params, zero_sum=zs,
                          team_spirit=float(rng.uniform(0, 1)) if zs else 0.0,
                          opponent_scale=float(rng.uniform(0, 1.5)) if zs else 1.0))
        out.append(t)
        ids = ids[k:]
    return out


def entries_array(table):
    arr = (E.TrainingRewardEntry * 16)()
    for k, e in enumerate(table):
        p = (list(e.get("params", ())) + [0.0] * 10)[:10]
        arr[k] = E.TrainingRewardEntry(e["id"], e["weight"], (R.C.c_float * 10)(*p), int(e["zero_sum"]),
                                       e["team_spirit"], e["opponent_scale"])
    return arr


def _car(rng, i, team, ids, tick, near=None):
    s = R.CarState()
    p = rng.uniform([-3000, -4000, 17], [3000, 4000, 600]) if near is None else near + rng.normal(0, 120, 3)
    s.pos.set(p)
    yaw = rng.uniform(-np.pi, np.pi)
    s.rot.forward.set((np.cos(yaw), np.sin(yaw), 0))
    s.rot.right.set((-np.sin(yaw), np.cos(yaw), 0))
    s.rot.up.set((0, 0, 1))
    s.vel.set(rng.normal(0, 900, 3))
    s.boost = float(rng.choice([0.0, float(rng.uniform(0, 100)), 100.0]))
    s.is_on_ground = int(rng.uniform() < 0.5)
    s.is_flipping = int(rng.uniform() < 0.4)
    s.is_demoed = int(rng.uniform() < 0.2)
    s.team = team
    s.id = ids[i]
    s.car_contact_other_id = int(rng.choice([0] + list(ids)))
    s.car_contact_cooldown = float(rng.choice([0.0, 0.1, 0.25, 0.3]))
    if rng.uniform() < 0.6:
        s.ball_hit_valid = 1
        s.ball_hit_tick = int(tick - 8 + rng.integers(-2, 3))
    else:
        s.ball_hit_valid = 0
        s.ball_hit_tick = 2 ** 64 - 1
    return s


def records(seed, n_blue, n_orange, tick=5000):
    """(cars, prev_cars, ball, prev_ball) for one arena: ctypes arrays of
    public records, ids 1..n in blue-first order."""
    rng = np.random.default_rng(seed)
    n = n_blue + n_orange
    ids = list(range(1, n + 1))
    teams = [0] * n_blue + [1] * n_orange
    ball, prev_ball = R.BallState(), R.BallState()
    centre = rng.uniform() < 0.25
    bp = np.array([rng.uniform(-0.9, 0.9), rng.uniform(-0.9, 0.9), 93.15]) if centre else \
        rng.uniform([-3000, -4000, 93.15], [3000, 4000, 1200])
    ball.pos.set(bp)
    ball.vel.set(rng.normal(0, 1500, 3) + [0, rng.choice([-2500, 0, 2500]), 0])
    prev_ball.pos.set(bp + rng.normal(0, 60, 3))
    prev_ball.vel.set(rng.normal(0, 800, 3))
    cars = (R.CarState * n)()
    prev = (R.CarState * n)()
    for i in range(n):
        near = bp if rng.uniform() < 0.5 else None
        cars[i] = _car(rng, i, teams[i], ids, tick, near)
        p = _car(rng, i, teams[i], ids, tick - 8, near)
        if rng.uniform() < 0.5:  # the same bump as at the start, or a cooldown that only ran down
            p.car_contact_other_id = cars[i].car_contact_other_id
            p.car_contact_cooldown = cars[i].car_contact_cooldown + float(rng.choice([0.0, 0.05, -0.05]))
        if rng.uniform() < 0.5:
            p.pos = prev_ball.pos
            p.pos.z = prev_ball.pos.z - 50
        cars[i].is_on_ground = int(cars[i].is_on_ground or (rng.uniform() < 0.3 and cars[i].pos.z > bp[2]))
        prev[i] = p
    return cars, prev, ball, prev_ball
