"""Runs the same arena case on the oracle (reference CPU RocketSim) and on an
engine build, and compares the results.

A case is a dict: {"blue": n, "orange": m, "tick": t, "cars": [CarState ...],
"ball": BallState, "pads": PADS or None, "controls": [Controls ...] or a
callable(tick_index) -> [Controls ...], "ticks": k}.

The oracle arena is first stepped once with its default kickoff state (so its
solver's time step is the 1/120 s it keeps for the rest of a run), then the
case's state is set and the per-wheel values that carry between ticks are
cleared on both sides."""
import ctypes as C

import numpy as np

import rsxlib as R

WHEEL_ZERO = (C.c_float * 48)()
WHEEL_ZERO6 = (C.c_float * 24)()


def default_car(pos=(0, 0, 17), yaw=0.0, vel=(0, 0, 0), ang=(0, 0, 0), boost=33.33, on_ground=True, team=0):
    s = R.CarState()
    s.pos.set(pos)
    cy, sy = np.cos(yaw), np.sin(yaw)
    s.rot.forward.set((cy, sy, 0))
    s.rot.right.set((-sy, cy, 0))
    s.rot.up.set((0, 0, 1))
    s.vel.set(vel)
    s.ang_vel.set(ang)
    s.boost = boost
    s.is_on_ground = 1 if on_ground else 0
    s.team = team
    s.ball_hit_tick = 2**64 - 1
    s.ball_hit_extra_impulse_tick = 2**64 - 1
    return s


def car_from_rot(pos, rot, vel=(0, 0, 0), ang=(0, 0, 0), boost=33.33, on_ground=False):
    """rot: 3x3 array whose columns are forward, right, up."""
    s = default_car(pos=pos, vel=vel, ang=ang, boost=boost, on_ground=on_ground)
    s.rot.forward.set(rot[:, 0])
    s.rot.right.set(rot[:, 1])
    s.rot.up.set(rot[:, 2])
    return s


def ball(pos=(0, 0, 93.15), vel=(0, 0, 0), ang=(0, 0, 0)):
    b = R.BallState()
    b.pos.set(pos)
    b.vel.set(vel)
    b.ang_vel.set(ang)
    return b


def controls(throttle=0., steer=0., pitch=0., yaw=0., roll=0., jump=0, boost=0, handbrake=0):
    c = R.Controls()
    c.throttle, c.steer, c.pitch, c.yaw, c.roll = throttle, steer, pitch, yaw, roll
    c.jump, c.boost, c.handbrake = jump, boost, handbrake
    return c


class Result:
    def __init__(self):
        self.cars = []     # per tick: list of CarState
        self.balls = []    # per tick: BallState
        self.pads = []
        self.scored = []
        self.contacts = []


def _ctrls(case, t):
    c = case["controls"]
    return c(t) if callable(c) else c


def run_oracle(lib, case):
    nb, no = case["blue"], case["orange"]
    h = C.c_void_p(lib.ref_arena_new(nb, no))
    try:
        lib.ref_step(h, 1)
        lib.ref_set_tick(h, case.get("tick", 1000))
        for i, s in enumerate(case["cars"]):
            lib.ref_set_car(h, i, C.byref(s))
            lib.ref_set_wheels(h, i, WHEEL_ZERO)
        lib.ref_set_ball(h, C.byref(case["ball"]))
        if case.get("pads") is not None:
            lib.ref_set_pads(h, case["pads"])
        else:
            p = R.PADS()
            for i in range(34):
                p[i].is_active = 1
            lib.ref_set_pads(h, p)
        res = Result()
        buf = (C.c_float * 4096)()
        for t in range(case["ticks"]):
            for i, c in enumerate(_ctrls(case, t)):
                lib.ref_set_controls(h, i, C.byref(c))
            lib.ref_step(h, 1)
            cars = []
            for i in range(nb + no):
                s = R.CarState()
                lib.ref_get_car(h, i, C.byref(s))
                cars.append(s)
            b = R.BallState()
            lib.ref_get_ball(h, C.byref(b))
            p = R.PADS()
            lib.ref_get_pads(h, p)
            res.cars.append(cars)
            res.balls.append(b)
            res.pads.append(p)
            res.scored.append(lib.ref_ball_scored(h))
            n = lib.ref_dump_contacts(h, buf, 4096)
            res.contacts.append(R.parse_contacts(buf, n))
        return res
    finally:
        lib.ref_arena_free(h)


def run_cpucheck(lib, world, case, boundary_every=0, fill=0):
    """The CPU build of the tick over the case; with boundary_every > 0 the
    arena goes through the SoA layout (pack, scribbled working state, load)
    every that many ticks, as at the start of each GPU launch."""
    nb, no = case["blue"], case["orange"]
    h = C.c_void_p(lib.cc_arena_new(world, nb, no))
    try:
        lib.cc_set_tick(h, case.get("tick", 1000))
        for i, s in enumerate(case["cars"]):
            lib.cc_set_car(h, i, C.byref(s))
            lib.cc_set_wheels(h, i, WHEEL_ZERO6)
        lib.cc_set_ball(h, C.byref(case["ball"]))
        p = case.get("pads")
        if p is None:
            p = R.PADS()
            for i in range(34):
                p[i].is_active = 1
        lib.cc_set_pads(h, p)
        res = Result()
        buf = (C.c_float * 4096)()
        for t in range(case["ticks"]):
            for i, c in enumerate(_ctrls(case, t)):
                lib.cc_set_controls(h, i, C.byref(c))
            if boundary_every and t % boundary_every == 0:
                lib.cc_soa_roundtrip(h, fill + t)
            sc = lib.cc_step(h, 1)
            cars = []
            for i in range(nb + no):
                s = R.CarState()
                lib.cc_get_car(h, i, C.byref(s))
                cars.append(s)
            b = R.BallState()
            lib.cc_get_ball(h, C.byref(b))
            pp = R.PADS()
            lib.cc_get_pads(h, pp)
            res.cars.append(cars)
            res.balls.append(b)
            res.pads.append(pp)
            res.scored.append(sc)
            n = lib.cc_dump_contacts(h, buf, 4096)
            res.contacts.append(R.parse_contacts(buf, n))
        return res
    finally:
        lib.cc_arena_free(h)


REL, ABS = 1e-4, 1e-3


def close(a, b, rel=REL, abs_=ABS):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.all(np.abs(a - b) <= abs_ + rel * np.abs(b))


def max_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b) / (ABS + REL * np.abs(b))))


def vec_err(a, b, abs_=ABS, rel=REL):
    """Error of a 3-vector quantity in units of its tolerance: the largest
    component difference over (abs + rel * |reference vector|)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / (abs_ + rel * np.linalg.norm(b)))


ROT_ABS = 1e-4  # rotation matrix entries are unitless


def phys_errs(av, bv, prefix, fails, rot=True):
    """av/bv: car_vec (18) or ball_vec (9) arrays; positions in uu, velocities
    in uu/s, angular velocities in rad/s are each checked as a vector within
    1e-3 absolute + 1e-4 relative to the reference vector's magnitude."""
    if len(av) == 18:
        parts = {"pos": (0, 3), "vel": (12, 15), "ang": (15, 18)}
    else:
        parts = {"pos": (0, 3), "vel": (3, 6), "ang": (6, 9)}
    for k, (lo, hi) in parts.items():
        e = vec_err(av[lo:hi], bv[lo:hi])
        if e > 1:
            fails[f"{prefix}.{k}"] = e
    if rot and len(av) == 18:
        e = float(np.max(np.abs(av[3:12] - bv[3:12]))) / ROT_ABS
        if e > 1:
            fails[f"{prefix}.rot"] = e


def compare_tick(ref, eng, t):
    """Returns a dict of failures for tick t (empty = parity).  Positions,
    velocities and angular velocities per vector within 1e-3 + 1e-4 *
    |reference|, rotation entries within 1e-4, continuous car scalars within
    the same relative bar; flags, wheel contacts, event timers and ids
    (CAR_EXACT_FIELDS), pad state and the goal flag bit for bit."""
    fails = {}
    for i, (rc, ec) in enumerate(zip(ref.cars[t], eng.cars[t])):
        phys_errs(R.car_vec(ec), R.car_vec(rc), f"car{i}", fails)
        for f in R.CAR_FLAG_FIELDS:
            if getattr(rc, f) != getattr(ec, f):
                fails[f"car{i}.{f}"] = (getattr(rc, f), getattr(ec, f))
        for f in R.CAR_FLOAT_FIELDS:
            if not close(getattr(ec, f), getattr(rc, f)):
                fails[f"car{i}.{f}"] = (getattr(rc, f), getattr(ec, f))
        for f in R.CAR_EXACT_FIELDS:
            if getattr(rc, f) != getattr(ec, f):
                fails[f"car{i}.{f}"] = (getattr(rc, f), getattr(ec, f))
        if list(rc.wheels_with_contact) != list(ec.wheels_with_contact):
            fails[f"car{i}.wheels"] = (list(rc.wheels_with_contact), list(ec.wheels_with_contact))
    phys_errs(R.ball_vec(eng.balls[t]), R.ball_vec(ref.balls[t]), "ball", fails)
    for p in range(34):
        a, b = ref.pads[t][p], eng.pads[t][p]
        # pad state is discrete: activity, the locking car and the cooldown
        # (BoostPad.cpp:51-105) agree bit for bit
        if a.is_active != b.is_active or a.prev_locked_car_id != b.prev_locked_car_id or a.cooldown != b.cooldown:
            fails[f"pad{p}"] = (a.is_active, b.is_active, a.cooldown, b.cooldown)
    if ref.scored[t] != eng.scored[t]:
        fails["scored"] = (ref.scored[t], eng.scored[t])
    return fails
