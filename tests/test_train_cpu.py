"""Pins of the CPU restatement of the GPU training path's rewards and reset
curriculum (oracle/train_spec.py, RocketSimCudaRewards.cuh:36-373,
RocketSimCudaApi.cuh:203-418) by closed forms and truth tables, then the
engine's own reward code (rsx_train.h, CPU build) against the restatement
on random records covering every reward id."""
import ctypes as C
import importlib.util
import math
import os
from types import SimpleNamespace as NS

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("train_spec", os.path.join(ROOT, "oracle", "train_spec.py"))
T = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(T)
F = np.float32


def v3(x=0.0, y=0.0, z=0.0):
    return NS(x=float(x), y=float(y), z=float(z))


def car(pos=(0, 0, 17), vel=(0, 0, 0), team=0, cid=1, fwd=(1, 0, 0), ground=True, flipping=False, boost=33.0,
        demoed=False, hit=None, contact=0, cd=0.0):
    return NS(pos=v3(*pos), vel=v3(*vel), team=team, id=cid, rot=NS(forward=v3(*fwd)), is_on_ground=ground,
              is_flipping=flipping, boost=boost, is_demoed=demoed, ball_hit_valid=hit is not None,
              ball_hit_tick=hit if hit is not None else 2 ** 64 - 1, car_contact_other_id=contact,
              car_contact_cooldown=cd)


def ball(pos=(0, 0, 93.15), vel=(0, 0, 0)):
    return NS(pos=v3(*pos), vel=v3(*vel))


def one(name, c, pc=None, b=None, pb=None, params=(), goal_team=-1, tick=1000, step=8, others=()):
    cars = [c] + list(others)
    prev = [pc or c] + list(others)
    return float(T.term({"id": name, "params": params}, cars, prev, b or ball(), pb or b or ball(), 0, goal_team,
                        tick, step))


def test_speed_velocity_air():
    assert one("SPEED", car(vel=(1150, 0, 0))) == 0.5
    assert one("VELOCITY", car(vel=(0, 2300, 0)), params=(2.0,)) == 2.0
    assert one("AIR", car(ground=False)) == 1.0 and one("AIR", car()) == 0.0


def test_directional_terms():
    b = ball(pos=(500, 0, 17))
    assert one("FACE_BALL", car(), b=b) == 1.0
    assert one("FACE_BALL", car(fwd=(-1, 0, 0)), b=b) == -1.0
    assert one("VELOCITY_PLAYER_TO_BALL", car(vel=(2300, 0, 0)), b=b) == 1.0
    assert abs(one("VELOCITY_PLAYER_TO_BALL", car(vel=(0, 2300, 0)), b=b)) < 1e-7
    # ball straight at the orange goal's centre height at max speed
    g = ball(pos=(0, 0, 321.3875), vel=(0, 6000, 0))
    assert one("VELOCITY_BALL_TO_GOAL", car(team=0), b=g) == pytest.approx(1.0, abs=1e-6)
    assert one("VELOCITY_BALL_TO_GOAL", car(team=1), b=g) == pytest.approx(-1.0, abs=1e-6)
    assert one("VELOCITY_BALL_TO_GOAL", car(team=0), b=g, params=(1.0,)) == pytest.approx(-1.0, abs=1e-6)
    # a zero direction normalizes to zero, not NaN
    assert one("FACE_BALL", car(pos=(0, 0, 93.15))) == 0.0


def test_touch_terms_and_the_step_boundary():
    tick, step = 1000, 8
    assert one("TOUCH_BALL", car(hit=tick - step), tick=tick, step=step) == 1.0
    assert one("TOUCH_BALL", car(hit=tick - step - 1), tick=tick, step=step) == 0.0
    mx = 110 * 100000 / 36000  # 305.55 uu/s: full reward speed
    b, pb = ball(vel=(mx / 2, 0, 0)), ball()
    assert one("TOUCH_ACCEL", car(hit=tick), b=b, pb=pb) == pytest.approx(0.5, rel=1e-6)
    assert one("TOUCH_ACCEL", car(hit=tick), b=pb, pb=b) == 0.0          # slowed down
    assert one("TOUCH_ACCEL", car(), b=b, pb=pb) == 0.0                  # no touch
    b2 = ball(vel=(300, 400, 0))                                         # change of 500 uu/s
    assert one("STRONG_TOUCH", car(hit=tick), b=b2, pb=pb, params=(400, 1000)) == 0.5
    assert one("STRONG_TOUCH", car(hit=tick), b=b2, pb=pb, params=(600, 1000)) == 0.0
    assert one("STRONG_TOUCH", car(hit=tick), b=b2, pb=pb, params=(0, 250)) == 1.0


def test_boost_and_wavedash():
    assert one("PICKUP_BOOST", car(boost=100), pc=car(boost=25)) == 0.5
    assert one("PICKUP_BOOST", car(boost=25), pc=car(boost=100)) == 0.0
    assert one("SAVE_BOOST", car(boost=25), params=(0.5,)) == 0.5
    assert one("SAVE_BOOST", car(boost=100), params=(3.0,)) == 1.0
    landed = car(ground=True)
    for prev_flip, prev_ground, want in ((True, False, 1.0), (True, True, 0.0), (False, False, 0.0)):
        assert one("WAVEDASH", landed, pc=car(flipping=prev_flip, ground=prev_ground)) == want
    assert one("WAVEDASH", car(ground=False), pc=car(flipping=True, ground=False)) == 0.0


def _bump_world(teammate=False, demo=False, old=False):
    """Car 1 (blue) bumps car 2 (orange, or blue when teammate) this step."""
    other_team = 0 if teammate else 1
    a = car(cid=1, team=0, contact=2, cd=0.25)
    b = car(cid=2, team=other_team, demoed=demo)
    pa = car(cid=1, team=0, contact=2 if old else 0, cd=0.3 if old else 0.0)
    pb = car(cid=2, team=other_team)
    return [a, b], [pa, pb]


@pytest.mark.parametrize("name,case,want", [
    ("BUMP", {}, [1, 0]), ("BUMP", {"teammate": True}, [0, 0]), ("BUMP", {"old": True}, [0, 0]),
    ("BUMPED_PENALTY", {}, [0, -1]), ("BUMPED_PENALTY", {"teammate": True}, [0, 0]),
    ("DEMO", {"demo": True}, [1, 0]), ("DEMO", {}, [0, 0]),
    ("DEMOED_PENALTY", {"demo": True}, [0, -1]),
    ("TEAMMATE_BUMP_PENALTY", {"teammate": True}, [-1, 0]), ("TEAMMATE_BUMP_PENALTY", {}, [0, 0]),
    ("TEAMMATE_BUMP_PENALTY", {"teammate": True, "old": True}, [0, 0]),
])
def test_contact_terms_truth_table(name, case, want):
    cars, prev = _bump_world(**case)
    got = [float(T.term({"id": name}, cars, prev, ball(), ball(), i, -1, 1000, 8)) for i in range(2)]
    assert got == want


def test_kickoff_proximity_and_goal():
    near, far = car(pos=(0, -1000, 17), cid=1, team=0), car(pos=(0, 2000, 17), cid=2, team=1)
    cars = [near, far]
    got = [float(T.term({"id": "KICKOFF_PROXIMITY"}, cars, cars, ball(), ball(), i, -1, 0, 8)) for i in range(2)]
    assert got == [1.0, -1.0]
    # equal distances: the tie test in RocketSimCudaRewards.cuh:185-188 sits
    # inside the strict-less branch and never fires, so both cars get -1
    tie = [car(pos=(0, -1000, 17), cid=1), car(pos=(0, 1000, 17), cid=2, team=1)]
    assert [float(T.term({"id": "KICKOFF_PROXIMITY"}, tie, tie, ball(), ball(), i, -1, 0, 8))
            for i in range(2)] == [-1, -1]
    off = ball(pos=(1.5, 0, 93.15))
    assert float(T.term({"id": "KICKOFF_PROXIMITY"}, cars, cars, off, off, 0, -1, 0, 8)) == 0.0
    g = [float(T.term({"id": "GOAL_REWARD", "params": [-0.5]}, cars, cars, ball(), ball(), i, 0, 0, 8))
         for i in range(2)]
    assert g == [1.0, -0.5]
    assert float(T.term({"id": "GOAL_REWARD"}, cars, cars, ball(), ball(), 0, -1, 0, 8)) == 0.0


def test_flick_carry_and_stack_terms():
    tick = 1000
    # mawkzy: ball sped up by 750 uu/s towards the orange goal on a flip touch
    c = car(hit=tick, flipping=True, ground=False)
    assert one("MAWKZY_FLICK", c, b=ball(vel=(0, 750, 0)), pb=ball()) == 0.5
    assert one("MAWKZY_FLICK", car(hit=tick), b=ball(vel=(0, 750, 0)), pb=ball()) == 0.0
    # pressure flick: ratio = 2920 / 2920 -> 1
    opp = car(pos=(0, 600, 17), cid=2, team=1)
    pc = car(pos=(0, 0, 17), vel=(0, 100, 0))
    pb = ball(pos=(0, 100, 93.15), vel=(0, 300, 0))
    got = one("PRESSURE_FLICK", car(hit=tick, flipping=True), pc=pc, b=ball(vel=(0, 2920, 0)), pb=pb, others=[opp])
    assert got == pytest.approx(1.0, rel=1e-6)
    far_opp = car(pos=(0, 900, 17), cid=2, team=1)
    assert one("PRESSURE_FLICK", car(hit=tick, flipping=True), pc=pc, b=ball(vel=(0, 2920, 0)), pb=pb,
               others=[far_opp]) == 0.0
    # air carry: close and slow relative to the ball, no touch
    cc = car(pos=(0, 0, 300), vel=(0, 600, 0), ground=False)
    assert one("AIR_CARRY", cc, b=ball(pos=(0, 0, 400), vel=(0, 0, 0))) == F(0.35)
    assert one("AIR_CARRY", cc, b=ball(pos=(0, 0, 400), vel=(0, 0, 600))) == pytest.approx(1 - math.sqrt(2) / 2,
                                                                                           rel=1e-6)
    assert one("AIR_CARRY", car(pos=(0, 0, 300), vel=(0, 1100, 0), ground=False),
               b=ball(pos=(0, 0, 400))) == pytest.approx(1 - 1100 / 1200, rel=1e-5)
    # stacked on the ball
    assert one("ANTI_BALL_STACK", car(pos=(50, 0, 200)), b=ball(pos=(0, 0, 93.15))) == -1.0
    assert one("ANTI_BALL_STACK", car(pos=(150, 0, 200)), b=ball(pos=(0, 0, 93.15))) == 0.0


@pytest.mark.parametrize("name", ["UNKNOWN", "PASS_INCENTIVE", "TEAM_PASSING", "MIXED_INITIATE", "MIXED_DOUBLETAP",
                                  "SPEED_AERIAL"])
def test_reserved_ids_score_zero(name):
    assert one(name, car(vel=(1000, 0, 0), hit=1000, ground=False)) == 0.0


def test_zero_sum_transform():
    """1v1, raw (1, 0): plain zero-sum gives (1, -1); 2v2 with team spirit s
    and opponent scale k gives raw(1 - s) + s * team_mean - k * opp_mean."""
    a, b = car(cid=1, team=0, vel=(2300, 0, 0)), car(cid=2, team=1)
    e = {"id": "SPEED", "weight": 2.0, "zero_sum": True}
    assert [float(x) for x in T.rewards([e], [a, b], [a, b], ball(), ball(), -1, 0, 8)] == [2.0, -2.0]
    cars = [car(cid=1, team=0, vel=(2300, 0, 0)), car(cid=2, team=0), car(cid=3, team=1, vel=(1150, 0, 0)),
            car(cid=4, team=1)]
    s, k = 0.25, 0.5
    e = {"id": "SPEED", "weight": 1.0, "zero_sum": True, "team_spirit": s, "opponent_scale": k}
    got = [float(x) for x in T.rewards([e], cars, cars, ball(), ball(), -1, 0, 8)]
    raw = [1.0, 0.0, 0.5, 0.0]
    tm = [0.5, 0.5, 0.25, 0.25]
    om = [0.25, 0.25, 0.5, 0.5]
    want = [raw[i] * (1 - s) + tm[i] * s - om[i] * k for i in range(4)]
    assert got == pytest.approx(want, abs=1e-7)
    # plain zero-sum over symmetric teams sums to zero
    e = {"id": "SPEED", "zero_sum": True}
    assert abs(sum(float(x) for x in T.rewards([e], cars, cars, ball(), ball(), -1, 0, 8))) < 1e-6
    # entries add in table order with their weights
    two = [{"id": "SPEED", "weight": 3.0}, {"id": "AIR", "weight": -1.0}]
    assert float(T.rewards(two, [a], [a], ball(), ball(), -1, 0, 8)[0]) == 3.0


def test_reset_checks_catch_misplacements():
    """The curriculum's property checks pass states built by the rules and
    fail each rule broken on its own."""
    import copy
    yaw = math.atan2(0 - 1000, 0 - 0)
    fwd = (math.cos(yaw), math.sin(yaw), 0)
    b = ball(pos=(0, 0, 93.15), vel=(100, -50, 20))
    good = [car(pos=(0, 1000, 17), vel=(300 * fwd[0], 300 * fwd[1], 0), fwd=fwd, boost=50, cid=1)]
    assert T.check_reset(3, good, b) == []
    for field, val in (("boost", 10.0), ("is_on_ground", False)):
        bad = copy.deepcopy(good)
        setattr(bad[0], field, val)
        assert T.check_reset(3, bad, b)
    bad = copy.deepcopy(good)
    bad[0].rot.forward = v3(1, 0, 0)
    assert T.check_reset(3, bad, b)
    bad = copy.deepcopy(good)
    bad[0].pos = v3(0, 2000, 17)
    assert T.check_reset(3, bad, b)
    kick = [car(pos=(-2048, -2560, 17), cid=1), car(pos=(2048, 2560, 17), cid=2, team=1)]
    assert T.check_reset(0, kick, ball()) == [] and T.reset_mode(kick, ball()) == 0
    moved = [car(pos=(-2048 + 30, -2560 - 30, 17), cid=1), car(pos=(2048, 2560, 17), cid=2, team=1)]
    assert T.check_reset(0, moved, ball()) and T.check_reset(1, moved, ball(pos=(10, 0, 93.15))) == []
    assert T.check_reset(1, moved, ball(pos=(500, 0, 93.15)))


@pytest.mark.parametrize("team", [1, 2, 3])
def test_engine_rewards_match_restatement_on_random_records(cpucheck, team):
    """The engine's reward code (CPU build of rsx_train.h) against the
    restatement: 300 random arenas x every reward table, every car, to float
    rounding (the engine stores Bullet units and converts back to uu)."""
    import train_harness as TH
    lib, world = cpucheck
    worst, nonzero = 0.0, 0
    for seed in range(300):
        cars, prev, b, pb = TH.records(seed, team, team)
        n = 2 * team
        goal_team = [-1, 0, 1][seed % 3]
        for table in TH.tables(seed):
            out = (C.c_float * n)()
            lib.cc_train_rewards(world, cars, prev, n, C.byref(b), C.byref(pb), TH.entries_array(table), len(table),
                                 goal_team, 5000, 8, out)
            want = T.rewards(table, list(cars), list(prev), b, pb, goal_team, 5000, 8)
            for i in range(n):
                err = abs(out[i] - float(want[i]))
                assert err <= 2e-5 * max(1.0, abs(float(want[i]))), (seed, i, out[i], float(want[i]), table)
                worst = max(worst, err)
                nonzero += abs(float(want[i])) > 0
    assert nonzero > 100
    print(f"{team}v{team}: worst |engine - restatement| {worst:.2e}")


def test_every_reward_id_fires_on_the_random_records():
    """Each id with a non-zero value in the reference gives a non-zero raw
    term on some random arena, so the comparison above covers its branch."""
    import train_harness as TH
    fired = set()
    for seed in range(400):
        cars, prev, b, pb = TH.records(seed, 2, 2)
        for rid in range(29):
            for i in range(4):
                if float(T.term({"id": rid, "params": [0.5, 800]}, list(cars), list(prev), b, pb, i, seed % 3 - 1,
                                5000, 8)) != 0:
                    fired.add(rid)
    reserved = {0, 20, 21, 22, 23, 24}
    assert fired == set(range(29)) - reserved, sorted(set(range(29)) - reserved - fired)


def test_curriculum_resets_of_the_cpu_build(cpucheck):
    """The engine's curriculum reset (CPU build): every reset passes its
    mode's placement rules, and the mode frequencies follow the weights."""
    import rsxlib as R
    lib, world = cpucheck
    w = (C.c_float * 4)(1.0, 2.0, 3.0, 4.0)
    counts = [0, 0, 0, 0]
    for seed in range(1, 2001):
        team = 1 + seed % 3
        cars = (R.CarState * (2 * team))()
        b = R.BallState()
        mode = lib.cc_curriculum_reset(world, team, team, seed * 2654435761 % (2 ** 32), w, cars, C.byref(b))
        counts[mode] += 1
        assert T.reset_mode(list(cars), b) == mode, (seed, mode)
        assert T.check_reset(mode, list(cars), b) == [], (seed, mode, T.check_reset(mode, list(cars), b))
    # multinomial(2000, (0.1, 0.2, 0.3, 0.4)): each count within 5 sigma
    for k, c in enumerate(counts):
        p = (k + 1) / 10
        assert abs(c - 2000 * p) < 5 * math.sqrt(2000 * p * (1 - p)), counts
    # kickoff-only weights never leave the kickoff placement
    w1 = (C.c_float * 4)(1.0, 0.0, 0.0, 0.0)
    for seed in range(1, 50):
        cars = (R.CarState * 2)()
        b = R.BallState()
        assert lib.cc_curriculum_reset(world, 1, 1, seed, w1, cars, C.byref(b)) == 0
        assert T.check_reset(0, list(cars), b) == []
