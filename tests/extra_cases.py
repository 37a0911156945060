"""Arena cases for the sub-paths the shared scenarios leave out, run with the
workload protocol (workload_harness: warm tick, state loaded again) and
checked teacher-forced against the oracle:

  car resting on car   the blue car's wheel rays land on the orange car's box
                       (btVehicleRL::rayCast against a dynamic body,
                       btVehicleRL.cpp:118-214) and its suspension pushes the
                       car below (btVehicleRL.cpp:277-417)
  car drives on car    the same, with both cars driving
  car on ball          the blue car's wheels rest on the ball
  autoflip             a car on its roof presses jump: the auto-flip
                       (Car::_UpdateAutoFlip, Car.cpp:763-800)
  demolition respawn   a supersonic car demolishes the other, which respawns
                       after 3 s (Car.cpp:43-79, Car::_PreTickUpdate
                       Car.cpp:58-82)

The respawn spot is a random draw on both sides (RocketSim's global RNG, the
engine's per-arena xorshift), so at the respawn tick the respawned car's
position is checked against the four spots of its team instead of the
oracle's (workload_harness.teacher_forced_check); from the next tick on the
oracle is re-fed the engine's state, so the run stays comparable."""
import numpy as np

import parity_harness as P
import rsxlib as R
import workload_harness as WH


def _case(name, cars, ball, controls, ticks):
    pads = R.PADS()
    for i in range(34):
        pads[i].is_active = 1
    return dict(blue=1, orange=1, cars=cars, ball=ball, pads=pads, tick=1000, source=name, controls=controls,
                ticks=ticks)


def cases():
    out = []
    stack = lambda: [P.default_car(pos=(0, 0, 75), yaw=0.3, on_ground=False),  # noqa: E731
                     P.default_car(pos=(10, 5, 17), yaw=0.35, team=1)]
    out.append(_case("car resting on car", stack(), P.ball((2000, 2000, 93.15)), [P.controls(), P.controls()], 120))
    out.append(_case("car drives on car", stack(), P.ball((2000, 2000, 93.15)),
                     [P.controls(throttle=0.5, steer=0.4), P.controls(throttle=-0.3)], 120))
    out.append(_case("car on ball", [P.default_car(pos=(0, 0, 200), yaw=0.2, on_ground=False),
                                     P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)],
                     P.ball((0, 0, 93.15)), [P.controls(), P.controls()], 90))
    roof = np.diag([1, -1, -1]).astype(float)
    out.append(_case("autoflip", [P.car_from_rot((-1000, -1000, 40), roof, on_ground=False),
                                  P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)],
                     P.ball(), lambda t: [P.controls(jump=1 if 30 <= t < 40 else 0), P.controls()], 120))
    out.append(_case("demolition respawn", [P.default_car(pos=(0, -600, 17), yaw=np.pi / 2, vel=(0, 2250, 0), boost=100),
                                            P.default_car(pos=(30, 300, 17), yaw=0.3, team=1)],
                     P.ball((2000, 0, 93)), lambda t: [P.controls(throttle=1, boost=1 if t < 60 else 0), P.controls()],
                     420))
    return out


def respawns(eng):
    """(tick, car) of every respawn in a run (is_demoed 1 -> 0)."""
    out = []
    for t in range(1, len(eng.cars)):
        for i, (a, b) in enumerate(zip(eng.cars[t - 1], eng.cars[t])):
            if a.is_demoed and not b.is_demoed:
                out.append((t, i))
    return out


def check(oracle, cases_, engs, lut):
    """workload_harness.teacher_forced_check per case over its own ticks.
    Returns (failures, exempt, worst, respawns as (case, tick, car))."""
    failures, exempt, worst, seen = [], [], 0.0, []
    for case, eng in zip(cases_, engs):
        f, e, w = WH.teacher_forced_check(oracle, [case], [eng], case["ticks"], lut)
        failures += f
        exempt += e
        worst = max(worst, w)
        seen += [(case["source"], t, i) for t, i in respawns(eng)]
    return failures, exempt, worst, seen
