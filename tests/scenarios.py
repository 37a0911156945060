"""Arena cases shared by the CPU and GPU parity tests.  Each runs the same
1v1 arena (two Octanes) on the oracle and on an engine build; the names say
what part of the tick they exercise."""
import numpy as np

import parity_harness as P


def _far():
    return [P.default_car(pos=(-3000, -4000, 17), team=0), P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)]


def _idle():
    return [P.controls(), P.controls()]


def _jump(t):
    return [P.controls(jump=1 if t < 10 else 0),
            P.controls(jump=1 if (t < 5 or 20 < t < 25) else 0, pitch=1 if t > 20 else 0)]


def _flip(t):
    # first jump, release, then a forward dodge; orange does a side dodge
    return [P.controls(jump=1 if (t < 6 or 14 <= t < 18) else 0, pitch=-1 if t >= 14 else 0),
            P.controls(jump=1 if (t < 6 or 14 <= t < 18) else 0, yaw=1 if t >= 14 else 0)]


def _air(t):
    return [P.controls(pitch=np.sin(t * 0.1), yaw=0.5, roll=-0.3, boost=1 if t % 20 < 10 else 0),
            P.controls(pitch=-1, roll=1, boost=1)]


def cases():
    """name -> case dict (see parity_harness)."""
    c = {}
    c["ball free fall"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((0, 0, 1000), (100, 200, 300), (1, 2, 3)),
                               controls=_idle(), ticks=60)
    c["ball floor bounce"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((0, 0, 120), (300, 200, -900)),
                                  controls=_idle(), ticks=60)
    c["ball rest"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball(), controls=_idle(), ticks=30)
    c["cars idle on ground"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball(), controls=_idle(), ticks=60)
    c["cars throttle"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball(),
                              controls=[P.controls(throttle=1, steer=0.3), P.controls(throttle=-1, steer=-1)],
                              ticks=120)
    c["cars boost"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball(),
                           controls=[P.controls(throttle=1, boost=1), P.controls(throttle=1, handbrake=1, steer=1)],
                           ticks=120)
    c["jump"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball(), controls=_jump, ticks=120)
    # Cars start yawed off the axes: with a car exactly axis-aligned, the corner
    # that first touches the floor after a front flip is a tie that the
    # reference breaks by SSE rounding residue (~1e-10 in the basis).
    dodge = [P.default_car(pos=(-3000, -4000, 17), yaw=0.4, team=0),
             P.default_car(pos=(3000, 4000, 17), yaw=np.pi - 0.7, team=1)]
    c["dodge"] = dict(blue=1, orange=1, cars=dodge, ball=P.ball(), controls=_flip, ticks=90)
    air = [P.default_car(pos=(-1000, -1000, 800), vel=(300, 100, 200), on_ground=False),
           P.default_car(pos=(1000, 1000, 600), yaw=2.0, vel=(-200, 0, 400), ang=(1, 0.5, -0.5), on_ground=False,
                         team=1)]
    c["air control"] = dict(blue=1, orange=1, cars=air, ball=P.ball(), controls=_air, ticks=90)
    c["ball side wall"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((3900, 0, 500), (1500, 0, 100)),
                               controls=_idle(), ticks=60)
    c["ball back wall"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((1500, 4900, 500), (0, 1500, 100), (1, 0, 0)),
                               controls=_idle(), ticks=60)
    c["ball corner"] = dict(blue=1, orange=1, cars=_far(),
                            ball=P.ball((3700, 4700, 300), (1200, 1500, -100), (0, 0, 2)), controls=_idle(), ticks=90)
    c["ball ramp roll"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((3800, 0, 120), (800, 0, 0)),
                               controls=_idle(), ticks=90)
    c["ball ceiling"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((0, 0, 1900), (0, 0, 2000)),
                             controls=_idle(), ticks=30)
    c["ball goal"] = dict(blue=1, orange=1, cars=_far(), ball=P.ball((0, 4800, 300), (0, 2000, 0)),
                          controls=_idle(), ticks=60)
    hit = [P.default_car(pos=(0, -400, 17), yaw=np.pi / 2, vel=(0, 1500, 0), team=0),
           P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)]
    c["car hits ball"] = dict(blue=1, orange=1, cars=hit, ball=P.ball((0, 0, 93.15)),
                              controls=[P.controls(throttle=1), P.controls()], ticks=40)
    pads = [P.default_car(pos=(-3072, -4096 + 300, 17), yaw=-np.pi / 2, vel=(0, -800, 0), team=0),
            P.default_car(pos=(0, 4240 - 400, 17), yaw=np.pi / 2, vel=(0, 1000, 0), team=1, boost=10)]
    c["boost pads"] = dict(blue=1, orange=1, cars=pads, ball=P.ball(),
                           controls=[P.controls(throttle=1), P.controls(throttle=1)], ticks=60)
    # car box against the arena mesh (GJK / EPA per triangle, internal edges)
    other = P.default_car(pos=(3000, 4000, 17), yaw=np.pi, team=1)
    drive = [P.controls(throttle=1), P.controls()]
    c["car into back wall"] = dict(blue=1, orange=1, ball=P.ball(), controls=drive, ticks=90, cars=[
        P.default_car(pos=(1500, -4500, 17), yaw=-np.pi / 2 + 0.2, vel=(300, -1400, 0)), other])
    c["car up side ramp"] = dict(blue=1, orange=1, ball=P.ball(), controls=drive, ticks=90, cars=[
        P.default_car(pos=(3300, 0, 17), yaw=0.1, vel=(1500, 100, 0)), other])
    c["car falls on corner"] = dict(blue=1, orange=1, ball=P.ball(), controls=_idle(), ticks=90, cars=[
        P.car_from_rot((3700, -4700, 300), np.eye(3), vel=(200, -200, -300), ang=(1, 2, 0.5)), other])
    c["car roof on ramp"] = dict(blue=1, orange=1, ball=P.ball(), controls=_idle(), ticks=90, cars=[
        P.car_from_rot((3500, 1000, 200), np.diag([1, -1, -1]).astype(float), vel=(600, 0, -200)), other])
    # car against car (box-box manifold, bump and demolition callback)
    far_ball = P.ball((2000, 0, 93))
    c["head-on bump"] = dict(blue=1, orange=1, ball=far_ball, ticks=60,
                             controls=[P.controls(throttle=1), P.controls(throttle=1)], cars=[
        P.default_car(pos=(0, -300, 17), yaw=np.pi / 2, vel=(0, 1000, 0)),
        P.default_car(pos=(0, 300, 17), yaw=-np.pi / 2 + 0.1, vel=(0, -600, 0), team=1)])
    c["side bump"] = dict(blue=1, orange=1, ball=P.ball((2000, 2000, 93)), ticks=60,
                          controls=[P.controls(throttle=1), P.controls(throttle=1)], cars=[
        P.default_car(pos=(-400, 0, 17), yaw=0.0, vel=(1500, 0, 0)),
        P.default_car(pos=(100, 30, 17), yaw=np.pi / 2, vel=(0, 300, 0), team=1)])
    c["demolition"] = dict(blue=1, orange=1, ball=far_ball, ticks=60,
                           controls=[P.controls(throttle=1, boost=1), P.controls()], cars=[
        P.default_car(pos=(0, -600, 17), yaw=np.pi / 2, vel=(0, 2250, 0), boost=100),
        P.default_car(pos=(30, 300, 17), yaw=0.3, team=1)])
    return c
