"""ORACLE / TEST INFRASTRUCTURE — CPU restatement (numpy float32 scalars, each
operation rounded as the reference's float code rounds it) of the GPU
training path's configurable rewards and property checks of its reset
curriculum.  Only tests use it; the engine (rsx_train.h) never does.

Rewards: training_eval_reward, one term per TrainingRewardID
(RocketSimCuda/src/RocketSimCudaRewards.cuh:36-287), summed over the table in
table order, a zero-sum entry's raw terms first turned into
raw * (1 - team_spirit) + team_mean * team_spirit - opponent_mean * opponent_scale
(buildRewardsAndTerminalsKernel, :320-355).  Inputs are the public records
(uu, uu/s) after the env step and at its start (SnapshotTrainingState,
RocketSimCuda.h:333).  The reference normalizes with rsqrtf; this restatement
divides by sqrtf, so the two agree to float rounding, not to the bit.

Curriculum: ResetArena's four placements (RocketSimCudaApi.cuh:268-375), as
checks a reset state must pass (`reset_mode`, `check_reset`): the draws
themselves come from a different random stream in the engine."""
import math

import numpy as np

F = np.float32

IDS = ["UNKNOWN", "GOAL_REWARD", "VELOCITY_BALL_TO_GOAL", "VELOCITY_PLAYER_TO_BALL", "FACE_BALL", "TOUCH_BALL",
       "TOUCH_ACCEL", "STRONG_TOUCH", "SPEED", "VELOCITY", "AIR", "WAVEDASH", "PICKUP_BOOST", "SAVE_BOOST", "BUMP",
       "BUMPED_PENALTY", "DEMO", "DEMOED_PENALTY", "KICKOFF_PROXIMITY", "TEAMMATE_BUMP_PENALTY", "PASS_INCENTIVE",
       "TEAM_PASSING", "MIXED_INITIATE", "MIXED_DOUBLETAP", "SPEED_AERIAL", "PRESSURE_FLICK", "MAWKZY_FLICK",
       "AIR_CARRY", "ANTI_BALL_STACK"]
ID = {n: i for i, n in enumerate(IDS)}

CAR_MAX_SPEED = F(2300)   # PhysConst::CAR_MAX_SPEED
BALL_MAX_SPEED = F(6000)  # PhysConst::BALL_MAX_SPEED


def vec(v):
    return np.array([v.x, v.y, v.z], dtype=F)


def dot(a, b):
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def length(v):
    return F(np.sqrt(dot(v, v)))


def safe_norm(v):
    """v3_safe_normalize (CudaMath.cuh:46-50): zero below |v|^2 = 1e-10."""
    l2 = dot(v, v)
    if l2 < F(1e-10):
        return np.zeros(3, F)
    return (v * F(F(1) / F(np.sqrt(l2)))).astype(F)


def dist2(a, b):
    d = (a - b).astype(F)
    return dot(d, d)


def touched(car, tick, step_ticks):
    """training_ball_touched_step (RocketSimCudaRewards.cuh:14-16)."""
    return bool(car.ball_hit_valid) and int(car.ball_hit_tick) >= int(tick) - int(step_ticks)


def new_contact(cur, prev):
    """training_is_new_contact (RocketSimCudaRewards.cuh:26-37)."""
    if not cur.car_contact_other_id or F(cur.car_contact_cooldown) <= 0:
        return False
    return (prev.car_contact_other_id != cur.car_contact_other_id or F(prev.car_contact_cooldown) <= 0
            or F(prev.car_contact_cooldown) < F(cur.car_contact_cooldown))


def _index_of(cars, cid):
    for i, c in enumerate(cars):
        if c.id == cid:
            return i
    return -1


def _entry(e):
    p = list(e.get("params", ())) + [0.0] * 10
    rid = ID[e["id"]] if isinstance(e["id"], str) else int(e["id"])
    return dict(id=rid, weight=F(e.get("weight", 1.0)), params=[F(x) for x in p[:10]],
                zero_sum=bool(e.get("zero_sum", False)), team_spirit=F(e.get("team_spirit", 0.0)),
                opponent_scale=F(e.get("opponent_scale", 1.0)))


def term(e, cars, prev_cars, ball, prev_ball, ci, goal_team, tick, step_ticks):
    """One raw reward term of car ci (training_eval_reward)."""
    e = _entry(e)
    c, pc = cars[ci], prev_cars[ci]
    cpos, cvel = vec(c.pos), vec(c.vel)
    bpos, bvel = vec(ball.pos), vec(ball.vel)
    tch = touched(c, tick, step_ticks)
    rid, prm = e["id"], e["params"]
    if rid == ID["GOAL_REWARD"]:
        if goal_team < 0:
            return F(0)
        return F(1) if c.team == goal_team else prm[0]
    if rid == ID["VELOCITY_BALL_TO_GOAL"]:
        orange = c.team == 0
        if prm[0] > F(0.5):
            orange = not orange
        ty = F(6000) if orange else F(-6000)
        d = safe_norm(np.array([-bpos[0], F(ty - bpos[1]), F(F(321.3875) - bpos[2])], F))
        return dot(d, (bvel / BALL_MAX_SPEED).astype(F))
    if rid == ID["VELOCITY_PLAYER_TO_BALL"]:
        return dot(safe_norm((bpos - cpos).astype(F)), (cvel / CAR_MAX_SPEED).astype(F))
    if rid == ID["FACE_BALL"]:
        return dot(vec(c.rot.forward), safe_norm((bpos - cpos).astype(F)))
    if rid == ID["TOUCH_BALL"]:
        return F(1) if tch else F(0)
    if rid == ID["TOUCH_ACCEL"]:
        if not tch:
            return F(0)
        mx = F(F(110) * F(F(100000) / F(36000)))
        pf = min(F(1), F(length(vec(prev_ball.vel)) / mx))
        cf = min(F(1), F(length(bvel) / mx))
        return F(cf - pf) if cf > pf else F(0)
    if rid == ID["STRONG_TOUCH"]:
        if not tch:
            return F(0)
        force = length((bvel - vec(prev_ball.vel)).astype(F))
        if force < prm[0]:
            return F(0)
        return min(F(1), F(force / max(prm[1], F(1e-6))))
    if rid == ID["SPEED"]:
        return F(length(cvel) / CAR_MAX_SPEED)
    if rid == ID["VELOCITY"]:
        return F(F(length(cvel) / CAR_MAX_SPEED) * prm[0])
    if rid == ID["AIR"]:
        return F(0) if c.is_on_ground else F(1)
    if rid == ID["WAVEDASH"]:
        return F(1) if (c.is_on_ground and pc.is_flipping and not pc.is_on_ground) else F(0)
    if rid == ID["PICKUP_BOOST"]:
        if F(c.boost) > F(pc.boost):
            return F(F(np.sqrt(F(F(c.boost) / F(100)))) - F(np.sqrt(F(F(pc.boost) / F(100)))))
        return F(0)
    if rid == ID["SAVE_BOOST"]:
        return F(min(max(F(np.power(F(F(c.boost) / F(100)), prm[0])), F(0)), F(1)))
    if rid == ID["BUMP"]:
        if not new_contact(c, pc):
            return F(0)
        o = _index_of(cars, c.car_contact_other_id)
        return F(0) if (o < 0 or cars[o].team == c.team) else F(1)
    if rid == ID["BUMPED_PENALTY"]:
        for o, oc in enumerate(cars):
            if o == ci or oc.team == c.team:
                continue
            if not new_contact(oc, prev_cars[o]):
                continue
            if oc.car_contact_other_id == c.id:
                return F(-1)
        return F(0)
    if rid == ID["DEMO"]:
        if not new_contact(c, pc):
            return F(0)
        o = _index_of(cars, c.car_contact_other_id)
        if o < 0 or cars[o].team == c.team:
            return F(0)
        return F(1) if (cars[o].is_demoed and not prev_cars[o].is_demoed) else F(0)
    if rid == ID["DEMOED_PENALTY"]:
        return F(-1) if (c.is_demoed and not pc.is_demoed) else F(0)
    if rid == ID["KICKOFF_PROXIMITY"]:
        if abs(bpos[0]) >= F(1) or abs(bpos[1]) >= F(1):
            return F(0)
        mine = dist2(cpos, bpos)
        opp, found = F(3.402823466e+38), False
        for oc in cars:
            if oc.id == c.id or oc.team == c.team:
                continue
            found = True
            opp = min(opp, dist2(vec(oc.pos), bpos))
        if not found or mine < opp:
            return F(0) if (found and mine == opp) else F(1)
        return F(-1)
    if rid == ID["TEAMMATE_BUMP_PENALTY"]:
        cid, cd = c.car_contact_other_id, F(c.car_contact_cooldown)
        if cd <= 0 or cid == 0:
            return F(0)
        pid, pcd = pc.car_contact_other_id, F(pc.car_contact_cooldown)
        if not ((cd > pcd) or (cid != pid and pcd <= 0)):
            return F(0)
        for oc in cars:
            if oc.id == cid:
                return F(-1) if oc.team == c.team else F(0)
        return F(0)
    if rid == ID["PRESSURE_FLICK"]:
        if not tch or not c.is_flipping:
            return F(0)
        if F(np.sqrt(dist2(vec(pc.pos), vec(prev_ball.pos)))) > F(250):
            return F(0)
        if abs(F(length(vec(pc.vel)) - length(vec(prev_ball.vel)))) > F(500):
            return F(0)
        closest = F(1e9)
        for oc in cars:
            if oc.team == c.team or oc.id == c.id:
                continue
            closest = min(closest, F(np.sqrt(dist2(cpos, vec(oc.pos)))))
        if closest > F(700):
            return F(0)
        ty = F(5120) if c.team == 0 else F(-5120)
        vg = dot(bvel, safe_norm(np.array([0, F(ty - bpos[1]), 0], F)))
        if vg <= F(1000):
            return F(0)
        ratio = F(vg / F(2920))
        return min(F(F(ratio * ratio) * F(np.sqrt(max(ratio, F(0))))), F(2))
    if rid == ID["MAWKZY_FLICK"]:
        if not tch or not c.is_flipping:
            return F(0)
        ty = F(5120) if c.team == 0 else F(-5120)
        acc = dot((bvel - vec(prev_ball.vel)).astype(F), safe_norm(np.array([0, F(ty - bpos[1]), 0], F)))
        return F(min(max(F(acc / F(1500)), F(0)), F(1)))
    if rid == ID["AIR_CARRY"]:
        if c.is_on_ground:
            return F(0)
        if tch:
            return F(1)
        if F(np.sqrt(dist2(cpos, bpos))) > F(300):
            return F(0)
        return F(min(max(F(F(1) - F(length((cvel - bvel).astype(F)) / F(1200))), F(0)), F(0.35)))
    if rid == ID["ANTI_BALL_STACK"]:
        dz = abs(F(cpos[2] - bpos[2]))
        dx, dy = F(cpos[0] - bpos[0]), F(cpos[1] - bpos[1])
        dxy = F(np.sqrt(F(F(dx * dx) + F(dy * dy))))
        return F(-1) if (dxy < F(120) and cpos[2] > bpos[2] and dz < F(180) and c.is_on_ground) else F(0)
    return F(0)  # UNKNOWN and the reserved team IDs (PASS_INCENTIVE .. SPEED_AERIAL)


def rewards(entries, cars, prev_cars, ball, prev_ball, goal_team, tick, step_ticks):
    """Every car's reward of one env step (the table loop of
    buildRewardsAndTerminalsKernel).  goal_team: -1, or the team that scored
    in this episode."""
    n = len(cars)
    total = [F(0)] * n
    for e in entries:
        e = _entry(e)
        raw = [term(e, cars, prev_cars, ball, prev_ball, i, goal_team, tick, step_ticks) for i in range(n)]
        for i in range(n):
            if e["zero_sum"]:
                ts, os_, tn, on = F(0), F(0), 0, 0
                for o in range(n):
                    if cars[o].team == cars[i].team:
                        ts, tn = F(ts + raw[o]), tn + 1
                    else:
                        os_, on = F(os_ + raw[o]), on + 1
                at = F(ts / F(tn)) if tn else F(0)
                ao = F(os_ / F(on)) if on else F(0)
                zs = F(F(F(raw[i] * F(F(1) - e["team_spirit"])) + F(at * e["team_spirit"])) - F(ao * e["opponent_scale"]))
                total[i] = F(total[i] + F(zs * e["weight"]))
            else:
                total[i] = F(total[i] + F(raw[i] * e["weight"]))
    return total


# ---- reset curriculum -------------------------------------------------------------

BALL_REST_Z = 93.15
CAR_REST_Z = 17.0
KICKOFF_XY = {(-2048.0, -2560.0), (2048.0, -2560.0), (-256.0, -3840.0), (256.0, -3840.0), (0.0, -4608.0)}


def _yaw_of(car):
    f = car.rot.forward
    return math.atan2(f.y, f.x)


def reset_mode(cars, ball):
    """Which of ResetArena's placements a reset state came from, told apart by
    the ball: 0 kickoff (at rest at centre), 1 fuzzed (near centre, moving),
    2 aerial (above 400 uu), 3 chase (on the ground, off centre or fast)."""
    b = ball.pos
    if b.x == 0 and b.y == 0 and abs(b.z - BALL_REST_Z) < 1e-3 and ball.vel.x == 0 and ball.vel.y == 0:
        return 0
    if b.z > 400:
        return 2
    fuzzed = abs(b.x) <= 180 and abs(b.y) <= 220 and abs(ball.vel.x) <= 90 and abs(ball.vel.y) <= 90 and \
        abs(ball.vel.z) <= 35
    if fuzzed and all(_near_kickoff(c, 40.5) for c in cars):
        return 1
    return 3


def _near_kickoff(car, tol):
    sgn = 1 if car.team == 0 else -1
    return any(abs(car.pos.x - sgn * x) <= tol and abs(car.pos.y - sgn * y) <= tol for x, y in KICKOFF_XY)


def check_reset(mode, cars, ball, tol=1e-2):
    """Violations of the placement rules of `mode` (ResetArena,
    RocketSimCudaApi.cuh:281-375) by a reset state; [] if it passes."""
    bad = []
    b, bv = ball.pos, ball.vel
    if mode == 0:
        for i, c in enumerate(cars):
            if not _near_kickoff(c, tol) or abs(c.pos.z - CAR_REST_Z) > tol:
                bad.append(f"car{i} not on a kickoff spot")
    elif mode == 1:
        if not (abs(b.x) <= 180 + tol and abs(b.y) <= 220 + tol and abs(bv.x) <= 90 + tol and abs(bv.y) <= 90 + tol
                and abs(bv.z) <= 35 + tol and abs(b.z - BALL_REST_Z) < tol):
            bad.append("ball outside the fuzzed-kickoff box")
        for i, c in enumerate(cars):
            if not _near_kickoff(c, 40 + tol):
                bad.append(f"car{i} more than 40 uu off a kickoff spot")
    elif mode == 2:
        if not (abs(b.x) <= 1400 + tol and abs(b.y) <= 1600 + tol and 420 - tol <= b.z <= 780 + tol):
            bad.append("ball outside the aerial box")
        if not (abs(bv.x) <= 100 + tol and abs(bv.y) <= 100 + tol and abs(bv.z) <= 50 + tol):
            bad.append("ball velocity outside the aerial range")
        setup = [i for i, c in enumerate(cars) if not c.is_on_ground]
        if len(setup) != 1:
            bad.append(f"{len(setup)} cars in the air, want 1")
        for i, c in enumerate(cars):
            if i in setup:
                dxy = math.hypot(b.x - c.pos.x, b.y - c.pos.y)
                if not (180 - 1 <= dxy <= 340 + 1):
                    bad.append(f"car{i} {dxy:.1f} uu from the ball, want 180-340")
                if not (c.pos.z < b.z - 40 + tol or abs(c.pos.z - (CAR_REST_Z + 80)) < tol):
                    bad.append(f"car{i} not below the ball")
                yaw_to_ball = math.atan2(b.y - c.pos.y, b.x - c.pos.x)
                if abs(math.remainder(_yaw_of(c) - yaw_to_ball, 2 * math.pi)) > 1e-3:
                    bad.append(f"car{i} not facing the ball")
                if not (55 - tol <= c.boost <= 100 + tol):
                    bad.append(f"car{i} boost {c.boost}")
            elif not _near_kickoff(c, tol):
                bad.append(f"car{i} not on a kickoff spot")
    else:
        if not (abs(b.x) <= 1900 + tol and abs(b.y) <= 2200 + tol and abs(b.z - BALL_REST_Z) < tol):
            bad.append("ball outside the chase box")
        if not (abs(bv.x) <= 500 + tol and abs(bv.y) <= 500 + tol and abs(bv.z) <= 80 + tol):
            bad.append("ball velocity outside the chase range")
        for i, c in enumerate(cars):
            d = math.hypot(b.x - c.pos.x, b.y - c.pos.y)
            clamped = abs(abs(c.pos.x) - 3500) < tol or abs(abs(c.pos.y) - 4500) < tol
            if not clamped and not (450 - 1 <= d <= 1850 + 1):
                bad.append(f"car{i} {d:.1f} uu from the ball, want 450-1850")
            if abs(c.pos.z - CAR_REST_Z) > tol or not c.is_on_ground:
                bad.append(f"car{i} not resting on the ground")
            yaw_to_ball = math.atan2(b.y - c.pos.y, b.x - c.pos.x)
            if abs(math.remainder(_yaw_of(c) - yaw_to_ball, 2 * math.pi)) > 1e-3:
                bad.append(f"car{i} not facing the ball")
            sp = math.hypot(c.vel.x, c.vel.y)
            if sp > 600 + tol or abs(c.vel.z) > tol:
                bad.append(f"car{i} speed {sp:.1f}")
            if sp > 1e-3 and abs(math.remainder(math.atan2(c.vel.y, c.vel.x) - yaw_to_ball, 2 * math.pi)) > 1e-3:
                bad.append(f"car{i} not rolling at the ball")
            if not (30 - tol <= c.boost <= 100 + tol):
                bad.append(f"car{i} boost {c.boost}")
    return bad
