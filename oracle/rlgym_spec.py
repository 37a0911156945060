"""ORACLE / TEST INFRASTRUCTURE — CPU restatement (numpy, float32) of the RLGym
pieces of an env step: DefaultObs rows, the four stub rewards, the terminal
conditions.  Only tests, the smoke check and the benchmark's CPU baseline may
use it; the engine never does.

The reference's own C++ for these pieces (GigaLearnCPP/RLGymCPP/src/RLGymCPP)
does not compile as shipped (CommonValues.h:2 includes "../Framework.h", a path
that does not exist), so this restatement is pinned by closed forms and
invariants in tests/test_spec_oracle.py instead of by a build of it.

Inputs are the public records (CarState / BallState / BoostPadState ctypes
structures, uu units) of one arena, in the arena's car order (the order cars
were added).  GameState::players follows Arena::_cars, which RocketSim visits
last-added first (the oracle's arenas pin that order, ref_car_order).

Reference lines followed:
  DefaultObs::BuildObs / AddPlayerToObs   ObsBuilders/DefaultObs.cpp:4-54, DefaultObs.h:11-17
  InvertPhys                              Gamestates/StateUtil.cpp:3-17
  pad order / inverted pad order          Gamestates/GameState.cpp:10-50, 96-109; CommonValues.h:45-80
  touched within the step                 Gamestates/Player.cpp:18-23
  HasFlipOrJump                           RocketSim/src/Sim/Car/Car.cpp:285-289
  Goal / Touch / VelBallToGoal / VelPlayerToBall
                                          RewardCore/CommonRewards.h:33-99
  GoalScoreCondition / NoTouchCondition   TerminalConditions/GoalScoreCondition.h, NoTouchCondition.h
  reward sum and terminal priority        EnvSet/EnvSet.cpp:169-242
  DefaultAction masks                     ActionParsers/DefaultAction.cpp:60-118
  AdvancedObs (GPU training layout)       RocketSimCuda/src/TrainingObs.cuh:20-127, fields of
                                          ObsBuilders/AdvancedObs.cpp:4-65
"""
import numpy as np

F = np.float32

SIDE_WALL_X, BACK_WALL_Y, CEILING_Z, BACK_NET_Y = F(4096), F(5120), F(2044), F(6000)
GOAL_HEIGHT = F(642.775)
CAR_MAX_SPEED, CAR_MAX_ANG_VEL, BALL_MAX_SPEED = F(2300), F(5.5), F(6000)
DOUBLEJUMP_MAX_DELAY = F(1.25)
GOAL_THRESHOLD_Y, BALL_RADIUS_RS = F(5124.25), F(91.25)  # RocketSim Arena::IsBallScored (Arena.cpp:949-957)

BOOST_LOCATIONS = np.array([
    (0, -4240, 70), (-1792, -4184, 70), (1792, -4184, 70), (-3072, -4096, 73), (3072, -4096, 73),
    (-940, -3308, 70), (940, -3308, 70), (0, -2816, 70), (-3584, -2484, 70), (3584, -2484, 70),
    (-1788, -2300, 70), (1788, -2300, 70), (-2048, -1036, 70), (0, -1024, 70), (2048, -1036, 70),
    (-3584, 0, 73), (-1024, 0, 70), (1024, 0, 70), (3584, 0, 73), (-2048, 1036, 70),
    (0, 1024, 70), (2048, 1036, 70), (-1788, 2300, 70), (1788, 2300, 70), (-3584, 2484, 70),
    (3584, 2484, 70), (0, 2816, 70), (-940, 3310, 70), (940, 3308, 70), (-3072, 4096, 73),
    (3072, 4096, 73), (-1792, 4184, 70), (1792, 4184, 70), (0, 4240, 70)], dtype=F)

# The arena's pad order (RocketSim Arena constructor: 6 big pads, then 28 small;
# RLConst.h BoostPads LOCS_BIG_SOCCAR / LOCS_SMALL_SOCCAR).
ARENA_PADS = np.array([
    (-3584, 0), (3584, 0), (-3072, 4096), (3072, 4096), (-3072, -4096), (3072, -4096),
    (0, -4240), (-1792, -4184), (1792, -4184), (-940, -3308), (940, -3308), (0, -2816), (-3584, -2484),
    (3584, -2484), (-1788, -2300), (1788, -2300), (-2048, -1036), (0, -1024), (2048, -1036), (-1024, 0),
    (1024, 0), (-2048, 1036), (0, 1024), (2048, 1036), (-1788, 2300), (1788, 2300), (-3584, 2484),
    (3584, 2484), (0, 2816), (-940, 3308), (940, 3308), (-1792, 4184), (1792, 4184), (0, 4240)], dtype=F)


def pad_index_map(arena_pads=ARENA_PADS):
    """BOOST_LOCATIONS slot i -> arena pad index (2D distance squared < 10)."""
    out = np.zeros(34, dtype=np.int64)
    for i, loc in enumerate(BOOST_LOCATIONS):
        d = ((arena_pads[:, 0] - loc[0]) ** 2 + (arena_pads[:, 1] - loc[1]) ** 2)
        out[i] = int(np.nonzero(d < 10)[0][0])
    return out


PAD_MAP = pad_index_map()
POS_COEF = np.array([F(1) / SIDE_WALL_X, F(1) / BACK_WALL_Y, F(1) / CEILING_Z], dtype=F)
VEL_COEF = F(1) / CAR_MAX_SPEED
ANG_COEF = F(1) / CAR_MAX_ANG_VEL
INV = np.array([-1, -1, 1], dtype=F)


def _v(v):
    return np.array([v.x, v.y, v.z], dtype=F)


def has_flip_or_jump(c):
    return bool(c.is_on_ground) or (not c.has_flipped and not c.has_double_jumped
                                    and F(c.air_time_since_jump) < DOUBLEJUMP_MAX_DELAY)


def player_block(c, inv):
    s = INV if inv else np.ones(3, dtype=F)
    pos, fwd, up = _v(c.pos) * s, _v(c.rot.forward) * s, _v(c.rot.up) * s
    vel, ang = _v(c.vel) * s, _v(c.ang_vel) * s
    return np.concatenate([pos * POS_COEF, fwd, up, vel * VEL_COEF, ang * ANG_COEF,
                           np.array([F(c.boost) / F(100), float(c.is_on_ground), float(has_flip_or_jump(c)),
                                     float(c.is_demoed)], dtype=F)]).astype(F)


def default_obs(cars, ball, pads, prev_actions, self_idx):
    """DefaultObs row of car self_idx.  prev_actions: (ncars, 8) throttle,
    steer, pitch, yaw, roll, jump, boost, handbrake."""
    me = cars[self_idx]
    inv = me.team != 0
    s = INV if inv else np.ones(3, dtype=F)
    row = [_v(ball.pos) * s * POS_COEF, _v(ball.vel) * s * VEL_COEF, _v(ball.ang_vel) * s * ANG_COEF,
           np.asarray(prev_actions[self_idx], dtype=F)]
    slots = PAD_MAP[::-1] if inv else PAD_MAP
    row.append(np.array([float(pads[j].is_active) for j in slots], dtype=F))
    row.append(player_block(me, inv))
    # other players in GameState::players order, which is Arena::_cars order:
    # last-added car first (GameState.cpp:68-78; DefaultObs.cpp:36-50)
    others = [(j, c) for j, c in enumerate(cars) if j != self_idx][::-1]
    mates = [player_block(c, inv) for j, c in others if c.team == me.team]
    opps = [player_block(c, inv) for j, c in others if c.team != me.team]
    return np.concatenate(row + mates + opps).astype(F)


def ball_scored(ball):
    return abs(F(ball.pos.y)) > GOAL_THRESHOLD_Y + BALL_RADIUS_RS


def touched_in_step(c, tick, tick_skip):
    """Player::UpdateFromCar (Player.cpp:18-23): a hit counts for the step when
    tickCountWhenHit >= tickCount - tickSkip, in uint64 arithmetic (the
    subtraction wraps when tickCount < tickSkip)."""
    return bool(c.ball_hit_valid) and c.ball_hit_tick >= (tick - tick_skip) % (1 << 64)


def _normalized(v):
    n = np.sqrt(F(np.dot(v, v)))
    return (v / n).astype(F) if n > F(1.1920929e-07) ** 2 else np.zeros(3, dtype=F)


def stub_rewards(cars, ball, goal, touched, w):
    """w: dict goal, concede, touch, vbtg, vptb.  Returns float32 per car."""
    bp, bv = _v(ball.pos), _v(ball.vel)
    out = []
    for c, t in zip(cars, touched):
        r = F(0)
        if w["goal"]:
            g = F(0)
            if goal:
                g = F(1) if c.team != (0 if bp[1] < 0 else 1) else F(w["concede"])
            r += g * F(w["goal"])
        if w["touch"]:
            r += F(float(t)) * F(w["touch"])
        if w["vbtg"]:
            target = np.array([0, BACK_NET_Y if c.team == 0 else -BACK_NET_Y, GOAL_HEIGHT / F(2)], dtype=F)
            r += F(np.dot(_normalized(target - bp), bv / BALL_MAX_SPEED)) * F(w["vbtg"])
        if w["vptb"]:
            r += F(np.dot(_normalized(bp - _v(c.pos)), _v(c.vel) / CAR_MAX_SPEED)) * F(w["vptb"])
        out.append(r)
    return np.array(out, dtype=F)


def step_dt(tick_skip):
    """GameState::deltaTime = tickSkip * (1 / 120.f) (GameState.cpp:64), float32."""
    return F(F(tick_skip) * (F(1) / F(120)))


def step_terminal(goal, any_touch, tick_skip, no_touch, timeout, goal_terminal=True):
    """Returns (terminal type, new no-touch timer): 0 none, 1 normal, 2
    truncated.  NoTouchCondition::IsTerminal (NoTouchCondition.h:18-28): the
    float32 timer is cleared by any touch in the step, otherwise advanced by
    the step's deltaTime and compared with maxTime (>=)."""
    t = 0
    if goal_terminal and goal:
        t = 1
    if timeout > 0:
        if any_touch:
            no_touch = F(0)
        else:
            no_touch = F(F(no_touch) + step_dt(tick_skip))
            if no_touch >= F(timeout) and t == 0:
                t = 2
    return t, no_touch


def default_action_masks(lut, ground_actions=24):
    """The four masks DefaultAction's constructor builds over the table `lut`
    (rows of throttle, steer, pitch, yaw, roll, jump, boost, handbrake):
    (ground, air, boost, jump), each a list of 0/1 per entry."""
    n = len(lut)
    ground, air, boost, jump = [0] * n, [0] * n, [0] * n, [0] * n
    for i, (thr, steer, pitch, yaw, roll, j, b, hb) in enumerate(lut):
        jump[i] = int(bool(j))
        boost[i] = int(bool(b))
        ground[i] = int(i < ground_actions)
        if i > ground_actions and not j:
            air[i] = 1
        if i < ground_actions and thr == b and float(yaw != 0) == hb:
            air[i] = 1
    return ground, air, boost, jump


def action_mask(c, masks):
    """DefaultAction::GetActionMask for car record c."""
    ground, air, boost, jump = masks
    out = list(ground if c.is_on_ground else air)
    if c.boost == 0:
        out = [m & (1 - bm) for m, bm in zip(out, boost)]
    turtled = bool(c.world_contact) and F(c.world_contact_normal.z) > F(0.9)
    if has_flip_or_jump(c) or turtled:
        out = [m | jm for m, jm in zip(out, jump)]
    return out


ADV_POS, ADV_VEL, ADV_ANG = F(1) / F(5000), F(1) / F(2300), F(1) / F(3)


def _inv(v, inv):
    return (v * INV).astype(F) if inv else v


def _dot3(a, b):
    """float32 x*x' + y*y' + z*z', summed left to right (v3_dot)."""
    return F(F(F(a[0] * b[0]) + F(a[1] * b[1])) + F(a[2] * b[2]))


def _rot_dot(fwd, right, up, v):
    """rotmat_dot_vec: v in the (forward, right, up) frame, float32."""
    return np.array([_dot3(fwd, v), _dot3(right, v), _dot3(up, v)], dtype=F)


def adv_player_block(c, inv, ball_pos, ball_vel):
    """The 29 floats of one car (add_advanced_obs_player, TrainingObs.cuh:52-79)."""
    pos, vel, ang = _inv(_v(c.pos), inv), _inv(_v(c.vel), inv), _inv(_v(c.ang_vel), inv)
    fwd, right, up = _inv(_v(c.rot.forward), inv), _inv(_v(c.rot.right), inv), _inv(_v(c.rot.up), inv)
    return np.concatenate([pos * ADV_POS, fwd, up, vel * ADV_VEL, ang * ADV_ANG,
                           _rot_dot(fwd, right, up, ang) * ADV_ANG,
                           _rot_dot(fwd, right, up, (ball_pos - pos).astype(F)) * ADV_POS,
                           _rot_dot(fwd, right, up, (ball_vel - vel).astype(F)) * ADV_VEL,
                           np.array([F(c.boost) / F(100), float(c.is_on_ground), float(has_flip_or_jump(c)),
                                     float(c.is_demoed), float(c.has_jumped)], dtype=F)]).astype(F)


def advanced_obs(cars, ball, pads, self_idx):
    """AdvancedObs row of car self_idx as the GPU training path builds it
    (device_build_advanced_obs, TrainingObs.cuh:85-125): ball, the car's last
    controls, per pad 1 when active else 1 / (1 + cooldown), then the car,
    its teammates and its opponents in slot order.  51 + 29 * cars floats."""
    me = cars[self_idx]
    inv = me.team != 0
    bp, bv, ba = _inv(_v(ball.pos), inv), _inv(_v(ball.vel), inv), _inv(_v(ball.ang_vel), inv)
    lc = me.last_controls
    row = [bp * ADV_POS, bv * ADV_VEL, ba * ADV_ANG,
           np.array([lc.throttle, lc.steer, lc.pitch, lc.yaw, lc.roll, float(bool(lc.jump)), float(bool(lc.boost)),
                     float(bool(lc.handbrake))], dtype=F)]
    slots = PAD_MAP[::-1] if inv else PAD_MAP
    row.append(np.array([F(1) if pads[j].is_active else F(F(1) / (F(1) + F(pads[j].cooldown))) for j in slots],
                        dtype=F))
    row.append(adv_player_block(me, inv, bp, bv))
    row += [adv_player_block(c, inv, bp, bv) for j, c in enumerate(cars) if j != self_idx and c.team == me.team]
    row += [adv_player_block(c, inv, bp, bv) for j, c in enumerate(cars) if j != self_idx and c.team != me.team]
    return np.concatenate(row).astype(F)
