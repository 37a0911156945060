"""ORACLE / TEST INFRASTRUCTURE — CPU restatement (numpy, float32) of the
VantaObs v3.5 observation as the reference's GPU path builds it.  Only tests
may use it; the engine never does.

The reference's definition is its GPU transcription,
RocketSimCuda/src/VantaObsGpu.cuh:1-750, launched by vantaArenaUpdateKernel /
vantaObsKernel (RocketSimCuda/src/RocketSimCudaObs.cuh:70-186) from
RocketSimCudaBatch::BuildVantaObs (RocketSimCudaApi.cuh:597-640).  The CPU
builder it transcribes (RLGymCPP/OBSBuilders/VantaObs.h) is not part of the
reference tree, so this restatement follows the GPU file and is pinned by
closed forms, symmetries and hand-derived rows in tests/test_vanta_cpu.py; the
ball-prediction slices are pinned against the compiled oracle (a ball-only
RocketSim arena stepped 120 / 240 / 360 ticks, the BallPredTracker arena of
BallPredTracker.cpp:5-73).

Inputs are the public records of one arena (CarState / BallState /
BoostPadState ctypes structures, uu units) in slot order, the arena's tick
count and the prediction slices.  Every arithmetic step is float32, in the
reference's order (Vec3 / s is a multiplication by 1 / s, CudaMath.cuh:24).

Reference lines followed:
  constants, row layout                 VantaObsGpu.cuh:22-43
  persistent state                      VantaObsGpu.cuh:47-81
  InvertPhys / MirrorPhysX / canonical  VantaObsGpu.cuh:98-140
  HasFlipOrJump / flip reset flags      VantaObsGpu.cuh:152-164
  small-pad influence and density       VantaObsGpu.cuh:167-188
  UpdateBoostPads / UpdateDemoTimers    VantaObsGpu.cuh:192-243
  coyote timers (UpdateAdditionalTimers) VantaObsGpu.cuh:246-338
  obs row (sections A-H)                VantaObsGpu.cuh:342-747
  last-touch car, empty slots           RocketSimCudaObs.cuh:145-186
  pad tables (locations, mirror map)    RocketSimCudaHostUtils.cuh:55-90
  reset clears the state                RocketSimCudaApi.cuh:265-274
"""
import importlib.util
import os

import numpy as np

F = np.float32

POS_STD = F(2300)
VEL_STD = F(2300)
BALL_VEL_STD = F(6000)
ANG_STD = F(5.5)
DEMO_TIMER_STD = F(3)
GOAL_DIST_COEF = F(1) / F(10240)
EGO_SIZE, BALL_SIZE, GOAL_SIZE, TEAM_CTX_SIZE, CAR_SIZE = 81, 29, 16, 6, 39
MAX_TEAMMATES, MAX_OPPONENTS, BOOST_SIZE = 2, 3, 34
MASK_SIZE = MAX_TEAMMATES + MAX_OPPONENTS
OBS_SIZE = EGO_SIZE + BALL_SIZE + GOAL_SIZE + TEAM_CTX_SIZE + CAR_SIZE * (MAX_TEAMMATES + MAX_OPPONENTS) \
    + MASK_SIZE + BOOST_SIZE
DOUBLEJUMP_MAX_DELAY = F(1.25)
FLT_MAX = F(3.402823466e+38)
PRED_TICKS = (120, 240, 360)


def _rlgym():
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "rlgym_spec.py")
    spec = importlib.util.spec_from_file_location("rlgym_spec_for_vanta", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


_R = _rlgym()


def pad_tables():
    """(pad_map, locations, mirror): RLGym pad slot -> arena pad index; each
    slot's location (arena pad positions, z 73 big / 70 small); the slot of
    its x-mirrored pad (RocketSimCudaHostUtils.cuh:68-86)."""
    pad_map = _R.PAD_MAP.copy()
    big = {(-3584, 0), (3584, 0), (-3072, 4096), (3072, 4096), (-3072, -4096), (3072, -4096)}
    locs = np.zeros((BOOST_SIZE, 3), dtype=F)
    for i in range(BOOST_SIZE):
        x, y = _R.ARENA_PADS[pad_map[i]]
        locs[i] = (x, y, 73 if (int(x), int(y)) in big else 70)
    mirror = np.arange(BOOST_SIZE)
    for i in range(BOOST_SIZE):
        for j in range(BOOST_SIZE):
            if abs(locs[j, 0] + locs[i, 0]) < 1 and abs(locs[j, 1] - locs[i, 1]) < 1 \
                    and abs(locs[j, 2] - locs[i, 2]) < 1:
                mirror[i] = j
                break
    return pad_map, locs, mirror


PAD_MAP, PAD_LOCS, PAD_MIRROR = pad_tables()


def v3(x, y, z):
    return np.array([x, y, z], dtype=F)


def vec(v):
    return v3(v.x, v.y, v.z)


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def length(v):
    return np.sqrt(dot(v, v))


def vdiv(v, s):
    """Vec3 / float: multiplication by the reciprocal (CudaMath.cuh:24)."""
    inv = F(1) / F(s)
    return v * inv


def safe_norm(v):
    n = length(v)
    return vdiv(v, n) if n > F(1e-6) else v3(0, 0, 0)


class Phys:
    """pos, vel, angVel and the (forward, right, up) basis (VantaPhys)."""

    def __init__(self, pos, vel, ang, fwd, right, up):
        self.pos, self.vel, self.ang = pos.copy(), vel.copy(), ang.copy()
        self.fwd, self.right, self.up = fwd.copy(), right.copy(), up.copy()

    @staticmethod
    def of_car(c):
        return Phys(vec(c.pos), vec(c.vel), vec(c.ang_vel), vec(c.rot.forward), vec(c.rot.right), vec(c.rot.up))

    @staticmethod
    def of_ball(pos, vel, ang):
        return Phys(pos, vel, ang, v3(1, 0, 0), v3(0, 1, 0), v3(0, 0, 1))

    def rot_dot(self, v):
        """rotMat.Dot(v): world -> body frame."""
        return v3(dot(self.fwd, v), dot(self.right, v), dot(self.up, v))


INV = v3(-1, -1, 1)


def invert(p, inv):
    if not inv:
        return p
    return Phys(p.pos * INV, p.vel * INV, p.ang * INV, p.fwd * INV, p.right * INV, p.up * INV)


def mirror_x(p):
    q = Phys(p.pos, p.vel, p.ang, p.fwd, p.right, p.up)
    q.pos[0] = -q.pos[0]
    q.vel[0] = -q.vel[0]
    q.ang[1] = -q.ang[1]
    q.ang[2] = -q.ang[2]
    q.fwd[0] = -q.fwd[0]
    q.right[0] = -q.right[0]
    q.up[0] = -q.up[0]
    return q


def canonical(p, inv, mx):
    q = invert(p, inv)
    return mirror_x(q) if mx else q


def canonical_normal(n, inv, mx):
    n = n.copy()
    if inv:
        n[0] = -n[0]
        n[1] = -n[1]
    if mx:
        n[0] = -n[0]
    return n


def has_flip_or_jump(c):
    return bool(c.is_on_ground) or (not c.has_flipped and not c.has_double_jumped
                                    and F(c.air_time_since_jump) < DOUBLEJUMP_MAX_DELAY)


def has_flip_reset(c):
    return (not c.is_on_ground) and has_flip_or_jump(c) and not c.has_jumped


def got_flip_reset(c):
    return (not c.is_on_ground) and not c.has_jumped


def small_pad_stats(arena_state, pos):
    max_inf, density = F(0), F(0)
    for i in range(BOOST_SIZE):
        if PAD_LOCS[i, 2] < F(72) and arena_state.avail[i]:
            dist = length(PAD_LOCS[i] - pos)
            inf = F(1) / (F(1) + dist / F(500))
            max_inf = max(max_inf, inf)  # fmaxf: neither operand is NaN here
            if dist < F(2000):
                density = density + (F(1) - dist / F(2000))
    return F(max_inf), density / F(4)


class Timers:
    """GpuVantaTimers (VantaObsGpu.cuh:47-67), all zero."""

    def __init__(self):
        self.boost_time = self.boost_hold = self.jump_time = self.flip_time = F(0)
        self.air_time = self.handbrake_time = self.demo_timer = F(0)
        self.is_jumping = self.has_jumped = self.has_double_jumped = self.has_flipped = False
        self.on_ground = self.was_demoed = self.init = False
        self.flip_dir = v3(0, 0, 0)
        self.pp = dict(throttle=F(0), steer=F(0), pitch=F(0), yaw=F(0), roll=F(0), jump=False, boost=False,
                       handbrake=False)
        self.last_tick = None  # ~0ULL

    def copy(self):
        t = Timers()
        t.__dict__.update({k: (v.copy() if isinstance(v, (np.ndarray, dict)) else v) for k, v in self.__dict__.items()})
        return t


class ArenaState:
    """GpuVantaArenaState (VantaObsGpu.cuh:69-81) plus the arena's per-car timers."""

    def __init__(self, ncars):
        self.timers = np.zeros(BOOST_SIZE, dtype=F)
        self.avail = np.zeros(BOOST_SIZE, dtype=bool)
        self.init = False
        self.last_tick = None
        self.pred_pos = np.zeros((3, 3), dtype=F)
        self.pred_vel = np.zeros((3, 3), dtype=F)
        self.has_pred = False
        self.cars = [Timers() for _ in range(ncars)]


def update_arena(va, cars, pads, tick, dt):
    """vanta_update_arena (VantaObsGpu.cuh:192-243); pads in arena order."""
    if not va.init:
        va.avail[:] = True
        va.timers[:] = 0
        va.init = True
        va.last_tick = None
    if va.last_tick == tick:
        return
    va.last_tick = tick
    for i in range(BOOST_SIZE):
        active = bool(pads[PAD_MAP[i]].is_active)
        if active != va.avail[i]:
            va.avail[i] = active
            va.timers[i] = F(0) if active else (F(10) if PAD_LOCS[i, 2] > F(72) else F(4))
        elif not va.avail[i]:
            va.timers[i] = max(F(0), F(va.timers[i] - dt))
    for c, pt in zip(cars, va.cars):
        if not pt.init:
            _init_timers(pt, c)
        if c.is_demoed:
            if not pt.was_demoed:
                pt.demo_timer = F(3)
                pt.was_demoed = True
            pt.demo_timer = max(F(0), F(pt.demo_timer - dt))
        else:
            pt.demo_timer = F(0)
            pt.was_demoed = False


def _init_timers(pt, c):
    fresh = Timers()
    pt.__dict__.update(fresh.__dict__)
    pt.on_ground = bool(c.is_on_ground)
    pt.last_tick = None
    pt.init = True


def update_player_timers(pt, c, tick, dt, deadzone):
    """vanta_update_player_timers (VantaObsGpu.cuh:246-338)."""
    if not pt.init:
        _init_timers(pt, c)
    if pt.last_tick == tick:
        return
    pt.last_tick = tick
    prev = c.last_controls
    step = dt * F(120)
    if prev.boost:
        pt.boost_time = min(pt.boost_time + step, F(12))
    elif pt.boost_time >= F(12):
        pt.boost_time = F(0)
    if prev.boost:
        pt.boost_hold = min(pt.boost_hold + step, F(12))
    else:
        pt.boost_hold = max(pt.boost_hold - F(2) * dt * F(120), F(0))
    if pt.on_ground and not pt.is_jumping:
        pt.has_jumped = False
    jump_held = bool(prev.jump)
    jump_prev_rel = not pt.pp["jump"]
    if pt.is_jumping:
        pt.is_jumping = pt.jump_time < F(3) or (jump_held and pt.jump_time < F(24))
    elif jump_held and jump_prev_rel and pt.on_ground:
        pt.is_jumping = True
        pt.jump_time = F(0)
    if pt.is_jumping:
        pt.has_jumped = True
        pt.jump_time = min(pt.jump_time + step, F(24))
    else:
        pt.jump_time = F(0)
    if c.is_on_ground:
        pt.is_jumping = pt.has_jumped = pt.has_double_jumped = pt.has_flipped = False
        pt.air_time = pt.flip_time = F(0)
        pt.flip_dir = v3(0, 0, 0)
        pt.on_ground = True
    else:
        pt.air_time = min(pt.air_time + step, F(150)) if (pt.has_jumped and not pt.is_jumping) else F(0)
        if jump_held and jump_prev_rel and (not pt.has_jumped or pt.air_time < F(150)):
            if not pt.has_double_jumped and not pt.has_flipped:
                ab = max(max(abs(F(prev.pitch)), abs(F(prev.yaw))), abs(F(prev.roll)))
                if ab >= deadzone:
                    pt.flip_time = F(0)
                    pt.has_flipped = True
                    fx = F(prev.yaw)
                    if abs(fx) < F(1e-6) and abs(F(prev.roll)) >= deadzone:
                        fx = F(prev.roll)
                    fy = -F(prev.pitch)
                    n = np.sqrt(fx * fx + fy * fy)
                    pt.flip_dir = v3(fx / n, fy / n, 0) if n > F(1e-6) else v3(0, 0, 0)
                else:
                    pt.has_double_jumped = True
        pt.on_ground = False
    if pt.has_flipped:
        pt.flip_time = min(pt.flip_time + step, F(78))
    if prev.handbrake:
        pt.handbrake_time = min(pt.handbrake_time + F(5) * dt, F(1))
    else:
        pt.handbrake_time = max(pt.handbrake_time - F(2) * dt, F(0))
    pt.pp = dict(throttle=F(prev.throttle), steer=F(prev.steer), pitch=F(prev.pitch), yaw=F(prev.yaw),
                 roll=F(prev.roll), jump=bool(prev.jump), boost=bool(prev.boost), handbrake=bool(prev.handbrake))


def last_touch_id(cars):
    """GameState.lastTouchCarID as the obs kernel derives it (RocketSimCudaObs.cuh:153-164)."""
    best, who = 0, 0
    for c in cars:
        if c.ball_hit_valid and c.ball_hit_tick != 0xFFFFFFFFFFFFFFFF and c.ball_hit_tick >= best:
            best, who = c.ball_hit_tick, c.id
    return who


def build_obs(va, cars, ball, idx, tick, dt, deadzone=F(0.5), only_closest_opp=False):
    """vanta_build_obs (VantaObsGpu.cuh:342-747) for car `idx`: 366 floats.
    Updates that car's coyote timers first, as the reference does."""
    deadzone = F(deadzone)
    player = cars[idx]
    update_player_timers(va.cars[idx], player, tick, dt, deadzone)
    pt = va.cars[idx]
    inv = player.team == 1
    mx = F(player.pos.x) < F(0)
    me = canonical(Phys.of_car(player), inv, mx)
    bc = canonical(Phys.of_ball(vec(ball.pos), vec(ball.vel), vec(ball.ang_vel)), inv, mx)
    o = []

    # A. ego (81)
    pos_diff = bc.pos - me.pos
    vel_diff = bc.vel - me.vel
    brp = me.rot_dot(pos_diff)
    brv = me.rot_dot(vel_diff)
    avb = me.rot_dot(me.ang)
    o += [me.pos[0] / POS_STD, me.pos[1] / POS_STD, me.pos[2] / POS_STD,
          me.vel[0] / VEL_STD, me.vel[1] / VEL_STD, me.vel[2] / VEL_STD,
          me.ang[0] / ANG_STD, me.ang[1] / ANG_STD, me.ang[2] / ANG_STD,
          brp[0] / POS_STD, brp[1] / POS_STD, brp[2] / POS_STD,
          brv[0] / BALL_VEL_STD, brv[1] / BALL_VEL_STD, brv[2] / BALL_VEL_STD,
          me.fwd[0], me.fwd[1], me.fwd[2], me.up[0], me.up[1], me.up[2],
          avb[0] / ANG_STD, avb[1] / ANG_STD, avb[2] / ANG_STD,
          length(me.vel) / VEL_STD,
          F(player.boost) / F(100), F(player.is_on_ground), F(has_flip_or_jump(player)), F(player.is_demoed),
          F(player.has_jumped), F(player.is_supersonic)]
    w = F(0)
    for k in range(4):
        w = w + (F(1) if player.wheels_with_contact[k] else F(0))
    o += [w / F(4), F(has_flip_reset(player)), F(got_flip_reset(player)), F(player.handbrake_val), F(mx), F(inv)]
    # threat frame
    own_goal = v3(0, -5120, 0)
    to_goal = own_goal - bc.pos
    ln = length(to_goal)
    danger = vdiv(to_goal, ln) if ln > F(1) else v3(0, -1, 0)
    lateral = v3(-danger[1], danger[0], 0)
    lat_len = length(lateral)
    if lat_len > F(1e-6):
        lateral = vdiv(lateral, lat_len)
    car_to_ball = bc.pos - me.pos
    o += [dot(bc.vel, danger) / BALL_VEL_STD, dot(me.vel, danger) / VEL_STD, dot(me.vel, lateral) / VEL_STD,
          dot(car_to_ball, danger) / POS_STD, dot(car_to_ball, lateral) / POS_STD,
          length(own_goal - bc.pos) / (F(2) * POS_STD)]
    # alignment scalars
    attack = v3(0, 5120, 0)
    vel_n = safe_norm(me.vel)
    to_ball_n = safe_norm(bc.pos - me.pos)
    to_goal_n = safe_norm(attack - me.pos)
    o += [dot(me.fwd, to_ball_n), dot(me.fwd, to_goal_n), dot(vel_n, to_ball_n), dot(me.up, v3(0, 0, 1))]
    # previous action, steer / yaw / roll mirrored with x
    pa = player.last_controls
    o += [F(pa.throttle), -F(pa.steer) if mx else F(pa.steer), F(pa.pitch), -F(pa.yaw) if mx else F(pa.yaw),
          -F(pa.roll) if mx else F(pa.roll), F(bool(pa.jump)), F(bool(pa.boost)), F(bool(pa.handbrake))]
    o += [F(bool(player.world_contact) and F(player.world_contact_normal.z) < F(0.7))]
    o += [pt.demo_timer / DEMO_TIMER_STD, pt.boost_time / F(12), pt.jump_time / F(24), pt.air_time / F(150),
          pt.flip_time / F(78), pt.handbrake_time, -pt.flip_dir[0] if mx else pt.flip_dir[0], pt.flip_dir[1],
          pt.boost_hold / F(12), F(pt.is_jumping), F(pt.has_flipped), F(pt.has_double_jumped)]
    # ego enrichments
    boost_frac = F(player.boost) / F(100)
    dist_ball = length(pos_diff)
    closing = ((-dot(vel_diff, pos_diff)) / dist_ball) / VEL_STD if dist_ball > F(1e-3) else F(0)
    o += [np.sqrt(max(F(0), boost_frac)), F(bc.pos[2] > me.pos[2] + F(200)), closing, dist_ball / POS_STD]
    has_contact = bool(player.world_contact)
    cn = canonical_normal(vec(player.world_contact_normal), inv, mx) if has_contact else v3(0, 0, 0)
    o += [F(has_contact), cn[0], cn[1], cn[2], F(player.is_flipping), F(0)]
    team_frame = invert(Phys.of_car(player), inv)
    o += [F(team_frame.pos[0] < F(0))]
    mi, de = small_pad_stats(va, vec(player.pos))
    o += [mi, de]
    assert len(o) == EGO_SIZE

    # B. ball (29)
    o += [bc.pos[0] / POS_STD, bc.pos[1] / POS_STD, bc.pos[2] / POS_STD,
          bc.vel[0] / BALL_VEL_STD, bc.vel[1] / BALL_VEL_STD, bc.vel[2] / BALL_VEL_STD,
          bc.ang[0] / ANG_STD, bc.ang[1] / ANG_STD, bc.ang[2] / ANG_STD,
          length(bc.vel) / BALL_VEL_STD, F(bc.pos[2] <= F(100))]
    if va.has_pred:
        for i in range(3):
            s = canonical(Phys.of_ball(va.pred_pos[i], va.pred_vel[i], v3(0, 0, 0)), inv, mx)
            o += [s.pos[0] / POS_STD, s.pos[1] / POS_STD, s.pos[2] / POS_STD,
                  s.vel[0] / BALL_VEL_STD, s.vel[1] / BALL_VEL_STD, s.vel[2] / BALL_VEL_STD]
    else:
        o += [F(0)] * 18
    assert len(o) == EGO_SIZE + BALL_SIZE

    # C. goal context (16)
    opp, own = v3(0, 5120, 0), v3(0, -5120, 0)
    for d, local in ((opp - me.pos, True), (own - me.pos, True), (opp - bc.pos, False), (own - bc.pos, False)):
        v = me.rot_dot(d) if local else d
        o += [v[0] * GOAL_DIST_COEF, v[1] * GOAL_DIST_COEF, v[2] * GOAL_DIST_COEF, length(d) * GOAL_DIST_COEF]

    # D. team context (6)
    lt_id = last_touch_id(cars)
    lt_team = -1
    if lt_id != 0:
        for c in cars:
            if c.id == lt_id:
                lt_team = 0 if c.id == player.id else (1 if c.team == player.team else 2)
                break
    o += [F(lt_team == 0), F(lt_team == 1), F(lt_team == 2)]
    blue = sum(1 for c in cars if c.team == 0)
    orange = len(cars) - blue
    mine, theirs = (orange, blue) if inv else (blue, orange)
    o += [F(mine) / F(3), F(theirs) / F(3), F(bc.pos[1] < F(0))]

    # E/F. other cars by canonical distance to the ball (stable insertion sort)
    mates, opps = [], []
    for i, c in enumerate(cars):
        if c.id == player.id:
            continue
        pc = canonical(Phys.of_car(c), inv, mx)
        dist = FLT_MAX if c.is_demoed else length(pc.pos - bc.pos)
        (mates if c.team == player.team else opps).append((dist, i))
    mates = _insertion_sorted(mates)
    opps = _insertion_sorted(opps)

    def packet(other, other_idx):
        oc = canonical(Phys.of_car(other), inv, mx)
        rp = oc.pos - me.pos
        rv = oc.vel - me.vel
        brp2, brv2, bav = me.rot_dot(rp), me.rot_dot(rv), me.rot_dot(oc.ang)
        out = [oc.pos[0] / POS_STD, oc.pos[1] / POS_STD, oc.pos[2] / POS_STD,
               oc.vel[0] / VEL_STD, oc.vel[1] / VEL_STD, oc.vel[2] / VEL_STD,
               oc.ang[0] / ANG_STD, oc.ang[1] / ANG_STD, oc.ang[2] / ANG_STD,
               brp2[0] / POS_STD, brp2[1] / POS_STD, brp2[2] / POS_STD,
               brv2[0] / VEL_STD, brv2[1] / VEL_STD, brv2[2] / VEL_STD,
               bav[0] / ANG_STD, bav[1] / ANG_STD, bav[2] / ANG_STD,
               oc.fwd[0], oc.fwd[1], oc.fwd[2], oc.up[0], oc.up[1], oc.up[2],
               F(other.boost) / F(100), F(other.is_on_ground), F(has_flip_or_jump(other)), F(other.is_demoed),
               F(other.is_supersonic), length(rp) / POS_STD, length(oc.pos - bc.pos) / POS_STD]
        demo_t = F(0)
        for j, c in enumerate(cars):
            if c.id == other.id:
                if va.cars[j].init:
                    demo_t = va.cars[j].demo_timer
                break
        out += [demo_t / DEMO_TIMER_STD, F(other.is_jumping), F(other.is_flipping), F(other.has_double_jumped),
                length(oc.vel) / VEL_STD, F(other.has_flipped)]
        mi2, de2 = small_pad_stats(va, vec(other.pos))
        return out + [mi2, de2]

    for i in range(MAX_TEAMMATES):
        o += packet(cars[mates[i]], mates[i]) if i < len(mates) else [F(0)] * CAR_SIZE
    opp_limit = 1 if only_closest_opp else MAX_OPPONENTS
    for i in range(MAX_OPPONENTS):
        o += packet(cars[opps[i]], opps[i]) if (i < opp_limit and i < len(opps)) else [F(0)] * CAR_SIZE
    o += [F(i < len(mates)) for i in range(MAX_TEAMMATES)]
    o += [F(i < opp_limit and i < len(opps)) for i in range(MAX_OPPONENTS)]

    # H. boost pads (34)
    for i in range(BOOST_SIZE):
        k = BOOST_SIZE - 1 - i if inv else i
        if mx:
            k = PAD_MIRROR[k]
        if va.avail[k]:
            o.append(F(1))
        else:
            t_max = F(10) if PAD_LOCS[k, 2] > F(72) else F(4)
            o.append(F(1) / (F(1) + va.timers[k] / t_max))
    row = np.array(o, dtype=F)
    assert row.shape == (OBS_SIZE,)
    return row


def _insertion_sorted(items):
    """Indices of (dist, idx) items after the reference's insertion sort
    (ties keep slot order)."""
    items = list(items)
    for i in range(1, len(items)):
        cur = items[i]
        j = i - 1
        while j >= 0 and items[j][0] > cur[0]:
            items[j + 1] = items[j]
            j -= 1
        items[j + 1] = cur
    return [idx for _, idx in items]


def build_all(va, cars, ball, pads, tick, dt, preds=None, deadzone=0.5, only_closest_opp=False):
    """BuildVantaObs for one arena: the arena update (pads, demo timers) and
    the prediction slices (preds: (pos[3][3], vel[3][3]) or None), then every
    car's row in slot order (vantaArenaUpdateKernel + vantaObsKernel)."""
    if preds is not None:
        va.pred_pos[:] = preds[0]
        va.pred_vel[:] = preds[1]
        va.has_pred = True
    else:
        va.has_pred = False
    update_arena(va, cars, pads, tick, F(dt))
    return np.stack([build_obs(va, cars, ball, i, tick, F(dt), deadzone, only_closest_opp)
                     for i in range(len(cars))])
