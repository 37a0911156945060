"""VantaObs on the CPU.

The restatement (oracle/vanta_spec.py) follows the reference's GPU
transcription (RocketSimCuda/src/VantaObsGpu.cuh); the CPU builder it
transcribes is not in the reference tree, so the restatement is pinned here by
a hand-derived row, the x-mirror symmetry the row is built around, and
hand-traced timer sequences.  The CPU build of the engine's Vanta code
(rsx_vanta.h) is then compared with it bit for bit on oracle trajectories, and
its ball-only rollout with the oracle's ball-only arena."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

V = R.load_vanta_spec()
F = np.float32
DT = F(8) / F(120)
NEVER = 2**64 - 1

# row offsets
EGO, BALL, GOAL, TEAM, MATES, OPPS, MASK, PADS = 0, 81, 110, 126, 132, 210, 327, 332


def car(pos, fwd=(1, 0, 0), right=(0, 1, 0), up=(0, 0, 1), vel=(0, 0, 0), ang=(0, 0, 0), team=0, cid=1,
        on_ground=True, boost=33.0):
    s = R.CarState()
    s.pos.set(pos)
    s.rot.forward.set(fwd)
    s.rot.right.set(right)
    s.rot.up.set(up)
    s.vel.set(vel)
    s.ang_vel.set(ang)
    s.team = team
    s.id = cid
    s.is_on_ground = 1 if on_ground else 0
    s.boost = boost
    s.ball_hit_tick = NEVER
    s.ball_hit_extra_impulse_tick = NEVER
    return s


def pads_all_active():
    p = R.PADS()
    for i in range(34):
        p[i].is_active = 1
    return p


def build(va, cars, ball, pads, tick, preds=None, only_closest=False):
    return V.build_all(va, cars, ball, pads, tick, DT, preds, 0.5, only_closest)


def test_row_layout():
    assert V.OBS_SIZE == 366
    assert (BALL, GOAL, TEAM, MATES, OPPS, MASK, PADS) == (81, 110, 126, 132, 210, 327, 332)
    assert PADS + 34 == V.OBS_SIZE
    # the pad tables: every slot maps to one arena pad, the mirror map is an
    # involution that negates x and keeps y, small pads at z 70, big at 73
    assert sorted(V.PAD_MAP) == list(range(34))
    for i in range(34):
        j = V.PAD_MIRROR[i]
        assert V.PAD_MIRROR[j] == i
        assert V.PAD_LOCS[j, 0] == -V.PAD_LOCS[i, 0] and V.PAD_LOCS[j, 1] == V.PAD_LOCS[i, 1]
    assert sorted(set(V.PAD_LOCS[:, 2].tolist())) == [70.0, 73.0]
    assert int((V.PAD_LOCS[:, 2] > 72).sum()) == 6


def test_orange_row_by_hand():
    """An orange car at (1000, 2000, 17) facing -y at 500 uu/s, the ball at
    rest at the centre: the team inversion turns its pose into (-1000, -2000)
    facing +y, x > 0 in the world so no mirror; every value below is derived
    by hand from that canonical pose."""
    me = car((1000, 2000, 17), fwd=(0, -1, 0), right=(1, 0, 0), vel=(0, -500, 0), team=1, cid=2)
    other = car((-1500, -3000, 17), team=0, cid=1)
    ball = P.ball((0, 0, 93.15))
    va = V.ArenaState(2)
    rows = build(va, [other, me], ball, pads_all_active(), 10)
    o = rows[1]
    S = F(2300)
    assert o[0] == F(-1000) / S and o[1] == F(-2000) / S and o[2] == F(17) / S
    assert o[3] == 0 and o[4] == F(500) / S and o[5] == 0
    assert tuple(o[15:18]) == (0, 1, 0) and tuple(o[18:21]) == (0, 0, 1)
    # ball in the car frame: ahead 2000, to the left (-right) 1000, up 76.15
    assert o[9] == F(2000) / S and o[10] == F(-1000) / S and o[11] == (F(93.15) - F(17)) / S
    assert o[24] == F(500) / S                      # speed
    assert o[35] == 0 and o[36] == 1                # not mirrored, inverted
    assert o[78] == 1                               # team-frame x < 0
    # goal context: the attacked goal (0, 5120, 0) seen from (-1000, -2000, 17)
    c = V.GOAL_DIST_COEF
    assert o[GOAL + 0] == F(7120) * c and o[GOAL + 1] == F(-1000) * c and o[GOAL + 2] == F(-17) * c
    assert o[GOAL + 3] == np.sqrt(F(1000) * F(1000) + F(7120) * F(7120) + F(17) * F(17)) * c
    # team context: one car each side; the ball at y = 0 is not on the own half
    assert tuple(o[TEAM + 3:TEAM + 6]) == (F(1) / F(3), F(1) / F(3), 0)
    # the one opponent: canonical (1500, 3000), 3000 uu ahead-ish of the car
    assert o[OPPS + 0] == F(1500) / S and o[OPPS + 1] == F(3000) / S
    assert tuple(o[MASK:MASK + 5]) == (0, 0, 1, 0, 0)
    assert (o[PADS:] == 1).all() and (o[MATES:OPPS] == 0).all()
    # the blue car is at x < 0: mirrored, not inverted
    b = rows[0]
    assert b[35] == 1 and b[36] == 0 and b[0] == F(1500) / S and b[1] == F(-3000) / S


def _random_rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    m = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    return m.astype(F)


def _random_car(rng, team, cid, x_sign=None):
    m = _random_rot(rng)
    x = rng.uniform(300, 3800) * (x_sign if x_sign else rng.choice([-1, 1]))
    c = car((x, rng.uniform(-4800, 4800), rng.uniform(17, 1800)), fwd=m[:, 0], right=m[:, 1], up=m[:, 2],
            vel=rng.uniform(-2000, 2000, 3), ang=rng.uniform(-5, 5, 3), team=team, cid=cid,
            on_ground=bool(rng.random() < 0.5), boost=float(rng.uniform(0, 100)))
    c.has_jumped = int(rng.random() < 0.5)
    c.has_flipped = int(rng.random() < 0.3)
    c.is_flipping = int(rng.random() < 0.2)
    c.is_supersonic = int(rng.random() < 0.3)
    c.air_time_since_jump = float(rng.uniform(0, 2))
    c.handbrake_val = float(rng.uniform(0, 1))
    c.world_contact = int(rng.random() < 0.5)
    c.world_contact_normal.set(m[:, 2])
    for k in range(4):
        c.wheels_with_contact[k] = int(rng.random() < 0.5)
    lc = c.last_controls
    lc.throttle, lc.steer, lc.pitch, lc.yaw, lc.roll = rng.uniform(-1, 1, 5)
    lc.jump, lc.boost, lc.handbrake = (int(rng.random() < 0.4) for _ in range(3))
    return c


def _mirror_car(c):
    """The car reflected in the x = 0 plane: positions and velocities flip x,
    the angular velocity (an axial vector) keeps x and flips y and z, each
    basis vector flips x, and the left-right controls change sign."""
    m = R.CarState()
    C.memmove(C.byref(m), C.byref(c), C.sizeof(c))
    m.pos.x, m.vel.x = -c.pos.x, -c.vel.x
    m.ang_vel.y, m.ang_vel.z = -c.ang_vel.y, -c.ang_vel.z
    for v, src in ((m.rot.forward, c.rot.forward), (m.rot.right, c.rot.right), (m.rot.up, c.rot.up)):
        v.x = -src.x
    m.world_contact_normal.x = -c.world_contact_normal.x
    m.last_controls.steer = -c.last_controls.steer
    m.last_controls.yaw = -c.last_controls.yaw
    m.last_controls.roll = -c.last_controls.roll
    return m


def _mirror_ball(b):
    return P.ball((-b.pos.x, b.pos.y, b.pos.z), (-b.vel.x, b.vel.y, b.vel.z), (b.ang_vel.x, -b.ang_vel.y,
                                                                                 -b.ang_vel.z))


def _mirror_pads(p):
    """Each arena pad takes the state of the pad at its x-mirror."""
    out = R.PADS()
    arena = V._R.ARENA_PADS
    for i in range(34):
        j = int(np.nonzero((arena[:, 0] == -arena[i, 0]) & (arena[:, 1] == arena[i, 1]))[0][0])
        out[i].is_active = p[j].is_active
        out[i].cooldown = p[j].cooldown
    return out


def test_mirror_symmetry_over_updates():
    """A blue car at x < 0 sees the x-mirror of its world: its row equals the
    row the mirrored car (at x > 0, not mirrored) gets in the mirrored world,
    except the mirror flag (35) and the team-frame kickoff side (78) -- over a
    sequence of updates, so pad timers, flip directions and the coyote timers
    are compared too."""
    rng = np.random.default_rng(7)
    va, vm = V.ArenaState(4), V.ArenaState(4)
    pads = pads_all_active()
    for step in range(12):
        cars = [_random_car(rng, 0, 1, x_sign=-1), _random_car(rng, 0, 2), _random_car(rng, 1, 3),
                _random_car(rng, 1, 4)]
        cars[2].is_demoed = int(step % 5 == 3)
        if step == 3:  # landed, jump released
            cars[0].last_controls.jump, cars[0].is_on_ground = 0, 1
        if step == 4:  # a flip press: a fresh jump in the air with a sideways stick
            cars[0].last_controls.jump, cars[0].last_controls.yaw, cars[0].is_on_ground = 1, 0.9, 0
        ball = P.ball(rng.uniform(-3000, 3000, 3) + np.array([0, 0, 3200]), rng.uniform(-3000, 3000, 3),
                      rng.uniform(-5, 5, 3))
        for i in range(34):
            if rng.random() < 0.2:
                pads[i].is_active = 1 - pads[i].is_active
        preds = (rng.uniform(-3000, 3000, (3, 3)).astype(F), rng.uniform(-3000, 3000, (3, 3)).astype(F))
        mpreds = (preds[0] * V.v3(-1, 1, 1), preds[1] * V.v3(-1, 1, 1))
        a = build(va, cars, ball, pads, 100 + 8 * step, preds)[0]
        b = build(vm, [_mirror_car(c) for c in cars], _mirror_ball(ball), _mirror_pads(pads), 100 + 8 * step,
                  mpreds)[0]
        assert a[35] == 1 and b[35] == 0, step
        # the small-pad densities sum the same terms in another order (pad
        # slots are not mirror-symmetric), so they agree to rounding only
        dens = [80] + [MATES + 39 * k + 38 for k in range(2)] + [OPPS + 39 * k + 38 for k in range(3)]
        keep = np.ones(366, dtype=bool)
        keep[[35, 78] + dens] = False
        np.testing.assert_array_equal(a[keep], b[keep], err_msg=f"update {step}")
        np.testing.assert_allclose(a[dens], b[dens], rtol=1e-6, atol=0)
        if step == 4:
            assert va.cars[0].has_flipped and va.cars[0].flip_dir[0] != 0 and a[62] == -va.cars[0].flip_dir[0]


def test_coyote_timers_by_hand():
    """Hand-traced updates of one car's timers (VantaObsGpu.cuh:246-338) at
    8 ticks per update: boost time grows 8 per update to 12 and returns to 0
    only from 12; boost hold falls 16 per update; a jump pressed on the ground
    and seen in the air starts the jump timer; a second press with a stick
    past the deadzone is a flip with a unit direction (yaw, -pitch)."""
    b = P.ball((0, 3000, 500))
    va = V.ArenaState(1)
    c = car((0, 0, 17), on_ground=True)
    tick = [0]

    def upd(**ctl):
        tick[0] += 8
        lc = c.last_controls
        lc.throttle = lc.steer = lc.pitch = lc.yaw = lc.roll = 0
        lc.jump = lc.boost = lc.handbrake = 0
        for k, v in ctl.items():
            setattr(lc, k, v)
        row = build(va, [c], b, pads_all_active(), tick[0])[0]
        return va.cars[0], row

    t, _ = upd(boost=1)
    assert abs(t.boost_time - 8) < 1e-5 and abs(t.boost_hold - 8) < 1e-5
    t, _ = upd(boost=0)
    assert abs(t.boost_time - 8) < 1e-5 and t.boost_hold == 0   # not at 12: kept
    t, _ = upd(boost=1)
    t, row = upd(boost=1)
    assert t.boost_time == 12 and t.boost_hold == 12 and row[57] == 1 and row[64] == 1
    t, _ = upd(boost=0)
    assert t.boost_time == 0 and t.boost_hold == 0              # reached 12: reset
    t, _ = upd(handbrake=1)
    assert abs(t.handbrake_time - F(5) * DT) < 1e-7
    # jump pressed: seen while the car is already in the air
    c.is_on_ground = 0
    t, row = upd(jump=1)
    assert t.is_jumping and t.has_jumped and abs(t.jump_time - 8) < 1e-5 and row[65] == 1
    assert t.has_double_jumped  # the press itself, with no stick, counts as the second jump
    t, _ = upd(jump=1)
    assert t.is_jumping and abs(t.jump_time - 16) < 1e-5
    t, _ = upd(jump=0)
    assert not t.is_jumping and t.jump_time == 0 and abs(t.air_time - 8) < 1e-5
    # landing clears the jump state
    c.is_on_ground = 1
    t, _ = upd()
    assert not (t.has_jumped or t.has_double_jumped or t.has_flipped) and t.air_time == 0 and t.on_ground
    # take off, press jump with yaw 0.6 and pitch -0.8: a flip towards (0.6, 0.8)
    c.is_on_ground = 0
    t, row = upd(jump=1, yaw=0.6, pitch=-0.8)
    assert t.has_flipped and not t.has_double_jumped
    assert abs(t.flip_dir[0] - 0.6) < 1e-6 and abs(t.flip_dir[1] - 0.8) < 1e-6 and abs(t.flip_time - 8) < 1e-5
    assert row[66] == 1 and abs(row[62] - 0.6) < 1e-6
    # the car at x < 0 sees its flip direction mirrored
    c.pos.x = -100
    t, row = upd()
    assert row[35] == 1 and abs(row[62] + 0.6) < 1e-6 and abs(t.flip_time - 16) < 1e-5


def test_pad_and_demo_timers_by_hand():
    """A pad that goes inactive starts at 4 s (small) or 10 s (big) and runs
    down by the update interval; its obs value is 1 / (1 + t / t_max).  A
    demolished car's timer starts at 3 s less one interval."""
    va = V.ArenaState(2)
    me, opp = car((500, -1000, 17), cid=1), car((0, 1000, 17), team=1, cid=2)
    b = P.ball()
    pads = pads_all_active()
    small = int(np.nonzero(V.PAD_LOCS[:, 2] < 72)[0][0])
    big = int(np.nonzero(V.PAD_LOCS[:, 2] > 72)[0][0])
    pads[V.PAD_MAP[small]].is_active = 0
    pads[V.PAD_MAP[big]].is_active = 0
    opp.is_demoed = 1
    r1 = build(va, [me, opp], b, pads, 8)
    assert va.timers[small] == 4 and va.timers[big] == 10
    assert r1[0][PADS + small] == F(1) / (F(1) + F(4) / F(4)) and r1[0][PADS + big] == F(1) / F(2)
    assert va.cars[1].demo_timer == F(3) - DT
    assert r1[0][OPPS + 31] == (F(3) - DT) / F(3)           # the opponent packet's demo timer
    r2 = build(va, [me, opp], b, pads, 16)
    assert va.timers[small] == F(4) - DT and va.cars[1].demo_timer == (F(3) - DT) - DT
    r3 = build(va, [me, opp], b, pads, 16)                   # same tick: nothing advances
    np.testing.assert_array_equal(r2, r3)
    opp.is_demoed = 0
    build(va, [me, opp], b, pads, 24)
    assert va.cars[1].demo_timer == 0 and not va.cars[1].was_demoed


def test_other_cars_sorted_by_distance_to_the_ball():
    """3v3: teammates and opponents each sorted by their canonical distance to
    the ball, demolished cars last; only_closest_opp keeps one packet; the
    last toucher's team is read from the hit ticks."""
    b = P.ball((0, 0, 93))
    me = car((-100, -4000, 17), cid=1)
    mates = [car((0, -3000, 17), cid=2), car((0, -500, 17), cid=3)]
    opps = [car((0, 2500, 17), team=1, cid=4), car((0, 600, 17), team=1, cid=5), car((0, 100, 17), team=1, cid=6)]
    opps[2].is_demoed = 1
    opps[0].ball_hit_valid, opps[0].ball_hit_tick = 1, 50
    mates[0].ball_hit_valid, mates[0].ball_hit_tick = 1, 40
    cars = [me] + mates + opps
    row = build(V.ArenaState(6), cars, b, pads_all_active(), 64)[0]
    S = F(2300)
    # mirrored (x < 0): canonical y is the world y
    assert row[MATES + 1] == F(-500) / S and row[MATES + 39 + 1] == F(-3000) / S
    assert row[OPPS + 1] == F(600) / S and row[OPPS + 39 + 1] == F(2500) / S and row[OPPS + 78 + 1] == F(100) / S
    assert tuple(row[TEAM:TEAM + 3]) == (0, 0, 1)            # an opponent touched last
    assert tuple(row[MASK:MASK + 5]) == (1, 1, 1, 1, 1)
    one = build(V.ArenaState(6), cars, b, pads_all_active(), 64, only_closest=True)[0]
    assert tuple(one[MASK:MASK + 5]) == (1, 1, 1, 0, 0) and (one[OPPS + 39:MASK] == 0).all()
    np.testing.assert_array_equal(one[:OPPS + 39], row[:OPPS + 39])


# ---- the CPU build of rsx_vanta.h --------------------------------------------

def _cc_arena(lib, w, cars, ball, pads, tick):
    nb = sum(1 for c in cars if c.team == 0)
    h = lib.cc_arena_new(w, nb, len(cars) - nb)
    for i, c in enumerate(cars):
        lib.cc_set_car(h, i, C.byref(c))
    lib.cc_set_ball(h, C.byref(ball))
    lib.cc_set_pads(h, pads)
    lib.cc_set_tick(h, tick)
    return h


def _cc_public(lib, h, n):
    cars = []
    for i in range(n):
        s = R.CarState()
        lib.cc_get_car(h, i, C.byref(s))
        cars.append(s)
    b = R.BallState()
    lib.cc_get_ball(h, C.byref(b))
    pads = R.PADS()
    lib.cc_get_pads(h, pads)
    return cars, b, pads


@pytest.mark.parametrize("team", [1, 2, 3])
def test_cpu_build_matches_the_restatement_on_oracle_trajectories(oracle, cpucheck, team):
    """Oracle arenas stepped with random controls from random states; after
    every env step the state goes into the CPU build, whose Vanta rows
    (prediction on) must equal the restatement's bit for bit, update after
    update, with the per-car state carried on both sides."""
    lib, w = cpucheck
    rng = np.random.default_rng(100 + team)
    n = 2 * team
    h = C.c_void_p(oracle.ref_arena_new(team, team))
    oracle.ref_step(h, 1)
    ids = (C.c_uint32 * 8)()
    oracle.ref_car_order(h, ids, 8)
    for i in range(n):
        s = _random_car(rng, 0 if i < team else 1, i + 1)
        s.pos.z, s.is_on_ground = 17.0, 1
        s.rot.forward.set((1, 0, 0)), s.rot.right.set((0, 1, 0)), s.rot.up.set((0, 0, 1))
        s.world_contact = 0
        oracle.ref_set_car(h, i, C.byref(s))
    oracle.ref_set_ball(h, C.byref(P.ball((rng.uniform(-2000, 2000), rng.uniform(-3000, 3000), 400),
                                          rng.uniform(-1500, 1500, 3))))
    vs = C.c_void_p(lib.cc_vanta_new())
    va = V.ArenaState(n)
    pred = (C.c_float * 18)()
    rows = (C.c_float * (366 * n))()
    for step in range(25):
        for i in range(n):
            ctl = P.controls(*rng.uniform(-1, 1, 5), *(int(rng.random() < 0.3) for _ in range(3)))
            oracle.ref_set_controls(h, i, C.byref(ctl))
        oracle.ref_step(h, 8)
        cars, ball = [], R.BallState()
        for i in range(n):
            s = R.CarState()
            oracle.ref_get_car(h, i, C.byref(s))
            cars.append(s)
        oracle.ref_get_ball(h, C.byref(ball))
        pads = R.PADS()
        oracle.ref_get_pads(h, pads)
        a = C.c_void_p(_cc_arena(lib, w, cars, ball, pads, oracle.ref_get_tick(h)))
        pc, pb, pp = _cc_public(lib, a, n)
        lib.cc_vanta_predict(a, pred)
        lib.cc_vanta_build(a, vs, pred, DT, F(0.5), 0, rows)
        tick = oracle.ref_get_tick(h)
        lib.cc_arena_free(a)
        p = np.array(pred, dtype=F).reshape(3, 6)
        want = V.build_all(va, pc, pb, pp, tick, DT, (p[:, :3], p[:, 3:]), 0.5, False)
        got = np.array(rows, dtype=F).reshape(n, 366)
        np.testing.assert_array_equal(got, want, err_msg=f"{team}v{team} step {step}")
    lib.cc_vanta_free(vs)
    oracle.ref_arena_free(h)


def _oracle_prediction(oracle, b):
    h = C.c_void_p(oracle.ref_arena_new(0, 0))
    oracle.ref_set_ball(h, C.byref(b))
    out = []
    for _ in range(3):
        oracle.ref_step(h, 120)
        s = R.BallState()
        oracle.ref_get_ball(h, C.byref(s))
        out.append([s.pos.x, s.pos.y, s.pos.z, s.vel.x, s.vel.y, s.vel.z])
    oracle.ref_arena_free(h)
    return np.array(out, dtype=np.float64)


def _prediction_states(rng, n):
    """Balls in flight, rolling, and thrown at walls, corners and goals."""
    out = []
    for k in range(n):
        if k % 3 == 0:
            pos = (rng.uniform(-3500, 3500), rng.uniform(-4500, 4500), rng.uniform(200, 1800))
        elif k % 3 == 1:
            pos = (rng.uniform(-3500, 3500), rng.uniform(-4500, 4500), 93.15)
        else:
            pos = (rng.choice([-1, 1]) * rng.uniform(3300, 3900), rng.choice([-1, 1]) * rng.uniform(3800, 4900),
                   rng.uniform(100, 900))
        vel = rng.uniform(-2500, 2500, 3)
        if k % 3 == 1:
            vel[2] = 0
        out.append(P.ball(pos, vel, rng.uniform(-4, 4, 3)))
    return out


def test_cpu_build_ball_prediction_against_the_oracle(oracle, cpucheck):
    """The CPU build's ball-only rollout against the oracle's ball-only arena
    (BallPredTracker's) on balls in flight, rolling and thrown at walls,
    corners and goals: every slice within 1e-4 of the obs scales (0.23 uu,
    0.6 uu/s) after 120, 240 and 360 ticks.  (On 300 such balls the largest
    differences were 0.009 uu and 0.003 uu/s; about half the 1 s slices are
    bit-identical.)"""
    lib, w = cpucheck
    rng = np.random.default_rng(11)
    pred = (C.c_float * 18)()
    for k, b in enumerate(_prediction_states(rng, 60)):
        h = C.c_void_p(lib.cc_arena_new(w, 0, 0))
        lib.cc_set_ball(h, C.byref(b))
        lib.cc_vanta_predict(h, pred)
        lib.cc_arena_free(h)
        got = np.array(pred, dtype=np.float64).reshape(3, 6)
        want = _oracle_prediction(oracle, b)
        assert np.abs(got[:, :3] - want[:, :3]).max() <= 1e-4 * 2300, (k, got, want)
        assert np.abs(got[:, 3:] - want[:, 3:]).max() <= 1e-4 * 6000, (k, got, want)


def test_ball_only_tick_equals_the_arena_tick_without_cars(cpucheck):
    """The prediction's tick (an arena with no car slots, whose car branches
    compile away) gives the same ball, bit for bit, as the full arena tick
    stepping an arena that holds no cars."""
    lib, w = cpucheck
    rng = np.random.default_rng(12)
    pred = (C.c_float * 18)()
    for b in _prediction_states(rng, 45):
        h = C.c_void_p(lib.cc_arena_new(w, 0, 0))
        lib.cc_set_ball(h, C.byref(b))
        lib.cc_vanta_predict(h, pred)
        full = []
        for _ in range(3):
            lib.cc_step(h, 120)
            s = R.BallState()
            lib.cc_get_ball(h, C.byref(s))
            full.append([s.pos.x, s.pos.y, s.pos.z, s.vel.x, s.vel.y, s.vel.z])
        lib.cc_arena_free(h)
        np.testing.assert_array_equal(np.array(pred, dtype=F).reshape(3, 6), np.array(full, dtype=F))


def test_rollout_continuation_equals_a_new_rollout(cpucheck):
    """The kernel continues the last rollout instead of starting a new one
    when the ball is, bit for bit, the one it passed after ts ticks: the
    slices (the states it passed at 120 + ts and 240 + ts, and its 360-tick
    state advanced ts ticks) and the advanced continuation states equal a new
    rollout's, bit for bit, step after step (balls in flight, rolling and at
    walls, ts = 8 and 1)."""
    lib, w = cpucheck
    rng = np.random.default_rng(13)
    for k, b in enumerate(_prediction_states(rng, 24)):
        h = C.c_void_p(lib.cc_arena_new(w, 0, 0))
        lib.cc_set_ball(h, C.byref(b))
        for ts in (8, 1):
            assert lib.cc_vanta_continue_check(h, ts, 6) == 6, (k, ts)
        lib.cc_arena_free(h)
