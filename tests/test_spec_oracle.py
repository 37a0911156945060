"""Pins the CPU restatement of the RLGym env pieces (oracle/rlgym_spec.py) with
closed forms and invariants, since the reference's own C++ for them does not
build as shipped."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

S = R.load_spec()


def test_pad_tables_match_the_oracle_arena(oracle):
    """ARENA_PADS is the oracle arena's pad order; every BOOST_LOCATIONS slot
    maps to exactly one pad; the inverted order is the point mirror."""
    h = C.c_void_p(oracle.ref_arena_new(1, 1))
    buf = (C.c_float * (34 * 4))()
    oracle.ref_pad_config(h, buf)
    oracle.ref_arena_free(h)
    cfg = np.array(buf, dtype=np.float32).reshape(34, 4)
    np.testing.assert_array_equal(cfg[:, :2], S.ARENA_PADS)
    assert list(cfg[:6, 3]) == [1] * 6 and not cfg[6:, 3].any()
    m = S.PAD_MAP
    assert sorted(m) == list(range(34))
    for i in range(34):
        a, b = S.ARENA_PADS[m[i]], S.ARENA_PADS[m[33 - i]]
        assert np.allclose(a, -b, atol=2.5), i


def _state():
    blue = P.default_car(pos=(-2048, -2560, 17), yaw=np.pi / 4, vel=(230, 460, 0), ang=(0, 0, 1.1), boost=50)
    orange = P.default_car(pos=(2048, 2560, 17), yaw=np.pi / 4 + np.pi, boost=25, team=1)
    ball = P.ball((100, -200, 300), (600, -1200, 50), (1, 2, 3))
    pads = R.PADS()
    for i in range(34):
        pads[i].is_active = 1 if i % 3 else 0
    return [blue, orange], ball, pads


def test_default_obs_closed_form():
    cars, ball, pads = _state()
    prev = np.zeros((2, 8), np.float32)
    prev[0] = [1, -1, 0, -1, 0, 0, 1, 0]
    row = S.default_obs(cars, ball, pads, prev, 0)
    assert row.shape == (89,)
    np.testing.assert_allclose(row[0:3], [100 / 4096, -200 / 5120, 300 / 2044], rtol=1e-6)
    np.testing.assert_allclose(row[3:6], [600 / 2300, -1200 / 2300, 50 / 2300], rtol=1e-6)
    np.testing.assert_allclose(row[6:9], [1 / 5.5, 2 / 5.5, 3 / 5.5], rtol=1e-6)
    np.testing.assert_array_equal(row[9:17], prev[0])
    np.testing.assert_array_equal(row[17:51], [float(pads[S.PAD_MAP[i]].is_active) for i in range(34)])
    me = row[51:70]
    c45 = np.cos(np.pi / 4)
    np.testing.assert_allclose(me[0:3], [-0.5, -0.5, 17 / 2044], rtol=1e-6)
    np.testing.assert_allclose(me[3:6], [c45, c45, 0], atol=1e-6)
    np.testing.assert_allclose(me[6:9], [0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(me[9:12], [0.1, 0.2, 0], rtol=1e-6)
    np.testing.assert_allclose(me[12:15], [0, 0, 0.2], rtol=1e-6)
    np.testing.assert_allclose(me[15:19], [0.5, 1, 1, 0])
    opp = row[70:89]
    np.testing.assert_allclose(opp[0:3], [0.5, 0.5, 17 / 2044], rtol=1e-6)
    np.testing.assert_allclose(opp[15], 0.25)


def _mirror(cars, ball, pads):
    """Point mirror through the z axis with teams swapped."""
    out = []
    for c in cars:
        m = R.CarState.from_buffer_copy(c)
        for v in (m.pos, m.rot.forward, m.rot.right, m.rot.up, m.vel, m.ang_vel):
            v.x, v.y = -v.x, -v.y
        m.team = 1 - c.team
        out.append(m)
    b = R.BallState.from_buffer_copy(ball)
    for v in (b.pos, b.vel, b.ang_vel):
        v.x, v.y = -v.x, -v.y
    mp = R.PADS()
    for i in range(34):  # arena pad j <-> the pad at the mirrored position
        j = S.PAD_MAP[i]
        k = S.PAD_MAP[33 - i]
        mp[k].is_active = pads[j].is_active
        mp[k].cooldown = pads[j].cooldown
    return out, b, mp


def test_orange_view_is_the_mirrored_blue_view():
    """Inversion invariant: an orange car sees the mirrored world exactly as
    the blue car sees the original one (catches sign and pad-order errors)."""
    cars, ball, pads = _state()
    mc, mb, mp = _mirror(cars, ball, pads)
    prev = np.arange(16, dtype=np.float32).reshape(2, 8) / 16
    for i in range(2):
        np.testing.assert_allclose(S.default_obs(cars, ball, pads, prev, i),
                                   S.default_obs(mc, mb, mp, prev, i), atol=1e-7)


def test_rewards_closed_form():
    w = dict(goal=1, concede=-1, touch=1, vbtg=1, vptb=1)
    blue = P.default_car(pos=(0, -1000, 17), vel=(0, 2300, 0))
    orange = P.default_car(pos=(0, 1000, 17), vel=(0, 0, 0), team=1)
    ball = P.ball((0, 0, 100), (0, 6000, 0))
    r = S.stub_rewards([blue, orange], ball, False, [False, True], w)
    vb = 6000 / np.hypot(6000, 642.775 / 2 - 100)        # ball -> orange goal back, ball speed 6000
    vp = 1000 / np.hypot(1000, 83)                        # car -> ball direction . (0, 1, 0)
    np.testing.assert_allclose(r[0], vb + vp, rtol=1e-5)
    np.testing.assert_allclose(r[1], 1 - vb, rtol=1e-5)  # touch, ball moving away from the blue goal target
    past = P.ball((0, 5300, 100), (0, 0, 0))
    assert S.ball_scored(past) and not S.ball_scored(P.ball((0, 5215, 100)))
    g = S.stub_rewards([blue, orange], past, True, [False, False], dict(w, touch=0, vbtg=0, vptb=0))
    np.testing.assert_array_equal(g, [1, -1])


def test_terminals():
    t, nt = S.step_terminal(False, False, 8, 0.0, 0.0)
    assert t == 0 and nt == 0.0
    nt, steps = 0.0, 0
    while True:
        steps += 1
        t, nt = S.step_terminal(False, False, 8, nt, 1.0)
        if t:
            break
    # float32: 8 * (1 / 120.f) = 0.06666667, and 15 of them sum to exactly 1.0f
    assert S.step_dt(8) == np.float32(0.06666667)
    assert t == 2 and steps == 15 and nt == np.float32(1.0)
    t, nt = S.step_terminal(True, False, 8, 0.99, 1.0)
    assert t == 1
    t, nt = S.step_terminal(False, True, 8, 0.99, 1.0)
    assert t == 0 and nt == 0.0
    # a goal takes priority over the truncation (EnvSet.cpp:178-200)
    t, _ = S.step_terminal(True, False, 8, 0.99, 1.0)
    assert t == 1


def _hit_car(tick_hit, valid=1):
    c = P.default_car()
    c.ball_hit_valid = valid
    c.ball_hit_tick = tick_hit
    return c


def test_touched_in_step_boundary():
    """Player.cpp:18-23: tickCountWhenHit >= tickCount - tickSkip.  A hit at
    tick - tick_skip counts, one tick earlier does not; an invalid hit never
    counts; with tickCount < tickSkip the uint64 subtraction wraps and only a
    hit tick at or past the wrapped value counts."""
    assert S.touched_in_step(_hit_car(1000 - 8), 1000, 8)
    assert not S.touched_in_step(_hit_car(1000 - 8 - 1), 1000, 8)
    assert S.touched_in_step(_hit_car(999), 1000, 8)
    assert not S.touched_in_step(_hit_car(1000 - 8, valid=0), 1000, 8)
    assert not S.touched_in_step(_hit_car(2), 3, 8)
    assert S.touched_in_step(_hit_car(2**64 - 5), 3, 8)
    assert not S.touched_in_step(_hit_car(2**64 - 6), 3, 8)


def test_has_flip_or_jump_boundary():
    """Car.cpp:285-289: on the ground, or no flip / double jump yet and
    airTimeSinceJump < 1.25 s (strict)."""
    c = P.default_car(on_ground=False)
    c.air_time_since_jump = 1.25
    assert not S.has_flip_or_jump(c)
    c.air_time_since_jump = float(np.nextafter(np.float32(1.25), np.float32(0)))
    assert S.has_flip_or_jump(c)
    c.has_flipped = 1
    assert not S.has_flip_or_jump(c)
    c.has_flipped, c.has_double_jumped = 0, 1
    assert not S.has_flip_or_jump(c)
    c.is_on_ground = 1
    c.air_time_since_jump = 5.0
    assert S.has_flip_or_jump(c)


def test_other_players_follow_the_oracles_car_order(oracle):
    """DefaultObs lists teammates, then opponents, in GameState::players
    order (DefaultObs.cpp:36-50), which follows Arena::_cars: the order the
    oracle's Arena::Step visits the cars (ref_car_order, last-added first)."""
    for team in (2, 3):
        h = C.c_void_p(oracle.ref_arena_new(team, team))
        buf = (C.c_uint32 * 8)()
        n = oracle.ref_car_order(h, buf, 8)
        oracle.ref_arena_free(h)
        order = [i - 1 for i in buf[:n]]  # car slots, visiting order
        cars = [P.default_car(pos=(100.0 * (i + 1), -50.0 * i, 17), team=0 if i < team else 1) for i in range(n)]
        ball = P.ball()
        pads = R.PADS()
        prev = np.zeros((n, 8), np.float32)
        for me in range(n):
            row = S.default_obs(cars, ball, pads, prev, me)
            assert row.shape == (51 + 19 * n,)
            xs = [round(float(row[51 + 19 * k]) * 4096 * (-1 if cars[me].team else 1)) for k in range(n)]
            mates = [j for j in order if j != me and cars[j].team == cars[me].team]
            opps = [j for j in order if cars[j].team != cars[me].team]
            assert xs == [100 * (j + 1) for j in [me] + mates + opps], (team, me, xs)


# ---- AdvancedObs ------------------------------------------------------------
# The five states of the reference's own AdvancedObs parity scenarios
# (RocketSimCuda/tests/cpu_parity_harness.cpp:667-714, builders :484-518):
# every car with boost 47.5, on the ground, last controls (0.7, -0.3, 0.15,
# -0.85, 1, jump 0, boost 1, handbrake 0), a yaw-only basis forward (c, s, 0),
# right (s, -c, 0), up (0, 0, 1); pad i active when i % 3 == 0, otherwise
# cooldown 0.25 (i + 1); viewer = car 0.
def _h_car(pos, yaw, vel, ang, team):
    c = R.CarState()
    c.pos.set(pos)
    cy, sy = np.float32(np.cos(np.float32(yaw))), np.float32(np.sin(np.float32(yaw)))
    c.rot.forward.set((cy, sy, 0))
    c.rot.right.set((sy, -cy, 0))
    c.rot.up.set((0, 0, 1))
    c.vel.set(vel)
    c.ang_vel.set(ang)
    c.boost, c.is_on_ground, c.team = 47.5, 1, team
    lc = c.last_controls
    lc.throttle, lc.steer, lc.pitch, lc.yaw, lc.roll, lc.jump, lc.boost, lc.handbrake = 0.7, -0.3, 0.15, -0.85, 1.0, 0, 1, 0
    return c


def _harness_states():
    pi = np.pi
    k = [_h_car((-2048, -2560, 17), pi / 4, (0, 0, 0), (0, 0, 0), 0), _h_car((2048, 2560, 17), pi * 1.25, (0, 0, 0), (0, 0, 0), 1)]
    m = [_h_car((120, -300, 950), 1.1, (600, -250, 420), (1.2, -0.7, 2.1), 0),
         _h_car((-800, 200, 60), 0.2, (-150, 300, 0), (0, 0, 0.4), 0),
         _h_car((900, 1200, 300), 2.7, (200, -400, 100), (0.3, 0.1, -1.0), 1),
         _h_car((1500, -900, 17), -1.4, (50, 80, 0), (0, 0, 0), 1)]
    m[0].is_on_ground, m[0].air_time_since_jump, m[0].has_jumped = 0, 0.4, 1
    m[2].is_on_ground, m[2].has_flipped, m[2].has_double_jumped = 0, 1, 1
    b = [_h_car((-400, 1800, 17), -0.6, (-300, 700, 0), (0, 0, 0), 1), _h_car((300, 1500, 17), 2.1, (100, -200, 0), (0, 0, 0), 0)]
    d = [_h_car((0, 0, 0), 0.0, (0, 0, 0), (0, 0, 0), 0), _h_car((1200, -600, 17), 0.9, (400, 0, 0), (0, 0, 0), 1)]
    d[0].rot.right.set((0, 1, 0))  # identity basis
    d[0].is_demoed, d[0].demo_respawn_timer, d[0].boost, d[0].is_on_ground = 1, 2.4, 0.0, 0
    g = [_h_car((200, 4600, 17), pi / 2, (300, 1400, 0), (0, 0, 0), 0), _h_car((-150, 4950, 30), -pi / 2, (0, -200, 0), (0, 0, 0), 1)]
    return [("kickoff_1v1", k, P.ball((0, 0, 93.15))),
            ("midair_2v2", m, P.ball((-50, 400, 1100), (300, 900, -200), (2.5, -1.1, 0.8))),
            ("post_bounce_inv", b, P.ball((-120, 1600, 140), (80, -260, 950), (5.9, -4.2, 3.3))),
            ("demo", d, P.ball((500, -200, 300), (-100, 150, 50), (0.4, 0.2, -0.1))),
            ("near_goal", g, P.ball((120, 5050, 320), (200, 1500, -120), (1.0, -2.0, 0.5)))]


def _harness_pads():
    """Arena pad records such that RLGym slot i holds the harness's pad i."""
    pads = R.PADS()
    for i in range(34):
        p = pads[S.PAD_MAP[i]]
        p.is_active = 1 if i % 3 == 0 else 0
        p.cooldown = 0.0 if i % 3 == 0 else 0.25 * (i + 1)
    return pads


def _closed_form_block(c, inv, bpos, bvel):
    """float64, for a car with a yaw-only basis (f = (cy, sy, 0) before the
    mirror): local coordinates by the 2D rotation formula."""
    sgn = np.array([-1.0, -1.0, 1.0]) if inv else np.ones(3)
    v = lambda x: np.array([x.x, x.y, x.z], dtype=np.float64) * sgn  # noqa: E731
    pos, vel, ang, f, r, u = v(c.pos), v(c.vel), v(c.ang_vel), v(c.rot.forward), v(c.rot.right), v(c.rot.up)
    cy, sy = f[0], f[1]
    rx, ry = r[0], r[1]

    def local(d):  # forward, right, up components; up = (0, 0, 1)
        return np.array([cy * d[0] + sy * d[1], rx * d[0] + ry * d[1], d[2]])
    hfj = bool(c.is_on_ground) or (not c.has_flipped and not c.has_double_jumped and c.air_time_since_jump < 1.25)
    return np.concatenate([pos / 5000, f, u, vel / 2300, ang / 3, local(ang) / 3, local(bpos - pos) / 5000,
                           local(bvel - vel) / 2300, [c.boost / 100, c.is_on_ground, hfj, c.is_demoed, c.has_jumped]])


@pytest.mark.parametrize("case", range(5))
def test_advanced_obs_on_the_reference_harness_states(case):
    """The restatement on the harness's five states against the closed form:
    row length 51 + 29 n, ball and pad blocks, then every player block with
    local vectors from the 2D rotation (teammates then opponents in slot
    order), within the harness's own 1e-6 tolerance."""
    name, cars, ball = _harness_states()[case]
    pads = _harness_pads()
    row = S.advanced_obs(cars, ball, pads, 0).astype(np.float64)
    n = len(cars)
    assert row.shape == (51 + 29 * n,), name
    inv = cars[0].team != 0
    sgn = np.array([-1.0, -1.0, 1.0]) if inv else np.ones(3)
    bp = np.array([ball.pos.x, ball.pos.y, ball.pos.z]) * sgn
    bv = np.array([ball.vel.x, ball.vel.y, ball.vel.z]) * sgn
    ba = np.array([ball.ang_vel.x, ball.ang_vel.y, ball.ang_vel.z]) * sgn
    pad = [1.0 if i % 3 == 0 else 1.0 / (1.0 + 0.25 * (i + 1)) for i in range(34)]
    # an orange viewer reads the slots mirrored (c_obsPadMapInv[i] = padMap[33 - i])
    want = [bp / 5000, bv / 2300, ba / 3, [0.7, -0.3, 0.15, -0.85, 1.0, 0.0, 1.0, 0.0], pad[::-1] if inv else pad]
    order = [0] + [j for j in range(1, n) if cars[j].team == cars[0].team] + \
            [j for j in range(1, n) if cars[j].team != cars[0].team]
    want += [_closed_form_block(cars[j], inv, bp, bv) for j in order]
    want = np.concatenate([np.asarray(w, np.float64) for w in want])
    np.testing.assert_allclose(row, want, rtol=1e-6, atol=1e-6, err_msg=name)


def test_advanced_obs_mirror_and_frame_invariants():
    """An orange viewer of the point-mirrored world sees what the blue viewer
    sees; local angular velocity keeps the angular velocity's length; the
    local ball offset keeps the distance to the ball."""
    name, cars, ball = _harness_states()[1]
    pads = _harness_pads()
    mc, mb, mp = _mirror(cars, ball, pads)
    for i in range(len(cars)):
        np.testing.assert_allclose(S.advanced_obs(cars, ball, pads, i), S.advanced_obs(mc, mb, mp, i), atol=1e-7)
    row = S.advanced_obs(cars, ball, pads, 0)
    me = row[51:80]
    np.testing.assert_allclose(np.linalg.norm(me[15:18]), np.linalg.norm(me[12:15]), rtol=1e-5)
    d = np.array([ball.pos.x - cars[0].pos.x, ball.pos.y - cars[0].pos.y, ball.pos.z - cars[0].pos.z]) / 5000
    np.testing.assert_allclose(np.linalg.norm(me[18:21]), np.linalg.norm(d), rtol=1e-5)
