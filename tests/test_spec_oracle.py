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
