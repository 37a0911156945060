"""Env outputs (DefaultObs rows, stub rewards, terminals) of the engine's own
env code compiled for the host (tests/cpu_check cc_env) against the CPU
restatement (oracle/rlgym_spec.py): on the engine's states to float rounding,
and on the oracle's states run in lockstep to the physics tolerance carried
through the obs scaling."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

S = R.load_spec()
W = dict(goal=1, concede=-1, touch=1, vbtg=1, vptb=1)


def _lut():
    return [(c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump, c.boost, c.handbrake)
            for c in R.engine.default_action_lut()]


@pytest.mark.parametrize("seed", [0, 1])
def test_env_outputs_match_spec(seed, oracle, cpucheck):
    if not __import__("os").path.exists(R.engine.LIB_PATH):
        pytest.skip("librsx.so is not built")
    lib, world = cpucheck
    lut = _lut()
    rng = np.random.default_rng(seed)
    cars = [P.default_car(pos=(-300, -900, 17), yaw=1.2, vel=(100, 800, 0), boost=60),
            P.default_car(pos=(500, 1200, 17), yaw=-2.0, vel=(-300, -900, 0), team=1, boost=20)]
    ball = P.ball((-150, -520, 93.15))  # in the blue car's path
    case = dict(blue=1, orange=1, cars=cars, ball=ball, ticks=0, controls=[P.controls(), P.controls()])
    # lockstep: cpucheck engine and oracle, same states, 8-tick steps
    h = C.c_void_p(lib.cc_arena_new(world, 1, 1))
    o = C.c_void_p(oracle.ref_arena_new(1, 1))
    oracle.ref_step(o, 1)
    tick0 = 1000
    lib.cc_set_tick(h, tick0)
    oracle.ref_set_tick(o, tick0)
    pads = R.PADS()
    for i in range(34):
        pads[i].is_active = 1
    for i, s in enumerate(cars):
        lib.cc_set_car(h, i, C.byref(s))
        lib.cc_set_wheels(h, i, P.WHEEL_ZERO6)
        oracle.ref_set_car(o, i, C.byref(s))
        oracle.ref_set_wheels(o, i, P.WHEEL_ZERO)
    lib.cc_set_ball(h, C.byref(ball))
    oracle.ref_set_ball(o, C.byref(ball))
    lib.cc_set_pads(h, pads)
    oracle.ref_set_pads(o, pads)
    nt_eng = C.c_float(0)
    nt_spec_e = nt_spec_o = 0.0
    obs = (C.c_float * (2 * 89))()
    rew = (C.c_float * 2)()
    try:
        seen = set()
        for step in range(24):
            idx = rng.integers(0, 90, size=2) if step > 3 else np.array([21, 20])  # drive at first
            prev = np.array([lut[k] for k in idx], dtype=np.float32)
            for i in range(2):
                ctl = R.Controls()
                ctl.throttle, ctl.steer, ctl.pitch, ctl.yaw, ctl.roll, ctl.jump, ctl.boost, ctl.handbrake = \
                    lut[idx[i]]
                lib.cc_set_controls(h, i, C.byref(ctl))
                oracle.ref_set_controls(o, i, C.byref(ctl))
            lib.cc_step(h, 8)
            oracle.ref_step(o, 8)
            pv = (C.c_float * 16)(*prev.ravel())
            term = lib.cc_env(h, pv, 8, 1.0, C.byref(nt_eng), obs, rew)
            ecars = [R.CarState() for _ in range(2)]
            ocars = [R.CarState() for _ in range(2)]
            for i in range(2):
                lib.cc_get_car(h, i, C.byref(ecars[i]))
                oracle.ref_get_car(o, i, C.byref(ocars[i]))
            eb, ob = R.BallState(), R.BallState()
            lib.cc_get_ball(h, C.byref(eb))
            oracle.ref_get_ball(o, C.byref(ob))
            ep, op = R.PADS(), R.PADS()
            lib.cc_get_pads(h, ep)
            oracle.ref_get_pads(o, op)
            tick = tick0 + 8 * (step + 1)
            got = np.array(obs, dtype=np.float32).reshape(2, 89)
            for states, cars_, ball_, pads_, tol in (("engine", ecars, eb, ep, 2e-6), ("oracle", ocars, ob, op, 2e-4)):
                touched = [S.touched_in_step(c, tick, 8) for c in cars_]
                goal = S.ball_scored(ball_)
                want_obs = np.stack([S.default_obs(cars_, ball_, pads_, prev, i) for i in range(2)])
                np.testing.assert_allclose(got, want_obs, atol=tol, err_msg=f"obs vs spec on {states} states, step {step}")
                want_rew = S.stub_rewards(cars_, ball_, goal, touched, W)
                np.testing.assert_allclose(np.array(rew), want_rew, atol=tol * 2, err_msg=f"rewards, {states}, step {step}")
                if states == "engine":
                    t, nt_spec_e = S.step_terminal(goal, any(touched), 8, nt_spec_e, 1.0)
                else:
                    t, nt_spec_o = S.step_terminal(goal, any(touched), 8, nt_spec_o, 1.0)
                assert term == t, (states, step)
            seen.add(("touch", any(S.touched_in_step(c, tick, 8) for c in ecars)))
            seen.add(("term", term))
        # the run covers a touch and a no-touch truncation
        assert ("touch", True) in seen and ("term", 2) in seen, seen
    finally:
        lib.cc_arena_free(h)
        oracle.ref_arena_free(o)


def _env_arena(lib, world, cars, ball, tick):
    h = C.c_void_p(lib.cc_arena_new(world, 1, 1))
    lib.cc_set_tick(h, tick)
    for i, s in enumerate(cars):
        lib.cc_set_car(h, i, C.byref(s))
    lib.cc_set_ball(h, C.byref(ball))
    pads = R.PADS()
    for i in range(34):
        pads[i].is_active = 1
    lib.cc_set_pads(h, pads)
    return h


def test_engine_touch_and_flip_boundaries(cpucheck):
    """The engine's own touch rule, no-touch timer and HasFlipOrJump
    (rsx_env.h, host build) at the boundaries the spec pins: a hit at
    tick - tick_skip counts and one a tick earlier does not (Player.cpp:18-23);
    a touch clears the no-touch timer; airTimeSinceJump < 1.25 s is strict
    (Car.cpp:285-289)."""
    lib, world = cpucheck
    tick = 5000
    ball = P.ball((0, 0, 93.15))
    obs = (C.c_float * (2 * 89))()
    rew = (C.c_float * 2)()
    prev = (C.c_float * 16)()
    for hit, want in ((tick - 8, True), (tick - 9, False), (tick - 1, True)):
        blue = P.default_car(pos=(-3000, -4000, 17))
        blue.ball_hit_valid, blue.ball_hit_tick = 1, hit
        orange = P.default_car(pos=(3000, 4000, 17), team=1)
        h = _env_arena(lib, world, [blue, orange], ball, tick)
        nt = C.c_float(0.5)
        term = lib.cc_env(h, prev, 8, 1.0, C.byref(nt), obs, rew)
        lib.cc_arena_free(h)
        assert term == 0
        assert (nt.value == 0.0) == want, (hit, nt.value)
        spec_t = S.touched_in_step(blue, tick, 8)
        assert spec_t == want
        _, spec_nt = S.step_terminal(False, spec_t, 8, np.float32(0.5), 1.0)
        assert np.float32(nt.value) == spec_nt
        assert abs(rew[0] - (1.0 if want else 0.0) - S.stub_rewards(
            [blue, orange], ball, False, [False, False], W)[0]) < 1e-6
    # no-touch truncation after exactly 15 steps of 8 ticks at a 1 s limit
    blue = P.default_car(pos=(-3000, -4000, 17))
    orange = P.default_car(pos=(3000, 4000, 17), team=1)
    h = _env_arena(lib, world, [blue, orange], ball, tick)
    nt = C.c_float(0.0)
    terms = [lib.cc_env(h, prev, 8, 1.0, C.byref(nt), obs, rew) for _ in range(15)]
    lib.cc_arena_free(h)
    assert terms == [0] * 14 + [2] and nt.value == 1.0
    # HasFlipOrJump in the self block (obs index 51 + 17)
    for t_air, want in ((1.25, 0.0), (float(np.nextafter(np.float32(1.25), np.float32(0))), 1.0)):
        blue = P.default_car(pos=(-3000, -4000, 600), on_ground=False)
        blue.air_time_since_jump = t_air
        blue.has_jumped = 1
        h = _env_arena(lib, world, [blue, orange], ball, tick)
        nt = C.c_float(0.0)
        lib.cc_env(h, prev, 8, 0.0, C.byref(nt), obs, rew)
        lib.cc_arena_free(h)
        assert obs[51 + 17] == want, t_air
        assert S.has_flip_or_jump(blue) == bool(want)


def test_oracle_env_step_matches_spec(oracle):
    """The oracle shim's C++ env step (ref_env_step: the CPU baseline's env
    half) against the spec restatement on the oracle's own post-step records:
    1v1, 2v2 and 3v3 arenas, 20 steps of random table actions, no-touch limit
    0.5 s, then a goal with the kickoff auto-reset."""
    lut = R.engine.default_action_lut()
    lt = _lut()
    pad_map = (C.c_int32 * 34)(*[int(x) for x in S.PAD_MAP])
    teams = [1, 2, 3]
    hs = []
    for k, t in enumerate(teams):
        h = oracle.ref_arena_new(t, t)
        oracle.ref_reset_kickoff(C.c_void_p(h), 5 + k)
        hs.append(h)
    arr = (C.c_void_p * len(hs))(*hs)
    ncars = [2 * t for t in teams]
    rows = [51 + 19 * n for n in ncars]
    nt = (C.c_float * len(hs))()
    obs = (C.c_float * sum(n * r for n, r in zip(ncars, rows)))()
    rew = (C.c_float * sum(ncars))()
    term = (C.c_uint8 * len(hs))()
    spec_nt = [np.float32(0)] * len(hs)
    rng = np.random.default_rng(2)
    seen = set()
    try:
        for step in range(20):
            act = rng.integers(0, 90, sum(ncars)).astype(np.int32)
            oracle.ref_env_step(arr, len(hs), act.ctypes.data_as(C.POINTER(C.c_int32)), 8, lut, len(lut), pad_map,
                                C.c_float(0.5), 0, nt, obs, rew, term, 100)
            car0 = of = 0
            for a, h in enumerate(hs):
                n = ncars[a]
                cars = [R.CarState() for _ in range(n)]
                for i in range(n):
                    oracle.ref_get_car(C.c_void_p(h), i, C.byref(cars[i]))
                ball = R.BallState()
                oracle.ref_get_ball(C.c_void_p(h), C.byref(ball))
                pads = R.PADS()
                oracle.ref_get_pads(C.c_void_p(h), pads)
                tick = oracle.ref_get_tick(C.c_void_p(h))
                prev = np.array([lt[j] for j in act[car0:car0 + n]], dtype=np.float32)
                touched = [S.touched_in_step(c, tick, 8) for c in cars]
                goal = S.ball_scored(ball)
                got = np.array(obs[of:of + n * rows[a]], dtype=np.float32).reshape(n, rows[a])
                want = np.stack([S.default_obs(cars, ball, pads, prev, i) for i in range(n)])
                np.testing.assert_array_equal(got, want, err_msg=f"arena {a} step {step}")
                np.testing.assert_allclose(np.array(rew[car0:car0 + n]), S.stub_rewards(cars, ball, goal, touched, W),
                                           atol=1e-6, err_msg=f"arena {a} step {step}")
                t, spec_nt[a] = S.step_terminal(goal, any(touched), 8, spec_nt[a], 0.5)
                assert term[a] == t and np.float32(nt[a]) == spec_nt[a], (a, step)
                seen.add(int(t))
                car0 += n
                of += n * rows[a]
        assert 2 in seen, seen
        # a goal: rewards from the scoring state, then the kickoff and its obs
        b = P.ball((0, 5150, 300), (0, 2500, 0))
        oracle.ref_set_ball(C.c_void_p(hs[0]), C.byref(b))
        act = np.zeros(sum(ncars), np.int32)
        oracle.ref_env_step(arr, 1, act.ctypes.data_as(C.POINTER(C.c_int32)), 8, lut, len(lut), pad_map,
                            C.c_float(0.5), 1, nt, obs, rew, term, 100)
        assert term[0] == 1 and rew[0] > 0.5 and rew[1] < -0.5
        cars = [R.CarState() for _ in range(2)]
        for i in range(2):
            oracle.ref_get_car(C.c_void_p(hs[0]), i, C.byref(cars[i]))
        ball = R.BallState()
        oracle.ref_get_ball(C.c_void_p(hs[0]), C.byref(ball))
        pads = R.PADS()
        oracle.ref_get_pads(C.c_void_p(hs[0]), pads)
        assert abs(ball.pos.y) < 1 and oracle.ref_get_tick(C.c_void_p(hs[0])) == 0
        want = np.stack([S.default_obs(cars, ball, pads, np.zeros((2, 8), np.float32), i) for i in range(2)])
        np.testing.assert_array_equal(np.array(obs[:2 * 89], dtype=np.float32).reshape(2, 89), want)
    finally:
        for h in hs:
            oracle.ref_arena_free(C.c_void_p(h))


def test_engine_advanced_obs_matches_spec(cpucheck):
    """The engine's AdvancedObs rows (rsx_env.h adv_obs_row, host build)
    against the spec restatement on the engine's own post-step records, for
    1v1, 2v2 and 3v3 arenas of the mixed workload after 8 ticks of random
    table actions."""
    import workload_harness as WH
    lib, world = cpucheck
    lut = R.engine.default_action_lut()
    cases = WH.sample("mixed", 24)
    rng = np.random.default_rng(4)
    seen = set()
    for case in cases:
        n = len(case["cars"])
        h = C.c_void_p(lib.cc_arena_new(world, case["blue"], case["orange"]))
        lib.cc_set_tick(h, case["tick"])
        for i, s in enumerate(case["cars"]):
            lib.cc_set_car(h, i, C.byref(s))
            lib.cc_set_wheels(h, i, P.WHEEL_ZERO6)
        lib.cc_set_ball(h, C.byref(case["ball"]))
        lib.cc_set_pads(h, case["pads"])
        for i, k in enumerate(rng.integers(0, 90, n)):
            lib.cc_set_controls(h, i, C.byref(lut[k]))
        lib.cc_step(h, 8)
        row = 51 + 29 * n
        obs = (C.c_float * (n * row))()
        rew = (C.c_float * n)()
        nt = C.c_float(0)
        lib.cc_env_kind(h, (C.c_float * (8 * n))(), 8, 0.0, C.byref(nt), obs, rew, 1)
        cars = [R.CarState() for _ in range(n)]
        for i in range(n):
            lib.cc_get_car(h, i, C.byref(cars[i]))
        ball = R.BallState()
        lib.cc_get_ball(h, C.byref(ball))
        pads = R.PADS()
        lib.cc_get_pads(h, pads)
        lib.cc_arena_free(h)
        got = np.array(obs, dtype=np.float32).reshape(n, row)
        for i in range(n):
            np.testing.assert_allclose(got[i], S.advanced_obs(cars, ball, pads, i), atol=2e-6,
                                       err_msg=f"arena {case['source']} car {i}")
        seen.add(n)
        assert any(not p.is_active for p in pads)  # the pad timer branch is exercised
    assert seen == {2, 4, 6}, seen
