"""The fused env step (rsx_step_env) on the GPU against the oracle stepped in
lockstep with the same actions, and against the CPU restatement of the RLGym
pieces (oracle/rlgym_spec.py): DefaultObs rows, stub rewards and terminals
each step, then the auto-reset of terminal arenas."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

pytestmark = pytest.mark.gpu
E = R.engine
S = R.load_spec()
W = dict(goal=1, concede=-1, touch=1, vbtg=1, vptb=1)


def _lut():
    return [(c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump, c.boost, c.handbrake)
            for c in E.default_action_lut()]


def _cases():
    a = [P.default_car(pos=(-300, -900, 17), yaw=1.2, vel=(100, 800, 0), boost=60),
         P.default_car(pos=(500, 1200, 17), yaw=-2.0, vel=(-300, -900, 0), team=1, boost=20)]
    b = [P.default_car(pos=(500, -1500, 17), yaw=0.3, boost=100),
         P.default_car(pos=(-2500, 3500, 17), yaw=2.8, vel=(0, -1500, 0), team=1)]
    c = [P.default_car(pos=(-1000, 2500, 17), yaw=0.2, vel=(300, 0, 0)),
         P.default_car(pos=(-3000, -3000, 17), yaw=0.7, team=1)]
    return [(a, P.ball((-150, -520, 93.15))),
            (b, P.ball((1000, 1000, 800), (-500, 300, -200), (1, 0, 2))),
            (c, P.ball((0, 4950, 200), (0, 900, 0)))]  # heading into the orange goal


def _oracle_arena(oracle, cars, ball, tick):
    o = C.c_void_p(oracle.ref_arena_new(1, 1))
    oracle.ref_step(o, 1)
    oracle.ref_set_tick(o, tick)
    for i, s in enumerate(cars):
        oracle.ref_set_car(o, i, C.byref(s))
        oracle.ref_set_wheels(o, i, P.WHEEL_ZERO)
    oracle.ref_set_ball(o, C.byref(ball))
    pads = R.PADS()
    for i in range(34):
        pads[i].is_active = 1
    oracle.ref_set_pads(o, pads)
    return o


def test_env_step_matches_oracle_and_spec(oracle):
    import torch
    lut = _lut()
    cases = _cases()
    n = len(cases)
    b = E.RocketSimBatch(n, 1, seed=3)
    b.ConfigureEnv(no_touch_timeout_s=1.0, auto_reset=False)
    tick0 = 1000
    oas = []
    for a, (cars, ball) in enumerate(cases):
        P_ = R.PADS()
        for i in range(34):
            P_[i].is_active = 1
        b.SetTickCount(a, tick0)
        for i, s in enumerate(cars):
            b.SetCarState(a, i, s)
        b.SetBallState(a, ball)
        b.SetBoostPadStates(a, P_)
        oas.append(_oracle_arena(oracle, cars, ball, tick0))
    off, ln = b.ObsLayout()
    assert list(ln) == [89] * (2 * n)
    rng = np.random.default_rng(5)
    nt = [0.0] * n
    seen_term = set()
    try:
        for step in range(20):
            idx = rng.integers(0, 90, size=2 * n).astype(np.int32)
            if step < 4:
                idx[:] = 21  # full-throttle ground action first
            d = torch.tensor(idx, device="cuda")
            b.StepEnv(d.data_ptr())
            b.Synchronize()
            obs = b.CopyBuiltObs()
            rew = b.CopyBuiltRewards()
            term = b.CopyBuiltTerminals()
            tick = tick0 + 8 * (step + 1)
            for a in range(n):
                o = oas[a]
                for i in range(2):
                    ctl = R.Controls()
                    (ctl.throttle, ctl.steer, ctl.pitch, ctl.yaw, ctl.roll, ctl.jump, ctl.boost,
                     ctl.handbrake) = lut[idx[2 * a + i]]
                    oracle.ref_set_controls(o, i, C.byref(ctl))
                oracle.ref_step(o, 8)
                prev = np.array([lut[idx[2 * a + i]] for i in range(2)], dtype=np.float32)
                ocars = [R.CarState() for _ in range(2)]
                for i in range(2):
                    oracle.ref_get_car(o, i, C.byref(ocars[i]))
                ob = R.BallState()
                oracle.ref_get_ball(o, C.byref(ob))
                op = R.PADS()
                oracle.ref_get_pads(o, op)
                ecars = [b.GetCarState(a, i) for i in range(2)]
                eb, ep = b.GetBallState(a), b.GetBoostPadStates(a)
                got = np.stack([obs[off[2 * a + i]:off[2 * a + i] + 89] for i in range(2)])
                for name, cars_, ball_, pads_, tol in (("engine", ecars, eb, ep, 2e-6), ("oracle", ocars, ob, op, 2e-4)):
                    touched = [S.touched_in_step(c, tick, 8) for c in cars_]
                    goal = S.ball_scored(ball_)
                    want = np.stack([S.default_obs(cars_, ball_, pads_, prev, i) for i in range(2)])
                    np.testing.assert_allclose(got, want, atol=tol, err_msg=f"arena {a} step {step} obs vs {name}")
                    np.testing.assert_allclose(rew[2 * a:2 * a + 2], S.stub_rewards(cars_, ball_, goal, touched, W),
                                               atol=2 * tol, err_msg=f"arena {a} step {step} rewards vs {name}")
                    if name == "oracle":
                        t, nt[a] = S.step_terminal(goal, any(touched), 8, nt[a], 1.0)
                        assert term[a] == t, (a, step)
                        seen_term.add(int(t))
        assert {1, 2} & seen_term, seen_term
    finally:
        for o in oas:
            oracle.ref_arena_free(o)
        b.Destroy()


def test_auto_reset_after_goal():
    import torch
    b = E.RocketSimBatch(4, [1, 2, 3, 1], seed=8)
    b.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=True)
    b.SetBallState(1, P.ball((0, 5150, 300), (0, 2500, 0)))  # scores for blue within a step
    d = torch.zeros(b.num_cars, dtype=torch.int32, device="cuda")
    b.StepEnv(d.data_ptr())
    b.Synchronize()
    term = b.CopyBuiltTerminals()
    rew = b.CopyBuiltRewards()
    assert list(term) == [0, 1, 0, 0]
    info = b.GetArenaInfo(1)
    assert info.tick_count == 0 and info.episode == 1 and info.goal_scored == 0
    ball = b.GetBallState(1)
    assert ball.pos.x == 0 and ball.pos.y == 0 and abs(ball.pos.z - 93.15) < 1e-3
    # arena 1 is 2v2 (cars 2..5): blue +1, orange -1 (goal term of the reward sum)
    r = rew[2:6]
    teams = [b.GetCarState(1, i).team for i in range(4)]
    for t, v in zip(teams, r):
        assert (v > 0.5) if t == 0 else (v < -0.5)
    # the obs rows of the reset arena are its kickoff with zero previous actions
    off, ln = b.ObsLayout()
    obs = b.CopyBuiltObs()
    cars = [b.GetCarState(1, i) for i in range(4)]
    pads = b.GetBoostPadStates(1)
    for i in range(4):
        want = S.default_obs(cars, ball, pads, np.zeros((4, 8), np.float32), i)
        assert ln[2 + i] == 127
        np.testing.assert_allclose(obs[off[2 + i]:off[2 + i] + 127], want, atol=2e-6)
    b.Destroy()


def test_action_masks_match_spec(torch_cuda):
    """DefaultAction masks written by the env step equal the restatement's
    GetActionMask on the same post-step car records (grounded and airborne
    cars, empty tanks, cars on their roofs)."""
    import torch
    from emy_s_framework_amd import workloads as WL
    n = 96
    b = E.RocketSimBatch(n, [1, 2, 3] * (n // 3), seed=5)
    WL.random_field_states(b, 5)
    cars = b.GetAllCarStates()
    for k in range(0, b.num_cars, 3):
        cars[k].boost = 0.0
    for k in range(1, b.num_cars, 5):  # airborne
        cars[k].pos.z = 600.0
        cars[k].is_on_ground = 0
    for k in range(2, b.num_cars, 7):  # on its roof
        cars[k].pos.z = 40.0
        cars[k].rot.up.z, cars[k].rot.right.x, cars[k].rot.right.y = -1.0, -cars[k].rot.right.x, -cars[k].rot.right.y
        cars[k].is_on_ground = 0
    b.SetAllCarStates(cars)
    lut = _lut()
    masks = S.default_action_masks(lut)
    b.SetActionMasks(True)
    g = np.random.default_rng(9)
    seen_roof = False
    for step in range(4):
        d = torch.tensor(g.integers(0, 90, b.num_cars).astype(np.int32), device="cuda")
        b.StepEnv(d.data_ptr())
        got = b.CopyBuiltDefaultActionMasks()
        recs = b.GetAllCarStates()
        for k in range(b.num_cars):
            want = S.action_mask(recs[k], masks)
            assert list(got[k]) == want, (step, k)
            seen_roof |= bool(recs[k].world_contact and recs[k].world_contact_normal.z > 0.9)
    b.BuildObs()
    got = b.CopyBuiltDefaultActionMasks()
    recs = b.GetAllCarStates()
    assert all(list(got[k]) == S.action_mask(recs[k], masks) for k in range(b.num_cars))
    assert seen_roof
    # with masks on, only the default table may be uploaded (the masks index it)
    with pytest.raises(E.RsxError):
        b.UploadDiscreteActionLut((E.CarControls * 40)(*E.default_action_lut()[:40]))
    b.UploadDiscreteActionLut(E.default_action_lut())
    b.SetActionMasks(False)
    b.UploadDiscreteActionLut((E.CarControls * 40)(*E.default_action_lut()[:40]))
    with pytest.raises(E.RsxError):
        b.SetActionMasks(True)
    b.Destroy()


def test_advanced_obs_matches_spec(torch_cuda):
    """The env step with the AdvancedObs builder (BuildAdvancedObsAndDefault-
    Masks' layout): every row of a mixed 1v1/2v2/3v3 batch equals the spec
    restatement on the post-step records, for several steps and after
    rsx_build_obs; the layout has 51 + 29 floats per car and switching back
    restores DefaultObs rows."""
    import torch
    from emy_s_framework_amd import workloads as WL
    n = 90
    b = E.RocketSimBatch(n, [1, 2, 3] * (n // 3), seed=6)
    WL.random_field_states(b, 6)
    b.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=True, obs_builder=E.OBS_ADVANCED)
    off, ln = b.ObsLayout()
    teams = [b.team_size[a] for a in range(n) for _ in range(2 * b.team_size[a])]
    assert list(ln) == [51 + 29 * 2 * t for t in teams]
    g = np.random.default_rng(3)
    for step in range(4):
        d = torch.tensor(g.integers(0, 90, b.num_cars).astype(np.int32), device="cuda")
        b.StepEnv(d.data_ptr())
        obs = b.CopyBuiltObs()
        recs = b.GetAllCarStates()
        balls = b.GetAllBallStates()
        allp = b.GetAllBoostPadStates()
        sz = C.sizeof(E.BoostPadState)
        raw = bytes(allp)
        o = 0
        for a in range(n):
            k = 2 * b.team_size[a]
            cars = [recs[o + i] for i in range(k)]
            pp = R.PADS.from_buffer_copy(raw[a * 34 * sz:(a + 1) * 34 * sz])
            for i in range(k):
                np.testing.assert_allclose(obs[off[o + i]:off[o + i] + ln[o + i]],
                                           S.advanced_obs(cars, balls[a], pp, i), atol=2e-6,
                                           err_msg=f"step {step} arena {a} car {i}")
            o += k
    b.BuildObs()
    obs = b.CopyBuiltObs()
    cars = [b.GetCarState(5, i) for i in range(6)]
    for i in range(6):
        c0 = sum(2 * b.team_size[x] for x in range(5))  # arena 5 is 3v3
        np.testing.assert_allclose(obs[off[c0 + i]:off[c0 + i] + 225],
                                   S.advanced_obs(cars, b.GetBallState(5), b.GetBoostPadStates(5), i), atol=2e-6)
    b.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=True, obs_builder=E.OBS_DEFAULT)
    off, ln = b.ObsLayout()
    assert list(ln) == [51 + 19 * 2 * t for t in teams]
    b.Destroy()
