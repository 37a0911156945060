"""The GPU training path's configurable rewards, reset curriculum and
per-car controls on the device, against oracle/train_spec.py:

  rewards      every reward id (two tables), contact-heavy 1v1 and mixed
               1v1/2v2/3v3 arenas, several env steps with the auto-reset off;
               each step's device rewards against the restatement evaluated on
               the records before and after the step (to float rounding);
  curriculum   rsx_reset_all_arenas, rsx_reset_arena and the env step's
               auto-reset under weights 1:2:3:4: every reset passes its
               mode's placement rules and the modes follow the weights; the
               no-touch stagger starts timers in [0, 0.9 x timeout);
  controls     rsx_set_car_controls car by car gives the states
               rsx_set_all_controls gives."""
import ctypes as C
import importlib.util
import math
import os

import numpy as np
import pytest

import rsxlib as R
import train_harness as TH

pytestmark = pytest.mark.gpu
E = R.engine
_spec = importlib.util.spec_from_file_location("train_spec", os.path.join(R.ROOT, "oracle", "train_spec.py"))
T = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(T)
W = importlib.import_module("emy_s_framework_amd.workloads")


def _copy(arr, cls):
    return [cls.from_buffer_copy(x) for x in arr]


def _arena_slices(b):
    o, out = 0, []
    for a in range(b.num_arenas):
        n = 2 * b.team_size[a]
        out.append((o, n))
        o += n
    return out


def _craft_contacts(b):
    """In every other arena: the ball at rest 300 uu up, car 0 flying into it
    at 2000 uu/s from 200 uu away, and car 1 (an opponent in 1v1) flying into
    car 0 from the side, so touches and bumps happen in the next env step."""
    cars, balls = b.GetAllCarStates(), b.GetAllBallStates()
    for a, (o, n) in enumerate(_arena_slices(b)):
        if a % 2:
            continue
        bp = balls[a].pos
        bp.x, bp.y, bp.z = max(-3000.0, min(3000.0, bp.x)), max(-4000.0, min(4000.0, bp.y)), 300.0
        balls[a].vel.set((0, 0, 0))
        balls[a].ang_vel.set((0, 0, 0))
        c0 = cars[o]
        c0.pos.set((bp.x - 200, bp.y, 300))
        c0.vel.set((2000, 0, 0))
        c0.is_on_ground = 0
        c1 = cars[o + 1]
        c1.pos.set((bp.x - 200, bp.y + 140, 300))
        c1.vel.set((0, -1800, 0))
        c1.is_on_ground = 0
    b.SetAllCarStates(cars)
    b.SetAllBallStates(balls)


@pytest.mark.parametrize("kind", ["stress", "mixed"])
def test_training_rewards_match_restatement(kind, torch_cuda):
    torch = torch_cuda
    if kind == "stress":
        b = E.RocketSimBatch(256, 1, seed=3)
        W.stress_states(b, 4)
    else:
        b = E.RocketSimBatch(192, W.mixed_team_sizes(192, 2), seed=3)
        W.random_field_states(b, 2)
    fired, worst, checked = set(), 0.0, 0
    try:
        b.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=False)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(5)
        sl = _arena_slices(b)
        for ti, table in enumerate(TH.tables(17)):
            _craft_contacts(b)
            b.ConfigureTrainingRewards(table)
            assert len(b.GetTrainingRewards()) == len(table)
            for step in range(4):
                print(f"{kind}: table {ti} step {step}", flush=True)
                pc, pb = _copy(b.GetAllCarStates(), R.CarState), _copy(b.GetAllBallStates(), R.BallState)
                d = torch.randint(0, 90, (b.num_cars,), dtype=torch.int32, device="cuda", generator=gen)
                b.StepEnv(d.data_ptr())
                b.Synchronize()
                rew = b.CopyBuiltRewards()
                cc, bb = _copy(b.GetAllCarStates(), R.CarState), _copy(b.GetAllBallStates(), R.BallState)
                infos = b.GetAllArenaInfos()
                for a, (o, n) in enumerate(sl):
                    inf = infos[a]
                    gt = inf.goal_team if inf.goal_scored else -1
                    want = T.rewards(table, cc[o:o + n], pc[o:o + n], bb[a], pb[a], gt, inf.tick_count, 8)
                    for i in range(n):
                        err = abs(float(rew[o + i]) - float(want[i]))
                        assert err <= 2e-5 * max(1.0, abs(float(want[i]))), (kind, ti, step, a, i, rew[o + i], want[i])
                        worst = max(worst, err)
                        checked += 1
                    for e in table:
                        for i in range(n):
                            if float(T.term(e, cc[o:o + n], pc[o:o + n], bb[a], pb[a], i, gt, inf.tick_count,
                                            8)) != 0:
                                fired.add(e["id"])
        b.ConfigureTrainingRewards([])
        assert b.GetTrainingRewards() == []
    finally:
        b.Destroy()
    names = sorted(T.IDS[k] for k in fired)
    print(f"{kind}: {checked} car-rewards, worst |device - restatement| {worst:.2e}; ids that fired: {names}")
    # every branch of every id is compared on random records with the CPU
    # build (test_train_cpu.py); here the ids are the ones real steps reach
    core = {"TOUCH_BALL", "BUMP", "BUMPED_PENALTY", "AIR", "SPEED", "FACE_BALL", "VELOCITY_PLAYER_TO_BALL",
            "VELOCITY_BALL_TO_GOAL"}
    assert core <= set(names), core - set(names)


def _check_resets(cars, balls, sl, arenas):
    counts = [0, 0, 0, 0]
    for a in arenas:
        o, n = sl[a]
        cs = cars[o:o + n]
        mode = T.reset_mode(cs, balls[a])
        bad = T.check_reset(mode, cs, balls[a])
        assert bad == [], (a, mode, bad)
        counts[mode] += 1
    return counts


def _within_weights(counts, w):
    tot, sw = sum(counts), sum(w)
    for c, x in zip(counts, w):
        p = x / sw
        assert abs(c - tot * p) < 5 * math.sqrt(tot * p * (1 - p)) + 1, (counts, w)


def test_reset_curriculum_on_the_device(torch_cuda):
    torch = torch_cuda
    n = 4096
    b = E.RocketSimBatch(n, W.mixed_team_sizes(n, 9), seed=21)
    try:
        sl = _arena_slices(b)
        w = [1.0, 2.0, 3.0, 4.0]
        b.SetResetCurriculum(*w)
        assert b.GetResetCurriculum() == w
        b.ConfigureEnv(no_touch_timeout_s=10.0, no_touch_stagger=True)
        b.ResetAllArenas()
        b.Synchronize()
        cars, balls = b.GetAllCarStates(), b.GetAllBallStates()
        counts = _check_resets(cars, balls, sl, range(n))
        print(f"rsx_reset_all_arenas: modes {counts}", flush=True)
        _within_weights(counts, w)
        nt = np.array([i.no_touch_time for i in b.GetAllArenaInfos()])
        assert nt.min() >= 0 and nt.max() < 9.0 and nt.std() > 1.0, (nt.min(), nt.max(), nt.std())
        # the host-side single-arena reset
        for a in range(0, n, 97):
            b.ResetArena(a)
        cars, balls = b.GetAllCarStates(), b.GetAllBallStates()
        _check_resets(cars, balls, sl, range(0, n, 97))
        # the env step's auto-reset: a 0.1 s no-touch limit ends most episodes
        # within two steps; terminal arenas come back reset
        b.ConfigureEnv(no_touch_timeout_s=0.1, no_touch_stagger=False)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(3)
        seen = [0, 0, 0, 0]
        for step in range(3):
            d = torch.randint(0, 90, (b.num_cars,), dtype=torch.int32, device="cuda", generator=gen)
            b.StepEnv(d.data_ptr())
            term = b.CopyBuiltTerminals()
            cars, balls = b.GetAllCarStates(), b.GetAllBallStates()
            reset = [a for a in range(n) if term[a]]
            c = _check_resets(cars, balls, sl, reset)
            seen = [x + y for x, y in zip(seen, c)]
            infos = b.GetAllArenaInfos()
            assert all(infos[a].tick_count == 0 and infos[a].no_touch_time == 0 for a in reset)
        print(f"auto-reset: modes {seen}", flush=True)
        assert sum(seen) > 1000
        _within_weights(seen, w)
        # back to kickoff only
        b.SetResetCurriculum(-1.0, float("nan"), 0.0, 0.0)
        assert b.GetResetCurriculum() == [1.0, 0.0, 0.0, 0.0]
        b.ResetAllArenas()
        cars, balls = b.GetAllCarStates(), b.GetAllBallStates()
        assert _check_resets(cars, balls, sl, range(n))[0] == n
    finally:
        b.Destroy()


def test_set_car_controls_matches_set_all(torch_cuda):
    n = 96
    teams = W.mixed_team_sizes(n, 4)
    b1, b2 = E.RocketSimBatch(n, teams, seed=8), E.RocketSimBatch(n, teams, seed=8)
    try:
        for b in (b1, b2):
            W.random_field_states(b, 6)
        rng = np.random.default_rng(1)
        ctrl = (E.CarControls * b1.num_cars)()
        for k in range(b1.num_cars):
            ctrl[k] = E.CarControls(*[float(x) for x in rng.uniform(-1, 1, 5)],
                                    *[int(x) for x in rng.integers(0, 2, 3)])
        b1.SetAllCarControls(ctrl)
        zero = (E.CarControls * b2.num_cars)()
        b2.SetAllCarControls(zero)
        o = 0
        for a in range(n):
            for i in range(2 * teams[a]):
                b2.SetCarControls(a, i, ctrl[o + i])
            o += 2 * teams[a]
        with pytest.raises(E.RsxError):
            b2.SetCarControls(0, 2 * teams[0], ctrl[0])
        with pytest.raises(E.RsxError):
            b2.SetCarControls(n, 0, ctrl[0])
        for b in (b1, b2):
            b.Step(8)
        assert bytes(b1.GetAllCarStates()) == bytes(b2.GetAllCarStates())
        assert bytes(b1.GetAllBallStates()) == bytes(b2.GetAllBallStates())
    finally:
        b1.Destroy()
        b2.Destroy()
