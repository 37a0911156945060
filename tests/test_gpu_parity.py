"""The gfx950 engine through its C-ABI (librsx.so) against the oracle.

Every shared case runs as one arena of a single batch, one physics tick per
launch, and is compared with the oracle tick by tick to the same bar as the
CPU build: positions, velocities and angular velocities per vector within
1e-3 + 1e-4*|reference|, rotation entries within 1e-4, flags, wheel contacts,
pads and the goal flag exact."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R
import scenarios as S

pytestmark = pytest.mark.gpu
E = R.engine


def _all_pads_active():
    p = E.PadStates()
    for i in range(34):
        p[i].is_active = 1
    return p


def load_case(batch, a, case, first_car=0):
    batch.SetTickCount(a, case.get("tick", 1000))
    for i, s in enumerate(case["cars"]):
        batch.SetCarState(a, first_car + i, s)
    batch.SetBallState(a, case["ball"])
    batch.SetBoostPadStates(a, case.get("pads") or _all_pads_active())


def run_engine(batch, cases, ticks_per_launch=1):
    """Runs cases (one per arena, arena i = cases[i]) and records every arena
    after each launch."""
    n = len(cases)
    for a, case in enumerate(cases):
        load_case(batch, a, case)
    res = [P.Result() for _ in cases]
    T = max(c["ticks"] for c in cases)
    ctrl = (E.CarControls * batch.num_cars)()
    for t in range(0, T, ticks_per_launch):
        o = 0
        for a, case in enumerate(cases):
            cs = P._ctrls(case, min(t, case["ticks"] - 1))
            for i in range(2 * batch.team_size[a]):
                ctrl[o + i] = cs[i] if i < len(cs) else E.CarControls()
            o += 2 * batch.team_size[a]
        batch.SetAllCarControls(ctrl)
        batch.Step(ticks_per_launch)
        batch.Synchronize()
        for a, case in enumerate(cases):
            if t >= case["ticks"]:
                continue
            r = res[a]
            r.cars.append([batch.GetCarState(a, i) for i in range(case["blue"] + case["orange"])])
            r.balls.append(batch.GetBallState(a))
            r.pads.append(batch.GetBoostPadStates(a))
            r.scored.append(int(batch.GetArenaInfo(a).goal_scored))
    return res


@pytest.fixture(scope="module")
def parity_runs(oracle):
    cases = S.cases()
    names = list(cases)
    batch = E.RocketSimBatch(len(names), 1, seed=7)
    eng = run_engine(batch, [cases[k] for k in names])
    batch.Destroy()
    out = {}
    for k, e in zip(names, eng):
        ref = P.run_oracle(oracle, cases[k])
        ref.scored = list(np.maximum.accumulate(ref.scored))  # the engine's goal flag is sticky
        out[k] = (ref, e)
    return out


@pytest.mark.parametrize("name", list(S.cases()))
def test_gpu_matches_oracle(name, parity_runs):
    ref, eng = parity_runs[name]
    for t in range(len(ref.balls)):
        fails = P.compare_tick(ref, eng, t)
        assert not fails, f"{name}: tick {t}: {fails}"


def test_tick_skip_launch_equals_single_ticks():
    """k ticks inside one launch give bit-identical state to k launches of one
    tick (cases whose controls hold still over the launch)."""
    cases = [S.cases()[k] for k in ("cars boost", "car hits ball", "ball corner", "cars throttle")]
    for c in cases:
        c["ticks"] = 32
    b1 = E.RocketSimBatch(len(cases), 1)
    b8 = E.RocketSimBatch(len(cases), 1)
    r1 = run_engine(b1, cases, 1)
    r8 = run_engine(b8, cases, 8)
    for a in range(len(cases)):
        x = [bytes(s) for s in r1[a].cars[7::8][-1]] + [bytes(r1[a].balls[7::8][-1])]
        y = [bytes(s) for s in r8[a].cars[-1]] + [bytes(r8[a].balls[-1])]
        assert x == y
    b1.Destroy()
    b8.Destroy()


@pytest.mark.parametrize("team", [2, 3])
def test_team_sizes_match_oracle(team, oracle):
    """2v2 and 3v3 arenas (the 4- and 6-car kernels) from the engine's own
    kickoff, stepped with per-car controls, against the oracle loaded with the
    same states."""
    batch = E.RocketSimBatch(3, [1, team, team], seed=11)
    rng = np.random.default_rng(team)
    n = 2 * team
    cars = [batch.GetCarState(1, i) for i in range(n)]
    case = dict(blue=team, orange=team, cars=cars, ball=batch.GetBallState(1), ticks=90,
                controls=[P.controls(throttle=float(rng.uniform(0.3, 1)), steer=float(rng.uniform(-1, 1)),
                                     boost=int(rng.integers(0, 2))) for _ in range(n)])
    case2 = dict(case, cars=[batch.GetCarState(2, i) for i in range(n)], ball=P.ball((200, -300, 600), (0, 0, -500)),
                 controls=lambda t: [P.controls(throttle=1, jump=1 if 20 < t < 30 else 0, yaw=0.5)] * n)
    eng = run_engine(batch, [S.cases()["ball rest"], case, case2])
    for c, e in ((case, eng[1]), (case2, eng[2])):
        ref = P.run_oracle(oracle, c)
        ref.scored = list(np.maximum.accumulate(ref.scored))
        for t in range(c["ticks"]):
            fails = P.compare_tick(ref, e, t)
            assert not fails, f"{team}v{team} tick {t}: {fails}"
    batch.Destroy()


def test_large_batch_deterministic_and_finite():
    """4096 1v1 arenas from seeded kickoffs, 240 ticks of random controls: two
    engines with the same seed agree bit for bit and states stay finite."""
    n = 4096
    rng = np.random.default_rng(3)
    outs = []
    for _ in range(2):
        b = E.RocketSimBatch(n, 1, seed=5)
        ctrl = (E.CarControls * b.num_cars)()
        for step in range(30):
            vals = rng.uniform(-1, 1, size=(b.num_cars, 5))
            for i in range(b.num_cars):
                c = ctrl[i]
                c.throttle, c.steer, c.pitch, c.yaw, c.roll = (float(v) for v in vals[i])
                c.jump = int(vals[i, 0] > 0.8)
                c.boost = int(vals[i, 1] > 0.3)
            b.SetAllCarControls(ctrl)
            b.Step(8)
        b.Synchronize()
        outs.append([bytes(b.GetCarState(a, i)) for a in range(0, n, 97) for i in range(2)])
        for a in range(0, n, 97):
            for i in range(2):
                s = b.GetCarState(a, i)
                v = R.car_vec(s)
                assert np.all(np.isfinite(v))
        b.Destroy()
        rng = np.random.default_rng(3)
    assert outs[0] == outs[1]


def test_reset_all_arenas_kickoff():
    b = E.RocketSimBatch(64, [1, 2, 3, 1] * 16, seed=9)
    b.Step(16)
    b.ResetAllArenas()
    b.Synchronize()
    kick = {(-2048, -2560), (2048, -2560), (-256, -3840), (256, -3840), (0, -4608)}
    for a in range(64):
        info = b.GetArenaInfo(a)
        assert info.tick_count == 0 and info.goal_scored == 0 and info.goal_team == -1 and info.episode == 1
        T = b.team_size[a]
        blue = [(round(b.GetCarState(a, i).pos.x), round(b.GetCarState(a, i).pos.y)) for i in range(T)]
        orange = [(round(b.GetCarState(a, T + i).pos.x), round(b.GetCarState(a, T + i).pos.y)) for i in range(T)]
        assert len(set(blue)) == T and set(blue) <= kick
        assert orange == [(-x, -y) for x, y in blue]
        ball = b.GetBallState(a)
        assert abs(ball.pos.z - 93.15) < 1e-3 and ball.pos.x == 0 and ball.pos.y == 0
    b.Destroy()
