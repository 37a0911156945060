"""VantaObs on the GPU (RSX_OBS_VANTA, BuildVantaObs of the reference's GPU
path): the rows rsx_step_env and rsx_build_obs write, against the
restatement (oracle/vanta_spec.py) bit for bit over env steps with
auto-resets, for 1v1, 2v2, 3v3 and a mixed batch; the ball-prediction slices
against the oracle's ball-only arena (BallPredTracker's); a reset clearing an
arena's state; graph replays against eager steps."""
import ctypes as C

import numpy as np
import pytest

import parity_harness as P
import rsxlib as R

pytestmark = pytest.mark.gpu
E = R.engine
V = R.load_vanta_spec()
F = np.float32


def _engine(teams, seed, predict=True, only_closest=False, timeout=1.5):
    from emy_s_framework_amd import workloads as WL
    b = E.RocketSimBatch(len(teams), teams, seed=seed)
    WL.random_field_states(b, seed)
    b.ConfigureEnv(no_touch_timeout_s=timeout, obs_builder=E.OBS_VANTA)
    b.ConfigureVanta(predict_ball=predict, only_closest_opp=only_closest)
    return b


def _public(b):
    """Per arena: (cars, ball, pads, tick) as the engine holds them now."""
    cars, balls, pads, infos = b.GetAllCarStates(), b.GetAllBallStates(), b.GetAllBoostPadStates(), \
        b.GetAllArenaInfos()
    out, c0 = [], 0
    for a in range(b.num_arenas):
        n = 2 * b.team_size[a]
        pa = R.PADS()
        for i in range(34):
            pa[i] = pads[34 * a + i]
        out.append(([cars[c0 + i] for i in range(n)], balls[a], pa, infos[a].tick_count))
        c0 += n
    return out


def _spec_rows(b, specs, dt, only_closest=False, predict=True):
    """The restatement's rows of every arena (updating its per-arena state)
    from the engine's records and prediction slices."""
    rows = []
    for a, (cars, ball, pads, tick) in enumerate(_public(b)):
        pred = b.GetVantaPrediction(a) if predict else None
        preds = None
        if pred is not None:
            preds = (np.array([[s.pos.x, s.pos.y, s.pos.z] for s in pred], F),
                     np.array([[s.vel.x, s.vel.y, s.vel.z] for s in pred], F))
        rows.append(V.build_all(specs[a], cars, ball, pads, tick, dt, preds, 0.5, only_closest))
    return np.concatenate(rows)


@pytest.mark.parametrize("teams,only_closest", [([1] * 24, False), ([2] * 16, False), ([3] * 12, True),
                                                ([1, 3, 2, 2, 1, 3, 1, 2] * 3, False)])
def test_rows_match_the_restatement_over_env_steps(teams, only_closest, torch_cuda):
    """Random-state arenas under random actions for 40 env steps (no-touch
    limit 1.5 s, so arenas reset along the way): after the first build and
    after every step, every row equals the restatement's bit for bit, the
    restatement keeping its own pad / coyote timers and starting over for an
    arena the step reset."""
    import torch
    b = _engine(teams, seed=21, only_closest=only_closest)
    dt = F(8) / F(120)
    specs = [V.ArenaState(2 * t) for t in teams]
    _, ln = b.ObsLayout()
    assert (ln == 366).all() and b.EnvBuffers()[3] == 366 * b.num_cars
    b.BuildObs()
    b.Synchronize()
    want = _spec_rows(b, specs, dt, only_closest)
    np.testing.assert_array_equal(b.CopyBuiltObs().reshape(-1, 366), want, err_msg="first build")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(5)
    resets = 0
    for step in range(40):
        d = torch.randint(0, 90, (b.num_cars,), dtype=torch.int32, device="cuda", generator=gen)
        b.StepEnv(d.data_ptr())
        b.Synchronize()
        term = np.array(b.CopyBuiltTerminals())
        for a in np.nonzero(term)[0]:
            specs[a] = V.ArenaState(2 * teams[a])  # ResetArena clears the arena's Vanta state
        resets += int((term != 0).sum())
        want = _spec_rows(b, specs, dt, only_closest)
        got = b.CopyBuiltObs().reshape(-1, 366)
        if not np.array_equal(got, want):
            bad = np.argwhere(got != want)
            raise AssertionError(f"step {step}: {len(bad)} floats differ, first (row, col) {bad[:5].tolist()}, "
                                 f"got {got[tuple(bad[0])]}, want {want[tuple(bad[0])]}")
    assert resets > 0
    print(f"teams {sorted(set(teams))}: 40 env steps, {resets} arena resets, rows bit-identical")
    b.Destroy()


def test_prediction_slices_match_the_oracle(oracle, torch_cuda):
    """The engine's ball-only rollouts on 128 random-state balls against the
    oracle's ball-only arena stepped 120, 240 and 360 ticks: every slice
    within 1e-4 of the obs scales (0.23 uu, 0.6 uu/s)."""
    b = _engine([1] * 128, seed=31)
    b.BuildObs()
    b.Synchronize()
    worst = 0.0
    for a in range(b.num_arenas):
        got = b.GetVantaPrediction(a)
        assert got is not None
        ball = b.GetBallState(a)
        h = C.c_void_p(oracle.ref_arena_new(0, 0))
        oracle.ref_set_ball(h, C.byref(ball))
        for k in range(3):
            oracle.ref_step(h, 120)
            s = R.BallState()
            oracle.ref_get_ball(h, C.byref(s))
            dp = np.abs(R.ball_vec(got[k])[:3] - R.ball_vec(s)[:3]).max()
            dv = np.abs(R.ball_vec(got[k])[3:6] - R.ball_vec(s)[3:6]).max()
            worst = max(worst, dp / 2300, dv / 6000)
            assert dp <= 1e-4 * 2300 and dv <= 1e-4 * 6000, (a, k, dp, dv)
        oracle.ref_arena_free(h)
    print(f"prediction slices of 128 balls: worst |diff| / obs scale {worst:.2e}")
    b.Destroy()


def test_reset_clears_the_state_and_prediction_can_be_off(torch_cuda):
    """rsx_reset_arena starts the arena's Vanta state over (pad timers and
    coyote timers as after a fresh build); with prediction off the 18
    prediction floats are zero and no slices are stored."""
    import torch
    b = _engine([1] * 8, seed=41)
    dt = F(8) / F(120)
    d = torch.randint(0, 90, (b.num_cars,), dtype=torch.int32, device="cuda")
    for _ in range(6):
        b.StepEnv(d.data_ptr())
    b.ResetArena(3)
    b.BuildObs()
    b.Synchronize()
    cars, ball, pads, tick = _public(b)[3]
    pred = b.GetVantaPrediction(3)
    preds = (np.array([[s.pos.x, s.pos.y, s.pos.z] for s in pred], F),
             np.array([[s.vel.x, s.vel.y, s.vel.z] for s in pred], F))
    want = V.build_all(V.ArenaState(2), cars, ball, pads, tick, dt, preds)
    off, _ = b.ObsLayout()
    got = b.CopyBuiltObs()[off[6]:off[6] + 2 * 366].reshape(2, 366)
    np.testing.assert_array_equal(got, want)
    b.ConfigureVanta(predict_ball=False)
    b.BuildObs()
    b.Synchronize()
    assert b.GetVantaPrediction(0) is None
    assert (b.CopyBuiltObs().reshape(-1, 366)[:, 92:110] == 0).all()
    b.Destroy()


def test_graph_replays_equal_eager_steps(torch_cuda):
    """A captured env step with VantaObs (fused step, then the Vanta kernel)
    replayed three times, with eager work and syncs between the replays,
    gives the rows, rewards and terminals of the same steps run eagerly."""
    import torch
    teams = [1, 2, 3, 1] * 4
    d = torch.randint(0, 90, (sum(2 * t for t in teams),), dtype=torch.int32, device="cuda",
                      generator=torch.Generator(device="cuda").manual_seed(3))
    outs = []
    for graph in (False, True):
        b = _engine(teams, seed=51)
        if graph:
            b.CaptureEnvGraph(d.data_ptr(), 1)
        rec = []
        for _ in range(3):
            if graph:
                b.LaunchEnvGraph()
            else:
                b.StepEnv(d.data_ptr())
            b.Synchronize()
            rec.append((b.CopyBuiltObs().tobytes(), b.CopyBuiltRewards().tobytes(), bytes(b.CopyBuiltTerminals())))
            torch.cuda.synchronize()
            _ = (d.float() * 2).sum().item()  # eager work between replays
        outs.append(rec)
        b.Destroy()
    assert outs[0] == outs[1]


def test_continued_rollouts_equal_new_ones(torch_cuda):
    """Two engines with the same arenas and actions: one builds VantaObs with
    prediction after every env step (so arenas whose ball moved as the last
    rollout predicted continue it), the other turns prediction on only at the
    end (every arena a new 360-tick rollout).  Slices and rows agree bit for
    bit."""
    import torch
    teams = [1] * 96 + [2] * 16
    gen = torch.Generator(device="cuda").manual_seed(9)
    acts = [torch.randint(0, 90, (sum(2 * t for t in teams),), dtype=torch.int32, device="cuda", generator=gen)
            for _ in range(12)]
    a = _engine(teams, seed=61, predict=True, timeout=100.0)
    b = _engine(teams, seed=61, predict=False, timeout=100.0)
    for d in acts:
        a.StepEnv(d.data_ptr())
        b.StepEnv(d.data_ptr())
    b.ConfigureVanta(predict_ball=True)
    b.BuildObs()
    a.BuildObs()  # nothing moved since its last build: slices kept
    a.Synchronize()
    b.Synchronize()
    for k in range(len(teams)):
        pa, pb = a.GetVantaPrediction(k), b.GetVantaPrediction(k)
        assert bytes(b"".join(bytes(x) for x in pa)) == bytes(b"".join(bytes(x) for x in pb)), k
    oa, ob = a.CopyBuiltObs().reshape(-1, 366), b.CopyBuiltObs().reshape(-1, 366)
    np.testing.assert_array_equal(oa[:, 81:110], ob[:, 81:110])
    a.Destroy()
    b.Destroy()
