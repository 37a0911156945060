"""The gfx950 engine through its C-ABI against the oracle on samples of the
three benchmark workloads, and full-size property runs of each.

Samples (workload_harness): 256 arenas of the random-state 1v1 workload
(4096 arenas, seed 1), 128 of the mixed 1v1/2v2/3v3 workload (16384 arenas,
seed 2, one batch holding all three team sizes) and 256 of the contact-heavy
1v1 workload (65536 arenas, seed 4).  Each arena runs under its own uniform
default-table actions: a warm tick, the workload state loaded again, then 8
launches of one tick (the default pooled launch shape), every arena read back
after each.

Gates, as for the CPU build (test_cpu_workloads.py):
  - ticks 0-7 teacher-forced: every arena-tick within the parity bar
    (positions, velocities and angular velocities per vector within
    1e-3 + 1e-4 |ref|, rotation entries 1e-4, flags, event timers, hit ticks,
    pads and the goal flag bit for bit), except where the oracle itself
    leaves the bar when its input moves by one float32 ulp (exempt with the
    arena's later ticks, reported; at most 1% of arenas at tick 0, 3% in all);
  - the learned cull attached (audited) and the unpooled launch shape give
    bit-identical records to the default run; the fused env step (8 ticks,
    DefaultObs, rewards, terminals in one launch) gives bit-identical state to
    the 8 single-tick launches, and its outputs match the spec restatement on
    those records;
  - no arena runs out of contact-manifold storage.

Full-size runs: 4096 random, 16384 mixed and 65536 contact-heavy arenas, 16
fused env steps with the cull attached and audited every tick: states finite,
a replay from the same states bit-identical, zero audit defects, no storage
overflow."""
import ctypes as C
import numpy as np
import pytest

import parity_harness as P
import rsxlib as R
import workload_harness as WH

pytestmark = pytest.mark.gpu
E = R.engine
S = R.load_spec()
SAMPLES = [("random", 256), ("mixed", 128), ("stress", 256)]
TICKS = 8
W = dict(goal=1, concede=-1, touch=1, vbtg=1, vptb=1)


def _records(res):
    return [b"".join(bytes(c) for c in r.cars[t]) + bytes(r.balls[t]) + bytes(r.pads[t])
            for r in res for t in range(len(r.balls))]


def _attach_cull(batch, audit=True, every=8):
    from emy_s_framework_amd.sdf import WEIGHTS_PATH, NeuralSdf, margins_for
    net = NeuralSdf(batch, seed=0)
    mb, mc = margins_for(net.load(WEIGHTS_PATH), every)
    net.attach(mb, every=every, audit=audit, margin_car_uu=mc)
    net.reset_stats()
    return net


@pytest.fixture(scope="module")
def samples(oracle, torch_cuda):
    """Each sample run once on a fresh batch.  Every comparison run below
    starts from a fresh batch too: a batch's per-wheel carried values and
    broadphase cells survive a state load, so a second run on the same batch
    starts its warm tick from different hidden state."""
    lut = WH.lut_controls()
    out = {}
    for kind, count in SAMPLES:
        cases = WH.sample(kind, count)
        batch = WH.engine_batch(cases)
        eng = WH.run_engine(batch, cases, TICKS)
        infos = batch.GetAllArenaInfos()
        batch.Destroy()
        assert not any(i.contact_overflow for i in infos), kind
        out[kind] = (cases, eng)
    return out


@pytest.mark.parametrize("kind,count", SAMPLES)
def test_workload_sample_matches_oracle(kind, count, samples, oracle):
    cases, eng = samples[kind]
    lut = WH.lut_controls()
    failures, exempt, worst = WH.teacher_forced_check(oracle, cases, eng, TICKS, lut)
    print(f"{kind}: {count} arenas x {TICKS} ticks on the GPU, worst passing error {worst:.3f} of the bar, "
          f"exempt {len(exempt)} arena-ticks {exempt[:4]}", flush=True)
    assert not failures, failures[:5]
    assert len([x for x in exempt if x[1] == 0]) <= 0.01 * count, exempt[:5]
    assert len({x[0] for x in exempt}) <= 0.03 * count, exempt


@pytest.mark.parametrize("kind", [k for k, _ in SAMPLES])
def test_cull_and_launch_shape_leave_results_unchanged(kind, samples):
    """The same sample with the learned cull attached (audited every tick) and
    with unpooled box-triangle queries: records bit-identical to the default
    exact run, zero audit defects."""
    cases, eng = samples[kind]
    ref = _records(eng)
    batch = WH.engine_batch(cases)
    net = _attach_cull(batch)
    try:
        got = WH.run_engine(batch, cases, TICKS)
        st = net.stats()
    finally:
        net.detach()
        net.destroy()
        batch.Destroy()
    assert st["defects"] == 0, st
    assert _records(got) == ref
    batch = WH.engine_batch(cases)
    batch.SetTuning(64, 64, 0)
    try:
        got = WH.run_engine(batch, cases, TICKS)
    finally:
        batch.Destroy()
    assert _records(got) == ref
    print(f"{kind}: cull culled {st['culled']} body-ticks of {st['queries']} network queries, 0 defects", flush=True)


@pytest.mark.parametrize("kind", [k for k, _ in SAMPLES])
def test_fused_env_step_equals_single_ticks(kind, samples):
    """rsx_step_env over the sample (same warm protocol and actions, auto-reset
    off) ends bit-identical to the 8 single-tick launches, and its DefaultObs
    rows, rewards and terminals equal the spec restatement on the post-step
    records (touch and terminal bit for bit)."""
    import torch
    cases, eng = samples[kind]
    lut = WH.lut_controls()
    batch = WH.engine_batch(cases)
    idx = np.concatenate([c["actions"] for c in cases]).astype(np.int32)
    batch.ConfigureEnv(no_touch_timeout_s=1.0, auto_reset=False)
    WH.load_engine(batch, cases)
    batch.SetDiscreteActionsFromIndices(idx)
    batch.Step(1)
    WH.load_engine(batch, cases)
    no_touch = [i.no_touch_time for i in batch.GetAllArenaInfos()]
    d = torch.tensor(idx, device="cuda")
    batch.StepEnv(d.data_ptr())
    batch.Synchronize()
    cars = batch.GetAllCarStates()
    balls = batch.GetAllBallStates()
    pads = batch.GetAllBoostPadStates()
    obs = batch.CopyBuiltObs()
    rew = batch.CopyBuiltRewards()
    term = batch.CopyBuiltTerminals()
    off, ln = batch.ObsLayout()
    sz = C.sizeof(E.BoostPadState)
    raw = bytes(pads)
    o = 0
    seen = set()
    for a, case in enumerate(cases):
        k = len(case["cars"])
        rc = [E.CarState.from_buffer_copy(cars[o + i]) for i in range(k)]
        rb = E.BallState.from_buffer_copy(balls[a])
        rp = R.PADS.from_buffer_copy(raw[a * 34 * sz:(a + 1) * 34 * sz])
        r = eng[a]
        assert [bytes(x) for x in rc] == [bytes(x) for x in r.cars[-1]], (kind, case["source"])
        assert bytes(rb) == bytes(r.balls[-1]) and bytes(rp) == bytes(r.pads[-1]), (kind, case["source"])
        tick = case["tick"] + TICKS
        prev = np.array([[lut[j].throttle, lut[j].steer, lut[j].pitch, lut[j].yaw, lut[j].roll, lut[j].jump,
                          lut[j].boost, lut[j].handbrake] for j in case["actions"]], dtype=np.float32)
        touched = [S.touched_in_step(c, tick, TICKS) for c in rc]
        goal = S.ball_scored(rb)
        for i in range(k):
            row = np.asarray(obs[off[o + i]:off[o + i] + ln[o + i]])
            np.testing.assert_allclose(row, S.default_obs(rc, rb, rp, prev, i), atol=2e-6)
        np.testing.assert_allclose(rew[o:o + k], S.stub_rewards(rc, rb, goal, touched, W), atol=4e-6)
        t, _ = S.step_terminal(goal, any(touched), TICKS, no_touch[a], 1.0)
        assert term[a] == t, (kind, case["source"])
        seen.add((int(goal), any(touched)))
        o += k
    batch.Destroy()
    print(f"{kind}: fused env step bit-identical to 8 single ticks; (goal, touch) outcomes seen {sorted(seen)}",
          flush=True)


def _full_size(kind):
    from emy_s_framework_amd import workloads as WL
    if kind == "random":
        b = E.RocketSimBatch(4096, 1, seed=1)
        WL.random_field_states(b, 1)
    elif kind == "mixed":
        b = E.RocketSimBatch(16384, WL.mixed_team_sizes(16384, 2), seed=2)
        WL.random_field_states(b, 2)
    else:
        b = E.RocketSimBatch(65536, 1, seed=4)
        WL.stress_states(b, 4)
    return b


def _full_size_run(kind, steps):
    """Fresh batch of the full workload, `steps` fused env steps (auto-reset
    on, no-touch limit 10 s) with the cull attached and audited every tick;
    returns the final records, every step's terminals and obs, the cull's
    counters and whether any arena ran out of contact storage."""
    import torch
    b = _full_size(kind)
    b.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True)
    net = _attach_cull(b)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    terms, obs = [], []
    try:
        for step in range(steps):
            d = torch.randint(0, 90, (b.num_cars,), dtype=torch.int32, device="cuda", generator=gen)
            b.StepEnv(d.data_ptr())
            terms.append(bytes(b.CopyBuiltTerminals()))
            obs.append(b.CopyBuiltObs().tobytes())
        b.Synchronize()
        st = net.stats()
    finally:
        net.detach()
        net.destroy()
    cars = b.GetAllCarStates()
    v = np.ctypeslib.as_array(cars)
    finite = all(np.all(np.isfinite(v[f][ax])) for f in ("pos", "vel", "ang_vel") for ax in "xyz")
    over = any(i.contact_overflow for i in b.GetAllArenaInfos())
    out = (bytes(cars), bytes(b.GetAllBallStates()), bytes(b.GetAllBoostPadStates()), b"".join(terms), b"".join(obs))
    n = b.num_arenas
    b.Destroy()
    return out, st, finite, over, n


@pytest.mark.parametrize("kind", ["random", "mixed", "stress"])
def test_full_size_runs(kind, torch_cuda):
    """The full workload, 16 fused env steps with the cull attached and
    audited every tick, twice from fresh batches: finite, bit-identical
    replays (states, terminals, obs), zero audit defects, no storage
    overflow."""
    r1, st, finite, over, n = _full_size_run(kind, 16)
    r2, _, _, _, _ = _full_size_run(kind, 16)
    assert finite and not over, (kind, finite, over)
    assert r1 == r2, kind
    assert st["defects"] == 0, st
    done = sum(x != 0 for x in r1[3])
    print(f"{kind}: {n} arenas x 16 env steps, replay bit-identical, {done} terminal arena-steps, cull {st}",
          flush=True)


def test_extra_cases_match_oracle(oracle, torch_cuda):
    """Wheel rays on another car and on the ball, the auto-flip and a
    demolition through its respawn (extra_cases.py) in one GPU batch, one
    tick per launch, teacher-forced against the oracle with the same rule."""
    import extra_cases as X
    cases = X.cases()
    lut = WH.lut_controls()
    batch = WH.engine_batch(cases)
    try:
        res = WH.run_engine(batch, cases, max(c["ticks"] for c in cases))
    finally:
        batch.Destroy()
    failures, exempt, worst, respawned = X.check(oracle, cases, res, lut)
    print(f"extra cases on the GPU: worst passing {worst:.3f}, "
          f"exempt {[(x[0], x[1]) for x in exempt if x[3] != 'after an exempt tick']}, respawns {respawned}",
          flush=True)
    assert not failures, failures[:5]
    assert [r[0] for r in respawned] == ["demolition respawn"]
    assert any(c[0].is_auto_flipping for c in res[3].cars)
