"""Launch shapes, the single launch over mixed team sizes, the autotuner and
the hipGraph of the env step.  None of them may change a result: every
comparison here is bit-exact against the default launch (64-lane
workgroups, one arena per lane, one bucket at a time, eager launches)."""
import os

import numpy as np
import pytest

import rsxlib as R

pytestmark = pytest.mark.gpu
E = R.engine


def _workloads():
    from emy_s_framework_amd import workloads
    return workloads


def _snapshot(b):
    b.Synchronize()
    return (bytes(b.GetAllCarStates()), bytes(b.GetAllBallStates()), bytes(b.GetAllBoostPadStates()),
            bytes(b.GetAllArenaInfos()), b.CopyBuiltObs().tobytes(), b.CopyBuiltRewards().tobytes(),
            b.CopyBuiltTerminals().tobytes())


def _actions(torch, n, seed, steps):
    g = np.random.default_rng(seed)
    return [torch.tensor(g.integers(0, 90, n).astype(np.int32), device="cuda") for _ in range(steps)]


def _stress_batch(n, seed=4):
    b = E.RocketSimBatch(n, 1, seed=seed)
    _workloads().stress_states(b, seed)
    b.ConfigureEnv(auto_reset=True)
    return b


def _mixed_teams(n, seed=2):
    return _workloads().mixed_team_sizes(n, seed)


def _run(b, acts, shape=None):
    if shape:
        b.SetTuning(*shape)
    out = []
    for d in acts:
        b.StepEnv(d.data_ptr())
        out.append(_snapshot(b))
    return out


def test_launch_shapes_are_bit_identical(torch_cuda):
    torch = torch_cuda
    n = 300  # partial last wavefront for every shape
    b = _stress_batch(n)
    acts = _actions(torch, b.num_cars, 11, 3)
    want = _run(b, acts)
    b.Destroy()
    for shape in ((64, 64, 0), (128, 64, 1), (128, 32, 0), (256, 16, 1), (256, 8, 1), (64, 8, 0), (512, 64, 0),
                  (512, 16, 0), (64, 4, 1), (128, 2, 1), (64, 1, 1)):
        b = _stress_batch(n)
        got = _run(b, acts, shape)
        assert b.GetTuning() == shape
        b.Destroy()
        for s, (w, g) in enumerate(zip(want, got)):
            assert w == g, f"shape {shape} differs at env step {s}"


def test_invalid_shapes_are_rejected(torch_cuda):
    b = E.RocketSimBatch(4, 1)
    for shape in ((1024, 64, 0), (96, 64, 1), (64, 48, 1), (64, 0, 1), (64, 64, 2), (512, 64, 1)):
        with pytest.raises(E.RsxError):
            b.SetTuning(*shape)
    assert b.GetTuning() == (64, 64, 1)
    b.Destroy()


def _copy_arena(src, a, dst, k):
    """Sets the same public records on both engines (the records' unit
    conversion is not an exact round trip, so each engine loads them itself)."""
    for i in range(2 * src.team_size[a]):
        rec = src.GetCarState(a, i)
        src.SetCarState(a, i, rec)
        dst.SetCarState(k, i, rec)
    ball = src.GetBallState(a)
    src.SetBallState(a, ball)
    dst.SetBallState(k, ball)
    dst.SetBoostPadStates(k, src.GetBoostPadStates(a))


def test_mixed_single_launch_matches_one_bucket_engines(torch_cuda):
    torch = torch_cuda
    n = 240
    teams = _mixed_teams(n)
    m = E.RocketSimBatch(n, teams, seed=2)
    _workloads().random_field_states(m, 2)
    m.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=False)
    singles = {}
    for t in (1, 2, 3):
        idx = [a for a in range(n) if teams[a] == t]
        s = E.RocketSimBatch(len(idx), t, seed=2)
        s.ConfigureEnv(no_touch_timeout_s=0.0, auto_reset=False)
        for k, a in enumerate(idx):
            _copy_arena(m, a, s, k)
        singles[t] = (s, idx)
    off_m, len_m = m.ObsLayout()
    car0 = np.concatenate([[0], np.cumsum([2 * t for t in teams])])
    g = np.random.default_rng(5)
    keep = []  # device action buffers stay alive until the launches that read them are done
    for step in range(3):
        act = g.integers(0, 90, m.num_cars).astype(np.int32)
        keep.append(torch.tensor(act, device="cuda"))
        m.StepEnv(keep[-1].data_ptr())
        for t, (s, idx) in singles.items():
            sa = np.concatenate([act[car0[a]:car0[a + 1]] for a in idx]).astype(np.int32)
            keep.append(torch.tensor(sa, device="cuda"))
            s.StepEnv(keep[-1].data_ptr())
        m.Synchronize()
        obs_m, rew_m, term_m = m.CopyBuiltObs(), m.CopyBuiltRewards(), m.CopyBuiltTerminals()
        for t, (s, idx) in singles.items():
            s.Synchronize()
            obs_s, rew_s, term_s = s.CopyBuiltObs(), s.CopyBuiltRewards(), s.CopyBuiltTerminals()
            off_s, len_s = s.ObsLayout()
            for k, a in enumerate(idx):
                for i in range(2 * t):
                    assert bytes(m.GetCarState(a, i)) == bytes(s.GetCarState(k, i)), (step, a, i)
                    cm, cs = car0[a] + i, 2 * t * k + i
                    assert len_m[cm] == len_s[cs]
                    assert obs_m[off_m[cm]:off_m[cm] + len_m[cm]].tobytes() == \
                        obs_s[off_s[cs]:off_s[cs] + len_s[cs]].tobytes(), (step, a, i)
                    assert rew_m[cm] == rew_s[cs]
                assert bytes(m.GetBallState(a)) == bytes(s.GetBallState(k)), (step, a)
                assert bytes(m.GetBoostPadStates(a)) == bytes(s.GetBoostPadStates(k)), (step, a)
                assert term_m[a] == term_s[k]
    for s, _ in singles.values():
        s.Destroy()
    m.Destroy()


def test_mixed_launch_shapes_are_bit_identical(torch_cuda):
    torch = torch_cuda
    n = 200
    teams = _mixed_teams(n, seed=7)

    def make():
        b = E.RocketSimBatch(n, teams, seed=7)
        _workloads().random_field_states(b, 7)
        return b

    b = make()
    acts = _actions(torch, b.num_cars, 3, 2)
    want = _run(b, acts)
    b.Destroy()
    # with 2 and 1 arenas per wavefront most lanes only take part in the pools
    for shape in ((256, 16, 1), (128, 8, 0), (64, 64, 0), (64, 2, 1), (128, 1, 1)):
        b = make()
        got = _run(b, acts, shape)
        b.Destroy()
        assert want == got, shape


def test_cull_with_multiwave_workgroups(torch_cuda):
    """The cull's network runs per wavefront; with several wavefronts per
    workgroup each uses its own LDS and waits for no other."""
    torch = torch_cuda
    from emy_s_framework_amd.sdf import WEIGHTS_PATH, NeuralSdf, margins_for
    if not os.path.exists(WEIGHTS_PATH):
        pytest.skip("no trained weights committed")
    n = 300
    b = _stress_batch(n)
    acts = _actions(torch, b.num_cars, 17, 3)
    want = _run(b, acts)
    b.Destroy()
    for shape in ((64, 64, 1), (256, 16, 1), (128, 64, 0)):
        b = _stress_batch(n)
        s = NeuralSdf(b, seed=0)
        meta = s.load(WEIGHTS_PATH)
        mb, mc = margins_for(meta, 2)
        s.attach(mb, every=2, audit=True, margin_car_uu=mc)
        got = _run(b, acts, shape)
        st = s.stats()
        s.detach()
        s.destroy()
        b.Destroy()
        assert st["defects"] == 0, (shape, st)
        assert st["culled"] > 0, (shape, st)
        assert want == got, shape


def test_env_graph_replays_match_eager(torch_cuda):
    torch = torch_cuda
    n = 200
    eager = _stress_batch(n)
    graph = _stress_batch(n)
    d_graph = torch.zeros(graph.num_cars, dtype=torch.int32, device="cuda")
    graph.CaptureEnvGraph(d_graph.data_ptr(), 1)
    stream = torch.cuda.current_stream().cuda_stream
    acts = _actions(torch, eager.num_cars, 23, 4)
    for k, d in enumerate(acts):
        d_graph.copy_(d)
        graph.LaunchEnvGraph(stream)
        eager.StepEnv(d.data_ptr())
        # eager work of the program's own between replays
        torch.cuda.synchronize()
        _ = (d.float() * 2).sum().item()
        assert _snapshot(graph) == _snapshot(eager), f"replay {k}"
    # a configuration change invalidates the capture until the next one
    graph.SetTuning(128, 32)
    with pytest.raises(E.RsxError):
        graph.LaunchEnvGraph(stream)
    graph.CaptureEnvGraph(d_graph.data_ptr(), 2)  # two env steps per replay
    for k in range(2):
        d_graph.copy_(acts[k])
        graph.LaunchEnvGraph(stream)
        eager.StepEnv(acts[k].data_ptr())
        eager.StepEnv(acts[k].data_ptr())
        assert _snapshot(graph) == _snapshot(eager), f"two-step replay {k}"
    graph.Destroy()
    eager.Destroy()


def test_autotune_restores_state_and_keeps_the_fastest(torch_cuda):
    torch = torch_cuda
    n = 300
    teams = _mixed_teams(n, seed=9)

    def make():
        b = E.RocketSimBatch(n, teams, seed=9)
        _workloads().random_field_states(b, 9)
        return b

    b = make()
    b.SetActionMasks(True)
    d = torch.tensor(np.random.default_rng(1).integers(0, 90, b.num_cars).astype(np.int32), device="cuda")
    b.StepEnv(d.data_ptr())
    before = _snapshot(b)
    masks_before = bytes(b.CopyBuiltDefaultActionMasks())
    res = b.Autotune(d.data_ptr(), env_steps=1)
    assert len(res) == 44
    assert any(r[0] == 512 and r[3] is not None for r in res)  # unpooled 512-lane workgroups fit
    feas = [r for r in res if r[3] is not None]
    assert feas and all(r[3] > 0 for r in feas)
    best = min(feas, key=lambda r: r[3])
    assert b.GetTuning() == best[:3]
    assert _snapshot(b) == before
    assert bytes(b.CopyBuiltDefaultActionMasks()) == masks_before  # the trials' masks are rolled back too
    b.StepEnv(d.data_ptr())
    got = _snapshot(b)
    b.Destroy()
    ref = make()
    ref.StepEnv(d.data_ptr())
    ref.StepEnv(d.data_ptr())
    want = _snapshot(ref)
    ref.Destroy()
    assert got == want
