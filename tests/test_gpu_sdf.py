"""The learned cull on the GPU: the exact mesh distance against a CPU
brute force, the network's matrix-core forward pass and the backward kernels
against numpy with the same BF16 rounding points, a short training run, and
the cull in the tick: bit-identical states with and without it, and zero
audit defects."""
import os

import numpy as np
import pytest

import rsxlib as R
import scenarios as SC
import sdf_ref as S
from test_gpu_parity import run_engine

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(torch_cuda):
    eng = R.import_engine()
    from emy_s_framework_amd.sdf import NeuralSdf
    b = eng.RocketSimBatch(64, 1, seed=1)
    s = NeuralSdf(b, seed=7)
    yield b, s
    s.destroy()
    b.Destroy()


def _scale(tris):
    v = tris.reshape(-1, 3).astype(np.float32)
    mn, mx = v.min(0), v.max(0)
    c = np.float32(0.5) * (mn + mx)
    h = np.float32(0.5) * (mx - mn)
    return (c[0], c[1], c[2], h[0], h[1], h[2])


def _random_params(seed, gain=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(S.NUM_PARAMS, np.float32)
    p[S.OFF_W1:S.OFF_B1] = rng.normal(0, gain, 384)
    p[S.OFF_B1:S.OFF_W2] = rng.normal(0, 0.3, 128)
    p[S.OFF_W2:S.OFF_B2] = rng.normal(0, gain / np.sqrt(128), 128 * 128)
    p[S.OFF_B2:S.OFF_W3] = rng.normal(0, 0.1, 128)
    p[S.OFF_W3:S.OFF_B3] = rng.normal(0, gain / np.sqrt(128), 128 * 128)
    p[S.OFF_B3:S.OFF_W4] = rng.normal(0, 0.1, 128)
    p[S.OFF_W4:S.OFF_B4] = rng.normal(0, 1 / np.sqrt(128), 128)
    p[S.OFF_B4] = 3.0
    return p


def test_exact_distance_matches_cpu(net, torch_cuda):
    from emy_s_framework_amd.sdf import sample_points
    _, s = net
    tris = s.mesh_triangles()
    pts = sample_points(tris, 300, seed=11)
    gpu = s.exact_distance(torch_cuda.from_numpy(pts).cuda()).cpu().numpy()
    ref = np.concatenate([S.exact_distance(pts[i:i + 25], tris) for i in range(0, len(pts), 25)])
    np.testing.assert_allclose(gpu, ref, rtol=1e-4, atol=2e-2)


def test_network_forward_matches_cpu(net, torch_cuda):
    """Pins the matrix-core operand layout: a transposed or mis-tiled operand
    changes every output by far more than the tolerance."""
    from emy_s_framework_amd.sdf import sample_points
    _, s = net
    tris = s.mesh_triangles()
    params = _random_params(3)
    s.set_params(params)
    pts = sample_points(tris, 1000, seed=12)
    gpu = s.eval(torch_cuda.from_numpy(pts).cuda()).cpu().numpy()
    y, _ = S.forward(params, S.scale_inputs(pts, _scale(tris)))
    ref = y * S.OUT_SCALE
    # FP32 accumulation order differs, and a bf16 activation can round the
    # other way: a few 1e-3 relative, not more
    np.testing.assert_allclose(gpu, ref, rtol=5e-3, atol=0.5)
    assert np.std(ref) > 10.0  # the check is not on a constant output


def test_backward_matches_cpu(net, torch_cuda):
    from emy_s_framework_amd.sdf import CLAMP_UU, sample_points
    _, s = net
    tris = s.mesh_triangles()
    params = _random_params(4)
    s.set_params(params)
    pts = sample_points(tris, 512, seed=13)
    tp = torch_cuda.from_numpy(pts).cuda()
    labels = s.exact_distance(tp)
    g, loss = s.gradients(tp, labels, CLAMP_UU, over_weight=3.0)
    rl, rg = S.loss_and_grads(params, S.scale_inputs(pts, _scale(tris)), labels.cpu().numpy(), CLAMP_UU,
                              over_weight=3.0)
    assert abs(loss - rl) <= 1e-4 * max(1.0, abs(rl))
    for name, lo, hi in (("w1", S.OFF_W1, S.OFF_B1), ("b1", S.OFF_B1, S.OFF_W2), ("w2", S.OFF_W2, S.OFF_B2),
                         ("b2", S.OFF_B2, S.OFF_W3), ("w3", S.OFF_W3, S.OFF_B3), ("b3", S.OFF_B3, S.OFF_W4),
                         ("w4", S.OFF_W4, S.OFF_B4), ("b4", S.OFF_B4, S.NUM_PARAMS)):
        a, b = g[lo:hi].astype(np.float64), rg[lo:hi].astype(np.float64)
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12)
        assert rel < 2e-2, (name, rel)


def test_backward_matches_finite_differences(net, torch_cuda):
    """Output-side parameters, central differences of the GPU loss (the L1
    loss is smooth away from its kinks).  Deeper weights move the loss by
    less than the BF16 rounding of the activations they feed, so their
    gradients are checked against the CPU backward instead, whose own
    derivation is checked by finite differences in float64 on the CPU."""
    from emy_s_framework_amd.sdf import CLAMP_UU, sample_points
    _, s = net
    tris = s.mesh_triangles()
    params = _random_params(5)
    pts = sample_points(tris, 256, seed=14)
    tp = torch_cuda.from_numpy(pts).cuda()
    labels = s.exact_distance(tp)
    s.set_params(params)
    g, _ = s.gradients(tp, labels, CLAMP_UU)
    for k in (S.OFF_B4, S.OFF_W4 + 5, S.OFF_W4 + 77, S.OFF_B3 + 17):
        h = 1e-2
        pp = params.copy()
        pp[k] += h
        s.set_params(pp)
        _, lp = s.gradients(tp, labels, CLAMP_UU)
        pp[k] -= 2 * h
        s.set_params(pp)
        _, lm = s.gradients(tp, labels, CLAMP_UU)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 0.05 * abs(g[k]) + 2e-4, (k, fd, g[k])
    s.set_params(params)


def test_short_training_reduces_error(net, torch_cuda):
    """A few hundred Adam steps lower the held-out error of the initial
    network and at least halve the safety margin it would need (the full run
    is train_sdf.py)."""
    from emy_s_framework_amd.sdf import sample_points, train, validate
    _, s = net
    s.set_params(_random_params(6, gain=0.5))
    tris = s.mesh_triangles()
    xv = torch_cuda.from_numpy(sample_points(tris, 65536, seed=21)).cuda()
    before = validate(s, xv, s.exact_distance(xv))
    rep = train(s, n_train=262144, n_val=65536, batch=4096, epochs=4, log=None)
    assert rep["mean_abs_err_uu"] < before["mean_abs_err_uu"], (before, rep)
    assert rep["margin_uu"] < 0.5 * before["margin_uu"], (before, rep)


def test_cull_in_tick_is_exact_and_audited(torch_cuda):
    eng = R.import_engine()
    from emy_s_framework_amd.sdf import WEIGHTS_PATH, NeuralSdf, margins_for
    if not os.path.exists(WEIGHTS_PATH):
        pytest.skip("no trained weights committed")
    cases = list(SC.cases().values())
    # a fresh batch per run: the wheels carry values from tick to tick that
    # setting a car's state leaves alone
    b = eng.RocketSimBatch(len(cases), 1, seed=3)
    exact = run_engine(b, cases)
    b.Destroy()
    b = eng.RocketSimBatch(len(cases), 1, seed=3)
    s = NeuralSdf(b, seed=0)
    meta = s.load(WEIGHTS_PATH)
    mb, mc = margins_for(meta, 2)
    s.attach(mb, every=2, audit=True, margin_car_uu=mc)
    culled = run_engine(b, cases)
    st = s.stats()
    s.detach()
    s.destroy()
    b.Destroy()
    assert st["defects"] == 0, st
    assert st["culled"] > 0, st
    for r0, r1 in zip(exact, culled):
        for t in range(len(r0.balls)):
            assert bytes(r0.balls[t]) == bytes(r1.balls[t])
            assert [bytes(c) for c in r0.cars[t]] == [bytes(c) for c in r1.cars[t]]


def test_cull_work_skipped_and_time_on_stress_states(torch_cuda):
    """What the trained network gains over the exact engine on the
    contact-heavy workload (65536 1v1 arenas, seed 4, the cull every 8 ticks
    with the committed margins): the same 16 fused env steps from the same
    states and actions with and without it give bit-identical arenas, an
    audited run of the culled engine finds no culled body within reach of the
    mesh, and the share of body-ticks culled and both step times are
    reported (measured, not gated: on this workload they are equal to within
    the box's run-to-run spread)."""
    torch = torch_cuda
    eng = R.import_engine()
    from emy_s_framework_amd.sdf import WEIGHTS_PATH, NeuralSdf, margins_for
    from emy_s_framework_amd.workloads import stress_states
    if not os.path.exists(WEIGHTS_PATH):
        pytest.skip("no trained weights committed")
    n, steps = 65536, 16
    gen = torch.Generator(device="cuda")
    gen.manual_seed(4)
    acts = [torch.randint(0, 90, (2 * n,), dtype=torch.int32, device="cuda", generator=gen) for _ in range(steps)]

    def run(audit=None):
        """A fresh batch (each arena's RNG stream and carried wheel values
        start alike), the workload, `steps` env steps; the cull attached
        unless audit is None."""
        b = eng.RocketSimBatch(n, 1, seed=4)
        net = None
        try:
            b.ConfigureEnv(no_touch_timeout_s=10.0, auto_reset=True)
            b.SetTuning(128, 64, 1)
            if audit is not None:
                net = NeuralSdf(b, seed=0)
                mb, mc = margins_for(net.load(WEIGHTS_PATH), 8)
                net.attach(mb, every=8, audit=audit, margin_car_uu=mc)
                net.reset_stats()
            stress_states(b, 4)
            b.Synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for a in acts:
                b.StepEnv(a.data_ptr())
            e1.record()
            torch.cuda.synchronize()
            st = net.stats() if net is not None else None
            return bytes(b.GetAllCarStates()) + bytes(b.GetAllBallStates()), e0.elapsed_time(e1) / steps, st
        finally:
            if net is not None:
                net.detach()
                net.destroy()
            b.Destroy()

    exact, t_exact, _ = run()
    culled, t_cull, st = run(audit=False)
    audited, _, st_audit = run(audit=True)
    assert culled == exact and audited == exact
    assert st_audit["defects"] == 0, st_audit
    assert st["culled"] > 0, st
    frac = st["culled"] / (steps * 8 * (n + 2 * n))
    print(f"65536 contact-heavy 1v1 arenas x {steps} env steps: exact {t_exact:.2f} ms/step, culled "
          f"{t_cull:.2f} ms/step; {100 * frac:.1f}% of body-ticks culled, {st['queries']} network queries; "
          f"audit {st_audit}", flush=True)
