"""The device GAE (rsx_gae) against the CPU restatement of the trainer's loop
(oracle/gae_spec.py, GAE.cpp:128-195): advantages, target values and returns
bit for bit, the reward clip portion to float rounding (its sum order differs).
Buffers: sizes around the 1024-step block, terminal mixes from none (one
segment over every block) to all, with and without reward standardisation and
clipping; a 16M-step buffer checked on windows of whole episodes."""
import importlib.util
import os

import numpy as np
import pytest

import rsxlib as R

pytestmark = pytest.mark.gpu
ROOT = R.ROOT
_spec = importlib.util.spec_from_file_location("gae_spec", os.path.join(ROOT, "oracle", "gae_spec.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)
F = np.float32


def _buffer(T, p_normal, p_trunc, seed, last=None):
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=T)
    term = np.where(u < p_normal, 1, np.where(u < p_normal + p_trunc, 2, 0)).astype(np.int8)
    if last is not None and T:
        term[-1] = last
    r = (rng.normal(size=T) * 2).astype(F)
    v = rng.normal(size=T).astype(F)
    tv = rng.normal(size=int((term == 2).sum())).astype(F)
    return r, term, v, tv


def _device(r, term, v, tv, **kw):
    import torch
    from emy_s_framework_amd.gae import ComputeGAEOnDevice
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    out = ComputeGAEOnDevice(t(r), t(term), t(v), t(tv) if len(tv) else None, **kw)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


# sizes around a lane's 16 steps and a workgroup's 8192-step region; no
# terminal at all (one segment over every region), all terminals, a last
# step with and without one
CASES = [(1, 0.0, 0.0, None), (5, 0.2, 0.2, 0), (16, 0.1, 0.1, 0), (17, 0.0, 0.0, 2), (1023, 0.01, 0.01, 1),
         (1024, 0.003, 0.003, 2), (1025, 0.0, 0.0, 0), (3000, 0.0, 0.0, 2), (4096, 0.5, 0.5, None),
         (8191, 0.001, 0.001, 0), (8192, 0.0, 0.0, 1), (8193, 0.002, 0.0, 0), (20000, 0.004, 0.002, 0),
         (3 * 8192 + 5, 0.0001, 0.0, 2), (100003, 0.001, 0.0005, 2)]


@pytest.mark.parametrize("T,pn,pt,last", CASES)
@pytest.mark.parametrize("std,clip", [(0.0, 0.0), (1.7, 0.0), (2.5, 1.0)])
def test_gae_matches_cpu_loop(T, pn, pt, last, std, clip, torch_cuda):
    r, term, v, tv = _buffer(T, pn, pt, seed=T + int(10 * std), last=last)
    want = G.gae(r, term, v, tv, 0.99, 0.95, return_std=std, clip_range=clip)
    got = _device(r, term, v, tv, gamma=0.99, lam=0.95, return_std=std, clip_range=clip)
    for k, name in enumerate(("advantages", "target values", "returns")):
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), (name, T, std, clip)
    assert abs(float(got[3][0]) - float(want[3])) <= 1e-5 * max(1.0, abs(float(want[3]))) + 1e-6


def test_large_buffer_on_episode_windows(torch_cuda):
    """16M steps, episodes of a few thousand steps: windows of whole episodes
    (their bootstrap values sliced by the truncations before them) equal the
    CPU loop run on the window alone."""
    import torch
    T = 1 << 24
    r, term, v, tv = _buffer(T, 0.0002, 0.0003, seed=5, last=2)
    got = _device(r, term, v, tv, gamma=0.995, lam=0.9)
    ends = np.nonzero(term)[0]
    trunc_before = np.concatenate([[0], np.cumsum(term == 2)])
    rng = np.random.default_rng(1)
    for k in rng.choice(len(ends) - 8, size=12, replace=False):
        a, b = ends[k] + 1, ends[k + 6]
        sl = slice(a, b + 1)
        t0, t1 = trunc_before[a], trunc_before[b + 1]
        want = G.gae(r[sl], term[sl], v[sl], tv[t0:t1], 0.995, 0.9)
        for j in range(3):
            assert np.array_equal(got[j][sl].view(np.uint32), want[j].view(np.uint32)), (a, b, j)
    # time it (reported, not gated)
    from emy_s_framework_amd.gae import ComputeGAEOnDevice
    d = [torch.from_numpy(x).cuda() for x in (r, term, v, tv)]
    ComputeGAEOnDevice(*d)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        ComputeGAEOnDevice(*d)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 5
    print(f"GAE over {T} steps: {ms:.3f} ms ({T * 17 / ms / 1e6:.1f} GB/s of inputs and outputs)", flush=True)

