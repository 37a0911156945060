"""rsx_gae_noalloc, the GAE entry point that allocates nothing (its scratch is
a fixed device array in the library), against rsx_gae through the C-ABI:
every output bit for bit on two calls in a row on one stream, the second with
a smaller buffer, so a stale scratch word from the first call would show."""
import ctypes as C

import numpy as np
import pytest

import rsxlib as R

pytestmark = pytest.mark.gpu


def _inputs(T, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=T)
    term = np.where(u < 0.002, 1, np.where(u < 0.004, 2, 0)).astype(np.int8)
    term[-1] = 2
    rews = (rng.normal(size=T) * 3).astype(np.float32)
    vals = rng.normal(size=T).astype(np.float32)
    tv = rng.normal(size=int((term == 2).sum())).astype(np.float32)
    return rews, term, vals, tv


def _run(fn, torch, d, T, params):
    outs = [torch.full((T,), np.nan, dtype=torch.float32, device="cuda") for _ in range(3)]
    clip = torch.full((1,), np.nan, dtype=torch.float32, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    st = fn(p(d[0]), p(d[1]), p(d[2]), p(d[3]), int(d[3].numel()), p(outs[0]), p(outs[1]), p(outs[2]), p(clip), T,
            C.byref(params), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, [o.cpu().numpy().view(np.uint32) for o in outs + [clip]]


def test_noalloc_gae_matches_gae_on_two_calls(torch_cuda):
    torch = torch_cuda
    from emy_s_framework_amd.gae import GAEParams
    lib = R.engine.load_library()
    for T, seed, params in ((600_011, 1, GAEParams(0.99, 0.95, 1.9, 1.0)), (70_001, 2, GAEParams(0.995, 0.9, 0.0, 0.0))):
        print(f"rsx_gae_noalloc vs rsx_gae: T={T} seed={seed}", flush=True)
        d = [torch.from_numpy(x).cuda() for x in _inputs(T, seed)]
        st_ref, want = _run(lib.rsx_gae, torch, d, T, params)
        st, got = _run(lib.rsx_gae_noalloc, torch, d, T, params)
        assert st_ref == 0 and st == 0, (st_ref, st)
        for k, name in enumerate(("advantages", "target values", "returns", "clip portion")):
            assert np.array_equal(got[k], want[k]), (T, name)
