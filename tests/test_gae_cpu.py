"""Pins the CPU restatement of the trainer's GAE (oracle/gae_spec.py,
GAE.cpp:128-195) with closed forms, so that the device GAE can be checked
against it bit for bit."""
import importlib.util
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_spec = importlib.util.spec_from_file_location("gae_spec", os.path.join(ROOT, "oracle", "gae_spec.py"))
G = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(G)
F = np.float32


def test_gamma_zero_is_reward_minus_value():
    r = np.array([1, 2, 3, 4], F)
    v = np.array([0.5, 0.25, 1, 2], F)
    adv, tgt, ret, _ = G.gae(r, [0, 0, 1, 0], v, [], 0.0, 0.95)
    np.testing.assert_array_equal(adv, r - v)
    np.testing.assert_array_equal(ret, r)
    np.testing.assert_array_equal(tgt, r)


def test_lambda_zero_is_the_td_residual():
    r = np.array([1, -1, 2, 0.5, 3], F)
    v = np.array([0.1, 0.2, 0.3, 0.4, 0.5], F)
    term = [0, 1, 0, 2, 0]
    tv = np.array([7.0], F)
    g = 0.9
    adv, _, _, _ = G.gae(r, term, v, tv, g, 0.0)
    want = [r[0] + g * v[1] - v[0], r[1] - v[1], r[2] + g * v[3] - v[2], r[3] + g * 7.0 - v[3], r[4] + 0.0 - v[4]]
    np.testing.assert_allclose(adv, want, rtol=1e-6)


def test_one_episode_closed_form():
    """Constant reward, zero values, normal end, lambda 1: the advantage and
    the return at step t are the geometric sum r (1 - g^(n-t)) / (1 - g)."""
    n, g, r0 = 50, 0.97, 0.5
    r = np.full(n, r0, F)
    term = np.zeros(n, np.int8)
    term[-1] = 1
    adv, tgt, ret, _ = G.gae(r, term, np.zeros(n, F), [], g, 1.0)
    k = n - np.arange(n)
    want = r0 * (1 - g ** k) / (1 - g)
    np.testing.assert_allclose(adv, want, rtol=2e-5)
    np.testing.assert_allclose(ret, want, rtol=2e-5)
    np.testing.assert_array_equal(tgt, adv)


def test_truncation_bootstraps_in_buffer_order():
    """Two truncated episodes: the first truncation in the buffer reads the
    first bootstrap value (GAE.cpp consumes them from the back while walking
    backwards); a truncation stops the return chain."""
    r = np.zeros(4, F)
    v = np.zeros(4, F)
    adv, _, ret, _ = G.gae(r, [0, 2, 0, 2], v, np.array([10, 20], F), 0.5, 1.0)
    np.testing.assert_allclose(adv, [0.5 * 0.5 * 10, 0.5 * 10, 0.5 * 0.5 * 20, 0.5 * 20])
    np.testing.assert_array_equal(ret, [0, 0, 0, 0])
    with pytest.raises(ValueError):
        G.gae(r, [0, 2, 0, 2], v, np.array([10], F), 0.5, 1.0)
    with pytest.raises(ValueError):
        G.gae(r, [0, 2, 0, 0], v, np.array([10, 20], F), 0.5, 1.0)


def test_last_step_without_terminal_bootstraps_zero():
    r = np.array([1, 1], F)
    v = np.array([3, 5], F)
    adv, _, ret, _ = G.gae(r, [0, 0], v, [], 0.5, 1.0)
    np.testing.assert_allclose(adv, [1 + 0.5 * 5 - 3 + 0.5 * (1 - 5), 1 - 5])
    np.testing.assert_allclose(ret, [1 + 0.5 * 1, 1])


def test_all_truncated_is_one_step_td():
    """Every step truncated (the trainer's fast path, GAE.cpp:86-126): the
    advantage is r + g * bootstrap - v at each step, returns are rewards."""
    rng = np.random.default_rng(0)
    n = 40
    r, v, tv = (rng.normal(size=n).astype(F) for _ in range(3))
    adv, tgt, ret, _ = G.gae(r, np.full(n, 2, np.int8), v, tv, 0.99, 0.95)
    np.testing.assert_array_equal(adv, (r + F(0.99) * tv - v).astype(F))
    np.testing.assert_array_equal(ret, r)
    np.testing.assert_array_equal(tgt, (v + adv).astype(F))


def test_clip_portion_closed_form():
    r = np.array([4, 1, -6], F)
    _, _, _, portion = G.gae(r, [0, 0, 1], np.zeros(3, F), [], 0.9, 0.9, return_std=2.0, clip_range=1.0)
    assert abs(portion - (5.5 - 2.5) / 5.5) < 1e-6


def test_argument_checks_without_device_work():
    """rsx_gae with no steps returns at once; null buffers, negative counts and
    null parameters are refused before any device call."""
    import ctypes as C
    import rsxlib as R
    from emy_s_framework_amd.gae import GAEParams
    lib = R.engine.load_library()
    p = GAEParams(0.99, 0.95, 0.0, 0.0)
    null = C.c_void_p(0)
    assert lib.rsx_gae(null, null, null, null, 0, null, null, null, null, 0, C.byref(p), null) == 0
    assert lib.rsx_gae(null, null, null, null, 0, null, null, null, null, 10, C.byref(p), null) != 0
    assert lib.rsx_gae(null, null, null, null, 0, null, null, null, null, -1, C.byref(p), null) != 0
    assert lib.rsx_gae(null, null, null, null, -2, null, null, null, null, 0, C.byref(p), null) != 0
    assert lib.rsx_gae(null, null, null, null, 0, null, null, null, null, 0, None, null) != 0


def test_noalloc_refuses_buffers_past_its_scratch():
    """rsx_gae_noalloc refuses T above RSX_GAE_NOALLOC_MAX_STEPS (read from the
    header) before any other check or device work; at the maximum the same
    call gets as far as the null-buffer check."""
    import ctypes as C
    import re
    import rsxlib as R
    from emy_s_framework_amd.gae import GAEParams
    hdr = open(os.path.join(ROOT, "include", "rsx_gae.h")).read()
    mx = eval(re.search(r"#define RSX_GAE_NOALLOC_MAX_STEPS \(([^)]*)\)", hdr).group(1))
    assert mx == 1 << 26
    lib = R.engine.load_library()
    p = GAEParams(0.99, 0.95, 0.0, 0.0)
    null = C.c_void_p(0)
    assert lib.rsx_gae_noalloc(null, null, null, null, 0, null, null, null, null, mx + 1, C.byref(p), null) != 0
    assert b"RSX_GAE_NOALLOC_MAX_STEPS" in lib.rsx_last_error()
    assert lib.rsx_gae_noalloc(null, null, null, null, 0, null, null, null, null, mx, C.byref(p), null) != 0
    assert b"null device buffer" in lib.rsx_last_error()
    assert lib.rsx_gae_noalloc(null, null, null, null, 0, null, null, null, null, 0, C.byref(p), null) == 0
