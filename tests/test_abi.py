"""The C-ABI library: it loads, exports every entry point include/*.h
declares, its records have the header's layout, and without a GPU it fails
loudly instead of falling back to anything."""
import ctypes as C
import os
import re
import subprocess

import pytest

import rsxlib as R

ROOT = R.ROOT
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def declared_functions():
    names = []
    for h in HEADERS:
        txt = open(h).read()
        names += re.findall(r"^(?:rsx_status|const char\*|void|int32_t)\s+(rsx_\w+)\s*\(", txt, re.M)
    return names


def _lib():
    if not os.path.exists(R.engine.LIB_PATH):
        pytest.skip("librsx.so is not built")
    return C.CDLL(R.engine.LIB_PATH)


def test_exports_every_declared_symbol():
    lib = _lib()
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(R.engine.EXPORTED)


def test_record_layouts_match_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "rsx.h"\nint main(void){printf("%zu %zu %zu '
                   '%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(rsx_car_state), sizeof(rsx_ball_state), '
                   'sizeof(rsx_pad_state), sizeof(rsx_controls), sizeof(rsx_arena_info), '
                   'offsetof(rsx_car_state, last_controls), offsetof(rsx_car_state, ball_hit_tick), '
                   'offsetof(rsx_arena_info, goal_team), offsetof(rsx_arena_info, contact_overflow), '
                   'sizeof(rsx_env_config), offsetof(rsx_env_config, obs_builder), '
                   'offsetof(rsx_env_config, no_touch_stagger), sizeof(rsx_training_reward_entry), '
                   'offsetof(rsx_training_reward_entry, is_zero_sum), '
                   'offsetof(rsx_training_reward_entry, opponent_scale), sizeof(rsx_vanta_config), '
                   'offsetof(rsx_vanta_config, tick_skip));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    E = R.engine
    want = [C.sizeof(E.CarState), C.sizeof(E.BallState), C.sizeof(E.BoostPadState), C.sizeof(E.CarControls),
            C.sizeof(E.ArenaInfo), E.CarState.last_controls.offset, E.CarState.ball_hit_tick.offset,
            E.ArenaInfo.goal_team.offset, E.ArenaInfo.contact_overflow.offset, C.sizeof(E.EnvConfig),
            E.EnvConfig.obs_builder.offset, E.EnvConfig.no_touch_stagger.offset, C.sizeof(E.TrainingRewardEntry),
            E.TrainingRewardEntry.is_zero_sum.offset, E.TrainingRewardEntry.opponent_scale.offset,
            C.sizeof(E.VantaConfig), E.VantaConfig.tick_skip.offset]
    assert got == want
    # TrainingRewardEntry's own layout (RocketSimCuda.h:235-242: uint8 id, float weight,
    # float params[10], uint8 isZeroSum, float teamSpirit, float opponentScale)
    assert got[-5:-2] == [60, 48, 56]


def test_invalid_config_is_rejected():
    R.engine.load_library()
    lib = R.engine.load_library()
    cfg = R.engine.records.Config()
    cfg.num_arenas = 0
    h = C.c_void_p()
    assert lib.rsx_create(C.byref(cfg), C.byref(h)) == 1
    assert b"num_arenas" in lib.rsx_last_error()


def test_no_device_fails_loudly():
    """On a machine without a HIP device the batch constructor raises; there is
    no CPU path behind it."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(R.engine.RsxError):
        R.engine.RocketSimBatch(4, 1)
