"""Host-side mirror of RocketSimCudaBatch (RocketSimCuda/include/RocketSimCuda.h:
107-262) over the gfx950 engine library librsx.so.

The method names are the reference's; each call goes straight to the C-ABI of
include/rsx.h.  There is no CPU fallback: constructing a batch without the
built library, or without a HIP device, raises."""
import ctypes as C
import os

from .records import (ArenaInfo, BallState, BoostPadState, CarControls, CarState, Config, EnvConfig, NUM_PADS,
                      PadStates, CARS_BLUE_FIRST, REWARD, MAX_TRAINING_REWARDS, TrainingRewardEntry, Tuning,
                      TuneResult, VantaConfig)

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "librsx.so")
MESH_DIR = os.path.join(PKG_DIR, "assets", "soccar")

_VP = C.c_void_p
_I = C.c_int32
_SIGS = [
    ("rsx_last_error", C.c_char_p, []),
    ("rsx_create", _I, [C.POINTER(Config), C.POINTER(_VP)]),
    ("rsx_destroy", _I, [_VP]),
    ("rsx_num_arenas", _I, [_VP, C.POINTER(_I)]),
    ("rsx_num_cars", _I, [_VP, C.POINTER(_I)]),
    ("rsx_team_size", _I, [_VP, _I, C.POINTER(_I)]),
    ("rsx_get_car_state", _I, [_VP, _I, _I, C.POINTER(CarState)]),
    ("rsx_set_car_state", _I, [_VP, _I, _I, C.POINTER(CarState)]),
    ("rsx_get_ball_state", _I, [_VP, _I, C.POINTER(BallState)]),
    ("rsx_set_ball_state", _I, [_VP, _I, C.POINTER(BallState)]),
    ("rsx_get_pad_states", _I, [_VP, _I, C.POINTER(BoostPadState)]),
    ("rsx_set_pad_states", _I, [_VP, _I, C.POINTER(BoostPadState)]),
    ("rsx_get_arena_info", _I, [_VP, _I, C.POINTER(ArenaInfo)]),
    ("rsx_set_tick_count", _I, [_VP, _I, C.c_uint64]),
    ("rsx_get_all_car_states", _I, [_VP, C.POINTER(CarState), _I]),
    ("rsx_set_all_car_states", _I, [_VP, C.POINTER(CarState), _I]),
    ("rsx_get_all_ball_states", _I, [_VP, C.POINTER(BallState), _I]),
    ("rsx_set_all_ball_states", _I, [_VP, C.POINTER(BallState), _I]),
    ("rsx_get_all_pad_states", _I, [_VP, C.POINTER(BoostPadState), _I]),
    ("rsx_set_all_pad_states", _I, [_VP, C.POINTER(BoostPadState), _I]),
    ("rsx_get_all_arena_infos", _I, [_VP, C.POINTER(ArenaInfo), _I]),
    ("rsx_upload_action_lut", _I, [_VP, C.POINTER(CarControls), _I]),
    ("rsx_set_actions_from_indices", _I, [_VP, C.POINTER(C.c_int32), _I]),
    ("rsx_set_actions_device", _I, [_VP, _VP, _I, _VP]),
    ("rsx_default_action_lut", _I, [C.POINTER(CarControls), _I, C.POINTER(_I)]),
    ("rsx_reset_arena", _I, [_VP, _I]),
    ("rsx_reset_all_arenas", _I, [_VP, _VP]),
    ("rsx_set_all_controls", _I, [_VP, C.POINTER(CarControls), _I]),
    ("rsx_set_car_controls", _I, [_VP, _I, _I, C.POINTER(CarControls)]),
    ("rsx_configure_training_rewards", _I, [_VP, C.POINTER(TrainingRewardEntry), _I]),
    ("rsx_get_training_rewards", _I, [_VP, C.POINTER(TrainingRewardEntry), _I, C.POINTER(C.c_int32)]),
    ("rsx_set_reset_curriculum", _I, [_VP, C.c_float, C.c_float, C.c_float, C.c_float]),
    ("rsx_get_reset_curriculum", _I, [_VP, C.POINTER(C.c_float)]),
    ("rsx_configure_env", _I, [_VP, C.POINTER(EnvConfig)]),
    ("rsx_get_env_config", _I, [_VP, C.POINTER(EnvConfig)]),
    ("rsx_step_env", _I, [_VP, _VP, _VP]),
    ("rsx_build_obs", _I, [_VP, _VP]),
    ("rsx_env_buffers", _I, [_VP, C.POINTER(_VP), C.POINTER(_VP), C.POINTER(_VP), C.POINTER(C.c_int64)]),
    ("rsx_obs_layout", _I, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int32), _I]),
    ("rsx_copy_obs", _I, [_VP, C.POINTER(C.c_float), C.c_int64]),
    ("rsx_copy_rewards", _I, [_VP, C.POINTER(C.c_float), _I]),
    ("rsx_copy_terminals", _I, [_VP, C.POINTER(C.c_uint8), _I]),
    ("rsx_step_ticks", _I, [_VP, _I, _VP]),
    ("rsx_synchronize", _I, [_VP]),
    ("rsx_set_action_masks", _I, [_VP, _I]),
    ("rsx_action_masks_buffer", _I, [_VP, C.POINTER(_VP)]),
    ("rsx_copy_action_masks", _I, [_VP, C.POINTER(C.c_uint8), C.c_int64]),
    ("rsx_set_tuning", _I, [_VP, C.POINTER(Tuning)]),
    ("rsx_get_tuning", _I, [_VP, C.POINTER(Tuning)]),
    ("rsx_autotune", _I, [_VP, _VP, _I, C.POINTER(TuneResult), _I, C.POINTER(_I)]),
    ("rsx_capture_env_graph", _I, [_VP, _VP, _I]),
    ("rsx_configure_vanta", _I, [_VP, C.POINTER(VantaConfig)]),
    ("rsx_get_vanta_config", _I, [_VP, C.POINTER(VantaConfig)]),
    ("rsx_get_vanta_prediction", _I, [_VP, _I, C.POINTER(BallState), C.POINTER(_I)]),
    ("rsx_launch_env_graph", _I, [_VP, _VP]),
    # learned broadphase cull (include/rsx_sdf.h)
    ("rsx_sdf_num_params", _I, []),
    ("rsx_sdf_create", _I, [_VP, C.c_uint64, C.POINTER(_VP)]),
    ("rsx_sdf_destroy", _I, [_VP]),
    ("rsx_sdf_mesh_triangles", _I, [_VP, C.POINTER(C.c_float), C.c_int64, C.POINTER(_I)]),
    ("rsx_sdf_exact_distance", _I, [_VP, _VP, _I, _VP, _VP]),
    ("rsx_sdf_eval", _I, [_VP, _VP, _I, _VP, _VP]),
    ("rsx_sdf_train_step", _I, [_VP, _VP, _VP, _I, C.c_float, C.c_float, C.c_float, _VP, _VP]),
    ("rsx_sdf_gradients", _I, [_VP, _VP, _VP, _I, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("rsx_sdf_get_params", _I, [_VP, C.POINTER(C.c_float)]),
    ("rsx_sdf_set_params", _I, [_VP, C.POINTER(C.c_float)]),
    ("rsx_attach_cull", _I, [_VP, _VP, C.c_float, C.c_float, _I, _I]),
    ("rsx_get_cull_stats", _I, [_VP, C.POINTER(C.c_uint64 * 3)]),
    ("rsx_reset_cull_stats", _I, [_VP]),
    ("rsx_cull_reach", _I, [_VP, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    # device GAE (include/rsx_gae.h)
    ("rsx_gae", _I, [_VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
    ("rsx_gae_noalloc", _I, [_VP, _VP, _VP, _VP, _I, _VP, _VP, _VP, _VP, _I, _VP, _VP]),
]
EXPORTED = [name for name, _, _ in _SIGS]

_lib = None


class RsxError(RuntimeError):
    pass


def default_action_lut():
    """RLGym DefaultAction's 90-entry table as built by the library."""
    lib = load_library()
    n = _I()
    _check(lib.rsx_default_action_lut(None, 0, C.byref(n)))
    out = (CarControls * n.value)()
    _check(lib.rsx_default_action_lut(out, n.value, C.byref(n)))
    return out


def load_library(path=LIB_PATH):
    """Loads librsx.so (once) and declares the signatures of include/rsx.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RsxError(f"{path} is missing: the gfx950 engine library must be built first "
                       "(__graft_entry__.build())")
    lib = C.CDLL(path)
    for name, res, args in _SIGS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(status):
    if status != 0:
        raise RsxError(f"rsx error {status}: {_lib.rsx_last_error().decode(errors='replace')}")


class RocketSimBatch:
    """N soccar arenas stepped on one GPU.

    team_sizes: int (every arena) or a sequence of per-arena team sizes 1..3.
    Arena a holds 2*T cars; car_order CARS_BLUE_FIRST puts blue in slots
    0..T-1 and orange in T..2T-1, CARS_ALTERNATING alternates blue and orange
    (CudaEnvSet's AddCar order)."""

    def __init__(self, num_arenas, team_sizes=1, tick_skip=8, seed=0, device=0, mesh_dir=MESH_DIR,
                 car_order=CARS_BLUE_FIRST):
        self._lib = load_library()
        cfg = Config()
        cfg.num_arenas = int(num_arenas)
        if isinstance(team_sizes, int):
            cfg.default_team_size = team_sizes
            self._teams = None
        else:
            arr = (C.c_uint8 * num_arenas)(*[int(t) for t in team_sizes])
            self._teams = arr
            cfg.team_size = C.cast(arr, C.POINTER(C.c_uint8))
        cfg.mesh_dir = mesh_dir.encode()
        cfg.tick_skip = tick_skip
        cfg.seed = seed
        cfg.device = device
        cfg.car_order = car_order
        h = _VP()
        _check(self._lib.rsx_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self.num_arenas = int(num_arenas)
        self.team_size = [self.GetTeamSize(a) for a in range(self.num_arenas)]
        n = _I()
        _check(self._lib.rsx_num_cars(self._h, C.byref(n)))
        self.num_cars = n.value
        assert self.num_cars == sum(2 * t for t in self.team_size)

    def Destroy(self):
        if getattr(self, "_h", None):
            self._lib.rsx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.Destroy()

    def GetTeamSize(self, arena):
        out = _I()
        _check(self._lib.rsx_team_size(self._h, arena, C.byref(out)))
        return out.value

    def GetCarState(self, arena, car):
        s = CarState()
        _check(self._lib.rsx_get_car_state(self._h, arena, car, C.byref(s)))
        return s

    def SetCarState(self, arena, car, state):
        _check(self._lib.rsx_set_car_state(self._h, arena, car, C.byref(state)))

    def GetBallState(self, arena):
        s = BallState()
        _check(self._lib.rsx_get_ball_state(self._h, arena, C.byref(s)))
        return s

    def SetBallState(self, arena, state):
        _check(self._lib.rsx_set_ball_state(self._h, arena, C.byref(state)))

    def GetBoostPadStates(self, arena):
        p = PadStates()
        _check(self._lib.rsx_get_pad_states(self._h, arena, p))
        return p

    def SetBoostPadStates(self, arena, pads):
        if len(pads) != NUM_PADS:
            raise ValueError("need 34 pad states")
        _check(self._lib.rsx_set_pad_states(self._h, arena, pads))

    def GetArenaInfo(self, arena):
        s = ArenaInfo()
        _check(self._lib.rsx_get_arena_info(self._h, arena, C.byref(s)))
        return s

    def SetTickCount(self, arena, tick):
        _check(self._lib.rsx_set_tick_count(self._h, arena, tick))

    def ResetArena(self, arena):
        _check(self._lib.rsx_reset_arena(self._h, arena))

    def ResetAllArenas(self, stream=None):
        _check(self._lib.rsx_reset_all_arenas(self._h, stream))

    def GetAllCarStates(self):
        out = (CarState * self.num_cars)()
        _check(self._lib.rsx_get_all_car_states(self._h, out, self.num_cars))
        return out

    def SetAllCarStates(self, states):
        _check(self._lib.rsx_set_all_car_states(self._h, states, len(states)))

    def GetAllBallStates(self):
        out = (BallState * self.num_arenas)()
        _check(self._lib.rsx_get_all_ball_states(self._h, out, self.num_arenas))
        return out

    def SetAllBallStates(self, states):
        _check(self._lib.rsx_set_all_ball_states(self._h, states, len(states)))

    def GetAllBoostPadStates(self):
        out = (BoostPadState * (NUM_PADS * self.num_arenas))()
        _check(self._lib.rsx_get_all_pad_states(self._h, out, len(out)))
        return out

    def SetAllBoostPadStates(self, states):
        _check(self._lib.rsx_set_all_pad_states(self._h, states, len(states)))

    def GetAllArenaInfos(self):
        out = (ArenaInfo * self.num_arenas)()
        _check(self._lib.rsx_get_all_arena_infos(self._h, out, self.num_arenas))
        return out

    def UploadDiscreteActionLut(self, lut):
        _check(self._lib.rsx_upload_action_lut(self._h, lut, len(lut)))

    def SetDiscreteActionsFromIndices(self, indices):
        """indices: sequence / int32 numpy array of rsx_num_cars action indices (host)."""
        import numpy as np
        arr = np.ascontiguousarray(indices, dtype=np.int32)
        _check(self._lib.rsx_set_actions_from_indices(self._h, arr.ctypes.data_as(C.POINTER(C.c_int32)), arr.size))

    def SetDiscreteActionsFromDevice(self, device_ptr, count, stream=None):
        """device_ptr: address of a device int32 array (e.g. tensor.data_ptr())."""
        _check(self._lib.rsx_set_actions_device(self._h, device_ptr, count, stream))

    # --- env half of the step -------------------------------------------------
    def ConfigureEnv(self, goal=1.0, concede=-1.0, touch=1.0, vel_ball_to_goal=1.0, vel_player_to_ball=1.0,
                     no_touch_timeout_s=0.0, goal_terminal=True, auto_reset=True, obs_builder=0,
                     no_touch_stagger=False):
        """ConfigureTrainingTerminals and the stub rewards' weights;
        obs_builder 0 (DefaultObs), 1 (AdvancedObs,
        BuildAdvancedObsAndDefaultMasks) or 2 (VantaObs, BuildVantaObs;
        see ConfigureVanta); no_touch_stagger: resets start the
        no-touch timer at a random age, as the reference's ResetArena."""
        c = EnvConfig(goal, concede, touch, vel_ball_to_goal, vel_player_to_ball, no_touch_timeout_s,
                      int(goal_terminal), int(auto_reset), int(obs_builder), int(no_touch_stagger))
        _check(self._lib.rsx_configure_env(self._h, C.byref(c)))

    def ConfigureTrainingRewards(self, entries):
        """ConfigureTrainingRewards (RocketSimCuda.h:329).  entries: a list of
        TrainingRewardEntry, or of dicts with "id" (an RSX_REWARD_* number or
        a TrainingRewardID name), "weight" and optional "params" (up to 10),
        "zero_sum", "team_spirit", "opponent_scale" (default 1).  An empty
        list returns to the stub rewards."""
        if len(entries) > MAX_TRAINING_REWARDS:
            raise ValueError("at most 16 reward entries")
        arr = (TrainingRewardEntry * max(1, len(entries)))()
        for k, e in enumerate(entries):
            if isinstance(e, TrainingRewardEntry):
                arr[k] = e
                continue
            rid = REWARD[e["id"]] if isinstance(e["id"], str) else int(e["id"])
            params = list(e.get("params", ())) + [0.0] * 10
            arr[k] = TrainingRewardEntry(rid, float(e.get("weight", 1.0)), (C.c_float * 10)(*params[:10]),
                                         int(bool(e.get("zero_sum", False))), float(e.get("team_spirit", 0.0)),
                                         float(e.get("opponent_scale", 1.0)))
        _check(self._lib.rsx_configure_training_rewards(self._h, arr, len(entries)))

    def GetTrainingRewards(self):
        n = C.c_int32(0)
        _check(self._lib.rsx_get_training_rewards(self._h, None, 0, C.byref(n)))
        arr = (TrainingRewardEntry * max(1, n.value))()
        _check(self._lib.rsx_get_training_rewards(self._h, arr, n.value, C.byref(n)))
        return [arr[k] for k in range(n.value)]

    def SetResetCurriculum(self, kickoff_w, fuzzed_w, aerial_w, chase_w=0.0):
        """SetResetCurriculum (RocketSimCuda.h:283): relative weights of the
        kickoff, fuzzed kickoff, aerial and ball-chase resets."""
        _check(self._lib.rsx_set_reset_curriculum(self._h, kickoff_w, fuzzed_w, aerial_w, chase_w))

    def GetResetCurriculum(self):
        w = (C.c_float * 4)()
        _check(self._lib.rsx_get_reset_curriculum(self._h, w))
        return list(w)

    def ConfigureVanta(self, predict_ball=False, only_closest_opp=False, dodge_deadzone=0.5, tick_skip=0):
        """The VantaObs parameters of BatchConfig (vantaPredictBall,
        vantaOnlyClosestOpp, vantaDodgeDeadzone, vantaTickSkip;
        RocketSimCuda.h:184-189); tick_skip 0 uses the engine's."""
        c = VantaConfig(int(bool(predict_ball)), int(bool(only_closest_opp)), float(dodge_deadzone), int(tick_skip))
        _check(self._lib.rsx_configure_vanta(self._h, C.byref(c)))

    def GetVantaConfig(self):
        c = VantaConfig()
        _check(self._lib.rsx_get_vanta_config(self._h, C.byref(c)))
        return c

    def GetVantaPrediction(self, arena):
        """The ball-prediction slices (3 BallStates, uu) of the arena's last
        Vanta build, or None when it had none."""
        out = (BallState * 3)()
        has = C.c_int32(0)
        _check(self._lib.rsx_get_vanta_prediction(self._h, arena, out, C.byref(has)))
        return list(out) if has.value else None

    def StepEnv(self, device_actions_ptr, stream=None):
        """Fused env step from a device int32 array of num_cars action indices."""
        _check(self._lib.rsx_step_env(self._h, device_actions_ptr, stream))

    def BuildObs(self, stream=None):
        _check(self._lib.rsx_build_obs(self._h, stream))

    def EnvBuffers(self):
        """(obs ptr, rewards ptr, terminals ptr, obs float count) device addresses."""
        o, r, t, n = _VP(), _VP(), _VP(), C.c_int64()
        _check(self._lib.rsx_env_buffers(self._h, C.byref(o), C.byref(r), C.byref(t), C.byref(n)))
        return o.value, r.value, t.value, n.value

    def ObsLayout(self):
        import numpy as np
        off = np.zeros(self.num_cars, dtype=np.int64)
        ln = np.zeros(self.num_cars, dtype=np.int32)
        _check(self._lib.rsx_obs_layout(self._h, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                        ln.ctypes.data_as(C.POINTER(C.c_int32)), self.num_cars))
        return off, ln

    def CopyBuiltObs(self):
        import numpy as np
        n = self.EnvBuffers()[3]
        out = np.zeros(n, dtype=np.float32)
        _check(self._lib.rsx_copy_obs(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), n))
        return out

    def CopyBuiltRewards(self):
        import numpy as np
        out = np.zeros(self.num_cars, dtype=np.float32)
        _check(self._lib.rsx_copy_rewards(self._h, out.ctypes.data_as(C.POINTER(C.c_float)), self.num_cars))
        return out

    def CopyBuiltTerminals(self):
        import numpy as np
        out = np.zeros(self.num_arenas, dtype=np.uint8)
        _check(self._lib.rsx_copy_terminals(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), self.num_arenas))
        return out

    def SetCarControls(self, arena, car, controls):
        """SetCarControls (RocketSimCuda.h:288)."""
        _check(self._lib.rsx_set_car_controls(self._h, arena, car, C.byref(controls)))

    def SetAllCarControls(self, controls):
        """controls: ctypes array of CarControls, arena-major, 2*T per arena."""
        _check(self._lib.rsx_set_all_controls(self._h, controls, len(controls)))

    def Step(self, ticks=1, stream=None):
        _check(self._lib.rsx_step_ticks(self._h, ticks, stream))

    def Synchronize(self):
        _check(self._lib.rsx_synchronize(self._h))

    def SetTuning(self, block_lanes=64, arenas_per_wave=64, pool_narrowphase=1):
        """Launch shape of the arena kernels (results do not depend on it)."""
        _check(self._lib.rsx_set_tuning(self._h, C.byref(Tuning(block_lanes, arenas_per_wave, pool_narrowphase))))

    def GetTuning(self):
        """(block_lanes, arenas_per_wave, pool_narrowphase)."""
        t = Tuning()
        _check(self._lib.rsx_get_tuning(self._h, C.byref(t)))
        return t.block_lanes, t.arenas_per_wave, t.pool_narrowphase

    def Autotune(self, device_actions_ptr, env_steps=3):
        """Times every launch shape on the current arenas, keeps the fastest and
        restores the arenas; returns [(block_lanes, arenas_per_wave, pool_narrowphase, ms or None)]."""
        res = (TuneResult * 44)()
        n = _I()
        _check(self._lib.rsx_autotune(self._h, device_actions_ptr, env_steps, res, 44, C.byref(n)))
        return [(r.tuning.block_lanes, r.tuning.arenas_per_wave, r.tuning.pool_narrowphase,
                 r.ms_per_env_step if r.feasible else None) for r in res[:n.value]]

    def CaptureEnvGraph(self, device_actions_ptr, env_steps=1):
        """Records env_steps StepEnv calls on device_actions_ptr into a hipGraph."""
        _check(self._lib.rsx_capture_env_graph(self._h, device_actions_ptr, env_steps))

    def LaunchEnvGraph(self, stream=None):
        _check(self._lib.rsx_launch_env_graph(self._h, stream))

    def SetActionMasks(self, enabled=True):
        """DefaultAction masks written by StepEnv / BuildObs (90 bytes per car)."""
        _check(self._lib.rsx_set_action_masks(self._h, int(enabled)))

    def ActionMasksBuffer(self):
        p = _VP()
        _check(self._lib.rsx_action_masks_buffer(self._h, C.byref(p)))
        return p.value

    def CopyBuiltDefaultActionMasks(self):
        import numpy as np
        out = np.zeros((self.num_cars, 90), dtype=np.uint8)
        _check(self._lib.rsx_copy_action_masks(self._h, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size))
        return out
