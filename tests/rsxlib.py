"""ctypes bindings used by the tests: the oracle (reference CPU RocketSim built
from source into oracle/_ref/librsx_ref.so), the CPU build of the engine's tick
code (tests/cpu_check/librsx_cpucheck.so), and the record types of
include/rsx_types.h."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "emy-s-framework_amd")
MESH_DIR = os.path.join(PKG_DIR, "assets", "soccar")


def import_engine():
    """Imports the engine package (its folder name is not an identifier)."""
    if "emy_s_framework_amd" in sys.modules:
        return sys.modules["emy_s_framework_amd"]
    spec = importlib.util.spec_from_file_location("emy_s_framework_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["emy_s_framework_amd"] = mod
    spec.loader.exec_module(mod)
    return mod



engine = import_engine()
Vec3, RotMat, CarState, BallState = engine.Vec3, engine.RotMat, engine.CarState, engine.BallState
Controls = engine.CarControls
PadState = engine.BoostPadState
PADS = engine.PadStates


CAR_FLOAT_FIELDS = ["boost", "time_spent_boosting", "jump_time", "flip_time", "air_time", "air_time_since_jump",
                    "supersonic_time", "handbrake_val", "auto_flip_timer", "auto_flip_torque_scale"]
# Timers and ids of discrete events, compared bit for bit: the demolition
# respawn countdown and the bump cooldown (Car.cpp:43-79, 133-193), the ticks
# of the last ball hit and of its extra impulse (Arena.cpp:283-334, the source
# of the touch flag) and the id of the last car bumped.
CAR_EXACT_FIELDS = ["demo_respawn_timer", "car_contact_cooldown", "ball_hit_tick", "ball_hit_extra_impulse_tick",
                    "car_contact_other_id"]
CAR_FLAG_FIELDS = ["is_on_ground", "has_jumped", "has_double_jumped", "has_flipped", "is_flipping", "is_jumping",
                   "is_supersonic", "is_auto_flipping", "is_demoed", "world_contact", "ball_hit_valid"]


def car_vec(s):
    """pos(3) fwd(3) right(3) up(3) vel(3) angvel(3) as float64."""
    return np.array([s.pos.x, s.pos.y, s.pos.z, s.rot.forward.x, s.rot.forward.y, s.rot.forward.z,
                     s.rot.right.x, s.rot.right.y, s.rot.right.z, s.rot.up.x, s.rot.up.y, s.rot.up.z,
                     s.vel.x, s.vel.y, s.vel.z, s.ang_vel.x, s.ang_vel.y, s.ang_vel.z], dtype=np.float64)


def ball_vec(s):
    return np.array([s.pos.x, s.pos.y, s.pos.z, s.vel.x, s.vel.y, s.vel.z, s.ang_vel.x, s.ang_vel.y,
                     s.ang_vel.z], dtype=np.float64)


def _lib(path, fns):
    if not os.path.exists(path):
        return None
    lib = C.CDLL(path)
    for name, res, args in fns:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


VP = C.c_void_p
ORACLE_PATH = os.path.join(ROOT, "oracle", "_ref", "librsx_ref.so")
CPUCHECK_PATH = os.path.join(ROOT, "tests", "cpu_check", "librsx_cpucheck.so")


def load_oracle():
    lib = _lib(ORACLE_PATH, [
        ("ref_init", C.c_int, [C.c_char_p]),
        ("ref_arena_new", VP, [C.c_int, C.c_int]),
        ("ref_arena_free", None, [VP]),
        ("ref_get_car", None, [VP, C.c_int, C.POINTER(CarState)]),
        ("ref_set_car", None, [VP, C.c_int, C.POINTER(CarState)]),
        ("ref_set_controls", None, [VP, C.c_int, C.POINTER(Controls)]),
        ("ref_get_ball", None, [VP, C.POINTER(BallState)]),
        ("ref_set_ball", None, [VP, C.POINTER(BallState)]),
        ("ref_get_pads", None, [VP, C.POINTER(PadState)]),
        ("ref_set_pads", None, [VP, C.POINTER(PadState)]),
        ("ref_step", None, [VP, C.c_int]),
        ("ref_get_tick", C.c_uint64, [VP]),
        ("ref_set_tick", None, [VP, C.c_uint64]),
        ("ref_ball_scored", C.c_int, [VP]),
        ("ref_reset_kickoff", None, [VP, C.c_int]),
        ("ref_get_wheels", None, [VP, C.c_int, C.POINTER(C.c_float)]),
        ("ref_set_wheels", None, [VP, C.c_int, C.POINTER(C.c_float)]),
        ("ref_num_contacts", C.c_int, [VP]),
        ("ref_dump_contacts", C.c_int, [VP, C.POINTER(C.c_float), C.c_int]),
        ("ref_get_aabb", None, [VP, C.c_int, C.POINTER(C.c_float)]),
        ("ref_pad_config", None, [VP, C.POINTER(C.c_float)]),
        ("ref_sizeof_car_state", C.c_uint32, []),
        ("ref_num_cars", C.c_int, [VP]),
        ("ref_ray_convex", None, [C.c_int] + [C.POINTER(C.c_float)] * 5),
        ("ref_mesh_tri_info", C.c_int, [C.c_int, C.POINTER(C.c_float), C.c_int]),
        ("ref_car_order", C.c_int, [VP, C.POINTER(C.c_uint32), C.c_int]),
        ("ref_mesh_query", C.c_int, [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int32), C.c_int]),
        ("ref_trace_contacts", None, [VP, C.c_int]),
        ("ref_trace_read", C.c_int, [C.POINTER(C.c_float), C.c_int]),
        ("ref_env_step", None, [C.POINTER(VP), C.c_int, C.POINTER(C.c_int32), C.c_int, C.POINTER(Controls), C.c_int,
                                C.POINTER(C.c_int32), C.c_float, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.c_int]),
        ("ref_env_run", C.c_int64, [C.POINTER(VP), C.c_int, C.c_int, C.c_int, C.c_uint32, C.POINTER(Controls),
                                    C.c_int, C.POINTER(C.c_int32), C.c_float]),
    ])
    if lib is not None:
        assert lib.ref_init(MESH_DIR.encode()) == 0
    return lib


def load_cpucheck():
    return _lib(CPUCHECK_PATH, [
        ("cc_world_new", VP, [C.c_char_p]),
        ("cc_world_free", None, [VP]),
        ("cc_mesh_hashes", None, [VP, C.POINTER(C.c_uint32)]),
        ("cc_arena_new", VP, [VP, C.c_int, C.c_int]),
        ("cc_arena_free", None, [VP]),
        ("cc_set_car", None, [VP, C.c_int, C.POINTER(CarState)]),
        ("cc_get_car", None, [VP, C.c_int, C.POINTER(CarState)]),
        ("cc_set_ball", None, [VP, C.POINTER(BallState)]),
        ("cc_get_ball", None, [VP, C.POINTER(BallState)]),
        ("cc_set_pads", None, [VP, C.POINTER(PadState)]),
        ("cc_get_pads", None, [VP, C.POINTER(PadState)]),
        ("cc_set_controls", None, [VP, C.c_int, C.POINTER(Controls)]),
        ("cc_set_tick", None, [VP, C.c_uint64]),
        ("cc_get_tick", C.c_uint64, [VP]),
        ("cc_set_wheels", None, [VP, C.c_int, C.POINTER(C.c_float)]),
        ("cc_get_wheels", None, [VP, C.c_int, C.POINTER(C.c_float)]),
        ("cc_step", C.c_int, [VP, C.c_int]),
        ("cc_soa_roundtrip", None, [VP, C.c_int]),
        ("cc_dump_contacts", C.c_int, [VP, C.POINTER(C.c_float), C.c_int]),
        ("cc_ray_convex", None, [C.c_int] + [C.POINTER(C.c_float)] * 5),
        ("cc_tri_info", C.c_int, [VP, C.POINTER(C.c_float), C.c_int]),
        ("cc_contact_overflow", C.c_int, [VP]),
        ("cc_env_kind", C.c_int, [VP, C.POINTER(C.c_float), C.c_int, C.c_float, C.POINTER(C.c_float),
                                  C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]),
        ("cc_tri_rank", C.c_int, [VP, C.POINTER(C.c_int32), C.c_int]),
        ("cc_point_tri_dist2", None, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        ("cc_env", C.c_int, [VP, C.POINTER(C.c_float), C.c_int, C.c_float, C.POINTER(C.c_float),
                             C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        ("cc_train_rewards", None, [VP, C.POINTER(CarState), C.POINTER(CarState), C.c_int, C.POINTER(BallState),
                                    C.POINTER(BallState), C.POINTER(engine.TrainingRewardEntry), C.c_int, C.c_int,
                                    C.c_uint64, C.c_int, C.POINTER(C.c_float)]),
        ("cc_curriculum_reset", C.c_int, [VP, C.c_int, C.c_int, C.c_uint32, C.POINTER(C.c_float),
                                          C.POINTER(CarState), C.POINTER(BallState)]),
        ("cc_vanta_new", VP, []),
        ("cc_vanta_free", None, [VP]),
        ("cc_vanta_clear", None, [VP]),
        ("cc_vanta_predict", None, [VP, C.POINTER(C.c_float)]),
        ("cc_vanta_continue_check", C.c_int, [VP, C.c_int, C.c_int]),
        ("cc_set_solver_lds", None, [VP, C.c_int]),
        ("cc_vanta_build", None, [VP, VP, C.POINTER(C.c_float), C.c_float, C.c_float, C.c_int,
                                  C.POINTER(C.c_float)]),
    ])


def load_spec():
    """The CPU restatement of the RLGym obs / reward / terminal pieces (oracle/rlgym_spec.py)."""
    if "rlgym_spec" in sys.modules:
        return sys.modules["rlgym_spec"]
    path = os.path.join(ROOT, "oracle", "rlgym_spec.py")
    spec = importlib.util.spec_from_file_location("rlgym_spec", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["rlgym_spec"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_vanta_spec():
    """The CPU restatement of the VantaObs rows (oracle/vanta_spec.py)."""
    if "vanta_spec" in sys.modules:
        return sys.modules["vanta_spec"]
    path = os.path.join(ROOT, "oracle", "vanta_spec.py")
    spec = importlib.util.spec_from_file_location("vanta_spec", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vanta_spec"] = mod
    spec.loader.exec_module(mod)
    return mod


def parse_contacts(buf, n):
    out = []
    i = 0
    while i < n:
        b0, b1, npt, brk = int(buf[i]), int(buf[i + 1]), int(buf[i + 2]), buf[i + 3]
        i += 4
        pts = []
        for _ in range(npt):
            pts.append(np.array(buf[i:i + 22], dtype=np.float64))
            i += 22
        out.append((b0, b1, brk, pts))
    return out
