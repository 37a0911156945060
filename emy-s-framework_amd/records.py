"""ctypes mirrors of the public C records of include/rsx_types.h.

Field meaning and units follow RocketSim's CarState / BallState /
BoostPadState / CarControls (RocketSim src/Sim/Car/Car.h:14-115,
src/Sim/Ball/Ball.h:13-40, src/Sim/BoostPad/BoostPad.h:23-35,
src/Sim/CarControls.h:6-30): positions in uu, velocities in uu/s, angular
velocities in rad/s, rotations as the forward/right/up axes."""
import ctypes as C

import numpy as np

NUM_PADS = 34
MAX_CARS = 6


class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def np(self):
        return np.array([self.x, self.y, self.z], dtype=np.float32)

    def set(self, v):
        self.x, self.y, self.z = (float(v[0]), float(v[1]), float(v[2]))


class RotMat(C.Structure):
    _fields_ = [("forward", Vec3), ("right", Vec3), ("up", Vec3)]


class CarControls(C.Structure):
    _fields_ = [("throttle", C.c_float), ("steer", C.c_float), ("pitch", C.c_float), ("yaw", C.c_float),
                ("roll", C.c_float), ("jump", C.c_uint8), ("boost", C.c_uint8), ("handbrake", C.c_uint8),
                ("_pad", C.c_uint8)]


class CarState(C.Structure):
    _fields_ = [("pos", Vec3), ("rot", RotMat), ("vel", Vec3), ("ang_vel", Vec3),
                ("boost", C.c_float), ("time_spent_boosting", C.c_float), ("jump_time", C.c_float),
                ("flip_time", C.c_float), ("air_time", C.c_float), ("air_time_since_jump", C.c_float),
                ("supersonic_time", C.c_float), ("handbrake_val", C.c_float), ("auto_flip_timer", C.c_float),
                ("auto_flip_torque_scale", C.c_float), ("demo_respawn_timer", C.c_float),
                ("car_contact_cooldown", C.c_float), ("flip_rel_torque", Vec3), ("world_contact_normal", Vec3),
                ("is_on_ground", C.c_uint8), ("has_jumped", C.c_uint8), ("has_double_jumped", C.c_uint8),
                ("has_flipped", C.c_uint8), ("is_flipping", C.c_uint8), ("is_jumping", C.c_uint8),
                ("is_supersonic", C.c_uint8), ("is_auto_flipping", C.c_uint8), ("is_demoed", C.c_uint8),
                ("world_contact", C.c_uint8), ("ball_hit_valid", C.c_uint8), ("team", C.c_uint8),
                ("wheels_with_contact", C.c_uint8 * 4), ("id", C.c_uint32), ("car_contact_other_id", C.c_uint32),
                ("ball_hit_tick", C.c_uint64), ("ball_hit_extra_impulse_tick", C.c_uint64),
                ("ball_hit_rel_pos", Vec3), ("ball_hit_ball_pos", Vec3), ("ball_hit_extra_vel", Vec3),
                ("last_controls", CarControls)]


class BallState(C.Structure):
    _fields_ = [("pos", Vec3), ("vel", Vec3), ("ang_vel", Vec3)]


class BoostPadState(C.Structure):
    _fields_ = [("is_active", C.c_uint8), ("_pad", C.c_uint8 * 3), ("cooldown", C.c_float),
                ("prev_locked_car_id", C.c_uint32)]


PadStates = BoostPadState * NUM_PADS


class ArenaInfo(C.Structure):
    _fields_ = [("tick_count", C.c_uint64), ("num_cars", C.c_int32), ("team_size", C.c_int32),
                ("goal_scored", C.c_uint8), ("contact_overflow", C.c_uint8), ("_pad", C.c_uint8 * 2),
                ("goal_team", C.c_int32),
                ("no_touch_time", C.c_float), ("episode", C.c_uint32)]


class Config(C.Structure):
    _fields_ = [("num_arenas", C.c_int32), ("team_size", C.POINTER(C.c_uint8)), ("default_team_size", C.c_int32),
                ("mesh_dir", C.c_char_p), ("tick_skip", C.c_int32), ("no_touch_timeout_s", C.c_float),
                ("seed", C.c_uint32), ("device", C.c_int32), ("car_order", C.c_int32)]


class EnvConfig(C.Structure):
    _fields_ = [("goal_weight", C.c_float), ("concede_scale", C.c_float), ("touch_weight", C.c_float),
                ("vel_ball_to_goal_weight", C.c_float), ("vel_player_to_ball_weight", C.c_float),
                ("no_touch_timeout_s", C.c_float), ("goal_terminal", C.c_int32), ("auto_reset", C.c_int32),
                ("obs_builder", C.c_int32), ("no_touch_stagger", C.c_int32)]


class VantaConfig(C.Structure):
    """rsx_vanta_config: BatchConfig's vanta* fields (RocketSimCuda.h:184-189)."""
    _fields_ = [("predict_ball", C.c_int32), ("only_closest_opp", C.c_int32), ("dodge_deadzone", C.c_float),
                ("tick_skip", C.c_int32)]


class TrainingRewardEntry(C.Structure):
    """rsx_training_reward_entry = TrainingRewardEntry (RocketSimCuda.h:235-242)."""
    _fields_ = [("id", C.c_uint8), ("weight", C.c_float), ("params", C.c_float * 10), ("is_zero_sum", C.c_uint8),
                ("team_spirit", C.c_float), ("opponent_scale", C.c_float)]


# TrainingRewardID (RocketSimCuda.h:200-231), the RSX_REWARD_* values
REWARD_IDS = ["UNKNOWN", "GOAL_REWARD", "VELOCITY_BALL_TO_GOAL", "VELOCITY_PLAYER_TO_BALL", "FACE_BALL", "TOUCH_BALL",
              "TOUCH_ACCEL", "STRONG_TOUCH", "SPEED", "VELOCITY", "AIR", "WAVEDASH", "PICKUP_BOOST", "SAVE_BOOST",
              "BUMP", "BUMPED_PENALTY", "DEMO", "DEMOED_PENALTY", "KICKOFF_PROXIMITY", "TEAMMATE_BUMP_PENALTY",
              "PASS_INCENTIVE", "TEAM_PASSING", "MIXED_INITIATE", "MIXED_DOUBLETAP", "SPEED_AERIAL",
              "PRESSURE_FLICK", "MAWKZY_FLICK", "AIR_CARRY", "ANTI_BALL_STACK"]
REWARD = {n: i for i, n in enumerate(REWARD_IDS)}
MAX_TRAINING_REWARDS = 16


OBS_DEFAULT = 0
OBS_ADVANCED = 1
OBS_VANTA = 2
VANTA_OBS_SIZE = 366


CARS_BLUE_FIRST = 0
CARS_ALTERNATING = 1


assert C.sizeof(CarControls) == 24
assert C.sizeof(TrainingRewardEntry) == 60
assert C.sizeof(CarState) == 248
assert C.sizeof(ArenaInfo) == 32


class Tuning(C.Structure):
    """rsx_tuning (include/rsx.h): launch shape of the arena kernels."""
    _fields_ = [("block_lanes", C.c_int32), ("arenas_per_wave", C.c_int32), ("pool_narrowphase", C.c_int32)]


class TuneResult(C.Structure):
    """rsx_tune_result (include/rsx.h)."""
    _fields_ = [("tuning", Tuning), ("ms_per_env_step", C.c_float), ("feasible", C.c_int32)]
