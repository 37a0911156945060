"""Batched soccar arena engine for AMD Instinct MI355X (gfx950).

The physics of every arena runs in hand-written HIP kernels (librsx.so, C-ABI
in include/rsx.h); this package is the host-side mirror of the reference's
RocketSimCudaBatch interface over that library."""
from .batch import (LIB_PATH, MESH_DIR, EXPORTED, RocketSimBatch, RsxError, default_action_lut,  # noqa: F401
                    load_library)
from .records import (CARS_ALTERNATING, CARS_BLUE_FIRST, OBS_ADVANCED, OBS_DEFAULT, OBS_VANTA,  # noqa: F401
                      ArenaInfo, BallState, BoostPadState, CarControls, CarState, EnvConfig, NUM_PADS, RotMat, Vec3, PadStates,
                      REWARD, REWARD_IDS, TrainingRewardEntry, VANTA_OBS_SIZE, VantaConfig)
