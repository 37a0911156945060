/*
 * rsx.h — C-ABI of the batched arena engine (librsx.so, gfx950).
 *
 * One engine holds N independent soccar arenas (1v1, 2v2 or 3v3 each) in GPU
 * memory and steps them with hand-written HIP kernels.  Every call returns an
 * rsx_status; rsx_last_error() gives the message of the last failure on the
 * calling thread.  One engine must be used from one host thread at a time.
 *
 * Each entry point names the reference interface it replaces
 * (RocketSimCuda/include/RocketSimCuda.h, class rsc::RocketSimCudaBatch,
 * lines 258-455).  The records are in rsx_types.h.
 *
 * Car order: arena-major; arena a holds 2*T cars (T = its team size) in the
 * slot order rsx_config.car_order names.  "All cars" arrays therefore have
 * sum_a 2*T_a entries, with no padding between arenas.
 */
#ifndef RSX_H
#define RSX_H

#include <stddef.h>
#include <stdint.h>

#include "rsx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rsx_engine rsx_engine;
typedef int32_t rsx_status;

#define RSX_OK 0
#define RSX_INVALID_ARG 1
#define RSX_OOM 2
#define RSX_HIP_ERROR 3
#define RSX_NOT_INITIALIZED 4
#define RSX_MESH_LOAD_FAILED 5

/* Engine configuration; replaces BatchConfig (RocketSimCuda.h:171-188). */
typedef struct rsx_config {
    int32_t num_arenas;          /* N > 0 */
    const uint8_t* team_size;    /* per-arena cars per team (1..3); NULL -> default_team_size */
    int32_t default_team_size;   /* used when team_size is NULL */
    const char* mesh_dir;        /* folder holding mesh_0.cmf .. mesh_15.cmf */
    int32_t tick_skip;           /* physics ticks per env step (8) */
    float no_touch_timeout_s;    /* env done after this long without a touch (0 = off) */
    uint32_t seed;               /* RNG seed for kickoffs and demo respawns */
    int32_t device;              /* HIP device ordinal */
    int32_t car_order;           /* RSX_CARS_BLUE_FIRST or RSX_CARS_ALTERNATING */
} rsx_config;

/* Team of each car slot.  BLUE_FIRST: slots 0..T-1 blue, T..2T-1 orange.
 * ALTERNATING: blue, orange, blue, orange, ... (the AddCar order of
 * CudaEnvSet.cpp:216-221). */
#define RSX_CARS_BLUE_FIRST 0
#define RSX_CARS_ALTERNATING 1

const char* rsx_last_error(void);

/* RocketSimCudaBatch::Init + AddCar x (2*T per arena) + ResetAllArenas
 * (RocketSimCuda.h:264, 272, 278; RocketSimCudaApi.cuh:18-201).  Arenas start
 * at a kickoff; the default discrete action table (RLGym DefaultAction, 90
 * entries) is loaded. */
rsx_status rsx_create(const rsx_config* cfg, rsx_engine** out);
/* RocketSimCudaBatch::Destroy (RocketSimCuda.h:267) */
rsx_status rsx_destroy(rsx_engine* e);

rsx_status rsx_num_arenas(const rsx_engine* e, int32_t* out);
/* Total cars over all arenas (the length of every "all cars" array). */
rsx_status rsx_num_cars(const rsx_engine* e, int32_t* out);
rsx_status rsx_team_size(const rsx_engine* e, int32_t arena, int32_t* out);

/* Per-arena state access; replace GetCarState (RocketSimCuda.h:350),
 * GetBallState (356), GetBoostPadStates (362), GetArenaInfo (366),
 * SetCarState (375), SetBallState (376), SetBoostPadStates (378). */
rsx_status rsx_get_car_state(rsx_engine* e, int32_t arena, int32_t car, rsx_car_state* out);
rsx_status rsx_set_car_state(rsx_engine* e, int32_t arena, int32_t car, const rsx_car_state* in);
rsx_status rsx_get_ball_state(rsx_engine* e, int32_t arena, rsx_ball_state* out);
rsx_status rsx_set_ball_state(rsx_engine* e, int32_t arena, const rsx_ball_state* in);
rsx_status rsx_get_pad_states(rsx_engine* e, int32_t arena, rsx_pad_state* out34);
rsx_status rsx_set_pad_states(rsx_engine* e, int32_t arena, const rsx_pad_state* in34);
rsx_status rsx_get_arena_info(rsx_engine* e, int32_t arena, rsx_arena_info* out);
rsx_status rsx_set_tick_count(rsx_engine* e, int32_t arena, uint64_t tick);

/* Bulk state access (synchronous, whole batch); replace GetAllCarStates
 * (RocketSimCuda.h:353), GetAllBallStates (359), GetAllBoostPadStates (363),
 * GetAllArenaInfos (367), SetAllBoostPadStates (380).  Counts: cars =
 * rsx_num_cars, balls and infos = N, pads = 34*N. */
rsx_status rsx_get_all_car_states(rsx_engine* e, rsx_car_state* out, int32_t count);
rsx_status rsx_set_all_car_states(rsx_engine* e, const rsx_car_state* in, int32_t count);
rsx_status rsx_get_all_ball_states(rsx_engine* e, rsx_ball_state* out, int32_t count);
rsx_status rsx_set_all_ball_states(rsx_engine* e, const rsx_ball_state* in, int32_t count);
rsx_status rsx_get_all_pad_states(rsx_engine* e, rsx_pad_state* out, int32_t count);
rsx_status rsx_set_all_pad_states(rsx_engine* e, const rsx_pad_state* in, int32_t count);
rsx_status rsx_get_all_arena_infos(rsx_engine* e, rsx_arena_info* out, int32_t count);

/* Reset of one arena / of every arena (RocketSimCuda.h:275, 278;
 * RocketSimCudaApi.cuh:203-330): with the default (kickoff-only) curriculum,
 * tick count 0, goal flag cleared, cars at a random kickoff permutation drawn
 * from the arena's own RNG stream, ball at rest at centre, every pad active;
 * otherwise the reset rsx_set_reset_curriculum draws. */
rsx_status rsx_reset_arena(rsx_engine* e, int32_t arena);
rsx_status rsx_reset_all_arenas(rsx_engine* e, void* stream);

/* Controls of every car (host array, rsx_num_cars records); replaces
 * SetAllCarControls (RocketSimCuda.h:291). */
rsx_status rsx_set_all_controls(rsx_engine* e, const rsx_controls* controls, int32_t count);
/* SetCarControls (RocketSimCuda.h:288): one car's controls (car index within
 * its arena). */
rsx_status rsx_set_car_controls(rsx_engine* e, int32_t arena, int32_t car, const rsx_controls* controls);

/* Discrete actions.  rsx_upload_action_lut replaces UploadDiscreteActionLut
 * (RocketSimCuda.h:294); rsx_set_actions_from_indices replaces
 * SetDiscreteActionsFromIndices (298, host int32 array); rsx_set_actions_device
 * replaces SetDiscreteActionsFromDevice (301) and is asynchronous on `stream`.
 * Indices are clamped into [0, lut size). */
rsx_status rsx_upload_action_lut(rsx_engine* e, const rsx_controls* lut, int32_t count);
rsx_status rsx_set_actions_from_indices(rsx_engine* e, const int32_t* indices, int32_t count);
rsx_status rsx_set_actions_device(rsx_engine* e, const int32_t* device_indices, int32_t count, void* stream);
/* The default table (RLGym DefaultAction, DefaultAction.cpp:3-89), 90 entries. */
rsx_status rsx_default_action_lut(rsx_controls* out, int32_t capacity, int32_t* count);

/* Env half of a step (the RLGym DefaultObs + stub rewards + terminals the
 * reference builds with BuildAdvancedObsAndDefaultMasks /
 * BuildRewardsAndTerminals, RocketSimCuda.h:316, 338).  Rewards are the
 * weighted sum, in this order, of GoalReward(concede_scale), TouchBallReward,
 * VelocityBallToGoalReward and VelocityPlayerToBallReward
 * (RLGymCPP RewardCore/CommonRewards.h:33-99).  Terminals per arena:
 * 0 none, 1 normal (GoalScoreCondition), 2 truncated (NoTouchCondition),
 * normal taking priority (EnvSet.cpp:178-200). */
typedef struct rsx_env_config {
    float goal_weight;                 /* default 1 */
    float concede_scale;               /* default -1 */
    float touch_weight;                /* default 1 */
    float vel_ball_to_goal_weight;     /* default 1 */
    float vel_player_to_ball_weight;   /* default 1 */
    float no_touch_timeout_s;          /* NoTouchCondition(maxTime); <= 0 off; default rsx_config's */
    int32_t goal_terminal;             /* GoalScoreCondition on; default 1 */
    int32_t auto_reset;                /* kickoff reset of terminal arenas inside rsx_step_env; default 1 */
    int32_t obs_builder;               /* RSX_OBS_DEFAULT (default), RSX_OBS_ADVANCED or RSX_OBS_VANTA */
    int32_t no_touch_stagger;          /* a reset starts the no-touch timer at a uniform draw in
                                          [0, 0.9) x the timeout (ResetArena, RocketSimCudaApi.cuh:396-406);
                                          default 0: at 0, as NoTouchCondition::Reset */
} rsx_env_config;

/* Observation builders.  RSX_OBS_DEFAULT: RLGym DefaultObs (DefaultObs.cpp:
 * 4-54), 51 + 19 * cars floats per car, the previous action being the step's
 * action.  RSX_OBS_ADVANCED: the AdvancedObs rows of the reference's GPU
 * training path (BuildAdvancedObsAndDefaultMasks, RocketSimCuda.h:316;
 * TrainingObs.cuh:20-127), 51 + 29 * cars floats per car (109 / 167 / 225),
 * the previous action being the car's last controls.  RSX_OBS_VANTA: the
 * VantaObs v3.5 rows of the reference's GPU path (ObsMode::VANTA,
 * RocketSimCuda.h:168; BuildVantaObs, 318-321; VantaObsGpu.cuh:342-747),
 * 366 floats per car whatever the team size, with per-arena pad timers,
 * per-car coyote timers and optional ball-prediction slices (see
 * rsx_configure_vanta).  rsx_step_env and rsx_build_obs then run the Vanta
 * kernel after the step; an arena's Vanta state is cleared whenever it is
 * reset (rsx_reset_arena, rsx_reset_all_arenas, the env step's auto-reset),
 * as ResetArena clears it (RocketSimCudaApi.cuh:265-274).  Changing the
 * builder changes rsx_obs_layout and obs_floats (the obs buffer's address
 * stays). */
#define RSX_OBS_DEFAULT 0
#define RSX_OBS_ADVANCED 1
#define RSX_OBS_VANTA 2

/* ConfigureTrainingTerminals (RocketSimCuda.h:330) and the stub rewards. */
rsx_status rsx_configure_env(rsx_engine* e, const rsx_env_config* c);
rsx_status rsx_get_env_config(const rsx_engine* e, rsx_env_config* c);

/* Training rewards of the reference's GPU path: TrainingRewardID
 * (RocketSimCuda.h:200-231), evaluated as training_eval_reward
 * (RocketSimCudaRewards.cuh:36-287) on the state after the env step and the
 * snapshot before it (SnapshotTrainingState, RocketSimCuda.h:333). */
#define RSX_REWARD_UNKNOWN 0
#define RSX_REWARD_GOAL 1                   /* params[0]: the scale of a conceded goal */
#define RSX_REWARD_VELOCITY_BALL_TO_GOAL 2  /* params[0] > 0.5: towards the own goal */
#define RSX_REWARD_VELOCITY_PLAYER_TO_BALL 3
#define RSX_REWARD_FACE_BALL 4
#define RSX_REWARD_TOUCH_BALL 5
#define RSX_REWARD_TOUCH_ACCEL 6
#define RSX_REWARD_STRONG_TOUCH 7           /* params[0]: least hit speed change, params[1]: full-reward change */
#define RSX_REWARD_SPEED 8
#define RSX_REWARD_VELOCITY 9               /* params[0]: scale */
#define RSX_REWARD_AIR 10
#define RSX_REWARD_WAVEDASH 11
#define RSX_REWARD_PICKUP_BOOST 12
#define RSX_REWARD_SAVE_BOOST 13            /* params[0]: exponent */
#define RSX_REWARD_BUMP 14
#define RSX_REWARD_BUMPED_PENALTY 15
#define RSX_REWARD_DEMO 16
#define RSX_REWARD_DEMOED_PENALTY 17
#define RSX_REWARD_KICKOFF_PROXIMITY 18
#define RSX_REWARD_TEAMMATE_BUMP_PENALTY 19
#define RSX_REWARD_PASS_INCENTIVE 20        /* 20-24: reserved team ids, score 0 as in the reference */
#define RSX_REWARD_TEAM_PASSING 21
#define RSX_REWARD_MIXED_INITIATE 22
#define RSX_REWARD_MIXED_DOUBLETAP 23
#define RSX_REWARD_SPEED_AERIAL 24
#define RSX_REWARD_PRESSURE_FLICK 25
#define RSX_REWARD_MAWKZY_FLICK 26
#define RSX_REWARD_AIR_CARRY 27
#define RSX_REWARD_ANTI_BALL_STACK 28
#define RSX_MAX_TRAINING_REWARDS 16
#define RSX_TRAINING_REWARD_PARAMS 10

/* TrainingRewardEntry (RocketSimCuda.h:235-242), same field order and
 * layout.  A zero-sum entry gives each car
 * raw * (1 - team_spirit) + team_mean * team_spirit - opponent_mean * opponent_scale. */
typedef struct rsx_training_reward_entry {
    uint8_t id;
    float weight;
    float params[RSX_TRAINING_REWARD_PARAMS];
    uint8_t is_zero_sum;
    float team_spirit;
    float opponent_scale;
} rsx_training_reward_entry;

/* ConfigureTrainingRewards (RocketSimCuda.h:329): from the next env step on,
 * each car's reward is the table's weighted sum, in table order, instead of
 * the stub rewards of rsx_env_config (count 0 returns to those).  The goal
 * term reads the episode's goal flag (rsx_arena_info.goal_scored / goal_team,
 * set once per episode as in the reference).  Unknown ids: RSX_INVALID_ARG. */
rsx_status rsx_configure_training_rewards(rsx_engine* e, const rsx_training_reward_entry* entries, int32_t count);
rsx_status rsx_get_training_rewards(const rsx_engine* e, rsx_training_reward_entry* out, int32_t capacity,
                                    int32_t* count);

/* SetResetCurriculum (RocketSimCuda.h:283; ResetArena,
 * RocketSimCudaApi.cuh:203-418): relative weights of the four resets every
 * later reset (rsx_reset_arena, rsx_reset_all_arenas, the env step's
 * auto-reset) chooses from: kickoff; fuzzed kickoff (ball within 180 x 220 uu
 * of centre with a mild velocity, cars +-40 uu off their spots); aerial (ball
 * at 420-780 uu height within 1400 x 1600 uu of centre, one car below and
 * behind it in the air, moving at it); ball chase (ball within 1900 x 2200
 * uu, every car 450-1850 uu from it on the ground, facing it).  Weights that
 * are negative, NaN or above 1e6 count as 0; all 0 means kickoff only (the
 * default).  The draws come from each arena's own RNG stream. */
rsx_status rsx_set_reset_curriculum(rsx_engine* e, float kickoff_w, float fuzzed_w, float aerial_w, float chase_w);
rsx_status rsx_get_reset_curriculum(const rsx_engine* e, float* weights4);

/* One fused env step for every arena, asynchronous on `stream`: device int32
 * action indices (rsx_num_cars, clamped into the action table) -> controls,
 * tick_skip physics ticks, rewards and terminals of the step, kickoff reset of
 * terminal arenas (when auto_reset), then each car's DefaultObs row with the
 * step's action as its previous action (zeros after a reset).  Replaces the
 * caller sequence SetDiscreteActionsFromDevice + Step(tickSkip) +
 * BuildAdvancedObsAndDefaultMasks + BuildRewardsAndTerminals +
 * ResetTerminalArenasGpu (CudaEnvSet.cpp:364-376, 182-200). */
rsx_status rsx_step_env(rsx_engine* e, const int32_t* device_actions, void* stream);

/* DefaultObs rows of the current state with zero previous actions (the first
 * observation after create or a reset). */
rsx_status rsx_build_obs(rsx_engine* e, void* stream);

/* Device pointers of the env outputs (GetBuilt*DevicePtr, RocketSimCuda.h:
 * 393-396): obs (obs_floats floats; car c's row at row_offset[c], length
 * row_len[c] = 51 + 19 * cars in its arena: 89 / 127 / 165 with DefaultObs,
 * 51 + 29 * cars: 109 / 167 / 225 with AdvancedObs), rewards
 * (rsx_num_cars floats), terminals (N bytes).  Stable for the engine's life. */
rsx_status rsx_env_buffers(rsx_engine* e, float** obs, float** rewards, uint8_t** terminals, int64_t* obs_floats);
rsx_status rsx_obs_layout(const rsx_engine* e, int64_t* row_offset, int32_t* row_len, int32_t count);

/* Synchronous copies to host; replace CopyBuiltAdvancedObs / CopyBuiltRewards /
 * CopyBuiltTerminals (RocketSimCuda.h:341-344). */
rsx_status rsx_copy_obs(rsx_engine* e, float* out, int64_t count);
rsx_status rsx_copy_rewards(rsx_engine* e, float* out, int32_t count);
rsx_status rsx_copy_terminals(rsx_engine* e, uint8_t* out, int32_t count);

/* DefaultAction masks (RLGym DefaultAction::GetActionMask,
 * DefaultAction.cpp:91-118), the masks half of BuildAdvancedObsAndDefaultMasks
 * (RocketSimCuda.h:316).  Off by default.  When on, rsx_step_env and
 * rsx_build_obs also write, per car, 90 bytes (1 = the action of that index
 * in the default table is allowed) from the state the obs rows describe.
 * Needs the default 90-entry action table.  rsx_action_masks_buffer gives the
 * device buffer (num_cars x 90, car-major, stable once allocated; replaces
 * GetBuiltDefaultActionMasksDevicePtr, RocketSimCuda.h:394);
 * rsx_copy_action_masks replaces CopyBuiltDefaultActionMasks (342). */
rsx_status rsx_set_action_masks(rsx_engine* e, int32_t enabled);
rsx_status rsx_action_masks_buffer(rsx_engine* e, uint8_t** d_masks);
rsx_status rsx_copy_action_masks(rsx_engine* e, uint8_t* out, int64_t count);

/* Step every arena by `ticks` physics ticks; replaces Step
 * (RocketSimCuda.h:310, RocketSimCudaApi.cuh:551-571).  Asynchronous on
 * `stream` (a hipStream_t, NULL = default stream). */
rsx_status rsx_step_ticks(rsx_engine* e, int32_t ticks, void* stream);

/* Wait for all work queued by this engine; replaces Synchronize
 * (RocketSimCuda.h:345). */
rsx_status rsx_synchronize(rsx_engine* e);

/* Launch shape of the arena kernels (rsx_step_env, rsx_step_ticks).  The
 * reference fixes it at 128 threads per block, one thread per arena
 * (RocketSimCudaApi.cuh:553-554).  Here:
 *   block_lanes       lanes per workgroup: 64, 128, 256 or 512 (512: two
 *                     wavefronts per SIMD with at most 256 registers a lane;
 *                     its LDS fits only without pooling; the tick path runs
 *                     it as 256-lane workgroups);
 *   arenas_per_wave   arenas per 64-lane wavefront: 64, 32, 16, 8, 4, 2 or 1
 *                     (below 64 the spare lanes store nothing: with pooled
 *                     queries they skip the tick's own work and take their
 *                     share of the wavefront's pooled work, so that each
 *                     arena's gathers, mesh queries and wheel-ray mesh tests
 *                     are spread over 64 / arenas_per_wave lanes; unpooled
 *                     they repeat a lane's arena);
 *   pool_narrowphase  1: every body's mesh gather and contact queries and
 *                     the wheel rays' mesh tests of all lanes of a wavefront
 *                     are spread over its 64 lanes; 0: each lane runs its
 *                     own.
 * Results do not depend on the shape.  Default {64, 64, 1}. */
typedef struct rsx_tuning {
    int32_t block_lanes;
    int32_t arenas_per_wave;
    int32_t pool_narrowphase;
} rsx_tuning;

typedef struct rsx_tune_result {
    rsx_tuning tuning;
    float ms_per_env_step;  /* mean over the timed env steps; < 0 when infeasible */
    int32_t feasible;       /* 0: the shape's LDS does not fit a workgroup */
} rsx_tune_result;

rsx_status rsx_set_tuning(rsx_engine* e, const rsx_tuning* t);
rsx_status rsx_get_tuning(const rsx_engine* e, rsx_tuning* t);

/* Autotuner: times `env_steps` env steps (after one untimed) of every shape
 * in {64, 128, 256, 512} x ({64, 32, 16, 8, 4, 2, 1} pooled, {64, 32, 16, 8}
 * unpooled) -- 44 shapes -- on the engine's current arenas with the given
 * device actions (shapes whose LDS does not fit a workgroup, or that need
 * more than 8 wavefronts per SIMD for these arenas, are reported infeasible
 * without a trial), keeps the fastest, and restores the arenas, the env
 * outputs and the cull's counters to what they were.  results (nullable,
 * capacity >= 44) receives every shape's time; *count the number written.
 * Blocks the host until done. */
rsx_status rsx_autotune(rsx_engine* e, const int32_t* device_actions, int32_t env_steps, rsx_tune_result* results,
                        int32_t capacity, int32_t* count);

/* VantaObs parameters; replace BatchConfig's vantaPredictBall,
 * vantaOnlyClosestOpp, vantaDodgeDeadzone and vantaTickSkip
 * (RocketSimCuda.h:184-189).  Defaults: no prediction, every opponent,
 * deadzone 0.5, the engine's tick skip. */
typedef struct rsx_vanta_config {
    int32_t predict_ball;      /* ball-prediction slices: the ball's position and velocity after 120, 240 and
                                  360 ticks of a ball-only rollout with the engine's own physics
                                  (vantaArenaUpdateKernel, RocketSimCudaObs.cuh:70-122) */
    int32_t only_closest_opp;  /* fill only the closest opponent's packet */
    float dodge_deadzone;      /* least |pitch|, |yaw| or |roll| a jump needs to count as a flip, in [0, 1] */
    int32_t tick_skip;         /* the coyote / pad timers advance tick_skip / 120 s per update; 0: the engine's */
} rsx_vanta_config;

rsx_status rsx_configure_vanta(rsx_engine* e, const rsx_vanta_config* c);
rsx_status rsx_get_vanta_config(const rsx_engine* e, rsx_vanta_config* c);
/* The ball-prediction slices an arena's last Vanta build used (positions in
 * uu, velocities in uu/s, angular velocities zero) and whether it had them;
 * no counterpart in the reference (its slices live only on the device). */
rsx_status rsx_get_vanta_prediction(rsx_engine* e, int32_t arena, rsx_ball_state* out3, int32_t* has_prediction);

/* hipGraph of the env step: rsx_capture_env_graph records `env_steps`
 * consecutive rsx_step_env calls reading `device_actions` (the caller
 * refreshes that buffer between launches); rsx_launch_env_graph replays them
 * on `stream`.  A change of launch shape, env config, action table or cull
 * after the capture makes the launch fail until the next capture. */
rsx_status rsx_capture_env_graph(rsx_engine* e, const int32_t* device_actions, int32_t env_steps);
rsx_status rsx_launch_env_graph(rsx_engine* e, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RSX_H */
