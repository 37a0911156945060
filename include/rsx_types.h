/*
 * rsx_types.h — plain-C state records of the batched arena engine.
 *
 * These are the public per-object records that cross the C-ABI (include/rsx.h).
 * They carry the same fields, units and meaning as the reference's public structs
 * (RocketSimCuda/include/RocketSimCuda.h:15-190: Vec3, RotMat, CarControls,
 * CarState, BallState, BoostPadState, ArenaInfo) and the CPU RocketSim's
 * CarState / BallState / BoostPadState (RocketSim/src/Sim/Car/Car.h:14-115,
 * Sim/Ball/Ball.h:14-40, Sim/BoostPad/BoostPad.h:23-35), with explicit integer
 * widths so the layout is identical for gcc, hipcc and ctypes.
 *
 * Units: positions in uu, velocities in uu/s, angular velocities in rad/s,
 * times in seconds.  Rotation matrices are stored by axis vector
 * (forward, right, up), each a world-space unit vector.
 */
#ifndef RSX_TYPES_H
#define RSX_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSX_NUM_PADS 34
#define RSX_NUM_ACTIONS 90
#define RSX_MAX_TEAM_SIZE 3
#define RSX_MAX_CARS 6

typedef struct rsx_vec3 {
    float x, y, z;
} rsx_vec3;

typedef struct rsx_rotmat {
    rsx_vec3 forward, right, up;
} rsx_rotmat;

/* RocketSimCuda.h:49-58, RocketSim Sim/CarControls.h:7-20 */
typedef struct rsx_controls {
    float throttle, steer, pitch, yaw, roll;
    uint8_t jump, boost, handbrake, _pad;
} rsx_controls; /* 24 bytes */

/* Full car record; RocketSim Car.h:14-115 (CarState) + BallHitInfo.h:7-25 */
typedef struct rsx_car_state {
    rsx_vec3 pos;
    rsx_rotmat rot;
    rsx_vec3 vel;
    rsx_vec3 ang_vel;

    float boost;
    float time_spent_boosting;
    float jump_time;
    float flip_time;
    float air_time;
    float air_time_since_jump;
    float supersonic_time;
    float handbrake_val;
    float auto_flip_timer;
    float auto_flip_torque_scale;
    float demo_respawn_timer;
    float car_contact_cooldown;

    rsx_vec3 flip_rel_torque;
    rsx_vec3 world_contact_normal;

    uint8_t is_on_ground;
    uint8_t has_jumped;
    uint8_t has_double_jumped;
    uint8_t has_flipped;
    uint8_t is_flipping;
    uint8_t is_jumping;
    uint8_t is_supersonic;
    uint8_t is_auto_flipping;
    uint8_t is_demoed;
    uint8_t world_contact;
    uint8_t ball_hit_valid;
    uint8_t team;           /* 0 blue, 1 orange */
    uint8_t wheels_with_contact[4];

    uint32_t id;            /* 1-based car id inside its arena */
    uint32_t car_contact_other_id;

    uint64_t ball_hit_tick;              /* tickCountWhenHit, ~0 if never */
    uint64_t ball_hit_extra_impulse_tick;/* tickCountWhenExtraImpulseApplied */
    rsx_vec3 ball_hit_rel_pos;
    rsx_vec3 ball_hit_ball_pos;
    rsx_vec3 ball_hit_extra_vel;

    rsx_controls last_controls;
} rsx_car_state;

/* RocketSim Ball.h:14-40 (noBallRot: rotation is not simulated) */
typedef struct rsx_ball_state {
    rsx_vec3 pos;
    rsx_vec3 vel;
    rsx_vec3 ang_vel;
} rsx_ball_state;

/* RocketSim BoostPad.h:23-35 */
typedef struct rsx_pad_state {
    uint8_t is_active;
    uint8_t _pad[3];
    float cooldown;
    uint32_t prev_locked_car_id;
} rsx_pad_state;

/* RocketSimCuda.h:157-163 */
typedef struct rsx_arena_info {
    uint64_t tick_count;
    int32_t num_cars;
    int32_t team_size;
    uint8_t goal_scored;    /* set once the ball is past a goal line (Arena::IsBallScored); cleared by reset */
    uint8_t contact_overflow; /* no counterpart in the reference: set when a tick had more contact manifolds than
                                 the engine stores (their contacts were dropped); cleared by rsx_reset_arena /
                                 rsx_reset_all_arenas, kept across the env step's auto-reset */
    uint8_t _pad[2];
    int32_t goal_team;      /* -1 none; 0 = blue scored (ball at +y), 1 = orange scored */
    float no_touch_time;    /* seconds since any car touched the ball */
    uint32_t episode;       /* number of resets this arena has seen */
} rsx_arena_info;

#ifdef __cplusplus
}
#endif

#endif /* RSX_TYPES_H */
