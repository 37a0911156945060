// rsx_world.h — per-arena working state and the read-only world data the
// per-tick code uses.  Everything here is plain data so the same structs live
// in a GPU lane's private memory and in host memory for the CPU build.
//
// Semantics of each field follow the CPU RocketSim + Bullet objects they stand
// for; the citations name the reference field.
#pragma once

#include "rsx_consts.h"

namespace rsx {

// ---------------------------------------------------------------------------
// Read-only geometry computed once at engine creation (host) with Bullet's own
// float expression order, then passed by value to the kernels.
// ---------------------------------------------------------------------------
struct CarGeom {
    // btBoxShape after setSafeMargin (btBoxShape.cpp:17-26): half extents without margin
    V3 box_implicit;
    float box_margin;
    V3 box_offset;            // child transform origin (Car.cpp:200-203)
    V3 compound_half;         // btCompoundShape local AABB half extents (btCompoundShape.cpp:159-179)
    V3 compound_center;
    V3 inv_inertia_local;     // 1/I of the box (btBoxShape.cpp:33-45)
    float inv_mass;
    float breaking;           // contact breaking threshold of the compound (angular motion disc * 0.02)
    V3 wheel_cs[4];           // chassis connection points (Car.cpp:243-264)
    float wheel_radius[4];
    float rest_len[4];        // m_suspensionRestLength1
    float susp_force_scale[4];
};

struct BallGeom {
    float radius;             // btSphereShape radius (BT)
    float inv_mass;
    float inv_inertia;        // 1/(0.4 m r^2)
    float breaking;           // contact breaking threshold of the sphere
};

struct PadGeom {
    V3 pos_bt;                // BoostPad::_posBT
    V3 box_min, box_max;      // BoostPad::_boxMinBT/_boxMaxBT (BoostPad.cpp:37-49)
    float cyl_rad_bt;
    int big;
    V3 pos_uu;
};

// One triangle's entry in a uniform-grid cell, with what the candidate
// gather tests so that a cell is scanned with one contiguous read per entry:
// the packed candidate key (mesh, tree rank, index), the grid cell of the
// triangle's AABB minimum (x | y << 10 | z << 20, clamped to the grid) and
// the AABB itself (BT units).
struct GridEntry {
    int key;
    int first;
    float mn[3];
    float mx[3];
};

// Triangle mesh of the arena plus the spatial indexes used by the queries.
struct MeshView {
    const float* tri;          // [num_tris][9] vertices, BT units
    const int* tri_mesh;       // mesh file index of each triangle (0..15)
    const int* tri_rank;       // traversal rank of the triangle inside its mesh's tree
    const int* cell_start;     // uniform grid CSR: [num_cells + 1]
    const int* cell_tris;      // triangle ids
    const GridEntry* cell_ent; // the same lists as gather entries
    const int* blk_start;      // per grid cell c: CSR of the 2 x 2 x 2 block of cells whose lowest corner is c
    const GridEntry* blk_ent;  // the block's distinct triangles, sorted by packed key
    const uint8_t* susp_world; // suspension grid cells with world geometry (bled), 32*56*8
    const uint16_t* bp_meshes; // broadphase grid: per cell, mask of meshes listed in the cell
    const float* mesh_aabb;    // [16][6] shape AABB min xyz, max xyz (BT)
    const float* tri_info;     // [num_tris][4] internal-edge angles v0v1, v1v2, v2v0 and flag bits
    V3 grid_min;               // uniform grid origin (BT)
    float grid_cell;           // cell size (BT)
    int gx, gy, gz;
    int num_tris;
    int bp_nx, bp_ny, bp_nz;
};

// ---------------------------------------------------------------------------
// Per-arena working state (lives in a lane's private memory during a kernel).
// ---------------------------------------------------------------------------
struct Controls {
    float throttle, steer, pitch, yaw, roll;
    bool jump, boost, handbrake;
};

struct Wheel {
    // carried from tick to tick (btWheelInfoRL fields set by Car::_UpdateWheels)
    float steer_angle, engine_force, brake, lat_friction, long_friction, extra_pushback;
    // recomputed every tick (btWheelInfo::RaycastInfo and friends)
    V3 hardpoint, dir_ws, axle_ws, contact_point, contact_normal;
    M3 wbasis;                 // wheel world basis (steering applied)
    V3 impulse;                // friction impulse (btVehicleRL::calcFrictionImpulses)
    float susp_len, susp_rel_vel, clipped_inv_dot, susp_force;
    int ground;                // -1 none, 0 static world, 1 + b for dynamic body b
    bool in_contact, in_contact_world;
};

struct Car {
    // rigid body (Bullet units)
    Xf xf;
    Xf interp;                 // m_interpolationWorldTransform (predicted)
    V3 vel, ang;
    M3 inv_inertia_world;
    V3 force, torque;          // m_totalForce / m_totalTorque
    V3 vel_cache;              // Car::_velocityImpulseCache
    Wheel wheel[4];
    Controls controls;
    // CarState fields (Car.h:14-115)
    bool on_ground, has_jumped, has_double_jumped, has_flipped, is_flipping, is_jumping;
    bool is_supersonic, is_auto_flipping, is_demoed, world_contact, ball_hit_valid;
    bool wheel_contact[4];
    float boost, time_spent_boosting, jump_time, flip_time, air_time, air_time_since_jump;
    float supersonic_time, handbrake_val, auto_flip_timer, auto_flip_torque_scale;
    float demo_respawn_timer, car_contact_cooldown;
    V3 flip_rel_torque, world_contact_normal;
    uint32_t id, car_contact_other_id;
    int team;
    uint64_t ball_hit_tick, ball_hit_extra_tick;
    V3 ball_hit_rel_pos, ball_hit_ball_pos, ball_hit_extra_vel;
    Controls last_controls;
    // The body responds to contacts (Bullet: no CF_NO_CONTACT_RESPONSE): set by
    // the car's own pre-tick (false while demolished, Car.cpp:68-82) and kept
    // until its next one, so the wheel rays of cars stepped before it in a
    // tick see last tick's value.  Stored with the car; true for a new car.
    bool active;
    // Broadphase registration (btRSBroadphase): a dynamic body is listed in the
    // 27 cells around the cell of its AABB minimum as of its last AABB update
    // (the previous world step, or when it was added), and a suspension ray
    // only meets bodies listed in the cell it starts in.  Setting a state
    // does not move the registration.  Packed cell index; -1: listed nowhere.
    int bp_reg;
};

struct Ball {
    Xf xf;                     // basis stays identity (noBallRot, ArenaConfig.h:33)
    Xf interp;
    V3 vel, ang;
    V3 force;
    V3 vel_cache;
    bool sleeping;
    int bp_reg;                // broadphase registration, as for Car::bp_reg
};

struct Pad {
    bool is_active;
    float cooldown;
    uint32_t prev_locked, cur_locked;   // car ids, 0 = none
};

// One contact point of a manifold (btManifoldPoint subset, btManifoldPoint.h).
struct Contact {
    V3 posA, posB, normal;     // normal on B, world
    V3 localA, localB;
    float dist;
    float friction, restitution;
    bool special;
};

constexpr int kMaxManifoldPoints = 4;

// Persistent-manifold record for one body pair; body codes: 0 ball, 1+i car i,
// 100+s static s (meshes 0..15, then planes 16..19).
struct Manifold {
    int body0, body1;
    int n;
    float breaking;
    Contact pt[kMaxManifoldPoints];
};

// Full working state of one arena.  MAXC is the car capacity of the arena
// (2, 4 or 6 for 1v1, 2v2, 3v3); ncars is the number in use.
template <int MAXC>
struct Arena {
    int ncars;
    uint64_t tick;
    Car car[MAXC > 0 ? MAXC : 1];  // MAXC = 0: the ball alone (one unused slot)
    Ball ball;
    Pad pad[kNumPads];
    uint32_t rng;              // per-arena RNG state (kickoff / demo respawn draws)
};

}  // namespace rsx
