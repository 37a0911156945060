// rsx_consts.h — game constants of the soccar arena, in the units the engine
// computes in.  Values and names follow RocketSim/src/RLConst.h (constants,
// curves, spawn tables, pad locations), Sim/Car/CarConfig/CarConfig.cpp
// (Octane preset), Sim/MutatorConfig/MutatorConfig.{h,cpp} (soccar defaults),
// and Bullet's defaults that RocketSim relies on (btContactSolverInfo.h:79-112,
// btCollisionMargin.h:22, btPersistentManifold.cpp gContactBreakingThreshold).
#pragma once

#include "rsx_math.h"

namespace rsx {

// Unit conversion (RocketSim/src/BulletLink.h:13-16)
constexpr float kBT2UU = 50.f;
constexpr float kUU2BT = 1.f / 50.f;

// Physics rate
constexpr float kTickRate = 120.f;

// RLConst.h:11-170
constexpr float kGravityZ = -650.f;
constexpr float kArenaExtentX = 4096.f;
constexpr float kArenaExtentY = 5120.f;
constexpr float kArenaHeight = 2048.f;
constexpr float kCarMassBT = 180.f;
constexpr float kBallMassBT = kCarMassBT / 6.f;
constexpr float kCarCollisionFriction = 0.3f;
constexpr float kCarCollisionRestitution = 0.1f;
constexpr float kCarBallFriction = 2.0f;
constexpr float kCarBallRestitution = 0.0f;
constexpr float kCarWorldFriction = 0.3f;
constexpr float kCarWorldRestitution = 0.3f;
constexpr float kCarCarFriction = 0.09f;
constexpr float kCarCarRestitution = 0.1f;
constexpr float kBallRestZ = 93.15f;
constexpr float kBallMaxAngSpeed = 6.f;
constexpr float kBallDrag = 0.03f;
constexpr float kBallFriction = 0.35f;
constexpr float kBallRestitution = 0.6f;
constexpr float kCarMaxSpeed = 2300.f;
constexpr float kBallMaxSpeed = 6000.f;
constexpr float kBoostMax = 100.f;
constexpr float kBoostUsedPerSecond = kBoostMax / 3;
constexpr float kBoostMinTime = 0.1f;
constexpr float kBoostAccelGround = 2975 / 3.f;
constexpr float kBoostAccelAir = 3175 / 3.f;
constexpr float kBoostSpawnAmount = kBoostMax / 3;
constexpr float kCarMaxAngSpeed = 5.5f;
constexpr float kSupersonicStartSpeed = 2200.f;
constexpr float kSupersonicMaintainMinSpeed = kSupersonicStartSpeed - 100.f;
constexpr float kSupersonicMaintainMaxTime = 1.f;
constexpr float kPowerslideRiseRate = 5;
constexpr float kPowerslideFallRate = 2;
constexpr float kThrottleTorqueAmount = kCarMassBT * 400.f;
constexpr float kBrakeTorqueAmount = kCarMassBT * (14.25f + (1.f / 3.f));
constexpr float kStoppingForwardVel = 25.f;
constexpr float kCoastingBrakeFactor = 0.15f;
constexpr float kBrakingNoThrottleSpeedThresh = 0.01f;
constexpr float kThrottleDeadzone = 0.001f;
constexpr float kThrottleAirAccel = 200 / 3.f;
constexpr float kJumpAccel = 4375.f / 3.f;
constexpr float kJumpImmediateForce = 875.f / 3.f;
constexpr float kJumpMinTime = 0.025f;
constexpr float kJumpResetTimePad = (1 / 40.f);
constexpr float kJumpMaxTime = 0.2f;
constexpr float kDoubleJumpMaxDelay = 1.25f;
constexpr float kFlipZDamp120 = 0.35f;
constexpr float kFlipZDampStart = 0.15f;
constexpr float kFlipZDampEnd = 0.21f;
constexpr float kFlipTorqueTime = 0.65f;
constexpr float kFlipPitchlockExtraTime = 0.3f;
constexpr float kFlipInitialVelScale = 500.f;
constexpr float kFlipTorqueX = 260.f;
constexpr float kFlipTorqueY = 224.f;
constexpr float kFlipForwardImpulseMaxSpeedScale = 1.f;
constexpr float kFlipSideImpulseMaxSpeedScale = 1.9f;
constexpr float kFlipBackwardImpulseMaxSpeedScale = 2.5f;
constexpr float kFlipBackwardImpulseScaleX = 16.f / 15.f;
constexpr float kBallRadius = 91.25f;
constexpr float kGoalThresholdY = 5124.25f;
// CAR_TORQUE_SCALE = 2 * M_PI / (1 << 16) * 1000 (double expression rounded to float)
constexpr float kCarTorqueScale = (float)(2 * 3.14159265358979323846 / (1 << 16) * 1000);
constexpr float kCarAutoflipImpulse = 200;
constexpr float kCarAutoflipTorque = 50;
constexpr float kCarAutoflipTime = 0.4f;
constexpr float kCarAutoflipNormzThresh = (float)0.70710678118654752440;  // M_SQRT1_2
constexpr float kCarAutoflipRollThresh = 2.8f;
constexpr float kCarAutorollForce = 100;
constexpr float kCarAutorollTorque = 80;
constexpr float kBallCarExtraImpulseZScale = 0.35f;
constexpr float kBallCarExtraImpulseForwardScale = 0.65f;
constexpr float kBallCarExtraImpulseMaxDeltaVel = 4600.f;
constexpr float kCarSpawnRestZ = 17.f;
constexpr float kCarRespawnZ = 36.f;
constexpr float kBumpCooldownTime = 0.25f;
constexpr float kBumpMinForwardDist = 64.5f;
constexpr float kDemoRespawnTime = 3.f;

// RLConst.h BTVehicle
constexpr float kSuspForceScaleFront = 36.f - (1.f / 4.f);
constexpr float kSuspForceScaleBack = 54.f + (1.f / 4.f) + (1.5f / 100.f);
constexpr float kSuspStiffness = 500.f;
constexpr float kWheelsDampingCompression = 25.f;
constexpr float kWheelsDampingRelaxation = 40.f;
constexpr float kMaxSuspensionTravel = 12.f;
constexpr float kSuspensionSubtraction = 0.05f;

// RLConst.h air control (PYR order)
constexpr float kAirTorqueP = 130, kAirTorqueY = 95, kAirTorqueR = 400;
constexpr float kAirDampP = 30, kAirDampY = 20, kAirDampR = 50;

// Boost pads (RLConst.h BoostPads)
constexpr float kPadCylHeight = 95;
constexpr float kPadCylRadBig = 208;
constexpr float kPadCylRadSmall = 144;
constexpr float kPadBoxHeight = 64;
constexpr float kPadBoxRadBig = 160;
constexpr float kPadBoxRadSmall = 120;
constexpr float kPadCooldownBig = 10;
constexpr float kPadCooldownSmall = 4;
constexpr float kPadAmountBig = 100;
constexpr float kPadAmountSmall = 12;
constexpr int kNumPads = 34;
constexpr int kNumBigPads = 6;

// Bullet defaults that matter
constexpr float kConvexMargin = 0.04f;             // CONVEX_DISTANCE_MARGIN
constexpr float kContactBreakingThreshold = 0.02f;  // gContactBreakingThreshold
constexpr int kSolverIterations = 10;
constexpr float kErp = 0.2f;           // m_erp (used by resolveSingleCollision)
constexpr float kErp2 = 0.8f;          // RocketSim sets solverInfo.m_erp2 = 0.8
constexpr float kWarmstartFactor = 0.85f;
constexpr float kSplitImpulseTurnErp = 0.1f;
constexpr float kRestitutionVelThreshold = 0.2f;
constexpr float kLinearSleepThreshold = 0.8f;
constexpr float kAngularSleepThreshold = 1.0f;

// Arena static planes (Arena.cpp:1060-1101): floor, ceiling, -X wall, +X wall.
// normal, constant 0, origin (BT units).
constexpr int kNumPlanes = 4;

// Suspension collision grid (SuspensionCollisionGrid.h:10-30, heavy variant)
constexpr int kSuspCellsX = 32, kSuspCellsY = 56, kSuspCellsZ = 8;
constexpr float kSuspExtentX = kArenaExtentX;
constexpr float kSuspExtentY = kArenaExtentY * 1.5f;
constexpr float kSuspHeight = kArenaHeight;
constexpr float kSuspCellSizeX = kSuspExtentX / (kSuspCellsX / 2);
constexpr float kSuspCellSizeY = kSuspExtentY / (kSuspCellsY / 2);
constexpr float kSuspCellSizeZ = kSuspHeight / kSuspCellsZ;

// RocketSim broadphase grid (ArenaConfig.h:19-30, btRSBroadphase.cpp:43-70)
constexpr float kBpMinX = -4500 * kUU2BT, kBpMinY = -6000 * kUU2BT, kBpMinZ = 0;
constexpr float kBpMaxX = 4500 * kUU2BT, kBpMaxY = 6000 * kUU2BT, kBpMaxZ = 2500 * kUU2BT;
constexpr float kBpCellSize = 370 * kUU2BT;

// Boost pad grid (BoostPadGrid.h:10-22)
constexpr float kPadGridExtentZ = kPadCylHeight + 250.f;
constexpr int kPadCellsX = 8, kPadCellsY = 10;
constexpr int kPadCellSizeX = (int)(4096.f / (kPadCellsX / 2));
constexpr int kPadCellSizeY = (int)(5120.f / (kPadCellsY / 2));

// Game mode kinds of arena-world statics: meshes 0..15, then the 4 planes.
constexpr int kNumMeshes = 16;

}  // namespace rsx
