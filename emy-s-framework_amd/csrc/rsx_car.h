// rsx_car.h — per-car work of one tick before and after the world step:
// the raycast vehicle (suspension rays, friction impulses, suspension forces),
// throttle/brake/steer/sticky forces, jump, air control, flips, auto-flip,
// auto-roll, boost, supersonic state and velocity limits.
//
// Reference (CPU RocketSim, followed step by step):
//   Car::_PreTickUpdate            Sim/Car/Car.cpp:58-131
//   Car::_UpdateWheels             Car.cpp:330-475
//   Car::_UpdateBoost              Car.cpp:477-505
//   Car::_UpdateJump               Car.cpp:507-554
//   Car::_UpdateAirTorque          Car.cpp:556-641
//   Car::_UpdateDoubleJumpOrFlip   Car.cpp:643-761
//   Car::_UpdateAutoFlip           Car.cpp:763-797
//   Car::_UpdateAutoRoll           Car.cpp:799-833
//   Car::_PostTickUpdate           Car.cpp:133-163
//   Car::_FinishPhysicsTick        Car.cpp:165-193
//   btVehicleRL (raycast vehicle)  Sim/btVehicleRL/btVehicleRL.cpp:64-417
//   resolveSingleCollision / resolveSingleBilateral
//                                  bullet3-3.24/BulletDynamics/ConstraintSolver/btContactConstraint.cpp
#pragma once

#include "rsx_gjk.h"
#include "rsx_query.h"

namespace rsx {

// LinearPieceCurve tables (RLConst.h:318-417)
RSX_HD Curve curve_steer() { return Curve{6, {0, 500, 1000, 1500, 1750, 3000}, {0.53356f, 0.31930f, 0.18203f, 0.10570f, 0.08507f, 0.03454f}}; }
RSX_HD Curve curve_powerslide_steer() { return Curve{2, {0, 2500}, {0.39235f, 0.12610f}}; }
RSX_HD Curve curve_drive_torque() { return Curve{3, {0, 1400, 1410}, {1.0f, 0.1f, 0.0f}}; }
RSX_HD Curve curve_non_sticky() { return Curve{3, {0, 0.7075f, 1}, {0.1f, 0.5f, 1.0f}}; }
RSX_HD Curve curve_lat_friction() { return Curve{2, {0, 1}, {1.0f, 0.2f}}; }
RSX_HD Curve curve_long_friction() { return Curve{0, {}, {}}; }
RSX_HD Curve curve_hb_lat() { return Curve{1, {0}, {0.1f}}; }
RSX_HD Curve curve_hb_long() { return Curve{2, {0, 1}, {0.5f, 0.9f}}; }
RSX_HD Curve curve_ball_car_extra() { return Curve{4, {0, 500.f, 2300.f, 4600.f}, {0.65f, 0.65f, 0.55f, 0.30f}}; }
RSX_HD Curve curve_bump_ground() { return Curve{3, {0.f, 1400.f, 2200.f}, {(5.f / 6.f), 1100.f, 1530.f}}; }
RSX_HD Curve curve_bump_air() { return Curve{3, {0.f, 1400.f, 2200.f}, {(5.f / 6.f), 1390.f, 1945.f}}; }
RSX_HD Curve curve_bump_up() { return Curve{3, {0.f, 1400.f, 2200.f}, {(2.f / 6.f), 278.f, 417.f}}; }

// --- rigid-body helpers (btRigidBody.h:277-352) ---------------------------
RSX_HD void car_apply_central_impulse(Car& c, const CarGeom& g, V3 imp) { c.vel += imp * g.inv_mass; }
RSX_HD void car_apply_impulse(Car& c, const CarGeom& g, V3 imp, V3 rel) {
    c.vel += imp * g.inv_mass;
    c.ang += c.inv_inertia_world * cross(rel, imp);
}
RSX_HD V3 car_vel_at(const Car& c, V3 rel) { return c.vel + cross(c.ang, rel); }
RSX_HD V3 car_fwd(const Car& c) { return col(c.xf.b, 0); }
RSX_HD V3 car_right(const Car& c) { return col(c.xf.b, 1); }
RSX_HD V3 car_up(const Car& c) { return col(c.xf.b, 2); }
RSX_HD float car_mass(const CarGeom& g) { return 1.0f / g.inv_mass; }

// btRigidBody::updateInertiaTensor
RSX_HD void car_update_inertia(Car& c, const CarGeom& g) {
    c.inv_inertia_world = scaled(c.xf.b, g.inv_inertia_local) * transpose(c.xf.b);
}

// --- raycast vehicle ---------------------------------------------------------
RSX_HD void wheel_update_ws(Car& c, const CarGeom& g, int i) {
    Wheel& w = c.wheel[i];
    w.in_contact = false;
    w.in_contact_world = false;
    w.hardpoint = xform(c.xf, g.wheel_cs[i]);
    w.dir_ws = c.xf.b * V3{0, 0, -1};
    w.axle_ws = c.xf.b * V3{0, -1, 0};
}

// btVehicleRL::updateWheelTransform (btVehicleRL.cpp:64-92)
RSX_HD void wheel_update_transform(Car& c, const CarGeom& g, int i) {
    wheel_update_ws(c, g, i);
    Wheel& w = c.wheel[i];
    V3 up = -w.dir_ws;
    V3 right = w.axle_ws;
    V3 fwd = normalized(cross(up, right));
    M3 steer = m3_from_quat(q_axis_angle(up, w.steer_angle));
    // coordinate system (right, up, forward) = (1, 2, 0)
    M3 b2 = m3(fwd.x, -right.x, up.x, fwd.y, -right.y, up.y, fwd.z, -right.z, up.z);
    w.wbasis = steer * b2;
}

// World transform of the car's compound child box (colObjWorldTransform *
// childTrans; the child has an identity basis and the hitbox offset).
RSX_HD Xf car_child_xf(const CarGeom& g, const Xf& car) {
    return Xf{car.b * m3_identity(), xform(car, g.box_offset)};
}

// Bullet ray test for a suspension ray (btDefaultVehicleRaycaster::castRay,
// btDefaultVehicleRaycaster.cpp:32-53): static world first, then the dynamic
// bodies (btRSBroadphase::rayTest visits a cell's static handles before its
// dynamic ones, btRSBroadphase.cpp:326-337), the ray's own chassis skipped
// (RayResultCallback::needsCollision, btCollisionWorld.cpp:1208-1214).  Convex
// bodies are cast with btSubsimplexConvexCast (btCollisionWorld.cpp:267-312):
// the ball's sphere, and each car's compound child box (the compound has no
// child tree, Car.cpp:198, so its one child is tested directly).  A closest
// body without contact response -- a car whose rigid body is disabled this
// tick -- gives no hit at all.
//
// Every car whose wheels are cast is live, so its own AABB marks the
// suspension grid cells around its hardpoints and the grid always defers to
// the full ray test (SuspensionCollisionGrid.cpp:131-133).  That test
// (btRSBroadphase::rayTest, btRSBroadphase.cpp:326-336) visits the bodies
// listed in the cell the ray starts in: a dynamic body is listed in the 3x3x3
// cells around the cell of its AABB minimum as of its last AABB update, which
// covers every body the ray can reach unless the body's state was set since
// the last world step (Car::bp_reg).
template <int MAXC>
RSX_HD void ray_world(const MeshView& mv, const Arena<MAXC>& A, const CarGeom& g, const BallGeom& bg, int self,
                      V3 from, V3 to, RayHit& hit, gjk::Simplex& sx, const RayHit& mesh_hit) {
    hit = mesh_hit;  // the mesh part, cast for the car's four rays together (ray_mesh4)
    ray_planes(from, to, hit);
    const int from_cell = bp_cell_index(mv, from);  // btRSBroadphase::rayTest: the start cell's list
    const V3 seg = to - from;
    const float seg2 = dot(seg, seg);
    // A cast only reports a hit where the ray comes within its stopping
    // distance (1e-2) of the body, so a body whose bounding sphere stays
    // farther than that from the whole segment is skipped; 0.25 keeps a wide
    // safety band around the stopping distance.
    auto far_from = [&](V3 centre, float radius) {
        float t = seg2 > 0.f ? clampf(dot(centre - from, seg) / seg2, 0.f, 1.f) : 0.f;
        V3 p = from + seg * t;
        float r = radius + 0.25f;
        return len2(centre - p) > r * r;
    };
    auto try_body = [&](const gjk::RayShape& s, int object) {
        float f;
        V3 n;
        if (far_from(s.xf.o, s.kind == 0 ? len(s.h) : s.h.x)) return;
        if (!gjk::ray_cast(s, from, to, f, n, sx)) return;
        if (!(len2(n) > 0.0001f)) return;
        float cur = hit.object >= 0 ? hit.fraction : 1.0f;
        if (!(f < cur)) return;
        hit.fraction = f;
        hit.normal = normalized(n);
        hit.object = object;
    };
    {
        gjk::RayShape s;
        s.kind = 1;
        s.h = V3{bg.radius, 0, 0};
        s.xf = A.ball.xf;
        if (bp_listed(mv, A.ball.bp_reg, from_cell)) try_body(s, 1);
    }
    for (int k = 0; k < A.ncars; k++) {
        if (k == self || !bp_listed(mv, A.car[k].bp_reg, from_cell)) continue;
        gjk::RayShape s;
        s.kind = 0;
        s.h = g.box_implicit + V3{g.box_margin, g.box_margin, g.box_margin};
        s.xf = car_child_xf(g, A.car[k].xf);
        try_body(s, 2 + k);
    }
    if (hit.object >= 2 && !A.car[hit.object - 2].active) hit.object = -1;
}

// The segment of wheel i's suspension ray (btVehicleRL::rayCast's source and
// target), with wheel_update_ws's hardpoint and direction.
RSX_HD void wheel_ray_segment(const Car& c, const CarGeom& g, int i, V3& source, V3& target) {
    const float travel = (kMaxSuspensionTravel * kUU2BT * 100) / 100;
    const float real_len = g.rest_len[i] + travel + g.wheel_radius[i] - kSuspensionSubtraction;
    source = xform(c.xf, g.wheel_cs[i]);
    target = source + (c.xf.b * V3{0, 0, -1}) * real_len;
}

// btVehicleRL::rayCast (btVehicleRL.cpp:118-212); mesh_hit is the ray's hit
// on the mesh triangles (object -1: none, or not cast).
template <int MAXC>
RSX_HD void wheel_raycast(Arena<MAXC>& A, int ci, const CarGeom& g, const BallGeom& bg, const MeshView& mv, int i,
                          float solver_dt, gjk::Simplex& sx, const RayHit& mesh_hit) {
    Car& c = A.car[ci];
    wheel_update_ws(c, g, i);
    Wheel& w = c.wheel[i];
    const float travel = (kMaxSuspensionTravel * kUU2BT * 100) / 100;
    const float rest = g.rest_len[i];
    const float radius = g.wheel_radius[i];
    float real_len = rest + travel + radius - kSuspensionSubtraction;
    V3 source = w.hardpoint;
    V3 target = source + w.dir_ws * real_len;
    w.contact_point = target;
    w.ground = -1;
    RayHit hit;
    ray_world(mv, A, g, bg, ci, source, target, hit, sx, mesh_hit);
    if (hit.object >= 0) {
        float s = 1.0f - hit.fraction;
        w.contact_point = V3{s * source.x + hit.fraction * target.x, s * source.y + hit.fraction * target.y,
                             s * source.z + hit.fraction * target.z};
        w.contact_normal = normalized(hit.normal);
        w.in_contact = true;
        w.in_contact_world = (hit.object == 0);
        w.ground = hit.object;
        V3 up = car_up(c);
        float trace = dot(w.hardpoint - w.contact_point, up);
        w.susp_len = trace - radius;
        w.susp_len = clampf(w.susp_len, rest - travel, rest + travel);  // std::clamp semantics agree here
        float denom = dot(w.contact_normal, up);
        V3 relpos = w.contact_point - c.xf.o;
        V3 vel_at = car_vel_at(c, relpos);
        float proj_vel = dot(w.contact_normal, vel_at);
        if (denom > 0.1f) {
            float inv = 1 / denom;
            w.susp_rel_vel = proj_vel * inv;
            w.clipped_inv_dot = inv;
        } else {
            w.susp_rel_vel = 0;
            w.clipped_inv_dot = 10;
        }
        if (hit.object == 0) {
            float thresh = (rest + radius) - kSuspensionSubtraction;
            if (trace < thresh) {
                float delta = trace - thresh;
                // resolveSingleCollision(chassis, static, hitPoint, hitNormal, info, delta, false)
                V3 n = w.contact_normal;
                V3 rel1 = w.contact_point - c.xf.o;
                V3 v1 = car_vel_at(c, rel1);
                float rel_vel = dot(n, v1);
                float restitution = 0.f * -rel_vel;
                float pos_err = kErp * -delta / solver_dt;
                float vel_err = -(1.0f + restitution) * rel_vel;
                V3 r0 = w.contact_point - c.xf.o;
                V3 c0 = cross(r0, n);
                V3 vec = cross(c0 * c.inv_inertia_world, r0);
                float denom0 = g.inv_mass + dot(n, vec);
                float jac_inv = 1.f / (denom0 + 0.f);
                float imp = pos_err * jac_inv + vel_err * jac_inv;
                imp = 0.f > imp ? 0.f : imp;
                w.extra_pushback = imp / 4;
            }
        }
    } else {
        w.susp_len = rest + travel;
        w.susp_rel_vel = 0;
        w.contact_normal = -w.dir_ws;
        w.clipped_inv_dot = 1;
        w.extra_pushback = 0;
    }
}

// Ground body of a wheel contact as resolveSingleBilateral sees it
// (btVehicleRL.cpp:321-343): velocity, inverse mass, local inverse inertia
// diagonal and centre of mass transform; the static world is the fixed body.
struct GroundBody {
    V3 vel, ang, inv_inertia_local, com;
    M3 basis;
    float inv_mass;
};

template <int MAXC>
RSX_HD GroundBody ground_body(const Arena<MAXC>& A, const CarGeom& g, const BallGeom& bg, int object) {
    GroundBody b;
    if (object == 1) {
        b.vel = A.ball.vel;
        b.ang = A.ball.ang;
        b.inv_mass = bg.inv_mass;
        b.inv_inertia_local = V3{bg.inv_inertia, bg.inv_inertia, bg.inv_inertia};
        b.com = A.ball.xf.o;
        b.basis = A.ball.xf.b;
    } else if (object >= 2) {
        const Car& k = A.car[object - 2];
        b.vel = k.vel;
        b.ang = k.ang;
        b.inv_mass = g.inv_mass;
        b.inv_inertia_local = g.inv_inertia_local;
        b.com = k.xf.o;
        b.basis = k.xf.b;
    } else {
        b.vel = b.ang = b.inv_inertia_local = b.com = V3{0, 0, 0};
        b.basis = m3_identity();
        b.inv_mass = 0.f;
    }
    return b;
}

// btVehicleRL::calcFrictionImpulses (btVehicleRL.cpp:313-387)
template <int MAXC>
RSX_HD void vehicle_friction_impulses(const Arena<MAXC>& A, Car& c, const CarGeom& g, float dt, const BallGeom& bg) {
    float friction_scale = car_mass(g) / 3;
    for (int i = 0; i < 4; i++) {
        Wheel& w = c.wheel[i];
        if (w.ground >= 0) {
            V3 axle = col(w.wbasis, 1);
            V3 n = w.contact_normal;
            float proj = dot(axle, n);
            axle -= n * proj;
            axle = safe_normalized(axle);
            V3 fwd = safe_normalized(cross(n, axle));
            // resolveSingleBilateral(chassis, p, ground, p, 0, axle, side, dt)
            const GroundBody gb = ground_body(A, g, bg, w.ground);
            float side;
            {
                V3 rel1 = w.contact_point - c.xf.o;
                V3 v1 = car_vel_at(c, rel1);
                V3 rel2 = w.contact_point - gb.com;
                V3 v2 = gb.vel + cross(gb.ang, rel2);
                V3 vel = v1 - v2;
                V3 aJ = transpose(c.xf.b) * cross(rel1, axle);
                V3 bJ = transpose(gb.basis) * cross(rel2, -axle);
                V3 m0 = g.inv_inertia_local * aJ;
                V3 m1 = gb.inv_inertia_local * bJ;
                float adiag = g.inv_mass + dot(m0, aJ) + gb.inv_mass + dot(m1, bJ);
                float jac_inv = 1.f / adiag;
                float rel_vel = dot(axle, vel);
                side = -0.2f * rel_vel * jac_inv;
            }
            float rolling;
            if (w.engine_force == 0) {
                if (w.brake != 0) {
                    V3 cp = w.contact_point;
                    V3 car_rel = cp - c.xf.o;
                    V3 v1 = car_vel_at(c, car_rel);
                    V3 v2 = V3{0, 0, 0};
                    if (w.ground >= 1) v2 = gb.vel + cross(gb.ang, car_rel);
                    V3 cv = v1 - v2;
                    float rv = dot(cv, fwd);
                    if (dt > (1 / 80.f)) {
                        float th = -(1 / (dt * 150.f)) + 0.8f;
                        if (fabsf(rv) < th) rv = 0;
                    }
                    const float kMagic = 113.73963f;
                    rolling = clampf(-rv * kMagic, -w.brake, w.brake);
                } else {
                    rolling = 0;
                }
            } else {
                rolling = -w.engine_force / friction_scale;
            }
            V3 total = (fwd * rolling * w.long_friction) + (axle * side * w.lat_friction);
            w.impulse = total * friction_scale;
        } else {
            w.impulse = V3{0, 0, 0};
        }
    }
}

// btVehicleRL::updateSuspension (btVehicleRL.cpp:277-310)
RSX_HD void vehicle_suspension(Car& c, const CarGeom& g, float dt) {
    for (int i = 0; i < 4; i++) {
        Wheel& w = c.wheel[i];
        if (w.in_contact) {
            float force = (g.rest_len[i] - w.susp_len) * kSuspStiffness * w.clipped_inv_dot;
            float damp = (w.susp_rel_vel < 0) ? kWheelsDampingCompression : kWheelsDampingRelaxation;
            w.susp_force = force - (damp * w.susp_rel_vel);
            w.susp_force *= g.susp_force_scale[i];
            if (w.susp_force < 0) w.susp_force = 0;
        } else {
            w.susp_force = 0;
        }
    }
    for (int i = 0; i < 4; i++) {
        Wheel& w = c.wheel[i];
        if (w.susp_force != 0) {
            V3 off = w.contact_point - c.xf.o;
            float base = (w.susp_force * dt) + w.extra_pushback;
            V3 f = w.contact_normal * base;
            car_apply_impulse(c, g, f, off);
        }
    }
}

// btVehicleRL::applyFrictionImpulses (btVehicleRL.cpp:390-402)
RSX_HD void vehicle_apply_friction(Car& c, const CarGeom& g, float dt) {
    V3 up = col(c.xf.b, 2);
    for (int i = 0; i < 4; i++) {
        Wheel& w = c.wheel[i];
        if (!is_zero(w.impulse)) {
            V3 off = w.contact_point - c.xf.o;
            float d = dot(up, off);
            V3 rel = off - up * d;
            car_apply_impulse(c, g, w.impulse * dt, rel);
        }
    }
}

RSX_HD V3 vehicle_up_from_contacts(const Car& c) {
    V3 sum = V3{0, 0, 0};
    for (int i = 0; i < 4; i++)
        if (c.wheel[i].in_contact) sum += c.wheel[i].contact_normal;
    if (sum == V3{0, 0, 0}) return col(c.xf.b, 2);
    return safe_normalized(sum);
}

// --- Car::_UpdateWheels -------------------------------------------------------
RSX_HD void car_update_wheels(Car& c, const CarGeom& g, float dt, int nwc, float fwd_speed) {
    float afs = fabsf(fwd_speed);
    bool world_contact = false;
    for (int i = 0; i < 4; i++) world_contact |= c.wheel[i].in_contact_world;
    if (c.controls.handbrake) c.handbrake_val += kPowerslideRiseRate * dt;
    else c.handbrake_val -= kPowerslideFallRate * dt;
    c.handbrake_val = clampf(c.handbrake_val, 0, 1);

    float real_throttle = c.controls.throttle;
    float real_brake = 0;
    if (c.controls.boost && c.boost > 0) real_throttle = 1;
    {
        float drive_scale = curve_eval(curve_drive_torque(), afs);
        float engine_throttle = real_throttle;
        if (c.controls.handbrake) {
        } else {
            float at = fabsf(real_throttle);
            if (at >= kThrottleDeadzone) {
                if (afs > kStoppingForwardVel && sgnf(real_throttle) != sgnf(fwd_speed)) {
                    real_brake = 1;
                    if (afs > kBrakingNoThrottleSpeedThresh) engine_throttle = 0;
                }
            } else {
                engine_throttle = 0;
                bool full_stop = (afs < kStoppingForwardVel);
                real_brake = full_stop ? 1 : kCoastingBrakeFactor;
            }
        }
        if (nwc < 3) drive_scale /= 4;
        float engine = engine_throttle * (kThrottleTorqueAmount * kUU2BT) * drive_scale;
        float brake = real_brake * (kBrakeTorqueAmount * kUU2BT);
        for (int i = 0; i < 4; i++) {
            c.wheel[i].engine_force = engine;
            c.wheel[i].brake = brake;
        }
    }
    {
        float steer = curve_eval(curve_steer(), afs);
        if (c.handbrake_val != 0) steer += (curve_eval(curve_powerslide_steer(), afs) - steer) * c.handbrake_val;
        steer *= c.controls.steer;
        c.wheel[0].steer_angle = steer;
        c.wheel[1].steer_angle = steer;
    }
    for (int i = 0; i < 4; i++) {
        Wheel& w = c.wheel[i];
        if (w.ground >= 0) {
            V3 vel = c.vel, ang = c.ang;
            V3 lat = col(w.wbasis, 1);
            V3 lon = cross(lat, w.contact_normal);
            float in = 0;
            V3 delta = w.hardpoint - c.xf.o;
            V3 cv = (cross(ang, delta) + vel) * kBT2UU;
            float base = fabsf(dot(cv, lat));
            if (base > 5) in = base / (fabsf(dot(cv, lon)) + base);
            float latf = curve_eval(curve_lat_friction(), in);
            float lonf = curve_eval(curve_long_friction(), in);
            if (c.handbrake_val != 0) {
                float hb = c.handbrake_val;
                latf *= (curve_eval(curve_hb_lat(), in) - 1) * hb + 1;
                lonf *= (curve_eval(curve_hb_long(), in) - 1) * hb + 1;
            } else {
                lonf = 1;
            }
            bool sticky = real_throttle != 0;
            if (!sticky) {
                float ns = curve_eval(curve_non_sticky(), w.contact_normal.z);
                latf *= ns;
                lonf *= ns;
            }
            w.lat_friction = latf;
            w.long_friction = lonf;
        }
    }
    if (world_contact) {
        V3 up = vehicle_up_from_contacts(c);
        bool full = (real_throttle != 0) || (afs > kStoppingForwardVel);
        float scale = 0.5f;
        if (full) scale += 1 - fabsf(up.z);
        c.force += up * scale * (kGravityZ * kUU2BT) * kCarMassBT;
    }
}

RSX_HD void car_update_boost(Car& c, const CarGeom& g, float dt) {
    if (c.time_spent_boosting > 0) {
        if (!c.controls.boost && c.time_spent_boosting >= kBoostMinTime) c.time_spent_boosting = 0;
        else c.time_spent_boosting += dt;
    } else {
        if (c.controls.boost) c.time_spent_boosting = dt;
    }
    if (c.boost > 0 && c.time_spent_boosting > 0) {
        c.boost = maxf(c.boost - kBoostUsedPerSecond * dt, 0);
        c.force += (c.on_ground ? kBoostAccelGround : kBoostAccelAir) * kUU2BT * car_fwd(c) * kCarMassBT;
    }
    c.boost = minf(c.boost, kBoostMax);
    (void)g;
}

RSX_HD void car_update_jump(Car& c, const CarGeom& g, float dt, bool jump_pressed) {
    if (c.on_ground && !c.is_jumping) {
        if (c.has_jumped && c.jump_time < kJumpMinTime + kJumpResetTimePad) {
        } else {
            c.has_jumped = false;
            c.jump_time = 0;
        }
    }
    if (c.is_jumping) {
        if (c.jump_time < kJumpMinTime || (c.controls.jump && c.jump_time < kJumpMaxTime)) c.is_jumping = true;
        else c.is_jumping = false;
    } else if (c.on_ground && jump_pressed) {
        c.is_jumping = true;
        c.jump_time = 0;
        V3 f = car_up(c) * kJumpImmediateForce * kUU2BT * kCarMassBT;
        car_apply_central_impulse(c, g, f);
    }
    if (c.is_jumping) {
        c.has_jumped = true;
        V3 total = car_up(c) * kJumpAccel;
        if (c.jump_time < kJumpMinTime) total *= 0.62f;
        c.force += total * kUU2BT * kCarMassBT;
    }
    if (c.is_jumping || c.has_jumped) c.jump_time += dt;
}

RSX_HD void car_update_air_torque(Car& c, const CarGeom& g, float dt, bool update_air_control) {
    V3 pitch_dir = -car_right(c);
    V3 yaw_dir = car_up(c);
    V3 roll_dir = -car_fwd(c);
    bool do_air = false;
    if (c.is_flipping) c.is_flipping = c.has_flipped && c.flip_time < kFlipTorqueTime;
    if (c.is_flipping) {
        V3 rel = c.flip_rel_torque;
        if (!is_zero(c.flip_rel_torque)) {
            float pitch_scale = 1;
            if (rel.y != 0 && c.controls.pitch != 0) {
                if (sgnf(rel.y) == sgnf(c.controls.pitch)) {
                    pitch_scale = 1 - minf(fabsf(c.controls.pitch), 1);
                    do_air = true;
                }
            }
            rel.y *= pitch_scale;
            V3 dodge = rel * V3{kFlipTorqueX, kFlipTorqueY, 0};
            c.torque += (inverse(c.inv_inertia_world) * c.xf.b) * dodge;
        } else {
            do_air = true;
        }
    } else {
        do_air = true;
    }
    do_air &= !c.is_auto_flipping;
    do_air &= update_air_control;
    if (do_air) {
        float pitch_ts = 1;
        V3 torque;
        if (c.controls.pitch != 0 || c.controls.yaw != 0 || c.controls.roll != 0) {
            if (c.is_flipping) pitch_ts = 0;
            else if (c.has_flipped) {
                if (c.flip_time < kFlipTorqueTime + kFlipPitchlockExtraTime) pitch_ts = 0;
            }
            torque = (pitch_dir * c.controls.pitch * pitch_ts * kAirTorqueP) + (yaw_dir * c.controls.yaw * kAirTorqueY) +
                     (roll_dir * c.controls.roll * kAirTorqueR);
        } else {
            torque = V3{0, 0, 0};
        }
        V3 av = c.ang;
        float dp = dot(pitch_dir, av) * kAirDampP * (1 - fabsf(do_air ? (c.controls.pitch * pitch_ts) : 0));
        float dy = dot(yaw_dir, av) * kAirDampY * (1 - fabsf(do_air ? c.controls.yaw : 0));
        float dr = dot(roll_dir, av) * kAirDampR;
        V3 damping = (yaw_dir * dy) + (pitch_dir * dp) + (roll_dir * dr);
        c.torque += inverse(c.inv_inertia_world) * (torque - damping) * kCarTorqueScale;
    }
    if (c.controls.throttle != 0) c.force += car_fwd(c) * c.controls.throttle * kThrottleAirAccel * kUU2BT * kCarMassBT;
    (void)g;
    (void)dt;
}

RSX_HD void car_update_double_jump_or_flip(Car& c, const CarGeom& g, float dt, bool jump_pressed, float fwd_speed) {
    if (c.on_ground) {
        c.has_double_jumped = false;
        c.has_flipped = false;
        c.air_time = 0;
        c.air_time_since_jump = 0;
        c.flip_time = 0;
    } else {
        c.air_time += dt;
        if (c.has_jumped && !c.is_jumping) c.air_time_since_jump += dt;
        else c.air_time_since_jump = 0;
        if (jump_pressed && c.air_time_since_jump < kDoubleJumpMaxDelay) {
            float mag = fabsf(c.controls.yaw) + fabsf(c.controls.pitch) + fabsf(c.controls.roll);
            bool is_flip = mag >= 0.5f;  // CarConfig::dodgeDeadzone
            bool can_use = (!c.has_double_jumped && !c.has_flipped);
            if (c.is_auto_flipping) can_use = false;
            if (can_use) {
                if (is_flip) {
                    c.flip_time = 0;
                    c.has_flipped = true;
                    c.is_flipping = true;
                    float ratio = fabsf(fwd_speed) / kCarMaxSpeed;
                    V3 dd = V3{-c.controls.pitch, c.controls.yaw + c.controls.roll, 0};
                    if (fabsf(c.controls.yaw + c.controls.roll) < 0.1f && fabsf(c.controls.pitch) < 0.1f) dd = V3{0, 0, 0};
                    else dd = safe_normalized(dd);
                    c.flip_rel_torque = V3{-dd.y, dd.x, 0};
                    if (fabsf(dd.x) < 0.1f) dd.x = 0;
                    if (fabsf(dd.y) < 0.1f) dd.y = 0;
                    if (!fuzzy_zero(dd)) {
                        bool back;
                        if (fabsf(fwd_speed) < 100.0f) back = dd.x < 0.0f;
                        else back = (dd.x >= 0.0f) != (fwd_speed >= 0.0f);
                        V3 iv = dd * kFlipInitialVelScale;
                        float msx = back ? kFlipBackwardImpulseMaxSpeedScale : kFlipForwardImpulseMaxSpeedScale;
                        iv.x *= ((msx - 1) * ratio) + 1.f;
                        iv.y *= ((kFlipSideImpulseMaxSpeedScale - 1) * ratio) + 1.f;
                        if (back) iv.x *= kFlipBackwardImpulseScaleX;
                        V3 f = car_fwd(c);
                        float a = atan2f(f.y, f.x);
                        V3 xd = V3{cosf(a), -sinf(a), 0.f};
                        V3 yd = V3{sinf(a), cosf(a), 0.f};
                        V3 fin = V3{dot(iv, xd), dot(iv, yd), 0.f};
                        car_apply_central_impulse(c, g, fin * kUU2BT * kCarMassBT);
                    }
                } else {
                    V3 f = car_up(c) * kJumpImmediateForce * kUU2BT * kCarMassBT;
                    car_apply_central_impulse(c, g, f);
                    c.has_double_jumped = true;
                }
            }
        }
    }
    if (c.is_flipping) {
        c.flip_time += dt;
        if (c.flip_time <= kFlipTorqueTime) {
            if (c.flip_time >= kFlipZDampStart && (c.vel.z < 0 || c.flip_time < kFlipZDampEnd))
                c.vel.z *= powf(1 - kFlipZDamp120, dt / (1 / 120.f));
        }
    } else if (c.has_flipped) {
        c.flip_time += dt;
    }
}

RSX_HD void car_update_auto_flip(Car& c, const CarGeom& g, float dt, bool jump_pressed) {
    if (jump_pressed && c.world_contact && c.world_contact_normal.z > kCarAutoflipNormzThresh) {
        // Angle::FromRotMat -> btMatrix3x3::getEulerYPR, roll negated (MathTypes.cpp:62-71)
        const M3& m = c.xf.b;
        float roll = atan2f(m.r[2].y, m.r[2].z);
        float pitch = asinf(-m.r[2].x);
        if (fabsf(pitch) == kHalfPi) {
            if (roll > 0) roll -= kPi;
            else roll += kPi;
        }
        roll *= -1;
        float ar = fabsf(roll);
        if (ar > kCarAutoflipRollThresh) {
            c.auto_flip_timer = (float)((double)kCarAutoflipTime * ((double)ar / 3.14159265358979323846));
            c.auto_flip_torque_scale = (roll > 0) ? 1 : -1;
            c.is_auto_flipping = true;
            car_apply_central_impulse(c, g, -car_up(c) * kCarAutoflipImpulse * kUU2BT * kCarMassBT);
        }
    }
    if (c.is_auto_flipping) {
        if (c.auto_flip_timer <= 0) {
            c.is_auto_flipping = false;
            c.auto_flip_timer = 0;
        } else {
            c.ang += car_fwd(c) * kCarAutoflipTorque * c.auto_flip_torque_scale * dt;
            c.auto_flip_timer -= dt;
        }
    }
}

RSX_HD void car_update_auto_roll(Car& c, const CarGeom& g, float dt, int nwc) {
    V3 gup = (nwc > 0) ? vehicle_up_from_contacts(c) : c.world_contact_normal;
    V3 gdown = -gup;
    V3 f = car_fwd(c), r = car_right(c);
    V3 cross_right = cross(gup, f);
    V3 cross_fwd = cross(gdown, cross_right);
    float rtf = 1 - clampf(dot(r, cross_right), 0, 1);
    float ftf = 1 - clampf(dot(f, cross_fwd), 0, 1);
    V3 tdr = f * (dot(r, gup) >= 0 ? -1.f : 1.f);
    V3 tdf = r * (dot(f, gup) >= 0 ? 1.f : -1.f);
    V3 tr = tdr * rtf;
    V3 tf = tdf * ftf;
    c.force += gdown * kCarAutorollForce * kUU2BT * kCarMassBT;
    c.torque += inverse(c.inv_inertia_world) * (tf + tr) * kCarAutorollTorque;
    (void)g;
    (void)dt;
}

// Car::_PreTickUpdate for a car that is not demoed (demo handling is done by the caller).
// `cast`: the mesh part of the four suspension rays, cast by the caller
// (null: cast here, unless rays_with_mesh is false).
template <int MAXC>
RSX_HD void car_pre_tick(Arena<MAXC>& A, int ci, const CarGeom& g, const BallGeom& bg, const MeshView& mv, float dt,
                         float solver_dt, gjk::Simplex& sx, bool rays_with_mesh = true, const RayHit* cast = nullptr) {
    Car& c = A.car[ci];
    // updateVehicleFirst
    for (int i = 0; i < 4; i++) wheel_update_transform(c, g, i);
    // the mesh part of the four suspension rays, in one pass (a car the
    // learned cull has shown to be out of the mesh's reach casts none)
    RayHit mesh_hit[4];
    if (!cast) {
        V3 from[4], to[4];
        for (int i = 0; i < 4; i++) {
            wheel_ray_segment(c, g, i, from[i], to[i]);
            mesh_hit[i].object = -1;
            mesh_hit[i].fraction = 1.0f;
        }
        if (rays_with_mesh) ray_mesh4(mv, from, to, mesh_hit);
        cast = mesh_hit;
    }
    for (int i = 0; i < 4; i++) wheel_raycast(A, ci, g, bg, mv, i, solver_dt, sx, cast[i]);
    vehicle_friction_impulses(A, c, g, dt, bg);

    bool jump_pressed = c.controls.jump && !c.last_controls.jump;
    int nwc = 0;
    for (int i = 0; i < 4; i++) {
        bool in = c.wheel[i].in_contact;
        c.wheel_contact[i] = in;
        nwc += in;
    }
    c.on_ground = nwc >= 3;
    float fwd_speed = dot(c.vel, col(c.xf.b, 0)) * kBT2UU;
    car_update_wheels(c, g, dt, nwc, fwd_speed);
    if (nwc < 3) car_update_air_torque(c, g, dt, nwc == 0);
    else c.is_flipping = false;
    car_update_jump(c, g, dt, jump_pressed);
    car_update_auto_flip(c, g, dt, jump_pressed);
    car_update_double_jump_or_flip(c, g, dt, jump_pressed, fwd_speed);
    if (c.controls.throttle != 0 && ((nwc > 0 && nwc < 4) || c.world_contact)) car_update_auto_roll(c, g, dt, nwc);
    c.world_contact = false;
    // updateVehicleSecond
    vehicle_suspension(c, g, dt);
    vehicle_apply_friction(c, g, dt);
    car_update_boost(c, g, dt);
}

RSX_HD void car_post_tick(Car& c, float dt) {
    float s2 = len2(c.vel * kBT2UU);
    if (c.is_supersonic && c.supersonic_time < kSupersonicMaintainMaxTime)
        c.is_supersonic = (s2 >= kSupersonicMaintainMinSpeed * kSupersonicMaintainMinSpeed);
    else
        c.is_supersonic = (s2 >= kSupersonicStartSpeed * kSupersonicStartSpeed);
    if (c.is_supersonic) c.supersonic_time += dt;
    else c.supersonic_time = 0;
    if (c.car_contact_cooldown > 0) c.car_contact_cooldown = maxf(c.car_contact_cooldown - dt, 0);
    c.last_controls = c.controls;
}

RSX_HD void car_finish_tick(Car& c) {
    if (!is_zero(c.vel_cache)) {
        c.vel += c.vel_cache;
        c.vel_cache = V3{0, 0, 0};
    }
    V3 v = c.vel, a = c.ang;
    if (len2(v) > (kCarMaxSpeed * kUU2BT) * (kCarMaxSpeed * kUU2BT)) v = normalized(v) * (kCarMaxSpeed * kUU2BT);
    if (len2(a) > (kCarMaxAngSpeed * kCarMaxAngSpeed)) a = normalized(a) * kCarMaxAngSpeed;
    c.vel = v;
    c.ang = a;
}

}  // namespace rsx
