// ref_driver.cpp — test-only C shim over the reference CPU RocketSim.
//
// ORACLE / TEST INFRASTRUCTURE.  This file is compiled together with the
// unmodified reference RocketSim + Bullet 3.24 sources (read in place from the
// reference tree by oracle/Makefile) into oracle/_ref/librsx_ref.so.  Only the
// tests, the smoke check and bench.py's cpu_baseline leg load it; the engine
// never links or calls it.
//
// It exposes the reference's Arena API (RocketSim/src/Sim/Arena/Arena.h:25-140)
// as plain C functions over the shared rsx_* records of include/rsx_types.h:
//   Arena::Create / AddCar            Arena.cpp:49-61, 568-570
//   Car::GetState / SetState          Car.cpp:10-36
//   Ball::GetState / SetState         Ball.cpp:27-49
//   BoostPad::GetState / SetState     BoostPad.h:44-45
//   Arena::Step                       Arena.cpp:716-812
//   Arena::IsBallScored               Arena.cpp:949-957
//   Arena::ResetToRandomKickoff       Arena.cpp:112-216
// Meshes are loaded through RocketSim::InitFromMem (RocketSim.cpp:102-212) in
// numeric file order mesh_0 … mesh_15, which fixes the static-body order that a
// directory scan would leave to the filesystem.

#include "RocketSim.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btTriangleInfoMap.h"
#include "../libsrc/bullet3-3.24/BulletCollision/CollisionShapes/btStridingMeshInterface.h"
#include "../include/rsx_types.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <vector>

using namespace RocketSim;

namespace {

struct RefArena {
    Arena* arena = nullptr;
    std::vector<Car*> cars;  // creation order: blue cars, then orange cars
};

Vec to_vec(const rsx_vec3& v) { return Vec(v.x, v.y, v.z); }
rsx_vec3 from_vec(const Vec& v) { return rsx_vec3{v.x, v.y, v.z}; }

CarControls to_controls(const rsx_controls& c) {
    CarControls r;
    r.throttle = c.throttle; r.steer = c.steer;
    r.pitch = c.pitch; r.yaw = c.yaw; r.roll = c.roll;
    r.jump = c.jump != 0; r.boost = c.boost != 0; r.handbrake = c.handbrake != 0;
    return r;
}

rsx_controls from_controls(const CarControls& c) {
    rsx_controls r{};
    r.throttle = c.throttle; r.steer = c.steer;
    r.pitch = c.pitch; r.yaw = c.yaw; r.roll = c.roll;
    r.jump = c.jump; r.boost = c.boost; r.handbrake = c.handbrake;
    return r;
}

void to_public(const CarState& s, const Car* car, rsx_car_state* o) {
    memset(o, 0, sizeof(*o));
    o->pos = from_vec(s.pos);
    o->rot.forward = from_vec(s.rotMat.forward);
    o->rot.right = from_vec(s.rotMat.right);
    o->rot.up = from_vec(s.rotMat.up);
    o->vel = from_vec(s.vel);
    o->ang_vel = from_vec(s.angVel);
    o->boost = s.boost;
    o->time_spent_boosting = s.timeSpentBoosting;
    o->jump_time = s.jumpTime;
    o->flip_time = s.flipTime;
    o->air_time = s.airTime;
    o->air_time_since_jump = s.airTimeSinceJump;
    o->supersonic_time = s.supersonicTime;
    o->handbrake_val = s.handbrakeVal;
    o->auto_flip_timer = s.autoFlipTimer;
    o->auto_flip_torque_scale = s.autoFlipTorqueScale;
    o->demo_respawn_timer = s.demoRespawnTimer;
    o->car_contact_cooldown = s.carContact.cooldownTimer;
    o->flip_rel_torque = from_vec(s.flipRelTorque);
    o->world_contact_normal = from_vec(s.worldContact.contactNormal);
    o->is_on_ground = s.isOnGround;
    o->has_jumped = s.hasJumped;
    o->has_double_jumped = s.hasDoubleJumped;
    o->has_flipped = s.hasFlipped;
    o->is_flipping = s.isFlipping;
    o->is_jumping = s.isJumping;
    o->is_supersonic = s.isSupersonic;
    o->is_auto_flipping = s.isAutoFlipping;
    o->is_demoed = s.isDemoed;
    o->world_contact = s.worldContact.hasContact;
    o->ball_hit_valid = s.ballHitInfo.isValid;
    o->team = (uint8_t)car->team;
    for (int i = 0; i < 4; i++) o->wheels_with_contact[i] = s.wheelsWithContact[i];
    o->id = car->id;
    o->car_contact_other_id = s.carContact.otherCarID;
    o->ball_hit_tick = s.ballHitInfo.tickCountWhenHit;
    o->ball_hit_extra_impulse_tick = s.ballHitInfo.tickCountWhenExtraImpulseApplied;
    o->ball_hit_rel_pos = from_vec(s.ballHitInfo.relativePosOnBall);
    o->ball_hit_ball_pos = from_vec(s.ballHitInfo.ballPos);
    o->ball_hit_extra_vel = from_vec(s.ballHitInfo.extraHitVel);
    o->last_controls = from_controls(s.lastControls);
}

CarState from_public(const rsx_car_state* o) {
    CarState s;
    s.pos = to_vec(o->pos);
    s.rotMat = RotMat(to_vec(o->rot.forward), to_vec(o->rot.right), to_vec(o->rot.up));
    s.vel = to_vec(o->vel);
    s.angVel = to_vec(o->ang_vel);
    s.boost = o->boost;
    s.timeSpentBoosting = o->time_spent_boosting;
    s.jumpTime = o->jump_time;
    s.flipTime = o->flip_time;
    s.airTime = o->air_time;
    s.airTimeSinceJump = o->air_time_since_jump;
    s.supersonicTime = o->supersonic_time;
    s.handbrakeVal = o->handbrake_val;
    s.autoFlipTimer = o->auto_flip_timer;
    s.autoFlipTorqueScale = o->auto_flip_torque_scale;
    s.demoRespawnTimer = o->demo_respawn_timer;
    s.carContact.cooldownTimer = o->car_contact_cooldown;
    s.flipRelTorque = to_vec(o->flip_rel_torque);
    s.worldContact.contactNormal = to_vec(o->world_contact_normal);
    s.isOnGround = o->is_on_ground;
    s.hasJumped = o->has_jumped;
    s.hasDoubleJumped = o->has_double_jumped;
    s.hasFlipped = o->has_flipped;
    s.isFlipping = o->is_flipping;
    s.isJumping = o->is_jumping;
    s.isSupersonic = o->is_supersonic;
    s.isAutoFlipping = o->is_auto_flipping;
    s.isDemoed = o->is_demoed;
    s.worldContact.hasContact = o->world_contact;
    s.ballHitInfo.isValid = o->ball_hit_valid;
    for (int i = 0; i < 4; i++) s.wheelsWithContact[i] = o->wheels_with_contact[i];
    s.carContact.otherCarID = o->car_contact_other_id;
    s.ballHitInfo.tickCountWhenHit = o->ball_hit_tick;
    s.ballHitInfo.tickCountWhenExtraImpulseApplied = o->ball_hit_extra_impulse_tick;
    s.ballHitInfo.relativePosOnBall = to_vec(o->ball_hit_rel_pos);
    s.ballHitInfo.ballPos = to_vec(o->ball_hit_ball_pos);
    s.ballHitInfo.extraHitVel = to_vec(o->ball_hit_extra_vel);
    s.lastControls = to_controls(o->last_controls);
    return s;
}

bool read_file(const std::string& path, std::vector<byte>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return !out.empty();
}

}  // namespace

extern "C" {

int ref_init(const char* mesh_dir) {
    if (RocketSim::GetStage() == RocketSimStage::INITIALIZED) return 0;
    std::map<GameMode, std::vector<FileData>> files;
    for (int i = 0; i < 16; i++) {
        std::vector<byte> data;
        std::string path = std::string(mesh_dir) + "/mesh_" + std::to_string(i) + ".cmf";
        if (!read_file(path, data)) return -1;
        files[GameMode::SOCCAR].push_back(data);
    }
    RocketSim::InitFromMem(files, true);
    return 0;
}

// Arena::Step visits Arena::_cars, a std::unordered_set<Car*>, whose order
// libstdc++ makes the reverse of insertion unless two car pointers share a
// bucket, which happens for an occasional 2v2 / 3v3 arena depending on where
// the cars were allocated.  The test arenas pin the usual order: when the set
// does not already visit the cars last-added first, it is rebuilt with one
// effective bucket (car pointers are 16-byte aligned and the table keeps two
// buckets at a load factor this high), where each insertion goes to the front.
static void pin_car_order(RefArena* h) {
    auto& set = h->arena->_cars;
    size_t k = h->cars.size();
    bool reversed = true;
    for (Car* c : set) reversed = reversed && k > 0 && c == h->cars[--k];
    if (reversed) return;
    set.clear();
    set.max_load_factor(1024.f);
    set.rehash(0);
    for (Car* c : h->cars) set.insert(c);
}

void* ref_arena_new(int n_blue, int n_orange) {
    auto* h = new RefArena();
    h->arena = Arena::Create(GameMode::SOCCAR);
    for (int i = 0; i < n_blue; i++) h->cars.push_back(h->arena->AddCar(Team::BLUE, CAR_CONFIG_OCTANE));
    for (int i = 0; i < n_orange; i++) h->cars.push_back(h->arena->AddCar(Team::ORANGE, CAR_CONFIG_OCTANE));
    pin_car_order(h);
    return h;
}

void ref_arena_free(void* p) {
    auto* h = (RefArena*)p;
    delete h->arena;
    delete h;
}

int ref_num_cars(void* p) { return (int)((RefArena*)p)->cars.size(); }

void ref_get_car(void* p, int i, rsx_car_state* out) {
    auto* h = (RefArena*)p;
    Car* car = h->cars[i];
    to_public(car->GetState(), car, out);
}

void ref_set_car(void* p, int i, const rsx_car_state* in) {
    auto* h = (RefArena*)p;
    h->cars[i]->SetState(from_public(in));
}

void ref_set_controls(void* p, int i, const rsx_controls* c) {
    auto* h = (RefArena*)p;
    h->cars[i]->controls = to_controls(*c);
}

void ref_get_ball(void* p, rsx_ball_state* out) {
    auto* h = (RefArena*)p;
    BallState s = h->arena->ball->GetState();
    out->pos = from_vec(s.pos);
    out->vel = from_vec(s.vel);
    out->ang_vel = from_vec(s.angVel);
}

void ref_set_ball(void* p, const rsx_ball_state* in) {
    auto* h = (RefArena*)p;
    BallState s;
    s.pos = to_vec(in->pos);
    s.vel = to_vec(in->vel);
    s.angVel = to_vec(in->ang_vel);
    h->arena->ball->SetState(s);
}

void ref_get_pads(void* p, rsx_pad_state* out) {
    auto* h = (RefArena*)p;
    auto& pads = h->arena->GetBoostPads();
    for (size_t i = 0; i < pads.size(); i++) {
        BoostPadState s = pads[i]->GetState();
        memset(&out[i], 0, sizeof(rsx_pad_state));
        out[i].is_active = s.isActive;
        out[i].cooldown = s.cooldown;
        out[i].prev_locked_car_id = s.prevLockedCarID;
    }
}

void ref_set_pads(void* p, const rsx_pad_state* in) {
    auto* h = (RefArena*)p;
    auto& pads = h->arena->GetBoostPads();
    for (size_t i = 0; i < pads.size(); i++) {
        BoostPadState s;
        s.isActive = in[i].is_active;
        s.cooldown = in[i].cooldown;
        s.curLockedCar = NULL;
        s.prevLockedCarID = in[i].prev_locked_car_id;
        pads[i]->SetState(s);
    }
}

void ref_pad_config(void* p, float* xyz_big) {  // 34 x (x, y, z, isBig)
    auto* h = (RefArena*)p;
    auto& pads = h->arena->GetBoostPads();
    for (size_t i = 0; i < pads.size(); i++) {
        xyz_big[i * 4 + 0] = pads[i]->config.pos.x;
        xyz_big[i * 4 + 1] = pads[i]->config.pos.y;
        xyz_big[i * 4 + 2] = pads[i]->config.pos.z;
        xyz_big[i * 4 + 3] = pads[i]->config.isBig ? 1.f : 0.f;
    }
}

void ref_step(void* p, int ticks) { ((RefArena*)p)->arena->Step(ticks); }

uint64_t ref_get_tick(void* p) { return ((RefArena*)p)->arena->tickCount; }
void ref_set_tick(void* p, uint64_t t) { ((RefArena*)p)->arena->tickCount = t; }

int ref_ball_scored(void* p) { return ((RefArena*)p)->arena->IsBallScored() ? 1 : 0; }

void ref_reset_kickoff(void* p, int seed) { ((RefArena*)p)->arena->ResetToRandomKickoff(seed); }

// Hidden per-wheel vehicle state (btWheelInfoRL, btVehicleRL.h) that carries
// from one tick to the next: 4 wheels x 12 floats.
void ref_get_wheels(void* p, int i, float* out) {
    auto* h = (RefArena*)p;
    Car* car = h->cars[i];
    for (int w = 0; w < 4; w++) {
        auto& wi = car->_bulletVehicle.m_wheelInfo[w];
        float* o = out + w * 12;
        o[0] = wi.m_steerAngle;
        o[1] = wi.m_engineForce;
        o[2] = wi.m_brake;
        o[3] = wi.m_latFriction;
        o[4] = wi.m_longFriction;
        o[5] = wi.m_raycastInfo.m_suspensionLength;
        o[6] = wi.m_suspensionRelativeVelocity;
        o[7] = wi.m_clippedInvContactDotSuspension;
        o[8] = wi.m_extraPushback;
        o[9] = wi.m_raycastInfo.m_isInContact ? 1.f : 0.f;
        o[10] = wi.m_wheelsSuspensionForce;
        o[11] = wi.m_isInContactWithWorld ? 1.f : 0.f;
    }
}

void ref_set_wheels(void* p, int i, const float* in) {
    auto* h = (RefArena*)p;
    Car* car = h->cars[i];
    for (int w = 0; w < 4; w++) {
        auto& wi = car->_bulletVehicle.m_wheelInfo[w];
        const float* o = in + w * 12;
        wi.m_steerAngle = o[0];
        wi.m_engineForce = o[1];
        wi.m_brake = o[2];
        wi.m_latFriction = o[3];
        wi.m_longFriction = o[4];
        wi.m_extraPushback = o[8];
    }
}

// Number of live contact points over all persistent manifolds of the arena's
// Bullet world (a diagnostic for which scenario cases touch contacts at all).
int ref_num_contacts(void* p) {
    auto* h = (RefArena*)p;
    auto* disp = h->arena->_bulletWorld.getDispatcher();
    int n = 0;
    for (int m = 0; m < disp->getNumManifolds(); m++) n += disp->getManifoldByIndexInternal(m)->getNumContacts();
    return n;
}

uint32_t ref_sizeof_car_state(void) { return (uint32_t)sizeof(rsx_car_state); }

// Identify a Bullet collision object of the arena: static bodies 0..N-1 in
// creation order (meshes, then planes), the ball 100, car i 200+i.
static int body_code(RefArena* h, const btCollisionObject* o) {
    if (o == &h->arena->ball->_rigidBody) return 100;
    for (size_t i = 0; i < h->cars.size(); i++)
        if (o == &h->cars[i]->_rigidBody) return 200 + (int)i;
    for (size_t i = 0; i < h->arena->_worldCollisionRBAmount; i++)
        if (o == &h->arena->_worldCollisionRBs[i]) return (int)i;
    return -1;
}

// Dump the manifolds Bullet built during the last tick, in dispatcher order.
// Record per manifold: body0, body1, numPoints, breakingThreshold, then per
// point 22 floats: posA(3) posB(3) normalB(3) dist friction restitution
// isSpecial appliedImpulse localA(3) localB(3) lateralImpulse1 lifeTime.
// Returns the number of floats written.
int ref_dump_contacts(void* p, float* out, int max_floats) {
    auto* h = (RefArena*)p;
    auto* disp = h->arena->_bulletWorld.getDispatcher();
    int n = 0;
    for (int m = 0; m < disp->getNumManifolds(); m++) {
        btPersistentManifold* mf = disp->getManifoldByIndexInternal(m);
        int np = mf->getNumContacts();
        if (n + 4 + np * 22 > max_floats) break;
        out[n++] = (float)body_code(h, mf->getBody0());
        out[n++] = (float)body_code(h, mf->getBody1());
        out[n++] = (float)np;
        out[n++] = mf->getContactBreakingThreshold();
        for (int i = 0; i < np; i++) {
            btManifoldPoint& cp = mf->getContactPoint(i);
            const btVector3 vals[] = {cp.m_positionWorldOnA, cp.m_positionWorldOnB, cp.m_normalWorldOnB};
            for (auto& v : vals) { out[n++] = v.x(); out[n++] = v.y(); out[n++] = v.z(); }
            out[n++] = cp.m_distance1;
            out[n++] = cp.m_combinedFriction;
            out[n++] = cp.m_combinedRestitution;
            out[n++] = cp.m_isSpecial ? 1.f : 0.f;
            out[n++] = cp.m_appliedImpulse;
            const btVector3 loc[] = {cp.m_localPointA, cp.m_localPointB};
            for (auto& v : loc) { out[n++] = v.x(); out[n++] = v.y(); out[n++] = v.z(); }
            out[n++] = cp.m_appliedImpulseLateral1;
            out[n++] = (float)cp.m_lifeTime;
        }
    }
    return n;
}

// World-space AABB (Bullet units) the broadphase holds for the ball (-1) or car i.
void ref_get_aabb(void* p, int i, float* out6) {
    auto* h = (RefArena*)p;
    btCollisionObject* o = i < 0 ? (btCollisionObject*)&h->arena->ball->_rigidBody
                                 : (btCollisionObject*)&h->cars[i]->_rigidBody;
    btBroadphaseProxy* px = o->getBroadphaseHandle();
    out6[0] = px->m_aabbMin.x(); out6[1] = px->m_aabbMin.y(); out6[2] = px->m_aabbMin.z();
    out6[3] = px->m_aabbMax.x(); out6[4] = px->m_aabbMax.y(); out6[5] = px->m_aabbMax.z();
}

}  // extern "C"

// World ray test (btCollisionWorld::rayTest with a closest-hit callback) from
// a to b in Bullet units, skipping car `ignore_car` (-1: none); out: hit(1/0),
// fraction, point(3), normal(3), body code.
extern "C" void ref_raytest(void* p, const float* a, const float* b, int ignore_car, float* out) {
    auto* h = (RefArena*)p;
    btVector3 from(a[0], a[1], a[2]), to(b[0], b[1], b[2]);
    const btCollisionObject* ign = ignore_car >= 0 ? &h->cars[ignore_car]->_rigidBody : nullptr;
    btCollisionWorld::ClosestRayResultCallback cb(from, to, ign);
    h->arena->_bulletWorld.rayTest(from, to, cb);
    out[0] = cb.hasHit() ? 1.f : 0.f;
    out[1] = cb.m_closestHitFraction;
    out[2] = cb.m_hitPointWorld.x(); out[3] = cb.m_hitPointWorld.y(); out[4] = cb.m_hitPointWorld.z();
    out[5] = cb.m_hitNormalWorld.x(); out[6] = cb.m_hitNormalWorld.y(); out[7] = cb.m_hitNormalWorld.z();
    out[8] = cb.hasHit() ? (float)body_code(h, cb.m_collisionObject) : -1.f;
}

// Internal-edge table of soccar collision mesh `mesh` as Bullet built it
// (btGenerateInternalEdgeInfo): per triangle 13 floats, the three vertices
// then the V0V1 / V1V2 / V2V0 angles and the flags (bit pattern).  Returns
// the triangle count; writes at most max_tris rows.
extern "C" int ref_mesh_tri_info(int mesh, float* out, int max_tris) {
    auto& shapes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (mesh < 0 || mesh >= (int)shapes.size()) return -1;
    btBvhTriangleMeshShape* sh = shapes[mesh];
    const btTriangleInfoMap* map = sh->getTriangleInfoMap();
    const unsigned char* verts; int nverts; int stride;
    const unsigned char* idx; int istride; int nfaces;
    sh->getMeshInterface()->getLockedReadOnlyVertexIndexBase(&verts, nverts, stride, &idx, istride, nfaces, 0);
    for (int t = 0; t < nfaces && t < max_tris; t++) {
        float* o = out + (size_t)t * 13;
        const int* tri = (const int*)(idx + (size_t)t * istride);
        for (int k = 0; k < 3; k++) {
            const float* v = (const float*)(verts + (size_t)tri[k] * stride);
            o[k * 3 + 0] = v[0]; o[k * 3 + 1] = v[1]; o[k * 3 + 2] = v[2];
        }
        btTriangleInfo info;
        const btTriangleInfo* f = map ? map->find(btHashInt(t)) : nullptr;
        if (f) info = *f;
        o[9] = info.m_edgeV0V1Angle; o[10] = info.m_edgeV1V2Angle; o[11] = info.m_edgeV2V0Angle;
        memcpy(&o[12], &info.m_flags, 4);
    }
    sh->getMeshInterface()->unLockReadOnlyVertexBase(0);
    return nfaces;
}

// Ray test of one convex body (btCollisionWorld::rayTestSingle with a
// closest-hit callback): kind 0 a box with implicit half extents dims[0..2]
// and margin dims[3], kind 1 a sphere of radius dims[0].  xf12: row-major
// basis (9) then origin (3).  out: hit(1/0), fraction, normal(3).
extern "C" void ref_ray_convex(int kind, const float* dims, const float* xf12, const float* a, const float* b,
                               float* out) {
    btBoxShape box(btVector3(dims[0] + 0.04f, dims[1] + 0.04f, dims[2] + 0.04f));
    btSphereShape sphere(dims[0]);
    btCollisionShape* shape = &sphere;
    if (kind == 0) {
        box.btConvexInternalShape::setMargin(dims[3]);
        shape = &box;
    }
    btTransform t;
    t.setBasis(btMatrix3x3(xf12[0], xf12[1], xf12[2], xf12[3], xf12[4], xf12[5], xf12[6], xf12[7], xf12[8]));
    t.setOrigin(btVector3(xf12[9], xf12[10], xf12[11]));
    btCollisionObject obj;
    obj.setCollisionShape(shape);
    obj.setWorldTransform(t);
    btVector3 from(a[0], a[1], a[2]), to(b[0], b[1], b[2]);
    btTransform tf, tt;
    tf.setIdentity();
    tf.setOrigin(from);
    tt.setIdentity();
    tt.setOrigin(to);
    btCollisionWorld::ClosestRayResultCallback cb(from, to, nullptr);
    btCollisionWorld::rayTestSingle(tf, tt, &obj, shape, t, cb);
    out[0] = cb.hasHit() ? 1.f : 0.f;
    out[1] = cb.m_closestHitFraction;
    out[2] = cb.m_hitNormalWorld.x(); out[3] = cb.m_hitNormalWorld.y(); out[4] = cb.m_hitNormalWorld.z();
}

// CPU baseline loop for the benchmark: `steps` env steps over the arenas
// [arenas, arenas + n): each step draws one action per car uniformly from the
// table (xorshift32 stream seeded by `seed`), then runs Arena::Step(tick_skip).
// A kickoff reset follows a goal, as the env loop does.  Returns the number of
// env steps done.
extern "C" int64_t ref_run_random(void** arenas, int n, int steps, int tick_skip, uint32_t seed,
                                  const rsx_controls* lut, int lut_n) {
    uint32_t r = seed ? seed : 1u;
    auto next = [&]() {
        r ^= r << 13;
        r ^= r >> 17;
        r ^= r << 5;
        return r;
    };
    int64_t done = 0;
    for (int s = 0; s < steps; s++) {
        for (int a = 0; a < n; a++) {
            auto* h = (RefArena*)arenas[a];
            for (Car* c : h->cars) c->controls = to_controls(lut[next() % (uint32_t)lut_n]);
            h->arena->Step(tick_skip);
            if (h->arena->IsBallScored()) h->arena->ResetToRandomKickoff((int)(next() & 0x7fffffff));
            done++;
        }
    }
    return done;
}

// Ids of the arena's cars in the order Arena::Step visits them (the
// iteration order of Arena::_cars, a std::unordered_set<Car*>).  Returns the
// number written.
extern "C" int ref_car_order(void* p, uint32_t* out, int cap) {
    auto* h = (RefArena*)p;
    int n = 0;
    for (Car* c : h->arena->GetCars())
        if (n < cap) out[n++] = c->id;
    return n;
}

// Contact tracing (a diagnostic): while on, every point Bullet adds to a
// manifold is logged as it leaves btManifoldResult::addContactPoint, before
// and after RocketSim's contact-added callback (which corrects normals against
// internal mesh edges): 16 floats per point, body codes of objects A and B,
// the triangle indices (index0, index1), position on B (3), normal before (3)
// and after (3) the callback, depth, and the point's slot in its manifold
// (-1 when not found).  Re-arm after creating an arena: Arena's constructor
// installs its own callback.
namespace {
std::vector<float> g_trace;
RefArena* g_trace_arena = nullptr;
bool trace_cb(btManifoldPoint& cp, const btCollisionObjectWrapper* a, int pa, int ia, const btCollisionObjectWrapper* b,
              int pb, int ib) {
    btVector3 n0 = cp.m_normalWorldOnB;
    bool r = Arena::_BulletContactAddedCallback(cp, a, pa, ia, b, pb, ib);
    RefArena* h = g_trace_arena;
    float slot = -1;
    auto* disp = h->arena->_bulletWorld.getDispatcher();
    for (int m = 0; m < disp->getNumManifolds(); m++) {
        btPersistentManifold* mf = disp->getManifoldByIndexInternal(m);
        for (int i = 0; i < mf->getNumContacts(); i++)
            if (&mf->getContactPoint(i) == &cp) slot = (float)i;
    }
    const float v[16] = {(float)body_code(h, a->getCollisionObject()), (float)body_code(h, b->getCollisionObject()),
                         (float)ia, (float)ib, cp.m_positionWorldOnB.x(), cp.m_positionWorldOnB.y(),
                         cp.m_positionWorldOnB.z(), n0.x(), n0.y(), n0.z(), cp.m_normalWorldOnB.x(),
                         cp.m_normalWorldOnB.y(), cp.m_normalWorldOnB.z(), cp.m_distance1, slot, 0.f};
    g_trace.insert(g_trace.end(), v, v + 16);
    return r;
}
}  // namespace

extern "C" void ref_trace_contacts(void* p, int on) {
    g_trace.clear();
    g_trace_arena = (RefArena*)p;
    gContactAddedCallback = on ? &trace_cb : &Arena::_BulletContactAddedCallback;
}

extern "C" int ref_trace_read(float* out, int cap) {
    int n = (int)g_trace.size() < cap ? (int)g_trace.size() : cap;
    memcpy(out, g_trace.data(), n * sizeof(float));
    return n;
}

// Triangle indices of soccar mesh `mesh` in the order of its quantized BVH's
// leaf nodes (btQuantizedBvh node array, which the stackless traversal walks
// in order).  Returns the number of leaves.
extern "C" int ref_mesh_leaf_order(int mesh, int32_t* out, int cap) {
    auto& shapes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (mesh < 0 || mesh >= (int)shapes.size()) return -1;
    btOptimizedBvh* bvh = shapes[mesh]->getOptimizedBvh();
    if (!bvh || !bvh->isQuantized()) return -2;
    const QuantizedNodeArray& nodes = bvh->getQuantizedNodeArray();
    int n = 0;
    for (int i = 0; i < nodes.size(); i++)
        if (nodes[i].isLeafNode()) {
            if (n < cap) out[n] = nodes[i].getTriangleIndex();
            n++;
        }
    return n;
}

// Triangle indices of soccar mesh `mesh` that btBvhTriangleMeshShape::
// processAllTriangles reports for the AABB (Bullet units), in report order.
extern "C" int ref_mesh_query(int mesh, const float* aabb6, int32_t* out, int cap) {
    struct Cb : btTriangleCallback {
        int32_t* out; int cap; int n = 0;
        void processTriangle(btVector3*, int, int idx) override {
            if (n < cap) out[n] = idx;
            n++;
        }
    } cb;
    cb.out = out;
    cb.cap = cap;
    auto& shapes = RocketSim::GetArenaCollisionShapes(GameMode::SOCCAR);
    if (mesh < 0 || mesh >= (int)shapes.size()) return -1;
    shapes[mesh]->processAllTriangles(&cb, btVector3(aabb6[0], aabb6[1], aabb6[2]),
                                      btVector3(aabb6[3], aabb6[4], aabb6[5]));
    return cb.n;
}

// ---- The RLGym env half of a step over the oracle, for the CPU baseline ----
// A C++ restatement of the pieces the engine's fused env step computes, run
// on the oracle's own states: DefaultObs rows (DefaultObs.cpp:4-54 with
// InvertPhys, StateUtil.cpp:3-17, and GameState's pad order,
// GameState.cpp:93-109), the four stub rewards (GoalReward,
// TouchBallReward, VelocityBallToGoalReward, VelocityPlayerToBallReward,
// CommonRewards.h:33-99, weights 1 and concede -1), GoalScoreCondition and
// NoTouchCondition (EnvSet.cpp:169-242), and the kickoff reset of terminal
// arenas.  pad_map[i] is the arena pad at RLGym BOOST_LOCATIONS slot i.
// tests/test_env_cpu.py checks it against oracle/rlgym_spec.py.
namespace {

struct EnvCtx {
    int tick_skip;
    const rsx_controls* lut;
    int lut_n;
    const int32_t* pad_map;
    float timeout;
    int auto_reset;
};

const Vec kInv(-1, -1, 1);

void obs_player(const CarState& s, bool inv, float*& o) {
    const Vec pos_coef(1 / 4096.f, 1 / 5120.f, 1 / 2044.f);
    const float vel_coef = 1 / 2300.f, ang_coef = 1 / 5.5f;
    Vec pos = s.pos, fwd = s.rotMat.forward, up = s.rotMat.up, vel = s.vel, ang = s.angVel;
    if (inv) {
        pos = pos * kInv; fwd = fwd * kInv; up = up * kInv; vel = vel * kInv; ang = ang * kInv;
    }
    Vec p = pos * pos_coef, v = vel * vel_coef, w = ang * ang_coef;
    const float vals[19] = {p.x, p.y, p.z, fwd.x, fwd.y, fwd.z, up.x, up.y, up.z, v.x, v.y, v.z, w.x, w.y, w.z,
                            s.boost / 100, (float)s.isOnGround, (float)s.HasFlipOrJump(), (float)s.isDemoed};
    for (float x : vals) *o++ = x;
}

// One env step of arena h (cars in creation order, act[i] for car i).
// Writes rew[i] and the DefaultObs row of car i at obs + i * row; returns the
// terminal type (0 none, 1 goal, 2 no-touch truncation).
int env_step_one(RefArena* h, const int32_t* act, const EnvCtx& c, float& no_touch, float* obs, float* rew,
                 int reset_seed) {
    const int n = (int)h->cars.size();
    float prev[RSX_MAX_CARS][8];
    for (int i = 0; i < n; i++) {
        int k = act[i] < 0 ? 0 : (act[i] >= c.lut_n ? c.lut_n - 1 : act[i]);
        const rsx_controls& ct = c.lut[k];
        h->cars[i]->controls = to_controls(ct);
        const float p[8] = {ct.throttle, ct.steer, ct.pitch, ct.yaw, ct.roll, (float)ct.jump, (float)ct.boost,
                            (float)ct.handbrake};
        memcpy(prev[i], p, sizeof(p));
    }
    Arena* a = h->arena;
    a->Step(c.tick_skip);
    const bool goal = a->IsBallScored();
    const BallState ball = a->ball->GetState();
    bool touched[RSX_MAX_CARS], any = false;
    CarState cs[RSX_MAX_CARS];
    for (int i = 0; i < n; i++) {
        cs[i] = h->cars[i]->GetState();
        touched[i] = cs[i].ballHitInfo.isValid &&
                     cs[i].ballHitInfo.tickCountWhenHit >= a->tickCount - (uint64_t)c.tick_skip;
        any |= touched[i];
    }
    int term = goal ? 1 : 0;
    if (c.timeout > 0) {
        if (any) {
            no_touch = 0;
        } else {
            no_touch += c.tick_skip * (1 / 120.f);
            if (no_touch >= c.timeout && term == 0) term = 2;
        }
    }
    for (int i = 0; i < n; i++) {
        const Team team = h->cars[i]->team;
        float r = 0;
        float g = 0;
        if (goal) g = (team != RS_TEAM_FROM_Y(ball.pos.y)) ? 1.f : -1.f;
        r += g * 1.f;
        r += (touched[i] ? 1.f : 0.f) * 1.f;
        const Vec target(0, team == Team::BLUE ? 6000.f : -6000.f, 642.775f / 2);
        r += (target - ball.pos).Normalized().Dot(ball.vel / 6000.f) * 1.f;
        r += (ball.pos - cs[i].pos).Normalized().Dot(cs[i].vel / 2300.f) * 1.f;
        rew[i] = r;
    }
    if (term && c.auto_reset) {
        a->ResetToRandomKickoff(reset_seed);
        a->tickCount = 0;
        no_touch = 0;
        for (int i = 0; i < n; i++) {
            cs[i] = h->cars[i]->GetState();
            for (int k = 0; k < 8; k++) prev[i][k] = 0;
        }
    }
    const BallState b = a->ball->GetState();
    const auto& pads = a->GetBoostPads();
    const int row = 51 + 19 * n;
    for (int i = 0; i < n; i++) {
        float* o = obs + (size_t)i * row;
        const bool inv = h->cars[i]->team == Team::ORANGE;
        const Vec pos_coef(1 / 4096.f, 1 / 5120.f, 1 / 2044.f);
        Vec bp = b.pos, bv = b.vel, ba = b.angVel;
        if (inv) { bp = bp * kInv; bv = bv * kInv; ba = ba * kInv; }
        Vec p = bp * pos_coef, v = bv * (1 / 2300.f), w = ba * (1 / 5.5f);
        const float head[9] = {p.x, p.y, p.z, v.x, v.y, v.z, w.x, w.y, w.z};
        for (float x : head) *o++ = x;
        for (int k = 0; k < 8; k++) *o++ = prev[i][k];
        for (int k = 0; k < 34; k++) {
            int slot = inv ? c.pad_map[33 - k] : c.pad_map[k];
            *o++ = pads[slot]->GetState().isActive ? 1.f : 0.f;
        }
        obs_player(cs[i], inv, o);
        // teammates then opponents, in GameState::players (Arena::_cars) order
        for (int pass = 0; pass < 2; pass++)
            for (int j = n - 1; j >= 0; j--) {
                if (j == i) continue;
                const bool mate = h->cars[j]->team == h->cars[i]->team;
                if (mate == (pass == 0)) obs_player(cs[j], inv, o);
            }
    }
    return term;
}

}  // namespace

// One env step of each arena [arenas, arenas + n) with the given action
// indices (the arenas' cars back to back); obs rows, rewards and terminals
// out in the same order; no_touch[a] carries each arena's NoTouchCondition
// timer.  Terminal arenas are reset to a kickoff drawn from reset_seed + a
// when auto_reset.
extern "C" void ref_env_step(void** arenas, int n, const int32_t* actions, int tick_skip, const rsx_controls* lut,
                             int lut_n, const int32_t* pad_map, float timeout, int auto_reset, float* no_touch,
                             float* obs, float* rew, uint8_t* term, int reset_seed) {
    EnvCtx c{tick_skip, lut, lut_n, pad_map, timeout, auto_reset};
    size_t car = 0, of = 0;
    for (int a = 0; a < n; a++) {
        auto* h = (RefArena*)arenas[a];
        const int nc = (int)h->cars.size();
        term[a] = (uint8_t)env_step_one(h, actions + car, c, no_touch[a], obs + of, rew + car, reset_seed + a);
        car += nc;
        of += (size_t)nc * (51 + 19 * nc);
    }
}

// CPU baseline loop: `steps` env steps over the arenas, one uniform table
// action per car per step from a xorshift32 stream seeded by `seed`, the env
// outputs of each step written to scratch owned by the call.  Returns the
// number of env steps done.
extern "C" int64_t ref_env_run(void** arenas, int n, int steps, int tick_skip, uint32_t seed, const rsx_controls* lut,
                               int lut_n, const int32_t* pad_map, float timeout) {
    EnvCtx c{tick_skip, lut, lut_n, pad_map, timeout, 1};
    uint32_t r = seed ? seed : 1u;
    auto next = [&]() {
        r ^= r << 13;
        r ^= r >> 17;
        r ^= r << 5;
        return r;
    };
    std::vector<float> no_touch(n, 0.f), obs(RSX_MAX_CARS * (51 + 19 * RSX_MAX_CARS)), rew(RSX_MAX_CARS);
    int32_t act[RSX_MAX_CARS];
    int64_t done = 0;
    for (int s = 0; s < steps; s++)
        for (int a = 0; a < n; a++) {
            auto* h = (RefArena*)arenas[a];
            for (size_t i = 0; i < h->cars.size(); i++) act[i] = (int32_t)(next() % (uint32_t)lut_n);
            env_step_one(h, act, c, no_touch[a], obs.data(), rew.data(), (int)(next() & 0x7fffffff));
            done++;
        }
    return done;
}
