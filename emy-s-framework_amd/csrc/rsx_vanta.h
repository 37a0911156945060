// rsx_vanta.h — the VantaObs observation (366 floats per car) of the
// reference's GPU path, with its per-arena and per-car state and the
// ball-prediction slices, on the engine's own records (host and device).
//
// Reference behaviour restated here (RocketSimCuda/src/):
//   layout, constants                    VantaObsGpu.cuh:22-43
//   persistent state                     VantaObsGpu.cuh:47-81
//   InvertPhys / MirrorPhysX / canonical VantaObsGpu.cuh:98-140
//   flip-reset flags, pad pooling        VantaObsGpu.cuh:152-188
//   UpdateBoostPads / UpdateDemoTimers   VantaObsGpu.cuh:192-243
//   coyote timers                        VantaObsGpu.cuh:246-338
//   the row (sections A-H)               VantaObsGpu.cuh:342-747
//   ball prediction, last-touch car      RocketSimCudaObs.cuh:70-186
//   pad tables                           RocketSimCudaHostUtils.cuh:55-90
//
// Where it runs: k_vanta (rsx_vanta.hip), one lane per arena, after the env
// step (or from rsx_build_obs).  The ball prediction is a ball-only copy of
// the arena stepped 360 ticks by the engine's own tick (arena_tick with no
// cars), so it is the same physics as the main step, and the ball-only
// RocketSim arena BallPredTracker steps (BallPredTracker.cpp:5-73).
#pragma once

#include "rsx_env.h"
#include "rsx_soa.h"

namespace rsx {

namespace vanta {

constexpr float kPosStd = 2300.f, kVelStd = 2300.f, kBallVelStd = 6000.f, kAngStd = 5.5f, kDemoStd = 3.f;
constexpr float kGoalDistCoef = 1.f / 10240.f;
constexpr int kEgo = 81, kBall = 29, kGoal = 16, kTeamCtx = 6, kCar = 39;
constexpr int kMaxMates = 2, kMaxOpps = 3, kMask = kMaxMates + kMaxOpps, kPads = 34;
constexpr int kRow = kEgo + kBall + kGoal + kTeamCtx + kCar * (kMaxMates + kMaxOpps) + kMask + kPads;  // 366
static_assert(kRow == 366, "VantaObs row");
constexpr int kPredTicks = 360;   // slices after 120, 240 and 360 ball-only ticks

// Per-arena words (field-major, like the arena words): pad timers, pad
// availability bits, flags, last updated tick, prediction slices; then per
// car slot its coyote timers.  All zero = no state (an arena after a reset).
constexpr int kVaTimers = 0;        // 34 floats, RLGym pad order
constexpr int kVaAvail = 34;        // pads 0-31
constexpr int kVaAvailHi = 35;      // pads 32-33 in bits 0-1; bit 8 initialized, bit 9 has prediction
constexpr int kVaTick = 36;         // 2 words; ~0: never updated
constexpr int kVaPredPos = 38;      // 3 x 3 floats (uu)
constexpr int kVaPredVel = 47;      // 3 x 3 floats (uu/s)
// The last rollout's continuation (bit 10 of kVaAvailHi: present): the arena
// tick it started at, the ball it started from and the ball states it passed
// at ts, 120 + ts, 240 + ts and 360 ticks (ts: physics ticks between builds),
// each as position, velocity, angular velocity in engine units.
constexpr int kVaT0 = 56;           // 2 words
constexpr int kVaStart = 58;        // 9 floats
constexpr int kVaChain = 67;        // 4 x 9 floats
constexpr int kVaWords = 103;
constexpr uint32_t kVaChainsBit = 1u << 10;
// per-car words
constexpr int kVcTimers = 0;        // boost, boost hold, jump, flip, air, handbrake, demo
constexpr int kVcFlags = 7;         // bits: jumping, jumped, double jumped, flipped, on ground, was demoed,
                                    // initialized, previous jump / boost / handbrake
constexpr int kVcFlip = 8;          // flip direction x, y
constexpr int kVcPrev = 10;         // previous-previous analog controls (5)
constexpr int kVcTick = 15;         // 2 words
constexpr int kVcWords = 17;

template <int MAXC>
struct VLayout {
    static constexpr int kWords = kVaWords + MAXC * kVcWords;
    static constexpr int car(int i) { return kVaWords + i * kVcWords; }
};

constexpr uint64_t kNever = ~0ull;

struct Timers {   // GpuVantaTimers
    float boost_time, boost_hold, jump_time, flip_time, air_time, handbrake_time, demo_timer;
    bool is_jumping, has_jumped, has_double_jumped, has_flipped, on_ground, was_demoed, init;
    float flip_x, flip_y;
    Controls pp;  // the controls of the previous update
    uint64_t last_tick;
};

template <int MAXC>
struct State {    // GpuVantaArenaState + the arena's timers
    float timers[kPads];
    uint64_t avail;   // bit i: pad slot i available
    bool init, has_pred;
    uint64_t last_tick;
    V3 pred_pos[3], pred_vel[3];
    Timers car[MAXC];
};

// Pad tables in RLGym order: each slot's arena pad, its location (uu; z 73
// for big pads, 70 for small) and the slot of its x-mirrored pad.
struct PadTables {
    int8_t map[kPads];
    int8_t mirror[kPads];
    V3 loc[kPads];
};

struct Params {
    float dt;              // time between updates: the env's tick skip / 120
    float deadzone;        // dodge deadzone of the flip detection
    int only_closest_opp;  // only the closest opponent's packet is filled
    int predict;           // ball-prediction slices on
    int ticks_between;     // physics ticks between two env steps (the rollout continuation's step)
};

template <int MAXC>
RSX_HD void load(const Col& c, State<MAXC>& s) {
    for (int i = 0; i < kPads; i++) s.timers[i] = c.getf(kVaTimers + i);
    const uint32_t hi = c.get(kVaAvailHi);
    s.avail = (uint64_t)c.get(kVaAvail) | ((uint64_t)(hi & 3u) << 32);
    s.init = (hi >> 8) & 1;
    s.has_pred = (hi >> 9) & 1;
    s.last_tick = c.get64(kVaTick);
    for (int k = 0; k < 3; k++) {
        s.pred_pos[k] = c.getv(kVaPredPos + 3 * k);
        s.pred_vel[k] = c.getv(kVaPredVel + 3 * k);
    }
    for (int i = 0; i < MAXC; i++) {
        const int b = VLayout<MAXC>::car(i);
        Timers& t = s.car[i];
        t.boost_time = c.getf(b + 0);
        t.boost_hold = c.getf(b + 1);
        t.jump_time = c.getf(b + 2);
        t.flip_time = c.getf(b + 3);
        t.air_time = c.getf(b + 4);
        t.handbrake_time = c.getf(b + 5);
        t.demo_timer = c.getf(b + 6);
        const uint32_t f = c.get(b + kVcFlags);
        t.is_jumping = f & 1;
        t.has_jumped = (f >> 1) & 1;
        t.has_double_jumped = (f >> 2) & 1;
        t.has_flipped = (f >> 3) & 1;
        t.on_ground = (f >> 4) & 1;
        t.was_demoed = (f >> 5) & 1;
        t.init = (f >> 6) & 1;
        t.pp.jump = (f >> 7) & 1;
        t.pp.boost = (f >> 8) & 1;
        t.pp.handbrake = (f >> 9) & 1;
        t.flip_x = c.getf(b + kVcFlip);
        t.flip_y = c.getf(b + kVcFlip + 1);
        t.pp.throttle = c.getf(b + kVcPrev);
        t.pp.steer = c.getf(b + kVcPrev + 1);
        t.pp.pitch = c.getf(b + kVcPrev + 2);
        t.pp.yaw = c.getf(b + kVcPrev + 3);
        t.pp.roll = c.getf(b + kVcPrev + 4);
        t.last_tick = c.get64(b + kVcTick);
    }
}

template <int MAXC>
RSX_HD void store(const Col& c, const State<MAXC>& s, int ncars) {
    for (int i = 0; i < kPads; i++) c.setf(kVaTimers + i, s.timers[i]);
    c.set(kVaAvail, (uint32_t)s.avail);
    c.set(kVaAvailHi, (uint32_t)((s.avail >> 32) & 3u) | ((uint32_t)s.init << 8) | ((uint32_t)s.has_pred << 9) |
                          (c.get(kVaAvailHi) & kVaChainsBit));
    c.set64(kVaTick, s.last_tick);
    for (int k = 0; k < 3; k++) {
        c.setv(kVaPredPos + 3 * k, s.pred_pos[k]);
        c.setv(kVaPredVel + 3 * k, s.pred_vel[k]);
    }
    for (int i = 0; i < ncars; i++) {
        const int b = VLayout<MAXC>::car(i);
        const Timers& t = s.car[i];
        c.setf(b + 0, t.boost_time);
        c.setf(b + 1, t.boost_hold);
        c.setf(b + 2, t.jump_time);
        c.setf(b + 3, t.flip_time);
        c.setf(b + 4, t.air_time);
        c.setf(b + 5, t.handbrake_time);
        c.setf(b + 6, t.demo_timer);
        uint32_t f = (uint32_t)t.is_jumping | ((uint32_t)t.has_jumped << 1) | ((uint32_t)t.has_double_jumped << 2) |
                     ((uint32_t)t.has_flipped << 3) | ((uint32_t)t.on_ground << 4) | ((uint32_t)t.was_demoed << 5) |
                     ((uint32_t)t.init << 6) | ((uint32_t)t.pp.jump << 7) | ((uint32_t)t.pp.boost << 8) |
                     ((uint32_t)t.pp.handbrake << 9);
        c.set(b + kVcFlags, f);
        c.setf(b + kVcFlip, t.flip_x);
        c.setf(b + kVcFlip + 1, t.flip_y);
        c.setf(b + kVcPrev, t.pp.throttle);
        c.setf(b + kVcPrev + 1, t.pp.steer);
        c.setf(b + kVcPrev + 2, t.pp.pitch);
        c.setf(b + kVcPrev + 3, t.pp.yaw);
        c.setf(b + kVcPrev + 4, t.pp.roll);
        c.set64(b + kVcTick, t.last_tick);
    }
}

template <int MAXC>
RSX_HD void clear(const Col& c) {
    for (int w = 0; w < VLayout<MAXC>::kWords; w++) c.set(w, 0u);
}

// A body's pose in uu (VantaPhys): position, velocity, angular velocity and
// the basis vectors forward, right, up.
struct Phys {
    V3 pos, vel, ang, f, r, u;
};

RSX_HD Phys phys_of(const Car& c) {
    return Phys{c.xf.o * kBT2UU, c.vel * kBT2UU, c.ang, col(c.xf.b, 0), col(c.xf.b, 1), col(c.xf.b, 2)};
}

RSX_HD Phys phys_of_ball(V3 pos, V3 vel, V3 ang) {
    return Phys{pos, vel, ang, V3{1, 0, 0}, V3{0, 1, 0}, V3{0, 0, 1}};
}

// InvertPhys: every vector times (-1, -1, 1).
RSX_HD Phys invert(const Phys& p, bool inv) {
    if (!inv) return p;
    const V3 k = V3{-1.f, -1.f, 1.f};
    return Phys{p.pos * k, p.vel * k, p.ang * k, p.f * k, p.r * k, p.u * k};
}

// MirrorPhysX: x of positions, velocities and basis vectors negated; y and
// z of the angular velocity negated.
RSX_HD Phys mirror_x(Phys p) {
    p.pos.x = -p.pos.x;
    p.vel.x = -p.vel.x;
    p.ang.y = -p.ang.y;
    p.ang.z = -p.ang.z;
    p.f.x = -p.f.x;
    p.r.x = -p.r.x;
    p.u.x = -p.u.x;
    return p;
}

RSX_HD Phys canonical(const Phys& w, bool inv, bool mx) {
    Phys p = invert(w, inv);
    return mx ? mirror_x(p) : p;
}

// rotMat.Dot(v): v in the (forward, right, up) frame.
RSX_HD V3 rot_dot(const Phys& p, V3 v) { return V3{dot(p.f, v), dot(p.r, v), dot(p.u, v)}; }

RSX_HD V3 safe_norm(V3 v) {
    const float l = len(v);
    return l > 1e-6f ? v / l : V3{0, 0, 0};
}

RSX_HD bool flip_reset(const Car& c) { return !c.on_ground && has_flip_or_jump(c) && !c.has_jumped; }
RSX_HD bool got_flip_reset(const Car& c) { return !c.on_ground && !c.has_jumped; }

// Small-pad influence (closest available pad) and density around `pos`.
template <int MAXC>
RSX_HD void small_pad_stats(const PadTables& P, const State<MAXC>& s, V3 pos, float& max_inf, float& density) {
    float mi = 0.f, de = 0.f;
    for (int i = 0; i < kPads; i++) {
        if (P.loc[i].z < 72.f && ((s.avail >> i) & 1)) {
            const float dist = len(P.loc[i] - pos);
            const float inf = 1.f / (1.f + dist / 500.f);
            mi = fmaxf(mi, inf);
            if (dist < 2000.f) de += (1.f - dist / 2000.f);
        }
    }
    max_inf = mi;
    density = de / 4.f;
}

RSX_HD void init_timers(Timers& t, const Car& c) {
    t = Timers{};
    t.on_ground = c.on_ground;
    t.last_tick = kNever;
    t.init = true;
}

// UpdateBoostPads + UpdateDemoTimers (vanta_update_arena), once per tick count.
template <int MAXC>
RSX_HD void update_arena(State<MAXC>& s, const Arena<MAXC>& A, const PadTables& P, float dt) {
    if (!s.init) {
        s.avail = (1ull << kPads) - 1;
        for (int i = 0; i < kPads; i++) s.timers[i] = 0.f;
        s.init = true;
        s.last_tick = kNever;
    }
    if (s.last_tick == A.tick) return;
    s.last_tick = A.tick;
    for (int i = 0; i < kPads; i++) {
        const bool active = A.pad[P.map[i]].is_active;
        const bool avail = (s.avail >> i) & 1;
        if (active != avail) {
            s.avail ^= 1ull << i;
            s.timers[i] = active ? 0.f : (P.loc[i].z > 72.f ? 10.f : 4.f);
        } else if (!avail) {
            s.timers[i] = fmaxf(0.f, s.timers[i] - dt);
        }
    }
    for (int c = 0; c < A.ncars; c++) {
        Timers& t = s.car[c];
        if (!t.init) init_timers(t, A.car[c]);
        if (A.car[c].is_demoed) {
            if (!t.was_demoed) {
                t.demo_timer = 3.f;
                t.was_demoed = true;
            }
            t.demo_timer = fmaxf(0.f, t.demo_timer - dt);
        } else {
            t.demo_timer = 0.f;
            t.was_demoed = false;
        }
    }
}

// The coyote timers driven by the car's last controls (UpdateAdditionalTimers).
RSX_HD void update_player_timers(Timers& t, const Car& c, uint64_t tick, float dt, float deadzone) {
    if (!t.init) init_timers(t, c);
    if (t.last_tick == tick) return;
    t.last_tick = tick;
    const Controls& prev = c.last_controls;
    if (prev.boost) t.boost_time = fminf(t.boost_time + dt * 120.f, 12.f);
    else if (t.boost_time >= 12.f) t.boost_time = 0.f;
    if (prev.boost) t.boost_hold = fminf(t.boost_hold + dt * 120.f, 12.f);
    else t.boost_hold = fmaxf(t.boost_hold - 2.f * dt * 120.f, 0.f);
    if (t.on_ground && !t.is_jumping) t.has_jumped = false;
    const bool held = prev.jump, released_before = !t.pp.jump;
    if (t.is_jumping) t.is_jumping = t.jump_time < 3.f || (held && t.jump_time < 24.f);
    else if (held && released_before && t.on_ground) {
        t.is_jumping = true;
        t.jump_time = 0.f;
    }
    if (t.is_jumping) {
        t.has_jumped = true;
        t.jump_time = fminf(t.jump_time + dt * 120.f, 24.f);
    } else {
        t.jump_time = 0.f;
    }
    if (c.on_ground) {
        t.is_jumping = t.has_jumped = t.has_double_jumped = t.has_flipped = false;
        t.air_time = t.flip_time = 0.f;
        t.flip_x = t.flip_y = 0.f;
        t.on_ground = true;
    } else {
        t.air_time = (t.has_jumped && !t.is_jumping) ? fminf(t.air_time + dt * 120.f, 150.f) : 0.f;
        if (held && released_before && (!t.has_jumped || t.air_time < 150.f) && !t.has_double_jumped &&
            !t.has_flipped) {
            const float ab = fmaxf(fmaxf(fabsf(prev.pitch), fabsf(prev.yaw)), fabsf(prev.roll));
            if (ab >= deadzone) {
                t.flip_time = 0.f;
                t.has_flipped = true;
                float fx = prev.yaw;
                if (fabsf(fx) < 1e-6f && fabsf(prev.roll) >= deadzone) fx = prev.roll;
                const float fy = -prev.pitch;
                const float n = sqrtf(fx * fx + fy * fy);
                if (n > 1e-6f) {
                    t.flip_x = fx / n;
                    t.flip_y = fy / n;
                } else {
                    t.flip_x = t.flip_y = 0.f;
                }
            } else {
                t.has_double_jumped = true;
            }
        }
        t.on_ground = false;
    }
    if (t.has_flipped) t.flip_time = fminf(t.flip_time + dt * 120.f, 78.f);
    if (prev.handbrake) t.handbrake_time = fminf(t.handbrake_time + 5.f * dt, 1.f);
    else t.handbrake_time = fmaxf(t.handbrake_time - 2.f * dt, 0.f);
    t.pp = prev;
}

// GameState.lastTouchCarID: the car with the latest valid hit tick, the last
// such car in slot order on ties; 0 when none.
template <int MAXC>
RSX_HD uint32_t last_touch_id(const Arena<MAXC>& A) {
    uint64_t best = 0;
    uint32_t who = 0;
    for (int i = 0; i < A.ncars; i++) {
        const Car& c = A.car[i];
        if (c.ball_hit_valid && c.ball_hit_tick != kNever && c.ball_hit_tick >= best) {
            best = c.ball_hit_tick;
            who = c.id;
        }
    }
    return who;
}

// The 39-float packet of another car.
template <int MAXC>
RSX_HD void car_packet(const Arena<MAXC>& A, const State<MAXC>& s, const PadTables& P, int j, const Phys& me,
                       const Phys& bc, bool inv, bool mx, float* o) {
    const Car& other = A.car[j];
    const Phys oc = canonical(phys_of(other), inv, mx);
    const V3 rp = oc.pos - me.pos, rv = oc.vel - me.vel;
    const V3 brp = rot_dot(me, rp), brv = rot_dot(me, rv), bav = rot_dot(me, oc.ang);
    int k = 0;
    o[k++] = oc.pos.x / kPosStd;
    o[k++] = oc.pos.y / kPosStd;
    o[k++] = oc.pos.z / kPosStd;
    o[k++] = oc.vel.x / kVelStd;
    o[k++] = oc.vel.y / kVelStd;
    o[k++] = oc.vel.z / kVelStd;
    o[k++] = oc.ang.x / kAngStd;
    o[k++] = oc.ang.y / kAngStd;
    o[k++] = oc.ang.z / kAngStd;
    o[k++] = brp.x / kPosStd;
    o[k++] = brp.y / kPosStd;
    o[k++] = brp.z / kPosStd;
    o[k++] = brv.x / kVelStd;
    o[k++] = brv.y / kVelStd;
    o[k++] = brv.z / kVelStd;
    o[k++] = bav.x / kAngStd;
    o[k++] = bav.y / kAngStd;
    o[k++] = bav.z / kAngStd;
    o[k++] = oc.f.x;
    o[k++] = oc.f.y;
    o[k++] = oc.f.z;
    o[k++] = oc.u.x;
    o[k++] = oc.u.y;
    o[k++] = oc.u.z;
    o[k++] = other.boost / 100.f;
    o[k++] = other.on_ground ? 1.f : 0.f;
    o[k++] = has_flip_or_jump(other) ? 1.f : 0.f;
    o[k++] = other.is_demoed ? 1.f : 0.f;
    o[k++] = other.is_supersonic ? 1.f : 0.f;
    o[k++] = len(rp) / kPosStd;
    o[k++] = len(oc.pos - bc.pos) / kPosStd;
    // LookupDemoTimer: by car id (ids are unique, so it is the car's own slot)
    float demo = 0.f;
    for (int i = 0; i < A.ncars; i++)
        if (A.car[i].id == other.id) {
            if (s.car[i].init) demo = s.car[i].demo_timer;
            break;
        }
    o[k++] = demo / kDemoStd;
    o[k++] = other.is_jumping ? 1.f : 0.f;
    o[k++] = other.is_flipping ? 1.f : 0.f;
    o[k++] = other.has_double_jumped ? 1.f : 0.f;
    o[k++] = len(oc.vel) / kVelStd;
    o[k++] = other.has_flipped ? 1.f : 0.f;
    float mi, de;
    small_pad_stats<MAXC>(P, s, other.xf.o * kBT2UU, mi, de);
    o[k++] = mi;
    o[k++] = de;
}

// The row of car `self` (vanta_build_obs); its coyote timers are updated
// first.  Writes kRow floats.
template <int MAXC>
RSX_HD void build_row(const Arena<MAXC>& A, State<MAXC>& s, const PadTables& P, const Params& Q, int self,
                      uint32_t last_touch, float* o) {
    const Car& pl = A.car[self];
    update_player_timers(s.car[self], pl, A.tick, Q.dt, Q.deadzone);
    const Timers& t = s.car[self];
    const bool inv = pl.team == 1;
    const V3 wpos = pl.xf.o * kBT2UU;
    const bool mx = wpos.x < 0.f;
    const Phys me = canonical(phys_of(pl), inv, mx);
    const Phys bc = canonical(phys_of_ball(A.ball.xf.o * kBT2UU, A.ball.vel * kBT2UU, A.ball.ang), inv, mx);
    int k = 0;

    // ---- A. ego (81) ----
    const V3 pos_diff = bc.pos - me.pos, vel_diff = bc.vel - me.vel;
    const V3 brp = rot_dot(me, pos_diff), brv = rot_dot(me, vel_diff), avb = rot_dot(me, me.ang);
    o[k++] = me.pos.x / kPosStd;
    o[k++] = me.pos.y / kPosStd;
    o[k++] = me.pos.z / kPosStd;
    o[k++] = me.vel.x / kVelStd;
    o[k++] = me.vel.y / kVelStd;
    o[k++] = me.vel.z / kVelStd;
    o[k++] = me.ang.x / kAngStd;
    o[k++] = me.ang.y / kAngStd;
    o[k++] = me.ang.z / kAngStd;
    o[k++] = brp.x / kPosStd;
    o[k++] = brp.y / kPosStd;
    o[k++] = brp.z / kPosStd;
    o[k++] = brv.x / kBallVelStd;
    o[k++] = brv.y / kBallVelStd;
    o[k++] = brv.z / kBallVelStd;
    o[k++] = me.f.x;
    o[k++] = me.f.y;
    o[k++] = me.f.z;
    o[k++] = me.u.x;
    o[k++] = me.u.y;
    o[k++] = me.u.z;
    o[k++] = avb.x / kAngStd;
    o[k++] = avb.y / kAngStd;
    o[k++] = avb.z / kAngStd;
    o[k++] = len(me.vel) / kVelStd;
    o[k++] = pl.boost / 100.f;
    o[k++] = pl.on_ground ? 1.f : 0.f;
    o[k++] = has_flip_or_jump(pl) ? 1.f : 0.f;
    o[k++] = pl.is_demoed ? 1.f : 0.f;
    o[k++] = pl.has_jumped ? 1.f : 0.f;
    o[k++] = pl.is_supersonic ? 1.f : 0.f;
    {
        float w = 0.f;
        for (int i = 0; i < 4; i++) w += pl.wheel_contact[i] ? 1.f : 0.f;
        o[k++] = w / 4.f;
    }
    o[k++] = flip_reset(pl) ? 1.f : 0.f;
    o[k++] = got_flip_reset(pl) ? 1.f : 0.f;
    o[k++] = pl.handbrake_val;
    o[k++] = mx ? 1.f : 0.f;
    o[k++] = inv ? 1.f : 0.f;
    {   // threat frame
        const V3 own = V3{0.f, -5120.f, 0.f};
        const V3 to_goal = own - bc.pos;
        const float l = len(to_goal);
        const V3 danger = l > 1.f ? to_goal / l : V3{0.f, -1.f, 0.f};
        V3 lateral = V3{-danger.y, danger.x, 0.f};
        const float ll = len(lateral);
        if (ll > 1e-6f) lateral = lateral / ll;
        const V3 car_to_ball = bc.pos - me.pos;
        o[k++] = dot(bc.vel, danger) / kBallVelStd;
        o[k++] = dot(me.vel, danger) / kVelStd;
        o[k++] = dot(me.vel, lateral) / kVelStd;
        o[k++] = dot(car_to_ball, danger) / kPosStd;
        o[k++] = dot(car_to_ball, lateral) / kPosStd;
        o[k++] = len(own - bc.pos) / (2.f * kPosStd);
    }
    {   // alignment
        const V3 attack = V3{0.f, 5120.f, 0.f};
        const V3 vn = safe_norm(me.vel), tb = safe_norm(bc.pos - me.pos), tg = safe_norm(attack - me.pos);
        o[k++] = dot(me.f, tb);
        o[k++] = dot(me.f, tg);
        o[k++] = dot(vn, tb);
        o[k++] = dot(me.u, V3{0.f, 0.f, 1.f});
    }
    {   // previous action, mirrored with x
        const Controls& pa = pl.last_controls;
        o[k++] = pa.throttle;
        o[k++] = mx ? -pa.steer : pa.steer;
        o[k++] = pa.pitch;
        o[k++] = mx ? -pa.yaw : pa.yaw;
        o[k++] = mx ? -pa.roll : pa.roll;
        o[k++] = pa.jump ? 1.f : 0.f;
        o[k++] = pa.boost ? 1.f : 0.f;
        o[k++] = pa.handbrake ? 1.f : 0.f;
    }
    o[k++] = (pl.world_contact && pl.world_contact_normal.z < 0.7f) ? 1.f : 0.f;
    o[k++] = t.demo_timer / kDemoStd;
    o[k++] = t.boost_time / 12.f;
    o[k++] = t.jump_time / 24.f;
    o[k++] = t.air_time / 150.f;
    o[k++] = t.flip_time / 78.f;
    o[k++] = t.handbrake_time;
    o[k++] = mx ? -t.flip_x : t.flip_x;
    o[k++] = t.flip_y;
    o[k++] = t.boost_hold / 12.f;
    o[k++] = t.is_jumping ? 1.f : 0.f;
    o[k++] = t.has_flipped ? 1.f : 0.f;
    o[k++] = t.has_double_jumped ? 1.f : 0.f;
    {   // enrichments
        const float boost_frac = pl.boost / 100.f;
        const float dist = len(pos_diff);
        const float closing = dist > 1e-3f ? (-dot(vel_diff, pos_diff) / dist) / kVelStd : 0.f;
        o[k++] = sqrtf(fmaxf(0.f, boost_frac));
        o[k++] = (bc.pos.z > me.pos.z + 200.f) ? 1.f : 0.f;
        o[k++] = closing;
        o[k++] = dist / kPosStd;
    }
    {   // surface contact
        const bool hc = pl.world_contact;
        V3 n = V3{0.f, 0.f, 0.f};
        if (hc) {
            n = pl.world_contact_normal;
            if (inv) {
                n.x = -n.x;
                n.y = -n.y;
            }
            if (mx) n.x = -n.x;
        }
        o[k++] = hc ? 1.f : 0.f;
        o[k++] = n.x;
        o[k++] = n.y;
        o[k++] = n.z;
    }
    o[k++] = pl.is_flipping ? 1.f : 0.f;
    o[k++] = 0.f;  // headroom slot
    o[k++] = invert(phys_of(pl), inv).pos.x < 0.f ? 1.f : 0.f;  // kickoff side, team frame only
    {
        float mi, de;
        small_pad_stats<MAXC>(P, s, wpos, mi, de);
        o[k++] = mi;
        o[k++] = de;
    }

    // ---- B. ball (29) ----
    o[k++] = bc.pos.x / kPosStd;
    o[k++] = bc.pos.y / kPosStd;
    o[k++] = bc.pos.z / kPosStd;
    o[k++] = bc.vel.x / kBallVelStd;
    o[k++] = bc.vel.y / kBallVelStd;
    o[k++] = bc.vel.z / kBallVelStd;
    o[k++] = bc.ang.x / kAngStd;
    o[k++] = bc.ang.y / kAngStd;
    o[k++] = bc.ang.z / kAngStd;
    o[k++] = len(bc.vel) / kBallVelStd;
    o[k++] = bc.pos.z <= 100.f ? 1.f : 0.f;
    for (int i = 0; i < 3; i++) {
        if (s.has_pred) {
            const Phys sl = canonical(phys_of_ball(s.pred_pos[i], s.pred_vel[i], V3{0, 0, 0}), inv, mx);
            o[k++] = sl.pos.x / kPosStd;
            o[k++] = sl.pos.y / kPosStd;
            o[k++] = sl.pos.z / kPosStd;
            o[k++] = sl.vel.x / kBallVelStd;
            o[k++] = sl.vel.y / kBallVelStd;
            o[k++] = sl.vel.z / kBallVelStd;
        } else {
            for (int q = 0; q < 6; q++) o[k++] = 0.f;
        }
    }

    // ---- C. goal context (16) ----
    {
        const V3 opp = V3{0.f, 5120.f, 0.f}, own = V3{0.f, -5120.f, 0.f};
        const V3 d[4] = {opp - me.pos, own - me.pos, opp - bc.pos, own - bc.pos};
        for (int q = 0; q < 4; q++) {
            const V3 v = q < 2 ? rot_dot(me, d[q]) : d[q];
            o[k++] = v.x * kGoalDistCoef;
            o[k++] = v.y * kGoalDistCoef;
            o[k++] = v.z * kGoalDistCoef;
            o[k++] = len(d[q]) * kGoalDistCoef;
        }
    }

    // ---- D. team context (6) ----
    {
        int lt = -1;
        if (last_touch != 0)
            for (int i = 0; i < A.ncars; i++)
                if (A.car[i].id == last_touch) {
                    lt = A.car[i].id == pl.id ? 0 : (A.car[i].team == pl.team ? 1 : 2);
                    break;
                }
        o[k++] = lt == 0 ? 1.f : 0.f;
        o[k++] = lt == 1 ? 1.f : 0.f;
        o[k++] = lt == 2 ? 1.f : 0.f;
        int blue = 0, orange = 0;
        for (int i = 0; i < A.ncars; i++) (A.car[i].team == 0 ? blue : orange)++;
        o[k++] = (inv ? orange : blue) / 3.f;
        o[k++] = (inv ? blue : orange) / 3.f;
        o[k++] = bc.pos.y < 0.f ? 1.f : 0.f;
    }

    // ---- E / F. other cars by canonical distance to the ball ----
    int mate[MAXC], opp[MAXC];
    float mate_d[MAXC], opp_d[MAXC];
    int nm = 0, no = 0;
    for (int i = 0; i < A.ncars; i++) {
        const Car& c = A.car[i];
        if (c.id == pl.id) continue;
        const float dd = c.is_demoed ? 3.402823466e+38f : len(canonical(phys_of(c), inv, mx).pos - bc.pos);
        if (c.team == pl.team) {
            mate[nm] = i;
            mate_d[nm++] = dd;
        } else {
            opp[no] = i;
            opp_d[no++] = dd;
        }
    }
    // insertion sorts (ties keep slot order)
    for (int i = 1; i < nm; i++) {
        const int id = mate[i];
        const float dv = mate_d[i];
        int j = i - 1;
        while (j >= 0 && mate_d[j] > dv) {
            mate[j + 1] = mate[j];
            mate_d[j + 1] = mate_d[j];
            j--;
        }
        mate[j + 1] = id;
        mate_d[j + 1] = dv;
    }
    for (int i = 1; i < no; i++) {
        const int id = opp[i];
        const float dv = opp_d[i];
        int j = i - 1;
        while (j >= 0 && opp_d[j] > dv) {
            opp[j + 1] = opp[j];
            opp_d[j + 1] = opp_d[j];
            j--;
        }
        opp[j + 1] = id;
        opp_d[j + 1] = dv;
    }
    for (int i = 0; i < kMaxMates; i++) {
        if (i < nm) car_packet<MAXC>(A, s, P, mate[i], me, bc, inv, mx, o + k);
        else
            for (int q = 0; q < kCar; q++) o[k + q] = 0.f;
        k += kCar;
    }
    const int opp_limit = Q.only_closest_opp ? 1 : kMaxOpps;
    for (int i = 0; i < kMaxOpps; i++) {
        if (i < opp_limit && i < no) car_packet<MAXC>(A, s, P, opp[i], me, bc, inv, mx, o + k);
        else
            for (int q = 0; q < kCar; q++) o[k + q] = 0.f;
        k += kCar;
    }
    // ---- G. validity mask (5) ----
    for (int i = 0; i < kMaxMates; i++) o[k++] = i < nm ? 1.f : 0.f;
    for (int i = 0; i < kMaxOpps; i++) o[k++] = (i < opp_limit && i < no) ? 1.f : 0.f;

    // ---- H. boost pads (34) ----
    for (int i = 0; i < kPads; i++) {
        int idx = inv ? (kPads - 1 - i) : i;
        if (mx) idx = P.mirror[idx];
        if ((s.avail >> idx) & 1) {
            o[k++] = 1.f;
        } else {
            const float tmax = P.loc[idx].z > 72.f ? 10.f : 4.f;
            o[k++] = 1.f / (1.f + s.timers[idx] / tmax);
        }
    }
}

// A ball of the ball-only rollout: position, velocity, angular velocity
// (engine units).  The ball-only tick is a function of these nine floats
// alone (the basis stays identity, forces and impulse caches are zero
// between ticks, the sleep state and the contact manifolds are rebuilt every
// tick), so two rollouts from equal balls pass through equal balls.
struct BallQ {
    V3 p, v, w;
};

RSX_HD BallQ ball_q(const Ball& b) { return BallQ{b.xf.o, b.vel, b.ang}; }

RSX_HD bool same_ball(const BallQ& a, const BallQ& b) {
    return f2u(a.p.x) == f2u(b.p.x) && f2u(a.p.y) == f2u(b.p.y) && f2u(a.p.z) == f2u(b.p.z) &&
           f2u(a.v.x) == f2u(b.v.x) && f2u(a.v.y) == f2u(b.v.y) && f2u(a.v.z) == f2u(b.v.z) &&
           f2u(a.w.x) == f2u(b.w.x) && f2u(a.w.y) == f2u(b.w.y) && f2u(a.w.z) == f2u(b.w.z);
}

RSX_HD BallQ get_q(const Col& c, int w) { return BallQ{c.getv(w), c.getv(w + 3), c.getv(w + 6)}; }
RSX_HD void set_q(const Col& c, int w, const BallQ& q) {
    c.setv(w, q.p);
    c.setv(w + 3, q.v);
    c.setv(w + 6, q.w);
}

// The ball-only arena's working state: P has no car slots, so the tick
// compiles to the ball's own path; W is its working set, T the tick context
// (no cull, no pooling: there are no cars).
struct Roller {
    const TickCtx& T;
    TickScratch<0>& W;
    Arena<0>& P;
    RSX_HD void load(const BallQ& q) {
        P.ncars = 0;
        P.tick = 0;
        P.rng = 0;
        P.ball.xf = Xf{m3_identity(), q.p};
        P.ball.interp = P.ball.xf;
        P.ball.vel = q.v;
        P.ball.ang = q.w;
        P.ball.force = V3{0, 0, 0};
        P.ball.vel_cache = V3{0, 0, 0};
        P.ball.sleeping = false;
        P.ball.bp_reg = -1;
    }
    RSX_HD void tick() { (void)arena_tick<0>(P, T, W); }
    RSX_HD BallQ q() const { return ball_q(P.ball); }
    RSX_HD BallQ advance(const BallQ& q, int ticks) {
        load(q);
        for (int t = 0; t < ticks; t++) tick();
        return this->q();
    }
};

// Whether rollouts of ts-tick steps keep a continuation (ts must leave its
// offsets inside the 360-tick rollout, in order).
RSX_HD bool chains_fit(int ts) { return ts >= 1 && ts <= 120; }

// Full 360-tick rollout from q0: the slices after 120, 240 and 360 ticks
// (uu, uu/s) and, when chains_fit(ts), the continuation states at ts,
// 120 + ts, 240 + ts and 360.
RSX_HD void rollout_full(Roller& R, const BallQ& q0, int ts, V3* pos, V3* vel, BallQ* chain) {
    R.load(q0);
    int slice = 0;
    for (int t = 1; t <= kPredTicks; t++) {
        R.tick();
        if (t == ts) chain[0] = R.q();
        if (t == 120 + ts) chain[1] = R.q();
        if (t == 240 + ts) chain[2] = R.q();
        if (t == 120 || t == 240 || t == 360) {
            pos[slice] = R.P.ball.xf.o * kBT2UU;
            vel[slice] = R.P.ball.vel * kBT2UU;
            slice++;
        }
    }
    chain[3] = R.q();
}

// The rollout from the ball the last one passed at ts ticks (chain[0]),
// from its continuation: the new slices are the states it passed at 120 + ts
// and 240 + ts and its 360-tick state advanced ts ticks; every continuation
// state advances ts ticks.  Equal to rollout_full from chain[0], bit for bit.
RSX_HD void rollout_continue(Roller& R, int ts, V3* pos, V3* vel, BallQ* chain) {
    pos[0] = chain[1].p * kBT2UU;
    vel[0] = chain[1].v * kBT2UU;
    pos[1] = chain[2].p * kBT2UU;
    vel[1] = chain[2].v * kBT2UU;
    for (int k = 0; k < 4; k++) chain[k] = R.advance(chain[k], ts);
    pos[2] = chain[3].p * kBT2UU;
    vel[2] = chain[3].v * kBT2UU;
}

// Ball-only rollout of `ball` (engine units): its position (uu) and
// velocity (uu/s) after 120, 240 and 360 ticks.
RSX_HD void predict_ball(const Ball& ball, const TickCtx& T, TickScratch<0>& W, Arena<0>& P, V3* pos, V3* vel) {
    Roller R{T, W, P};
    BallQ chain[4];
    rollout_full(R, ball_q(ball), 0, pos, vel, chain);
}

}  // namespace vanta

}  // namespace rsx
