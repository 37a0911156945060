// rsx_soa.h — the engine's structure-of-arrays state layout in HBM (ABI v1).
//
// All persistent per-arena state is stored as 32-bit words in one buffer laid
// out field-major: word w of arena a lives at base[w * N + a], so the 64 lanes
// of a wavefront (64 consecutive arenas) read each field as one coalesced
// 256-byte transaction.  Floats are stored by bit pattern; 64-bit tick counts
// take two words (low, high).
//
//   car slot c (0..MAXC-1):  words [c*kCarWords, (c+1)*kCarWords)
//   ball:                    kBallBase .. +kBallWords
//   boost pads:              kPadBase  .. +34*3  (active, cooldown, prev locked id)
//   arena record:            kArenaBase .. +kArenaWords
//
// Car words (Bullet units for pose/velocities, seconds for timers):
//   0-2 pos, 3-11 basis rows, 12-14 vel, 15-17 ang vel, 18-29 timers/scalars,
//   30-32 flip torque, 33-35 world contact normal, 36 flag bits (bit 19: the
//   body responds to contacts as of its last pre-tick), 37 id,
//   38 contact other id, 39-42 ball-hit ticks, 43-51 ball-hit vectors,
//   52-56 last controls (analog), 57-80 per-wheel carried values (4 x 6),
//   81 broadphase registration cell.
// Arena words: tick (2), RNG, car count, no-touch time, episode, goal flags,
//   ball broadphase registration cell.
#pragma once

#include "rsx_step.h"

namespace rsx {

constexpr int kCarWords = 84;
constexpr int kBallWords = 9;
constexpr int kPadWords = 3;
constexpr int kArenaWords = 8;

template <int MAXC>
struct Layout {
    static constexpr int kBallBase = MAXC * kCarWords;
    static constexpr int kPadBase = kBallBase + kBallWords;
    static constexpr int kArenaBase = kPadBase + kNumPads * kPadWords;
    static constexpr int kWords = kArenaBase + kArenaWords;
};

// Arena record words
constexpr int kArTickLo = 0, kArTickHi = 1, kArRng = 2, kArNcars = 3, kArNoTouch = 4, kArEpisode = 5, kArFlags = 6,
              kArBallReg = 7;

RSX_HD uint32_t f2u(float f) {
    union {
        float f;
        uint32_t u;
    } c;
    c.f = f;
    return c.u;
}
RSX_HD float u2f(uint32_t u) {
    union {
        float f;
        uint32_t u;
    } c;
    c.u = u;
    return c.f;
}

// Accessor over one arena's column in the SoA buffer.
struct Col {
    uint32_t* base;
    int64_t stride;
    RSX_HD uint32_t get(int w) const { return base[(int64_t)w * stride]; }
    RSX_HD void set(int w, uint32_t v) const { base[(int64_t)w * stride] = v; }
    RSX_HD float getf(int w) const { return u2f(get(w)); }
    RSX_HD void setf(int w, float v) const { set(w, f2u(v)); }
    RSX_HD V3 getv(int w) const { return V3{getf(w), getf(w + 1), getf(w + 2)}; }
    RSX_HD void setv(int w, V3 v) const {
        setf(w, v.x);
        setf(w + 1, v.y);
        setf(w + 2, v.z);
    }
    RSX_HD uint64_t get64(int w) const { return (uint64_t)get(w) | ((uint64_t)get(w + 1) << 32); }
    RSX_HD void set64(int w, uint64_t v) const {
        set(w, (uint32_t)v);
        set(w + 1, (uint32_t)(v >> 32));
    }
};

RSX_HD void unpack_car(const Col& c, int b, Car& k, const CarGeom& g) {
    k.xf.o = c.getv(b + 0);
    k.xf.b.r[0] = c.getv(b + 3);
    k.xf.b.r[1] = c.getv(b + 6);
    k.xf.b.r[2] = c.getv(b + 9);
    k.vel = c.getv(b + 12);
    k.ang = c.getv(b + 15);
    k.boost = c.getf(b + 18);
    k.time_spent_boosting = c.getf(b + 19);
    k.jump_time = c.getf(b + 20);
    k.flip_time = c.getf(b + 21);
    k.air_time = c.getf(b + 22);
    k.air_time_since_jump = c.getf(b + 23);
    k.supersonic_time = c.getf(b + 24);
    k.handbrake_val = c.getf(b + 25);
    k.auto_flip_timer = c.getf(b + 26);
    k.auto_flip_torque_scale = c.getf(b + 27);
    k.demo_respawn_timer = c.getf(b + 28);
    k.car_contact_cooldown = c.getf(b + 29);
    k.flip_rel_torque = c.getv(b + 30);
    k.world_contact_normal = c.getv(b + 33);
    uint32_t f = c.get(b + 36);
    k.on_ground = f & 1;
    k.has_jumped = (f >> 1) & 1;
    k.has_double_jumped = (f >> 2) & 1;
    k.has_flipped = (f >> 3) & 1;
    k.is_flipping = (f >> 4) & 1;
    k.is_jumping = (f >> 5) & 1;
    k.is_supersonic = (f >> 6) & 1;
    k.is_auto_flipping = (f >> 7) & 1;
    k.is_demoed = (f >> 8) & 1;
    k.world_contact = (f >> 9) & 1;
    k.ball_hit_valid = (f >> 10) & 1;
    for (int i = 0; i < 4; i++) k.wheel_contact[i] = (f >> (11 + i)) & 1;
    k.team = (f >> 15) & 1;
    k.last_controls.jump = (f >> 16) & 1;
    k.last_controls.boost = (f >> 17) & 1;
    k.last_controls.handbrake = (f >> 18) & 1;
    k.active = (f >> 19) & 1;
    k.id = c.get(b + 37);
    k.car_contact_other_id = c.get(b + 38);
    k.ball_hit_tick = c.get64(b + 39);
    k.ball_hit_extra_tick = c.get64(b + 41);
    k.ball_hit_rel_pos = c.getv(b + 43);
    k.ball_hit_ball_pos = c.getv(b + 46);
    k.ball_hit_extra_vel = c.getv(b + 49);
    k.last_controls.throttle = c.getf(b + 52);
    k.last_controls.steer = c.getf(b + 53);
    k.last_controls.pitch = c.getf(b + 54);
    k.last_controls.yaw = c.getf(b + 55);
    k.last_controls.roll = c.getf(b + 56);
    for (int w = 0; w < 4; w++) {
        Wheel& wh = k.wheel[w];
        int o = b + 57 + w * 6;
        wh.steer_angle = c.getf(o + 0);
        wh.engine_force = c.getf(o + 1);
        wh.brake = c.getf(o + 2);
        wh.lat_friction = c.getf(o + 3);
        wh.long_friction = c.getf(o + 4);
        wh.extra_pushback = c.getf(o + 5);
    }
    k.bp_reg = (int)c.get(b + 81);
    // derived
    car_update_inertia(k, g);
    k.force = k.torque = k.vel_cache = V3{0, 0, 0};
    k.interp = k.xf;
}

RSX_HD void pack_car(const Col& c, int b, const Car& k) {
    c.setv(b + 0, k.xf.o);
    c.setv(b + 3, k.xf.b.r[0]);
    c.setv(b + 6, k.xf.b.r[1]);
    c.setv(b + 9, k.xf.b.r[2]);
    c.setv(b + 12, k.vel);
    c.setv(b + 15, k.ang);
    c.setf(b + 18, k.boost);
    c.setf(b + 19, k.time_spent_boosting);
    c.setf(b + 20, k.jump_time);
    c.setf(b + 21, k.flip_time);
    c.setf(b + 22, k.air_time);
    c.setf(b + 23, k.air_time_since_jump);
    c.setf(b + 24, k.supersonic_time);
    c.setf(b + 25, k.handbrake_val);
    c.setf(b + 26, k.auto_flip_timer);
    c.setf(b + 27, k.auto_flip_torque_scale);
    c.setf(b + 28, k.demo_respawn_timer);
    c.setf(b + 29, k.car_contact_cooldown);
    c.setv(b + 30, k.flip_rel_torque);
    c.setv(b + 33, k.world_contact_normal);
    uint32_t f = 0;
    f |= (uint32_t)k.on_ground;
    f |= (uint32_t)k.has_jumped << 1;
    f |= (uint32_t)k.has_double_jumped << 2;
    f |= (uint32_t)k.has_flipped << 3;
    f |= (uint32_t)k.is_flipping << 4;
    f |= (uint32_t)k.is_jumping << 5;
    f |= (uint32_t)k.is_supersonic << 6;
    f |= (uint32_t)k.is_auto_flipping << 7;
    f |= (uint32_t)k.is_demoed << 8;
    f |= (uint32_t)k.world_contact << 9;
    f |= (uint32_t)k.ball_hit_valid << 10;
    for (int i = 0; i < 4; i++) f |= (uint32_t)k.wheel_contact[i] << (11 + i);
    f |= (uint32_t)(k.team & 1) << 15;
    f |= (uint32_t)k.last_controls.jump << 16;
    f |= (uint32_t)k.last_controls.boost << 17;
    f |= (uint32_t)k.last_controls.handbrake << 18;
    f |= (uint32_t)k.active << 19;
    c.set(b + 36, f);
    c.set(b + 37, k.id);
    c.set(b + 38, k.car_contact_other_id);
    c.set64(b + 39, k.ball_hit_tick);
    c.set64(b + 41, k.ball_hit_extra_tick);
    c.setv(b + 43, k.ball_hit_rel_pos);
    c.setv(b + 46, k.ball_hit_ball_pos);
    c.setv(b + 49, k.ball_hit_extra_vel);
    c.setf(b + 52, k.last_controls.throttle);
    c.setf(b + 53, k.last_controls.steer);
    c.setf(b + 54, k.last_controls.pitch);
    c.setf(b + 55, k.last_controls.yaw);
    c.setf(b + 56, k.last_controls.roll);
    for (int w = 0; w < 4; w++) {
        const Wheel& wh = k.wheel[w];
        int o = b + 57 + w * 6;
        c.setf(o + 0, wh.steer_angle);
        c.setf(o + 1, wh.engine_force);
        c.setf(o + 2, wh.brake);
        c.setf(o + 3, wh.lat_friction);
        c.setf(o + 4, wh.long_friction);
        c.setf(o + 5, wh.extra_pushback);
    }
    c.set(b + 81, (uint32_t)k.bp_reg);
}

template <int MAXC>
RSX_HD void unpack_arena(const Col& c, Arena<MAXC>& A, const CarGeom& g) {
    using L = Layout<MAXC>;
    A.tick = c.get64(L::kArenaBase + kArTickLo);
    A.rng = c.get(L::kArenaBase + kArRng);
    A.ncars = (int)c.get(L::kArenaBase + kArNcars);
    for (int i = 0; i < MAXC; i++)
        if (i < A.ncars) unpack_car(c, i * kCarWords, A.car[i], g);
    A.ball.xf.b = m3_identity();
    A.ball.xf.o = c.getv(L::kBallBase + 0);
    A.ball.vel = c.getv(L::kBallBase + 3);
    A.ball.ang = c.getv(L::kBallBase + 6);
    A.ball.force = A.ball.vel_cache = V3{0, 0, 0};
    A.ball.interp = A.ball.xf;
    A.ball.bp_reg = (int)c.get(L::kArenaBase + kArBallReg);
    for (int p = 0; p < kNumPads; p++) {
        int o = L::kPadBase + p * kPadWords;
        A.pad[p].is_active = c.get(o) != 0;
        A.pad[p].cooldown = c.getf(o + 1);
        A.pad[p].prev_locked = c.get(o + 2);
        A.pad[p].cur_locked = 0;
    }
}

template <int MAXC>
RSX_HD void pack_arena(const Col& c, const Arena<MAXC>& A) {
    using L = Layout<MAXC>;
    c.set64(L::kArenaBase + kArTickLo, A.tick);
    c.set(L::kArenaBase + kArRng, A.rng);
    for (int i = 0; i < MAXC; i++)
        if (i < A.ncars) pack_car(c, i * kCarWords, A.car[i]);
    c.setv(L::kBallBase + 0, A.ball.xf.o);
    c.setv(L::kBallBase + 3, A.ball.vel);
    c.setv(L::kBallBase + 6, A.ball.ang);
    c.set(L::kArenaBase + kArBallReg, (uint32_t)A.ball.bp_reg);
    for (int p = 0; p < kNumPads; p++) {
        int o = L::kPadBase + p * kPadWords;
        c.set(o, A.pad[p].is_active ? 1u : 0u);
        c.setf(o + 1, A.pad[p].cooldown);
        c.set(o + 2, A.pad[p].prev_locked);
    }
}

}  // namespace rsx
