// rsx_train.h — the GPU training path's configurable rewards and its reset
// curriculum (host and device), evaluated per arena inside the env step.
//
// Rewards: a table of up to 16 weighted terms over the 28 reward IDs of the
// reference's GPU training path (TrainingRewardID, RocketSimCuda/include/
// RocketSimCuda.h:198-252; per-ID definitions training_eval_reward,
// src/RocketSimCudaRewards.cuh:36-287; table sum with the zero-sum /
// team-spirit transform, buildRewardsAndTerminalsKernel :290-373).  Each term
// reads the state after the env step and a snapshot of the state before it
// (SnapshotTrainingState, RocketSimCuda.h:333), all in uu.  The IDs that
// approximate RLGym reward classes follow those classes where the GPU code
// does (GoalReward, VelocityBallToGoal, VelocityPlayerToBall, FaceBall,
// TouchBall, TouchAccel, StrongTouch, Speed, Velocity, Air, Wavedash,
// PickupBoost, SaveBoost, Bump, BumpedPenalty, Demo, DemoedPenalty,
// KickoffProximity, TeammateBumpPenalty: RLGymCPP Rewards/CommonRewards.h,
// GigaLearnCPP's reward headers); the reserved team IDs score 0, as there.
//
// Curriculum: RocketSimCudaBatch::SetResetCurriculum / ResetArena
// (src/RocketSimCudaApi.cuh:203-418): each reset picks kickoff, fuzzed
// kickoff, aerial or ball chase with the given relative weights.  The draws
// come from the arena's own RNG stream (the reference draws from one host
// LCG in arena order), so resets follow the same distributions, not the same
// numbers.
#pragma once

#include "rsx_env.h"

namespace rsx {

constexpr int kMaxRewardEntries = 16;
constexpr int kRewardParams = 10;

// TrainingRewardID (RocketSimCuda.h:200-231)
enum RewardId : int32_t {
    kRwUnknown = 0,
    kRwGoal,
    kRwVelBallToGoal,
    kRwVelPlayerToBall,
    kRwFaceBall,
    kRwTouchBall,
    kRwTouchAccel,
    kRwStrongTouch,
    kRwSpeed,
    kRwVelocity,
    kRwAir,
    kRwWavedash,
    kRwPickupBoost,
    kRwSaveBoost,
    kRwBump,
    kRwBumpedPenalty,
    kRwDemo,
    kRwDemoedPenalty,
    kRwKickoffProximity,
    kRwTeammateBumpPenalty,
    kRwPassIncentive,
    kRwTeamPassing,
    kRwMixedInitiate,
    kRwMixedDoubletap,
    kRwSpeedAerial,
    kRwPressureFlick,
    kRwMawkzyFlick,
    kRwAirCarry,
    kRwAntiBallStack,
    kRwCount
};

struct RewardEntry {
    int32_t id;
    float weight;
    float params[kRewardParams];
    int32_t zero_sum;
    float team_spirit;
    float opp_scale;
};

// The state a car's rewards compare against: the car at the start of the env
// step (uu, uu/s).
struct PrevCar {
    V3 pos, vel;
    float boost, contact_cd;
    uint32_t contact_id;
    bool on_ground, is_flipping, is_demoed;
};

template <int MAXC>
struct PrevSnap {
    PrevCar car[MAXC];
    V3 ball_pos, ball_vel;
};

template <int MAXC>
RSX_HD void snapshot_prev(const Arena<MAXC>& A, PrevSnap<MAXC>& P) {
    for (int i = 0; i < A.ncars; i++) {
        const Car& c = A.car[i];
        PrevCar& p = P.car[i];
        p.pos = c.xf.o * kBT2UU;
        p.vel = c.vel * kBT2UU;
        p.boost = c.boost;
        p.contact_cd = c.car_contact_cooldown;
        p.contact_id = c.car_contact_other_id;
        p.on_ground = c.on_ground;
        p.is_flipping = c.is_flipping;
        p.is_demoed = c.is_demoed;
    }
    P.ball_pos = A.ball.xf.o * kBT2UU;
    P.ball_vel = A.ball.vel * kBT2UU;
}

// v3_safe_normalize of the GPU training code: zero below |v|^2 = 1e-10.
RSX_HD V3 train_safe_norm(V3 v) {
    const float l2 = v.x * v.x + v.y * v.y + v.z * v.z;
    if (l2 < 1e-10f) return V3{0, 0, 0};
    return v * (1.f / sqrtf(l2));
}
RSX_HD float train_len(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }
RSX_HD float train_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RSX_HD float train_dist2(V3 a, V3 b) {
    const V3 d = a - b;
    return d.x * d.x + d.y * d.y + d.z * d.z;
}

// A bump that began during the step (training_is_new_contact,
// RocketSimCudaRewards.cuh:26-37).
RSX_HD bool new_contact(uint32_t id, float cd, const PrevCar& p) {
    if (!id || cd <= 0.f) return false;
    return p.contact_id != id || p.contact_cd <= 0.f || p.contact_cd < cd;
}

template <int MAXC>
RSX_HD int car_index_by_id(const Arena<MAXC>& A, uint32_t id) {
    for (int i = 0; i < A.ncars; i++)
        if (A.car[i].id == id) return i;
    return -1;
}

// One reward term for car ci (training_eval_reward, RocketSimCudaRewards.cuh:36-287).
// goal_team: -1 no goal this episode, else the scoring team.
template <int MAXC>
RSX_HD float train_term(const RewardEntry& e, const Arena<MAXC>& A, const PrevSnap<MAXC>& P, int ci, int goal_team,
                        bool touched) {
    const Car& c = A.car[ci];
    const PrevCar& pc = P.car[ci];
    const V3 cpos = c.xf.o * kBT2UU, cvel = c.vel * kBT2UU;
    const V3 bpos = A.ball.xf.o * kBT2UU, bvel = A.ball.vel * kBT2UU;
    switch (e.id) {
    case kRwGoal:
        if (goal_team < 0) return 0.f;
        return (int)c.team == goal_team ? 1.f : e.params[0];
    case kRwVelBallToGoal: {
        bool orange_goal = c.team == 0;
        if (e.params[0] > 0.5f) orange_goal = !orange_goal;
        const float ty = orange_goal ? 6000.f : -6000.f;
        const V3 dir = train_safe_norm(V3{-bpos.x, ty - bpos.y, 321.3875f - bpos.z});
        return train_dot(dir, bvel * (1.f / kRLBallMaxSpeed));
    }
    case kRwVelPlayerToBall:
        return train_dot(train_safe_norm(bpos - cpos), cvel * (1.f / kRLCarMaxSpeed));
    case kRwFaceBall:
        return train_dot(col(c.xf.b, 0), train_safe_norm(bpos - cpos));
    case kRwTouchBall:
        return touched ? 1.f : 0.f;
    case kRwTouchAccel: {
        if (!touched) return 0.f;
        const float max_speed = 110.f * (100000.f / 36000.f);
        const float pf = fminf(1.f, train_len(P.ball_vel) / max_speed);
        const float cf = fminf(1.f, train_len(bvel) / max_speed);
        return cf > pf ? cf - pf : 0.f;
    }
    case kRwStrongTouch: {
        if (!touched) return 0.f;
        const float force = train_len(bvel - P.ball_vel);
        if (force < e.params[0]) return 0.f;
        return fminf(1.f, force / fmaxf(e.params[1], 1e-6f));
    }
    case kRwSpeed:
        return train_len(cvel) / kRLCarMaxSpeed;
    case kRwVelocity:
        return (train_len(cvel) / kRLCarMaxSpeed) * e.params[0];
    case kRwAir:
        return c.on_ground ? 0.f : 1.f;
    case kRwWavedash:
        return (c.on_ground && pc.is_flipping && !pc.on_ground) ? 1.f : 0.f;
    case kRwPickupBoost:
        if (c.boost > pc.boost) return sqrtf(c.boost / 100.f) - sqrtf(pc.boost / 100.f);
        return 0.f;
    case kRwSaveBoost:
        return clampf(powf(c.boost / 100.f, e.params[0]), 0.f, 1.f);
    case kRwBump: {
        if (!new_contact(c.car_contact_other_id, c.car_contact_cooldown, pc)) return 0.f;
        const int o = car_index_by_id(A, c.car_contact_other_id);
        return (o < 0 || A.car[o].team == c.team) ? 0.f : 1.f;
    }
    case kRwBumpedPenalty:
        for (int o = 0; o < A.ncars; o++) {
            if (o == ci || A.car[o].team == c.team) continue;
            if (!new_contact(A.car[o].car_contact_other_id, A.car[o].car_contact_cooldown, P.car[o])) continue;
            if (A.car[o].car_contact_other_id == c.id) return -1.f;
        }
        return 0.f;
    case kRwDemo: {
        if (!new_contact(c.car_contact_other_id, c.car_contact_cooldown, pc)) return 0.f;
        const int o = car_index_by_id(A, c.car_contact_other_id);
        if (o < 0 || A.car[o].team == c.team) return 0.f;
        return (A.car[o].is_demoed && !P.car[o].is_demoed) ? 1.f : 0.f;
    }
    case kRwDemoedPenalty:
        return (c.is_demoed && !pc.is_demoed) ? -1.f : 0.f;
    case kRwKickoffProximity: {
        if (fabsf(bpos.x) >= 1.f || fabsf(bpos.y) >= 1.f) return 0.f;
        const float mine = train_dist2(cpos, bpos);
        float opp = 3.402823466e+38f;
        bool found = false;
        for (int o = 0; o < A.ncars; o++) {
            if (A.car[o].id == c.id || A.car[o].team == c.team) continue;
            found = true;
            opp = fminf(opp, train_dist2(A.car[o].xf.o * kBT2UU, bpos));
        }
        if (!found || mine < opp) return (found && mine == opp) ? 0.f : 1.f;
        return -1.f;
    }
    case kRwTeammateBumpPenalty: {
        const uint32_t id = c.car_contact_other_id;
        const float cd = c.car_contact_cooldown;
        if (cd <= 0.f || id == 0) return 0.f;
        const bool fresh = (cd > pc.contact_cd) || (id != pc.contact_id && pc.contact_cd <= 0.f);
        if (!fresh) return 0.f;
        for (int o = 0; o < A.ncars; o++)
            if (A.car[o].id == id) return A.car[o].team == c.team ? -1.f : 0.f;
        return 0.f;
    }
    case kRwPressureFlick: {
        if (!touched || !c.is_flipping) return 0.f;
        if (sqrtf(train_dist2(pc.pos, P.ball_pos)) > 250.f) return 0.f;
        if (fabsf(train_len(pc.vel) - train_len(P.ball_vel)) > 500.f) return 0.f;
        float closest = 1e9f;
        for (int o = 0; o < A.ncars; o++) {
            if (A.car[o].team == c.team || A.car[o].id == c.id) continue;
            closest = fminf(closest, sqrtf(train_dist2(cpos, A.car[o].xf.o * kBT2UU)));
        }
        if (closest > 700.f) return 0.f;
        const float ty = c.team == 0 ? 5120.f : -5120.f;
        const float vg = train_dot(bvel, train_safe_norm(V3{0.f, ty - bpos.y, 0.f}));
        if (vg <= 1000.f) return 0.f;
        const float ratio = vg / 2920.f;
        return fminf(ratio * ratio * sqrtf(fmaxf(ratio, 0.f)), 2.f);
    }
    case kRwMawkzyFlick: {
        if (!touched || !c.is_flipping) return 0.f;
        const float ty = c.team == 0 ? 5120.f : -5120.f;
        const float acc = train_dot(bvel - P.ball_vel, train_safe_norm(V3{0.f, ty - bpos.y, 0.f}));
        return clampf(acc / 1500.f, 0.f, 1.f);
    }
    case kRwAirCarry: {
        if (c.on_ground) return 0.f;
        if (touched) return 1.f;
        if (sqrtf(train_dist2(cpos, bpos)) > 300.f) return 0.f;
        return clampf(1.f - train_len(cvel - bvel) / 1200.f, 0.f, 0.35f);
    }
    case kRwAntiBallStack: {
        const float dz = fabsf(cpos.z - bpos.z);
        const float dx = cpos.x - bpos.x, dy = cpos.y - bpos.y;
        const float dxy = sqrtf(dx * dx + dy * dy);
        return (dxy < 120.f && cpos.z > bpos.z && dz < 180.f && c.on_ground) ? -1.f : 0.f;
    }
    default:  // UNKNOWN and the reserved team IDs (PassIncentive .. SpeedAerial)
        return 0.f;
    }
}

// Every car's reward from the table, in table order, each term through the
// zero-sum transform when the entry asks for it.
template <int MAXC>
RSX_HD void train_rewards(const Arena<MAXC>& A, const PrevSnap<MAXC>& P, const RewardEntry* table, int n,
                          int goal_team, uint64_t tick, int step_ticks, float* out) {
    float total[MAXC], raw[MAXC];
    bool touched[MAXC];
    for (int i = 0; i < A.ncars; i++) {
        total[i] = 0.f;
        const Car& c = A.car[i];
        touched[i] = c.ball_hit_valid && c.ball_hit_tick >= tick - (uint64_t)step_ticks;
    }
    for (int k = 0; k < n; k++) {
        const RewardEntry e = table[k];
        for (int i = 0; i < A.ncars; i++) raw[i] = train_term<MAXC>(e, A, P, i, goal_team, touched[i]);
        if (e.zero_sum) {
            for (int i = 0; i < A.ncars; i++) {
                float team_sum = 0.f, opp_sum = 0.f;
                int team_n = 0, opp_n = 0;
                for (int o = 0; o < A.ncars; o++) {
                    if (A.car[o].team == A.car[i].team) {
                        team_sum += raw[o];
                        team_n++;
                    } else {
                        opp_sum += raw[o];
                        opp_n++;
                    }
                }
                const float avg_team = team_n > 0 ? team_sum / (float)team_n : 0.f;
                const float avg_opp = opp_n > 0 ? opp_sum / (float)opp_n : 0.f;
                const float zs = raw[i] * (1.f - e.team_spirit) + avg_team * e.team_spirit - avg_opp * e.opp_scale;
                total[i] += zs * e.weight;
            }
        } else {
            for (int i = 0; i < A.ncars; i++) total[i] += raw[i] * e.weight;
        }
    }
    for (int i = 0; i < A.ncars; i++) out[i] = total[i];
}

// ---- reset curriculum ---------------------------------------------------------

constexpr int kResetKickoff = 0, kResetFuzzed = 1, kResetAerial = 2, kResetChase = 3;

// Uniform in [0, 1) from the arena's RNG (24 bits).
RSX_HD float rng_u01(uint32_t& s) { return (float)(rng_next(s) >> 8) * (1.f / 16777216.f); }
RSX_HD float rng_signed(uint32_t& s, float amp) { return (rng_u01(s) * 2.f - 1.f) * amp; }

// The reset mode for relative weights w (ResetArena's draw, RocketSimCudaApi.cuh:235-246).
// Kickoff-only weights draw nothing, so the default resets keep the RNG
// stream of the plain kickoff reset.
RSX_HD int pick_reset_mode(uint32_t& s, const float* w) {
    if (!(w[1] > 0.f) && !(w[2] > 0.f) && !(w[3] > 0.f)) return kResetKickoff;
    const float total = w[0] + w[1] + w[2] + w[3];
    if (!(total > 0.f)) return kResetKickoff;
    const float r = rng_u01(s) * total;
    if (r < w[0]) return kResetKickoff;
    if (r < w[0] + w[1]) return kResetFuzzed;
    if (r < w[0] + w[1] + w[2]) return kResetAerial;
    return kResetChase;
}

// A unit direction in the xy plane from two signed draws (ResetArena's
// chase and aerial placement).
RSX_HD void rng_dir2(uint32_t& s, float& dx, float& dy) {
    dx = rng_signed(s, 1.f);
    dy = rng_signed(s, 1.f);
    float l = sqrtf(dx * dx + dy * dy);
    if (l < 1e-3f) {
        dx = 1.f;
        dy = 0.f;
        l = 1.f;
    }
    dx /= l;
    dy /= l;
}

// A reset under the curriculum weights w (kickoff, fuzzed, aerial, chase):
// the kickoff reset, then the chosen mode's ball and car placement
// (RocketSimCudaApi.cuh:281-375).  Returns the mode.
template <int MAXC>
RSX_HD int arena_reset_curriculum(Arena<MAXC>& A, const CarGeom& g, const float* w) {
    const int mode = pick_reset_mode(A.rng, w);
    arena_reset_kickoff<MAXC>(A, g);
    if (mode == kResetKickoff) return mode;
    V3 bp = V3{0, 0, kBallRestZ}, bv = V3{0, 0, 0};
    if (mode == kResetFuzzed) {
        bp.x = rng_signed(A.rng, 180.f);
        bp.y = rng_signed(A.rng, 220.f);
        bv = V3{rng_signed(A.rng, 90.f), rng_signed(A.rng, 90.f), rng_signed(A.rng, 35.f)};
    } else if (mode == kResetAerial) {
        bp.x = rng_signed(A.rng, 1400.f);
        bp.y = rng_signed(A.rng, 1600.f);
        bp.z = 420.f + rng_u01(A.rng) * 360.f;
        bv = V3{rng_signed(A.rng, 100.f), rng_signed(A.rng, 100.f), rng_signed(A.rng, 50.f)};
    } else {
        bp.x = rng_signed(A.rng, 1900.f);
        bp.y = rng_signed(A.rng, 2200.f);
        bv = V3{rng_signed(A.rng, 500.f), rng_signed(A.rng, 500.f), rng_signed(A.rng, 80.f)};
    }
    A.ball.xf.o = bp * kUU2BT;
    A.ball.vel = bv * kUU2BT;
    A.ball.interp = A.ball.xf;
    const int setup = (mode == kResetAerial && A.ncars > 0) ? (int)(rng_u01(A.rng) * (float)A.ncars) % A.ncars : -1;
    for (int i = 0; i < A.ncars; i++) {
        Car& c = A.car[i];
        if (mode == kResetFuzzed) {
            // kickoff spot with a mild positional fuzz, same yaw
            V3 p = c.xf.o * kBT2UU;
            p.x += rng_signed(A.rng, 40.f);
            p.y += rng_signed(A.rng, 40.f);
            c.xf.o = p * kUU2BT;
        } else if (mode == kResetChase) {
            // on the ground near the ball, facing it, rolling towards it
            float dx, dy;
            rng_dir2(A.rng, dx, dy);
            const float dist = 450.f + rng_u01(A.rng) * 1400.f;
            const float x = clampf(bp.x + dx * dist, -3500.f, 3500.f);
            const float y = clampf(bp.y + dy * dist, -4500.f, 4500.f);
            const float yaw = atan2f(bp.y - y, bp.x - x);
            const float spd = rng_u01(A.rng) * 600.f;
            const float boost = 30.f + rng_u01(A.rng) * 70.f;
            car_set_spawn<MAXC>(c, g, x, y, kCarSpawnRestZ, yaw, boost);
            const float tx = bp.x - x, ty = bp.y - y;
            const float tl = sqrtf(tx * tx + ty * ty);
            if (tl > 1.f) c.vel = V3{(tx / tl) * spd, (ty / tl) * spd, 0.f} * kUU2BT;
        } else if (i == setup) {
            // the aerial setup car: below and behind the ball, in the air, moving at it
            float dx, dy;
            rng_dir2(A.rng, dx, dy);
            const float dist = 180.f + rng_u01(A.rng) * 160.f;
            const float x = bp.x - dx * dist, y = bp.y - dy * dist;
            float z = bp.z - (40.f + rng_u01(A.rng) * 120.f);
            if (z < kCarSpawnRestZ) z = kCarSpawnRestZ + 80.f;
            const float yaw = atan2f(bp.y - y, bp.x - x);
            const float spd = 500.f + rng_u01(A.rng) * 400.f;
            const float vz = 80.f + rng_u01(A.rng) * 120.f;
            const bool jumped = rng_u01(A.rng) < 0.55f;
            const float boost = 55.f + rng_u01(A.rng) * 45.f;
            car_set_spawn<MAXC>(c, g, x, y, z, yaw, boost);
            c.vel = V3{dx * spd, dy * spd, vz} * kUU2BT;
            c.on_ground = false;
            c.has_jumped = jumped;
        }
    }
    return mode;
}

// The no-touch timer after a reset: 0, or with the stagger on a uniform
// draw in [0, 0.9) x the timeout, so that arenas reset together do not all
// time out on the same step (ResetArena, RocketSimCudaApi.cuh:396-406).
RSX_HD float reset_no_touch(uint32_t& s, const EnvConfig& E) {
    if (!E.no_touch_stagger || !(E.no_touch_timeout > 0.f)) return 0.f;
    return rng_u01(s) * E.no_touch_timeout * 0.9f;
}

}  // namespace rsx
