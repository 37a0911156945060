// rsx_env.h — the env half of a step: DefaultObs rows, the four stub rewards,
// the terminal conditions and the auto-reset, evaluated per arena on the
// post-step state (host and device).
//
// Reference behaviour restated here (RLGymCPP, src/RLGymCPP/):
//   DefaultObs::BuildObs / AddPlayerToObs   ObsBuilders/DefaultObs.cpp:4-54, DefaultObs.h:11-17
//   InvertPhys                              Gamestates/StateUtil.cpp:3-17
//   GameState boost-pad order (+ inverted)  Gamestates/GameState.cpp:10-50, 96-109
//   Player::UpdateFromCar (touch in step)   Gamestates/Player.cpp:8-24
//   CarState::HasFlipOrJump                 RocketSim/src/Sim/Car/Car.cpp:285-289
//   GoalReward, VelocityBallToGoalReward,
//   VelocityPlayerToBallReward, TouchBall   RewardCore/CommonRewards.h:33-99
//   GoalScoreCondition, NoTouchCondition    TerminalConditions/*.h
//   reward sum / terminal priority / reset  EnvSet/EnvSet.cpp:169-242, 392-420
#pragma once

#include "rsx_step.h"

namespace rsx {

// RLGym CommonValues (CommonValues.h:12-40)
constexpr float kSideWallX = 4096.f, kBackWallY = 5120.f, kCeilingZ = 2044.f, kBackNetY = 6000.f;
constexpr float kGoalHeightRL = 642.775f;
constexpr float kRLCarMaxSpeed = 2300.f, kRLCarMaxAngVel = 5.5f, kRLBallMaxSpeed = 6000.f;
constexpr int kActionFloats = 8;
constexpr int kPlayerObs = 19;
constexpr int kObsHead = 9 + kActionFloats + kNumPads;  // ball, previous action, pads

RSX_HD int obs_row_len(int ncars) { return kObsHead + kPlayerObs * ncars; }

// Observation builders of the env step.
constexpr int kObsDefault = 0;   // RLGym DefaultObs: 51 + 19 per car
constexpr int kObsAdvanced = 1;  // GigaLearn AdvancedObs as the GPU path builds it: 51 + 29 per car
constexpr int kObsVanta = 2;     // VantaObs v3.5: 366 per car whatever the team size (rsx_vanta.h)
constexpr int kAdvPlayerObs = 29;
constexpr int kVantaRow = 366;
RSX_HD int obs_row_len_of(int kind, int ncars) {
    if (kind == kObsVanta) return kVantaRow;
    return kObsHead + (kind == kObsAdvanced ? kAdvPlayerObs : kPlayerObs) * ncars;
}

// Terminal codes (TerminalCondition.h TerminalType)
constexpr uint8_t kNotTerminal = 0, kTerminalNormal = 1, kTerminalTruncated = 2;

struct EnvConfig {
    float goal_w, concede_scale, touch_w, vbtg_w, vptb_w;
    float no_touch_timeout;   // seconds; <= 0 disables NoTouchCondition
    int goal_terminal;        // GoalScoreCondition enabled
    int auto_reset;           // kickoff reset of terminal arenas inside the step
    int obs_kind;             // kObsDefault or kObsAdvanced
    int8_t pad_map[kNumPads]; // obs pad slot i -> engine pad index (GameState boostPadIndexMap)
    float reset_w[4];         // reset curriculum weights: kickoff, fuzzed kickoff, aerial, ball chase
    int no_touch_stagger;     // a reset starts the no-touch timer at U(0, 0.9) x the timeout
};

struct StepOut {
    bool goal_scored;
    uint8_t terminal;
};

RSX_HD bool has_flip_or_jump(const Car& c) {
    return c.on_ground || (!c.has_flipped && !c.has_double_jumped && c.air_time_since_jump < kDoubleJumpMaxDelay);
}

constexpr int kNumDefaultActions = 90;

// DefaultAction::GetActionMask (DefaultAction.cpp:91-118) for one car:
// tables holds the ground, air, boost and jump masks of the default table.
RSX_HD void action_mask_row(const Car& c, const uint8_t* tables, uint8_t* out) {
    const uint8_t* base = c.on_ground ? tables : tables + kNumDefaultActions;
    const uint8_t* boost = tables + 2 * kNumDefaultActions;
    const uint8_t* jump = tables + 3 * kNumDefaultActions;
    const bool no_boost = c.boost == 0;
    const bool turtled = c.world_contact && c.world_contact_normal.z > 0.9f;
    const bool can_jump = has_flip_or_jump(c) || turtled;
    for (int i = 0; i < kNumDefaultActions; i++) {
        uint8_t m = base[i];
        if (no_boost && boost[i]) m = 0;
        if (can_jump && jump[i]) m = 1;
        out[i] = m;
    }
}

RSX_HD bool ball_scored_now(const Ball& b) {
    return fabsf(b.xf.o.y * kBT2UU) > (kGoalThresholdY + kBallRadius);
}

RSX_HD float rl_dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// One player's 19-float block (AddPlayerToObs), inverted for orange viewers.
template <int MAXC>
RSX_HD void obs_player(const Car& c, bool inv, float* o) {
    const float px = 1.f / kSideWallX, py = 1.f / kBackWallY, pz = 1.f / kCeilingZ;
    const float vc = 1.f / kRLCarMaxSpeed, ac = 1.f / kRLCarMaxAngVel;
    const float s = inv ? -1.f : 1.f;
    V3 pos = c.xf.o * kBT2UU, vel = c.vel * kBT2UU, ang = c.ang;
    V3 fwd = col(c.xf.b, 0), up = col(c.xf.b, 2);
    o[0] = pos.x * s * px;
    o[1] = pos.y * s * py;
    o[2] = pos.z * pz;
    o[3] = fwd.x * s;
    o[4] = fwd.y * s;
    o[5] = fwd.z;
    o[6] = up.x * s;
    o[7] = up.y * s;
    o[8] = up.z;
    o[9] = vel.x * s * vc;
    o[10] = vel.y * s * vc;
    o[11] = vel.z * vc;
    o[12] = ang.x * s * ac;
    o[13] = ang.y * s * ac;
    o[14] = ang.z * ac;
    o[15] = c.boost / 100.f;
    o[16] = c.on_ground ? 1.f : 0.f;
    o[17] = has_flip_or_jump(c) ? 1.f : 0.f;
    o[18] = c.is_demoed ? 1.f : 0.f;
}

// DefaultObs row of car `self`; prev holds each car's previous action
// (8 floats per car).  Writes obs_row_len(ncars) floats.
template <int MAXC>
RSX_HD void obs_row(const Arena<MAXC>& A, int self, const float (*prev)[kActionFloats], const EnvConfig& E, float* o) {
    const Car& me = A.car[self];
    const bool inv = me.team != 0;
    const float s = inv ? -1.f : 1.f;
    const float px = 1.f / kSideWallX, py = 1.f / kBackWallY, pz = 1.f / kCeilingZ;
    const float vc = 1.f / kRLCarMaxSpeed, ac = 1.f / kRLCarMaxAngVel;
    V3 bp = A.ball.xf.o * kBT2UU, bv = A.ball.vel * kBT2UU, ba = A.ball.ang;
    o[0] = bp.x * s * px;
    o[1] = bp.y * s * py;
    o[2] = bp.z * pz;
    o[3] = bv.x * s * vc;
    o[4] = bv.y * s * vc;
    o[5] = bv.z * vc;
    o[6] = ba.x * s * ac;
    o[7] = ba.y * s * ac;
    o[8] = ba.z * ac;
    for (int k = 0; k < kActionFloats; k++) o[9 + k] = prev[self][k];
    for (int i = 0; i < kNumPads; i++) {
        int slot = inv ? E.pad_map[kNumPads - 1 - i] : E.pad_map[i];
        o[9 + kActionFloats + i] = A.pad[slot].is_active ? 1.f : 0.f;
    }
    int w = kObsHead;
    obs_player<MAXC>(me, inv, o + w);
    w += kPlayerObs;
    // Teammates, then opponents (DefaultObs.cpp:36-50), each in the order of
    // GameState::players, which follows Arena::_cars: last-added car first
    // (GameState.cpp:68-78), i.e. descending slot.
    for (int pass = 0; pass < 2; pass++) {
        for (int j = A.ncars - 1; j >= 0; j--) {
            if (j == self) continue;
            bool mate = A.car[j].team == me.team;
            if (mate != (pass == 0)) continue;
            obs_player<MAXC>(A.car[j], inv, o + w);
            w += kPlayerObs;
        }
    }
}

// ---- AdvancedObs ----------------------------------------------------------
// The ContinuousV2 + attention observation of the reference's GPU training
// path (RocketSimCuda/src/TrainingObs.cuh:20-127, consumed through
// buildAdvancedObsAndDefaultMasksKernel, RocketSimCudaObs.cuh:13-64), whose
// fields are GigaLearn's AdvancedObs (AdvancedObs.cpp:4-65): per row the ball
// (position / 5000, velocity / 2300, angular velocity / 3, orange viewers
// mirrored in x and y), the car's last controls, per pad 1 when active else
// 1 / (1 + cooldown) (same pad order and mirror as DefaultObs), then 29 floats
// for the car itself, its teammates and its opponents, each in slot order.
// Where the CPU AdvancedObs differs (its pad timers come from the mirrored
// pad, GameState.cpp:64-66, and it lists other players in Arena::_cars
// order), the GPU path's layout is the one restated.
RSX_HD V3 obs_inv(V3 v, bool inv) { return inv ? V3{-v.x, -v.y, v.z} : v; }

RSX_HD void adv_obs_player(const Car& c, bool inv, V3 ball_pos, V3 ball_vel, float* o) {
    const float pc = 1.f / 5000.f, vc = 1.f / 2300.f, ac = 1.f / 3.f;
    const V3 pos = obs_inv(c.xf.o * kBT2UU, inv), vel = obs_inv(c.vel * kBT2UU, inv), ang = obs_inv(c.ang, inv);
    const V3 f = obs_inv(col(c.xf.b, 0), inv), r = obs_inv(col(c.xf.b, 1), inv), u = obs_inv(col(c.xf.b, 2), inv);
    // rotmat_dot_vec: the vector in the car's (forward, right, up) frame
    auto local = [&](V3 v) { return V3{rl_dot(f, v), rl_dot(r, v), rl_dot(u, v)}; };
    const V3 la = local(ang), lp = local(ball_pos - pos), lv = local(ball_vel - vel);
    const float vals[kAdvPlayerObs] = {pos.x * pc, pos.y * pc, pos.z * pc, f.x, f.y, f.z, u.x, u.y, u.z,
                                       vel.x * vc, vel.y * vc, vel.z * vc, ang.x * ac, ang.y * ac, ang.z * ac,
                                       la.x * ac, la.y * ac, la.z * ac, lp.x * pc, lp.y * pc, lp.z * pc,
                                       lv.x * vc, lv.y * vc, lv.z * vc, c.boost / 100.f, c.on_ground ? 1.f : 0.f,
                                       has_flip_or_jump(c) ? 1.f : 0.f, c.is_demoed ? 1.f : 0.f,
                                       c.has_jumped ? 1.f : 0.f};
    for (int k = 0; k < kAdvPlayerObs; k++) o[k] = vals[k];
}

// AdvancedObs row of car `self`: obs_row_len_of(kObsAdvanced, ncars) floats.
template <int MAXC>
RSX_HD void adv_obs_row(const Arena<MAXC>& A, int self, const EnvConfig& E, float* o) {
    const Car& me = A.car[self];
    const bool inv = me.team != 0;
    const float pc = 1.f / 5000.f, vc = 1.f / 2300.f, ac = 1.f / 3.f;
    const V3 bp = obs_inv(A.ball.xf.o * kBT2UU, inv), bv = obs_inv(A.ball.vel * kBT2UU, inv);
    const V3 ba = obs_inv(A.ball.ang, inv);
    o[0] = bp.x * pc;
    o[1] = bp.y * pc;
    o[2] = bp.z * pc;
    o[3] = bv.x * vc;
    o[4] = bv.y * vc;
    o[5] = bv.z * vc;
    o[6] = ba.x * ac;
    o[7] = ba.y * ac;
    o[8] = ba.z * ac;
    const Controls& lc = me.last_controls;
    o[9] = lc.throttle;
    o[10] = lc.steer;
    o[11] = lc.pitch;
    o[12] = lc.yaw;
    o[13] = lc.roll;
    o[14] = lc.jump ? 1.f : 0.f;
    o[15] = lc.boost ? 1.f : 0.f;
    o[16] = lc.handbrake ? 1.f : 0.f;
    for (int i = 0; i < kNumPads; i++) {
        const Pad& pad = A.pad[inv ? E.pad_map[kNumPads - 1 - i] : E.pad_map[i]];
        o[9 + kActionFloats + i] = pad.is_active ? 1.f : (1.f / (1.f + pad.cooldown));
    }
    int w = kObsHead;
    adv_obs_player(me, inv, bp, bv, o + w);
    w += kAdvPlayerObs;
    for (int pass = 0; pass < 2; pass++)  // teammates, then opponents, in slot order
        for (int j = 0; j < A.ncars; j++) {
            if (j == self || (A.car[j].team == me.team) != (pass == 0)) continue;
            adv_obs_player(A.car[j], inv, bp, bv, o + w);
            w += kAdvPlayerObs;
        }
}

// The configured observation row of car `self`; prev is DefaultObs's
// previous-action block (AdvancedObs reads the car's last controls instead).
// VantaObs rows are written by their own kernel after the step (they carry
// per-arena state and a ball-only rollout), so nothing is written here.
template <int MAXC>
RSX_HD void env_obs_row(const Arena<MAXC>& A, int self, const float (*prev)[kActionFloats], const EnvConfig& E,
                        float* o) {
    if (E.obs_kind == kObsAdvanced) adv_obs_row<MAXC>(A, self, E, o);
    else if (E.obs_kind == kObsDefault) obs_row<MAXC>(A, self, prev, E, o);
}

RSX_HD bool touched_in_step(const Car& c, uint64_t tick, int tick_skip) {
    return c.ball_hit_valid && c.ball_hit_tick >= tick - (uint64_t)tick_skip;
}

// Weighted sum of the four stub rewards for car i (reward-list order: goal,
// touch, velocity ball->goal, velocity player->ball).
template <int MAXC>
RSX_HD float stub_reward(const Arena<MAXC>& A, int i, bool goal, bool touched, const EnvConfig& E) {
    const Car& c = A.car[i];
    V3 bp = A.ball.xf.o * kBT2UU, bv = A.ball.vel * kBT2UU;
    V3 cp = c.xf.o * kBT2UU, cv = c.vel * kBT2UU;
    float r = 0.f;
    if (E.goal_w != 0.f) {
        float g = 0.f;
        if (goal) {
            int ball_side = bp.y < 0 ? 0 : 1;  // RS_TEAM_FROM_Y
            g = (c.team != ball_side) ? 1.f : E.concede_scale;
        }
        r += g * E.goal_w;
    }
    if (E.touch_w != 0.f) r += (touched ? 1.f : 0.f) * E.touch_w;
    if (E.vbtg_w != 0.f) {
        V3 target = c.team == 0 ? V3{0, kBackNetY, kGoalHeightRL / 2} : V3{0, -kBackNetY, kGoalHeightRL / 2};
        V3 dir = rl_norm(target - bp);
        V3 nv = V3{bv.x / kRLBallMaxSpeed, bv.y / kRLBallMaxSpeed, bv.z / kRLBallMaxSpeed};
        r += rl_dot(dir, nv) * E.vbtg_w;
    }
    if (E.vptb_w != 0.f) {
        V3 dir = rl_norm(bp - cp);
        V3 nv = V3{cv.x / kRLCarMaxSpeed, cv.y / kRLCarMaxSpeed, cv.z / kRLCarMaxSpeed};
        r += rl_dot(dir, nv) * E.vptb_w;
    }
    return r;
}

// Terminal type after a step of `step_dt` seconds; updates the no-touch timer.
RSX_HD uint8_t step_terminal(bool goal, bool any_touch, float step_dt, float& no_touch, const EnvConfig& E) {
    uint8_t t = kNotTerminal;
    if (E.goal_terminal && goal) t = kTerminalNormal;
    if (E.no_touch_timeout > 0.f) {
        if (any_touch) {
            no_touch = 0.f;
        } else {
            no_touch += step_dt;
            if (no_touch >= E.no_touch_timeout && t == kNotTerminal) t = kTerminalTruncated;
        }
    }
    return t;
}

}  // namespace rsx
