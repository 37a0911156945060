// rsx_cpucheck.cpp — TEST INFRASTRUCTURE: a host (CPU) build of the engine's
// per-arena tick code (the same headers the gfx950 kernels compile), exposed as
// plain C so the CPU-side tests can compare it with the oracle without a GPU.
// Nothing in the engine's product path loads this library.
#include "../../emy-s-framework_amd/csrc/rsx_host.h"
#include "../../emy-s-framework_amd/csrc/rsx_soa.h"
#include "../../emy-s-framework_amd/csrc/rsx_train.h"
#include "../../emy-s-framework_amd/csrc/rsx_vanta.h"
#include "../../include/rsx.h"

#include <cstring>
#include <string>

using namespace rsx;

struct CcWorld {
    HostWorld w;
    MeshView mv;
};

struct CcArena {
    CcWorld* world;
    Arena<6> A;
    TickScratch<6> W;
    int last_scored;
    // when set, the solver runs its iterations on this buffer standing in
    // for the lane's LDS column (solver_iterate_lds), as the GPU kernels do
    std::vector<float> slv;
};

extern "C" {

void* cc_world_new(const char* dir) {
    auto* w = new CcWorld();
    std::string err;
    if (!load_world(dir, w->w, err)) {
        delete w;
        return nullptr;
    }
    w->mv = w->w.view();
    return w;
}

void cc_world_free(void* p) { delete (CcWorld*)p; }

void cc_mesh_hashes(void* p, uint32_t* out16) {
    auto* w = (CcWorld*)p;
    for (int i = 0; i < 16; i++) out16[i] = w->w.mesh_hash[i];
}

void* cc_arena_new(void* world, int n_blue, int n_orange) {
    auto* a = new CcArena();
    memset(&a->A, 0, sizeof(a->A));
    a->world = (CcWorld*)world;
    a->A.ncars = n_blue + n_orange;
    a->A.rng = 0x9E3779B9u;
    for (int i = 0; i < a->A.ncars; i++) {
        Car& c = a->A.car[i];
        c.team = i < n_blue ? 0 : 1;
        c.id = (uint32_t)(i + 1);
        c.active = true;
        c.bp_reg = -1;  // registered where a state set before the first tick cannot see it
        car_set_spawn<6>(c, a->world->w.cg, 0, 0, kCarSpawnRestZ, 0, kBoostSpawnAmount);
    }
    a->A.ball.xf = xf_identity();
    a->A.ball.xf.o = V3{0, 0, kBallRestZ * kUU2BT};
    a->A.ball.bp_reg = -1;
    for (int p = 0; p < kNumPads; p++) a->A.pad[p] = Pad{true, 0.f, 0, 0};
    return a;
}

void cc_arena_free(void* p) { delete (CcArena*)p; }

void cc_set_car(void* p, int i, const rsx_car_state* s) {
    auto* a = (CcArena*)p;
    Car& c = a->A.car[i];
    uint32_t id = c.id;
    int team = c.team;
    car_from_public(c, *s, a->world->w.cg);
    c.id = id;
    c.team = team;
}

void cc_get_car(void* p, int i, rsx_car_state* s) {
    auto* a = (CcArena*)p;
    car_to_public(a->A.car[i], *s);
}

void cc_set_ball(void* p, const rsx_ball_state* s) { ball_from_public(((CcArena*)p)->A.ball, *s); }
void cc_get_ball(void* p, rsx_ball_state* s) { ball_to_public(((CcArena*)p)->A.ball, *s); }

void cc_set_pads(void* p, const rsx_pad_state* s) {
    auto* a = (CcArena*)p;
    for (int i = 0; i < kNumPads; i++) a->A.pad[i] = Pad{s[i].is_active != 0, s[i].cooldown, s[i].prev_locked_car_id, 0};
}
void cc_get_pads(void* p, rsx_pad_state* s) {
    auto* a = (CcArena*)p;
    for (int i = 0; i < kNumPads; i++) {
        memset(&s[i], 0, sizeof(s[i]));
        s[i].is_active = a->A.pad[i].is_active;
        s[i].cooldown = a->A.pad[i].cooldown;
        s[i].prev_locked_car_id = a->A.pad[i].prev_locked;
    }
}

void cc_set_controls(void* p, int i, const rsx_controls* c) {
    ((CcArena*)p)->A.car[i].controls = controls_from_public(*c);
}

void cc_set_tick(void* p, uint64_t t) { ((CcArena*)p)->A.tick = t; }
uint64_t cc_get_tick(void* p) { return ((CcArena*)p)->A.tick; }

// Per-wheel persistent values: steer, engine, brake, lat, long, extra pushback.
void cc_set_wheels(void* p, int i, const float* in) {
    auto* a = (CcArena*)p;
    for (int w = 0; w < 4; w++) {
        Wheel& wh = a->A.car[i].wheel[w];
        wh.steer_angle = in[w * 6 + 0];
        wh.engine_force = in[w * 6 + 1];
        wh.brake = in[w * 6 + 2];
        wh.lat_friction = in[w * 6 + 3];
        wh.long_friction = in[w * 6 + 4];
        wh.extra_pushback = in[w * 6 + 5];
    }
}
void cc_get_wheels(void* p, int i, float* out) {
    auto* a = (CcArena*)p;
    for (int w = 0; w < 4; w++) {
        const Wheel& wh = a->A.car[i].wheel[w];
        float* o = out + w * 12;
        o[0] = wh.steer_angle; o[1] = wh.engine_force; o[2] = wh.brake; o[3] = wh.lat_friction;
        o[4] = wh.long_friction; o[5] = wh.susp_len; o[6] = wh.susp_rel_vel; o[7] = wh.clipped_inv_dot;
        o[8] = wh.extra_pushback; o[9] = wh.in_contact; o[10] = wh.susp_force; o[11] = wh.in_contact_world;
    }
}

int cc_step(void* p, int ticks) {
    auto* a = (CcArena*)p;
    TickCtx T;
    T.cg = a->world->w.cg;
    T.bg = a->world->w.bg;
    T.mv = a->world->mv;
    T.pads = a->world->w.pads;
    T.dt = 1.f / kTickRate;
    gjk::Simplex sx;
    T.sx = &sx;
    T.cull = sdf::Cull{};
    T.lds_act = nullptr;
    T.own = true;
    T.pool = nullptr;
    T.slv = a->slv.empty() ? nullptr : a->slv.data();
    int scored = 0;
    for (int t = 0; t < ticks; t++) scored = arena_tick(a->A, T, a->W) ? 1 : 0;
    a->last_scored = scored;
    return scored;
}

// Stores the arena into the engine's SoA column layout and loads it back, as
// a launch boundary of the GPU kernels does.
void cc_soa_roundtrip(void* p, int fill) {
    auto* a = (CcArena*)p;
    std::vector<uint32_t> col(Layout<6>::kWords, 0);
    Col c{col.data(), 1};
    c.set(Layout<6>::kArenaBase + kArNcars, (uint32_t)a->A.ncars);
    pack_arena<6>(c, a->A);
    Controls keep[6];
    for (int i = 0; i < a->A.ncars; i++) keep[i] = a->A.car[i].controls;
    // fill: seed of the pseudo-random bytes written over the working state and
    // the tick scratch before the load, so that anything read before it is
    // written shows up.
    uint32_t r = (uint32_t)fill * 2654435761u + 12345u;
    auto scribble = [&](void* dst, size_t n) {
        uint8_t* b = (uint8_t*)dst;
        for (size_t i = 0; i < n; i++) {
            r ^= r << 13; r ^= r >> 17; r ^= r << 5;
            b[i] = (uint8_t)r;
        }
    };
    scribble(&a->A, sizeof(a->A));
    scribble(&a->W, sizeof(a->W));
    // flags the load must restore, set to the value that breaks a missing load
    for (int i = 0; i < 6; i++) a->A.car[i].active = (fill >> i) & 1;
    unpack_arena<6>(c, a->A, a->world->w.cg);
    for (int i = 0; i < a->A.ncars; i++) a->A.car[i].controls = keep[i];
}

// Manifolds with contacts from the last tick, same record layout as the
// oracle's ref_dump_contacts (body codes: static s -> s, ball 100, car i 200+i).
int cc_dump_contacts(void* p, float* out, int max_floats) {
    auto* a = (CcArena*)p;
    const MfList<6>& L = a->W.L;
    int n = 0;
    auto code = [](int b) { return b >= kBodyStatic ? b - kBodyStatic : (b == 0 ? 100 : 200 + b - 1); };
    for (int k = 0; k < L.count; k++) {
        int s = L.slot[k];
        if (s < 0) continue;
        const Manifold& m = L.live[s];
        if (n + 4 + m.n * 22 > max_floats) break;
        out[n++] = (float)code(m.body0);
        out[n++] = (float)code(m.body1);
        out[n++] = (float)m.n;
        out[n++] = m.breaking;
        for (int i = 0; i < m.n; i++) {
            const Contact& c = m.pt[i];
            const V3 vals[] = {c.posA, c.posB, c.normal};
            for (auto& v : vals) { out[n++] = v.x; out[n++] = v.y; out[n++] = v.z; }
            out[n++] = c.dist; out[n++] = c.friction; out[n++] = c.restitution; out[n++] = c.special;
            out[n++] = 0.f;
            const V3 loc[] = {c.localA, c.localB};
            for (auto& v : loc) { out[n++] = v.x; out[n++] = v.y; out[n++] = v.z; }
            out[n++] = 0.f; out[n++] = 0.f;
        }
    }
    return n;
}

}  // extern "C"

extern "C" void cc_raytest_static(void* w, const float* a, const float* b, float* out) {
    auto* W = (CcWorld*)w;
    RayHit hit;
    hit.object = -1;
    hit.fraction = 1.f;
    ray_static_world(W->mv, V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}, hit);
    out[0] = hit.object >= 0;
    out[1] = hit.fraction;
    out[5] = hit.normal.x; out[6] = hit.normal.y; out[7] = hit.normal.z;
}

// The engine's triangle table: per triangle 14 floats, the three vertices,
// the internal-edge angles and flags (bit pattern), and the mesh index.
extern "C" int cc_tri_info(void* w, float* out, int max_tris) {
    auto* W = (CcWorld*)w;
    const HostWorld& h = W->w;
    int n = (int)h.tri_mesh.size();
    for (int t = 0; t < n && t < max_tris; t++) {
        float* o = out + (size_t)t * 14;
        memcpy(o, &h.tri[(size_t)t * 9], 9 * sizeof(float));
        memcpy(o + 9, &h.tri_info[(size_t)t * 4], 4 * sizeof(float));
        o[13] = (float)h.tri_mesh[t];
    }
    return n;
}

// Squared point-triangle distance of the learned cull's exact query, for n
// (point, triangle) pairs: pts 3 floats each, tris 9 floats each.
extern "C" void cc_point_tri_dist2(const float* pts, const float* tris, int n, float* out) {
    for (int i = 0; i < n; i++) {
        const float* p = pts + 3 * i;
        const float* t = tris + 9 * i;
        out[i] = sdf::point_tri_dist2(V3{p[0], p[1], p[2]}, V3{t[0], t[1], t[2]}, V3{t[3], t[4], t[5]},
                                      V3{t[6], t[7], t[8]});
    }
}

// Ray test of one convex body, as the suspension ray test does it for the
// ball and the car boxes (same arguments and outputs as the oracle's
// ref_ray_convex).
extern "C" void cc_ray_convex(int kind, const float* dims, const float* xf12, const float* a, const float* b,
                              float* out) {
    gjk::RayShape s;
    s.kind = kind;
    s.h = kind == 0 ? V3{dims[0], dims[1], dims[2]} + V3{dims[3], dims[3], dims[3]} : V3{dims[0], 0, 0};
    s.xf.b = m3(xf12[0], xf12[1], xf12[2], xf12[3], xf12[4], xf12[5], xf12[6], xf12[7], xf12[8]);
    s.xf.o = V3{xf12[9], xf12[10], xf12[11]};
    float f;
    V3 n;
    gjk::Simplex sx;
    out[0] = 0.f;
    out[1] = 1.f;
    if (gjk::ray_cast(s, V3{a[0], a[1], a[2]}, V3{b[0], b[1], b[2]}, f, n, sx) && len2(n) > 0.0001f && f < 1.0f) {
        n = normalized(n);
        out[0] = 1.f;
        out[1] = f;
        out[2] = n.x; out[3] = n.y; out[4] = n.z;
    }
}

// Env outputs of the current state after a step of tick_skip ticks, with the
// default stub-reward weights (1, concede -1, 1, 1, 1): each car's DefaultObs
// row into obs (rows back to back), rewards, and the terminal type (return
// value) with the no-touch timer carried through *no_touch.
extern "C" int cc_env_kind(void* p, const float* prev, int tick_skip, float timeout, float* no_touch, float* obs,
                           float* rew, int obs_kind) {
    auto* a = (CcArena*)p;
    EnvConfig E;
    E.goal_w = 1.f;
    E.concede_scale = -1.f;
    E.touch_w = 1.f;
    E.vbtg_w = 1.f;
    E.vptb_w = 1.f;
    E.no_touch_timeout = timeout;
    E.goal_terminal = 1;
    E.auto_reset = 0;
    E.obs_kind = obs_kind;
    if (!obs_pad_map(a->world->w.pads, E.pad_map)) return -1;
    const Arena<6>& A = a->A;
    float pa[6][kActionFloats];
    for (int i = 0; i < A.ncars; i++)
        for (int k = 0; k < kActionFloats; k++) pa[i][k] = prev[i * kActionFloats + k];
    bool goal = ball_scored_now(A.ball);
    bool any = false;
    bool touched[6];
    for (int i = 0; i < A.ncars; i++) {
        touched[i] = touched_in_step(A.car[i], A.tick, tick_skip);
        any |= touched[i];
    }
    uint8_t t = step_terminal(goal, any, (float)tick_skip * (1.f / kTickRate), *no_touch, E);
    for (int i = 0; i < A.ncars; i++) rew[i] = stub_reward<6>(A, i, goal, touched[i], E);
    int row = obs_row_len_of(obs_kind, A.ncars);
    for (int i = 0; i < A.ncars; i++) env_obs_row<6>(A, i, pa, E, obs + i * row);
    return t;
}

// cc_env with DefaultObs rows.
extern "C" int cc_env(void* p, const float* prev, int tick_skip, float timeout, float* no_touch, float* obs,
                      float* rew) {
    return cc_env_kind(p, prev, tick_skip, timeout, no_touch, obs, rew, kObsDefault);
}

// Rank of every triangle within its mesh's BVH traversal order (the engine's
// reconstruction of btQuantizedBvh's leaf order); returns the triangle count.
extern "C" int cc_tri_rank(void* w, int32_t* out, int cap) {
    const HostWorld& h = ((CcWorld*)w)->w;
    int n = (int)h.tri_rank.size();
    for (int t = 0; t < n && t < cap; t++) out[t] = h.tri_rank[t];
    return n;
}

// Whether any tick since the arena was created ran out of contact-manifold
// storage (MfList::overflow).
extern "C" int cc_contact_overflow(void* p) { return ((CcArena*)p)->W.L.overflow ? 1 : 0; }

// The training rewards of one arena (rsx_train.h train_rewards) from public
// records: n cars after the step and at its start, the ball after and at the
// start, the reward table, the episode's scoring team (-1 none), the tick
// count after the step and the step's ticks.  out: n rewards.
extern "C" void cc_train_rewards(void* w, const rsx_car_state* cars, const rsx_car_state* prev, int n,
                                 const rsx_ball_state* ball, const rsx_ball_state* prev_ball,
                                 const rsx_training_reward_entry* entries, int ne, int goal_team, uint64_t tick,
                                 int step_ticks, float* out) {
    const CarGeom& g = ((CcWorld*)w)->w.cg;
    auto load = [&](Arena<6>& A, const rsx_car_state* cs, const rsx_ball_state* b) {
        memset(&A, 0, sizeof(A));
        A.ncars = n;
        for (int i = 0; i < n; i++) {
            car_from_public(A.car[i], cs[i], g);
            A.car[i].id = cs[i].id;
            A.car[i].team = cs[i].team;
        }
        ball_from_public(A.ball, *b);
        A.tick = tick;
    };
    static Arena<6> A, B;
    load(A, cars, ball);
    load(B, prev, prev_ball);
    PrevSnap<6> P;
    snapshot_prev<6>(B, P);
    RewardEntry t[kMaxRewardEntries];
    for (int k = 0; k < ne && k < kMaxRewardEntries; k++) {
        t[k].id = entries[k].id;
        t[k].weight = entries[k].weight;
        for (int q = 0; q < kRewardParams; q++) t[k].params[q] = entries[k].params[q];
        t[k].zero_sum = entries[k].is_zero_sum;
        t[k].team_spirit = entries[k].team_spirit;
        t[k].opp_scale = entries[k].opponent_scale;
    }
    train_rewards<6>(A, P, t, ne < kMaxRewardEntries ? ne : kMaxRewardEntries, goal_team, tick, step_ticks, out);
}

// One curriculum reset of a fresh arena (rsx_train.h arena_reset_curriculum)
// with the given RNG seed and weights; returns the mode and writes the cars'
// and the ball's public records.
extern "C" int cc_curriculum_reset(void* w, int n_blue, int n_orange, uint32_t seed, const float* weights,
                                   rsx_car_state* cars, rsx_ball_state* ball) {
    auto* a = (CcArena*)cc_arena_new(w, n_blue, n_orange);
    a->A.rng = seed ? seed : 1u;
    const int mode = arena_reset_curriculum<6>(a->A, a->world->w.cg, weights);
    for (int i = 0; i < a->A.ncars; i++) car_to_public(a->A.car[i], cars[i]);
    ball_to_public(a->A.ball, *ball);
    cc_arena_free(a);
    return mode;
}

// ---- VantaObs (rsx_vanta.h) on the host ----------------------------------
// A Vanta state is one arena's Vanta words (6 car slots, stride 1), all zero
// when new or cleared.
extern "C" void* cc_vanta_new() { return new std::vector<uint32_t>(vanta::VLayout<6>::kWords, 0u); }
extern "C" void cc_vanta_free(void* v) { delete (std::vector<uint32_t>*)v; }
extern "C" void cc_vanta_clear(void* v) { vanta::clear<6>(Col{((std::vector<uint32_t>*)v)->data(), 1}); }

// The ball-only rollout of the arena's ball: 3 slices of (pos, vel), uu.
extern "C" void cc_vanta_predict(void* p, float* out18) {
    auto* a = (CcArena*)p;
    TickCtx T;
    T.cg = a->world->w.cg;
    T.bg = a->world->w.bg;
    T.mv = a->world->mv;
    T.pads = a->world->w.pads;
    T.dt = 1.f / kTickRate;
    gjk::Simplex sx;
    T.sx = &sx;
    T.cull = sdf::Cull{};
    T.lds_act = nullptr;
    T.own = true;
    T.pool = nullptr;
    auto* W = new TickScratch<0>();
    W->cull_age = 1 << 30;
    W->L.overflow = false;
    auto* P = new Arena<0>();
    V3 pos[3], vel[3];
    vanta::predict_ball(a->A.ball, T, *W, *P, pos, vel);
    for (int k = 0; k < 3; k++) {
        out18[6 * k + 0] = pos[k].x;
        out18[6 * k + 1] = pos[k].y;
        out18[6 * k + 2] = pos[k].z;
        out18[6 * k + 3] = vel[k].x;
        out18[6 * k + 4] = vel[k].y;
        out18[6 * k + 5] = vel[k].z;
    }
    delete P;
    delete W;
}

// BuildVantaObs for the arena with its Vanta state: pred18 (null: no
// prediction) holds the slices; rows of every car to obs (ncars x 366).
extern "C" void cc_vanta_build(void* p, void* v, const float* pred18, float dt, float deadzone, int only_closest,
                               float* obs) {
    auto* a = (CcArena*)p;
    const Col c{((std::vector<uint32_t>*)v)->data(), 1};
    EnvConfig E;
    obs_pad_map(a->world->w.pads, E.pad_map);
    vanta::PadTables P{};
    for (int i = 0; i < vanta::kPads; i++) {
        P.map[i] = E.pad_map[i];
        P.loc[i] = a->world->w.pads[E.pad_map[i]].pos_uu;
    }
    for (int i = 0; i < vanta::kPads; i++) {
        int m = i;
        for (int j = 0; j < vanta::kPads; j++)
            if (fabsf(P.loc[j].x + P.loc[i].x) < 1.f && fabsf(P.loc[j].y - P.loc[i].y) < 1.f &&
                fabsf(P.loc[j].z - P.loc[i].z) < 1.f) {
                m = j;
                break;
            }
        P.mirror[i] = (int8_t)m;
    }
    vanta::Params Q{dt, deadzone, only_closest, pred18 != nullptr, 8};
    vanta::State<6> S;
    vanta::load<6>(c, S);
    if (pred18) {
        for (int k = 0; k < 3; k++) {
            S.pred_pos[k] = V3{pred18[6 * k], pred18[6 * k + 1], pred18[6 * k + 2]};
            S.pred_vel[k] = V3{pred18[6 * k + 3], pred18[6 * k + 4], pred18[6 * k + 5]};
        }
        S.has_pred = true;
    } else {
        S.has_pred = false;
    }
    vanta::update_arena<6>(S, a->A, P, dt);
    const uint32_t lt = vanta::last_touch_id<6>(a->A);
    for (int i = 0; i < a->A.ncars; i++) vanta::build_row<6>(a->A, S, P, Q, i, lt, obs + (int64_t)i * vanta::kRow);
    vanta::store<6>(c, S, a->A.ncars);
}

// The rollout continuation against new rollouts: a full rollout from the
// arena's ball keeps its continuation (ts ticks between builds); then, `steps`
// times, the rollout from the ball it passed at ts ticks is made both by
// continuation and anew.  Returns the number of steps whose slices and new
// continuation states agree bit for bit.
extern "C" int cc_vanta_continue_check(void* p, int ts, int steps) {
    auto* a = (CcArena*)p;
    TickCtx T;
    T.cg = a->world->w.cg;
    T.bg = a->world->w.bg;
    T.mv = a->world->mv;
    T.pads = a->world->w.pads;
    T.dt = 1.f / kTickRate;
    gjk::Simplex sx;
    T.sx = &sx;
    T.cull = sdf::Cull{};
    T.lds_act = nullptr;
    T.own = true;
    T.pool = nullptr;
    auto* W = new TickScratch<0>();
    W->cull_age = 1 << 30;
    W->L.overflow = false;
    auto* P = new Arena<0>();
    vanta::Roller R{T, *W, *P};
    V3 pos[3], vel[3];
    vanta::BallQ chain[4];
    vanta::rollout_full(R, vanta::ball_q(a->A.ball), ts, pos, vel, chain);
    int ok = 0;
    for (int s = 0; s < steps; s++) {
        const vanta::BallQ next = chain[0];
        V3 pc[3], vc[3], pf[3], vf[3];
        vanta::BallQ cc[4] = {chain[0], chain[1], chain[2], chain[3]}, cf[4];
        vanta::rollout_continue(R, ts, pc, vc, cc);
        vanta::rollout_full(R, next, ts, pf, vf, cf);
        bool same = true;
        for (int k = 0; k < 3; k++)
            same = same && vanta::same_ball(vanta::BallQ{pc[k], vc[k], V3{0, 0, 0}}, vanta::BallQ{pf[k], vf[k], V3{0, 0, 0}});
        for (int k = 0; k < 4; k++) same = same && vanta::same_ball(cc[k], cf[k]);
        if (!same) break;
        ok++;
        for (int k = 0; k < 4; k++) chain[k] = cc[k];
    }
    delete P;
    delete W;
    return ok;
}

// Solver iterations on an LDS-like column (1) or on the private bodies (0).
extern "C" void cc_set_solver_lds(void* p, int on) {
    auto* a = (CcArena*)p;
    if (on) a->slv.assign(solver_lds_bytes<6>() / sizeof(float), 0.f);
    else a->slv.clear();
}
