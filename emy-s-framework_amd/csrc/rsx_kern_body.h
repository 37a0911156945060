// rsx_kern_body.h — device bodies of the arena kernels (one wavefront's
// share of a launch) and the shape arithmetic shared by their launchers.
// Included by the kernel translation units only.
#pragma once

#include <mutex>

#include "rsx_kern.h"
#include "rsx_soa.h"

namespace rsx {

// A wavefront's LDS: its lanes' simplices, then the pooled queries.
constexpr int kSimplexBytes = (64 * (int)sizeof(gjk::Simplex) + 15) & ~15;
constexpr int kPoolOffset = kSimplexBytes;

// The lane's arena within a wavefront whose first arena is `first`:
// lanes past arenas_per_wave mirror lane (l % arenas_per_wave); lanes past
// the bucket's end run its last arena.  Neither stores.
struct LaneArena {
    int a;
    bool own;
};

__device__ inline LaneArena lane_arena(int first, int n, int apw) {
    const int l = threadIdx.x & 63;
    const int a0 = first + (l & (apw - 1));
    return LaneArena{a0 < n ? a0 : n - 1, l < apw && a0 < n};
}

__device__ inline void tick_ctx(TickCtx& T, const WorldArgs& G, uint8_t* lds_wave, bool own, int pool_off) {
    T.cg = G.cg;
    T.bg = G.bg;
    T.mv = G.mv;
    T.pads = G.pads;
    T.dt = 1.f / kTickRate;
    T.sx = (gjk::Simplex*)lds_wave + (threadIdx.x & 63);
    T.lds_act = (uint16_t*)lds_wave;
    T.cull = G.cull;
    T.own = own;
    T.pool = G.pool ? lds_wave + pool_off : nullptr;
    T.slv = nullptr;
}

// The solver keeps its bodies' hot state in the wavefront's LDS share when
// the share holds it (the simplices and pooled queries are idle meanwhile).
template <int MAXC>
__device__ inline void solver_lds(TickCtx& T, const WorldArgs& G, uint8_t* lds_wave) {
    T.slv = G.lds_bytes >= solver_lds_bytes<MAXC>() ? (float*)lds_wave + (threadIdx.x & 63) : nullptr;
}

// `ticks` ticks of the bucket's arenas under their stored controls
// (RocketSimCudaBatch::Step).  wave: the wavefront's index within the bucket.
template <int MAXC>
__device__ __forceinline__ void ticks_wave(const BucketArgs& B, const WorldArgs& G, int ticks, int wave, int apw,
                                           uint8_t* lds_wave) {
    const int first = wave * apw;
    if (first >= B.n) return;
    const LaneArena la = lane_arena(first, B.n, apw);
    const int a = la.a;
    Col col{B.words + a, B.n};
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, G.cg);
    for (int i = 0; i < A.ncars; i++) {
        const rsx_controls& c = B.ctrl[(int64_t)a * MAXC + i];
        A.car[i].controls = Controls{c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump != 0, c.boost != 0,
                                     c.handbrake != 0};
    }
    TickCtx T;
    tick_ctx(T, G, lds_wave, la.own, kPoolOffset);
    solver_lds<MAXC>(T, G, lds_wave);
    TickScratch<MAXC> W;
    W.cull_age = 1 << 30;  // the first tick of a launch runs the cull's network
    W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
    W.L.overflow = false;
    uint32_t flags = col.get(Layout<MAXC>::kArenaBase + kArFlags);
    for (int t = 0; t < ticks; t++) {
        if (arena_tick(A, T, W) && !(flags & kFlagGoal))
            flags |= kFlagGoal | (A.ball.xf.o.y > 0.f ? 0u : kFlagGoalOrange);
    }
    if (W.L.overflow) flags |= kFlagContactOverflow;
    cull_flush(T, W);
    if (!la.own) return;
    pack_arena<MAXC>(col, A);
    col.set(Layout<MAXC>::kArenaBase + kArFlags, flags);
}

// Fused env step of the bucket's arenas: action indices -> controls, the tick
// skip, then the stub rewards, the terminal conditions, the kickoff
// auto-reset of terminal arenas, the DefaultObs rows and (when requested)
// the DefaultAction mask rows, all in the lane that owns the arena.
template <int MAXC>
__device__ __forceinline__ void env_wave(const BucketArgs& B, const WorldArgs& G, const EnvArgs& V, int wave, int apw,
                                         uint8_t* lds_wave) {
    using L = Layout<MAXC>;
    const int first = wave * apw;
    if (first >= B.n) return;
    const LaneArena la = lane_arena(first, B.n, apw);
    const int a = la.a;
    Col col{B.words + a, B.n};
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, G.cg);
    float prev[MAXC][kActionFloats];
    const int c0 = B.car_off[a];
    for (int i = 0; i < A.ncars; i++) {
        int k = V.act[c0 + i];
        k = k < 0 ? 0 : (k >= V.lut_n ? V.lut_n - 1 : k);
        const rsx_controls c = V.lut[k];
        A.car[i].controls = Controls{c.throttle, c.steer, c.pitch, c.yaw, c.roll, c.jump != 0, c.boost != 0,
                                     c.handbrake != 0};
        prev[i][0] = c.throttle;
        prev[i][1] = c.steer;
        prev[i][2] = c.pitch;
        prev[i][3] = c.yaw;
        prev[i][4] = c.roll;
        prev[i][5] = c.jump ? 1.f : 0.f;
        prev[i][6] = c.boost ? 1.f : 0.f;
        prev[i][7] = c.handbrake ? 1.f : 0.f;
    }
    // the state the training rewards compare against (SnapshotTrainingState)
    PrevSnap<MAXC> P;
    if (V.rt_n > 0) snapshot_prev<MAXC>(A, P);
    TickCtx T;
    tick_ctx(T, G, lds_wave, la.own, kPoolOffset);
    solver_lds<MAXC>(T, G, lds_wave);
    TickScratch<MAXC> W;
    W.cull_age = 1 << 30;  // the first tick of a launch runs the cull's network
    W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
    W.L.overflow = false;
    uint32_t flags = col.get(L::kArenaBase + kArFlags);
    for (int t = 0; t < V.ticks; t++) {
        if (arena_tick(A, T, W) && !(flags & kFlagGoal))
            flags |= kFlagGoal | (A.ball.xf.o.y > 0.f ? 0u : kFlagGoalOrange);
    }
    if (W.L.overflow) flags |= kFlagContactOverflow;
    cull_flush(T, W);
    if (!la.own) return;
    const EnvConfig& E = V.E;
    const bool goal = ball_scored_now(A.ball);
    bool touched[MAXC];
    bool any_touch = false;
    for (int i = 0; i < A.ncars; i++) {
        touched[i] = touched_in_step(A.car[i], A.tick, V.ticks);
        any_touch |= touched[i];
    }
    float no_touch = col.getf(L::kArenaBase + kArNoTouch);
    const uint8_t tt = step_terminal(goal, any_touch, (float)V.ticks * (1.f / kTickRate), no_touch, E);
    if (V.rt_n > 0) {
        // the configured training rewards (BuildRewardsAndTerminals); the
        // goal term reads the episode's goal flag and scoring team
        float r[MAXC];
        const int goal_team = (flags & kFlagGoal) ? ((flags & kFlagGoalOrange) ? 1 : 0) : -1;
        train_rewards<MAXC>(A, P, V.rt, V.rt_n, goal_team, A.tick, V.ticks, r);
        for (int i = 0; i < A.ncars; i++) V.rew[c0 + i] = r[i];
    } else {
        for (int i = 0; i < A.ncars; i++) V.rew[c0 + i] = stub_reward<MAXC>(A, i, goal, touched[i], E);
    }
    uint32_t episode = col.get(L::kArenaBase + kArEpisode);
    if (tt != kNotTerminal && E.auto_reset) {
        arena_reset_curriculum<MAXC>(A, G.cg, E.reset_w);
        A.tick = 0;
        flags &= kFlagContactOverflow;
        no_touch = reset_no_touch(A.rng, E);
        episode++;
        for (int i = 0; i < A.ncars; i++)
            for (int k = 0; k < kActionFloats; k++) prev[i][k] = 0.f;
    }
    const int row = obs_row_len_of(E.obs_kind, A.ncars);
    float* o = V.obs + B.obs_off[a];
    for (int i = 0; i < A.ncars; i++) env_obs_row<MAXC>(A, i, prev, E, o + (int64_t)i * row);
    if (V.masks)
        for (int i = 0; i < A.ncars; i++)
            action_mask_row(A.car[i], V.mask_tables, V.masks + (int64_t)(c0 + i) * kNumDefaultActions);
    pack_arena<MAXC>(col, A);
    col.set(L::kArenaBase + kArFlags, flags);
    col.setf(L::kArenaBase + kArNoTouch, no_touch);
    col.set(L::kArenaBase + kArEpisode, episode);
    V.term[B.global[a]] = tt;
}

// One wavefront of the launch over every team-size bucket: the wavefront's
// index picks its bucket, 3v3 wavefronts first, then 2v2, then 1v1.
__device__ __forceinline__ void mixed_wave(const BucketArgs& b2, const BucketArgs& b4, const BucketArgs& b6, int w6,
                                           int w4, const WorldArgs& G, const EnvArgs& V, int apw, int lds_wave);

// The workgroup's first wavefront index and this wavefront's LDS.
__device__ inline int wave_index() { return blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); }

inline int blocks_for_waves(int waves, const LaunchShape& s) {
    const int wpb = s.block / 64;
    return (waves + wpb - 1) / wpb;
}

// hipFuncSetAttribute once per kernel and size.
template <typename F>
hipError_t set_lds_limit(F kernel, size_t bytes, size_t& done) {
    static std::mutex m;
    std::lock_guard<std::mutex> g(m);
    if (bytes <= done) return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) done = bytes;
    return e;
}

__device__ __forceinline__ void mixed_wave(const BucketArgs& b2, const BucketArgs& b4, const BucketArgs& b6, int w6,
                                           int w4, const WorldArgs& G, const EnvArgs& V, int apw, int lds_wave) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint8_t* mine = lds + (threadIdx.x >> 6) * lds_wave;
    int w = wave_index();
    if (w < w6) {
        env_wave<6>(b6, G, V, w, apw, mine);
        return;
    }
    w -= w6;
    if (w < w4) {
        env_wave<4>(b4, G, V, w, apw, mine);
        return;
    }
    env_wave<2>(b2, G, V, w - w4, apw, mine);
}

}  // namespace rsx
