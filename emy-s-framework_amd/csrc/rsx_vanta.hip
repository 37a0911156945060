// rsx_vanta.hip — the VantaObs kernels (BuildVantaObs, RocketSimCudaApi.cuh:
// 597-640): per arena the ball-prediction slices (k_vanta_pred), then the
// pad and demo timers, every car's coyote timers and its 366-float row
// (k_vanta).
//
// One lane per arena, 64-lane workgroups.  The arena and its Vanta words are
// read field-major (coalesced over the wavefront), the 360-tick ball-only
// rollout runs in the lane with the engine's own tick compiled for an arena
// without cars, and each row is written by the lane that owns the arena.
// The reference splits the rows over a player kernel
// (RocketSimCudaObs.cuh:125-186); one lane per arena keeps the per-arena
// state in registers between the arena update and the rows.
#include "rsx_kern_body.h"

namespace rsx {

// The tick context of a ball-only arena: no cars, so no simplex work, no
// pooled queries and no cull.
__device__ inline void ball_only_ctx(TickCtx& T, const VantaArgs& V, gjk::Simplex* sx) {
    T.cg = V.cg;
    T.bg = V.bg;
    T.mv = V.mv;
    T.pads = V.pads;
    T.dt = 1.f / kTickRate;
    T.sx = sx;
    T.cull = sdf::Cull{};
    T.lds_act = nullptr;
    T.own = true;
    T.pool = nullptr;
}

// The ball-prediction rollouts (vantaArenaUpdateKernel's first half,
// RocketSimCudaObs.cuh:92-121): one lane per arena clears the Vanta words of
// an arena that was reset, reads its ball from the arena words, runs the
// 360-tick ball-only rollout and stores the three slices.  A kernel of its
// own: the ball-only tick needs far fewer registers and far less private
// memory than the rows, so more wavefronts are in flight to hide its
// latency.
template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta_pred(BucketArgs B, uint32_t* vwords, const uint8_t* reset, VantaArgs V) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= B.n) return;
    const Col col{B.words + a, B.n};
    const Col vc{vwords + a, B.n};
    if (reset && reset[B.global[a]]) vanta::clear<MAXC>(vc);
    Ball ball;
    ball.xf = Xf{m3_identity(), col.getv(Layout<MAXC>::kBallBase + 0)};
    ball.vel = col.getv(Layout<MAXC>::kBallBase + 3);
    ball.ang = col.getv(Layout<MAXC>::kBallBase + 6);
    ball.force = ball.vel_cache = V3{0, 0, 0};
    ball.interp = ball.xf;
    ball.sleeping = false;
    ball.bp_reg = -1;
    gjk::Simplex sx;
    TickCtx T;
    ball_only_ctx(T, V, &sx);
    TickScratch<0> W;
    W.cull_age = 1 << 30;
    W.cull_counts[0] = W.cull_counts[1] = W.cull_counts[2] = 0;
    W.L.overflow = false;
    Arena<0> P;
    V3 pos[3], vel[3];
    vanta::predict_ball(ball, T, W, P, pos, vel);
    for (int k = 0; k < 3; k++) {
        vc.setv(vanta::kVaPredPos + 3 * k, pos[k]);
        vc.setv(vanta::kVaPredVel + 3 * k, vel[k]);
    }
}

// The rows: the arena update (pads, demo timers), then every car's coyote
// timers and row, reading the slices k_vanta_pred stored when V.Q.predict.
template <int MAXC>
__global__ void __launch_bounds__(64) k_vanta(BucketArgs B, uint32_t* vwords, const uint8_t* reset, VantaArgs V) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= B.n) return;
    const Col col{B.words + a, B.n};
    const Col vc{vwords + a, B.n};
    if (reset && reset[B.global[a]]) vanta::clear<MAXC>(vc);
    Arena<MAXC> A;
    unpack_arena<MAXC>(col, A, V.cg);
    vanta::State<MAXC> S;
    vanta::load<MAXC>(vc, S);
    S.has_pred = V.Q.predict != 0;
    vanta::update_arena<MAXC>(S, A, V.P, V.Q.dt);
    const uint32_t lt = vanta::last_touch_id<MAXC>(A);
    float* o = V.obs + B.obs_off[a];
    for (int i = 0; i < A.ncars; i++) vanta::build_row<MAXC>(A, S, V.P, V.Q, i, lt, o + (int64_t)i * vanta::kRow);
    vanta::store<MAXC>(vc, S, A.ncars);
}

template <int MAXC>
__global__ void __launch_bounds__(256) k_vanta_clear(uint32_t* vwords, int n, int only) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n || (only >= 0 && a != only)) return;
    vanta::clear<MAXC>(Col{vwords + a, n});
}

hipError_t launch_vanta(int maxc, const BucketArgs& b, uint32_t* vwords, const uint8_t* reset, const VantaArgs& v,
                        hipStream_t st) {
    if (!b.n) return hipSuccess;
    dim3 grid((b.n + 63) / 64), block(64);
    if (v.Q.predict) {
        // the rollouts clear the reset arenas' words first
        if (maxc == 2) hipLaunchKernelGGL(k_vanta_pred<2>, grid, block, 0, st, b, vwords, reset, v);
        else if (maxc == 4) hipLaunchKernelGGL(k_vanta_pred<4>, grid, block, 0, st, b, vwords, reset, v);
        else hipLaunchKernelGGL(k_vanta_pred<6>, grid, block, 0, st, b, vwords, reset, v);
        const hipError_t r = hipGetLastError();
        if (r != hipSuccess) return r;
        reset = nullptr;
    }
    if (maxc == 2) hipLaunchKernelGGL(k_vanta<2>, grid, block, 0, st, b, vwords, reset, v);
    else if (maxc == 4) hipLaunchKernelGGL(k_vanta<4>, grid, block, 0, st, b, vwords, reset, v);
    else hipLaunchKernelGGL(k_vanta<6>, grid, block, 0, st, b, vwords, reset, v);
    return hipGetLastError();
}

hipError_t launch_vanta_clear(int maxc, uint32_t* vwords, int n, int only, hipStream_t st) {
    if (!n) return hipSuccess;
    dim3 grid((n + 255) / 256), block(256);
    if (maxc == 2) hipLaunchKernelGGL(k_vanta_clear<2>, grid, block, 0, st, vwords, n, only);
    else if (maxc == 4) hipLaunchKernelGGL(k_vanta_clear<4>, grid, block, 0, st, vwords, n, only);
    else hipLaunchKernelGGL(k_vanta_clear<6>, grid, block, 0, st, vwords, n, only);
    return hipGetLastError();
}

}  // namespace rsx
